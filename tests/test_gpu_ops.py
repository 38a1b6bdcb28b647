"""GPU parity of the field operators against the FP64 oracle (C ABI via
maxlq_amd.lq).  FP64 path: 1e-10 relative, most
checks land at ~1e-14.  Mixed path: stated per test."""
import numpy as np
import pytest

from oracle.lq_oracle import Oracle, Params
from oracle.seeded import seeded_field
from tests.fields import abc_field, bandlimit, rel

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _ctx(n, q=4.0, B=1.0, precision="fp64"):
    from maxlq_amd.lq import MaxLq
    return MaxLq(n, q=q, B=B, precision=precision)


@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_fft_r2c_c2r(n):
    g = _ctx(n)
    u = np.random.default_rng(n).standard_normal((n, n, n))
    F = g.fftfwdr(u)
    assert rel(F, np.fft.rfftn(u)) < 1e-14
    assert rel(g.fftbwdr(F), u) < 1e-14


@pytest.mark.parametrize("n", [16, 32])
def test_fft_c2c(n):
    g = _ctx(n)
    rng = np.random.default_rng(1)
    u = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    F = g.fftfwd(u)
    assert rel(F, np.fft.fftn(u)) < 1e-14
    assert rel(g.fftbwd(F), u) < 1e-14


@pytest.mark.parametrize("n", [16, 64])
def test_fft_mixed_precision(n):
    g = _ctx(n, precision="mixed")
    u = np.random.default_rng(3).standard_normal((n, n, n))
    F = g.fftfwdr(u)
    assert rel(F, np.fft.rfftn(u)) < 2e-6
    assert rel(g.fftbwdr(F), u) < 2e-6


def test_spectral_primitives_vs_oracle():
    n = 16
    g = _ctx(n, q=6.0)
    orc = Oracle(Params(n=n, q=6.0))
    rng = np.random.default_rng(4)
    f = rng.standard_normal((n, n, n))
    v = rng.standard_normal((3, n, n, n))
    assert rel(g.gradient(f), orc.gradient(f)) < TOL64
    assert rel(g.divergence(v), orc.divergence(v)) < TOL64
    assert rel(g.laplacian(v), orc.laplacian(v)) < TOL64
    assert rel(g.solve_poisson(f, 1.0), orc.solve_poisson(f, 1.0)) < TOL64
    assert rel(g.solve_poisson(f, 0.5), orc.solve_poisson(f, 0.5)) < TOL64
    for p in (2.0, 1.5, 1.0, 3.0):
        assert rel(g.dealias_scalar(f, p), orc.dealias_scalar(f, p)) < TOL64
    assert rel(g.divAvg_free(v), orc.divAvg_free(v)) < TOL64
    s = 1.5 - 1.0 / 6.0
    assert rel(g.HilbertGradient(v, s), orc.HilbertGradient(v, s)) < TOL64
    m = g.calc0thFourierModes(v)
    assert np.max(np.abs(m - orc.calc0thFourierModes(v))) < 1e-10 * np.sum(np.abs(v))


def test_inner_products_vs_oracle():
    n = 16
    g = _ctx(n, q=9.0)
    orc = Oracle(Params(n=n, q=9.0))
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal((2, 3, n, n, n))
    for t in ("L2", "H_l^(3/2-1/q)", "H^(3/2-1/q)"):
        x = g.global_summed_field_inner_product(a, b, t)
        y = orc.global_summed_field_inner_product(a, b, t)
        assert abs(x - y) < TOL64 * abs(y), t
        x1 = g.inner_product(a[0], b[1], t)
        assert abs(x1 - orc.inner_product(a[0], b[1], t)) < 1e-12 * np.sqrt(
            orc.inner_product(a[0], a[0], t) * orc.inner_product(b[1], b[1], t))
    from maxlq_amd.capi import MlqError, MLQ_EBADKIND
    with pytest.raises(MlqError) as e:
        g.inner_product(a[0], b[0], "H_l^((3q-1)/(2q))")
    assert e.value.code == MLQ_EBADKIND


@pytest.mark.parametrize("k", [0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 4.5, 7.0, -1.0])
def test_calc_uk_vs_oracle(k):
    n = 16
    g = _ctx(n)
    orc = Oracle(Params(n=n))
    u = seeded_field(n, 3, "gauss", 16.0)
    assert rel(g.calc_uk(u, k), orc.calc_uk(u, k)) < TOL64


@pytest.mark.parametrize("q", [4.0, 6.0, 9.0])
def test_lq_norm_rescale(q):
    n = 16
    g = _ctx(n, q=q)
    orc = Oracle(Params(n=n, q=q))
    u = seeded_field(n, 4, "gauss", 16.0, q=q, B=2.0) * 1.3
    L = g.calc_global_Lq_norm(u)
    assert abs(L - orc.calc_global_Lq_norm(u)) < TOL64 * L
    r = g.rescaleLqNorm(u, 2.0)
    assert rel(r, orc.rescaleLqNorm(u, 2.0)) < TOL64


@pytest.mark.parametrize("q", [2.0, 3.0, 4.0, 6.0, 9.0])
def test_R_parts_vs_oracle(q):
    n = 16
    g = _ctx(n, q=q)
    orc = Oracle(Params(n=n, q=q))
    u = seeded_field(n, 5, "gauss", 16.0, q=q)
    a = g.calc_local_dLqdt_inclParts(u)
    b = orc.calc_local_dLqdt_inclParts(u)
    scale = abs(b[1]) + abs(b[2])
    assert np.max(np.abs(a - b)) < TOL64 * scale, (a, b)


@pytest.mark.parametrize("q", [3.0, 4.0, 6.0, 9.0])
def test_gradL2_vs_oracle(q):
    n = 16
    g = _ctx(n, q=q)
    orc = Oracle(Params(n=n, q=q))
    u = seeded_field(n, 6, "gauss", 16.0, q=q)
    assert rel(g.GradL2ForLq(u), orc.GradL2ForLq(u)) < TOL64


@pytest.mark.parametrize("q", [4.0, 9.0])
def test_constraint_and_transport_vs_oracle(q):
    n = 16
    B = 1.5
    g = _ctx(n, q=q, B=B)
    orc = Oracle(Params(n=n, q=q, B=B))
    u = seeded_field(n, 7, "gauss", 16.0, q=q, B=B)
    eta = 0.1 * seeded_field(n, 8, "gauss", 16.0, q=q)
    xi = seeded_field(n, 9, "gauss", 16.0, q=q)
    assert rel(g.calcConstraintDerivativeL2(u), orc.calcConstraintDerivativeL2(u)) < TOL64
    assert rel(g.vectorTransport(u, q, B, eta, xi), orc.vectorTransport(u, q, B, eta, xi)) < TOL64


@pytest.mark.parametrize("q", [4.0, 6.0])
def test_eval_J_vs_oracle(q):
    n = 16
    g = _ctx(n, q=q, B=2.0)
    orc = Oracle(Params(n=n, q=q, B=2.0))
    u = seeded_field(n, 10, "gauss", 16.0, q=q) * 0.7
    for s in ("maxdLqdt", "LineMin", "||u||_q"):
        a, b = g.eval_J(u, s), orc.eval_J(u, s)
        assert abs(a - b) < TOL64 * abs(b), s


def test_abc_closed_form_on_gpu():
    """R = -4 pi^2 ||u||_q^q and R3 = 0 for the Beltrami ABC flow."""
    n = 32
    u = abc_field(n)
    for q in (4.0, 6.0):
        g = _ctx(n, q=q)
        R = g.calc_local_dLqdt_inclParts(u)
        s = (u * u).sum(axis=0)
        Lqq = np.mean(s ** (q / 2))
        assert abs(R[1] + 4 * np.pi ** 2 * Lqq) < 1e-11 * 4 * np.pi ** 2 * Lqq
        assert abs(R[2]) < 1e-10 * abs(R[1])


def test_seeded_field_matches_host_generator():
    n = 16
    g = _ctx(n, q=4.0, B=1.0)
    h = g.field(3)
    g.seeded_field(h, 1, "gauss", 16.0)
    a = g.download(h)
    b = seeded_field(n, 1, "gauss", 16.0, q=4.0, B=1.0)
    assert rel(a, b) < 1e-12


def test_eval_J_batch_matches_single():
    n = 16
    g = _ctx(n, q=4.0, B=1.0)
    orc = Oracle(Params(n=n, q=4.0, B=1.0))
    u = seeded_field(n, 11, "gauss", 16.0)
    d = seeded_field(n, 12, "gauss", 16.0)
    hu, hd = g.field(3), g.field(3)
    g.upload(hu, u)
    g.upload(hd, d)
    taus = [0.0, 1e-3, 3e-2, 0.5]
    J = g.eval_J_batch(hu, hd, taus)
    for t, j in zip(taus, J):
        ref = orc.eval_J(u + t * d, "LineMin")
        assert abs(j - ref) < TOL64 * abs(ref)


@pytest.mark.parametrize("q", [4.0, 9.0])
def test_mixed_precision_R(q):
    n = 32
    g = _ctx(n, q=q, precision="mixed")
    orc = Oracle(Params(n=n, q=q))
    u = seeded_field(n, 13, "gauss", 16.0, q=q)
    a = g.calc_local_dLqdt_inclParts(u)
    b = orc.calc_local_dLqdt_inclParts(u)
    assert abs(a[0] - b[0]) < 1e-5 * abs(b[0])


def test_fortran_caller_smoke():
    """The Fortran program maxlq_amd/fortran/mlq_fsmoke (built by build())
    calls the library through the iso_c_binding module."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "maxlq_amd", "fortran", "mlq_fsmoke")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr


def test_transform_counts():
    """mlq_transform_counts: one scalar R2C/C2R pair counts one field each
    way."""
    n = 16
    g = _ctx(n)
    f0, i0 = g.transform_counts()
    u = np.random.default_rng(7).standard_normal((n, n, n))
    F = g.fftfwdr(u)
    g.fftbwdr(F)
    f1, i1 = g.transform_counts()
    assert (f1 - f0, i1 - i0) == (1, 1)


@pytest.mark.gpu
def test_time_column_copy():
    """mlq_time_column_copy: both access patterns, every tile; algorithmic
    bytes = read + write of nfields half spectra (padding excluded)."""
    n = 128
    g = _ctx(n)
    for axis in (1, 2):
        for tile in range(8):
            ms, by = g.time_column_copy(axis, tile, nfields=2, reps=2)
            assert ms > 0.0
            assert by == 2 * 2 * 2 * 8 * (n // 2 + 1) * n * n
    with pytest.raises(RuntimeError):
        g.time_column_copy(3, 0)


@pytest.mark.gpu
def test_nan_reports_enan():
    """A NaN in the field: fftfwd stops with MLQ_ENAN (fftwfunction.f90:78-89),
    and so do R(u) (function_ops.f90:1843-1854 for e_u; the RCG's J) and the
    solve; a finite field afterwards works again."""
    from maxlq_amd.capi import MlqError, MLQ_ENAN
    n = 16
    g = _ctx(n)
    u = np.random.default_rng(3).standard_normal((3, n, n, n))
    bad = u.copy()
    bad[1, 2, 3, 4] = np.nan
    with pytest.raises(MlqError) as e:
        g.fftfwd(bad[0] + 1j * bad[1])
    assert e.value.code == MLQ_ENAN
    with pytest.raises(MlqError) as e:
        g.calc_global_dLqdt_inclParts(bad)
    assert e.value.code == MLQ_ENAN
    h = g.field(3)
    g.upload(h, bad)
    with pytest.raises(MlqError) as e:
        g.maxdLqdt(h, max_iter=1, optim_tol=0.0)
    assert e.value.code == MLQ_ENAN
    assert np.isfinite(g.calc_global_dLqdt_inclParts(u)).all()
