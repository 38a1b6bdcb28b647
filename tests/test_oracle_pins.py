"""Pins for the FP64 oracle: every check compares the oracle with something
that is not the oracle's own formula -- a brute-force DFT, a closed form, an
integration-by-parts identity, or finite differences of another oracle
function (CPU only)."""
import math

import numpy as np
import pytest

from oracle.lq_oracle import (INNER_H, INNER_HL, INNER_L2, Grid, Oracle,
                              OracleStop, Params, fftbwd, fftfwd)
from oracle.seeded import seeded_field
from tests.fields import abc_field, bandlimit, grid_xyz, mode_index, rel

TP = 2.0 * np.pi


def _dft3(u, sign):
    n = u.shape[-1]
    idx = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(idx, idx) / n)
    out = np.einsum("az,by,cx,zyx->abc", W, W, W, u.astype(np.complex128))
    return out


def test_fft_bruteforce_dft():
    rng = np.random.default_rng(0)
    for n in (4, 6, 8):
        u = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
        F = fftfwd(u)
        assert rel(F, _dft3(u, -1.0)) < 1e-13
        assert rel(fftbwd(F), _dft3(F, +1.0) / n ** 3) < 1e-13


def test_fft_nan_stops():
    u = np.zeros((4, 4, 4))
    u[1, 2, 3] = np.nan
    with pytest.raises(OracleStop):
        fftfwd(u)


def test_wavenumbers():
    g = Grid(8)
    expect = TP * np.array([0, 1, 2, 3, 4, -3, -2, -1], dtype=float)
    assert np.array_equal(g.k1d, expect)          # Nyquist is +pi*n
    assert g.k1d[4] == np.pi * 8
    assert g.dV == 1.0 / 512


@pytest.mark.parametrize("p,mcut", [(2.0, 2), (1.0, 4), (1.5, 3), (0.0, 4), (3.0, 2)])
def test_dealias_bruteforce(p, mcut):
    n = 8
    rng = np.random.default_rng(1)
    u = rng.standard_normal((n, n, n))
    orc = Oracle(Params(n=n))
    F = _dft3(u, -1.0)
    m = np.abs(mode_index(n))
    keep = np.maximum(np.maximum(m[:, None, None], m[None, :, None]), m[None, None, :]) <= mcut
    expect = (_dft3(F * keep, +1.0) / n ** 3).real
    assert rel(orc.dealias_scalar(u, p), expect) < 1e-13


def test_gradient_closed_form_and_nyquist():
    n = 16
    X, Y, Z = grid_xyz(n)
    orc = Oracle(Params(n=n))
    a, b, c = 1, -2, 3
    ph = TP * (a * X + b * Y + c * Z)
    gu = orc.gradient(np.sin(ph))
    for j, kj in enumerate((a, b, c)):
        assert rel(gu[j], TP * kj * np.cos(ph)) < 1e-12
    # Nyquist rule: REAL(ifft(iK fft)) kills the own-dimension Nyquist mode
    nyq = np.cos(np.pi * n * X)                      # (-1)^i
    g2 = orc.gradient(nyq * np.sin(TP * Y))
    assert np.max(np.abs(g2[0])) < 1e-12
    assert rel(g2[1], TP * nyq * np.cos(TP * Y)) < 1e-12
    assert np.max(np.abs(g2[2])) < 1e-12


def test_divergence_closed_form():
    n = 16
    X, Y, Z = grid_xyz(n)
    orc = Oracle(Params(n=n))
    u = np.stack([np.sin(TP * X), np.cos(TP * 2 * Z) * np.sin(TP * Y), np.cos(TP * Z)])
    expect = TP * np.cos(TP * X) + TP * np.cos(TP * 2 * Z) * np.cos(TP * Y) - TP * np.sin(TP * Z)
    assert rel(orc.divergence(u), expect) < 1e-12


def test_laplacian_abc():
    n = 16
    u = abc_field(n, 1.0, 0.7, 0.3)
    orc = Oracle(Params(n=n))
    assert rel(orc.laplacian(u), -4 * np.pi ** 2 * u) < 1e-12


def test_poisson_closed_form():
    n = 16
    X, Y, Z = grid_xyz(n)
    orc = Oracle(Params(n=n))
    f = np.cos(TP * X) + 2.0 * np.sin(TP * (Y + 2 * Z)) + 3.0
    expect = -np.cos(TP * X) / (4 * np.pi ** 2) - 2.0 * np.sin(TP * (Y + 2 * Z)) / (20 * np.pi ** 2)
    assert rel(orc.solve_poisson(f), expect) < 1e-12


def test_leray_invariants():
    n = 16
    orc = Oracle(Params(n=n))
    u = bandlimit(np.random.default_rng(2).standard_normal((3, n, n, n)), 5)
    Pu = orc.divAvg_free(u)
    assert np.max(np.abs(orc.divergence(Pu))) < 1e-10 * np.max(np.abs(u)) * TP * n
    assert rel(orc.divAvg_free(Pu), Pu) < 1e-12              # idempotent
    assert np.max(np.abs(Pu.mean(axis=(1, 2, 3)))) < 1e-14   # mean removed
    psi = bandlimit(np.random.default_rng(3).standard_normal((n, n, n)), 5)
    assert np.max(np.abs(orc.divAvg_free(orc.gradient(psi)))) < 1e-11 * np.max(np.abs(orc.gradient(psi)))
    w = abc_field(n)                                          # div-free, mean 0
    assert rel(orc.divAvg_free(w), w) < 1e-12
    # u = w + grad psi + const  ->  P u = w
    uu = w + orc.gradient(psi) + 0.5
    assert rel(orc.divAvg_free(uu), w) < 1e-11


def test_hilbert_single_mode():
    n = 16
    X, Y, Z = grid_xyz(n)
    q = 4.0
    orc = Oracle(Params(n=n, q=q))
    s = 1.5 - 1.0 / q
    ph = TP * (X + 2 * Y)
    u = np.stack([np.sin(ph), 0 * ph, np.cos(ph)])
    k = TP * math.sqrt(5.0)
    fac = 1.0 / (1.0 + (0.1 * k) ** (2 * s))
    assert rel(orc.HilbertGradient(u, s), fac * u) < 1e-12


def test_inner_products_closed_form():
    n = 16
    X, Y, Z = grid_xyz(n)
    q = 6.0
    orc = Oracle(Params(n=n, q=q))
    s = 1.5 - 1.0 / q
    u = np.stack([np.sin(TP * X), np.cos(TP * Y), 0 * X])
    v = np.stack([np.sin(TP * X), 2.0 * np.cos(TP * Y), np.ones_like(X)])
    ip = orc.global_summed_field_inner_product
    assert abs(ip(u, v, INNER_L2) - (0.5 + 1.0)) < 1e-14
    wl = (1.0 + (0.1 * TP) ** s) ** 2
    assert abs(ip(u, v, INNER_HL) - wl * 1.5) < 1e-13
    wh = (1.0 + TP) ** (2 * s)
    assert abs(ip(u, v, INNER_H) - wh * 1.5) < 1e-12
    # constant (k = 0) parts: weight 1 in both Sobolev products
    c = np.stack([np.ones_like(X)] * 3)
    assert abs(ip(c, c, INNER_HL) - 3.0) < 1e-14
    with pytest.raises(OracleStop):
        orc.inner_product(u[0], u[0], "H_l^((3q-1)/(2q))")


@pytest.mark.parametrize("k", [0.0, 0.5, 1.0, 2.0, 3.0, 5.0, 7.0])
def test_gk_order2_polynomial(k):
    n = 32
    X, Y, Z = grid_xyz(n)
    g = 2.0 + np.cos(TP * X) + 0.5 * np.sin(TP * Y)          # positive, modes <= 1
    orc = Oracle(Params(n=n))
    assert rel(orc.calc_gk_order2(g, k), g ** k) < 1e-12


@pytest.mark.parametrize("k", [0.0, 1.0, 2.0, 2.5, 4.5, 7.0])
def test_calc_uk_constant_modulus(k):
    n = 16
    X, Y, Z = grid_xyz(n)
    c = 1.7
    u = c * np.stack([np.sin(TP * Z), np.cos(TP * Z), 0 * Z])   # |u| = c
    orc = Oracle(Params(n=n))
    assert rel(orc.calc_uk(u, k), np.full_like(X, c ** k)) < 1e-12


@pytest.mark.parametrize("q", [4.0, 6.0, 9.0])
def test_lq_norm_closed_form(q):
    n = 16
    X, Y, Z = grid_xyz(n)
    c = 0.8
    u = c * np.stack([np.sin(TP * Z), np.cos(TP * Z), 0 * Z])
    orc = Oracle(Params(n=n, q=q))
    assert abs(orc.calc_global_Lq_norm(u) - c) < 1e-13
    B = 2.0
    assert abs(orc.calc_global_Lq_norm(orc.rescaleLqNorm(u, B)) - B) < 1e-13


def test_lq_norm_abc_q4():
    n = 32
    u = abc_field(n)
    orc = Oracle(Params(n=n, q=4.0))
    s = (u * u).sum(axis=0)                     # (u.u)^2 has modes <= 4 < n/3
    assert abs(orc.calc_global_Lq_norm(u) - np.mean(s * s) ** 0.25) < 1e-13


def _spectral_grad(f):
    n = f.shape[-1]
    K = TP * mode_index(n).astype(float)
    K[n // 2] = 0.0  # band-limited fields carry no Nyquist content
    fh = np.fft.fftn(f)
    return np.stack([np.fft.ifftn(1j * K[None, None, :] * fh).real,
                     np.fft.ifftn(1j * K[None, :, None] * fh).real,
                     np.fft.ifftn(1j * K[:, None, None] * fh).real])


def test_R_q2_closed_form():
    n = 16
    u = seeded_field(n, 7, "gauss", 16.0, q=2.0, B=1.0)
    orc = Oracle(Params(n=n, q=2.0))
    g = Grid(n)
    uh = np.fft.fftn(u, axes=(-3, -2, -1))
    Keff = g.k1d.copy()
    Keff[n // 2] = 0.0
    k2 = (Keff[None, None, :] ** 2 + Keff[None, :, None] ** 2 + Keff[:, None, None] ** 2)
    expect = -np.sum(k2 * np.abs(uh) ** 2) / n ** 6
    R = orc.calc_local_dLqdt_inclParts(u)
    assert abs(R[0] - expect) < 1e-12 * abs(expect)
    assert R[2] == 0.0


@pytest.mark.parametrize("q", [4.0, 6.0])
def test_R_abc_closed_form(q):
    n = 32
    u = abc_field(n)
    orc = Oracle(Params(n=n, q=q))
    s = (u * u).sum(axis=0)
    Lqq = np.mean(s ** (q / 2.0))
    R = orc.calc_local_dLqdt_inclParts(u)
    assert abs(R[1] - (-4 * np.pi ** 2 * Lqq)) < 1e-11 * 4 * np.pi ** 2 * Lqq
    assert abs(R[2]) < 1e-10 * abs(R[1])      # Beltrami: pressure part vanishes


@pytest.mark.parametrize("q,M", [(4.0, 2), (6.0, 1)])
def test_R_integration_by_parts_identity(q, M):
    """R1+R2 = nu int |u|^(q-2) u.Lap u and R3 = -int |u|^(q-2) u.grad p for a
    band-limited divergence-free field (different formulas from the
    reference's function_ops.f90:1304-1358)."""
    n = 32
    orc = Oracle(Params(n=n, q=q))
    u = orc.divAvg_free(bandlimit(seeded_field(n, 11, "gauss", 16.0, q=q), M))
    s = (u * u).sum(axis=0)
    A = s ** ((q - 2.0) / 2.0)
    gu = np.stack([_spectral_grad(u[i]) for i in range(3)])      # gu[i, j] = d_j u_i
    lap = np.stack([sum(_spectral_grad(gu[i, j])[j] for j in range(3)) for i in range(3)])
    src = sum(gu[i, j] * gu[j, i] for i in range(3) for j in range(3))
    K = TP * mode_index(n).astype(float)
    ksq = K[None, None, :] ** 2 + K[None, :, None] ** 2 + K[:, None, None] ** 2
    ksq[0, 0, 0] = 1.0
    ph = np.fft.fftn(src) / ksq
    ph[0, 0, 0] = 0.0
    p = np.fft.ifftn(ph).real                                    # p = -Lap^-1 src
    gp = _spectral_grad(p)
    R12 = np.mean(A * (u * lap).sum(axis=0))
    R3 = -np.mean(A * (u * gp).sum(axis=0))
    R = orc.calc_local_dLqdt_inclParts(u)
    assert abs(R[1] - R12) < 1e-11 * abs(R12)
    assert abs(R[2] - R3) < 1e-9 * abs(R12)
    assert abs(R3) > 1e-6 * abs(R12)           # the pressure part is exercised


@pytest.mark.parametrize("q,M,tol", [(4.0, 2, 1e-7), (6.0, 1, 1e-6), (9.0, 1, 2e-3)])
def test_gradient_fd_bandlimited(q, M, tol):
    n = 32
    orc = Oracle(Params(n=n, q=q))
    u = orc.divAvg_free(bandlimit(seeded_field(n, 21, "gauss", 16.0, q=q), M))
    g = orc.GradL2ForLq(u)
    h = 1e-4
    for seed in (1001, 1002):
        phi = orc.divAvg_free(bandlimit(seeded_field(n, seed, "gauss", 16.0, q=q), M))
        fd = (orc.eval_J(u + h * phi, "maxdLqdt") - orc.eval_J(u - h * phi, "maxdLqdt")) / (2 * h)
        an = orc.global_summed_field_inner_product(g, phi, INNER_L2)
        assert abs(fd - an) < tol * abs(an), (fd, an)


def test_gradient_q2_closed_form():
    """q = 2: P(GradL2ForLq(u)) = 2 nu Lap u for band-limited div-free u."""
    n = 16
    orc = Oracle(Params(n=n, q=2.0))
    u = orc.divAvg_free(bandlimit(seeded_field(n, 5, "gauss", 16.0, q=2.0), 3))
    Pg = orc.divAvg_free(orc.GradL2ForLq(u))
    assert rel(Pg, 2.0 * orc.laplacian(u)) < 1e-11


def test_constraint_derivative_fd():
    n = 32
    q = 4.0
    orc = Oracle(Params(n=n, q=q))
    u = orc.divAvg_free(bandlimit(seeded_field(n, 31, "gauss", 16.0, q=q), 2))
    phi = orc.divAvg_free(bandlimit(seeded_field(n, 32, "gauss", 16.0, q=q), 2))
    nrm = orc.calcConstraintDerivativeL2(u)
    h = 1e-5
    fd = (orc.calc_global_Lq_norm(u + h * phi) - orc.calc_global_Lq_norm(u - h * phi)) / (2 * h)
    an = orc.global_summed_field_inner_product(nrm, phi, INNER_L2)
    assert abs(fd - an) < 1e-8 * abs(an)


def test_vector_transport_tangency():
    n = 32
    q, B = 4.0, 1.5
    orc = Oracle(Params(n=n, q=q, B=B))
    u = orc.rescaleLqNorm(orc.divAvg_free(bandlimit(seeded_field(n, 41, "gauss", 16.0, q=q), 2)), B)
    eta = 0.3 * orc.divAvg_free(bandlimit(seeded_field(n, 42, "gauss", 16.0, q=q), 2))
    xi1 = orc.divAvg_free(bandlimit(seeded_field(n, 43, "gauss", 16.0, q=q), 2))
    xi2 = orc.divAvg_free(bandlimit(seeded_field(n, 44, "gauss", 16.0, q=q), 2))
    T1 = orc.vectorTransport(u, q, B, eta, xi1)
    T2 = orc.vectorTransport(u, q, B, eta, xi2)
    T12 = orc.vectorTransport(u, q, B, eta, 2.0 * xi1 - 3.0 * xi2)
    assert rel(T12, 2.0 * T1 - 3.0 * T2) < 1e-12                  # linear in xi
    upe = u + eta
    nrm = (upe * upe).sum(axis=0) * upe                            # |u+eta|^2 (u+eta)
    ip = orc.global_summed_field_inner_product
    assert abs(ip(nrm, T1)) < 1e-11 * math.sqrt(ip(nrm, nrm) * ip(T1, T1))
    # with eta = 0 on the sphere the transport is the tangent projection
    T0 = orc.vectorTransport(u, q, B, 0 * u, xi1)
    n0 = (u * u).sum(axis=0) * u
    c = ip(n0, xi1) / np.mean(((u * u).sum(axis=0)) ** 2)
    assert rel(T0, xi1 - c * u) < 1e-11


# ---- kappa test (optimization.f90:957-1129) ---------------------------------------
def test_kappa_adjust_hand_case():
    """The flattest 4-ratio window (rows 7..11, 1-based) is constant 2 -> frac 2;
    a constant negative window gives a negative frac (sign of its first row)."""
    from oracle.lq_oracle import kappa_adjust
    k = np.array([50., -3., 9., 0.1, 7., 30., 2., 2., 2., 2., 2., 0.5, -8., 4., 60., 0.01])
    adj, frac = kappa_adjust(k)
    assert frac == 2.0
    assert np.array_equal(adj, k / 2.0)
    k2 = k.copy()
    k2[6:11] = -3.0
    adj2, frac2 = kappa_adjust(k2)
    assert abs(frac2 + 3.0) < 1e-15
    assert np.allclose(adj2[6:11], 1.0, rtol=1e-15, atol=0)


def test_kappa_best_criterion_quirk():
    """The reference compares max(|k-1|, |1/k-1|) against max(|best-1|, |1/k-1|)
    (the current row's 1/k on both sides): 0.5 is kept over 1.8, which a
    symmetric criterion would have preferred."""
    from oracle.lq_oracle import kappa_best
    assert kappa_best([0.5, 1.8]) == (0.5, 1)
    assert kappa_best([10.0, 0.5, 1.2, 0.9, 1.05]) == (1.05, 5)


def test_kappa_tends_to_one_for_exact_gradient():
    """For band-limited fields GradL2ForLq is the exact L2 gradient of R, so
    kappa -> 1 at moderate eps (a wrong sign or scale of the table fails)."""
    from oracle.lq_oracle import kappa_test
    n, q = 32, 4.0
    orc = Oracle(Params(n=n, q=q))
    u = orc.divAvg_free(bandlimit(seeded_field(n, 21, "gauss", 16.0, q=q), 5))
    g = orc.GradL2ForLq(u)
    pert = orc.divAvg_free(bandlimit(seeded_field(n, 1001, "gauss", 16.0, q=q), 5))
    tab, frac, best, row = kappa_test(orc, u, g, pert, "maxdLqdt")
    assert np.allclose(tab[:, 0], 10.0 ** (np.arange(1, 17) - 15.0), rtol=1e-15, atol=0)
    for ii in range(5, 9):           # eps = 1e-9 .. 1e-6
        assert abs(tab[ii, 3] - 1.0) < 1e-4, (tab[ii, 0], tab[ii, 3])
    assert tab[13, 3] < 0.0          # eps = 0.1: far outside the linear range
    assert abs(best - 1.0) < 1e-4 and 1 <= row <= 16
    assert abs(frac - 1.0) < 1e-3


# ---- calculate_spectral_data (function_ops.f90:2614-2677) -------------------

@pytest.mark.parametrize("n", [8, 12])
def test_spectrum_parseval_every_mode_once(n):
    """The shells 2 pi i < |K| <= 2 pi (i+1), i < n, hold every mode but k = 0
    exactly once: shell sums + |u_hat(0)|^2 = n^3 sum |u|^2 (Parseval of the
    unnormalised DFT).  A missing or doubled shell edge fails this."""
    u = np.random.default_rng(n).standard_normal((3, n, n, n))
    orc = Oracle(Params(n=n, q=4.0))
    sp = orc.calculate_spectral_data(u, normalize=False)
    zero = sum(abs(u[c].sum()) ** 2 for c in range(3))
    assert rel(sp[:, 1].sum() + zero, n ** 3 * float((u * u).sum())) < 1e-12
    assert np.allclose(sp[:, 0], TP * np.arange(1, n + 1), rtol=1e-15, atol=0)


def test_spectrum_single_mode_closed_form():
    """cos(2 pi (x + y)): two modes +-(1, 1, 0) of |u_hat|^2 = (n^3/2)^2 at
    |m| = sqrt 2, i.e. shell 1; normalised, 2 pi sum = ||u||_2^2 = 1/2."""
    n = 16
    X, Y, Z = grid_xyz(n)
    u = np.stack([np.cos(TP * (X + Y)), 0 * X, 0 * X])
    orc = Oracle(Params(n=n, q=4.0))
    sp = orc.calculate_spectral_data(u, normalize=False)
    want = np.zeros(n)
    want[1] = 2 * (n ** 3 / 2.0) ** 2
    assert np.allclose(sp[:, 1], want, rtol=0, atol=1e-12 * want[1])
    spn = orc.calculate_spectral_data(u, normalize=True)
    assert abs(spn[1, 1] - 1.0 / (4.0 * math.pi)) < 1e-14
    assert abs(TP * spn[:, 1].sum() - 0.5) < 1e-14


def test_spectrum_shell_edges():
    """|m| = 3 exactly (m = (3, 0, 0): |K| = SQRT(K1**2) = K1 = 2 pi 3, the
    upper bound of shell 2) lies in shell 2, not 3: the upper edge belongs to
    the shell, the lower edge does not."""
    n = 16
    X, Y, Z = grid_xyz(n)
    u = np.stack([0 * X, np.cos(TP * 3 * X), 0 * X])
    sp = Oracle(Params(n=n, q=4.0)).calculate_spectral_data(u, normalize=False)
    assert sp[2, 1] > 0.49 * n ** 6 and sp[3, 1] < 1e-12 * n ** 6


# ---- enstrophy and its rate (function_ops.f90:1417-1481, 2422-2467) ---------

def test_enstrophy_closed_forms():
    """u = (0, sin(2 pi k x), 0): w = (0, 0, 2 pi k cos(2 pi k x)), enstrophy
    pi^2 k^2; the ABC flow: 2 pi^2 (A^2 + B^2 + C^2)."""
    n = 16
    X, Y, Z = grid_xyz(n)
    orc = Oracle(Params(n=n, q=4.0))
    for k in (1, 3):
        u = np.stack([0 * X, np.sin(TP * k * X), 0 * X])
        w = orc.vel2vort(u)
        assert np.abs(w[2] - TP * k * np.cos(TP * k * X)).max() < 1e-11 * k
        assert abs(orc.calc_global_summed_enstrophy(u) - math.pi ** 2 * k ** 2) < 1e-11
    A, B, C = 1.0, 0.7, 0.4
    u = abc_field(n, A, B, C)
    assert abs(orc.calc_global_summed_enstrophy(u) - 2 * math.pi ** 2 * (A * A + B * B + C * C)) < 1e-11


def test_denstrophy_single_mode():
    """u = (0, sin(2 pi k x), 0): (u.grad) u = 0, Lap u = -(2 pi k)^2 u, so
    dE/dt = -nu (2 pi k)^4 / 2."""
    n = 16
    X, Y, Z = grid_xyz(n)
    nu = 0.7
    orc = Oracle(Params(n=n, q=4.0, visc=nu))
    u = np.stack([0 * X, np.sin(TP * 2 * X), 0 * X])
    assert abs(orc.calc_global_dEnstrophydt(u) + nu * (TP * 2) ** 4 / 2) < 1e-9


def test_denstrophy_vortex_stretching_identity():
    """For divergence-free band-limited u (|m| <= n/3, where the dealias mask
    does not act on the product's overlap with Lap u):
    int (u.grad)u . Lap u = int w_i w_j d_j u_i (vortex stretching), computed
    here from numpy spectral derivatives, and R_1 = -nu ||Lap u||^2 from
    Parseval."""
    n = 24
    orc = Oracle(Params(n=n, q=4.0, visc=1.0))
    u = bandlimit(orc.divAvg_free(np.random.default_rng(5).standard_normal((3, n, n, n))), n // 3 - 1)
    K = TP * mode_index(n).astype(float)
    ks = [K.reshape(1, 1, -1), K.reshape(1, -1, 1), K.reshape(-1, 1, 1)]
    uh = [np.fft.fftn(u[c]) for c in range(3)]
    d = [[np.fft.ifftn(1j * ks[j] * uh[i]).real for j in range(3)] for i in range(3)]  # d[i][j] = d_j u_i
    w = np.stack([d[2][1] - d[1][2], d[0][2] - d[2][0], d[1][0] - d[0][1]])
    stretch = sum(float(np.sum(w[i] * w[j] * d[i][j])) for i in range(3) for j in range(3)) / n ** 3
    ksq = ks[0] ** 2 + ks[1] ** 2 + ks[2] ** 2
    lap2 = sum(float(np.sum(np.abs(ksq * uh[c]) ** 2)) for c in range(3)) / n ** 6
    got = orc.calc_global_dEnstrophydt(u)
    assert rel(got, -lap2 + stretch) < 1e-10


# ---------------------------------------------------------------------------
# The 2/3 mask active inside the nonlinear chain.  For fields whose dealiased
# products stay alias-free (band-limited inside the cube, n a power of two)
# every dealiased product has an exact value -- the cube mask of the
# unaliased product, formed on a 2n grid (tests/fields.py) -- which fixes
# calc_uk, calc_gk_order2, the gradient products and R(u) at q = 4 and 6
# independently of the oracle's transforms and dealias code.
# ---------------------------------------------------------------------------

from tests.fields import (cube_mcut, exact_R, mask_down, masked_chain_field,  # noqa: E402
                          spec_deriv, up2)


def _scalar_bandlimited(n, seed=11, c=1.5, eps=0.4):
    """c + eps w, w band-limited to max_j |m_j| <= mcut - 1, max |w| = 1."""
    M = cube_mcut(n) - 1
    rng = np.random.default_rng(seed)
    w = bandlimit(rng.standard_normal((n, n, n)), M)
    return c + eps * w / np.max(np.abs(w))


def _exact_gk(g, k):
    """calc_gk_order2(g, k) for a band-limited g (function_ops.f90:2281-2319):
    gk(0) = 1, gk(1) = g, gk(k) = D(D(g g) gk(k-2)); every product alias-free."""
    n = g.shape[-1]
    mc = cube_mcut(n)
    G2 = mask_down(up2(g) * up2(g), n, mc)
    if k == 0:
        return np.ones_like(g)
    if k == 1:
        return g
    return mask_down(up2(G2) * up2(_exact_gk(g, k - 2)), n, mc)


_CHAIN_CACHE = {}


def _chain(n):
    if n not in _CHAIN_CACHE:
        u = masked_chain_field(n)
        _CHAIN_CACHE[n] = (u, exact_R(u, 4.0)[1], _scalar_bandlimited(n))
    return _CHAIN_CACHE[n]


def _chain_pins(orc, n):
    """(name, oracle value, exact value) of every chain quantity pinned."""
    u, E, g = _chain(n)
    out = [("calc_uk(u,1)", orc.calc_uk(u, 1.0), np.sqrt(E["S"])),
           ("calc_uk(u,2)", orc.calc_uk(u, 2.0), E["S"]),
           ("calc_uk(u,4)", orc.calc_uk(u, 4.0), E["SS"]),
           ("calc_vgraduvgradu(u,u)", orc.calc_vgraduvgradu(u, u), E["vgraduvgradu"]),
           ("calc_vvgradu(u,u)", orc.calc_vvgradu(u, u), E["vvgradu"]),
           ("calc_nablaModSquared(u)", orc.calc_nablaModSquared(u), E["gradsq"]),
           ("calc_nablaUnablaUt(u)", orc.calc_nablaUnablaUt(u), E["ggT"])]
    for k in (2, 3, 4, 5):
        out.append((f"calc_gk_order2(g,{k})", orc.calc_gk_order2(g, float(k)), _exact_gk(g, k)))
    return out


def test_masked_chain_field_exercises_the_mask():
    """The pin fields: D(u.u) > 0 (no clipping), and u.u has modes beyond the
    cube (the mask truncates) and beyond n/2 (the n-grid product aliases)."""
    n = 32
    u, E, g = _chain(n)
    mc = cube_mcut(n)
    assert E["S"].min() > 0.5
    m = np.abs(mode_index(2 * n))
    mm = np.maximum(np.maximum(m[:, None, None], m[None, :, None]), m[None, None, :])
    for f2 in (sum(up2(u[i]) ** 2 for i in range(3)), up2(g) ** 2):
        F = np.abs(np.fft.fftn(f2))
        F[0, 0, 0] = 0.0                                   # fractions of the fluctuation
        tot = np.sqrt(np.sum(F ** 2))
        assert np.sqrt(np.sum(F[mm > mc] ** 2)) > 0.02 * tot        # truncated by the mask
        assert np.sqrt(np.sum(F[mm > n // 2] ** 2)) > 0.005 * tot   # aliased on the n grid


@pytest.mark.parametrize("n", [32, 64])
def test_chain_quantities_exact_with_active_mask(n):
    orc = Oracle(Params(n=n, q=4.0))
    for name, got, ref in _chain_pins(orc, n):
        assert rel(got, ref) < 1e-12, name


@pytest.mark.parametrize("n", [32, 64])
@pytest.mark.parametrize("q", [4.0, 6.0])
def test_R_exact_with_active_mask(n, q):
    """R, R1+R2 and R3 of calc_local_dLqdt_inclParts with every dealiased
    product evaluated exactly (|u|^(q-2) = D(u.u) at q = 4, D(S S) at q = 6)."""
    u = masked_chain_field(n)
    ref, _ = exact_R(u, q)
    got = Oracle(Params(n=n, q=q)).calc_local_dLqdt_inclParts(u)
    assert abs(got[0] - ref[0]) <= 1e-12 * abs(ref[0])
    assert abs(got[1] - ref[1]) <= 1e-12 * abs(ref[1])
    assert abs(got[2] - ref[2]) <= 1e-11 * abs(ref[2])


def test_R_q4_closed_forms_with_active_mask():
    """q = 4, S = D(u.u): R1 = -nu <S, |grad u|^2> and R2 = -(nu/2) ||grad S||^2
    = -(nu/2) sum_k |K|^2 |S_hat_k|^2 / N^6 (D is an orthogonal projection and
    D(u_i d_j u_i) = d_j S / 2)."""
    n = 32
    u = masked_chain_field(n)
    _, E = exact_R(u, 4.0)
    S = E["S"]
    gsq = sum(spec_deriv(u[i], j) ** 2 for i in range(3) for j in range(3))
    R1 = -np.sum(S * gsq) / n ** 3
    K = TP * mode_index(n)
    ksq = K[None, None, :] ** 2 + K[None, :, None] ** 2 + K[:, None, None] ** 2
    R2 = -0.5 * np.sum(ksq * np.abs(np.fft.fftn(S)) ** 2) / n ** 6
    got = Oracle(Params(n=n, q=4.0)).calc_local_dLqdt_inclParts(u)
    assert abs(got[1] - (R1 + R2)) <= 1e-12 * abs(R1 + R2)


def _dealias_sites(orc, run):
    """Source lines of the oracle that call dealias_scalar with a mask that is
    not the identity, during run()."""
    import sys
    sites = set()
    orig = orc.dealias_scalar

    def spy(f, nonlinOrder=2.0):
        if not orc.dealias_mask(nonlinOrder).all():
            fr = sys._getframe(1)
            while fr.f_code.co_name in ("_D", "dealias_scalar"):
                fr = fr.f_back
            sites.add((fr.f_code.co_name, fr.f_lineno))
        return orig(f, nonlinOrder)

    orc.dealias_scalar = spy
    try:
        run()
    finally:
        del orc.dealias_scalar
    return sites


def test_every_chain_dealias_is_pinned():
    """Mutations planted in the oracle on the CPU: deleting any one dealias
    call (every call from one source line) of calc_uk, calc_gk_order2,
    calc_vgraduvgradu, calc_vvgradu, calc_nablaModSquared or
    calc_nablaUnablaUt changes at least one pinned quantity by far more than
    round-off."""
    import sys
    n = 32
    orc = Oracle(Params(n=n, q=4.0))
    sites = _dealias_sites(orc, lambda: _chain_pins(orc, n))
    targets = {"calc_uk", "calc_gk_order2", "calc_vgraduvgradu", "calc_vvgradu",
               "calc_nablaModSquared", "calc_nablaUnablaUt"}
    assert {s[0] for s in sites} >= targets - {"calc_nablaModSquared", "calc_nablaUnablaUt"}
    orig = orc.dealias_scalar
    missed = []
    for site in sorted(sites):
        def mutant(f, nonlinOrder=2.0, site=site):
            fr = sys._getframe(1)
            while fr.f_code.co_name in ("_D", "dealias_scalar"):
                fr = fr.f_back
            if (fr.f_code.co_name, fr.f_lineno) == site:
                return f                                   # the deleted dealias
            return orig(f, nonlinOrder)
        orc.dealias_scalar = mutant
        try:
            worst = max(rel(got, ref) for _, got, ref in _chain_pins(orc, n))
        finally:
            del orc.dealias_scalar
        if worst < 1e-9:
            missed.append(site)
    print("dealias sites pinned:", sorted(sites))
    assert not missed, missed
