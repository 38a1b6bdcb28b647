"""GPU RCG solve (maxdLqdt) and batched line-search J against the oracle."""
import numpy as np
import pytest

from oracle import lq_oracle as L
from oracle.seeded import seeded_field
from tests.fields import rel

pytestmark = pytest.mark.gpu

REF = 1  # MLQ_RCG_REFERENCE
LEAN = 8  # MLQ_RCG_LEAN: line search from coefficient fields


def _solve_gpu(u0, n, q, B, iters, flags=0, batch=0, precision="fp64"):
    import ctypes as C
    from maxlq_amd import capi
    from maxlq_amd.lq import MaxLq
    g = MaxLq(n, q=q, B=B, precision=precision)
    h = g.field(3)
    g.upload(h, u0)
    tr = {"iter": [], "lm": [], "br": []}

    def cb(user, kind, it, i0, i1, d):
        if kind == capi.MLQ_LOG_ITER:
            tr["iter"].append((i0, d[0], d[1], d[2]))
        elif kind == capi.MLQ_LOG_LINEMIN:
            tr["lm"].append((it, i0, i1) + tuple(d[k] for k in range(6)))
        elif kind == capi.MLQ_LOG_BRENT:
            tr["br"].append((it, i0, d[0], d[1]))

    cfg = capi.RcgCfg(iters, 1.0e-5, 0, batch, 25, 10, flags)
    res = capi.RcgResult()
    g._chk(g.lib.mlq_rcg_solve(g.ctx, h, C.byref(cfg), C.byref(res), capi.LOG_CB(cb), None))
    return res, tr, g.download(h)


def _oracle_solve(u0, n, q, B, iters):
    orc = L.Oracle(L.Params(n=n, q=q, B=B, MAX_ITER=iters))
    brent_log = []
    return L.maxdLqdt(orc, u0)


@pytest.mark.parametrize("q,flags", [(4.0, 0), (4.0, REF), (6.0, 0), (9.0, 0), (4.0, LEAN),
                                     (6.0, LEAN), (9.0, LEAN)])
def test_rcg_trajectory_fp64(q, flags):
    n, B, iters = 16, 1.0, 4
    u0 = seeded_field(n, 1, "gauss", 16.0, q=q, B=B)
    ref = _oracle_solve(u0, n, q, B, iters)
    res, tr, u = _solve_gpu(u0, n, q, B, iters, flags=flags)
    assert res.iters == ref.iters
    got = tr["iter"][1:]
    assert len(got) == len(ref.records)
    for (it, tau, beta, J), r in zip(got, ref.records):
        assert it == r.iter
        assert abs(tau - r.tau) < 1e-8 * abs(r.tau)
        assert abs(beta - r.beta) <= 1e-8 * abs(r.beta)
        assert abs(J - r.J1) < 1e-10 * abs(r.J1)
    assert rel(u, ref.u) < 1e-8
    assert abs(res.J1 - ref.J1) < 1e-10 * abs(ref.J1)


def test_rcg_fast_matches_reference_path_32():
    """Fast (resident-state, tau-polynomial) vs reference-structured GPU solve,
    including the mnbrak / brent records."""
    n, q, B, iters = 32, 4.0, 1.0, 3
    u0 = seeded_field(n, 2, "gauss", 16.0, q=q, B=B)
    r1, t1, u1 = _solve_gpu(u0, n, q, B, iters, flags=REF)
    r2, t2, u2 = _solve_gpu(u0, n, q, B, iters, flags=0, batch=3)
    assert r1.iters == r2.iters
    assert len(t1["br"]) == len(t2["br"]) and len(t1["lm"]) == len(t2["lm"])
    for a, b in zip(t1["br"], t2["br"]):
        assert a[:2] == b[:2]
        assert abs(a[2] - b[2]) <= 1e-9 * max(abs(a[2]), 1e-300)
        assert abs(a[3] - b[3]) <= 1e-10 * abs(a[3])
    assert rel(u2, u1) < 1e-9
    assert r2.n_J == r1.n_J
    assert r2.n_J_computed <= r2.n_J + 2 * r2.iters


@pytest.mark.parametrize("q,precision", [(6.0, "fp64"), (9.0, "fp64"), (9.0, "mixed")])
def test_rcg_lean_linesearch_matches_reference_path_32(q, precision):
    """The line search from coefficient fields (the path of grids where the
    resident state does not fit) against the reference-structured solve, which
    evaluates every J(tau) from the field: same iterations, mnbrak / brent
    step sizes and values, final field (FP64 to rounding; mixed within the
    rounding of FP32 fields)."""
    n, B, iters = 32, 1.0, 3
    u0 = seeded_field(n, 4, "gauss", 16.0, q=q, B=B)
    r1, t1, u1 = _solve_gpu(u0, n, q, B, iters, flags=REF, precision=precision)
    r2, t2, u2 = _solve_gpu(u0, n, q, B, iters, flags=LEAN, precision=precision)
    assert r1.iters == r2.iters
    if precision == "fp64":
        assert r1.n_J == r2.n_J and len(t1["br"]) == len(t2["br"])
        for a, b in zip(t1["br"], t2["br"]):
            assert a[:2] == b[:2]
            assert abs(a[2] - b[2]) <= 1e-8 * max(abs(a[2]), 1e-300)
            assert abs(a[3] - b[3]) <= 1e-10 * abs(a[3])
    # mixed: both evaluate J from FP32 fields in different orders (the mixed
    # path's J agrees with FP64 to 1e-4 on this case, test_mixed_vs_fp64_rcg_32)
    tj, tu = (1e-10, 1e-9) if precision == "fp64" else (1e-4, 1e-3)
    for a, b in zip(t1["iter"], t2["iter"]):
        print(precision, "J per iteration, reference-structured vs lean:", a[3], b[3])
        assert abs(a[3] - b[3]) <= tj * abs(a[3])
    assert rel(u2, u1) < tu


def test_rcg_fast_unfused_matches_fused_32(monkeypatch):
    """The producer x passes (pointwise work inside the forward transform)
    against the pointwise-kernel + transform form (MLQ_FUSED=0): same solve,
    FP64, to rounding."""
    n, q, B, iters = 32, 4.0, 1.0, 3
    u0 = seeded_field(n, 2, "gauss", 16.0, q=q, B=B)
    r1, t1, u1 = _solve_gpu(u0, n, q, B, iters, flags=0, batch=3)
    monkeypatch.setenv("MLQ_FUSED", "0")
    r2, t2, u2 = _solve_gpu(u0, n, q, B, iters, flags=0, batch=3)
    assert r1.iters == r2.iters and r1.n_J == r2.n_J
    for a, b in zip(t1["br"], t2["br"]):
        assert a[:2] == b[:2]
        assert abs(a[3] - b[3]) <= 1e-12 * abs(a[3])
    assert rel(u2, u1) < 1e-12
    assert abs(r2.J1 - r1.J1) <= 1e-12 * abs(r1.J1)


def test_rcg_fast_with_clipping_matches_reference_path():
    """A rough field (flat spectrum up to |m| = n/2): the 2/3 mask truncates
    u.u and D(phi.phi) goes negative at some points, so the line search mixes
    the exact tau-polynomials with the batched-transform evaluation wherever a
    point clips.  The resident-state solve must still follow the
    reference-structured one."""
    n, q, B, iters = 32, 4.0, 1.0, 3
    u0 = seeded_field(n, 7, "gauss", 1.0e4, q=q, B=B)
    r1, t1, u1 = _solve_gpu(u0, n, q, B, iters, flags=REF)
    r2, t2, u2 = _solve_gpu(u0, n, q, B, iters, flags=0, batch=3)
    print("J requests", r2.n_J, "computed", r2.n_J_computed, "by polynomials", r2.n_J_poly)
    assert r1.iters == r2.iters
    for a, b in zip(t1["iter"], t2["iter"]):
        assert a[0] == b[0]
        assert abs(a[3] - b[3]) <= 1e-10 * abs(a[3])
    assert rel(u2, u1) < 1e-9
    assert r2.n_J_poly < r2.n_J_computed     # some step sizes went through the transform


@pytest.mark.parametrize("q", [4.0, 6.0, 9.0])
def test_eval_J_batch_fast_vs_oracle(q):
    from maxlq_amd.lq import MaxLq
    n, B = 16, 2.0
    orc = L.Oracle(L.Params(n=n, q=q, B=B))
    u = seeded_field(n, 21, "gauss", 16.0, q=q, B=B)
    d = seeded_field(n, 22, "gauss", 16.0, q=q, B=B)
    g = MaxLq(n, q=q, B=B)
    hu, hd = g.field(3), g.field(3)
    g.upload(hu, u)
    g.upload(hd, d)
    taus = 10.0 ** np.linspace(-4, -1, 8)
    J = g.eval_J_batch(hu, hd, taus)
    for t, j in zip(taus, J):
        r = orc.eval_J(u + t * d, "LineMin")
        assert abs(j - r) < 1e-10 * abs(r), (t, j, r)
    J1 = np.array([g.eval_J_batch(hu, hd, [t])[0] for t in taus])
    assert np.max(np.abs(J1 - J)) <= 1e-12 * np.max(np.abs(J))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-11)])
def test_q6_linesearch_polynomials_vs_chain(precision, tol):
    """q = 6: J(tau) from the polynomial R sums and the two-round-trip norm
    against the full |phi|^k chain evaluation on the same fields, at step
    sizes where D(phi.phi) does not clip (both FP64: the mixed path
    evaluates J on its FP64 twin)."""
    from maxlq_amd.lq import MaxLq
    n, q, B = 64, 6.0, 1.0
    u = seeded_field(n, 31, "gauss", 16.0, q=q, B=B)
    d = seeded_field(n, 32, "gauss", 16.0, q=q, B=B)
    g = MaxLq(n, q=q, B=B, precision=precision)
    hu, hd = g.field(3), g.field(3)
    g.upload(hu, u)
    g.upload(hd, 0.1 * d)
    taus = 10.0 ** np.linspace(-4, -1, 8)
    Jp, _, _ = g.time_linesearch(hu, hd, taus, -1)   # polynomials where exact
    Jc, _, _ = g.time_linesearch(hu, hd, taus, 3)    # every J through the chain
    rel = np.max(np.abs(Jp - Jc) / np.abs(Jc))
    print(precision, "q=6 polynomial vs chain J: max rel", rel)
    assert rel < tol
    # the solve takes the polynomials for its step sizes on this kind of field
    r, _, _ = _solve_gpu(u, n, q, B, 2, batch=3, precision=precision)
    print(precision, "q=6 solve: J requests", r.n_J, "computed", r.n_J_computed,
          "by polynomials", r.n_J_poly)
    assert r.n_J_poly > 0


def test_mixed_vs_fp64_rcg_32():
    n, q, B, iters = 32, 4.0, 1.0, 3
    u0 = seeded_field(n, 3, "gauss", 16.0, q=q, B=B)
    r64, t64, u64 = _solve_gpu(u0, n, q, B, iters)
    r32, t32, u32 = _solve_gpu(u0, n, q, B, iters, precision="mixed", batch=3)
    assert r32.iters == r64.iters
    for a, b in zip(t32["iter"], t64["iter"]):
        print("mixed vs fp64 J per iteration:", a[0], a[3], b[3], abs(a[3] - b[3]) / abs(b[3]))
        assert abs(a[3] - b[3]) < 1e-4 * abs(b[3])


def test_rcg_spectra_records():
    """save_spectraEveryXiteration inside the solve (optimization.f90:119-121,
    377-381): uvec at the start and uvec and d every 2nd iteration; the last
    uvec record is the spectrum of the solve's result, and the resident-state
    and reference-structured solves deliver the same spectra."""
    from maxlq_amd.lq import MaxLq
    n, q, B = 32, 4.0, 1.0
    u0 = seeded_field(n, 2, "gauss", 16.0, q=q, B=B)
    orc = L.Oracle(L.Params(n=n, q=q, B=B))
    ustart = orc.divAvg_free(orc.rescaleLqNorm(u0, B))      # optimization.f90:61-62
    runs = {}
    for fl in (0, REF):
        g = MaxLq(n, q=q, B=B, precision="fp64")
        h = g.field(3)
        g.upload(h, u0)
        res, tr = g.maxdLqdt(h, max_iter=4, optim_tol=0.0, spectra_every=2, flags=fl,
                             batch_linesearch=3 if fl == 0 else 0)
        assert res.iters == 4
        assert [(it, w) for it, w, _ in tr.spectra] == [(0, "uvec"), (2, "uvec"), (2, "d"),
                                                         (4, "uvec"), (4, "d")]
        u = g.download(h)
        last = tr.spectra[3][2]
        assert np.allclose(last, g.calculate_spectral_data(u, True), rtol=1e-12, atol=0)
        assert abs(2 * np.pi * last[:, 1].sum() - float((u * u).sum()) / n ** 3) < 1e-12
        runs[fl] = tr.spectra
        # raw spectra (normalizeSpectrumByL2Norm = .false.) of the start field
        g.upload(h, u0)
        res, tr = g.maxdLqdt(h, max_iter=1, optim_tol=0.0, spectra_every=1, flags=fl | 2)
        assert np.allclose(tr.spectra[0][2], orc.calculate_spectral_data(ustart, False),
                           rtol=1e-10, atol=1e-10 * n ** 6)
    for (a_it, a_w, a), (b_it, b_w, b) in zip(runs[0], runs[REF]):
        assert (a_it, a_w) == (b_it, b_w)
        assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max()
