"""The mixed path (FP32 fields and transforms, FP64 reductions, scalars and
line-search chain) against the FP64 path on the same fields: R(u) to
|dR|/R < 1e-7 on the 256^3 q = 6 B-continuation and on the 512^3 q = 4
line-search field, and a line-search J(tau) that is as smooth in tau as the
FP64 one (brent resolves tau to TOL = 1e-6, optimization.f90:804-806)."""
import ctypes as C
import time

import numpy as np
import pytest

from oracle.seeded import seeded_field

pytestmark = pytest.mark.gpu

DR_TOL = 1.0e-7   # |dR|/R of the mixed path against FP64 on the same field


def _solve(g, h, iters, batch):
    from maxlq_amd import capi
    cfg = capi.RcgCfg(iters, 0.0, 0, batch, 25, 10, 0)
    res = capi.RcgResult()
    t0 = time.perf_counter()
    g._chk(g.lib.mlq_rcg_solve(g.ctx, h, C.byref(cfg), C.byref(res), C.cast(None, capi.LOG_CB),
                               None))
    return res, time.perf_counter() - t0


@pytest.mark.timeout(600)
def test_mixed_q6_continuation_256():
    """256^3, q = 6, seed 3, |m|^2 exp(-|m|^2/64): B_j = 4^(j/9), j = 0..9,
    30 RCG iterations per B (OPTIM_TOL 0), the mixed and the FP64 solve each
    continuing from its own result.  On every mixed iterate R(u) from the
    mixed path (reference-structured evaluation and the solver's line-search
    J at tau = 0) agrees with FP64 to 1e-7, and the mixed solve requests no
    more J values than the FP64 one (+15%)."""
    from maxlq_amd.lq import MaxLq
    n, q = 256, 6.0
    Bs = [4.0 ** (j / 9.0) for j in range(10)]
    gm = MaxLq(n, q=q, B=Bs[0], precision="mixed")
    g64 = MaxLq(n, q=q, B=Bs[0], precision="fp64")
    hm, h64 = gm.field(3), g64.field(3)
    gm.seeded_field(hm, 3, "k2gauss", 64.0)
    g64.seeded_field(h64, 3, "k2gauss", 64.0)
    hu, hd = g64.field(3), g64.field(3)
    hmu, hmd = gm.field(3), gm.field(3)
    nJm = nJ64 = 0
    tm = t64 = 0.0
    worst = 0.0
    for j, B in enumerate(Bs):
        gm.set_params(B=B)
        g64.set_params(B=B)
        rm, sm = _solve(gm, hm, 30, 3)
        r64, s64 = _solve(g64, h64, 30, 0)
        nJm += rm.n_J
        nJ64 += r64.n_J
        tm += sm
        t64 += s64
        u = gm.download(hm)
        Pm = gm.calc_local_dLqdt_inclParts(u)
        P64 = g64.calc_local_dLqdt_inclParts(u)
        dR = abs(Pm[0] - P64[0]) / abs(P64[0])
        # the solver's own evaluation: J_LineMin at tau = 0 (d: the iterate itself)
        gm.upload(hmu, u)
        gm.upload(hmd, u)
        g64.upload(hu, u)
        g64.upload(hd, u)
        Jm = gm.eval_J_batch(hmu, hmd, [0.0, 1.0e-3])
        J64 = g64.eval_J_batch(hu, hd, [0.0, 1.0e-3])
        dJ = float(np.max(np.abs(Jm - J64) / np.abs(J64)))
        print(f"B={B:.4f} J1 mixed {rm.J1:.10e} fp64 {r64.J1:.10e} same-field dR/R {dR:.2e} "
              f"dJ/J {dJ:.2e} n_J mixed {rm.n_J} fp64 {r64.n_J} s {sm:.2f} / {s64:.2f}")
        worst = max(worst, dR, dJ)
        assert dR < DR_TOL, (B, dR)
        assert dJ < DR_TOL, (B, dJ)
        assert rm.iters == 30 and r64.iters == 30
    print(f"max same-field |dR|/R {worst:.2e}; J requests mixed {nJm} fp64 {nJ64}; "
          f"mixed {300 / tm:.1f} iter/s, fp64 {300 / t64:.1f} iter/s")
    assert nJm <= 1.15 * nJ64, (nJm, nJ64)
    gm.close()
    g64.close()


@pytest.mark.timeout(600)
def test_mixed_q4_512_same_field():
    """512^3, q = 4, seed 4, exp(-|m|^2/256), B = 2 (the line-search
    workload): after two mixed RCG iterations, R(u) of the iterate and
    J_LineMin at the 8 step sizes tau_i = 10^(-4 + 3i/7) along a seeded
    direction (seed 14) from the mixed path agree with FP64 to 1e-7."""
    from maxlq_amd.lq import MaxLq
    n, q, B = 512, 4.0, 2.0
    taus = 10.0 ** (-4.0 + 3.0 * np.arange(8) / 7.0)
    gm = MaxLq(n, q=q, B=B, precision="mixed")
    hm, hd = gm.field(3), gm.field(3)
    gm.seeded_field(hm, 4, "gauss", 256.0)
    gm.seeded_field(hd, 14, "gauss", 256.0)
    _solve(gm, hm, 2, 3)
    u = gm.download(hm)
    d = gm.download(hd)
    Pm = gm.calc_local_dLqdt_inclParts(u)
    gm.upload(hm, u)
    Jm = gm.eval_J_batch(hm, hd, taus)
    gm.close()
    g64 = MaxLq(n, q=q, B=B, precision="fp64")
    hu, hd64 = g64.field(3), g64.field(3)
    g64.upload(hu, u)
    g64.upload(hd64, d)
    P64 = g64.calc_local_dLqdt_inclParts(u)
    J64 = g64.eval_J_batch(hu, hd64, taus)
    g64.close()
    dR = abs(Pm[0] - P64[0]) / abs(P64[0])
    dJ = np.abs(Jm - J64) / np.abs(J64)
    print(f"R fp64 {P64} mixed {Pm}: |dR|/R {dR:.2e}; J line search |dJ|/J max {dJ.max():.2e}")
    assert dR < DR_TOL
    assert np.all(dJ < DR_TOL), dJ


@pytest.mark.parametrize("q", [6.0, 9.0])
def test_mixed_linesearch_J_smooth_in_tau(q):
    """J_LineMin(tau) of the mixed path has no rounding that depends on tau:
    its second differences over steps of 1e-4 tau (FP64 round-off is ~1e-4
    of them) are those of the FP64 J up to the fields' FP32 rounding, where
    FP32 per-step chains (~1e-8 noise in J) would swamp them."""
    from maxlq_amd.lq import MaxLq
    n, B = 64, 2.0
    u = seeded_field(n, 31, "gauss", 16.0, q=q, B=B)
    d = seeded_field(n, 32, "gauss", 16.0, q=q, B=0.3 * B)
    t0 = 0.37
    taus = t0 * (1.0 + 1.0e-4 * np.arange(-3, 4))
    J = {}
    for prec in ("fp64", "mixed"):
        g = MaxLq(n, q=q, B=B, precision=prec)
        hu, hd = g.field(3), g.field(3)
        g.upload(hu, u)
        g.upload(hd, d)
        J[prec] = g.eval_J_batch(hu, hd, taus)
        g.close()
    d2 = {p: J[p][2:] - 2.0 * J[p][1:-1] + J[p][:-2] for p in J}
    scale = np.max(np.abs(d2["fp64"]))
    print("second differences fp64", d2["fp64"], "mixed", d2["mixed"])
    assert scale > 0.0
    assert np.max(np.abs(d2["mixed"] - d2["fp64"])) < 2.0e-3 * scale
    assert np.max(np.abs(J["mixed"] - J["fp64"]) / np.abs(J["fp64"])) < DR_TOL
