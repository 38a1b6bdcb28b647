"""GPU cases at the workload sizes: the 64^3 RCG solve and the 256^3
R / gradient / finite-difference case against oracle fixtures, the batched
line search against one-at-a-time and reference-structured evaluation, and
the kappa test against the oracle."""
import json
import math
import os

import numpy as np
import pytest

from oracle import lq_oracle as L
from oracle.seeded import seeded_field
from tests.fields import rel

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} not generated (tests/golden/make_golden.py)")
    with open(path) as f:
        return json.load(f)


def test_kappa_test_vs_oracle():
    """kappa_test (optimization.f90:957-1129) on the GPU, FP64, against the
    oracle on the same phi, gradient and perturbation."""
    from maxlq_amd.lq import MaxLq
    n, q = 16, 4.0
    u = seeded_field(n, 3, "gauss", 16.0, q=q)
    pert = seeded_field(n, 1003, "gauss", 16.0, q=q)
    orc = L.Oracle(L.Params(n=n, q=q))
    grad = orc.GradL2ForLq(u)
    ref, rfrac, rbest, rrow = L.kappa_test(orc, u, grad, pert, "maxdLqdt")
    g = MaxLq(n, q=q, precision="fp64")
    hu, hg, hp = g.field(3), g.field(3), g.field(3)
    g.upload(hu, u)
    g.upload(hg, grad)
    g.upload(hp, pert)
    tab, frac, best, row = g.kappa_test(hu, hg, hp, "maxdLqdt", "L2")
    print("kappa gpu", tab[:, 3], "oracle", ref[:, 3])
    assert np.allclose(tab[:, 0], ref[:, 0], rtol=1e-15, atol=0)
    assert np.allclose(tab[:, 1], ref[:, 1], rtol=1e-12, atol=0)
    # rows whose J difference is far above round-off (eps >= 1e-6)
    assert np.allclose(tab[8:, 2], ref[8:, 2], rtol=1e-6, atol=0)
    assert np.allclose(tab[8:, 3], ref[8:, 3], rtol=1e-6, atol=0)
    # the adjustment follows the same rule on the GPU's own kappa column
    adj, f2 = L.kappa_adjust(tab[:, 3])
    assert abs(f2 - frac) <= 1e-14 * abs(frac)
    assert np.allclose(tab[:, 4], adj, rtol=1e-14, atol=0)
    assert (best, row) == L.kappa_best(tab[:, 3])
    g.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_linesearch_batched_vs_sequential(precision, tol):
    """8 step sizes tau_i = 10^(-4 + 3i/7) evaluated in one batch, one at a
    time, and reference-structured (every J from the field): same J."""
    from maxlq_amd.lq import MaxLq
    n, q, B = 32, 4.0, 2.0
    g = MaxLq(n, q=q, B=B, precision=precision)
    hu, hd = g.field(3), g.field(3)
    g.seeded_field(hu, 4, "gauss", 16.0)
    g.seeded_field(hd, 14, "gauss", 16.0)
    taus = 10.0 ** (-4.0 + 3.0 * np.arange(8) / 7.0)
    Jb, tp, tb = g.time_linesearch(hu, hd, taus, 8)
    Js, _, ts = g.time_linesearch(hu, hd, taus, 1)
    Jr, _, tr = g.time_linesearch(hu, hd, taus, 0)
    Jp, tpp, tpe = g.time_linesearch(hu, hd, taus, -1)   # polynomials where no point clips
    print(precision, "batched", tb, "ms, sequential", ts, "ms, reference", tr, "ms, precompute", tp,
          "polynomial", tpe, "ms after", tpp)
    assert np.allclose(Jb, Js, rtol=1e-14, atol=0)
    assert np.allclose(Jb, Jr, rtol=tol, atol=0)
    assert np.allclose(Jp, Jb, rtol=1e-12 if precision == "fp64" else 1e-6, atol=0)
    if precision == "fp64":
        orc = L.Oracle(L.Params(n=n, q=q, B=B))
        u, d = g.download(hu), g.download(hd)
        Jo = np.array([orc.eval_J(u + t * d, "LineMin") for t in taus])
        assert np.allclose(Jb, Jo, rtol=1e-10, atol=0)
    g.close()


def test_rcg_solve_64_against_oracle_fixture():
    """64^3, q = 4, seed 1, exp(-|m|^2/16), B = 1: RCG solve with MAX_ITER 50,
    OPTIM_TOL 1e-8 (FP64) against the oracle's trajectory and final field."""
    import ctypes as C
    from maxlq_amd import capi
    from maxlq_amd.lq import MaxLq
    gold = _golden("rcg64_q4.json")
    n, q, B = gold["n"], gold["q"], gold["B"]
    g = MaxLq(n, q=q, B=B, precision="fp64")
    h = g.field(3)
    g.seeded_field(h, gold["seed"], gold["envelope"][0], gold["envelope"][1])
    recs = []
    msgs = {}

    def cb(user, kind, it, i0, i1, d):
        if kind == capi.MLQ_LOG_ITER:
            recs.append((i0, d[0], d[1], d[2], d[3], d[4], int(d[5])))
        elif kind == capi.MLQ_LOG_MSG:
            msgs.setdefault(it, []).append(i0)

    cfg = capi.RcgCfg(gold["max_iter"], gold["optim_tol"], 0, 0, 25, 10, 0)
    res = capi.RcgResult()
    g._chk(g.lib.mlq_rcg_solve(g.ctx, h, C.byref(cfg), C.byref(res), capi.LOG_CB(cb), None))
    got = recs[1:]
    ref = gold["records"]
    print("iterations gpu", res.iters, "oracle", gold["iters"])
    assert res.iters == gold["iters"]
    assert len(got) == len(ref)
    # brent stops within TOL = 1e-6 (relative, optimization.f90:790) of the
    # minimiser, so tau agrees to a few TOL; J is second order in that, and
    # the differences of the iterates accumulate over the 50 steps (observed
    # 2e-10).  Parity of every evaluation on identical inputs is 1e-10
    # (the oracle tests at 16^3-256^3); the trajectory bound is 1e-9.
    worst, worst_tau, worst_br = 0.0, 0.0, 0.0
    for (it, tau, beta, J, ta, tc, nj), r in zip(got, ref):
        assert it == r["iter"]
        worst_tau = max(worst_tau, abs(tau - r["tau"]) / abs(r["tau"]))
        worst = max(worst, abs(J - r["J1"]) / abs(r["J1"]))
        # mnbrak's bracket (tA, tC) and the number of J requests of mnbrak
        # and brent, and the message codes of the iteration (the final
        # "optimization finished" code 1 is outside the oracle's records)
        worst_br = max(worst_br, abs(ta - r["tau_brack"][0]) / max(abs(r["tau_brack"][1]), 1e-300),
                       abs(tc - r["tau_brack"][1]) / abs(r["tau_brack"][1]))
        assert nj == r["nJ"], (it, nj, r["nJ"])
        m = [c for c in msgs.get(it, []) if c != 1 or it != res.iters]
        assert m == r["msgs"], (it, m, r["msgs"])
    print("max |dJ|/|J| over the trajectory", worst, "max |dtau|/tau", worst_tau,
          "max bracket difference", worst_br)
    assert worst_tau < 1e-4
    assert worst_br < 1e-4
    assert worst < 1e-9
    assert abs(res.J1 - gold["J1"]) < 1e-9 * abs(gold["J1"])
    u = g.download(h)
    orc = L.Oracle(L.Params(n=n, q=q, B=B))
    fin = gold["final"]
    assert abs(orc.calc_global_Lq_norm(u) - fin["Lq"]) < 1e-12
    assert abs(orc.global_summed_field_inner_product(u, u, L.INNER_L2) - fin["L2sq"]) < 1e-8 * fin["L2sq"]
    for c, z, y, x, v in fin["points"]:
        assert abs(u[c, z, y, x] - v) < 1e-7 * max(1.0, abs(v))
    g.close()


def test_R_gradient_fd_256_against_oracle_fixture():
    """256^3, q = 4, seed 2, |m|^2 exp(-|m|^2/64), B = 1 (FP64): R parts,
    <GradL2ForLq(u), phi_i>, <HilbertGradient(G), phi_i> and central
    differences of R (h = 1e-6) along 8 seeded directions, against the oracle."""
    from maxlq_amd.lq import MaxLq
    gold = _golden("fd256_q4.json")
    n, q, B, h = gold["n"], gold["q"], gold["B"], gold["h"]
    g = MaxLq(n, q=q, B=B, precision="fp64")
    hu = g.field(3)
    g.seeded_field(hu, gold["seed"], gold["envelope"][0], gold["envelope"][1])
    u = g.download(hu)
    Rp = g.calc_local_dLqdt_inclParts(u)
    assert np.allclose(Rp, gold["R_parts"], rtol=1e-10, atol=0)
    G = g.GradL2ForLq(u)
    H = g.HilbertGradient(G.copy(), 1.5 - 1.0 / q)
    ip = lambda a, b: float(np.sum(a * b)) / float(n) ** 3
    uu = ip(u, u)
    assert abs(uu - gold["L2sq_u"]) < 1e-12 * uu
    assert abs(ip(G, G) - gold["G_L2sq"]) < 1e-10 * gold["G_L2sq"]
    assert abs(ip(H, H) - gold["H_L2sq"]) < 1e-10 * gold["H_L2sq"]
    hp = g.field(3)
    for dref in gold["directions"]:
        g.seeded_field(hp, dref["seed"], gold["envelope"][0], gold["envelope"][1])
        phi = g.download(hp)
        phi *= math.sqrt(uu / ip(phi, phi))
        Gp, Hp = ip(G, phi), ip(H, phi)
        fd = (g.eval_J(u + h * phi, "maxdLqdt") - g.eval_J(u - h * phi, "maxdLqdt")) / (2.0 * h)
        print("dir", dref["seed"], "<G,phi>", Gp, "fd", fd, "oracle fd", dref["fd"])
        assert abs(Gp - dref["G_phi"]) < 1e-10 * abs(dref["G_phi"])
        assert abs(Hp - dref["H_phi"]) < 1e-10 * abs(dref["H_phi"])
        # both central differences carry the cancellation error of R(u +- h phi)
        # evaluated to ~1e-16 |R| each: |d fd| <~ 2 * 1e-16 |R| / (2h) per side
        noise = 8.0 * 2.2e-16 * abs(gold["R_parts"][0]) / h
        assert abs(fd - dref["fd"]) < noise + 1e-9 * abs(dref["fd"])
        assert abs(fd - Gp) < 1e-6 * abs(Gp)
    g.close()
