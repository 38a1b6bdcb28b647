"""GPU parity of diagnosticScalars (function_ops.f90:2503-2602),
kappa_test_pert (function_ops.f90:613-977) and the kappa tests inside the
solve (optimization.f90:69-72, 147-149, 457-469) against the oracle, through
the C ABI."""
import numpy as np
import pytest

from maxlq_amd import capi
from oracle import diag as D
from oracle import lq_oracle as L
from oracle.seeded import seeded_field
from tests.fields import grid_xyz, rel

pytestmark = pytest.mark.gpu

TP = 2.0 * np.pi


def _ctx(n, precision="fp64", q=4.0, B=1.0):
    from maxlq_amd.lq import MaxLq
    return MaxLq(n, q=q, B=B, precision=precision)


def _check_diag(g, orc, u, tol):
    n = u.shape[-1]
    d, fields, ring = D.diagnosticScalars(orc, u)
    got = g.diagnosticScalars(u, fields=True, ring_cap=n ** 3)
    for k in ("K", "E", "Umax", "Wmax"):
        assert rel(got[k], d[k]) < tol, k
    for k in ("magUmax", "magWmax", "H", "maxHel", "minHel"):
        assert abs(got[k] - d[k]) <= tol * max(abs(d[k]), abs(d["maxHel"]), 1e-300), k
    assert abs(got["divU_L2"] - d["divU_L2"]) <= tol * max(d["divU_L2"], 2 * d["E"].sum())
    assert abs(got["vorCore"][1] - d["vorCore"][1]) <= tol * d["vorCore"][1]
    assert got["vorCore"][2] == 0.0
    assert rel(got["fields"][:3], fields[:3]) < tol
    assert np.max(np.abs(got["fields"][3] - fields[3])) < 10 * tol
    # points whose Q or |w| ratio sits within the tolerance of the threshold may
    # fall either way; elsewhere the GPU's classification is the oracle's
    W = L.Oracle(orc.p).vel2vort(u)
    ratio = np.linalg.norm(W, axis=0) / np.linalg.norm(W, axis=0).max()
    Qf = fields[3]
    n_amb_q = np.count_nonzero((Qf > 0) & (Qf < 10 * tol))
    assert abs(got["vorCore"][0] - d["vorCore"][0]) * n ** 3 <= n_amb_q
    amb = np.abs(ratio - 0.95) < 10 * tol
    if not amb.any():
        assert got["ring_points"] == len(ring)
        assert np.array_equal(got["ring"], ring)
    else:
        assert abs(got["ring_points"] - len(ring)) <= np.count_nonzero(amb)


@pytest.mark.parametrize("n", [16, 32])
def test_diag_seeded_fp64(n):
    u = seeded_field(n, 7, "gauss", 16.0, q=4.0, B=1.0)
    _check_diag(_ctx(n), L.Oracle(L.Params(n=n, q=4.0)), u, 1e-12)


def test_diag_random_nonsolenoidal_fp64():
    n = 32
    u = np.random.default_rng(5).standard_normal((3, n, n, n))
    _check_diag(_ctx(n), L.Oracle(L.Params(n=n, q=4.0)), u, 1e-11)


def test_diag_seeded_mixed():
    n = 64
    u = seeded_field(n, 8, "gauss", 16.0, q=4.0, B=1.0)
    _check_diag(_ctx(n, "mixed"), L.Oracle(L.Params(n=n, q=4.0)), u, 2e-5)


def test_diag_beltrami_closed_form():
    n = 32
    X, Y, Z = grid_xyz(n)
    u = np.stack([np.sin(TP * Z), np.cos(TP * Z), np.zeros_like(Z)])
    got = _ctx(n).diagnosticScalars(u, ring_cap=4)
    assert np.allclose(got["K"], [0.25, 0.25, 0.0], atol=1e-14)
    assert np.allclose(got["E"], [np.pi ** 2, np.pi ** 2, 0.0], atol=1e-11)
    assert abs(got["H"] - TP) < 1e-12 and got["divU_L2"] < 1e-22
    assert got["ring_points"] == n ** 3
    assert np.allclose(got["ring"], [[0, 0, 0], [1 / n, 0, 0], [2 / n, 0, 0], [3 / n, 0, 0]])


def _loaded(n, seed=3):
    rng = np.random.default_rng(seed)
    from tests.fields import bandlimit
    return bandlimit(rng.standard_normal((3, n, n, n)), n // 3) + 0.1


KINDS = [
    (D.PERT_SINE, (4.0, 2.0, 5.0), False),          # function_ops.f90:54 arguments
    (D.PERT_SINE, (2.0, 3.0, 1.0), False),
    (D.PERT_DIVFREE_SINE, (2.0, 0.0, 3.0), False),
    (D.PERT_LOAD_RAW, (0.0, 0.0, 0.0), True),
    (D.PERT_LOAD, (0.0, 0.0, 0.0), True),
    (D.PERT_LOAD_SMOOTH, (0.0, 0.0, 0.0), True),
    (D.PERT_LOAD_SMOOTH_RAW, (0.0, 0.0, 0.0), True),
    (D.PERT_LOAD_K, (1000.0, 0.0, 0.0), True),      # :89
    (D.PERT_LOAD_POLY, (-3.0, 0.0, 0.0), True),     # :86
    (D.PERT_LOAD_POLY, (-5.0, 0.0, 0.0), True),     # optimization.f90:1025
    (D.PERT_LOAD_EXP, (0.0, 0.0, 0.0), True),       # :83
    (D.PERT_RANDOM_POLY, (-5.0, 0.0, 0.0), False),
    (D.PERT_RANDOM, (0.0, 0.0, 0.0), False),
]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_kappa_test_pert_all_kinds(precision, tol):
    n = 32
    orc = L.Oracle(L.Params(n=n, q=4.0))
    g = _ctx(n, precision)
    h = g.field(3)
    ld = _loaded(n)
    for kind, m, needs in KINDS:
        print(f"kappa_test_pert kind {kind} m {m}", flush=True)
        want = D.kappa_test_pert(orc, kind, *m, seed=11, loaded=ld if needs else None)
        g.kappa_test_pert(h, kind, *m, seed=11, loaded=ld if needs else None)
        got = g.download(h)
        assert rel(got, want) < tol, (kind, m, rel(got, want))
    g.free(h)


def test_kappa_test_pert_rejects():
    g = _ctx(16)
    h = g.field(3)
    with pytest.raises(capi.MlqError):
        g.kappa_test_pert(h, D.PERT_LOAD, loaded=None)
    with pytest.raises(capi.MlqError):
        g.kappa_test_pert(h, 99)
    s = g.field(1)
    with pytest.raises(capi.MlqError):
        g.kappa_test_pert(s, D.PERT_SINE, 1.0, 1.0, 1.0)


def _start_field(n, B, seed=2):
    return seeded_field(n, seed, "gauss", 16.0, q=4.0, B=B)


@pytest.mark.parametrize("flags", [0, 1])
def test_kappa_tests_inside_solve(flags):
    """kappaTest and kappaTestEveryXiteration = 2 on a 4-iteration solve
    that stops at MAX_ITER: records initial_grad_L2 (iter 0), gradL2_iter at
    2 and 4, then end_d1 and end_gradJ (unsuccessful solve); the initial and
    end_gradJ tables against the oracle's kappa_test of the same fields."""
    n, B, seed = 32, 1.0, 21
    orc = L.Oracle(L.Params(n=n, q=4.0, B=B))
    g = _ctx(n, "fp64", B=B)
    u0 = _start_field(n, B)
    h = g.field(3)
    g.upload(h, u0)
    res, tr = g.maxdLqdt(h, max_iter=4, optim_tol=0.0, flags=flags, kappa=True, kappa_every=2,
                         kappa_seed=seed, save_every=1,
                         batch_linesearch=3 if flags == 0 else 0)
    assert res.iters == 4 and res.successful == 0
    names = [(k[0], k[1], k[2]) for k in tr.kappa]
    assert names == [(0, "initial_grad_L2", False), (2, "gradL2_iter", False),
                     (4, "gradL2_iter", False), (4, "end_d1", True), (4, "end_gradJ", True)]
    pert = D.kappa_test_pert(orc, D.PERT_RANDOM_POLY, -5.0, 0.0, seed=seed)
    phi = orc.divAvg_free(orc.rescaleLqNorm(u0, B))
    tab, frac, best, row = L.kappa_test(orc, phi, orc.GradL2ForLq(phi), pert, "maxdLqdt",
                                        L.INNER_L2)
    _, _, _, gt, gfrac, gbest, grow = tr.kappa[0]
    _compare_tables(gt, tab, orc.eval_J(phi, "maxdLqdt"))
    # best kappa of the L2 gradient: within 1e-3 of 1 (the discrete gradient
    # of the dealiased R is not the exact one; the oracle's kappa levels at
    # 0.9999 for this field)
    assert abs(gbest - best) < 1e-5 and abs(best - 1.0) < 1e-3 and grow == row
    # end_gradJ: at uvec0 = the iterate before the last update (after
    # iteration 3) with the H_l gradient there
    it3 = [f for (i, f) in tr.fields if i == 3][0]
    gH = orc.divAvg_free(orc.HilbertGradient(orc.GradL2ForLq(it3), 1.5 - 1.0 / 4.0))
    tab2, _, _, _ = L.kappa_test(orc, it3, gH, pert, "maxdLqdt", L.INNER_HL)
    _compare_tables(tr.kappa[4][3], tab2, orc.eval_J(it3, "maxdLqdt"))
    # the periodic test at iteration 2 is taken at the iterate after iteration 1
    it1 = [f for (i, f) in tr.fields if i == 1][0]
    tab3, _, _, _ = L.kappa_test(orc, it1, orc.GradL2ForLq(it1), pert, "maxdLqdt", L.INNER_L2)
    _compare_tables(tr.kappa[1][3], tab3, orc.eval_J(it1, "maxdLqdt"))


def _compare_tables(got, want, J0):
    """eps equal; <grad, pert> to 1e-10; deltaJ = J(phi + eps pert) - J(phi) to
    1e-6 plus 64 ulps of J(phi) (the two sides sum R in different orders),
    and kappa = deltaJ / (eps <grad, pert>) to the same bound, for
    1e-6 <= eps <= 1e-2."""
    assert np.array_equal(got[:, 0], want[:, 0])
    assert rel(got[:, 1], want[:, 1]) < 1e-10
    sel = (want[:, 0] >= 1e-6) & (want[:, 0] <= 1e-2)
    bound = 1e-6 * np.abs(want[sel, 2]) + 64 * 2.2e-16 * abs(J0)
    assert np.all(np.abs(got[sel, 2] - want[sel, 2]) <= bound)
    kb = bound / (want[sel, 0] * np.abs(want[sel, 1]))
    assert np.all(np.abs(got[sel, 3] - want[sel, 3]) <= kb)
