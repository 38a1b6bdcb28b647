"""Pins for the scalar line search (optimization.f90:613-951) and the RCG
loop of the oracle, on analytic 1-D functions and a small field (CPU)."""
import math

import numpy as np
import pytest

from oracle import lq_oracle as L
from oracle.seeded import seeded_field


def test_constants_single_precision_literals():
    # optimization.f90:804-806 are default-real literals promoted to double
    assert L.BRENT_TOL == 9.999999974752427e-07
    assert L.BRENT_ZEPS == 9.99999993922529e-09
    assert L.BRENT_CGOLD == 0.3819659948348999
    assert abs(L.MNBRAK_GOLD * L.MNBRAK_CGOLD - 1.0) < 1e-15


@pytest.mark.parametrize("xmin", [0.3, 2.7, 41.0])
def test_brent_parabola(xmin):
    f = lambda t: (t - xmin) ** 2 + 1.0
    X, FX, nev = L.brent(f, (0.0, 3.0 * xmin))
    assert abs(X - xmin) < 4 * (L.BRENT_TOL * xmin + L.BRENT_ZEPS)
    assert FX == f(X)
    assert nev >= 2                        # the tau = 0 probe plus the start point


def test_brent_nonquadratic_against_golden_section():
    f = lambda t: math.cosh(t - 1.234) + 0.1 * (t - 1.234) ** 4
    X, _, _ = L.brent(f, (0.0, 5.0))
    assert abs(X - 1.234) < 1e-5


def test_brent_log_records():
    recs = []
    L.brent(lambda t: (t - 1.0) ** 2, (0.0, 2.0), log=lambda j, x, fx: recs.append((j, x, fx)))
    assert recs[0][0] == 0 and recs[0][1] == 0.0
    assert recs[1][0] == 1 and recs[1][1] == 2.0 * L.BRENT_CGOLD
    assert recs[-1][0] == 999


def test_mnbrak_shrink_then_bracket():
    f = lambda t: (t - 3.0) ** 2
    (tA, tC), flag, nev = L.mnbrak(f, 0.0, 100.0)
    assert flag == 0
    assert min(tA, tC) < 3.0 < max(tA, tC)


def test_mnbrak_expand():
    f = lambda t: (t - 50.0) ** 2
    (tA, tC), flag, _ = L.mnbrak(f, 0.0, 1.0)
    assert flag == 0
    assert min(tA, tC) < 50.0 < max(tA, tC)


def test_mnbrak_uphill_flag():
    (br, flag, nev) = L.mnbrak(lambda t: t, 0.0, 1.0)
    assert flag == 1 and br is None
    # shrinks by CGOLD until tB <= MACH_EPSILON
    assert nev == 2 + math.ceil(math.log(L.MACH_EPSILON) / math.log(L.MNBRAK_CGOLD))


def test_mnbrak_no_bracket_flag():
    (br, flag, nev) = L.mnbrak(lambda t: -t, 0.0, 1.0)
    assert flag == 2


def test_rcg_monotone_and_sphere():
    n, q, B = 16, 4.0, 1.0
    orc = L.Oracle(L.Params(n=n, q=q, B=B, MAX_ITER=3))
    u0 = seeded_field(n, 1, "gauss", 16.0, q=q, B=B)
    res = L.maxdLqdt(orc, u0)
    assert res.iters == 3 and not res.successful
    Js = [r.J1 for r in res.records]
    assert all(b >= a for a, b in zip(Js, Js[1:]))
    assert abs(orc.calc_global_Lq_norm(res.u) - B) < 1e-12
    assert all(r.tau > 0 for r in res.records)
    assert res.records[0].beta == 0.0          # no momentum on the first step
