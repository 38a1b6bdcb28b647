"""GPU parity of calc_global_summed_enstrophy and calc_global_dEnstrophydt
(function_ops.f90:1417-1481, 2422-2467) against the oracle, through the C
ABI, and their values in the solve result (optimization.f90:491-493)."""
import numpy as np
import pytest

from oracle import lq_oracle as L
from oracle.seeded import seeded_field
from tests.fields import rel

pytestmark = pytest.mark.gpu


def _ctx(n, precision="fp64", **kw):
    from maxlq_amd.lq import MaxLq
    return MaxLq(n, q=4.0, B=1.0, precision=precision, **kw)


@pytest.mark.parametrize("n", [16, 32])
def test_enstrophy_and_rate_random_field(n):
    u = np.random.default_rng(200 + n).standard_normal((3, n, n, n))
    orc = L.Oracle(L.Params(n=n, q=4.0))
    g = _ctx(n)
    assert rel(g.calc_global_summed_enstrophy(u), orc.calc_global_summed_enstrophy(u)) < 1e-12
    assert rel(g.calc_global_dEnstrophydt(u), orc.calc_global_dEnstrophydt(u)) < 1e-11


def test_enstrophy_rate_seeded_field_mixed():
    n = 32
    u = seeded_field(n, 3, "gauss", 16.0, q=4.0, B=1.0)
    orc = L.Oracle(L.Params(n=n, q=4.0))
    g = _ctx(n, "mixed")
    assert rel(g.calc_global_summed_enstrophy(u), orc.calc_global_summed_enstrophy(u)) < 1e-6
    assert rel(g.calc_global_dEnstrophydt(u), orc.calc_global_dEnstrophydt(u)) < 1e-5


@pytest.mark.parametrize("flags", [0, 1])
def test_solve_result_enstrophy(flags):
    """mlq_rcg_result.enstrophy / dEnstrophydt are those of the result field."""
    n = 32
    u0 = seeded_field(n, 2, "gauss", 16.0, q=4.0, B=1.0)
    g = _ctx(n)
    h = g.field(3)
    g.upload(h, u0)
    res, _ = g.maxdLqdt(h, max_iter=2, optim_tol=0.0, flags=flags,
                        batch_linesearch=3 if flags == 0 else 0)
    u = g.download(h)
    orc = L.Oracle(L.Params(n=n, q=4.0))
    assert rel(res.enstrophy, orc.calc_global_summed_enstrophy(u)) < 1e-12
    assert rel(res.dEnstrophydt, orc.calc_global_dEnstrophydt(u)) < 1e-11
