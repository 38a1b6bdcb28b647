"""GPU RCG solve (maxdLqdt) against the oracle's trajectory."""
import numpy as np
import pytest

from oracle import lq_oracle as L
from oracle.seeded import seeded_field
from tests.fields import rel

pytestmark = pytest.mark.gpu


def _solve_both(n, q, B, iters, seed=1, batch=0, precision="fp64"):
    from maxlq_amd.lq import MaxLq
    u0 = seeded_field(n, seed, "gauss", 16.0, q=q, B=B)
    orc = L.Oracle(L.Params(n=n, q=q, B=B, MAX_ITER=iters))
    ref = L.maxdLqdt(orc, u0)
    g = MaxLq(n, q=q, B=B, precision=precision)
    h = g.field(3)
    g.upload(h, u0)
    res, tr = g.maxdLqdt(h, max_iter=iters, batch_linesearch=batch)
    return ref, res, tr, g.download(h)


@pytest.mark.parametrize("q", [4.0, 6.0])
def test_rcg_trajectory_fp64(q):
    ref, res, tr, u = _solve_both(16, q, 1.0, 4)
    assert res.iters == ref.iters
    # iteration-info records (iter, tau, beta, J) after the initial line
    got = [r for r in tr.iteration_info[1:]]
    assert len(got) == len(ref.records)
    for (it, tau, beta, J), r in zip(got, ref.records):
        assert it == r.iter
        assert abs(tau - r.tau) < 1e-8 * abs(r.tau)
        assert abs(beta - r.beta) < 1e-8 * max(abs(r.beta), 1e-300) or (beta == 0 and r.beta == 0)
        assert abs(J - r.J1) < 1e-10 * abs(r.J1)
    assert rel(u, ref.u) < 1e-8
    assert abs(res.J1 - ref.J1) < 1e-10 * abs(ref.J1)
