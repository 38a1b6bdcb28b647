"""The GPU's nonlinear chain with the 2/3 mask active, FP64, against exact
dealiased products (the cube mask of the unaliased product, formed on a 2n
grid in tests/fields.py): calc_uk, calc_gk_order2 and R(u) at q = 4 and 6 on
fields whose products the mask truncates and the n grid aliases."""
import numpy as np
import pytest

from tests.fields import exact_R, masked_chain_field, rel
from tests.test_oracle_pins import _chain, _exact_gk

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [32, 64])
def test_gpu_chain_exact_with_active_mask(n):
    from maxlq_amd.lq import MaxLq
    u, E, g = _chain(n)
    G = MaxLq(n, q=4.0, precision="fp64")
    assert rel(G.calc_uk(u, 1.0), np.sqrt(E["S"])) < 1e-12
    assert rel(G.calc_uk(u, 2.0), E["S"]) < 1e-12
    assert rel(G.calc_uk(u, 4.0), E["SS"]) < 1e-12
    for k in (2, 3, 4, 5):
        assert rel(G.calc_gk_order2(g, float(k)), _exact_gk(g, k)) < 1e-12, k
    G.close()


@pytest.mark.parametrize("n", [32, 64])
@pytest.mark.parametrize("q", [4.0, 6.0])
def test_gpu_R_exact_with_active_mask(n, q):
    from maxlq_amd.lq import MaxLq
    u = masked_chain_field(n)
    ref, _ = exact_R(u, q)
    G = MaxLq(n, q=q, precision="fp64")
    got = G.calc_local_dLqdt_inclParts(u)
    G.close()
    print(f"n={n} q={q}: R exact {ref} gpu {got}")
    assert abs(got[0] - ref[0]) <= 1e-12 * abs(ref[0])
    assert abs(got[1] - ref[1]) <= 1e-12 * abs(ref[1])
    assert abs(got[2] - ref[2]) <= 1e-11 * abs(ref[2])
