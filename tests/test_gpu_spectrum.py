"""GPU parity of calculate_spectral_data (function_ops.f90:2614-2677) against
the oracle, through the C ABI."""
import math

import numpy as np
import pytest

from oracle.lq_oracle import Oracle, Params
from tests.fields import grid_xyz

TP = 2.0 * np.pi


def _ctx(n, precision="fp64"):
    from maxlq_amd.lq import MaxLq
    return MaxLq(n, q=4.0, B=1.0, precision=precision)


def _check(sp_gpu, sp_orc, tol):
    assert np.array_equal(sp_gpu[:, 0], sp_orc[:, 0])
    scale = np.abs(sp_orc[:, 1]).max()
    assert np.abs(sp_gpu[:, 1] - sp_orc[:, 1]).max() <= tol * scale


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 32, 64])
@pytest.mark.parametrize("normalize", [False, True])
def test_spectrum_random_field(n, normalize):
    u = np.random.default_rng(100 + n).standard_normal((3, n, n, n))
    g = _ctx(n)
    _check(g.calculate_spectral_data(u, normalize),
           Oracle(Params(n=n, q=4.0)).calculate_spectral_data(u, normalize), 1e-12)


@pytest.mark.gpu
def test_spectrum_shell_boundary_modes():
    """Modes with integer |m| sit exactly on a shell edge: the GPU must put
    each in the oracle's shell (a mode in the neighbouring shell moves its
    whole energy, far beyond the tolerance)."""
    n = 32
    X, Y, Z = grid_xyz(n)
    modes = [(3, 4, 0), (1, 2, 2), (2, 3, 6), (0, 0, 5), (4, 4, 7), (2, 6, 9), (6, 8, 0),
             (2, 10, 11), (12, 0, 5), (8, 9, 12), (1, 4, 8), (16, 0, 0)]
    u = np.zeros((3, n, n, n))
    for j, (a, b, c) in enumerate(modes):
        u[j % 3] += (1.0 + 0.1 * j) * np.cos(TP * (a * X + b * Y + c * Z) + 0.3 * j)
    g = _ctx(n)
    orc = Oracle(Params(n=n, q=4.0))
    _check(g.calculate_spectral_data(u, False), orc.calculate_spectral_data(u, False), 1e-12)


@pytest.mark.gpu
def test_spectrum_half_spectrum_weights():
    """kx = 0 and kx = n/2 (Nyquist) modes appear once in the full spectrum,
    interior kx twice (m and -m): a single-mode field of each kind."""
    n = 16
    X, Y, Z = grid_xyz(n)
    g = _ctx(n)
    orc = Oracle(Params(n=n, q=4.0))
    for f in (np.cos(TP * 3 * Z), np.cos(TP * (n // 2) * X), np.cos(TP * (2 * X + 5 * Y)),
              np.cos(TP * (n // 2) * X + TP * 3 * Y)):
        u = np.stack([f, 0.5 * f, 0 * f])
        _check(g.calculate_spectral_data(u, False), orc.calculate_spectral_data(u, False), 1e-12)


@pytest.mark.gpu
def test_spectrum_mixed_precision():
    n = 32
    u = np.random.default_rng(7).standard_normal((3, n, n, n))
    g = _ctx(n, "mixed")
    _check(g.calculate_spectral_data(u, True),
           Oracle(Params(n=n, q=4.0)).calculate_spectral_data(u, True), 2e-6)


@pytest.mark.gpu
def test_spectrum_device_field_matches_host():
    n = 32
    u = np.random.default_rng(9).standard_normal((3, n, n, n))
    g = _ctx(n)
    h = g.field(3)
    g.upload(h, u)
    a = g.spectral_data_field(h, True)
    b = g.calculate_spectral_data(u, True)
    assert np.allclose(a, b, rtol=1e-13, atol=0)
    assert abs(TP * a[:, 1].sum() - float((u * u).sum()) / n ** 3) < 1e-12 * float((u * u).sum())
