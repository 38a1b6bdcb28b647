"""Analytic and band-limited test fields shared by the oracle and GPU tests."""
import numpy as np


def grid_xyz(n):
    """Physical coordinates on the unit box (initialize.f90:162, dx = 1/n)."""
    x = np.arange(n) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z


def abc_field(n, A=1.0, B=1.0, C=1.0):
    """The iguess=50 ABC flow (function_ops.f90:100-116) with amplitudes."""
    X, Y, Z = grid_xyz(n)
    tp = 2.0 * np.pi
    ux = A * np.cos(tp * Z) + C * np.cos(tp * Y)
    uy = B * np.cos(tp * X) + A * np.cos(tp * Z)
    uz = C * np.cos(tp * Y) + B * np.cos(tp * X)
    return np.stack([ux, uy, uz])


def mode_index(n):
    m = np.arange(n)
    return np.where(m <= n // 2, m, m - n)


def bandlimit(u, M):
    """Keep only modes with max_j |m_j| <= M."""
    n = u.shape[-1]
    m = np.abs(mode_index(n))
    mask = (m[:, None, None] <= M) & (m[None, :, None] <= M) & (m[None, None, :] <= M)
    return np.fft.ifftn(np.fft.fftn(u, axes=(-3, -2, -1)) * mask, axes=(-3, -2, -1)).real


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
