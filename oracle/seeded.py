"""Seeded divergence-free velocity fields (numpy restatement of the project's
generator; TEST INFRASTRUCTURE ONLY).

The reference has no seeded generator (its random fields are loaded from
NetCDF files on cluster paths, initialize.f90:12-45, or drawn from an
unseeded ``random_number``, function_ops.f90:795-844), so the build defines
one, identically on the host (here) and on the device (mlq_seeded_field):

1. white noise: for component c and linear point index i (x fastest), two
   uniforms from the SplitMix64 finaliser keyed by
   ``seed*0xD1B54A32D192ED03 + 2*(c*N^3 + i) + j`` (j = 0, 1), combined by
   Box-Muller: ``sqrt(-2 ln u1) cos(2 pi u2)``;
2. forward transform, multiply by the envelope E(|m|) of the integer mode
   index m (Nyquist counted as +N/2), zero the k = 0 mode;
3. Leray projection with the reference's Nyquist rule
   (divAvg_free, function_ops.f90:2843-2918): P u_j = u_j - K'_j (K'.u)/|K|^2
   where K' has the own-dimension Nyquist entry zeroed;
4. inverse transform, then rescale to ||u||_q = B with the reference Lq norm
   (calc_global_Lq_norm, function_ops.f90:1675-1692).
"""
from __future__ import annotations

import numpy as np

from .lq_oracle import Oracle, Params

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
KEYMUL = np.uint64(0xD1B54A32D192ED03)
C1 = np.uint64(0xBF58476D1CE4E5B9)
C2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    """SplitMix64 output function of a uint64 array (Steele et al. 2014)."""
    with np.errstate(over="ignore"):
        z = x + GOLDEN
        z = (z ^ (z >> np.uint64(30))) * C1
        z = (z ^ (z >> np.uint64(27))) * C2
        return z ^ (z >> np.uint64(31))


def white_noise(n: int, seed: int, ncomp: int = 3):
    """Physical-space N(0,1) noise, shape (ncomp, n, n, n)."""
    npts = n ** 3
    out = np.empty((ncomp, n, n, n))
    with np.errstate(over="ignore"):
        base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * KEYMUL
    idx = np.arange(npts, dtype=np.uint64)
    for c in range(ncomp):
        lin = (np.uint64(c) * np.uint64(npts) + idx) * np.uint64(2)
        with np.errstate(over="ignore"):
            h1 = splitmix64(base + lin)
            h2 = splitmix64(base + lin + np.uint64(1))
        u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
        u2 = (h2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
        out[c] = z.reshape(n, n, n)
    return out


def mode_index(n: int):
    m = np.arange(n)
    return np.where(m <= n // 2, m, m - n).astype(np.float64)


def envelope(n: int, kind: str, a: float):
    """E(|m|) on the half spectrum, shape (n, n, n//2+1)."""
    m = mode_index(n)
    mx = m[: n // 2 + 1].reshape(1, 1, -1)
    my = m.reshape(1, -1, 1)
    mz = m.reshape(-1, 1, 1)
    m2 = mx * mx + my * my + mz * mz
    if kind == "gauss":
        return np.exp(-m2 / a)
    if kind == "k2gauss":
        return m2 * np.exp(-m2 / a)
    raise ValueError(kind)


def seeded_field(n: int, seed: int, kind: str = "gauss", a: float = 16.0,
                 q: float = 4.0, B: float = 1.0):
    """Seeded divergence-free field rescaled to ||u||_q = B; shape (3,n,n,n)."""
    w = white_noise(n, seed)
    E = envelope(n, kind, a)
    uh = np.fft.rfftn(w, axes=(-3, -2, -1)) * E
    uh[:, 0, 0, 0] = 0.0
    pi = np.pi
    m = mode_index(n)
    K = 2.0 * pi * m
    Keff = np.where(np.arange(n) == n // 2, 0.0, K)   # own-dimension Nyquist -> 0
    Kx = K[: n // 2 + 1].reshape(1, 1, -1)
    Ky = K.reshape(1, -1, 1)
    Kz = K.reshape(-1, 1, 1)
    Kxe = Keff[: n // 2 + 1].reshape(1, 1, -1)
    Kye = Keff.reshape(1, -1, 1)
    Kze = Keff.reshape(-1, 1, 1)
    ksq = Kx * Kx + Ky * Ky + Kz * Kz
    ksq[0, 0, 0] = 1.0
    dot = Kxe * uh[0] + Kye * uh[1] + Kze * uh[2]
    uh = np.stack([uh[0] - Kxe * dot / ksq, uh[1] - Kye * dot / ksq,
                   uh[2] - Kze * dot / ksq])
    u = np.fft.irfftn(uh, s=(n, n, n), axes=(-3, -2, -1))
    orc = Oracle(Params(n=n, q=q, B=B))
    L = orc.calc_global_Lq_norm(u)
    return B / L * u
