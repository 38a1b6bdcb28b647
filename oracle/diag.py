"""diagnosticScalars and kappa_test_pert of the reference (numpy restatement;
TEST INFRASTRUCTURE ONLY -- imported by tests/, never by the product package).

Layout as in lq_oracle: a vector field is (3, n3, n2, n1), x fastest.

diagnosticScalars (function_ops.f90:2503-2602) returns the values it hands
to save_diagnosticFields_global (data_ops.f90:125-155), the four fields it
hands to save_diagnosticScalars and the ring points of calculate_ring_data.
kappa_test_pert (function_ops.f90:613-977) returns the perturbation; the
reference's random_number is replaced by the project's seeded uniform
generator (``uniform_pm1``), and its file loads by a field the caller passes.

Parity pins: ``PINS`` at the bottom.
"""
from __future__ import annotations

import numpy as np

from .lq_oracle import MACH_EPSILON, Oracle, fftbwd, fftfwd
from .seeded import GOLDEN, splitmix64

PERT_SINE, PERT_DIVFREE_SINE, PERT_LOAD_RAW, PERT_LOAD = 0, 1, 2, 3
PERT_LOAD_SMOOTH, PERT_LOAD_SMOOTH_RAW, PERT_LOAD_K = 4, 5, 6
PERT_LOAD_POLY, PERT_LOAD_EXP, PERT_RANDOM_POLY, PERT_RANDOM = 7, 8, 9, 10


# ---------------------------------------------------------------------------
# diagnosticScalars
# ---------------------------------------------------------------------------

def energy(orc: Oracle, U):
    """Energy(U), function_ops.f90:2394-2412: (1/2 sum U_c^2 dV)_c."""
    return np.array([float(np.sum(0.5 * U[c] ** 2 * orc.g.dV)) for c in range(3)])


def helicity(orc: Oracle, U, W):
    """Helicity(U, W), function_ops.f90:2479-2497: sum U.W dV."""
    return float(np.sum((U[0] * W[0] + U[1] * W[1] + U[2] * W[2]) * orc.g.dV))


def geometric_Q(orc: Oracle, U, W):
    """calculate_geometric_data(U, W, vortexCore, "Q"), function_ops.f90:
    2681-2832: the Q criterion from the component gradients (:2711-2716) and
    W (:2727-2731), clipped at 0; returns (vortexCore / max, [volume where
    Q > 0, max Q, 0]).  The third entry is never assigned by the reference."""
    gx, gy, gz = (orc.gradient(U[c]) for c in range(3))    # g?[j] = d_j u_?
    Q = np.maximum(0.25 * (W[0] ** 2 + W[1] ** 2 + W[2] ** 2)
                   - 0.25 * (gx[1] + gy[0]) ** 2
                   - 0.25 * (gx[2] + gz[0]) ** 2
                   - 0.25 * (gy[2] + gz[1]) ** 2
                   - 0.50 * (gx[0] ** 2 + gy[1] ** 2 + gz[2] ** 2), 0.0)
    vol = float(np.count_nonzero(Q > 0.0)) * orc.g.dV          # :2733-2735
    qmax = float(np.max(Q))                                     # :2746-2748
    with np.errstate(invalid="ignore", divide="ignore"):        # Q = 0 everywhere: 0/0
        return Q / qmax, np.array([vol, qmax, 0.0])


def ring_points(orc: Oracle, magW):
    """calculate_ring_data, function_ops.f90:2837-2900: coordinates
    ((i-1) dx, (j-1) dx, (k-1) dx) of the points with |w| / max |w| > 0.95 in
    the loop order (x fastest)."""
    n = orc.g.n
    dx = 1.0 / float(n)
    mask = magW / np.max(magW) > 0.95
    k, j, i = np.nonzero(mask)          # C order of (z, y, x): x fastest
    return np.stack([i * dx, j * dx, k * dx], axis=1).astype(np.float64)


def diagnosticScalars(orc: Oracle, U):
    """function_ops.f90:2503-2602.  Returns (dict, fields (4, n3, n2, n1),
    ring points (m, 3))."""
    W = orc.vel2vort(U)                                          # :2520
    magU = np.sqrt(U[0] ** 2 + U[1] ** 2 + U[2] ** 2)            # :2538
    magW = np.sqrt(W[0] ** 2 + W[1] ** 2 + W[2] ** 2)            # :2539
    hel = U[0] * W[0] + U[1] * W[1] + U[2] * W[2]                # :2541
    div = orc.divergence(U)                                      # :2545
    d = {
        "divU_L2": orc.inner_product(div, div),                  # :2546
        "K": energy(orc, U),                                     # :2549
        "E": energy(orc, W),                                     # :2551
        "Umax": np.array([float(np.max(U[c])) for c in range(3)]),   # :2554-2557
        "Wmax": np.array([float(np.max(W[c])) for c in range(3)]),   # :2559-2562
        "H": helicity(orc, U, W),                                # :2564
        "magUmax": float(np.max(magU)),                          # :2567-2569
        "magWmax": float(np.max(magW)),                          # :2570-2572
        "maxHel": float(np.max(hel)),                            # :2573-2575
        "minHel": float(np.min(hel)),                            # :2576-2577
    }
    ring = ring_points(orc, magW)                                # :2579
    core, d["vorCore"] = geometric_Q(orc, U, W)                  # :2581
    fields = np.stack([magU, magW, hel, core])                   # :2583
    return d, fields, ring


# ---------------------------------------------------------------------------
# kappa_test_pert
# ---------------------------------------------------------------------------

def uniform_pm1(n: int, seed: int):
    """2 x - 1, x uniform on [0, 1): for component c and linear point index i
    (x fastest) x = (SplitMix64(seed*0x9E3779B97F4A7C15 + c n^3 + i) >> 11)
    * 2^-53.  The role of ``call random_number(phi_pert); phi_pert =
    2 phi_pert - 1`` (function_ops.f90:798-799)."""
    npts = n ** 3
    with np.errstate(over="ignore"):
        base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * GOLDEN
    idx = np.arange(npts, dtype=np.uint64)
    out = np.empty((3, n, n, n))
    for c in range(3):
        with np.errstate(over="ignore"):
            h = splitmix64(base + np.uint64(c) * np.uint64(npts) + idx)
        x = (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        out[c] = (2.0 * x - 1.0).reshape(n, n, n)
    return out


def _reshape(orc: Oracle, phi, factor):
    """fftfwdv, faux *= factor(|K|) for |K| >= MACH_EPSILON and 0 below,
    fftbwdv, REAL (function_ops.f90:719-740 and the like)."""
    g = orc.g
    nk = g.kabs
    small = nk < MACH_EPSILON
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.where(small, 0.0, factor(np.where(small, 1.0, nk)))
    return np.stack([fftbwd(fftfwd(phi[c]) * f).real for c in range(3)])


def _poly(orc, m1, m2):
    n = orc.g.n
    return lambda nk: (10.0 ** m2) * (nk / (float(n) / 4.0)) ** m1      # :788


def _random_normalised(orc: Oracle, seed):
    """:797-823: uniform [-1, 1), zero mode removed, divAvg_free, divided by
    100 sum_c <phi_c, phi_c>_L2."""
    phi = uniform_pm1(orc.g.n, seed)
    phi = _reshape(orc, phi, lambda nk: np.ones_like(nk))
    phi = orc.divAvg_free(phi)
    c = 100.0 * orc.global_summed_field_inner_product(phi, phi)
    return phi / c


def kappa_test_pert(orc: Oracle, kind, m1=0.0, m2=0.0, m3=0.0, seed=0, loaded=None):
    """kappa_test_pert(phi_pert, mytype, m1, m2, m3), function_ops.f90:613-977."""
    g = orc.g
    n = g.n
    if PERT_LOAD_RAW <= kind <= PERT_LOAD_EXP:
        if loaded is None:
            raise ValueError("this kind needs the loaded field")
        phi = np.array(loaded, dtype=np.float64)
    dx = 1.0 / float(n)
    if kind in (PERT_SINE, PERT_DIVFREE_SINE):                 # :660-689
        i = np.arange(n, dtype=np.float64)
        X = (i * dx).reshape(1, 1, n)
        Y = (i * dx).reshape(1, n, 1)
        Z = (i * dx).reshape(n, 1, 1)
        b = m2 if kind == PERT_SINE else m1 + m3
        tp = 2.0 * g.pi
        sx, cx = np.sin(tp * m1 * X), np.cos(tp * m1 * X)
        sy, cy = np.sin(tp * b * Y), np.cos(tp * b * Y)
        sz, cz = np.sin(tp * m3 * Z), np.cos(tp * m3 * Z)
        phi = np.stack([sx * cy * cz, -cx * sy * cz, cx * cy * sz])
        return orc.divAvg_free(phi) if kind == PERT_SINE else phi
    if kind == PERT_LOAD_RAW:                                  # :655-658
        return phi
    if kind == PERT_LOAD:                                      # :692-702
        return orc.divAvg_free(phi)
    if kind in (PERT_LOAD_SMOOTH, PERT_LOAD_SMOOTH_RAW):       # :704-713, 927-935
        phi = np.stack([orc.dealias_scalar(phi[c], 20.0) for c in range(3)])
        return orc.divAvg_free(phi) if kind == PERT_LOAD_SMOOTH else phi
    if kind == PERT_LOAD_K:                                    # :715-741
        return orc.divAvg_free(_reshape(orc, phi, lambda nk: np.full_like(nk, m1)))
    if kind == PERT_LOAD_POLY:                                 # :769-795
        return orc.divAvg_free(_reshape(orc, phi, _poly(orc, m1, m2)))
    if kind == PERT_LOAD_EXP:                                  # :873-899
        return orc.divAvg_free(_reshape(orc, phi, lambda nk: 10.0 ** (-nk / 5.0)))
    if kind == PERT_RANDOM_POLY:                               # :797-844
        phi = _random_normalised(orc, seed)
        return orc.divAvg_free(_reshape(orc, phi, _poly(orc, m1, m2)))
    if kind == PERT_RANDOM:                                    # :937-966
        return _random_normalised(orc, seed)
    raise ValueError(f"unknown kind {kind}")


PINS = {
    "energy/helicity/diagnosticScalars (K, E, H, maxima, |u|, |w|, div)":
        "tests/test_oracle_diag.py::test_diag_beltrami_closed_form, "
        "test_diag_parseval, test_diag_divergent_field",
    "geometric_Q": "tests/test_oracle_diag.py::test_q_equals_minus_half_trace_a2, "
                   "test_q_taylor_green",
    "ring_points": "tests/test_oracle_diag.py::test_ring_arccos_count",
    "uniform_pm1": "tests/test_oracle_diag.py::test_uniform_known_answers",
    "kappa_test_pert": "tests/test_oracle_diag.py::test_pert_*",
}
