"""Pins for the oracle's diagnosticScalars and kappa_test_pert: closed forms,
Parseval sums, the tensor identity Q = -tr(A A)/2, an arccos count and
SplitMix64's published outputs -- none of them the oracle's own formula
(CPU only)."""
import math

import numpy as np

from oracle import diag as D
from oracle.lq_oracle import Oracle, Params
from oracle.seeded import splitmix64
from tests.fields import bandlimit, grid_xyz, mode_index, rel, spec_deriv

TP = 2.0 * np.pi


def _orc(n):
    return Oracle(Params(n=n, q=4.0))


def _random_band(n, M, seed):
    rng = np.random.default_rng(seed)
    return bandlimit(rng.standard_normal((3, n, n, n)), M)


def test_diag_beltrami_closed_form():
    # u = (sin 2 pi z, cos 2 pi z, 0): w = curl u = 2 pi u, div u = 0
    n = 16
    X, Y, Z = grid_xyz(n)
    u = np.stack([np.sin(TP * Z), np.cos(TP * Z), np.zeros_like(Z)])
    d, fields, ring = D.diagnosticScalars(_orc(n), u)
    pi2 = math.pi ** 2
    assert np.allclose(d["K"], [0.25, 0.25, 0.0], atol=1e-14)
    assert np.allclose(d["E"], [pi2, pi2, 0.0], rtol=1e-13, atol=1e-12)
    assert np.allclose(d["Umax"], [1.0, 1.0, 0.0], atol=1e-14)
    assert np.allclose(d["Wmax"], [TP, TP, 0.0], rtol=1e-13, atol=1e-12)
    assert abs(d["magUmax"] - 1.0) < 1e-14 and abs(d["magWmax"] - TP) < 1e-12
    assert abs(d["H"] - TP) < 1e-12
    assert abs(d["maxHel"] - TP) < 1e-12 and abs(d["minHel"] - TP) < 1e-12
    assert d["divU_L2"] < 1e-24
    assert np.allclose(fields[0], 1.0, atol=1e-14) and np.allclose(fields[1], TP, atol=1e-12)
    assert np.allclose(fields[2], TP, atol=1e-12)
    # |w| is constant: every point is a ring point, x fastest
    assert ring.shape == (n ** 3, 3)
    assert np.allclose(ring[:3], [[0, 0, 0], [1 / n, 0, 0], [2 / n, 0, 0]])
    assert np.allclose(ring[n], [0, 1 / n, 0]) and np.allclose(ring[n * n], [0, 0, 1 / n])


def test_diag_parseval():
    # K, E, H and ||div u||^2 as spectral sums with w_hat = i K x u_hat
    n, M = 16, 5
    u = _random_band(n, M, 3)
    d, fields, ring = D.diagnosticScalars(_orc(n), u)
    U = np.fft.fftn(u, axes=(-3, -2, -1))
    K = TP * mode_index(n)
    k1, k2, k3 = K[None, None, :], K[None, :, None], K[:, None, None]
    W = np.stack([1j * (k2 * U[2] - k3 * U[1]), 1j * (k3 * U[0] - k1 * U[2]),
                  1j * (k1 * U[1] - k2 * U[0])])
    N6 = float(n) ** 6
    assert rel(d["K"], 0.5 * np.sum(np.abs(U) ** 2, axis=(1, 2, 3)) / N6) < 1e-12
    assert rel(d["E"], 0.5 * np.sum(np.abs(W) ** 2, axis=(1, 2, 3)) / N6) < 1e-12
    assert abs(d["H"] - np.sum((U * np.conj(W)).real) / N6) < 1e-12 * np.sum(np.abs(U * W)) / N6
    divh = 1j * (k1 * U[0] + k2 * U[1] + k3 * U[2])
    assert abs(d["divU_L2"] - np.sum(np.abs(divh) ** 2) / N6) < 1e-11 * d["divU_L2"]
    w = np.fft.ifftn(W, axes=(-3, -2, -1)).real
    assert rel(d["Umax"], u.max(axis=(1, 2, 3))) < 1e-15
    assert rel(d["Wmax"], w.max(axis=(1, 2, 3))) < 1e-12
    assert abs(d["magUmax"] - np.linalg.norm(u, axis=0).max()) < 1e-14
    assert abs(d["magWmax"] - np.linalg.norm(w, axis=0).max()) < 1e-11 * d["magWmax"]
    h = np.einsum("c...,c...->...", u, w)
    assert abs(d["maxHel"] - h.max()) < 1e-11 * abs(h).max()
    assert abs(d["minHel"] - h.min()) < 1e-11 * abs(h).max()
    assert rel(fields[1], np.linalg.norm(w, axis=0)) < 1e-12


def test_diag_divergent_field():
    # u = (sin 2 pi x, 0, cos 2 pi x): div u = 2 pi cos 2 pi x, w = (0, 2 pi sin 2 pi x, 0)
    n = 16
    X, Y, Z = grid_xyz(n)
    u = np.stack([np.sin(TP * X), np.zeros_like(X), np.cos(TP * X)])
    with np.errstate(invalid="ignore", divide="ignore"):
        d, fields, ring = D.diagnosticScalars(_orc(n), u)
    assert abs(d["divU_L2"] - 2.0 * math.pi ** 2) < 1e-12
    assert np.allclose(d["E"], [0.0, math.pi ** 2, 0.0], atol=1e-12)
    assert abs(d["H"]) < 1e-13


def test_q_equals_minus_half_trace_a2():
    # Q = (|Omega|^2 - |S|^2) / 2 = -tr(A A) / 2 with A_ij = d_j u_i
    n, M = 16, 5
    u = _random_band(n, M, 11)
    d, fields, ring = D.diagnosticScalars(_orc(n), u)
    A = np.stack([np.stack([spec_deriv(u[i], j) for j in range(3)]) for i in range(3)])
    Qref = np.maximum(-0.5 * np.einsum("ij...,ji...->...", A, A), 0.0)
    qmax = Qref.max()
    assert abs(d["vorCore"][1] - qmax) < 1e-11 * qmax
    assert d["vorCore"][0] == np.count_nonzero(Qref > 0.0) / n ** 3
    assert d["vorCore"][2] == 0.0
    assert np.max(np.abs(fields[3] - Qref / qmax)) < 1e-11
    assert 0.05 < d["vorCore"][0] < 0.95


def test_q_taylor_green():
    # u = (sin a cos b, -cos a sin b, 0), a = 2 pi (x + 1/(3n)), b = 2 pi y:
    # Q = 4 pi^2 (sin^2 a sin^2 b - cos^2 a cos^2 b) = -4 pi^2 cos(a + b) cos(a - b),
    # never 0 on the grid; Q > 0 where the two cosines differ in sign
    n = 16
    X, Y, Z = grid_xyz(n)
    a, b = TP * (X + 1.0 / (3 * n)), TP * Y
    u = np.stack([np.sin(a) * np.cos(b), -np.cos(a) * np.sin(b), np.zeros_like(X)])
    d, fields, ring = D.diagnosticScalars(_orc(n), u)
    i = np.arange(n)
    s = TP * (i[None, :] + i[:, None] + 1.0 / 3.0) / n       # a + b on the (y, x) plane
    t = TP * (i[None, :] - i[:, None] + 1.0 / 3.0) / n       # a - b
    vol = np.count_nonzero(np.cos(s) * np.cos(t) < 0) / n ** 2
    assert d["vorCore"][0] == vol
    Qgrid = -4.0 * math.pi ** 2 * np.cos(s) * np.cos(t)
    assert abs(d["vorCore"][1] - Qgrid.max()) < 1e-11 * Qgrid.max()
    # unshifted, n % 4 == 0: the maximum 4 pi^2 sits on the grid (x = y = 1/4)
    u0 = np.stack([np.sin(TP * X) * np.cos(TP * Y), -np.cos(TP * X) * np.sin(TP * Y),
                   np.zeros_like(X)])
    d0, _, _ = D.diagnosticScalars(_orc(n), u0)
    assert abs(d0["vorCore"][1] - 4.0 * math.pi ** 2) < 1e-11


def test_ring_arccos_count():
    # u = (0, 0, sin 2 pi x): |w| = 2 pi |cos 2 pi x|; |cos t| > 0.95 within
    # arccos(0.95) of a multiple of pi
    n = 32
    X, Y, Z = grid_xyz(n)
    u = np.stack([np.zeros_like(X), np.zeros_like(X), np.sin(TP * X)])
    with np.errstate(invalid="ignore", divide="ignore"):
        d, fields, ring = D.diagnosticScalars(_orc(n), u)
    th = math.acos(0.95) / TP
    xs = [i for i in range(n) if min((i / n) % 0.5, 0.5 - (i / n) % 0.5) < th]
    assert xs == [0, 1, 15, 16, 17, 31]
    want = [(i / n, j / n, k / n) for k in range(n) for j in range(n) for i in xs]
    assert ring.shape == (len(want), 3)
    assert np.allclose(ring, want, atol=0)


def test_uniform_known_answers():
    # SplitMix64 (Steele, Lea, Flood 2014) from state 0: the first outputs are
    # 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F; output k is
    # splitmix64(k * golden) here
    g = 0x9E3779B97F4A7C15
    xs = np.array([0, g, (2 * g) & (2 ** 64 - 1)], dtype=np.uint64)
    assert [int(v) for v in splitmix64(xs)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                               0x06C45D188009454F]
    # seed 1: component 0 point i draws splitmix64(golden + i); seed 0 draws splitmix64(i)
    u1 = D.uniform_pm1(4, 1)
    x = (0x6E789E6AA1B965F4 >> 11) * 2.0 ** -53
    assert u1[0, 0, 0, 0] == 2.0 * x - 1.0
    u0 = D.uniform_pm1(4, 0)
    assert u0[0, 0, 0, 0] == 2.0 * ((0xE220A8397B1DCDAF >> 11) * 2.0 ** -53) - 1.0
    big = D.uniform_pm1(16, 7)
    assert big.min() >= -1.0 and big.max() < 1.0
    assert abs(big.mean()) < 0.03 and abs(big.var() - 1.0 / 3.0) < 0.02
    assert not np.array_equal(big[0], big[1])


def _sine(n, m1, m2, m3):
    X, Y, Z = grid_xyz(n)
    sx, cx = np.sin(TP * m1 * X), np.cos(TP * m1 * X)
    sy, cy = np.sin(TP * m2 * Y), np.cos(TP * m2 * Y)
    sz, cz = np.sin(TP * m3 * Z), np.cos(TP * m3 * Z)
    return sx, cx, sy, cy, sz, cz


def _spec_div(f):
    return sum(spec_deriv(f[j], j) for j in range(3))


def _no_nyquist(f):
    # the Leray projection with the reference's +pi N Nyquist wavenumber
    # leaves a real field's Nyquist planes not divergence-free; drop them
    n = f.shape[-1]
    F = np.fft.fftn(f, axes=(-3, -2, -1))
    F[..., n // 2, :, :] = 0
    F[..., :, n // 2, :] = 0
    F[..., n // 2] = 0
    return np.fft.ifftn(F, axes=(-3, -2, -1)).real


def test_pert_sine_closed_form():
    # "sine" then divAvg_free: the Leray projection of a single mode triple in
    # closed form, s = m1 - m2 + m3, |m|^2 = m1^2 + m2^2 + m3^2:
    # (sx cy cz (1 - m1 s/|m|^2), -cx sy cz (1 + m2 s/|m|^2), cx cy sz (1 - m3 s/|m|^2))
    n = 16
    orc = _orc(n)
    for m in ((2, 3, 1), (1, 4, 0), (3, 1, 2)):
        m1, m2, m3 = m
        sx, cx, sy, cy, sz, cz = _sine(n, m1, m2, m3)
        s, mm = m1 - m2 + m3, m1 * m1 + m2 * m2 + m3 * m3
        want = np.stack([sx * cy * cz * (1 - m1 * s / mm), -cx * sy * cz * (1 + m2 * s / mm),
                         cx * cy * sz * (1 - m3 * s / mm)])
        got = D.kappa_test_pert(orc, D.PERT_SINE, *map(float, m))
        assert np.max(np.abs(got - want)) < 1e-13
    # "divfree-sine": m2 = m1 + m3, the field itself, divergence-free
    sx, cx, sy, cy, sz, cz = _sine(n, 2, 5, 3)
    got = D.kappa_test_pert(orc, D.PERT_DIVFREE_SINE, 2.0, 99.0, 3.0)
    assert np.max(np.abs(got - np.stack([sx * cy * cz, -cx * sy * cz, cx * cy * sz]))) < 1e-14
    assert np.max(np.abs(_spec_div(got))) < 1e-11


def test_pert_load_kinds():
    n = 32
    orc = _orc(n)
    rng = np.random.default_rng(4)
    L = bandlimit(rng.standard_normal((3, n, n, n)), 6) + np.array([0.3, -0.2, 0.1])[:, None, None, None]
    assert np.array_equal(D.kappa_test_pert(orc, D.PERT_LOAD_RAW, loaded=L), L)
    P = D.kappa_test_pert(orc, D.PERT_LOAD, loaded=L)
    assert np.max(np.abs(_spec_div(_no_nyquist(P)))) < 1e-10 and np.max(np.abs(P.mean(axis=(1, 2, 3)))) < 1e-15
    # Leray projection is idempotent and removes only a gradient: P(L) - L is curl-free
    assert np.max(np.abs(D.kappa_test_pert(orc, D.PERT_LOAD, loaded=P) - P)) < 1e-13
    # "load-k-random": factor m1 on every mode but k = 0, then the projection
    K7 = D.kappa_test_pert(orc, D.PERT_LOAD_K, 7.0, loaded=L)
    assert np.max(np.abs(K7 - 7.0 * P)) < 1e-12
    # single-mode field (divergence-free, |m| = sqrt(6)): poly and exp
    sx, cx, sy, cy, sz, cz = _sine(n, 1, 2, 1)
    f = np.stack([sx * cy * cz, -cx * sy * cz, cx * cy * sz])
    kk = TP * math.sqrt(6.0)
    poly = D.kappa_test_pert(orc, D.PERT_LOAD_POLY, -2.0, 1.0, loaded=f)
    assert np.max(np.abs(poly - 10.0 * (kk / (n / 4.0)) ** -2.0 * f)) < 1e-13
    ex = D.kappa_test_pert(orc, D.PERT_LOAD_EXP, loaded=f)
    assert np.max(np.abs(ex - 10.0 ** (-kk / 5.0) * f)) < 1e-14
    # order-20 dealias: k_cut = 2 pi n / 21 keeps max |m_j| <= 1 at n = 32
    X, Y, Z = grid_xyz(n)
    low = np.stack([np.sin(TP * X) * np.cos(TP * Y), -np.cos(TP * X) * np.sin(TP * Y), 0 * X])
    grad_low = np.stack([np.cos(TP * X), 0 * X, 0 * X])         # a gradient, kept by the mask
    high = np.stack([0 * X, np.sin(TP * 2 * X), np.cos(TP * 3 * Z)])
    c = np.array([0.5, 0.0, -0.25])[:, None, None, None]
    src = low + grad_low + high + c
    sm = D.kappa_test_pert(orc, D.PERT_LOAD_SMOOTH, loaded=src)
    assert np.max(np.abs(sm - low)) < 1e-13
    smr = D.kappa_test_pert(orc, D.PERT_LOAD_SMOOTH_RAW, loaded=src)
    assert np.max(np.abs(smr - (low + grad_low + c))) < 1e-13


def test_pert_random_kinds():
    n = 16
    orc = _orc(n)
    R = D.kappa_test_pert(orc, D.PERT_RANDOM, seed=5)
    # the projection alone already removes the zero mode
    phi0 = orc.divAvg_free(D.uniform_pm1(n, 5))
    s0 = float(np.sum(phi0 ** 2)) / n ** 3
    assert np.max(np.abs(R - phi0 / (100.0 * s0))) < 1e-15 * np.max(np.abs(R)) * 10
    assert abs(float(np.sum(R ** 2)) / n ** 3 * s0 - 1e-4) < 1e-15
    RP = D.kappa_test_pert(orc, D.PERT_RANDOM_POLY, -5.0, 0.0, seed=5)
    assert np.max(np.abs(RP - D.kappa_test_pert(orc, D.PERT_LOAD_POLY, -5.0, 0.0, loaded=R))) == 0.0
    assert np.max(np.abs(_spec_div(_no_nyquist(RP)))) < 1e-12 * np.max(np.abs(RP)) * n
    assert np.max(np.abs(RP.mean(axis=(1, 2, 3)))) < 1e-18
    # the poly factor steepens the spectrum: the top band holds far less
    # energy after it than before
    Fr = np.abs(np.fft.fftn(R, axes=(-3, -2, -1))) ** 2
    Fp = np.abs(np.fft.fftn(RP, axes=(-3, -2, -1))) ** 2
    m = np.abs(mode_index(n))
    top = (m[:, None, None] >= 6) | (m[None, :, None] >= 6) | (m[None, None, :] >= 6)
    assert Fp[:, top].sum() / Fp.sum() < 1e-3 * Fr[:, top].sum() / Fr.sum()
