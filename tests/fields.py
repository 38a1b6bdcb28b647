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


# ---------------------------------------------------------------------------
# Exact dealiased products.  For fields band-limited inside the 2/3 cube
# (max_j |m_j| <= mcut = floor(n/3), n a power of two), the product of two
# of them has modes up to 2 mcut < n - mcut, so on the n grid its aliases land
# outside the cube and the cube mask leaves exactly the mask of the
# unaliased product.  These helpers form that product on a 2n grid (no
# aliasing at all), apply the cube mask and return the field on the n grid:
# an evaluation independent of the oracle's dealias code and transforms.
# ---------------------------------------------------------------------------

def cube_mcut(n):
    """Largest m with 2 pi m <= 2 pi n/3 (the order-2 mask, function_ops.f90:3374)."""
    return int(np.floor(n / 3.0 + 1e-12))


def up2(f):
    """Values of a field with no Nyquist content on the 2n grid (spectral
    zero padding)."""
    n = f.shape[-1]
    F = np.fft.fftn(f, axes=(-3, -2, -1))
    m = mode_index(n)
    if np.any(F[..., n // 2, :, :]) or np.any(F[..., :, n // 2, :]) or np.any(F[..., n // 2]):
        assert np.max(np.abs(F[..., n // 2, :, :])) < 1e-9 * np.max(np.abs(F))
    p = np.where(m >= 0, m, m + 2 * n)
    G = np.zeros(f.shape[:-3] + (2 * n,) * 3, np.complex128)
    G[..., p[:, None, None], p[None, :, None], p[None, None, :]] = F
    return np.fft.ifftn(G, axes=(-3, -2, -1)).real * 8.0


def mask_down(h2, n, mcut):
    """Cube mask max_j |m_j| <= mcut of a field on the 2n grid, returned on the n grid."""
    H = np.fft.fftn(h2, axes=(-3, -2, -1))
    m = mode_index(n)
    m2 = np.where(m >= 0, m, m + 2 * n)
    keep = (np.abs(m) <= mcut)
    F = H[..., m2[:, None, None], m2[None, :, None], m2[None, None, :]] / 8.0
    F = F * (keep[:, None, None] & keep[None, :, None] & keep[None, None, :])
    return np.fft.ifftn(F, axes=(-3, -2, -1)).real


def spec_deriv(f, j):
    """d/dx_j of a band-limited field (no Nyquist content), K = 2 pi m; j = 0
    is x (the last axis), 2 is z (the first)."""
    n = f.shape[-1]
    F = np.fft.fftn(f, axes=(-3, -2, -1))
    K = 2.0 * np.pi * mode_index(n)
    shape = [1, 1, 1]
    shape[2 - j] = n
    return np.fft.ifftn(1j * K.reshape(shape) * F, axes=(-3, -2, -1)).real


def masked_chain_field(n, seed=5, eps=0.25):
    """u = (sin 2 pi z, cos 2 pi z, 0) + eps w: |(sin, cos, 0)| = 1, w a seeded
    divergence-free zero-mean field band-limited to max_j |m_j| <= mcut - 1
    (mcut = floor(n/3)) with max |w| = 1.  u.u = 1 + 2 eps (..).w + eps^2 w.w
    has modes beyond the cube (truncated) and beyond n/2 (aliased on the n
    grid) while D(u.u) stays positive."""
    M = cube_mcut(n) - 1
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((3, n, n, n))
    F = np.fft.fftn(w, axes=(-3, -2, -1))
    m = mode_index(n)
    band = np.abs(m) <= M
    F *= (band[:, None, None] & band[None, :, None] & band[None, None, :])
    K = 2.0 * np.pi * m
    K1, K2, K3 = K[None, None, :], K[None, :, None], K[:, None, None]
    ksq = K1 ** 2 + K2 ** 2 + K3 ** 2
    ksq[0, 0, 0] = 1.0
    div = K1 * F[0] + K2 * F[1] + K3 * F[2]
    F[0] -= K1 * div / ksq
    F[1] -= K2 * div / ksq
    F[2] -= K3 * div / ksq
    F[:, 0, 0, 0] = 0.0
    w = np.fft.ifftn(F, axes=(-3, -2, -1)).real
    w /= np.max(np.abs(w))
    X, Y, Z = grid_xyz(n)
    u0 = np.stack([np.sin(2 * np.pi * Z), np.cos(2 * np.pi * Z), np.zeros_like(Z)])
    return u0 + eps * w


def exact_chain(u):
    """Exact values of the dealiased quantities of R's chain for a field whose
    products stay alias-free under the 2/3 mask (masked_chain_field):
    S = D(u.u), the gradient, vg_j = D(sum_i u_i d_j u_i),
    calc_vgraduvgradu(u, u) = D(sum_j vg_j^2), calc_vvgradu(u, u) = D(sum_j u_j vg_j),
    D(|grad u|^2),
    D(grad u : grad u^T), D(S S) and the pressure p = -Lap^-1 D(grad u : grad u^T)."""
    n = u.shape[-1]
    mc = cube_mcut(n)
    uu = [up2(u[i]) for i in range(3)]
    G = [[spec_deriv(u[i], j) for j in range(3)] for i in range(3)]   # G[i][j] = d_j u_i
    GG = [[up2(G[i][j]) for j in range(3)] for i in range(3)]
    S = mask_down(sum(uu[i] * uu[i] for i in range(3)), n, mc)
    vg = [mask_down(sum(uu[i] * GG[i][j] for i in range(3)), n, mc) for j in range(3)]
    vgu = [up2(v) for v in vg]
    vgvg = mask_down(sum(v * v for v in vgu), n, mc)
    # calc_vvgradu (function_ops.f90:2027-2056) dealiases aux_j = sum_i v_i d_j u_i
    # (gradU(:,:,:,ii,jj) = d_j u_i), not (v.grad) u_j: with v = u, aux_j = vg_j
    vvgradu = mask_down(sum(uu[j] * vgu[j] for j in range(3)), n, mc)
    gradsq = mask_down(sum(GG[i][j] * GG[i][j] for i in range(3) for j in range(3)), n, mc)
    ggT = mask_down(sum(GG[j][i] * GG[i][j] for i in range(3) for j in range(3)), n, mc)
    Su = up2(S)
    SS = mask_down(Su * Su, n, mc)
    K = 2.0 * np.pi * mode_index(n)
    ksq = K[None, None, :] ** 2 + K[None, :, None] ** 2 + K[:, None, None] ** 2
    Fg = np.fft.fftn(ggT)
    inv = np.where(ksq > 0, 1.0 / np.where(ksq > 0, ksq, 1.0), 0.0)
    p = np.fft.ifftn(Fg * inv).real                    # p = -Lap^-1 (ggT), zero mean
    return dict(S=S, G=G, vg=vg, vgraduvgradu=vgvg, vvgradu=vvgradu, gradsq=gradsq, ggT=ggT,
                SS=SS, p=p, mcut=mc)


def exact_R(u, q, nu=1.0):
    """(R, R1+R2, R3) of calc_local_dLqdt_inclParts (function_ops.f90:1304-1358)
    for q = 4 or 6 from the exact chain: |u|^(q-2) = S (q = 4) or D(S S)
    (q = 6), |u|^(q-4) = 1 or S, and the dealiased products above."""
    n = u.shape[-1]
    E = exact_chain(u)
    dV = 1.0 / n ** 3
    if abs(q - 4.0) < 1e-14:
        A, Bq = E["S"], np.ones_like(E["S"])
    elif abs(q - 6.0) < 1e-14:
        A, Bq = E["SS"], E["S"]
    else:
        raise ValueError("q = 4 or 6")
    R1 = -nu * np.sum(A * E["gradsq"]) * dV
    R2 = -nu * (q - 2.0) * np.sum(Bq * E["vgraduvgradu"]) * dV
    pB = mask_down(up2(E["p"]) * up2(Bq), n, E["mcut"])
    R3 = (q - 2.0) * np.sum(pB * E["vvgradu"]) * dV
    return np.array([R1 + R2 + R3, R1 + R2, R3]), E
