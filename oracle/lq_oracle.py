"""FP64 numpy restatement of the reference RCG hot path (test oracle).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Layout: a scalar field is a C-ordered array of shape (n3, n2, n1), i.e. the
reference's Fortran column-major ``(1:n(1),1:n(2),1:local_N)`` with x fastest;
a vector field is (3, n3, n2, n1), the reference's ``(...,1:3)``.

Every routine follows the reference operation order: all transforms are
complex-to-complex FP64 (fftwfunction.f90:62-113), every inverse is taken as
``REAL(ifft(.))``, and each ``dealias_scalar`` is a full forward/backward round
trip (function_ops.f90:3332-3350).  Citations are to the reference files.

Parity pins (tests that tie each function to something other than itself)
are listed in ``PINS`` at the bottom of this file.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

try:  # pocketfft with threads; numpy's pocketfft is the fallback
    import scipy.fft as _sfft

    def _fftn(a):
        return _sfft.fftn(a, axes=(-3, -2, -1), workers=_WORKERS)

    def _ifftn(a):
        return _sfft.ifftn(a, axes=(-3, -2, -1), workers=_WORKERS)
except ImportError:  # pragma: no cover
    def _fftn(a):
        return np.fft.fftn(a, axes=(-3, -2, -1))

    def _ifftn(a):
        return np.fft.ifftn(a, axes=(-3, -2, -1))

_WORKERS = int(os.environ.get("MLQ_ORACLE_THREADS", "1"))


def set_threads(n: int) -> None:
    """Number of host threads the oracle's FFTs may use."""
    global _WORKERS
    _WORKERS = max(1, int(n))


# global_variables.f90:27-29
MACH_EPSILON = 2.0e-16
TAU_MAX = 1.0e2

# optimization.f90:803-806: REAL(pr), PARAMETER :: TOL = 1E-6 etc. are
# default-real (single precision) literals promoted to double.
BRENT_ITMAX = 300
BRENT_TOL = float(np.float32(1e-6))
BRENT_ZEPS = float(np.float32(1e-8))
BRENT_CGOLD = float(np.float32(0.381966))
# optimization.f90:630-634 (double precision)
MNBRAK_GOLD = (1.0 + math.sqrt(5.0)) / 2.0
MNBRAK_CGOLD = 1.0 / MNBRAK_GOLD
MNBRAK_GLIMIT = 10.0
MNBRAK_ITMAX = 100

INNER_L2 = "L2"
INNER_H = "H^(3/2-1/q)"
INNER_HL = "H_l^(3/2-1/q)"
_INNER_OLD = ("H^((3q-1)/(2q))", "H_l^((3q-1)/(2q))")


class OracleStop(RuntimeError):
    """The reference would ``stop 1`` / ``call exit`` here."""


@dataclass
class Params:
    """The hot-path subset of global_variables.f90 (values as shipped)."""

    n: int
    q: float = 4.0                      # lebesgueQ
    B: float = 1.0                      # constraintB (global_variables.f90:115)
    visc: float = 1.0                   # :8
    lambda1: float = 0.1                # :31
    toDealias: bool = True              # :68
    # set_e_u_or_just_u (initialize.f90:75-86) with use_e_u_auto_for_q_less_4
    use_e_u: bool | None = None
    dealiase_if_mult_by_e_u: bool = True  # :40
    viscCoefficient: float = 1.0        # :113
    pressureCoefficient: float = 1.0    # :114
    # RCG switches (global_variables.f90:26-36)
    MAX_ITER: int = 99999
    OPTIM_TOL: float | None = None      # initialize.f90:95-100 when None
    useTangentialGradient: bool = True
    useConjugateGradient: bool = True
    useRiemannianGeometry: bool = True
    resetMomentumTermEveryXiterations: int = 25
    toleranceRollingAverageSize: int = 10
    normalizeGradient: bool = True
    checkDivergenceTolerance: float = float(np.float32(1.0e-15))
    checkNormalTolerance: float = float(np.float32(1.0e-12))
    checkAverageTolerance: float = float(np.float32(1.0e-3))

    def __post_init__(self):
        if self.use_e_u is None:
            # initialize.f90:79-85
            self.use_e_u = self.q < 4.0 - MACH_EPSILON
        if self.OPTIM_TOL is None:
            self.OPTIM_TOL = 1.0e-4 if self.n >= 1024 else 1.0e-5


class Grid:
    """Wavenumbers and cell volume of initialize.f90:131-203 (single rank)."""

    def __init__(self, n: int):
        self.n = int(n)
        pi = 4.0 * math.atan2(1.0, 1.0)                 # initialize.f90:164
        i = np.arange(self.n, dtype=np.float64)
        k = np.where(np.arange(self.n) <= self.n // 2, 2.0 * pi * i,
                     2.0 * pi * (i - self.n))           # :170-192
        self.pi = pi
        self.k1d = k
        self.K1 = k.reshape(1, 1, -1)
        self.K2 = k.reshape(1, -1, 1)
        self.K3 = k.reshape(-1, 1, 1)
        self.ksq = self.K1 ** 2 + self.K2 ** 2 + self.K3 ** 2
        self.kabs = np.sqrt(self.ksq)
        self.kmaxabs = np.maximum(np.maximum(np.abs(self.K1), np.abs(self.K2)),
                                  np.abs(self.K3))
        self.dV = 1.0 / float(self.n) ** 3            # :162
        self.shape = (self.n, self.n, self.n)


# ---------------------------------------------------------------------------
# fftwfunction.f90
# ---------------------------------------------------------------------------

def fftfwd(u):
    """fftwfunction.f90:62-92: unnormalised forward C2C, NaN scan -> stop 1."""
    fu = _fftn(np.asarray(u, dtype=np.complex128))
    if np.isnan(fu.real).any():                    # :78-89
        raise OracleStop("nan in fftwfunction")
    return fu


def fftbwd(fu):
    """fftwfunction.f90:97-113: backward C2C divided by product(n)."""
    return _ifftn(fu)


# ---------------------------------------------------------------------------
# function_ops.f90: spectral primitives
# ---------------------------------------------------------------------------

class Oracle:
    """Reference field routines bound to one grid and parameter set."""

    def __init__(self, params: Params):
        self.p = params
        self.g = Grid(params.n)

    # -- dealiasing ---------------------------------------------------------
    def dealias_mask(self, nonlinOrder: float):
        """dealiasing_cutoff_scalar_complex, function_ops.f90:3358-3398."""
        g = self.g
        k_cut = 2.0 * g.pi * float(g.n) / (nonlinOrder + 1.0)   # :3374
        return ~(g.kmaxabs > k_cut)                              # :3391

    def dealias_scalar(self, f, nonlinOrder=2.0):
        """function_ops.f90:3332-3350."""
        fh = fftfwd(f)
        fh = fh * self.dealias_mask(nonlinOrder)
        return fftbwd(fh).real

    def _D(self, f, order=2.0):
        return self.dealias_scalar(f, order) if self.p.toDealias else f

    # -- derivatives ----------------------------------------------------------
    def gradient(self, u):
        """function_ops.f90:3147-3195: REAL(ifft(iK_j fft(u)))."""
        g = self.g
        fu = fftfwd(u)
        return np.stack([fftbwd(1j * g.K1 * fu).real,
                         fftbwd(1j * g.K2 * fu).real,
                         fftbwd(1j * g.K3 * fu).real])

    def divergence(self, u):
        """function_ops.f90:3200-3243."""
        g = self.g
        uh = [fftfwd(u[i]) for i in range(3)]
        dh = 1j * g.K1 * uh[0] + 1j * g.K2 * uh[1] + 1j * g.K3 * uh[2]
        return fftbwd(dh).real

    def laplacian(self, f):
        """function_ops.f90:3248-3282 (vector, returns a new array)."""
        g = self.g
        return np.stack([fftbwd(-g.ksq * fftfwd(f[i])).real for i in range(3)])

    def solve_poisson(self, f, l=1.0):
        """function_ops.f90:1951-1996: l^2 Delta phi = f, k=0 mode -> 0."""
        g = self.g
        fh = fftfwd(f)
        with np.errstate(divide="ignore", invalid="ignore"):
            ph = np.where(g.ksq == 0, 0.0, -fh / ((l ** 2.0) * g.ksq))
        return fftbwd(ph).real

    def divAvg_free(self, f):
        """function_ops.f90:2843-2918: Leray projection + mean removal."""
        g = self.g
        divU = self.divergence(f)
        dh = fftfwd(divU)
        Uh = [fftfwd(f[i]) for i in range(3)]
        big = g.ksq > MACH_EPSILON                                  # :2883
        with np.errstate(divide="ignore", invalid="ignore"):
            gp = [np.where(big, -(1j * Kj) * dh / g.ksq, 0.0)
                  for Kj in (g.K1, g.K2, g.K3)]
        out = []
        for i in range(3):
            Ui = np.where(big, Uh[i], 0.0)                          # :2891-2893
            out.append(fftbwd(Ui).real - fftbwd(gp[i]).real)         # :2899-2909
        return np.stack(out)

    def HilbertGradient(self, grad, order):
        """function_ops.f90:3583-3627: divide by 1 + (lambda|K|)^(2 order)."""
        g = self.g
        w = 1.0 / (1.0 + (self.p.lambda1 * g.kabs) ** (2.0 * order))
        return np.stack([fftbwd(fftfwd(grad[i]) * w).real for i in range(3)])

    def calc0thFourierModes(self, v):
        """function_ops.f90:3048-3094: unnormalised k=0 coefficient (real)."""
        return np.array([fftfwd(v[i])[0, 0, 0].real for i in range(3)])

    # -- inner products -------------------------------------------------------
    def inner_product(self, f, g_, mytype=INNER_L2):
        """function_ops.f90:3404-3503."""
        G = self.g
        if mytype == INNER_L2:
            return float(np.sum(f * g_)) * G.dV                    # :3427-3437
        if mytype in (INNER_H, INNER_HL):
            order = 1.5 - 1.0 / self.p.q
            if mytype == INNER_H:
                factor = (1.0 + G.kabs) ** order                   # :3450
            else:
                factor = 1.0 + (self.p.lambda1 * G.kabs) ** order  # :3472
            fa = fftbwd(factor * fftfwd(f)).real
            ga = fftbwd(factor * fftfwd(g_)).real
            return self.inner_product(fa, ga, INNER_L2)
        if mytype in _INNER_OLD:
            raise OracleStop("old inner product exponent")         # :3485-3493
        raise ValueError(f"inner product type {mytype!r} not defined")

    def global_inner_product(self, f, g_, mytype=INNER_L2):
        """function_ops.f90:3508-3521 (single rank: the allreduce is a copy)."""
        return self.inner_product(f, g_, mytype)

    def global_summed_field_inner_product(self, f, g_, mytype=INNER_L2):
        """function_ops.f90:3557-3577: sum of the 3 component products."""
        s = 0.0
        for i in range(3):
            s = s + self.global_inner_product(f[i], g_[i], mytype)
        return s

    # -- diagnostics ------------------------------------------------------------
    def vel2vort(self, u):
        """vel2vort, function_ops.f90:2982-3042: REAL(ifft(i K x u_hat))."""
        g = self.g
        ux, uy, uz = (fftfwd(u[c]) for c in range(3))                  # :3004-3012
        wx = 1j * g.K2 * uz - 1j * g.K3 * uy                             # :3017
        wy = 1j * g.K3 * ux - 1j * g.K1 * uz                             # :3018
        wz = 1j * g.K1 * uy - 1j * g.K2 * ux                             # :3019
        return np.stack([fftbwd(wx).real, fftbwd(wy).real, fftbwd(wz).real])   # :3026-3031

    def calc_global_summed_enstrophy(self, u):
        """calc_global_summed_enstrophy, function_ops.f90:2455-2467, over
        Enstrophy (:2422-2442): sum_c 1/2 sum W_c^2 dV."""
        w = self.vel2vort(u)
        return float(sum(0.5 * float(np.sum(w[c] ** 2)) * self.g.dV for c in range(3)))

    def calc_global_dEnstrophydt(self, u):
        """calc_global_dEnstrophydt, function_ops.f90:1417-1481."""
        p = self.p
        lap = self.laplacian(u)                                           # :1429-1430
        r1 = -p.viscCoefficient * p.visc * self.global_summed_field_inner_product(lap, lap)  # :1433
        gradU = [self.gradient(u[i]) for i in range(3)]                  # :1440-1446, [i][j] = d_j u_i
        adv = np.zeros_like(u)
        for i in range(3):                                                # :1456-1461
            for j in range(3):
                adv[i] = adv[i] + u[j] * gradU[i][j]
        if p.toDealias:                                                   # :1465-1469
            adv = np.stack([self.dealias_scalar(adv[i], 2.0) for i in range(3)])
        r2 = self.global_summed_field_inner_product(adv, lap)             # :1472
        return r1 + r2

    def calculate_spectral_data(self, u, normalize=True):
        """calculate_spectral_data, function_ops.f90:2614-2677: the (n, 2)
        array of shell upper bounds 2 pi (i+1) and shell sums of
        |fftfwd(u_c)|^2 over the three components and the modes with
        kk_min < |K| <= kk_max (:2640-2659); with normalize
        (normalizeSpectrumByL2Norm, global_variables.f90:76) the sums are
        scaled by ||u||_2^2 / (2 pi sum) (:2667-2671)."""
        g = self.g
        e = sum(np.abs(fftfwd(u[c])) ** 2 for c in range(3))      # :2634-2638, :2653
        out = np.zeros((g.n, 2))
        for i in range(g.n):
            kk_min = 2.0 * g.pi * float(i)                          # :2641
            kk_max = 2.0 * g.pi * float(i + 1)                      # :2642
            out[i, 0] = kk_max                                      # :2643
            sel = (kk_min < g.kabs) & (g.kabs <= kk_max)            # :2651-2652
            out[i, 1] = float(e[sel].sum())
        if normalize:                                               # :2667-2671
            l2 = self.global_summed_field_inner_product(u, u, INNER_L2)
            spectral_ener = 2.0 * g.pi * float(out[:, 1].sum())
            out[:, 1] = l2 / spectral_ener * out[:, 1]
        return out

    # -- pointwise nonlinear helpers -----------------------------------------
    def calc_pointwiseVectorNorm(self, u):
        """function_ops.f90:1775-1791."""
        nrm = u[0] * u[0] + u[1] * u[1] + u[2] * u[2]
        nrm = self._D(nrm, 2.0)
        return np.sqrt(np.abs(nrm))

    def calc_unitVectorInUdirection(self, u):
        """function_ops.f90:1822-1858 (NaN -> stop 1)."""
        un = self.calc_pointwiseVectorNorm(u)
        small = un < MACH_EPSILON
        with np.errstate(divide="ignore", invalid="ignore"):
            e = np.stack([np.where(small, 0.0, u[i] / un) for i in range(3)])
        if np.isnan(e).any():
            raise OracleStop("unitvector is nan")
        return e

    def _gradU(self, u):
        """gradU[i][j] = partial_j u_i (function_ops.f90:1927-1931 etc.)."""
        return np.stack([self.gradient(u[i]) for i in range(3)])

    def calc_nablaUnablaUt(self, u, gradU=None):
        """function_ops.f90:1912-1945: sum_ij d_i u_j d_j u_i, dealiased."""
        G = self._gradU(u) if gradU is None else gradU
        acc = np.zeros_like(u[0])
        for ii in range(3):
            for jj in range(3):
                acc = acc + G[jj, ii] * G[ii, jj]
        return self._D(acc, 2.0)

    def _inner_dealias(self):
        p = self.p
        if not p.toDealias:
            return False
        return (not p.use_e_u) or p.dealiase_if_mult_by_e_u

    def calc_vvgradu(self, u, v, gradU=None):
        """function_ops.f90:2001-2064: v.(v.grad)u."""
        G = self._gradU(u) if gradU is None else gradU
        aux = []
        for jj in range(3):
            a = np.zeros_like(u[0])
            for ii in range(3):
                a = a + v[ii] * G[ii, jj]
            aux.append(a)
        if self._inner_dealias():
            aux = [self.dealias_scalar(a, 2.0) for a in aux]
        res = np.zeros_like(u[0])
        for jj in range(3):
            res = res + v[jj] * aux[jj]
        if self._inner_dealias():
            res = self.dealias_scalar(res, 2.0)
        return res

    def calc_vgraduvgradu(self, u, v, gradU=None):
        """function_ops.f90:2165-2218: (v_i d_j u_i)(v_k d_j u_k)."""
        G = self._gradU(u) if gradU is None else gradU
        vg = []
        for jj in range(3):
            a = np.zeros_like(u[0])
            for ii in range(3):
                a = a + v[ii] * G[ii, jj]
            vg.append(a)
        if self._inner_dealias():
            vg = [self.dealias_scalar(a, 2.0) for a in vg]
        res = np.zeros_like(u[0])
        for jj in range(3):
            res = res + vg[jj] * vg[jj]
        return self._D(res, 2.0)

    def calc_nablaModSquared(self, u, gradU=None):
        """function_ops.f90:2366-2389: |grad u|^2, dealiased."""
        G = self._gradU(u) if gradU is None else gradU
        acc = np.zeros_like(u[0])
        for ii in range(3):
            for jj in range(3):
                acc = acc + G[ii, jj] * G[ii, jj]
        return self._D(acc, 2.0)

    def calc_gk_order2(self, g, k):
        """function_ops.f90:2281-2319: g^k with products of order <= 2."""
        if k + MACH_EPSILON < 0:                                       # :2291
            with np.errstate(divide="ignore", invalid="ignore"):
                return np.where(np.abs(g) < MACH_EPSILON, 0.0, g ** k)
        if k + MACH_EPSILON < 2.0:                                     # :2302
            gk = g ** k
            return self._D(gk, k)
        g2 = self._D(g * g, 2.0)                                       # :2309-2310
        gkm2 = self.calc_gk_order2(g, k - 2.0)
        return self._D(g2 * gkm2, 2.0)                                 # :2312-2313

    def calc_uk(self, u, k):
        """function_ops.f90:2324-2343: |u|^k via sqrt(dealias(u.u))."""
        uk = u[0] ** 2 + u[1] ** 2 + u[2] ** 2
        uk = self._D(uk, 2.0)
        uk = np.where(uk <= 0.0, 0.0, np.sqrt(np.maximum(uk, 0.0)))   # :2335-2339
        return self.calc_gk_order2(uk, k)

    # -- constraint geometry ----------------------------------------------------
    def calc_global_Lq_norm(self, u):
        """function_ops.f90:1675-1692."""
        aux = self.calc_uk(u, self.p.q / 2.0)
        ip = self.global_inner_product(aux, aux, INNER_L2)
        return ip ** (1.0 / self.p.q)

    def rescaleLqNorm(self, u, B):
        """function_ops.f90:1699-1709 (returns the rescaled field)."""
        L = self.calc_global_Lq_norm(u)
        return B / L * u

    def calcConstraintDerivativeL2(self, u):
        """function_ops.f90:984-1006: ||u||_q^(1-q) dealias(|u|^(q-2) u)."""
        q = self.p.q
        nrm = self.calc_global_Lq_norm(u)
        a = self.calc_uk(u, q - 2.0)
        res = []
        for i in range(3):
            r = a * u[i]
            if self.p.toDealias and (q - 2.0 > MACH_EPSILON):
                r = self.dealias_scalar(r, 2.0)
            res.append(r)
        return (nrm ** (1.0 - q)) * np.stack(res)

    def vectorTransport(self, u, q, B, eta, xi):
        """function_ops.f90:1716-1768."""
        upe = u + eta
        a = self.calc_uk(upe, q - 2.0)
        aux = []
        for i in range(3):
            r = a * upe[i]
            if self.p.toDealias and (q - 2.0 > MACH_EPSILON):
                r = self.dealias_scalar(r, 2.0)
            aux.append(r)
        aux = np.stack(aux)
        intResult = self.global_summed_field_inner_product(aux, xi, INNER_L2)
        L = self.calc_global_Lq_norm(upe)
        res = xi - L ** (-q) * intResult * upe                     # :1746
        return B / L * res                                         # :1747

    # -- objective ----------------------------------------------------------------
    def calc_local_dLqdt_inclParts(self, u, q=None):
        """function_ops.f90:1304-1358: returns (R, R1+R2, R3)."""
        p = self.p
        q = p.q if q is None else q
        nu = p.visc
        A = self.calc_uk(u, q - 2.0)                               # :1314
        gradU = self._gradU(u)
        aux = self.calc_nablaModSquared(u, gradU)                  # :1316
        R1 = -p.viscCoefficient * nu * self.inner_product(A, aux)  # :1318
        e_u = self.calc_unitVectorInUdirection(u)                  # :1323
        if p.use_e_u:
            aux = self.calc_vgraduvgradu(u, e_u, gradU)
            R2 = -p.viscCoefficient * (q - 2.0) * nu * self.inner_product(A, aux)
        else:
            Bq = self.calc_uk(u, q - 4.0)                          # :1331
            aux = self.calc_vgraduvgradu(u, u, gradU)
            R2 = -p.viscCoefficient * (q - 2.0) * nu * self.inner_product(Bq, aux)
        aux = self.calc_nablaUnablaUt(u, gradU)                    # :1339
        pr = self.solve_poisson(-aux, 1.0)                         # :1340
        if p.use_e_u:
            aux = self._D(pr * A, 2.0)
            v = self.calc_vvgradu(u, e_u, gradU)
        else:
            aux = self._D(pr * Bq, 2.0)                            # :1349-1350
            v = self.calc_vvgradu(u, u, gradU)                     # :1351
        R3 = p.pressureCoefficient * (q - 2.0) * self.inner_product(aux, v)
        return np.array([R1 + R2 + R3, R1 + R2, R3])

    def calc_local_dLqdt(self, u, q=None):
        """function_ops.f90:1367-1378."""
        return float(self.calc_local_dLqdt_inclParts(u, q)[0])

    def calc_global_dLqdt_inclParts(self, u, q=None):
        """function_ops.f90:1392-1405."""
        return self.calc_local_dLqdt_inclParts(u, q)

    def GradL2ForLq(self, u):
        """function_ops.f90:1066-1211: L2 gradient of R."""
        p = self.p
        q = p.q
        nu = p.visc
        vc, pc = p.viscCoefficient, p.pressureCoefficient
        A = self.calc_uk(u, q - 2.0)                               # :1102
        gradU = self._gradU(u)
        aux = self.calc_nablaUnablaUt(u, gradU)                    # :1105
        pr = self.solve_poisson(-aux, 1.0)                         # :1107
        gradP = self.gradient(pr)                                  # :1109
        DeltaU = self.laplacian(u)                                 # :1111-1112
        if p.use_e_u:
            e_u = self.calc_unitVectorInUdirection(u)              # :1117
            aux2 = np.zeros_like(u[0])
            for i in range(3):
                aux2 = aux2 + e_u[i] * (vc * nu * DeltaU[i] - pc * gradP[i])
            if p.toDealias and p.dealiase_if_mult_by_e_u:
                aux2 = self.dealias_scalar(aux2, 2.0)
            aux2 = self._D(A * aux2, 2.0)
            aux3v = []
            for i in range(3):
                r = aux2 * e_u[i]
                if p.toDealias and p.dealiase_if_mult_by_e_u:
                    r = self.dealias_scalar(r, 2.0)
                aux3v.append(r)
        else:
            aux2 = np.zeros_like(u[0])
            for i in range(3):                                     # :1131-1133
                aux2 = aux2 + u[i] * (vc * nu * DeltaU[i] - pc * gradP[i])
            aux2 = self._D(aux2, 2.0)                              # :1134
            Bq = self.calc_uk(u, q - 4.0)                          # :1136
            aux2 = self._D(Bq * aux2, 2.0)                         # :1137-1138
            aux3v = [self._D(aux2 * u[i], 2.0) for i in range(3)]  # :1140-1143
        uq2u = np.stack([self._D(A * u[i], 2.0) for i in range(3)])  # :1150-1153
        aux3 = self.divergence(uq2u)                               # :1155
        aux4 = self.solve_poisson(aux3, 1.0)                       # :1156
        aux4v = self.gradient(aux4)                                # :1158
        aux5v = []
        for i in range(3):                                         # :1165-1169
            r = (gradU[0, i] * aux4v[0] + gradU[1, i] * aux4v[1]
                 + gradU[2, i] * aux4v[2])
            aux5v.append(self._D(r, 2.0))
        aux7v = self.laplacian(uq2u)                               # :1172-1173
        V = []
        for i in range(3):                                         # :1178-1204
            a4 = self._D(A * gradP[i], 2.0)
            a6 = self._D(A * DeltaU[i], 2.0)
            V.append((q - 2.0) * aux3v[i] - pc * a4 + pc * 2.0 * aux5v[i]
                     + vc * nu * a6 + vc * nu * aux7v[i])
        return np.stack(V)

    # -- optimization.f90 -----------------------------------------------------------
    def eval_J(self, field, mysystem):
        """optimization.f90:520-547."""
        aux1 = np.array(field, copy=True)
        if mysystem == "maxdLqdt":
            return self.calc_local_dLqdt(aux1, self.p.q)
        if mysystem == "LineMin":
            aux1 = self.rescaleLqNorm(aux1, self.p.B)
            return -1.0 * self.eval_J(aux1, "maxdLqdt")
        if mysystem == "||u||_q":
            return self.calc_global_Lq_norm(aux1)
        return 0.0


# ---------------------------------------------------------------------------
# Line search: optimization.f90:613-951, as scalar routines over a callable J
# ---------------------------------------------------------------------------

def _sign(a, b):
    """Fortran SIGN(a, b)."""
    return math.copysign(abs(a), b)


def mnbrak(J, tA0, tB0, log=None):
    """optimization.f90:613-778 (mnbra_calcSaveAllJvalues = .false.).

    ``J(tau)`` returns eval_J(phi + tau*grad, "LineMin").  Returns
    ``(tau_brack, flag, nevals)``; with flag 1 the bracket is undefined in the
    reference (it returns early) and ``None`` is returned for it.
    ``log(kind, iter, tA, tB, tC, FA, FB, FC)`` receives the save_linemin_data
    records (mode "replace" for the first, "append" afterwards).
    """
    GOLD, CGOLD, GLIMIT, ITMAX = MNBRAK_GOLD, MNBRAK_CGOLD, MNBRAK_GLIMIT, MNBRAK_ITMAX
    ev = [0]

    def f(t):
        ev[0] += 1
        return J(t)

    brakIter = 0
    tA = tA0
    tB = max(tB0, MACH_EPSILON)
    tC = 0.0
    FC = 0.0
    FA = f(tA)
    FB = f(tB)
    if log:
        log("replace", brakIter, tA, tB, tC, FA, FB, FC)
    while FB > FA and tB > MACH_EPSILON:                     # :677-683
        tB = CGOLD * tB
        FB = f(tB)
        if log:
            log("append", brakIter, tA, tB, tC, FA, FB, FC)
    if tB <= MACH_EPSILON:                                   # :685-688
        return None, 1, ev[0]
    tC = GOLD * tB                                           # :690
    FC = f(tC)
    if log:
        log("append", brakIter, tA, tB, tC, FA, FB, FC)
    while FB >= FC and brakIter < ITMAX:                     # :697
        brakIter += 1
        tC = GOLD * tC
        FC = f(tC)
        R = (tB - tA) * (FB - FC)
        Q = (tB - tC) * (FB - FA)
        tP = tB - 0.5 * ((tB - tC) * Q - (tB - tA) * R) / _sign(
            max(abs(Q - R), MACH_EPSILON), Q - R)
        Pmax = tB + GLIMIT * (tC - tB)
        if (tB - tP) * (tP - tC) > 0:                        # :710
            FP = f(tP)
            if FP < FC:
                tA, FA, tB, FB = tB, FB, tP, FP
                break
            elif FP > FB:
                tC, FC = tP, FP
                break
            tP = tC + GOLD * (tC - tB)
            FP = f(tP)
        elif (tC - tP) * (tP - Pmax) > 0:                    # :730
            FP = f(tP)
            if FP < FC:
                tB, tC, FB, FC = tC, tP, FC, FP
                tP = tC + GOLD * (tC - tB)
                FP = f(tP)
        elif (tP - Pmax) * (Pmax - tC) >= 0:                 # :744
            tP = Pmax
            FP = f(tP)
        else:
            tP = tC + GOLD * (tC - tB)
            FP = f(tP)
        tA, tB, tC = tB, tC, tP
        FA, FB, FC = FB, FC, FP
        if log:
            log("append", brakIter, tA, tB, tC, FA, FB, FC)
    flag = 2 if brakIter >= ITMAX else 0
    return (tA, tC), flag, ev[0]


def brent(J, tau_brack, log=None):
    """optimization.f90:784-951.  Returns ``(X, FX, nevals)``.

    ``log(j, X, FX)`` receives the brent-info records (j = 0 for the tau = 0
    probe, 999 for the final line).
    """
    TOL, ZEPS, CGOLD, ITMAX = BRENT_TOL, BRENT_ZEPS, BRENT_CGOLD, BRENT_ITMAX
    ev = [0]

    def f(t):
        ev[0] += 1
        return J(t)

    D = 0.0
    A = min(tau_brack[0], tau_brack[1])
    Bb = max(tau_brack[0], tau_brack[1])
    V = tau_brack[1] * CGOLD
    W = V
    X = V
    E = 0.0
    FX = f(D)                                          # :833-834
    if log:
        log(0, D, FX)
    FX = f(X)                                          # :844-845
    FV = FX
    FW = FX
    P = Q = R = 0.0
    for j in range(1, ITMAX + 1):
        if log:
            log(j, X, FX)
        XM = 0.5 * (A + Bb)
        TOL1 = TOL * abs(X) + ZEPS
        TOL2 = 2.0 * TOL1
        if abs(X - XM) <= (TOL2 - 0.5 * (Bb - A)):     # :865
            break
        FLAG = 1
        if abs(E) > TOL1:
            R = (X - W) * (FX - FV)
            Q = (X - V) * (FX - FW)
            P = (X - V) * Q - (X - W) * R
            Q = 2.0 * (Q - R)
            if Q > 0.0:
                P = -P
            Q = abs(Q)
            ETEMP = E
            E = D
            if (abs(P) >= abs(0.5 * Q * ETEMP)) or (P <= Q * (A - X)) or (P >= Q * (Bb - X)):
                FLAG = 1
            else:
                FLAG = 2
        if FLAG == 1:
            E = (A - X) if X >= XM else (Bb - X)
            D = CGOLD * E
        else:
            D = P / Q
            U = X + D
            if (U - A < TOL2) or (Bb - U < TOL2):
                D = _sign(TOL1, XM - X)
        if abs(D) >= TOL1:
            U = X + D
        else:
            U = X + _sign(TOL1, D)
        FU = f(U)
        if FU <= FX:
            if U >= X:
                A = X
            else:
                Bb = X
            V, FV, W, FW, X, FX = W, FW, X, FX, U, FU
        else:
            if U < X:
                A = U
            else:
                Bb = U
            if (FU <= FW) or (W == X):
                V, FV, W, FW = W, FW, U, FU
            elif (FU <= FV) or (V == X) or (V == W):
                V, FV = U, FU
    if log:
        log(999, X, FX)
    return X, FX, ev[0]


# ---------------------------------------------------------------------------
# maxdLqdt: optimization.f90:7-514 (algorithmic body; diagnostics-only prints,
# spectra, field saves and kappa tests are not part of it)
# ---------------------------------------------------------------------------

@dataclass
class IterRecord:
    iter: int
    tau_brack: tuple
    tau: float
    beta: float
    J0: float
    J1: float
    deltaJ: float
    nJ: int
    msgs: list = field(default_factory=list)


@dataclass
class SolveResult:
    u: np.ndarray
    J1: float
    iters: int
    successful: bool
    records: list
    while_flag: int
    stop_reason: str


def maxdLqdt(orc: Oracle, u0, startIter=0, on_iter=None, on_start=None):
    """One RCG solve for the current B (optimization.f90:7-514).  on_start()
    is called once the loop's initial values are set (timing hook)."""
    P = orc.p
    q, B = P.q, P.B
    HL = INNER_HL
    ip = orc.global_summed_field_inner_product
    u = orc.rescaleLqNorm(np.array(u0, dtype=np.float64, copy=True), B)   # :61
    u = orc.divAvg_free(u)                                                # :62
    J0 = orc.eval_J(u, "maxdLqdt")                                        # :78
    J1 = J0
    it = startIter
    zeros = np.zeros_like(u)
    gradJused0 = zeros.copy()
    vT_g0 = zeros.copy()
    vT_d0 = zeros.copy()
    tau0 = 100.0                                                          # :92
    tau1 = tau0
    tol_n = max(1, P.toleranceRollingAverageSize)
    rel = np.zeros(tol_n)
    avg = 1.0
    while_flag = 0
    records = []
    stop_reason = ""
    s = 1.5 - 1.0 / q
    if on_start:
        on_start()
    while abs(avg) > P.OPTIM_TOL and it < P.MAX_ITER and while_flag < 1:
        it += 1
        msgs = []
        g1 = orc.GradL2ForLq(u)                                           # :146
        g1 = orc.HilbertGradient(g1, s)                                   # :151
        g1 = orc.divAvg_free(g1)                                          # :153
        nL2 = orc.calcConstraintDerivativeL2(u)                           # :162
        nH = orc.HilbertGradient(nL2, s)                                  # :164
        nH = orc.divAvg_free(nH)                                          # :169
        alpha = ip(nH, g1, HL) / ip(nH, nH, HL)                           # :170-171
        gproj = g1 - alpha * nH                                           # :177
        t1 = abs(ip(gproj, nH, HL) / ip(gproj + nH, gproj + nH, HL))      # :180
        if t1 > P.checkNormalTolerance:
            msgs.append(42)
        fs = orc.divergence(gproj)                                        # :188
        t1 = orc.global_inner_product(fs, fs, INNER_L2) / ip(gproj, gproj, HL)
        if t1 > P.checkDivergenceTolerance:
            msgs.append(43)
        gused1 = gproj if P.useTangentialGradient else g1
        if P.normalizeGradient:                                           # :211-214
            gused1 = math.sqrt(ip(u, u, HL) / ip(gused1, gused1, HL)) * gused1
        if P.useConjugateGradient:
            beta = ip(gused1, gused1 - vT_g0, HL)                         # :226
            g0n = ip(gradJused0, gradJused0, HL)                          # :227
            if g0n < MACH_EPSILON or beta < MACH_EPSILON:
                beta = 0.0
            else:
                beta = beta / g0n
            if P.resetMomentumTermEveryXiterations > 0 and \
                    it % P.resetMomentumTermEveryXiterations == 0:        # :235
                beta = 0.0
            d1 = gused1 + beta * vT_d0                                    # :240
        else:
            beta = 0.0
            d1 = gused1
        tv = orc.calc0thFourierModes(d1)                                  # :251
        if np.sum(np.abs(tv)) > P.checkAverageTolerance * B:
            msgs.append(44)
        tv = orc.calc0thFourierModes(u)                                   # :257
        if np.sum(np.abs(tv)) > P.checkAverageTolerance * B:
            msgs.append(44)
        nJ = [0]

        def Jline(t, u=u, d1=d1):
            nJ[0] += 1
            return orc.eval_J(u + t * d1, "LineMin")

        msgs.append(20)
        tau_brack, flag, _ = mnbrak(Jline, 0.0, tau0)                     # :269
        if flag != 0:
            if flag == 1:
                msgs.append(22)
            records.append(IterRecord(it, tau_brack, tau1, beta, J0, J0, 0.0, nJ[0], msgs))
            while_flag = 1
            stop_reason = f"mnbrak flag {flag}"
            break
        msgs.append(21)
        msgs.append(30)
        tau1, FX, _ = brent(Jline, tau_brack)                             # :297
        if tau1 > tau_brack[1]:
            if tau1 > 2.0 * tau_brack[1]:
                msgs.append(32)
        else:
            msgs.append(31)
        if tau1 > TAU_MAX:                                                # :310-314
            msgs.append(32)
            while_flag = 1
            tau1 = 0.0
            stop_reason = "tau > TAU_MAX"
        elif tau1 > math.sqrt(TAU_MAX):
            msgs.append(41)
        if P.useRiemannianGeometry:                                       # :324-327
            vT_g0 = orc.vectorTransport(u, q, B, tau1 * d1, gused1)
            vT_d0 = orc.vectorTransport(u, q, B, tau1 * d1, d1)
        else:
            vT_g0 = gused1
            vT_d0 = d1
        u = orc.rescaleLqNorm(u + tau1 * d1, B)                           # :345-346
        J1 = orc.eval_J(u, "maxdLqdt")                                    # :349
        deltaJ = (J1 - J0) / abs(J0)                                      # :350
        rec = IterRecord(it, tau_brack, tau1, beta, J0, J1, deltaJ, nJ[0], msgs)
        records.append(rec)
        if on_iter:
            on_iter(rec, u)
        if deltaJ < -MACH_EPSILON:                                        # :351-365
            msgs.append(0)
            while_flag = 1
            stop_reason = "cost functional not increasing"
            break
        J0 = J1                                                           # :432-436
        gradJused0 = gused1
        tau0 = tau1
        rel[1:] = rel[:-1].copy()                                         # :439-445
        rel[0] = deltaJ
        avg = float(np.sum(rel)) / float(it - startIter)
    successful = (it < P.MAX_ITER) and while_flag < 1                    # :451
    if not stop_reason:
        stop_reason = "converged" if successful else "max iter"
    return SolveResult(u, J1, it, successful, records, while_flag, stop_reason)


def kappa_fd_check(orc: Oracle, u, grad, directions, h=1e-6):
    """Central finite differences of R along ``directions`` against
    <grad, phi>_L2 (the config-2 variant of optimization.f90:1044-1069)."""
    out = []
    for phi in directions:
        Rp = orc.eval_J(u + h * phi, "maxdLqdt")
        Rm = orc.eval_J(u - h * phi, "maxdLqdt")
        fd = (Rp - Rm) / (2.0 * h)
        an = orc.global_summed_field_inner_product(grad, phi, INNER_L2)
        out.append((fd, an))
    return out


def kappa_best(kappas):
    """Best row by the reference's criterion (optimization.f90:1035): row ii
    replaces the best when max(|k-1|, |1/k-1|) < max(|best-1|, |1/k-1|),
    both sides using the current row's 1/k; best starts at huge(pr).
    Returns (best kappa, 1-based row)."""
    best, bestii = np.finfo(np.float64).max, 0
    for ii, k in enumerate(kappas, start=1):
        if max(abs(k - 1.0), abs(1.0 / k - 1.0)) < max(abs(best - 1.0), abs(1.0 / k - 1.0)):
            best, bestii = k, ii
    return best, bestii


def kappa_adjust(kappas, adj=4):
    """Adjusted kappa (optimization.f90:1075-1097): the window of `adj`
    consecutive rows (start ii = 1..n-adj) with the smallest sum of
    |1 - k[i]/k[i+1]| gives frac = sign(k[start]) * (prod |k|)^(1/adj);
    returns (kappas / frac, frac)."""
    k = np.asarray(kappas, dtype=np.float64)
    n = len(k)
    imin, fmin = 0, 9999.0
    for i in range(n - (adj - 1) - 1):
        fr = sum(abs(1.0 - k[i + j] / k[i + j + 1]) for j in range(adj))
        if fr < fmin:
            fmin, imin = fr, i
    frac = 1.0
    for j in range(adj):
        frac *= abs(k[imin + j])
    frac = math.copysign(frac ** (1.0 / adj), k[imin])
    return k / frac, frac


def kappa_test(orc: Oracle, phi, gradJ, pert, mysystem="maxdLqdt", space=INNER_L2):
    """kappa_test (optimization.f90:957-1129) with a given perturbation.
    Returns (table[16][5] = {eps, <grad, pert>, J1 - J0, kappa, kappa adj},
    frac, best kappa, best row)."""
    nk = 16
    ip = orc.global_summed_field_inner_product(gradJ, pert, space)
    J0 = orc.eval_J(phi, mysystem)
    tab = np.zeros((nk, 5))
    for ii in range(1, nk + 1):
        eps = 10.0 ** (-7.0 - nk / 2.0 + ii)
        J1 = orc.eval_J(phi + eps * pert, mysystem)
        kap = (J1 - J0) / (eps * ip)
        tab[ii - 1] = (eps, ip, J1 - J0, kap, kap)
    best, bestii = kappa_best(tab[:, 3])
    tab[:, 4], frac = kappa_adjust(tab[:, 4])
    return tab, frac, best, bestii


# ---------------------------------------------------------------------------
# Parity pins: oracle function -> the tests that check it against something
# other than this file (golden vectors, closed forms, brute force, identities).
# ---------------------------------------------------------------------------
PINS = {
    "fftfwd/fftbwd": "tests/test_oracle_pins.py::test_fft_bruteforce_dft",
    "Grid (K, Nyquist +pi N, dV)": "tests/test_oracle_pins.py::test_wavenumbers",
    "dealias_scalar": "tests/test_oracle_pins.py::test_dealias_bruteforce",
    "gradient": "tests/test_oracle_pins.py::test_gradient_closed_form_and_nyquist",
    "divergence": "tests/test_oracle_pins.py::test_divergence_closed_form",
    "laplacian": "tests/test_oracle_pins.py::test_laplacian_abc",
    "solve_poisson": "tests/test_oracle_pins.py::test_poisson_closed_form",
    "divAvg_free": "tests/test_oracle_pins.py::test_leray_invariants",
    "HilbertGradient": "tests/test_oracle_pins.py::test_hilbert_single_mode",
    "inner_product": "tests/test_oracle_pins.py::test_inner_products_closed_form",
    "calc_gk_order2/calc_uk": "tests/test_oracle_pins.py::test_gk_order2_polynomial",
    "calc_global_Lq_norm": "tests/test_oracle_pins.py::test_lq_norm_closed_form",
    "calc_local_dLqdt_inclParts": "tests/test_oracle_pins.py::test_R_*",
    "GradL2ForLq": "tests/test_oracle_pins.py::test_gradient_fd_*",
    "calcConstraintDerivativeL2": "tests/test_oracle_pins.py::test_constraint_derivative_fd",
    "vectorTransport": "tests/test_oracle_pins.py::test_vector_transport_tangency",
    "mnbrak/brent": "tests/test_oracle_linesearch.py",
    "kappa_test/kappa_best/kappa_adjust": "tests/test_oracle_pins.py::test_kappa_*",
    "maxdLqdt": "tests/test_oracle_linesearch.py::test_rcg_monotone_and_sphere",
    "calculate_spectral_data": "tests/test_oracle_pins.py::test_spectrum_*",
    "vel2vort/calc_global_summed_enstrophy": "tests/test_oracle_pins.py::test_enstrophy_*",
    "calc_global_dEnstrophydt": "tests/test_oracle_pins.py::test_denstrophy_*",
}
