"""Host-side mirror of the reference's Fortran interface for the RCG hot path.

``MaxLq`` owns one device context (one GPU, one HIP stream).  Its methods
carry the reference routine names and argument meaning (fftwfunction.f90,
function_ops.f90, optimization.f90); arrays use the reference layout: a
scalar field is (n3, n2, n1) C-ordered float64 (= Fortran (n1,n2,n3), x
fastest), a vector field (3, n3, n2, n1).  Routines that are ``intent(inout)``
in the reference update the numpy array in place.  Every call runs on the GPU
through libmaxlq.so; errors raise ``MlqError`` (NaN -> MLQ_ENAN where the
reference stops with ``stop 1``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .capi import MlqError, _p

_IP = {"L2": capi.MLQ_IP_L2, "H_l^(3/2-1/q)": capi.MLQ_IP_HL, "H^(3/2-1/q)": capi.MLQ_IP_H,
       "H^((3q-1)/(2q))": capi.MLQ_IP_OLD_H, "H_l^((3q-1)/(2q))": capi.MLQ_IP_OLD_HL}
_SYS = {"maxdLqdt": capi.MLQ_J_MAXDLQDT, "LineMin": capi.MLQ_J_LINEMIN,
        "||u||_q": capi.MLQ_J_LQNORM}


def _f64(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _c128(a, shape):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


@dataclass
class RcgTrace:
    """Records the reference writes from inside maxdLqdt (data_ops.f90)."""
    iteration_info: list = field(default_factory=list)   # (iter, tau, beta, J)
    brackets: list = field(default_factory=list)         # (iter, tau_brack a, b, J requests)
    linemin: list = field(default_factory=list)          # (iter, brakIter, mode, tA..FC)
    brent: list = field(default_factory=list)            # (iter, j, X, FX)
    msgs: list = field(default_factory=list)             # (iter, code)
    errors: list = field(default_factory=list)           # (iter, code)
    spectra: list = field(default_factory=list)          # (iter, "uvec" | "d", (n, 2) array)
    fields: list = field(default_factory=list)           # (iter, (3, n, n, n) array): save_every
    kappa: list = field(default_factory=list)            # (iter, name, adjusted, table, frac,
                                                         #  best kappa, best row)


KAPPA_NAMES = ("initial_grad_L2", "gradL2_iter", "end_gradJ", "end_d1")


class MaxLq:
    """One device context for an n^3 grid (replaces init_fft / initialize)."""

    def __init__(self, n: int, q: float = 4.0, B: float = 1.0, precision: str = "fp64", **kw):
        self.lib = capi.load()
        self.n = int(n)
        prec = capi.MLQ_MIXED if precision in ("mixed", "fp32") else capi.MLQ_FP64
        self.params = capi.default_params(q, B=float(B), precision=prec, **kw)
        self.ctx = C.c_void_p()
        rc = self.lib.mlq_init(C.byref(self.ctx), self.n, self.n, self.n, C.byref(self.params))
        if rc != capi.MLQ_OK:
            raise MlqError(rc, "mlq_init failed")
        self.shape = (self.n, self.n, self.n)
        self.vshape = (3,) + self.shape
        self.hshape = (self.n, self.n, self.n // 2 + 1)

    # -- plumbing ---------------------------------------------------------------
    def _chk(self, rc):
        if rc != capi.MLQ_OK:
            raise MlqError(rc, self.lib.mlq_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.lib.mlq_finalize(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        self._chk(self.lib.mlq_set_params(self.ctx, C.byref(self.params)))

    @property
    def q(self):
        return self.params.q

    # -- fftwfunction ---------------------------------------------------------------
    def fftfwd(self, u):
        u = _c128(u, self.shape)
        out = np.empty(self.shape, np.complex128)
        self._chk(self.lib.mlq_fftfwd(self.ctx, _p(u.view(np.float64)), _p(out.view(np.float64))))
        return out

    def fftbwd(self, fu):
        fu = _c128(fu, self.shape)
        out = np.empty(self.shape, np.complex128)
        self._chk(self.lib.mlq_fftbwd(self.ctx, _p(fu.view(np.float64)), _p(out.view(np.float64))))
        return out

    def fftfwdr(self, u):
        u = _f64(u, self.shape)
        out = np.empty(self.hshape, np.complex128)
        self._chk(self.lib.mlq_fftfwdr(self.ctx, _p(u), _p(out.view(np.float64))))
        return out

    def fftbwdr(self, fu):
        fu = _c128(fu, self.hshape)
        out = np.empty(self.shape, np.float64)
        self._chk(self.lib.mlq_fftbwdr(self.ctx, _p(fu.view(np.float64)), _p(out)))
        return out

    # -- function_ops -----------------------------------------------------------------
    def gradient(self, u):
        u = _f64(u, self.shape)
        out = np.empty(self.vshape)
        self._chk(self.lib.mlq_gradient(self.ctx, _p(u), _p(out)))
        return out

    def divergence(self, u):
        u = _f64(u, self.vshape)
        out = np.empty(self.shape)
        self._chk(self.lib.mlq_divergence(self.ctx, _p(u), _p(out)))
        return out

    def laplacian(self, f):
        a = _f64(f, self.vshape).copy()
        self._chk(self.lib.mlq_laplacian(self.ctx, _p(a)))
        return a

    def solve_poisson(self, f, l=1.0):
        f = _f64(f, self.shape)
        out = np.empty(self.shape)
        self._chk(self.lib.mlq_solve_poisson(self.ctx, _p(f), float(l), _p(out)))
        return out

    def dealias_scalar(self, g, nonlinOrder=2.0):
        a = _f64(g, self.shape).copy()
        self._chk(self.lib.mlq_dealias_scalar(self.ctx, _p(a), float(nonlinOrder)))
        return a

    def divAvg_free(self, f):
        a = _f64(f, self.vshape).copy()
        self._chk(self.lib.mlq_divAvg_free(self.ctx, _p(a)))
        return a

    def HilbertGradient(self, grad, order):
        a = _f64(grad, self.vshape).copy()
        self._chk(self.lib.mlq_HilbertGradient(self.ctx, _p(a), float(order)))
        return a

    def calc0thFourierModes(self, v):
        v = _f64(v, self.vshape)
        out = np.empty(3)
        self._chk(self.lib.mlq_calc0thFourierModes(self.ctx, _p(v), _p(out)))
        return out

    def calc_uk(self, u, k):
        u = _f64(u, self.vshape)
        out = np.empty(self.shape)
        self._chk(self.lib.mlq_calc_uk(self.ctx, _p(u), float(k), _p(out)))
        return out

    def calc_gk_order2(self, g, k):
        g = _f64(g, self.shape)
        out = np.empty(self.shape)
        self._chk(self.lib.mlq_calc_gk_order2(self.ctx, _p(g), float(k), _p(out)))
        return out

    def calc_global_Lq_norm(self, u):
        u = _f64(u, self.vshape)
        r = C.c_double()
        self._chk(self.lib.mlq_calc_global_Lq_norm(self.ctx, _p(u), C.byref(r)))
        return r.value

    def calculate_spectral_data(self, u, normalize=True):
        """calculate_spectral_data (function_ops.f90:2614-2677): (n, 2) array,
        column 0 the shell bounds 2 pi (i+1), column 1 the shell energies."""
        u = _f64(u, self.vshape)
        out = np.zeros(2 * self.n)
        self._chk(self.lib.mlq_calculate_spectral_data(self.ctx, _p(u), int(bool(normalize)),
                                                       _p(out)))
        return out.reshape(2, self.n).T.copy()

    def calc_global_summed_enstrophy(self, u):
        """function_ops.f90:2455-2467: 1/2 ||curl u||_2^2."""
        u = _f64(u, self.vshape)
        r = C.c_double()
        self._chk(self.lib.mlq_calc_global_summed_enstrophy(self.ctx, _p(u), C.byref(r)))
        return r.value

    def calc_global_dEnstrophydt(self, u):
        """function_ops.f90:1417-1481: -nu ||Lap u||^2 + <D((u.grad)u), Lap u>."""
        u = _f64(u, self.vshape)
        r = C.c_double()
        self._chk(self.lib.mlq_calc_global_dEnstrophydt(self.ctx, _p(u), C.byref(r)))
        return r.value

    def spectral_data_field(self, h, normalize=True):
        """The same on a device-resident vector field handle."""
        out = np.zeros(2 * self.n)
        self._chk(self.lib.mlq_spectral_data_field(self.ctx, h, int(bool(normalize)), _p(out)))
        return out.reshape(2, self.n).T.copy()

    def rescaleLqNorm(self, u, B):
        a = _f64(u, self.vshape).copy()
        self._chk(self.lib.mlq_rescaleLqNorm(self.ctx, _p(a), float(B)))
        return a

    def calcConstraintDerivativeL2(self, u):
        u = _f64(u, self.vshape)
        out = np.empty(self.vshape)
        self._chk(self.lib.mlq_calcConstraintDerivativeL2(self.ctx, _p(u), _p(out)))
        return out

    def vectorTransport(self, u, q, B, eta, xi):
        u, eta, xi = (_f64(a, self.vshape) for a in (u, eta, xi))
        out = np.empty(self.vshape)
        self._chk(self.lib.mlq_vectorTransport(self.ctx, _p(u), float(q), float(B), _p(eta),
                                               _p(xi), _p(out)))
        return out

    def inner_product(self, f, g, mytype="L2"):
        f, g = _f64(f, self.shape), _f64(g, self.shape)
        r = C.c_double()
        self._chk(self.lib.mlq_inner_product(self.ctx, _p(f), _p(g), _IP[mytype], C.byref(r)))
        return r.value

    global_inner_product = inner_product

    def global_summed_field_inner_product(self, f, g, mytype="L2"):
        f, g = _f64(f, self.vshape), _f64(g, self.vshape)
        r = C.c_double()
        self._chk(self.lib.mlq_global_summed_field_inner_product(self.ctx, _p(f), _p(g),
                                                                 _IP[mytype], C.byref(r)))
        return r.value

    def calc_local_dLqdt_inclParts(self, u, q=None):
        u = _f64(u, self.vshape)
        out = np.empty(3)
        qq = self.q if q is None else float(q)
        self._chk(self.lib.mlq_calc_local_dLqdt_inclParts(self.ctx, _p(u), qq, _p(out)))
        return out

    calc_global_dLqdt_inclParts = calc_local_dLqdt_inclParts

    def calc_local_dLqdt(self, u, q=None):
        return float(self.calc_local_dLqdt_inclParts(u, q)[0])

    def GradL2ForLq(self, U):
        U = _f64(U, self.vshape)
        out = np.empty(self.vshape)
        self._chk(self.lib.mlq_GradL2ForLq(self.ctx, _p(U), _p(out)))
        return out

    def eval_J(self, myfield, mysystem):
        a = _f64(myfield, self.vshape)
        r = C.c_double()
        self._chk(self.lib.mlq_eval_J(self.ctx, _p(a), _SYS[mysystem], C.byref(r)))
        return r.value

    # -- device-resident fields ----------------------------------------------------------
    def field(self, ncomp=3):
        h = C.c_int()
        self._chk(self.lib.mlq_field_alloc(self.ctx, ncomp, C.byref(h)))
        return h.value

    def free(self, h):
        self._chk(self.lib.mlq_field_free(self.ctx, h))

    def upload(self, h, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(self.lib.mlq_upload(self.ctx, h, _p(a)))

    def download(self, h, ncomp=3):
        out = np.empty(self.vshape if ncomp == 3 else self.shape)
        self._chk(self.lib.mlq_download(self.ctx, h, _p(out)))
        return out

    def seeded_field(self, h, seed, kind="gauss", a=16.0):
        k = {"gauss": 0, "k2gauss": 1}[kind]
        self._chk(self.lib.mlq_seeded_field(self.ctx, h, C.c_ulonglong(seed), k, float(a)))

    def eval_J_batch(self, hu, hd, taus):
        taus = np.ascontiguousarray(taus, dtype=np.float64)
        out = np.empty(len(taus))
        self._chk(self.lib.mlq_eval_J_batch(self.ctx, hu, hd, _p(taus), len(taus), _p(out)))
        return out

    # -- optimization: maxdLqdt ------------------------------------------------------------
    def rcg_path(self):
        """The iteration mlq_rcg_solve runs: 1 resident state, 2 reference-structured
        with the line search from coefficient fields, 0 reference-structured."""
        r = C.c_int()
        self._chk(self.lib.mlq_rcg_path(self.ctx, C.byref(r)))
        return r.value

    def uses_resident_state(self):
        """True when mlq_rcg_solve runs the resident-state iteration."""
        return self.rcg_path() == 1

    def time_linesearch(self, hu, hd, taus, batch):
        """J_LineMin at `taus` after one coefficient precompute, in launches of
        `batch` step sizes (0: reference-structured); returns (J, t_precompute_ms,
        t_eval_ms)."""
        taus = np.ascontiguousarray(taus, dtype=np.float64)
        J = np.empty(len(taus))
        t = np.zeros(2)
        self._chk(self.lib.mlq_time_linesearch(self.ctx, hu, hd, _p(taus), len(taus), int(batch),
                                               _p(J), _p(t)))
        return J, t[0], t[1]

    def diagnosticScalars(self, u, fields=False, ring_cap=0):
        """diagnosticScalars (function_ops.f90:2503-2602): dict of the values
        save_diagnosticFields_global writes; with fields, also the (4, n, n, n)
        fields magU, magW, Helicity, VortexCore; with ring_cap, the ring points."""
        u = _f64(u, self.vshape)
        out = np.zeros(21)
        fl = np.zeros((4,) + self.shape) if fields else None
        ring = np.zeros((max(ring_cap, 1), 3))
        cnt = C.c_longlong()
        self._chk(self.lib.mlq_diagnosticScalars(self.ctx, _p(u), _p(out),
                                                 _p(fl) if fields else None,
                                                 _p(ring) if ring_cap else None, int(ring_cap),
                                                 C.byref(cnt)))
        d = {"K": out[0:3], "E": out[3:6], "Umax": out[6:9], "Wmax": out[9:12],
             "vorCore": out[12:15], "magUmax": out[15], "magWmax": out[16], "H": out[17],
             "maxHel": out[18], "minHel": out[19], "divU_L2": out[20], "ring_points": cnt.value}
        if fields:
            d["fields"] = fl
        if ring_cap:
            d["ring"] = ring[:min(ring_cap, cnt.value)]
        return d

    def kappa_test_pert(self, h, kind, m1=0.0, m2=0.0, m3=0.0, seed=0, loaded=None):
        """kappa_test_pert (function_ops.f90:613-977) into device field h."""
        ld = None if loaded is None else _f64(loaded, self.vshape)
        self._chk(self.lib.mlq_kappa_test_pert(self.ctx, h, int(kind), float(m1), float(m2),
                                               float(m3), C.c_ulonglong(seed),
                                               None if ld is None else _p(ld)))

    def kappa_test(self, hphi, hgrad, hpert, mysystem="maxdLqdt", inner_product_space="L2"):
        """kappa_test (optimization.f90:957-1129) on device fields; returns
        (table[16][5], frac, best_kappa, best_row)."""
        tab = np.zeros(80)
        frac, best = C.c_double(), C.c_double()
        row = C.c_int()
        self._chk(self.lib.mlq_kappa_test(self.ctx, hphi, hgrad, hpert, _SYS[mysystem],
                                          _IP[inner_product_space], _p(tab), C.byref(frac),
                                          C.byref(best), C.byref(row)))
        return tab.reshape(16, 5), frac.value, best.value, row.value

    def maxdLqdt(self, hu, max_iter=99999, optim_tol=None, start_iter=0, batch_linesearch=0,
                 reset_every=25, rolling=10, trace=True, spectra_every=0, flags=0, save_every=0,
                 kappa=False, kappa_every=0, kappa_seed=0, kappa_field=None):
        """One RCG solve for the current B on the device-resident field ``hu``
        (optimization.f90:7-514).  Returns (RcgResult, RcgTrace).  kappa:
        kappaTest (tests at the start and end); kappa_every:
        kappaTestEveryXiteration; the perturbation is load-random-poly-b of
        kappa_field when given, else create-random-poly-b with kappa_seed."""
        if optim_tol is None:
            optim_tol = 1.0e-4 if self.n >= 1024 else 1.0e-5       # initialize.f90:95-100
        kf = None if kappa_field is None else _f64(kappa_field, self.vshape)
        cfg = capi.RcgCfg(int(max_iter), float(optim_tol), int(start_iter), int(batch_linesearch),
                          int(reset_every), int(rolling),
                          int(flags) | (capi.MLQ_RCG_KAPPA if kappa else 0), int(spectra_every),
                          int(save_every), int(kappa_every), C.c_ulonglong(kappa_seed),
                          None if kf is None else _p(kf))
        res = capi.RcgResult()
        tr = RcgTrace()

        def cb(user, kind, it, i0, i1, d):
            if kind == capi.MLQ_LOG_ITER:
                tr.iteration_info.append((i0, d[0], d[1], d[2]))
                tr.brackets.append((i0, d[3], d[4], int(d[5])))
            elif kind == capi.MLQ_LOG_LINEMIN:
                tr.linemin.append((it, i0, i1) + tuple(d[k] for k in range(6)))
            elif kind == capi.MLQ_LOG_BRENT:
                tr.brent.append((it, i0, d[0], d[1]))
            elif kind == capi.MLQ_LOG_MSG:
                tr.msgs.append((it, i0))
            elif kind == capi.MLQ_LOG_ERR:
                tr.errors.append((it, i0))
            elif kind == capi.MLQ_LOG_FIELD:
                a = np.ctypeslib.as_array(d, shape=(3 * self.n ** 3,)).copy()
                tr.fields.append((i0, a.reshape(self.vshape)))
            elif kind == capi.MLQ_LOG_KAPPA:
                a = np.ctypeslib.as_array(d, shape=(83,)).copy()
                tr.kappa.append((it, KAPPA_NAMES[i0], bool(i1), a[:80].reshape(16, 5), a[80],
                                 a[81], int(a[82])))
            elif kind == capi.MLQ_LOG_SPECTRUM:
                sp = np.ctypeslib.as_array(d, shape=(2 * self.n,)).copy()
                tr.spectra.append((it, "uvec" if i0 == 0 else "d", sp.reshape(2, self.n).T.copy()))

        cbf = capi.LOG_CB(cb) if trace else C.cast(None, capi.LOG_CB)
        self._chk(self.lib.mlq_rcg_solve(self.ctx, hu, C.byref(cfg), C.byref(res), cbf, None))
        return res, tr

    # -- measurement ------------------------------------------------------------------------
    def time_fft(self, batch=3, reps=5, with_batch=False):
        """Per-pass average launch durations (ms) of the transform pair of a
        batch of fields; with_batch: also the batch timed (at most the
        transform scratch)."""
        out = np.zeros(7)
        self._chk(self.lib.mlq_time_fft(self.ctx, batch, reps, _p(out)))
        return (out[:6], int(out[6])) if with_batch else out[:6]

    def autotune(self, batch=3):
        """Time the strided-pass tile candidates; returns (times[8][8] ms,
        choice[8]): passes {y fwd, z fwd, z inv, y inv} for full transforms
        (0-3) and band-limited ones (4-7); candidate = shape (0-3) + 4 *
        (XCD-contiguous order)."""
        t = np.zeros(64)
        ch = (C.c_int * 8)()
        self._chk(self.lib.mlq_autotune(self.ctx, batch, _p(t), ch))
        return t.reshape(8, 8), list(ch)

    def set_tiles(self, choice):
        """Set the tile of each strided pass (8 values as autotune returns them)."""
        ch = (C.c_int * 8)(*[int(x) for x in choice])
        self._chk(self.lib.mlq_set_tiles(self.ctx, ch))

    def time_copy(self, nfields=3, reps=5):
        r = C.c_double()
        self._chk(self.lib.mlq_time_copy(self.ctx, nfields, reps, C.byref(r)))
        return r.value

    def time_column_copy(self, axis, tile=0, nfields=3, reps=5):
        """Copy of nfields half spectra in the access pattern of the y (axis 1)
        or z (axis 2) transform pass with tile `tile`: (ms per launch, bytes)."""
        r, b = C.c_double(), C.c_double()
        self._chk(self.lib.mlq_time_column_copy(self.ctx, axis, tile, nfields, reps,
                                                C.byref(r), C.byref(b)))
        return r.value, b.value

    def sizes(self):
        a, b = C.c_double(), C.c_double()
        self._chk(self.lib.mlq_sizes(self.ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def prof_enable(self, on=True):
        self._chk(self.lib.mlq_prof_enable(self.ctx, 1 if on else 0))

    def prof(self):
        out = {}
        for cls in range(11):
            cnt, ms, by = C.c_longlong(), C.c_double(), C.c_double()
            self._chk(self.lib.mlq_prof_get(self.ctx, cls, C.byref(cnt), C.byref(ms), C.byref(by)))
            out[self.lib.mlq_prof_name(cls).decode()] = (cnt.value, ms.value, by.value)
        return out

    def mem_info(self):
        v = [C.c_double() for _ in range(4)]
        self._chk(self.lib.mlq_mem_info(self.ctx, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def mem_live(self):
        """(bytes in use, their peak): allocations minus the free space of the
        slabs that temporaries are carved from."""
        a, b = C.c_double(), C.c_double()
        self._chk(self.lib.mlq_mem_live(self.ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def transform_counts(self):
        """(forward, inverse) fields transformed since the context was created."""
        f, i = C.c_longlong(), C.c_longlong()
        self._chk(self.lib.mlq_transform_counts(self.ctx, C.byref(f), C.byref(i)))
        return f.value, i.value


# -- field files and result-file names (data_ops.f90, function_ops.f90:4590-4674) ------------

def save_field_nc(path, u):
    """save_field_R3toR3_ncdf (data_ops.f90:390-447): Ux, Uy, Uz of a vector
    field (3, n3, n2, n1) to a NetCDF classic file (library code, no GPU)."""
    lib = capi.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    if u.ndim != 4 or u.shape[0] != 3:
        raise ValueError(f"expected a vector field (3, n3, n2, n1), got {u.shape}")
    n3, n2, n1 = u.shape[1:]
    rc = lib.mlq_save_field_nc(str(path).encode(), n1, n2, n3, _p(u))
    if rc != capi.MLQ_OK:
        raise MlqError(rc, f"cannot write {path}")


def read_field_nc(path, n):
    """read_field_R3toR3_ncdf (data_ops.f90:1198-1397) of an n^3 grid."""
    lib = capi.load()
    out = np.empty((3, n, n, n))
    rc = lib.mlq_read_field_nc(str(path).encode(), n, n, n, _p(out))
    if rc != capi.MLQ_OK:
        raise MlqError(rc, f"cannot read a {n}^3 Ux/Uy/Uz field from {path}")
    return out


def result_filename(q, n, b_iter, iter_, field_name="u_result", dir_=""):
    """The reference's field-file name (data_ops.f90:31); iter_ < 0 is "end"."""
    lib = capi.load()
    buf = C.create_string_buffer(4096)
    rc = lib.mlq_result_filename(buf, len(buf), dir_.encode(), field_name.encode(), float(q),
                                 int(n), int(b_iter), int(iter_))
    if rc != capi.MLQ_OK:
        raise MlqError(rc, "result_filename")
    return buf.value.decode()


def q_text(q):
    """lebesgueQTxt (setLebesgueQtext, initialize.f90:49-70)."""
    lib = capi.load()
    buf = C.create_string_buffer(64)
    if lib.mlq_q_text(float(q), buf, len(buf)) != capi.MLQ_OK:
        raise MlqError(capi.MLQ_EINVAL, "q_text")
    return buf.value.decode()


def parse_result_filename(name):
    """set_q_resol_bIterOffset_optimIterOffsets (function_ops.f90:4635-4674):
    (q, resol, bIterOffset, optimizationIterOffset)."""
    lib = capi.load()
    q = C.c_double()
    r, b, i = C.c_int(), C.c_int(), C.c_int()
    rc = lib.mlq_parse_result_filename(str(name).encode(), C.byref(q), C.byref(r), C.byref(b),
                                       C.byref(i))
    if rc != capi.MLQ_OK:
        raise MlqError(rc, f"cannot parse {name}")
    return q.value, r.value, b.value, i.value
