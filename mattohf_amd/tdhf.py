"""Python host interface of libtdhf.so (ctypes over the C-ABI of include/tdhf.h).

Method names mirror the reference routines they stand in for, so that a
caller of the reference's Fortran path finds the same operations:

    ubkwbk_hwbk   src/Functions.f90:187-209      H[u] w
    ubk_Ecomps    src/Functions.f90:265-286      (Ekin, Eion, EH, Ex)
    ubk_Etot      src/Functions.f90:74-91
    ubk_Je        src/Functions.f90:289-318
    RT_RK4_ubk    src/RT_calculation.f90:154-173
    RT_calculation src/RT_calculation.f90:2-69 (the time loop)
    GS_calculation src/GS_calculation.f90:2-121
    Gram_Schmidt  src/mATTOHF.f90:134-157

Arrays are complex128 in the reference layout ubk(NG,NB,NK) (Fortran order).
The compute path is the HIP library only: if libtdhf.so is missing or the
device is not a gfx950 GPU, the calls raise TDHFError.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libtdhf.so")
_lib = None


class TDHFError(RuntimeError):
    pass


PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_long, C.c_long)


class Params(C.Structure):
    _fields_ = [("ng", C.c_int), ("nk", C.c_int), ("nb", C.c_int), ("nbocc", C.c_int),
                ("acell", C.c_double), ("dt", C.c_double)]


# Every entry point of include/tdhf.h with its ctypes signature.
_vp, _d, _i = C.c_void_p, C.c_double, C.c_int
_PP = C.POINTER(Params)
SIGNATURES = {
    "tdhf_create": (_i, [_PP, _vp, _vp, _vp, C.POINTER(_vp)]),
    "tdhf_destroy": (_i, [_vp]),
    "tdhf_last_error": (C.c_char_p, [_vp]),
    "tdhf_ion_potential": (_i, [_PP, _i, _vp, _vp, _vp, _vp]),
    "tdhf_get_tables": (_i, [_vp, _vp, _vp]),
    "tdhf_hpsi": (_i, [_vp, _vp, _vp, _d, _vp]),
    "tdhf_observables": (_i, [_vp, _vp, _d, _vp, _vp, _vp, _vp]),
    "tdhf_rk4_step": (_i, [_vp, _vp, _d, _d]),
    "tdhf_rk4_step_noalloc": (_i, [_vp, _vp, _d, _d, _vp]),
    "tdhf_rt_run": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tdhf_state_upload": (_i, [_vp, _vp]),
    "tdhf_state_download": (_i, [_vp, _vp]),
    "tdhf_rt_propagate": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tdhf_rt_begin": (_i, [_vp, _i, _vp, _vp]),
    "tdhf_rt_advance": (_i, [_vp, _i]),
    "tdhf_rt_end": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "tdhf_gs_run": (_i, [_vp, _vp, _i, _i, _d, _vp, _vp, _vp, C.POINTER(_i)]),
    "tdhf_gram_schmidt": (_i, [_vp, _vp]),
    "tdhf_set_kernel_timing": (_i, [_vp, _i]),
    "tdhf_kernel_stats": (_i, [_vp, C.POINTER(C.c_long), C.POINTER(_d)]),
    "tdhf_fwd": (_i, [_vp, _vp, _i, _vp, _d, _d, C.POINTER(_d), _vp, _vp, _vp]),
    "tdhf_bwd": (_i, [_vp, _vp, _vp]),
    "tdhf_adjoint_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "tdhf_set_progress": (_i, [_vp, _vp, _vp]),
    "tdhf_fwd_batch": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _d, _vp, _vp, _vp, _vp]),
    "tdhf_te_step": (_i, [_vp, _vp, _d, _i]),
    "tdhf_gs_run_diag": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp, _vp]),
    "tdhf_make_dos": (_i, [_i, _i, _i, _vp, _i, _vp, _vp, _vp]),
    "tdhf_fourier": (_i, [_i, _d, _vp, _vp, _d, _vp, _vp, _i, _d, _i, _d, _i, _vp]),
    "tdhf_fp64_fma_probe": (_i, [_i, C.POINTER(_d), C.POINTER(_d)]),
    "tdhf_bwd_batch": (_i, [_vp, _vp, _vp]),
    "tdhf_batch_create": (_i, [_PP, _i, _vp, _vp, C.POINTER(_vp)]),
    "tdhf_batch_destroy": (_i, [_vp]),
    "tdhf_batch_last_error": (C.c_char_p, [_vp]),
    "tdhf_batch_upload": (_i, [_vp, _i, _vp]),
    "tdhf_batch_download": (_i, [_vp, _i, _vp]),
    "tdhf_batch_propagate": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tdhf_batch_run": (_i, [_PP, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tdhf_batch_set_waves": (_i, [_vp, _i]),
    "tdhf_batch_get_waves": (_i, [_vp]),
    "tdhf_batch_autotune": (_i, [_vp, _i, _vp, _vp, C.POINTER(_i)]),
    "tdhf_batch_set_schedule": (_i, [_vp, _i]),
    "tdhf_batch_get_schedule": (_i, [_vp]),
    "tdhf_batch_set_exchange": (_i, [_vp, _i]),
    "tdhf_batch_get_exchange": (_i, [_vp]),
    "tdhf_batch_set_kernel_timing": (_i, [_vp, _i]),
    "tdhf_batch_kernel_stats": (_i, [_vp, C.POINTER(C.c_long), C.POINTER(_d)]),
}


def load(path: str | None = None):
    """Loads libtdhf.so and declares the signatures of include/tdhf.h."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise TDHFError(f"{p} is missing: build it with `python -m mattohf_amd.build`")
    L = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _chk(rc, h=None):
    if rc != 0:
        msg = load().tdhf_last_error(h)
        raise TDHFError(f"libtdhf error {rc}: {msg.decode() if msg else ''}")


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


# ---------------------------------------------------------------------------
# set-up helpers (host side of the driver, as in the reference's Program Main)
# ---------------------------------------------------------------------------
def grids(ng, nk, acell):
    """rx, Gx, kx of src/Initialization.f90:93-112."""
    bx = 2.0 * math.pi / acell
    rx = np.array([i / ng * acell for i in range(ng)])
    gx = np.array([(i if i < ng // 2 else i - ng) * bx for i in range(ng)])
    kx = np.array([(-0.5 + (2.0 * (i + 1) - 1.0) / (2 * nk)) * bx for i in range(nk)])
    return rx, gx, kx


def init_wf(ng, nb, nk, acell):
    """Init_wf + quickrnd (src/mATTOHF.f90:159-196): LCG seed 123, k outer."""
    u = np.zeros((ng, nb, nk), dtype=np.complex128, order="F")
    rx = np.array([i / ng * acell for i in range(ng)])
    iseed = 123
    for ik in range(nk):
        for ib in range(nb):
            iseed = (iseed * 106 + 1283) % 6075
            rxc = (iseed / 6075.0) * acell
            u[:, ib, ik] = [math.exp(-0.5 * ((x - rxc) * (x - rxc))) for x in rx]
    return u


def ion_potential(ng, nk, acell, z, alpha, rion):
    """Prep_periodic_ion_potential on the device (tdhf_ion_potential)."""
    z, alpha, rion = (np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
                      for x in (z, alpha, rion))
    p = Params(ng, nk, 1, 1, float(acell), 0.0)
    v = np.empty(ng)
    _chk(load().tdhf_ion_potential(C.byref(p), len(z), _ptr(z), _ptr(alpha), _ptr(rion),
                                   _ptr(v)))
    return v


HARTREE = 27.2116
ATOMTIME = 0.02419
ATOMFIELD = 514.2


def define_afield_lp(nt, dt, omegac_ev, tpulse_fs, e0_vnm, phicep, nenvelope):
    """Define_Afield, field_option 'LP' (src/RT_calculation.f90:186-234)."""
    time = dt * np.arange(nt, dtype=np.float64)
    A = np.zeros(nt)
    omegac = omegac_ev / HARTREE
    tpulse = tpulse_fs / ATOMTIME
    e0 = e0_vnm / ATOMFIELD
    aamp = e0 / omegac
    ntp = int(min(tpulse / dt, float(nt)))
    for it in range(ntp):
        env = math.sin(math.pi * time[it] / tpulse)
        p = 1.0
        for _ in range(nenvelope):
            p *= env
        A[it] = aamp * math.sin(omegac * (time[it] - 0.5 * tpulse) + 2.0 * math.pi * phicep) * p
    Aave = np.empty(nt)
    E = np.empty(nt)
    prev = 0.0
    for it in range(nt):
        Aave[it] = 0.5 * (A[it] + prev)
        E[it] = -(A[it] - prev) / dt
        prev = A[it]
    return time, A, Aave, E


def define_afield_lr(nt, dt, dAc):
    """Define_Afield, field_option 'LR': impulsive kick A = dAc."""
    time = dt * np.arange(nt, dtype=np.float64)
    A = np.full(nt, float(dAc))
    Aave = np.empty(nt)
    E = np.empty(nt)
    prev = 0.0
    for it in range(nt):
        Aave[it] = 0.5 * (A[it] + prev)
        E[it] = -(A[it] - prev) / dt
        prev = A[it]
    return time, A, Aave, E


def pulse_lp_params(A0, omega, T):
    """Maps A(t) = A0 sin^2(pi t/T) sin(omega t) onto Define_Afield's 'LP'
    inputs: (omegac [eV], Tpulse [fs], E0 [V/nm], phiCEP, nenvelope=2)."""
    return omega * HARTREE, T * ATOMTIME, A0 * omega * ATOMFIELD, omega * T / (4 * math.pi), 2


# ---------------------------------------------------------------------------
# the context
# ---------------------------------------------------------------------------
class TDHF:
    """One system on the GPU (a tdhf_ctx)."""

    def __init__(self, ng, nk, nb, nbocc, acell, dt, vion, vcg=None, vcgkk=None):
        self.ng, self.nk, self.nb, self.nbocc = int(ng), int(nk), int(nb), int(nbocc)
        self.acell, self.dt = float(acell), float(dt)
        self.params = Params(self.ng, self.nk, self.nb, self.nbocc, self.acell, self.dt)
        vion = _f64(vion, self.ng)
        vcg = None if vcg is None else _f64(vcg, self.ng)
        vcgkk = None if vcgkk is None else _f64(vcgkk, self.ng * self.nk * self.nk)
        h = C.c_void_p()
        _chk(load().tdhf_create(C.byref(self.params), _ptr(vion), _ptr(vcg), _ptr(vcgkk),
                                C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().tdhf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (self.ng, self.nb, self.nk)

    def _u(self, u, copy=False):
        u = np.asarray(u, dtype=np.complex128)
        if u.shape != self.shape:
            u = u.reshape(self.shape, order="F")
        u = np.asfortranarray(u)
        return u.copy(order="F") if copy else u

    def tables(self):
        vcg = np.empty(self.ng)
        vq = np.empty(self.ng * (2 * self.nk - 1))
        _chk(load().tdhf_get_tables(self.h, _ptr(vcg), _ptr(vq)), self.h)
        return vcg, vq.reshape(2 * self.nk - 1, self.ng)

    def ubkwbk_hwbk(self, ubk, wbk, A):
        u, w = self._u(ubk), self._u(wbk)
        out = np.empty(self.shape, dtype=np.complex128, order="F")
        _chk(load().tdhf_hpsi(self.h, _ptr(u), _ptr(w), float(A), _ptr(out)), self.h)
        return out

    def observables(self, ubk, A):
        u = self._u(ubk)
        e, n, j = C.c_double(), C.c_double(), C.c_double()
        ec = np.empty(4)
        _chk(load().tdhf_observables(self.h, _ptr(u), float(A), C.byref(e), _ptr(ec),
                                     C.byref(n), C.byref(j)), self.h)
        return e.value, ec, n.value, j.value

    def ubk_Ecomps(self, ubk, A):
        return self.observables(ubk, A)[1]

    def ubk_Etot(self, ubk, A):
        return self.observables(ubk, A)[0]

    def ubk_Je(self, ubk, A):
        return self.observables(ubk, A)[3]

    def nele(self, ubk):
        return self.observables(ubk, 0.0)[2]

    def RT_RK4_ubk(self, ubk, Aave, A):
        u = self._u(ubk, copy=True)
        _chk(load().tdhf_rk4_step(self.h, _ptr(u), float(Aave), float(A)), self.h)
        return u

    def RT_calculation(self, ubk, A, Aave):
        """nt RK4 steps; returns (ubk, Etot, Nele, Jtot, (E0, N0, J0))."""
        u = self._u(ubk, copy=True)
        A, Aave = _f64(A), _f64(Aave, np.size(A))
        nt = A.size
        e, n, j, i3 = np.empty(nt), np.empty(nt), np.empty(nt), np.empty(3)
        _chk(load().tdhf_rt_run(self.h, _ptr(u), nt, _ptr(A), _ptr(Aave), _ptr(e), _ptr(n),
                                _ptr(j), _ptr(i3)), self.h)
        return u, e, n, j, i3

    def upload(self, ubk):
        u = self._u(ubk)
        _chk(load().tdhf_state_upload(self.h, _ptr(u)), self.h)

    def download(self):
        u = np.empty(self.shape, dtype=np.complex128, order="F")
        _chk(load().tdhf_state_download(self.h, _ptr(u)), self.h)
        return u

    def propagate(self, A, Aave, want_obs=True):
        A, Aave = _f64(A), _f64(Aave, np.size(A))
        nt = A.size
        e, n, j, i3 = (np.empty(nt), np.empty(nt), np.empty(nt), np.empty(3)) if want_obs \
            else (None, None, None, None)
        _chk(load().tdhf_rt_propagate(self.h, nt, _ptr(A), _ptr(Aave), _ptr(e), _ptr(n),
                                      _ptr(j), _ptr(i3)), self.h)
        return e, n, j, i3

    def rt_begin(self, A, Aave):
        """Uploads the field of a run (tdhf_rt_begin); the resident state is the
        one last uploaded or propagated."""
        A, Aave = _f64(A), _f64(Aave, np.size(A))
        _chk(load().tdhf_rt_begin(self.h, A.size, _ptr(A), _ptr(Aave)), self.h)
        self._rt_nt, self._rt_done = A.size, 0

    def rt_advance(self, n):
        """The next n steps of the run opened by rt_begin (tdhf_rt_advance)."""
        _chk(load().tdhf_rt_advance(self.h, int(n)), self.h)
        self._rt_done += int(n)

    def rt_end(self):
        """Observables of the steps advanced since rt_begin: (etot, nele, jtot, init3)."""
        n = getattr(self, "_rt_done", 0)
        e, nn, j, i3 = np.empty(n), np.empty(n), np.empty(n), np.empty(3)
        _chk(load().tdhf_rt_end(self.h, _ptr(e), _ptr(nn), _ptr(j), _ptr(i3)), self.h)
        return e, nn, j, i3

    def GS_calculation(self, ubk, nscf, ncg, tol=0.0):
        """Returns (ubk, Etot history[0..niter], epsbk(NB,NK), epserrorbk(NB,NK))."""
        u = self._u(ubk, copy=True)
        eh = np.empty(nscf + 1)
        eps = np.empty((self.nb, self.nk), order="F")
        err = np.empty((self.nb, self.nk), order="F")
        nit = C.c_int()
        _chk(load().tdhf_gs_run(self.h, _ptr(u), int(nscf), int(ncg), float(tol), _ptr(eh),
                                _ptr(eps), _ptr(err), C.byref(nit)), self.h)
        return u, eh[: nit.value + 1], eps, err

    def Gram_Schmidt(self, ubk):
        u = self._u(ubk, copy=True)
        _chk(load().tdhf_gram_schmidt(self.h, _ptr(u)), self.h)
        return u

    def fwd(self, ubk, A, eref=0.0, mem_budget=0.0, want_obs=True):
        """Forward pass of the adjoint pair: L = Etot(u_N, A[N-1]) - eref.
        Returns (L, u_final, etot, nele, jtot)."""
        u = self._u(ubk, copy=True)
        A = _f64(A)
        nt = A.size
        e, n, j = (np.empty(nt), np.empty(nt), np.empty(nt)) if want_obs else (None,) * 3
        L = C.c_double()
        _chk(load().tdhf_fwd(self.h, _ptr(u), nt, _ptr(A), float(eref), float(mem_budget),
                             C.byref(L), _ptr(e), _ptr(n), _ptr(j)), self.h)
        self._fwd_nt = nt
        self._fwd_ntraj = 1
        return L.value, u, e, n, j

    def bwd(self, want_u0=True):
        """Backward pass after fwd: (dL/dA[0..N-1], dL/du0 as a Euclidean cotangent
        or None)."""
        nt = getattr(self, "_fwd_nt", None)
        if not nt:
            raise TDHFError("bwd needs a preceding fwd")
        g = np.empty(nt)
        gu = np.empty(self.shape, dtype=np.complex128, order="F") if want_u0 else None
        _chk(load().tdhf_bwd(self.h, _ptr(g), _ptr(gu)), self.h)
        return g, gu

    def fwd_batch(self, us, A, objective="energy", eref=None, mem_budget=0.0, want_obs=True):
        """Forward pass for len(us) trajectories of this cell in one launch per
        stage (tdhf_fwd_batch).  A: (ntraj, nt).  objective "energy"
        (L = Etot(u_N) - eref) or "jsq" (L = dt sum_n J_n^2).
        Returns (L[ntraj], finals[list], etot, nele, jtot) with (ntraj, nt) arrays."""
        ntraj = len(us)
        A = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(ntraj, -1))
        nt = A.shape[1]
        u = np.concatenate([self._u(x).ravel(order="F") for x in us])
        obj = {"energy": 0, "jsq": 1}[objective]
        er = None if eref is None else _f64(np.broadcast_to(np.asarray(eref, dtype=np.float64),
                                                            (ntraj,)).copy())
        L = np.empty(ntraj)
        e, n, j = ((np.empty((ntraj, nt)), np.empty((ntraj, nt)), np.empty((ntraj, nt)))
                   if want_obs else (None,) * 3)
        _chk(load().tdhf_fwd_batch(self.h, ntraj, _ptr(u), nt, _ptr(A), obj, _ptr(er),
                                   float(mem_budget), _ptr(L), _ptr(e), _ptr(n), _ptr(j)), self.h)
        self._fwd_nt = nt
        self._fwd_ntraj = ntraj
        ns = int(np.prod(self.shape))
        finals = [u[t * ns:(t + 1) * ns].reshape(self.shape, order="F") for t in range(ntraj)]
        return L, finals, e, n, j

    def bwd_batch(self, want_u0=True):
        """Backward pass after fwd_batch: (dL/dA (ntraj, nt), list of dL/du0 or None)."""
        nt, ntraj = getattr(self, "_fwd_nt", None), getattr(self, "_fwd_ntraj", 1)
        if not nt:
            raise TDHFError("bwd_batch needs a preceding fwd_batch")
        g = np.empty((ntraj, nt))
        ns = int(np.prod(self.shape))
        gu = np.empty(ntraj * ns, dtype=np.complex128) if want_u0 else None
        _chk(load().tdhf_bwd_batch(self.h, _ptr(g), _ptr(gu)), self.h)
        gus = ([gu[t * ns:(t + 1) * ns].reshape(self.shape, order="F") for t in range(ntraj)]
               if want_u0 else None)
        return g, gus

    def TE_step(self, ubk, Aave, variant=0):
        """RT_TE_ubk (0), RT_TEPC_ubk (1), RT_TEPC2_ubk (2): one Taylor step."""
        u = self._u(ubk, copy=True)
        _chk(load().tdhf_te_step(self.h, _ptr(u), float(Aave), int(variant)), self.h)
        return u

    def GS_calculation_diag(self, ubk, nscf, ncg):
        """GS_calculation with the per-iteration _GSinfo.out columns:
        (u, etot_hist, epsbk, epserr, diag(nscf, 5))."""
        u = self._u(ubk, copy=True)
        eh = np.empty(nscf + 1)
        eps = np.empty((self.nb, self.nk), order="F")
        err = np.empty((self.nb, self.nk), order="F")
        dg = np.empty((max(nscf, 1), 5))
        _chk(load().tdhf_gs_run_diag(self.h, _ptr(u), int(nscf), int(ncg), _ptr(eh), _ptr(eps),
                                     _ptr(err), _ptr(dg)), self.h)
        return u, eh, eps, err, dg[:nscf]

    def set_progress(self, fn):
        """fn(phase, done, total), called after each checkpoint segment of the
        checkpointed adjoint calls (phase 0 forward, 1 backward); None: off."""
        if fn is None:
            self._progress = None
            _chk(load().tdhf_set_progress(self.h, None, None), self.h)
            return
        self._progress = PROGRESS_FN(lambda user, phase, done, total: fn(phase, done, total))
        _chk(load().tdhf_set_progress(self.h, C.cast(self._progress, C.c_void_p), None), self.h)

    def adjoint_info(self):
        f, s = C.c_int(), C.c_int()
        _chk(load().tdhf_adjoint_info(self.h, C.byref(f), C.byref(s)), self.h)
        return bool(f.value), s.value

    def set_kernel_timing(self, on, block=False):
        """on: time the fused apply kernel with HIP events; block: one event pair
        around the steps of each propagate call (4 launches per step) instead of
        one per launch."""
        _chk(load().tdhf_set_kernel_timing(self.h, (2 if block else 1) if on else 0), self.h)

    def kernel_stats(self):
        n, ms = C.c_long(), C.c_double()
        _chk(load().tdhf_kernel_stats(self.h, C.byref(n), C.byref(ms)), self.h)
        return n.value, ms.value


def _bchk(rc, h):
    if rc != 0:
        msg = load().tdhf_batch_last_error(h)
        raise TDHFError(f"libtdhf error {rc}: {msg.decode() if msg else ''}")


class Batch:
    """Independent trajectories of one cell advanced together (tdhf_batch).

    nks[t] is the k-mesh of trajectory t; vions is (ntraj, NG).  Orbitals are
    passed as a list of (NG, NB, nk_t) arrays."""

    def __init__(self, ng, nb, nbocc, acell, dt, nks, vions):
        self.ng, self.nb, self.nbocc = int(ng), int(nb), int(nbocc)
        self.acell, self.dt = float(acell), float(dt)
        self.nks = np.ascontiguousarray(nks, dtype=np.int32)
        self.ntraj = len(self.nks)
        v = np.ascontiguousarray(vions, dtype=np.float64).reshape(self.ntraj, self.ng)
        self.params = Params(self.ng, 0, self.nb, self.nbocc, self.acell, self.dt)
        h = C.c_void_p()
        rc = load().tdhf_batch_create(C.byref(self.params), self.ntraj, _ptr(self.nks), _ptr(v),
                                      C.byref(h))
        if rc != 0:
            _chk(rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().tdhf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def shapes(self):
        return [(self.ng, self.nb, int(k)) for k in self.nks]

    def upload(self, us):
        flat = np.concatenate([np.asfortranarray(u, dtype=np.complex128).ravel(order="F")
                               for u in us])
        _bchk(load().tdhf_batch_upload(self.h, -1, _ptr(flat)), self.h)

    def download(self):
        n = sum(int(np.prod(s)) for s in self.shapes())
        flat = np.empty(n, dtype=np.complex128)
        _bchk(load().tdhf_batch_download(self.h, -1, _ptr(flat)), self.h)
        out, o = [], 0
        for s in self.shapes():
            m = int(np.prod(s))
            out.append(flat[o:o + m].reshape(s, order="F"))
            o += m
        return out

    def propagate(self, A, Aave, want_obs=True):
        """A, Aave: (ntraj, nt).  Returns (etot, nele, jtot) each (ntraj, nt) and init3."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        Aave = np.ascontiguousarray(Aave, dtype=np.float64)
        nt = A.shape[1]
        if want_obs:
            e, n, j = (np.empty((self.ntraj, nt)) for _ in range(3))
            i3 = np.empty((self.ntraj, 3))
        else:
            e = n = j = i3 = None
        _bchk(load().tdhf_batch_propagate(self.h, nt, _ptr(A), _ptr(Aave), _ptr(e), _ptr(n),
                                          _ptr(j), _ptr(i3)), self.h)
        return e, n, j, i3

    def set_waves(self, nw):
        _bchk(load().tdhf_batch_set_waves(self.h, int(nw)), self.h)

    def waves(self):
        return load().tdhf_batch_get_waves(self.h)

    def set_schedule(self, sched):
        _bchk(load().tdhf_batch_set_schedule(self.h, int(sched)), self.h)

    def schedule(self):
        return load().tdhf_batch_get_schedule(self.h)

    def set_exchange(self, mode):
        """0: row-wise exchange; 1, 2, 3: pair tiles of 64, 32, 16 rows (NG=64,
        NB == NBocc)."""
        _bchk(load().tdhf_batch_set_exchange(self.h, int(mode)), self.h)

    def exchange(self):
        return load().tdhf_batch_get_exchange(self.h)

    @staticmethod
    def exchange_name(mode):
        return "row-wise" if mode == 0 else f"pair tiles of {128 >> mode} rows"

    def autotune(self, A, Aave):
        A = np.ascontiguousarray(A, dtype=np.float64)
        Aave = np.ascontiguousarray(Aave, dtype=np.float64)
        best = C.c_int()
        _bchk(load().tdhf_batch_autotune(self.h, A.shape[1], _ptr(A), _ptr(Aave), C.byref(best)),
              self.h)
        return best.value

    def set_kernel_timing(self, on):
        _bchk(load().tdhf_batch_set_kernel_timing(self.h, 1 if on else 0), self.h)

    def kernel_stats(self):
        n, ms = C.c_long(), C.c_double()
        _bchk(load().tdhf_batch_kernel_stats(self.h, C.byref(n), C.byref(ms)), self.h)
        return n.value, ms.value


# ---------------------------------------------------------------------------
# pulse parameterisation by piecewise-linear knots (the adjoint workload)
# ---------------------------------------------------------------------------
def make_dos(epsbk, nbocc, ndos=10000):
    """Make_DoS on GS_calculation's energy grid (tdhf_make_dos): (epsdos, dos, dosocc)."""
    e = np.asfortranarray(np.asarray(epsbk, dtype=np.float64))
    nb, nk = e.shape
    x, d, q = np.empty(ndos), np.empty(ndos), np.empty(ndos)
    _chk(load().tdhf_make_dos(nb, nk, int(nbocc), _ptr(e), int(ndos), _ptr(x), _ptr(d), _ptr(q)))
    return x, d, q


def fourier(time, J, j0, E, A, dt, nomega, omegam, window=0, tpulse=0.0, nenvelope=2):
    """FourierTr (window 0) / WindowFourierTr (window 1) on the device
    (tdhf_fourier): (nomega, 7) rows omega, Re/Im JF, Re/Im EF, Re/Im AF."""
    t, J, E, A = (np.ascontiguousarray(v, dtype=np.float64) for v in (time, J, E, A))
    out = np.empty((int(nomega), 7))
    _chk(load().tdhf_fourier(t.size, float(dt), _ptr(t), _ptr(J), float(j0), _ptr(E), _ptr(A),
                             int(nomega), float(omegam), int(window), float(tpulse),
                             int(nenvelope), _ptr(out)))
    return out


def fp64_fma_probe(iters=400000):
    """Sustained FP64 FMA rate of the device in TFLOP/s and the timed launch's
    milliseconds (tdhf_fp64_fma_probe)."""
    tf, ms = C.c_double(), C.c_double()
    _chk(load().tdhf_fp64_fma_probe(int(iters), C.byref(tf), C.byref(ms)))
    return tf.value, ms.value


def knot_matrix(nt, dt, T, nknots):
    """P (nt x nknots): A[it] = sum_m P[it,m] c_m is the piecewise-linear
    interpolation of knots c_m at tau_m = m T/(nknots-1), times the sin^2
    envelope sin^2(pi t/T) (zero for t >= T), at t = dt*it (the time of A[it]
    in Define_Afield)."""
    t = dt * np.arange(nt)
    tau = np.linspace(0.0, T, nknots)
    P = np.zeros((nt, nknots))
    for m in range(nknots):
        e = np.zeros(nknots)
        e[m] = 1.0
        P[:, m] = np.interp(t, tau, e, right=0.0)
    env = np.where(t < T, np.sin(np.pi * t / T) ** 2, 0.0)
    return P * env[:, None]
