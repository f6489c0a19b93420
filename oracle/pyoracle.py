"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.  Function names
follow the reference routines they restate (src/Functions.f90,
src/RT_calculation.f90, src/GS_calculation.f90, src/mATTOHF.f90).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class OrcSys(C.Structure):
    _fields_ = [("ng", C.c_int), ("nk", C.c_int), ("nb", C.c_int), ("nbocc", C.c_int),
                ("acell", C.c_double)]


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_Z = np.ctypeslib.ndpointer(dtype=np.complex128, flags="F_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, dp, ip, d, i = C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_double, C.c_int
        sp = C.POINTER(OrcSys)
        sig = {
            "orc_grids": (None, [sp, vp, vp, vp]),
            "orc_kvck_nsteps": (C.c_long, [sp]),
            "orc_k_vck": (d, [sp, d]),
            "orc_vcg": (None, [sp, vp]),
            "orc_vcgkk_full": (None, [sp, vp]),
            "orc_ion_potential": (None, [sp, i, vp, vp, vp, vp]),
            "orc_init_wf": (None, [sp, vp]),
            "orc_gram_schmidt": (None, [sp, vp]),
            "orc_fft": (None, [i, i, vp, vp]),
            "orc_create": (vp, [sp, vp, vp]),
            "orc_destroy": (None, [vp]),
            "orc_set_threads": (None, [i]),
            "orc_dns": (None, [vp, vp, vp]),
            "orc_vh": (None, [vp, vp, vp]),
            "orc_tpsi": (None, [vp, vp, d, vp]),
            "orc_vfpsi": (None, [vp, vp, vp, vp]),
            "orc_hpsi": (None, [vp, vp, vp, d, vp]),
            "orc_ecomps": (None, [vp, vp, d, vp]),
            "orc_etot": (d, [vp, vp, d]),
            "orc_je": (d, [vp, vp, d]),
            "orc_nele": (d, [vp, vp]),
            "orc_epsbk": (None, [vp, vp, vp, vp]),
            "orc_rk4_step": (None, [vp, vp, d, d, d]),
            "orc_rt_run": (None, [vp, vp, d, i, vp, vp, vp, vp, vp, vp]),
            "orc_define_afield_lp": (None, [i, d, d, d, d, d, i, vp, vp, vp, vp]),
            "orc_define_afield_lr": (None, [i, d, d, vp, vp, vp, vp]),
            "orc_heev": (i, [i, vp, vp]),
            "orc_ks_krylov": (None, [vp, vp, i]),
            "orc_subspace_diag": (None, [vp, vp]),
            "orc_gs_run": (None, [vp, vp, i, i, vp, vp, vp]),
            "orc_adjoint_energy": (d, [vp, vp, d, i, vp, d, vp, vp]),
            "orc_objective_energy": (d, [vp, vp, d, i, vp, d]),
            "orc_adjoint_jsq": (d, [vp, vp, d, i, vp, vp, vp]),
            "orc_objective_jsq": (d, [vp, vp, d, i, vp]),
            "orc_gs_run_diag": (None, [vp, vp, i, i, vp, vp, vp, vp]),
            "orc_make_dos": (None, [i, i, i, vp, i, vp, vp, vp]),
            "orc_fourier": (None, [i, d, vp, vp, d, vp, vp, i, d, i, d, i, vp]),
            "orc_te_step": (None, [vp, vp, d, d, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _fz(a, shape=None):
    a = np.asarray(a, dtype=np.complex128)
    if shape is not None:
        a = a.reshape(shape, order="F")
    return np.asfortranarray(a)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def fft(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_fft(x.size, sign, _p(x), _p(out))
    return out


class System:
    """Module-global state of src/Global_variables.f90 for one system."""

    def __init__(self, ng, nk, nb, nbocc, acell):
        self.ng, self.nk, self.nb, self.nbocc, self.acell = ng, nk, nb, nbocc, float(acell)
        self.s = OrcSys(ng, nk, nb, nbocc, float(acell))

    @property
    def shape(self):
        return (self.ng, self.nb, self.nk)

    def grids(self):
        rx, gx, kx = np.empty(self.ng), np.empty(self.ng), np.empty(self.nk)
        lib().orc_grids(C.byref(self.s), _p(rx), _p(gx), _p(kx))
        return rx, gx, kx

    def k_vCk(self, q):
        return lib().orc_k_vck(C.byref(self.s), float(q))

    def kvck_nsteps(self):
        return lib().orc_kvck_nsteps(C.byref(self.s))

    def vCG(self):
        v = np.empty(self.ng)
        lib().orc_vcg(C.byref(self.s), _p(v))
        return v

    def vCGkk(self):
        v = np.empty(self.ng * self.nk * self.nk)
        lib().orc_vcgkk_full(C.byref(self.s), _p(v))
        return v.reshape(self.nk * self.nk, self.ng)  # [ikk][ig]

    def ion_potential(self, z, alpha, rion):
        z, alpha, rion = (np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64)
                          for a in (z, alpha, rion))
        v = np.empty(self.ng)
        lib().orc_ion_potential(C.byref(self.s), len(z), _p(z), _p(alpha), _p(rion), _p(v))
        return v

    def init_wf(self):
        u = np.zeros(self.shape, dtype=np.complex128, order="F")
        lib().orc_init_wf(C.byref(self.s), _p(u))
        return u

    def gram_schmidt(self, u):
        u = _fz(u, self.shape).copy(order="F")
        lib().orc_gram_schmidt(C.byref(self.s), _p(u))
        return u


class Oracle:
    """CPU restatement bound to one system and ion potential."""

    def __init__(self, system: System, vion, vcgkk=None):
        self.sys = system
        self.vion = np.ascontiguousarray(vion, dtype=np.float64)
        vk = None if vcgkk is None else np.ascontiguousarray(vcgkk, dtype=np.float64)
        self.h = lib().orc_create(C.byref(system.s), _p(self.vion), _p(vk))

    def __del__(self):
        try:
            if self.h:
                lib().orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _u(self, u):
        return _fz(u, self.sys.shape)

    def ubk_dns(self, u):
        d = np.empty(self.sys.ng)
        lib().orc_dns(self.h, _p(self._u(u)), _p(d))
        return d

    def dns_vH(self, dns):
        dns = np.ascontiguousarray(dns, dtype=np.float64)
        v = np.empty(self.sys.ng)
        lib().orc_vh(self.h, _p(dns), _p(v))
        return v

    def ubk_tubk(self, w, A):
        w = self._u(w)
        out = np.empty_like(w, order="F")
        lib().orc_tpsi(self.h, _p(w), float(A), _p(out))
        return out

    def ubkwbk_vFwbk(self, u, w):
        u, w = self._u(u), self._u(w)
        out = np.empty_like(w, order="F")
        lib().orc_vfpsi(self.h, _p(u), _p(w), _p(out))
        return out

    def ubkwbk_hwbk(self, u, w, A):
        u, w = self._u(u), self._u(w)
        out = np.empty_like(w, order="F")
        lib().orc_hpsi(self.h, _p(u), _p(w), float(A), _p(out))
        return out

    def ubk_Ecomps(self, u, A):
        ec = np.empty(4)
        lib().orc_ecomps(self.h, _p(self._u(u)), float(A), _p(ec))
        return ec

    def ubk_Etot(self, u, A):
        return lib().orc_etot(self.h, _p(self._u(u)), float(A))

    def ubk_Je(self, u, A):
        return lib().orc_je(self.h, _p(self._u(u)), float(A))

    def nele(self, u):
        return lib().orc_nele(self.h, _p(self._u(u)))

    def ubk_epsbk(self, u):
        eps = np.empty((self.sys.nb, self.sys.nk), order="F")
        err = np.empty((self.sys.nb, self.sys.nk), order="F")
        lib().orc_epsbk(self.h, _p(self._u(u)), _p(eps), _p(err))
        return eps, err

    def RT_RK4_ubk(self, u, dt, Aave, A):
        u = self._u(u).copy(order="F")
        lib().orc_rk4_step(self.h, _p(u), float(dt), float(Aave), float(A))
        return u

    def RT_calculation(self, u, dt, A, Aave):
        """Returns (u_final, etot, nele, jtot, init3=(E0,N0,J0 at A=0))."""
        u = self._u(u).copy(order="F")
        A = np.ascontiguousarray(A, dtype=np.float64)
        Aave = np.ascontiguousarray(Aave, dtype=np.float64)
        nt = len(A)
        e, n, j, i3 = np.empty(nt), np.empty(nt), np.empty(nt), np.empty(3)
        lib().orc_rt_run(self.h, _p(u), float(dt), nt, _p(A), _p(Aave), _p(e), _p(n), _p(j),
                         _p(i3))
        return u, e, n, j, i3

    def KS_krylov(self, u, ncg):
        u = self._u(u).copy(order="F")
        lib().orc_ks_krylov(self.h, _p(u), int(ncg))
        return u

    def Subspace_diag(self, u):
        u = self._u(u).copy(order="F")
        lib().orc_subspace_diag(self.h, _p(u))
        return u

    def objective_energy(self, u0, dt, A, eref=0.0):
        """L = Etot(u_N, A[N-1]) - eref after len(A) RK4 steps (Aave from A)."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        return lib().orc_objective_energy(self.h, _p(self._u(u0)), float(dt), A.size, _p(A),
                                          float(eref))

    def adjoint_energy(self, u0, dt, A, eref=0.0):
        """(L, dL/dA[0..N-1], dL/du0 as a Euclidean cotangent)."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        g = np.empty(A.size)
        gu = np.empty(self.sys.shape, dtype=np.complex128, order="F")
        L = lib().orc_adjoint_energy(self.h, _p(self._u(u0)), float(dt), A.size, _p(A),
                                     float(eref), _p(g), _p(gu))
        return L, g, gu

    def objective_jsq(self, u0, dt, A):
        """L = dt * sum_it J(u_{it+1}, A[it])^2 over len(A) RK4 steps (Aave from A)."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        return lib().orc_objective_jsq(self.h, _p(self._u(u0)), float(dt), A.size, _p(A))

    def adjoint_jsq(self, u0, dt, A):
        """(L, dL/dA[0..N-1], dL/du0) for L = dt * sum_it J_it^2."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        g = np.empty(A.size)
        gu = np.empty(self.sys.shape, dtype=np.complex128, order="F")
        L = lib().orc_adjoint_jsq(self.h, _p(self._u(u0)), float(dt), A.size, _p(A), _p(g), _p(gu))
        return L, g, gu

    def GS_calculation(self, u, nscf, ncg):
        """Returns (u, etot_hist[0..nscf], epsbk(NB,NK), epserrorbk(NB,NK))."""
        u = self._u(u).copy(order="F")
        eh = np.empty(nscf + 1)
        eps = np.empty((self.sys.nb, self.sys.nk), order="F")
        err = np.empty((self.sys.nb, self.sys.nk), order="F")
        lib().orc_gs_run(self.h, _p(u), int(nscf), int(ncg), _p(eh), _p(eps), _p(err))
        return u, eh, eps, err


    def GS_calculation_diag(self, u, nscf, ncg):
        """GS_calculation with the per-iteration columns of _GSinfo.out:
        (u, etot_hist, epsbk, epserr, diag(nscf, 5)) with diag = (mean |d dns|,
        mean err, max err, fundamental gap, optical gap)."""
        u = self._u(u).copy(order="F")
        eh = np.empty(nscf + 1)
        eps = np.empty((self.sys.nb, self.sys.nk), order="F")
        err = np.empty((self.sys.nb, self.sys.nk), order="F")
        dg = np.empty((max(nscf, 1), 5))
        lib().orc_gs_run_diag(self.h, _p(u), int(nscf), int(ncg), _p(eh), _p(eps), _p(err), _p(dg))
        return u, eh, eps, err, dg[:nscf]

    def TE_step(self, u, dt, Aave, variant=0):
        """RT_TE_ubk (0), RT_TEPC_ubk (1), RT_TEPC2_ubk (2): one step."""
        u = self._u(u).copy(order="F")
        lib().orc_te_step(self.h, _p(u), float(dt), float(Aave), int(variant))
        return u


def make_dos(epsbk, nbocc, ndos=10000):
    """Make_DoS on the GS_calculation energy grid: (epsdos, dos, dosocc)."""
    e = np.asfortranarray(np.asarray(epsbk, dtype=np.float64))
    nb, nk = e.shape
    x, d, q = np.empty(ndos), np.empty(ndos), np.empty(ndos)
    lib().orc_make_dos(nb, nk, int(nbocc), _p(e), int(ndos), _p(x), _p(d), _p(q))
    return x, d, q


def fourier(time, J, j0, E, A, dt, nomega, omegam, window=0, tpulse=0.0, nenvelope=2):
    """FourierTr (window 0) / WindowFourierTr (window 1): (nomega, 7) rows
    omega, Re/Im JF, Re/Im EF, Re/Im AF."""
    t, J, E, A = (np.ascontiguousarray(v, dtype=np.float64) for v in (time, J, E, A))
    out = np.empty((int(nomega), 7))
    lib().orc_fourier(t.size, float(dt), _p(t), _p(J), float(j0), _p(E), _p(A), int(nomega),
                      float(omegam), int(window), float(tpulse), int(nenvelope), _p(out))
    return out


def heev(a):
    a = np.asfortranarray(np.array(a, dtype=np.complex128))
    m = a.shape[0]
    w = np.empty(m)
    rc = lib().orc_heev(m, _p(a), _p(w))
    return w, a, rc


def define_afield_lp(nt, dt, omegac_ev, tpulse_fs, e0_vnm, phicep, nenvelope):
    t, A, Aave, E = (np.empty(nt) for _ in range(4))
    lib().orc_define_afield_lp(nt, float(dt), float(omegac_ev), float(tpulse_fs),
                               float(e0_vnm), float(phicep), int(nenvelope), _p(t), _p(A),
                               _p(Aave), _p(E))
    return t, A, Aave, E


def define_afield_lr(nt, dt, dAc):
    t, A, Aave, E = (np.empty(nt) for _ in range(4))
    lib().orc_define_afield_lr(nt, float(dt), float(dAc), _p(t), _p(A), _p(Aave), _p(E))
    return t, A, Aave, E
