"""tdhf_rk4_step_noalloc (the RK4 step that allocates nothing, stage buffers in
the library's fixed device array) against tdhf_rk4_step on the same inputs:
two calls in a row on different systems, the second one smaller, and the
size limit.  The device buffers of the test come from the HIP runtime
through ctypes."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from mattohf_amd import tdhf as T

pytestmark = pytest.mark.gpu


def hip():
    L = C.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipDeviceSynchronize.argtypes = []
    return L


H2D, D2H = 1, 2


def random_orbitals(S, seed):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    return S.gram_schmidt(np.asfortranarray(u))


def step_noalloc(G, u, Aave, A):
    h = hip()
    d = C.c_void_p()
    nbytes = u.nbytes
    assert h.hipMalloc(C.byref(d), nbytes) == 0
    try:
        uu = np.asfortranarray(u).copy(order="F")
        assert h.hipMemcpy(d, uu.ctypes.data_as(C.c_void_p), nbytes, H2D) == 0
        rc = T.load().tdhf_rk4_step_noalloc(G.h, d, float(Aave), float(A), None)
        assert rc == 0, T.load().tdhf_last_error(G.h)
        assert h.hipDeviceSynchronize() == 0
        out = np.empty_like(uu, order="F")
        assert h.hipMemcpy(out.ctypes.data_as(C.c_void_p), d, nbytes, D2H) == 0
        return out
    finally:
        h.hipFree(d)


@pytest.mark.parametrize("sys_", [(64, 8, 4, 10.0, 8.0), (32, 4, 2, 8.0, 4.0)],
                         ids=["ng64_nk8_nb4", "ng32_nk4_nb2_smaller"])
def test_noalloc_step_equals_rk4_step(sys_):
    ng, nk, nb, a, z = sys_
    S = po.System(ng, nk, nb, nb, a)
    vion = S.ion_potential([z], [1.0], [0.0])
    u = random_orbitals(S, 7 + ng)
    with T.TDHF(ng, nk, nb, nb, a, 0.01, vion) as G:
        ref = G.RT_RK4_ubk(u, 0.05, 0.11)
        got = step_noalloc(G, u, 0.05, 0.11)
        got2 = step_noalloc(G, got, 0.07, 0.02)
        ref2 = G.RT_RK4_ubk(ref, 0.07, 0.02)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got2, ref2)


def test_noalloc_rejects_states_beyond_its_array():
    ng, nk, nb = 128, 160, 8  # 163840 complex > 131072
    S = po.System(ng, nk, nb, nb, 12.0)
    with T.TDHF(ng, nk, nb, nb, 12.0, 0.004, np.zeros(ng)) as G:
        rc = T.load().tdhf_rk4_step_noalloc(G.h, C.c_void_p(16), 0.0, 0.0, None)
        assert rc == -1
        assert b"exceeds" in T.load().tdhf_last_error(G.h)
