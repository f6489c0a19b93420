"""C-ABI and host-side logic, CPU only (no kernel is launched).

* libtdhf.so exports every entry point declared in include/tdhf.h, and the
  ctypes signatures of mattohf_amd.tdhf cover exactly that set.
* Without a usable gfx950 device the library refuses loudly (error code and
  message), i.e. there is no silent CPU path.
* Host set-up helpers (Init_wf, Define_Afield, grids) equal the oracle's.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as po
from mattohf_amd import tdhf as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tdhf.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tdhf_[a-z0-9_]+)\s*\(", src)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_header_symbols_exported():
    names = declared()
    assert len(names) >= 15
    syms = exported(T.LIB_PATH)
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    assert set(names) == set(T.SIGNATURES), set(names) ^ set(T.SIGNATURES)


def test_library_loads_and_reports_errors():
    L = T.load()
    assert L.tdhf_last_error(None) is not None
    p = T.Params(33, 4, 2, 2, 8.0, 0.01)  # unsupported NG: rejected before any device call
    import ctypes as C
    h = C.c_void_p()
    v = np.zeros(33)
    rc = L.tdhf_create(C.byref(p), v.ctypes.data_as(C.c_void_p), None, None, C.byref(h))
    assert rc == -1 and b"NG=33" in L.tdhf_last_error(None)


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_no_device_fails_loudly():
    vion = np.zeros(32)
    with pytest.raises(T.TDHFError) as e:
        T.TDHF(32, 4, 2, 2, 8.0, 0.02, vion)
    assert "device" in str(e.value).lower() or "hip" in str(e.value).lower()


def test_init_wf_and_grids_match_oracle():
    S = po.System(32, 5, 3, 2, 8.0)
    assert np.array_equal(T.init_wf(32, 3, 5, 8.0), S.init_wf())
    for a, b in zip(T.grids(32, 5, 8.0), S.grids()):
        np.testing.assert_allclose(a, b, rtol=1e-15, atol=1e-15)


def test_define_afield_matches_oracle():
    lp = T.pulse_lp_params(0.3, 0.057, 400.0)
    for nt, dt in ((5000, 0.1), (120, 0.01)):
        a = T.define_afield_lp(nt, dt, *lp)
        b = po.define_afield_lp(nt, dt, *lp)
        for x, y in zip(a, b):
            np.testing.assert_allclose(x, y, rtol=1e-15, atol=1e-15)
    a = T.define_afield_lr(50, 0.02, 1e-4)
    b = po.define_afield_lr(50, 0.02, 1e-4)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=0, atol=0)
