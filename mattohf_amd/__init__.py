"""mattohf_amd -- MI355X (gfx950) real-time TDHF path of mATTOHF.

The compute path is libtdhf.so (HIP kernels in mattohf_amd/csrc, C-ABI in
include/tdhf.h).  mattohf_amd.tdhf is the Python host interface; a Fortran
driver binds the same ABI through iso_c_binding (mattohf_amd/fortran).
"""
from .tdhf import TDHF, TDHFError, load  # noqa: F401

__all__ = ["TDHF", "TDHFError", "load"]
