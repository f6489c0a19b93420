"""The batched-trajectory extension called from Fortran (iso_c_binding,
mattohf_amd/fortran/tdhf_capi.f90): batch_check.f90 runs three trajectories
of uneven k-meshes through tdhf_batch_* and each through tdhf_rt_run on its
own context, and fails unless E(t), J(t) and the final orbitals agree to
1e-12."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mattohf_amd", "fortran", "batch_check")


def test_fortran_batch_equals_rt_runs():
    if not os.path.exists(EXE):
        pytest.skip("Fortran programs not built (amdflang missing)")
    r = subprocess.run([EXE], text=True, capture_output=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    m = re.search(r"max\|dE\|, max\|dJ\|/max\|J\|, max\|du\| =\s*(\S+)\s+(\S+)\s+(\S+)", r.stdout)
    assert m and all(float(x) <= 1e-12 for x in m.groups())
    assert "batch_check: OK" in r.stdout
