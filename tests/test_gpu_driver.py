"""The Fortran driver (mattohf_amd/fortran, iso_c_binding over libtdhf.so)
run end to end on a small system (one empty band), its output files checked
against the oracle: _GSinfo.out (all per-iteration columns), _band.out,
_DoS.out, _EtotNJEAt.out, _JEAF.out and _JEAF_window.out.

Input keywords and output formats are the reference's (src/mATTOHF.f90:198-272,
src/GS_calculation.f90:22-24,48-51,93-117, src/RT_calculation.f90:22-24,52-53,
246-351).
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from mattohf_amd import tdhf as T

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "mattohf_amd", "fortran", "mattohf_drv")

INPUT = """sys_name
  'drv'
acell
  8.0d0
NB
  3
NBocc
  2
NG
  32
NK
  4
Nscf
  150
NCG
  4
dt
  0.025
NT
  300
field_option
  'LP'
omegac
  {omegac}
Tpulse
  {tpulse}
E0
  {e0}
phiCEP
  {phicep}
nenvelope
  2
atoms
  1             : Nspecies
  4.0d0         : Zatom(1:Nspecies)
  1.0d0         : alphaatom(1:Nspecies)
  1             : Nion
  0.0d0  1      : Rion(ii=1), Kion(ii=1)
"""


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(DRV):
        pytest.skip("Fortran driver not built (amdflang missing)")
    d = tmp_path_factory.mktemp("drv")
    w, A0, Tp = 0.057, 0.2, 7.5
    om, tp, e0, ph, _ = T.pulse_lp_params(A0, w, Tp)
    (d / "input.dat").write_text(INPUT.format(omegac=repr(om), tpulse=repr(tp), e0=repr(e0),
                                              phicep=repr(ph)))
    r = subprocess.run([DRV], input='"input.dat"\n', text=True, capture_output=True, cwd=d,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    return d, (om, tp, e0, ph)


def test_driver_outputs_match_oracle(run):
    d, lp = run
    S = po.System(32, 4, 3, 2, 8.0)
    vion = S.ion_potential([4.0], [1.0], [0.0])
    O = po.Oracle(S, vion)
    u0 = S.gram_schmidt(S.init_wf())
    ugs, eh, eps, err, dg = O.GS_calculation_diag(u0, 150, 4)
    band = np.loadtxt(d / "drv_band.out", comments="#")
    np.testing.assert_allclose(band[:, 3:6].T, eps, rtol=0, atol=1e-9)
    gsinfo = np.loadtxt(d / "drv_GSinfo.out", comments="#")
    assert gsinfo.shape == (150, 8)
    np.testing.assert_allclose(gsinfo[:, 1], eh[1:], rtol=0, atol=1e-9)
    # density change and gaps at every iteration; the residual columns are
    # compared tightly while they are large (the first 20 iterations): later,
    # near convergence, their size is set by rounding-level differences of
    # the two implementations' orbitals, so only their magnitude is checked
    np.testing.assert_allclose(gsinfo[:, [3, 6, 7]], dg[:, [0, 3, 4]], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(gsinfo[:20, 4:6], dg[:20, 1:3], rtol=1e-6, atol=1e-9)
    ratio = gsinfo[20:, 4:6] / dg[20:, 1:3]
    assert np.all((ratio > 0.1) & (ratio < 10.0))
    dos = np.loadtxt(d / "drv_DoS.out", comments="#")
    # oracle on the 12-digit energies of _band.out: the Gaussians (width
    # 5 de) magnify that rounding to ~1e-8 relative
    x, dd, dq = po.make_dos(band[:, 3:6].T, 2, 10000)
    want = np.stack([x, dd, dq], axis=1)
    np.testing.assert_allclose(dos, want, rtol=1e-7, atol=1e-10 * np.abs(want).max())
    rt = np.loadtxt(d / "drv_EtotNJEAt.out", comments="#")
    assert rt.shape == (300, 8)
    _, A, Aave, E = po.define_afield_lp(300, 0.025, *lp, 2)
    np.testing.assert_allclose(rt[:, 7], A, rtol=0, atol=1e-11)
    np.testing.assert_allclose(rt[:, 6], E, rtol=1e-9, atol=1e-11)
    _, e, n, j, i3 = O.RT_calculation(ugs, 0.025, A, Aave)
    np.testing.assert_allclose(rt[:, 1], e, rtol=0, atol=1e-9)
    np.testing.assert_allclose(rt[:, 3], n, rtol=0, atol=1e-9)
    assert np.abs(rt[:, 5] - j).max() / np.abs(j).max() < 1e-6
    # spectra of the driver's own J, E, A (as written, 12 digits)
    t = rt[:, 0]
    jeaf = np.loadtxt(d / "drv_JEAF.out", comments="#")
    ref = po.fourier(t, rt[:, 5], i3[2], rt[:, 6], rt[:, 7], 0.025, 4000, 4.0, window=0)
    np.testing.assert_allclose(jeaf, ref, rtol=0, atol=1e-9 * np.abs(ref[:, 1:]).max())
    jw = np.loadtxt(d / "drv_JEAF_window.out", comments="#")
    ref = po.fourier(t, rt[:, 5], i3[2], rt[:, 6], rt[:, 7], 0.025, 4000, 4.0, window=1,
                     tpulse=lp[1] / 0.02419, nenvelope=2)
    np.testing.assert_allclose(jw, ref, rtol=0, atol=1e-9 * np.abs(ref[:, 1:]).max())
