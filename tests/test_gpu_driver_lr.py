"""The linear-response ('LR') workflow of the Fortran driver end to end:
ground state, impulsive kick A(t) = dAc, RT propagation, and the spectra
FourierTr writes for LR runs, _JEAF.out and _sigmaeps.out with
sigma = -JF/dAc and eps = 4 pi i sigma/omega (src/RT_calculation.f90:201-204,
246-303).  WindowFourierTr writes nothing for LR runs (:315).

Two inputs:
  * a small cell (NG=32, NK=4, two doubly occupied bands);
  * the reference's own example system (its work/input.dat: acell 3.6, NB=8,
    NBocc=1, NG=16, NK=40, Nscf=100, NCG=2, dt=0.02, LR kick dAc=1e-4, two
    ions Z=1, alpha=1 at 0.277778 and 0.722222, no sys_name) with NT cut from
    100000 to 200 steps, in the same keyword layout (including the keyword
    blocks the reader skips).
Both are checked against the oracle: GS eigenvalues, E(t) and J(t) of every
step, and the spectra of the driver's own J (as written, 12 digits).
"""
import math
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "mattohf_amd", "fortran", "mattohf_drv")

SMALL = """sys_name
  'lr'
acell
  8.0d0
NB
  2
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
  400
Nomega
  500
omegaM
  2.0
field_option
  'LR'
dAc
  1.0d-3
atoms
  1             : Nspecies
  4.0d0         : Zatom(1:Nspecies)
  1.0d0         : alphaatom(1:Nspecies)
  1             : Nion
  0.0d0  1      : Rion(ii=1), Kion(ii=1)
"""

EXAMPLE = """Band_option
  .false.
  .true.

functional
  'HF'

acell
  3.6d0

NB
  8

NBocc
  1

NG
  16
  128

NK
  40
  100

Nscf
  100
  1

NCG
  2

dt
  0.02

NT
  200
  400000

field_option
  'LR'

dAc
  0.0001


atoms
  1             : Nspecies
  1             : Zatom(1:Nspecies)
  1.0d0         : alphaatom(1:Nspecies)
  2             : Nion
  0.277778d0  1 : Rion(ii=1), Kion(ii=1)
  0.722222d0  1 : Rion(ii=2), Kion(ii=2)
"""


def run_driver(tmp_path, text):
    if not os.path.exists(DRV):
        pytest.skip("Fortran driver not built (amdflang missing)")
    (tmp_path / "input.dat").write_text(text)
    r = subprocess.run([DRV], input='"input.dat"\n', text=True, capture_output=True,
                       cwd=tmp_path, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    return tmp_path


def check_lr_outputs(d, prefix, S, vion, nscf, ncg, nt, dt, dAc, nomega, omegam, eps_tol):
    O = po.Oracle(S, vion)
    u0 = S.gram_schmidt(S.init_wf())
    ugs, eh, eps, err = O.GS_calculation(u0, nscf, ncg)
    band = np.loadtxt(d / f"{prefix}_band.out", comments="#")
    got = band[:, 3:3 + S.nb].T
    print(f"{prefix or 'example'}: max |d eps| occupied {np.abs(got - eps)[:S.nbocc].max():.2e}, "
          f"all {np.abs(got - eps).max():.2e}")
    np.testing.assert_allclose(got[:S.nbocc], eps[:S.nbocc], rtol=0, atol=1e-10)
    np.testing.assert_allclose(got, eps, rtol=0, atol=eps_tol)
    rt = np.loadtxt(d / f"{prefix}_EtotNJEAt.out", comments="#")
    assert rt.shape == (nt, 8)
    _, A, Aave, E = po.define_afield_lr(nt, dt, dAc)
    np.testing.assert_allclose(rt[:, 7], A, rtol=0, atol=1e-15)
    _, e, n, j, i3 = O.RT_calculation(ugs, dt, A, Aave)
    np.testing.assert_allclose(rt[:, 1], e, rtol=0, atol=1e-9)
    np.testing.assert_allclose(rt[:, 3], n, rtol=0, atol=1e-9)
    assert np.abs(rt[:, 5] - j).max() / np.abs(j).max() < 1e-8
    # the spectra of the driver's own J, E, A (as written)
    t = rt[:, 0]
    ref = po.fourier(t, rt[:, 5], i3[2], rt[:, 6], rt[:, 7], dt, nomega, omegam, window=0)
    jeaf = np.loadtxt(d / f"{prefix}_JEAF.out", comments="#")
    np.testing.assert_allclose(jeaf, ref, rtol=0, atol=1e-9 * np.abs(ref[:, 1:]).max())
    se = np.loadtxt(d / f"{prefix}_sigmaeps.out", comments="#")
    assert se.shape == (nomega, 5)
    om = ref[:, 0]
    sigma = -(ref[:, 1] + 1j * ref[:, 2]) / dAc
    epsw = 4 * math.pi * 1j * sigma / om
    np.testing.assert_allclose(se[:, 0], om, rtol=1e-11)
    np.testing.assert_allclose(se[:, 1] + 1j * se[:, 2], sigma, rtol=0,
                               atol=1e-9 * np.abs(sigma).max())
    np.testing.assert_allclose(se[:, 3] + 1j * se[:, 4], epsw, rtol=0,
                               atol=1e-9 * np.abs(epsw).max())
    assert not (d / f"{prefix}_JEAF_window.out").exists()


def test_lr_small_cell(tmp_path):
    d = run_driver(tmp_path, SMALL)
    S = po.System(32, 4, 2, 2, 8.0)
    vion = S.ion_potential([4.0], [1.0], [0.0])
    check_lr_outputs(d, "lr", S, vion, 150, 4, 400, 0.025, 1e-3, 500, 2.0, 1e-10)


def test_lr_reference_example_input(tmp_path):
    d = run_driver(tmp_path, EXAMPLE)
    S = po.System(16, 40, 8, 1, 3.6)
    vion = S.ion_potential([1.0, 1.0], [1.0, 1.0], [0.277778, 0.722222])
    # the dAc line of the input is read as a double (1e-4 exactly); the
    # default of the reference, 0.0001 in single precision, is not used
    check_lr_outputs(d, "", S, vion, 100, 2, 200, 0.02, 1e-4, 4000, 4.0, 1e-8)
