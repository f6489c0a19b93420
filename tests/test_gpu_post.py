"""Taylor propagators, GS diagnostics, Make_DoS and the spectra through the
C-ABI (tdhf_te_step, tdhf_gs_run_diag, tdhf_make_dos, tdhf_fourier) against
the oracle (oracle/tdhf_oracle.c: orc_te_step, orc_gs_run_diag, orc_make_dos,
orc_fourier), on seeded inputs; and the FP64 FMA probe the benchmark reads
(tdhf_fp64_fma_probe)."""
import numpy as np
import pytest

import pyoracle as po
from mattohf_amd import tdhf as T

pytestmark = pytest.mark.gpu


def cell(ng=32, nk=4, nb=3, nbocc=2, a=8.0, z=4.0):
    S = po.System(ng, nk, nb, nbocc, a)
    vion = S.ion_potential([z], [1.0], [0.0])
    return S, po.Oracle(S, vion), vion


@pytest.mark.parametrize("variant", [0, 1, 2], ids=["TE", "TEPC", "TEPC2"])
def test_te_step_matches_oracle(variant):
    S, O, vion = cell()
    u = S.gram_schmidt(S.init_wf())
    with T.TDHF(32, 4, 3, 2, 8.0, 0.02, vion) as G:
        ug, uo = u.copy(), u.copy()
        for aave in (0.05, 0.12, -0.07):
            ug = G.TE_step(ug, aave, variant)
            uo = O.TE_step(uo, 0.02, aave, variant)
        np.testing.assert_allclose(ug, uo, rtol=0, atol=1e-12)


def test_gs_diag_matches_oracle():
    S, O, vion = cell()
    u0 = S.gram_schmidt(S.init_wf())
    uo, eho, epso, erro, dgo = O.GS_calculation_diag(u0, 12, 4)
    with T.TDHF(32, 4, 3, 2, 8.0, 0.02, vion) as G:
        ug, ehg, epsg, errg, dgg = G.GS_calculation_diag(u0, 12, 4)
    np.testing.assert_allclose(ehg, eho, rtol=0, atol=1e-10)
    np.testing.assert_allclose(epsg, epso, rtol=0, atol=1e-10)
    np.testing.assert_allclose(errg, erro, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(dgg, dgo, rtol=1e-6, atol=1e-9)


def test_gs_diag_without_empty_bands():
    S, O, vion = cell(nb=2)
    u0 = S.gram_schmidt(S.init_wf())
    with T.TDHF(32, 4, 2, 2, 8.0, 0.02, vion) as G:
        dg = G.GS_calculation_diag(u0, 3, 4)[4]
    assert np.all(dg[:, 3] == np.finfo(float).max) and np.all(dg[:, 4] == np.finfo(float).max)


def test_make_dos_matches_oracle():
    rng = np.random.default_rng(4)
    eps = np.sort(rng.uniform(-3.0, 2.0, (4, 16)), axis=0)
    for nbocc in (2, 4):
        got = T.make_dos(eps, nbocc, 10000)
        want = po.make_dos(eps, nbocc, 10000)
        for g, w in zip(got, want):
            np.testing.assert_allclose(g, w, rtol=1e-12, atol=1e-14 * np.abs(w).max())


@pytest.mark.parametrize("window", [0, 1])
def test_fourier_matches_oracle(window):
    rng = np.random.default_rng(6)
    nt, dt = 3000, 0.02
    t = dt * np.arange(nt)
    J, E, A = rng.standard_normal(nt), rng.standard_normal(nt), rng.standard_normal(nt)
    got = T.fourier(t, J, 0.2, E, A, dt, 700, 4.0, window=window, tpulse=37.3, nenvelope=4)
    want = po.fourier(t, J, 0.2, E, A, dt, 700, 4.0, window=window, tpulse=37.3, nenvelope=4)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_post_argument_errors():
    with pytest.raises(T.TDHFError):
        T.fourier(np.zeros(3), np.zeros(3), 0.0, np.zeros(3), np.zeros(3), 0.1, 5, 1.0, window=1,
                  tpulse=0.0)
    with pytest.raises(T.TDHFError):
        T.make_dos(np.zeros((2, 2)), 3, 100)


def test_fp64_fma_probe():
    # the sustained FP64 FMA rate: positive, below the 78.6 TFLOP/s the
    # published peak gives, and the timed launch long enough to be measured
    tf, ms = T.fp64_fma_probe(20000)
    assert 1.0 < tf < 80.0 and ms > 0.0
    with pytest.raises(T.TDHFError):
        T.fp64_fma_probe(0)
