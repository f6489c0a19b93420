"""Batched trajectories (tdhf_batch_*): every trajectory of a batched launch
equals its own single-trajectory run (tdhf_rt_run) and the oracle.

The batch mixes k-meshes of uneven size (NK = 2, 4, 8) and different pulses,
with more steps than one captured graph chunk (25) so that graph replay and
the remainder path are both exercised.
"""
import numpy as np
import pytest

import pyoracle as po
from mattohf_amd import tdhf as T

pytestmark = pytest.mark.gpu

NG, NB, ACELL, DT, NT = 32, 2, 8.0, 0.025, 60
NKS = [8, 2, 4, 8, 4, 2, 8]


def orbitals(nk, seed):
    S = po.System(NG, nk, NB, NB, ACELL)
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    return S, S.gram_schmidt(np.asfortranarray(u))


@pytest.fixture(scope="module")
def batch_case():
    vions, us, systems, fields = [], [], [], []
    for t, nk in enumerate(NKS):
        S, u = orbitals(nk, 100 + t)
        systems.append(S)
        us.append(u)
        vions.append(S.ion_potential([4.0], [1.0], [0.0]))
        _, A, Aave, _ = T.define_afield_lp(NT, DT, *T.pulse_lp_params(0.05 + 0.05 * t,
                                                                      0.057, 1.5))
        fields.append((A, Aave))
    A = np.array([f[0] for f in fields])
    Aave = np.array([f[1] for f in fields])
    return systems, us, np.array(vions), A, Aave


@pytest.mark.parametrize("nw", [1, 2, 4, 8])
def test_batch_matches_single_runs(batch_case, nw):
    systems, us, vions, A, Aave = batch_case
    with T.Batch(NG, NB, NB, ACELL, DT, NKS, vions) as B:
        B.set_waves(nw)
        B.upload(us)
        e, n, j, i3 = B.propagate(A, Aave)
        finals = B.download()
    for t, nk in enumerate(NKS):
        with T.TDHF(NG, nk, NB, NB, ACELL, DT, vions[t]) as G:
            u1, e1, n1, j1, i31 = G.RT_calculation(us[t], A[t], Aave[t])
        np.testing.assert_allclose(e[t], e1, rtol=0, atol=1e-12)
        np.testing.assert_allclose(n[t], n1, rtol=0, atol=1e-12)
        np.testing.assert_allclose(j[t], j1, rtol=0, atol=1e-12 * max(1.0, np.abs(j1).max()))
        np.testing.assert_allclose(i3[t], i31, rtol=0, atol=1e-12)
        np.testing.assert_allclose(finals[t], u1, rtol=0, atol=1e-12)


def test_batch_matches_oracle(batch_case):
    systems, us, vions, A, Aave = batch_case
    with T.Batch(NG, NB, NB, ACELL, DT, NKS, vions) as B:
        B.upload(us)
        e, n, j, i3 = B.propagate(A, Aave)
    for t in (0, 1, 2):
        O = po.Oracle(systems[t], vions[t])
        _, eo, no, jo, i3o = O.RT_calculation(us[t], DT, A[t], Aave[t])
        assert np.abs(e[t] - eo).max() < 1e-10
        assert np.abs(n[t] - no).max() < 1e-10
        assert np.abs(j[t] - jo).max() / np.abs(jo).max() < 1e-8


def test_autotune_keeps_state(batch_case):
    systems, us, vions, A, Aave = batch_case
    with T.Batch(NG, NB, NB, ACELL, DT, NKS, vions) as B:
        B.upload(us)
        best = B.autotune(A[:, :30], Aave[:, :30])
        assert best in (1, 2, 4, 8) and B.waves() == best
        assert B.schedule() in (0, 1, 2)
        assert B.exchange() == 0  # NG = 32: the pair tiles are for NG = 64
        after = B.download()
        for a, b in zip(after, us):
            assert np.array_equal(a, b)


def test_autotune_ng64_keeps_state_and_results():
    """NG=64 batches: the autotuner also tries the pair-tile exchange; the
    chosen configuration reproduces the single runs."""
    ng, nb, a, dt, nt, nks = 64, 2, 10.0, 0.01, 30, [4, 8, 4]
    us, vions, fields = [], [], []
    for t, nk in enumerate(nks):
        S = po.System(ng, nk, nb, nb, a)
        rng = np.random.default_rng(700 + t)
        u = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
        us.append(S.gram_schmidt(np.asfortranarray(u)))
        vions.append(S.ion_potential([8.0], [1.0], [0.0]))
        _, A, Aave, _ = T.define_afield_lp(nt, dt, *T.pulse_lp_params(0.1 + 0.1 * t, 0.057, 0.3))
        fields.append((A, Aave))
    A = np.array([f[0] for f in fields])
    Aave = np.array([f[1] for f in fields])
    with T.Batch(ng, nb, nb, a, dt, nks, np.array(vions)) as B:
        B.upload(us)
        B.autotune(A[:, :25], Aave[:, :25])
        assert B.exchange() in (0, 1)
        for x, y in zip(B.download(), us):
            assert np.array_equal(x, y)
        e, n, j, i3 = B.propagate(A, Aave)
    for t, nk in enumerate(nks):
        with T.TDHF(ng, nk, nb, nb, a, dt, vions[t]) as G:
            u1, e1, n1, j1, _ = G.RT_calculation(us[t], A[t], Aave[t])
        np.testing.assert_allclose(e[t], e1, rtol=0, atol=1e-12 * max(1.0, np.abs(e1).max()))
        np.testing.assert_allclose(j[t], j1, rtol=0, atol=1e-12 * max(1.0, np.abs(j1).max()))


@pytest.mark.parametrize("nw", [1, 2])
def test_batch_ng64_matches_single_runs(nw):
    # NG=64: the groups of a wave are summed in registers; with one wave per
    # workgroup that is the whole reduction
    ng, nb, a, dt, nt, nks = 64, 2, 10.0, 0.01, 30, [4, 2, 4]
    us, vions, fields = [], [], []
    for t, nk in enumerate(nks):
        S = po.System(ng, nk, nb, nb, a)
        rng = np.random.default_rng(300 + t)
        u = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
        us.append(S.gram_schmidt(np.asfortranarray(u)))
        vions.append(S.ion_potential([8.0], [1.0], [0.0]))
        _, A, Aave, _ = T.define_afield_lp(nt, dt, *T.pulse_lp_params(0.1 + 0.1 * t, 0.057, 0.3))
        fields.append((A, Aave))
    A = np.array([f[0] for f in fields])
    Aave = np.array([f[1] for f in fields])
    with T.Batch(ng, nb, nb, a, dt, nks, np.array(vions)) as B:
        B.set_waves(nw)
        B.upload(us)
        e, n, j, i3 = B.propagate(A, Aave)
        finals = B.download()
    for t, nk in enumerate(nks):
        with T.TDHF(ng, nk, nb, nb, a, dt, vions[t]) as G:
            u1, e1, n1, j1, _ = G.RT_calculation(us[t], A[t], Aave[t])
        np.testing.assert_allclose(e[t], e1, rtol=0, atol=1e-12 * max(1.0, np.abs(e1).max()))
        np.testing.assert_allclose(n[t], n1, rtol=0, atol=1e-12)
        np.testing.assert_allclose(j[t], j1, rtol=0, atol=1e-12 * max(1.0, np.abs(j1).max()))
        np.testing.assert_allclose(finals[t], u1, rtol=0, atol=1e-12)
