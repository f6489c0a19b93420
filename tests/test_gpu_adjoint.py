"""Adjoint (tdhf_fwd / tdhf_bwd): gradient of L = Etot(u_N, A[N-1]) - E_ref
with respect to the per-step field A and the initial orbitals.

Checked against (i) the oracle's adjoint (oracle/tdhf_oracle.c, three literal
pair transforms per exchange pair, independent of the GPU's two-transform
form with the Nyquist correction), (ii) central finite differences of the GPU
forward objective (h = 1e-5) on random directions, relative agreement 1e-6,
and (iii) checkpointed (sqrt(N) segments) vs fully taped backward runs.
"""
import numpy as np
import pytest

import pyoracle as po
from mattohf_amd import tdhf as T

pytestmark = pytest.mark.gpu

CASES = [  # ng, nk, nb, a, z, nt, dt
    (16, 3, 2, 5.0, 2.0, 12, 0.02),
    (32, 4, 2, 8.0, 4.0, 30, 0.025),
    (64, 3, 2, 10.0, 8.0, 20, 0.01),
    (128, 2, 2, 12.0, 4.0, 6, 0.004),
]


def setup_case(ng, nk, nb, a, z, nt, dt, seed=0):
    S = po.System(ng, nk, nb, nb, a)
    vion = S.ion_potential([z], [1.0], [0.0])
    O = po.Oracle(S, vion)
    u0 = S.gram_schmidt(S.init_wf())
    u0 = O.GS_calculation(u0, 60, 3)[0]
    rng = np.random.default_rng(seed)
    A = 0.1 * rng.standard_normal(nt)
    return S, O, vion, u0, A


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"ng{c[0]}_nk{c[1]}")
def test_adjoint_matches_oracle(case):
    ng, nk, nb, a, z, nt, dt = case
    S, O, vion, u0, A = setup_case(*case)
    Lo, go, guo = O.adjoint_energy(u0, dt, A)
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        L, uf, e, n, j = G.fwd(u0, A)
        g, gu = G.bwd()
    assert abs(L - Lo) < 1e-11
    assert np.abs(g - go).max() < 1e-9 * max(1.0, np.abs(go).max())
    assert np.abs(gu - guo).max() < 1e-9 * max(1.0, np.abs(guo).max())


def test_adjoint_pair_tiles_match_oracle():
    """NG = 128 with every band occupied runs the reverse stages through the
    pair tiles (32-row tiles): 48 rows make one off-diagonal tile and two
    diagonal ones, the second block partly empty."""
    ng, nk, nb, a, z, nt, dt = 128, 8, 6, 12.0, 12.0, 4, 0.004
    S = po.System(ng, nk, nb, nb, a)
    vion = S.ion_potential([z], [1.0], [0.0])
    O = po.Oracle(S, vion)
    u0 = O.GS_calculation(S.gram_schmidt(S.init_wf()), 8, 2)[0]
    A = 0.1 * np.random.default_rng(3).standard_normal(nt)
    Lo, go, guo = O.adjoint_energy(u0, dt, A)
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        L, uf, e, n, j = G.fwd(u0, A)
        g, gu = G.bwd()
    assert abs(L - Lo) < 1e-11
    assert np.abs(g - go).max() < 1e-9 * max(1.0, np.abs(go).max())
    assert np.abs(gu - guo).max() < 1e-9 * max(1.0, np.abs(guo).max())


@pytest.mark.parametrize("case", CASES[:2], ids=lambda c: f"ng{c[0]}_nk{c[1]}")
def test_adjoint_finite_differences(case):
    ng, nk, nb, a, z, nt, dt = case
    S, O, vion, u0, A = setup_case(*case)
    rng = np.random.default_rng(7)
    h = 1e-5
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        G.fwd(u0, A)
        g, gu = G.bwd()
        for _ in range(4):
            d = rng.standard_normal(nt)
            d /= np.linalg.norm(d)
            lp = G.fwd(u0, A + h * d, want_obs=False)[0]
            lm = G.fwd(u0, A - h * d, want_obs=False)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g @ d) <= 1e-6 * abs(fd) + 1e-12, (fd, g @ d)
        du = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
        du /= np.linalg.norm(du)
        lp = G.fwd(u0 + h * du, A, want_obs=False)[0]
        lm = G.fwd(u0 - h * du, A, want_obs=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - np.real(np.vdot(gu, du))) <= 1e-6 * abs(fd) + 1e-12


def test_checkpointed_equals_full_tape():
    ng, nk, nb, a, z, nt, dt = 32, 4, 2, 8.0, 4.0, 37, 0.025
    S, O, vion, u0, A = setup_case(ng, nk, nb, a, z, nt, dt)
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        L1, _, _, _, j1 = G.fwd(u0, A)
        assert G.adjoint_info()[0]
        g1, gu1 = G.bwd()
        state = ng * nb * nk * 16
        L2, _, _, _, j2 = G.fwd(u0, A, mem_budget=40 * state)
        full, seg = G.adjoint_info()
        assert not full and 1 < seg < nt
        g2, gu2 = G.bwd()
    assert L1 == L2
    np.testing.assert_array_equal(j1, j2)
    np.testing.assert_allclose(g2, g1, rtol=0, atol=1e-14 * np.abs(g1).max())
    np.testing.assert_allclose(gu2, gu1, rtol=0, atol=1e-14 * np.abs(gu1).max())


def test_forward_matches_rt_run():
    ng, nk, nb, a, z, nt, dt = 32, 4, 2, 8.0, 4.0, 30, 0.025
    S, O, vion, u0, A = setup_case(ng, nk, nb, a, z, nt, dt)
    Aave = 0.5 * (A + np.concatenate([[0.0], A[:-1]]))
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        L, uf, e, n, j = G.fwd(u0, A)
        u1, e1, n1, j1, _ = G.RT_calculation(u0, A, Aave)
    np.testing.assert_array_equal(e, e1)
    np.testing.assert_array_equal(j, j1)
    np.testing.assert_array_equal(uf, u1)


# ---- current-squared objective and many trajectories per launch ----

@pytest.mark.parametrize("case", CASES[:2], ids=lambda c: f"ng{c[0]}_nk{c[1]}")
def test_jsq_matches_oracle(case):
    """L = dt sum_n J_n^2: GPU pair vs the oracle's adjoint (pinned by finite
    differences in test_oracle.py)."""
    ng, nk, nb, a, z, nt, dt = case
    S, O, vion, u0, A = setup_case(*case)
    Lo, go, guo = O.adjoint_jsq(u0, dt, A)
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        L, _, _, _, j = G.fwd_batch([u0], A[None, :], objective="jsq")
        g, gu = G.bwd_batch()
    assert abs(L[0] - Lo) < 1e-11 * max(1.0, abs(Lo))
    assert np.abs(g[0] - go).max() < 1e-9 * max(1.0, np.abs(go).max())
    assert np.abs(gu[0] - guo).max() < 1e-9 * max(1.0, np.abs(guo).max())
    assert abs(L[0] - dt * np.sum(j[0] ** 2)) < 1e-14 * max(1.0, abs(L[0]))


def test_jsq_finite_differences():
    ng, nk, nb, a, z, nt, dt = CASES[1]
    S, O, vion, u0, A = setup_case(*CASES[1])
    rng = np.random.default_rng(11)
    h = 1e-5
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        G.fwd_batch([u0], A[None, :], objective="jsq")
        g, gu = G.bwd_batch()
        f = lambda uu, AA: G.fwd_batch([uu], AA[None, :], objective="jsq", want_obs=False)[0][0]
        for _ in range(3):
            d = rng.standard_normal(nt)
            d /= np.linalg.norm(d)
            fd = (f(u0, A + h * d) - f(u0, A - h * d)) / (2 * h)
            assert abs(fd - g[0] @ d) <= 1e-6 * abs(fd) + 1e-13, (fd, g[0] @ d)
        du = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
        du /= np.linalg.norm(du)
        fd = (f(u0 + h * du, A) - f(u0 - h * du, A)) / (2 * h)
        assert abs(fd - np.real(np.vdot(gu[0], du))) <= 1e-6 * abs(fd) + 1e-13


# the second case runs its reverse stages through the pair tiles (NG = 128,
# 48 rows: three tiles per trajectory), so the per-trajectory offsets of the
# tile kernel and its partial sums are checked against single runs
BATCH_CASES = [CASES[1], (128, 8, 6, 12.0, 12.0, 4, 0.004)]


@pytest.mark.parametrize("case", BATCH_CASES, ids=lambda c: f"ng{c[0]}_nk{c[1]}")
@pytest.mark.parametrize("objective", ["energy", "jsq"])
def test_batch_equals_single_runs(objective, case):
    """Three trajectories (own pulse, own initial state) in one launch per stage
    give each trajectory's single-run objective and gradient."""
    ng, nk, nb, a, z, nt, dt = case
    S, O, vion, u0, A0 = setup_case(*case)
    rng = np.random.default_rng(5)
    As = np.stack([A0, 0.1 * rng.standard_normal(nt), 0.05 * rng.standard_normal(nt)])
    us = [u0, S.gram_schmidt(u0 + 0.05 * rng.standard_normal(S.shape)), u0]
    er = np.array([0.1, -0.2, 0.3])
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        L, fin, e, n, j = G.fwd_batch(us, As, objective=objective, eref=er)
        g, gu = G.bwd_batch()
        for t in range(3):
            L1, fin1, e1, n1, j1 = G.fwd_batch([us[t]], As[t:t + 1], objective=objective,
                                               eref=er[t:t + 1])
            g1, gu1 = G.bwd_batch()
            assert abs(L[t] - L1[0]) <= 1e-13 * max(1.0, abs(L1[0]))
            np.testing.assert_allclose(j[t], j1[0], rtol=0, atol=1e-13 * np.abs(j1).max())
            np.testing.assert_allclose(fin[t], fin1[0], rtol=0, atol=1e-13)
            np.testing.assert_allclose(g[t], g1[0], rtol=0, atol=1e-12 * np.abs(g1).max())
            np.testing.assert_allclose(gu[t], gu1[0], rtol=0, atol=1e-12 * np.abs(gu1[0]).max())


def test_batch_jsq_checkpointed_equals_full_tape():
    ng, nk, nb, a, z, nt, dt = 32, 4, 2, 8.0, 4.0, 29, 0.025
    S, O, vion, u0, A0 = setup_case(ng, nk, nb, a, z, nt, dt)
    rng = np.random.default_rng(9)
    As = np.stack([A0, 0.1 * rng.standard_normal(nt)])
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        L1, _, _, _, _ = G.fwd_batch([u0, u0], As, objective="jsq")
        assert G.adjoint_info()[0]
        g1, gu1 = G.bwd_batch()
        state = 2 * ng * nb * nk * 16
        L2, _, _, _, _ = G.fwd_batch([u0, u0], As, objective="jsq", mem_budget=30 * state)
        full, seg = G.adjoint_info()
        assert not full and 1 < seg < nt
        g2, gu2 = G.bwd_batch()
        with pytest.raises(T.TDHFError):
            G.bwd_batch()  # the checkpointed record is spent
    np.testing.assert_array_equal(L1, L2)
    np.testing.assert_allclose(g2, g1, rtol=0, atol=1e-14 * np.abs(g1).max())
    for a_, b_ in zip(gu2, gu1):
        np.testing.assert_allclose(a_, b_, rtol=0, atol=1e-14 * np.abs(b_).max())


def test_progress_reports_every_segment():
    """tdhf_set_progress: one report per checkpoint segment, forward then
    backward, counted 1..total; the gradient is the same with reports on."""
    ng, nk, nb, a, z, nt, dt = 32, 4, 2, 8.0, 4.0, 29, 0.025
    S, O, vion, u0, A0 = setup_case(ng, nk, nb, a, z, nt, dt)
    As = np.stack([A0, 0.5 * A0])
    state = 2 * ng * nb * nk * 16
    with T.TDHF(ng, nk, nb, nb, a, dt, vion) as G:
        G.fwd_batch([u0, u0], As, objective="jsq", mem_budget=30 * state)
        g0, _ = G.bwd_batch(want_u0=False)
        seen = []
        G.set_progress(lambda phase, done, total: seen.append((phase, done, total)))
        G.fwd_batch([u0, u0], As, objective="jsq", mem_budget=30 * state)
        full, seg = G.adjoint_info()
        g1, _ = G.bwd_batch(want_u0=False)
        G.set_progress(None)
        G.fwd_batch([u0, u0], As, objective="jsq", mem_budget=30 * state)
        G.bwd_batch(want_u0=False)
    nseg = (nt + seg - 1) // seg
    assert not full and nseg > 1
    assert seen == [(0, i, nseg) for i in range(1, nseg + 1)] + \
        [(1, i, nseg) for i in range(1, nseg + 1)]
    np.testing.assert_array_equal(g1, g0)
