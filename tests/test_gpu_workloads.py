"""GPU parity at the benchmark workload shapes, through the C-ABI, against the
CPU oracle (and, for the adjoint, against central finite differences).

  ng64_nk64     plain      a=10, NG=64, NK=64, NB=NBocc=4, Z=8 (8-wave launches)
  mixed_batch   batched    NG=64, NB=4, NK in {16, 32, 64} in one batch, seeded
                           pulses (A0 log-uniform [0.01, 0.5], w uniform
                           [0.03, 0.10] Ha, CEP uniform), every launch shape
  adjoint_ng32  adjoint    a=8, NG=32, NK=32, NB=2, Z=4, 4000 steps dt=0.025,
                           pulse = 200 piecewise-linear knots ~ N(0, 0.05^2) x sin^2
  ng128_nk128   at scale   a=12, NG=128, NK=128, NB=NBocc=8, Z=16, objective
                           dt sum J^2

Tolerances: operators max|d| / max|ref| < 1e-12; trajectories |dE| < 1e-10 Ha
and max_t|dJ| / max_t|J| < 1e-8; batched trajectory = its single run to
1e-12; adjoint vs oracle 1e-9 of the gradient's max, vs central differences
(h = 1e-5) 1e-6 relative plus the difference quotient's rounding floor
(4 eps |E| / h for an objective whose terms are of size |E|).

The time steps are the stable ones of each grid (0.01 au at NG=64, a=10 and
0.004 au at NG=128, a=12; the specified 0.02 au exceeds RK4's stability bound
there).  Prefixes start mid-pulse, where the field and the current are large.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from mattohf_amd import tdhf as T

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -52


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def random_orbitals(S, seed):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    return S.gram_schmidt(np.asfortranarray(u))


def seeded_pulses(ntraj, nt, dt, T_, seed):
    """mixed_batch pulse set: A0 log-uniform [0.01, 0.5], w uniform [0.03, 0.10] Ha,
    CEP uniform [0, 2 pi], A(t) = A0 sin^2(pi t/T) sin(w t + CEP)."""
    rng = np.random.default_rng(seed)
    A0 = np.exp(rng.uniform(math.log(0.01), math.log(0.5), ntraj))
    om = rng.uniform(0.03, 0.10, ntraj)
    cep = rng.uniform(0.0, 2 * math.pi, ntraj)
    A, Aave = np.empty((ntraj, nt)), np.empty((ntraj, nt))
    for t in range(ntraj):
        oc, tp, e0, ph, ne = T.pulse_lp_params(A0[t], om[t], T_)
        _, A[t], Aave[t], _ = T.define_afield_lp(nt, dt, oc, tp, e0, ph + cep[t] / (2 * math.pi), ne)
    return A, Aave


# --------------------------------------------------------------------------
# ng64_nk64: a=10, NG=64, NK=64, NB=NBocc=4, Z=8
# --------------------------------------------------------------------------
C2 = dict(ng=64, nk=64, nb=4, a=10.0, z=8.0, dt=0.01)


@pytest.fixture(scope="module")
def ng64_nk64():
    S = po.System(C2["ng"], C2["nk"], C2["nb"], C2["nb"], C2["a"])
    vion = S.ion_potential([C2["z"]], [1.0], [0.0])
    O = po.Oracle(S, vion)
    G = T.TDHF(C2["ng"], C2["nk"], C2["nb"], C2["nb"], C2["a"], C2["dt"], vion)
    yield S, O, G, vion
    G.close()


def test_ng64_nk64_tables(ng64_nk64):
    S, O, G, _ = ng64_nk64
    vcg, vq = G.tables()
    assert rel(vcg, S.vCG()) < 1e-12
    full, nk = S.vCGkk(), S.nk
    for d in range(-(nk - 1), nk):
        ikr, ikc = (d, 0) if d >= 0 else (0, -d)
        assert rel(vq[d + nk - 1], full[ikr * nk + ikc]) < 1e-12


def test_ng64_nk64_hpsi_and_observables(ng64_nk64):
    S, O, G, _ = ng64_nk64
    u, w = random_orbitals(S, 201), random_orbitals(S, 202)
    for A in (0.0, 0.3):
        assert rel(G.ubkwbk_hwbk(u, w, A), O.ubkwbk_hwbk(u, w, A)) < 1e-12
        e, ec, n, j = G.observables(u, A)
        assert rel(ec, O.ubk_Ecomps(u, A)) < 1e-12
        assert abs(e - O.ubk_Etot(u, A)) < 1e-10
        assert abs(n - O.nele(u)) < 1e-12
        assert abs(j - O.ubk_Je(u, A)) < 1e-12 * max(1.0, abs(j))


def test_ng64_nk64_rk4_prefix_from_ground_state(ng64_nk64):
    """16 steps of the benchmark pulse (A0=0.3, w=0.057, T=400) from step
    20000 (mid-pulse), starting from a GPU ground state."""
    S, O, G, _ = ng64_nk64
    u0 = G.Gram_Schmidt(T.init_wf(C2["ng"], C2["nb"], C2["nk"], C2["a"]))
    ugs = G.GS_calculation(u0, 20, 4)[0]
    _, A, Aave, _ = T.define_afield_lp(20016, C2["dt"], *T.pulse_lp_params(0.3, 0.057, 400.0))
    A, Aave = A[20000:], Aave[20000:]
    ug, eg, ng_, jg, i3g = G.RT_calculation(ugs, A, Aave)
    uo, eo, no, jo, i3o = O.RT_calculation(ugs, C2["dt"], A, Aave)
    assert np.abs(eg - eo).max() < 1e-10
    assert np.abs(ng_ - no).max() < 1e-10
    assert rel(jg, jo) < 1e-8
    assert rel(ug, uo) < 1e-10
    assert np.abs(i3g - i3o).max() < 1e-10
    # the same steps advanced in pieces on the resident state and field
    G.upload(ugs)
    G.rt_begin(A, Aave)
    G.rt_advance(5)
    G.rt_advance(11)
    e2, n2, j2, i32 = G.rt_end()
    np.testing.assert_array_equal(e2, eg)
    np.testing.assert_array_equal(j2, jg)
    np.testing.assert_array_equal(G.download(), ug)


def test_ng64_nk64_advance_rejects_steps_beyond_the_field(ng64_nk64):
    S, O, G, _ = ng64_nk64
    A = np.full(3, 0.01)
    G.upload(random_orbitals(S, 203))
    G.rt_begin(A, A)
    G.rt_advance(2)
    with pytest.raises(T.TDHFError):
        G.rt_advance(2)
    e, n, j, _ = G.rt_end()
    assert e.size == 2 and np.isfinite(e).all()
    with pytest.raises(T.TDHFError):
        G.rt_advance(1)  # the run is closed


# --------------------------------------------------------------------------
# mixed_batch: batch of NK in {16, 32, 64}, NG=64, NB=4
# --------------------------------------------------------------------------
MIXED = dict(ng=64, nb=4, a=10.0, z=8.0, dt=0.01, nt=30, nks=[16, 32, 64, 64, 32, 16])


@pytest.fixture(scope="module")
def mixed_batch():
    us, vions, systems = [], [], []
    for t, nk in enumerate(MIXED["nks"]):
        S = po.System(MIXED["ng"], nk, MIXED["nb"], MIXED["nb"], MIXED["a"])
        systems.append(S)
        us.append(random_orbitals(S, 300 + t))
        vions.append(S.ion_potential([MIXED["z"]], [1.0], [0.0]))
    A, Aave = seeded_pulses(len(MIXED["nks"]), 20000 + MIXED["nt"], MIXED["dt"], 400.0, 20261202)
    A, Aave = np.ascontiguousarray(A[:, 20000:]), np.ascontiguousarray(Aave[:, 20000:])
    singles = []
    for t, nk in enumerate(MIXED["nks"]):
        with T.TDHF(MIXED["ng"], nk, MIXED["nb"], MIXED["nb"], MIXED["a"], MIXED["dt"], vions[t]) as G:
            singles.append(G.RT_calculation(us[t], A[t], Aave[t]))
    return systems, us, np.array(vions), A, Aave, singles


@pytest.mark.parametrize("nw", [1, 2, 4, 8])
def test_mixed_batch_batch_equals_single_runs(mixed_batch, nw):
    systems, us, vions, A, Aave, singles = mixed_batch
    with T.Batch(MIXED["ng"], MIXED["nb"], MIXED["nb"], MIXED["a"], MIXED["dt"], MIXED["nks"], vions) as B:
        B.set_waves(nw)
        B.upload(us)
        e, n, j, i3 = B.propagate(A, Aave)
        finals = B.download()
    for t in range(len(MIXED["nks"])):
        u1, e1, n1, j1, i31 = singles[t]
        np.testing.assert_allclose(e[t], e1, rtol=0, atol=1e-12 * np.abs(e1).max())
        np.testing.assert_allclose(n[t], n1, rtol=0, atol=1e-12)
        np.testing.assert_allclose(j[t], j1, rtol=0, atol=1e-12 * max(1.0, np.abs(j1).max()))
        np.testing.assert_allclose(i3[t], i31, rtol=0, atol=1e-12 * np.abs(i31).max())
        np.testing.assert_allclose(finals[t], u1, rtol=0, atol=1e-12)


@pytest.mark.parametrize("sched", [1, 2])
def test_mixed_batch_batch_row_schedules_equal_single_runs(mixed_batch, sched):
    systems, us, vions, A, Aave, singles = mixed_batch
    with T.Batch(MIXED["ng"], MIXED["nb"], MIXED["nb"], MIXED["a"], MIXED["dt"], MIXED["nks"], vions) as B:
        B.set_schedule(sched)
        assert B.schedule() == sched
        B.upload(us)
        e, n, j, i3 = B.propagate(A, Aave)
        finals = B.download()
    for t in range(len(MIXED["nks"])):
        u1, e1, n1, j1, i31 = singles[t]
        np.testing.assert_allclose(e[t], e1, rtol=0, atol=1e-12 * np.abs(e1).max())
        np.testing.assert_allclose(j[t], j1, rtol=0, atol=1e-12 * max(1.0, np.abs(j1).max()))
        np.testing.assert_allclose(finals[t], u1, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode,nw,sched", [(1, 1, 0), (1, 2, 1), (1, 8, 2), (2, 1, 2),
                                           (3, 1, 0), (3, 4, 2)])
def test_mixed_batch_pair_tiles_equal_single_runs(mixed_batch, mode, nw, sched):
    """The exchange from pair tiles (one transform pair per unordered pair of
    rows, Nyquist correction for the reversed pair) against the row-wise
    single runs, for NK = 16, 32, 64, with tiles of 64, 32 and 16 rows (modes
    1, 2, 3), the epilogue launch at 1 to 8 waves per workgroup, and tile
    lists pinned to XCDs per trajectory (row schedules 0, 1) or not (2)."""
    systems, us, vions, A, Aave, singles = mixed_batch
    with T.Batch(MIXED["ng"], MIXED["nb"], MIXED["nb"], MIXED["a"], MIXED["dt"], MIXED["nks"], vions) as B:
        B.set_waves(nw)
        B.set_exchange(mode)
        B.set_schedule(sched)
        assert B.exchange() == mode and B.schedule() == sched
        B.upload(us)
        e, n, j, i3 = B.propagate(A, Aave)
        finals = B.download()
    for t in range(len(MIXED["nks"])):
        u1, e1, n1, j1, i31 = singles[t]
        np.testing.assert_allclose(e[t], e1, rtol=0, atol=1e-12 * np.abs(e1).max())
        np.testing.assert_allclose(n[t], n1, rtol=0, atol=1e-12)
        np.testing.assert_allclose(j[t], j1, rtol=0, atol=1e-12 * max(1.0, np.abs(j1).max()))
        np.testing.assert_allclose(i3[t], i31, rtol=0, atol=1e-12 * np.abs(i31).max())
        np.testing.assert_allclose(finals[t], u1, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_pair_tiles_partial_blocks_match_oracle(mode):
    """Row counts that are not multiples of the row block (NK = 10: 40 rows;
    NK = 20: 80 rows; NK = 3: 12 rows), against the oracle, for each tile
    size; the same batch object also switches between tile sizes."""
    ng, nb, a, dt, nt, nks = 64, 4, 10.0, 0.01, 8, [10, 20, 3]
    us, vions, systems = [], [], []
    for t, nk in enumerate(nks):
        S = po.System(ng, nk, nb, nb, a)
        systems.append(S)
        us.append(random_orbitals(S, 600 + t))
        vions.append(S.ion_potential([8.0], [1.0], [0.0]))
    A, Aave = seeded_pulses(len(nks), 20000 + nt, dt, 400.0, 20261206)
    A, Aave = np.ascontiguousarray(A[:, 20000:]), np.ascontiguousarray(Aave[:, 20000:])
    with T.Batch(ng, nb, nb, a, dt, nks, np.array(vions)) as B:
        B.set_exchange(1 + mode % 3)  # another tile size first: buffers are rebuilt
        B.upload(us)
        B.propagate(A[:, :2], Aave[:, :2])
        B.set_exchange(mode)
        B.upload(us)
        e, n, j, _ = B.propagate(A, Aave)
        finals = B.download()
    for t in range(len(nks)):
        O = po.Oracle(systems[t], vions[t])
        uo, eo, no, jo, _ = O.RT_calculation(us[t], dt, A[t], Aave[t])
        assert np.abs(e[t] - eo).max() < 1e-10
        assert np.abs(n[t] - no).max() < 1e-10
        assert rel(j[t], jo) < 1e-8
        assert rel(finals[t], uo) < 1e-10


def test_pair_tiles_need_full_occupation():
    S = po.System(64, 4, 3, 2, 10.0)
    with T.Batch(64, 3, 2, 10.0, 0.01, [4], np.array([np.zeros(64)])) as B:
        with pytest.raises(T.TDHFError):
            B.set_exchange(1)
        assert B.exchange() == 0


def test_mixed_batch_batch_matches_oracle(mixed_batch):
    """One trajectory of each k-mesh of the batch against the oracle, over an
    8-step prefix."""
    systems, us, vions, A, Aave, _ = mixed_batch
    with T.Batch(MIXED["ng"], MIXED["nb"], MIXED["nb"], MIXED["a"], MIXED["dt"], MIXED["nks"], vions) as B:
        B.upload(us)
        e, n, j, _ = B.propagate(A[:, :8], Aave[:, :8])
    for t in (0, 1, 2):  # NK = 16, 32, 64
        O = po.Oracle(systems[t], vions[t])
        _, eo, no, jo, _ = O.RT_calculation(us[t], MIXED["dt"], A[t, :8], Aave[t, :8])
        assert np.abs(e[t] - eo).max() < 1e-10
        assert np.abs(n[t] - no).max() < 1e-10
        assert rel(j[t], jo) < 1e-8


# --------------------------------------------------------------------------
# adjoint_ng32: adjoint, a=8, NG=32, NK=32, NB=2, Z=4, 4000 steps dt=0.025
# --------------------------------------------------------------------------
C4 = dict(ng=32, nk=32, nb=2, a=8.0, z=4.0, dt=0.025, nt=4000, nknots=200, sigma=0.05)


@pytest.fixture(scope="module")
def adjoint_ng32():
    S = po.System(C4["ng"], C4["nk"], C4["nb"], C4["nb"], C4["a"])
    vion = S.ion_potential([C4["z"]], [1.0], [0.0])
    G = T.TDHF(C4["ng"], C4["nk"], C4["nb"], C4["nb"], C4["a"], C4["dt"], vion)
    u0 = G.Gram_Schmidt(T.init_wf(C4["ng"], C4["nb"], C4["nk"], C4["a"]))
    ugs, eh, _, _ = G.GS_calculation(u0, 200, 4, tol=1e-13)
    rng = np.random.default_rng(4)
    P = T.knot_matrix(C4["nt"], C4["dt"], C4["nt"] * C4["dt"], C4["nknots"])
    knots = C4["sigma"] * rng.standard_normal(C4["nknots"])
    yield S, G, vion, ugs, eh[-1], P, knots, rng
    G.close()


def test_adjoint_ng32_adjoint_matches_oracle_mid_pulse(adjoint_ng32):
    """The gradient of E(T) - E_GS over 40 steps from the state at step 2000
    of the pulse, against the oracle's adjoint of the same steps."""
    S, G, vion, ugs, egs, P, knots, _ = adjoint_ng32
    A = P @ knots
    Aave = 0.5 * (A + np.concatenate([[0.0], A[:-1]]))
    G.upload(ugs)
    G.rt_begin(A[:2000], Aave[:2000])
    G.rt_advance(2000)
    G.rt_end()
    umid = G.download()
    # the adjoint pair takes A only and derives Aave with A(-1) = 0, as the
    # oracle's adjoint does
    Aseg = A[2000:2040].copy()
    L, _, _, _, _ = G.fwd(umid, Aseg, eref=egs)
    g, gu = G.bwd()
    O = po.Oracle(S, vion)
    Lo, go, guo = O.adjoint_energy(umid, C4["dt"], Aseg, egs)
    assert abs(L - Lo) < 1e-11 * max(1.0, abs(Lo))
    assert np.abs(g - go).max() < 1e-9 * np.abs(go).max()
    assert np.abs(gu - guo).max() < 1e-9 * np.abs(guo).max()


def test_adjoint_ng32_full_run_finite_differences(adjoint_ng32):
    """The full 4000-step gradient along 8 random knot directions against
    central differences of the forward objective (h = 1e-5)."""
    S, G, vion, ugs, egs, P, knots, rng = adjoint_ng32
    L, _, e, _, _ = G.fwd(ugs, P @ knots, eref=egs)
    gA, _ = G.bwd(want_u0=False)
    gk = P.T @ gA
    h = 1e-5
    floor = 4 * EPS * abs(egs) / h
    worst = 0.0
    for _ in range(8):
        d = rng.standard_normal(C4["nknots"])
        d /= np.linalg.norm(d)
        lp = G.fwd(ugs, P @ (knots + h * d), eref=egs, want_obs=False)[0]
        lm = G.fwd(ugs, P @ (knots - h * d), eref=egs, want_obs=False)[0]
        fd = (lp - lm) / (2 * h)
        worst = max(worst, abs(fd - gk @ d) / (1e-6 * abs(fd) + floor))
        assert abs(fd - gk @ d) <= 1e-6 * abs(fd) + floor, (fd, gk @ d, floor)
    assert np.isfinite(e).all() and L > 0  # the pulse excites the system
    print(f"adjoint_ng32 FD: worst |fd - g.d| / (1e-6 |fd| + floor) = {worst:.3f}, floor {floor:.2e}")


# --------------------------------------------------------------------------
# ng128_nk128: a=12, NG=128, NK=128, NB=NBocc=8, Z=16
# --------------------------------------------------------------------------
C5 = dict(ng=128, nk=128, nb=8, a=12.0, z=16.0, dt=0.004)


@pytest.fixture(scope="module")
def ng128_nk128():
    S = po.System(C5["ng"], C5["nk"], C5["nb"], C5["nb"], C5["a"])
    vion = S.ion_potential([C5["z"]], [1.0], [0.0])
    G = T.TDHF(C5["ng"], C5["nk"], C5["nb"], C5["nb"], C5["a"], C5["dt"], vion)
    yield S, G, vion
    G.close()


def test_ng128_nk128_tables_and_hpsi(ng128_nk128):
    S, G, vion = ng128_nk128
    vcg, vq = G.tables()
    assert rel(vcg, S.vCG()) < 1e-12
    full, nk = S.vCGkk(), S.nk
    for d in range(-(nk - 1), nk):
        ikr, ikc = (d, 0) if d >= 0 else (0, -d)
        assert rel(vq[d + nk - 1], full[ikr * nk + ikc]) < 1e-12
    O = po.Oracle(S, vion, full)
    u, w = random_orbitals(S, 501), random_orbitals(S, 502)
    assert rel(G.ubkwbk_hwbk(u, w, 0.2), O.ubkwbk_hwbk(u, w, 0.2)) < 1e-12
    e, ec, n, j = G.observables(u, -0.1)
    assert rel(ec, O.ubk_Ecomps(u, -0.1)) < 1e-12
    assert abs(n - O.nele(u)) < 1e-12
    assert abs(j - O.ubk_Je(u, -0.1)) < 1e-12 * max(1.0, abs(j))


def test_ng128_nk128_jsq_gradient_finite_differences(ng128_nk128):
    """Three mixed_batch-style pulses in one launch, 4 steps from mid-pulse,
    objective dt sum J^2: each pulse's gradient against central differences
    along a random direction (one pair of batched forwards)."""
    S, G, vion = ng128_nk128
    u0 = random_orbitals(S, 503)
    A, _ = seeded_pulses(3, 50004, C5["dt"], 400.0, 20261205)
    A = np.ascontiguousarray(A[:, 50000:])
    us = [u0] * 3
    L, _, _, _, j = G.fwd_batch(us, A, objective="jsq")
    g, _ = G.bwd_batch(want_u0=False)
    rng = np.random.default_rng(7)
    d = rng.standard_normal(A.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    h = 1e-5
    lp = G.fwd_batch(us, A + h * d, objective="jsq", want_obs=False)[0]
    lm = G.fwd_batch(us, A - h * d, objective="jsq", want_obs=False)[0]
    fd = (lp - lm) / (2 * h)
    ad = np.einsum("tn,tn->t", g, d)
    floor = 4 * EPS * np.abs(L) / h
    assert np.all(np.abs(fd - ad) <= 1e-6 * np.abs(fd) + floor), (fd, ad)
    np.testing.assert_allclose(L, C5["dt"] * np.sum(j ** 2, axis=1), rtol=1e-13)
