"""Pins the CPU oracle (oracle/tdhf_oracle.c) -- CPU only, no GPU.

The reference ships no tests, fixtures or golden vectors (its tree has no test
files), and its Fortran build needs FFTW3's fftw3.f include file, which this
image does not have (MPI and MKL's FFTW3/LAPACK symbols are present; a
stand-in header would not make an acceptable oracle), so it is not run here.  Each oracle
function is therefore checked against something independent of its code:

  orc_fft            numpy.fft (FFTW sign / normalisation) incl. a non-2^n size
  orc_grids          numpy.fft.fftfreq (Nyquist at -NG/2), midpoint k mesh
  orc_k_vck          scipy.integrate.quad of the same integral + Euler-Maclaurin
                     bound; q = 0 against 2*asinh(L); loop count in [20 NG NK,
                     20 NG NK + 1]; per-pair vs per-difference table at small NK
  orc_ion_potential  max = 0, mirror symmetry, linearity in the ions
  orc_init_wf        the LCG's first value by hand
  orc_gram_schmidt   numpy QR (positive-diagonal R) in the dvcell metric
  orc_dns            N = 2*NBocc for orthonormal orbitals
  orc_vh             cosine density -> vCG(G1) cos(G1 x)
  orc_tpsi           plane wave -> (G+k+A)^2/2
  orc_vfpsi          dense numpy formulation (explicit DFT matrices, no FFT);
                     plane-wave orbitals closed form (pins the k-q sign);
                     single-band NK=1 self-interaction identity; hermiticity
  orc_ecomps         HF identity sum(occ*eps) = Ekin + Eion + 2 EH + 2 Ex;
                     Ekin by Parseval
  orc_je             diamagnetic identity J(A) = J(0) - A N / a; J(GS) = 0
  orc_rk4_step       4th-order convergence (A=0 energy drift ratio ~ 2^4);
                     the stage fields (2 Aave - A, Aave, Aave, A): with
                     plane-wave orbitals and no ions every stage state stays a
                     plane wave per (k, band), so one step is a scalar
                     recursion over the amplitudes with the closed-form
                     kinetic and exchange eigenvalues
  orc_epsbk          band energy Re<u|H u> dv and residual sqrt(sum |r| dv)
                     (sum of moduli, not of squares) from H|u> on a
                     non-eigenstate
  orc_heev           numpy.linalg.eigh
  orc_ks_krylov      Ritz subspace invariant under rotation of the input bands
  orc_gs_run         converged residuals; eigenvalues independent of NCG
  orc_define_afield  closed-form sin^2 pulse
  orc_adjoint_*      central finite differences of the oracle's own forward
                     objective (energy at the end and dt*sum J^2), in the
                     field and in the initial orbitals
  orc_gs_run_diag    stepwise recomposition from one-iteration GS runs,
                     ubk_dns and ubk_epsbk
  orc_make_dos       Gaussian normalisation integrals, single-level peak,
                     energy-grid ends
  orc_fourier        numpy FFT on the transform's own frequency grid, both
                     filters
  orc_te_step        4th-order Taylor polynomial of dense H[u] matrices built
                     column by column (TE, TEPC, TEPC2)
"""
import math

import numpy as np
import pytest
from scipy import integrate

import pyoracle as po


def sysm(ng=32, nk=4, nb=2, nbocc=2, a=8.0):
    return po.System(ng, nk, nb, nbocc, a)


def make(ng=32, nk=4, nb=2, nbocc=2, a=8.0, z=4.0):
    S = sysm(ng, nk, nb, nbocc, a)
    vion = S.ion_potential([z], [1.0], [0.0])
    return S, po.Oracle(S, vion), vion


def rand_orb(S, seed):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    return S.gram_schmidt(np.asfortranarray(u))


# ----------------------------------------------------------------------------
def test_fft_matches_numpy():
    rng = np.random.default_rng(0)
    for n in (2, 4, 8, 16, 32, 64, 128, 12):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        np.testing.assert_allclose(po.fft(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * n)
        np.testing.assert_allclose(po.fft(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * n)


def test_grids():
    S = sysm(16, 5, 1, 1, 3.6)
    rx, gx, kx = S.grids()
    b = 2 * math.pi / 3.6
    np.testing.assert_allclose(rx, np.arange(16) * 3.6 / 16, rtol=1e-15)
    np.testing.assert_allclose(gx, np.fft.fftfreq(16) * 16 * b, rtol=1e-14, atol=1e-14)
    assert gx[8] < 0  # Nyquist at -NG/2 only
    np.testing.assert_allclose(kx, (-0.5 + (np.arange(5) + 0.5) / 5) * b, rtol=1e-14)


@pytest.mark.parametrize("ng,nk,a", [(32, 8, 8.0), (16, 3, 3.6), (64, 2, 10.0)])
def test_k_vck_against_quadrature(ng, nk, a):
    S = sysm(ng, nk, 1, 1, a)
    M = S.kvck_nsteps()
    assert M in (20 * ng * nk, 20 * ng * nk + 1)
    dx = a / ng / 20.0
    Lp = M * dx
    fp = lambda x, q: (-q * math.sin(q * x) * (x * x + 1.0) - x * math.cos(q * x)) / (x * x + 1.0) ** 1.5
    for q in (0.0, 0.3, 1.7, 5.0, 12.0):
        # QUADPACK's cos-weighted rule (QAWO) for the oscillatory integrand
        exact = 2.0 * integrate.quad(lambda x: 1.0 / math.sqrt(x * x + 1.0), 0.0, Lp,
                                     weight="cos", wvar=q, limit=2000, epsabs=1e-12,
                                     epsrel=1e-13)[0]
        # trapezoid = integral + dx^2/12 (f'(L) - f'(0)) + O(dx^4), times 2
        em = 2.0 * dx * dx / 12.0 * (fp(Lp, q) - fp(0.0, q))
        got = S.k_vCk(q)
        tol = 2.0 * dx ** 4 / 720.0 * (q ** 4 + 10.0) * Lp + 1e-11
        assert abs(got - (exact + em)) < tol, (q, got, exact + em)


def test_k_vck_q0_closed_form():
    S = sysm(32, 8, 1, 1, 8.0)
    M = S.kvck_nsteps()
    dx = 8.0 / 32 / 20.0
    Lp = M * dx
    em = 2.0 * dx * dx / 12.0 * (-Lp / (Lp * Lp + 1.0) ** 1.5)
    assert abs(S.k_vCk(0.0) - (2.0 * math.asinh(Lp) + em)) < 1e-10


def test_vcgkk_per_difference_equals_per_pair():
    S = sysm(16, 3, 1, 1, 3.6)
    full = S.vCGkk()
    _, gx, kx = S.grids()
    for ikr in range(3):
        for ikc in range(3):
            lit = np.array([S.k_vCk(gx[g] + kx[ikr] - kx[ikc]) for g in range(16)])
            np.testing.assert_allclose(full[ikr * 3 + ikc], lit, rtol=1e-13, atol=1e-14)
    vcg = S.vCG()
    assert vcg[0] == 0.0
    np.testing.assert_allclose(vcg[1:], [S.k_vCk(g) for g in gx[1:]], rtol=1e-15)


def test_ion_potential_properties():
    S = sysm(32, 4, 1, 1, 8.0)
    v = S.ion_potential([4.0], [1.0], [0.0])
    assert v.max() == 0.0
    # mirror symmetry about x=0 holds up to the truncated image window |n| <= 4 NK
    edge = 4.0 * 8.0 / (4 * 4 * 8.0) ** 2
    np.testing.assert_allclose(v[1:], v[1:][::-1], rtol=0, atol=2 * edge)
    # the window itself: images n = -4NK..4NK, softened -Z/sqrt(d^2+alpha^2)
    rx = np.arange(32) / 32 * 8.0
    n = np.arange(-16, 17)
    d = rx[:, None] - 8.0 * n[None, :]
    ref = (-4.0 / np.sqrt(d * d + 1.0)).sum(axis=1)
    np.testing.assert_allclose(v, ref - ref.max(), rtol=0, atol=1e-12)
    v1 = S.ion_potential([1.0], [1.0], [0.25])
    v2 = S.ion_potential([2.0], [0.7], [0.6])
    v12 = S.ion_potential([1.0, 2.0], [1.0, 0.7], [0.25, 0.6])
    d = v12 - (v1 + v2)
    assert np.ptp(d) < 1e-12  # linear in the ions up to the constant shift


def test_init_wf_first_orbital():
    S = sysm(32, 2, 2, 2, 8.0)
    u = S.init_wf()
    rnd = ((123 * 106 + 1283) % 6075) / 6075.0
    rx = np.arange(32) / 32 * 8.0
    np.testing.assert_allclose(u[:, 0, 0].real, np.exp(-0.5 * (rx - rnd * 8.0) ** 2), rtol=1e-15)
    assert np.all(u.imag == 0)


def test_gram_schmidt_matches_qr():
    S = sysm(32, 3, 4, 4, 8.0)
    rng = np.random.default_rng(1)
    u = np.asfortranarray(rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape))
    g = S.gram_schmidt(u)
    dv = 8.0 / 32
    for ik in range(3):
        q, r = np.linalg.qr(u[:, :, ik])
        ph = np.diag(r) / np.abs(np.diag(r))
        q = q * ph[None, :] / math.sqrt(dv)
        np.testing.assert_allclose(g[:, :, ik], q, rtol=0, atol=1e-12)
        np.testing.assert_allclose(g[:, :, ik].conj().T @ g[:, :, ik] * dv, np.eye(4), atol=1e-13)


def test_density_counts_electrons():
    S, O, _ = make(32, 4, 3, 2)
    u = rand_orb(S, 2)
    assert abs(O.ubk_dns(u).sum() * 8.0 / 32 - 4.0) < 1e-12


def test_hartree_of_cosine_density():
    S, O, _ = make(32, 4, 2, 2)
    rx, gx, _ = S.grids()
    vcg = S.vCG()
    for g1 in (1, 3, 7):
        v = O.dns_vH(np.cos(gx[g1] * rx))
        np.testing.assert_allclose(v, vcg[g1] * np.cos(gx[g1] * rx), rtol=0, atol=1e-13)


def test_kinetic_of_plane_wave():
    S, O, _ = make(32, 4, 1, 1)
    rx, gx, kx = S.grids()
    w = np.zeros(S.shape, dtype=complex, order="F")
    gsel = [2, 30, 16, 5]
    for ik in range(4):
        w[:, 0, ik] = np.exp(1j * gx[gsel[ik]] * rx)
    A = 0.31
    t = O.ubk_tubk(w, A)
    for ik in range(4):
        lam = 0.5 * (gx[gsel[ik]] + kx[ik] + A) ** 2
        np.testing.assert_allclose(t[:, 0, ik], lam * w[:, 0, ik], rtol=0, atol=1e-11)


def dense_exchange(S, vcgkk, u, w):
    """V_F w with explicit DFT matrices (no FFT): the same sum written as
    sum_q sum_j u_qj(x) * (1/NG) sum_G v(G+k-q) e^{iGx} sum_x' e^{-iGx'} rho(x')."""
    ng, nk, nb, nbo = S.ng, S.nk, S.nb, S.nbocc
    x = np.arange(ng)
    F = np.exp(-2j * np.pi * np.outer(x, x) / ng)
    Fi = np.conj(F)
    out = np.zeros(S.shape, dtype=complex, order="F")
    for ik in range(nk):
        for iq in range(nk):
            v = vcgkk[ik * nk + iq]
            K = (Fi @ np.diag(-v) @ F) / (ng * nk)
            for ib in range(nb):
                for jb in range(nbo):
                    rho = w[:, ib, ik] * np.conj(u[:, jb, iq])
                    out[:, ib, ik] += (K @ rho) * u[:, jb, iq]
    return out


def test_exchange_dense_formulation():
    S, O, _ = make(16, 3, 3, 2, 5.0)
    u, w = rand_orb(S, 3), rand_orb(S, 4)
    ref = dense_exchange(S, S.vCGkk(), u, w)
    np.testing.assert_allclose(O.ubkwbk_vFwbk(u, w), ref, rtol=0, atol=1e-12)


def test_exchange_plane_wave_closed_form():
    ng, nk, a = 16, 4, 5.0
    S = po.System(ng, nk, 2, 2, a)
    vion = np.zeros(ng)
    O = po.Oracle(S, vion)
    rx, gx, kx = S.grids()
    vk = S.vCGkk()
    gi = np.array([[1, 14], [3, 0], [15, 6], [8, 2]])  # FFT index of each (k, band)
    u = np.zeros(S.shape, dtype=complex, order="F")
    for ik in range(nk):
        for ib in range(2):
            u[:, ib, ik] = np.exp(2j * np.pi * gi[ik, ib] * np.arange(ng) / ng) / math.sqrt(a)
    vf = O.ubkwbk_vFwbk(u, u)
    for ik in range(nk):
        for ib in range(2):
            lam = 0.0
            for iq in range(nk):
                for jb in range(2):
                    g = (gi[ik, ib] - gi[iq, jb]) % ng  # pair density lives at this G index
                    lam += vk[ik * nk + iq][g]
            lam = -lam / (nk * a)
            np.testing.assert_allclose(vf[:, ib, ik], lam * u[:, ib, ik], rtol=0, atol=1e-13)


def test_exchange_self_interaction_nk1():
    S = po.System(32, 1, 1, 1, 8.0)
    O = po.Oracle(S, np.zeros(32))
    u = rand_orb(S, 5)
    vf = O.ubkwbk_vFwbk(u, u)
    vh = O.dns_vH(O.ubk_dns(u))
    v0 = S.vCGkk()[0][0]
    np.testing.assert_allclose(vf[:, 0, 0], (-0.5 * vh - v0 / 8.0) * u[:, 0, 0], rtol=0,
                               atol=1e-12)


def test_exchange_hermitian():
    S, O, _ = make(32, 3, 2, 2)
    u = rand_orb(S, 6)
    a, b = rand_orb(S, 7), rand_orb(S, 8)
    va, vb = O.ubkwbk_vFwbk(u, a), O.ubkwbk_vFwbk(u, b)
    for ik in range(3):
        for ib in range(2):
            lhs = np.vdot(a[:, ib, ik], vb[:, ib, ik])
            rhs = np.conj(np.vdot(b[:, ib, ik], va[:, ib, ik]))
            assert abs(lhs - rhs) < 1e-12


def test_hpsi_is_sum_of_terms():
    S, O, vion = make(32, 3, 2, 2)
    u, w = rand_orb(S, 9), rand_orb(S, 10)
    vh = O.dns_vH(O.ubk_dns(u))
    ref = O.ubk_tubk(w, 0.2) + (vh + vion)[:, None, None] * w + O.ubkwbk_vFwbk(u, w)
    np.testing.assert_allclose(O.ubkwbk_hwbk(u, w, 0.2), ref, rtol=0, atol=1e-12)


def test_energy_hf_identity_and_parseval():
    S, O, vion = make(32, 4, 2, 2)
    u = rand_orb(S, 11)
    ek, ei, eh, ex = O.ubk_Ecomps(u, 0.0)
    h = O.ubkwbk_hwbk(u, u, 0.0)
    dv = 8.0 / 32
    occ = 2.0 / 4
    s = sum(occ * np.vdot(u[:, ib, ik], h[:, ib, ik]).real * dv for ik in range(4)
            for ib in range(2))
    assert abs(s - (ek + ei + 2 * eh + 2 * ex)) < 1e-11
    _, gx, kx = S.grids()
    A = 0.13
    e2 = 0.0
    for ik in range(4):
        for ib in range(2):
            uh = np.fft.fft(u[:, ib, ik])
            e2 += occ * np.sum(0.5 * (gx + kx[ik] + A) ** 2 * np.abs(uh) ** 2) / 32 * dv
    assert abs(O.ubk_Ecomps(u, A)[0] - e2) < 1e-11


def test_current_diamagnetic_identity():
    S, O, _ = make(32, 4, 2, 2)
    u = rand_orb(S, 12)
    n = O.nele(u)
    for A in (0.1, -0.37):
        assert abs(O.ubk_Je(u, A) - (O.ubk_Je(u, 0.0) - A * n / 8.0)) < 1e-12


def test_heev_matches_numpy():
    rng = np.random.default_rng(13)
    for m in (1, 2, 5, 12):
        a = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
        a = a + a.conj().T
        w, v, rc = po.heev(a)
        assert rc == 0
        np.testing.assert_allclose(w, np.linalg.eigvalsh(a), rtol=0, atol=1e-12)
        np.testing.assert_allclose(a @ v, v * w, rtol=0, atol=1e-11)
        np.testing.assert_allclose(v.conj().T @ v, np.eye(m), atol=1e-13)


def test_krylov_step_invariant_under_band_rotation():
    S, O, _ = make(32, 3, 2, 2)
    u = rand_orb(S, 14)
    rng = np.random.default_rng(15)
    ur = u.copy(order="F")
    for ik in range(3):
        q, _ = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))
        ur[:, :, ik] = u[:, :, ik] @ q
    a, b = O.KS_krylov(u, 3), O.KS_krylov(ur, 3)
    np.testing.assert_allclose(O.ubk_dns(a), O.ubk_dns(b), rtol=0, atol=1e-10)


@pytest.fixture(scope="module")
def gs_ng32_nk8():
    S, O, vion = make(32, 8, 2, 2, 8.0, 4.0)
    u0 = S.gram_schmidt(S.init_wf())
    u, eh, eps, err = O.GS_calculation(u0, 250, 4)
    return S, O, u, eh, eps, err


def test_gs_converged(gs_ng32_nk8):
    S, O, u, eh, eps, err = gs_ng32_nk8
    assert abs(eh[-1] - eh[-2]) < 1e-12
    assert err.max() < 1e-5
    assert abs(O.ubk_Je(u, 0.0)) < 1e-10  # no current in the ground state


def test_gs_eigenvalues_independent_of_ncg(gs_ng32_nk8):
    S, O, u, eh, eps, err = gs_ng32_nk8
    u0 = S.gram_schmidt(S.init_wf())
    _, eh5, eps5, _ = O.GS_calculation(u0, 250, 5)
    assert abs(eh5[-1] - eh[-1]) < 1e-10
    np.testing.assert_allclose(eps5, eps, rtol=0, atol=1e-9)


def test_rk4_fourth_order(gs_ng32_nk8):
    S, O, u, *_ = gs_ng32_nk8
    up = u.copy(order="F")
    up[:, 0, :] *= np.exp(0.3j * np.sin(2 * np.pi * np.arange(32) / 32))[:, None]  # excite
    e0 = O.ubk_Etot(up, 0.0)
    drift = []
    for dt in (0.02, 0.01):
        n = int(round(0.4 / dt))
        z = np.zeros(n)
        _, e, _, _, _ = O.RT_calculation(up, dt, z, z)
        drift.append(abs(e[-1] - e0))
    assert drift[0] > 0 and 8.0 < drift[0] / drift[1] < 40.0


def test_rk4_stage_fields_plane_waves():
    """RT_RK4_ubk (src/RT_calculation.f90:154-173) evaluates its four stages at
    A(it-1) = 2 Aave - A, Aave, Aave and A.  Orbitals u_kb = c_kb e^{i G_kb x}/sqrt(a)
    with vion = 0 have a uniform density (v_H = 0 since vCG(0) = 0), and the
    exchange maps e^{i G_kb x} onto itself with the eigenvalue
    -(1/(NK a)) sum_qj |c_qj|^2 v(G_kb - G_qj + k - q), so every stage state is a
    plane wave per (k, b) and the step is a recursion over the amplitudes c_kb.
    Any other assignment of the stage fields moves the result by O(dt dA)."""
    ng, nk, nb, a = 16, 3, 2, 5.0
    S = po.System(ng, nk, nb, nb, a)
    O = po.Oracle(S, np.zeros(ng))
    _, gx, kx = S.grids()
    vk = S.vCGkk()
    gi = np.array([[1, 14], [8, 0], [15, 6]])  # FFT index of each (k, band); 8 = Nyquist
    rng = np.random.default_rng(21)
    c0 = (0.8 + 0.4 * rng.random((nk, nb))) * np.exp(2j * np.pi * rng.random((nk, nb)))
    x = np.arange(ng)
    u = np.zeros(S.shape, dtype=complex, order="F")
    for ik in range(nk):
        for ib in range(nb):
            u[:, ib, ik] = c0[ik, ib] * np.exp(2j * np.pi * gi[ik, ib] * x / ng) / math.sqrt(a)

    def h(c, A):
        out = np.empty_like(c)
        for ik in range(nk):
            for ib in range(nb):
                lam = 0.0
                for iq in range(nk):
                    for jb in range(nb):
                        lam += abs(c[iq, jb]) ** 2 * vk[ik * nk + iq][(gi[ik, ib] - gi[iq, jb]) % ng]
                out[ik, ib] = 0.5 * (gx[gi[ik, ib]] + kx[ik] + A) ** 2 - lam / (nk * a)
        return out

    def step(c, dt, A1, A2, A3, A4):
        k1 = -1j * h(c, A1) * c
        y = c + 0.5 * dt * k1
        k2 = -1j * h(y, A2) * y
        y = c + 0.5 * dt * k2
        k3 = -1j * h(y, A3) * y
        y = c + dt * k3
        k4 = -1j * h(y, A4) * y
        return c + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)

    dt, aave, A = 0.05, 0.1, 0.3
    got = O.RT_RK4_ubk(u, dt, aave, A)
    cg = np.array([[np.vdot(u[:, ib, ik] / c0[ik, ib], got[:, ib, ik]) * a / ng
                    for ib in range(nb)] for ik in range(nk)])
    want = step(c0, dt, 2 * aave - A, aave, aave, A)
    np.testing.assert_allclose(cg, want, rtol=0, atol=1e-13)
    # the state stays a plane wave per (k, b)
    for ik in range(nk):
        for ib in range(nb):
            pw = cg[ik, ib] * np.exp(2j * np.pi * gi[ik, ib] * x / ng) / math.sqrt(a)
            np.testing.assert_allclose(got[:, ib, ik], pw, rtol=0, atol=1e-13)
    # the recursion separates the conventions: each alternative is far off
    for alt in ((A, aave, aave, A), (2 * aave - A, A, A, A), (aave, aave, aave, aave),
                (2 * aave - A, aave, aave, aave)):
        assert np.abs(step(c0, dt, *alt) - want).max() > 1e-5


def test_epsbk_residual_is_sum_of_moduli():
    """ubk_epserrorbk (src/Functions.f90:340-358) takes sqrt(sum |r| dv) of the
    residual r = H u - eps u, with eps = Re<u|H u> dv (ubk_epsbk, :321-337):
    the sum of moduli, not of squared moduli.  On a non-eigenstate the two
    differ by orders of magnitude."""
    S, O, _ = make(32, 3, 3, 2)
    u = rand_orb(S, 22)
    eps, err = O.ubk_epsbk(u)
    hu = O.ubkwbk_hwbk(u, u, 0.0)
    dv = 8.0 / 32
    for ik in range(3):
        for ib in range(3):
            e = np.vdot(u[:, ib, ik], hu[:, ib, ik]).real * dv
            r = hu[:, ib, ik] - e * u[:, ib, ik]
            assert abs(eps[ib, ik] - e) < 1e-12 * max(1.0, abs(e))
            assert abs(err[ib, ik] - math.sqrt(np.abs(r).sum() * dv)) < 1e-12 * err[ib, ik]
            sq = math.sqrt((np.abs(r) ** 2).sum() * dv)
            assert abs(err[ib, ik] - sq) > 0.1 * err[ib, ik]


def test_afield_lp_closed_form():
    nt, dt, A0, w, T = 3000, 0.05, 0.2, 0.057, 100.0
    t, A, Aave, E = po.define_afield_lp(nt, dt, w * 27.2116, T * 0.02419, A0 * w * 514.2,
                                        w * T / (4 * math.pi), 2)
    tt = dt * np.arange(nt)
    ref = np.where(tt < T, A0 * np.sin(math.pi * tt / T) ** 2 * np.sin(w * tt), 0.0)
    np.testing.assert_allclose(A, ref, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Aave[1:], 0.5 * (A[1:] + A[:-1]), rtol=0, atol=0)
    np.testing.assert_allclose(E[1:], -(A[1:] - A[:-1]) / dt, rtol=1e-15, atol=0)


def _adjoint_case(seed=3):
    S = po.System(16, 2, 1, 1, 5.0)
    vion = S.ion_potential([2.0], [1.0], [0.0])
    O = po.Oracle(S, vion)
    u0 = S.gram_schmidt(S.init_wf())
    rng = np.random.default_rng(seed)
    A = 0.2 * rng.standard_normal(6)
    return S, O, u0, A, rng


@pytest.mark.parametrize("kind", ["energy", "jsq"])
def test_oracle_adjoint_finite_differences(kind):
    S, O, u0, A, rng = _adjoint_case()
    dt, h = 0.02, 1e-5
    if kind == "energy":
        L, g, gu = O.adjoint_energy(u0, dt, A, eref=0.0)
        f = lambda uu, AA: O.objective_energy(uu, dt, AA)
    else:
        L, g, gu = O.adjoint_jsq(u0, dt, A)
        f = lambda uu, AA: O.objective_jsq(uu, dt, AA)
    assert abs(L - f(u0, A)) < 1e-13 * max(1.0, abs(L))
    for _ in range(3):
        d = rng.standard_normal(A.size)
        fd = (f(u0, A + h * d) - f(u0, A - h * d)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 * abs(fd) + 1e-11, (fd, g @ d)
    du = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    fd = (f(u0 + h * du, A) - f(u0 - h * du, A)) / (2 * h)
    assert abs(fd - np.real(np.vdot(gu, du))) <= 1e-6 * abs(fd) + 1e-11


# ---- widening: GS diagnostics, Make_DoS, spectra, Taylor propagators ----

def test_gs_diag_matches_stepwise_recomposition():
    """orc_gs_run_diag's columns against the same quantities rebuilt from
    one-iteration GS runs, ubk_dns and ubk_epsbk (mixing 0.5 in numpy)."""
    S = po.System(16, 2, 3, 2, 5.0)
    vion = S.ion_potential([2.0], [1.0], [0.0])
    O = po.Oracle(S, vion)
    u0 = S.gram_schmidt(S.init_wf())
    u, eh, eps, err, dg = O.GS_calculation_diag(u0, 4, 3)
    uu = u0.copy()
    dmix = O.ubk_dns(uu)
    for it in range(4):
        uu = O.GS_calculation(uu, 1, 3)[0]
        e, r = O.ubk_epsbk(uu)
        dn = 0.5 * O.ubk_dns(uu) + 0.5 * dmix
        want = [np.abs(dn - dmix).sum() / 16, r.mean(), r.max(),
                e[2:, :].min() - e[:2, :].max(), (e[2, :] - e[1, :]).min()]
        dmix = dn
        np.testing.assert_allclose(dg[it], want, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(u, uu)


def test_make_dos_normalisation_and_peak():
    """Gaussians normalised to 1: sum(DoS) de = NB NK occmax, sum(DoSocc) de =
    NBocc NK occ; a single level peaks at its energy with height
    occ/(w sqrt(pi))."""
    rng = np.random.default_rng(2)
    eps = np.sort(rng.uniform(-2.0, 1.0, (3, 4)), axis=0)
    x, d, q = po.make_dos(eps, 2, 10000)
    de = x[1] - x[0]
    assert abs(d.sum() * de - 3 * 4 * 0.5) < 1e-9
    assert abs(q.sum() * de - 2 * 4 * 0.5) < 1e-9
    one = np.array([[0.25, 1.25]]).T  # NB=2, NK=1: range [0.25, 1.25]
    x, d, q = po.make_dos(one, 1, 1000)
    w = 5 * (x[1] - x[0])
    i = np.argmin(abs(x - 0.25))
    assert abs(q[i] - 2.0 / (w * math.sqrt(math.pi)) * math.exp(-((0.25 - x[i]) / w) ** 2)) < 1e-12
    assert abs(x[0] - (0.25 - 0.1)) < 1e-15 and abs(x[-1] + (x[1] - x[0]) - (1.25 + 0.1)) < 1e-12


def test_fourier_matches_fft_on_its_grid():
    """At omega_i = 2 pi i / (NT dt) the transform is NT * ifft of the
    filtered signal times dt (numpy's FFT, an independent algorithm)."""
    nt, dt = 256, 0.05
    t = dt * np.arange(nt)
    rng = np.random.default_rng(1)
    J, E, A = rng.standard_normal(nt), rng.standard_normal(nt), rng.standard_normal(nt)
    nomega = 20
    omegam = nomega * 2 * math.pi / (nt * dt)
    out = po.fourier(t, J, 0.3, E, A, dt, nomega, omegam, window=0)
    f = 1 - 3 * (t / t[-1]) ** 2 + 2 * (t / t[-1]) ** 3
    for sig, col in (((J - 0.3) * f, 1), (E * f, 3), (A * f, 5)):
        ref = np.fft.ifft(sig)[1:nomega + 1] * nt * dt
        np.testing.assert_allclose(out[:, col] + 1j * out[:, col + 1], ref, rtol=0, atol=1e-12)
    # window 1: sin^2 envelope over the first tpulse/dt steps only
    tp = 7.0
    out1 = po.fourier(t, J, 0.0, E, A, dt, nomega, omegam, window=1, tpulse=tp, nenvelope=2)
    g = np.where(np.arange(nt) < int(tp / dt), np.sin(np.pi * t / tp) ** 2, 0.0)
    ref = np.fft.ifft(J * g)[1:nomega + 1] * nt * dt
    np.testing.assert_allclose(out1[:, 1] + 1j * out1[:, 2], ref, rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[:, 0], (np.arange(nomega) + 1) * omegam / nomega, rtol=1e-15)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_taylor_step_is_the_matrix_series(variant):
    """One RT_TE/TEPC/TEPC2 step equals the 4th-order Taylor polynomial of
    the dense matrices H[u] built column by column from the oracle's H|w>."""
    S = po.System(16, 2, 2, 2, 5.0)
    vion = S.ion_potential([2.0], [1.0], [0.0])
    O = po.Oracle(S, vion)
    u0 = S.gram_schmidt(S.init_wf())
    dt, aave = 0.05, 0.1
    n = 16 * 2 * 2

    def hmat(u):
        cols = []
        for k in range(n):
            e = np.zeros(n, complex)
            e[k] = 1.0
            cols.append(O.ubkwbk_hwbk(u, e.reshape(S.shape, order="F"), aave).ravel(order="F"))
        return np.array(cols).T

    def series(H, x):
        out, w = x.copy(), x.copy()
        for k in range(1, 5):
            w = -1j * dt * (H @ w) / k
            out = out + w
        return out

    x0 = u0.ravel(order="F")
    H0 = hmat(u0)
    if variant == 0:
        want = series(H0, x0)
    else:
        x1 = series(H0, x0)
        if variant == 1:
            want = series(0.5 * (H0 + hmat(x1.reshape(S.shape, order="F"))), x0)
        else:
            want = series(hmat((0.5 * (x0 + x1)).reshape(S.shape, order="F")), x0)
    got = O.TE_step(u0, dt, aave, variant).ravel(order="F")
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
