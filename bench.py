"""Benchmark of the plain FP64 TDHF RK4 path on one MI355X.

Workload (one trajectory, 4 fused launches per RK4 step): 1D periodic cell
a=10 bohr, NG=64, NK=64, NB=NBocc=4 (Z=8, alpha=1), soft-Coulomb e-e and
e-ion, RK4 with E(t), N(t), J(t) evaluated after every step, sin^2 pulse
A0=0.3, w=0.057 Ha, T=400 au.  dt = 0.01 au: the specified 0.02 au exceeds the
RK4 stability bound of this grid (|lambda_max| dt ~ 4.4 > 2*sqrt(2)); the
per-step cost does not depend on dt.  Orbitals start from a ground state
computed on the GPU (not timed).

A step = one RK4 step of the trajectory (4 fused H-apply/stage launches) plus
the observables.  value = K / wall seconds (traj-steps/s) of one
tdhf_rt_advance(K) call, with the state and the field resident in HBM (the
call returns after the device finished; graph-replayed chunks).  The roofline
entry uses the algorithmic FP64 flop count per launch (F_step/4, with F_step
from the reference algorithm's formula) over the average duration of the
fused apply kernel, measured with HIP events on the stream it runs on.  The
whole pulse (40000 steps) is then run once; the FP64 C oracle (same
algorithm, OpenMP over output k-points) takes the device state at mid-pulse,
runs a bounded sample of steps (the CPU baseline), and its E and J are
compared with the device's (parity entry).  The batched and adjoint legs run
afterwards, each in a child process, and are merged into the same line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TDHF traj-RK4-steps/s batched vs plain FP64 (rel dJ<1e-8); adjoint/fwd cost"
PEAK_FP64_TFLOPS = 78.6
# HBM bytes per apply_kernel<64> launch of the plain workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE per dispatch scaled by a plain 1 GiB device copy
# (FETCH x2.0, WRITE x1.0): profiles/apply_kernel_ng64_nk64_nb4_counters.md
TRAFFIC_KIB = {"fetch_kib": 1568.4, "fetch_scale": 2.0, "write_kib": 449.8, "write_scale": 1.0}

CFG = dict(acell=10.0, ng=64, nk=64, nb=4, nbocc=4, z=8.0, alpha=1.0, dt=0.01, A0=0.3,
           omega=0.057, T=400.0)


def flops_per_step(ng, nk, nb, nbocc):
    """Algorithmic FP64 flops of one RK4 step (reference algorithm, no symmetry):
    F = 5 N log2 N per c2c FFT, P = NK^2 NB NBocc exchange pairs,
    F_H = P(16N + 2F) + NB NK (2F + 12N) + (2F + 4N) + NK NBocc 4N + NB NK 4N,
    F_step = 4 F_H + 40 NG NB NK."""
    n = ng
    f = 5.0 * n * math.log2(n)
    p = nk * nk * nb * nbocc
    fh = p * (16 * n + 2 * f) + nb * nk * (2 * f + 12 * n) + (2 * f + 4 * n) + nk * nbocc * 4 * n \
        + nb * nk * 4 * n
    return 4 * fh + 40 * ng * nb * nk


def setup(cfg, gs_iters=60):
    from mattohf_amd import tdhf as T

    ng, nk, nb, nbo, a = cfg["ng"], cfg["nk"], cfg["nb"], cfg["nbocc"], cfg["acell"]
    vion = T.ion_potential(ng, nk, a, [cfg["z"]], [cfg["alpha"]], [0.0])
    G = T.TDHF(ng, nk, nb, nbo, a, cfg["dt"], vion)
    u = G.Gram_Schmidt(T.init_wf(ng, nb, nk, a))
    eh = None
    if gs_iters > 0:
        u, eh, eps, _ = G.GS_calculation(u, gs_iters, 4, tol=1e-11)
    return G, vion, u, eh


def host_threads():
    """Host threads for the CPU legs: the CPU set this process may run on,
    capped by OMP_NUM_THREADS when that is set (the GPU box allots each GPU a
    share of its host cores and sets OMP_NUM_THREADS to it)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def oracle_for(cfg, vion, vq):
    """The CPU oracle of the benchmark cell, fed with the exchange table rows
    the device built (vq[d], d = ikr - ikc + NK - 1; the tables themselves are
    compared with the oracle's in tests/test_gpu_parity.py)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    ng, nk, nb, nbo, a = cfg["ng"], cfg["nk"], cfg["nb"], cfg["nbocc"], cfg["acell"]
    S = po.System(ng, nk, nb, nbo, a)
    full = np.empty((nk * nk, ng))
    for ikr in range(nk):
        for ikc in range(nk):
            full[ikr * nk + ikc] = vq[ikr - ikc + nk - 1]
    return po, po.Oracle(S, vion, full)


def cpu_baseline(cfg, vion, vq, u_start, A, Aave, s0, budget_s=12.0):
    """The FP64 C oracle (same algorithm: RK4 with E, N, J after every step,
    src/RT_calculation.f90:38-47; OpenMP over output k-points) on a bounded
    sample of steps s0, s0+1, ... of the benchmark trajectory, started from
    the device state after step s0-1.  Returns (cpu_baseline entry, oracle E,
    oracle J) of the sample."""
    threads = host_threads()
    po, O = oracle_for(cfg, vion, vq)
    po.set_threads(threads)
    uu = np.asfortranarray(u_start).copy(order="F")
    es, js = [], []
    t0 = time.perf_counter()
    n = 0
    while s0 + n < len(A):
        it = s0 + n
        uu = O.RT_RK4_ubk(uu, cfg["dt"], Aave[it], A[it])
        O.nele(uu)
        es.append(O.ubk_Etot(uu, A[it]))
        js.append(O.ubk_Je(uu, A[it]))
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    entry = {"value": n / dt, "unit": "traj-steps/s", "cores": threads, "kind": "port",
             "sample": f"{n} RK4 steps (+E/N/J each step) of the benchmark trajectory from step "
                       f"{s0} (mid-pulse), FP64 C oracle (same algorithm), OpenMP {threads} "
                       f"threads, {dt:.1f} s"}
    return entry, np.array(es), np.array(js)


def copy_probe(nbytes=1 << 30, reps=4):
    """A plain device-to-device copy of nbytes, reps times (torch's copy kernel):
    under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE it gives the counter reading
    of a kernel whose HBM traffic is known (nbytes read + nbytes written)."""
    import torch

    n = nbytes // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)


def batch_pulses(ntraj, nt, dt, T, seed):
    """Seeded pulse set of the batched workload: A0 log-uniform in [0.01, 0.5],
    omega uniform in [0.03, 0.10] Ha, CEP uniform in [0, 2 pi] (splitmix-seeded
    numpy generator), A(t) = A0 sin^2(pi t/T) sin(omega t + CEP)."""
    from mattohf_amd import tdhf as T_

    rng = np.random.default_rng(seed)
    A0 = np.exp(rng.uniform(math.log(0.01), math.log(0.5), ntraj))
    om = rng.uniform(0.03, 0.10, ntraj)
    cep = rng.uniform(0.0, 2 * math.pi, ntraj)
    A = np.empty((ntraj, nt))
    Aave = np.empty((ntraj, nt))
    for t in range(ntraj):
        oc, tp, e0, ph, ne = T_.pulse_lp_params(A0[t], om[t], T)
        _, A[t], Aave[t], _ = T_.define_afield_lp(nt, dt, oc, tp, e0, ph + cep[t] / (2 * math.pi), ne)
    return A, Aave


def batch_bench(cfg, nks, steps, warmup, seed, gs_iters=60, autotune=True, exchange=None):
    """Batched run of len(nks) trajectories (NG, NB, a, Z of cfg; NK per
    trajectory).  Each trajectory starts from the ground state of its NK.
    exchange (0 row-wise; 1, 2, 3 pair tiles of 64, 32, 16 rows) fixes the
    exchange evaluation, with one wave per workgroup and the interleaved row
    schedule, instead of the autotuner's choice (for profiling runs)."""
    from mattohf_amd import tdhf as T

    ng, nb, a = cfg["ng"], cfg["nb"], cfg["acell"]
    gs, vi = {}, {}
    for nk in sorted(set(nks)):
        vion = T.ion_potential(ng, nk, a, [cfg["z"]], [cfg["alpha"]], [0.0])
        with T.TDHF(ng, nk, nb, cfg["nbocc"], a, cfg["dt"], vion) as G:
            u = G.Gram_Schmidt(T.init_wf(ng, nb, nk, a))
            gs[nk] = G.GS_calculation(u, gs_iters, 4, tol=1e-11)[0]
        vi[nk] = vion
    ntraj = len(nks)
    A, Aave = batch_pulses(ntraj, warmup + steps + 25, cfg["dt"], cfg["T"], seed)
    B = T.Batch(ng, nb, cfg["nbocc"], a, cfg["dt"], nks, np.array([vi[k] for k in nks]))
    B.upload([gs[k] for k in nks])
    if exchange is not None:
        B.set_waves(1)
        B.set_schedule(1)
        B.set_exchange(exchange)
    elif autotune:
        B.autotune(A[:, :25], Aave[:, :25])
    if warmup:
        B.propagate(A[:, :warmup], Aave[:, :warmup])
    t0 = time.perf_counter()
    e, n, j, _ = B.propagate(A[:, warmup:warmup + steps], Aave[:, warmup:warmup + steps])
    wall = time.perf_counter() - t0
    B.set_kernel_timing(True)
    B.propagate(A[:, warmup + steps:warmup + steps + 10], Aave[:, warmup + steps:warmup + steps + 10])
    launches, ms = B.kernel_stats()
    B.set_kernel_timing(False)
    waves = B.waves()
    sched = B.schedule()
    xmode = B.exchange()
    B.close()
    f_launch = sum(flops_per_step(ng, k, nb, cfg["nbocc"]) for k in nks) / 4.0
    avg_s = ms / launches / 1e3
    return {
        "traj_steps_per_s": ntraj * steps / wall,
        "ntraj": ntraj,
        "nk_counts": {str(k): int(sum(1 for x in nks if x == k)) for k in sorted(set(nks))},
        "steps": steps,
        "ms_per_batch_step": wall / steps * 1e3,
        "waves_per_workgroup": waves,
        "row_schedule": sched,
        "exchange": T.Batch.exchange_name(xmode),
        "avg_launch_us": avg_s * 1e6,
        "avg_launch_note": "HIP events around one RK4 stage of the whole batch (with pair "
                           "tiles: the pair-tile kernel and the fused stage kernel together)",
        "achieved_tflops": f_launch / avg_s / 1e12,
        "frac": f_launch / avg_s / 1e12 / PEAK_FP64_TFLOPS,
        "flop_basis": "the reference algorithm's count (all NK^2 NB NBocc ordered exchange "
                      "pairs per H application); pair tiles run one transform pair per "
                      "unordered pair, so they execute fewer FP64 operations than this basis",
        "observables_finite": bool(np.isfinite(e).all() and np.isfinite(j).all()),
    }


ADJ_CFG = dict(acell=8.0, ng=32, nk=32, nb=2, nbocc=2, z=4.0, alpha=1.0, dt=0.025, nt=4000,
               nknots=200, knot_sigma=0.05, fd_h=1e-5, fd_dirs=8, cpu_steps=150)


def adjoint_bench(seed=4, cpu_check=True):
    """Adjoint workload: the CPU-runnable cell at NK=32, 4000 RK4 steps,
    A(t) = sin^2(pi t/T) * piecewise-linear(200 knots ~ N(0, 0.05^2)),
    objective L = E(T) - E_GS.  Reports the cost of one gradient (forward
    with tape + backward) relative to a plain forward run (the same steps
    with E/N/J every step), central finite differences (h=1e-5) of L along 8
    random knot directions, and, on a bounded 150-step prefix, the CPU
    oracle's adjoint gradient."""
    from mattohf_amd import tdhf as T

    c = ADJ_CFG
    ng, nk, nb, a, dt, nt = c["ng"], c["nk"], c["nb"], c["acell"], c["dt"], c["nt"]
    vion = T.ion_potential(ng, nk, a, [c["z"]], [c["alpha"]], [0.0])
    G = T.TDHF(ng, nk, nb, c["nbocc"], a, dt, vion)
    u = G.Gram_Schmidt(T.init_wf(ng, nb, nk, a))
    ugs, eh, _, _ = G.GS_calculation(u, 200, 4, tol=1e-13)
    egs = eh[-1]
    rng = np.random.default_rng(seed)
    P = T.knot_matrix(nt, dt, nt * dt, c["nknots"])
    knots = c["knot_sigma"] * rng.standard_normal(c["nknots"])
    A = P @ knots
    Aave = 0.5 * (A + np.concatenate([[0.0], A[:-1]]))
    # plain forward run (graph-replayed steps, observables every step)
    G.upload(ugs)
    G.propagate(A[:50], Aave[:50])  # warm-up
    G.upload(ugs)
    t0 = time.perf_counter()
    G.propagate(A, Aave)
    t_plain = time.perf_counter() - t0
    # gradient: forward with tape + backward
    t0 = time.perf_counter()
    L, _, _, _, _ = G.fwd(ugs, A, eref=egs)
    t1 = time.perf_counter()
    gA, _ = G.bwd(want_u0=False)
    t2 = time.perf_counter()
    full, seg = G.adjoint_info()
    gk = P.T @ gA
    h = c["fd_h"]
    rels, rels_norm, rels4, rels_h4, dirs = [], [], [], [], []
    gn = np.linalg.norm(gk)
    for _ in range(c["fd_dirs"]):
        d = rng.standard_normal(c["nknots"])
        d /= np.linalg.norm(d)
        lv = {k: G.fwd(ugs, P @ (knots + k * h * d), eref=egs, want_obs=False)[0]
              for k in (-2, -1, 1, 2)}
        fd = (lv[1] - lv[-1]) / (2 * h)
        # fourth-order central difference on the same points (truncation O(h^4))
        fd4 = (8.0 * (lv[1] - lv[-1]) - (lv[2] - lv[-2])) / (12 * h)
        # central difference with 10 h: a tenth of the rounding noise of L / (2 h)
        hb = 10.0 * h
        fdb = (G.fwd(ugs, P @ (knots + hb * d), eref=egs, want_obs=False)[0] -
               G.fwd(ugs, P @ (knots - hb * d), eref=egs, want_obs=False)[0]) / (2 * hb)
        rels.append(abs(fd - gk @ d) / abs(fd))
        rels4.append(abs(fd4 - gk @ d) / abs(fd4))
        rels_h4.append(abs(fdb - gk @ d) / abs(fdb))
        rels_norm.append(abs(fd - gk @ d) / gn)
        dirs.append({"g_dot_d_over_g_norm": float(gk @ d / gn), "rel_err": float(rels[-1]),
                     "rel_err_h1e-4": float(rels_h4[-1])})
    out = {
        "workload": "adjoint of E(T)-E_GS: a=8, NG=32, NK=32, NB=2, 4000 RK4 steps dt=0.025, "
                    "200 piecewise-linear knots x sin^2 envelope",
        "L": L,
        "t_forward_s": t_plain,
        "t_fwd_taped_s": t1 - t0,
        "t_bwd_s": t2 - t1,
        "gradient_over_forward": (t2 - t0) / t_plain,
        "tape": "full" if full else f"sqrt(N) checkpoints every {seg} steps",
        "fd_h": h,
        "fd_max_rel_err": max(rels),
        "fd4_max_rel_err": max(rels4),
        "fd_max_err_over_grad_norm": max(rels_norm),
        "fd_roundoff_floor": 1e-16 * abs(egs) / h / gn,
        "fd_roundoff_note": "L = E(T) - E_GS carries the rounding of two ~|E_GS| totals; "
                            "divided by 2h it bounds |fd - g.d| / |g| from below, so a "
                            "direction with small |g.d| / |g| has a large relative error",
        "fd_h1e-4_max_rel_err": max(rels_h4),
        "fd_dirs_detail": dirs,
        "fd_dirs": len(rels),
    }
    if cpu_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po

        po.set_threads(host_threads())
        ns = c["cpu_steps"]
        S = po.System(ng, nk, nb, c["nbocc"], a)
        vcg, vq = G.tables()
        full_t = np.empty((nk * nk, ng))
        for ikr in range(nk):
            for ikc in range(nk):
                full_t[ikr * nk + ikc] = vq[ikr - ikc + nk - 1]
        O = po.Oracle(S, vion, full_t)
        t0 = time.perf_counter()
        Lo, go, _ = O.adjoint_energy(ugs, dt, A[:ns], egs)
        tcpu = time.perf_counter() - t0
        G.fwd(ugs, A[:ns], eref=egs, want_obs=False)
        gg, _ = G.bwd(want_u0=False)
        out["cpu_gradient_check"] = {
            "steps": ns, "max_rel_diff": float(np.abs(gg - go).max() / np.abs(go).max()),
            "cpu_seconds": tcpu}
    G.close()
    return out


SCALE_CFG = dict(acell=12.0, ng=128, nk=128, nb=8, nbocc=8, z=16.0, alpha=1.0, dt=0.004,
                 ntraj=32, nt_full=8000, T=400.0, fd_h=1e-5)


def adjoint_scale_bench(nt=8, seed=20261205, gs_iters=40):
    """Many-pulse adjoint at scale: a=12, NG=128, NK=128, NB=NBocc=8 (Z=16), 32
    pulses (A0 log-uniform [0.01, 0.5], omega uniform [0.03, 0.10] Ha, CEP
    uniform, sin^2 envelope T=400 au), objective L = dt sum_n J_n^2, gradients
    of all 32 pulses per forward/backward pair.  A bounded sample of nt steps
    is timed (the specified 8000 steps take ~1e3 x longer); the memory budget
    is set so the sample uses sqrt(nt) checkpoints, the spacing the full run
    uses under its budget.  dt = 0.004 au: 0.02 au exceeds the RK4 stability
    bound of this grid; the per-step cost does not depend on dt.
    Reports the plain batched forward rate (tdhf_batch, same pulses), the
    gradient cost relative to it, and central finite differences of every
    pulse's L along one random direction each (one pair of batched forwards)."""
    from mattohf_amd import tdhf as T

    c = SCALE_CFG
    ng, nk, nb, a, dt, ntr = c["ng"], c["nk"], c["nb"], c["acell"], c["dt"], c["ntraj"]
    vion = T.ion_potential(ng, nk, a, [c["z"]], [c["alpha"]], [0.0])
    G = T.TDHF(ng, nk, nb, c["nbocc"], a, dt, vion)
    t0 = time.perf_counter()
    u = G.Gram_Schmidt(T.init_wf(ng, nb, nk, a))
    ugs = G.GS_calculation(u, gs_iters, 4, tol=1e-11)[0]
    t_gs = time.perf_counter() - t0
    A, Aave = batch_pulses(ntr, nt, dt, c["T"], seed)
    us = [ugs] * ntr
    # plain batched forward of the same pulses (E/N/J every step)
    B = T.Batch(ng, nb, c["nbocc"], a, dt, [nk] * ntr, np.array([vion] * ntr))
    B.upload(us)
    B.propagate(A[:, :2], Aave[:, :2])  # warm-up (the state is re-uploaded)
    B.upload(us)
    t0 = time.perf_counter()
    B.propagate(A, Aave)
    t_plain = time.perf_counter() - t0
    B.close()
    state = ng * nb * nk * 16 * ntr
    S = int(math.ceil(math.sqrt(nt)))
    budget = ((nt + S - 1) // S + 4 * S) * state
    G.fwd_batch(us, A[:, :2], objective="jsq", want_obs=False)  # warm-up
    t0 = time.perf_counter()
    L, _, _, _, j = G.fwd_batch(us, A, objective="jsq", mem_budget=budget)
    t1 = time.perf_counter()
    g, _ = G.bwd_batch(want_u0=False)
    t2 = time.perf_counter()
    full, seg = G.adjoint_info()
    rng = np.random.default_rng(seed + 1)
    d = rng.standard_normal(A.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    h = c["fd_h"]
    lp = G.fwd_batch(us, A + h * d, objective="jsq", want_obs=False)[0]
    lm = G.fwd_batch(us, A - h * d, objective="jsq", want_obs=False)[0]
    G.close()
    fd = (lp - lm) / (2 * h)
    ad = np.einsum("tn,tn->t", g, d)
    rel = np.abs(fd - ad) / np.maximum(np.abs(fd), 1e-300)
    S_full = int(math.ceil(math.sqrt(c["nt_full"])))
    plan_bytes = ((c["nt_full"] + S_full - 1) // S_full + 4 * S_full) * state
    fstep = flops_per_step(ng, nk, nb, c["nbocc"])
    plain_rate = ntr * nt / t_plain
    return {
        "workload": "a=12, NG=128, NK=128, NB=8, 32 pulses per launch, objective dt*sum J^2, "
                    f"{nt}-step sample of the {c['nt_full']}-step run, dt={dt}",
        "gs_seconds": t_gs,
        "plain_batched_traj_steps_per_s": plain_rate,
        "plain_batched_achieved_tflops": plain_rate * fstep / 1e12,
        "plain_batched_frac": plain_rate * fstep / 1e12 / PEAK_FP64_TFLOPS,
        "t_fwd_taped_s": t1 - t0,
        "t_bwd_s": t2 - t1,
        "gradient_over_forward": (t2 - t0) / t_plain,
        "tape": "full" if full else f"sqrt(N) checkpoints every {seg} steps",
        "full_run_checkpoint_plan": {"spacing": S_full, "bytes": plan_bytes},
        "fd_h": h,
        "fd_max_rel_err": float(rel.max()),
        "fd_pulses": int(ntr),
        "L_range": [float(L.min()), float(L.max())],
    }


def adjoint_full_bench(npulse=8, nt=8000, seed=20261206, gs_iters=40, fd=True):
    """The batched adjoint at its full length: the a=12, NG=128, NK=128, NB=8
    cell, npulse pulses of the many-pulse set (A0 log-uniform [0.01, 0.5],
    omega uniform [0.03, 0.10] Ha, CEP uniform, sin^2 envelope T=400 au) over
    all nt = 8000 steps (dt = 0.004 au), objective L = dt sum_n J_n^2.  The
    forward keeps sqrt(nt)-spaced checkpoints under the default memory budget
    (a quarter of the free device memory); the backward recomputes each
    segment.  Reports the plain batched forward of the same pulses over the
    same steps, the gradient cost relative to it, and central finite
    differences of every pulse's L along one random direction (two more
    checkpointed forwards).  Opt-in (--only adjoint-full): ~10 minutes at
    npulse = 8; progress lines go to stderr."""
    from mattohf_amd import tdhf as T

    c = SCALE_CFG
    ng, nk, nb, a, dt = c["ng"], c["nk"], c["nb"], c["acell"], c["dt"]

    def say(msg):
        print(f"adjoint-full: {msg}", file=sys.stderr, flush=True)

    def progress(phase, done, total):
        if done % 10 == 0 or done == total:
            say(f"{'forward' if phase == 0 else 'backward'} segment {done}/{total}")

    vion = T.ion_potential(ng, nk, a, [c["z"]], [c["alpha"]], [0.0])
    G = T.TDHF(ng, nk, nb, c["nbocc"], a, dt, vion)
    u = G.Gram_Schmidt(T.init_wf(ng, nb, nk, a))
    ugs = G.GS_calculation(u, gs_iters, 4, tol=1e-11)[0]
    A, Aave = batch_pulses(npulse, nt, dt, c["T"], seed)
    us = [ugs] * npulse
    B = T.Batch(ng, nb, c["nbocc"], a, dt, [nk] * npulse, np.array([vion] * npulse))
    B.upload(us)
    B.propagate(A[:, :2], Aave[:, :2])  # warm-up (the state is re-uploaded)
    B.upload(us)
    chunk, t_plain = 1000, 0.0
    for i in range(0, nt, chunk):
        t0 = time.perf_counter()
        B.propagate(A[:, i:i + chunk], Aave[:, i:i + chunk])
        t_plain += time.perf_counter() - t0
        say(f"plain forward {min(nt, i + chunk)}/{nt} steps")
    B.close()
    G.set_progress(progress)
    G.fwd_batch(us, A[:, :2], objective="jsq", want_obs=False)  # warm-up
    t0 = time.perf_counter()
    L, _, _, _, j = G.fwd_batch(us, A, objective="jsq")
    t1 = time.perf_counter()
    g, _ = G.bwd_batch(want_u0=False)
    t2 = time.perf_counter()
    full, seg = G.adjoint_info()
    out = {
        "workload": f"a=12, NG=128, NK=128, NB=8, {npulse} pulses per launch, objective "
                    f"dt*sum J^2, all {nt} steps, dt={dt}",
        "seed": seed,
        "plain_batched_s": t_plain,
        "plain_batched_traj_steps_per_s": npulse * nt / t_plain,
        "plain_note": "tdhf_batch of the same pulses, E/N/J every step, in 1000-step calls",
        "t_fwd_checkpointed_s": t1 - t0,
        "t_bwd_s": t2 - t1,
        "gradient_over_forward": (t2 - t0) / t_plain,
        "tape": "full" if full else f"sqrt(N) checkpoints every {seg} steps",
        "checkpoint_bytes": int(((nt + seg - 1) // seg + 4 * seg) * ng * nb * nk * 16 * npulse),
        "L_range": [float(L.min()), float(L.max())],
        "grad_finite": bool(np.isfinite(g).all()),
    }
    if fd:
        rng = np.random.default_rng(seed + 1)
        d = rng.standard_normal(A.shape)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        h = c["fd_h"]
        say("finite differences: forward at A + h d")
        lp = G.fwd_batch(us, A + h * d, objective="jsq", want_obs=False)[0]
        say("finite differences: forward at A - h d")
        lm = G.fwd_batch(us, A - h * d, objective="jsq", want_obs=False)[0]
        fdv = (lp - lm) / (2 * h)
        ad = np.einsum("tn,tn->t", g, d)
        rel = np.abs(fdv - ad) / np.maximum(np.abs(fdv), 1e-300)
        out.update({"fd_h": h, "fd_max_rel_err": float(rel.max()),
                    "fd_rel_err": [float(x) for x in rel]})
    G.set_progress(None)
    G.close()
    return out


def batched_full_bench(nt=10000, ntraj=256, seed=20261202, gs_iters=60):
    """The batched workload at its specified length: ntraj trajectories with NK
    drawn from {16, 32, 64} (a=10, NG=64, NB=NBocc=4, Z=8), the seeded pulse
    set of the batched leg, nt RK4 steps (dt = 0.01 au: 0.02 exceeds the RK4
    stability bound of this grid), each from the GPU ground state of its NK,
    autotuned launch.  The first trajectory of each NK is run again alone
    (tdhf_rt_run on its own context) and its E(t), J(t) and final orbitals are
    compared with the batch's over the whole run.  Opt-in (--only
    batched-full); progress lines go to stderr."""
    from mattohf_amd import tdhf as T

    cfg = dict(CFG)
    ng, nb, a, dt = cfg["ng"], cfg["nb"], cfg["acell"], cfg["dt"]

    def say(msg):
        print(f"batched-full: {msg}", file=sys.stderr, flush=True)

    rng = np.random.default_rng(seed)
    nks = [int(k) for k in rng.choice([16, 32, 64], size=ntraj)]
    gs, vi = {}, {}
    for nk in sorted(set(nks)):
        vion = T.ion_potential(ng, nk, a, [cfg["z"]], [cfg["alpha"]], [0.0])
        with T.TDHF(ng, nk, nb, cfg["nbocc"], a, dt, vion) as G:
            u = G.Gram_Schmidt(T.init_wf(ng, nb, nk, a))
            gs[nk] = G.GS_calculation(u, gs_iters, 4, tol=1e-11)[0]
        vi[nk] = vion
    A, Aave = batch_pulses(ntraj, nt, dt, cfg["T"], seed)
    B = T.Batch(ng, nb, cfg["nbocc"], a, dt, nks, np.array([vi[k] for k in nks]))
    B.upload([gs[k] for k in nks])
    B.autotune(A[:, :25], Aave[:, :25])
    say(f"autotuned: {B.waves()} wave(s) per workgroup, row schedule {B.schedule()}, "
        f"exchange {T.Batch.exchange_name(B.exchange())}")
    chunk, wall, es, ns, js = 1000, 0.0, [], [], []
    for i in range(0, nt, chunk):
        t0 = time.perf_counter()
        e, n, j, _ = B.propagate(A[:, i:i + chunk], Aave[:, i:i + chunk])
        wall += time.perf_counter() - t0
        es.append(e)
        ns.append(n)
        js.append(j)
        say(f"{min(nt, i + chunk)}/{nt} steps")
    finals = B.download()
    waves, sched, xmode = B.waves(), B.schedule(), B.exchange()
    B.close()
    e, n, j = np.hstack(es), np.hstack(ns), np.hstack(js)
    checks = {}
    for nk in sorted(set(nks)):
        t = nks.index(nk)
        with T.TDHF(ng, nk, nb, cfg["nbocc"], a, dt, vi[nk]) as G:
            u1, e1, n1, j1, _ = G.RT_calculation(gs[nk], A[t], Aave[t])
        checks[str(nk)] = {
            "trajectory": t,
            "rel_dJ": float(np.abs(j[t] - j1).max() / np.abs(j1).max()),
            "abs_dE_Ha": float(np.abs(e[t] - e1).max()),
            "max_abs_du": float(np.abs(finals[t] - u1).max()),
        }
        say(f"NK={nk}: trajectory {t} checked against its single run")
    return {
        "workload": f"{ntraj} trajectories, NK from {{16, 32, 64}}, NG=64, NB=4, all {nt} "
                    f"steps, dt={dt}, hipGraph-replayed stage launches",
        "seed": seed,
        "nk_counts": {str(k): int(sum(1 for x in nks if x == k)) for k in sorted(set(nks))},
        "traj_steps_per_s": ntraj * nt / wall,
        "wall_s": wall,
        "waves_per_workgroup": waves,
        "row_schedule": sched,
        "exchange": T.Batch.exchange_name(xmode),
        "observables_finite": bool(np.isfinite(e).all() and np.isfinite(j).all()),
        "N_rel_drift_max": float(np.abs(n - n[:, :1]).max() / np.abs(n[:, :1]).min()),
        "vs_single_runs": checks,
        "vs_single_note": "one trajectory per NK through tdhf_rt_run on its own context "
                          "(row-wise exchange), the whole run",
    }


LEGS = ("batched", "adjoint", "adjoint-scale")


def run_leg(leg, args, timeout_s):
    """One secondary leg in a child process (its failure cannot take the
    headline line with it); returns its JSON object or an error entry."""
    import subprocess

    cmd = [sys.executable, os.path.abspath(__file__), "--only", leg,
           "--adjoint-scale-steps", str(args.adjoint_scale_steps),
           "--batch-traj", str(args.batch_traj), "--batch-steps", str(args.batch_steps)]
    if args.no_cpu:
        cmd.append("--no-cpu")
    t0 = time.perf_counter()
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout_s} s"}
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"exit {r.returncode}", "stderr_tail": r.stderr[-800:]}
    out = json.loads(lines[-1])
    key = leg.replace("-", "_")
    if isinstance(out.get(key), dict):
        out[key]["leg_wall_s"] = time.perf_counter() - t0
    return out


def headline(args):
    """The plain-path measurement, its kernel timing, the whole-pulse run with
    its mid-pulse parity sample against the oracle, and the CPU baseline."""
    from mattohf_amd import tdhf as T

    cfg = dict(CFG)
    G, vion, u, eh = setup(cfg, gs_iters=args.gs_iters)
    W, K, ts, te = args.warmup, args.steps, args.timing_steps, 25
    nfull = int(round(cfg["T"] / cfg["dt"]))
    lp = T.pulse_lp_params(cfg["A0"], cfg["omega"], cfg["T"])
    nrun = W + K + ts + te
    _, A, Aave, _ = T.define_afield_lp(max(nrun, nfull), cfg["dt"], *lp)

    # the whole pulse (T = 400 au: 40000 steps at dt = 0.01) from the ground
    # state in one run, with the state taken at mid-pulse for the parity
    # sample; it runs first, so the timed steps below find the device at its
    # steady clock
    s0 = nfull // 2
    full = not args.no_full_run
    if full:
        G.upload(u)
        G.rt_begin(A[:nfull], Aave[:nfull])
        f0 = time.perf_counter()
        G.rt_advance(s0)
        f1 = time.perf_counter()
        u_mid = G.download()
        f2 = time.perf_counter()
        G.rt_advance(nfull - s0)
        f3 = time.perf_counter()
        ef, nf, jf, i3f = G.rt_end()
    vcg, vq = G.tables()

    # timed region: K steps of the resident state with the field resident too
    G.upload(u)
    G.rt_begin(A[:nrun], Aave[:nrun])
    G.rt_advance(W)  # warm-up; the first advance captures the step graphs
    t0 = time.perf_counter()
    G.rt_advance(K)  # synchronous: returns after the device finished
    wall = time.perf_counter() - t0
    value = K / wall

    # duration of the fused apply kernel: HIP events on its stream around ts
    # graph-replayed steps (4 launches per step, no per-launch events)
    G.set_kernel_timing(True, block=True)
    G.rt_advance(ts)
    launches, total_ms = G.kernel_stats()
    # the same with an event pair around every launch (graph replay off): an
    # upper bound, each pair adds its own overhead to the interval
    G.set_kernel_timing(True)
    G.rt_advance(te)
    launches_ev, total_ms_ev = G.kernel_stats()
    G.set_kernel_timing(False)
    e, n, j, i3 = G.rt_end()
    G.close()
    if launches < 1:
        raise SystemExit("bench.py: no timed apply launches (--timing-steps must be >= 1)")
    avg_s = total_ms / launches / 1e3
    f_step = flops_per_step(cfg["ng"], cfg["nk"], cfg["nb"], cfg["nbocc"])
    f_launch = f_step / 4.0
    achieved = f_launch / avg_s / 1e12
    fma_tflops, _ = T.fp64_fma_probe()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "traj-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Init_wf + GPU ground state; sin^2 pulse)",
        "config": {
            "workload": "plain FP64 TDHF RK4, one trajectory, 4 fused launches per step, "
                        "E/N/J every step; a=10 NG=64 NK=64 NB=NBocc=4 Z=8",
            "acell": cfg["acell"], "NG": cfg["ng"], "NK": cfg["nk"], "NB": cfg["nb"],
            "NBocc": cfg["nbocc"], "Z": cfg["z"], "dt": cfg["dt"],
            "dt_note": "0.02 au exceeds the RK4 stability bound of NG=64, a=10",
            "pulse": {"A0": cfg["A0"], "omega": cfg["omega"], "T": cfg["T"]},
            "flop_per_traj_step": f_step,
            "timed_call": "tdhf_rt_advance(K) on the device-resident state and field, "
                          "graph-replayed chunks",
            "observables_finite": bool(np.isfinite(e).all() and np.isfinite(j).all()),
        },
        "roofline": {
            "bound": "fp64-valu",
            "bound_detail": "FP64 vector ALU (no dense contraction on the path, no MFMA); "
                            "peak = MI355X FP64 vector",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "peak_measured": fma_tflops,
            "peak_measured_note": "FP64 FMA probe on this device (tdhf_fp64_fma_probe)",
            "frac_of_measured": achieved / fma_tflops,
            "traffic": (TRAFFIC_KIB["fetch_kib"] * TRAFFIC_KIB["fetch_scale"]
                        + TRAFFIC_KIB["write_kib"] * TRAFFIC_KIB["write_scale"]) * 1024.0,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
            "kernel": "apply_kernel<64> (fused exchange + H + RK4 stage)",
            "avg_launch_us": avg_s * 1e6,
            "avg_launch_us_method": "HIP events on the kernel's stream around the "
                                    f"{ts} graph-replayed steps of one call, 4 launches per step",
            "avg_launch_us_per_launch_events": total_ms_ev / max(launches_ev, 1) * 1e3,
            "flop_per_launch": f_launch,
            "timed_launches": launches,
        },
        "cpu_baseline": None,
    }

    if not full:
        return out
    tw = (f1 - f0) + (f3 - f2)
    out["full_run"] = {
        "steps": nfull,
        "duration_au": nfull * cfg["dt"],
        "wall_s": tw,
        "traj_steps_per_s": nfull / tw,
        "N_rel_drift_max": float(np.abs(nf - i3f[1]).max() / i3f[1]),
        "N_note": "RK4 (RT_RK4_ubk) is not unitary: this is the reference algorithm's "
                  "own norm change",
        "E_absorbed_Ha": float(ef[-1] - i3f[0]),
        "J_max_abs": float(np.abs(jf).max()),
        "observables_finite": bool(np.isfinite(ef).all() and np.isfinite(jf).all()),
    }
    if not args.no_cpu:
        entry, eo, jo = cpu_baseline(cfg, vion, vq, u_mid, A[:nfull], Aave[:nfull], s0,
                                     budget_s=args.cpu_seconds)
        out["cpu_baseline"] = entry
        m = len(eo)
        dj = float(np.abs(jf[s0:s0 + m] - jo).max() / np.abs(jo).max())
        de = float(np.abs(ef[s0:s0 + m] - eo).max())
        out["parity"] = {
            "steps": m, "from_step": s0,
            "rel_dJ": dj, "abs_dE_Ha": de,
            "tolerance": {"rel_dJ": 1e-8, "abs_dE_Ha": 1e-10},
            "ok": bool(dj < 1e-8 and de < 1e-10),
            "method": "device observables of the whole-pulse run vs the FP64 oracle from the "
                      "same mid-pulse state, same field; rel_dJ = max|dJ|/max|J|",
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--timing-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--batch-traj", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=30)
    ap.add_argument("--no-legs", action="store_true",
                    help="headline only (no batched / adjoint legs)")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the whole-pulse run (and with it the parity sample and the CPU "
                         "baseline); for profiling runs")
    ap.add_argument("--adjoint-scale-steps", type=int, default=8)
    ap.add_argument("--leg-timeout", type=float, default=300.0)
    ap.add_argument("--only", choices=list(LEGS) + ["adjoint-full", "batched-full"],
                    default=None,
                    help="run one secondary leg alone and print its JSON (adjoint-full: "
                         "the 8000-step batched gradient; batched-full: the 10000-step "
                         "mixed batch; neither is part of the default run)")
    ap.add_argument("--adjoint-full-pulses", type=int, default=8)
    ap.add_argument("--batch-case", choices=["both", "mixed", "uniform"], default="both",
                    help="batched leg: which batch to run (profiling runs)")
    ap.add_argument("--batch-exchange", choices=["auto", "rows", "pairs", "pairs32", "pairs16"],
                    default="auto",
                    help="batched leg: the autotuner's exchange choice, or a fixed one "
                         "(profiling runs)")
    ap.add_argument("--gs-iters", type=int, default=60,
                    help="ground-state iterations before the timed run (0: orthonormalised "
                         "initial guess; used for profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of known size (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py: every GPU box of this build has one GPU and the path has no "
              "multi-GPU mode; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.only == "batched-full":
        print(json.dumps({"batched_full": batched_full_bench()}), flush=True)
        return
    if args.only == "adjoint-full":
        print(json.dumps({"adjoint_full": adjoint_full_bench(npulse=args.adjoint_full_pulses)}),
              flush=True)
        return
    if args.only == "adjoint-scale":
        print(json.dumps({"adjoint_scale": adjoint_scale_bench(nt=args.adjoint_scale_steps)}),
              flush=True)
        return
    if args.only == "adjoint":
        print(json.dumps({"adjoint": adjoint_bench(cpu_check=not args.no_cpu)}), flush=True)
        return
    if args.only == "batched":
        print(json.dumps({"batched": batched_leg(args)}), flush=True)
        return

    out = headline(args)
    print(f"bench: headline {out['value']:.1f} traj-steps/s, "
          f"{out['roofline']['avg_launch_us']:.2f} us/launch", file=sys.stderr, flush=True)
    if not args.no_legs:
        for leg in LEGS:
            r = run_leg(leg, args, args.leg_timeout)
            key = leg.replace("-", "_")
            out[key] = r.get(key, r)
            print(f"bench: leg {leg} done", file=sys.stderr, flush=True)
        b = out.get("batched", {})
        if "uniform_nk64" in b:
            b["batched_vs_plain_at_nk64"] = b["uniform_nk64"]["traj_steps_per_s"] / out["value"]
    print(json.dumps(out), flush=True)


def batched_leg(args):
    cfg = dict(CFG)
    rng = np.random.default_rng(20261202)
    mixed = [int(k) for k in rng.choice([16, 32, 64], size=args.batch_traj)]
    x = {"auto": None, "rows": 0, "pairs": 1, "pairs32": 2, "pairs16": 3}[args.batch_exchange]
    out = {
        "workload": "independent trajectories, one launch per RK4 stage for all, "
                    "hipGraph-replayed steps, autotuned waves/workgroup, row schedule "
                    "and exchange evaluation",
        "seed": 20261202,
    }
    if args.batch_case in ("both", "mixed"):
        out["mixed_nk"] = batch_bench(cfg, mixed, args.batch_steps, 5, seed=20261202,
                                      exchange=x)
    if args.batch_case in ("both", "uniform"):
        out["uniform_nk64"] = batch_bench(cfg, [64] * max(1, args.batch_traj // 4),
                                          args.batch_steps, 5, seed=20261203, exchange=x)
    return out


if __name__ == "__main__":
    main()
