"""The benchmark's host-side legs on the CPU (no GPU): the CPU-baseline /
parity sample of bench.py runs to the end of its field arrays without reading
past them, and its observables are the oracle's RT_calculation of the same
steps."""
import os
import sys

import numpy as np

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _cell():
    cfg = dict(acell=8.0, ng=32, nk=4, nb=2, nbocc=2, z=4.0, alpha=1.0, dt=0.025)
    S = po.System(32, 4, 2, 2, 8.0)
    vion = S.ion_potential([4.0], [1.0], [0.0])
    full = S.vCGkk()
    nk = 4
    vq = np.array([full[(d if d >= 0 else 0) * nk + (0 if d >= 0 else -d)]
                   for d in range(-(nk - 1), nk)])
    return cfg, S, vion, vq


def test_cpu_baseline_stops_at_the_end_of_the_field():
    cfg, S, vion, vq = _cell()
    u = S.gram_schmidt(S.init_wf())
    lp = (0.057 * 27.2116, 100.0 * 0.02419, 0.2 * 0.057 * 514.2, 0.057 * 100.0 / (4 * np.pi), 2)
    _, A, Aave, _ = po.define_afield_lp(12, cfg["dt"], *lp)
    entry, e, j = bench.cpu_baseline(cfg, vion, vq, u, A, Aave, s0=9, budget_s=60.0)
    assert len(e) == 3 and len(j) == 3  # steps 9, 10, 11 of a 12-step field
    assert entry["kind"] == "port" and entry["cores"] >= 1 and entry["value"] > 0
    O = po.Oracle(S, vion)
    _, eo, _, jo, _ = O.RT_calculation(u, cfg["dt"], A[9:], Aave[9:])
    np.testing.assert_allclose(e, eo, rtol=0, atol=1e-12)
    np.testing.assert_allclose(j, jo, rtol=0, atol=1e-13)


def test_host_threads_follows_omp_num_threads(monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.host_threads() == 1
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.host_threads() >= 1


def test_flop_formula_matches_published_table():
    # per traj-step algorithmic flops of the reference algorithm
    assert abs(bench.flops_per_step(64, 64, 4, 4) / 1.281e9 - 1) < 1e-3
    assert abs(bench.flops_per_step(32, 8, 2, 2) / 2.33e6 - 1) < 5e-3
    assert abs(bench.flops_per_step(128, 128, 8, 8) / 4.62e10 - 1) < 1e-3
