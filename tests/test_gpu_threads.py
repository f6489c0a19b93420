"""Contexts and batches used from different host threads at once (one thread
per context, as include/tdhf.h allows): the process-wide records they share
(the adjoint's forward records, the batch autotuner's cache) are guarded, and
every result equals the same run made alone, bit for bit."""
import threading

import numpy as np
import pytest

import pyoracle as po
from mattohf_amd import tdhf as T

pytestmark = pytest.mark.gpu

NG, NK, NB, A_CELL, DT, NT = 32, 4, 2, 8.0, 0.025, 24


def setup():
    S = po.System(NG, NK, NB, NB, A_CELL)
    vion = S.ion_potential([4.0], [1.0], [0.0])
    rng = np.random.default_rng(11)
    u = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    u = S.gram_schmidt(np.asfortranarray(u))
    return vion, u


def adjoint_run(vion, u, scale):
    _, A, _, _ = T.define_afield_lp(NT, DT, *T.pulse_lp_params(0.1 * scale, 0.057, 1.0))
    with T.TDHF(NG, NK, NB, NB, A_CELL, DT, vion) as G:
        out = []
        for _ in range(3):
            L, *_ = G.fwd_batch([u, u], np.stack([A, 0.5 * A]), objective="jsq",
                                mem_budget=30 * 2 * NG * NB * NK * 16)
            g, _ = G.bwd_batch(want_u0=False)
            out.append((L.copy(), g.copy()))
    return out


def batch_run(vion, u, nks, scale):
    _, A1, Aave1, _ = T.define_afield_lp(NT, DT, *T.pulse_lp_params(0.1 * scale, 0.057, 1.0))
    A = np.tile(A1, (len(nks), 1))
    Aave = np.tile(Aave1, (len(nks), 1))
    us = []
    for nk in nks:
        S = po.System(NG, nk, NB, NB, A_CELL)
        us.append(S.gram_schmidt(S.init_wf()))
    with T.Batch(NG, NB, NB, A_CELL, DT, nks, np.array([vion] * len(nks))) as B:
        B.upload(us)
        B.autotune(A[:, :8], Aave[:, :8])
        e, n, j, _ = B.propagate(A, Aave)
    return e, j


def run_threads(fns):
    res = [None] * len(fns)
    errs = []

    def wrap(i, f):
        try:
            res[i] = f()
        except Exception as ex:  # noqa: BLE001 - reported below
            errs.append(ex)

    th = [threading.Thread(target=wrap, args=(i, f)) for i, f in enumerate(fns)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs, errs
    return res


def test_two_contexts_adjoint_on_two_threads():
    vion, u = setup()
    alone = [adjoint_run(vion, u, s) for s in (1, 2)]
    both = run_threads([lambda: adjoint_run(vion, u, 1), lambda: adjoint_run(vion, u, 2)])
    for a, b in zip(alone, both):
        for (La, ga), (Lb, gb) in zip(a, b):
            np.testing.assert_array_equal(La, Lb)
            np.testing.assert_array_equal(ga, gb)


def test_two_batches_autotuned_on_two_threads():
    vion = po.System(NG, 2, NB, NB, A_CELL).ion_potential([4.0], [1.0], [0.0])
    nks = [2, 4, 2]
    alone = [batch_run(vion, None, nks, s) for s in (1, 2)]
    both = run_threads([lambda: batch_run(vion, None, nks, 1), lambda: batch_run(vion, None, nks, 2)])
    for (ea, ja), (eb, jb) in zip(alone, both):
        # the autotuner may choose another launch shape under contention; the
        # results agree to rounding
        np.testing.assert_allclose(eb, ea, rtol=0, atol=1e-12 * np.abs(ea).max())
        np.testing.assert_allclose(jb, ja, rtol=0, atol=1e-12 * max(1.0, np.abs(ja).max()))
