"""The batched uneven workload at its full size on the GPU: 4096 permanents,
n ~ U{8..24}, row and column multiplicities 1..3, selections of a seeded
48-mode Haar unitary (piquasso_amd.workloads.batched_uneven(seed=4)).

Every item is checked against its double-double oracle value, committed in
tests/golden/batched_uneven_seed4.npz (with the unitary the selections are
taken from), at the engine's 1e-10 relative-error bar, in both launch modes
of a batch: direct launches and the hipGraph replay.  The two modes run the
same kernels on the same plan, so their results must agree bit for bit, on
the first replay and on later ones.
"""

import os

import numpy as np
import pytest

import piquasso_amd as pa
from piquasso_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def workload():
    fx = np.load(os.path.join(ROOT, "tests", "golden", "batched_uneven_seed4.npz"))
    mats, rows, cols = W.batched_uneven(seed=4, unitary=fx["unitary"])
    dd = fx["dd"]
    ref = (dd[:, 0] + dd[:, 1]) + 1j * (dd[:, 2] + dd[:, 3])
    assert len(mats) == 4096 and ref.shape == (4096,)
    return mats, rows, cols, ref


def test_workload_shape(workload):
    mats, rows, cols, _ = workload
    ns = np.array([int(np.sum(r)) for r in rows])
    assert ns.min() == 8 and ns.max() == 24
    for r, c in zip(rows, cols):
        assert int(np.sum(r)) == int(np.sum(c))
        assert set(np.unique(r[r > 0])) <= {1, 2, 3} and set(np.unique(c[c > 0])) <= {1, 2, 3}


@pytest.mark.parametrize("persistent", [False, True])
def test_all_4096_items_in_both_launch_modes(gpu, workload, persistent):
    mats, rows, cols, ref = workload
    print(f"batched workload, persistent={persistent}", flush=True)
    with pa.PermanentBatch(mats, rows, cols, persistent=persistent) as b:
        b.run("eager")
        b.sync()
        eager = b.results()
        err = np.abs(eager - ref) / np.abs(ref)
        print(f"eager: max rel.err {err.max():.3e} (item {int(err.argmax())}), "
              f"median {np.median(err):.3e}", flush=True)
        assert err.max() <= 1e-10, (int(err.argmax()), eager[err.argmax()], ref[err.argmax()])
        for replay in range(3):
            b.run("graph")
            b.sync()
            graph = b.results()
            assert np.array_equal(graph.view(np.float64), eager.view(np.float64)), replay


def test_both_orientations_agree(gpu, workload):
    """perm(A; r, c) = perm(A^T; c, r): the planner sweeps some items
    transposed; with as_given=True every item is swept with its rows as the
    Gray digits, as the reference does.  Both plans meet the oracle, the
    as-given plan executes the reference's FLOP count, and the default plan
    fewer."""
    mats, rows, cols, ref = workload
    stats, vals = {}, {}
    for as_given in (False, True):
        print(f"batched workload, as_given={as_given}", flush=True)
        with pa.PermanentBatch(mats, rows, cols, as_given=as_given) as b:
            b.run("eager")
            b.sync()
            vals[as_given] = b.results()
            stats[as_given] = b.stats()
        err = np.abs(vals[as_given] - ref) / np.abs(ref)
        assert err.max() <= 1e-10, (as_given, int(err.argmax()))
    given = sum(W.glynn_flops(r, c) for r, c in zip(rows, cols))
    assert stats[True]["flops"] == given
    assert stats[False]["flops"] < 0.8 * given
    assert not np.array_equal(vals[True], vals[False])  # some items took the other sweep


def test_timing_modes_leave_results_exact(workload):
    """pqp_batch_time in every mode gives positive finite times, the
    sweep-only time is at most the whole evaluation's, and the results of
    the last timed evaluation are the oracle's values."""
    mats, rows, cols, ref = workload
    k = slice(0, 512)
    with pa.PermanentBatch(mats[k], rows[k], cols[k]) as b:
        times = {}
        for mode in ("eager", "graph", "kernel"):
            ms = b.time(mode, reps=6)
            assert ms.shape == (6,) and np.all(np.isfinite(ms)) and np.all(ms > 0), mode
            times[mode] = float(np.median(ms))
            got = b.results()
            rel = np.abs(got - ref[k]) / np.abs(ref[k])
            assert rel.max() <= 1e-10, (mode, float(rel.max()))
        assert times["kernel"] <= times["eager"] * 1.05, times
        assert b.time("eager", reps=0).shape == (0,)
        # one pair of events around back-to-back evaluations
        for mode in ("graph", "eager"):
            span = b.time_span(mode, 8)
            assert np.isfinite(span) and span > 0, mode
            rel = np.abs(b.results() - ref[k]) / np.abs(ref[k])
            assert rel.max() <= 1e-10, (mode, float(rel.max()))
        with pytest.raises(ValueError):
            b.time_span("kernel", 4)
