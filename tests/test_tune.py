"""The autotuner's table lookup (CPU): which tuned workload class a batch
gets, by item count and mean Gray term count, and how its launch settings
scale with the device's CU count."""

import numpy as np

from piquasso_amd import tune
from piquasso_amd import workloads as W

TABLE = tune.load_tuning(tune.DEFAULT_TABLE)
ARCH = "gfx950:sramecc+:xnack-"


def _single(n):
    return tune.best_config(1, 256, ARCH, log2_terms=tune.mean_log2_terms([np.ones(n, int)]))


def test_single_permanents_take_the_class_of_their_size():
    assert TABLE, "the committed gfx950 table is missing"
    wl = TABLE["workloads"]
    for n, name in ((12, "single_n24"), (20, "single_n24"), (24, "single_n24"), (30, "single_n32"),
                    (32, "single_n32")):
        assert tune.choose_class(TABLE, 1, tune.mean_log2_terms([np.ones(n, int)])) == name, n
        assert _single(n) == {k: v for k, v in wl[name]["best"].items()}, n
    # no size given: the largest single class, as before the size lookup
    assert tune.choose_class(TABLE, 1) == "single_n32"


def test_batches_take_the_class_of_their_items():
    _, rows, _ = W.batched_uneven(count=4096, seed=4)
    _, grows, _ = W.gradient_batch(count=256, seed=5)
    assert tune.choose_class(TABLE, 4096, tune.mean_log2_terms(rows)) == "batched_uneven"
    assert tune.choose_class(TABLE, 256, tune.mean_log2_terms(grows)) == "uniform_n12_20"
    # (a 256-item sample of the 4096: an estimate)
    assert abs(tune.mean_log2_terms(rows) - tune.CLASS_LOG2_TERMS["batched_uneven"]) < 0.5
    assert abs(tune.mean_log2_terms(grows) - tune.CLASS_LOG2_TERMS["uniform_n12_20"]) < 0.05


def test_units_scale_with_the_device_and_other_devices_get_nothing():
    full = tune.best_config(4096, 256, ARCH, log2_terms=14.9)
    half = tune.best_config(4096, 128, ARCH, log2_terms=14.9)
    assert half["total_units"] * 2 == full["total_units"]
    assert half["persistent"] == full["persistent"]
    assert tune.best_config(4096, 304, "gfx942", log2_terms=14.9) == {}
    assert tune.choose_class({"workloads": {}}, 1, 20.0) is None


def test_tuned_inner_length_passes_through(tmp_path):
    # a table whose batch class tuned an inner-sweep length returns it with
    # the other settings; the committed table's entries carry only keys the
    # batch constructor takes
    import json

    table = {"device": "gfx950", "cus": 256, "workloads": {
        "batched_uneven": {"best": {"total_units": 5120, "persistent": False, "target_inner": 256},
                           "log2_terms": 15.2}}}
    path = tmp_path / "t.json"
    path.write_text(json.dumps(table))
    cfg = tune.best_config(4096, 128, ARCH, path=str(path), log2_terms=15.0)
    assert cfg == {"total_units": 2560, "persistent": False, "target_inner": 256}
    for name, rec in TABLE["workloads"].items():
        assert set(rec["best"]) <= {"total_units", "persistent", "target_inner"}, name
        assert set(rec["best"]) <= set(tune.GRIDS[name]), name
