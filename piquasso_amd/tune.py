"""Autotuner for the permanent engine's launch parameters.

The planner splits each permanent into wave work units (64 top Gray states
per wave, an inner Gray sweep per lane).  How many units a batch gets
(total_units, the waves per permanent in proportion to its cost), how long
each lane's inner sweep is (target_inner, the chunk of Gray states per lane;
0 leaves the split to the planner's cost model) and whether to use the
persistent work-stealing launch or one block per unit are workload
dependent.  This module measures candidate settings on the device (HIP-event
timing of the sweep kernels, after untimed evaluations that bring the clock
to its sustained value, in interleaved passes) and writes the fastest per
workload class to a JSON table that :func:`load_tuning` reads back.

CLI:  python -m piquasso_amd.tune [--out piquasso_amd/tuning/gfx950.json] [--reps 5]
"""

from __future__ import annotations

import argparse
import itertools
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning", "gfx950.json")


def measure(mats, rows, cols, reps: int = 5, mode: str = "kernel", warm_s: float = 0.1, **cfg) -> dict:
    """Median time of the batch under one configuration (PermanentBatch
    kwargs), after about `warm_s` seconds (at least 3 evaluations) of
    untimed evaluations: planning leaves the device idle, and its clock
    drops meanwhile."""
    from .engine import PermanentBatch

    t0 = time.perf_counter()
    with PermanentBatch(mats, rows, cols, **cfg) as b:
        plan_s = time.perf_counter() - t0
        t1 = time.perf_counter()
        k = 0
        while k < 3 or time.perf_counter() - t1 < warm_s:
            b.run("eager")
            b.sync()
            k += 1
        ms = b.time(mode, reps=reps)
        st = b.stats()
    return {"config": cfg, "ms": float(np.median(ms)), "ms_all": ms.tolist(),
            "plan_s": plan_s, "units": st["units"], "flops": st["flops"],
            "tflops": st["flops"] / (float(np.median(ms)) * 1e-3) / 1e12}


def sweep(mats, rows, cols, grid: dict, reps: int = 5, mode: str = "kernel", log=print,
          passes: int = 2, warm_s: float = 0.3):
    """Measure every combination of `grid` (dict of PermanentBatch kwarg ->
    candidate values) in `passes` interleaved passes, after about `warm_s`
    seconds of untimed evaluations of the first combination (the device's
    clock rises under sustained load, and the first combinations would
    otherwise run at a lower one); each record's time is the median over
    the passes.  Returns the records sorted fastest first."""
    from .engine import PermanentBatch

    keys = sorted(grid)
    cfgs = [dict(zip(keys, values)) for values in itertools.product(*[grid[k] for k in keys])]
    with PermanentBatch(mats, rows, cols, **cfgs[0]) as b:
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < warm_s:
            b.run("eager")
            b.sync()
    recs = [None] * len(cfgs)
    for p in range(max(passes, 1)):
        for i, cfg in enumerate(cfgs):
            rec = measure(mats, rows, cols, reps=reps, mode=mode, **cfg)
            log(json.dumps({"pass": p, "config": cfg, "ms": round(rec["ms"], 4),
                            "tflops": round(rec["tflops"], 2), "units": rec["units"]}))
            if recs[i] is None:
                rec["ms_passes"] = [rec["ms"]]
                recs[i] = rec
            else:
                recs[i]["ms_passes"].append(rec["ms"])
    for rec in recs:
        rec["ms"] = float(np.median(rec["ms_passes"]))
        rec["tflops"] = rec["flops"] / (rec["ms"] * 1e-3) / 1e12
    recs.sort(key=lambda r: r["ms"])
    return recs


def _workloads():
    from . import workloads as W

    mats, rows, cols = W.batched_uneven(count=4096, seed=4)  # the batched benchmark workload
    a32, r32, c32 = W.large_single(32, seed=3)
    a24, r24, c24 = W.large_single(24, seed=3)
    g_m, g_r, g_c = W.gradient_batch(count=256, seed=5)
    return {
        "batched_uneven": (mats, rows, cols),
        "single_n32": ([a32], [r32], [c32]),
        "single_n24": ([a24], [r24], [c24]),
        "uniform_n12_20": (g_m, g_r, g_c),
    }


# target_inner > 0 sets every item's inner sweep to the largest Gray
# sub-space of at most that many states and overrides the split total_units
# asks for; 0 lets total_units and the cost model choose it per item.
GRIDS = {
    "batched_uneven": {"total_units": [4096, 5120, 6144, 8192, 12288], "persistent": [False, True],
                       "target_inner": [0, 64, 256, 1024]},
    "single_n32": {"total_units": [16384, 32768, 65536], "persistent": [False, True],
                   "target_inner": [0, 16384]},
    "single_n24": {"total_units": [1024, 2048, 4096, 8192], "persistent": [False, True],
                   "target_inner": [0, 1024]},
    "uniform_n12_20": {"total_units": [2048, 4096, 8192], "persistent": [False, True],
                       "target_inner": [0, 128, 512]},
}


# log2 of the mean Gray terms per item of each tuned workload class (the
# tuner writes the same figure into the table as "log2_terms"; these are the
# fallback for tables written before it did)
CLASS_LOG2_TERMS = {"single_n32": 31.0, "single_n24": 23.0, "batched_uneven": 15.16,
                    "uniform_n12_20": 16.84}
SINGLE_CLASSES = ("single_n24", "single_n32")
BATCH_CLASSES = ("batched_uneven", "uniform_n12_20")


def choose_class(table: dict, count: int, log2_terms=None) -> str:
    """The tuned workload class for `count` permanents whose mean Gray term
    count is 2^log2_terms: among the single-permanent classes for one
    permanent, the batch classes otherwise, the one of nearest log2 term
    count that the table holds (the first present one when log2_terms is
    None); None when the table holds none of them."""
    wl = table.get("workloads", {})
    names = [k for k in (SINGLE_CLASSES if count == 1 else BATCH_CLASSES) if k in wl]
    if not names:
        return None
    if log2_terms is None:
        return names[-1] if count == 1 else names[0]

    def dist(name):
        return abs(float(wl[name].get("log2_terms", CLASS_LOG2_TERMS[name])) - float(log2_terms))

    return min(names, key=dist)


def best_config(count: int, cus: int, arch: str, path: str = None, log2_terms=None) -> dict:
    """Launch settings for a batch of `count` permanents of mean Gray term
    count 2^log2_terms from the tuned table of this device (PermanentBatch
    kwargs: total_units, persistent, target_inner), or {} when no table was written for
    it.  The entry is the tuned class of nearest size (choose_class); unit
    counts scale with the CU count of the table's device."""
    table = load_tuning(path or DEFAULT_TABLE)
    if not table or not arch.split(":")[0] == str(table.get("device", "")).split(":")[0]:
        return {}
    name = choose_class(table, count, log2_terms)
    if name is None:
        return {}
    rec = table["workloads"][name]
    cfg = dict(rec["best"])
    if "total_units" in cfg and table.get("cus"):
        cfg["total_units"] = int(round(cfg["total_units"] / table["cus"] * cus))
    return {k: v for k, v in cfg.items() if k in ("total_units", "persistent", "target_inner")}


def mean_log2_terms(rows, sample: int = 256) -> float:
    """log2 of the mean Gray term count over (up to `sample` evenly spaced)
    items' row multiplicities."""
    from .workloads import glynn_terms

    n = len(rows)
    if n == 0:
        return 0.0
    idx = np.unique(np.linspace(0, n - 1, min(n, sample)).astype(int))
    return float(np.log2(np.mean([float(glynn_terms(rows[k])) for k in idx])))


def tune(out: str = DEFAULT_TABLE, reps: int = 5, only=None, log=print) -> dict:
    from . import _capi

    n, cus, arch = _capi.device_info()
    table = {"device": arch, "cus": cus, "workloads": {}}
    for name, (mats, rows, cols) in _workloads().items():
        if only and name not in only:
            continue
        log(f"== {name}")
        recs = sweep(mats, rows, cols, GRIDS[name], reps=reps, log=log)
        best = recs[0]
        table["workloads"][name] = {
            "best": best["config"], "ms": best["ms"], "tflops": best["tflops"],
            "log2_terms": mean_log2_terms(rows),
            "units_per_cu": best["config"].get("total_units", 0) / cus,
            "all": [{"config": r["config"], "ms": r["ms"], "ms_passes": r["ms_passes"],
                     "units": r["units"]} for r in recs],
        }
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        json.dump(table, f, indent=1)
    log(f"wrote {out}")
    return table


def load_tuning(path: str = DEFAULT_TABLE) -> dict:
    """The tuned table, or {} when none has been written for this device."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--only", nargs="*", default=None)
    args = ap.parse_args()
    tune(args.out, args.reps, args.only, log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    main()
