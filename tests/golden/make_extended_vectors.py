"""Write tests/golden/extended_vectors.json: extended-precision (double-double)
oracle values for the measured workloads that are too large to recompute
inside a test, together with their inputs.

Inputs come from piquasso_amd.workloads (seeded); the values from the
double-double oracle (oracle/glynn_dd.c), which tests/test_oracle.py pins
against exact rational brute force, Ryser's formula and the reference's own
vectors.  The n = 32 case takes a few minutes on 8 cores.

With --reference the script instead adds, to the cases already in the file,
the output of the reference's own pybind ``permanent()`` built from its
sources into oracle/_ref (``make -C oracle ref``), and the number of host
threads it ran with: its chunking, hence its last bits, depends on that
count (src/permanent.cpp:144-165).  The GPU tests compare against both
values.

Run:  python tests/golden/make_extended_vectors.py [--skip-32]
      python tests/golden/make_extended_vectors.py --reference
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from piquasso_amd import workloads as W  # noqa: E402


def enc(a):
    a = np.asarray(a, dtype=np.complex128)
    return {"shape": list(a.shape), "re": a.real.ravel().tolist(), "im": a.imag.ravel().tolist()}


def entry(name, a, rows, cols, note):
    t0 = time.time()
    parts = O.perm_dd_parts(a, rows, cols)
    dt = time.time() - t0
    print(f"{name}: {parts[0] + parts[1]:+.17e} {parts[2] + parts[3]:+.17e}i  ({dt:.1f} s)", flush=True)
    return {
        "id": name,
        "note": note,
        "matrix": enc(a),
        "rows": [int(x) for x in rows],
        "cols": [int(x) for x in cols],
        "dd": [float(x) for x in parts],
        "oracle_seconds": dt,
    }


def add_reference_outputs(path):
    """Record the reference build's permanent() beside each stored dd value."""
    mod = O.ref_module()
    with open(path) as f:
        data = json.load(f)
    threads = os.cpu_count()
    for case in data["cases"]:
        m = case["matrix"]
        a = (np.asarray(m["re"]) + 1j * np.asarray(m["im"])).reshape(m["shape"])
        a = np.ascontiguousarray(a, dtype=np.complex128)
        rows = np.asarray(case["rows"], dtype=np.int64)
        cols = np.asarray(case["cols"], dtype=np.int64)
        t0 = time.time()
        val = complex(mod.permanent(a, rows=rows, cols=cols))
        dt = time.time() - t0
        dd = case["dd"]
        exact = complex(dd[0] + dd[1], dd[2] + dd[3])
        err = abs(val - exact) / abs(exact)
        print(f"{case['id']}: reference {val:.17e}  |ref - dd| / |dd| = {err:.2e}  ({dt:.1f} s)",
              flush=True)
        case["reference"] = {"re": val.real, "im": val.imag, "host_threads": threads,
                             "seconds": dt, "rel_err_vs_dd": err}
    data["reference"] = ("the reference's pybind permanent() compiled from its sources "
                         "(oracle/Makefile ref), run in the build container")
    with open(path, "w") as f:
        json.dump(data, f)
    print("wrote", path)


def main():
    if "--reference" in sys.argv:
        add_reference_outputs(os.path.join(HERE, "extended_vectors.json"))
        return
    skip32 = "--skip-32" in sys.argv
    cases = []
    a, r, c = W.plain_single(seed=1)
    cases.append(entry("plain_single_n12", a, r, c, "workloads.plain_single(seed=1)"))
    u, r, c = W.boson_sampling(seed=2)
    cases.append(entry("boson_sampling_m40_n20", u, r, c, "workloads.boson_sampling(seed=2)"))
    for n in (20, 24, 28):
        a, r, c = W.large_single(n=n, seed=3)
        cases.append(entry(f"large_single_n{n}", a, r, c, f"workloads.large_single(n={n}, seed=3)"))
    if not skip32:
        a, r, c = W.large_single(n=32, seed=3)
        cases.append(entry("large_single_n32", a, r, c, "workloads.large_single(n=32, seed=3)"))
    path = os.path.join(HERE, "extended_vectors.json")
    if skip32 and os.path.exists(path):
        with open(path) as f:
            old = {x["id"]: x for x in json.load(f)["cases"]}
        if "large_single_n32" in old:
            cases.append(old["large_single_n32"])
    with open(path, "w") as f:
        json.dump({"oracle": "double-double Glynn (oracle/glynn_dd.c)", "cases": cases}, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
