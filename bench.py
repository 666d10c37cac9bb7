"""Benchmark of the MI355X permanent engine.

Headline workload (one "step" = one evaluation of the whole batch): 4096
permanents of uneven size, n ~ U{8..24}, each a row/column selection with
multiplicities 1..3 of a 48-mode Haar unitary, complex128, double-double
accumulation, evaluated by one launch per column-width bucket (the buckets
run concurrently on forked streams), either replayed from a hipGraph or
launched directly; `value` is the faster mode, both are reported.  Inputs are
resident in device memory before the timed region (planning and upload happen
once, outside).

Printed as ONE JSON line (rank 0):
  value           permanents/s of the batched engine (whole job, 1 GPU)
  plain           the same batch through the plain FP64 kernel, one launch per
                  permanent (speedup_vs_plain = value / plain rate)
  n32             one n = 32 permanent (2^31 Gray terms): time, % of FP64 peak
  fp32_products_batch  the same batch with FP32 row sums and products
  gradient        gradients/s of 256 permanents n ~ U{12..20} (d perm / dA)
  tensor          tensor permanent of 18 partially distinguishable photons
                  (detection probabilities): time and % of FP64 peak; the CPU
                  figure is a numpy restatement of the reference's loop at n = 9
  output_probabilities  a probability table (77 520 output patterns) in one call
  sampling        Clifford & Clifford boson samples/s (40 modes, 20 photons,
                  1024 shots in lockstep: one batched column-minor launch per
                  photon round), checked shot for shot against the reference's
                  own permanent_laplace driven sequentially on the host
  rel_err_max     vs the double-double oracle values committed with the inputs
  error_bound     the per-result error bounds the sweep kernels report, relative to
                  |perm|, and whether every result lies within its bound of the oracle
  roofline        the sweep kernels: algorithmic FLOP / kernel time vs 78.6 TFLOP/s;
                  traffic = HBM bytes per evaluation from the committed rocprofv3
                  FETCH_SIZE + WRITE_SIZE summary in profiles/
  cpu_baseline    the reference's own C++ (oracle/_ref) on the host cores, a
                  bounded sample of the same batch

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak, AMD specification
# Untimed evaluations of the batch, for about this many seconds, before its W
# warm-up steps: the device's clock rises under sustained load (≈2.0 GHz
# after idle, ≈2.3 GHz after some tens of milliseconds of sweeps), and the
# CPU baseline that runs first leaves the device idle for seconds.  With this
# settle phase the K timed steps measure the sustained rate whatever W is.
CLOCK_SETTLE_S = 0.3
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_FILE = os.path.join(PROFILES, "hbm_traffic_batched_uneven.txt")
METRIC = "permanents/s batched n\u2208[8,24] and n=32 %FP64 peak, vs plain, rel.err\u22641e-10"
GOLDEN = os.path.join(ROOT, "tests", "golden")
SAMPLING_SHOTS = 1024
SAMPLING_SEED = 1234
SAMPLING_CPU_SECONDS = 8.0
TENSOR_N = 18
GRADIENT_CPU_SECONDS = 5.0
PROB_CPU_SECONDS = 4.0
TENSOR_CPU_N = 9
# lossy partially distinguishable tables: (modes, photons); the first is
# also run by the numpy port of the reference loop on LOSSY_PD_CPU_SUBSETS
# of its 63 particle subsets
LOSSY_PD_CONFIGS = ((7, 6), (10, 10))
LOSSY_PD_CPU_SUBSETS = 8


def _load_batched():
    from piquasso_amd import workloads as W

    fx = np.load(os.path.join(GOLDEN, "batched_uneven_seed4.npz"))
    mats, rows, cols = W.batched_uneven(seed=4, unitary=fx["unitary"])
    dd = fx["dd"]
    ref = (dd[:, 0] + dd[:, 1]) + 1j * (dd[:, 2] + dd[:, 3])
    return mats, rows, cols, ref


def _load_extended(cid):
    with open(os.path.join(GOLDEN, "extended_vectors.json")) as f:
        cases = {c["id"]: c for c in json.load(f)["cases"]}
    c = cases[cid]
    m = c["matrix"]
    a = (np.asarray(m["re"]) + 1j * np.asarray(m["im"])).reshape(m["shape"])
    dd = c["dd"]
    return a, np.asarray(c["rows"]), np.asarray(c["cols"]), complex(dd[0] + dd[1], dd[2] + dd[3])


def _load_n32():
    return _load_extended("large_single_n32")


# the single-permanent configurations, timed through the drop-in entry
SINGLE_CONFIGS = {
    "plain_single_n12": "n = 12 block of a 24-mode Haar unitary, multiplicities 1",
    "boson_sampling_m40_n20": "40-mode Haar unitary, 20 photons in modes 0..19, output pattern "
                              "with multiplicities up to 4 (probability |perm|^2 / prod r! c!)",
    "large_single_n20": "n = 20 i.i.d. N(0, 1/n) complex128",
}


def _median_call_us(fn, reps):
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)) * 1e6


# ------------------------------------------------------------ CPU leg --

def cpu_baseline_child(seconds: float) -> None:
    """Runs in a child process (before the parent touches the GPU): time the
    reference's own permanent (oracle/_ref, built from its sources) on a
    the batch, cycling over its items in order, until `seconds` of wall time
    are used."""
    from oracle import oracle as O
    from piquasso_amd import workloads as W

    mats, rows, cols, _ = _load_batched()
    kind = "reference"
    try:
        fn = O.ref_module().permanent
    except O.OracleUnavailable:
        kind = "port"
        fn = lambda a, r, c: O.perm_port(a, r, c)  # noqa: E731
    done = 0
    flops = 0.0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        k = done % len(mats)
        fn(mats[k], rows[k], cols[k])
        flops += W.glynn_flops(rows[k], cols[k])
        done += 1
    dt = time.perf_counter() - t0
    singles = {}
    for cid in SINGLE_CONFIGS:
        a, r, c, _ = _load_extended(cid)
        reps = 200 if a.shape[0] <= 12 else 5
        singles[cid] = _median_call_us(lambda: fn(a, r, c), reps)
    print(json.dumps({"kind": kind, "items": done, "seconds": dt, "flops": flops,
                      "threads": int(os.environ.get("OMP_THREAD_LIMIT", "0") or 0),
                      "singles_us": singles, "sampling": _cpu_sampling(SAMPLING_CPU_SECONDS),
                      "tensor": _cpu_tensor(), "gradient": _cpu_gradient(GRADIENT_CPU_SECONDS),
                      "laplace_us": _cpu_laplace(), "probabilities": _cpu_probabilities(PROB_CPU_SECONDS),
                      "lossy_pd": _cpu_lossy_pd()}))


def _cpu_lossy_pd():
    """The reference's lossy partially distinguishable probability loop
    (numpy restatement, oracle/lossy_pd_port.py) on the first subsets of the
    7-mode, 6-photon table: the box polynomial of (6 + 1)^7 coefficients per
    subset, so the full table costs (63 / subsets) times the sample."""
    from oracle.lossy_pd_port import lossy_pd_port
    from piquasso_amd import workloads as W

    d, n = LOSSY_PD_CONFIGS[0]
    t, inp, g, outs = W.lossy_pd(d, n)
    t0 = time.perf_counter()
    lossy_pd_port(outs, t, inp, g, subsets=LOSSY_PD_CPU_SUBSETS)
    return {"modes": d, "photons": n, "subsets": LOSSY_PD_CPU_SUBSETS, "of": (1 << n) - 1,
            "seconds": time.perf_counter() - t0}


def _cpu_probabilities(seconds: float):
    """The reference's way to a probability table: its pybind permanent()
    once per output pattern (piquasso/_simulators/passive/probabilities.py:26-54),
    patterns in order until `seconds` are used."""
    from oracle import oracle as O
    from piquasso_amd import workloads as W

    u, inp, outs = W.output_patterns()
    try:
        fn, kind = O.ref_module().permanent, "reference"
    except O.OracleUnavailable:
        fn, kind = O.perm_port, "port"
    u = np.ascontiguousarray(u)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and done < len(outs):
        fn(u, outs[done], inp)
        done += 1
    return {"kind": kind, "patterns": done, "seconds": time.perf_counter() - t0}


def _cpu_gradient(seconds: float):
    """The reference's grad_perm (oracle/_ref) on the gradient workload's
    items in order until `seconds` are used."""
    from oracle import oracle as O
    from piquasso_amd import workloads as W

    mats, rows, cols = W.gradient_batch(count=256, seed=5)
    try:
        O.ref_lib()
        fn, kind = O.ref_grad, "reference"
    except O.OracleUnavailable:
        fn, kind = O.grad_dd, "port"
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and done < len(mats):
        fn(mats[done], rows[done], cols[done])
        done += 1
    return {"kind": kind, "items": done, "seconds": time.perf_counter() - t0}


def _laplace_config():
    """Column minors of the boson-sampling configuration's last photon: the
    output pattern with one photon removed (19) against the 20 inputs."""
    a, r, c, _ = _load_extended("boson_sampling_m40_n20")
    r = np.array(r, dtype=np.int32)
    r[int(np.flatnonzero(r)[0])] -= 1
    return a, r, np.array(c, dtype=np.int32)


def _cpu_laplace():
    from oracle import oracle as O

    a, r, c = _laplace_config()
    try:
        O.ref_lib()
        return _median_call_us(lambda: O.ref_laplace(a, r, c), 5)
    except O.OracleUnavailable:
        return None


def _cpu_tensor():
    """The reference's tensor-permanent loop (numpy restatement) at n = 9."""
    from oracle.tensor_port import tensor_perm_port
    from piquasso_amd import workloads as W

    a = W.tensor(TENSOR_CPU_N)
    t0 = time.perf_counter()
    tensor_perm_port(a, 0.5)
    dt = time.perf_counter() - t0
    return {"n": TENSOR_CPU_N, "seconds": dt, "pairs": 4.0 ** TENSOR_CPU_N}


def _cpu_sampling(seconds: float):
    """The sampling workload on the host: the reference's permanent_laplace
    (oracle/_ref) called shot after shot by the oracle's restatement of the
    reference's sampler loop, shots 0, 1, ... until `seconds` are used."""
    from oracle import oracle as O
    from oracle import sampling as S
    from piquasso_amd import workloads as W

    u, occ = W.sampling()
    try:
        O.ref_lib()
        lap, kind = O.ref_laplace, "reference"
    except O.OracleUnavailable:
        lap, kind = O.laplace_dd, "port"
    samples = []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        samples += S.generate_samples(occ, 1, u, SAMPLING_SEED + len(samples), laplace=lap)
    return {"kind": kind, "shots": len(samples), "seconds": time.perf_counter() - t0,
            "samples": samples}


def run_cpu_baseline(seconds: float):
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, cores or 1)
    # every host core this process may run on, unless the host's CPU share
    # for this GPU is set through OMP_NUM_THREADS (16 threads per GPU on the
    # MI355X hosts, whatever the affinity mask shows)
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(cores, int(share)) if share.isdigit() and int(share) > 0 else cores
    env = dict(os.environ)
    # the reference asks for 4 x hardware_concurrency OpenMP threads per call
    # (src/permanent.cpp:144-153); cap them at the threads above
    env["OMP_THREAD_LIMIT"] = str(threads)
    env["OMP_NUM_THREADS"] = str(threads)
    try:
        out = subprocess.run(
            [sys.executable, os.path.abspath(__file__), "--cpu-child", str(seconds)],
            env=env, capture_output=True, text=True, timeout=seconds * 4 + 120,
        )
        rec = json.loads(out.stdout.strip().splitlines()[-1])
    except Exception as e:  # the GPU numbers are still worth reporting
        return {"value": None, "unit": "permanents/s", "cores": threads, "nproc": cores,
                "kind": "reference", "sample": f"failed: {e}"}
    rate = rec["items"] / rec["seconds"] if rec["seconds"] > 0 else None
    return {
        "value": rate,
        "unit": "permanents/s",
        "cores": threads,
        "nproc": cores,
        "omp_num_threads": share or None,
        "kind": rec["kind"],
        "sample": (f"{rec['items']} permanents of the batched workload, in order, "
                   f"cycling over its 4096 items ({rec['flops']:.3e} algorithmic FLOP) in "
                   f"{rec['seconds']:.1f} s, reference pybind permanent(), "
                   f"OMP_THREAD_LIMIT={threads}"),
        "gflops": rec["flops"] / rec["seconds"] / 1e9 if rec["seconds"] > 0 else None,
        "singles_us": rec.get("singles_us"),
        "sampling": rec.get("sampling"),
        "tensor": rec.get("tensor"),
        "gradient": rec.get("gradient"),
        "laplace_us": rec.get("laplace_us"),
        "probabilities": rec.get("probabilities"),
        "lossy_pd": rec.get("lossy_pd"),
    }


# ------------------------------------------------------------ GPU leg --

def _traffic_from_profile(key: str = "traffic_bytes_per_launch"):
    """HBM bytes per evaluation of the batch from the committed rocprofv3
    summary, or None when it is absent.  traffic_bytes_per_launch: reads
    counted by request size (TCC_EA0_RDREQ, _32B, _128B) plus WRITE_SIZE;
    traffic_bytes_per_launch_copy_scaled: FETCH_SIZE x 2 + WRITE_SIZE x 1,
    the factors a copy kernel's counts give."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.split("=")[0].strip() == key:
                    return float(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


def _gradient_leg(pa, W, reps: int, cpu=None, steps: int = 20):
    """Gradients of 256 permanents, n ~ U{12..20}, unit multiplicities, Haar
    48-mode submatrices: one Gray sweep per permanent gives the permanent and
    all n^2 entries of its gradient, in one hipGraph.  Euler's identity
    sum_ij A_ij G_ij = n perm(A) checks every gradient (relative to
    sum_ij |A_ij G_ij|).  TFLOP/s by the model F = sum T (8 m_c + 18 (n - 1)
    + 4) per gradient (row update, exclusion products, accumulation, one row
    flush and the weight per Gray term)."""
    mats, rows, cols = W.gradient_batch(count=256, seed=5)
    with pa.PermanentBatch(mats, rows, cols, gradient=True) as gb:
        gb.run("graph")
        gb.sync()
        perms = gb.results()
        grads = gb.gradients()
        # both launch modes of the same sweeps, timed alike after untimed
        # evaluations of their own (the clock rises under sustained load); the
        # faster is quoted
        # (device time per evaluation: one pair of HIP events around `steps`
        # back-to-back evaluations, as for the batched permanents; the
        # evaluations timed one at a time are kept beside it)
        ms_by, iso_by = {}, {}
        for m in ("graph", "eager"):
            for _ in range(max(3, reps)):
                gb.run(m)
            gb.sync()
            ms_by[m] = gb.time_span(m, steps) / steps
            iso_by[m] = float(np.median(gb.time(m, reps=reps)))
        st = gb.stats()
    launch = min(ms_by, key=ms_by.get)
    ms = ms_by[launch]
    resid = 0.0
    for a, r, p, g in zip(mats, rows, perms, grads):
        n = int(np.sum(r))
        resid = max(resid, abs(np.sum(a * g) - n * p) / np.sum(np.abs(a * g)))
    med = float(np.median(ms))
    cg = (cpu or {}).get("gradient")
    ref_cpu = None
    if cg and cg.get("items"):
        ref_cpu = {"gradients_per_s": cg["items"] / cg["seconds"], "kind": cg["kind"],
                   "cores": (cpu or {}).get("cores"),
                   "sample": f"the first {cg['items']} items of the 256, in order, "
                             f"{cg['seconds']:.1f} s, grad_perm (n^2 permanents each)"}
    return {
        "reference_cpu": ref_cpu,
        "gradients_per_s": len(mats) / (med * 1e-3),
        "ms": med,
        "tflops": st["flops"] / (med * 1e-3) / 1e12,
        "euler_residual_max": float(resid),
        "fp64_peak_frac": st["flops"] / (med * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        "flop_model": "sum over items of T * (8 m_c + 18 (n - 1) + 4), T = Gray terms",
        "launch": launch,
        "ms_by_launch": {m: float(v) for m, v in ms_by.items()},
        "ms_isolated_by_launch": iso_by,
        "what": "256 gradients d perm/dA (n~U{12..20}, 48-mode Haar submatrices): one Gray sweep "
                "per permanent for the permanent and all n^2 gradient entries; hipGraph replay and "
                "direct launches timed alike (HIP events around back-to-back evaluations), the "
                "faster quoted",
    }


def _tensor_leg(pa, W, cpu):
    """One tensor permanent of n = 18 (x = 0.5): 4^17 / 2 sign-vector pairs,
    timed around the whole call (planning and transfers are microseconds)."""
    a = W.tensor(TENSOR_N)
    pa.tensor_permanent(a, 0.5)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        (v,), (asum,) = pa.tensor_permanent_batch([a], [0.5], with_abs_sum=True)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    fl = W.tensor_flops(TENSOR_N)
    rec = {
        "n": TENSOR_N,
        "ms": t * 1e3,
        "tflops": fl / t / 1e12,
        "frac_fp64_peak": fl / t / 1e12 / FP64_PEAK_TFLOPS,
        "value": v,
        "error_bound": 8 * (TENSOR_N + 2) * np.finfo(float).eps * asum,
        "what": "tensor permanent (uniform overlap x = 0.5) of an 18 x 18 complex Gaussian matrix, "
                "sign-vector form, 4^17 / 2 pairs",
    }
    ct = (cpu or {}).get("tensor")
    if ct:
        rec["reference_cpu"] = {
            "pairs_per_s": ct["pairs"] / ct["seconds"],
            "kind": "port",
            "sample": f"numpy restatement of the reference's subset-pair loop, n = {ct['n']} "
                      f"({ct['pairs']:.0f} pairs) in {ct['seconds']:.2f} s, one thread",
        }
        rec["gpu_pairs_per_s"] = 4.0 ** (TENSOR_N - 1) / 2 / t
    return rec


def _probabilities_leg(pa, W, cpu):
    """A probability table: every output pattern of 7 photons in 14 modes
    (77 520 permanents, n = 7) in one batched call, timed around the whole
    call (planning, upload, kernels, download); the probabilities of a
    unitary must sum to 1."""
    u, inp, outs = W.output_patterns()
    pa.output_probabilities(u, inp, outs[:1000])  # warm-up
    t0 = time.perf_counter()
    p = pa.output_probabilities(u, inp, outs)
    t = time.perf_counter() - t0
    rec = {
        "patterns": len(outs),
        "seconds": t,
        "probabilities_per_s": len(outs) / t,
        "sum_minus_one": float(p.sum() - 1.0),
        "what": "output_probabilities: all 77520 output patterns of 7 photons in 14 modes of a Haar "
                "unitary, one batched call",
    }
    cp = (cpu or {}).get("probabilities")
    if cp and cp.get("patterns"):
        rec["reference_cpu"] = {
            "probabilities_per_s": cp["patterns"] / cp["seconds"],
            "kind": cp["kind"],
            "sample": f"the reference's permanent() per pattern, first {cp['patterns']} patterns in "
                      f"{cp['seconds']:.1f} s",
        }
    return rec


def _lossy_pd_leg(pa, W, cpu):
    """Lossy, partially distinguishable probability tables (every output with
    at most n photons; random Gram matrix; losses before a Haar unitary):
    the matrix-polynomial permanent over all particle subsets in one call,
    timed around the whole call (upload, kernels, download).  `ms` is the
    median over new Gram matrices for the same output set (the monomial plan
    is reused), `ms_first_call` a call that plans the closure afresh.
    FLOP: the kernel's algorithmic count per subset (complex multiply 6, FMA
    8 per link) times the subsets."""
    from piquasso_amd import distinguishability as D

    recs = {}
    for d, n in LOSSY_PD_CONFIGS:
        t, inp, g, outs = W.lossy_pd(d, n)
        b0, bms, _ = D._lossy_matrices(t, inp, g)
        # a first call with another output set, so the timed first call plans afresh
        D.perm_poly_coefficients(b0, bms, outs[:1])
        t0 = time.perf_counter()
        pa.lossy_detection_probabilities(outs, t, inp, g)
        cold = time.perf_counter() - t0
        _, info = D.perm_poly_coefficients(b0, bms, outs, with_info=True)
        ts = []
        for k in range(3):
            # new overlaps each time: the same output set, so the plan is reused
            gk = W.lossy_pd(d, n, seed=20 + k)[2]
            t0 = time.perf_counter()
            p = pa.lossy_detection_probabilities(outs, t, inp, gk)
            ts.append(time.perf_counter() - t0)
        sec = float(np.median(ts))
        fl = float(info["flops_per_subset"]) * info["subsets"]
        rec = {
            "outputs": len(outs),
            "ms": sec * 1e3,
            "ms_first_call": cold * 1e3,
            "tflops": fl / sec / 1e12,
            "flops": fl,
            "monomials": info["monomials"],
            "polynomial_in": "LDS" if info["lds"] else "global memory",
            "workgroups": info["workgroups"],
            "sum_minus_one": float(p.sum() - 1.0),
            "what": f"lossy_detection_probabilities: all outputs with <= {n} photons in {d} modes, "
                    f"{n} photons, random {n} x {n} Gram matrix",
        }
        c = (cpu or {}).get("lossy_pd")
        if c and (c["modes"], c["photons"]) == (d, n):
            full = c["seconds"] * c["of"] / c["subsets"]
            rec["reference_cpu"] = {
                "seconds_full_table_est": full,
                "kind": "port",
                "sample": f"numpy restatement of the reference's subset loop (box polynomial), "
                          f"{c['subsets']} of {c['of']} subsets in {c['seconds']:.2f} s, one thread",
            }
            rec["speedup_vs_reference_cpu"] = full / sec
        recs[f"modes{d}_photons{n}"] = rec
    return recs


def _sampling_leg(pa, W, cpu):
    """Clifford & Clifford sampling of 20 photons in 40 modes, 1024 shots in
    lockstep (20 rounds, each one batched column-minor launch over all
    shots); samples/s over the whole call, host bookkeeping included.  The
    shots the CPU leg also drew must come out identical."""
    from piquasso_amd import sampling as PS

    u, occ = W.sampling()
    # warm-up at full size (device and pinned buffers grow to the round size)
    pa.generate_samples(occ, SAMPLING_SHOTS, u, seed_sequence=SAMPLING_SEED + SAMPLING_SHOTS)
    t0 = time.perf_counter()
    samples = pa.generate_samples(occ, SAMPLING_SHOTS, u, seed_sequence=SAMPLING_SEED)
    wall = time.perf_counter() - t0
    st = dict(PS.last_run)
    # a second run counts the minors' FLOP (host bookkeeping kept out of the timed run)
    PS.count_flops = True
    pa.generate_samples(occ, SAMPLING_SHOTS, u, seed_sequence=SAMPLING_SEED)
    PS.count_flops = False
    st["minors_flops"] = PS.last_run["minors_flops"]
    rec = {
        "samples_per_s": SAMPLING_SHOTS / wall,
        "seconds": wall,
        "shots": SAMPLING_SHOTS,
        "photons": int(occ.sum()),
        "modes": len(occ),
        "rounds": st["rounds"],
        "minor_problems": st["problems"],
        "minors_seconds": st["minors_seconds"],
        "minors_tflops": st["minors_flops"] / st["minors_seconds"] / 1e12 if st["minors_seconds"] else None,
        "what": "boson sampling (Clifford & Clifford B), 40-mode Haar unitary, one photon in each of "
                "modes 0..19, shots in lockstep: one pqp_laplace_batch_c128 call per photon round",
    }
    cs = (cpu or {}).get("sampling")
    if cs and cs.get("shots"):
        k = cs["shots"]
        rec["reference_cpu"] = {
            "samples_per_s": k / cs["seconds"],
            "kind": cs["kind"],
            "cores": cpu.get("cores"),
            "sample": f"shots 0..{k - 1} one after another in {cs['seconds']:.1f} s",
        }
        rec["shots_identical_to_reference"] = [tuple(x) for x in cs["samples"]] == samples[:k]
        rec["shots_compared"] = k
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 evaluations of the batch take ~40 ms; 20 untimed ones before each
    # launch mode bring the device to its sustained clock first
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-child", type=float, default=None)
    ap.add_argument("--profile", choices=["batch", "n32", "grad", "sampling", "tensor", "lossy_pd",
                                          "single12", "exact32"],
                    default=None,
                    help="profiling runs: only launch this workload's kernels --steps times")
    ap.add_argument("--replay", choices=["eager", "graph"], default="eager",
                    help="profiling runs: direct launches or hipGraph replays")
    ap.add_argument("--copy-calibration", type=int, default=None, metavar="MIB",
                    help="profiling runs: one device-to-device copy of MIB MiB (the runtime's "
                         "copy kernel) and one elementwise kernel reading and writing MIB MiB "
                         "with ordinary cached loads and stores (torch add), to calibrate "
                         "FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()

    if args.cpu_child is not None:
        cpu_baseline_child(args.cpu_child)
        return 0
    if args.copy_calibration is not None:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        torch.add(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n, "add_bytes_read": 4 * n, "add_bytes_written": 4 * n}))
        return 0
    if args.profile is not None:
        import piquasso_amd as pa

        if args.profile == "tensor":
            from piquasso_amd import workloads as W

            a = W.tensor(TENSOR_N)
            for _ in range(args.steps):
                pa.tensor_permanent(a, 0.5)
            print(json.dumps({"profile": "tensor", "n": TENSOR_N}))
            return 0
        if args.profile == "lossy_pd":
            from piquasso_amd import workloads as W

            d, n = LOSSY_PD_CONFIGS[1]
            t, inp, g, outs = W.lossy_pd(d, n)
            for _ in range(args.steps):
                pa.lossy_detection_probabilities(outs, t, inp, g)
            print(json.dumps({"profile": "lossy_pd", "modes": d, "photons": n}))
            return 0
        if args.profile == "single12":
            # back-to-back drop-in calls of the n = 12 single configuration
            a, r, c, _ = _load_extended("plain_single_n12")
            for _ in range(args.steps):
                pa.permanent(a, r, c)
            print(json.dumps({"profile": "single12", "calls": args.steps}))
            return 0
        if args.profile == "exact32":
            # the exact integer permanent of the 32 x 32 all-ones matrix (32!)
            one = np.ones(32, dtype=np.int64)
            vals = [pa.permanent_exact(np.ones((32, 32), dtype=np.int64), one, one)
                    for _ in range(args.steps)]
            print(json.dumps({"profile": "exact32", "exact": all(v == math.factorial(32) for v in vals)}))
            return 0
        if args.profile == "sampling":
            from piquasso_amd import workloads as W

            u, occ = W.sampling()
            for _ in range(args.steps):
                pa.generate_samples(occ, SAMPLING_SHOTS, u, seed_sequence=SAMPLING_SEED)
            print(json.dumps({"profile": "sampling", "last_run": pa.sampling.last_run}))
            return 0
        if args.profile == "batch":
            mats, rows, cols, _ = _load_batched()
            b = pa.PermanentBatch(mats, rows, cols)
        elif args.profile == "grad":
            from piquasso_amd import workloads as W

            mats, rows, cols = W.gradient_batch(count=256, seed=5)
            b = pa.PermanentBatch(mats, rows, cols, gradient=True)
        else:
            a32, r32, c32, _ = _load_n32()
            from piquasso_amd import _capi
            from piquasso_amd.tune import best_config

            _, cus, arch = _capi.device_info()
            cfg32 = best_config(1, cus, arch, log2_terms=31.0) or {"persistent": False, "total_units": 128 * cus}
            b = pa.PermanentBatch([a32], [r32], [c32], **cfg32)
        # --warmup untimed evaluations first (traced too: the summaries
        # average the last --steps dispatches), at the sustained clock
        for _ in range(args.warmup):
            b.run(args.replay)
        b.sync()
        ms = b.time(args.replay, reps=args.steps)
        print(json.dumps({"profile": args.profile, "warmup": args.warmup, "ms": ms.tolist()}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "here (every measurement is of the single-GPU engine)", file=sys.stderr)
        return 2

    cpu = None if args.skip_cpu else run_cpu_baseline(args.cpu_seconds)

    import piquasso_amd as pa
    from piquasso_amd import _capi, workloads as W

    n_dev, cus, arch = _capi.device_info()
    mats, rows, cols, ref = _load_batched()
    flops_batch = sum(W.glynn_flops(r, c) for r, c in zip(rows, cols))

    batch = pa.PermanentBatch(mats, rows, cols)
    settle_n, t_settle = 0, time.perf_counter()
    while time.perf_counter() - t_settle < CLOCK_SETTLE_S:
        for _ in range(10):
            batch.run("eager")
        batch.sync()
        settle_n += 10
    settle_s = time.perf_counter() - t_settle
    # Both launch modes of the product, timed the same way (W untimed, then
    # K back-to-back evaluations between two synchronisations): the
    # hipGraph replay of the forked bucket launches, and the same launches
    # issued directly.  `value` is the faster; both are reported.
    # The K evaluations are issued by pqp_batch_time_span between one pair
    # of HIP events on the batch stream, which also gives the device time of
    # the timed region (the roofline's time per evaluation).
    walls, spans, rels = {}, {}, {}
    for mode in ("graph", "eager"):
        for _ in range(args.warmup):
            batch.run(mode)
        batch.sync()
        t0 = time.perf_counter()
        spans[mode] = batch.time_span(mode, args.steps)
        batch.sync()
        walls[mode] = time.perf_counter() - t0
        rels[mode] = np.abs(batch.results() - ref) / np.abs(ref)
    # per-result error bounds of the last evaluation (term magnitudes summed
    # by the static sweep kernels) against the oracle's distance
    bounds = batch.error_bounds()
    abs_err = np.abs(batch.results() - ref)
    error_bound = {
        "rel_max": float(np.max(bounds / np.abs(ref))),
        "rel_median": float(np.median(bounds / np.abs(ref))),
        "all_hold": bool(np.all(abs_err <= bounds)),
        "what": "4 (n+2) u sum|terms| / 2^(n-1) + final rounding per result, "
                "from the sweep kernels' term magnitudes; all_hold: |result - oracle| <= bound",
    }
    launch = min(walls, key=walls.get)
    wall = walls[launch]
    rel = np.maximum(rels["graph"], rels["eager"])
    step_ms = wall / args.steps * 1e3
    value = len(mats) * args.steps / wall
    graph_ms = batch.time("graph", reps=max(3, min(args.steps, 10)))
    kernel_ms = batch.time("kernel", reps=max(3, min(args.steps, 10)))
    plain_ms = batch.time("plain", reps=2)
    stats = batch.stats()
    # device time per evaluation over the timed region (HIP events on the
    # batch stream around the K evaluations of the faster launch mode)
    kernel_avg = spans[launch] / args.steps
    # the FLOP the kernels execute: each item's sweep in the orientation the
    # planner ran (perm(A; r, c) or perm(A^T; c, r), whichever is cheaper);
    # flops_batch is the same count with every item swept as given, as the
    # reference sweeps it
    flops_exec = float(stats["flops"])
    achieved = flops_exec / (kernel_avg * 1e-3) / 1e12

    # single large permanent, n = 32 (2^31 Gray terms)
    a32, r32, c32, ref32 = _load_n32()
    # one large permanent: the autotuned launch for this device, else 128
    # work units per CU with one block per unit (the fastest setting
    # measured for n = 32 on MI355X)
    from piquasso_amd.tune import best_config

    cfg32 = best_config(1, cus, arch, log2_terms=31.0) or {"persistent": False, "total_units": 128 * cus}
    b32 = pa.PermanentBatch([a32], [r32], [c32], **cfg32)
    b32.run("eager")
    v32 = b32.results()[0]
    t32 = b32.time("kernel", reps=5)
    f32 = W.glynn_flops(r32, c32)
    t32_ms = float(np.median(t32))
    _, _, bound32 = pa.permanent_ex(a32, r32, c32, with_bound=True)
    n32 = {
        "ms": t32_ms,
        "tflops": f32 / (t32_ms * 1e-3) / 1e12,
        "frac_fp64_peak": f32 / (t32_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        "rel_err": abs(v32 - ref32) / abs(ref32),
        "rel_error_bound": bound32 / abs(ref32),
        "flops": f32,
    }

    # the same batch with FP32 row sums and products (double-double accumulation)
    b32b = pa.PermanentBatch(mats, rows, cols, precision="fp32")
    b32b.run("graph")
    b32b.sync()
    v32b = b32b.results()
    bounds32b = b32b.error_bounds()
    g32 = b32b.time("graph", reps=max(3, min(args.steps, 10)))
    k32 = b32b.time("kernel", reps=max(3, min(args.steps, 10)))
    flops32b = float(b32b.stats()["flops"])
    b32b.close()
    rel32 = np.abs(v32b - ref) / np.abs(ref)
    fp32_batch = {
        "permanents_per_s": len(mats) / (float(np.median(g32)) * 1e-3),
        "graph_ms": float(np.median(g32)),
        "kernel_ms": float(np.median(k32)),
        "rel_err_max": float(rel32.max()),
        "rel_err_median": float(np.median(rel32)),
        "error_bound_rel_max": float(np.max(bounds32b / np.abs(ref))),
        "error_bounds_hold": bool(np.all(np.abs(v32b - ref) <= bounds32b)),
        "tflops": flops32b / (float(np.median(k32)) * 1e-3) / 1e12,
        "frac_fp32_peak": flops32b / (float(np.median(k32)) * 1e-3) / 1e12 / 157.3,
        "what": "the batched workload with FP32 row sums and products (FP64 + double-double "
                "accumulation), hipGraph replay; MI355X FP32 vector peak 157.3 TFLOP/s assumes "
                "packed FP32, which this build does not emit",
    }

    grad = _gradient_leg(pa, W, reps=max(3, min(args.steps, 10)), cpu=cpu, steps=max(args.steps, 1))
    sampling = _sampling_leg(pa, W, cpu)
    tensor = _tensor_leg(pa, W, cpu)
    probabilities = _probabilities_leg(pa, W, cpu)
    lossy_pd = _lossy_pd_leg(pa, W, cpu)

    # single permanents through the drop-in entry (host buffers in, result
    # out: planning, upload, sweep, download); the reference's CPU time for
    # the same call comes from the CPU leg
    singles = {}
    for cid, what in SINGLE_CONFIGS.items():
        a, r, c, ref_c = _load_extended(cid)
        reps = 200 if a.shape[0] <= 12 else 20
        got = complex(pa.permanent(a, r, c))
        rec_s = {"what": what, "gpu_call_us": _median_call_us(lambda: pa.permanent(a, r, c), reps),
                 "rel_err": abs(got - ref_c) / abs(ref_c)}
        if cid == "boson_sampling_m40_n20":
            norm = float(np.prod([math.factorial(int(k)) for k in r]) *
                         np.prod([math.factorial(int(k)) for k in c]))
            rec_s["probability"] = abs(got) ** 2 / norm
        if cid == "large_single_n20":
            v32, _ = pa.permanent_ex(a, r, c, mode="fp32")
            rec_s["fp32_products"] = {
                "gpu_call_us": _median_call_us(lambda: pa.permanent_ex(a, r, c, mode="fp32"), reps),
                "rel_err": abs(v32 - ref_c) / abs(ref_c)}
        if cpu and cpu.get("singles_us"):
            rec_s["reference_cpu_us"] = cpu["singles_us"].get(cid)
        singles[cid] = rec_s
    # column minors (permanent_laplace) of the boson-sampling configuration
    la, lr, lc = _laplace_config()
    from oracle import oracle as O

    lap = pa.permanent_laplace(la, lr, lc)
    lap_ref = O.laplace_dd(la, lr, lc)
    singles["laplace_m40_n20"] = {
        "what": "permanent_laplace: the boson-sampling configuration with one output photon "
                "removed (19 rows, 20 input columns): all 20 column minors",
        "gpu_call_us": _median_call_us(lambda: pa.permanent_laplace(la, lr, lc), 20),
        "rel_err": float(np.abs(lap - lap_ref).max() / np.abs(lap_ref).max()),
        "reference_cpu_us": (cpu or {}).get("laplace_us"),
    }

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "permanents/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "clock_settle": {"evaluations": settle_n, "seconds": round(settle_s, 3),
                         "what": "untimed evaluations of the batch before the W warm-up steps, "
                                 "so that the timed steps run at the sustained clock"},
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 48-mode Haar unitary committed with the oracle values)",
        "config": {
            "workload": "batched_uneven: 4096 permanents, n~U{8..24}, row/col multiplicities "
                        "1..3 of a 48-mode Haar unitary, complex128, dd accumulation, "
                        "one launch per column-width bucket (concurrent streams), the faster of "
                        "hipGraph replay and direct launches",
            "items": len(mats),
            "n_range": [8, 24],
            "modes": 48,
            "seed": 4,
            "device": f"{arch} ({cus} CUs)",
        },
        "rel_err_max": float(rel.max()),
        "error_bound": error_bound,
        "launch": launch,
        "permanents_per_s_by_launch": {m: len(mats) * args.steps / w for m, w in walls.items()},
        "graph_ms_median": float(np.median(graph_ms)),
        "plain": {
            "permanents_per_s": len(mats) / (float(np.median(plain_ms)) * 1e-3),
            "ms": float(np.median(plain_ms)),
            "what": "plain FP64 kernel, one launch + finalize per permanent, same device inputs",
        },
        "speedup_vs_plain": float(np.median(plain_ms)) / float(np.median(graph_ms)),
        "n32": n32,
        "fp32_products_batch": fp32_batch,
        "gradient": grad,
        "sampling": sampling,
        "tensor_permanent": tensor,
        "lossy_pd_probabilities": lossy_pd,
        "output_probabilities": probabilities,
        "single_permanents": singles,
        "roofline": {
            "bound": "fp64-valu",
            "pipe": "FP64 vector ALU (no MFMA: the sweep is a product of row sums, not a "
                    "contraction); the peak is MI355X's FP64 vector peak, 78.6 TFLOP/s",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": _traffic_from_profile(),
            "traffic_copy_scaled": _traffic_from_profile("traffic_bytes_per_launch_copy_scaled"),
            "traffic_method": "reads counted by request size (TCC_EA0_RDREQ, _32B, _128B: the sweep "
                              "kernels issue 64-byte requests only) + WRITE_SIZE; FETCH_SIZE counts "
                              "128-byte requests as 64 bytes on gfx950, so the copy-kernel factor 2 "
                              "applies to 128-byte streams only (traffic_copy_scaled)",
            "kernel_ms": kernel_avg,
            "kernel_ms_what": "device time per evaluation over the timed region: HIP events on the "
                              "batch stream around its K back-to-back evaluations (every kernel of an "
                              "evaluation: the bucket sweeps and the finalize), faster launch mode",
            "sweep_ms_isolated": float(np.median(kernel_ms)),
            "flops_per_launch": flops_exec,
            "flops_per_launch_as_given": flops_batch,
            "flop_model": "sum over items of T * (2 m_c + 6 (n - 1) + 4), T = Gray terms, "
                          "m_c = product-side columns, of the sweep the planner runs "
                          "(as given or transposed); _as_given: every item swept as given",
            "units": stats["units"],
        },
        "cpu_baseline": ({k: v for k, v in cpu.items()
                          if k not in ("sampling", "tensor", "gradient", "laplace_us", "probabilities",
                                       "lossy_pd")}
                         if cpu else cpu),
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
