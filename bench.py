"""Benchmark: Schrodinger-VM sampling throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one sampling pass over the workload's shots, with the bit-packed
records written to device memory (inputs and outputs resident in HBM).
Prints one JSON line.

Default workload (batch4096): 4096 random Clifford+T programs (16-40
qubits, peak rank 4-16) x 1000 shots, run as a batch -- one persistent
launch per launch group, the groups side by side -- against one plain launch
per program on the same stream ("speedup_vs_per_program").  The other
workloads time one program's sampling launch (surface_d3, clifford_t20,
large_k34, clifford_t24_k16), the per-program launches alone (programs4096)
or the adjoint gradient (rz24_grad).

roofline: HBM bytes of the dominant kernel's launch / that kernel's duration
from HIP events on the stream it runs on, against the 8 TB/s HBM peak.  The
bytes (`traffic`) are the ones measured with rocprofv3 FETCH_SIZE /
WRITE_SIZE counters for that kernel and workload (profiles/hbm_traffic.json,
scaled by a copy-kernel calibration; "achieved_from": "PMC traffic").  The
per-shot byte model (16 B per amplitude read or written by each array opcode
at its static rank, plus the record bytes) rides along as
algorithmic_bytes_per_launch / model_GBps; fused passes move fewer bytes
than it charges ("effective": model above the peak).  Without a measurement
for the kernel, the model stands in ("achieved_from": "byte model").
cpu_baseline: the C oracle built with -O3 -ffast-math -fopenmp, single PRNG
stream, shots in sequence (the reference's scheduling), on a bounded sample;
the shot-parallel OpenMP figure on all host cores rides along as "mode_b".
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (description, shots per step)
    "surface_d3": ("noisy d=3 rotated surface-code memory, 3 rounds, p=1e-3, 1 T/round (17 qubits)", 1_000_000),
    "clifford_t20": ("20 qubits x 40 layers random Clifford+T, 6 T, final M (peak k=6)", 100_000),
    "large_k34": ("34 qubits x 40 layers, 26 T on distinct qubits, MR on 4 qubits / 10 layers", 256),
    "clifford_t24_k16": ("24 qubits x 40 layers random Clifford+T, 16 T, final M (peak k=16)", 2000),
    "clifford_t24_k8": ("24 qubits x 40 layers random Clifford+T, 8 T on distinct qubits, final M (peak k=8)", 20000),
    "clifford_t24_k10": ("24 qubits x 40 layers random Clifford+T, 10 T on distinct qubits, final M (peak k=10)",
                         20000),
    "clifford_t24_k13": ("24 qubits x 40 layers random Clifford+T, 13 T, final M (peak k=13)", 20000),
    "rz24_grad": ("24 qubits, 30 Clifford layers, 16 R_Z(theta_i) slots (peak k=16), 32 random Pauli probes "
                  "with nonzero expectation; "
                  "<O> and d<O>/dtheta for 512 theta vectors by the adjoint kernel", 512),
    "programs4096": ("the 4096 programs of batch4096, 1000 shots each, one plain launch per program", 1000),
    "batch4096": ("4096 random Clifford+T programs, n ~ U[16,40], T-count ~ U[4,16], 40 layers, final M; "
                  "one persistent launch per launch group", 1000),
}


def build_program(name):
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as ccompile
    if name == "surface_d3":
        return ccompile(W.surface_code_d3())
    if name == "clifford_t20":
        return ccompile(W.random_clifford_t())
    if name == "large_k34":
        return ccompile(W.large_k())
    if name == "clifford_t24_k16":
        return ccompile(W.random_clifford_t(n=24, layers=40, t_count=16, seed=42))
    if name in ("clifford_t24_k8", "clifford_t24_k10", "clifford_t24_k13"):
        t = int(name.rsplit("k", 1)[1])
        # T on distinct qubits reaches the peak for 8 and 10 T; 13 T reach it as drawn
        return ccompile(W.random_clifford_t(n=24, layers=40, t_count=t, seed=42, t_distinct_qubits=t != 13))
    raise SystemExit(f"unknown workload {name}")


def measured_peaks():
    """LDS (read + write) and HBM copy bandwidth measured on this device with
    the probes in clifft_amd/libclifft_probe.so (GB/s), or {} if unbuilt."""
    import ctypes
    path = os.path.join(ROOT, "clifft_amd", "libclifft_probe.so")
    if not os.path.exists(path):
        return {}
    L = ctypes.CDLL(path)
    L.clifft_probe_lds.restype = ctypes.c_double
    L.clifft_probe_lds.argtypes = [ctypes.c_int, ctypes.c_int]
    L.clifft_probe_stream.restype = ctypes.c_double
    L.clifft_probe_stream.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    # 4 GiB streams, 16384 workgroups of 256 threads, 8 loads in flight per lane
    return {"lds_probe_GBps": L.clifft_probe_lds(256 * 8, 200),
            "hbm_read_probe_GBps": L.clifft_probe_stream(0, 4 << 30, 16384, 5),
            "hbm_write_probe_GBps": L.clifft_probe_stream(1, 4 << 30, 16384, 5),
            "hbm_copy_probe_GBps": L.clifft_probe_stream(2, 4 << 30, 16384, 5)}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _num(x):
    """A finite float, or None (JSON has no NaN)."""
    return float(x) if np.isfinite(x) else None


def _timed(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(prog, run_seconds=2.5, max_shots=20_000_000):
    """The C oracle built like the reference Release build (-O3 -ffast-math
    -fopenmp), timed on this host on a bounded sample of the same program:
    mode A = one PRNG stream, shots in sequence (the reference's scheduling,
    src/clifft/svm/svm.cc:191-255), median of 5 runs after a warm-up;
    mode B = OpenMP over shots with per-shot seeds on every host core,
    median of 3.  About 20 s of CPU time in all."""
    from oracle import oracle
    probe = 100 if prog.peak_rank < 20 else 1
    while True:
        t0 = time.perf_counter()
        oracle.sample(prog, probe, seed=7, per_shot=False, fast=True)
        dt = time.perf_counter() - t0
        if dt > 0.25 or prog.peak_rank >= 20:
            break
        probe *= 4
    per_shot = max(dt / probe, 1e-9)
    n_a = int(min(max(run_seconds / per_shot, 1), max_shots))
    oracle.sample(prog, min(n_a, 1000), seed=7, per_shot=False, fast=True)
    t_a = _timed(lambda: oracle.sample(prog, n_a, seed=7, per_shot=False, fast=True), 5)
    threads = oracle.lib(True).orc_max_threads()
    n_b = int(min(max(run_seconds / per_shot * threads * 0.6, threads), max_shots))
    mem_b = n_b * (prog.num_measurements + prog.num_detectors + prog.num_observables + 8 * prog.num_exp_vals)
    if mem_b > (8 << 30):
        n_b = max(threads, int(n_b * (8 << 30) / mem_b))
    t_b = _timed(lambda: oracle.sample_parallel(prog, n_b, seed=7, threads=threads, fast=True), 3)
    host = f"{os.cpu_count()} logical CPUs, {_cpu_model()}"
    return {"value": n_a / t_a, "unit": "shots/s", "cores": 1, "kind": "port",
            "sample": f"mode A: {n_a} shots of the same program, one stream, median of 5 = {t_a:.2f} s; "
                      f"host {host}",
            "mode_b": {"value": n_b / t_b, "unit": "shots/s", "cores": threads,
                       "sample": f"{n_b} shots, OpenMP over shots, per-shot seeds, median of 3 = {t_b:.2f} s"}}


def run_grad(args):
    """Config: 512 theta vectors of the 24-qubit / 16-parameter / 32-probe
    circuit; <O> and its gradient by one forward + one backward sweep per
    vector, against GPU parameter-shift (2P+1 forward evaluations per vector
    with the same kernel in forward-only mode)."""
    import torch
    from clifft_amd import _lib, grad
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as ccompile

    desc, B = WORKLOADS["rz24_grad"]
    if args.shots:
        B = args.shots
    build, coeffs, _ = W.rz_exp_val()
    rng = np.random.default_rng(42)
    thetas = rng.uniform(0, 2 * np.pi, (B, 16))
    prog = ccompile(build(thetas[0]), keep_rotations=True)
    adj = grad.Adjoint(prog)
    P, nev = adj.num_params, adj.num_exp_vals
    dev = torch.device("cuda:0")
    th = torch.tensor(thetas, dtype=torch.float64, device=dev)
    c = torch.tensor(coeffs, dtype=torch.float64, device=dev)
    ev = torch.empty((B, nev), dtype=torch.float64, device=dev)
    g = torch.empty((B, P), dtype=torch.float64, device=dev)
    # parameter-shift inputs: for each vector, theta and theta +- pi/2 e_i
    shift = np.repeat(thetas[:, None, :], 2 * P + 1, axis=1)
    for i in range(P):
        shift[:, 1 + 2 * i, i] += np.pi / 2
        shift[:, 2 + 2 * i, i] -= np.pi / 2
    th_ps = torch.tensor(shift.reshape(-1, P), dtype=torch.float64, device=dev)
    ev_ps = torch.empty((th_ps.shape[0], nev), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    def adjoint_step():
        adj.run_device(th.data_ptr(), B, c.data_ptr(), ev.data_ptr(), g.data_ptr(), stream=sh)

    def shift_step():
        adj.run_device(th_ps.data_ptr(), th_ps.shape[0], None, ev_ps.data_ptr(), None, stream=sh)

    def timed(fn, steps):
        with torch.cuda.stream(stream):
            for _ in range(args.warmup):
                fn()
            torch.cuda.synchronize()
            s0 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
            s1 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
            t0 = time.perf_counter()
            for i in range(steps):
                s0[i].record(stream)
                fn()
                s1[i].record(stream)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        return wall, float(np.mean([a.elapsed_time(b) for a, b in zip(s0, s1)]))

    wall, kern_ms = timed(adjoint_step, args.steps)
    wall_ps, kern_ps = timed(shift_step, max(1, args.steps // 4))
    steps_ps = max(1, args.steps // 4)
    F = (ev @ c).cpu().numpy()
    G = g.cpu().numpy()
    Fps = (ev_ps @ c).cpu().numpy().reshape(B, 2 * P + 1)
    Gps = 0.5 * (Fps[:, 1::2] - Fps[:, 2::2])
    grads_per_s = B * args.steps / wall
    ps_per_s = B * steps_ps / wall_ps
    # algorithmic bytes per theta vector: forward state traffic, lambda
    # construction (read phi + read/write lambda per probe), backward sweeps
    # over phi and lambda, and one inner product per parameter
    info = adj.dp.info()
    fwd = info["bytes_per_shot"]
    kfin = 16 * (1 << prog.peak_rank)
    alg_vec = fwd + 48 * (kfin // 16) * nev + 2 * (fwd - 16 * (kfin // 16) * nev) + 32 * (kfin // 16) * P
    achieved = alg_vec * B / (kern_ms * 1e-3) / 1e9
    line = {
        "metric": "grads/s", "value": grads_per_s, "unit": "gradients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "rz24_grad", "description": desc, "theta_vectors": B, "params": P, "probes": nev,
                   "peak_rank": prog.peak_rank, "instructions": int(len(prog.bytecode)),
                   "param_shift_grads_per_s": ps_per_s, "speedup_vs_param_shift": grads_per_s / ps_per_s,
                   "max_abs_diff_vs_param_shift": float(np.abs(G - Gps).max()),
                   "max_abs_diff_value": float(np.abs(F - Fps[:, 0]).max()),
                   "param_shift_kernel_ms": kern_ps, "parallelism": "theta vectors"},
        "roofline": _roofline(achieved, None, kern_ms, alg_vec * B),
        "backend": _lib.backend(),
    }
    print(json.dumps(line), flush=True)


def _roofline(model_gbs, traffic, kernel_ms, model_bytes, **extra):
    """HBM roofline of the dominant kernel.  `achieved` and `frac` come from
    the PMC-measured HBM bytes of that kernel (profiles/hbm_traffic.json)
    when they were profiled for it; the byte model (one sweep per opcode,
    an upper bound: fused passes move fewer bytes) is kept as
    `model_GBps` and stands in, flagged, only without a measurement."""
    measured = traffic is not None
    achieved = traffic / (kernel_ms * 1e-3) / 1e9 if measured else model_gbs
    r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
         "traffic": traffic, "achieved_from": "PMC traffic" if measured else "byte model",
         "kernel_ms": kernel_ms, "algorithmic_bytes_per_launch": model_bytes, "model_GBps": model_gbs,
         "effective": model_gbs > HBM_PEAK_GBS}
    r.update(extra)
    return r


def _measure_batch(progs, shots, seeds0, steps, warmup, tune_shots, batch_only=False, fixed_tuning=None):
    """The batch (autotuned on a `tune_shots`-per-program sample) against one
    plain launch per program, on the same stream and outputs.  Returns
    timings in seconds and the per-group kernel times of one eager run with
    the groups in sequence (side by side, a group's events would also count
    the time it waits for the others).  batch_only: default tuning, no
    per-program launches and no graph replays (a short run to profile)."""
    import ctypes

    import torch
    from clifft_amd import _lib, svm
    from clifft_amd import batch as B
    count = len(progs)
    dev = torch.device("cuda:0")
    nb = lambda n: (n + 7) // 8
    bufs, outs = [], []
    for p in progs:
        m = torch.empty((shots, nb(p.num_measurements)), dtype=torch.uint8, device=dev)
        bufs.append(m)
        outs.append({"measurements": m.data_ptr()})
    handles = [svm.device_program(p) for p in progs]
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    bt = B.Batch(progs, [shots] * count, outs, packed=True)
    t_tune = time.perf_counter()
    if batch_only and fixed_tuning:
        bt.set_tuning(**fixed_tuning)
    tuning = bt.tuning if batch_only else bt.tune(sample_shots=tune_shots, stream=sh)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t_tune
    groups = bt.groups()
    recs = (_lib.Records * count)(*[_lib.Records(o["measurements"], None, None, None, None) for o in outs])
    harr = (ctypes.c_void_p * count)(*[h.handle.value for h in handles])
    shots_arr = np.full(count, shots, dtype=np.uint64)
    flags_pp = _lib.REC_DEVICE | _lib.REC_PACKED | _lib.BATCH_PER_PROGRAM

    def per_program(seeds):
        _lib.check(_lib.lib().clifft_svm_batch_sample(harr, shots_arr.ctypes.data, seeds.ctypes.data, count,
                                                      flags_pp, recs, sh), "per-program")

    t_pp = float("nan")
    if not batch_only:
        per_program(seeds0)
        torch.cuda.synchronize()
        reps_pp = max(1, min(3, steps))
        t0 = time.perf_counter()
        for r in range(reps_pp):
            per_program(seeds0 + np.uint64(r + 1))
        torch.cuda.synchronize()
        t_pp = (time.perf_counter() - t0) / reps_pp
    # per-group kernel times: timing events around each group's launch, the
    # groups in sequence for this one run
    bt.set_tuning(concurrent=0)
    bt.time_groups(True)
    bt.run(seeds0 + np.uint64(99), stream=sh)
    torch.cuda.synchronize()
    group_ms = bt.group_times()
    bt.time_groups(False)
    bt.set_tuning(**tuning)
    for i in range(warmup):
        bt.run(seeds0 + np.uint64(100 + i), stream=sh)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    with torch.cuda.stream(stream):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            starts[i].record(stream)
            bt.run(seeds0 + np.uint64(1000 + i), stream=sh)
            ends[i].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        wall_graph = float("nan")
        if not batch_only:
            bt.run_graph(seeds0, stream=sh)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for i in range(steps):
                bt.run_graph(seeds0 + np.uint64(2000 + i), stream=sh)
            torch.cuda.synchronize()
            wall_graph = time.perf_counter() - t0
    step_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    per_prog_bytes = [(h.info()["bytes_per_shot"] + nb(p.num_measurements)) * shots for h, p in zip(handles, progs)]
    bt.close()
    return {"t_batch": wall / steps, "t_graph": wall_graph / steps, "t_per_program": t_pp, "step_ms": step_ms,
            "per_prog_bytes": per_prog_bytes, "tuning": tuning, "groups": groups, "group_ms": group_ms,
            "tune_s": t_tune}


def batch_cpu_baseline(progs, shots_total, run_seconds=12.0):
    """The C oracle (-O3 -ffast-math -fopenmp) on a bounded sample of the
    batch: every 64th program (the same peak-rank mix), with a per-program
    shot count sized to about `run_seconds` of single-core work; mode A = one
    stream per program, shots in sequence (the reference's scheduling), mode
    B = OpenMP over shots on every host core."""
    from oracle import oracle
    sample = progs[::64]
    t0 = time.perf_counter()
    for i, p in enumerate(sample):
        oracle.sample(p, 4, seed=7 + i, per_shot=False, fast=True)
    per4 = time.perf_counter() - t0
    n = int(max(4, min(shots_total, 4 * run_seconds / max(per4, 1e-9))))
    t0 = time.perf_counter()
    for i, p in enumerate(sample):
        oracle.sample(p, n, seed=7 + i, per_shot=False, fast=True)
    t_a = time.perf_counter() - t0
    threads = oracle.lib(True).orc_max_threads()
    nb_ = max(threads, n * max(1, threads // 2))
    t0 = time.perf_counter()
    for i, p in enumerate(sample):
        oracle.sample_parallel(p, nb_, seed=7 + i, threads=threads, fast=True)
    t_b = time.perf_counter() - t0
    host = f"{os.cpu_count()} logical CPUs, {_cpu_model()}"
    return {"value": len(sample) * n / t_a, "unit": "shots/s", "cores": 1, "kind": "port",
            "sample": f"mode A: every 64th program of the batch ({len(sample)} programs, the same peak-rank mix), "
                      f"{n} shots each, one stream per program, {t_a:.2f} s; host {host}",
            "mode_b": {"value": len(sample) * nb_ / t_b, "unit": "shots/s", "cores": threads,
                       "sample": f"{nb_} shots per sampled program, OpenMP over shots, {t_b:.2f} s"}}


def run_programs(args):
    """Config: the batch4096 programs (4096 x 1000 shots) run with one plain
    launch per program, back to back on one stream, records in device
    memory."""
    import ctypes

    import torch
    from clifft_amd import _lib, svm
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as ccompile

    desc, shots = WORKLOADS["programs4096"]
    count = int(os.environ.get("CLIFFT_BENCH_BATCH_COUNT", "4096"))
    if args.shots:
        shots = args.shots
    progs = [ccompile(text) for text, _ in W.batch_programs(count=count, seed=42, shots=shots)]
    seeds0 = np.array([42 ^ i for i in range(count)], dtype=np.uint64)
    nb = lambda n: (n + 7) // 8
    dev = torch.device("cuda:0")
    bufs = [torch.empty((shots, nb(p.num_measurements)), dtype=torch.uint8, device=dev) for p in progs]
    handles = [svm.device_program(p) for p in progs]
    recs = (_lib.Records * count)(*[_lib.Records(b.data_ptr(), None, None, None, None) for b in bufs])
    harr = (ctypes.c_void_p * count)(*[h.handle.value for h in handles])
    shots_arr = np.full(count, shots, dtype=np.uint64)
    flags = _lib.REC_DEVICE | _lib.REC_PACKED | _lib.BATCH_PER_PROGRAM
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    def step(i):
        _lib.check(_lib.lib().clifft_svm_batch_sample(harr, shots_arr.ctypes.data, (seeds0 + np.uint64(i)).ctypes.data,
                                                      count, flags, recs, sh), "per-program launches")

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    alg = sum((h.info()["bytes_per_shot"] + nb(p.num_measurements)) * shots for h, p in zip(handles, progs))
    achieved = alg / wall / 1e9
    peaks = np.bincount([p.peak_rank for p in progs], minlength=17)
    line = {
        "metric": "shots/s", "value": shots * count / wall, "unit": "shots/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "programs4096", "description": desc, "programs": count, "shots_per_program": shots,
                   "peak_rank_histogram": peaks.tolist(), "records": "bit-packed measurements, device memory",
                   "parallelism": "shots within each program"},
        "roofline": _roofline(achieved, None, wall * 1e3, alg, note="time is wall time of all launches of one step"),
        "backend": _lib.backend(),
    }
    print(json.dumps(line), flush=True)


def run_batch(args):
    """Config: 4096 programs x 1000 shots (16-40 qubits, T-count = peak rank
    4-16, 40 layers, final M on every qubit) as one batch: one persistent
    launch per launch group, groups side by side, autotuned; against one
    plain launch per program on the same stream and outputs."""
    from clifft_amd import _lib
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as ccompile

    desc, shots = WORKLOADS["batch4096"]
    count = int(os.environ.get("CLIFFT_BENCH_BATCH_COUNT", "4096"))
    if args.shots:
        shots = args.shots
    progs = [ccompile(text) for text, _ in W.batch_programs(count=count, seed=42, shots=shots)]
    seeds0 = np.array([42 ^ i for i in range(count)], dtype=np.uint64)
    r = _measure_batch(progs, shots, seeds0, args.steps, args.warmup, tune_shots=max(1, shots // 8),
                       batch_only=args.batch_only,
                       fixed_tuning=json.loads(args.batch_tuning) if args.batch_tuning else None)
    total = shots * count
    value = total / r["t_batch"]
    # dominant kernel: the launch group with the longest kernel time
    gi = int(np.argmax(r["group_ms"]))
    g = r["groups"][gi]
    dom_bytes = g["algorithmic_bytes"]
    dom_ms = r["group_ms"][gi]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        # profiled per group membership (the tuning moves programs between groups)
        w = tr["workloads"].get(f"batch4096_group_{g['kind']}_{g['programs']}")
        if w and w["shots"] == shots and w.get("programs") == count and w.get("group_programs") == g["programs"]:
            traffic = (w["fetch_kib"] * tr["fetch_scale"] + w["write_kib"] * tr["write_scale"]) * 1024.0
    peaks = np.bincount([p.peak_rank for p in progs], minlength=17)
    line = {
        "metric": "shots/s", "value": value, "unit": "shots/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["t_batch"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batch4096", "description": desc, "programs": count, "shots_per_program": shots,
                   "peak_rank_histogram": peaks.tolist(),
                   "mean_instructions": float(np.mean([len(p.bytecode) for p in progs])),
                   "tuning": r["tuning"], "tune_seconds": r["tune_s"], "launch_groups": r["groups"],
                   "group_kernel_ms": r["group_ms"], "records": "bit-packed measurements, device memory",
                   "per_program_launches_shots_per_s": _num(total / r["t_per_program"]),
                   "speedup_vs_per_program": _num(r["t_per_program"] / r["t_batch"]),
                   "graph_replay_shots_per_s": _num(total / r["t_graph"]),
                   "speedup_graph_vs_per_program": _num(r["t_per_program"] / r["t_graph"]),
                   "parallelism": "programs x shots"},
        "roofline": _roofline(achieved, traffic, dom_ms, dom_bytes,
                              kernel=f"launch group {gi} ({g['kind']}, {g['programs']} programs, block {g['block']})"),
        "backend": _lib.backend(),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = batch_cpu_baseline(progs, shots)
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batch4096", choices=sorted(WORKLOADS))
    ap.add_argument("--shots", type=int, default=0, help="override shots per step")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="batch4096: default tuning, no per-program launches or graph replays (to profile)")
    ap.add_argument("--batch-tuning", default="", metavar="JSON",
                    help="with --batch-only: tuning fields over the default, e.g. '{\"wave_rank_max\": 9}'")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this executor runs on one GPU; sharding shots over several GPUs is "
                                   "not implemented", "n_gpus": args.gpus}))
        raise SystemExit(2)

    import torch
    from clifft_amd import _lib, svm

    assert torch.cuda.is_available(), "needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_calibrate:
        n = args.copy_calibrate << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": args.steps}), flush=True)
        return
    if args.workload == "batch4096":
        return run_batch(args)
    if args.workload == "programs4096":
        return run_programs(args)
    if args.workload == "rz24_grad":
        return run_grad(args)
    desc, shots = WORKLOADS[args.workload]
    if args.shots:
        shots = args.shots
    prog = build_program(args.workload)
    dp = svm.device_program(prog)
    info = dp.info()

    nb = lambda n: (n + 7) // 8
    dev = torch.device("cuda:0")
    meas = torch.empty((shots, nb(prog.num_measurements)), dtype=torch.uint8, device=dev)
    det = torch.empty((shots, nb(prog.num_detectors)), dtype=torch.uint8, device=dev)
    obs = torch.empty((shots, nb(prog.num_observables)), dtype=torch.uint8, device=dev)
    out = {"measurements": meas.data_ptr() if meas.numel() else None,
           "detectors": det.data_ptr() if det.numel() else None,
           "observables": obs.data_ptr() if obs.numel() else None}
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    def step(i):
        svm.sample_device(prog, shots, seed=1000 + i, out=out, packed=True, stream=sh)

    with torch.cuda.stream(stream):
        for i in range(args.warmup):
            step(i)
        torch.cuda.synchronize()
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        t0 = time.perf_counter()
        for i in range(args.steps):
            starts[i].record(stream)
            step(args.warmup + i)
            ends[i].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    kern_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    value = shots * args.steps / wall

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        w = tr["workloads"].get(args.workload)
        if w and w["shots"] == shots:
            traffic = (w["fetch_kib"] * tr["fetch_scale"] + w["write_kib"] * tr["write_scale"]) * 1024.0
    rec_bytes = nb(prog.num_measurements) + nb(prog.num_detectors) + nb(prog.num_observables)
    bytes_per_launch = (info["bytes_per_shot"] + rec_bytes) * shots
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    line = {
        "metric": "shots/s", "value": value, "unit": "shots/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "shots_per_step": shots,
                   "num_qubits": prog.num_qubits, "peak_rank": prog.peak_rank,
                   "instructions": int(len(prog.bytecode)), "array_ops": int(info["num_array_ops"]),
                   "tier": ["team-lds", "team-global", "lane"][info["tier"]], "team_waves": info["team_waves"],
                   "shots_per_block": info["shots_per_block"], "records": "bit-packed, device memory",
                   "parallelism": "shots"},
        "roofline": _roofline(achieved, traffic, kern_ms, bytes_per_launch,
                              limiter=("HBM" if info["tier"] == 1 else
                                       "instruction issue: amplitudes stay in LDS/registers, HBM carries the records"),
                              measured_peaks=measured_peaks()),
        "backend": _lib.backend(),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(prog)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
