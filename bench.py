"""Benchmark of the Hermite-4 block step on one MI355X.

Workload (default): Plummer sphere N = 262144 (seeded, Henon units, equal
masses), eps = 1e-4, eta = 0.01.  One timed step is one fused block step
(h4h_step: predictor of all N, active selection, acc/jerk update, Hermite
corrector + Aarseth/normalize_dt timestep, next block time) in which all N
particles are active: the context runs with dt_min = dt_max = 2^-10, so the
block quantisation returns the shared step every time and every block step
is a full N x N step.  value = N*N / (seconds per step) pair interactions/s.

Also reported: the update kernel's fraction of the FP64 vector peak
(60 flops per pair, HIP events on the context's stream), the realistic
block-timestep schedule rate (first block steps from the same IC with the
reference's dt bounds 2^-23..2^-3), an active-set sweep of the update, the
mixed-precision path, the CPU oracle timed on the host cores over a
bounded sample of the same update, whole runs (N = 1024 to t = 1 beside
the oracle's run, N = 65536 to t = 0.5), the ensemble of 64 clusters to
t = 1 (the oracle over the same window for its smallest clusters), and the
N = 262144 and N = 1048577 (Salpeter + central mass) configurations: the
same first block steps on the oracle and on the GPU (FP64 and mixed), and
a GPU time window of each.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector, published
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector, published (packed)
FLOPS_PER_PAIR = 60        # convention for Hermite acc+jerk
METRIC = "pair interactions/s & % FP64 vector peak at N=256k; mixed-vs-FP64 speedup"


def measured_traffic(kernel, workload, precision):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a plain copy), or
    None when no such measurement is on file."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_plummer_n262144_traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or t.get("precision") != precision:
        return None
    return t["bytes_per_launch"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def full_step_rate(h4, n, steps, warmup):
    """Runs warmup + steps fused full block steps; returns (s/step, update
    ms/launch, launches)."""
    h4.init_acc_jrk()
    T = h4.init_dt()
    for _ in range(warmup):
        nact, T = h4.step(T)
        assert nact == n, (nact, n)
    h4.synchronize()
    h4.stats(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        nact, T = h4.step(T)
        assert nact == n, (nact, n)
    h4.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st = h4.stats(reset=True)
    upd = st["update_ms"] / max(1, st["update_launches"])
    return dt, upd, st


def block_schedule_rate(h4, n, steps, device_loop):
    """First `steps` block steps of the real schedule from the IC, either
    replayed on the device (h4h_run) or one h4h_step call per step."""
    h4.init_acc_jrk()
    T = h4.init_dt()
    h4.synchronize()
    h4.stats(reset=True)
    t0 = time.perf_counter()
    if device_loop:
        info, _, nacts = h4.run(0.0, T, float("inf"), float("inf"), steps)
        pairs = info["interactions"]
        T = info["atime"]
    else:
        pairs = 0
        nacts = []
        for _ in range(steps):
            nact, T = h4.step(T)
            pairs += nact * n
            nacts.append(nact)
    h4.synchronize()
    el = time.perf_counter() - t0
    return pairs / el, el, np.asarray(nacts), T


def nact_sweep(h4, n, nacts, reps=3):
    out = {}
    rng = np.random.default_rng(7)
    perm = rng.permutation(n).astype(np.uint32)
    for k in nacts:
        move = np.sort(perm[:k])
        h4.set_active(move)
        h4.update_acc_jrk(k)  # warm
        h4.stats(reset=True)
        for _ in range(reps):
            h4.update_acc_jrk(k)
        st = h4.stats(reset=True)
        ms = st["update_ms"] / max(1, st["update_launches"])
        out[str(k)] = {"pairs_per_s": k * n / (ms * 1e-3), "update_ms": ms,
                       "frac_fp64_peak": FLOPS_PER_PAIR * k * n / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
    return out


def cpu_baseline(r, v, e2, sample_nact, seed=7):
    """Oracle (CPU restatement of Hermite4CPU::update_acc_jrk) over the first
    sample_nact particles of a seeded permutation against all N."""
    from oracle import Oracle
    o = Oracle()
    n = r.shape[0]
    rng = np.random.default_rng(seed)
    move = np.sort(rng.permutation(n)[:sample_nact]).astype(np.uint32)
    p = np.concatenate([r[:, :3], v[:, :3]], axis=1)
    pm = r[:, 3].astype(np.float32)
    f = np.zeros((n, 6))
    t0 = time.perf_counter()
    o.update_acc_jrk(move, p, pm, e2, f)
    el = time.perf_counter() - t0
    cpu = cpu_model()
    return {"value": sample_nact * n / el, "unit": "pair interactions/s",
            "cores": o.num_threads(), "kind": "port",
            "sample": f"Hermite4CPU::update_acc_jrk restated (oracle, -O3 -fopenmp "
                      f"-ffp-contract=off), nact={sample_nact} seeded-random active particles "
                      f"x N={n} j, {el:.2f} s on {o.num_threads()} threads of {cpu}",
            "seconds": el}


def run_leg(e2, n, seed, t_end, precision, cpu=False, graphs=True):
    """A whole integration (device loop) of a seeded Plummer sphere: block
    steps, interactions and the rate; with cpu=True the oracle runs the same
    integration on the host cores and its time is reported beside it."""
    from gravidy_amd import Hermite4HIP, ic
    r, v = ic.plummer(n, seed)
    with Hermite4HIP(r, v, e2, precision=precision) as h:
        h.set_graphs(graphs)
        t0 = time.perf_counter()
        run = h.integration(t_end)
        el = time.perf_counter() - t0
        small = h.stats()["small_steps"]
    e = run["energy"]
    out = {"n": n, "seed": seed, "t_end": t_end, "precision": precision,
           "block_steps": int(len(run["step_time"])), "interactions": int(run["interactions"]),
           "seconds": el, "pairs_per_s": run["interactions"] / el,
           "cum_energy_error": float(abs((e[-1] - e[0]) / e[0])),
           "one_workgroup_steps": int(small)}
    if small:
        # the same run with every block step on the multi-workgroup kernels
        with Hermite4HIP(r, v, e2, precision=precision) as h:
            h.set_graphs(graphs)
            h.set_small_path(0)
            t0 = time.perf_counter()
            h.integration(t_end)
            out["seconds_multi_workgroup_only"] = time.perf_counter() - t0
    if cpu:
        from oracle import Oracle
        o = Oracle()
        t0 = time.perf_counter()
        ref = o.integrate(r, v, e2, t_end)
        cel = time.perf_counter() - t0
        out["cpu_baseline"] = {"seconds": cel, "pairs_per_s": ref["interactions"] / cel,
                               "cores": o.num_threads(), "kind": "port",
                               "block_steps": int(len(ref["step_time"])),
                               "same_schedule": bool(np.array_equal(ref["step_time"],
                                                                    run["step_time"]))}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def same_steps_leg(name, r, v, e2, dt_min, cpu_budget_s, t_window, interval):
    """One configuration over the same block steps on the CPU and the GPU,
    plus a longer GPU time window.

    The initial forces and steps come from the FP64 kernel (h4h_init_acc_jrk,
    h4h_init_dt); from that common state the oracle (Hermite4CPU's loop body,
    all host cores) runs block steps in chunks until cpu_budget_s is spent,
    K steps in all, and the GPU then runs exactly those K steps (device
    loop, FP64 and mixed, timed after an untimed pass over the same steps).
    Rates are sum(nact * N) / time over the same steps; same_schedule says
    whether the GPU's block times and active counts equal the oracle's.
    Then each precision integrates the configuration from t = 0 to t_window
    on its own (energies at t = 0, interval, ...), for a whole-run rate and
    the cumulative energy error."""
    from gravidy_amd import Hermite4HIP
    from oracle import Oracle
    n = r.shape[0]
    out = {"workload": name, "n": n, "dt_min": dt_min}
    with Hermite4HIP(r, v, e2, precision="fp64", dt_min=dt_min) as h:
        t0 = time.perf_counter()
        h.init_acc_jrk()
        atime0 = h.init_dt()
        s0 = h.download()
        out["init_seconds"] = time.perf_counter() - t0
    start = {k: np.ascontiguousarray(s0[k]) for k in ("r", "v", "f", "t", "dt")}

    o = Oracle()
    st = {k: start[k].copy() for k in start}
    st["a2"] = np.zeros((n, 3))
    st["a3"] = np.zeros((n, 3))
    times, nacts, cpu_pairs = [], [], 0
    at = atime0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < cpu_budget_s and len(times) < 100000:
        at, stt, snn, inter = o.block_steps(st, e2, at, 8, dt_min=dt_min)
        times.append(stt)
        nacts.append(snn)
        cpu_pairs += inter
    cel = time.perf_counter() - t0
    times, nacts = np.concatenate(times), np.concatenate(nacts)
    K = int(times.size)
    log(f"[bench] {name}: oracle {K} block steps in {cel:.2f} s on {o.num_threads()} threads")
    out["block_steps"] = K
    out["pairs"] = int(cpu_pairs)
    out["t_reached"] = float(times[-1])
    out["mean_nact"] = float(nacts.mean())
    out["median_nact"] = float(np.median(nacts))
    out["cpu_baseline"] = {"value": cpu_pairs / cel, "unit": "pair interactions/s",
                           "seconds": cel, "cores": o.num_threads(), "kind": "port",
                           "sample": f"oracle (Hermite4CPU loop body restated) over the first {K} "
                                     f"block steps from the FP64 initial forces, "
                                     f"{o.num_threads()} threads of {cpu_model()}"}
    for precision in ("fp64", "mixed"):
        with Hermite4HIP(r, v, e2, precision=precision, dt_min=dt_min) as h:
            h.upload(start["r"], start["v"], start["f"], start["t"], start["dt"])
            h.run(0.0, atime0, float("inf"), float("inf"), K, record=False)  # graph capture
            h.upload(start["r"], start["v"], start["f"], start["t"], start["dt"])
            h.synchronize()
            t0 = time.perf_counter()
            info, gst, gsn = h.run(0.0, atime0, float("inf"), float("inf"), K)
            el = time.perf_counter() - t0
        out[precision] = {"seconds": el, "pairs_per_s": info["interactions"] / el,
                          "us_per_step": el / K * 1e6,
                          "speedup_vs_cpu": (info["interactions"] / el) / (cpu_pairs / cel),
                          "same_schedule": bool(np.array_equal(gst, times)
                                                and np.array_equal(gsn, nacts))}
        log(f"[bench] {name} {precision}: {K} block steps in {el * 1e3:.1f} ms, "
            f"{info['interactions'] / el:.3e} pairs/s, same schedule "
            f"{out[precision]['same_schedule']}")
    out["mixed"]["speedup_vs_fp64"] = out["mixed"]["pairs_per_s"] / out["fp64"]["pairs_per_s"]

    win = {"t_end": t_window, "interval": interval}
    for precision in ("fp64", "mixed"):
        with Hermite4HIP(r, v, e2, precision=precision, dt_min=dt_min) as h:
            t0 = time.perf_counter()
            run = h.integration(t_window, interval=interval)
            el = time.perf_counter() - t0
        e = run["energy"]
        win[precision] = {"seconds": el, "block_steps": int(len(run["step_time"])),
                          "interactions": int(run["interactions"]),
                          "pairs_per_s": run["interactions"] / el,
                          "mean_nact": float(run["step_nact"].mean()),
                          "cum_energy_error": float(abs((e[-1] - e[0]) / e[0]))}
        log(f"[bench] {name} {precision} to t={t_window}: {len(run['step_time'])} block steps "
            f"in {el:.2f} s, {run['interactions'] / el:.3e} pairs/s")
    win["mixed_speedup"] = win["fp64"]["seconds"] / win["mixed"]["seconds"]
    out["window"] = win
    return out


def plummer_run(t_end, n=262144):
    """The large Plummer configuration as a whole run: n equal-mass
    Plummer stars (seed n), eps = 1e-4, dt in [2^-23, 2^-3], to t_end, FP64
    and mixed from identical initial conditions (device loop, energy every
    0.125)."""
    from gravidy_amd import Hermite4HIP, ic, softening_e2
    r, v = ic.plummer(n, n)
    e2 = softening_e2(1e-4)
    out = {"workload": f"plummer_n{n}_eps1e-4_eta0.01", "t_end": t_end}
    for precision in ("mixed", "fp64"):
        t0 = time.perf_counter()
        with Hermite4HIP(r, v, e2, precision=precision) as h:
            run = h.integration(t_end, interval=0.125,
                                on_output=lambda _h, t: log(f"[bench] N={n} {precision} t={t:.6f} "
                                                            f"{time.perf_counter() - t0:.1f} s"))
        el = time.perf_counter() - t0
        e = run["energy"]
        out[precision] = {"seconds": el, "block_steps": int(len(run["step_time"])),
                          "interactions": int(run["interactions"]),
                          "pairs_per_s": run["interactions"] / el,
                          "mean_nact": float(run["step_nact"].mean()),
                          "cum_energy_error": float(abs((e[-1] - e[0]) / e[0]))}
    out["mixed_speedup"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    return out


def salpeter_bh_run(t_end):
    """The large mixed-precision configuration as a whole run: 1048576
    Salpeter-mass Plummer stars plus a central particle of mass 0.01,
    eps = 1e-5, block steps down to 2^-30, to t_end, FP64 and mixed from
    identical initial conditions (device loop, outputs every 2^-8)."""
    from gravidy_amd import Hermite4HIP, ic, softening_e2
    r, v = ic.plummer_salpeter_bh(1048576, 1048576)
    e2 = softening_e2(1e-5)
    out = {"workload": "salpeter_plummer_bh_n1048577_eps1e-5_dtmin2^-30", "t_end": t_end}
    for precision in ("mixed", "fp64"):
        t0 = time.perf_counter()
        with Hermite4HIP(r, v, e2, precision=precision, dt_min=2.0 ** -30) as h:
            run = h.integration(t_end, interval=2.0 ** -8,
                                on_output=lambda _h, t: log(f"[bench] N=1048577 {precision} t={t:.6f} "
                                                            f"{time.perf_counter() - t0:.1f} s"))
        el = time.perf_counter() - t0
        e = run["energy"]
        out[precision] = {"seconds": el, "block_steps": int(len(run["step_time"])),
                          "interactions": int(run["interactions"]),
                          "pairs_per_s": run["interactions"] / el,
                          "mean_nact": float(run["step_nact"].mean()),
                          "cum_energy_error": float(abs((e[-1] - e[0]) / e[0]))}
    out["mixed_speedup"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    return out


def copy_probe(mib, reps=3):
    """Plain device-to-device copy of `mib` MiB (hipMemcpy, the runtime's
    copy kernel), the reference point that scales rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters to bytes.  Returns GB/s of the last copy."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = int(mib) << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMemset failed")
        hip.hipDeviceSynchronize()
        el = 0.0
        for _ in range(reps):
            t0 = time.perf_counter()
            if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # device to device
                raise RuntimeError("hipMemcpy failed")
            hip.hipDeviceSynchronize()
            el = time.perf_counter() - t0
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return {"copy_mib": int(mib), "bytes_read": nbytes, "bytes_written": nbytes,
            "gb_per_s": 2 * nbytes / el / 1e9}


def small_clusters_leg(e2, precision, count=256, seed=256, t_end=0.25, graphs=True):
    """An ensemble of many small Plummer clusters (N log-uniform in
    [128, 1024], seeded), every cluster to t_end: the clusters' block steps
    in workgroups of their own (the default) and, for comparison, every
    step on the multi-workgroup kernels."""
    from gravidy_amd import EnsembleHIP, ic
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(128), np.log(1024), count)).astype(int)
    systems = [ic.plummer(int(nk), 5000 + k) for k, nk in enumerate(sizes)]
    out = {"clusters": count, "particles": int(sizes.sum()), "n_range": [128, 1024],
           "seed": seed, "t_end": t_end, "precision": precision}
    for label, limit in (("", None), ("_multi_workgroup_only", 0)):
        with EnsembleHIP(systems, e2, precision=precision) as ens:
            ens.set_graphs(graphs)
            if limit is not None:
                ens.set_small_path(limit)
            t0 = time.perf_counter()
            res = ens.integration(t_end)
            el = time.perf_counter() - t0
            st = ens.stats()
        out["seconds" + label] = el
        if not label:
            out["cluster_steps"] = int(st["block_steps"])
            out["own_workgroup_steps"] = int(st["small_steps"])
            out["interactions"] = int(res["interactions"].sum())
            out["pairs_per_s"] = out["interactions"] / el
    return out


def ensemble_leg(e2, precision, count=64, seed=64, t_end=1.0, cpu_clusters=4, graphs=True,
                 tune_steps=64, launch=None):
    """64 independent Plummer clusters, N_k log-uniform in [1024, 32768]
    (seeded), each in its own Henon units, all integrated to t_end in one
    context (one launch sequence per block step, graph-replayed), with the
    launch (persistent workgroups, j split, lane-group rule) picked by
    h4h_tune_launch over `tune_steps` device steps from the initial state
    (timed separately)."""
    from gravidy_amd import EnsembleHIP, ic
    sizes = ic.log_uniform_sizes(count, seed)
    systems = [ic.plummer(int(n), seed * 1000 + k) for k, n in enumerate(sizes)]
    with EnsembleHIP(systems, e2, precision=precision) as ens:
        ens.set_graphs(graphs)
        choice, tune_s = None, 0.0
        if launch:
            ens.set_launch_choice(*launch)
            choice = {"workgroups": launch[0], "max_jsplit": launch[1], "iw_fill16": launch[2],
                      "given": True}
        elif tune_steps:
            t0 = time.perf_counter()
            ens.init()
            choice = ens.tune_launch(tune_steps)
            tune_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        res = ens.integration(t_end)
        el = time.perf_counter() - t0
    e = res["energy"]
    cum = np.abs((e[-1] - e[0]) / e[0])
    out = {"clusters": count, "seed": seed, "n_total": int(sizes.sum()),
           "n_min": int(sizes.min()), "n_max": int(sizes.max()), "t_end": t_end,
           "seconds": el, "interactions": int(res["interactions"].sum()),
           "pairs_per_s": float(res["interactions"].sum() / el),
           "block_steps_max": int(res["steps"].max()), "block_steps_total": int(res["steps"].sum()),
           "cum_energy_error_median": float(np.median(cum)), "cum_energy_error_max": float(cum.max()),
           "launch": choice, "tune_seconds": tune_s}
    if cpu_clusters:
        # The oracle over the same time window for the smallest clusters,
        # each cluster's run compared with the GPU's: block steps,
        # interactions and the cumulative energy error to 2 digits.
        from oracle import Oracle
        o = Oracle()
        order = np.argsort(sizes)[:cpu_clusters]
        t0 = time.perf_counter()
        pairs = 0
        checks = []
        for k in order:
            r, v = systems[k]
            ref = o.integrate(r, v, e2, t_end)
            pairs += ref["interactions"]
            eo = ref["energy"]
            cum_ref = abs((eo[-1] - eo[0]) / eo[0])
            checks.append({"cluster": int(k), "n": int(sizes[k]),
                           "steps": int(res["steps"][k]), "steps_cpu": int(len(ref["step_time"])),
                           "interactions_equal": int(res["interactions"][k]) == ref["interactions"],
                           "cum_energy_error": float(cum[k]), "cum_energy_error_cpu": cum_ref,
                           "cum_energy_error_2_digits": f"{cum[k]:.1e}" == f"{cum_ref:.1e}"})
        cel = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": pairs / cel, "unit": "pair interactions/s",
                               "cores": o.num_threads(), "kind": "port",
                               "sample": f"oracle integration of the {cpu_clusters} smallest "
                                         f"clusters (n={sorted(int(sizes[k]) for k in order)}) "
                                         f"to t={t_end}, the GPU's window, {cel:.2f} s",
                               "clusters": checks}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--seed", type=int, default=262144)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "exact", "mixed"])
    ap.add_argument("--block-steps", type=int, default=200,
                    help="real-schedule block steps to time (0 = skip)")
    ap.add_argument("--cpu-sample", type=int, default=131072,
                    help="active particles in the CPU baseline sample (0 = skip)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-t-end", type=float, default=1.0)
    ap.add_argument("--tune-steps", type=int, default=64,
                    help="device steps per candidate of the ensemble's launch autotuner (0 = off)")
    ap.add_argument("--ensemble-launch", default="", metavar="W,S,F",
                    help="ensemble launch (workgroups, largest j split, lane-group fill in 16ths) "
                         "instead of the autotuner")
    ap.add_argument("--no-graphs", action="store_true",
                    help="device loops launch kernels directly (for kernel tracers)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-runs", action="store_true",
                    help="skip the whole-run legs (N=1024 to t=1 with the CPU oracle beside it, "
                         "N=65536 to t=0.5)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the N=262144 and N=1048577 configuration legs (same block steps "
                         "on CPU and GPU, FP64 and mixed, and a GPU time window)")
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of oracle block steps per configuration leg")
    ap.add_argument("--salpeter-bh-run", type=float, default=0.0, metavar="T_END",
                    help="only run the N=1048577 Salpeter + central-mass configuration to T_END "
                         "in mixed and FP64 (whole runs, minutes)")
    ap.add_argument("--plummer-run", type=float, default=0.0, metavar="T_END",
                    help="only run the N (default 262144) Plummer configuration to T_END in "
                         "mixed and FP64 (whole runs, energy every 0.125)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "a multi-GPU j split is designed but not built.", file=sys.stderr)
        sys.exit(2)
    if os.environ.get("WORLD_SIZE", "1") not in ("", "1"):
        print("bench.py: one process, one GPU only", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe > 0:
        print(json.dumps({"copy_probe": copy_probe(args.copy_probe)}), flush=True)
        return
    if args.salpeter_bh_run > 0:
        print(json.dumps({"salpeter_bh_run": salpeter_bh_run(args.salpeter_bh_run)}), flush=True)
        return
    if args.plummer_run > 0:
        print(json.dumps({"plummer_run": plummer_run(args.plummer_run, args.n)}), flush=True)
        return

    from gravidy_amd import Hermite4HIP, ic, softening_e2

    n = args.n
    e2 = softening_e2(1e-4)
    log(f"[bench] generating Plummer N={n} seed={args.seed}")
    r, v = ic.plummer(n, args.seed)
    dt_shared = 2.0 ** -10

    log(f"[bench] full block steps, precision={args.precision}")
    with Hermite4HIP(r, v, e2, precision=args.precision, dt_min=dt_shared, dt_max=dt_shared) as h:
        h.set_timing(True)
        h.init_acc_jrk()
        f_ref = h.download()["f"]
        sec, upd_ms, st = full_step_rate(h, n, args.steps, args.warmup)
        sweep = None
        if not args.no_sweep:
            log("[bench] active-set sweep of the update")
            sweep = nact_sweep(h, n, [n, n // 8, n // 64, 4096, 1024, 256, 32])
    pairs = float(n) * n
    value = pairs / sec
    upd_rate = pairs / (upd_ms * 1e-3)
    achieved = FLOPS_PER_PAIR * upd_rate / 1e12
    log(f"[bench] {value:.4e} pairs/s per step; update {upd_ms:.3f} ms = {achieved:.2f} TFLOP/s")

    mixed = None
    if not args.no_mixed and args.precision == "fp64":
        log("[bench] mixed precision full block steps")
        with Hermite4HIP(r, v, e2, precision="mixed", dt_min=dt_shared, dt_max=dt_shared) as h:
            h.set_timing(True)
            h.init_acc_jrk()
            f_mix = h.download()["f"]
            msec, mupd, _ = full_step_rate(h, n, max(3, args.steps // 2), 1)
        err = {}
        for c, name in ((0, "acc"), (3, "jerk")):
            rel = (np.linalg.norm(f_mix[:, c:c + 3] - f_ref[:, c:c + 3], axis=1)
                   / np.linalg.norm(f_ref[:, c:c + 3], axis=1))
            err[f"{name}_rel_err_rms"] = float(np.sqrt(np.mean(rel ** 2)))
            err[f"{name}_rel_err_max"] = float(rel.max())
        mixed = {"value": pairs / msec, "ms_per_step": msec * 1e3, "update_ms": mupd,
                 "vs_fp64_initial_forces": err,
                 "speedup_vs_fp64": (pairs / msec) / value,
                 "tflops_60": FLOPS_PER_PAIR * pairs / (mupd * 1e-3) / 1e12,
                 "frac_fp32_peak": FLOPS_PER_PAIR * pairs / (mupd * 1e-3) / 1e12 / FP32_PEAK_TFLOPS}

    sched = None
    if args.block_steps > 0:
        sched = {}
        for mode, dev in (("device_loop", True), ("host_loop", False)):
            log(f"[bench] real block-timestep schedule, {args.block_steps} block steps, {mode}")
            with Hermite4HIP(r, v, e2, precision=args.precision) as h:
                h.set_graphs(not args.no_graphs)
                # untimed pass over the same steps (graph capture, first
                # launches), then the IC again and the timed pass
                block_schedule_rate(h, n, args.block_steps, dev)
                h.upload(r, v)
                rate, el, nacts, T = block_schedule_rate(h, n, args.block_steps, dev)
            sched[mode] = {"pairs_per_s": rate, "block_steps": args.block_steps, "seconds": el,
                           "us_per_step": el / args.block_steps * 1e6, "t_reached": T,
                           "mean_nact": float(np.mean(nacts)),
                           "median_nact": float(np.median(nacts))}

    ens = None
    if not args.no_ensemble:
        log(f"[bench] ensemble: 64 clusters, N in [1024, 32768], to t={args.ensemble_t_end}")
        ens = ensemble_leg(e2, args.precision, t_end=args.ensemble_t_end,
                           cpu_clusters=4 if args.cpu_sample > 0 else 0,
                           graphs=not args.no_graphs, tune_steps=args.tune_steps,
                           launch=[int(x) for x in args.ensemble_launch.split(",")]
                           if args.ensemble_launch else None)
        log(f"[bench] ensemble {ens['pairs_per_s']:.3e} pairs/s in {ens['seconds']:.2f} s")
        log("[bench] ensemble of 256 small clusters, N in [128, 1024], to t=0.25")
        ens["small_clusters"] = small_clusters_leg(e2, args.precision, graphs=not args.no_graphs)
        log(f"[bench] small clusters {ens['small_clusters']['seconds']:.3f} s "
            f"({ens['small_clusters']['seconds_multi_workgroup_only']:.3f} s without "
            f"own-workgroup steps)")

    runs = None
    if not args.no_runs:
        log("[bench] whole runs: Plummer N=1024 to t=1 (+ CPU oracle), N=65536 to t=0.5")
        g = not args.no_graphs
        runs = {"plummer_n1024_t1": run_leg(e2, 1024, 12345, 1.0, args.precision,
                                            cpu=args.cpu_sample > 0, graphs=g),
                "plummer_n65536_t0.5": run_leg(e2, 65536, 65536, 0.5, args.precision, graphs=g)}

    configs = None
    if not args.no_configs:
        configs = {}
        log("[bench] configuration legs: Plummer N=262144 (the whole run to t=1/4), "
            "Salpeter + central mass N=1048577 (to t=2^-8)")
        configs["plummer_n262144"] = same_steps_leg(
            "plummer_n262144_eps1e-4", r, v, e2, 2.0 ** -23, args.cpu_budget, 0.25, 0.125)
        from gravidy_amd import ic as _ic
        rb, vb = _ic.plummer_salpeter_bh(1048576, 1048576)
        configs["salpeter_bh_n1048577"] = same_steps_leg(
            "salpeter_plummer_bh_n1048577_eps1e-5_dtmin2^-30", rb, vb, softening_e2(1e-5),
            2.0 ** -30, args.cpu_budget, 2.0 ** -8, 2.0 ** -9)
        del rb, vb

    cpu = None
    if args.cpu_sample > 0:
        log(f"[bench] CPU oracle baseline, nact={args.cpu_sample}")
        cpu = cpu_baseline(r, v, e2, args.cpu_sample)

    kernel = {"fp64": "k_update_fp64", "exact": "k_update_exact", "mixed": "k_update_mixed"}[args.precision]
    workload = f"plummer_n{n}_eps1e-4_eta0.01_full_block_step"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "exact": "f64", "mixed": "f32+f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": workload,
                   "n": n, "seed": args.seed, "precision": args.precision,
                   "shared_dt": dt_shared, "pairs_per_step": int(pairs)},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic(kernel, workload, args.precision),
                     "kernel": kernel, "flops_per_pair": FLOPS_PER_PAIR, "update_ms": upd_ms},
        "cpu_baseline": cpu,
        "mixed": mixed,
        "block_schedule": sched,
        "ensemble": ens,
        "runs": runs,
        "configs": configs,
        "nact_sweep": sweep,
        "host": platform.node(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
