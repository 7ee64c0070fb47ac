#!/usr/bin/env python3
"""Wide-Fibonacci 2^20 x 100 proof on one MI355X through the rocm C-ABI.

One step = one whole proof (channel init -> proof bytes; the reference's
prove_timer, src/prover/native/runner.zig:620-663) of the wide-Fibonacci AIR
with log_n_rows=20, sequence_len=100 under the functional protocol (pow 10,
3 queries, blowup 1), the configuration of the reference's published
2^20 x 100 number.  The trace is generated on the device before the timed
region (the reference times trace generation separately as input_seconds).
Every timed proof is checked byte for byte against the reference's golden
proof (sha256 e6609d05...).

Prints one JSON line.  `value` is the median prove time in ms (lower is
better).  `roofline` describes the dominant kernel class (the class with the
most device time in separate profiled proofs), timed with HIP events on the
context stream around its launches during the timed proofs; no other launch of
the timed proofs carries events.  `kernels` is the per-class breakdown of the
profiled proofs.  `cpu_baseline` times the C
oracle (a multithreaded restatement of the reference CPU backend) on this
host on one proof of the same workload.
"""
import argparse
import hashlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "stwo-zig_amd"))

METRIC = "wide-Fib 2^20x100 prove ms; batched proofs/s + p99; LDE/Merkle HBM GB/s"
HBM_PEAK_GBS = 8000.0
# wave64 VALU instructions per second of the whole MI355X: 256 CUs x 2.4 GHz, one per CU-cycle
# (four SIMDs per CU, each issuing a wave64 instruction over four cycles)
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9
PUBLISHED_CPU_MS = 900.664792  # conformance/evidence/metrics-v2/w1-log20x100/cpu-report.json (M5 Max, 18 threads)
GOLDEN_SHA = {(20, 100, "functional"): "e6609d0564a47192212bec7973e2660c2eea88bef90c573c3df09569cc3c7e86"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log-n-rows", type=int, default=20)
    ap.add_argument("--sequence-len", type=int, default=100)
    ap.add_argument("--protocol", choices=["functional", "secure"], default="functional")
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--batched", choices=["on", "off"], default="on",
                    help="also run the batched primitive pipeline (64 columns, logs 14..22) into the line")
    ap.add_argument("--serving", choices=["on", "off"], default="on",
                    help="also run the serving comparison (256 requests, 1 slot vs 16 slots at equal p99)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain streaming kernel of known size (HBM counter calibration)")
    ap.add_argument("--workload", choices=["widefib", "logup", "batched", "serving", "poseidon", "xor", "many"],
                    default="widefib",
                    help="widefib: the 2^20x100 proof (default; with the batched and serving parts); logup: "
                         "XOR-lookup LogUp workload at 2^18; batched: 64 uneven columns LDE + Merkle + 3 FRI folds "
                         "as one hipGraph; serving: 256 Poisson requests (wide-Fib, XOR, Poseidon, XOR-lookup AIR), 1 slot "
                         "vs continuous batching; poseidon, xor: the reference's example AIRs at the size of their "
                         "golden proofs; many: batches of small wide-Fib proofs in single launches vs one proof "
                         "per call")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the backend has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serving" or (args.workload == "widefib" and args.serving == "on" and not args.copy_probe):
        # Every serving slot owns a HIP stream; HIP spreads a process's streams over
        # GPU_MAX_HW_QUEUES hardware queues (4 by default), so with more slots than queues
        # streams share a queue and their proofs serialize.  Set before the runtime starts.
        os.environ["GPU_MAX_HW_QUEUES"] = "16"
    import stwo_rocm

    if args.copy_probe:
        copy_probe(stwo_rocm)
        return
    if args.workload == "logup":
        bench_logup(stwo_rocm, args)
        return
    if args.workload == "batched":
        bench_batched(stwo_rocm, args)
        return
    if args.workload == "serving":
        bench_serving(stwo_rocm, args)
        return
    if args.workload in ("poseidon", "xor"):
        bench_example_air(stwo_rocm, args)
        return
    if args.workload == "many":
        bench_many(stwo_rocm, args)
        return
    bench_widefib(stwo_rocm, args)


def lde_unit_classes(prof, n, W, B):
    """Kernel classes of the trace's interpolation and LDE (every launch over the W trace columns
    on the 2^n trace domain and the 2^(n+B) evaluation domain)."""
    out = []
    for k in prof:
        if not k.endswith(" cols=%d" % W):
            continue
        if k.startswith("fft_ifft_pass m=%d " % n) or k.startswith("fft_lde_pass m=%d " % (n + B)) or \
                k.startswith("fft_ilde_pass m=%d " % n):
            out.append(k)
    return sorted(out)


def unit_figures(name, classes, alg_bytes, per_proof_ms, counters, share=None):
    """Roofline of a unit of work done by several launches: algorithmic bytes over the summed
    device time of its launches, the HBM bytes that rocprof counted for them, and the VALU
    bound (wave64 VALU instructions / (CUs x clock), one per CU-cycle)."""
    share = share or {}
    t_ms = sum(per_proof_ms[k] * share.get(k, 1.0) for k in classes)
    achieved = alg_bytes / (t_ms * 1e-3) / 1e9
    traffic = [counters["traffic"].get(k) for k in classes]
    traffic = sum(t * share.get(k, 1.0) for k, t in zip(classes, traffic)) if all(t is not None for t in traffic) \
        else None
    valu = [counters["valu"].get(k) for k in classes]
    valu = sum(v * share.get(k, 1.0) for k, v in zip(classes, valu)) if all(v is not None for v in valu) else None
    out = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": int(traffic) if traffic else None,
           "traffic_ratio": round(traffic / alg_bytes, 3) if traffic else None,
           "algorithmic_bytes": int(alg_bytes), "device_ms": round(t_ms, 4),
           "launches": {k: round(per_proof_ms[k] * share.get(k, 1.0), 4) for k in classes}}
    if valu:
        bound_ms = valu / VALU_PEAK_WAVE_INSTR_PER_S * 1e3
        out["valu"] = {"wave_instructions": int(valu), "peak_wave_instructions_per_s": VALU_PEAK_WAVE_INSTR_PER_S,
                       "bound_ms": round(bound_ms, 4), "frac": round(bound_ms / t_ms, 4)}
    return out


def bench_widefib(stwo_rocm, args):
    n, W = args.log_n_rows, args.sequence_len
    proto = dict(stwo_rocm.PROTOCOLS[args.protocol])
    B = proto["log_blowup_factor"]
    ctx = stwo_rocm.Context(0)
    ident = ctx.identity()
    ctx.prepare_twiddles(n + B + 1)
    d_trace = ctx.alloc(4 * W << n)
    ctx.fill_quadratic_recurrence_trace(n, W, d_trace.ptr, mode=0)
    kw = dict(log_n_rows=n, sequence_len=W, trace_mode=0, d_trace=d_trace, **proto)

    for _ in range(args.warmup):
        ctx.prove_wide_fibonacci(**kw)

    golden = GOLDEN_SHA.get((n, W, args.protocol))
    # Per-kernel breakdown from separate profiled proofs (every launch bracketed by events).
    prof_steps = 3
    ctx.profile_select(None)
    ctx.profile_enable(True)
    for _ in range(prof_steps):
        ctx.prove_wide_fibonacci(**kw)
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    lde = lde_unit_classes(prof, n, W, B)
    if not lde:
        raise SystemExit("no interpolation/LDE launches found in the profile")

    # Timed proofs: events only around the launches of the LDE unit, sampled (each event
    # costs GPU time), the stride coprime with the unit's launches per proof.
    ctx.profile_select(lde)
    ctx.profile_sample(_sample_every(sum(prof[k]["launches"] for k in lde) / prof_steps, target=1))
    ctx.profile_enable(True)
    ctx.synchronize()
    times, shas = [], set()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        proof, _ = ctx.prove_wide_fibonacci(**kw, stats=False)
        times.append((time.perf_counter() - t0) * 1e3)
        shas.add(hashlib.sha256(proof).hexdigest())
    ctx.synchronize()
    wall_ms = (time.perf_counter() - t_start) * 1e3
    dprof = ctx.profile_read()
    ctx.profile_enable(False)
    ctx.profile_select(None)
    ctx.profile_sample(1)
    # per-stage device times from untimed proofs (stage events are kept out of the timed ones)
    stats_all = []
    for _ in range(3):
        proof_s, st = ctx.prove_wide_fibonacci(**kw)
        shas.add(hashlib.sha256(proof_s).hexdigest())
        stats_all.append(st)
    if len(shas) != 1:
        raise SystemExit("non-deterministic proofs across steps")
    sha = shas.pop()
    if golden is not None and sha != golden:
        raise SystemExit("proof differs from the reference golden proof: %s" % sha)
    try:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import stark_verifier
        stark_verifier.verify_proof(proof, "wide_fibonacci", dict(log_n_rows=n, sequence_len=W), proto)
        verified = True
    except ImportError:
        verified = None

    median_ms = statistics.median(times)
    counters = profiled_counters()
    # LDE unit (live, timed proofs): per launch average of each class x its launches per proof
    live_ms = {k: dprof[k]["ms"] / dprof[k]["launches"] * prof[k]["launches"] / prof_steps for k in lde if k in dprof}
    if set(live_ms) != set(lde):
        raise SystemExit("a launch of the LDE unit was never sampled in the timed proofs")
    lde_bytes = 4.0 * W * ((1 << n) + (1 << n) + (1 << (n + B)))  # trace in, coefficients out, LDE out
    roofline = unit_figures("LDE unit: trace interpolation + LDE (%d launches)" % len(lde), lde, lde_bytes, live_ms,
                            counters)
    # Merkle unit of the main tree (profiled proofs): its leaf launch, one parent launch and one
    # tree-top launch (the trace, composition and first FRI trees have the same shape and share
    # the parent class; the trace and composition trees share the top class)
    per_proof = {k: v["ms"] / prof_steps for k, v in prof.items()}
    L = n + B
    leaf_cls = [k for k in prof if k == "merkle_leaves L=%d cols=%d" % (L, W)]
    par_cls = [k for k in prof if k.startswith("merkle_parents L=%d " % L)]
    top_cls = [k for k in prof if k.startswith("merkle_top L=") and "+channel" not in k]
    # one launch's worth of each class (the classes are shared by trees of the same shape)
    share = {k: prof_steps / prof[k]["launches"] for k in par_cls + top_cls}
    merkle_bytes = 4.0 * W * (1 << L) + 32.0 * ((2 << L) - 1)
    merkle = unit_figures("Merkle unit: main-tree leaves + parents + top", leaf_cls + par_cls + top_cls,
                          merkle_bytes, per_proof, counters, share) if leaf_cls else None

    kernels = {k: {"launches_per_proof": v["launches"] / prof_steps, "ms_per_proof": round(v["ms"] / prof_steps, 4),
                   "GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])}
    stage_gpu = {k: statistics.median(s["stage_gpu_ms"][k] for s in stats_all) for k in stats_all[0]["stage_gpu_ms"]}

    line = {
        "metric": METRIC,
        "value": round(median_ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms / args.steps, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(median_ms / PUBLISHED_CPU_MS, 6) if (n, W, args.protocol) == (20, 100, "functional")
        else None,
        "dtype": "u32 (exact M31/CM31/QM31)",
        "data": "synthetic: reference wide-Fibonacci trace seeds (1, i), generated on device before timing",
        "config": {"workload": "wide_fibonacci log_n_rows=%d sequence_len=%d protocol=%s (pow_bits=%d, n_queries=%d,"
                               " log_blowup=%d), one proof" % (n, W, args.protocol, proto["pow_bits"],
                                                               proto["n_queries"], proto["log_blowup_factor"]),
                   "log_n_rows": n, "sequence_len": W, "protocol": args.protocol, "device": ident},
        "roofline": roofline,
        "merkle_unit": merkle,
        "cpu_baseline": None,
        "proof_sha256": sha,
        "bit_exact_vs_reference_golden": golden is not None and sha == golden,
        "verifier_accepts": verified,
        "prove_ms": {"median": round(median_ms, 3), "min": round(min(times), 3), "max": round(max(times), 3)},
        "device_ms_per_proof": round(statistics.median(s["total_gpu_ms"] for s in stats_all), 3),
        "stage_device_ms": {k: round(v, 3) for k, v in stage_gpu.items()},
        "kernels": kernels,
        "published_cpu_ms_m5max_18t": PUBLISHED_CPU_MS,
        "speedup_vs_published_cpu": round(PUBLISHED_CPU_MS / median_ms, 1),
    }
    d_trace.free()
    if args.batched == "on":
        line["batched"] = batched_part(stwo_rocm, ctx, steps=10, warmup=2, check_oracle=args.cpu_baseline == "on")
        print(json.dumps({"batched": line["batched"]}), file=sys.stderr, flush=True)
    if args.batched == "on":
        line["many_small_proofs"] = many_proofs_part(stwo_rocm, ctx, check_oracle=args.cpu_baseline == "on")
        ctx.release_cached()
    if args.serving == "on":
        ctx.release_cached()
        line["serving"] = serving_part(stwo_rocm, ctx, check_oracle=args.cpu_baseline == "on")
        print(json.dumps({"serving": {k: v for k, v in line["serving"].items() if k != "sweep"}}), file=sys.stderr,
              flush=True)
    ctx.close()
    if args.cpu_baseline == "on":
        line["cpu_baseline"] = cpu_baseline(n, W, proto, sha)
        line["speedup_vs_cpu_baseline"] = round(line["cpu_baseline"]["value"] / median_ms, 1)
        if line["cpu_baseline"].get("one_thread_ms"):
            line["speedup_vs_cpu_one_thread"] = round(line["cpu_baseline"]["one_thread_ms"] / median_ms, 1)
    print(json.dumps(line), flush=True)


def _dominant(prof, steps):
    """(class, per-launch ms, per-launch algorithmic bytes, kernels table) of a profile."""
    dom, dv = max(prof.items(), key=lambda kv: kv[1]["ms"])
    kernels = {k: {"launches_per_step": v["launches"] / steps, "ms_per_step": v["ms"] / steps,
                   "GBs": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])}
    return dom, dv["ms"] / dv["launches"], dv["bytes"] / dv["launches"], kernels


def _sample_every(launches_per_step, target=4):
    """Stride for the timed run's events on the selected classes: about `target` bracketed
    launches per step (each event costs a few microseconds of GPU time), the stride coprime with
    the launches of a step so that over k steps every launch position is sampled once."""
    import math

    n = int(round(launches_per_step))
    k = max(1, n // target)
    while k > 1 and math.gcd(n, k) != 1:
        k += 1
    return k


def _roofline(dom, avg_ms, bytes_per_launch, traffic):
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic.get(dom),
            "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(bytes_per_launch)}


def bench_logup(stwo_rocm, args):
    """XOR-lookup LogUp workload, log 18: trace -> commit -> cumulative columns -> commit -> GKR ->
    sampled values -> quotients -> commit, one call; checked word for word against the oracle."""
    log_n, seed = 18, 0x5EED
    ctx = stwo_rocm.Context(0)
    ident = ctx.identity()
    for _ in range(max(args.warmup, 1)):
        words, _ = ctx.logup_xor_pipeline(log_n, seed)
    prof_steps = 2
    ctx.profile_select(None)
    ctx.profile_enable(True)
    for _ in range(prof_steps):
        ctx.logup_xor_pipeline(log_n, seed)
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    dom, _, _, kernels = _dominant(prof, prof_steps)
    ctx.profile_select(dom)
    ctx.profile_sample(_sample_every(kernels[dom]["launches_per_step"]))
    ctx.profile_enable(True)
    times = []
    shas = set()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        words, ms = ctx.logup_xor_pipeline(log_n, seed)
        times.append(ms)
        shas.add(hashlib.sha256(words.tobytes()).hexdigest())
    wall_ms = (time.perf_counter() - t_start) * 1e3
    dprof = ctx.profile_read()
    ctx.profile_enable(False)
    ctx.profile_select(None)
    if len(shas) != 1:
        raise SystemExit("non-deterministic LogUp workload")
    dv = dprof[dom]
    cpu, ref_sha = None, None
    if args.cpu_baseline == "on":
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t0 = time.perf_counter()
        ref = oracle.logup_xor_pipeline(log_n, seed)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        ref_sha = hashlib.sha256(ref.tobytes()).hexdigest()
        if ref_sha not in shas:
            raise SystemExit("LogUp workload differs from the oracle")
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        cpu = {"value": round(cpu_ms, 1), "unit": "ms", "cores": threads, "kind": "port",
               "sample": "1 full run of the same workload (XOR LogUp 2^18) by the C oracle"}
    med = statistics.median(times)
    line = {"metric": "LogUp XOR-lookup 2^18 workload ms (trace, commits, cumulative columns, GKR, quotients)",
            "value": round(med, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(wall_ms / args.steps, 3), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32 (exact M31/QM31)",
            "data": "synthetic: seeded 8-bit a, b (splitmix64), c = a xor b",
            "config": {"workload": "xor_logup log_n=18 (lookups 2^18, table 2^16 padded), blowup 2", "device": ident},
            "roofline": _roofline(dom, dv["ms"] / dv["launches"], dv["bytes"] / dv["launches"], profiled_traffic()),
            "cpu_baseline": cpu, "result_sha256": shas.pop(), "bit_exact_vs_oracle": True if ref_sha is not None else None,
            "kernels": kernels}
    if cpu:
        line["speedup_vs_cpu_baseline"] = round(cpu["value"] / med, 1)
    print(json.dumps(line), flush=True)
    ctx.close()


BATCHED_SPEC = (64, 14, 22, 2, 22, 3, 1)  # columns, log min, log max, log blowup, FRI line log, folds, seed


def _splitmix64(x):
    M = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def batched_column_logs(spec):
    """Column log sizes of the batched pipeline (the rule of batched_pipeline.hip / the oracle)."""
    n, lo, hi, _, _, _, seed = spec
    M = (1 << 64) - 1
    return [lo + _splitmix64(seed ^ 0xA5A5A5A5 ^ ((i * 0x9E3779B97F4A7C15) & M)) % (hi - lo + 1) for i in range(n)]


def batched_part(stwo_rocm, ctx, steps, warmup, check_oracle, spec=BATCHED_SPEC):
    """Config C on `ctx`: autotuned launch shapes, graph replays timed (device ms per replay),
    eager launches profiled for the LDE and Merkle units, result checked against the oracle."""
    logs = batched_column_logs(spec)
    B = spec[3]
    bp = ctx.batched(*spec)
    try:
        tile, levels, table = bp.autotune(reps=2)
        bp.run(max(warmup, 1), graph=True)
        eager_words, eager_ms = bp.run(3, graph=False)
        # per-kernel profile of eager runs (a graph replay carries no events)
        prof_steps = 2
        ctx.profile_select(None)
        ctx.profile_enable(True)
        bp.run(prof_steps, graph=False)
        prof = ctx.profile_read()
        ctx.profile_enable(False)
        times = []
        t_start = time.perf_counter()
        for _ in range(steps):
            words, ms = bp.run(1, graph=True)
            times.append(ms)
        wall_ms = (time.perf_counter() - t_start) * 1e3
    finally:
        bp.set_tuning(13, 10)
        bp.close()
    if not np.array_equal(words, eager_words):
        raise SystemExit("batched pipeline: graph and eager results differ")
    ok, cpu = None, None
    if check_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t0 = time.perf_counter()
        ref = oracle.batched_pipeline(*spec)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        ok = bool(np.array_equal(ref, words))
        if not ok:
            raise SystemExit("batched pipeline differs from the oracle")
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        cpu = {"value": round(cpu_ms, 1), "unit": "ms", "cores": threads, "kind": "port",
               "sample": "1 full run of the same pipeline (64 columns, logs 14..22) by the C oracle"}
    per_run = {k: v["ms"] / prof_steps for k, v in prof.items()}
    counters = profiled_counters()
    fft = sorted(k for k in prof if k.startswith("fft_"))
    lde_bytes = 4.0 * sum((1 << l) + (1 << l) + (1 << (l + B)) for l in logs)
    L = max(logs) + B
    mk = sorted(k for k in prof if k.startswith("merkle_leaves") or k.startswith("merkle_parents L=%d " % L) or
                (k.startswith("merkle_top") and "+channel" not in k))
    merkle_bytes = 4.0 * sum(1 << (l + B) for l in logs) + 32.0 * ((2 << L) - 1)
    med = statistics.median(times)
    return {"value": round(med, 3), "unit": "ms per graph replay", "steps": steps, "warmup": warmup,
            "ms_per_step": round(wall_ms / steps, 3),
            "workload": "64 columns (log sizes 14..22, seeded), interpolate + LDE at blowup 4, one lifted Blake2s "
                        "Merkle tree, 3 FRI fold_line rounds on a log-22 QM31 line with a device channel; one hipGraph",
            "tuning": {"fft_tile_log": tile, "merkle_levels": levels}, "autotune_table": table,
            "eager_ms": round(eager_ms, 3), "graph_speedup_vs_eager": round(eager_ms / med, 3),
            "roofline": unit_figures("LDE unit: interpolation + LDE of 64 columns (eager launches)", fft, lde_bytes,
                                     per_run, counters),
            "merkle_unit": unit_figures("Merkle unit: lifted tree of 64 LDE columns (eager launches)", mk,
                                        merkle_bytes, per_run, counters),
            "bit_exact_vs_oracle": ok, "cpu_baseline": cpu,
            "kernels": {k: round(v, 4) for k, v in sorted(per_run.items(), key=lambda kv: -kv[1])}}


def bench_batched(stwo_rocm, args):
    """64 seeded columns, log sizes 14..22, LDE at blowup 4, one lifted Merkle tree, 3 FRI fold_line
    rounds on a log-22 line; autotuned launch shapes; one hipGraph replay per step."""
    ctx = stwo_rocm.Context(0)
    ident = ctx.identity()
    part = batched_part(stwo_rocm, ctx, args.steps, args.warmup, args.cpu_baseline == "on")
    ctx.close()
    line = {"metric": "batched primitive pipeline ms (64 cols log 14..22, LDE x4, lifted Merkle, 3 FRI folds)",
            "value": part["value"], "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": part["ms_per_step"], "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32 (exact M31/QM31)",
            "data": "synthetic: seeded uniform M31 columns and QM31 line (splitmix64)",
            "config": {"workload": "batched: 64 columns, log sizes 14..22, log_blowup 2, fri_log 22, 3 folds",
                       "device": ident, "tuning": part["tuning"]},
            "roofline": part["roofline"], "merkle_unit": part["merkle_unit"], "cpu_baseline": part["cpu_baseline"],
            "bit_exact_vs_oracle": part["bit_exact_vs_oracle"], "eager_ms": part["eager_ms"],
            "autotune_table": part["autotune_table"], "kernels": part["kernels"]}
    if part["cpu_baseline"]:
        line["speedup_vs_cpu_baseline"] = round(part["cpu_baseline"]["value"] / part["value"], 1)
    print(json.dumps(line), flush=True)


def bench_example_air(stwo_rocm, args):
    """The reference's Poseidon (2^13 instances) or periodic XOR (2^16 rows, step 4, offset 3)
    example AIR, functional protocol, at the size of its committed golden proof: every timed
    proof must equal the golden bytes."""

    poseidon = args.workload == "poseidon"
    gold = json.load(open(os.path.join(ROOT, "tests", "golden",
                                       "poseidon_goldens.json" if poseidon else "xor_goldens.json")))
    g = gold["poseidon_log13_pow10" if poseidon else "xor_log16_step2_offset3_pow10"]
    ctx = stwo_rocm.Context(0)
    ident = ctx.identity()

    def prove(stats=False):
        if poseidon:
            return ctx.prove_poseidon(13, stats=stats)
        return ctx.prove_xor(16, 2, 3, stats=stats)

    for _ in range(max(args.warmup, 1)):
        prove()
    prof_steps = 3
    ctx.profile_select(None)
    ctx.profile_enable(True)
    for _ in range(prof_steps):
        prove()
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    dom, _, _, kernels = _dominant(prof, prof_steps)
    ctx.profile_select(dom)
    ctx.profile_sample(_sample_every(kernels[dom]["launches_per_step"]))
    ctx.profile_enable(True)
    ctx.synchronize()
    times, shas = [], set()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        proof, _ = prove()
        times.append((time.perf_counter() - t0) * 1e3)
        shas.add(hashlib.sha256(proof).hexdigest())
    wall_ms = (time.perf_counter() - t_start) * 1e3
    dprof = ctx.profile_read()
    ctx.profile_enable(False)
    ctx.profile_select(None)
    ctx.profile_sample(1)
    if shas != {g["sha256"]}:
        raise SystemExit("example AIR proof differs from the reference golden proof")
    dv = dprof[dom]
    median_ms = statistics.median(times)
    cpu = None
    if args.cpu_baseline == "on":
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        k = 5
        t0 = time.perf_counter()
        for _ in range(k):
            ref = oracle.prove_poseidon(13) if poseidon else oracle.prove_xor(16, 2, 3)
        cpu_ms = (time.perf_counter() - t0) * 1e3 / k
        if hashlib.sha256(ref).hexdigest() != g["sha256"]:
            raise SystemExit("oracle proof differs from the reference golden proof")
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        cpu = {"value": round(cpu_ms, 2), "unit": "ms", "cores": threads, "kind": "port",
               "sample": "%d full proofs of the same statement by the C oracle (mean)" % k}
    name = "Poseidon AIR log_n_instances=13 (1264 columns x 2^10 rows)" if poseidon else \
        "periodic XOR AIR log_size=16 log_step=2 offset=3"
    line = {"metric": "%s prove ms" % name, "value": round(median_ms, 3), "unit": "ms", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall_ms / args.steps, 3),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u32 (exact M31/CM31/QM31)",
            "data": "the reference example's own deterministic trace, generated on device",
            "config": {"workload": name + ", functional protocol (pow 10, 3 queries, blowup 2)", "device": ident},
            "roofline": _roofline(dom, dv["ms"] / dv["launches"], dv["bytes"] / dv["launches"], profiled_traffic()),
            "cpu_baseline": cpu, "proof_sha256": g["sha256"], "bit_exact_vs_reference_golden": True,
            "prove_ms": {"median": round(median_ms, 3), "min": round(min(times), 3), "max": round(max(times), 3)},
            "kernels": kernels}
    if cpu:
        line["speedup_vs_cpu_baseline"] = round(cpu["value"] / median_ms, 1)
    print(json.dumps(line), flush=True)
    ctx.close()


MANY_SHAPES = ((10, 8, 256), (14, 32, 64))


def many_proofs_part(stwo_rocm, ctx, check_oracle, shapes=MANY_SHAPES, reps=5):
    """Many small proofs in single launches against the one-at-a-time GPU path: for each shape, K
    seeded-uniform wide-Fibonacci proofs (functional protocol) proven one call per proof, as one
    batch call (stwo_rocm_prove_wide_fibonacci_batch), and from a batch plan whose device part is
    replayed from a hipGraph (the headline figure); proofs/s by wall clock, the device and host
    phase times, and the check that every batched proof is the one-at-a-time proof (with
    check_oracle, the first and last also the oracle's)."""
    pcs = dict(stwo_rocm.PROTOCOLS["functional"])
    out = {}
    for n, w, k in shapes:
        seeds = [1000 * n + w + i for i in range(k)]
        ctx.prove_wide_fibonacci(n, w, trace_mode=1, trace_seed=seeds[0], stats=False, **pcs)
        for _ in range(2):
            ctx.prove_wide_fibonacci_batch(n, w, seeds=seeds, trace_mode=1, **pcs)
        one_s, ones = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            ones = [ctx.prove_wide_fibonacci(n, w, trace_mode=1, trace_seed=sd, stats=False, **pcs)[0] for sd in seeds]
            one_s.append(time.perf_counter() - t0)
        bat_s, stats_all, proofs = [], [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            proofs, st = ctx.prove_wide_fibonacci_batch(n, w, seeds=seeds, trace_mode=1, **pcs)
            bat_s.append(time.perf_counter() - t0)
            stats_all.append(st)
        if proofs != ones:
            raise SystemExit("batched proofs differ from the one-at-a-time proofs (%dx%d)" % (n, w))
        # the same batches from a plan set up once, its device part replayed from a hipGraph
        plan_s, plan_stats = [], []
        with ctx.batch_prover(n, w, k, trace_mode=1, **pcs) as bp:
            bp.run(seeds, graph=True)  # records the graph
            for _ in range(reps):
                t0 = time.perf_counter()
                pproofs, st = bp.run(seeds, graph=True)
                plan_s.append(time.perf_counter() - t0)
                plan_stats.append(st)
        if pproofs != ones:
            raise SystemExit("graph-replayed batch proofs differ from the one-at-a-time proofs (%dx%d)" % (n, w))
        ok = None
        if check_oracle:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            ok = all(proofs[i] == oracle.prove_wide_fibonacci(n, w, trace_mode=1, trace_seed=seeds[i], **pcs)
                     for i in (0, k - 1))
            if not ok:
                raise SystemExit("batched proofs differ from the oracle (%dx%d)" % (n, w))
        one_ps, bat_ps = k / statistics.median(one_s), k / statistics.median(bat_s)
        plan_ps = k / statistics.median(plan_s)
        out["%dx%d" % (n, w)] = {
            "proofs": k, "one_at_a_time_proofs_per_s": round(one_ps, 1), "batched_proofs_per_s": round(plan_ps, 1),
            "speedup": round(plan_ps / one_ps, 2),
            "graph_plan": {"proofs_per_s": round(plan_ps, 1), "wall_ms": round(statistics.median(plan_s) * 1e3, 3),
                           "device_ms": round(statistics.median(x["device_ms"] for x in plan_stats), 3),
                           "host_phase_ms": {ph: round(statistics.median(x["phase_ms"][ph] for x in plan_stats), 3)
                                             for ph in plan_stats[0]["phase_ms"]}},
            "one_call_proofs_per_s": round(bat_ps, 1), "batch_wall_ms": round(statistics.median(bat_s) * 1e3, 3),
            "batch_device_ms": round(statistics.median(x["device_ms"] for x in stats_all), 3),
            "batch_host_tail_ms": round(statistics.median(x["tail_ms"] for x in stats_all), 3),
            "batch_host_phase_ms": {ph: round(statistics.median(x["phase_ms"][ph] for x in stats_all), 3)
                                    for ph in stats_all[0]["phase_ms"]},
            "device_bytes_per_proof": stats_all[0]["slice_bytes"], "equal_to_one_at_a_time": True,
            "oracle_first_last_equal": ok}
        print(json.dumps({"many_small_proofs": {"%dx%d" % (n, w): out["%dx%d" % (n, w)]}}), file=sys.stderr,
              flush=True)
    return out


def bench_many(stwo_rocm, args):
    ctx = stwo_rocm.Context(0)
    ident = ctx.identity()
    part = many_proofs_part(stwo_rocm, ctx, args.cpu_baseline == "on")
    ctx.close()
    head = part["%dx%d" % MANY_SHAPES[0][:2]]
    line = {"metric": "many small proofs in single launches: wide-Fib 2^10x8 proofs/s (batch of 256)",
            "value": head["batched_proofs_per_s"], "unit": "proofs/s", "n_gpus": 1, "steps": 5, "warmup": 2,
            "ms_per_step": head["batch_wall_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32 (exact M31/CM31/QM31)", "data": "synthetic: seeded-uniform traces",
            "config": {"workload": "wide_fibonacci seeded-uniform, functional protocol; shapes %s" %
                                   ", ".join("2^%dx%d (K=%d)" % s for s in MANY_SHAPES), "device": ident},
            "cpu_baseline": None, "shapes": part}
    print(json.dumps(line), flush=True)


def serving_requests(n=256, rate_per_s=1500.0, seed=2024):
    """Seeded request stream: Poisson arrivals at `rate_per_s`; half the requests are wide-Fibonacci
    proofs (log_n_rows 14..20, sequence_len 16..100, seeded uniform trace), the rest periodic XOR
    AIR proofs (log_size 12..18, log_step 0..4, offset < 2^20), Poseidon AIR proofs
    (log_n_instances 12..16) and XOR-lookup AIR proofs (the LogUp AIR, log 16..18), in shares
    2:1.5:1.5."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n):
        t += rng.exponential(1000.0 / rate_per_s)
        u = rng.random()
        if u < 0.5:
            jobs.append((0, int(rng.integers(14, 21)), int(rng.integers(16, 101)), int(rng.integers(1, 1 << 40)), t))
        elif u < 0.7:
            jobs.append((3, int(rng.integers(12, 19)), int(rng.integers(0, 5)), int(rng.integers(0, 1 << 20)), t))
        elif u < 0.85:
            jobs.append((2, int(rng.integers(12, 17)), 0, 0, t))
        else:
            jobs.append((4, int(rng.integers(16, 19)), 0, int(rng.integers(1, 1 << 40)), t))
    return jobs


def _serve_ref_digest(oracle, job):
    kind, n, w, seed, _ = job
    if kind == 0:
        out = oracle.prove_wide_fibonacci(n, w, trace_mode=1, trace_seed=seed)
    elif kind == 2:
        out = oracle.prove_poseidon(n)
    elif kind == 3:
        out = oracle.prove_xor(n, w, seed)
    elif kind == 4:
        out = oracle.prove_logup_xor(n, seed)
    else:
        out = oracle.logup_xor_pipeline(n, seed).astype(np.uint32).tobytes()
    return hashlib.blake2s(out).hexdigest()


def serving_oracle_subset(jobs):
    """Requests whose served result is recomputed by the oracle: the first request of every kind
    in the stream, and four spread over it."""
    first = {next(i for i, j in enumerate(jobs) if j[0] == k) for k in {j[0] for j in jobs}}
    return sorted(first | {i for i in (0, 64, 128, 192) if i < len(jobs)})


def serving_part(stwo_rocm, ctx, check_oracle, slot_counts=(1, 16)):
    """Config E on `ctx`: the 256-request stream served by one slot (the unbatched GPU path, one
    proof at a time) and by 16 slots (continuous batching); proofs/s and p50/p99 at 1500/s, and
    the equal-p99 comparison: the offered Poisson rate swept for both, keeping for each the highest
    achieved proofs/s whose p99 stays within 50 ms.  Every run's results must be identical."""
    jobs = serving_requests()
    burst = serving_requests(rate_per_s=1e9)  # the same requests, all arriving at once
    # warm-up: the widest service proves the whole stream once, so the resident slot contexts
    # hold warm arenas and towers for every size in the mix
    stwo_rocm.serve(ctx, burst, n_slots=16)
    digests = None

    def rounded(st):
        return {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}

    def check(res, what):
        nonlocal digests
        d = [r["digest"] for r in res]
        if digests is None:
            digests = d
        elif d != digests:
            raise SystemExit("serving results differ (%s)" % what)

    runs, capacity = {}, {}
    for slots in slot_counts:
        res, st = stwo_rocm.serve(ctx, jobs, n_slots=slots)
        check(res, "slots %d" % slots)
        runs[slots] = rounded(st)
        print(json.dumps({"serving_slots": slots, "rate_per_s": 1500, **runs[slots]}), file=sys.stderr, flush=True)
    for slots in (1, 16):
        res, st = stwo_rocm.serve(ctx, burst, n_slots=slots)
        check(res, "burst, slots %d" % slots)
        capacity[slots] = rounded(st)
    p99_bound_ms = 50.0
    sweep = {1: [], 16: []}
    for slots, rates in ((1, (200, 300, 400, 500, 600)), (16, (1000, 1500, 2000, 2500, 3000))):
        for rate in rates:
            res, st = stwo_rocm.serve(ctx, serving_requests(rate_per_s=float(rate)), n_slots=slots)
            check(res, "rate %d, slots %d" % (rate, slots))
            sweep[slots].append({"rate_per_s": rate, "jobs_per_s": round(st["jobs_per_s"], 1),
                                 "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2)})

    def sustained(rows):
        ok = [r for r in rows if r["p99_ms"] <= p99_bound_ms]
        return max(ok, key=lambda r: r["jobs_per_s"]) if ok else None

    one, many = sustained(sweep[1]), sustained(sweep[16])
    equal_p99 = {"p99_bound_ms": p99_bound_ms, "one_slot": one, "slots_16": many, "sweep": sweep}
    if one and many:
        equal_p99["speedup"] = round(many["jobs_per_s"] / one["jobs_per_s"], 3)
    ok = None
    if check_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        subset = serving_oracle_subset(jobs)
        ok = all(_serve_ref_digest(oracle, jobs[i]) == digests[i] for i in subset)
        if not ok:
            raise SystemExit("served results differ from the oracle")
    return {"value": many["jobs_per_s"] if many else None, "unit": "proofs/s at p99 <= 50 ms (16 slots)",
            "p50_ms": many["p50_ms"] if many else None, "p99_ms": many["p99_ms"] if many else None,
            "unbatched_one_slot": one, "speedup_vs_unbatched_at_equal_p99": equal_p99.get("speedup"),
            "workload": "256 requests, Poisson arrivals (seeded): 50% wide-Fibonacci (log 14..20 x 16..100, "
                        "seeded-uniform traces), 20% periodic XOR AIR (log 12..18), 15% Poseidon AIR (log 12..16), "
                        "15% XOR-lookup LogUp (log 16..18); functional protocol",
            "at_1500_per_s": runs, "capacity_burst": capacity, "equal_p99": equal_p99,
            "results_equal_across_runs": True, "oracle_subset_equal": ok}


def bench_serving(stwo_rocm, args):
    """256 requests with Poisson arrivals: the one-slot service (the unbatched GPU path: one proof
    at a time) against continuous batching over several slots; proofs/s and p50/p99 latency."""
    ctx = stwo_rocm.Context(0)
    ident = ctx.identity()
    part = serving_part(stwo_rocm, ctx, args.cpu_baseline == "on", slot_counts=(1, 2, 4, 8, 16))
    ctx.close()
    cpu = None
    if args.cpu_baseline == "on":
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        jobs = serving_requests()
        k = 12
        t0 = time.perf_counter()
        for job in jobs[:k]:
            _serve_ref_digest(oracle, job)
        el = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        cpu = {"value": round(k / el, 2), "unit": "proofs/s", "cores": threads, "kind": "port",
               "sample": "the first %d requests of the same stream run back to back by the C oracle" % k}
    line = {"metric": "serving proofs/s at equal p99 (256 Poisson requests, wide-Fib 14..20 x 16..100 + XOR 12..18 "
                      "+ Poseidon 12..16 + LogUp 16..18) + p50/p99",
            "value": part["value"], "unit": "proofs/s", "n_gpus": 1, "steps": 256, "warmup": 256,
            "ms_per_step": round(1e3 / part["value"], 4) if part["value"] else None,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32 (exact M31/QM31)",
            "data": "synthetic: seeded request stream and seeded traces",
            "config": {"workload": part["workload"], "device": ident, "slots": 16},
            "cpu_baseline": cpu, **{k: v for k, v in part.items() if k not in ("value", "unit", "workload")}}
    print(json.dumps(line), flush=True)


def profiled_counters():
    """Per-launch rocprof figures per kernel class from the committed summaries in profiles/:
    lines `traffic_bytes_per_launch[<class>]: <bytes>` (FETCH_SIZE + WRITE_SIZE, calibrated on a
    copy kernel) and `valu_wave_instructions_per_launch[<class>]: <count>` (SQ_INSTS_VALU)."""
    import glob
    import re

    out = {"traffic": {}, "valu": {}}
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            m = re.match(r"(traffic_bytes|valu_wave_instructions)_per_launch\[(.+)\]: ([0-9.e+]+)$", line.strip())
            if m:
                out["traffic" if m.group(1) == "traffic_bytes" else "valu"][m.group(2)] = float(m.group(3))
    return out


def profiled_traffic():
    return profiled_counters()["traffic"]


def copy_probe(stwo_rocm):
    """acc[i] += src[i] over 2^26 words: reads 512 MiB, writes 256 MiB per launch (5 launches)."""
    ctx = stwo_rocm.Context(0)
    n = 1 << 26
    a, b = ctx.alloc(4 * n), ctx.alloc(4 * n)
    for _ in range(5):
        ctx.accumulate(a.ptr, b.ptr, n)
    ctx.synchronize()
    print(json.dumps({"copy_probe": {"launches": 5, "read_bytes": 8 * n, "write_bytes": 4 * n}}), flush=True)
    ctx.close()


def cpu_baseline(n, W, proto, gpu_sha=None, samples=10, warmups=2):
    """The oracle (C + OpenMP restatement of the reference CPU backend) on this host: `warmups`
    excluded proofs, then `samples` timed proofs at all host threads (median and MAD of the
    prove time, trace generation excluded as in the reference's prove_seconds), then one proof
    on one thread.  Every sample's proof must equal the GPU's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle
    except Exception as e:  # oracle not built
        return {"error": str(e)}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    oracle.set_threads(threads)
    shas = set()
    for _ in range(warmups):
        shas.add(hashlib.sha256(oracle.prove_wide_fibonacci(n, W, **proto)).hexdigest())
    runs = []
    for _ in range(samples):
        proof = oracle.prove_wide_fibonacci(n, W, **proto)
        runs.append(oracle.last_timings()["prove_total"])
        shas.add(hashlib.sha256(proof).hexdigest())
    oracle.set_threads(1)
    proof = oracle.prove_wide_fibonacci(n, W, **proto)
    one = oracle.last_timings()["prove_total"]
    shas.add(hashlib.sha256(proof).hexdigest())
    oracle.set_threads(threads)
    med = statistics.median(runs)
    mad = statistics.median(abs(r - med) for r in runs)
    return {"value": round(med, 1), "unit": "ms", "cores": threads, "kind": "port",
            "mad_ms": round(mad, 1), "min_ms": round(min(runs), 1), "samples": samples, "warmups": warmups,
            "one_thread_ms": round(one, 1),
            "proofs_equal_gpu": (shas == {gpu_sha}) if gpu_sha else None,
            "sample": "%d warmups, then %d full proofs of the same workload (wide_fibonacci %dx%d) at %d threads "
                      "(median, MAD) and 1 proof at 1 thread; prove time excludes trace generation, as the "
                      "reference's prove_seconds; every proof sha256 %s"
                      % (warmups, samples, n, W, threads, (gpu_sha or shas.pop())[:16])}


if __name__ == "__main__":
    main()
