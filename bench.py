#!/usr/bin/env python3
"""Benchmark of the compaction path on one MI355X.

A step is one whole compaction of the medium workload -- 50 Mbp seeded
synthetic genome with 5% repeats (2-10 kbp copies, 0.1% substitutions)
and 1% error k-mers, every distinct canonical 31-mer (~48M), l = 11,
64 minimizer buckets -- from packed k-mers resident in HBM to the unitig
text in HBM: bucketing, edge finding, contraction, emission.

Prints one JSON line.  value = k-mers compacted per second (whole job).
roofline = the dominant kernel's algorithmic bytes per launch over its
average launch time (HIP events on the library's stream).  cpu_baseline =
the serial restatement of the reference pipeline (oracle/refpipe.c) on a
bounded sample of the same model, on one host core, checked against the
GPU's output for that sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "k-mers compacted/sec on 1 MI355X (bit-exact unitigs), batched vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ensure_built():
    import unitigcoarsening_amd as uc

    if not (os.path.exists(uc.LIB_PATH) and os.path.exists(uc.SYNTH_PATH)):
        from unitigcoarsening_amd.build import build_all

        build_all()


def kernel_bytes(name, n, m, x, W, n_unitigs, text_bytes, k):
    """Algorithmic HBM bytes of one launch of `name` over the whole job:
    every array the kernel must touch, once, at its natural width.
    n k-mers, m slots (k-mers + crossing duplicates), x crossing k-mers,
    8W bytes per packed k-mer."""
    kb = 8 * W
    ins = m - x          # slots that insert their prefix (not borrowed)
    look = m - x         # slots that look up their suffix (not lent)
    table = {
        # read k-mer, write bucket code
        "k_lmin_count": n * (kb + 4),
        # read k-mer + code; write slot k-mer, flag, source id; pair list
        "k_lmin_scatter": n * (kb + 4) + m * (kb + 1 + 4) + x * 8,
        # read flag + k-mer; claim one 8-byte table slot per insert
        "k_edge_insert": m * (1 + kb) + ins * 8,
        # read flag + k-mer; read the hit slot, read the owner k-mer, add to
        # the slot; write the hit index
        "k_edge_lookup": m * (1 + kb + 4) + look * (8 + kb + 8),
        # read hit, read the table slot, write succ and pred
        "k_edge_resolve": m * (4 + 4) + look * 8 + n * 4,
        # read flag + pred per slot, succ per k-mer; write length and next
        # splitter per head, unitig length per head, one sum per tile
        "k_seg_count": m * (1 + 4) + n * 4 + n_unitigs * (4 + 4 + 4),
        # read flag + pred per slot, length per head; succ + k-mer per
        # k-mer; write the text
        "k_head_emit": m * (1 + 4) + n_unitigs * (4 + 4 + 4) + n * (4 + kb) + text_bytes,
        # move the borrowed out-edge onto the home copy
        "k_splice": x * (4 + 4 + 4 + 4 + 4),
    }
    return table.get(name)


def traffic_for(name):
    """Measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration) when a summary was committed."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            t = json.load(f)
        return t.get("kernels", {}).get(name, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(sample_bp, seed, k, l, B, gpu_ctx_kwargs):
    """Serial restatement of the reference pipeline on a bounded sample."""
    import numpy as np  # noqa: F401

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import unitigcoarsening_amd as uc

    kmers = uc.medium_model(seed, sample_bp, k)
    order = oracle.generate_permutation(l, seed)
    t0 = time.perf_counter()
    text, stats = oracle.refpipe(kmers, k, l, B, order)
    wall = time.perf_counter() - t0
    with uc.Compactor(**gpu_ctx_kwargs) as c:
        got = c.compact(kmers, k, l)
    parity = uc.sorted_lines(got) == uc.sorted_lines(text)
    return {
        "value": kmers.shape[0] / wall,
        "unit": "k-mers/s",
        "cores": 1,
        "kind": "port",
        "sample": f"medium model at {sample_bp / 1e6:.0f} Mbp (seed {seed}): {kmers.shape[0]} canonical {k}-mers, "
                  f"l={l}, {B} buckets; oracle/refpipe.c (serial restatement of the reference pipeline), "
                  f"{wall:.1f} s",
        "phases_s": {kk: round(stats[kk], 3) for kk in ("bucketing", "assemble", "coarsen", "write_intra",
                                                          "minitigs", "cross_finish")},
        "parity_with_gpu": parity,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--genome-bp", type=int, default=50_000_000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--k", type=int, default=31)
    ap.add_argument("--l", type=int, default=11)
    ap.add_argument("--buckets", type=int, default=64)
    ap.add_argument("--mode", choices=["batched", "plain"], default="batched")
    ap.add_argument("--load-pct", type=int, default=50, help="hashmap load factor (percent)")
    ap.add_argument("--sample-log2", type=int, default=5, help="splitter sampling rate 2^-s")
    ap.add_argument("--xcd-map", action="store_true", help="XCD-contiguous tile order")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    ap.add_argument("--cpu-sample-bp", type=int, default=4_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the path shards by minimizer bucket "
              "across GPUs by design but this build runs and reports a single MI355X.", file=sys.stderr)
        sys.exit(2)

    ensure_built()
    import unitigcoarsening_amd as uc

    k, l, B = args.k, args.l, args.buckets
    t = time.time()
    kmers = uc.medium_model(args.seed, args.genome_bp, k)
    n = kmers.shape[0]
    W = 1 if k <= 32 else 2
    log(f"workload: {n} canonical {k}-mers from {args.genome_bp} bp in {time.time() - t:.1f}s")

    mode = uc.UC_MODE_PLAIN if args.mode == "plain" else uc.UC_MODE_BATCHED
    tun = {"load_pct": args.load_pct, "sample_log2": args.sample_log2}
    c = uc.Compactor(n_buckets=B, mode=mode, profile=True, xcd_map=args.xcd_map, **tun)
    c.load(kmers, k)  # one H2D copy; every step starts from HBM
    for i in range(args.warmup):
        c.compact_loaded(l)
        log(f"warmup {i + 1}/{args.warmup}: {c.timing().total_ms:.2f} ms")
    c.sync()
    c.kernel_stats_reset()
    phase = {"bucketize_ms": 0.0, "edges_ms": 0.0, "coarsen_ms": 0.0, "emit_ms": 0.0}
    t0 = time.perf_counter()
    for i in range(args.steps):
        c.compact_loaded(l)
        tm = c.timing()
        for key in phase:
            phase[key] += getattr(tm, key)
    c.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    ms_per_step = 1e3 * wall / args.steps
    value = n * args.steps / wall
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, {value / 1e6:.1f} M k-mers/s")

    b = c.bucket_info()
    ci = c.coarsen_info()
    ks = c.kernel_stats()
    dom = max(ks.items(), key=lambda kv: kv[1][0])
    dom_name, (dom_ms, dom_launches) = dom
    avg_ms = dom_ms / max(dom_launches, 1)
    # algorithmic bytes per launch: per-job bytes spread over the launches of
    # one job (batched: one launch per job; plain: one per bucket)
    per_job = kernel_bytes(dom_name, n, b.m_placements, b.n_cross, W, ci.n_unitigs, ci.text_bytes, k)
    launches_per_job = dom_launches / args.steps
    roof = None
    if per_job is not None:
        per_launch = per_job / launches_per_job
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for(dom_name),
                "kernel": dom_name, "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes_per_launch": int(per_launch)}
    kernels = {name: {"ms_per_step": round(ms / args.steps, 4), "launches_per_step": nl / args.steps}
               for name, (ms, nl) in sorted(ks.items(), key=lambda kv: -kv[1][0])}
    text_bytes = c.text_size()
    c.close()

    plain = None
    if not args.no_plain and mode == uc.UC_MODE_BATCHED:
        with uc.Compactor(n_buckets=B, mode=uc.UC_MODE_PLAIN, **tun) as cp:
            cp.load(kmers, k)
            cp.compact_loaded(l)
            cp.sync()
            tp = time.perf_counter()
            reps = max(1, min(args.steps, 3))
            for _ in range(reps):
                cp.compact_loaded(l)
            cp.sync()
            pw = (time.perf_counter() - tp) / reps
        plain = {"value": n / pw, "ms_per_step": round(1e3 * pw, 3), "batched_over_plain": round(pw / (wall / args.steps), 3)}
        log(f"plain form: {1e3 * pw:.2f} ms/step")

    base = None
    if not args.no_cpu_baseline:
        log("cpu baseline (serial restatement of the reference pipeline) ...")
        base = cpu_baseline(args.cpu_sample_bp, args.seed, k, l, B, {"n_buckets": B})
        log(f"cpu baseline: {base['value'] / 1e6:.3f} M k-mers/s, parity with GPU: {base['parity_with_gpu']}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"medium: {args.genome_bp / 1e6:.0f} Mbp seeded genome (seed {args.seed}), 5% repeats of "
                        f"2-10 kbp copies with 0.1% substitutions, 1% error k-mers; distinct canonical {k}-mers",
            "n_kmers": n, "k": k, "l": l, "buckets": B, "mode": args.mode, **tun,
            "m_placements": b.m_placements, "n_cross": b.n_cross, "n_unitigs": ci.n_unitigs,
            "text_bytes": text_bytes,
        },
        "phases_ms_per_step": {kk: round(v / args.steps, 3) for kk, v in phase.items()},
        "kernels": kernels,
        "roofline": roof,
        "plain": plain,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
