#!/usr/bin/env python3
"""Benchmark of the compaction path on one MI355X.

A step is one whole compaction of the medium workload -- 50 Mbp seeded
synthetic genome with 5% repeats (2-10 kbp copies, 0.1% substitutions)
and 1% error k-mers, every distinct canonical 31-mer (~48M), l = 11,
64 minimizer buckets -- from packed k-mers resident in HBM to the unitig
text in HBM: bucketing, edge finding, contraction, emission.

Prints one JSON line.  value = k-mers compacted per second (whole job).
roofline = the dominant kernel's algorithmic bytes per launch over its
average launch time (HIP events on the library's stream, around that kernel
only, during the timed steps).  The per-kernel breakdown ("kernels") comes
from the warmup steps, where every launch is timed.  cpu_baseline =
the serial restatement of the reference pipeline (oracle/refpipe.c) on a
bounded sample of the same model, on one host core, checked against the
GPU's output for that sample.  Beside it: "plain" (the same job with one
launch per bucket), "sets" (256 uneven sets, batched form vs per-set plain
runs, both from HBM) and "serve" (the request loop under Poisson arrivals:
p50 / p99 latency and k-mers/s).
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


def kernel_bytes(name, n, m, x, W, n_unitigs, text_bytes, k, runs):
    """Algorithmic HBM bytes of one launch of `name` over the whole job:
    every array the kernel must touch, once, at its natural width.
    n k-mers, m slots (k-mers + crossing duplicates), x crossing k-mers,
    8W bytes per packed k-mer, U unitigs, T text bytes, R runs."""
    kb = 8 * W
    table = {
        # read k-mer; write both fine bucket codes
        "k_lmin_count": n * (kb + 8),
        # read k-mer + codes; write slot k-mer, flag, fine index; pair list
        "k_lmin_scatter": n * (kb + 8) + m * (kb + 1 + 2) + x * 8,
        # read the fine index of every slot
        "k_sub_count": m * 2,
        # read fine index, k-mer, flag; write k-mer, flag, slot permutation
        "k_sub_scatter": m * (2 + kb + 1) + m * (kb + 1 + 4),
        # read both pair entries and their permuted slots; write both
        "k_pair_remap": x * (8 + 8 + 8),
        # read k-mer + flag per slot; write link, flag and run character per
        # slot, the has-predecessor bitmap and an 8-byte record per run
        "k_edge_lds": m * (kb + 1) + m * (4 + 1 + 1) + m // 8 + runs * 8,
        # fallback tables for oversized fine buckets (16-byte entries): the
        # same three steps over HBM
        "k_edge_insert": m * (1 + kb) + (m - x) * 16,
        "k_edge_lookup": m * (1 + kb + 4) + (m - x) * (16 + kb),
        "k_edge_resolve": m * (4 + 8) + (m - x) * 16 + m // 8,
        # pair entries, the duplicate's link, the lent k-mer's link (its run
        # head); write that link and its run record's next run
        "k_splice": x * (4 + 4 + 4 + 4 + 4 + 4),
        # read flag + has-predecessor bit per slot, one run record per run
        # (the list entries of heads of several runs are not counted)
        "k_seg_count": m * 1 + m // 8 + runs * 8,
        # read flag + has-predecessor bit per slot, head k-mer per head, one
        # record and the characters per run; write the text (list entries
        # of heads of several runs are not counted)
        "k_head_emit": m * 1 + m // 8 + n_unitigs * kb + runs * 8 + n + text_bytes,
    }
    return table.get(name)


def phase_roofline(n, m, k, text_bytes, ms):
    """Compulsory HBM bytes of each phase by the reference's data layout
    (W = 4 * ceil(k / 16) bytes per packed k-mer, T text bytes) over the
    phase's time: bucketing B1 = nW + mW, edge finding B2 = m(4W + 4),
    contraction B3 = 12m, emission B4 = mW + T; contraction together with
    emission is B3 + B4."""
    W = 4 * ((k + 15) // 16)
    rows = {"bucketing": (n * W + m * W, ms["bucketize_ms"]), "edge_finding": (m * (4 * W + 4), ms["edges_ms"]),
            "contraction": (12 * m, ms["coarsen_ms"]), "emission": (m * W + text_bytes, ms["emit_ms"]),
            "contraction_and_emission": (12 * m + m * W + text_bytes, ms["coarsen_ms"] + ms["emit_ms"])}
    out = {}
    for name, (nb, t) in rows.items():
        gbs = nb / (t * 1e-3) / 1e9 if t > 0 else None
        out[name] = {"bytes": int(nb), "ms": round(t, 3), "achieved_gbs": round(gbs, 1) if gbs else None,
                     "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
    return out


def traffic_for(name, n, k, l, buckets):
    """Measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration) when a summary of this same
    workload was committed, else None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    w = t.get("workload", {})
    if (w.get("n_kmers"), w.get("k"), w.get("l"), w.get("buckets")) != (n, k, l, buckets):
        return None
    return t.get("kernels", {}).get(name, {}).get("bytes_per_launch")


def workload(args, genome_bp=None):
    """(k-mers, description) of the configured workload."""
    import unitigcoarsening_amd as uc

    bp = genome_bp or args.genome_bp
    if args.large:
        return (uc.large_model(args.seed, bp, args.k),
                f"large instance at reduced size: {bp / 1e6:.0f} Mbp seeded genome (seed {args.seed}), 45% repeat "
                f"families (300 bp - 6 kbp segments, copy number 2-1000, divergence 0.5-15%), 0.5% error k-mers; "
                f"distinct canonical {args.k}-mers")
    return (uc.medium_model(args.seed, bp, args.k),
            f"medium: {bp / 1e6:.0f} Mbp seeded genome (seed {args.seed}), 5% repeats of 2-10 kbp copies with 0.1% "
            f"substitutions, 1% error k-mers; distinct canonical {args.k}-mers")


def cpu_threads():
    """Host threads of the CPU baseline: every core of the GPU's share."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else n


def refpipe_leg(oracle, uc, kmers, k, l, B, seed, threads, gpu_text=None, gpu_kwargs=None):
    """One run of the restated reference pipeline (OpenMP, `threads`) with its
    parity against the GPU's text of the same k-mers (sorted lines)."""
    order = oracle.generate_permutation(l, seed)
    t0 = time.perf_counter()
    text, st = oracle.refpipe(kmers, k, l, B, order, threads=threads)
    wall = time.perf_counter() - t0
    if gpu_text is None:
        with uc.Compactor(**(gpu_kwargs or {"n_buckets": B})) as c:
            gpu_text = c.compact(kmers, k, l)
    parity = uc.sorted_lines(gpu_text) == uc.sorted_lines(text)
    return wall, st, parity


def cpu_baseline(args, l, B, gpu_ctx_kwargs, sets_data=None):
    """The restated reference pipeline (oracle/refpipe.c) on the host's cores:
    orc_refpipe_par, OpenMP threads over the minimizers and over whole
    buckets, timed at the reference's phase boundaries (driver.cpp:443,506-525),
    on the bench workload's sample, on the 1 Mbp uniform workload and on the
    256-set batch run set after set; beside it the serial form on the sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import unitigcoarsening_amd as uc

    threads = cpu_threads()
    sample_bp, seed, k = args.cpu_sample_bp, args.seed, args.k
    kmers, desc = workload(args, sample_bp)
    wall, st, parity = refpipe_leg(oracle, uc, kmers, k, l, B, seed, threads, gpu_kwargs=gpu_ctx_kwargs)
    log(f"cpu baseline, {threads} threads: {kmers.shape[0] / wall / 1e6:.2f} M k-mers/s, parity {parity}")
    t0 = time.perf_counter()
    stext, sst = oracle.refpipe(kmers, k, l, B, oracle.generate_permutation(l, seed))
    swall = time.perf_counter() - t0
    serial = {"value": kmers.shape[0] / swall, "unit": "k-mers/s", "cores": 1, "wall_s": round(swall, 3),
              "phases_s": {kk: round(sst[kk], 3) for kk in ("bucketing", "assemble", "coarsen", "write_intra",
                                                              "minitigs", "cross_finish")}}
    del stext
    out = {
        "value": kmers.shape[0] / wall,
        "unit": "k-mers/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{desc}: {kmers.shape[0]} k-mers, l={l}, {B} buckets; oracle/refpipe.c, the restated "
                  f"reference pipeline on {threads} OpenMP threads (orc_refpipe_par), {wall:.2f} s",
        "phases_s": {"bucketing": round(st["bucketing"], 3), "buckets": round(st["buckets_wall"], 3),
                     "cross_finish": round(st["cross_finish"], 3), "total": round(st["total"], 3)},
        "bucket_thread_s": {kk: round(st[kk], 3) for kk in ("assemble", "coarsen", "write_intra", "minitigs")},
        "parity_with_gpu": parity,
        "serial": serial,
        "host": host_info(),
    }
    if not args.large:
        # the 1 Mbp uniform workload: 31-mers, l = 11, 16 buckets
        c1 = uc.synth_kmers(1, 1_000_000, 31)
        w1, _, p1 = refpipe_leg(oracle, uc, c1, 31, 11, 16, 1, threads)
        out["uniform_1mbp"] = {"value": c1.shape[0] / w1, "n_kmers": int(c1.shape[0]), "wall_s": round(w1, 3),
                               "cores": threads, "parity_with_gpu": p1}
        log(f"cpu baseline 1 Mbp uniform: {c1.shape[0] / w1 / 1e6:.2f} M k-mers/s, parity {p1}")
    if sets_data is not None:
        # the 256-set batch, one set after another, against the GPU's batched texts
        sets, texts = sets_data
        order = oracle.generate_permutation(l, seed)
        tot, wall3, same = 0, 0.0, 0
        for km, gt in zip(sets, texts):
            t0 = time.perf_counter()
            text, _ = oracle.refpipe(km, k, l, B, order, threads=threads)
            wall3 += time.perf_counter() - t0
            tot += km.shape[0]
            same += uc.sorted_lines(text) == uc.sorted_lines(gt)
        out["sets_256"] = {"value": tot / wall3, "n_kmers": tot, "wall_s": round(wall3, 3), "cores": threads,
                           "parity_with_gpu": f"{same} of {len(sets)} sets"}
        log(f"cpu baseline 256 sets: {tot / wall3 / 1e6:.2f} M k-mers/s, {same} of {len(sets)} sets match")
    return out


def host_info():
    """nproc, CPU model and RAM of the host the CPU baseline ran on."""
    info = {"nproc": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            info["cpu_model"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
        with open("/proc/meminfo") as f:
            kb = next((int(ln.split()[1]) for ln in f if ln.startswith("MemTotal")), 0)
        info["ram_gib"] = round(kb / 2 ** 20, 1)
    except OSError:
        pass
    return info


def sets_workload(n_sets, seed0, k, median=200_000, sigma=1.0, lo=10_000, hi=5_000_000):
    """256 independent sets, sizes ~ LogNormal(ln median, sigma) clipped to
    [lo, hi] k-mers, each from the medium repeat/error model (seeds seed0..)."""
    import numpy as np
    import unitigcoarsening_amd as uc

    rng = np.random.default_rng(seed0)
    sizes = np.clip(np.exp(rng.normal(np.log(median), sigma, n_sets)), lo, hi).astype(np.int64)
    # ~0.97 distinct canonical k-mers per genome base under the medium model
    return [uc.medium_model(seed0 + i, int(sz / 0.97) + k, k) for i, sz in enumerate(sizes)]


def bench_sets(args, keep=None):
    """Batched form (one run over all sets) vs the plain form run set by set,
    both from k-mers already resident in HBM: the sets are loaded once, the
    plain runs take device pointers into that buffer."""
    import ctypes as C

    import unitigcoarsening_amd as uc

    k, l, B = args.k, args.l, args.buckets
    t = time.time()
    sets = sets_workload(args.n_sets, 1000, k)
    sizes = [s.shape[0] for s in sets]
    n = sum(sizes)
    log(f"sets workload: {len(sets)} sets, {n} k-mers in {time.time() - t:.1f}s")
    import numpy as np

    allk = np.ascontiguousarray(np.concatenate(sets, axis=0))
    keep_sets = sets if keep is not None else None
    del sets
    tun = {"load_pct": args.load_pct, "sample_log2": args.sample_log2}
    L = uc.lib()
    W = 1 if k <= 32 else 2
    with uc.Compactor(n_buckets=B, **tun) as cb:
        cb.load(allk, k, sizes)
        for _ in range(max(args.warmup, 1)):
            cb.compact_loaded(l)
        cb.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            cb.compact_loaded(l)
        cb.sync()
        batched = (time.perf_counter() - t0) / args.steps
        log(f"batched: {batched * 1e3:.2f} ms for all sets")
        base = L.uc_loaded_device_ptr(cb._ctx)
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        with uc.Compactor(n_buckets=B, mode=uc.UC_MODE_PLAIN, **tun) as cp:
            def run_plain():
                for i, sz in enumerate(sizes):
                    rc = L.uc_compact(cp._ctx, C.c_void_p(base + int(offs[i]) * 8 * W), sz, k, l, 1)
                    if rc != 0:
                        raise uc.UcError(rc, L.uc_last_error(cp._ctx).decode())
                cp.sync()
            run_plain()
            t0 = time.perf_counter()
            reps = max(1, min(args.steps, 2))
            for _ in range(reps):
                run_plain()
            plain = (time.perf_counter() - t0) / reps
            # the outputs, untimed: every set's batched text against its plain run
            batched_text = cb.text()
            bounds = [L.uc_set_text_offset(cb._ctx, i) for i in range(len(sizes) + 1)]
            equal = 0
            for i, sz in enumerate(sizes):
                rc = L.uc_compact(cp._ctx, C.c_void_p(base + int(offs[i]) * 8 * W), sz, k, l, 1)
                if rc != 0:
                    raise uc.UcError(rc, L.uc_last_error(cp._ctx).decode())
                equal += cp.text() == batched_text[bounds[i]:bounds[i + 1]]
    log(f"plain per-bucket, set by set: {plain * 1e3:.2f} ms; {equal} of {len(sizes)} set texts byte-equal")
    if keep is not None:  # the sets and their batched texts, for the CPU baseline's parity
        keep.append((keep_sets, [batched_text[bounds[i]:bounds[i + 1]] for i in range(len(sizes))]))
    if equal != len(sizes):
        raise RuntimeError(f"batched and plain texts differ on {len(sizes) - equal} sets")
    return {"workload": f"{len(sizes)} sets, log-normal sizes (median 200k, sigma 1, clipped 10k-5M k-mers), "
                        f"medium repeat/error model, seeds 1000..{1000 + len(sizes) - 1}, k={k}, l={l}, {B} buckets",
            "n_kmers": n, "batched_ms": round(batched * 1e3, 3), "plain_ms": round(plain * 1e3, 3),
            "batched_value": n / batched, "plain_value": n / plain, "batched_over_plain": round(plain / batched, 3),
            "texts_byte_equal": f"{equal} of {len(sizes)} sets",
            "note": "inputs resident in HBM for both forms; the plain form runs uc_compact per set on device "
                    "pointers with one launch per bucket; after timing, every set's batched text is compared "
                    "byte for byte with its plain run"}


def serve_pool(n_inputs, seed0, lo=100_000, hi=50_000_000):
    """Distinct job inputs for the request loop: sizes log-uniform in
    [lo, hi] k-mers, k drawn from {31, 63}, medium repeat/error model."""
    import numpy as np
    import unitigcoarsening_amd as uc

    rng = np.random.default_rng(seed0)
    pool = []
    for i in range(n_inputs):
        k = int(rng.choice([31, 63]))
        size = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        pool.append((uc.medium_model(seed0 + i, int(size / 0.97) + k, k), k, 11 if k == 31 else 13))
    return pool


def copy_ceiling():
    """Stream-copy ceiling of the device: the best of five 1 GiB copies by
    uc_copy_probe (a plain copy kernel), read + write bytes per second."""
    import subprocess

    import unitigcoarsening_amd as uc

    exe = os.path.join(os.path.dirname(uc.LIB_PATH), "uc_copy_probe")
    try:
        r = subprocess.run([exe, "1024", "5"], capture_output=True, text=True, timeout=60)
    except (OSError, subprocess.SubprocessError):
        return None
    rates = [float(line.split()[-2]) for line in r.stdout.splitlines() if line.startswith("copy ")]
    return max(rates[1:] or rates) if r.returncode == 0 and rates else None


def bench_file(args, kmers, l, B):
    """End to end, file to file: the workload as reference k-mer text in a
    file, read and parsed on the device (pinned chunks), compacted, and the
    unitig text appended to an output file."""
    import tempfile

    import unitigcoarsening_amd as uc

    k = args.k
    text = uc.format_text(kmers, k)
    d = tempfile.mkdtemp(prefix="uc_bench_")
    src, dst = os.path.join(d, "kmers.txt"), os.path.join(d, "unitigs.txt")
    try:
        with open(src, "wb") as f:
            f.write(text)
        in_bytes = len(text)
        del text
        with uc.Compactor(n_buckets=B) as c:
            c.load_file(src, k)  # warm: page cache, pinned buffers, device buffers
            c.compact_loaded(l)
            c.sync()
            t0 = time.perf_counter()
            c.load_file(src, k)
            c.compact_loaded(l)
            c.append_text_file(dst)
            c.sync()
            wall = time.perf_counter() - t0
        out_bytes = os.path.getsize(dst)
        return {"ms": round(wall * 1e3, 2), "kmers_per_s": kmers.shape[0] / wall, "input_bytes": in_bytes,
                "output_bytes": out_bytes,
                "note": "reference text in, unitig text out, on the device's host; the input file is in the page "
                        "cache (read once before); the output is appended to an empty file"}
    except (OSError, RuntimeError) as e:
        log(f"file to file skipped: {e}")
        return None
    finally:
        for f in (src, dst):
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(d)


SERVE_CLIENTS = 4  # client threads that submit the arrivals (each submission copies its input)


def serve_run(args, pool, sizes, n_slots, load, n_jobs, mean_svc, seed):
    """Poisson arrivals at `load` x one worker's capacity onto a server of
    n_slots workers, submitted by SERVE_CLIENTS client threads; latency =
    submit to result of every job."""
    import queue
    import threading

    import numpy as np
    import unitigcoarsening_amd as uc

    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(pool), n_jobs)
    rate = load / mean_svc  # jobs per second offered
    gaps = rng.exponential(1.0 / rate, n_jobs)
    lat = []
    with uc.Server(n_slots=n_slots, hbm_budget=int(args.serve_budget_gb * 1e9), n_buckets=args.buckets,
                   sample_log2=args.sample_log2) as srv:
        # steady state: every input once per slot, largest first, so each
        # worker context has grown its buffers before the timed arrivals
        order = sorted(pool, key=lambda p: -uc.job_bytes_estimate(p[0].shape[0], p[1]))
        warm = [srv.submit(km, k, l) for km, k, l in order for _ in range(n_slots)]
        for i in warm:
            srv.result(i, text=False)
        ids = [None] * n_jobs
        waiter_err = []

        def waiter():
            i = 0
            while i < n_jobs:
                while ids[i] is None:
                    time.sleep(0.0002)
                try:
                    # the text stays in the server: this measures the server, not
                    # the client's copy of each result
                    _, qms, rms = srv.result(ids[i], text=False)
                    lat.append(qms + rms)
                except Exception as e:  # reported after the run
                    waiter_err.append(repr(e))
                i += 1

        arrivals = queue.Queue()

        def client():
            while True:
                j = arrivals.get()
                if j is None:
                    return
                km, k, l = pool[picks[j]]
                ids[j] = srv.submit(km, k, l)  # (the copy runs with the GIL released)

        clients = [threading.Thread(target=client) for _ in range(SERVE_CLIENTS)]
        th = threading.Thread(target=waiter)
        for c in clients:
            c.start()
        th.start()
        t0 = time.perf_counter()
        nxt = t0
        for j in range(n_jobs):
            nxt += gaps[j]
            dt = nxt - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            arrivals.put(j)
        for _ in clients:
            arrivals.put(None)
        for c in clients:
            c.join()
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
    if waiter_err:
        raise RuntimeError(f"serving jobs failed: {waiter_err[:3]}")
    lat = np.array(lat)
    served = int(sizes[picks].sum())
    log(f"serve {n_slots} slot(s) at {load:.2f}x: {n_jobs} jobs in {wall:.2f}s, {served / wall / 1e6:.1f} M k-mers/s, "
        f"p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms")
    return {"n_slots": n_slots, "offered_load": load, "offered_jobs_per_s": round(rate, 2), "jobs": n_jobs,
            "kmers_per_s": served / wall, "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "max_ms": round(float(lat.max()), 3),
            "peak_held_gib": round(st["peak_held_bytes"] / 2 ** 30, 3)}


def bench_serve(args):
    """Request loop (uc_server) at the serving configuration: inputs of
    log-uniform 100k-50M k-mers with k drawn from {31, 63}, an HBM budget
    of 288 GB, Poisson arrivals at a fixed fraction of one worker's
    capacity; then saturation points (1.5x and 3x that capacity) with one
    worker slot and with four."""
    import numpy as np
    import unitigcoarsening_amd as uc

    t = time.time()
    pool = serve_pool(args.serve_inputs, 5000)
    sizes = np.array([p[0].shape[0] for p in pool])
    log(f"serve pool: {len(pool)} inputs, {sizes.sum()} k-mers, sizes {sizes.min()}..{sizes.max()} "
        f"in {time.time() - t:.1f}s")
    # one worker's capacity: a one-slot server, jobs back to back after one
    # warm-up pass over the inputs, each job's run time (the worker's copy of
    # the input to the device, the job, the text copied out to host memory)
    svc = []
    with uc.Server(n_slots=1, hbm_budget=int(args.serve_budget_gb * 1e9), n_buckets=args.buckets,
                   sample_log2=args.sample_log2) as srv:
        for km, k, l in pool:
            srv.result(srv.submit(km, k, l), text=False)
        for km, k, l in pool:
            _, _, rms = srv.result(srv.submit(km, k, l), text=False)
            svc.append(rms / 1e3)
    mean_svc = float(np.mean(svc))
    main_run = serve_run(args, pool, sizes, args.serve_slots, args.serve_load, args.serve_jobs, mean_svc, 6)
    sat = None
    if args.serve_saturation_jobs:
        sat = {}
        for load in (1.5, 3.0):
            pt = {f"slots_{ns}": serve_run(args, pool, sizes, ns, load, args.serve_saturation_jobs, mean_svc, 7)
                  for ns in (1, args.serve_slots)}
            pt["slots_ratio"] = round(pt[f"slots_{args.serve_slots}"]["kmers_per_s"] / pt["slots_1"]["kmers_per_s"], 3)
            sat[f"load_{load:g}x"] = pt
    return {"workload": f"{args.serve_jobs} jobs, Poisson arrivals at {args.serve_load:.2f} x one worker's capacity "
                        f"(submitted by {SERVE_CLIENTS} client threads), "
                        f"drawn from {len(pool)} inputs of log-uniform 100k-50M k-mers (medium model, k drawn from "
                        f"{{31, 63}}, seeds 5000..{5000 + len(pool) - 1}), {args.serve_slots} worker slots, budget "
                        f"{args.serve_budget_gb:g} GB",
            "pool_kmers": [int(x) for x in sizes], "pool_k": [p[1] for p in pool],
            "mean_service_ms_one_worker": round(1e3 * mean_svc, 3),
            **{kk: v for kk, v in main_run.items() if kk not in ("n_slots",)},
            "saturation": sat,
            "note": "latency is submit to finish and includes the server's copies of the input to the device and of "
                    "the result text to host memory (the client does not fetch the texts here); the saturation points "
                    f"offer 1.5x and 3x one worker's capacity to one worker slot and to {args.serve_slots}: served "
                    "k-mers/s above the one-slot figure is work overlapped across the workers (device, copies and host "
                    "threads)"}


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
    ap.add_argument("--sample-log2", type=int, default=0, help="splitter sampling rate 2^-s (0: library default)")
    ap.add_argument("--xcd-map", action="store_true", help="XCD-contiguous tile order")
    ap.add_argument("--autotune", action="store_true", help="tune the settings on the workload before timing")
    ap.add_argument("--tune-buckets", action="store_true", help="let the tuner change the bucket count too")
    ap.add_argument("--direct-scatter", action="store_true", help="bucketing scatters without LDS staging")
    ap.add_argument("--fine-slots", type=int, default=0)
    ap.add_argument("--group-slots", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    ap.add_argument("--cpu-sample-bp", type=int, default=50_000_000,
                    help="genome size of the CPU baseline's sample (default: the whole medium workload, ~10 s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sets", action="store_true", help="skip the 256-set batched vs plain comparison")
    ap.add_argument("--n-sets", type=int, default=256)
    ap.add_argument("--no-serve", action="store_true", help="skip the request-loop latency benchmark")
    ap.add_argument("--no-file", action="store_true", help="skip the file-to-file run")
    ap.add_argument("--eager", action="store_true", help="launch the contraction pass kernel by kernel, no graph")
    ap.add_argument("--no-ceiling", action="store_true", help="skip the stream-copy ceiling measurement")
    ap.add_argument("--serve-jobs", type=int, default=300)
    ap.add_argument("--serve-inputs", type=int, default=16, help="distinct job inputs drawn from")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-load", type=float, default=0.7, help="offered load, x one worker's capacity")
    ap.add_argument("--serve-saturation-jobs", type=int, default=300,
                    help="jobs of each saturation run (1.5x load, 1 slot and --serve-slots); 0 skips them")
    ap.add_argument("--serve-budget-gb", type=float, default=288.0)
    ap.add_argument("--large", action="store_true",
                    help="the large instance's model (k=63, l=13, 256 buckets, repeat families) at --genome-bp")
    ap.add_argument("--low-memory", action="store_true",
                    help="UC_FLAG_LOW_MEMORY: free every buffer after its last use (the input is consumed, so it is "
                         "loaded again before each step, outside the timed region); on by default with --large")
    args = ap.parse_args()
    if args.large:
        # defaults of the large instance unless given explicitly
        if "--k" not in sys.argv:
            args.k = 63
        if "--l" not in sys.argv:
            args.l = 13
        if "--buckets" not in sys.argv:
            args.buckets = 256
        if "--seed" not in sys.argv:
            args.seed = 4
        if "--genome-bp" not in sys.argv:
            args.genome_bp = 300_000_000
        args.no_sets = args.no_serve = True
        args.low_memory = True

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the path shards by minimizer bucket "
              "across GPUs by design but this library runs and reports a single MI355X.", file=sys.stderr)
        sys.exit(2)

    ensure_built()
    import unitigcoarsening_amd as uc

    k, l, B = args.k, args.l, args.buckets
    t = time.time()
    kmers, desc = workload(args)
    n = kmers.shape[0]
    W = 1 if k <= 32 else 2
    log(f"workload: {n} canonical {k}-mers from {args.genome_bp} bp in {time.time() - t:.1f}s")

    mode = uc.UC_MODE_PLAIN if args.mode == "plain" else uc.UC_MODE_BATCHED
    tun = {"load_pct": args.load_pct, "sample_log2": args.sample_log2, "fine_slots": args.fine_slots,
           "group_slots": args.group_slots, "direct_scatter": args.direct_scatter}
    lowmem = args.low_memory
    c = uc.Compactor(n_buckets=B, mode=mode, profile=True, xcd_map=args.xcd_map, eager=args.eager,
                     low_memory=lowmem, **tun)
    c.load(kmers, k)  # one H2D copy; every step starts from HBM

    fresh = [True]  # the loaded input has not been consumed by a job yet

    def reload():
        # a low-memory job consumes its loaded input: copy it in again, untimed
        if lowmem and not fresh[0]:
            c.load(kmers, k)
            c.sync()
        fresh[0] = not lowmem
    tune = None
    if args.autotune:
        t = time.time()
        tune = c.autotune(l, keep_buckets=not args.tune_buckets)
        log(f"autotune: {tune} in {time.time() - t:.1f}s")
        tun = {"load_pct": tune["load_pct"], "sample_log2": tune["sample_log2"], "fine_slots": tune["fine_slots"],
               "group_slots": tune["group_slots"], "direct_scatter": args.direct_scatter}
        B = tune["n_buckets"]
    # warmup: every kernel timed, for the per-kernel breakdown and to find the
    # dominant kernel; the timed steps then carry events around that kernel
    # only (events around all ~35 launches cost about 5% of a step)
    n_warm = max(args.warmup, 1)
    for i in range(n_warm):
        if i == 1:
            c.kernel_stats_reset()  # the first step carries one-time costs (code loading, allocation)
        reload()
        c.compact_loaded(l)
        log(f"warmup {i + 1}/{n_warm}: {c.timing().total_ms:.2f} ms")
    c.sync()
    ks = c.kernel_stats()
    warm_steps = n_warm - 1 if n_warm > 1 else 1
    dom_name = max(ks.items(), key=lambda kv: kv[1][0])[0]
    c.kernel_stats_select(dom_name)
    reload()
    c.compact_loaded(l)  # one more untimed step: the contraction pass is captured as a graph here
    c.sync()
    c.kernel_stats_reset()
    phase = {"bucketize_ms": 0.0, "edges_ms": 0.0, "coarsen_ms": 0.0, "emit_ms": 0.0}
    peak = 0
    if not lowmem:
        t0 = time.perf_counter()
        for i in range(args.steps):
            c.compact_loaded(l)
            tm = c.timing()
            peak = max(peak, tm.device_peak_bytes)
            for key in phase:
                phase[key] += getattr(tm, key)
        c.sync()
        wall = time.perf_counter() - t0
    else:
        # each step times its own job; the input copy before it is not timed
        wall = 0.0
        for i in range(args.steps):
            reload()
            t0 = time.perf_counter()
            c.compact_loaded(l)
            c.sync()
            wall += time.perf_counter() - t0
            tm = c.timing()
            peak = max(peak, tm.device_peak_bytes)
            for key in phase:
                phase[key] += getattr(tm, key)
    ms_per_step = 1e3 * wall / args.steps
    value = n * args.steps / wall
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, {value / 1e6:.1f} M k-mers/s")

    b = c.bucket_info()
    ci = c.coarsen_info()
    edges_ms = phase["edges_ms"] / args.steps
    dom_ms, dom_launches = c.kernel_stats()[dom_name]
    avg_ms = dom_ms / max(dom_launches, 1)
    # algorithmic bytes per launch: per-job bytes spread over the launches of
    # one job (batched: one launch per job; plain: one per bucket)
    per_job = kernel_bytes(dom_name, n, b.m_placements, b.n_cross, W, ci.n_unitigs, ci.text_bytes, k, ci.n_runs)
    launches_per_job = dom_launches / args.steps
    roof = None
    if per_job is not None:
        per_launch = per_job / launches_per_job
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for(dom_name, n, k, l, B),
                "kernel": dom_name, "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes_per_launch": int(per_launch)}
    kernels = {name: {"ms_per_step": round(ms / warm_steps, 4), "launches_per_step": nl / warm_steps}
               for name, (ms, nl) in sorted(ks.items(), key=lambda kv: -kv[1][0])}
    text_bytes = c.text_size()
    device_gib = round(c.held_bytes() / 2 ** 30, 2)
    device_peak_gib = round(peak / 2 ** 30, 2)
    c.close()

    # hash-table probes of the same job, counted in a step of its own (the
    # timed steps count nothing)
    with uc.Compactor(n_buckets=B, mode=mode, xcd_map=args.xcd_map, eager=args.eager, count_probes=True,
                      low_memory=lowmem, **tun) as cp:
        cp.load(kmers, args.k)
        cp.compact_loaded(l)
        probes = cp.edge_probes()
    hashmap = {"probes_per_step": probes, "keys_per_step": 2 * (b.m_placements - b.n_cross),
               "probes_per_s": probes / (edges_ms * 1e-3) if edges_ms > 0 else None,
               "note": "slot inspections of the insert and lookup probes over every hash table of the edge phase, "
                       "counted in one extra step, over the timed steps' edge-phase time (HIP events); keys = "
                       "inserts (all but borrowed copies) + lookups (all but lent copies)"}

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

    sets_keep = [] if not args.no_cpu_baseline else None
    sets = None if args.no_sets else bench_sets(args, sets_keep)
    serve = None if args.no_serve else bench_serve(args)
    f2f = None if args.no_file else bench_file(args, kmers, l, B)
    if f2f:
        log(f"file to file: {f2f['ms']:.1f} ms, {f2f['kmers_per_s'] / 1e6:.1f} M k-mers/s")
    ceiling = None if args.no_ceiling else copy_ceiling()
    if roof is not None and ceiling:
        roof["copy_ceiling"] = {"achieved_copy_gbs": round(ceiling, 1), "frac": round(roof["achieved"] / ceiling, 4),
                                "note": "stream-copy ceiling of this device (uc_copy_probe, 1 GiB read + write)"}

    base = None
    if not args.no_cpu_baseline:
        log("cpu baseline (serial restatement of the reference pipeline) ...")
        base = cpu_baseline(args, l, B, {"n_buckets": B}, sets_keep[0] if sets_keep else None)
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
            "workload": desc,
            "n_kmers": n, "k": k, "l": l, "buckets": B, "mode": args.mode, **tun, "autotuned": tune,
            "m_placements": b.m_placements, "n_cross": b.n_cross, "n_unitigs": ci.n_unitigs,
            "text_bytes": text_bytes, "n_fine_buckets": b.n_fine_buckets,
            "largest_fine_bucket": b.largest_fine_bucket, "bucket_tile": b.bucket_tile, "fine_tile": b.fine_tile,
            "lds_limit": b.lds_limit, "n_oversized_fine": b.n_oversized, "oversized_slots": b.oversized_slots,
            "n_splitters": ci.n_splitters, "chain_rounds": ci.chain_rounds, "device_gib_held": device_gib,
            "device_gib_peak": device_peak_gib, "low_memory": lowmem,
            "n_runs": ci.n_runs,
        },
        "phases_ms_per_step": {kk: round(v / args.steps, 3) for kk, v in phase.items()},
        "kernels": kernels,
        "kernels_note": f"per-kernel times from {warm_steps} warmup step(s) after the first, every launch timed; the "
                        f"timed steps carry events around {dom_name} only",
        "roofline": roof,
        "plain": plain,
        "sets": sets,
        "phase_roofline": phase_roofline(n, b.m_placements, args.k, text_bytes,
                                         {kk: v / args.steps for kk, v in phase.items()}),
        "hashmap": hashmap,
        "serve": serve,
        "file_to_file": f2f,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
