"""Multi-core timing of the serial restatement of the reference pipeline
(refpipe.c): one independent sample per worker process, all released
together by a barrier.  Benchmark infrastructure only: nothing in the
product imports it, and it never touches a GPU.

The reference's own CPU path (Kokkos OpenMP) parallelises inside one job;
this measures what the host's cores do over as many independent jobs, the
form of the batched workload.  Usage:

    python oracle/parallel_baseline.py WORKERS GENOME_BP K L BUCKETS SEED MODEL

prints one JSON object: k-mers, wall seconds from the first start to the
last finish, and k-mers per second.
"""
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _worker(seed, genome_bp, k, l, buckets, model, barrier, out):
    sys.path.insert(0, HERE)
    sys.path.insert(0, ROOT)
    import oracle
    import unitigcoarsening_amd as uc

    gen = uc.large_model if model == "large" else uc.medium_model
    kmers = gen(seed, genome_bp, k)
    order = oracle.generate_permutation(l, seed)
    barrier.wait(timeout=600)
    t0 = time.perf_counter()
    oracle.refpipe(kmers, k, l, buckets, order)
    out.put((int(kmers.shape[0]), t0, time.perf_counter()))


def main():
    workers, genome_bp, k, l, buckets, seed = (int(x) for x in sys.argv[1:7])
    model = sys.argv[7] if len(sys.argv) > 7 else "medium"
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(workers)
    out = ctx.Queue()
    procs = [ctx.Process(target=_worker, args=(seed + i, genome_bp, k, l, buckets, model, barrier, out))
             for i in range(workers)]
    for p in procs:
        p.start()
    res = [out.get(timeout=1200) for _ in procs]
    for p in procs:
        p.join()
        if p.exitcode != 0:
            sys.exit(f"worker exited with {p.exitcode}")
    n = sum(r[0] for r in res)
    wall = max(r[2] for r in res) - min(r[1] for r in res)
    print(json.dumps({"kmers": n, "wall_s": wall, "kmers_per_s": n / wall, "workers": workers}))


if __name__ == "__main__":
    main()
