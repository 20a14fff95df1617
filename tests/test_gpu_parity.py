"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Bar: bit-exact.  Unitig text is compared as sorted lines (the reference's
line order is nondeterministic, driver.cpp:512,617-639); per-phase results
are compared as exact sets.
"""
import random

import numpy as np
import pytest

import oracle
import unitigcoarsening_amd as uc
from brute import brute_unitigs, check_partition

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = uc.Compactor(n_buckets=64)
    yield c
    c.close()


def textbook_sorted(kmers, k):
    text, _, _ = oracle.textbook(kmers, k)
    return uc.sorted_lines(text)


def run(kmers, k, l, **kw):
    with uc.Compactor(**kw) as c:
        return c.compact(kmers, k, l)


def rand_set(rng, n, k, alphabet="ACGT"):
    s = set()
    cap = len(alphabet) ** k
    n = min(n, cap)
    while len(s) < n:
        s.add("".join(rng.choice(alphabet) for _ in range(k)))
    return sorted(s)


# ---------------------------------------------------------------- hand cases


@pytest.mark.parametrize("kmers,expect", [
    (["ACGT", "CGTA", "GTAC"], ["ACGTAC"]),
    (["ACGT", "CGTA", "CGTC"], ["ACGT", "CGTA", "CGTC"]),
    (["ACGT", "TCGT", "CGTA"], ["ACGT", "CGTA", "TCGT"]),
    (["AAAC"], ["AAAC"]),
    (["GACG", "CGAC", "ACGA"], ["ACGACG"]),          # isolated 3-cycle, from its smallest k-mer
    (["AAAA"], ["AAAA"]),                            # self-loop
    (["AAAA", "CAAA"], ["AAAA", "CAAA"]),            # self-edge pruned by a 2nd predecessor
])
def test_hand_cases(kmers, expect):
    log(f"hand case {kmers}")
    for B in (1, 4):
        got = run(uc.encode(kmers, 4), 4, 2, n_buckets=B)
        assert uc.sorted_lines(got) == sorted(expect)


def test_empty_input(ctx):
    got = ctx.compact(np.zeros((0, 2), np.uint32), 31, 11)
    assert got == b""


# ---------------------------------------------------------------- random sets


@pytest.mark.parametrize("seed", range(12))
def test_random_small_alphabets(seed):
    # dense branching, bubbles and cycles
    rng = random.Random(seed)
    k = rng.choice([3, 4, 5, 6, 8])
    alphabet = rng.choice(["AC", "ACG", "ACGT"])
    kmers = rand_set(rng, rng.randint(1, 300), k, alphabet)
    l = rng.randint(1, k - 1)
    B = rng.choice([1, 2, 4, 16, 64])
    log(f"random seed={seed} k={k} l={l} B={B} n={len(kmers)}")
    got = run(uc.encode(kmers, k), k, l, n_buckets=B, seed=seed + 1)
    lines = got.decode().splitlines()
    assert sorted(lines) == sorted(brute_unitigs(kmers))
    check_partition(kmers, lines, k)


@pytest.mark.parametrize("k", [2, 7, 15, 16, 17, 20, 31, 32, 33, 40, 47, 48, 49, 63, 64])
def test_genome_all_k(k):
    rng = np.random.default_rng(k)
    kmers = uc.synth_kmers(int(rng.integers(1 << 30)), 40_000, k, repeat_frac=0.05, seg_min=200, seg_max=800,
                           repeat_sub=0.01, error_frac=0.01)
    l = min(11, k - 1)
    log(f"genome k={k} l={l} n={kmers.shape[0]}")
    got = run(kmers, k, l)
    assert uc.sorted_lines(got) == textbook_sorted(kmers, k)


@pytest.mark.parametrize("k,l", [(16, 15), (64, 15), (33, 15), (3, 2), (2, 1)])
def test_extreme_minimizer_lengths(k, l):
    # the largest l (15: 30-bit l-mer codes) and l = k - 1 (two l-mers per
    # k-mer, so both minimizers come from the end l-mers)
    kmers = uc.synth_kmers(k * 7 + l, 30_000, k, repeat_frac=0.05, seg_min=200, seg_max=800, repeat_sub=0.01,
                           error_frac=0.01)
    log(f"extreme l: k={k} l={l} n={kmers.shape[0]}")
    for B in (1, 64):
        got = run(kmers, k, l, n_buckets=B)
        assert uc.sorted_lines(got) == textbook_sorted(kmers, k)


@pytest.mark.parametrize("k,l,B", [(31, 11, 64), (31, 11, 16), (63, 11, 256), (15, 7, 64), (47, 9, 32)])
def test_medium_model_vs_textbook_and_refpipe(k, l, B):
    kmers = uc.medium_model(k + B, 400_000, k)
    log(f"medium k={k} l={l} B={B} n={kmers.shape[0]}")
    got = run(kmers, k, l, n_buckets=B)
    tb = textbook_sorted(kmers, k)
    assert uc.sorted_lines(got) == tb
    if k == 31 and B == 64:
        ref, _ = oracle.refpipe(kmers, k, l, 64, oracle.generate_permutation(l, 5))
        assert uc.sorted_lines(ref) == tb


def test_canonical_sorted_equal(ctx):
    kmers = uc.medium_model(3, 200_000, 31)
    got = ctx.compact(kmers, 31, 11)
    tb, _, _ = oracle.textbook(kmers, 31)
    assert uc.canonical_sorted(got) == uc.canonical_sorted(tb)


# ---------------------------------------------------------------- modes / forms


def test_plain_equals_batched():
    kmers = uc.medium_model(11, 300_000, 31)
    a = run(kmers, 31, 11, mode=uc.UC_MODE_BATCHED)
    b = run(kmers, 31, 11, mode=uc.UC_MODE_PLAIN)
    assert a == b  # same slots, same order: byte-identical


def test_deterministic(ctx):
    kmers = uc.medium_model(12, 200_000, 31)
    a = ctx.compact(kmers, 31, 11)
    b = ctx.compact(kmers, 31, 11)
    assert a == b


def test_result_independent_of_seed_buckets_sampling():
    kmers = uc.medium_model(13, 200_000, 31)
    ref = textbook_sorted(kmers, 31)
    for seed, B, s, load in ((1, 64, 5, 50), (2, 1, 1, 30), (3, 256, 0, 80), (4, 1024, 12, 50)):
        log(f"seed={seed} B={B} sample_log2={s} load={load}")
        got = run(kmers, 31, 11, seed=seed, n_buckets=B, sample_log2=s, load_pct=load)
        assert uc.sorted_lines(got) == ref


@pytest.mark.parametrize("k,l", [(31, 11), (63, 13)])
def test_capped_fine_bucket_count(k, l):
    # one coarse bucket and about 600 k k-mers want more than 1024 fine
    # buckets: the batched form takes the fewest that average at most 1700
    # slots (512 here), the plain form keeps 1024; both against the textbook
    kmers = uc.medium_model(17, 600_000, k)
    ref = textbook_sorted(kmers, k)
    for mode, want in ((uc.UC_MODE_BATCHED, 512), (uc.UC_MODE_PLAIN, 1024)):
        with uc.Compactor(n_buckets=1, mode=mode) as c:
            got = c.compact(kmers, k, l)
            fine = c.bucket_info().n_fine_buckets
        log(f"k={k} mode={mode}: {kmers.shape[0]} k-mers, {fine} fine buckets")
        assert fine == want
        assert uc.sorted_lines(got) == ref


def test_sets_equal_individual():
    sets = [uc.medium_model(100 + i, int(n), 31) for i, n in enumerate([5000, 120_000, 40, 0, 70_000, 31])]
    with uc.Compactor(n_buckets=16) as c:
        texts = c.compact_sets(sets, 31, 11)
        for s, t in zip(sets, texts):
            if s.shape[0] == 0:
                assert t == b""
                continue
            assert uc.sorted_lines(t) == textbook_sorted(s, 31)


def test_phase_by_phase_per_bucket():
    kmers = uc.medium_model(21, 150_000, 31)
    with uc.Compactor(n_buckets=16) as c:
        c.load(kmers, 31)
        c.bucketize(11)
        info = c.bucket_info()
        for b in range(info.n_buckets):
            c.find_edges(b)
        c.coarsen()
        c.emit()
        got = c.text()
    assert uc.sorted_lines(got) == textbook_sorted(kmers, 31)


# ---------------------------------------------------------------- phases vs oracle


def test_order_table_matches_oracle(ctx):
    for l in (1, 3, 6, 8):
        t = ctx.lmer_order_table(l)
        assert np.array_equal(t, oracle.lmer_order_table(l, ctx.opts.seed))


@pytest.mark.parametrize("use_table", [False, True])
@pytest.mark.parametrize("genome", [3_000, 100_000])
def test_bucketize_matches_oracle(use_table, genome):
    # The device buckets by B*S fine buckets (S per set, from its size): fine
    # bucket f is the reference's bucket f with B*S buckets, and coarse bucket
    # b holds the fine buckets f = j*B + b in order of j.  At S = 1 this is
    # the reference's B-bucket layout itself.
    k, l, B = 31, 9, 64
    kmers = uc.medium_model(31, genome, k)
    with uc.Compactor(n_buckets=B, seed=77, keep_src=True) as c:
        c.load(kmers, k)
        if use_table:
            order = oracle.generate_permutation(l, 123)  # the reference's kind of table
            c.bucketize(l, order)
        else:
            order = oracle.lmer_order_table(l, 77)       # the inline bijection
            c.bucketize(l)
        rows, src, flags, ph, pb = c.bucket_arrays()
        info = c.bucket_info()
    S = info.n_fine_buckets // B
    Bf = B * S
    log(f"bucketize n={kmers.shape[0]} S={S}")
    assert (S == 1) == (genome == 3_000)
    ob = oracle.bucketize(kmers, k, l, Bf, order)
    orow = ob["bucket_rows"].astype(np.int64)
    mins = np.array([oracle.double_lmin(kmers[i], k, l, order) for i in range(kmers.shape[0])], dtype=np.int64)
    # fine bucket of every slot: prefix minimizer for home copies, suffix for borrowed ones
    side = (flags == uc.F_BORROW).astype(np.int64)
    fine = mins[src.astype(np.int64), side] % Bf
    for b in range(B):
        fs = [j * B + b for j in range(S)]
        assert rows[b + 1] - rows[b] == sum(orow[f + 1] - orow[f] for f in fs), b
        row_fine = fine[rows[b]:rows[b + 1]]
        assert np.all(row_fine % B == b)
        assert np.all(np.diff(row_fine // B) >= 0), b            # ordered by fine index
        for f in fs:
            g = np.sort(src[rows[b]:rows[b + 1]][row_fine == f])
            o = np.sort(ob["kmer_src"][orow[f]:orow[f + 1]])
            assert np.array_equal(g, o.astype(np.uint32)), (b, f)
    # every k-mer placed once, crossing k-mers twice
    cnt = np.bincount(src.astype(np.int64), minlength=kmers.shape[0])
    assert cnt.max() <= 2 and (cnt >= 1).all()
    # pairs: same k-mer, home copy flagged lent, duplicate flagged borrowed
    assert ph.shape[0] == (flags == uc.F_BORROW).sum() == (flags == uc.F_LENT).sum()
    # the reference lists every crossing k-mer twice: in the (home, suffix)
    # and the (suffix, home) cross list at the same offset
    assert 2 * ph.shape[0] == ob["cross_rows"][-1]
    assert np.array_equal(src[ph], src[pb])
    assert np.all(flags[ph] == uc.F_LENT) and np.all(flags[pb] == uc.F_BORROW)
    assert np.all(mins[src[ph], 0] % Bf != mins[src[ph], 1] % Bf)
    bucket_of = np.repeat(np.arange(B), np.diff(rows).astype(np.int64))
    assert np.array_equal(bucket_of[ph], mins[src[ph], 0] % B)
    assert np.array_equal(bucket_of[pb], mins[src[pb], 1] % B)


@pytest.mark.parametrize("k,l,genome", [(31, 11, 200_000), (63, 13, 120_000), (15, 13, 0)])
def test_staged_and_direct_scatter_agree(k, l, genome):
    # LDS-staged scatters (the default) and direct stores from registers
    # give the same bucket CSR, pairs and text.  k = 15, l = 13 on random
    # k-mers makes about two thirds of them cross, more than a pass-1 tile's
    # stage holds, so those tiles take the direct path inside the staged kernel.
    if genome:
        kmers = uc.medium_model(50 + k, genome, k)
    else:
        kmers = uc.encode(rand_set(random.Random(9), 60_000, k), k)
    out = []
    for direct in (False, True):
        with uc.Compactor(n_buckets=64, seed=5, keep_src=True, direct_scatter=direct) as c:
            c.load(kmers, k)
            c.bucketize(l)
            arrays = c.bucket_arrays()
            log(f"k={k} l={l} direct={direct} m={c.bucket_info().m_placements} x={c.bucket_info().n_cross}")
            text = c.compact(kmers, k, l)
        out.append((arrays, text))
    (a0, t0), (a1, t1) = out
    for x, y in zip(a0, a1):
        assert np.array_equal(x, y)
    assert t0 == t1
    assert uc.sorted_lines(t0) == textbook_sorted(kmers, k)
    if not genome:
        assert a0[3].shape[0] > kmers.shape[0] // 2  # most k-mers crossed


def test_edges_match_textbook_succ():
    k = 31
    kmers = uc.medium_model(41, 120_000, k)
    with uc.Compactor(n_buckets=64, keep_src=True) as c:
        c.load(kmers, k)
        c.bucketize(11)
        c.find_edges()
        rows, src, flags, ph, pb = c.bucket_arrays()
        succ = c.succ()
    ts = oracle.textbook_succ(kmers, k)
    n = kmers.shape[0]
    # the out-edge of k-mer i lives on its duplicate when it crosses buckets,
    # else on its only slot
    out_slot = np.full(n, -1, np.int64)
    for s in range(src.shape[0]):
        if not (flags[s] & uc.F_LENT):
            out_slot[src[s]] = s
    got = np.full(n, np.iinfo(np.uint64).max, np.uint64)
    has = succ[out_slot] != 0xFFFFFFFF
    got[has] = src[succ[out_slot[has]]].astype(np.uint64)
    assert np.array_equal(got, ts)
    # lent copies never carry an edge; borrowed copies are never a target
    assert np.all(succ[flags == uc.F_LENT] == 0xFFFFFFFF)
    tgt = succ[succ != 0xFFFFFFFF]
    assert not np.any(flags[tgt] & uc.F_BORROW)


# ---------------------------------------------------------------- cycles


def circular_kmers(rng, length, k):
    g = "".join(rng.choice("ACGT") for _ in range(length))
    c = g + g[:k - 1]
    return sorted(set(c[i:i + k] for i in range(length)))


@pytest.mark.parametrize("length,k,s", [(50, 15, 5), (2000, 31, 5), (20000, 31, 0), (300, 63, 2)])
def test_isolated_cycles(length, k, s):
    rng = random.Random(length)
    kmers = circular_kmers(rng, length, k) + circular_kmers(rng, length // 2 + 7, k)
    kmers = sorted(set(kmers))
    log(f"cycles length={length} k={k} sample_log2={s}")
    with uc.Compactor(n_buckets=8, sample_log2=s) as c:
        got = c.compact(uc.encode(kmers, k), k, min(11, k - 1))
        info = c.coarsen_info()
    lines = got.decode().splitlines()
    tb, _, nc = oracle.textbook(uc.encode(kmers, k), k)
    assert sorted(lines) == uc.sorted_lines(tb)
    assert info.n_cycles == nc
    check_partition(kmers, lines, k)


def test_long_path_chain_ranking():
    # one long unitig (a random non-repeating sequence, no canonicalisation):
    # many sampled splitters, segment chains ranked by pointer jumping
    rng = random.Random(9)
    k = 31
    g = "".join(rng.choice("ACGT") for _ in range(200_000))
    kmers = sorted(set(g[i:i + k] for i in range(len(g) - k + 1)))
    with uc.Compactor(n_buckets=64, sample_log2=4) as c:
        got = c.compact(uc.encode(kmers, k), k, 11)
        info = c.coarsen_info()
    # chains of thousands of segments: more than the two speculative rounds,
    # so the host-checked round loop ran
    assert info.chain_rounds > 2
    assert uc.sorted_lines(got) == textbook_sorted(uc.encode(kmers, k), k)


@pytest.mark.parametrize("l,B", [(11, 64), (1, 2), (4, 4)])
def test_edge_probe_count_bounds(l, B):
    # every insert and every lookup inspects at least one slot; at the load
    # factors used (at most one half) the average stays far below four
    kmers = uc.medium_model(91, 200_000, 31)
    with uc.Compactor(n_buckets=B, count_probes=True) as c:
        c.compact(kmers, 31, l)
        b = c.bucket_info()
        probes = c.edge_probes()
    with uc.Compactor(n_buckets=B) as c:
        c.compact(kmers, 31, l)
        with pytest.raises(uc.UcError):
            c.edge_probes()  # not counted without the flag
    keys = 2 * (b.m_placements - b.n_cross)
    log(f"l={l} B={B} probes={probes} keys={keys} oversized={b.n_oversized}")
    assert keys <= probes <= 4 * keys


def test_pass_graph_replays_match_eager():
    # the contraction pass is captured once and replayed while sizes and
    # buffers repeat; every replay, with eager work on another context in
    # between, must give the eager context's text byte for byte
    jobs = [(uc.medium_model(31, 120_000, 31), 31, 11), (uc.medium_model(32, 60_000, 63), 63, 13),
            (uc.medium_model(33, 3_000, 31), 31, 11)]
    with uc.Compactor() as g, uc.Compactor(eager=True) as e:
        for km, k, l in jobs:
            want = e.compact(km, k, l)
            g.load(km, k)
            for rep in range(3):
                log(f"n={km.shape[0]} k={k} replay {rep}")
                g.compact_loaded(l)
                assert g.text() == want
                e.compact(jobs[2][0], 31, 11)
            tb, n_unitigs, _ = oracle.textbook(km, k)
            assert uc.sorted_lines(want) == uc.sorted_lines(tb)


@pytest.mark.parametrize("case", ["cycles", "long_chains", "both"])
def test_deferred_counters_match_checked_order(case):
    # small whole jobs emit before the host reads the contraction counters;
    # when the pass did not settle (isolated cycles, chains longer than the
    # speculative rounds) the emission kernels write nothing and the job runs
    # again in the checked order, which the eager form always takes: the two
    # must agree byte for byte, and with the textbook
    rng = random.Random(2024)
    k = 31
    kmers = []
    if case in ("cycles", "both"):
        kmers += circular_kmers(rng, 1500, k) + circular_kmers(rng, 433, k)
    if case in ("long_chains", "both"):
        g = "".join(rng.choice("ACGT") for _ in range(60_000))
        kmers += [g[i:i + k] for i in range(len(g) - k + 1)]
    packed = uc.encode(sorted(set(kmers)), k)
    with uc.Compactor(n_buckets=16, sample_log2=3) as d, uc.Compactor(n_buckets=16, sample_log2=3, eager=True) as e:
        for rep in range(2):  # twice: the second job reuses the first one's buffers
            got = d.compact(packed, k, 11)
            info = d.coarsen_info()
            want = e.compact(packed, k, 11)
            log(f"deferred {case} rep {rep}: cycles {info.n_cycles} chain rounds {info.chain_rounds}")
            assert got == want
    if case != "long_chains":
        assert info.n_cycles > 0
    if case != "cycles":
        assert info.chain_rounds > 2
    assert uc.sorted_lines(got) == textbook_sorted(packed, k)


@pytest.mark.parametrize("l,B,mode", [(1, 2, uc.UC_MODE_BATCHED), (1, 2, uc.UC_MODE_PLAIN), (3, 2, uc.UC_MODE_BATCHED),
                                      (4, 4, uc.UC_MODE_PLAIN), (5, 8, uc.UC_MODE_BATCHED)])
def test_oversized_fine_buckets(l, B, mode):
    # few distinct minimizers -> fine buckets above the LDS group size: those
    # of at most 4096 slots run alone in a larger LDS area, those of at most
    # 65535 in LDS hash slices, bigger ones (l = 1: four minimizers) through
    # the HBM-table kernels; the rest through LDS groups
    kmers = uc.medium_model(77 + l, 300_000, 31)
    with uc.Compactor(n_buckets=B, mode=mode) as c:
        got = c.compact(kmers, 31, l)
        info = c.bucket_info()
    log(f"l={l} B={B} largest fine bucket {info.largest_fine_bucket} oversized {info.n_oversized}")
    assert info.largest_fine_bucket > 2048 and info.n_oversized > 0
    if l == 1:
        assert info.largest_fine_bucket > 65535
    assert uc.sorted_lines(got) == textbook_sorted(kmers, 31)


@pytest.mark.parametrize("k,l,mode", [(31, 11, uc.UC_MODE_BATCHED), (63, 13, uc.UC_MODE_PLAIN), (31, 3, uc.UC_MODE_BATCHED),
                                      (63, 2, uc.UC_MODE_BATCHED)])
def test_small_groups_route_fine_buckets_to_every_path(k, l, mode):
    # with 256-slot LDS groups most fine buckets are "oversized": those up to
    # 4096 slots take the single-bucket LDS kernel, larger ones (l <= 3) the
    # LDS hash slices; the text is the textbook one either way
    kmers = uc.medium_model(91 + k + l, 200_000, k)
    with uc.Compactor(n_buckets=16, mode=mode, group_slots=256) as c:
        got = c.compact(kmers, k, l)
        info = c.bucket_info()
    log(f"k={k} l={l} largest fine bucket {info.largest_fine_bucket} oversized {info.n_oversized} "
        f"slots {info.oversized_slots}")
    assert info.n_oversized > 0
    if l <= 3:
        assert info.largest_fine_bucket > 4096
    assert uc.sorted_lines(got) == textbook_sorted(kmers, k)


@pytest.mark.parametrize("low_memory", [False, True])
def test_slice_overflow_takes_hbm_tables(low_memory):
    # one probe per slice-table insert: the first collision marks its slice
    # as overflowed, and the edge phase runs again with every sliced fine
    # bucket on an HBM table.  The whole-job form reads the flag after
    # emission (deferred counters), the low-memory form after contraction.
    kmers = uc.medium_model(93, 200_000, 31)
    with uc.Compactor(n_buckets=16, group_slots=256, low_memory=low_memory) as c:
        c.set_slice_probes(1)
        got = c.compact(kmers, 31, 3)
        b = c.bucket_info()
    log(f"slice overflow low_memory={low_memory}: largest fine {b.largest_fine_bucket} slices {b.n_slice_tasks} "
        f"hbm {b.n_hbm_tables}")
    assert b.largest_fine_bucket > 4096
    assert b.n_slice_tasks == 0 and b.n_hbm_tables > 0
    assert uc.sorted_lines(got) == textbook_sorted(kmers, 31)
    with uc.Compactor(n_buckets=16, group_slots=256) as c:
        c.compact(kmers, 31, 3)
        assert c.bucket_info().n_slice_tasks > 0  # the same job with the default bound keeps its slices


# ---------------------------------------------------------------- the configured workloads, reduced


def shuffled(kmers, seed):
    """The same k-mer set in a seeded random order (the generator and the
    golden files are sorted; slot order, LDS groups, runs and the stable
    scatters all follow input order)."""
    return np.ascontiguousarray(kmers[np.random.default_rng(seed).permutation(kmers.shape[0])])


@pytest.fixture(scope="module")
def c1_case():
    # 1 Mbp uniform ACGT (seed 1), every distinct canonical 31-mer, l = 11
    kmers = uc.synth_kmers(1, 1_000_000, 31)
    tb = textbook_sorted(kmers, 31)
    ref, _ = oracle.refpipe(kmers, 31, 11, 16, oracle.generate_permutation(11, 1))
    assert uc.sorted_lines(ref) == tb
    return kmers, tb


@pytest.mark.parametrize("mode", [uc.UC_MODE_BATCHED, uc.UC_MODE_PLAIN])
def test_config_uniform_1mbp_text_16_buckets(c1_case, mode, tmp_path):
    # as k-mer text (count line, one k-mer per line) through the file reader
    kmers, tb = c1_case
    path = tmp_path / "c1.txt"
    path.write_bytes(uc.format_text(kmers, 31))
    log(f"uniform 1 Mbp: {kmers.shape[0]} k-mers, mode {mode}")
    with uc.Compactor(n_buckets=16, mode=mode) as c:
        c.load_file(str(path), 31)
        c.compact_loaded(11)
        got = c.text()
    assert uc.sorted_lines(got) == tb


def test_config_uniform_1mbp_shuffled(c1_case):
    kmers, tb = c1_case
    with uc.Compactor(n_buckets=16) as c:
        got = c.compact(shuffled(kmers, 1), 31, 11)
    assert uc.sorted_lines(got) == tb


def uneven_sets(n_sets, seed0, median=2000, lo=100, hi=50_000):
    # sizes ~ LogNormal(ln median, 1) clipped to [lo, hi] k-mers, each from the
    # medium repeat/error model; every other set shuffled
    rng = np.random.default_rng(seed0)
    sizes = np.clip(np.exp(rng.normal(np.log(median), 1.0, n_sets)), lo, hi).astype(np.int64)
    sets = []
    for i, sz in enumerate(sizes):
        s = uc.medium_model(seed0 + i, int(sz / 0.97) + 31, 31)
        sets.append(shuffled(s, i) if i % 2 else s)
    return sets


def test_config_256_uneven_sets_batched():
    # 256 sets in one batched run: every set's text equals the textbook's
    # lines and, byte for byte, a plain per-bucket run of that set alone
    sets = uneven_sets(256, 1000)
    log(f"256 sets, {sum(s.shape[0] for s in sets)} k-mers, largest {max(s.shape[0] for s in sets)}")
    with uc.Compactor(n_buckets=64) as c:
        texts = c.compact_sets(sets, 31, 11)
    assert len(texts) == 256
    with uc.Compactor(n_buckets=64, mode=uc.UC_MODE_PLAIN) as p:
        for i, (s, t) in enumerate(zip(sets, texts)):
            assert t == p.compact(s, 31, 11), i
            assert uc.sorted_lines(t) == textbook_sorted(s, 31), i


@pytest.fixture(scope="module")
def c4_case():
    # the large instance's repeat-family model at 5 Mbp (45% repeats, copy
    # number 2-1000, divergence 0.5-15%, 0.5% errors), canonical 63-mers
    kmers = uc.large_model(4, 5_000_000, 63)
    return kmers, textbook_sorted(kmers, 63)


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
def test_config_repeat_families_k63_256_buckets(c4_case, order):
    # l = 13, 256 buckets: repeat families make single minimizers frequent,
    # so fine buckets above the LDS group size go alone into the larger LDS
    # area and, above 4096 slots, into LDS hash slices
    kmers, tb = c4_case
    km = kmers if order == "sorted" else shuffled(kmers, 4)
    with uc.Compactor(n_buckets=256) as c:
        got = c.compact(km, 63, 13)
        b = c.bucket_info()
    log(f"repeat families {order}: n={b.n_kmers} largest fine {b.largest_fine_bucket} big {b.n_big_lds} "
        f"slices {b.n_slice_tasks} hbm {b.n_hbm_tables}")
    assert b.n_big_lds > 0 and b.n_slice_tasks > 0
    assert uc.sorted_lines(got) == tb
    if order == "sorted":
        ref, _ = oracle.refpipe(kmers, 63, 13, 256, oracle.generate_permutation(13, 4))
        assert uc.sorted_lines(ref) == tb


def test_config_repeat_families_hbm_tables(c4_case):
    # l = 5 leaves a few hundred distinct minimizers: fine buckets of more
    # than 65535 slots take the HBM hash tables
    kmers, tb = c4_case
    with uc.Compactor(n_buckets=256) as c:
        got = c.compact(shuffled(kmers, 9), 63, 5)
        b = c.bucket_info()
    log(f"repeat families l=5: largest fine {b.largest_fine_bucket} hbm tables {b.n_hbm_tables}")
    assert b.n_hbm_tables > 0 and b.largest_fine_bucket > 65535
    assert uc.sorted_lines(got) == tb


@pytest.mark.parametrize("k", [15, 31, 33, 47, 63])
def test_genome_shuffled_input(k):
    rng = np.random.default_rng(k + 500)
    kmers = uc.synth_kmers(int(rng.integers(1 << 30)), 60_000, k, repeat_frac=0.05, seg_min=200, seg_max=800,
                           repeat_sub=0.01, error_frac=0.01)
    km = shuffled(kmers, k)
    log(f"shuffled genome k={k} n={km.shape[0]}")
    got = run(km, k, min(11, k - 1))
    assert uc.sorted_lines(got) == textbook_sorted(kmers, k)
    # the same multiset of lines as the sorted input's run
    assert uc.sorted_lines(got) == uc.sorted_lines(run(kmers, k, min(11, k - 1)))


@pytest.mark.parametrize("k,l,B", [(31, 11, 64), (63, 13, 256)])
def test_medium_model_shuffled_vs_textbook_and_refpipe(k, l, B):
    kmers = shuffled(uc.medium_model(k + B + 1, 400_000, k), 7)
    log(f"medium shuffled k={k} l={l} B={B} n={kmers.shape[0]}")
    got = run(kmers, k, l, n_buckets=B)
    tb = textbook_sorted(kmers, k)
    assert uc.sorted_lines(got) == tb
    ref, _ = oracle.refpipe(kmers, k, l, 64, oracle.generate_permutation(l, 5))
    assert uc.sorted_lines(ref) == tb
    # plain form byte-identical to batched on the same (unsorted) slots
    assert run(kmers, k, l, n_buckets=B, mode=uc.UC_MODE_PLAIN) == got


@pytest.mark.parametrize("k,l,B,model", [(31, 11, 64, "medium"), (63, 13, 256, "large")])
def test_low_memory_mode_same_bytes_smaller_peak(k, l, B, model):
    # UC_FLAG_LOW_MEMORY frees every buffer after its last use: the text is
    # byte for byte the normal mode's, the peak is lower, and the loaded
    # input is consumed by the job
    kmers = uc.medium_model(31, 600_000, k) if model == "medium" else uc.large_model(8, 1_500_000, k)
    with uc.Compactor(n_buckets=B) as c:
        c.load(kmers, k)
        c.compact_loaded(l)
        want, peak_full = c.text(), c.timing().device_peak_bytes
    with uc.Compactor(n_buckets=B, low_memory=True) as c:
        for _ in range(2):  # twice: the freed buffers come back for the next job
            c.load(kmers, k)
            c.compact_loaded(l)
            got, peak_low = c.text(), c.timing().device_peak_bytes
            assert got == want
            log(f"low memory {model}: peak {peak_low / 2**20:.1f} MiB against {peak_full / 2**20:.1f} MiB")
            assert 0 < peak_low < 0.8 * peak_full
            info = c.coarsen_info()
            assert info.text_bytes == len(want)
            # the input is consumed, and the head lists (which hold the
            # text offsets of multi-segment unitigs now) are spent
            with pytest.raises(uc.UcError) as e:
                c.compact_loaded(l)
            assert e.value.status == uc.UC_E_STATE
            with pytest.raises(uc.UcError) as e:
                c.emit()
            assert e.value.status == uc.UC_E_STATE
            assert c.text() == want
        # a whole job from host memory works as usual, and so does one from
        # the k-mer text parsed on the device
        assert c.compact(kmers, k, l) == want
        c.load_text(uc.format_text(kmers, k), k)
        c.compact_loaded(l)
        assert c.text() == want
        held = c.held_bytes()
        c.load(kmers, k)  # the pool stays: loading and running again allocates nothing new
        c.compact_loaded(l)
        assert c.text() == want and c.held_bytes() == held
    assert uc.sorted_lines(want) == textbook_sorted(kmers, k)


def test_low_memory_mode_with_isolated_cycles():
    # cycles make contraction run twice with the cycle cut between; a
    # low-memory job gives the pass's arrays back around the cut
    rng = random.Random(77)
    cyc = circular_kmers(rng, 3000, 31) + circular_kmers(rng, 1207, 31)
    base = uc.decode(uc.medium_model(77, 200_000, 31), 31)
    kmers = uc.encode(sorted(set(cyc) | set(base)), 31)
    want = run(kmers, 31, 11, n_buckets=16)
    with uc.Compactor(n_buckets=16, low_memory=True) as c:
        got = c.compact(kmers, 31, 11)
        info = c.coarsen_info()
    assert info.n_cycles >= 2
    assert got == want
    assert uc.sorted_lines(got) == textbook_sorted(kmers, 31)
