"""Autotuner (uc_autotune): it picks settings from its candidate lists, keeps
them in the context, and the unitig set stays the textbook's."""
import pytest

import oracle
import unitigcoarsening_amd as uc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,l", [(31, 11), (63, 13)])
def test_autotune_keeps_results(k, l):
    kmers = uc.medium_model(3, 300_000, k)
    want = uc.sorted_lines(oracle.textbook(kmers, k)[0])
    with uc.Compactor() as c:
        c.load(kmers, k)
        r = c.autotune(l, reps=1)
        print(r, flush=True)
        assert r["trials"] >= 10
        assert r["n_buckets"] in (32, 64, 128, 256)
        assert r["fine_slots"] in (0, 256, 384, 512, 768, 1024)  # 0: the library's default choice
        assert r["group_slots"] in (1024, 2048)
        assert 4 <= r["sample_log2"] <= 12
        assert r["best_ms"] <= r["default_ms"]
        assert uc.sorted_lines(c.compact(kmers, k, l)) == want
        c.load(kmers, k)
        c.compact_loaded(l)
        assert uc.sorted_lines(c.text()) == want


def test_autotune_before_load_is_an_error():
    with uc.Compactor() as c:
        with pytest.raises(uc.UcError) as e:
            c.autotune(11)
        assert e.value.status == uc.UC_E_STATE


@pytest.mark.parametrize("fine,group", [(64, 256), (4096, 2048), (256, 512)])
def test_fine_and_group_settings(fine, group):
    kmers = uc.medium_model(4, 200_000, 31)
    with uc.Compactor(fine_slots=fine, group_slots=group) as c:
        got = c.compact(kmers, 31, 11)
        info = c.bucket_info()
    print(fine, group, info.n_fine_buckets, info.largest_fine_bucket, flush=True)
    assert uc.sorted_lines(got) == uc.sorted_lines(oracle.textbook(kmers, 31)[0])


def test_autotune_keeps_bucket_count_when_asked():
    kmers = uc.medium_model(5, 100_000, 31)
    with uc.Compactor(n_buckets=16) as c:
        c.load(kmers, 31)
        r = c.autotune(11, reps=1, keep_buckets=True)
        assert r["n_buckets"] == 16
        assert uc.sorted_lines(c.compact(kmers, 31, 11)) == uc.sorted_lines(oracle.textbook(kmers, 31)[0])


def test_kernel_stats_select_times_one_kernel():
    # every launch timed, then only the selected kernel; the text is unchanged
    kmers = uc.medium_model(4, 200_000, 31)
    with uc.Compactor(profile=True) as c:
        c.load(kmers, 31)
        c.compact_loaded(11)
        want = c.text()
        full = c.kernel_stats()
        assert {"k_lmin_count", "k_edge_lds", "k_head_emit"} <= set(full)
        c.kernel_stats_select("k_head_emit")
        c.kernel_stats_reset()
        for _ in range(2):
            c.compact_loaded(11)
        one = c.kernel_stats()
        assert one["k_head_emit"][1] == 2 and one["k_head_emit"][0] > 0
        assert all(nl == 0 for name, (ms, nl) in one.items() if name != "k_head_emit")
        assert c.text() == want
        c.kernel_stats_select(None)
        c.kernel_stats_reset()
        c.compact_loaded(11)
        assert c.kernel_stats()["k_lmin_count"][1] == 1
