"""Host-side checks that need no GPU: the C-ABI library loads and exports
every declared symbol, the host text parser matches the oracle's parser,
and the CLI keeps the reference's usage and I/O failure behaviour."""
import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle
import unitigcoarsening_amd as uc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(uc_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared("uc_api.h")
    assert len(names) >= 25
    L = C.CDLL(uc.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python mirror binds exactly the declared set
    assert sorted(uc.API) == names


def test_synth_library_exports():
    L = C.CDLL(uc.SYNTH_PATH)
    for n in declared("uc_synth.h"):
        assert hasattr(L, n), n


def test_strerror_codes():
    L = uc.lib()
    for code in (0, -1, -2, -3, -4, -5, -6, -7):
        assert L.uc_strerror(code)
    assert L.uc_strerror(-99) == b"unknown status"


@pytest.mark.parametrize("k", [4, 15, 16, 17, 31, 33, 47, 63, 64])
def test_parser_matches_oracle(k):
    kmers = uc.synth_kmers(k, 5000, k)
    text = uc.format_text(kmers, k)
    a = uc.parse_text(text, k)
    b = oracle.parse_text(text, k)
    assert np.array_equal(a, kmers)
    assert np.array_equal(a, b)


def test_parser_edge_cases():
    # trailing fields are skipped, lowercase and N read as A; every line is
    # read whatever the count says (the reference reads on past it)
    t = b"3\nACGTN 17\nacgtT\nTTTT\nGGGG\n"
    a = uc.parse_text(t, 4)
    assert uc.decode(a, 4) == ["ACGT", "AAAA", "TTTT", "GGGG"]
    assert np.array_equal(a, oracle.parse_text(t, 4))
    # CRLF line ends
    t = b"2\r\nACGT\r\nCCCC\r\n"
    assert uc.decode(uc.parse_text(t, 4), 4) == ["ACGT", "CCCC"]
    # short line -> format error
    with pytest.raises(uc.UcError) as e:
        uc.parse_text(b"2\nACGT\nAC\n", 4)
    assert e.value.status == -6
    # missing count line
    with pytest.raises(uc.UcError):
        uc.parse_text(b"ACGT\n", 4)
    # no newline at end, more count than lines
    assert uc.decode(uc.parse_text(b"9\nACGT\nTTTT", 4), 4) == ["ACGT", "TTTT"]
    # empty body
    assert uc.parse_text(b"0\n", 4).shape == (0, 1)
    # newlines and carriage returns at the end end the text; an empty line
    # before another line is a short line
    assert uc.decode(uc.parse_text(b"1\nACGT\n\n\r\n", 4), 4) == ["ACGT"]
    with pytest.raises(uc.UcError):
        uc.parse_text(b"2\nACGT\n\nCCCC\n", 4)


def test_placement_limit_check():
    # every k-mer has one home slot and crossing k-mers one more: n + x
    # placements must stay below 2^32 - 1 (32-bit slot ids)
    L = uc.lib()
    assert L.uc_check_placements(10, 3) == uc.UC_OK
    assert L.uc_check_placements(2**32 - 2, 0) == uc.UC_OK
    assert L.uc_check_placements(2**31, 2**31 - 2) == uc.UC_OK
    assert L.uc_check_placements(2**31, 2**31 - 1) == uc.UC_E_LIMIT
    # a job of 3.9e9 k-mers with 10% crossing: the sum wraps a 32-bit count
    assert L.uc_check_placements(3_900_000_000, 400_000_000) == uc.UC_E_LIMIT
    assert L.uc_check_placements(5, 6) == uc.UC_E_STATE


def test_job_estimate_bounds_every_crossing_rate():
    # the estimate assumes up to two placements per k-mer; at any size it
    # must be at least that many slots' worth of the per-slot buffers
    for n, k in ((0, 31), (1000, 15), (50_000_000, 31), (50_000_000, 63)):
        est = uc.job_bytes_estimate(n, k)
        W = 1 if k <= 32 else 2
        assert est >= 2 * n * (16 * W + 100) + n * (8 * W + k)


def test_parser_large_multithreaded():
    kmers = uc.synth_kmers(1, 300_000, 31)
    text = uc.format_text(kmers, 31)
    a = uc.parse_text(text, 31)
    assert np.array_equal(a, kmers)


def test_cli_usage():
    r = subprocess.run([uc.CLI_PATH, "a", "b"], capture_output=True, text=True)
    assert r.returncode == 1
    assert "You input 3 args" in r.stdout
    assert "Usage:" in r.stderr and "<k-mer file> k l <output file>" in r.stderr


def test_cli_missing_input(tmp_path):
    r = subprocess.run([uc.CLI_PATH, str(tmp_path / "nope.txt"), "31", "11", str(tmp_path / "o.txt")],
                       capture_output=True, text=True)
    assert r.returncode == 1
    assert "Error: Could not open input file. Exiting ..." in r.stdout


def test_packed_file_roundtrip(tmp_path):
    for k in (4, 17, 31, 33, 63, 64):
        kmers = uc.synth_kmers(k, 3000, k)
        p = tmp_path / f"k{k}.ucb"
        uc.write_packed_file(str(p), kmers, k)
        h = uc.read_packed_header(str(p))
        assert h == {"k": k, "words": uc.words_per_kmer(k), "n": kmers.shape[0]}
        raw = p.read_bytes()
        assert raw[:8] == b"UCKMERS1" and len(raw) == 32 + kmers.size * 4
        assert np.array_equal(np.frombuffer(raw[32:], np.uint32).reshape(kmers.shape), kmers)
    # text is not a packed file; a missing file is an I/O error
    t = tmp_path / "t.txt"
    t.write_bytes(b"1\nACGT\n")
    with pytest.raises(uc.UcError) as e:
        uc.read_packed_header(str(t))
    assert e.value.status == uc.UC_E_FORMAT
    with pytest.raises(uc.UcError) as e:
        uc.read_packed_header(str(tmp_path / "none.ucb"))
    assert e.value.status == uc.UC_E_IO


def test_cpu_threads_follow_the_host_share(monkeypatch):
    # the CPU baseline runs on every core of the GPU's share (OMP_NUM_THREADS
    # when the host sets it, else every core)
    import bench

    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    assert bench.cpu_threads() == min(3, os.cpu_count())
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.cpu_threads() == os.cpu_count()


def test_bench_refuses_more_than_one_gpu():
    # the bench line is one GPU's; a request for more exits with a message
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
