"""The k-mer text parsed on the device (uc_load_text / uc_load_file) equals
the host parser and the oracle's parser word for word, with the same
errors; the device-parsed job compacts to the same text; the pinned-chunk
writer appends the same bytes."""
import json
import os

import numpy as np
import pytest

import oracle
import unitigcoarsening_amd as uc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def log(msg):
    print(msg, flush=True)


def device_parse(text, k):
    with uc.Compactor() as c:
        c.load_text(text, k)
        return c.loaded(k)


def expect(text, k):
    a = uc.parse_text(text, k)
    assert np.array_equal(a, oracle.parse_text(text, k))
    return a


CASES = [
    (b"3\nACGTN 17\nacgtT\nTTTT\nGGGG\n", 4),      # trailing fields, lowercase and N read as A, lines past the count
    (b"2\r\nACGT\r\nCCCC\r\n", 4),                  # CRLF
    (b"9\nACGT\nTTTT", 4),                          # no final newline, count above the lines
    (b"0\n", 4),                                    # empty body
    (b"5", 4),                                      # count line only, no newline
    (b"  2\t\nACGT\nCCCC\n", 4),                    # blanks around the count
    (b"1\nACGT\nCCCC\n\n\r\n\n", 4),               # newlines and CRs at the end end the text
    (b"2\nACGT\r\nCCCCG\r", 4),                     # '\r' at the end without a newline
    (b"2\nAC\rT\nCCCC\n", 4),                       # '\r' inside a line reads as A
    (b"2\nAC\n", 2),                                # k = 2
    (b"0\n" + b"\n" * 5000, 4),                     # only newlines after the count, over a 4 KiB step
]


@pytest.mark.parametrize("text,k", CASES)
def test_edge_cases(text, k):
    assert np.array_equal(device_parse(text, k), expect(text, k))


@pytest.mark.parametrize("text,k", [
    (b"2\nACGT\nAC\n", 4),          # short line
    (b"2\nACGT\nACG\r\n", 4),       # k - 1 characters and a '\r'
    (b"2\nACGT\n\nCCCC\n", 4),      # empty line before another line
    (b"1\nACGT\nAC\n", 4),          # a short line past the count is read too
    (b"3\nACGT\nTTTT\nGG", 4),      # short last line without newline
    (b"ACGT\n", 4),                 # no count line
])
def test_format_errors(text, k):
    with pytest.raises(uc.UcError) as e:
        uc.parse_text(text, k)
    host = e.value.status
    with pytest.raises(uc.UcError) as e:
        device_parse(text, k)
    assert e.value.status == host == uc.UC_E_FORMAT


@pytest.mark.parametrize("seed", range(4))
def test_device_parse_matches_reference_loop(seed):
    # the reference's own chunked read loop (driver.cpp:257-354, restated in
    # the oracle): wherever it ends and does not stop at a chunk end on a
    # count met by chance, its k-mers are the device parser's; where it
    # stops early, ours start with its k-mers
    import random

    rng = random.Random(100 + seed)
    k = rng.choice([4, 15, 31, 63])
    n = rng.randint(200, 3000)
    for count in (n, rng.randint(0, n - 1), n + 7):
        for trailing in (True, False):
            lines = ["".join(rng.choice("ACGT") for _ in range(k)) + ("\t%d" % rng.randint(1, 99) if rng.random() < 0.3
                                                                        else "") for _ in range(n)]
            text = (f"{count}\n" + "\n".join(lines) + ("\n" if trailing else "")).encode()
            got = device_parse(text, k)
            assert np.array_equal(got, expect(text, k))
            assert got.shape[0] == n
            try:
                ref = oracle.ref_load_kmers(text, k)
            except oracle.OracleError as e:
                assert e.rc == -6 and count != n  # the reference would not end
                continue
            log(f"k={k} lines={n} count={count} trailing={trailing}: reference read {ref.shape[0]}")
            assert np.array_equal(ref, got[:ref.shape[0]])
            if count == n:
                assert ref.shape[0] == n


@pytest.mark.parametrize("k", [2, 4, 15, 16, 17, 31, 32, 33, 47, 63, 64])
def test_all_k_across_tiles(k):
    # lines of uneven length so that lines straddle the 4 KiB parse tiles
    rng = np.random.default_rng(k)
    kmers = uc.synth_kmers(k + 3, 20_000, k) if k > 8 else uc.encode(
        sorted({"".join(rng.choice(list("ACGT"), k)) for _ in range(200)}), k)
    lines = [s + " " * int(rng.integers(0, 40)) + ("\t%d" % rng.integers(1, 99) if rng.integers(0, 2) else "")
             for s in uc.decode(kmers, k)]
    text = (f"{len(lines)}\n" + "\n".join(lines) + "\n").encode()
    got = device_parse(text, k)
    assert np.array_equal(got, expect(text, k))
    assert np.array_equal(got, kmers)


def test_large_text_and_compaction_match_host_path():
    kmers = uc.medium_model(12, 2_000_000, 31)
    text = uc.format_text(kmers, 31)
    log(f"text {len(text)} bytes, {kmers.shape[0]} k-mers")
    with uc.Compactor() as c:
        c.load_text(text, 31)
        assert np.array_equal(c.loaded(31), kmers)
        c.compact_loaded(11)
        dev = c.text()
        host = c.compact(uc.parse_text(text, 31), 31, 11)
    assert dev == host


def test_device_buffer_input():
    import ctypes as C

    # plain HIP runtime calls for the device buffer
    hip = C.CDLL("libamdhip64.so")
    kmers = uc.synth_kmers(5, 30_000, 47)
    text = uc.format_text(kmers, 47)
    ptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(ptr), C.c_size_t(len(text) + 3)) == 0
    try:
        # an odd offset: the library copies the text into its own padded buffer
        dst = C.c_void_p(ptr.value + 3)
        assert hip.hipMemcpy(dst, C.c_char_p(text), C.c_size_t(len(text)), 1) == 0  # host to device
        with uc.Compactor() as c:
            rc = uc.lib().uc_load_text(c._ctx, C.cast(dst, C.c_char_p), len(text), 47, 1)
            assert rc == uc.UC_OK
            assert np.array_equal(c.loaded(47), kmers)
    finally:
        hip.hipFree(ptr)


@pytest.mark.parametrize("name", ["hand_k4", "genome_k31", "genome_k63", "repeats_errors_k31"])
def test_golden_files_through_the_file_reader(name, tmp_path):
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        m = {e["name"]: e for e in json.load(f)}[name]
    path = os.path.join(GOLDEN, name + ".kmers.txt")
    with open(path, "rb") as f:
        text = f.read()
    k, l = m["k"], m["l"]
    out = tmp_path / "o.txt"
    out.write_bytes(b"existing line\n")
    with uc.Compactor(n_buckets=16) as c:
        c.load_file(path, k)
        assert np.array_equal(c.loaded(k), uc.parse_text(text, k))
        c.compact_loaded(l)
        want = c.text()
        _check_append(c, out)
    got = out.read_bytes()
    assert got == b"existing line\n" + want
    with open(os.path.join(GOLDEN, name + ".unitigs.txt"), "rb") as f:
        assert uc.sorted_lines(want) == uc.sorted_lines(f.read())


def _check_append(c, out):
    rc = uc.lib().uc_text_append_file(c._ctx, os.fsencode(str(out)))
    assert rc == uc.UC_OK


def test_file_reader_errors(tmp_path):
    with uc.Compactor() as c:
        rc = uc.lib().uc_load_file(c._ctx, os.fsencode(str(tmp_path / "missing.txt")), 31)
        assert rc == uc.UC_E_IO
        p = tmp_path / "bad.txt"
        p.write_bytes(b"2\nACGT\n")
        assert uc.lib().uc_load_file(c._ctx, os.fsencode(str(p)), 31) == uc.UC_E_FORMAT


def test_file_reader_many_pieces(tmp_path):
    # files of several reader pieces (16 MiB each, read by four threads, two
    # pieces each for some): every byte lands in its place, from the start of
    # a text file and from after a packed file's header, and a packed file cut
    # short in its last piece is a format error
    kmers = uc.medium_model(33, 2_600_000, 31)
    text = uc.format_text(kmers, 31)
    log(f"text {len(text)} bytes")
    assert len(text) > 4 * (16 << 20) and len(text) % (16 << 20)
    p = tmp_path / "big.txt"
    p.write_bytes(text)
    with uc.Compactor() as c:
        c.load_file(str(p), 31)
        assert np.array_equal(c.loaded(31), kmers)
    k63 = uc.medium_model(34, 1_500_000, 63)
    q = tmp_path / "big.ucb"
    uc.write_packed_file(str(q), k63, 63)
    assert q.stat().st_size > (16 << 20)
    with uc.Compactor() as c:
        assert c.load_packed_file(str(q)) == 63
        assert np.array_equal(c.loaded(63), k63)
    cut = tmp_path / "cut.ucb"
    cut.write_bytes(q.read_bytes()[:-4])
    with uc.Compactor() as c:
        with pytest.raises(uc.UcError) as e:
            c.load_packed_file(str(cut))
        assert e.value.status == uc.UC_E_FORMAT


def test_writer_multi_chunk(tmp_path):
    # a text above one 64 MiB pinned chunk: copy/write overlap across chunks
    kmers = uc.synth_kmers(8, 16_000_000, 63, error_frac=0.1)
    with uc.Compactor() as c:
        c.compact(kmers, 63, 13)
        want = c.text()
        log(f"text {len(want)} bytes")
        assert len(want) > (64 << 20)
        out = tmp_path / "big.txt"
        _check_append(c, out)
    assert out.read_bytes() == want


def test_text_copy_in_pinned_pieces_matches_writer(tmp_path):
    # uc_text_copy moves texts of 1 MiB and more through two pinned 4 MiB
    # pieces; a text of a few pieces with a partial last one must equal
    # what the file writer (64 MiB chunks) appends, and the oracle's lines
    kmers = uc.medium_model(21, 700_000, 31)
    with uc.Compactor() as c:
        got = c.compact(kmers, 31, 11)
        log(f"text {len(got)} bytes")
        assert len(got) > (8 << 20) and len(got) % (4 << 20)
        out = tmp_path / "o.txt"
        _check_append(c, out)
    assert out.read_bytes() == got
    want, _, _ = oracle.textbook(kmers, 31)
    assert uc.sorted_lines(got) == uc.sorted_lines(want)


@pytest.mark.parametrize("k", [15, 31, 47, 63])
def test_packed_file_load_and_cli(k, tmp_path):
    import subprocess

    kmers = uc.medium_model(40 + k, 60_000, k)
    p = tmp_path / "in.ucb"
    uc.write_packed_file(str(p), kmers, k)
    l = 7 if k == 15 else 11
    with uc.Compactor(n_buckets=64) as c:
        assert c.load_packed_file(str(p)) == k
        assert np.array_equal(c.loaded(k), kmers)
        c.compact_loaded(l)
        want = c.text()
        assert want == c.compact(kmers, k, l)
    out = tmp_path / "o.txt"
    r = subprocess.run([uc.CLI_PATH, str(p), str(k), str(l), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert uc.sorted_lines(out.read_bytes()) == uc.sorted_lines(want)
    # a truncated file is a format error
    raw = p.read_bytes()
    q = tmp_path / "cut.ucb"
    q.write_bytes(raw[:-4])
    with uc.Compactor() as c:
        with pytest.raises(uc.UcError) as e:
            c.load_packed_file(str(q))
        assert e.value.status == uc.UC_E_FORMAT
