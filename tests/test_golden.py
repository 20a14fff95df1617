"""Golden fixtures (tests/golden, written by make_golden.py).

CPU: the oracle, the restated reference pipeline and the library's host
parser reproduce every fixture.  GPU: the uc_main drop-in, file to file,
reproduces every fixture (its output appended to an existing file, like
the reference's std::ofstream::app)."""
import hashlib
import json
import os
import subprocess

import pytest

import oracle
import unitigcoarsening_amd as uc

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(HERE, "manifest.json")))
IDS = [m["name"] for m in MANIFEST]


def load(m):
    inp = open(os.path.join(HERE, m["name"] + ".kmers.txt"), "rb").read()
    exp = open(os.path.join(HERE, m["name"] + ".unitigs.txt"), "rb").read()
    return inp, exp


@pytest.mark.parametrize("m", MANIFEST, ids=IDS)
def test_fixture_integrity(m):
    _, exp = load(m)
    assert hashlib.sha256(exp).hexdigest() == m["sha256_unitigs"]
    assert len(exp.splitlines()) == m["n_unitigs"]


@pytest.mark.parametrize("m", MANIFEST, ids=IDS)
def test_oracle_reproduces_fixture(m):
    inp, exp = load(m)
    k = m["k"]
    packed = oracle.parse_text(inp, k)
    assert packed.shape[0] == m["n_kmers"]
    text, nu, nc = oracle.textbook(packed, k)
    assert uc.sorted_lines(text) == exp.decode().splitlines()
    assert nc == m["n_cycles"]
    if k % 16 == 15 and nc == 0:
        ref, _ = oracle.refpipe(packed, k, m["l"], 64, oracle.generate_permutation(m["l"], 9))
        assert uc.sorted_lines(ref) == exp.decode().splitlines()


@pytest.mark.parametrize("m", MANIFEST, ids=IDS)
def test_host_parser_on_fixture(m):
    inp, _ = load(m)
    assert (uc.parse_text(inp, m["k"]) == oracle.parse_text(inp, m["k"])).all()


@pytest.mark.gpu
@pytest.mark.parametrize("m", MANIFEST, ids=IDS)
def test_cli_file_to_file(m, tmp_path):
    _, exp = load(m)
    out = tmp_path / "unitigs.txt"
    out.write_text("PREVIOUS\n")  # the output file is appended to, never truncated
    r = subprocess.run([uc.CLI_PATH, os.path.join(HERE, m["name"] + ".kmers.txt"), str(m["k"]), str(m["l"]),
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = out.read_text().splitlines()
    assert lines[0] == "PREVIOUS"
    assert sorted(lines[1:]) == exp.decode().splitlines()
