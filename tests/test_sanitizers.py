"""The host parts under the sanitizers, on the CPU.

- the request loop (csrc/uc_server.cpp) with many client threads, over
  host-only stand-in contexts (tests/sanitize/stub_ctx.cpp), under
  ThreadSanitizer and under AddressSanitizer + UndefinedBehaviorSanitizer;
- the host k-mer text parser (csrc/uc_text.cpp) against the oracle's parser
  on random texts, under ASan + UBSan;
- the oracle itself (textbook and the reference-pipeline restatement) on
  random genomes, under ASan + UBSan.

Each binary exits non-zero on a mismatch; a sanitizer report aborts it.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_DIR = os.path.join(ROOT, "tests", "sanitize")
CSRC = os.path.join(ROOT, "unitigcoarsening_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")
INC = [f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}", f"-I{ORACLE}"]
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
TSAN = ["-fsanitize=thread"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs gcc/g++")


def build(tmp, name, cmd):
    out = os.path.join(tmp, name)
    r = subprocess.run(cmd + ["-o", out], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def run(exe, *args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300, env=e)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    return r.stdout


def server_cmd(flags):
    return ["g++", "-std=c++17", "-O1", "-g", "-pthread", *flags, *INC, os.path.join(CSRC, "uc_server.cpp"),
            os.path.join(SAN_DIR, "stub_ctx.cpp"), os.path.join(SAN_DIR, "server_stress.cpp")]


def test_server_under_tsan(tmp_path):
    exe = build(str(tmp_path), "srv_tsan", server_cmd(TSAN))
    out = run(exe, 8, 60, env={"TSAN_OPTIONS": "halt_on_error=1"})
    assert "bad 0" in out


def test_server_under_asan(tmp_path):
    exe = build(str(tmp_path), "srv_asan", server_cmd(ASAN))
    out = run(exe, 8, 60)
    assert "bad 0" in out


def test_text_parser_under_asan(tmp_path):
    t = str(tmp_path)
    obj = os.path.join(t, "common.o")
    r = subprocess.run(["gcc", "-std=gnu11", "-O1", "-g", *ASAN, *INC, "-c", os.path.join(ORACLE, "common.c"), "-o",
                        obj], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    exe = build(t, "text_asan", ["g++", "-std=c++17", "-O1", "-g", "-pthread", *ASAN, *INC,
                                 os.path.join(CSRC, "uc_text.cpp"), os.path.join(SAN_DIR, "text_fuzz.cpp"), obj])
    out = run(exe, 4000)
    assert "bad 0" in out


def test_oracle_under_asan(tmp_path):
    exe = build(str(tmp_path), "oracle_asan", ["gcc", "-std=gnu11", "-O1", "-g", *ASAN, *INC,
                                               *(os.path.join(ORACLE, f) for f in ("common.c", "textbook.c",
                                                                                   "refpipe.c")),
                                               os.path.join(SAN_DIR, "oracle_fuzz.c")])
    out = run(exe, 60)
    assert "bad 0" in out
