"""Build recipe for the native parts (in-tree, gfx950).

Outputs under unitigcoarsening_amd/_lib/:
  libuc.so        the compaction library (HIP kernels + C ABI, include/uc_api.h)
  uc_main         the command-line drop-in for the reference's `main`
  libuc_synth.so  the seeded synthetic k-mer generator (host only)
  uc_copy_probe   a plain device copy of known size (traffic-counter calibration)

Run ``python -m unitigcoarsening_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "_lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed fp32 instructions off for all HIP code (no float math here, but the
# flag keeps the compiler from emitting them anywhere)
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    f"-I{INCLUDE}",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def _atomic_move(tmp, dst):
    os.replace(tmp, dst)


def build_lib(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    objs = []
    core_o = os.path.join(LIB, "uc_core.o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "uc_core.hip"), "-o", core_o], verbose)
    objs.append(core_o)
    text_o = os.path.join(LIB, "uc_text.o")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", "-c",
          os.path.join(CSRC, "uc_text.cpp"), "-o", text_o], verbose)
    objs.append(text_o)
    srv_o = os.path.join(LIB, "uc_server.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{INCLUDE}", "-c",
          os.path.join(CSRC, "uc_server.cpp"), "-o", srv_o], verbose)
    objs.append(srv_o)
    out = os.path.join(LIB, "libuc.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread"], verbose)
    _atomic_move(tmp, out)
    for o in objs:
        os.remove(o)
    return out


def build_cli(verbose=False):
    out = os.path.join(LIB, "uc_main")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", os.path.join(CSRC, "uc_main.cpp"),
          f"-L{LIB}", "-luc", "-Wl,-rpath,$ORIGIN", "-o", tmp], verbose)
    _atomic_move(tmp, out)
    return out


def build_probe(verbose=False):
    """Plain copy kernel of known size (counter calibration)."""
    out = os.path.join(LIB, "uc_copy_probe")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "uc_probe.hip"), "-o", tmp], verbose)
    _atomic_move(tmp, out)
    return out


def build_synth(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libuc_synth.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", f"-I{INCLUDE}",
          os.path.join(CSRC, "synth.cpp"), "-o", tmp], verbose)
    _atomic_move(tmp, out)
    return out


def build_oracle(verbose=False):
    """The CPU oracle (test infrastructure) under oracle/."""
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")], verbose)
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all(verbose=False):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found")
    paths = [build_lib(verbose), build_cli(verbose), build_synth(verbose), build_probe(verbose),
             build_oracle(verbose)]
    return paths


if __name__ == "__main__":
    for p in build_all(verbose="-v" in sys.argv):
        print(p)
