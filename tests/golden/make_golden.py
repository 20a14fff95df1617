#!/usr/bin/env python3
"""Writes the golden fixtures of this directory.

Each fixture is a k-mer file in the reference's text format (count line,
then one k-mer per line, driver.cpp:251-256) and the expected unitigs, one
per line, sorted.  Expected outputs come from the textbook oracle and are
cross-checked here against the brute-force definition (small fixtures) and
the restated reference pipeline (k = 15 mod 16) before anything is written.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import oracle  # noqa: E402
import unitigcoarsening_amd as uc  # noqa: E402
from brute import brute_unitigs, check_partition  # noqa: E402

COMP = str.maketrans("ACGT", "TGCA")


def canonical_genome_kmers(rng, glen, k):
    g = "".join(rng.choice("ACGT") for _ in range(glen))
    return sorted({min(g[i:i + k], g[i:i + k].translate(COMP)[::-1]) for i in range(glen - k + 1)})


def circular_kmers(rng, length, k):
    g = "".join(rng.choice("ACGT") for _ in range(length))
    c = g + g[:k - 1]
    return sorted({c[i:i + k] for i in range(length)})


def fixtures():
    rng = random.Random(20261202)
    out = []
    # hand-built: a path, a branch, a merge, an isolated 3-cycle, a self-loop
    out.append(("hand_k4", 4, 2, ["ACGT", "CGTA", "GTAC", "TACC", "TACG", "GGCA", "TGGC", "AGGC",
                                  "GACG", "CGAC", "ACGA", "TTTT", "CCAT"]))
    # every other binary-alphabet 4-mer: dense branching
    b4 = ["".join(p) for p in __import__("itertools").product("AC", repeat=4)]
    out.append(("binary_k4", 4, 2, b4[::2] + b4[1::5]))
    out.append(("genome_k15", 15, 7, canonical_genome_kmers(rng, 3000, 15)))
    out.append(("genome_k31", 31, 11, canonical_genome_kmers(rng, 6000, 31)))
    out.append(("genome_k47", 47, 11, canonical_genome_kmers(rng, 4000, 47)))
    out.append(("genome_k63", 63, 11, canonical_genome_kmers(rng, 4000, 63)))
    out.append(("circular_k31", 31, 11, circular_kmers(rng, 700, 31) + circular_kmers(rng, 45, 31)))
    rep = uc.medium_model(4242, 12000, 31)
    out.append(("repeats_errors_k31", 31, 11, uc.decode(rep, 31)))
    return out


def main():
    manifest = []
    for name, k, l, kmers in fixtures():
        kmers = sorted(set(kmers))
        packed = uc.encode(kmers, k)
        text, n_unitigs, n_cycles = oracle.textbook(packed, k)
        lines = sorted(text.decode().splitlines())
        check_partition(kmers, lines, k)
        if len(kmers) <= 3000:
            assert lines == sorted(brute_unitigs(kmers)), name
        if k % 16 == 15 and n_cycles == 0:
            ref, _ = oracle.refpipe(packed, k, l, 64, oracle.generate_permutation(l, 1))
            assert sorted(ref.decode().splitlines()) == lines, name
        inp = f"{len(kmers)}\n" + "".join(x + "\n" for x in kmers)
        exp = "".join(x + "\n" for x in lines)
        with open(os.path.join(HERE, f"{name}.kmers.txt"), "w") as f:
            f.write(inp)
        with open(os.path.join(HERE, f"{name}.unitigs.txt"), "w") as f:
            f.write(exp)
        manifest.append({"name": name, "k": k, "l": l, "n_kmers": len(kmers), "n_unitigs": n_unitigs,
                         "n_cycles": n_cycles, "sha256_unitigs": hashlib.sha256(exp.encode()).hexdigest()})
        print(f"{name}: k={k} n={len(kmers)} unitigs={n_unitigs} cycles={n_cycles}")
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
