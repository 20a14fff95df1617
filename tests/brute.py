"""Pure-Python definitional checks for tiny inputs (test helper).

brute_unitigs: O(n^2) string-level restatement of the definition -- edges
by comparing every pair of k-mers, degrees by counting, unitigs as maximal
non-branching paths, cycles from their smallest k-mer.

check_partition: validates any unitig text against the definition without
constructing it: every k-mer exactly once, every adjacent pair a kept
edge, every line maximal.  Linear time; usable at moderate sizes.
"""
from collections import Counter


def brute_unitigs(kmers):
    k = len(kmers[0]) if kmers else 0
    succs = {u: [v for v in kmers if v[:-1] == u[1:]] for u in kmers}
    preds = {v: [u for u in kmers if u[1:] == v[:-1]] for v in kmers}
    edge = {}
    for u in kmers:
        if len(succs[u]) == 1 and len(preds[succs[u][0]]) == 1:
            edge[u] = succs[u][0]
    has_in = set(edge.values())
    seen = set()
    lines = []
    for u in kmers:
        if u in has_in:
            continue
        s, v = u, edge.get(u)
        seen.add(u)
        while v is not None:
            s += v[-1]
            seen.add(v)
            v = edge.get(v)
        lines.append(s)
    for u in kmers:
        if u in seen:
            continue
        cyc = [u]
        v = edge[u]
        while v != u:
            cyc.append(v)
            v = edge[v]
        start = cyc.index(min(cyc))
        cyc = cyc[start:] + cyc[:start]
        seen.update(cyc)
        lines.append(cyc[0] + "".join(x[-1] for x in cyc[1:]))
    return lines


def check_partition(kmers, lines, k):
    """Raises AssertionError when `lines` is not the unitig set of `kmers`."""
    kset = set(kmers)
    assert len(kset) == len(kmers), "duplicate k-mers in the input"
    pref = Counter(x[:-1] for x in kmers)
    suf = Counter(x[1:] for x in kmers)
    seen = Counter()

    def kept(u, v):
        return u[1:] == v[:-1] and pref[u[1:]] == 1 and suf[v[:-1]] == 1

    for line in lines:
        assert len(line) >= k, f"line shorter than k: {line!r}"
        win = [line[i:i + k] for i in range(len(line) - k + 1)]
        for w in win:
            assert w in kset, f"{w} is not an input k-mer"
            seen[w] += 1
        for a, b in zip(win, win[1:]):
            assert kept(a, b), f"{a}->{b} is not a kept edge"
        first, last = win[0], win[-1]
        cyc = kept(last, first) and len(set(win)) == len(win)
        if not cyc:
            # maximal: no kept edge into the first or out of the last k-mer
            for c in "ACGT":
                p = c + first[:-1]
                assert not (p in kset and kept(p, first)), f"line not maximal at its start: {line}"
                q = last[1:] + c
                assert not (q in kset and kept(last, q)), f"line not maximal at its end: {line}"
    missing = [x for x in kmers if seen[x] != 1]
    assert not missing, f"k-mers not covered exactly once: {missing[:5]}"
