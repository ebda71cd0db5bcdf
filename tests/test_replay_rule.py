"""The insertion rule k_sample_replay applies in parallel (one list position per
lane) equals the reference's sequential insertion (insertSample,
sequentialPoissonRadixTree.cuh:15-41: strict < on the keys, the displaced entry
carried down until an empty slot takes it or it leaves the list) on 20,000
random key streams with heavy ties and partly filled lists.

The rule: inserting c puts it at s0, the count of keys <= c's; a position s > s0
inside a run of equal keys keeps its entry; any other s > s0 takes the first
entry of the run ending at s - 1 (found as a prefix maximum of run starts); the
empty slots are one run, so a carried empty slot ends the insertion."""
import math
import random

INF = math.inf
EMPTY = (INF, -1)


def reference_insert(lst, c):
    key, idx = c
    for s in range(len(lst)):
        if key < lst[s][0]:
            lst[s], (key, idx) = (key, idx), lst[s]
        if idx == -1:
            break


def parallel_rule(v, c):
    k = len(v)
    kc = c[0]
    s0 = sum(1 for pos in range(k) if not (kc < v[pos][0]))
    if s0 >= k:
        return
    prev = [v[pos - 1] if pos >= 1 else EMPTY for pos in range(k)]
    start, m = [], -1
    for pos in range(k):
        if pos >= s0 and (pos == s0 or v[pos][0] != prev[pos][0]):
            m = max(m, pos)
        start.append(m)
    nv = []
    for pos in range(k):
        sp = start[pos - 1] if pos >= 1 else -1
        inside = pos > s0 and v[pos][0] == prev[pos][0]
        nv.append(v[pos] if pos < s0 or inside else c if pos == s0 else (v[sp] if sp >= 0 else EMPTY))
    v[:] = nv


def test_parallel_rule_equals_reference_insertion():
    rng = random.Random(1)
    for _ in range(20000):
        k = rng.randint(1, 12)
        n = rng.randint(0, 40)
        nkeys = rng.choice([2, 3, 5, 10, 1000])
        a = [EMPTY] * k
        b = [EMPTY] * k
        for i in range(n):
            c = (float(rng.randint(0, nkeys)), i)
            reference_insert(a, c)
            parallel_rule(b, c)
            assert a == b, (k, i, a, b)
