"""Pins the oracle's nearestsamples restatement (oracle.nearest_samples,
nearestSamplesCuda.cu:290-404): a hand-checked case (surface choice by the first
largest weight, self and empty slots excluded, duplicates merged, ties by index,
-1 padding) and invariants on random inputs against a brute force over all
(surface, candidate) pairs."""
import numpy as np

from oracle import oracle as O


def test_hand_case():
    pts = np.array([[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 0.5], [-1.0, 0.0, 0.0]]],
                   np.float32)
    # pixel A: surface 0 (weight 0.9), neighbours 1, 2, 3; pixel B: surface 0 again
    # (tie 0.5 / 0.5: the first slot wins), neighbours 4 and 1 (duplicate); pixel C:
    # surface 2, neighbour 2 itself only; pixel D: empty
    idx = np.array([[[[1, 0, 2, 3], [0, 4, 1, -1]], [[2, 2, -1, -1], [-1, -1, -1, -1]]]], np.int32)
    w = np.array([[[[0.1, 0.9, 0.2, 0.3], [0.5, 0.5, 0.1, 0.0]], [[0.4, 0.1, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0]]]],
                 np.float32)
    out = O.nearest_samples(idx, w, pts, 3)
    # splat 0: distances^2 to 1, 2, 3, 4 = 1, 4, 0.25, 1 -> 3, then 1 and 4 tied at 1 (index order)
    assert out[0, 0].tolist() == [3, 1, 4]
    assert out[0, 2].tolist() == [-1, -1, -1]
    assert (out[0, [1, 3, 4]] == -1).all()


def test_random_against_brute_force():
    rng = np.random.default_rng(1)
    N, K = 60, 5
    pts = rng.normal(size=(1, N, 3)).astype(np.float32)
    idx = rng.integers(-1, N, size=(2, 6, 5, 7)).astype(np.int32)
    w = rng.uniform(size=idx.shape).astype(np.float32)
    out = O.nearest_samples(idx, w, pts, K)
    pairs = {}
    for b in range(2):
        for i in range(6):
            for j in range(5):
                row, ww = idx[b, i, j], w[b, i, j]
                s = row[int(np.argmax(ww))]
                if s < 0:
                    continue
                for c in row:
                    if c >= 0 and c != s:
                        d = pts[0, s].astype(np.float64) - pts[0, c].astype(np.float64)
                        pairs.setdefault(int(s), {})[int(c)] = float((d * d).sum())
    for n in range(N):
        got = [x for x in out[0, n] if x >= 0]
        ref = sorted(pairs.get(n, {}).items(), key=lambda kv: (kv[1], kv[0]))[:K]
        assert len(got) == len(ref)
        assert len(set(got)) == len(got) and n not in got
        # the float32 distance order agrees with FP64 except within rounding
        dg = [pairs[n][c] for c in got]
        assert all(dg[q] <= dg[q + 1] * (1 + 1e-6) for q in range(len(dg) - 1))
        assert set(got) == {c for c, _ in ref} or abs(max(dg) - ref[-1][1]) <= 1e-6 * ref[-1][1]
