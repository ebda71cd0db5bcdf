"""Pins the oracle's scalar building blocks against what the mathematics and
published references fix: exp/log accuracy, Philox4x32-10 known-answer
vectors (Salmon et al., SC'11, Random123 kat_vectors) plus an independent
implementation (rocRAND), the (0,1] uniform mapping, the Hermite truncation
bound (hermitefmm.cuh:184-195), Morton codes (buildRadixTree.cuh:18-54) and
the launch count of the sequential sampler (sequentialPoissonDynamicTree.cuh:198-287)."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------ exp / log
def test_expf_accuracy():
    xs = np.concatenate([np.linspace(-87.0, 88.0, 20001), np.random.default_rng(0).uniform(-30, 30, 5000)])
    got = np.array([O.expf(x) for x in xs.astype(np.float32)], np.float64)
    ref = np.exp(xs.astype(np.float32).astype(np.float64))
    rel = np.abs(got - ref) / ref
    assert rel.max() < 4e-7, rel.max()          # < 4 ulp of FP32 over the whole range


def test_expf_special_values():
    assert O.expf(0.0) == 1.0
    assert O.expf(-100.0) == 0.0
    assert math.isinf(O.expf(100.0))
    assert math.isnan(O.expf(float("nan")))
    assert abs(O.expf(1.0) - math.e) / math.e < 2.4e-7


def test_logf_accuracy():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-80, 80, 20000)), np.linspace(0.5, 2.0, 5001),
                         [1e-40, 3e-39, 1.17549435e-38]]).astype(np.float32)
    got = np.array([O.logf(x) for x in xs], np.float64)
    ref = np.log(xs.astype(np.float64))
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    assert err.max() < 3e-7, err.max()


def test_logf_special_values():
    assert O.logf(1.0) == 0.0
    assert O.logf(0.0) == -math.inf
    assert math.isnan(O.logf(-1.0))
    assert O.logf(math.inf) == math.inf


# ------------------------------------------------------------------ Philox
# Random123 kat_vectors: philox4x32 10 rounds, counter / key / expected output
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(O.philox4x32_10(ctr, key)) == out


def test_philox_matches_rocrand(tmp_path):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None or not os.path.exists("/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"):
        pytest.skip("rocRAND headers / hipcc not available")
    exe = str(tmp_path / "philox_rocrand")
    subprocess.check_call([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1", "-o", exe,
                           os.path.join(HERE, "philox_rocrand.cpp")])
    rng = np.random.default_rng(7)
    rows = rng.integers(0, 2 ** 32, size=(256, 6), dtype=np.uint64)
    txt = "\n".join(" ".join(str(int(v)) for v in r) for r in rows) + "\n"
    res = subprocess.run([exe], input=txt, capture_output=True, text=True, check=True).stdout.split("\n")
    for r, line in zip(rows, res):
        ref = tuple(int(v) for v in line.split())
        got = tuple(O.philox4x32_10(tuple(int(v) for v in r[:4]), tuple(int(v) for v in r[4:])))
        assert got == ref


def test_uniform_mapping():
    """u = ((x >> 8) + 1) 2^-24 with x the (ordinal & 3)-th word of Philox at counter
    (ordinal >> 2, launch, pixel, view) and key (seed lo, seed hi)."""
    seed = 0x1234_5678_9abc_def0
    for (view, pixel, launch, ordinal) in [(0, 0, 0, 0), (1, 77, 3, 5), (2, 4095, 6, 39), (0, 262143, 0, 2)]:
        words = O.philox4x32_10((ordinal >> 2, launch, pixel, view), (seed & 0xFFFFFFFF, seed >> 32))
        x = words[ordinal & 3]
        assert O.uniform(seed, view, pixel, launch, ordinal) == ((x >> 8) + 1) * 2.0 ** -24


def test_uniform_distribution():
    """(0,1] support and uniformity (Kolmogorov-Smirnov, n = 20000)."""
    from scipy import stats
    u = np.array([O.uniform(99, 0, p, 0, o) for p in range(500) for o in range(40)])
    assert u.min() > 0.0 and u.max() <= 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    # independent streams per launch
    v = np.array([O.uniform(99, 0, p, 1, o) for p in range(500) for o in range(40)])
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.03


# ------------------------------------------------------------------ Hermite truncation bound
def _bound64(r, n, p):
    if r >= 1.0:
        return r
    b = r ** p / math.sqrt(math.factorial(p))
    return n / (1.0 - r) ** 2 * (b * b + 2.0 * (1.0 - r ** p) * b)


@pytest.mark.parametrize("p", [0, 1, 2, 3, 5, 8, 12, 16])
def test_hermite_error_closed_form(p):
    for r in [0.0, 0.01, 0.1, 0.3, 0.5, 0.7, 0.9, 0.99]:
        for n in [1.0, 7.0, 1000.0, 1e6]:
            got = O.hermite_error(r, n, p)
            ref = _bound64(np.float32(r), np.float32(n), p)
            assert abs(got - ref) <= 2e-5 * abs(ref) + 1e-30, (r, n, p, got, ref)


def test_hermite_error_special_cases():
    # p = 0: b = 1, bound = n / (1 - r)^2
    assert abs(O.hermite_error(0.5, 3.0, 0) - 12.0) < 1e-6
    # r = 0: exact expansion, zero error for p >= 1
    assert O.hermite_error(0.0, 10.0, 4) == 0.0
    # r >= 1: not convergent, returns r (never accepted)
    assert O.hermite_error(1.5, 10.0, 4) == 1.5
    # monotone decreasing in p for r < 1
    vals = [O.hermite_error(0.4, 100.0, p) for p in range(1, 17)]
    assert all(a > b for a, b in zip(vals, vals[1:]))


# ------------------------------------------------------------------ Morton codes
def _interleave(v):
    out = 0
    for bit in range(10):
        out |= ((v >> bit) & 1) << (3 * bit)
    return out


def test_morton_known_values():
    assert O.morton2(0.0, 0.0) == 0
    # expandBits(1023) = 0x09249249; key = 2 E(x) + E(y)
    assert O.morton2(1.0, 1.0) == 3 * 0x09249249
    assert O.morton2(1.0, 0.0) == 2 * 0x09249249
    assert O.morton2(1.0 / 1024, 0.0) == 2
    assert O.morton2(0.0, 1.0 / 1024) == 1
    assert O.morton2(2.0 / 1024, 3.0 / 1024) == 2 * 8 + 9
    # clamping outside [0, 1]
    assert O.morton2(-0.5, 2.0) == 0x09249249


def test_morton_brute_interleave():
    rng = np.random.default_rng(3)
    for ux, uy in rng.uniform(-0.1, 1.1, size=(2000, 2)).astype(np.float32):
        cx = int(min(max(np.float32(ux) * np.float32(1024.0), 0.0), 1023.0))
        cy = int(min(max(np.float32(uy) * np.float32(1024.0), 0.0), 1023.0))
        assert O.morton2(ux, uy) == 2 * _interleave(cx) + _interleave(cy)


# ------------------------------------------------------------------ launches
def test_num_launches_is_ceil_sqrt():
    for k in range(1, 300):
        assert O.num_launches(k) == math.ceil(math.sqrt(k)), k
