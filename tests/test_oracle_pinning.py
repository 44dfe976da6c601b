"""Pins the CPU oracle before it is trusted as the parity checker.

Each check is independent of the oracle's own formulas:
* the reference's Python viewer states the path geometry a second way
  (python/viewer.py:101-130, restated below, not imported);
* brute-force path walks in FP64 on integer data (exact) pin the tree's
  indexing and association against the path definition;
* the integer ramp of taylor_test.cu:14;
* the known answer of dedoppler_test.cpp:9-33;
* data/voyager.hits (real data): five hits whose path sums, recomputed from
  their stored slices, must satisfy one affine SNR relation
  snr = (P - median) / std (dedoppler.cu:308);
* statistics against a full sort in FP64.
A planted geometry mistake (in a Python copy of the tree, on the CPU) must fail
the voyager check, showing that check can tell right from wrong.
"""

import math
import os

import numpy as np
import pytest

from conftest import GOLDEN


# --- second statement of the path geometry (python/viewer.py:101-130) ------
def _viewer_interpolate_drift(total_drift, timesteps):
    rounded = 1
    while rounded < timesteps:
        rounded *= 2
    if rounded > timesteps:
        return _viewer_interpolate_drift(total_drift, rounded)[:timesteps]
    if timesteps == 2:
        return [0, total_drift]
    if total_drift < 0:
        return [-x for x in _viewer_interpolate_drift(-total_drift, timesteps)]
    shift = total_drift // (timesteps - 1)
    if shift > 0:
        post = _viewer_interpolate_drift(total_drift % (timesteps - 1), timesteps)
        return [i * shift + x for (i, x) in enumerate(post)]
    parity = total_drift % 2
    half = _viewer_interpolate_drift(total_drift // 2, timesteps // 2)
    start = half[-1] + parity
    return half + [x + start for x in half]


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64, 128, 256])
def test_path_geometry_matches_viewer(oracle, T):
    # Non-negative drift blocks: total drift = b*(T-1) + p, channel at step t is
    # t*b + offset(p, t). (The viewer mirrors negative totals, which is not how
    # the tree builds negative drift blocks, so only b >= 0 is compared.)
    for b in range(0, 3):
        for p in range(T):
            total = b * (T - 1) + p
            want = _viewer_interpolate_drift(total, T)
            got = [t * b + oracle.path_offset(T, p, t) for t in range(T)]
            assert got == want, (T, b, p)


def _brute_force_sums(data, b):
    """Every valid path walked cell by cell (float64, exact on small integers)."""
    from oracle import oracle as o
    T, N = data.shape
    out = np.full((T, N), np.nan)
    offs = np.array([[o.path_offset(T, p, t) for t in range(T)] for p in range(T)])
    tt = np.arange(T)
    for p in range(T):
        for c in range(N):
            last = c + (T - 1) * b + p
            if last < 0 or last >= N:
                continue
            cols = c + tt * b + offs[p]
            out[p, c] = data[tt, cols].astype(np.float64).sum()
    return out


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
def test_tree_matches_brute_force_paths(oracle, T):
    rng = np.random.default_rng(T)
    N = max(160, 5 * T)
    data = rng.integers(0, 16, size=(T, N)).astype(np.float32)
    for b in range(-3, 4):
        tree = oracle.basic_taylor_tree(data, b)
        bf = _brute_force_sums(data, b)
        mask = ~np.isnan(bf)
        assert mask.any()
        np.testing.assert_array_equal(tree[mask], bf[mask].astype(np.float32))


@pytest.mark.parametrize("T", [4, 8, 16, 32, 64, 128])
def test_tree_integer_ramp(oracle, T):
    # taylor_test.cu:12-16: input 1 + time + chan; sums are exact integers with
    # closed form T*(1+c) + (1+b)*T*(T-1)/2 + sum_t offset(p, t).
    N = 1000
    t = np.arange(T)[:, None]
    c = np.arange(N)[None, :]
    data = (1.0 * (1 + t + c)).astype(np.float32)
    for b in range(-2, 3):
        tree = oracle.basic_taylor_tree(data, b)
        for p in range(T):
            offsum = sum(oracle.path_offset(T, p, k) for k in range(T))
            cs = np.arange(N)
            last = cs + (T - 1) * b + p
            valid = (last >= 0) & (last < N)
            want = T * (1 + cs) + (1 + b) * T * (T - 1) // 2 + offsum
            np.testing.assert_array_equal(tree[p][valid], want[valid].astype(np.float32))


def _brute_force_top(data, bmin, bmax):
    """Max over valid (b, p), strict '>' in ascending (b, p) with the early
    return of dedoppler.cu:43-46, starting from 0."""
    T, N = data.shape
    sums = {b: _brute_force_sums(data, b) for b in range(bmin, bmax + 1)}
    top = np.zeros(N, np.float32)
    blk = np.full(N, -999, np.int32)
    off = np.full(N, -999, np.int32)
    for c in range(N):
        for b in range(bmin, bmax + 1):
            for p in range(T):
                last = (T - 1) * b + p + c
                if last < 0 or last >= N:
                    break
                s = np.float32(sums[b][p, c])
                if s > top[c]:
                    top[c], blk[c], off[c] = s, b, p
    return top, blk, off


def test_top_path_sums_brute_force(oracle):
    rng = np.random.default_rng(7)
    T, N = 8, 120
    # Small integers: many exact ties across blocks and offsets.
    data = rng.integers(0, 3, size=(T, N)).astype(np.float32)
    top = np.zeros(N, np.float32)
    blk = np.full(N, -999, np.int32)
    off = np.full(N, -999, np.int32)
    for b in range(-3, 4):
        oracle.find_top_path_sums(oracle.basic_taylor_tree(data, b), b, top, blk, off)
    wt, wb, wo = _brute_force_top(data, -3, 3)
    np.testing.assert_array_equal(top, wt)
    np.testing.assert_array_equal(blk, wb)
    np.testing.assert_array_equal(off, wo)
    # The early return: channels c < (T-1)*3 never see block -3.
    assert not np.any((blk == -3) & (np.arange(N) < 21))


def test_dedoppler_known_answer(oracle):
    # dedoppler_test.cpp:9-33
    T, N = 8, 1000
    data = np.zeros((T, N), np.float32)
    data[0, :] = np.array([0.1 * c / N for c in range(N)], np.float64).astype(np.float32)
    for t in range(8):
        data[t, 70 + t // 2] = 1.0
    res = oracle.search(data, T, foff=1.0, tsamp=1.0, max_drift=0.01, min_drift=0.01,
                        snr_threshold=200.0, beam=-1, coarse_channel=555)
    assert len(res.hits) == 1
    h = res.hits[0]
    assert (h["index"], h["drift_steps"], h["beam"], h["coarse_channel"]) == (70, 3, -1, 555)


def test_drift_parameters_match_survey_table(oracle):
    # dedoppler.cu:87,195-198,274 at the benchmark configurations.
    assert oracle.drift_params(16, -2.79e-6, 18.25, 4.0)[:3] == (-27, 26, 786)
    assert oracle.drift_params(16, -2.79e-6, 18.25, 10.0)[:3] == (-66, 65, 1964)
    # The known-answer search: blocks -1..0, window 2.
    assert oracle.drift_params(8, 1.0, 1.0, 0.01)[:3] == (-1, 0, 2)


def test_stats_against_full_sort(oracle):
    rng = np.random.default_rng(3)
    for n in (1000, 4097, 65536):
        v = (16 + 4 * rng.standard_normal(n)).astype(np.float32)
        med, sd = oracle.host_stats(v)
        s = np.sort(v.astype(np.float64))
        assert med == np.float32(s[n // 2])  # upper median, exact
        first, last = math.ceil(0.05 * n), math.floor(0.95 * n)
        trimmed = s[first:last + 1]
        want = math.sqrt(((trimmed - trimmed.mean()) ** 2).mean())
        assert abs(sd - want) <= 2e-4 * want


def test_threaded_search_equals_single(oracle):
    rng = np.random.default_rng(11)
    T, N = 16, 20000
    data = rng.exponential(1.0, size=(T, N)).astype(np.float32)
    a = oracle.search(data, 16, -2.79e-6, 18.25, 2.0, 0.0, 5.0, num_threads=1)
    b = oracle.search(data, 16, -2.79e-6, 18.25, 2.0, 0.0, 5.0, num_threads=5)
    np.testing.assert_array_equal(a.top_sums, b.top_sums)
    np.testing.assert_array_equal(a.top_blocks, b.top_blocks)
    np.testing.assert_array_equal(a.top_offsets, b.top_offsets)
    assert a.hits == b.hits


def test_max_drift_zero_rejected(oracle):
    data = np.ones((4, 16), np.float32)
    with pytest.raises(ValueError):
        oracle.search(data, 4, 1.0, 1.0, 0.0, 0.0, 10.0)


# --- data/voyager.hits: real data, affine SNR relation ----------------------
def _voyager():
    from seticore_amd.hitfile import read_hits
    return read_hits(os.path.join(GOLDEN, "voyager.hits"))


def _voyager_path_sums(tree_fn):
    out = []
    for h in _voyager():
        s, f = h["signal"], h["filterbank"]
        T, W = f["num_timesteps"], f["num_channels"]
        data = f["data"].reshape(T, W)
        b = s["drift_steps"] // (T - 1)
        p = s["drift_steps"] - b * (T - 1)
        sums = tree_fn(data, b)
        out.append((float(sums[p, s["index"]]), float(s["snr"])))
    return out


def _affine_residual(pairs):
    P = np.array([p for p, _ in pairs])
    snr = np.array([s for _, s in pairs])
    # snr = (P - median) / std  <=>  P = median + std * snr: least squares.
    A = np.stack([np.ones_like(snr), snr], axis=1)
    (median, std), *_ = np.linalg.lstsq(A, P, rcond=None)
    return np.max(np.abs(A @ np.array([median, std]) - P) / P), median, std


def test_voyager_fixture_contents():
    hits = _voyager()
    assert len(hits) == 5
    assert [h["signal"]["drift_steps"] for h in hits] == [36, 39, 37, 38, 39]
    for h in hits:
        s, f = h["signal"], h["filterbank"]
        assert s["index"] == 40 and f["num_timesteps"] == 16
        assert f["data"].size == 16 * f["num_channels"]
        # frequency is the slice's fch1 plus index * foff (dedoppler_hit.cpp:20-22)
        assert abs(s["frequency"] - (f["fch1"] + 40 * f["foff"])) < 1e-9


def test_voyager_affine_snr(oracle):
    resid, median, std = _affine_residual(_voyager_path_sums(oracle.basic_taylor_tree))
    assert resid < 1e-6, resid
    assert 1.7e11 < median < 1.8e11 and 1.8e10 < std < 2.0e10


def _mutated_tree(data, b):
    """The tree with a plausible mistake: the second half shifted by p//2
    instead of (p+1)//2."""
    T, N = data.shape
    src = data.astype(np.float32).copy()
    L = 2
    while L <= T:
        dst = np.zeros_like(src)
        for tb in range(T // L):
            for p in range(L):
                shift = p // 2 + b * L // 2
                for c in range(N):
                    if 0 <= c + shift < N:
                        dst[tb * L + p, c] = src[tb * L + p // 2, c] + \
                            src[tb * L + p // 2 + L // 2, c + shift]
        src = dst
        L *= 2
    return src


def test_voyager_check_rejects_planted_mistake():
    resid, _, _ = _affine_residual(_voyager_path_sums(_mutated_tree))
    assert resid > 1e-4, resid
