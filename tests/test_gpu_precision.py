"""Reduced-precision storage: FP16/BF16 samples after per-coarse-channel
normalisation on the GPU ((x - median column sum / T) / FP64 deviation), FP32
tree accumulation.

* Tree parity: the half-width search equals the oracle run on the same
  samples widened to float32, bit for bit (the tree converts each sample once
  and then adds in float32 in the reference's association).
* Against the FP32 search of the raw data: the strong-hit sets agree and the
  SNR error against the FP64 oracle is bounded (reported per storage type).
Workload: a reduced form of the reduced-precision configuration - exponential
noise, drifting tones, stationary narrowband RFI up to 1e6 x noise and
broadband bursts on 3 % of the time steps at 100 x noise.
"""

import math

import numpy as np
import pytest

import seticore_amd as sc

pytestmark = pytest.mark.gpu

T, N = 16, 1 << 18
FOFF, TSAMP = -2.79e-6, 18.25


def _rfi_waterfall(torch, seed):
    rng = np.random.default_rng(seed)
    tones = []
    for _ in range(12):
        drift = rng.uniform(-5, 5)
        slope = drift * TSAMP / (FOFF * 1e6)
        tones.append((float(rng.uniform(2000, N - 2000)), float(slope),
                      float(rng.uniform(10, 60) / 4.0)))
    # Stationary narrowband RFI up to 1e6 x the noise mean.
    for _ in range(6):
        tones.append((float(rng.uniform(1000, N - 1000)), 0.0,
                      float(10 ** rng.uniform(2, 6))))
    x = torch.empty((T, N), dtype=torch.float32, device="cuda")
    sc.gen_waterfall(x, T, seed=seed, tones=tones)
    bursts = rng.choice(T, size=max(1, round(0.03 * T)), replace=False)
    for t in bursts:
        x[int(t)] *= 101.0  # broadband burst: 100 x noise added on the row
    return x


@pytest.mark.parametrize("T_r", [32, 64, 128, 256])
@pytest.mark.parametrize("dtype", [sc.SC_F16, sc.SC_BF16])
def test_half_storage_parity_longer_windows(gpu, oracle, dtype, T_r):
    # The default one-pass kernels at 32..256 time steps (LDS-staged; at 256
    # one workgroup per offset group) on half-width samples equal the oracle
    # on the same samples widened to float32.
    import torch
    N = 1 << 15 if T_r <= 64 else 1 << 13
    print("half T_r=%d dtype=%d N=%d" % (T_r, dtype, N), flush=True)
    x = torch.empty((T_r, N), dtype=torch.float32, device="cuda")
    tones = ([(9000.0, 0.7, 3.0), (20000.0, -0.4, 2.0)] if T_r <= 64 else
             [(2000.0, 0.7, 3.0), (5000.0, -0.4, 2.0)])
    sc.gen_waterfall(x, T_r, seed=T_r, tones=tones)
    xn, _, _ = sc.normalize(x, dtype)
    widened = xn.float().cpu().numpy()
    dd = sc.Dedopplerer(T_r, N, FOFF, TSAMP)
    hits = []
    dd.search(sc.FilterbankBuffer(T_r, N, data=xn, dtype=dtype),
              sc.FilterbankMetadata(coarse_channel_size=N), -1, 0, 2.0, 0.0, 6.0, hits)
    st = dd.state()
    want = oracle.search(widened, T_r, FOFF, TSAMP, 2.0, 0.0, 6.0, num_threads=8)
    np.testing.assert_array_equal(st["column_sums"], want.column_sums)
    np.testing.assert_array_equal(st["top_sums"], want.top_sums)
    assert [(h.index, h.drift_steps, h.snr) for h in hits] == \
        [(h["index"], h["drift_steps"], h["snr"]) for h in want.hits]


@pytest.mark.parametrize("dtype", [sc.SC_F16, sc.SC_BF16])
def test_half_storage_tree_parity(gpu, oracle, dtype):
    import torch
    x = _rfi_waterfall(torch, 3)
    xn, loc, sd = sc.normalize(x, dtype)
    widened = xn.float().cpu().numpy()
    # loc is the median (rank N/2) of the raw column sums, summed in the
    # reference's order, over T; std the FP64 deviation of all samples.
    raw = x.cpu().numpy()
    assert loc == float(np.sort(oracle.sum_columns(raw))[N // 2]) / T
    ref = raw.astype(np.float64)
    assert abs(sd - ref.std()) <= 1e-9 * ref.std()
    dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
    hits = []
    dd.search(sc.FilterbankBuffer(T, N, data=xn, dtype=dtype),
              sc.FilterbankMetadata(coarse_channel_size=N), -1, 0, 5.0, 0.0, 8.0, hits)
    st = dd.state()
    want = oracle.search(widened, T, FOFF, TSAMP, 5.0, 0.0, 8.0, num_threads=8)
    np.testing.assert_array_equal(st["column_sums"], want.column_sums)
    np.testing.assert_array_equal(st["top_sums"], want.top_sums)
    assert [(h.index, h.drift_steps, h.snr) for h in hits] == \
        [(h["index"], h["drift_steps"], h["snr"]) for h in want.hits]
    # The normalised column sums are centred on zero, so the SNR threshold is a
    # positive path sum. (Around the mean, which the RFI pulls up, the median
    # column sum would lie this many deviations below zero:)
    print("mean-centred median: %.1f deviations" % (-T * (ref.mean() - loc) / sd / st["std"]))
    assert abs(st["median"]) <= 0.05 * st["std"]


@pytest.mark.parametrize("dtype", [sc.SC_F16, sc.SC_BF16])
def test_half_storage_hits_and_snr_error(gpu, oracle, dtype):
    import torch
    x = _rfi_waterfall(torch, 5)
    raw = x.cpu().numpy()
    dd32 = sc.Dedopplerer(T, N, FOFF, TSAMP)
    h32 = []
    dd32.search(sc.FilterbankBuffer(T, N, data=x), sc.FilterbankMetadata(coarse_channel_size=N),
                -1, 0, 5.0, 0.0, 8.0, h32)
    xn, _, _ = sc.normalize(x, dtype)
    dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
    hh = []
    dd.search(sc.FilterbankBuffer(T, N, data=xn, dtype=dtype),
              sc.FilterbankMetadata(coarse_channel_size=N), -1, 0, 5.0, 0.0, 8.0, hh)
    strong32 = {(h.index, h.drift_steps) for h in h32 if h.snr >= 20}
    got = {(h.index, h.drift_steps) for h in hh}
    assert strong32, "the workload must have strong hits"
    # Every strong FP32 hit is found; quantisation may move a near-tie to the
    # neighbouring start channel or drift offset.
    for i, d in strong32:
        assert any(abs(i - gi) <= 1 and abs(d - gd) <= 1 for gi, gd in got), (i, d)
    exact = len(strong32 & got) / len(strong32)
    print("strong hits found exactly: %.3f" % exact)
    # SNR error against the FP64 oracle (FP64 column sums, exact ranks).
    med64, _, std64 = oracle.stats_fp64(raw)
    errs = []
    for h in hh:
        if (h.index, h.drift_steps) not in strong32 or h.snr < 20:
            continue
        b = h.drift_steps // 15
        p = h.drift_steps - 15 * b
        snr64 = (oracle.path_sum_fp64(raw, h.index, b, p) - med64) / std64
        errs.append(abs(h.snr - snr64) / abs(snr64))
    bound = 2e-3 if dtype == sc.SC_F16 else 2e-2
    print("max relative SNR error vs FP64 (%s): %.3g" % (
        "fp16" if dtype == sc.SC_F16 else "bf16", max(errs)))
    assert max(errs) < bound
