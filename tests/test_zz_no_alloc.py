"""The allocation-free search (sc_dd_search_static) against sc_dd_search.

Two calls in a row on one GPU, the second with a smaller input (fewer
channels, more time steps): every hit of each call equals, field by field,
the hits of the handle-based device-statistics search of the same data
(dedoppler.cu:186-320 semantics; the same kernels run on the library's fixed
per-GPU array instead of a handle's allocations). Inputs beyond the array's
limits are rejected before anything runs.
"""

import math

import numpy as np
import pytest

import seticore_amd as sc

pytestmark = pytest.mark.gpu

FOFF, TSAMP = -2.79e-6, 18.25
FIELDS = ["frequency", "index", "drift_steps", "drift_rate", "snr", "coarse_channel", "beam",
          "num_timesteps", "power", "incoherent_power"]


def _data(Tu, N, seed, max_drift, gain=1.0, dc_spike=0.0):
    import torch
    T = sc.round_up_to_power_of_two(Tu)
    rng = np.random.default_rng(seed)
    tones = []
    for _ in range(12):
        drift = float(rng.uniform(-max_drift, max_drift))
        tones.append((float(rng.uniform(0.06 * N, 0.94 * N)), drift * TSAMP / (FOFF * 1e6),
                      gain * float(rng.uniform(10.0, 60.0)) / math.sqrt(Tu)))
    x = torch.empty((T, N), dtype=torch.float32, device="cuda")
    sc.gen_waterfall(x, Tu, seed=seed, gain=gain, taper_frac=0.05, dc_spike=dc_spike,
                     tones=tones)
    torch.cuda.synchronize()
    return x


def _handle_search(x, Tu, meta, beam, cc, max_drift, snr):
    T, N = x.shape
    dd = sc.Dedopplerer(Tu, N, FOFF, TSAMP)
    dd.set_stats_mode(sc.SC_STATS_DEVICE)
    hits = []
    dd.search(sc.FilterbankBuffer(T, N, data=x), meta, beam, cc, max_drift, 0.0, snr, hits)
    dd.close()
    return hits


def _fields(hits):
    return [tuple(getattr(h, f) for f in FIELDS) for h in hits]


def test_static_search_equals_handle_search(gpu):
    cases = [
        # (unrounded time steps, channels, seed, max drift, beam, coarse channel)
        (16, 1 << 20, 11, 10.0, 1, 5),   # 132 drift blocks, fused kernel
        (50, 1 << 17, 12, 5.0, -1, 2),   # padded to 64 rows, staged kernel
    ]
    for Tu, N, seed, max_drift, beam, cc in cases:
        print("case", Tu, N, seed, flush=True)
        x = _data(Tu, N, seed, max_drift, gain=1.3, dc_spike=50.0)
        T = x.shape[0]
        meta = sc.FilterbankMetadata(fch1=8419.0, foff=FOFF, coarse_channel_size=N,
                                     source_names=["a", "b", "c"])
        want = _handle_search(x, Tu, meta, beam, cc, max_drift, 10.0)
        got = sc.search_static(sc.FilterbankBuffer(T, N, data=x), Tu, FOFF, TSAMP, meta, beam,
                               cc, max_drift, 0.0, 10.0)
        assert len(want) >= 5
        assert _fields(got) == _fields(want)
        assert all(h.num_timesteps == Tu and h.coarse_channel == cc for h in got)


def test_static_search_limits(gpu):
    import torch
    meta = sc.FilterbankMetadata(fch1=8419.0, foff=FOFF, coarse_channel_size=1 << 21)
    big = torch.zeros((16, (1 << 20) + 4096), dtype=torch.float32, device="cuda")
    with pytest.raises(sc.SeticoreError) as e:
        sc.search_static(sc.FilterbankBuffer(16, (1 << 20) + 4096, data=big), 16, FOFF, TSAMP,
                         meta, -1, 0, 4.0, 0.0, 10.0)
    assert e.value.status == sc.SC_ERR_INVALID_ARGUMENT and "channels" in str(e.value)
    long_ = torch.zeros((512, 4096), dtype=torch.float32, device="cuda")
    with pytest.raises(sc.SeticoreError) as e:
        sc.search_static(sc.FilterbankBuffer(512, 4096, data=long_), 300, FOFF, TSAMP, meta, -1,
                         0, 4.0, 0.0, 10.0)
    assert e.value.status == sc.SC_ERR_UNSUPPORTED
