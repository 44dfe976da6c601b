"""Batched search (one graph launch over instances of uneven shapes) and the
autotuner: every instance's hits must equal a separate device-statistics
search of the same coarse channel, with default and with tuned shapes, on
repeated runs."""

import numpy as np
import pytest

import seticore_amd as sc
from seticore_amd import batch as B

pytestmark = pytest.mark.gpu

FOFF, TSAMP = -2.79e-6, 18.25


def _instances(torch):
    rng = np.random.default_rng(3)
    shapes = [(16, 1 << 17), (32, 1 << 17), (64, 1 << 16), (256, 1 << 15), (16, 1 << 18),
              (13, 1 << 17), (16, 1 << 17), (128, 1 << 15)]
    data, inst = [], []
    for i, (T, N) in enumerate(shapes):
        T_r = sc.round_up_to_power_of_two(T)
        x = torch.zeros((T_r, N), dtype=torch.float32, device="cuda")
        tones = []
        for _ in range(6):
            drift = rng.uniform(-4, 4)
            tones.append((rng.uniform(0.1 * N, 0.9 * N), drift * TSAMP / (FOFF * 1e6) * 15 / (T_r - 1),
                          rng.uniform(12, 40) / np.sqrt(T_r)))
        sc.gen_waterfall(x, T, seed=100 + i, tones=tones)
        data.append(x)
        inst.append(B.instance(x, T, FOFF, TSAMP, 5.0, 0.0, 10.0, beam=-1, coarse_channel=i))
    return data, inst


def _reference_hits(data, inst):
    out = []
    for x, it in zip(data, inst):
        T_r, N = x.shape
        dd = sc.Dedopplerer(it.num_timesteps, N, FOFF, TSAMP)
        dd.set_stats_mode(sc.SC_STATS_DEVICE)
        hits = []
        dd.search(sc.FilterbankBuffer(T_r, N, data=x), sc.FilterbankMetadata(
            coarse_channel_size=N, foff=FOFF), -1, it.coarse_channel, 5.0, 0.0, 10.0, hits)
        out.append([h.key() + (h.snr, h.power, h.num_timesteps) for h in hits])
    return out


def _batch_hits(b):
    return [[h.key() + (h.snr, h.power, h.num_timesteps) for h in b.hits(i)]
            for i in range(b.count)]


def test_batch_equals_separate_searches(gpu):
    import torch
    data, inst = _instances(torch)
    want = _reference_hits(data, inst)
    assert sum(len(w) for w in want) > 0
    b = B.Batch(inst)
    info = b.info()
    assert info["shapes"] == 6 and info["graph_nodes"] > len(inst)
    for _ in range(3):  # replays, with eager work in between
        b.run()
        assert _batch_hits(b) == want
        torch.cuda.synchronize()


def test_autotuned_batch_equals_separate_searches(gpu):
    import torch
    data, inst = _instances(torch)
    want = _reference_hits(data, inst)
    tuning = []
    seen = set()
    for x in data:
        shape = tuple(x.shape)
        if shape in seen:
            continue
        seen.add(shape)
        t = B.autotune(shape[0], shape[1], FOFF, TSAMP, 5.0, reps=1)
        assert t.rounded_timesteps == shape[0] and t.num_channels == shape[1]
        tuning.append(t)
    b = B.Batch(inst, tuning)
    b.run()
    assert _batch_hits(b) == want
