"""Batched search (one graph launch over instances of uneven shapes) and the
autotuner: every instance's hits must equal a separate device-statistics
search of the same coarse channel, with default and with tuned shapes, on
repeated runs."""

import numpy as np
import pytest

import seticore_amd as sc
from seticore_amd import batch as B
from parity_helpers import device_hits_match_oracle

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


def test_two_pass_and_fused32_tunings_in_a_batch(gpu):
    # Shapes tuned to the kernels that are not the defaults - the two-pass
    # search at 64..256 time steps (full keys, one search handle per shape) and
    # the register-parked one-pass kernel at 32 - must give the default
    # (staged, block-key) searches' hits.
    import torch
    data, inst = _instances(torch)
    want = _reference_hits(data, inst)
    tuning = []
    for T_r, N in sorted({tuple(x.shape) for x in data}):
        if T_r == 32:
            tuning.append(B.ScTuning(T_r, N, 2, 4, 2, 0, 0))
        elif T_r >= 64:
            tuning.append(B.ScTuning(T_r, N, 4, 8, 0, 4, 2))
    b = B.Batch(inst, tuning)
    for _ in range(2):
        b.run()
        assert _batch_hits(b) == want


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


def test_multibeam_regions_in_place(gpu, oracle):
    # The beamformer feed (beamforming_pipeline.cpp:230-245): every (beam,
    # coarse channel) region of a multibeam buffer searched in place through
    # its row stride equals searching a contiguous copy of the region
    # (MultibeamBuffer::copyRegionAsync, multibeam_buffer.cu:65-76), and the
    # oracle's search of that region (statistics differences only between the
    # two thresholds, every common hit bit-equal).
    import torch
    beams, T, ccs, ncc = 3, 16, 1 << 15, 4
    multibeam = torch.empty((beams, T, ncc * ccs), dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(8)
    for b in range(beams):
        tones = [(float(rng.uniform(0, ncc * ccs)), float(rng.uniform(-1, 1)), 4.0)
                 for _ in range(10)]
        sc.gen_waterfall(multibeam[b], T, seed=70 + b, tones=tones)
    meta = sc.FilterbankMetadata(fch1=8000.0, foff=FOFF, coarse_channel_size=ccs,
                                 source_names=["b%d" % b for b in range(beams)])
    inst = B.multibeam_instances(multibeam, T, ccs, FOFF, TSAMP, 2.0, 0.0, 8.0, metadata=meta)
    assert len(inst) == beams * ncc and inst[5].row_stride == ncc * ccs
    b = B.Batch(inst)
    b.run()
    dd = sc.Dedopplerer(T, ccs, FOFF, TSAMP)
    dd.set_stats_mode(sc.SC_STATS_DEVICE)
    total = same = 0
    for i, it in enumerate(inst):
        beam, cc = divmod(i, ncc)
        region = multibeam[beam, :, cc * ccs:(cc + 1) * ccs].contiguous()
        want = []
        dd.search(sc.FilterbankBuffer(T, ccs, data=region), meta, beam, cc, 2.0, 0.0, 8.0, want)
        got = b.hits(i)
        assert [h.key() + (h.snr, h.frequency) for h in got] == \
            [h.key() + (h.snr, h.frequency) for h in want]
        assert all(h.beam == beam for h in got)
        data = region.cpu().numpy()
        ref = oracle.search(data, T, FOFF, TSAMP, 2.0, 0.0, 8.0, fch1=8000.0,
                            coarse_channel=cc, beam=beam, num_threads=8)
        summary = device_hits_match_oracle(got, ref, data, 8.0, T - 1)
        same += summary["same"]
        total += len(got)
    assert total > 0 and same >= total - 2


def test_runs_of_one_shape_share_a_launch(gpu, oracle):
    # Consecutive instances of one 16-step shape are searched by one launch of
    # the fused kernel over up to eight coarse channels (runs of 4 and 5 here;
    # the instance of another shape in between ends a run). Each equals
    # its own per-call search, and the first its oracle search (statistics
    # differences only between the two thresholds).
    import torch
    N, T = 1 << 17, 16
    rng = np.random.default_rng(21)
    data, inst = [], []
    for i in range(10):
        n = N if i != 4 else N // 2
        x = torch.empty((T, n), dtype=torch.float32, device="cuda")
        tones = [(rng.uniform(0.1 * n, 0.9 * n), rng.uniform(-1.5, 1.5), rng.uniform(3, 9))
                 for _ in range(8)]
        sc.gen_waterfall(x, T, seed=300 + i, tones=tones)
        data.append(x)
        inst.append(B.instance(x, T, FOFF, TSAMP, 5.0, 0.0, 8.0, coarse_channel=i))
    want = _reference_hits(data, [B.instance(x, T, FOFF, TSAMP, 5.0, 0.0, 10.0,
                                              coarse_channel=i) for i, x in enumerate(data)])
    b = B.Batch([B.instance(x, T, FOFF, TSAMP, 5.0, 0.0, 10.0, coarse_channel=i)
                 for i, x in enumerate(data)])
    for _ in range(2):
        b.run()
        assert _batch_hits(b) == want
    assert sum(len(w) for w in want) > 0
    # Kernel timing: the same sequence launched directly, three path-sum
    # launches (4, 1 and 5 coarse channels), the same hits.
    b.set_timing(True)
    b.run()
    assert _batch_hits(b) == want
    ms, launches, channels = b.kernel_time()
    assert (launches, channels) == (3, 10) and ms > 0
    b.set_timing(False)
    b.run()
    assert _batch_hits(b) == want and b.kernel_time() == (0.0, 0, 0)
    # Against the oracle, at SNR 8 (more hits).
    b8 = B.Batch(inst)
    b8.run()
    x0 = data[0].cpu().numpy()
    ref = oracle.search(x0, T, FOFF, TSAMP, 5.0, 0.0, 8.0, num_threads=8)
    summary = device_hits_match_oracle(b8.hits(0), ref, x0, 8.0, T - 1)
    assert summary["same"] > 0


@pytest.mark.parametrize("dtype", [sc.SC_F16, sc.SC_BF16])
def test_half_width_runs_share_a_launch(gpu, dtype):
    # Five normalised coarse channels of one 16-step shape in FP16 or BF16
    # storage: one set launch, each channel equal to its per-call search on
    # the same samples.
    import torch
    N, T = 1 << 16, 16
    rng = np.random.default_rng(40 + dtype)
    data, inst, want = [], [], []
    for i in range(5):
        x = torch.empty((T, N), dtype=torch.float32, device="cuda")
        tones = [(rng.uniform(0.1 * N, 0.9 * N), rng.uniform(-1, 1), rng.uniform(4, 9))
                 for _ in range(6)]
        sc.gen_waterfall(x, T, seed=500 + i, tones=tones)
        xh = sc.normalize(x, dtype)[0]
        data.append(xh)
        inst.append(B.instance(xh, T, FOFF, TSAMP, 5.0, 0.0, 8.0, coarse_channel=i, dtype=dtype))
        dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
        dd.set_stats_mode(sc.SC_STATS_DEVICE)
        hits = []
        dd.search(sc.FilterbankBuffer(T, N, data=xh, dtype=dtype),
                  sc.FilterbankMetadata(coarse_channel_size=N, foff=FOFF), -1, i, 5.0, 0.0,
                  8.0, hits)
        want.append([h.key() + (h.snr, h.power, h.num_timesteps) for h in hits])
    b = B.Batch(inst)
    b.set_timing(True)
    b.run()
    assert b.kernel_time()[1:] == (1, 5)  # one launch over the five channels
    assert _batch_hits(b) == want
    assert sum(len(w) for w in want) > 0
