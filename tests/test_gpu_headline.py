"""The benchmarked configurations against the CPU oracle, hit for hit.

* File scale (16 time steps x 2^20 channels, per-coarse-channel gain, 5 %
  cosine taper, 50x DC spike at the centre bin, tones at up to +-10 Hz/s,
  max drift 10 Hz/s = drift blocks -66..65, SNR 10): host-exact searches are
  bit-exact on every hit field and on the per-channel state; the batched
  device-statistics search the benchmark times differs from the oracle only by
  windows whose largest top sum lies between the two thresholds.
* Batched uneven search (time steps 16..256, widths 2^17..2^20, max drift
  5 Hz/s = blocks -33..32): one batch of mixed shapes against the oracle per
  instance, under the same rule.
* The multibeam feed: every (beam, coarse channel) region searched in place
  against the oracle on a copy of the region.
"""

import math

import numpy as np
import pytest

import seticore_amd as sc
from seticore_amd import batch as B

from parity_helpers import device_hits_match_oracle, host_exact_hits_equal

pytestmark = pytest.mark.gpu

FOFF, TSAMP = -2.79e-6, 18.25


def _tones(rng, n, T, max_drift, snr_lo, snr_hi, N, gain=1.0):
    out = []
    for _ in range(n):
        drift = float(rng.uniform(-max_drift, max_drift))
        slope = drift * TSAMP / (FOFF * 1e6)
        snr = float(rng.uniform(snr_lo, snr_hi))
        out.append((float(rng.uniform(0.06 * N, 0.94 * N)), slope, gain * snr / math.sqrt(T)))
    return out


def _file_scale(num_cc, seed):
    import torch
    T, N = 16, 1 << 20
    rng = np.random.default_rng(seed)
    x = torch.empty((num_cc, T, N), dtype=torch.float32, device="cuda")
    for cc in range(num_cc):
        gain = float(rng.uniform(0.5, 2.0)) if cc else 1.7
        sc.gen_waterfall(x[cc], T, seed=seed * 100 + cc, gain=gain, taper_frac=0.05,
                         dc_spike=50.0, tones=_tones(rng, 6, T, 10.0, 8.0, 200.0, N, gain))
    torch.cuda.synchronize()
    return x


def test_file_scale_coarse_channel_host_exact(gpu, oracle):
    # One coarse channel at gain 1.7 with taper, DC spike and tones at
    # +-10 Hz/s, 132 drift blocks: every field of every hit and the column
    # sums, top sums, blocks, offsets, median and deviation equal the oracle's.
    T, N = 16, 1 << 20
    x = _file_scale(1, 41)[0]
    data = x.cpu().numpy()
    dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
    meta = sc.FilterbankMetadata(fch1=8419.0, foff=FOFF, coarse_channel_size=N)
    hits = []
    dd.search(sc.FilterbankBuffer(T, N, data=x), meta, sc.NO_BEAM, 0, 10.0, 0.0, 10.0, hits)
    st = dd.state()
    ref = oracle.search(data, T, FOFF, TSAMP, 10.0, 0.0, 10.0, fch1=8419.0, num_threads=8)
    assert oracle.drift_params(T, FOFF, TSAMP, 10.0)[:2] == (-66, 65)
    np.testing.assert_array_equal(st["column_sums"], ref.column_sums)
    np.testing.assert_array_equal(st["top_sums"], ref.top_sums)
    pos = ref.top_sums > 0
    np.testing.assert_array_equal(st["top_blocks"][pos], ref.top_blocks[pos])
    np.testing.assert_array_equal(st["top_offsets"][pos], ref.top_offsets[pos])
    assert st["median"] == ref.median and st["std"] == ref.std
    host_exact_hits_equal(hits, ref.hits)
    assert len(hits) >= 4


def test_file_scale_batch_against_the_oracle(gpu, oracle):
    # The benchmark's step on four coarse channels: one batched graph launch
    # (device statistics), each coarse channel's hits against the oracle.
    T, N, ncc = 16, 1 << 20, 4
    x = _file_scale(ncc, 42)
    meta = sc.FilterbankMetadata(fch1=8419.0, foff=FOFF, coarse_channel_size=N)
    batch = B.Batch([B.instance(x[cc], T, FOFF, TSAMP, 10.0, 0.0, 10.0, beam=sc.NO_BEAM,
                                coarse_channel=cc, metadata=meta) for cc in range(ncc)])
    batch.run()
    totals = {"same": 0, "only_gpu": 0, "only_oracle": 0}
    for cc in range(ncc):
        data = x[cc].cpu().numpy()
        ref = oracle.search(data, T, FOFF, TSAMP, 10.0, 0.0, 10.0, fch1=8419.0,
                            coarse_channel=cc, num_threads=8)
        got = batch.hits(cc)
        assert all(h.coarse_channel == cc for h in got)
        r = device_hits_match_oracle(got, ref, data, 10.0, T - 1)
        for k in totals:
            totals[k] += r[k]
    print("file-scale batch vs oracle:", totals)
    assert totals["same"] >= 4 * ncc
    assert totals["only_gpu"] + totals["only_oracle"] <= 2 * ncc


# (time steps, channels): a 2^20-wide 32-step instance, 256-step instances at
# 2^17 and 2^18 (one padded from 200 steps), and the other shapes between.
MIXED = [(16, 1 << 20), (32, 1 << 20), (64, 1 << 18), (128, 1 << 17), (200, 1 << 17),
         (256, 1 << 18), (32, 1 << 17), (16, 1 << 18)]


def test_mixed_time_steps_batch_against_the_oracle(gpu, oracle):
    # One batch over mixed (time steps, width) instances, max drift 5 Hz/s
    # (66 drift blocks at every time step count), 20 tones per instance at
    # SNR 10-50: each instance's hits against the oracle.
    import torch
    rng = np.random.default_rng(77)
    xs, inst = [], []
    for i, (Tu, N) in enumerate(MIXED):
        T = sc.round_up_to_power_of_two(Tu)
        x = torch.empty((T, N), dtype=torch.float32, device="cuda")
        sc.gen_waterfall(x, Tu, seed=500 + i, tones=_tones(rng, 20, Tu, 5.0, 10.0, 50.0, N))
        xs.append(x)
        meta = sc.FilterbankMetadata(fch1=8000.0, foff=FOFF, coarse_channel_size=N)
        inst.append(B.instance(x, Tu, FOFF, TSAMP, 5.0, 0.0, 10.0, beam=sc.NO_BEAM,
                               coarse_channel=i, metadata=meta))
    torch.cuda.synchronize()
    batch = B.Batch(inst)
    batch.run()
    same = 0
    for i, (Tu, N) in enumerate(MIXED):
        T = sc.round_up_to_power_of_two(Tu)
        print("instance", i, Tu, N, flush=True)
        data = xs[i].cpu().numpy()
        ref = oracle.search(data, Tu, FOFF, TSAMP, 5.0, 0.0, 10.0, fch1=8000.0,
                            coarse_channel=i, num_threads=8)
        assert oracle.drift_params(Tu, FOFF, TSAMP, 5.0)[:2] == (-33, 32)
        got = batch.hits(i)
        assert all(h.num_timesteps == Tu for h in got)
        r = device_hits_match_oracle(got, ref, data, 10.0, T - 1)
        print("  vs oracle:", r, flush=True)
        # Not vacuous: every instance has hits in common with the oracle (some
        # of its 20 tones start within a drift span of an edge or fall below
        # the threshold, so fewer than 20 are found).
        assert r["same"] >= 4, (Tu, N, r)
        same += r["same"]
    assert same >= 5 * len(MIXED)


def test_multibeam_regions_against_the_oracle(gpu, oracle):
    # beamforming_pipeline.cpp:230-245: each (beam, coarse channel) region of
    # a multibeam buffer, searched in place through its row stride in one
    # batch, against the oracle on a copy of the region
    # (MultibeamBuffer::copyRegionAsync, multibeam_buffer.cu:65-76).
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
    batch = B.Batch(inst)
    batch.run()
    host = multibeam.cpu().numpy()
    same = 0
    for i in range(len(inst)):
        beam, cc = divmod(i, ncc)
        region = np.ascontiguousarray(host[beam, :, cc * ccs:(cc + 1) * ccs])
        ref = oracle.search(region, T, FOFF, TSAMP, 2.0, 0.0, 8.0, fch1=8000.0,
                            coarse_channel_size=ccs, num_coherent_beams=beams, beam=beam,
                            coarse_channel=cc, num_threads=8)
        got = batch.hits(i)
        assert all(h.beam == beam and h.coarse_channel == cc for h in got)
        same += device_hits_match_oracle(got, ref, region, 8.0, T - 1)["same"]
    assert same > 0


# --- data/voyager.hits: the reference's real data through the HIP path ------
def _voyager():
    import os
    from conftest import GOLDEN
    from seticore_amd.hitfile import read_hits
    return read_hits(os.path.join(GOLDEN, "voyager.hits"))


def test_voyager_slices_on_the_gpu(gpu, oracle):
    # The five real hits (index 40, drift 36..39 steps = drift block 2) of the
    # fixture: the HIP tree of each slice equals the oracle's on every valid
    # path, and the path sums at the hits satisfy one affine relation
    # snr = (P - median) / std (dedoppler.cu:308), as the oracle's do.
    import torch
    pairs = []
    for h in _voyager():
        s, f = h["signal"], h["filterbank"]
        T, W = f["num_timesteps"], f["num_channels"]
        data = np.ascontiguousarray(f["data"].reshape(T, W), np.float32)
        b = s["drift_steps"] // (T - 1)
        p = s["drift_steps"] - b * (T - 1)
        assert b == 2
        dd = sc.Dedopplerer(T, W, f["foff"], f["tsamp"])
        got = dd.taylor_tree(sc.FilterbankBuffer(T, W, data=torch.from_numpy(data).cuda()),
                             b).cpu().numpy()
        want = oracle.basic_taylor_tree(data, b)
        for q in range(T):
            last = np.arange(W) + (T - 1) * b + q
            ok = last < W
            np.testing.assert_array_equal(got[q, ok], want[q, ok])
        pairs.append((float(got[p, s["index"]]), float(s["snr"])))
    P = np.array([a for a, _ in pairs])
    snr = np.array([z for _, z in pairs])
    A = np.stack([np.ones_like(snr), snr], axis=1)
    coef, *_ = np.linalg.lstsq(A, P, rcond=None)
    assert np.max(np.abs(A @ coef - P) / P) < 1e-6
    assert 1.7e11 < coef[0] < 1.8e11 and 1.8e10 < coef[1] < 2.0e10


def test_voyager_slice_searches_match_the_oracle(gpu, oracle):
    # Each real slice searched end to end (host-exact statistics, drift blocks
    # -4..3): every hit field bit-equal to the oracle's, and the one hit of
    # each slice is the fixture's own (index 40 and its drift).
    import torch
    for h in _voyager():
        s, f = h["signal"], h["filterbank"]
        T, W = f["num_timesteps"], f["num_channels"]
        data = np.ascontiguousarray(f["data"].reshape(T, W), np.float32)
        dd = sc.Dedopplerer(T, W, f["foff"], f["tsamp"])
        hits = []
        meta = sc.FilterbankMetadata(fch1=f["fch1"], foff=f["foff"], coarse_channel_size=W)
        dd.search(sc.FilterbankBuffer(T, W, data=torch.from_numpy(data).cuda()), meta,
                  sc.NO_BEAM, 0, 0.5, 0.0, 3.0, hits)
        ref = oracle.search(data, T, f["foff"], f["tsamp"], 0.5, 0.0, 3.0, fch1=f["fch1"])
        host_exact_hits_equal(hits, ref.hits)
        np.testing.assert_array_equal(dd.state()["top_sums"], ref.top_sums)
        assert [(x.index, x.drift_steps) for x in hits] == [(40, s["drift_steps"])]
