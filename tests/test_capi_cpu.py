"""CPU checks of the C ABI library: it loads, exports every declared entry
point, and its host-only parts (argument validation, the .dat and .hits
writers) behave like the reference's. No GPU compute runs here."""

import ctypes
import os
import re

import numpy as np
import pytest

import seticore_amd as sc
from seticore_amd.hitfile import read_hits

from conftest import GOLDEN, ROOT


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "seticore_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = sc.lib()
    names = _declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name


def test_abi_version():
    assert sc.lib().sc_abi_version() == 1


def test_null_arguments_are_rejected():
    L = sc.lib()
    assert L.sc_dd_create(16, 1000, 1.0, 1.0, 0, None) == sc.SC_ERR_INVALID_ARGUMENT
    h = ctypes.c_void_p()
    # num_timesteps must be > 1 (dedoppler.cu:83): rejected before any device call.
    assert L.sc_dd_create(1, 1000, 1.0, 1.0, 0, ctypes.byref(h)) == sc.SC_ERR_INVALID_ARGUMENT
    assert b"num_timesteps" in L.sc_last_error_string()
    assert L.sc_dd_create(16, 0, 1.0, 1.0, 0, ctypes.byref(h)) == sc.SC_ERR_INVALID_ARGUMENT
    assert L.sc_dd_search(None, None, 0, 16, 16, 16, None, 0, 0, 1.0, 0.0, 10.0, None, 0,
                          None, None) == sc.SC_ERR_INVALID_ARGUMENT
    assert L.sc_writer_record_hit(None, None, None) == sc.SC_ERR_IO


@pytest.mark.parametrize("T", [2, 3, 4, 8, 13, 16, 32, 33, 64, 100, 128, 200, 256])
def test_default_kernel_shape_is_compiled(T):
    # sc_dd_create checks, before any device call, that the default one-pass
    # shape for T_r <= 256 is a compiled kernel (otherwise the search would
    # fall back to the global-memory staged path). Without a GPU the create
    # then fails at its first allocation; with one it succeeds.
    L = sc.lib()
    h = ctypes.c_void_p()
    status = L.sc_dd_create(T, 4096, -2.79e-6, 18.25, 0, ctypes.byref(h))
    assert status in (sc.SC_OK, sc.SC_ERR_HIP, sc.SC_ERR_OUT_OF_MEMORY), (
        status, L.sc_last_error_string())
    if status == sc.SC_OK:
        L.sc_dd_destroy(h)


def _meta():
    return sc.FilterbankMetadata(fch1=8419.319795910269, foff=-2.7939677238464355e-06,
                                 coarse_channel_size=1 << 20, source_name="Voyager1",
                                 tstart=57650.78209490741, tsamp=18.253611008,
                                 src_raj=17.17, src_dej=12.18, num_timesteps=16,
                                 num_channels=1 << 26, telescope_id=6)


def _hit(index, drift_steps, snr=12.5, cc=3, beam=-1, power=123.0):
    h = sc.DedopplerHit.__new__(sc.DedopplerHit)
    h.index, h.drift_steps, h.snr, h.coarse_channel, h.beam = index, drift_steps, snr, cc, beam
    h.drift_rate = drift_steps * -0.010204
    h.num_timesteps = 16
    h.power = power
    h.incoherent_power = 0.0
    m = _meta()
    h.frequency = m.fch1 + (cc * m.coarse_channel_size + index) * m.foff
    return h


def test_dat_writer_format(tmp_path):
    # dat_file_writer.cpp:23-47 (header) and 58-73 (rows), restated.
    path = str(tmp_path / "out.dat")
    m = _meta()
    w = sc.DatFileWriter(path, m, 10.0)
    hits = [_hit(1234, 36), _hit(77, -5, snr=245.8251)]
    for h in hits:
        w.recordHit(h)
    w.close()
    lines = open(path).read().split("\n")
    assert lines[0] == "# -------------------------- o --------------------------"
    assert lines[1] == "# File ID: out.dat "
    assert lines[3] == "# Source:Voyager1"
    assert lines[4] == "# MJD: %18.12f\tRA: %.6f\tDEC: %.6f" % (m.tstart, m.src_raj, m.src_dej)
    assert lines[5] == ("# DELTAT: %10.6f\tDELTAF(Hz): %10.6f\tmax_drift_rate: %10.6f\t"
                        "obs_length: %10.6f" % (m.tsamp, m.foff * 1e6, 10.0, 16 * m.tsamp))
    assert lines[7].startswith("# Top_Hit_# \tDrift_Rate \tSNR \t")
    for k, h in enumerate(hits):
        f = m.fch1 + (h.coarse_channel * m.coarse_channel_size + h.index) * m.foff
        want = "%d\t%10.6f\t%10.6f\t%14.6f\t%14.6f\t%d\t%14.6f\t%14.6f\t0.0\t0.000000\t%d\t1\t" % (
            k + 1, h.drift_rate, np.float32(h.snr), f, f, h.index, f, f, h.coarse_channel)
        assert lines[9 + k] == want


def test_hits_writer_round_trip(tmp_path):
    path = str(tmp_path / "out.hits")
    m = _meta()
    m.coarse_channel_size = 4096
    m.source_names = ["beam0", "beam1"]
    m.ras = [1.5, 2.5]
    m.decs = [-3.0, 4.0]
    rng = np.random.default_rng(0)
    data = rng.random((16, 4096), dtype=np.float32)
    hits = [_hit(100, 36, beam=1), _hit(10, -7), _hit(2000, 1500), _hit(4090, 3)]
    w = sc.HitFileWriter(path, m)
    for h in hits:
        w.recordHit(h, data)
    w.close()
    assert not os.path.exists(path + ".tmp")
    dec = read_hits(path)
    assert len(dec) == len(hits)
    for h, d in zip(hits, dec):
        s, f = d["signal"], d["filterbank"]
        assert s["index"] == h.index and s["drift_steps"] == h.drift_steps
        assert s["frequency"] == h.frequency and s["drift_rate"] == h.drift_rate
        assert s["snr"] == np.float32(h.snr) and s["power"] == np.float32(h.power)
        assert s["beam"] == h.beam and s["num_timesteps"] == 16
        # hit_file_writer.cpp:74-97: slice [max(low-40,0), min(high+40,N)-1).
        begin = max(h.lowIndex() - 40, 0)
        end = min(h.highIndex() + 40, 4096) - 1
        assert f["start_channel"] == begin and f["num_channels"] == end - begin
        np.testing.assert_array_equal(f["data"].reshape(16, end - begin), data[:, begin:end])
        assert f["source_name"] == ("beam1" if h.beam == 1 else "Voyager1")
        assert f["ra"] == (2.5 if h.beam == 1 else m.src_raj)
        assert f["fch1"] == m.fch1 + (h.coarse_channel * 4096 + begin) * m.foff
        assert f["telescope_id"] == 6 and f["beam"] == h.beam
    # A 16 x ~1580 slice does not fit the 1024-word first segment: two segments
    # (far pointer + landing pad), the small ones fit in one.
    assert len(dec[2]["layout"]["segments"]) == 2
    assert len(dec[0]["layout"]["segments"]) == 1


def test_hits_layout_matches_voyager_structure(tmp_path):
    """Re-encode the fixture's first hit: same object order and pointer
    structure as the reference's file, with the current schema's larger
    Signal (6 words) and Filterbank (9 words) data sections."""
    ref = open(os.path.join(GOLDEN, "voyager.hits"), "rb").read()
    dec = read_hits(ref)[0]
    s, f = dec["signal"], dec["filterbank"]
    m = sc.FilterbankMetadata(fch1=f["fch1"] - 0 * f["foff"], foff=f["foff"],
                              coarse_channel_size=f["num_channels"] + 1,
                              source_name=f["source_name"], tstart=f["tstart"],
                              tsamp=f["tsamp"], src_raj=f["ra"], src_dej=f["dec"],
                              num_timesteps=f["num_timesteps"], telescope_id=f["telescope_id"])
    h = sc.DedopplerHit.__new__(sc.DedopplerHit)
    h.frequency, h.index, h.drift_steps = s["frequency"], 40, s["drift_steps"]
    h.drift_rate, h.snr, h.coarse_channel = s["drift_rate"], s["snr"], 0
    h.beam, h.num_timesteps, h.power, h.incoherent_power = -1, 16, 0.0, 0.0
    data = f["data"].reshape(16, -1)
    full = np.zeros((16, m.coarse_channel_size), np.float32)
    full[:, :data.shape[1]] = data
    path = str(tmp_path / "v.hits")
    w = sc.HitFileWriter(path, m)
    w.recordHit(h, full)
    w.close()
    ours = open(path, "rb").read()
    d2 = read_hits(ours)[0]
    np.testing.assert_array_equal(d2["filterbank"]["data"], f["data"])
    for k in ("frequency", "index", "drift_steps", "drift_rate", "snr"):
        assert d2["signal"][k] == s[k]
    # Word layout: reference words = 1 root + 2 Hit + 4 Signal + 10 Filterbank
    # + 2 text + data; ours has 2 more Signal words and 1 more Filterbank word.
    assert d2["layout"]["segments"][0] == dec["layout"]["segments"][0] + 3
    q = lambda b, i: int.from_bytes(b[8 + 8 * i:16 + 8 * i], "little")
    assert q(ours, 0) == q(ref, 0)                 # root -> Hit{0 data, 2 ptrs}
    assert q(ours, 1) == (1 << 2) | (6 << 32)      # Signal at word 3, 6 data words
    assert q(ref, 1) == (1 << 2) | (4 << 32)       # (reference schema: 4)
    assert q(ours, 2) == (6 << 2) | ((9 | (2 << 16)) << 32)  # Filterbank at word 9
    assert q(ref, 2) == (4 << 2) | ((8 | (2 << 16)) << 32)   # (reference schema: word 7)


def test_make_hit_recorder_rejects_unknown_suffix(tmp_path):
    with pytest.raises(sc.SeticoreError):
        sc.makeHitRecorder(str(tmp_path / "x.txt"), _meta(), 1.0)


def test_hit_to_string():
    # dedoppler_hit.cpp:25-30 and util.cpp:103-108 (pluralize)
    h = _hit(566231, 0, snr=10.069624)
    h.coarse_channel = 23
    h.drift_rate = -0.0
    assert h.toString() == ("coarse channel = 23, index = 566231, snr = 10.06962, "
                            "drift rate = -0.00000 (0 bins)")
    h.drift_steps = -1
    h.drift_rate = 0.0102
    assert h.toString().endswith("(-1 bin)")


def test_batch_instance_checks_its_tensor():
    # The batch graph reads T_r rows of N samples through the raw pointer, so
    # the instance must be [roundUpToPowerOfTwo(T)][N] of the stated sample
    # type, with unit channel stride, on the GPU (checked before any GPU call).
    import torch
    from seticore_amd import batch as B
    args = (-2.79e-6, 18.25, 2.0, 0.0, 10.0)
    with pytest.raises(ValueError, match="rows"):
        B.instance(torch.zeros((15, 8192)), 15, *args)       # 15 steps need 16 rows
    with pytest.raises(ValueError, match="rows"):
        B.instance(torch.zeros((16, 8192)), 32, *args)
    with pytest.raises(ValueError, match="dtype"):
        B.instance(torch.zeros((16, 8192), dtype=torch.float16), 16, *args)  # default SC_F32
    with pytest.raises(ValueError, match="dtype"):
        B.instance(torch.zeros((16, 8192)), 16, *args, dtype=sc.SC_BF16)
    with pytest.raises(ValueError, match="CUDA"):
        B.instance(torch.zeros((16, 8192)), 16, *args)      # host memory
    with pytest.raises(ValueError, match="2-D"):
        B.instance(torch.zeros((2, 16, 8192)), 16, *args)
