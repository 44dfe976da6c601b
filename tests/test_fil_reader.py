"""Sigproc filterbank input (fil_reader.cpp:31-131), metadata inference
(filterbank_metadata.cpp:53-93) and the loader conditioning
(h5_reader.cpp:195-209), on files written here. Host-only calls: no GPU."""

import struct

import numpy as np
import pytest

import seticore_amd as sc


def _s(text):
    b = text.encode()
    return struct.pack("<I", len(b)) + b


def write_fil(path, data, nchans=None, nsamples=True, telescope_id=None, extra=b"", nbits=32,
              src_raj=123456.7, src_dej=-301530.0):
    T, N = data.shape
    h = _s("HEADER_START")
    if telescope_id is not None:
        h += _s("telescope_id") + struct.pack("<i", telescope_id)
    h += _s("machine_id") + struct.pack("<i", 10)
    h += _s("source_name") + _s("VOYAGER1")
    h += _s("src_raj") + struct.pack("<d", src_raj)
    h += _s("src_dej") + struct.pack("<d", src_dej)
    h += _s("az_start") + struct.pack("<d", 0.0)
    h += _s("tstart") + struct.pack("<d", 57650.78)
    h += _s("tsamp") + struct.pack("<d", 18.253611)
    h += _s("fch1") + struct.pack("<d", 8421.38)
    h += _s("foff") + struct.pack("<d", -2.7939677e-06)
    h += _s("nchans") + struct.pack("<i", N if nchans is None else nchans)
    h += _s("nifs") + struct.pack("<i", 1)
    h += _s("nbits") + struct.pack("<i", nbits)
    if nsamples:
        h += _s("nsamples") + struct.pack("<i", T)
    h += extra
    h += _s("HEADER_END")
    with open(path, "wb") as f:
        f.write(h)
        f.write(np.ascontiguousarray(data, np.float32).tobytes())


def test_header_and_angles(tmp_path):
    p = tmp_path / "a.fil"
    data = np.zeros((16, 1 << 20), dtype=np.float32)
    write_fil(p, data)
    r = sc.FilReader(str(p))
    assert r.coarse_channel_size == 1 << 20 and r.num_coarse_channels == 1
    r.set_coarse_channel_size(4096)
    assert r.source_name == "VOYAGER1"
    assert r.num_timesteps == 16 and r.num_channels == 1 << 20
    assert r.coarse_channel_size == 4096 and r.num_coarse_channels == 256
    assert r.fch1 == 8421.38 and r.foff == -2.7939677e-06 and r.tsamp == 18.253611
    # fil_reader.cpp:16-24: 12h34m56.7s; -30d15m30s
    assert abs(r.src_raj - (12 + 34 / 60 + 56.7 / 3600)) < 1e-12
    assert abs(r.src_dej + (30 + 15 / 60 + 30 / 3600)) < 1e-12


def test_green_bank_inference_and_conditioning(tmp_path):
    # 16 x 2^20 with no nfpc: Green Bank (coarse channel 2^20, DC spike).
    T, N = 16, 1 << 20
    rng = np.random.default_rng(1)
    data = rng.exponential(1.0, size=(T, N)).astype(np.float32)
    data[:, N // 2] = 50.0
    p = tmp_path / "gbt.fil"
    write_fil(p, data, nsamples=False)  # time steps inferred from the size
    r = sc.FilReader(str(p))
    assert r.num_timesteps == 16 and r.coarse_channel_size == N and r.has_dc_spike
    got = r.read_coarse_channel(0, 16)
    want = data.copy()
    want[:, N // 2] = (want[:, N // 2 - 1] + want[:, N // 2 + 1]) / np.float32(2.0)
    np.testing.assert_array_equal(got, want)


def test_padding_rows_are_zero(tmp_path):
    T, N = 15, 1 << 20  # 15 rows also looks like Green Bank
    data = np.ones((T, N), np.float32)
    p = tmp_path / "pad.fil"
    write_fil(p, data)
    r = sc.FilReader(str(p))
    got = r.read_coarse_channel(0)
    assert got.shape == (16, N)
    assert np.all(got[15] == 0) and np.all(got[:15] == 1)


def test_uninferable_shape(tmp_path):
    # Neither a Green Bank nor an ATA shape, and .fil headers carry no nfpc:
    # the reference exits (filterbank_metadata.cpp:84-89); here it is an error.
    p = tmp_path / "x.fil"
    write_fil(p, np.zeros((4, 3 * 4096), np.float32), telescope_id=6)
    with pytest.raises(sc.SeticoreError) as e:
        sc.FilReader(str(p))
    assert "nfpc" in str(e.value)


@pytest.mark.parametrize("case", ["bad_start", "attr", "nbits", "size", "nsamples"])
def test_errors_are_status_codes(tmp_path, case):
    p = tmp_path / "bad.fil"
    data = np.zeros((16, 1 << 20), np.float32)
    if case == "bad_start":
        p.write_bytes(_s("HEADER_STOP") + b"\0" * 64)
    elif case == "attr":
        write_fil(p, data, extra=_s("mystery") + struct.pack("<i", 1))
    elif case == "nbits":
        write_fil(p, data, nbits=8)
    elif case == "size":
        write_fil(p, data)
        with open(p, "ab") as f:
            f.write(b"\0\0")
    elif case == "nsamples":
        write_fil(p, data[:8])
        raw = p.read_bytes()
        i = raw.index(b"nsamples") + len(b"nsamples")
        p.write_bytes(raw[:i] + struct.pack("<i", 9) + raw[i + 4:])
    with pytest.raises(sc.SeticoreError):
        sc.FilReader(str(p))


def test_small_coarse_channels_with_dc_spike_are_rejected(tmp_path):
    # The DC fix reads the centre channel's two neighbours (h5_reader.cpp:202-209):
    # a Green Bank file (DC spike inferred) cannot be cut into coarse channels
    # of one or two channels; three channels work and fix channel 1.
    T, N = 16, 1 << 20
    data = np.arange(T * N, dtype=np.float32).reshape(T, N) % 7
    p = tmp_path / "gbt.fil"
    write_fil(p, data)
    r = sc.FilReader(str(p))
    assert r.has_dc_spike
    for bad in (1, 2):
        with pytest.raises(sc.SeticoreError):
            r.set_coarse_channel_size(bad)
    assert r.coarse_channel_size == N  # unchanged after a rejected size
    r.set_coarse_channel_size(4)
    got = r.read_coarse_channel(5, 16)
    want = data[:, 20:24].copy()
    want[:, 2] = (want[:, 1] + want[:, 3]) / np.float32(2.0)
    np.testing.assert_array_equal(got, want)
