"""Cadence events: the .events writer (event_file_writer.cpp:35-69, hit.capnp
Event) decoded by the pure-Python reader (host only), and the cadence search
(find_events.cpp:66-227) on synthetic ABACAD .fil files (GPU)."""

import numpy as np
import pytest

import seticore_amd as sc
from seticore_amd import hitfile

from test_fil_reader import write_fil


def _hit(index, drift, snr, cc=0, T=16):
    h = sc.DedopplerHit()
    h.index, h.drift_steps, h.snr, h.coarse_channel, h.beam = index, drift, snr, cc, -1
    h.num_timesteps, h.power, h.frequency, h.drift_rate = T, 10.0 * snr, 8000.0 + index, 0.1
    return h


def test_event_writer_round_trip(tmp_path):
    T, N = 16, 4096
    metas, inputs = [], []
    for i in range(4):
        metas.append(sc.FilterbankMetadata(fch1=8000.0 + i, foff=-1e-6, coarse_channel_size=N,
                                           source_name="ON" if i % 2 == 0 else "OFF%d" % i,
                                           tstart=57000.0 + i, tsamp=18.0, num_timesteps=T,
                                           telescope_id=6))
        inputs.append(np.arange(T * N, dtype=np.float32).reshape(T, N) + 1000 * i)
    hits = [_hit(1000, 12, 30.0), None, _hit(1020, -8, 25.0), None]
    path = tmp_path / "x.events"
    with sc.EventFileWriter(str(path), metas) as w:
        w.write(hits, inputs)
        w.write([_hit(3000, 3, 12.0), None, None, None], inputs)
    events = hitfile.read_events(str(path))
    assert len(events) == 2
    ev = events[0]
    assert len(ev) == 4
    assert ev[0]["signal"]["index"] == 1000 and ev[0]["signal"]["drift_steps"] == 12
    assert ev[2]["signal"]["index"] == 1020 and ev[1]["signal"] is None and ev[3]["signal"] is None
    # One channel range over the event: lowIndex/highIndex of the present hits,
    # +-40 columns, one short on the right (hit_file_writer.cpp:80-81).
    low, high = min(1000, 1012, 1020, 1012), max(1012, 1020)
    begin, end = low - 40, high + 40 - 1
    for i, h in enumerate(ev):
        fb = h["filterbank"]
        assert fb["start_channel"] == begin and fb["num_channels"] == end - begin
        assert fb["source_name"] == metas[i].source_name
        assert fb["tstart"] == metas[i].tstart and fb["fch1"] == pytest.approx(
            metas[i].fch1 + begin * metas[i].foff)
        np.testing.assert_array_equal(fb["data"].reshape(T, -1), inputs[i][:, begin:end])


def _cadence(tmp_path, tag, signal_in_off=False):
    """Six 16 x 2^20 scans, ABACAD, 300 s apart: a signal drifting 0.5 channel
    per step in the "on" scans (and optionally in an "off" scan too)."""
    T, N = 16, 1 << 20
    tsamp = 18.253611
    files = []
    for i in range(6):
        rng = np.random.default_rng(100 + i)
        data = rng.exponential(1.0, size=(T, N)).astype(np.float32)
        on = i % 2 == 0
        if on or (signal_in_off and i == 3):
            for t in range(T):
                data[t, int(round(1040000 + 0.5 * (i * 300.0 / tsamp + t)))] += 12.0
        p = tmp_path / ("%s%d.fil" % (tag, i))
        write_fil(p, data)
        raw = p.read_bytes()
        k = raw.index(b"tstart") + len(b"tstart")  # MJD of the scan
        raw = raw[:k] + np.float64(57650.0 + i * 300.0 / 86400.0).tobytes() + raw[k + 8:]
        k = raw.index(b"VOYAGER1")
        raw = raw[:k] + (b"TARGET01" if on else b"OFFSRC%02d" % i) + raw[k + 8:]
        p.write_bytes(raw)
        files.append(str(p))
    return files


@pytest.mark.gpu
def test_find_events_on_a_synthetic_cadence(gpu, tmp_path):
    # find_events.cpp:66-227 with its matching rules: the scan offsets are
    # computed from tstart (MJD) as if it were seconds, so the expected index
    # stays the first hit's, and the range lookup accepts the first hit at or
    # after the range start; a signal near the top of the coarse channel keeps
    # noise hits of the "off" scans out of its range.
    files = _cadence(tmp_path, "a")
    out = tmp_path / "a.events"
    n = sc.findEvents(files, str(out), 1.0, 12.0, 12.0)
    events = hitfile.read_events(str(out))
    assert n == len(events) == 1
    ev = events[0]
    assert [h["signal"] is not None for h in ev] == [True, False, True, False, True, False]
    assert abs(ev[0]["signal"]["index"] - 1040000) <= 1
    assert ev[2]["signal"]["index"] > ev[0]["signal"]["index"]  # the signal kept drifting
    assert ev[0]["filterbank"]["source_name"] == "TARGET01"
    assert ev[1]["filterbank"]["source_name"] == "OFFSRC01"
    # The same signal in an "off" scan: no event.
    files = _cadence(tmp_path, "b", signal_in_off=True)
    assert sc.findEvents(files, str(tmp_path / "b.events"), 1.0, 12.0, 12.0) == 0


def _cadence_several(tmp_path):
    """Six 16 x 2^20 ABACAD scans, 300 s apart, with four drifting signals:
    one also in "off" scan 3, two only in the "on" scans, one only in the
    first scan. Returns (file names, [(tstart, data)] for the oracle)."""
    T, N = 16, 1 << 20
    tsamp = 18.253611
    signals = [  # (start channel, channels per step, amplitude, scans)
        (1000000, 0.5, 12.0, (0, 2, 3, 4)),
        (1020000, -0.4, 12.0, (0, 2, 4)),
        (1040000, 0.5, 12.0, (0, 2, 4)),
        (1046000, 0.5, 12.0, (0,)),
    ]
    files, scans = [], []
    for i in range(6):
        rng = np.random.default_rng(300 + i)
        data = rng.exponential(1.0, size=(T, N)).astype(np.float32)
        for c0, slope, amp, which in signals:
            if i in which:
                for t in range(T):
                    data[t, int(round(c0 + slope * (i * 300.0 / tsamp + t)))] += amp
        on = i % 2 == 0
        tstart = 57650.0 + i * 300.0 / 86400.0
        p = tmp_path / ("c%d.fil" % i)
        write_fil(p, data)
        raw = p.read_bytes()
        k = raw.index(b"tstart") + len(b"tstart")
        raw = raw[:k] + np.float64(tstart).tobytes() + raw[k + 8:]
        k = raw.index(b"VOYAGER1")
        raw = raw[:k] + (b"TARGET01" if on else b"OFFSRC%02d" % i) + raw[k + 8:]
        p.write_bytes(raw)
        files.append(str(p))
        scans.append((tstart, data))
    return files, scans


@pytest.mark.gpu
def test_find_events_matches_the_oracle(gpu, oracle, tmp_path):
    # sc_find_events against the restatement of find_events.cpp:124-217
    # (oracle/events_oracle.py) fed by the oracle's search: the same events,
    # each with the same hit (index, drift, SNR bit for bit: both use the
    # reference's host statistics) or none per scan. Of the four candidates
    # two are rejected, one by an "off" scan's hit in its range, one because
    # the later "on" scans have no hit at or after its window.
    from oracle import events_oracle
    files, scans = _cadence_several(tmp_path)
    out = tmp_path / "c.events"
    n = sc.findEvents(files, str(out), 1.0, 12.0, 12.0)
    got = hitfile.read_events(str(out))
    want, outcomes = events_oracle.find_events(scans, 1 << 20, True, -2.7939677e-06, 18.253611,
                                               1.0, 12.0, 12.0)
    assert [k for k, _ in outcomes[0]] == [1000000, 1020000, 1040000, 1046000]
    assert [v for _, v in outcomes[0]] == ["range", "event", "event", "on"]
    assert n == len(got) == len(want[0]) == 2
    for ev_got, ev_want in zip(got, want[0]):
        assert len(ev_got) == len(ev_want) == 6
        for g, w in zip(ev_got, ev_want):
            if w is None:
                assert g["signal"] is None
            else:
                s = g["signal"]
                assert (s["index"], s["drift_steps"], s["snr"], s["power"]) == \
                    (w["index"], w["drift_steps"], w["snr"], w["power"])
