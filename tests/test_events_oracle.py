"""Pins the cadence-event oracle (oracle/events_oracle.py, the restatement of
find_events.cpp:159-217) on hand-built hit lists whose outcomes follow from the
reference's rules by hand, one rule per case, plus a planted mistake that one
case must catch. CPU only."""

import pytest

from oracle import events_oracle as E

T = 16


def hit(index, drift):
    return {"index": index, "drift_steps": drift, "num_timesteps": T}


def run(lists, tstarts=None, tsamp=18.25, **kw):
    tstarts = tstarts or [0.0] * len(lists)
    return E.match_events(lists, tstarts, tsamp, **kw)


def six(first, on2, on4, off1=(), off3=(), off5=()):
    return [list(first), list(off1), list(on2), list(off3), list(on4), list(off5)]


def test_clean_event():
    # On scans: 1000..1005, 1003..1008, 998..1002 -> low 998, high 1008;
    # range [948, 958]: each "on" file's first hit at or after 948 is its only
    # one, the "off" files have none.
    lists = six([hit(1000, 5)], [hit(1003, 5)], [hit(998, 4)])
    events, outcomes = run(lists)
    assert outcomes == [(1000, "event")]
    assert [h["index"] if h else None for h in events[0]] == [1000, None, 1003, None, 998, None]


def test_missing_on_hit_rejects():
    lists = six([hit(1000, 5)], [], [hit(998, 4)])
    assert run(lists)[1] == [(1000, "on")]


def test_two_hits_in_an_on_window_reject():
    lists = six([hit(1000, 5)], [hit(1002, 5), hit(1006, 1)], [hit(998, 4)])
    assert run(lists)[1] == [(1000, "on")]


def test_off_hit_in_range_rejects():
    # 960 is the first "off" hit at or after low - 50 = 948: count 1, not 0.
    lists = six([hit(1000, 5)], [hit(1003, 5)], [hit(998, 4)], off3=[hit(960, 2)])
    assert run(lists)[1] == [(1000, "range")]


def test_find_hit_counts_the_first_hit_past_the_range():
    # find_events.cpp:27-46: lower_bound(990) is 1030, beyond 1010, yet it is
    # counted once (no second hit follows), so the "on" check passes; a
    # strict in-range lookup would reject this candidate.
    lists = six([hit(1000, 5)], [hit(1030, 2)], [hit(998, 4)])
    events, outcomes = run(lists)
    assert outcomes == [(1000, "event")]
    assert events[0][2]["index"] == 1030


def test_range_ends_at_high_minus_50():
    # A second hit at 1030 in on scan 2 lies past high - 50 = 958 (so the range
    # check still counts one hit) but before high + 50: the reference's range
    # (find_events.cpp:206-207) accepts the event; a symmetric range does not.
    lists = six([hit(1000, 5)], [hit(1003, 5), hit(1030, 3)], [hit(998, 4)])
    assert run(lists)[1] == [(1000, "event")]
    # Planted mistake: [low - 50, high + 50].
    assert run(lists, range_end_shift=+50)[1] == [(1000, "range")]


def test_later_hit_at_an_index_replaces_the_earlier():
    # find_events.cpp:159-165: hitmaps[i][hit.index] = hit, so the second hit at
    # 1000 (drift -3) is the candidate, and its low index 997 sets the range.
    lists = six([hit(1000, 5), hit(1000, -3)], [hit(1003, 5)], [hit(998, 4)])
    events, outcomes = run(lists)
    assert outcomes == [(1000, "event")]
    assert events[0][0]["drift_steps"] == -3


def test_expected_index_follows_the_scan_offset():
    # Scans 600 s and 1200 s after the first at tsamp 18.25 s: round(32.88) = 33
    # and round(65.75) = 66 time steps; 15 drift steps over 15 steps = 1 per
    # step, so on scans 2 and 4 are expected at +33 and +66.
    ts = [0.0, 300.0, 600.0, 900.0, 1200.0, 1500.0]
    assert E.expected_index(hit(1000, 15), 33) == 1033
    assert E.expected_index(hit(1000, -15), 66) == 934
    lists = six([hit(1000, 15)], [hit(1033, 15)], [hit(1066, 15)])
    assert run(lists, ts)[1] == [(1000, "event")]
    # 1022 is below 1033 - 10 and nothing follows it: no hit at or after 1023.
    # (Without the offset, the window [990, 1010] would find 1022 and pass.)
    lists = six([hit(1000, 15)], [hit(1022, 15)], [hit(1066, 15)])
    assert run(lists, ts)[1] == [(1000, "on")]


def test_expected_index_rounds_half_away_from_zero():
    # C round(): 2.5 -> 3 and -2.5 -> -3 (Python's round gives 2 and -2).
    h = {"index": 100, "drift_steps": 5, "num_timesteps": 3}   # 2.5 per step
    assert E.expected_index(h, 1) == 103
    h["drift_steps"] = -5
    assert E.expected_index(h, 1) == 97


def test_off_hit_anywhere_above_rejects():
    # The same lookup in an "off" file: its first hit at or after low - 50
    # counts even thousands of channels higher, so it rejects the candidate.
    lists = six([hit(1000, 5)], [hit(1003, 5)], [hit(998, 4)], off5=[hit(8000, 2)])
    assert run(lists)[1] == [(1000, "range")]
    # An "off" hit below low - 50 does not.
    lists = six([hit(1000, 5)], [hit(1003, 5)], [hit(998, 4)], off5=[hit(900, 2)])
    assert run(lists)[1] == [(1000, "event")]


def test_several_candidates_in_one_channel():
    # 1000: an event. 5000: two hits (5002, 5004) in on scan 2's window
    # [4990, 5010]. 9000: low 8998, high 9001, range [8948, 8951]; on scan 4's
    # first hit at or after 8948 is 8949 and 8950 follows within the range.
    lists = six([hit(1000, 5), hit(5000, 4), hit(9000, -2)],
                [hit(1003, 5), hit(5002, 4), hit(5004, 1), hit(9001, -2)],
                [hit(998, 4), hit(8949, 1), hit(8950, 1), hit(9000, -1)])
    events, outcomes = run(lists)
    assert outcomes == [(1000, "event"), (5000, "on"), (9000, "range")]
    assert len(events) == 1


def test_condition_matches_the_loader():
    import numpy as np
    rows = np.arange(3 * 8, dtype=np.float32).reshape(3, 8)
    out = E.condition(rows, 4, True)
    assert out.shape == (4, 8) and np.all(out[3] == 0)
    np.testing.assert_array_equal(out[:3, 4], (rows[:, 3] + rows[:, 5]) / np.float32(2))
    np.testing.assert_array_equal(np.delete(out[:3], 4, axis=1), np.delete(rows, 4, axis=1))
