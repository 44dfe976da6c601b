"""CPU oracle of the cadence event search (test infrastructure only).

Imported by tests/ as the checker of sc_find_events; the product library never
loads it. A plain-Python restatement of lacker/seticore's findEvents
(find_events.cpp:66-227), fed by the oracle's dedoppler search
(oracle.search, the restatement of dedoppler.cu:186-320) on coarse channels
conditioned as the reference's loader conditions them (h5_reader.cpp:195-209).

Functions, each with the reference lines it follows:

* `find_hit`          find_events.cpp:27-46 (range lookup, with its quirk: the
                      first hit at or after `low` counts once even when it lies
                      beyond `high`).
* `expected_index`    dedoppler_hit.cpp:40-44.
* `low_index`, `high_index`  dedoppler_hit.cpp:32-38.
* `match_events`      find_events.cpp:159-217 on per-file hit lists: hit maps
                      keyed by start index (a later hit at an index replaces an
                      earlier one), the "on" check (+-10 around the expected
                      index, exactly one hit each) and the range check
                      ([low - 50, high - 50]: one hit in each "on" file, none in
                      each "off" file).
* `condition`         h5_reader.cpp:195-209 (zero rows up to T_r, DC fix).
* `find_events`       find_events.cpp:124-152 + match_events, per coarse channel.

How it is pinned (tests/test_events_oracle.py): hand-derived outcomes of
`match_events` on hit lists built to exercise each rule (a clean event, an
"on" scan missing its hit, two hits in an "on" window, a hit in an "off"
scan's range, the findHit at-or-after quirk, the high - 50 range end, the
replacing hit map), and a planted mistake (a symmetric [low - 50, high + 50]
range) that one of those cases must catch.
"""

import bisect
import math

import numpy as np

from . import oracle


def low_index(h):
    return min(h["index"], h["index"] + h["drift_steps"])


def high_index(h):
    return max(h["index"], h["index"] + h["drift_steps"])


def expected_index(h, timesteps):
    """dedoppler_hit.cpp:40-44: index + round(timesteps * drift per step), the
    drift per step over the unrounded time steps minus one."""
    per_step = float(h["drift_steps"]) / float(h["num_timesteps"] - 1)
    x = timesteps * per_step
    # C's round(): halves away from zero.
    r = math.floor(abs(x) + 0.5) * (1 if x >= 0 else -1)
    return h["index"] + int(r)


def find_hit(keys, hitmap, low, high):
    """find_events.cpp:27-46. keys: sorted start indices of hitmap. Returns
    (count, hit): (0, None) if no hit starts at or after `low`; (1, that hit)
    if the next one does not also start at or before `high`; else (2, None)."""
    i = bisect.bisect_left(keys, low)
    if i == len(keys):
        return 0, None
    if i + 1 == len(keys) or keys[i + 1] > high:
        return 1, hitmap[keys[i]]
    return 2, None


def match_events(hit_lists, tstarts, tsamp, range_end_shift=-50):
    """find_events.cpp:159-217 for one coarse channel. hit_lists[i]: file i's
    hits (dicts with index, drift_steps, num_timesteps), in search order.
    Returns (events, outcomes): each event is a list with one hit (or None)
    per file; outcomes lists (first hit index, "event" | "on" | "range") for
    every candidate from the first file. range_end_shift is the reference's
    -50 (a parameter only so a test can plant a mistake)."""
    maps = []
    for hits in hit_lists:
        m = {}
        for h in hits:
            m[h["index"]] = h
        maps.append(m)
    keys = [sorted(m) for m in maps]
    events, outcomes = [], []
    t0 = tstarts[0]
    for k in keys[0]:
        first = maps[0][k]
        ok = True
        low, high = 2 ** 31 - 1, -1
        for i in range(0, len(maps), 2):
            delta = tstarts[i] - t0
            x = delta / tsamp
            timesteps = int(math.floor(abs(x) + 0.5) * (1 if x >= 0 else -1))
            e = expected_index(first, timesteps)
            count, hit = find_hit(keys[i], maps[i], e - 10, e + 10)
            if count != 1:
                ok = False
                break
            low = min(low, low_index(hit))
            high = max(high, high_index(hit))
        if not ok:
            outcomes.append((k, "on"))
            continue
        candidate = []
        for i in range(len(maps)):
            ideal = 1 if i % 2 == 0 else 0
            count, hit = find_hit(keys[i], maps[i], low - 50, high + range_end_shift)
            if count != ideal:
                ok = False
                break
            candidate.append(hit)
        if not ok:
            outcomes.append((k, "range"))
            continue
        outcomes.append((k, "event"))
        events.append(candidate)
    return events, outcomes


def condition(rows, rounded_timesteps, has_dc_spike):
    """h5_reader.cpp:195-209 on a host [T][N] coarse channel: rows T..T_r-1
    zero; with a DC spike the centre channel of each loaded row becomes the
    mean of its two neighbours (float add, then halved)."""
    T, N = rows.shape
    out = np.zeros((rounded_timesteps, N), np.float32)
    out[:T] = rows
    if has_dc_spike:
        mid = N // 2
        out[:T, mid] = (out[:T, mid - 1] + out[:T, mid + 1]) / np.float32(2.0)
    return out


def find_events(scans, coarse_channel_size, has_dc_spike, foff, tsamp, max_drift, snr_on,
                snr_off, num_threads=8):
    """find_events.cpp:124-217 over in-memory scans. scans[i] = (tstart,
    [T][num_channels] float32 array). Returns (events per coarse channel,
    outcomes per coarse channel)."""
    T = scans[0][1].shape[0]
    T_r = oracle.round_up_pow2(T)
    nch = scans[0][1].shape[1]
    all_events, all_outcomes = [], []
    for cc in range(nch // coarse_channel_size):
        lo, hi = cc * coarse_channel_size, (cc + 1) * coarse_channel_size
        hit_lists = [[] for _ in scans]

        def search(i, snr, drop_zero):
            data = condition(scans[i][1][:, lo:hi], T_r, has_dc_spike)
            res = oracle.search(data, T, foff, tsamp, max_drift, 0.0, snr, coarse_channel=cc,
                                num_threads=num_threads)
            hits = res.hits
            if drop_zero:  # find_events.cpp:48-54
                hits = [h for h in hits if h["drift_steps"] != 0]
            return hits

        hit_lists[0] = search(0, snr_on, True)
        if not hit_lists[0]:
            all_events.append([])
            all_outcomes.append([])
            continue
        for i in range(1, len(scans)):
            on = i % 2 == 0
            hit_lists[i] = search(i, snr_on if on else snr_off, on)
        events, outcomes = match_events(hit_lists, [s[0] for s in scans], tsamp)
        all_events.append(events)
        all_outcomes.append(outcomes)
    return all_events, all_outcomes
