"""Shared checks of GPU search results against the CPU oracle.

A device-statistics search (median by exact rank selection, deviation in
FP64) may differ from the reference's host statistics (float accumulation,
dedoppler.cu:241-262) only through the SNR threshold: a merge window whose
largest top path sum lies between the two thresholds is a hit in one and not
in the other (dedoppler.cu:273-303). Every other hit must agree on start
channel, drift and path sum bit for bit, and the SNR must agree with the FP64
statistics to the stated tolerance.
"""

import numpy as np

# Relative SNR tolerance of a device-statistics hit against the FP64 oracle.
SNR_TOL = 1e-4
# Relative deviation of the device's FP64 deviation from the oracle's FP64
# value (the device sums in a fixed order; bound taken from the device
# statistics tests).
STD_TOL = 1e-6


def threshold_band(ref, data, snr_threshold):
    """[lo, hi] between the reference's float threshold and the FP64 one
    (widened by the device deviation's tolerance)."""
    med64, _, std64 = _stats64(data)
    thr_ref = np.float32(snr_threshold * ref.std + ref.median)
    cands = [float(thr_ref)]
    for f in (1 - STD_TOL, 1 + STD_TOL):
        for m in (med64, ref.median):
            cands.append(float(np.float32(snr_threshold * std64 * f + m)))
    return min(cands), max(cands), med64, std64


def _stats64(data):
    from oracle import oracle
    return oracle.stats_fp64(data)


def device_hits_match_oracle(hits, ref, data, snr_threshold, drift_timesteps):
    """hits: GPU hits (DedopplerHit) of a device-statistics search of `data`;
    ref: oracle.search result of the same data. Returns a summary dict
    {same, only_gpu, only_oracle, all_diffs_between_thresholds} and raises
    AssertionError on any difference the threshold band does not explain."""
    from oracle import oracle
    lo, hi, med64, std64 = threshold_band(ref, data, snr_threshold)
    got = {h.index: h for h in hits}
    want = {h["index"]: h for h in ref.hits}
    assert len(got) == len(hits) and len(want) == len(ref.hits)
    only_gpu = sorted(set(got) - set(want))
    only_ref = sorted(set(want) - set(got))
    for idx in only_gpu + only_ref:
        s = float(ref.top_sums[idx])
        assert lo <= s <= hi, ("hit at %d (top sum %r) outside the threshold band [%r, %r]"
                               % (idx, s, lo, hi))
    for idx in set(got) & set(want):
        g, w = got[idx], want[idx]
        assert (g.drift_steps, g.power) == (w["drift_steps"], w["power"]), (idx, g, w)
        assert g.drift_rate == w["drift_rate"] and g.frequency == w["frequency"]
        # The reference's own (block, offset) for this channel: a drift of a
        # whole multiple of T_r - 1 is ambiguous between two paths.
        b, p = int(ref.top_blocks[idx]), int(ref.top_offsets[idx])
        assert b * drift_timesteps + p == g.drift_steps
        snr64 = (oracle.path_sum_fp64(data, idx, b, p) - med64) / std64
        assert abs(g.snr - snr64) <= SNR_TOL * max(1.0, abs(snr64)), (idx, g.snr, snr64)
    return {"same": len(set(got) & set(want)), "only_gpu": len(only_gpu),
            "only_oracle": len(only_ref), "all_diffs_between_thresholds": True}


def host_exact_hits_equal(hits, ref_hits):
    """Every field of every hit bit-equal (host-exact statistics)."""
    names = ["frequency", "index", "drift_steps", "drift_rate", "snr", "coarse_channel", "beam",
             "num_timesteps", "power", "incoherent_power"]
    got = [tuple(getattr(h, n) for n in names) for h in hits]
    want = [tuple(h[n] for n in names) for h in ref_hits]
    assert got == want
