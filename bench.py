#!/usr/bin/env python3
"""Benchmark: plain Taylor-tree dedoppler search on one MI355X.

Workload (file scale): 64 coarse channels x 2^20 fine channels x 16 time
steps, float32 (4 GiB, resident in HBM before timing), foff -2.79e-6 MHz,
tsamp 18.25 s. Exponential noise with a per-coarse-channel gain U[0.5, 2], a
cosine taper over the outer 5 % of each coarse channel, a DC spike (50x mean)
at each centre bin, and 200 linear tones at drift +-10 Hz/s with SNR 8-200.
Max drift 10 Hz/s (drift blocks -66..65), SNR threshold 10.

One step = one search of every coarse channel (statistics and merge on the GPU,
hits back on the host), run as one batched graph launch (sc_batch) whose hits
must equal per-call Dedopplerer::search results, which are timed as well
(--loop times only those). The value is path sums per second: one path sum =
one (drift block, path offset, start channel); 132 * 16 * 2^20 per coarse
channel.

Prints one JSON line.
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Dedoppler Gpath-sums/s & % HBM roofline; learned vs plain at ≥99.5% hit recall"
HBM_PEAK_GBS = 8000.0
# Vector ALU issue of one MI355X without packed FP32: 256 CUs x 4 SIMDs x 16
# lanes per clock at 2.4 GHz, in lane-operations per second (the tree is adds
# and maxima on the VALU; packed FP32 forms are not used).
VALU_PEAK_OPS = 256 * 4 * 16 * 2.4e9
# The box's CPU share for one GPU: the CPU baseline never runs more threads.
HOST_THREAD_SHARE = 16

# Workload constants.
NUM_CC = 64
N = 1 << 20
T = 16
FOFF = -2.79e-6
TSAMP = 18.25
MAX_DRIFT = 10.0
SNR_THRESHOLD = 10.0
NUM_TONES = 200
SEED = 20261202


def make_tones(rng, num_cc):
    """200 tones spread over the coarse channels: (cc, start, slope, amplitude)."""
    tones = []
    for _ in range(NUM_TONES):
        cc = int(rng.integers(0, num_cc))
        drift = float(rng.uniform(-10.0, 10.0))
        slope = drift * TSAMP / (FOFF * 1e6)  # channels per time step
        snr = float(rng.uniform(8.0, 200.0))
        amp = snr / math.sqrt(T)
        start = float(rng.uniform(0.06 * N, 0.94 * N))
        tones.append((cc, start, slope, amp))
    return tones


def generate(num_cc, torch, sc):
    rng = np.random.default_rng(SEED)
    gains = rng.uniform(0.5, 2.0, size=num_cc)
    tones = make_tones(rng, num_cc)
    data = torch.empty((num_cc, T, N), dtype=torch.float32, device="cuda")
    for cc in range(num_cc):
        mine = [(s, sl, a * gains[cc]) for (c, s, sl, a) in tones if c == cc]
        sc.gen_waterfall(data[cc], T, seed=SEED * 1000 + cc, gain=float(gains[cc]),
                         taper_frac=0.05, dc_spike=50.0, tones=mine)
    torch.cuda.synchronize()
    return data


def path_sums_per_cc():
    import seticore_amd as sc
    bmin, bmax, _, _ = sc.drift_parameters(T, FOFF, TSAMP, MAX_DRIFT)
    return (bmax - bmin + 1) * T * N, bmax - bmin + 1


def cpu_baseline(data_cc, threads):
    """The oracle (reference-order CPU restatement) on a bounded sample."""
    from oracle import oracle
    t0 = time.perf_counter()
    res = oracle.search(data_cc, T, FOFF, TSAMP, MAX_DRIFT, 0.0, SNR_THRESHOLD,
                        num_threads=threads)
    dt = time.perf_counter() - t0
    return dt, res


# Committed rocprofv3 counter summaries (profiles/): HBM bytes per launch of
# the dominant kernel, by (time steps, max drift) of the measured workload.
TRAFFIC_FILES = {("fused_tree_top_set_kernel", 16, 10.0): "fused_tree_set_cfg2_traffic.json",
                 ("fused_tree_top_kernel", 16, 10.0): "fused_tree_cfg2_traffic.json",
                 ("staged_tree_kernel", 256, 5.0): "staged_t256_traffic.json"}


def load_traffic(kernel, T_r, max_drift):
    """HBM bytes per launch of `kernel` on this workload (committed counter
    summaries), or None."""
    name = TRAFFIC_FILES.get((kernel, T_r, max_drift))
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def alg_valu_ops_per_path_sum(T_r):
    """Algorithmic vector operations per path sum: log2(T_r) adds (one per
    tree level, taylor.h:53-54) and half a max (a three-input max folds two
    path sums into the running best, dedoppler.cu:47-52)."""
    return math.log2(T_r) + 0.5


def host_description():
    """nproc, the CPUs this process may run on, and lscpu's model name."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"nproc": os.cpu_count(), "affinity_cpus": affinity, "lscpu_model": model}


def hits_vs_oracle(gpu_hits, ref, data_cc, snr_threshold):
    """Hit-set difference of a device-statistics search against the oracle's
    (reference-order) search of the same coarse channel: hits keyed by start
    channel; a difference is allowed only for a window whose largest top sum
    lies between the reference's float threshold and the FP64 one
    (dedoppler.cu:273-303). Common hits must agree on drift and path sum."""
    from oracle import oracle
    med64, _, std64 = oracle.stats_fp64(data_cc)
    cands = [float(np.float32(snr_threshold * ref.std + ref.median))]
    for f in (1 - 1e-6, 1 + 1e-6):
        for m in (med64, ref.median):
            cands.append(float(np.float32(snr_threshold * std64 * f + m)))
    lo, hi = min(cands), max(cands)
    got = {h.index: h for h in gpu_hits}
    want = {h["index"]: h for h in ref.hits}
    diffs = sorted(set(got) ^ set(want))
    between = all(lo <= float(ref.top_sums[i]) <= hi for i in diffs)
    common = set(got) & set(want)
    agree = all((got[i].drift_steps, got[i].power) == (want[i]["drift_steps"], want[i]["power"])
                for i in common)
    return {"same": len(common), "only_gpu": len(set(got) - set(want)),
            "only_oracle": len(set(want) - set(got)), "all_diffs_between_thresholds": between,
            "common_hits_identical": agree}


def main():
    global T, MAX_DRIFT
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--coarse-channels", type=int, default=NUM_CC)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-cc", type=int, default=2)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--group", type=int, default=0)
    ap.add_argument("--prefetch", type=int, default=0)
    ap.add_argument("--loop", action="store_true",
                    help="search coarse channel by coarse channel (Dedopplerer::search per "
                         "call) instead of one batched graph launch per step")
    ap.add_argument("--graphs", type=int, default=1,
                    help="replay each coarse channel's search from a cached hipGraph (1) or "
                         "launch it directly (0)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run device-to-device copies of this size (for scaling "
                         "profiler byte counters); not timed")
    ap.add_argument("--time-steps", type=int, default=T,
                    help="time steps per coarse channel for the plain workload (default 16)")
    ap.add_argument("--max-drift", type=float, default=MAX_DRIFT,
                    help="max drift in Hz/s for the plain workload (default 10)")
    ap.add_argument("--workload", default="plain",
                    choices=["plain", "half", "batched", "learned"],
                    help="plain: the file-scale search (default); half: the same with "
                         "FP16/BF16 storage; batched: uneven beams x coarse channels in one "
                         "graph launch; learned: proposal-gated search")
    ap.add_argument("--snr-threshold", type=float, default=SNR_THRESHOLD,
                    help="learned workload: SNR threshold of the plain search that labels the "
                         "blocks and of the searches timed (default 10)")
    ap.add_argument("--storage", default="f32", choices=["f32", "bf16"],
                    help="batched workload: FP32 input in waves, or BF16 (normalised) input "
                         "searched in one launch")
    ap.add_argument("--wave-gib", type=int, default=0,
                    help="batched workload: GiB of input per wave (default 96 for FP32, "
                         "everything for BF16)")
    args = ap.parse_args()
    T, MAX_DRIFT = args.time_steps, args.max_drift
    if T not in (2, 4, 8, 16, 32, 64, 128, 256, 512, 1024):
        print("--time-steps must be a power of two from 2 to 1024", file=sys.stderr)
        sys.exit(2)

    if args.gpus != 1:
        print("bench.py measures one GPU: the search shards by coarse channel with no "
              "collective, so --gpus N > 1 is not run here", file=sys.stderr)
        sys.exit(2)

    import torch
    import seticore_amd as sc

    if args.workload != "plain":
        {"half": run_half, "batched": run_batched, "learned": run_learned}[args.workload](
            args, torch, sc)
        return

    num_cc = args.coarse_channels
    data = generate(num_cc, torch, sc)
    per_cc, nblocks = path_sums_per_cc()

    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib << 18  # float32 elements
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    dd = sc.Dedopplerer(T, N, FOFF, TSAMP, False)
    dd.set_stats_mode(sc.SC_STATS_DEVICE)
    dd.set_graphs(bool(args.graphs))
    if args.lanes:
        dd.set_tuning(args.lanes, args.group, args.prefetch)
    meta = sc.FilterbankMetadata(fch1=8419.0, foff=FOFF, coarse_channel_size=N)
    bufs = [sc.FilterbankBuffer(T, N, data=data[cc]) for cc in range(num_cc)]

    def loop_step():
        hits = []
        for cc in range(num_cc):
            dd.search(bufs[cc], meta, sc.NO_BEAM, cc, MAX_DRIFT, 0.0, SNR_THRESHOLD, hits)
        return hits

    def timed(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = step()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    loop_elapsed, loop_hits = timed(loop_step)
    set_ms, set_launches, set_channels = 0.0, 0, 0
    if args.loop:
        elapsed, hits = loop_elapsed, loop_hits
        mode = "one Dedopplerer::search call per coarse channel"
    else:
        # The same searches as one batched graph launch (sc_batch): the
        # statistics/merge tail of each coarse channel overlaps the next one's
        # path sums. Hits must equal the per-call searches'.
        from seticore_amd import batch as B
        tuning = []
        if args.lanes:
            tuning = [B.ScTuning(T, N, args.lanes, args.group, args.prefetch, 4 if T <= 16 else 0,
                                 0)]
        batch = B.Batch([B.instance(data[cc], T, FOFF, TSAMP, MAX_DRIFT, 0.0, SNR_THRESHOLD,
                                    beam=sc.NO_BEAM, coarse_channel=cc, metadata=meta)
                         for cc in range(num_cc)], tuning)

        def batch_step():
            batch.run()
            return [h for cc in range(num_cc) for h in batch.hits(cc)]

        elapsed, hits = timed(batch_step)
        key = lambda h: h.key() + (h.snr, h.power)
        if sorted(map(key, hits)) != sorted(map(key, loop_hits)):
            print("batched hits differ from the per-call searches", file=sys.stderr)
            sys.exit(1)
        mode = "one batched graph launch over the coarse channels (hits equal to per-call searches)"
        batch_hits_cc = {cc: batch.hits(cc) for cc in range(num_cc)}
        # The timed step's path-sum kernel, measured live: one more run of the
        # batch's launch sequence, launched directly with HIP events around
        # each path-sum launch (at 16 time steps one launch covers up to eight
        # coarse channels: fused_tree_top_set_kernel).
        batch.set_timing(True)
        batch.run()
        set_ms, set_launches, set_channels = batch.kernel_time()
        batch.set_timing(False)
        del batch
    if args.loop:
        batch_hits_cc = {cc: [h for h in hits if h.coarse_channel == cc] for cc in range(num_cc)}
    # The per-call kernel's duration, measured live with HIP events on the
    # search stream around the path-sum kernel (one extra pass, launched
    # eagerly).
    dd.set_timing(True)
    loop_step()
    torch.cuda.synchronize()
    kernel_ms, launches = dd.kernel_time()
    dd.set_timing(False)
    per_call_kernel_ms = kernel_ms / max(1, launches)
    channels_per_launch = 1
    kernel_name = "fused_tree_top_kernel" if T <= 16 else "staged_tree_kernel"
    if not args.loop and set_launches and set_channels > set_launches:
        kernel_ms, launches = set_ms, set_launches
        channels_per_launch = set_channels / set_launches
        kernel_name = "fused_tree_top_set_kernel"

    ms_per_step = 1000.0 * elapsed / args.steps
    total_ps = per_cc * num_cc
    value = total_ps / (elapsed / args.steps) / 1e9

    # Roofline of the dominant kernel of the timed step (tree + top-path
    # tracking over all drift blocks; at 16 time steps in the batch, one launch
    # per run of up to eight coarse channels).
    # * T_r <= 16: bound by vector ALU issue. The input is read from HBM about
    #   once per launch and the drift blocks are served from LDS and L2.
    #   achieved = algorithmic VALU lane-operations per launch (path sums x
    #   (log2(T_r) adds + 1/2 max)) / launch time, peak = 256 CUs x 4 SIMDs x
    #   16 lanes x 2.4 GHz; beside it the HBM traffic the counters measured
    #   per launch (profiles/) as a fraction of 8 TB/s, and the
    #   reference-dataflow-equivalent byte rate, which is not a fraction of a
    #   roof here.
    # * T_r >= 32: bound by memory. The shear moves a row by T_r - 1 channels
    #   per drift block, more than a tile that fits in LDS, so the input is
    #   streamed once per drift block: the reference's dataflow bytes
    #   (D*T*N*4 + T*N*4 + 12*N per launch, dedoppler.cu:214-227) are the
    #   floor. achieved = those bytes / launch time, peak 8 TB/s; the VALU
    #   fraction is reported beside it.
    avg_ms = kernel_ms / max(1, launches)
    ops_per_ps = alg_valu_ops_per_path_sum(T)
    ps_per_launch = int(round(per_cc * channels_per_launch))
    alg_ops = ps_per_launch * ops_per_ps
    achieved_tops = alg_ops / (avg_ms * 1e-3) / 1e12
    peak_tops = VALU_PEAK_OPS / 1e12
    ref_dataflow_bytes = int(round((nblocks * T * N * 4 + T * N * 4 + 12 * N) *
                                   channels_per_launch))
    traffic = load_traffic(kernel_name, T, MAX_DRIFT)
    ref_gbs = ref_dataflow_bytes / (avg_ms * 1e-3) / 1e9

    rl_common = {
        "kernel": kernel_name,
        "kernel_ms_per_launch": round(avg_ms, 4),
        "coarse_channels_per_launch": round(channels_per_launch, 3),
        "kernel_ms_per_coarse_channel": round(avg_ms / channels_per_launch, 4),
        "per_call_kernel_ms": round(per_call_kernel_ms, 4),
        "path_sums_per_launch": ps_per_launch,
        "hbm_traffic_gbs": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
        "hbm_traffic_frac": (round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                             if traffic else None),
    }
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpath-sums/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded exponential noise + 200 tones, generated on the GPU)",
        "config": {
            "workload": "plain search, file scale: %d coarse channels x 2^20 fine channels x "
                        "%d time steps float32, max drift %g Hz/s (%d drift blocks), "
                        "SNR 10, device statistics" % (num_cc, T, MAX_DRIFT, nblocks),
            "coarse_channels": num_cc, "fine_channels": N, "time_steps": T,
            "drift_blocks": nblocks, "path_sums_per_step": total_ps,
            "hits_per_step": len(hits), "step": mode,
            "per_call_loop_gpath_per_s": round(total_ps / (loop_elapsed / args.steps) / 1e9, 3),
        },
        "roofline": ({
            "bound": "valu",
            "achieved": round(achieved_tops, 3),
            "peak": round(peak_tops, 3),
            "unit": "TOP/s (VALU lane-operations, non-packed)",
            "frac": round(achieved_tops / peak_tops, 4),
            "traffic": traffic,
            "alg_ops_per_path_sum": ops_per_ps,
            "alg_ops_per_launch": alg_ops,
            "reference_dataflow_equiv_gbs": round(ref_gbs, 1),
            "reference_dataflow_bytes_per_launch": ref_dataflow_bytes,
        } if T <= 16 else {
            "bound": "hbm",
            "achieved": round(ref_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ref_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_bytes_per_launch": ref_dataflow_bytes,
            "valu_frac": round(achieved_tops / peak_tops, 4),
        }),
    }

    line["roofline"].update(rl_common)

    if not args.no_cpu_baseline:
        host = host_description()
        avail = host["affinity_cpus"] or host["nproc"] or 1
        threads = max(1, min(HOST_THREAD_SHARE, avail))
        k = max(1, min(args.cpu_sample_cc, num_cc))
        total_dt = 0.0
        diff = {"same": 0, "only_gpu": 0, "only_oracle": 0,
                "all_diffs_between_thresholds": True, "common_hits_identical": True}
        for cc in range(k):
            sample = data[cc].cpu().numpy()
            dt, ref = cpu_baseline(sample, threads)
            total_dt += dt
            d = hits_vs_oracle(batch_hits_cc[cc], ref, sample, SNR_THRESHOLD)
            for key_ in ("same", "only_gpu", "only_oracle"):
                diff[key_] += d[key_]
            for key_ in ("all_diffs_between_thresholds", "common_hits_identical"):
                diff[key_] = diff[key_] and d[key_]
        line["cpu_baseline"] = {
            "value": round(per_cc * k / total_dt / 1e9, 4),
            "unit": "Gpath-sums/s",
            "cores": threads,
            "kind": "port",
            "sample": "%d coarse channel(s) of the same workload (%d x 2^20, %d drift "
                      "blocks), full search incl. host statistics, %d threads (the box's "
                      "CPU share for one GPU is %d)" % (k, T, nblocks, threads, HOST_THREAD_SHARE),
            "host": host,
            "hits_vs_oracle": diff,
        }

    # The same step after the loader's DC-spike fix (h5_reader.cpp:202-209:
    # the centre bin of every coarse channel replaced by the mean of its two
    # neighbours), applied on the GPU in place: hits and per-call throughput
    # with the fix, beside the unconditioned figures above.
    dc = N // 2
    before = sum(1 for h in loop_hits if abs(h.index - dc) <= 1)
    for cc in range(num_cc):
        sc.condition_coarse_channel(data[cc], T, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fixed_hits = loop_step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    line["config"]["dc_spike_fix"] = {
        "hits_per_step": len(fixed_hits),
        "hits_at_the_centre_bin": sum(1 for h in fixed_hits if abs(h.index - dc) <= 1),
        "hits_at_the_centre_bin_without_fix": before,
        "per_call_loop_gpath_per_s": round(total_ps / dt / 1e9, 3),
    }
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# Further workloads (not the headline): each prints one JSON line.

def _line(metric_value, unit, ms_per_step, args, dtype, data, config, **extra):
    line = {
        "metric": METRIC, "value": round(metric_value, 3), "unit": unit, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": data, "config": config,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


def _add_rfi_and_bursts(data, torch, seed):
    """Stationary narrowband RFI up to 1e6 x the noise and broadband bursts on
    3 % of the time steps at 100 x the noise, per coarse channel."""
    rng = np.random.default_rng(seed)
    for cc in range(data.shape[0]):
        gain = float(data[cc, :, N // 4:N // 4 + 4096].mean())
        for _ in range(4):
            c = int(rng.integers(int(0.06 * N), int(0.94 * N)))
            data[cc, :, c] += gain * float(10 ** rng.uniform(1.0, 6.0))
        for t in range(T):
            if rng.uniform() < 0.03:
                data[cc, t] += 100.0 * gain
    torch.cuda.synchronize()


def _timed_searches(torch, fn, num_cc, steps, warmup):
    for _ in range(warmup):
        for cc in range(num_cc):
            fn(cc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for cc in range(num_cc):
            fn(cc)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def run_half(args, torch, sc):
    """Reduced-precision storage: the file-scale layout plus RFI and bursts,
    normalised per coarse channel (FP64 moments on the GPU) and stored as FP16
    and BF16; FP32 tree accumulation. Compares throughput and hit sets with
    the FP32 search of the raw data."""
    num_cc = args.coarse_channels
    data = generate(num_cc, torch, sc)
    _add_rfi_and_bursts(data, torch, SEED + 1)
    per_cc, nblocks = path_sums_per_cc()
    meta = sc.FilterbankMetadata(fch1=8419.0, foff=FOFF, coarse_channel_size=N)
    results = {}
    hitsets = {}
    for name, dtype in (("f32", sc.SC_F32), ("f16", sc.SC_F16), ("bf16", sc.SC_BF16)):
        if dtype == sc.SC_F32:
            store = data
        else:
            store = torch.empty((num_cc, T, N), dtype=torch.float16 if dtype == sc.SC_F16
                                else torch.bfloat16, device="cuda")
            for cc in range(num_cc):
                sc.normalize(data[cc], dtype, out=store[cc])
        dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
        dd.set_stats_mode(sc.SC_STATS_DEVICE)
        bufs = [sc.FilterbankBuffer(T, N, data=store[cc], dtype=dtype) for cc in range(num_cc)]
        hits = {}

        def one(cc):
            out = []
            dd.search(bufs[cc], meta, sc.NO_BEAM, cc, MAX_DRIFT, 0.0, SNR_THRESHOLD, out)
            hits[cc] = out

        # At least two warm-up passes: a search's graph is captured the second
        # time it comes, so the timed passes replay graphs as a steady stream does.
        sec = _timed_searches(torch, one, num_cc, args.steps, max(2, args.warmup))
        results[name] = sec
        hitsets[name] = {(cc, h.index, h.drift_steps): h.snr for cc in hits for h in hits[cc]}
        if dtype != sc.SC_F32:
            del store
        del dd, bufs
    ref = hitsets["f32"]
    strong = {k for k, v in ref.items() if v >= 20}
    diffs = {}
    for name in ("f16", "bf16"):
        got = hitsets[name]
        common = set(ref) & set(got)
        rel = [abs(got[k] - ref[k]) / abs(ref[k]) for k in common if abs(ref[k]) > 0]
        diffs[name] = {
            "hits": len(got), "only_in_f32": len(set(ref) - set(got)),
            "only_in_half": len(set(got) - set(ref)),
            "strong_f32_hits_found": "%d/%d" % (len(strong & set(got)), len(strong)),
            "median_rel_snr_diff_vs_f32": float(np.median(rel)) if rel else None,
        }
    value = per_cc * num_cc / results["bf16"] / 1e9
    _line(value, "Gpath-sums/s", 1000 * results["bf16"], args, "bf16 storage, f32 accumulate",
          "synthetic (file-scale layout + narrowband RFI up to 1e6x + 3 % burst rows)",
          {"workload": "reduced-precision storage: %d coarse channels x 2^20 x 16, normalised "
                       "per coarse channel, max drift 10 Hz/s, SNR 10, device statistics"
                       % num_cc,
           "f32_gpath_per_s": round(per_cc * num_cc / results["f32"] / 1e9, 3),
           "f16_gpath_per_s": round(per_cc * num_cc / results["f16"] / 1e9, 3),
           "bf16_gpath_per_s": round(value, 3),
           "f32_hits": len(ref), "vs_f32": diffs})


def run_batched(args, torch, sc):
    """64 beams x 32 coarse channels: time steps per beam from {16, 32, 64, 256},
    fine channels per coarse channel from {2^17, 2^18, 2^20}, 20 tones per beam
    at +-5 Hz/s and SNR 10-50, max drift 5 Hz/s (66 drift blocks at every time
    step count), autotuned kernel shapes. The full input is ~353 GB as float32,
    more than one GPU holds, so the beams are searched in waves: each wave is
    as many consecutive beams (mixed time steps) as fit in WAVE_BYTES, generated
    on the GPU (untimed) and searched by one batch graph launch over all its
    (beam, coarse channel) instances. With --storage bf16 every instance is
    normalised to BF16 (FP64 moments on the GPU) and the whole input (~177 GB)
    is searched by a single graph launch. The same data is searched instance
    by instance with Dedopplerer::search (the per-instance plain loop, as
    run_dedoppler.cpp:43-52 and beamforming_pipeline.cpp:230-255 call it);
    hits must agree instance by instance."""
    from seticore_amd import batch as B
    rng = np.random.default_rng(SEED + 3)
    beams, ccs = 64, 32
    if args.coarse_channels != NUM_CC:
        beams = max(1, args.coarse_channels)
    bf16 = args.storage == "bf16"
    beam_T = [int(rng.choice([16, 32, 64, 256])) for _ in range(beams)]
    cc_N = [int(rng.choice([1 << 17, 1 << 18, 1 << 20])) for _ in range(ccs)]
    max_drift = 5.0
    wave_bytes = (args.wave_gib << 30) if args.wave_gib else (1 << 62 if bf16 else 96 << 30)
    bytes_per = 2 if bf16 else 4
    waves, cur, cur_bytes = [], [], 0
    for beam in range(beams):
        nb = beam_T[beam] * sum(cc_N) * bytes_per
        if cur and cur_bytes + nb > wave_bytes:
            waves.append(cur)
            cur, cur_bytes = [], 0
        cur.append(beam)
        cur_bytes += nb
    waves.append(cur)
    # Tones per beam, drawn before any timing so every run sees the same data.
    beam_tones = []
    for beam in range(beams):
        tones = {}
        for _ in range(20):
            c = int(rng.integers(0, ccs))
            drift = float(rng.uniform(-5.0, 5.0))
            tones.setdefault(c, []).append((float(rng.uniform(0.06, 0.94)), drift,
                                            float(rng.uniform(10.0, 50.0))))
        beam_tones.append(tones)
    t_tune = time.perf_counter()
    tuning = []
    for Tb in sorted(set(beam_T)):
        for Nc in sorted(set(cc_N)):
            tuning.append(B.autotune(Tb, Nc, FOFF, TSAMP, max_drift, reps=2))
    t_tune = time.perf_counter() - t_tune
    metas = {Nc: sc.FilterbankMetadata(fch1=8419.0, foff=FOFF, coarse_channel_size=Nc)
             for Nc in set(cc_N)}
    dtype = sc.SC_BF16 if bf16 else sc.SC_F32
    plain = {}
    t_batch = t_plain = t_gen = 0.0
    total_ps = mismatched = total_hits = 0
    graph_nodes = 0
    for wave in waves:
        t0 = time.perf_counter()
        bufs, inst = [], []
        for beam in wave:
            Tb = beam_T[beam]
            for c, Nc in enumerate(cc_N):
                x = torch.empty((Tb, Nc), dtype=torch.float32, device="cuda")
                tl = [(f * Nc, d * TSAMP / (FOFF * 1e6), snr / math.sqrt(Tb))
                      for f, d, snr in beam_tones[beam].get(c, [])]
                sc.gen_waterfall(x, Tb, seed=SEED * 7 + beam * 1000 + c, tones=tl)
                if bf16:
                    x = sc.normalize(x, sc.SC_BF16)[0]
                bufs.append((beam, c, Tb, Nc, x))
                inst.append(B.instance(x, Tb, FOFF, TSAMP, max_drift, 0.0, SNR_THRESHOLD,
                                       beam=-1, coarse_channel=c, dtype=dtype,
                                       metadata=metas[Nc]))
                nb = sc.drift_parameters(Tb, FOFF, TSAMP, max_drift)
                total_ps += (nb[1] - nb[0] + 1) * Tb * Nc
        torch.cuda.synchronize()
        batch = B.Batch(inst, tuning)  # plan and capture: once per wave, untimed
        graph_nodes += batch.info()["graph_nodes"]
        t_gen += time.perf_counter() - t0
        batch.run()  # warm-up launch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.run()
        t_batch += time.perf_counter() - t0
        batch_hits = [batch.hits(i) for i in range(len(inst))]
        # Per-instance plain loop over the same data.
        t0 = time.perf_counter()
        loop_hits = []
        for beam, c, Tb, Nc, x in bufs:
            key = (Tb, Nc)
            if key not in plain:
                dd = sc.Dedopplerer(Tb, Nc, FOFF, TSAMP)
                dd.set_stats_mode(sc.SC_STATS_DEVICE)
                dd.set_graphs(False)
                plain[key] = dd
            out = []
            plain[key].search(sc.FilterbankBuffer(Tb, Nc, data=x, dtype=dtype), metas[Nc],
                              sc.NO_BEAM, c, max_drift, 0.0, SNR_THRESHOLD, out)
            loop_hits.append(out)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        for a, b in zip(batch_hits, loop_hits):
            total_hits += len(b)
            mismatched += [h.key() + (h.snr,) for h in a] != [h.key() + (h.snr,) for h in b]
        del batch, bufs, inst, batch_hits, loop_hits
        torch.cuda.empty_cache()
    value = total_ps / t_batch / 1e9
    _line(value, "Gpath-sums/s", 1000 * t_batch, args,
          "bf16 storage, f32 accumulate" if bf16 else "f32",
          "synthetic (seeded exponential noise + 20 tones per beam, generated on the GPU "
          "wave by wave; generation, normalisation and graph capture untimed)",
          {"workload": "batched uneven search: %d beams x %d coarse channels, time steps per "
                       "beam from {16,32,64,256}, channels per coarse channel from "
                       "{2^17,2^18,2^20}, max drift 5 Hz/s, SNR 10, %s storage; %d wave(s) "
                       "of mixed-time-step beams, one graph launch per wave"
                       % (beams, ccs, "BF16" if bf16 else "FP32", len(waves)),
           "path_sums": total_ps, "beam_time_steps": beam_T, "coarse_channel_widths": cc_N,
           "waves": [len(w) for w in waves], "graph_nodes": graph_nodes,
           "plain_loop_gpath_per_s": round(total_ps / t_plain / 1e9, 3),
           "speedup_vs_plain_loop": round(t_plain / t_batch, 3),
           "instances_with_different_hits": mismatched, "hits": total_hits,
           "autotune_seconds": round(t_tune, 2), "setup_seconds": round(t_gen, 2),
           "tuning": [t.as_dict() for t in tuning]},
          steps_note="one timed launch per wave after one warm-up launch (the steps/warmup "
                     "flags do not apply)")


def run_learned(args, torch, sc):
    """Proposal-gated search on the file-scale layout with the learned-proposal
    generator (tones at SNR 8-50, drift +-10 Hz/s, narrowband RFI, bursts).
    The network is trained here on 2^16 blocks (256 coarse channels) labelled
    by the plain search, thresholds are fixed on 32 held-out coarse channels
    for block recalls of 0.9, 0.99, 0.995 and 1.0, and the gated search is
    timed against the plain search on 64 further coarse channels (FP32 and
    BF16 storage), with hit recall versus the plain search - per call, and
    batched (runs of eight coarse channels: one launch per front step -
    statistics, pooling, network, tile lists - and one tree launch over the
    tile lists, the next run's front beside the current run's tree) against
    the plain batch. The headline value is the combined form: gated, batched,
    BF16 storage, at 0.995 block recall."""
    from seticore_amd import proposal as P
    snr = args.snr_threshold
    t0 = time.perf_counter()
    train = P.make_dataset(256, N, seed=11, snr_threshold=snr)
    val = P.make_dataset(32, N, seed=12, snr_threshold=snr)
    net = sc.ProposalNet(max_batch=256, seed=13)
    hist = P.train(net, train)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    z = P.scores(net, val.tiles)
    zh = P.heuristic_scores(val.tiles)
    recalls = (0.9, 0.99, 0.995, 1.0)
    thresholds = {r: P.threshold_for_recall(z, val.labels, r) for r in recalls}
    # Held-out block fractions each method flags at the same block recall.
    val_flagged = {}
    for r in recalls:
        th, hh = thresholds[r], P.threshold_for_recall(zh, val.labels, r)
        val_flagged[str(r)] = {"network": round(float((z > th).float().mean()), 4),
                               "brightest_pixel": round(float((zh > hh).float().mean()), 4)}
    del train, val
    num_cc = args.coarse_channels
    rng = np.random.default_rng(SEED + 5)
    recipes = [P.coarse_channel_recipe(rng, N, num_tones=4, num_rfi=1) for _ in range(num_cc)]
    data = torch.empty((num_cc, T, N), dtype=torch.float32, device="cuda")
    for cc in range(num_cc):
        P.synth(data[cc], SEED * 31 + cc, recipes[cc])
    half = torch.empty((num_cc, T, N), dtype=torch.bfloat16, device="cuda")
    for cc in range(num_cc):
        sc.normalize(data[cc], sc.SC_BF16, out=half[cc])
    per_cc, nblocks = path_sums_per_cc()
    meta = sc.FilterbankMetadata(fch1=8419.0, foff=FOFF, coarse_channel_size=N)
    dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
    dd.set_stats_mode(sc.SC_STATS_DEVICE)
    dd.set_graphs(False)
    b32 = [sc.FilterbankBuffer(T, N, data=data[cc]) for cc in range(num_cc)]
    b16 = [sc.FilterbankBuffer(T, N, data=half[cc], dtype=sc.SC_BF16) for cc in range(num_cc)]
    plain_hits = {}

    def plain(cc):
        out = []
        dd.search(b32[cc], meta, sc.NO_BEAM, cc, MAX_DRIFT, 0.0, snr, out)
        plain_hits[cc] = {(h.index, h.drift_steps) for h in out}

    t_plain = _timed_searches(torch, plain, num_cc, args.steps, args.warmup)
    nhits = sum(len(v) for v in plain_hits.values())
    from seticore_amd import batch as B
    inst = {name: [B.instance(src[cc], T, FOFF, TSAMP, MAX_DRIFT, 0.0, snr, coarse_channel=cc,
                              dtype=dt, metadata=meta) for cc in range(num_cc)]
            for name, src, dt in (("f32", data, sc.SC_F32), ("bf16", half, sc.SC_BF16))}

    def timed_batch(b):
        for _ in range(args.warmup):
            b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    t_plain_batch = {}
    for name in ("f32", "bf16"):
        pb = B.Batch(inst[name])
        t_plain_batch[name] = timed_batch(pb)
        if name == "f32":
            assert all({(h.index, h.drift_steps) for h in pb.hits(cc)} == plain_hits[cc]
                       for cc in range(num_cc)), "plain batch differs from the per-call search"
        del pb
    points = []
    for rec, thr in sorted(thresholds.items()):
        for name, bufs in (("f32", b32), ("bf16", b16)):
            found = [0]
            flagged = [0]

            def gated(cc):
                out = []
                flagged[0] += dd.search_proposed(bufs[cc], net, meta, sc.NO_BEAM, cc, MAX_DRIFT,
                                                 0.0, snr, thr, out)
                if name == "f32":
                    found[0] += len(plain_hits[cc] & {(h.index, h.drift_steps) for h in out})

            _timed_searches(torch, gated, num_cc, 1, 0)  # recall and flags; warm-up
            frac = flagged[0] / (num_cc * (N // 4096))
            recall = found[0] / max(nhits, 1)
            sec = _timed_searches(torch, gated, num_cc, args.steps, 0)
            gb = B.Batch(inst[name], proposal=net, logit_threshold=thr)
            sec_b = timed_batch(gb)
            bflag = sum(gb.flagged(cc)[1] for cc in range(num_cc))
            assert bflag == round(frac * num_cc * (N // 4096)), \
                "gated batch flags differ from the per-call search"
            brec = (sum(len(plain_hits[cc] & {(h.index, h.drift_steps) for h in gb.hits(cc)})
                        for cc in range(num_cc)) / max(nhits, 1)) if name == "f32" else None
            del gb
            points.append({"val_block_recall": rec, "storage": name,
                           "threshold": round(thr, 4), "flagged_fraction": round(frac, 4),
                           "hit_recall_vs_plain": round(recall, 4) if name == "f32" else None,
                           "equiv_gpath_per_s": round(per_cc * num_cc / sec / 1e9, 3),
                           "speedup_vs_plain": round(t_plain / sec, 3),
                           "batched": {
                               "equiv_gpath_per_s": round(per_cc * num_cc / sec_b / 1e9, 3),
                               "speedup_vs_plain_per_call_f32": round(t_plain / sec_b, 3),
                               "speedup_vs_plain_batch_f32": round(t_plain_batch["f32"] / sec_b,
                                                                   3),
                               "speedup_vs_plain_batch_same_storage":
                                   round(t_plain_batch[name] / sec_b, 3),
                               "hit_recall_vs_plain": (round(brec, 4) if brec is not None
                                                       else None)}})
    main_pt = [p for p in points if p["val_block_recall"] == 0.995 and p["storage"] == "bf16"][0]
    main_b = main_pt["batched"]
    _line(main_b["equiv_gpath_per_s"], "Gpath-sums/s (equivalent: full-search path sums / time)",
          1000 * t_plain / main_b["speedup_vs_plain_per_call_f32"], args,
          "bf16 storage and network, f32 tree accumulation",
          "synthetic (learned-proposal generator: tones SNR 8-50, +-10 Hz/s, RFI, bursts)",
          {"workload": "learned proposal + batched + BF16 storage: %d coarse channels x 2^20 x "
                       "16, 4096-channel blocks, SNR threshold %g, logit threshold for 0.995 "
                       "block recall fixed on held-out data; against the plain per-call FP32 "
                       "search" % (num_cc, snr),
           "snr_threshold": snr,
           "combined_speedup_vs_plain_per_call_f32": main_b["speedup_vs_plain_per_call_f32"],
           "combined_speedup_vs_plain_batch_f32": main_b["speedup_vs_plain_batch_f32"],
           "plain_gpath_per_s": round(per_cc * num_cc / t_plain / 1e9, 3),
           "plain_batch_gpath_per_s": {k: round(per_cc * num_cc / v / 1e9, 3)
                                       for k, v in t_plain_batch.items()},
           "plain_hits": nhits, "points": points,
           "held_out_blocks_flagged_at_block_recall": val_flagged,
           "train_seconds_incl_data": round(t_train, 2),
           "final_train_loss": round(hist[-1], 4)})


if __name__ == "__main__":
    main()
