"""Batched search of many coarse channels / beams with uneven shapes, and the
autotuner that picks kernel shapes per (time steps, channels).

No counterpart in lacker/seticore (it searches one coarse channel per call,
run_dedoppler.cpp:35-51). A batch is planned once - one search handle per
distinct shape, every instance's launch sequence captured into one hipGraph -
and each run is a single graph launch (sc_batch_* in include/seticore_amd.h).

The autotuner times every compiled kernel shape for a (T_r, N) on synthetic
exponential noise with HIP events around the path-sum kernels, and keeps the
fastest; results are cached in memory and optionally in a JSON file.
"""

import ctypes
import json
import os

from . import (SC_BF16, SC_F16, SC_F32, SC_OK, SC_STATS_DEVICE, Dedopplerer, FilterbankMetadata,
               ScHit, ScMeta, SeticoreError, _stream_ptr, check, gen_waterfall, lib,
               round_up_to_power_of_two)


def _torch_mod():
    import torch
    return torch


class ScInstance(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dtype", ctypes.c_int),
                ("num_timesteps", ctypes.c_int), ("num_channels", ctypes.c_int),
                ("row_stride", ctypes.c_long), ("foff", ctypes.c_double),
                ("tsamp", ctypes.c_double), ("meta", ScMeta), ("beam", ctypes.c_int),
                ("coarse_channel", ctypes.c_int), ("max_drift", ctypes.c_double),
                ("min_drift", ctypes.c_double), ("snr_threshold", ctypes.c_double)]


class ScTuning(ctypes.Structure):
    _fields_ = [("rounded_timesteps", ctypes.c_int), ("num_channels", ctypes.c_int),
                ("channels_per_lane", ctypes.c_int), ("drift_group", ctypes.c_int),
                ("prefetch", ctypes.c_int), ("pass_group", ctypes.c_int),
                ("pass_blocks", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_SIGS = {
    "sc_batch_create": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                        ctypes.POINTER(ctypes.c_void_p)],
    "sc_batch_create_proposed": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_float,
                                 ctypes.POINTER(ctypes.c_void_p)],
    "sc_batch_flagged_blocks": [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int)],
    "sc_batch_run": [ctypes.c_void_p, ctypes.c_void_p],
    "sc_batch_set_timing": [ctypes.c_void_p, ctypes.c_int],
    "sc_batch_kernel_time": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)],
    "sc_batch_copy_hits": [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ScHit), ctypes.c_size_t,
                           ctypes.POINTER(ctypes.c_size_t)],
    "sc_batch_info": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                      ctypes.POINTER(ctypes.c_int)],
    "sc_batch_destroy": [ctypes.c_void_p],
}


def _lib():
    L = lib()
    if not getattr(L, "_batch_sigs", False):
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L._batch_sigs = True
    return L


def instance(data, num_timesteps, foff, tsamp, max_drift, min_drift, snr_threshold, beam=-1,
             coarse_channel=0, dtype=SC_F32, metadata=None):
    """An ScInstance for a CUDA tensor [T_r][N], T_r = roundUpToPowerOfTwo(
    num_timesteps), of the sample type `dtype`, unit channel stride (the checks
    FilterbankBuffer makes: the batch graph reads T_r rows of N samples through
    the raw pointer). The instance holds a reference to the tensor, and a
    Batch holds its instances', so the memory outlives the captured graph."""
    torch = _torch_mod()
    if data.dim() != 2:
        raise ValueError("instance data must be a 2-D [T_r][N] tensor")
    T_r, N = data.shape
    if T_r != round_up_to_power_of_two(int(num_timesteps)):
        raise ValueError("instance data has %d rows; %d time steps need %d" %
                         (T_r, num_timesteps, round_up_to_power_of_two(int(num_timesteps))))
    want = {SC_F32: torch.float32, SC_F16: torch.float16, SC_BF16: torch.bfloat16}.get(dtype)
    if want is None or data.dtype != want:
        raise ValueError("instance data dtype %s does not match sample type %r" %
                         (data.dtype, dtype))
    if not data.is_cuda or data.stride(1) != 1:
        raise ValueError("instance data must be a CUDA tensor with unit channel stride")
    meta = (metadata or FilterbankMetadata(coarse_channel_size=N, foff=foff))._sc_meta()
    inst = ScInstance(data.data_ptr(), int(dtype), int(num_timesteps), int(N),
                      int(data.stride(0)), float(foff), float(tsamp), meta, int(beam),
                      int(coarse_channel), float(max_drift), float(min_drift),
                      float(snr_threshold))
    inst._source = data
    return inst


def multibeam_instances(multibeam, num_timesteps, coarse_channel_size, foff, tsamp, max_drift,
                        min_drift, snr_threshold, metadata=None, beams=None):
    """Instances for every (beam, coarse channel) region of a CUDA multibeam
    buffer [beams][T_r][num_channels] - the regions MultibeamBuffer::
    copyRegionAsync copies out one by one (multibeam_buffer.cu:65-76,
    beamforming_pipeline.cpp:230-245); here each instance reads its region in
    place (row stride = num_channels), so no copy is made. Beam numbers are
    the buffer's indices (or `beams`)."""
    B_, T_r, n = multibeam.shape
    if n % coarse_channel_size:
        raise ValueError("coarse_channel_size must divide the channel count")
    out = []
    for beam in range(B_):
        for cc in range(n // coarse_channel_size):
            region = multibeam[beam, :, cc * coarse_channel_size:(cc + 1) * coarse_channel_size]
            out.append(instance(region, num_timesteps, foff, tsamp, max_drift, min_drift,
                                snr_threshold, beam=beams[beam] if beams else beam,
                                coarse_channel=cc, metadata=metadata))
    return out


class Batch:
    """sc_batch: plan once (graph capture; a gated batch: its launch plan), run
    many times.

    With `proposal` (a ProposalNet), every instance is searched as
    Dedopplerer.search_proposed does, gated at `logit_threshold`; runs of up
    to eight consecutive instances of one shape share each front launch
    (statistics, pooling, network, tile lists) and one tree launch, and the
    next run's front runs beside the current run's tree."""

    def __init__(self, instances, tuning=(), proposal=None, logit_threshold=0.0):
        L = _lib()
        # A run reads the instances' memory (and the network's) through raw
        # pointers: keep them alive as long as the batch.
        self._sources = [getattr(i, "_source", None) for i in instances]
        self._proposal = proposal
        self._inst = (ScInstance * len(instances))(*instances)
        self._tuning = (ScTuning * max(1, len(tuning)))(*tuning)
        h = ctypes.c_void_p()
        args = (ctypes.cast(self._inst, ctypes.c_void_p), len(instances),
                ctypes.cast(self._tuning, ctypes.c_void_p), len(tuning))
        if proposal is None:
            check(L.sc_batch_create(*args, ctypes.byref(h)))
        else:
            check(L.sc_batch_create_proposed(*args, proposal._h, float(logit_threshold),
                                             ctypes.byref(h)))
        self._h = h
        self.count = len(instances)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None:
            try:
                _lib().sc_batch_destroy(h)
            except Exception:  # interpreter shutdown: the library is already gone
                pass
            self._h = None

    def run(self, stream=None):
        check(_lib().sc_batch_run(self._h, _stream_ptr(stream)))

    def hits(self, i):
        L = _lib()
        n = ctypes.c_size_t()
        check(L.sc_batch_copy_hits(self._h, i, None, 0, ctypes.byref(n)))
        buf = (ScHit * max(1, n.value))()
        check(L.sc_batch_copy_hits(self._h, i, buf, n.value, ctypes.byref(n)))
        return list(buf)[:n.value]

    def flagged(self, i):
        """Gated batch, last run: (channel tiles searched, blocks flagged) of instance i."""
        a, b = ctypes.c_int(), ctypes.c_int()
        check(_lib().sc_batch_flagged_blocks(self._h, i, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_timing(self, enable):
        """Plain batches: time the path-sum launches of later runs (direct
        launches instead of the graph); resets the totals."""
        check(_lib().sc_batch_set_timing(self._h, int(bool(enable))))

    def kernel_time(self):
        """(total kernel ms, path-sum launches, coarse channels covered) since set_timing."""
        ms, n, c = ctypes.c_double(), ctypes.c_long(), ctypes.c_long()
        check(_lib().sc_batch_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(n),
                                          ctypes.byref(c)))
        return ms.value, n.value, c.value

    def info(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(_lib().sc_batch_info(self._h, ctypes.byref(a), ctypes.byref(b)))
        return {"shapes": a.value, "graph_nodes": b.value}


# ---------------------------------------------------------------------------
# Autotuner.

FUSED_CANDIDATES = [(4, 4, 0), (4, 8, 0), (4, 16, 0), (4, 4, 1), (4, 8, 1), (4, 4, 2),
                    (4, 8, 2), (4, 16, 2), (8, 8, 0), (8, 16, 0), (8, 4, 0)]
FUSED32_CANDIDATES = [(4, 4, 8), (4, 8, 8), (4, 16, 8), (2, 4, 2), (2, 8, 2), (4, 8, 2)]
STAGED_CANDIDATES = [(4, 4, 0), (4, 8, 0), (4, 16, 0), (4, 32, 0), (4, 8, 1), (4, 16, 1),
                     (4, 32, 1), (4, 256, 1)]
TWO_PASS_CANDIDATES = [(4, 1), (4, 2), (4, 4), (8, 4), (8, 8)]

_cache = {}


def _device_name():
    import torch
    return torch.cuda.get_device_name(0)


def _time(dd, buf, meta, max_drift, reps):
    hits = []
    dd.search(buf, meta, -1, 0, max_drift, 0.0, 1e9, hits)  # warm-up (and allocation)
    dd.set_timing(True)
    for _ in range(reps):
        dd.search(buf, meta, -1, 0, max_drift, 0.0, 1e9, hits)
    ms, n = dd.kernel_time()
    dd.set_timing(False)
    return ms / max(n, 1)


def autotune(T_r, N, foff, tsamp, max_drift, reps=3, cache_file=None, log=None):
    """The fastest compiled kernel shape for (T_r, N) as an ScTuning."""
    import torch
    key = "%s|%d|%d|%.6g|%.6g|%.6g" % (_device_name(), T_r, N, foff, tsamp, max_drift)
    if key not in _cache and cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            _cache.update({k: v for k, v in json.load(f).items()})
    if key in _cache:
        return ScTuning(**_cache[key])
    x = torch.empty((T_r, N), dtype=torch.float32, device="cuda")
    gen_waterfall(x, T_r, seed=12345)
    from . import FilterbankBuffer
    buf = FilterbankBuffer(T_r, N, data=x)
    meta = FilterbankMetadata(coarse_channel_size=N, foff=foff)
    dd = Dedopplerer(T_r, N, foff, tsamp)
    dd.set_stats_mode(SC_STATS_DEVICE)
    dd.set_graphs(False)
    best = None
    if T_r <= 256:
        for c, g, p in (FUSED_CANDIDATES if T_r <= 16 else
                        FUSED32_CANDIDATES if T_r == 32 else STAGED_CANDIDATES):
            try:
                dd.set_tuning(c, g, p)
            except SeticoreError:
                continue
            ms = _time(dd, buf, meta, max_drift, reps)
            if log:
                log("T=%d N=%d fused C=%d G=%d P=%d: %.3f ms" % (T_r, N, c, g, p, ms))
            if best is None or ms < best[0]:
                best = (ms, ScTuning(T_r, N, c, g, p, 4 if T_r <= 16 else 0, 0))
    if T_r >= 32:
        for g, nb in TWO_PASS_CANDIDATES:
            try:
                dd.set_two_pass(g, nb)
            except SeticoreError:
                continue
            ms = _time(dd, buf, meta, max_drift, reps)
            if log:
                log("T=%d N=%d two-pass G=%d blocks=%d: %.3f ms" % (T_r, N, g, nb, ms))
            if best is None or ms < best[0]:
                best = (ms, ScTuning(T_r, N, 4, 8, 0, g, nb))
    if best is None:  # no tunable kernel (T_r > 256): default shapes
        best = (0.0, ScTuning(T_r, N, 4, 8, 0, 4, 0))
    _cache[key] = best[1].as_dict()
    if cache_file:
        with open(cache_file, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)
    del dd, x
    return best[1]
