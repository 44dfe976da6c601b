"""seticore_amd: MI355X-native Taylor-tree dedoppler search.

Python mirror of lacker/seticore's dedoppler interface, calling the HIP
kernels through the C ABI in ``include/seticore_amd.h``:

* :class:`Dedopplerer` - dedoppler.h:15-79 (``search``, ``addIncoherentPower``,
  ``memoryUsage``, ``print_hits``).
* :class:`FilterbankBuffer` - filterbank_buffer.h:11-48 (``[time][channel]``
  float32 on the GPU, owning or a view).
* :class:`FilterbankMetadata` - the fields of filterbank_metadata.h:15-62 that
  the search and the writers read.
* :class:`DedopplerHit` - dedoppler_hit.h:11-66.
* :class:`DatFileWriter`, :class:`HitFileWriter`, :func:`makeHitRecorder` -
  hit_recorder.h, dat_file_writer.h, hit_file_writer.h.

Device memory and streams come from PyTorch (ROCm); every computation runs in
``libseticore_amd.so``. There is no CPU fallback: if the library or a GPU is
missing the calls raise.
"""

import ctypes
import math
import os

__version__ = "0.1.0"

_HERE = os.path.dirname(os.path.abspath(__file__))
# The in-tree build; SETICORE_AMD_LIBRARY selects another build of the same
# sources (for comparing kernel variants).
LIBRARY_PATH = os.environ.get("SETICORE_AMD_LIBRARY") or os.path.join(_HERE, "libseticore_amd.so")

NO_BEAM = -1  # dedoppler_hit.h:9

SC_OK = 0
SC_ERR_INVALID_ARGUMENT = 1
SC_ERR_HIP = 2
SC_ERR_OUT_OF_MEMORY = 3
SC_ERR_UNSUPPORTED = 4
SC_ERR_IO = 5
SC_ERR_INTERNAL = 6

SC_F32, SC_F16, SC_BF16 = 0, 1, 2
SC_STATS_HOST_EXACT, SC_STATS_DEVICE = 0, 1


class SeticoreError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("seticore_amd error %d: %s" % (status, message))
        self.status = status


class ScMeta(ctypes.Structure):
    _fields_ = [("fch1", ctypes.c_double), ("foff", ctypes.c_double),
                ("coarse_channel_size", ctypes.c_long),
                ("num_coherent_beams", ctypes.c_int)]


class DedopplerHit(ctypes.Structure):
    """dedoppler_hit.h:11-66. The structure is the C ABI's sc_hit, so hits come
    back from the library with one buffer copy."""

    _fields_ = [("frequency", ctypes.c_double), ("index", ctypes.c_int),
                ("drift_steps", ctypes.c_int), ("drift_rate", ctypes.c_double),
                ("snr", ctypes.c_float), ("coarse_channel", ctypes.c_int),
                ("beam", ctypes.c_int), ("num_timesteps", ctypes.c_int),
                ("power", ctypes.c_float), ("incoherent_power", ctypes.c_float)]

    def lowIndex(self):
        return min(self.index, self.index + self.drift_steps)

    def highIndex(self):
        return max(self.index, self.index + self.drift_steps)

    def toString(self):
        """dedoppler_hit.cpp:25-30 (fmt "{:.5f}" == printf "%.5f")."""
        n = self.drift_steps
        bins = "%d bin" % n if n in (1, -1) else "%d bins" % n
        return "coarse channel = %d, index = %d, snr = %.5f, drift rate = %.5f (%s)" % (
            self.coarse_channel, self.index, self.snr, self.drift_rate, bins)

    def __repr__(self):
        return "DedopplerHit(%s)" % self.toString()

    def key(self):
        return (self.index, self.drift_steps, self.coarse_channel, self.beam)


ScHit = DedopplerHit


class ScFileMeta(ctypes.Structure):
    _fields_ = [("source_name", ctypes.c_char_p), ("fch1", ctypes.c_double),
                ("foff", ctypes.c_double), ("tstart", ctypes.c_double),
                ("tsamp", ctypes.c_double), ("src_raj", ctypes.c_double),
                ("src_dej", ctypes.c_double), ("num_timesteps", ctypes.c_long),
                ("coarse_channel_size", ctypes.c_long), ("telescope_id", ctypes.c_int)]


class ScFilMeta(ctypes.Structure):
    _fields_ = [("source_name", ctypes.c_char * 256), ("fch1", ctypes.c_double),
                ("foff", ctypes.c_double), ("tstart", ctypes.c_double),
                ("tsamp", ctypes.c_double), ("src_raj", ctypes.c_double),
                ("src_dej", ctypes.c_double), ("telescope_id", ctypes.c_int),
                ("num_timesteps", ctypes.c_long), ("num_channels", ctypes.c_long),
                ("coarse_channel_size", ctypes.c_long), ("num_coarse_channels", ctypes.c_long),
                ("has_dc_spike", ctypes.c_int)]


_lib_handle = None


def lib():
    """Loads libseticore_amd.so (built in-tree by seticore_amd.build)."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not os.path.exists(LIBRARY_PATH):
        raise SeticoreError(SC_ERR_UNSUPPORTED,
                            "%s is not built; run python -m seticore_amd.build" % LIBRARY_PATH)
    L = ctypes.CDLL(LIBRARY_PATH)
    vp, i, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    lng = ctypes.c_long
    P = ctypes.POINTER
    sig = {
        "sc_dd_create": [i, i, d, d, i, P(vp)],
        "sc_dd_destroy": [vp],
        "sc_dd_memory_usage": [vp, P(sz)],
        "sc_dd_set_stats_mode": [vp, i],
        "sc_dd_set_tuning": [vp, i, i, i],
        "sc_dd_set_two_pass": [vp, i, i],
        "sc_dd_set_timing": [vp, i],
        "sc_dd_set_graphs": [vp, i],
        "sc_dd_kernel_time": [vp, P(d), P(lng)],
        "sc_dd_search": [vp, vp, i, i, i, lng, P(ScMeta), i, i, d, d, d, P(ScHit), sz,
                         P(sz), vp],
        "sc_dd_search_static": [i, i, i, d, d, vp, i, lng, P(ScMeta), i, i, d, d, d,
                                P(ScHit), sz, P(sz), vp],
        "sc_dd_copy_hits": [vp, P(ScHit), sz, P(sz)],
        "sc_dd_copy_state": [vp, vp, vp, vp, vp, P(ctypes.c_float), P(ctypes.c_float)],
        "sc_dd_add_incoherent_power": [vp, vp, i, i, i, lng, P(ScHit), sz, vp],
        "sc_dd_taylor_tree": [vp, vp, i, i, i, lng, i, vp, vp],
        "sc_normalize": [vp, i, i, lng, i, ctypes.c_float, vp, P(d), P(d), vp],
        "sc_gen_waterfall": [vp, i, i, i, ctypes.c_uint64, ctypes.c_float, ctypes.c_float,
                             ctypes.c_float, i, P(ctypes.c_float), P(ctypes.c_float),
                             P(ctypes.c_float), vp],
        "sc_dat_writer_open": [ctypes.c_char_p, P(ScFileMeta), d, P(vp)],
        "sc_hits_writer_open": [ctypes.c_char_p, P(ScFileMeta), i, P(ctypes.c_char_p),
                                P(d), P(d), P(vp)],
        "sc_writer_record_hit": [vp, P(ScHit), P(ctypes.c_float)],
        "sc_writer_close": [vp],
        "sc_events_writer_open": [ctypes.c_char_p, P(ScFileMeta), i, P(vp)],
        "sc_events_writer_record": [vp, P(P(ScHit)), P(P(ctypes.c_float))],
        "sc_find_events": [P(ctypes.c_char_p), i, ctypes.c_char_p, d, d, d, P(lng), vp],
        "sc_pn_create": [i, P(vp)],
        "sc_pn_destroy": [vp],
        "sc_pn_memory_usage": [vp, P(sz)],
        "sc_pn_init": [vp, ctypes.c_uint64],
        "sc_pn_set_params": [vp, P(ctypes.c_float), sz],
        "sc_pn_get_params": [vp, P(ctypes.c_float), sz],
        "sc_pn_pool": [vp, i, i, i, lng, vp, vp, vp],
        "sc_pn_forward": [vp, vp, i, vp, vp],
        "sc_pn_gradient": [vp, vp, vp, i, ctypes.c_float, vp, vp, vp],
        "sc_pn_train_step": [vp, vp, vp, i, ctypes.c_float, ctypes.c_float, vp, vp],
        "sc_pn_activations": [vp, P(vp), P(vp), P(vp), P(vp), P(vp)],
        "sc_fil_open": [ctypes.c_char_p, P(vp)],
        "sc_fil_metadata": [vp, P(ScFilMeta)],
        "sc_fil_set_coarse_channel_size": [vp, lng],
        "sc_fil_read_coarse_channel": [vp, i, i, P(ctypes.c_float)],
        "sc_fil_load_coarse_channel": [vp, i, i, vp, vp],
        "sc_fil_close": [vp],
        "sc_condition_coarse_channel": [vp, i, i, i, lng, i, vp],
        "sc_dd_search_proposed": [vp, vp, vp, i, i, i, lng, P(ScMeta), i, i, d, d, d,
                                  ctypes.c_float, P(ScHit), sz, P(sz), P(i), vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.sc_last_error_string.argtypes = []
    L.sc_last_error_string.restype = ctypes.c_char_p
    L.sc_abi_version.argtypes = []
    L.sc_abi_version.restype = ctypes.c_int
    L.sc_pn_num_params.argtypes = []
    L.sc_pn_num_params.restype = ctypes.c_int
    _lib_handle = L
    return L


def check(status):
    if status != SC_OK:
        msg = lib().sc_last_error_string().decode("utf-8", "replace")
        raise SeticoreError(status, msg)


def round_up_to_power_of_two(n):
    """util.cpp:16-22"""
    r = 1
    while r < n:
        r *= 2
    return r


roundUpToPowerOfTwo = round_up_to_power_of_two


def _torch():
    import torch
    return torch


def _stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


# ---------------------------------------------------------------------------
class FilterbankMetadata:
    """The metadata fields a search and the writers read
    (filterbank_metadata.h:15-62)."""

    def __init__(self, fch1=0.0, foff=0.0, coarse_channel_size=0, source_name="",
                 source_names=None, ras=None, decs=None, tstart=0.0, tsamp=0.0,
                 src_raj=0.0, src_dej=0.0, num_timesteps=0, num_channels=0,
                 telescope_id=-1, has_dc_spike=False):
        self.fch1 = fch1
        self.foff = foff
        self.coarse_channel_size = coarse_channel_size
        self.source_name = source_name
        self.source_names = list(source_names or [])
        self.ras = list(ras or [])
        self.decs = list(decs or [])
        self.tstart = tstart
        self.tsamp = tsamp
        self.src_raj = src_raj
        self.src_dej = src_dej
        self.num_timesteps = num_timesteps
        self.num_channels = num_channels
        self.telescope_id = telescope_id
        self.has_dc_spike = has_dc_spike

    @property
    def num_coarse_channels(self):
        return self.num_channels // self.coarse_channel_size if self.coarse_channel_size else 0

    def isCoherentBeam(self, beam):
        """filterbank_metadata.cpp:116-118"""
        return 0 <= beam < len(self.source_names)

    def _sc_meta(self):
        key = (self.fch1, self.foff, self.coarse_channel_size, len(self.source_names))
        cached = getattr(self, "_meta_cache", None)
        if cached is None or cached[0] != key:
            cached = (key, ScMeta(self.fch1, self.foff, int(self.coarse_channel_size),
                                  len(self.source_names)))
            self._meta_cache = cached
        return cached[1]

    def _sc_file_meta(self):
        return ScFileMeta(self.source_name.encode(), self.fch1, self.foff, self.tstart,
                          self.tsamp, self.src_raj, self.src_dej, int(self.num_timesteps),
                          int(self.coarse_channel_size), int(self.telescope_id))


class FilterbankBuffer:
    """Row-major ``data[time][channel]`` power on the GPU
    (filterbank_buffer.h:11-48). ``data`` is a torch tensor; pass ``data=`` to
    make a view on an existing ``[num_timesteps][>= num_channels]`` tensor."""

    def __init__(self, num_timesteps, num_channels, data=None, dtype=SC_F32):
        torch = _torch()
        self.num_timesteps = int(num_timesteps)
        self.num_channels = int(num_channels)
        self.dtype = dtype
        tdtype = {SC_F32: torch.float32, SC_F16: torch.float16, SC_BF16: torch.bfloat16}[dtype]
        if data is None:
            self.managed = True
            self.data = torch.zeros((self.num_timesteps, self.num_channels), dtype=tdtype,
                                    device="cuda")
        else:
            self.managed = False
            if data.dim() != 2 or data.shape[0] != self.num_timesteps or \
                    data.shape[1] < self.num_channels or data.stride(1) != 1:
                raise ValueError("view must be [num_timesteps][>= num_channels], unit stride")
            if data.dtype != tdtype or not data.is_cuda:
                raise ValueError("view must be a CUDA tensor of the buffer dtype")
            self.data = data
        self.size = self.num_timesteps * self.num_channels

    @property
    def row_stride(self):
        return self.data.stride(0)

    def zero(self):
        self.data.zero_()

    def set(self, time, channel, value):
        assert 0 <= time < self.num_timesteps and 0 <= channel < self.num_channels
        self.data[time, channel] = value

    def get(self, time, channel):
        return float(self.data[time, channel].item())


def makeNoisyBuffer(num_timesteps, num_channels):
    """filterbank_buffer.cu:71-78: row 0 = 0.1*chan/num_channels, rest 0."""
    torch = _torch()
    buf = FilterbankBuffer(num_timesteps, num_channels)
    buf.zero()
    row = torch.tensor([0.1 * c / num_channels for c in range(num_channels)],
                       dtype=torch.float64).to(torch.float32)
    buf.data[0, :] = row.to("cuda")
    return buf


class Dedopplerer:
    """dedoppler.h:15-79, running on the GPU through libseticore_amd.so."""

    def __init__(self, num_timesteps, num_channels, foff, tsamp, has_dc_spike=False):
        L = lib()
        self.num_timesteps = int(num_timesteps)
        self.num_channels = int(num_channels)
        self.foff = float(foff)
        self.tsamp = float(tsamp)
        self.has_dc_spike = bool(has_dc_spike)
        self.print_hits = False
        self.rounded_num_timesteps = round_up_to_power_of_two(self.num_timesteps)
        h = ctypes.c_void_p()
        _torch().cuda.init()
        check(L.sc_dd_create(self.num_timesteps, self.num_channels, self.foff, self.tsamp,
                             int(self.has_dc_spike), ctypes.byref(h)))
        self._h = h
        self._cap = 1024
        self._hits = (ScHit * self._cap)()
        self._n = ctypes.c_size_t()

    def close(self):
        if getattr(self, "_h", None):
            lib().sc_dd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def memoryUsage(self):
        n = ctypes.c_size_t()
        check(lib().sc_dd_memory_usage(self._h, ctypes.byref(n)))
        return n.value

    def set_stats_mode(self, mode):
        check(lib().sc_dd_set_stats_mode(self._h, int(mode)))

    def set_tuning(self, channels_per_lane, drift_group, prefetch=0):
        check(lib().sc_dd_set_tuning(self._h, int(channels_per_lane), int(drift_group),
                                     int(prefetch)))

    def set_two_pass(self, drift_group, blocks_per_launch=0):
        """Selects the T_r = 32..256 two-pass search and its shape (speed only)."""
        check(lib().sc_dd_set_two_pass(self._h, int(drift_group), int(blocks_per_launch)))

    def set_graphs(self, enable):
        check(lib().sc_dd_set_graphs(self._h, int(bool(enable))))

    def set_timing(self, enable):
        check(lib().sc_dd_set_timing(self._h, int(bool(enable))))

    def kernel_time(self):
        """(total ms of the path-sum kernels, number of timed searches)."""
        ms = ctypes.c_double()
        n = ctypes.c_long()
        check(lib().sc_dd_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def search(self, input, metadata, beam, coarse_channel, max_drift, min_drift,
               snr_threshold, output, stream=None):
        """dedoppler.cu:186-320: appends this coarse channel's hits to `output`."""
        L = lib()
        n = self._n
        status = L.sc_dd_search(self._h, input.data.data_ptr(), input.dtype,
                                input.num_timesteps, input.num_channels, input.row_stride,
                                ctypes.byref(metadata._sc_meta()), beam, coarse_channel,
                                max_drift, min_drift, snr_threshold, self._hits, self._cap,
                                ctypes.byref(n), _stream_ptr(stream))
        if status != SC_OK:
            check(status)
        count = n.value
        if count > self._cap:
            self._cap = count
            self._hits = (ScHit * count)()
            check(L.sc_dd_copy_hits(self._h, self._hits, count, ctypes.byref(n)))
        new = list((ScHit * count).from_buffer_copy(self._hits)) if count else []
        if self.print_hits:
            for h in new:
                print("hit: " + h.toString())
        output.extend(new)
        return new

    def search_proposed(self, input, net, metadata, beam, coarse_channel, max_drift, min_drift,
                        snr_threshold, logit_threshold, output, stream=None):
        """Gated search: the full tree only on blocks the proposal network flags
        (device statistics). Appends hits to `output`; returns the number of
        flagged 4096-channel blocks."""
        L = lib()
        n = self._n
        flagged = ctypes.c_int()
        status = L.sc_dd_search_proposed(self._h, net._h, input.data.data_ptr(), input.dtype,
                                         input.num_timesteps, input.num_channels,
                                         input.row_stride, ctypes.byref(metadata._sc_meta()),
                                         beam, coarse_channel, max_drift, min_drift,
                                         snr_threshold, float(logit_threshold), self._hits,
                                         self._cap, ctypes.byref(n), ctypes.byref(flagged),
                                         _stream_ptr(stream))
        if status != SC_OK:
            check(status)
        count = n.value
        if count > self._cap:
            self._cap = count
            self._hits = (ScHit * count)()
            check(L.sc_dd_copy_hits(self._h, self._hits, count, ctypes.byref(n)))
        new = list((ScHit * count).from_buffer_copy(self._hits)) if count else []
        output.extend(new)
        return flagged.value

    def state(self):
        """Column sums, top sums/blocks/offsets, median and std of the last search."""
        import numpy as np
        N = self.num_channels
        cs = np.empty(N, np.float32)
        ts = np.empty(N, np.float32)
        tb = np.empty(N, np.int32)
        to = np.empty(N, np.int32)
        med = ctypes.c_float()
        sd = ctypes.c_float()
        check(lib().sc_dd_copy_state(self._h, cs.ctypes.data, ts.ctypes.data, tb.ctypes.data,
                                     to.ctypes.data, ctypes.byref(med), ctypes.byref(sd)))
        return {"column_sums": cs, "top_sums": ts, "top_blocks": tb, "top_offsets": to,
                "median": med.value, "std": sd.value}

    def addIncoherentPower(self, input, hits, stream=None):
        """dedoppler.cu:139-176: sorts `hits` by drift steps, fills incoherent_power."""
        arr = (ScHit * max(1, len(hits)))(*hits)
        check(lib().sc_dd_add_incoherent_power(
            self._h, ctypes.c_void_p(input.data.data_ptr()), input.dtype, input.num_timesteps,
            input.num_channels, input.row_stride, arr, len(hits), _stream_ptr(stream)))
        hits[:] = [ScHit.from_buffer_copy(arr[k]) for k in range(len(hits))]

    def taylor_tree(self, input, drift_block, stream=None):
        """[T_r][N] path sums of one drift block (optimizedTaylorTree, taylor.cu:333-350)."""
        torch = _torch()
        out = torch.empty((self.rounded_num_timesteps, self.num_channels), dtype=torch.float32,
                          device="cuda")
        check(lib().sc_dd_taylor_tree(self._h, ctypes.c_void_p(input.data.data_ptr()),
                                      input.dtype, input.num_timesteps, input.num_channels,
                                      input.row_stride, int(drift_block),
                                      ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)))
        return out


def drift_parameters(num_timesteps, foff, tsamp, max_drift):
    """dedoppler.cu:87,195-198,274: (min_block, max_block, window, resolution)."""
    T = round_up_to_power_of_two(num_timesteps)
    res = 1e6 * foff / ((T - 1) * tsamp)
    nmd = max_drift / abs(res * (T - 1))
    return (int(math.floor(-nmd)), int(math.floor(nmd)), int(2 * math.ceil(nmd * (T - 1))), res)


def search_static(input, num_timesteps, foff, tsamp, metadata, beam, coarse_channel, max_drift,
                  min_drift, snr_threshold, stream=None):
    """The allocation-free search (``sc_dd_search_static``): the device-statistics
    search of :meth:`Dedopplerer.search` without a handle, its intermediates in
    the library's fixed per-GPU array. ``input`` holds the rounded time steps;
    ``num_timesteps`` is the unrounded count. Returns the hits. The output
    array (host memory, sized for one hit per merge window) is allocated here,
    by the caller of the entry point."""
    L = lib()
    window = drift_parameters(num_timesteps, foff, tsamp, max_drift)[2]
    cap = max(1, -(-input.num_channels // max(window, 1)))
    out = (ScHit * cap)()
    n = ctypes.c_size_t(0)
    check(L.sc_dd_search_static(int(num_timesteps), input.num_timesteps, input.num_channels,
                                foff, tsamp, input.data.data_ptr(), input.dtype,
                                input.row_stride, ctypes.byref(metadata._sc_meta()), beam,
                                coarse_channel, max_drift, min_drift, snr_threshold, out, cap,
                                ctypes.byref(n), _stream_ptr(stream)))
    if n.value > cap:
        raise SeticoreError(SC_ERR_INTERNAL, "more hits than merge windows")
    return list(out[:n.value])


# ---------------------------------------------------------------------------
class _Writer:
    def __init__(self, handle):
        self._h = handle

    def recordHit(self, hit, input=None):
        """hit_recorder.h:10-19. `input` is the host float32 [T][N] coarse channel
        (numpy array or torch tensor) - the .hits writer copies a slice of it."""
        ptr = None
        keep = None
        if input is not None:
            import numpy as np
            if hasattr(input, "detach"):
                input = input.detach().float().cpu().numpy()
            keep = np.ascontiguousarray(input, dtype=np.float32)
            ptr = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        check(lib().sc_writer_record_hit(self._h, ctypes.byref(hit), ptr))

    def close(self):
        if self._h:
            h, self._h = self._h, None
            check(lib().sc_writer_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DatFileWriter(_Writer):
    """dat_file_writer.h:15-27"""

    def __init__(self, filename, metadata, max_drift):
        h = ctypes.c_void_p()
        fm = metadata._sc_file_meta()
        check(lib().sc_dat_writer_open(filename.encode(), ctypes.byref(fm), float(max_drift),
                                       ctypes.byref(h)))
        super().__init__(h)


class HitFileWriter(_Writer):
    """hit_file_writer.h:16-31"""

    def __init__(self, filename, metadata):
        h = ctypes.c_void_p()
        fm = metadata._sc_file_meta()
        nb = len(metadata.source_names)
        names = (ctypes.c_char_p * max(1, nb))(*[s.encode() for s in metadata.source_names])
        ras = (ctypes.c_double * max(1, nb))(*metadata.ras)
        decs = (ctypes.c_double * max(1, nb))(*metadata.decs)
        check(lib().sc_hits_writer_open(filename.encode(), ctypes.byref(fm), nb, names, ras,
                                        decs, ctypes.byref(h)))
        super().__init__(h)


def makeHitRecorder(filename, metadata, max_drift):
    """hit_recorder.cpp:10-23: picks the writer by suffix."""
    if filename.endswith(".dat"):
        return DatFileWriter(filename, metadata, max_drift)
    if filename.endswith(".hits"):
        return HitFileWriter(filename, metadata)
    raise SeticoreError(SC_ERR_INVALID_ARGUMENT,
                        "could not recognize the file type of %s" % filename)


def gen_waterfall(out, num_timesteps, seed, gain=1.0, taper_frac=0.0, dc_spike=0.0,
                  tones=(), stream=None):
    """Fills a float32 CUDA tensor [T_r][N] with seeded synthetic power:
    exponential(1) noise * gain (cosine taper on the outer taper_frac of the
    channels, optional DC spike) plus linear tones (start, slope, amplitude)."""
    T_r, N = out.shape
    n = len(tones)
    F = ctypes.c_float * max(1, n)
    st = F(*[t[0] for t in tones])
    sl = F(*[t[1] for t in tones])
    am = F(*[t[2] for t in tones])
    check(lib().sc_gen_waterfall(ctypes.c_void_p(out.data_ptr()), int(num_timesteps), int(T_r),
                                 int(N), ctypes.c_uint64(seed), float(gain), float(taper_frac),
                                 float(dc_spike), n, st, sl, am, _stream_ptr(stream)))
    return out


def normalize(data, out_dtype, clamp=65504.0, stream=None, out=None):
    """Per-coarse-channel normalisation (x - loc) / std on the GPU, stored as
    `out_dtype` (into `out` if given, a contiguous [T][N] CUDA tensor of that
    type): loc = median column sum / T (so the normalised column sums have
    median ~0 and a positive SNR threshold is a positive path sum), std = the
    FP64 deviation of all samples. Returns (tensor, loc, std)."""
    torch = _torch()
    T, N = data.shape
    tdtype = {SC_F32: torch.float32, SC_F16: torch.float16, SC_BF16: torch.bfloat16}[out_dtype]
    if out is None:
        out = torch.empty((T, N), dtype=tdtype, device=data.device)
    elif out.dtype != tdtype or tuple(out.shape) != (T, N) or not out.is_contiguous():
        raise SeticoreError(SC_ERR_INVALID_ARGUMENT, "normalize: out must be a contiguous "
                                                     "[T][N] tensor of the output type")
    loc = ctypes.c_double()
    sd = ctypes.c_double()
    check(lib().sc_normalize(ctypes.c_void_p(data.data_ptr()), int(T), int(N),
                             int(data.stride(0)), int(out_dtype), float(clamp),
                             ctypes.c_void_p(out.data_ptr()), ctypes.byref(loc),
                             ctypes.byref(sd), _stream_ptr(stream)))
    return out, loc.value, sd.value


# ---------------------------------------------------------------------------
# Learned block proposal network (no counterpart in the reference).

BLOCK_CHANNELS = 4096
TILE_ROWS = 16
TILE_COLUMNS = 1024


def pn_num_params():
    return lib().sc_pn_num_params()


def _check_tiles(tiles, blocks=None):
    """Pooled tiles as the network's kernels read them: a contiguous CUDA
    bfloat16 tensor [blocks][16][1024]."""
    torch = _torch()
    if (not tiles.is_cuda or tiles.dtype != torch.bfloat16 or not tiles.is_contiguous() or
            tuple(tiles.shape[1:]) != (TILE_ROWS, TILE_COLUMNS) or
            (blocks is not None and tiles.shape[0] != blocks)):
        raise ValueError("tiles must be a contiguous CUDA bfloat16 tensor [%s][%d][%d]" %
                         ("blocks" if blocks is None else blocks, TILE_ROWS, TILE_COLUMNS))


def _check_labels(labels, blocks):
    torch = _torch()
    if not labels.is_cuda or labels.dtype != torch.float32 or labels.numel() < blocks:
        raise ValueError("labels must be a CUDA float32 tensor of one 0/1 value per block")


class ProposalNet:
    """The block proposal network: BF16 weights/activations, FP32 master
    weights and Adam state on the device. Tiles are BF16 CUDA tensors
    [B][16][1024]; labels and logits float32 CUDA tensors [B]."""

    def __init__(self, max_batch=256, seed=None):
        h = ctypes.c_void_p()
        check(lib().sc_pn_create(int(max_batch), ctypes.byref(h)))
        self._h = h
        self.max_batch = int(max_batch)
        if seed is not None:
            self.init(seed)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib_handle is not None:
            _lib_handle.sc_pn_destroy(h)
            self._h = None

    def init(self, seed):
        check(lib().sc_pn_init(self._h, ctypes.c_uint64(seed)))

    def get_params(self):
        import numpy as np
        p = np.empty(pn_num_params(), np.float32)
        check(lib().sc_pn_get_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     p.size))
        return p

    def set_params(self, params):
        import numpy as np
        p = np.ascontiguousarray(params, np.float32)
        check(lib().sc_pn_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     p.size))

    def memoryUsage(self):
        n = ctypes.c_size_t()
        check(lib().sc_pn_memory_usage(self._h, ctypes.byref(n)))
        return n.value

    @staticmethod
    def pool(data, stats, dtype=None, out=None, stream=None):
        """BF16 tiles [N/4096][16][1024] of a CUDA [16][N] coarse channel;
        stats = CUDA float32 tensor {median, std} of its column sums. The
        sample type is the tensor's (dtype, if given, must agree)."""
        torch = _torch()
        codes = {torch.float32: SC_F32, torch.float16: SC_F16, torch.bfloat16: SC_BF16}
        if data.dim() != 2 or not data.is_cuda or data.stride(1) != 1 or data.dtype not in codes:
            raise ValueError("pool data must be a CUDA [16][N] float32/float16/bfloat16 tensor "
                             "with unit channel stride")
        if dtype is not None and dtype != codes[data.dtype]:
            raise ValueError("pool data dtype %s does not match sample type %r" % (data.dtype,
                                                                                   dtype))
        if not stats.is_cuda or stats.dtype != torch.float32 or stats.numel() < 2:
            raise ValueError("stats must be a CUDA float32 tensor {median, std}")
        T, N = data.shape
        if out is None:
            out = torch.empty((N // BLOCK_CHANNELS, TILE_ROWS, TILE_COLUMNS), dtype=torch.bfloat16,
                              device=data.device)
        _check_tiles(out, N // BLOCK_CHANNELS)
        check(lib().sc_pn_pool(data.data_ptr(), int(codes[data.dtype]), T, N, data.stride(0),
                               stats.data_ptr(), out.data_ptr(), _stream_ptr(stream)))
        return out

    def forward(self, tiles, stream=None):
        torch = _torch()
        _check_tiles(tiles)
        B = tiles.shape[0]
        logits = torch.empty(B, dtype=torch.float32, device=tiles.device)
        check(lib().sc_pn_forward(self._h, tiles.data_ptr(), B, logits.data_ptr(),
                                  _stream_ptr(stream)))
        return logits

    def gradient(self, tiles, labels, pos_weight=1.0, stream=None):
        """(grad tensor [num_params], loss tensor [1])."""
        torch = _torch()
        _check_tiles(tiles)
        _check_labels(labels, tiles.shape[0])
        grad = torch.empty(pn_num_params(), dtype=torch.float32, device=tiles.device)
        loss = torch.empty(1, dtype=torch.float32, device=tiles.device)
        check(lib().sc_pn_gradient(self._h, tiles.data_ptr(), labels.data_ptr(), tiles.shape[0],
                                   float(pos_weight), grad.data_ptr(), loss.data_ptr(),
                                   _stream_ptr(stream)))
        return grad, loss

    def train_step(self, tiles, labels, lr, pos_weight=1.0, loss=None, stream=None):
        _check_tiles(tiles)
        _check_labels(labels, tiles.shape[0])
        check(lib().sc_pn_train_step(self._h, tiles.data_ptr(), labels.data_ptr(),
                                     tiles.shape[0], float(lr), float(pos_weight),
                                     loss.data_ptr() if loss is not None else None,
                                     _stream_ptr(stream)))

    def activations(self, B):
        """Views of the last forward's activations and the last gradient call's
        logits/argmax: (h1, h2, h3, logits, argmax) as CUDA tensors."""
        torch = _torch()
        ptrs = [ctypes.c_void_p() for _ in range(5)]
        check(lib().sc_pn_activations(self._h, *[ctypes.byref(p) for p in ptrs]))
        shapes = [(B, 16, 512, 16), (B, 8, 256, 32), (B, 4, 128, 32), (B,), (B,)]
        dtypes = [torch.bfloat16] * 3 + [torch.float32, torch.int32]
        return [_device_view(p.value, sh, dt) for p, sh, dt in zip(ptrs, shapes, dtypes)]


def _device_view(ptr, shape, dtype):
    """A CUDA tensor viewing device memory owned by the library (copied)."""
    torch = _torch()
    out = torch.empty(shape, dtype=dtype, device="cuda")
    n = out.numel() * out.element_size()
    torch.cuda.synchronize()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    if hip.hipMemcpy(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ptr), n, 3) != 0:
        raise SeticoreError(SC_ERR_HIP, "hipMemcpy of activations failed")
    return out


# ---------------------------------------------------------------------------
# Input files (fil_reader.cpp, filterbank_metadata.cpp:53-93, h5_reader.cpp:195-209).

class FilReader:
    """Sigproc filterbank reader: header and inferred metadata on open,
    coarse channels loaded with zero padding and the DC-spike fix."""

    def __init__(self, filename):
        h = ctypes.c_void_p()
        check(lib().sc_fil_open(os.fsencode(filename), ctypes.byref(h)))
        self._h = h
        self._refresh()

    def _refresh(self):
        m = ScFilMeta()
        check(lib().sc_fil_metadata(self._h, ctypes.byref(m)))
        self.meta = m
        for f, _ in ScFilMeta._fields_:
            v = getattr(m, f)
            setattr(self, f, v.decode("utf-8", "replace") if isinstance(v, bytes) else v)
        self.has_dc_spike = bool(m.has_dc_spike)

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().sc_fil_close(self._h)
            self._h = None

    __del__ = close

    def set_coarse_channel_size(self, n):
        check(lib().sc_fil_set_coarse_channel_size(self._h, int(n)))
        self._refresh()

    def getMetadata(self):
        return FilterbankMetadata(fch1=self.fch1, foff=self.foff,
                                  coarse_channel_size=self.coarse_channel_size,
                                  source_name=self.source_name, tstart=self.tstart,
                                  tsamp=self.tsamp, src_raj=self.src_raj, src_dej=self.src_dej,
                                  num_timesteps=self.num_timesteps,
                                  num_channels=self.num_channels,
                                  telescope_id=self.telescope_id,
                                  has_dc_spike=self.has_dc_spike)

    def read_coarse_channel(self, i, rounded_timesteps=None):
        """Host numpy float32 [T_r][coarse_channel_size]."""
        import numpy as np
        T_r = rounded_timesteps or round_up_to_power_of_two(self.num_timesteps)
        out = np.empty((T_r, self.coarse_channel_size), np.float32)
        check(lib().sc_fil_read_coarse_channel(
            self._h, int(i), int(T_r), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def loadCoarseChannel(self, i, buffer, stream=None):
        """Into a FilterbankBuffer (device float32 [T_r][coarse_channel_size])."""
        if buffer.num_channels != self.coarse_channel_size or buffer.row_stride != buffer.num_channels:
            raise SeticoreError(SC_ERR_INVALID_ARGUMENT, "buffer shape mismatch")
        check(lib().sc_fil_load_coarse_channel(self._h, int(i), buffer.num_timesteps,
                                               ctypes.c_void_p(buffer.data.data_ptr()),
                                               _stream_ptr(stream)))


def condition_coarse_channel(data, num_timesteps, has_dc_spike, stream=None):
    """Zero rows num_timesteps.. and apply the DC-spike fix in place on a CUDA
    float32 [T_r][N] tensor (h5_reader.cpp:195-209)."""
    T_r, N = data.shape
    check(lib().sc_condition_coarse_channel(ctypes.c_void_p(data.data_ptr()), int(num_timesteps),
                                            int(T_r), int(N), int(data.stride(0)),
                                            int(bool(has_dc_spike)), _stream_ptr(stream)))
    return data


# ---------------------------------------------------------------------------
# Cadence events (find_events.cpp, event_file_writer.cpp).

class EventFileWriter:
    """EventFileWriter (event_file_writer.cpp:19-69): one Event message per
    cadence event. `metadatas` are FilterbankMetadata objects, one per file."""

    def __init__(self, filename, metadatas):
        self._keep = [m._sc_file_meta() for m in metadatas]  # keeps the name bytes alive
        metas = (ScFileMeta * len(metadatas))(*self._keep)
        h = ctypes.c_void_p()
        check(lib().sc_events_writer_open(os.fsencode(filename), metas, len(metadatas),
                                          ctypes.byref(h)))
        self._h = h
        self.n = len(metadatas)

    def write(self, hits, inputs):
        """hits: one DedopplerHit or None per file; inputs: host float32
        [T][coarse_channel_size] arrays, one per file."""
        import numpy as np
        hp = (ctypes.POINTER(ScHit) * self.n)(*[ctypes.pointer(h) if h is not None else None
                                               for h in hits])
        arrs = [np.ascontiguousarray(a, np.float32) for a in inputs]
        ip = (ctypes.POINTER(ctypes.c_float) * self.n)(
            *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for a in arrs])
        check(lib().sc_events_writer_record(self._h, hp, ip))

    def close(self):
        if getattr(self, "_h", None) is not None:
            check(lib().sc_writer_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def findEvents(input_filenames, output_filename, max_drift, snr_on, snr_off, stream=None):
    """find_events.cpp:66-227 over .fil files; returns the number of events."""
    names = (ctypes.c_char_p * len(input_filenames))(*[os.fsencode(f) for f in input_filenames])
    n = ctypes.c_long()
    check(lib().sc_find_events(names, len(input_filenames), os.fsencode(output_filename),
                               float(max_drift), float(snr_on), float(snr_off), ctypes.byref(n),
                               _stream_ptr(stream)))
    return n.value
