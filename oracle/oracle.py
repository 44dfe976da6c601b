"""ctypes wrapper for the CPU oracle (liboracle.so, built by ``make -C oracle``).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg. The product library never
loads it.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


class OracleHit(ctypes.Structure):
    _fields_ = [("frequency", ctypes.c_double), ("index", ctypes.c_int),
                ("drift_steps", ctypes.c_int), ("drift_rate", ctypes.c_double),
                ("snr", ctypes.c_float), ("coarse_channel", ctypes.c_int),
                ("beam", ctypes.c_int), ("num_timesteps", ctypes.c_int),
                ("power", ctypes.c_float), ("incoherent_power", ctypes.c_float)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        i, d, lng = ctypes.c_int, ctypes.c_double, ctypes.c_long
        L.oracle_basic_taylor_tree.argtypes = [fp, fp, fp, i, i, i]
        L.oracle_basic_taylor_tree.restype = i
        L.oracle_sum_columns.argtypes = [fp, fp, i, i]
        L.oracle_find_top_path_sums.argtypes = [fp, i, i, i, fp, ip, ip]
        L.oracle_host_stats.argtypes = [fp, i, ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(ctypes.c_float)]
        L.oracle_search.argtypes = [fp, i, i, d, d, d, d, d, d, lng, i, i, i, fp, ip, ip,
                                    ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_float), i,
                                    ctypes.POINTER(OracleHit), lng]
        L.oracle_search.restype = lng
        L.oracle_path_offset.argtypes = [i, i, i]
        L.oracle_path_offset.restype = i
        L.oracle_path_sum_fp64.argtypes = [fp, i, i, i, i, i]
        L.oracle_path_sum_fp64.restype = d
        L.oracle_stats_fp64.argtypes = [dp, i, ctypes.POINTER(d), ctypes.POINTER(d),
                                        ctypes.POINTER(d)]
        L.oracle_sum_columns_fp64.argtypes = [fp, i, i, dp]
        L.oracle_drift_params.argtypes = [i, d, d, d, ctypes.POINTER(i), ctypes.POINTER(i),
                                          ctypes.POINTER(i), ctypes.POINTER(d)]
        L.oracle_round_up_pow2.argtypes = [i]
        L.oracle_round_up_pow2.restype = i
        _lib = L
    return _lib


def round_up_pow2(n):
    return lib().oracle_round_up_pow2(int(n))


def basic_taylor_tree(data, drift_block):
    """taylor.cu:116-156 on a [T][N] float32 array; returns the [T][N] sums
    (entries of invalid paths are 0 here, stale in the reference)."""
    data = np.ascontiguousarray(data, np.float32)
    T, N = data.shape
    b1 = np.zeros_like(data)
    b2 = np.zeros_like(data)
    which = lib().oracle_basic_taylor_tree(data, b1, b2, T, N, int(drift_block))
    return b1 if which == 1 else b2


def sum_columns(data):
    data = np.ascontiguousarray(data, np.float32)
    T, N = data.shape
    out = np.empty(N, np.float32)
    lib().oracle_sum_columns(data, out, T, N)
    return out


def find_top_path_sums(path_sums, drift_block, top, blocks, offsets):
    T, N = path_sums.shape
    lib().oracle_find_top_path_sums(np.ascontiguousarray(path_sums, np.float32), T, N,
                                    int(drift_block), top, blocks, offsets)


def host_stats(colsums):
    v = np.array(colsums, np.float32)
    m = ctypes.c_float()
    s = ctypes.c_float()
    lib().oracle_host_stats(v, len(v), ctypes.byref(m), ctypes.byref(s))
    return m.value, s.value


def drift_params(num_timesteps, foff, tsamp, max_drift):
    a, b, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    r = ctypes.c_double()
    lib().oracle_drift_params(int(num_timesteps), float(foff), float(tsamp), float(max_drift),
                              ctypes.byref(a), ctypes.byref(b), ctypes.byref(w), ctypes.byref(r))
    return a.value, b.value, w.value, r.value


class SearchResult:
    def __init__(self, hits, top_sums, top_blocks, top_offsets, colsums, median, std):
        self.hits = hits
        self.top_sums = top_sums
        self.top_blocks = top_blocks
        self.top_offsets = top_offsets
        self.column_sums = colsums
        self.median = median
        self.std = std


def search(data, num_timesteps, foff, tsamp, max_drift, min_drift, snr_threshold, fch1=0.0,
           coarse_channel_size=None, num_coherent_beams=0, beam=-1, coarse_channel=0,
           num_threads=1, top_blocks=None, top_offsets=None):
    """dedoppler.cu:186-320 for one coarse channel held in `data` [T_r][N]."""
    data = np.ascontiguousarray(data, np.float32)
    T, N = data.shape
    if coarse_channel_size is None:
        coarse_channel_size = N
    top = np.zeros(N, np.float32)
    blk = top_blocks if top_blocks is not None else np.zeros(N, np.int32)
    off = top_offsets if top_offsets is not None else np.zeros(N, np.int32)
    cs = np.empty(N, np.float32)
    med = ctypes.c_float()
    sd = ctypes.c_float()
    cap = 1 << 16
    hits = (OracleHit * cap)()
    n = lib().oracle_search(data, int(num_timesteps), N, float(foff), float(tsamp),
                            float(max_drift), float(min_drift), float(snr_threshold),
                            float(fch1), int(coarse_channel_size), int(num_coherent_beams),
                            int(beam), int(coarse_channel), top, blk, off,
                            cs.ctypes.data, ctypes.byref(med), ctypes.byref(sd),
                            int(num_threads), hits, cap)
    if n < 0:
        raise ValueError("window size 0 (max_drift must be > 0)")
    out = [{name: getattr(hits[k], name) for name, _ in OracleHit._fields_}
           for k in range(min(n, cap))]
    return SearchResult(out, top, blk, off, cs, med.value, sd.value)


def path_offset(L, p, t):
    return lib().oracle_path_offset(int(L), int(p), int(t))


def path_sum_fp64(data, chan, drift_block, p):
    data = np.ascontiguousarray(data, np.float32)
    T, N = data.shape
    return lib().oracle_path_sum_fp64(data, T, N, int(chan), int(drift_block), int(p))


def stats_fp64(data):
    """FP64 column sums, exact median, trimmed mean and std over sorted ranks."""
    data = np.ascontiguousarray(data, np.float32)
    T, N = data.shape
    cs = np.empty(N, np.float64)
    lib().oracle_sum_columns_fp64(data, T, N, cs)
    m, mean, sd = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().oracle_stats_fp64(cs, N, ctypes.byref(m), ctypes.byref(mean), ctypes.byref(sd))
    return m.value, mean.value, sd.value
