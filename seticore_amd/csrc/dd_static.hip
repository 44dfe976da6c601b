// Allocation-free dedoppler search: sc_dd_search_static.
//
// The device-statistics search of sc_dd_search (dedoppler.cu:186-320 semantics)
// for callers that cannot allocate memory: no handle, no hipMalloc/hipFree, no
// pinned host buffers. Every intermediate (top keys, column sums, unpacked top
// arrays, per-window candidates, statistics state and histograms) lives in one
// file-scope device array, `scratch`, that the loader places in each GPU's
// memory once per process. Its address differs per GPU and is only known
// through hipGetSymbolAddress; the array's name in host code is a host-side
// stand-in, never passed to a kernel. Window candidates come back to the host
// through a fixed buffer on the caller's stack, chunk by chunk.
//
// Calls that overlap on one GPU race: they share that GPU's array, which keeps
// the last call's contents (the search clears what must start at zero - the top
// keys and the statistics histograms - with a kernel before using them).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/seticore_amd.h"
#include "dd_internal.h"
#include "sc_error.h"

namespace {

constexpr long kMaxChannels = SC_STATIC_MAX_CHANNELS;
constexpr long kAlign = 256;
constexpr long align_up(long n) { return (n + kAlign - 1) / kAlign * kAlign; }

// Layout of the array (byte offsets). One candidate per window and window >= 2.
constexpr long kMaxWindows = (kMaxChannels + 1) / 2 + 1;
constexpr long kOffKey = 0;
constexpr long kOffColsum = align_up(kOffKey + 8 * kMaxChannels);
constexpr long kOffTopSum = align_up(kOffColsum + 4 * kMaxChannels);
constexpr long kOffTopBlk = align_up(kOffTopSum + 4 * kMaxChannels);
constexpr long kOffTopOff = align_up(kOffTopBlk + 4 * kMaxChannels);
constexpr long kOffCand = align_up(kOffTopOff + 4 * kMaxChannels);
constexpr long kOffStatsOut = align_up(kOffCand + 16 * kMaxWindows);
constexpr long kOffStats = align_up(kOffStatsOut + 4 * sizeof(float));
constexpr long kStatsBytes = 64 * 1024;  // statistics state + histograms (checked at run time)
constexpr long kScratchBytes = align_up(kOffStats + kStatsBytes);
static_assert(kScratchBytes == SC_STATIC_SCRATCH_BYTES, "header and layout disagree");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

sc_status fail(sc_status code, const std::string& msg) {
  sc::set_last_error(msg);
  return code;
}

sc_status hip_fail(hipError_t e, const char* where) {
  return fail(SC_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define SC_HIP(call, where)                            \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return hip_fail(_e, where);  \
  } while (0)

// This device's copy of `scratch`, looked up once per device with that device
// current (the caller's current device is the one the search runs on).
sc_status device_scratch(unsigned char** out) {
  int dev = 0;
  SC_HIP(hipGetDevice(&dev), "current device");
  if (dev < 0 || dev >= kMaxDevices) return fail(SC_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
    if (!p) return fail(SC_ERR_HIP, "scratch symbol address is null");
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return SC_OK;
}

int round_up_pow2(int n) {
  int r = 1;
  while (r < n) r *= 2;
  return r;
}

}  // namespace

extern "C" sc_status sc_dd_search_static(int num_timesteps, int rounded_timesteps,
                                         int num_channels, double foff_mhz, double tsamp_s,
                                         const void* data, int dtype, long row_stride,
                                         const sc_meta* meta, int beam, int coarse_channel,
                                         double max_drift, double min_drift,
                                         double snr_threshold, sc_hit* hits, size_t capacity,
                                         size_t* num_hits, void* stream_ptr) {
  // Preconditions of the constructor (dedoppler.cu:83) and of search
  // (dedoppler.cu:191-192), then this entry point's own limits.
  if (num_timesteps <= 1) return fail(SC_ERR_INVALID_ARGUMENT, "num_timesteps must be > 1");
  if (!data) return fail(SC_ERR_INVALID_ARGUMENT, "null data");
  if (!meta) return fail(SC_ERR_INVALID_ARGUMENT, "null metadata");
  if (dtype != SC_F32 && dtype != SC_F16 && dtype != SC_BF16)
    return fail(SC_ERR_INVALID_ARGUMENT, "unknown dtype");
  if (!(tsamp_s > 0) || foff_mhz == 0 || !std::isfinite(foff_mhz) || !std::isfinite(tsamp_s))
    return fail(SC_ERR_INVALID_ARGUMENT, "foff must be nonzero and tsamp positive");
  const int T = round_up_pow2(num_timesteps);
  if (rounded_timesteps != T)
    return fail(SC_ERR_INVALID_ARGUMENT, "input.num_timesteps != rounded_num_timesteps (" +
                                             std::to_string(rounded_timesteps) + " vs " +
                                             std::to_string(T) + ")");
  if (T > 256)
    return fail(SC_ERR_UNSUPPORTED,
                "the allocation-free search handles at most 256 (rounded) time steps");
  const int N = num_channels;
  if (N < 4096 || N > kMaxChannels)
    return fail(SC_ERR_INVALID_ARGUMENT,
                "the allocation-free search handles 4096 to " + std::to_string(kMaxChannels) +
                    " channels (got " + std::to_string(N) + ")");
  if (row_stride < N) return fail(SC_ERR_INVALID_ARGUMENT, "row_stride < num_channels");
  if (!(max_drift > 0) || !std::isfinite(max_drift))
    return fail(SC_ERR_INVALID_ARGUMENT,
                "max_drift must be > 0 (the reference loops forever at 0, dedoppler.cu:275)");
  if (sc::device_stats_scratch_bytes(N) + sc::device_stats_hist_bytes() > (size_t)kStatsBytes)
    return fail(SC_ERR_INTERNAL, "statistics state exceeds its part of the scratch array");

  const int drift_timesteps = T - 1;
  // dedoppler.cu:87, 195-198, 274
  const double resolution = 1e6 * foff_mhz / (drift_timesteps * tsamp_s);
  const sc::DriftRange r = sc::drift_range_of(resolution, drift_timesteps, max_drift);
  if (r.window <= 0) return fail(SC_ERR_INVALID_ARGUMENT, "window size is 0");
  const int nblocks = r.bmax - r.bmin + 1;
  if ((long)nblocks * T > 0x1FFFFFFFL) return fail(SC_ERR_UNSUPPORTED, "drift range too large");
  int lanes = 0, group = 0, prefetch = 0;
  sc::default_tuning(T, &lanes, &group, &prefetch);
  if (!sc::fused_supported(T, lanes, group, prefetch))
    return fail(SC_ERR_UNSUPPORTED, "no compiled default kernel shape for this T_r");
  const bool block_keys = T <= 16 || sc::staged_shape(T, lanes, group, prefetch);

  unsigned char* base = nullptr;
  sc_status s = device_scratch(&base);
  if (s != SC_OK) return s;
  sc::top_key_t* key = (sc::top_key_t*)(base + kOffKey);
  float* colsum = (float*)(base + kOffColsum);
  float* top_sum = (float*)(base + kOffTopSum);
  int* top_blk = (int*)(base + kOffTopBlk);
  int* top_off = (int*)(base + kOffTopOff);
  int4* cand = (int4*)(base + kOffCand);
  float* stats_out = (float*)(base + kOffStatsOut);
  void* stats_state = base + kOffStats;
  unsigned* stats_hist =
      (unsigned*)(base + kOffStats + align_up((long)sc::device_stats_scratch_bytes(N)));

  hipStream_t stream = (hipStream_t)stream_ptr;
  // dedoppler.cu:205 (top sums) and the histograms, by a kernel.
  SC_HIP(sc::launch_reset(key, N, stats_hist, stream), "reset");
  // Tree and top-path tracking over every drift block, with the column sums.
  SC_HIP(sc::launch_fused_tree_top(data, dtype, row_stride, T, N, r.bmin, nblocks, lanes, group,
                                   prefetch, key, colsum, stream, nullptr, nullptr, block_keys),
         "fused tree");
  SC_HIP(sc::launch_device_stats(colsum, N, stats_state, stats_hist, stats_out, stream),
         "device stats");
  SC_HIP(sc::launch_window_merge(key, N, T, r.bmin, r.window, snr_threshold, stats_out, top_sum,
                                 top_blk, top_off, cand, block_keys, stream),
         "window merge");
  const int nwin = (int)(((long)N + r.window - 1) / r.window);
  if (block_keys)
    SC_HIP(sc::launch_candidate_offsets(data, dtype, row_stride, T, N, nwin, cand, top_off,
                                        stream),
           "candidate offsets");

  // {median, std, mean, number of non-finite column sums}
  float stats[4];
  SC_HIP(hipMemcpyAsync(stats, stats_out, sizeof(stats), hipMemcpyDeviceToHost, stream),
         "copy stats");
  SC_HIP(hipStreamSynchronize(stream), "search");
  if (stats[3] != 0.0f)
    return fail(SC_ERR_INVALID_ARGUMENT,
                "input has NaN or infinite samples (" + std::to_string((long)stats[3]) +
                    " non-finite column sums)");
  const float median = stats[0], std_dev = stats[1];

  // Hits in ascending channel order (dedoppler.cu:305-317), from the window
  // candidates copied back a chunk at a time.
  constexpr int kChunk = 1024;
  int4 buf[kChunk];
  size_t total = 0;
  for (int w0 = 0; w0 < nwin; w0 += kChunk) {
    const int n = std::min(kChunk, nwin - w0);
    SC_HIP(hipMemcpyAsync(buf, cand + w0, n * sizeof(int4), hipMemcpyDeviceToHost, stream),
           "copy candidates");
    SC_HIP(hipStreamSynchronize(stream), "copy candidates");
    for (int i = 0; i < n; ++i) {
      const int4 c = buf[i];
      if (c.x < 0) continue;
      if (c.w < 0 && c.y == 0)
        return fail(SC_ERR_UNSUPPORTED,
                    "the threshold is below zero and channel " + std::to_string(c.x) +
                        " has no positive path: its hit takes the path offset of an earlier "
                        "search (dedoppler.cu:305-306), which this search does not keep");
      if (c.w < 0)
        return fail(SC_ERR_INTERNAL, "no path offset reproduces the top sum at channel " +
                                         std::to_string(c.x));
      float cand_sum;
      std::memcpy(&cand_sum, &c.y, sizeof(float));
      const int drift_bins = c.z * drift_timesteps + c.w;
      const double drift_rate = drift_bins * resolution;
      const float snr = (cand_sum - median) / std_dev;
      if (std::fabs(drift_rate) >= min_drift) {
        if (hits && total < capacity)
          hits[total] = sc::make_hit(meta, c.x, drift_bins, drift_rate, snr, beam,
                                     coarse_channel, num_timesteps, cand_sum);
        ++total;
      }
    }
  }
  if (num_hits) *num_hits = total;
  return SC_OK;
}
