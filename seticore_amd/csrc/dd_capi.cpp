// C-ABI implementation of the dedoppler search (host side).
//
// Replaces lacker/seticore's Dedopplerer (dedoppler.h:15-79, dedoppler.cu:79-320).
// The orchestration differs from the reference: the whole drift range is one
// fused launch for T_r <= 16 (the reference launches a tree and a top-path
// kernel per drift block), and the statistics/merge can stay on the GPU. The
// host-exact statistics reproduce dedoppler.cu:241-319 step for step.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/seticore_amd.h"
#include "dd_internal.h"
#include "pn_internal.h"
#include "pn_state.h"
#include "sc_error.h"

namespace sc {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }
}  // namespace sc

namespace {

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

// util.cpp:16-22
int round_up_pow2(int n) {
  int r = 1;
  while (r < n) r *= 2;
  return r;
}

}  // namespace

namespace sc {
// Default fused-kernel shape per T_r: channels per lane and drift blocks per
// LDS window.
void default_tuning(int T, int* lanes, int* group, int* prefetch) {
  // Measured on MI355X: at 16 x 2^20 (132 drift blocks) C=4, G=4 with chunked
  // windows (mode 2) is the fastest T=16 shape; at 32..256 x 2^20 (66 blocks)
  // the LDS-staged one-pass kernel with 8, 16, 8, 32 drift blocks per
  // workgroup (1.29, 2.83, 5.46, 20.1 ms against 2.25 ms for fused32_kernel at
  // 32 and 4.25, 7.83, 16.5, 33.7 ms for the two-pass search); at 256 with one
  // workgroup per offset group (20.1 ms against 21.9 for one workgroup over
  // all four groups at 8 blocks; at 128 the split is slower, 5.58 vs 5.46)
  // and all drift blocks (up to 256) in one chunk: 18.1 vs 18.8 ms at 32.
  // Below 16 time steps: 8 channels per lane, 16 drift blocks per window.
  const bool staged = T >= 32 && T <= 256;
  *lanes = (T == 16 || staged) ? 4 : 8;
  *group = T == 16 ? 4 : T == 64 ? 16 : T == 256 ? 256 : staged ? 8 : 16;
  *prefetch = T == 16 ? 2 : T == 32 ? sc::kStaged32 : T == 256 ? 1 : 0;
}

}  // namespace sc

struct sc_dd {
  int num_timesteps = 0;
  int rounded = 0;
  int num_channels = 0;
  double foff = 0, tsamp = 0;
  bool has_dc_spike = false;
  int drift_timesteps = 0;
  double resolution = 0;
  int stats_mode = SC_STATS_HOST_EXACT;
  int lanes = 4;     // fused kernel: channels per lane
  int group = 8;     // fused kernel: drift blocks per LDS window
  int prefetch = 0;  // fused kernel: register prefetch of the next window
  int device = 0;

  // Device state.
  sc::top_key_t* top_key = nullptr;
  float* colsum = nullptr;
  float* top_sum = nullptr;
  int* top_blk = nullptr;
  int* top_off = nullptr;
  void* stats_scratch = nullptr;
  unsigned* stats_hist = nullptr;  // level-1 and level-2 histograms
  float* stats_out = nullptr;  // {median, std, mean, count}
  int4* cand = nullptr;        // per-window candidates (+ one slot: the statistics)
  size_t cand_capacity = 0;
  float* tree_a = nullptr;     // T_r*N scratch for the staged path and tree API
  float* tree_b = nullptr;
  // Two-pass search (T_r = 32..256): scratch of pass_blocks * T_r * pass_pitch
  // floats, pass_group drift blocks per LDS window in the first pass.
  float* pass_scratch = nullptr;
  long pass_pitch = 0;
  int pass_group = 4;
  int pass_blocks = 0;  // drift blocks per launch pair; 0 = from the size
  int pass_alloc_blocks = 0;
  bool use_two_pass = false;  // T_r = 32: two passes instead of the one-pass kernel
  // Gated search (learned block proposal): channel-tile list and counters.
  int* gate_list = nullptr;    // one entry per channel tile
  int* gate_count = nullptr;   // {tile count, flagged blocks}
  int* h_gate = nullptr;       // pinned mirror of gate_count
  int* gather_idx = nullptr;   // 2*capacity ints
  float* gather_out = nullptr;
  size_t gather_capacity = 0;
  size_t device_bytes = 0;

  // Pinned host mirrors.
  float* h_colsum = nullptr;
  float* h_top_sum = nullptr;
  int* h_top_blk = nullptr;
  int* h_top_off = nullptr;
  int4* h_cand = nullptr;
  float* h_stats = nullptr;

  // Graph replay of the device-statistics search (keyed by input and
  // parameters), captured on a private stream and launched on the caller's.
  struct GraphEntry {
    const void* data;
    int dtype;
    long stride;
    int bmin, nblocks, window;
    double snr;
    int lanes, group, prefetch;
    int pass_group, pass_blocks;
    bool use_two_pass;
    hipGraphExec_t exec;
  };
  bool graphs = true;
  hipStream_t capture_stream = nullptr;
  // Second stream and fork/join events: the statistics run beside the path sums.
  hipStream_t aux_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::vector<GraphEntry> graph_cache;
  // Searches seen once (input pointer and parameters, as a hash): a graph is
  // captured only when the same search comes again, so callers that hand a
  // fresh buffer to every search never pay for captures.
  std::vector<size_t> seen_once;
  static constexpr size_t kMaxGraphs = 256;

  // Optional kernel timing (HIP events on the search stream).
  bool timing = false;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  double kernel_ms = 0.0;
  long kernel_launches = 0;

  float last_median = 0, last_std = 0;
  bool state_valid = false;
  std::vector<sc_hit> hits;
};

namespace {

sc_status dev_alloc(sc_dd* dd, void** p, size_t bytes, const char* what) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? SC_ERR_OUT_OF_MEMORY : SC_ERR_HIP,
                std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
  }
  dd->device_bytes += bytes;
  return SC_OK;
}

sc_status host_alloc(void** p, size_t bytes, const char* what) {
  hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(SC_ERR_OUT_OF_MEMORY, std::string("hipHostMalloc ") + what + ": " +
                                          hipGetErrorString(e));
  }
  return SC_OK;
}

void release(sc_dd* dd) {
  (void)hipFree(dd->top_key);
  (void)hipFree(dd->colsum);
  (void)hipFree(dd->top_sum);
  (void)hipFree(dd->top_blk);
  (void)hipFree(dd->top_off);
  (void)hipFree(dd->stats_scratch);
  (void)hipFree(dd->stats_hist);
  (void)hipFree(dd->stats_out);
  (void)hipFree(dd->cand);
  (void)hipFree(dd->tree_a);
  (void)hipFree(dd->tree_b);
  (void)hipFree(dd->pass_scratch);
  (void)hipFree(dd->gate_list);
  (void)hipFree(dd->gate_count);
  (void)hipHostFree(dd->h_gate);
  (void)hipFree(dd->gather_idx);
  (void)hipFree(dd->gather_out);
  (void)hipHostFree(dd->h_colsum);
  (void)hipHostFree(dd->h_top_sum);
  (void)hipHostFree(dd->h_top_blk);
  (void)hipHostFree(dd->h_top_off);
  (void)hipHostFree(dd->h_cand);
  (void)hipHostFree(dd->h_stats);
  if (dd->ev_start) (void)hipEventDestroy(dd->ev_start);
  if (dd->ev_stop) (void)hipEventDestroy(dd->ev_stop);
  for (auto& g : dd->graph_cache) (void)hipGraphExecDestroy(g.exec);
  dd->graph_cache.clear();
  if (dd->capture_stream) (void)hipStreamDestroy(dd->capture_stream);
  if (dd->aux_stream) (void)hipStreamDestroy(dd->aux_stream);
  if (dd->ev_fork) (void)hipEventDestroy(dd->ev_fork);
  if (dd->ev_join) (void)hipEventDestroy(dd->ev_join);
}

using sc::DriftRange;

DriftRange drift_range(const sc_dd* dd, double max_drift) {
  return sc::drift_range_of(dd->resolution, dd->drift_timesteps, max_drift);
}

sc_status ensure_tree_scratch(sc_dd* dd) {
  if (dd->tree_a) return SC_OK;
  size_t bytes = sizeof(float) * (size_t)dd->rounded * dd->num_channels;
  sc_status s = dev_alloc(dd, (void**)&dd->tree_a, bytes, "tree scratch");
  if (s != SC_OK) return s;
  return dev_alloc(dd, (void**)&dd->tree_b, bytes, "tree scratch");
}

// Drift blocks per launch pair of the two-pass search: as set, or as many as
// keep the scratch within 4.5 GB (at least one, at most the pass-1 group;
// more blocks per launch share each input window across blocks).
int pass_blocks(const sc_dd* dd) {
  if (dd->pass_blocks > 0) return dd->pass_blocks;
  const double per_block = 4.0 * dd->rounded * (double)dd->pass_pitch;
  int nb = (int)(4.5e9 / per_block);
  return std::max(1, std::min(dd->pass_group, nb));
}

sc_status ensure_pass_scratch(sc_dd* dd, int nb) {
  if (dd->pass_scratch && dd->pass_alloc_blocks >= nb) return SC_OK;
  if (dd->pass_scratch) {
    (void)hipFree(dd->pass_scratch);
    dd->device_bytes -= sizeof(float) * (size_t)dd->pass_alloc_blocks * dd->rounded *
                        (size_t)dd->pass_pitch;
    dd->pass_scratch = nullptr;
    dd->pass_alloc_blocks = 0;
  }
  const size_t bytes = sizeof(float) * (size_t)nb * dd->rounded * (size_t)dd->pass_pitch;
  sc_status s = dev_alloc(dd, (void**)&dd->pass_scratch, bytes, "two-pass scratch");
  if (s == SC_OK) dd->pass_alloc_blocks = nb;
  return s;
}

// Path sums of one drift block into a [T_r][N] buffer; returns which buffer.
sc_status run_tree(sc_dd* dd, const void* data, int dtype, long stride, int drift,
                   float** result, hipStream_t stream) {
  sc_status s = ensure_tree_scratch(dd);
  if (s != SC_OK) return s;
  const int T = dd->rounded, N = dd->num_channels;
  SC_HIP(sc::launch_tree_blocks(data, dtype, stride, T, N, drift, dd->tree_a, stream),
         "tree blocks");
  float* src = dd->tree_a;
  float* dst = dd->tree_b;
  for (int L = 32; L <= T; L *= 2) {
    SC_HIP(sc::launch_one_step(src, dst, T, N, L, drift, stream), "one step");
    std::swap(src, dst);
  }
  *result = src;
  return SC_OK;
}

// dedoppler.cu:241-262, same calls on a copy of the column sums.
void host_exact_stats(const float* colsums, int n, float* median_out, float* std_out) {
  std::vector<float> v(colsums, colsums + n);
  float* sums = v.data();
  int mid = n / 2;
  float* end = sums + n;
  std::nth_element(sums, sums + mid, end);
  int first = (int)std::ceil(0.05 * n);
  int last = (int)std::floor(0.95 * n);
  std::nth_element(sums, sums + first, sums + mid - 1);
  std::nth_element(sums + mid + 1, sums + last, end);
  float median = sums[mid];
  float sum = std::accumulate(sums + first, sums + last + 1, 0.0);
  float m = sum / (last + 1 - first);
  float accum = 0.0;
  std::for_each(sums + first, sums + last + 1,
                [&](const float f) { accum += (f - m) * (f - m); });
  float std_dev = std::sqrt(accum / (last + 1 - first));
  *median_out = median;
  *std_out = std_dev;
}

}  // namespace

// DedopplerHit constructor (dedoppler_hit.cpp:11-23).
sc_hit sc::make_hit(const sc_meta* meta, int index, int drift_steps, double drift_rate,
                    float snr, int beam, int coarse_channel, int num_timesteps, float power) {
  sc_hit h;
  h.index = index;
  h.drift_steps = drift_steps;
  h.drift_rate = drift_rate;
  h.snr = snr;
  h.coarse_channel = coarse_channel;
  h.beam = (beam >= 0 && beam < meta->num_coherent_beams) ? beam : -1;
  h.num_timesteps = num_timesteps;
  h.power = power;
  h.incoherent_power = 0.0f;
  int coarse_offset = (int)(coarse_channel * meta->coarse_channel_size);
  int global_index = coarse_offset + index;
  h.frequency = meta->fch1 + global_index * meta->foff;
  return h;
}

namespace {

// dedoppler.cu:264-319 on host arrays.
void host_merge(sc_dd* dd, const DriftRange& r, const sc_meta* meta, int beam,
                int coarse_channel, double min_drift, double snr_threshold, float median,
                float std_dev) {
  const int N = dd->num_channels;
  const float* top = dd->h_top_sum;
  float threshold = snr_threshold * std_dev + median;
  const int window = r.window;
  for (int i = 0; (long)i * window < N; ++i) {
    int cand = -1;
    float cand_sum = threshold;
    for (int j = 0; j < window; ++j) {
      int f = i * window + j;
      if (f >= N) break;
      if (top[f] > cand_sum) {
        cand = f;
        cand_sum = top[f];
      }
    }
    if (cand < 0) continue;
    int wend = std::min(N, cand + window);
    bool larger = false;
    for (int f = std::max(0, cand - window + 1); f < wend; ++f) {
      if (top[f] > cand_sum) {
        larger = true;
        break;
      }
    }
    if (larger) continue;
    int drift_bins = dd->h_top_blk[cand] * dd->drift_timesteps + dd->h_top_off[cand];
    double drift_rate = drift_bins * dd->resolution;
    float snr = (cand_sum - median) / std_dev;
    if (std::fabs(drift_rate) >= min_drift) {
      dd->hits.push_back(sc::make_hit(meta, cand, drift_bins, drift_rate, snr, beam,
                                      coarse_channel, dd->num_timesteps, cand_sum));
    }
  }
}

sc_status validate_input(const sc_dd* dd, const void* data, int dtype, int rounded, int N,
                         long stride) {
  if (!dd) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (!data) return fail(SC_ERR_INVALID_ARGUMENT, "null data");
  if (dtype != SC_F32 && dtype != SC_F16 && dtype != SC_BF16)
    return fail(SC_ERR_INVALID_ARGUMENT, "unknown dtype");
  // dedoppler.cu:191-192
  if (rounded != dd->rounded)
    return fail(SC_ERR_INVALID_ARGUMENT, "input.num_timesteps != rounded_num_timesteps (" +
                                             std::to_string(rounded) + " vs " +
                                             std::to_string(dd->rounded) + ")");
  if (N != dd->num_channels)
    return fail(SC_ERR_INVALID_ARGUMENT, "input.num_channels != num_channels");
  if (stride < N) return fail(SC_ERR_INVALID_ARGUMENT, "row_stride < num_channels");
  return SC_OK;
}

}  // namespace

extern "C" {

const char* sc_last_error_string(void) { return sc::last_error(); }

int sc_abi_version(void) { return SC_ABI_VERSION; }

sc_status sc_dd_create(int num_timesteps, int num_channels, double foff_mhz, double tsamp_s,
                       int has_dc_spike, sc_dd** out) {
  if (!out) return fail(SC_ERR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  // dedoppler.cu:83
  if (num_timesteps <= 1) return fail(SC_ERR_INVALID_ARGUMENT, "num_timesteps must be > 1");
  if (num_channels <= 0) return fail(SC_ERR_INVALID_ARGUMENT, "num_channels must be > 0");
  if (!(tsamp_s > 0) || foff_mhz == 0 || !std::isfinite(foff_mhz))
    return fail(SC_ERR_INVALID_ARGUMENT, "foff must be nonzero and tsamp positive");
  int rounded = round_up_pow2(num_timesteps);
  if (rounded > 4096) return fail(SC_ERR_UNSUPPORTED, "more than 4096 time steps");
  if ((long)rounded * num_channels > (1L << 31))
    return fail(SC_ERR_UNSUPPORTED, "T_r * N above 2^31 samples");

  sc_dd* dd = new sc_dd();
  dd->num_timesteps = num_timesteps;
  dd->rounded = rounded;
  dd->num_channels = num_channels;
  dd->foff = foff_mhz;
  dd->tsamp = tsamp_s;
  dd->has_dc_spike = has_dc_spike != 0;
  dd->drift_timesteps = rounded - 1;
  // dedoppler.cu:87
  dd->resolution = 1e6 * foff_mhz / (dd->drift_timesteps * tsamp_s);
  sc::default_tuning(rounded, &dd->lanes, &dd->group, &dd->prefetch);
  if (rounded <= 256 && !sc::fused_supported(rounded, dd->lanes, dd->group, dd->prefetch)) {
    // Every T_r up to 256 has a compiled one-pass default; never fall back
    // to the staged global-memory path silently.
    delete dd;
    return fail(SC_ERR_UNSUPPORTED, "no compiled default kernel shape for this T_r");
  }
  (void)hipGetDevice(&dd->device);

  const size_t N = num_channels;
  sc_status s = SC_OK;
  if (s == SC_OK) s = dev_alloc(dd, (void**)&dd->top_key, N * sizeof(sc::top_key_t), "top key");
  if (s == SC_OK) s = dev_alloc(dd, (void**)&dd->colsum, N * sizeof(float), "column sums");
  if (s == SC_OK) s = dev_alloc(dd, (void**)&dd->top_sum, N * sizeof(float), "top sums");
  if (s == SC_OK) s = dev_alloc(dd, (void**)&dd->top_blk, N * sizeof(int), "top blocks");
  if (s == SC_OK) s = dev_alloc(dd, (void**)&dd->top_off, N * sizeof(int), "top offsets");
  if (s == SC_OK)
    s = dev_alloc(dd, &dd->stats_scratch, sc::device_stats_scratch_bytes(num_channels),
                  "stats scratch");
  if (s == SC_OK)
    s = dev_alloc(dd, (void**)&dd->stats_hist, sc::device_stats_hist_bytes(), "stats hist");
  if (s == SC_OK) s = dev_alloc(dd, (void**)&dd->stats_out, 4 * sizeof(float), "stats");
  // At most one candidate per window and window >= 2 (max_drift > 0).
  dd->cand_capacity = (N + 1) / 2 + 1;
  if (s == SC_OK)
    s = dev_alloc(dd, (void**)&dd->cand, dd->cand_capacity * sizeof(int4), "candidates");
  if (s == SC_OK) s = host_alloc((void**)&dd->h_colsum, N * sizeof(float), "column sums");
  if (s == SC_OK) s = host_alloc((void**)&dd->h_top_sum, N * sizeof(float), "top sums");
  if (s == SC_OK) s = host_alloc((void**)&dd->h_top_blk, N * sizeof(int), "top blocks");
  if (s == SC_OK) s = host_alloc((void**)&dd->h_top_off, N * sizeof(int), "top offsets");
  if (s == SC_OK)
    s = host_alloc((void**)&dd->h_cand, dd->cand_capacity * sizeof(int4), "candidates");
  if (s == SC_OK) s = host_alloc((void**)&dd->h_stats, 4 * sizeof(float), "stats");
  if (sc::two_pass_supported(rounded, dd->pass_group)) {
    // The one-pass kernels (default up to 256 time steps) need no scratch;
    // sc_dd_set_two_pass allocates the two-pass search's.
    dd->pass_pitch = sc::two_pass_row_pitch(rounded, num_channels);
  } else if (s == SC_OK && rounded > 16) {
    s = ensure_tree_scratch(dd);
  }
  if (s == SC_OK) {
    hipError_t e = hipStreamCreateWithFlags(&dd->aux_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&dd->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&dd->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) s = hip_fail(e, "statistics stream");
  }
  if (s == SC_OK) {
    // Block/offset arrays persist across searches like the reference's
    // (dedoppler.cu:205 only resets the sums); start them at zero.
    hipError_t e = hipMemset(dd->top_blk, 0, N * sizeof(int));
    if (e == hipSuccess) e = hipMemset(dd->top_off, 0, N * sizeof(int));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) s = hip_fail(e, "initialise top arrays");
  }
  if (s != SC_OK) {
    release(dd);
    delete dd;
    return s;
  }
  *out = dd;
  return SC_OK;
}

sc_status sc_dd_destroy(sc_dd* dd) {
  if (!dd) return SC_OK;
  release(dd);
  delete dd;
  return SC_OK;
}

sc_status sc_dd_memory_usage(const sc_dd* dd, size_t* bytes) {
  if (!dd || !bytes) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  *bytes = dd->device_bytes;
  return SC_OK;
}

sc_status sc_dd_set_stats_mode(sc_dd* dd, int mode) {
  if (!dd) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (mode != SC_STATS_HOST_EXACT && mode != SC_STATS_DEVICE)
    return fail(SC_ERR_INVALID_ARGUMENT, "unknown stats mode");
  dd->stats_mode = mode;
  return SC_OK;
}

sc_status sc_dd_set_tuning(sc_dd* dd, int channels_per_lane, int drift_group, int prefetch) {
  if (!dd) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (dd->rounded > 256) return fail(SC_ERR_UNSUPPORTED, "no one-pass kernel above 256 time steps");
  if (!sc::fused_supported(dd->rounded, channels_per_lane, drift_group, prefetch))
    return fail(SC_ERR_UNSUPPORTED, "no fused kernel for this (lanes, group, prefetch) shape");
  dd->lanes = channels_per_lane;
  dd->group = drift_group;
  dd->prefetch = prefetch;
  dd->use_two_pass = false;
  return SC_OK;
}

sc_status sc_dd_set_two_pass(sc_dd* dd, int drift_group, int blocks_per_launch) {
  if (!dd) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (!sc::two_pass_supported(dd->rounded, drift_group))
    return fail(SC_ERR_UNSUPPORTED, "no two-pass kernel for this (T_r, group)");
  if (blocks_per_launch < 0 || blocks_per_launch > 64)
    return fail(SC_ERR_INVALID_ARGUMENT, "blocks_per_launch must be in [0, 64]");
  dd->pass_group = drift_group;
  dd->pass_blocks = blocks_per_launch;
  dd->use_two_pass = true;
  if (!dd->pass_pitch) dd->pass_pitch = sc::two_pass_row_pitch(dd->rounded, dd->num_channels);
  const int nb = pass_blocks(dd);
  if (nb > dd->pass_alloc_blocks) {
    // Cached graphs hold the old scratch address.
    for (auto& g : dd->graph_cache) (void)hipGraphExecDestroy(g.exec);
    dd->graph_cache.clear();
    return ensure_pass_scratch(dd, nb);
  }
  return SC_OK;
}

}  // extern "C"

namespace {

// Reset, then the path sums of every drift block into the top keys, plus the
// column sums (dedoppler.cu:205-227).
//
// With overlap_stats (device statistics, one search at a time): the column
// sums (their own kernel) and the statistics run on the handle's second
// stream while the path-sum kernels run on `stream` (the statistics need only
// the column sums); `stream` waits for them at the end, so the tail is then
// just the window merge.
//
// Searches that end in the device tail carry block keys when the one-pass
// kernel is the T_r <= 16 or the staged one (block_key_search): the tail
// recovers the candidates' offsets.
bool block_key_search(const sc_dd* dd) {
  return dd->rounded <= 16 ||
         (!dd->use_two_pass && sc::staged_shape(dd->rounded, dd->lanes, dd->group, dd->prefetch));
}

sc_status enqueue_path_sums(sc_dd* dd, const void* data, int dtype, long row_stride,
                            const DriftRange& r, hipStream_t stream,
                            bool overlap_stats = false, bool device_tail = false) {
  const int T = dd->rounded, N = dd->num_channels;
  const int nblocks = r.bmax - r.bmin + 1;
  // dedoppler.cu:205 (the reference zeroes the float sums; the key 0 means "no
  // positive path"), plus the statistics histograms.
  SC_HIP(sc::launch_reset(dd->top_key, N, dd->stats_hist, stream), "reset");
  const bool one_pass = (T <= 16 || !dd->use_two_pass) &&
                        sc::fused_supported(T, dd->lanes, dd->group, dd->prefetch);
  // The fork point is right after the reset, but the side stream's launches
  // are issued after the path sums': a captured graph submits its nodes in
  // the order they were captured, and with the statistics first the tree
  // started ~100 us after the reset (after the column sums and most of the
  // statistics).
  if (overlap_stats) SC_HIP(hipEventRecord(dd->ev_fork, stream), "fork");
  if (dd->timing) SC_HIP(hipEventRecord(dd->ev_start, stream), "timing event");
  if (one_pass) {
    SC_HIP(sc::launch_fused_tree_top(data, dtype, row_stride, T, N, r.bmin, nblocks, dd->lanes,
                                     dd->group, dd->prefetch, dd->top_key,
                                     overlap_stats ? nullptr : dd->colsum, stream, nullptr,
                                     nullptr, device_tail && block_key_search(dd)),
           "fused tree");
  } else if (sc::two_pass_supported(T, dd->pass_group)) {
    if (!overlap_stats)
      SC_HIP(sc::launch_colsum(data, dtype, row_stride, T, N, dd->colsum, stream),
             "column sums");
    const int nb = pass_blocks(dd);
    sc_status s = ensure_pass_scratch(dd, nb);
    if (s != SC_OK) return s;
    for (int b = r.bmin; b <= r.bmax; b += nb) {
      SC_HIP(sc::launch_two_pass(data, dtype, row_stride, T, N, r.bmin, b,
                                 std::min(nb, r.bmax - b + 1), dd->pass_group,
                                 dd->pass_scratch, dd->pass_pitch, dd->top_key, stream),
             "two-pass tree");
    }
  } else {
    if (!overlap_stats)
      SC_HIP(sc::launch_colsum(data, dtype, row_stride, T, N, dd->colsum, stream),
             "column sums");
    for (int b = r.bmin; b <= r.bmax; ++b) {
      float* sums = nullptr;
      sc_status s = run_tree(dd, data, dtype, row_stride, b, &sums, stream);
      if (s != SC_OK) return s;
      SC_HIP(sc::launch_top_update(sums, T, N, b, r.bmin, dd->top_key, stream), "top update");
    }
  }
  if (dd->timing) SC_HIP(hipEventRecord(dd->ev_stop, stream), "timing event");
  if (overlap_stats) {
    SC_HIP(hipStreamWaitEvent(dd->aux_stream, dd->ev_fork, 0), "fork");
    SC_HIP(sc::launch_colsum(data, dtype, row_stride, T, N, dd->colsum, dd->aux_stream),
           "column sums");
    SC_HIP(sc::launch_device_stats(dd->colsum, N, dd->stats_scratch, dd->stats_hist,
                                   dd->stats_out, dd->aux_stream),
           "device stats");
    SC_HIP(hipEventRecord(dd->ev_join, dd->aux_stream), "join");
    SC_HIP(hipStreamWaitEvent(stream, dd->ev_join, 0), "join");
  }
  return SC_OK;
}

// Device statistics, window merge (with the candidates' offsets after a
// block-key search of `data`) and the copies of the per-window candidates.
sc_status enqueue_device_tail(sc_dd* dd, const DriftRange& r, double snr_threshold,
                              const void* data, int dtype, long row_stride, hipStream_t stream,
                              int4* h_cand = nullptr, float* h_stats = nullptr,
                              bool stats_done = false, bool gated = false) {
  // The handle's own buffers (a single search): the statistics ride in the
  // candidates' copy, at h_cand[nwin] (one copy instead of two).
  const bool own = !h_cand && !h_stats;
  if (!h_cand) h_cand = dd->h_cand;
  if (!h_stats) h_stats = dd->h_stats;
  const int T = dd->rounded, N = dd->num_channels;
  if (!stats_done)
    SC_HIP(sc::launch_device_stats(dd->colsum, N, dd->stats_scratch, dd->stats_hist,
                                   dd->stats_out, stream),
           "device stats");
  const bool bk = block_key_search(dd);
  SC_HIP(sc::launch_window_merge(dd->top_key, N, T, r.bmin, r.window, snr_threshold,
                                 dd->stats_out, dd->top_sum, dd->top_blk, dd->top_off, dd->cand,
                                 bk, stream, gated, own),
         "window merge");
  const int nwin = (int)(((long)N + r.window - 1) / r.window);
  if (bk)
    SC_HIP(sc::launch_candidate_offsets(data, dtype, row_stride, T, N, nwin, dd->cand,
                                        dd->top_off, stream),
           "candidate offsets");
  SC_HIP(hipMemcpyAsync(h_cand, dd->cand, (nwin + (own ? 1 : 0)) * sizeof(int4),
                        hipMemcpyDeviceToHost, stream),
         "copy candidates");
  if (!own)
    SC_HIP(hipMemcpyAsync(h_stats, dd->stats_out, 4 * sizeof(float), hipMemcpyDeviceToHost,
                          stream), "copy stats");
  return SC_OK;
}

// The cached graph for this search, capturing it on first use.
sc_status search_graph(sc_dd* dd, const void* data, int dtype, long row_stride,
                       const DriftRange& r, double snr_threshold, hipGraphExec_t* out) {
  const int nblocks = r.bmax - r.bmin + 1;
  for (auto& g : dd->graph_cache) {
    if (g.data == data && g.dtype == dtype && g.stride == row_stride && g.bmin == r.bmin &&
        g.nblocks == nblocks && g.window == r.window && g.snr == snr_threshold &&
        g.lanes == dd->lanes && g.group == dd->group && g.prefetch == dd->prefetch &&
        g.pass_group == dd->pass_group && g.pass_blocks == pass_blocks(dd) &&
        g.use_two_pass == dd->use_two_pass) {
      *out = g.exec;
      return SC_OK;
    }
  }
  if (!dd->capture_stream)
    SC_HIP(hipStreamCreateWithFlags(&dd->capture_stream, hipStreamNonBlocking), "stream");
  SC_HIP(hipStreamBeginCapture(dd->capture_stream, hipStreamCaptureModeThreadLocal),
         "begin capture");
  sc_status s =
      enqueue_path_sums(dd, data, dtype, row_stride, r, dd->capture_stream, true, true);
  if (s == SC_OK)
    s = enqueue_device_tail(dd, r, snr_threshold, data, dtype, row_stride, dd->capture_stream,
                            nullptr, nullptr, true);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(dd->capture_stream, &graph);
  if (s != SC_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return s;
  }
  if (e != hipSuccess) return hip_fail(e, "end capture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(e, "instantiate graph");
  if (dd->graph_cache.size() >= sc_dd::kMaxGraphs) {
    (void)hipGraphExecDestroy(dd->graph_cache.front().exec);
    dd->graph_cache.erase(dd->graph_cache.begin());
  }
  dd->graph_cache.push_back({data, dtype, row_stride, r.bmin, nblocks, r.window, snr_threshold,
                             dd->lanes, dd->group, dd->prefetch, dd->pass_group,
                             pass_blocks(dd), dd->use_two_pass, exec});
  *out = exec;
  return SC_OK;
}

// Hits from the per-window candidates (dedoppler.cu:305-317). h_stats holds
// {median, std, mean, number of non-finite column sums}. Fails on input with
// NaN or infinite samples, and if a candidate's offset could not be recovered.
sc_status device_hits(sc_dd* dd, const DriftRange& r, const sc_meta* meta, int beam,
                      int coarse_channel, double min_drift, const int4* h_cand = nullptr,
                      const float* h_stats = nullptr, std::vector<sc_hit>* out = nullptr,
                      int num_timesteps = 0) {
  if (!h_cand) h_cand = dd->h_cand;
  if (!h_stats) h_stats = dd->h_stats;
  if (!out) out = &dd->hits;
  if (num_timesteps <= 0) num_timesteps = dd->num_timesteps;
  const int N = dd->num_channels;
  const int nwin = (int)(((long)N + r.window - 1) / r.window);
  if (h_stats[3] != 0.0f)
    return fail(SC_ERR_INVALID_ARGUMENT,
                "input has NaN or infinite samples (" + std::to_string((long)h_stats[3]) +
                    " non-finite column sums)");
  const float median = h_stats[0], std_dev = h_stats[1];
  dd->last_median = median;
  dd->last_std = std_dev;
  for (int i = 0; i < nwin; ++i) {
    const int4 c = h_cand[i];
    if (c.x < 0) continue;
    if (c.w < 0 && c.y == 0)
      return fail(SC_ERR_UNSUPPORTED,
                  "the threshold is below zero and channel " + std::to_string(c.x) +
                      " has no positive path: its hit takes the path offset of an earlier "
                      "search (dedoppler.cu:305-306), which block-key searches do not keep; "
                      "use SC_STATS_HOST_EXACT");
    if (c.w < 0)
      return fail(SC_ERR_INTERNAL, "no path offset reproduces the top sum at channel " +
                                       std::to_string(c.x));
    float cand_sum;
    std::memcpy(&cand_sum, &c.y, sizeof(float));
    int drift_bins = c.z * dd->drift_timesteps + c.w;
    double drift_rate = drift_bins * dd->resolution;
    float snr = (cand_sum - median) / std_dev;
    if (std::fabs(drift_rate) >= min_drift)
      out->push_back(sc::make_hit(meta, c.x, drift_bins, drift_rate, snr, beam,
                                  coarse_channel, num_timesteps, cand_sum));
  }
  return SC_OK;
}

// Checks of a gated search's shape (16 time steps, whole 4096-channel blocks,
// a network batch that holds them, a compiled fused kernel), and the tile
// list buffers, allocated on first use.
sc_status prepare_gate(sc_dd* dd, const sc_pn* pn) {
  if (!pn) return fail(SC_ERR_INVALID_ARGUMENT, "null proposal network");
  const int T = dd->rounded, N = dd->num_channels;
  if (T != sc::pn::kRows || N % sc::pn::kBlockChannels != 0)
    return fail(SC_ERR_UNSUPPORTED,
                "the gated search takes 16 time steps and a multiple of 4096 channels");
  if (N / sc::pn::kBlockChannels > pn->max_batch)
    return fail(SC_ERR_INVALID_ARGUMENT, "proposal network max_batch below the block count");
  if (!sc::fused_supported(T, dd->lanes, dd->group, dd->prefetch))
    return fail(SC_ERR_UNSUPPORTED, "no fused kernel for the current tuning");
  const int width = sc::fused_tile_width(T, dd->lanes);
  const long ntiles = ((long)N + width - 1) / width;
  if (ntiles > (N + 63) / 64 + 1) return fail(SC_ERR_UNSUPPORTED, "tile list too long");
  if (!dd->gate_list) {
    sc_status s = dev_alloc(dd, (void**)&dd->gate_list, sizeof(int) * (size_t)((N + 63) / 64 + 1),
                            "tile list");
    if (s == SC_OK) s = dev_alloc(dd, (void**)&dd->gate_count, 2 * sizeof(int), "tile count");
    if (s == SC_OK) s = host_alloc((void**)&dd->h_gate, 2 * sizeof(int), "tile count");
    if (s != SC_OK) return s;
  }
  return SC_OK;
}

// The gated search up to the tree: column sums and statistics of the whole
// coarse channel (they do not depend on the search), pooling, the network,
// and the ascending list of channel tiles that overlap a flagged block.
sc_status enqueue_gate_front(sc_dd* dd, sc_pn* pn, const void* data, int dtype, long row_stride,
                             float logit_threshold, hipStream_t stream) {
  const int T = dd->rounded, N = dd->num_channels;
  const int nblk = N / sc::pn::kBlockChannels;
  const int width = sc::fused_tile_width(T, dd->lanes);
  SC_HIP(sc::launch_colsum(data, dtype, row_stride, T, N, dd->colsum, stream), "column sums");
  SC_HIP(sc::launch_device_stats(dd->colsum, N, dd->stats_scratch, dd->stats_hist,
                                 dd->stats_out, stream),
         "device stats");
  SC_HIP(sc::pn::launch_pool(data, dtype, row_stride, nblk, dd->stats_out, pn->z, stream),
         "pool");
  SC_HIP(sc::pn::launch_forward(pn->w, pn->z, nblk, pn->a, pn->logits, nullptr, stream),
         "proposal forward");
  SC_HIP(sc::pn::launch_tile_list(pn->logits, nblk, logit_threshold, N, width, dd->gate_list,
                                  dd->gate_count, dd->gate_count + 1, stream),
         "tile list");
  return SC_OK;
}

// The tree over the listed tiles only; unsearched channels keep key 0 (no
// path), so they never form or suppress a candidate.
sc_status enqueue_gated_tree(sc_dd* dd, const void* data, int dtype, long row_stride,
                             const DriftRange& r, hipStream_t stream) {
  SC_HIP(sc::launch_fused_tree_top(data, dtype, row_stride, dd->rounded, dd->num_channels, r.bmin,
                                   r.bmax - r.bmin + 1, dd->lanes, dd->group, dd->prefetch,
                                   dd->top_key, nullptr, stream, dd->gate_list, dd->gate_count,
                                   true),
         "gated tree");
  return SC_OK;
}

}  // namespace

extern "C" {

sc_status sc_dd_search(sc_dd* dd, const void* data, int dtype, int rounded_timesteps,
                       int num_channels, long row_stride, const sc_meta* meta, int beam,
                       int coarse_channel, double max_drift, double min_drift,
                       double snr_threshold, sc_hit* hits, size_t capacity, size_t* num_hits,
                       void* stream_ptr) {
  sc_status s = validate_input(dd, data, dtype, rounded_timesteps, num_channels, row_stride);
  if (s != SC_OK) return s;
  if (!meta) return fail(SC_ERR_INVALID_ARGUMENT, "null metadata");
  if (!(max_drift > 0) || !std::isfinite(max_drift))
    return fail(SC_ERR_INVALID_ARGUMENT,
                "max_drift must be > 0 (the reference loops forever at 0, dedoppler.cu:275)");
  hipStream_t stream = (hipStream_t)stream_ptr;
  const int T = dd->rounded, N = dd->num_channels;
  const DriftRange r = drift_range(dd, max_drift);
  if (r.window <= 0) return fail(SC_ERR_INVALID_ARGUMENT, "window size is 0");
  const int nblocks = r.bmax - r.bmin + 1;
  // Drift codes must fit the 29 key bits (dd_internal.h).
  if ((long)nblocks * T > 0x1FFFFFFFL) return fail(SC_ERR_UNSUPPORTED, "drift range too large");
  dd->hits.clear();
  dd->state_valid = false;

  // Tiny coarse channels take the host statistics: the device selection
  // assumes 0 < ceil(.05N) <= floor(.95N) (the reference's own nth_element
  // calls have empty ranges below a few channels).
  const bool host_stats = dd->stats_mode == SC_STATS_HOST_EXACT || N < 4096;
  if (host_stats) {
    s = enqueue_path_sums(dd, data, dtype, row_stride, r, stream);
    if (s != SC_OK) return s;
    SC_HIP(sc::launch_unpack_top(dd->top_key, N, T, r.bmin, dd->top_sum, dd->top_blk,
                                 dd->top_off, stream),
           "unpack top");
    // dedoppler.cu:232-240
    SC_HIP(hipMemcpyAsync(dd->h_colsum, dd->colsum, N * sizeof(float), hipMemcpyDeviceToHost,
                          stream), "copy column sums");
    SC_HIP(hipMemcpyAsync(dd->h_top_sum, dd->top_sum, N * sizeof(float), hipMemcpyDeviceToHost,
                          stream), "copy top sums");
    SC_HIP(hipMemcpyAsync(dd->h_top_blk, dd->top_blk, N * sizeof(int), hipMemcpyDeviceToHost,
                          stream), "copy top blocks");
    SC_HIP(hipMemcpyAsync(dd->h_top_off, dd->top_off, N * sizeof(int), hipMemcpyDeviceToHost,
                          stream), "copy top offsets");
    SC_HIP(hipStreamSynchronize(stream), "search");
    // NaN or infinite samples make their column's sum non-finite (bit test:
    // exponent all ones). The key arithmetic assumes finite samples.
    long bad = 0;
    for (int c = 0; c < N; ++c) {
      uint32_t u;
      std::memcpy(&u, &dd->h_colsum[c], sizeof(u));
      bad += (u & 0x7F800000u) == 0x7F800000u;
    }
    if (bad)
      return fail(SC_ERR_INVALID_ARGUMENT, "input has NaN or infinite samples (" +
                                               std::to_string(bad) + " non-finite column sums)");
    float median, std_dev;
    host_exact_stats(dd->h_colsum, N, &median, &std_dev);
    dd->last_median = median;
    dd->last_std = std_dev;
    host_merge(dd, r, meta, beam, coarse_channel, min_drift, snr_threshold, median, std_dev);
  } else {
    bool use_graph = dd->graphs && !dd->timing;
    if (use_graph) {
      size_t h = std::hash<const void*>()(data);
      for (size_t v : {(size_t)dtype, (size_t)row_stride, (size_t)(long)r.bmin,
                       (size_t)(long)r.bmax, (size_t)r.window,
                       std::hash<double>()(snr_threshold), (size_t)dd->lanes,
                       (size_t)dd->group, (size_t)dd->prefetch, (size_t)dd->pass_group,
                       (size_t)pass_blocks(dd), (size_t)dd->use_two_pass})
        h = h * 1000003u ^ v;
      bool cached = false;
      for (const auto& g : dd->graph_cache)
        cached = cached || (g.data == data && g.dtype == dtype && g.stride == row_stride &&
                            g.bmin == r.bmin && g.window == r.window && g.snr == snr_threshold);
      if (!cached) {
        auto it = std::find(dd->seen_once.begin(), dd->seen_once.end(), h);
        if (it == dd->seen_once.end()) {
          use_graph = false;  // first time: launch directly, capture next time
          dd->seen_once.push_back(h);
          if (dd->seen_once.size() > 4 * sc_dd::kMaxGraphs)
            dd->seen_once.erase(dd->seen_once.begin());
        } else {
          dd->seen_once.erase(it);
        }
      }
    }
    if (use_graph) {
      hipGraphExec_t exec = nullptr;
      s = search_graph(dd, data, dtype, row_stride, r, snr_threshold, &exec);
      if (s != SC_OK) return s;
      SC_HIP(hipGraphLaunch(exec, stream), "graph launch");
    } else {
      s = enqueue_path_sums(dd, data, dtype, row_stride, r, stream, true, true);
      if (s == SC_OK)
        s = enqueue_device_tail(dd, r, snr_threshold, data, dtype, row_stride, stream, nullptr,
                                nullptr, true);
      if (s != SC_OK) return s;
    }
    SC_HIP(hipStreamSynchronize(stream), "search");
    // The statistics came with the candidates (enqueue_device_tail).
    const int nwin = (int)(((long)N + r.window - 1) / r.window);
    s = device_hits(dd, r, meta, beam, coarse_channel, min_drift, nullptr,
                    (const float*)(dd->h_cand + nwin));
    if (s != SC_OK) return s;
  }
  if (dd->timing) {
    float ms = 0.0f;
    SC_HIP(hipEventElapsedTime(&ms, dd->ev_start, dd->ev_stop), "timing");
    dd->kernel_ms += ms;
    dd->kernel_launches += 1;
  }
  dd->state_valid = true;
  if (num_hits) *num_hits = dd->hits.size();
  if (hits && capacity) {
    size_t n = std::min(capacity, dd->hits.size());
    std::memcpy(hits, dd->hits.data(), n * sizeof(sc_hit));
  }
  return SC_OK;
}

sc_status sc_dd_search_proposed(sc_dd* dd, sc_pn* pn, const void* data, int dtype,
                                int rounded_timesteps, int num_channels, long row_stride,
                                const sc_meta* meta, int beam, int coarse_channel,
                                double max_drift, double min_drift, double snr_threshold,
                                float logit_threshold, sc_hit* hits, size_t capacity,
                                size_t* num_hits, int* flagged_blocks, void* stream_ptr) {
  sc_status s = validate_input(dd, data, dtype, rounded_timesteps, num_channels, row_stride);
  if (s != SC_OK) return s;
  if (!meta) return fail(SC_ERR_INVALID_ARGUMENT, "null metadata");
  if (!(max_drift > 0) || !std::isfinite(max_drift))
    return fail(SC_ERR_INVALID_ARGUMENT,
                "max_drift must be > 0 (the reference loops forever at 0, dedoppler.cu:275)");
  s = prepare_gate(dd, pn);
  if (s != SC_OK) return s;
  const int T = dd->rounded, N = dd->num_channels;
  hipStream_t stream = (hipStream_t)stream_ptr;
  const DriftRange r = drift_range(dd, max_drift);
  if (r.window <= 0) return fail(SC_ERR_INVALID_ARGUMENT, "window size is 0");
  if ((long)(r.bmax - r.bmin + 1) * T > 0x1FFFFFFFL)
    return fail(SC_ERR_UNSUPPORTED, "drift range too large");
  dd->hits.clear();
  dd->state_valid = false;

  SC_HIP(sc::launch_reset(dd->top_key, N, dd->stats_hist, stream), "reset");
  if (dd->timing) SC_HIP(hipEventRecord(dd->ev_start, stream), "timing event");
  s = enqueue_gate_front(dd, pn, data, dtype, row_stride, logit_threshold, stream);
  if (s == SC_OK) s = enqueue_gated_tree(dd, data, dtype, row_stride, r, stream);
  if (s != SC_OK) return s;
  SC_HIP(sc::launch_window_merge(dd->top_key, N, T, r.bmin, r.window, snr_threshold,
                                 dd->stats_out, dd->top_sum, dd->top_blk, dd->top_off, dd->cand,
                                 true, stream, true),
         "window merge");
  const int nwin = (int)(((long)N + r.window - 1) / r.window);
  SC_HIP(sc::launch_candidate_offsets(data, dtype, row_stride, T, N, nwin, dd->cand, dd->top_off,
                                      stream),
         "candidate offsets");
  if (dd->timing) SC_HIP(hipEventRecord(dd->ev_stop, stream), "timing event");
  SC_HIP(hipMemcpyAsync(dd->h_cand, dd->cand, nwin * sizeof(int4), hipMemcpyDeviceToHost,
                        stream), "copy candidates");
  SC_HIP(hipMemcpyAsync(dd->h_stats, dd->stats_out, 4 * sizeof(float), hipMemcpyDeviceToHost,
                        stream), "copy stats");
  SC_HIP(hipMemcpyAsync(dd->h_gate, dd->gate_count, 2 * sizeof(int), hipMemcpyDeviceToHost,
                        stream), "copy tile count");
  SC_HIP(hipStreamSynchronize(stream), "gated search");
  s = device_hits(dd, r, meta, beam, coarse_channel, min_drift);
  if (s != SC_OK) return s;
  if (dd->timing) {
    float ms = 0.0f;
    SC_HIP(hipEventElapsedTime(&ms, dd->ev_start, dd->ev_stop), "timing");
    dd->kernel_ms += ms;
    dd->kernel_launches += 1;
  }
  dd->state_valid = true;
  if (flagged_blocks) *flagged_blocks = dd->h_gate[1];
  if (num_hits) *num_hits = dd->hits.size();
  if (hits && capacity) {
    size_t n = std::min(capacity, dd->hits.size());
    std::memcpy(hits, dd->hits.data(), n * sizeof(sc_hit));
  }
  return SC_OK;
}

sc_status sc_dd_set_graphs(sc_dd* dd, int enable) {
  if (!dd) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  dd->graphs = enable != 0;
  return SC_OK;
}

sc_status sc_dd_set_timing(sc_dd* dd, int enable) {
  if (!dd) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (enable && !dd->ev_start) {
    SC_HIP(hipEventCreate(&dd->ev_start), "event");
    SC_HIP(hipEventCreate(&dd->ev_stop), "event");
  }
  dd->timing = enable != 0;
  dd->kernel_ms = 0.0;
  dd->kernel_launches = 0;
  return SC_OK;
}

sc_status sc_dd_kernel_time(const sc_dd* dd, double* total_ms, long* searches) {
  if (!dd) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (total_ms) *total_ms = dd->kernel_ms;
  if (searches) *searches = dd->kernel_launches;
  return SC_OK;
}

sc_status sc_dd_copy_hits(const sc_dd* dd, sc_hit* hits, size_t capacity, size_t* num_hits) {
  if (!dd) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (num_hits) *num_hits = dd->hits.size();
  if (hits && capacity) {
    size_t n = std::min(capacity, dd->hits.size());
    std::memcpy(hits, dd->hits.data(), n * sizeof(sc_hit));
  }
  return SC_OK;
}

sc_status sc_dd_copy_state(const sc_dd* dd, float* column_sums, float* top_sums,
                           int* top_drift_blocks, int* top_path_offsets, float* median,
                           float* std_dev) {
  if (!dd) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (!dd->state_valid) return fail(SC_ERR_INVALID_ARGUMENT, "no completed search");
  const size_t N = dd->num_channels;
  if (column_sums) SC_HIP(hipMemcpy(column_sums, dd->colsum, N * 4, hipMemcpyDeviceToHost), "copy");
  if (top_sums) SC_HIP(hipMemcpy(top_sums, dd->top_sum, N * 4, hipMemcpyDeviceToHost), "copy");
  if (top_drift_blocks)
    SC_HIP(hipMemcpy(top_drift_blocks, dd->top_blk, N * 4, hipMemcpyDeviceToHost), "copy");
  if (top_path_offsets)
    SC_HIP(hipMemcpy(top_path_offsets, dd->top_off, N * 4, hipMemcpyDeviceToHost), "copy");
  if (median) *median = dd->last_median;
  if (std_dev) *std_dev = dd->last_std;
  return SC_OK;
}

sc_status sc_dd_taylor_tree(sc_dd* dd, const void* data, int dtype, int rounded_timesteps,
                            int num_channels, long row_stride, int drift_block, float* out,
                            void* stream_ptr) {
  sc_status s = validate_input(dd, data, dtype, rounded_timesteps, num_channels, row_stride);
  if (s != SC_OK) return s;
  if (!out) return fail(SC_ERR_INVALID_ARGUMENT, "null output");
  hipStream_t stream = (hipStream_t)stream_ptr;
  float* sums = nullptr;
  s = run_tree(dd, data, dtype, row_stride, drift_block, &sums, stream);
  if (s != SC_OK) return s;
  SC_HIP(hipMemcpyAsync(out, sums, sizeof(float) * (size_t)dd->rounded * dd->num_channels,
                        hipMemcpyDeviceToDevice, stream),
         "copy tree");
  SC_HIP(hipStreamSynchronize(stream), "tree");
  return SC_OK;
}

sc_status sc_dd_add_incoherent_power(sc_dd* dd, const void* data, int dtype,
                                     int rounded_timesteps, int num_channels, long row_stride,
                                     sc_hit* hits, size_t num_hits, void* stream_ptr) {
  sc_status s = validate_input(dd, data, dtype, rounded_timesteps, num_channels, row_stride);
  if (s != SC_OK) return s;
  if (num_hits && !hits) return fail(SC_ERR_INVALID_ARGUMENT, "null hits");
  hipStream_t stream = (hipStream_t)stream_ptr;
  // dedoppler.cu:144: sort by drift steps (same comparator, same std::sort).
  std::sort(hits, hits + num_hits, [](const sc_hit& a, const sc_hit& b) {
    return a.drift_steps < b.drift_steps;
  });
  if (num_hits > dd->gather_capacity) {
    (void)hipFree(dd->gather_idx);
    (void)hipFree(dd->gather_out);
    dd->gather_idx = nullptr;
    dd->gather_out = nullptr;
    s = dev_alloc(dd, (void**)&dd->gather_idx, 2 * num_hits * sizeof(int), "gather");
    if (s == SC_OK) s = dev_alloc(dd, (void**)&dd->gather_out, num_hits * sizeof(float), "gather");
    if (s != SC_OK) return s;
    dd->gather_capacity = num_hits;
  }
  const int drift_shift = dd->rounded - 1;
  size_t i = 0;
  while (i < num_hits) {
    // dedoppler.cu:156-157: block by float floor division.
    int block = (int)std::floor((float)hits[i].drift_steps / drift_shift);
    size_t j = i;
    std::vector<int> idx;
    while (j < num_hits &&
           (int)std::floor((float)hits[j].drift_steps / drift_shift) == block) {
      int offset = hits[j].drift_steps - block * drift_shift;
      if (offset < 0 || offset >= drift_shift || hits[j].index < 0 ||
          hits[j].index >= dd->num_channels)
        return fail(SC_ERR_INVALID_ARGUMENT, "hit outside the searched range");
      idx.push_back(offset);
      ++j;
    }
    const int n = (int)(j - i);
    for (int k = 0; k < n; ++k) idx.push_back(hits[i + k].index);
    float* sums = nullptr;
    s = run_tree(dd, data, dtype, row_stride, block, &sums, stream);
    if (s != SC_OK) return s;
    SC_HIP(hipMemcpyAsync(dd->gather_idx, idx.data(), 2 * n * sizeof(int),
                          hipMemcpyHostToDevice, stream), "copy indices");
    SC_HIP(sc::launch_gather(sums, dd->num_channels, dd->gather_idx, dd->gather_idx + n, n,
                             dd->gather_out, stream), "gather");
    std::vector<float> vals(n);
    SC_HIP(hipMemcpyAsync(vals.data(), dd->gather_out, n * sizeof(float), hipMemcpyDeviceToHost,
                          stream), "copy values");
    SC_HIP(hipStreamSynchronize(stream), "incoherent power");
    for (int k = 0; k < n; ++k) hits[i + k].incoherent_power = vals[k];
    i = j;
  }
  return SC_OK;
}

sc_status sc_normalize(const float* data, int rounded_timesteps, int num_channels,
                       long row_stride, int out_dtype, float clamp, void* out, double* loc_out,
                       double* std_out, void* stream_ptr) {
  if (!data || !out) return fail(SC_ERR_INVALID_ARGUMENT, "null pointer");
  if (rounded_timesteps <= 0 || num_channels <= 0 || row_stride < num_channels)
    return fail(SC_ERR_INVALID_ARGUMENT, "bad shape");
  if (out_dtype != SC_F32 && out_dtype != SC_F16 && out_dtype != SC_BF16)
    return fail(SC_ERR_INVALID_ARGUMENT, "unknown dtype");
  hipStream_t stream = (hipStream_t)stream_ptr;
  void* scratch = nullptr;
  double* moments = nullptr;
  SC_HIP(hipMalloc(&scratch, sc::normalize_scratch_bytes(rounded_timesteps, num_channels)),
         "normalize scratch");
  hipError_t e = hipMalloc((void**)&moments, 3 * sizeof(double));
  if (e != hipSuccess) {
    (void)hipFree(scratch);
    return hip_fail(e, "normalize moments");
  }
  e = sc::launch_normalize(data, rounded_timesteps, num_channels, row_stride, out_dtype, clamp,
                           out, moments, scratch, stream);
  double h[3] = {0, 0, 0};
  if (e == hipSuccess)
    e = hipMemcpyAsync(h, moments, sizeof(h), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(scratch);
  (void)hipFree(moments);
  if (e != hipSuccess) return hip_fail(e, "normalize");
  if (loc_out) *loc_out = h[2];
  if (std_out) *std_out = h[1];
  return SC_OK;
}

sc_status sc_gen_waterfall(float* out, int num_timesteps, int rounded_timesteps,
                           int num_channels, uint64_t seed, float gain, float taper_frac,
                           float dc_spike, int num_tones, const float* tone_start,
                           const float* tone_slope, const float* tone_amp, void* stream_ptr) {
  if (!out) return fail(SC_ERR_INVALID_ARGUMENT, "null output");
  if (num_timesteps <= 0 || rounded_timesteps < num_timesteps || num_channels <= 0)
    return fail(SC_ERR_INVALID_ARGUMENT, "bad shape");
  if (num_tones < 0 || (num_tones > 0 && (!tone_start || !tone_slope || !tone_amp)))
    return fail(SC_ERR_INVALID_ARGUMENT, "bad tone list");
  hipStream_t stream = (hipStream_t)stream_ptr;
  float* dev = nullptr;
  if (num_tones > 0) {
    SC_HIP(hipMalloc((void**)&dev, 3 * sizeof(float) * num_tones), "tone params");
    std::vector<float> h(3 * (size_t)num_tones);
    std::memcpy(h.data(), tone_start, sizeof(float) * num_tones);
    std::memcpy(h.data() + num_tones, tone_slope, sizeof(float) * num_tones);
    std::memcpy(h.data() + 2 * num_tones, tone_amp, sizeof(float) * num_tones);
    hipError_t e = hipMemcpy(dev, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(dev);
      return hip_fail(e, "tone params");
    }
  }
  hipError_t e = sc::launch_gen_waterfall(out, num_timesteps, rounded_timesteps, num_channels,
                                          seed, gain, taper_frac, dc_spike, num_tones, dev,
                                          dev ? dev + num_tones : nullptr,
                                          dev ? dev + 2 * num_tones : nullptr, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(dev);
  if (e != hipSuccess) return hip_fail(e, "generate waterfall");
  return SC_OK;
}

}  // extern "C"

extern "C" sc_status sc_condition_coarse_channel(float* data, int num_timesteps,
                                                 int rounded_timesteps, int num_channels,
                                                 long row_stride, int has_dc_spike,
                                                 void* stream) {
  if (!data) return fail(SC_ERR_INVALID_ARGUMENT, "null data");
  if (num_timesteps < 0 || rounded_timesteps < num_timesteps || num_channels <= 0 ||
      row_stride < num_channels)
    return fail(SC_ERR_INVALID_ARGUMENT, "bad shape");
  if (has_dc_spike && num_channels < 3)
    return fail(SC_ERR_INVALID_ARGUMENT,
                "coarse channels of fewer than 3 channels cannot take the DC-spike fix");
  SC_HIP(sc::launch_condition(data, num_timesteps, rounded_timesteps, num_channels, row_stride,
                              has_dc_spike, (hipStream_t)stream),
         "condition");
  return SC_OK;
}

// ---------------------------------------------------------------------------
// Batched search: many coarse channels / beams of uneven shapes in one graph.
//
// The reference searches one coarse channel per call (run_dedoppler.cpp:35-51,
// beamforming_pipeline.cpp:145-253). Here every instance's launch sequence
// (reset, path sums, device statistics, window merge, candidate copy) is
// captured once into a single hipGraph and the whole batch runs as one graph
// launch. The path sums of successive instances form one chain; each
// instance's statistics/merge/copy tail hangs off its path sums on a second
// chain, so it can overlap the next instance's path sums. Instances of one
// shape (rounded time steps, channels, foff, tsamp) share search handles:
// two, used alternately (measured +10 % at 16 x 2^20 and +1.6 % on the uneven
// 16..256 batch), 2 x kMaxFusedSet at 16 time steps (runs of consecutive
// instances share one launch of the fused kernel, each on its own handle), or
// one for a shape tuned to the two-pass search (whose scratch is large); an
// instance's path sums wait for the tail of the previous instance on the
// same handle.
// Each instance's candidates land in its own pinned host area.
// A gated batch (with a proposal network) is not captured: it runs as direct
// launches on three streams (enqueue_gated).

struct sc_batch {
  struct Instance {
    sc_instance in;
    DriftRange r;
    int shape = 0;   // index into shapes (the handle the instance runs on)
    size_t cand_offset = 0;  // into h_cand
    std::vector<sc_hit> hits;
  };
  std::vector<Instance> inst;
  std::vector<sc_dd*> shapes;   // handles; a shape's alternates are consecutive
  std::vector<int> alternates;  // per handle: handles of its shape (1 or 2)
  int4* h_cand = nullptr;
  float* h_stats = nullptr;  // 4 per instance
  int* h_gate = nullptr;     // gated batch: {tiles searched, flagged blocks} per instance
  sc_pn* pn = nullptr;       // gated batch: the proposal network (not owned)
  float logit_threshold = 0.0f;
  hipStream_t capture_stream = nullptr;
  hipStream_t tail_stream = nullptr;
  hipGraphExec_t exec = nullptr;
  int nodes = 0;
  // Gated batch: the fronts' stream; events {start, done, then three per
  // instance: front and tree done (used for a run's first instance) and
  // tail done}; per instance the previous instance on its handle (-1 if
  // none).
  hipStream_t front_stream = nullptr;
  std::vector<hipEvent_t> gate_events;
  std::vector<long> prev_on_handle;
  // Gated batch: the network's input tiles, activations and logits for the
  // blocks of one run (net_blocks: the most blocks of any run).
  unsigned short* net_z = nullptr;
  sc::pn::Acts net_a;
  float* net_logits = nullptr;
  long net_blocks = 0;
  // Kernel timing (plain batches): runs launch directly, with events around
  // each path-sum launch.
  bool timing = false;
  double kernel_ms = 0.0;
  long kernel_launches = 0, kernel_channels = 0;
};

namespace {

void batch_release(sc_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  if (b->tail_stream) (void)hipStreamDestroy(b->tail_stream);
  if (b->front_stream) (void)hipStreamDestroy(b->front_stream);
  for (sc_dd* dd : b->shapes) sc_dd_destroy(dd);
  (void)hipHostFree(b->h_cand);
  (void)hipHostFree(b->h_stats);
  (void)hipHostFree(b->h_gate);
  (void)hipFree(b->net_z);
  (void)hipFree(b->net_a.h1);
  (void)hipFree(b->net_a.h2);
  (void)hipFree(b->net_a.h3);
  (void)hipFree(b->net_logits);
  for (hipEvent_t ev : b->gate_events)
    if (ev) (void)hipEventDestroy(ev);
}

// The length of the run of consecutive instances from i that one launch of
// the fused kernel searches (T_r = 16, the default shape): up to
// kMaxFusedSet instances of the same shape and launch parameters, each on its
// own handle; else 1.
size_t set_run(const sc_batch* b, size_t i) {
  const sc_batch::Instance& a = b->inst[i];
  const sc_dd* da = b->shapes[a.shape];
  if (da->rounded != 16 || da->use_two_pass ||
      !sc::fused_set_supported(16, da->lanes, da->group, da->prefetch))
    return 1;
  size_t n = 1;
  for (; n < (size_t)sc::kMaxFusedSet && i + n < b->inst.size(); ++n) {
    const sc_batch::Instance& c = b->inst[i + n];
    const sc_dd* dc = b->shapes[c.shape];
    bool ok = da->rounded == dc->rounded && da->num_channels == dc->num_channels &&
              da->foff == dc->foff && da->tsamp == dc->tsamp && a.in.dtype == c.in.dtype &&
              a.in.row_stride == c.in.row_stride && a.r.bmin == c.r.bmin &&
              a.r.bmax == c.r.bmax && da->lanes == dc->lanes && da->group == dc->group &&
              da->prefetch == dc->prefetch;
    for (size_t m = 0; ok && m < n; ++m) ok = b->inst[i + m].shape != c.shape;
    if (!ok) break;
  }
  return n;
}

// The front of the run [a, e) of a gated batch, one launch per step for all
// its channels: reset, column sums, statistics, pooling, the network over
// all the run's blocks, the tile lists.
sc_status enqueue_front_set(sc_batch* b, size_t a, size_t e, hipStream_t stream) {
  sc::FrontSet fs = {};
  for (size_t j = a; j < e; ++j) {
    sc_dd* dj = b->shapes[b->inst[j].shape];
    fs.x[j - a] = b->inst[j].in.data;
    fs.key[j - a] = dj->top_key;
    fs.hist[j - a] = dj->stats_hist;
    fs.colsum[j - a] = dj->colsum;
    fs.stats_scratch[j - a] = dj->stats_scratch;
    fs.stats_out[j - a] = dj->stats_out;
    fs.list[j - a] = dj->gate_list;
    fs.count[j - a] = dj->gate_count;
  }
  const sc_batch::Instance& b0 = b->inst[a];
  const sc_dd* d0 = b->shapes[b0.shape];
  const int n = (int)(e - a), T = d0->rounded, N = d0->num_channels;
  const int nblk = N / sc::pn::kBlockChannels;
  if ((long)n * nblk > b->net_blocks) return fail(SC_ERR_INTERNAL, "gated batch: run too long");
  SC_HIP(sc::launch_front_set_stats(fs, n, b0.in.dtype, b0.in.row_stride, T, N, stream),
         "front statistics");
  SC_HIP(sc::pn::launch_pool_set(fs, n, b0.in.dtype, b0.in.row_stride, nblk, b->net_z, stream),
         "pool");
  SC_HIP(sc::pn::launch_forward(b->pn->w, b->net_z, n * nblk, b->net_a, b->net_logits, nullptr,
                                stream),
         "proposal forward");
  SC_HIP(sc::pn::launch_tile_list_set(b->net_logits, nblk, b->logit_threshold, N,
                                      sc::fused_tile_width(T, d0->lanes), fs, n, stream),
         "tile lists");
  return SC_OK;
}

// Gated batch run, as direct launches on three streams joined by events
// (device-side waits): the trees on one in instance order, the fronts (reset,
// column sums, statistics, pooling, network, tile lists) on the second, the
// tails (merge, offsets, copies) on the third. Runs of consecutive instances
// of one shape (set_run) share one launch per step of the front and one tree
// launch over their tile lists. The next run's front is enqueued right after
// the current run's tree, so it runs beside it, once the tails of the
// previous instances on its handles are done. (Captured into a graph, the
// same sequence ran each tree only after the previous instance's tail; with
// the tails on the fronts' stream the fronts became the longer chain. One
// front per channel, eight of them took longer beside an eight-channel tree
// than the tree itself.)
sc_status enqueue_gated(sc_batch* b, hipStream_t stream) {
  const size_t n = b->inst.size();
  hipStream_t M = b->capture_stream, F = b->front_stream, Q = b->tail_stream;
  hipEvent_t* ev = b->gate_events.data();
  // Per run (indexed by its first instance): front and tree done; per
  // instance: tail done.
  auto front_ev = [&](size_t a) { return ev[2 + 3 * a]; };
  auto sums_ev = [&](size_t a) { return ev[3 + 3 * a]; };
  auto tail_ev = [&](size_t j) { return ev[4 + 3 * j]; };
  std::vector<std::pair<size_t, size_t>> runs;
  for (size_t a = 0; a < n; a += runs.back().second - a) runs.push_back({a, a + set_run(b, a)});
  SC_HIP(hipEventRecord(ev[0], stream), "batch start");
  SC_HIP(hipStreamWaitEvent(M, ev[0], 0), "batch start");
  SC_HIP(hipStreamWaitEvent(F, ev[0], 0), "batch start");
  SC_HIP(hipStreamWaitEvent(Q, ev[0], 0), "batch start");
  size_t fronts = 0;  // runs whose front is enqueued
  // Enqueues the front of run r if the tails of the instances below `tails`
  // cover the previous instance on each of its handles.
  auto issue_front = [&](size_t r, long tails) -> sc_status {
    if (r >= runs.size() || fronts != r) return SC_OK;
    for (size_t j = runs[r].first; j < runs[r].second; ++j)
      if (b->prev_on_handle[j] >= tails) return SC_OK;
    for (size_t j = runs[r].first; j < runs[r].second; ++j)
      if (b->prev_on_handle[j] >= 0)
        SC_HIP(hipStreamWaitEvent(F, tail_ev((size_t)b->prev_on_handle[j]), 0), "wait tail");
    sc_status st = enqueue_front_set(b, runs[r].first, runs[r].second, F);
    if (st != SC_OK) return st;
    SC_HIP(hipEventRecord(front_ev(runs[r].first), F), "front done");
    ++fronts;
    return SC_OK;
  };
  sc_status s = SC_OK;
  for (size_t r = 0; r < runs.size() && s == SC_OK; ++r) {
    const size_t a = runs[r].first, e = runs[r].second;
    // The run's instances are on distinct handles, so each one's previous
    // instance on its handle comes before the run: the front is issued.
    s = issue_front(r, (long)a);
    if (s != SC_OK) break;
    if (fronts <= r) return fail(SC_ERR_INTERNAL, "gated batch: front not issued");
    SC_HIP(hipStreamWaitEvent(M, front_ev(a), 0), "wait front");
    const sc_batch::Instance& b0 = b->inst[a];
    sc_dd* d0 = b->shapes[b0.shape];
    if (e - a == 1) {
      s = enqueue_gated_tree(d0, b0.in.data, b0.in.dtype, b0.in.row_stride, b0.r, M);
      if (s != SC_OK) break;
    } else {
      sc::FusedSet set = {};
      for (size_t j = a; j < e; ++j) {
        const sc_dd* dj = b->shapes[b->inst[j].shape];
        set.x[j - a] = b->inst[j].in.data;
        set.key[j - a] = dj->top_key;
        set.list[j - a] = dj->gate_list;
        set.count[j - a] = dj->gate_count;
      }
      SC_HIP(sc::launch_fused16_set(set, (int)(e - a), b0.in.dtype, b0.in.row_stride, 16,
                                    d0->num_channels, b0.r.bmin, b0.r.bmax - b0.r.bmin + 1,
                                    d0->lanes, d0->group, d0->prefetch, true, M, true),
             "gated set tree");
    }
    SC_HIP(hipEventRecord(sums_ev(a), M), "tree done");
    s = issue_front(r + 1, (long)a);
    if (s != SC_OK) break;
    SC_HIP(hipStreamWaitEvent(Q, sums_ev(a), 0), "wait tree");
    if (e - a > 1) {
      // The run's window merges (channels without a key skipped) and
      // candidate offsets, one launch each.
      sc::TailSet ts = {};
      for (size_t j = a; j < e; ++j) {
        const sc_batch::Instance& bj = b->inst[j];
        sc_dd* dj = b->shapes[bj.shape];
        ts.x[j - a] = bj.in.data;
        ts.key[j - a] = dj->top_key;
        ts.stats[j - a] = dj->stats_out;
        ts.top_sum[j - a] = dj->top_sum;
        ts.top_blk[j - a] = dj->top_blk;
        ts.top_off[j - a] = dj->top_off;
        ts.cand[j - a] = dj->cand;
        ts.window[j - a] = bj.r.window;
        ts.snr[j - a] = bj.in.snr_threshold;
      }
      SC_HIP(sc::launch_window_merge_set(ts, (int)(e - a), d0->num_channels, 16, b0.r.bmin, Q,
                                         true),
             "window merges");
      SC_HIP(sc::launch_candidate_offsets_set(ts, (int)(e - a), b0.in.dtype, b0.in.row_stride,
                                              d0->num_channels, Q),
             "candidate offsets");
    }
    for (size_t i = a; i < e; ++i) {
      sc_batch::Instance& bi = b->inst[i];
      sc_dd* dd = b->shapes[bi.shape];
      if (e - a > 1) {
        const int nwin = (int)(((long)dd->num_channels + bi.r.window - 1) / bi.r.window);
        SC_HIP(hipMemcpyAsync(b->h_cand + bi.cand_offset, dd->cand, nwin * sizeof(int4),
                              hipMemcpyDeviceToHost, Q),
               "copy candidates");
        SC_HIP(hipMemcpyAsync(b->h_stats + 4 * i, dd->stats_out, 4 * sizeof(float),
                              hipMemcpyDeviceToHost, Q),
               "copy stats");
      } else {
        s = enqueue_device_tail(dd, bi.r, bi.in.snr_threshold, bi.in.data, bi.in.dtype,
                                bi.in.row_stride, Q, b->h_cand + bi.cand_offset,
                                b->h_stats + 4 * i, true, true);
      }
      if (s != SC_OK) break;
      SC_HIP(hipMemcpyAsync(b->h_gate + 2 * i, dd->gate_count, 2 * sizeof(int),
                            hipMemcpyDeviceToHost, Q),
             "copy tile count");
      SC_HIP(hipEventRecord(tail_ev(i), Q), "tail done");
    }
  }
  if (s != SC_OK) return s;
  // Each run's tails follow its tree, and its front precedes it: the tails'
  // stream ends last.
  SC_HIP(hipEventRecord(ev[1], Q), "batch done");
  SC_HIP(hipStreamWaitEvent(stream, ev[1], 0), "batch done");
  return SC_OK;
}

struct BatchTimer {
  hipEvent_t start, stop;
  int channels;
};

// The plain batch's launch sequence, into streams M (path sums) and Q
// (statistics, merge, copies), captured or launched directly: every
// instance's tail hangs off its path sums on Q; an instance's path sums wait
// for the tail of the previous instance on the same handle. Runs of
// consecutive instances of one 16-step shape share one launch of the fused
// kernel. With `timers`, a pair of timing events brackets each path-sum
// launch (with its channel count). new_event creates an event the caller
// destroys; *tail_done receives the last tail event, *err the first HIP error.
template <typename NewEvent>
sc_status enqueue_plain(sc_batch* b, hipStream_t M, hipStream_t Q, NewEvent&& new_event,
                        std::vector<BatchTimer>* timers,
                        hipEvent_t* tail_done_out, hipError_t* err) {
  hipError_t& e = *err;
  sc_status s = SC_OK;
  hipEvent_t tail_done = nullptr;
  // last_tail[h]: the tail event of the previous instance on handle h.
  std::vector<hipEvent_t> last_tail(b->shapes.size(), nullptr);
  // The tails of run [i, i + n), after its path sums; a set run's statistics
  // (one launch per pass), window merges and candidate offsets in one launch
  // each for all its channels.
  auto enqueue_tails = [&](size_t i, size_t n, hipEvent_t sums) {
    e = hipStreamWaitEvent(Q, sums, 0);
    if (e == hipSuccess && n > 1) {
      sc::FrontSet fs = {};
      for (size_t j = 0; j < n; ++j) {
        sc_dd* dj = b->shapes[b->inst[i + j].shape];
        fs.colsum[j] = dj->colsum;
        fs.hist[j] = dj->stats_hist;
        fs.stats_scratch[j] = dj->stats_scratch;
        fs.stats_out[j] = dj->stats_out;
      }
      const sc_batch::Instance& b0 = b->inst[i];
      const int N = b->shapes[b0.shape]->num_channels;
      e = sc::launch_device_stats_set(fs, (int)n, N, Q);
      // Window merges and candidate offsets (block keys), one launch each.
      sc::TailSet ts = {};
      for (size_t j = 0; j < n; ++j) {
        const sc_batch::Instance& bj = b->inst[i + j];
        sc_dd* dj = b->shapes[bj.shape];
        ts.x[j] = bj.in.data;
        ts.key[j] = dj->top_key;
        ts.stats[j] = dj->stats_out;
        ts.top_sum[j] = dj->top_sum;
        ts.top_blk[j] = dj->top_blk;
        ts.top_off[j] = dj->top_off;
        ts.cand[j] = dj->cand;
        ts.window[j] = bj.r.window;
        ts.snr[j] = bj.in.snr_threshold;
      }
      if (e == hipSuccess) e = sc::launch_window_merge_set(ts, (int)n, N, 16, b0.r.bmin, Q);
      if (e == hipSuccess)
        e = sc::launch_candidate_offsets_set(ts, (int)n, b0.in.dtype, b0.in.row_stride, N, Q);
    }
    for (size_t j = i; j < i + n && s == SC_OK && e == hipSuccess; ++j) {
      sc_batch::Instance& bj = b->inst[j];
      sc_dd* dj = b->shapes[bj.shape];
      if (n > 1) {
        // The candidates and statistics to the host (merged above).
        const int nwin = (int)(((long)dj->num_channels + bj.r.window - 1) / bj.r.window);
        e = hipMemcpyAsync(b->h_cand + bj.cand_offset, dj->cand, nwin * sizeof(int4),
                           hipMemcpyDeviceToHost, Q);
        if (e == hipSuccess)
          e = hipMemcpyAsync(b->h_stats + 4 * j, dj->stats_out, 4 * sizeof(float),
                             hipMemcpyDeviceToHost, Q);
        if (e != hipSuccess) break;
      } else {
        s = enqueue_device_tail(dj, bj.r, bj.in.snr_threshold, bj.in.data, bj.in.dtype,
                                bj.in.row_stride, Q, b->h_cand + bj.cand_offset,
                                b->h_stats + 4 * j);
      }
      if (s != SC_OK) break;
      hipEvent_t tail = nullptr;
      e = new_event(&tail);
      if (e == hipSuccess) e = hipEventRecord(tail, Q);
      last_tail[bj.shape] = tail;
      tail_done = tail;
    }
  };
  // A run's tails are enqueued after the next run's path sums (unless that
  // run needs one of their handles): a captured graph submits its nodes in
  // capture order, so the next tree then starts right after the previous
  // one instead of after the previous run's tails were submitted.
  size_t pend_i = 0, pend_n = 0;
  hipEvent_t pend_sums = nullptr;
  for (size_t i = 0; i < b->inst.size() && s == SC_OK && e == hipSuccess;) {
    const size_t n = set_run(b, i);
    bool shares = false;
    for (size_t j = i; j < i + n; ++j)
      for (size_t k = pend_i; k < pend_i + pend_n; ++k)
        shares = shares || b->inst[j].shape == b->inst[k].shape;
    if (pend_n && shares) {
      enqueue_tails(pend_i, pend_n, pend_sums);
      pend_n = 0;
      if (s != SC_OK || e != hipSuccess) break;
    }
    for (size_t j = i; j < i + n && e == hipSuccess; ++j)
      if (last_tail[b->inst[j].shape])
        e = hipStreamWaitEvent(M, last_tail[b->inst[j].shape], 0);
    if (e == hipSuccess && timers) {
      hipEvent_t t0 = nullptr, t1 = nullptr;
      e = hipEventCreate(&t0);
      if (e == hipSuccess) e = hipEventCreate(&t1);
      if (e == hipSuccess) timers->push_back({t0, t1, (int)n});
    }
    if (e != hipSuccess) break;
    if (n == 1) {
      sc_batch::Instance& bi = b->inst[i];
      if (timers) e = hipEventRecord(timers->back().start, M);  // (bracket includes the reset)
      if (e == hipSuccess)
        s = enqueue_path_sums(b->shapes[bi.shape], bi.in.data, bi.in.dtype, bi.in.row_stride,
                              bi.r, M, false, true);
    } else {
      // dedoppler.cu:205 per channel, then the set launch (column sums
      // written by the workgroups that hold drift block 0).
      sc::FusedSet set = {};
      for (size_t j = 0; j < n && e == hipSuccess; ++j) {
        sc_batch::Instance& bj = b->inst[i + j];
        sc_dd* dj = b->shapes[bj.shape];
        e = sc::launch_reset(dj->top_key, dj->num_channels, dj->stats_hist, M);
        set.x[j] = bj.in.data;
        set.key[j] = dj->top_key;
        set.colsum[j] = dj->colsum;
      }
      const sc_batch::Instance& b0 = b->inst[i];
      const sc_dd* d0 = b->shapes[b0.shape];
      if (e == hipSuccess && timers) e = hipEventRecord(timers->back().start, M);
      if (e == hipSuccess)
        e = sc::launch_fused16_set(set, (int)n, b0.in.dtype, b0.in.row_stride, 16,
                                   d0->num_channels, b0.r.bmin, b0.r.bmax - b0.r.bmin + 1,
                                   d0->lanes, d0->group, d0->prefetch, true,
                                   M);
      if (e != hipSuccess) s = hip_fail(e, "set launch");
    }
    if (s == SC_OK && e == hipSuccess && timers) e = hipEventRecord(timers->back().stop, M);
    if (s != SC_OK) break;
    hipEvent_t sums = nullptr;
    e = new_event(&sums);
    if (e == hipSuccess) e = hipEventRecord(sums, M);
    if (e != hipSuccess) break;
    if (pend_n) enqueue_tails(pend_i, pend_n, pend_sums);
    pend_i = i;
    pend_n = n;
    pend_sums = sums;
    i += n;
  }
  if (pend_n && s == SC_OK && e == hipSuccess) enqueue_tails(pend_i, pend_n, pend_sums);
  *tail_done_out = tail_done;
  if (e != hipSuccess && s == SC_OK) s = hip_fail(e, "batch launches");
  return s;
}

// A plain batch run as direct launches with timing events (sc_batch_set_timing):
// the same sequence as the captured graph, on the same two streams.
sc_status run_plain_timed(sc_batch* b, hipStream_t stream) {
  std::vector<hipEvent_t> events;
  auto new_event = [&](hipEvent_t* ev) {
    hipError_t e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e == hipSuccess) events.push_back(*ev);
    return e;
  };
  std::vector<BatchTimer> timers;
  hipEvent_t start = nullptr, tail_done = nullptr;
  hipError_t e = new_event(&start);
  if (e == hipSuccess) e = hipEventRecord(start, stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->capture_stream, start, 0);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->tail_stream, start, 0);
  sc_status s = e == hipSuccess ? enqueue_plain(b, b->capture_stream, b->tail_stream, new_event,
                                                &timers, &tail_done, &e)
                                : hip_fail(e, "batch start");
  // The last tail follows the last path sums: waiting for it joins both streams.
  if (s == SC_OK && tail_done) {
    e = hipStreamWaitEvent(stream, tail_done, 0);
    if (e != hipSuccess) s = hip_fail(e, "batch done");
  }
  (void)hipStreamSynchronize(b->capture_stream);
  (void)hipStreamSynchronize(b->tail_stream);
  (void)hipStreamSynchronize(stream);
  for (const BatchTimer& t : timers) {
    float ms = 0.0f;
    if (s == SC_OK) {
      e = hipEventElapsedTime(&ms, t.start, t.stop);
      if (e != hipSuccess) s = hip_fail(e, "kernel timing");
    }
    b->kernel_ms += ms;
    b->kernel_launches += 1;
    b->kernel_channels += t.channels;
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
  return s;
}

sc_status batch_create(const sc_instance* instances, int count, const sc_tuning* tuning,
                       int num_tuning, sc_pn* pn, float logit_threshold, sc_batch** out) {
  if (!out) return fail(SC_ERR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  if (!instances || count <= 0) return fail(SC_ERR_INVALID_ARGUMENT, "empty batch");
  if (num_tuning > 0 && !tuning) return fail(SC_ERR_INVALID_ARGUMENT, "null tuning table");
  if (pn && !std::isfinite(logit_threshold) && !(logit_threshold < 0))
    return fail(SC_ERR_INVALID_ARGUMENT, "logit_threshold must not be NaN or +inf");
  sc_batch* b = new sc_batch();
  b->pn = pn;
  b->logit_threshold = logit_threshold;
  sc_status s = SC_OK;
  size_t total_windows = 0;
  for (int i = 0; i < count && s == SC_OK; ++i) {
    const sc_instance& in = instances[i];
    sc_batch::Instance bi;
    bi.in = in;
    if (!in.data || in.num_channels < 4096 || in.num_timesteps <= 1 ||
        !(in.max_drift > 0) || !std::isfinite(in.max_drift) || in.row_stride < in.num_channels) {
      s = fail(SC_ERR_INVALID_ARGUMENT,
               "instance " + std::to_string(i) +
                   ": needs data, >= 4096 channels, > 1 time step, max_drift > 0");
      break;
    }
    const int T = round_up_pow2(in.num_timesteps);
    int shape = -1;
    for (size_t k = 0; k < b->shapes.size(); ++k) {
      const sc_dd* dd = b->shapes[k];
      if (dd->rounded == T && dd->num_channels == in.num_channels && dd->foff == in.foff &&
          dd->tsamp == in.tsamp) {
        shape = (int)k;
        break;
      }
    }
    if (shape < 0) {
      shape = (int)b->shapes.size();
      // Two handles per shape, used alternately; at 16 time steps two per
      // channel of a set launch (runs of consecutive instances of the shape
      // share one launch of the fused kernel: enqueue_plain, enqueue_gated).
      int copies = T == 16 ? 2 * sc::kMaxFusedSet : 2;
      for (int k = 0; k < num_tuning; ++k)
        if (tuning[k].rounded_timesteps == T && tuning[k].num_channels == in.num_channels &&
            T > 16 && tuning[k].pass_group != 0)
          copies = 1;
      for (int c = 0; c < copies && s == SC_OK; ++c) {
        sc_dd* dd = nullptr;
        s = sc_dd_create(in.num_timesteps, in.num_channels, in.foff, in.tsamp, 0, &dd);
        if (s != SC_OK) break;
        dd->stats_mode = SC_STATS_DEVICE;
        b->shapes.push_back(dd);
        b->alternates.push_back(copies);
        for (int k = 0; k < num_tuning; ++k) {
          const sc_tuning& t = tuning[k];
          if (t.rounded_timesteps != T || t.num_channels != in.num_channels) continue;
          if (T <= 16 || t.pass_group == 0)
            s = sc_dd_set_tuning(dd, t.channels_per_lane, t.drift_group, t.prefetch);
          else if (sc::two_pass_supported(T, t.pass_group))
            s = sc_dd_set_two_pass(dd, t.pass_group, t.pass_blocks);
          break;
        }
      }
      if (s != SC_OK) break;
    }
    // Alternate over the shape's handles in instance order.
    int used = 0;
    for (const auto& prev : b->inst)
      if (prev.shape >= shape && prev.shape < shape + b->alternates[shape]) ++used;
    shape += used % b->alternates[shape];
    bi.shape = shape;
    bi.r = drift_range(b->shapes[shape], in.max_drift);
    if (bi.r.window <= 0 || (long)(bi.r.bmax - bi.r.bmin + 1) * T > 0x1FFFFFFFL) {
      s = fail(SC_ERR_UNSUPPORTED, "instance " + std::to_string(i) + ": drift range");
      break;
    }
    if (pn) {
      s = prepare_gate(b->shapes[shape], pn);
      if (s != SC_OK) {
        s = fail(s, "instance " + std::to_string(i) + ": " + sc::last_error());
        break;
      }
    }
    bi.cand_offset = total_windows;
    total_windows += ((size_t)in.num_channels + bi.r.window - 1) / bi.r.window;
    b->inst.push_back(bi);
  }
  if (s == SC_OK) s = host_alloc((void**)&b->h_cand, total_windows * sizeof(int4), "candidates");
  if (s == SC_OK)
    s = host_alloc((void**)&b->h_stats, 4 * sizeof(float) * (size_t)count, "statistics");
  if (s == SC_OK && pn)
    s = host_alloc((void**)&b->h_gate, 2 * sizeof(int) * (size_t)count, "tile counts");
  if (s == SC_OK) {
    hipError_t e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->tail_stream, hipStreamNonBlocking);
    if (e == hipSuccess && pn)
      e = hipStreamCreateWithFlags(&b->front_stream, hipStreamNonBlocking);
    if (e != hipSuccess) s = hip_fail(e, "stream");
  }
  std::vector<hipEvent_t> events;
  auto new_event = [&](hipEvent_t* ev) {
    hipError_t e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e == hipSuccess) events.push_back(*ev);
    return e;
  };
  if (s == SC_OK && pn) {
    // Gated batch: run as direct launches (enqueue_gated); its events and the
    // previous instance on each handle are set up here.
    b->prev_on_handle.assign(b->inst.size(), -1);
    std::vector<long> last(b->shapes.size(), -1);
    for (size_t j = 0; j < b->inst.size(); ++j) {
      b->prev_on_handle[j] = last[b->inst[j].shape];
      last[b->inst[j].shape] = (long)j;
    }
    b->gate_events.assign(2 + 3 * b->inst.size(), nullptr);
    for (hipEvent_t& ev : b->gate_events) {
      hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e != hipSuccess) {
        ev = nullptr;
        s = hip_fail(e, "event");
        break;
      }
    }
    // The network's buffers for the blocks of the longest run (448 KiB per
    // block: tiles and three activation maps, BF16).
    for (size_t a = 0, len = 0; a < b->inst.size(); a += len) {
      len = set_run(b, a);
      const long blocks =
          (long)len * (b->shapes[b->inst[a].shape]->num_channels / sc::pn::kBlockChannels);
      b->net_blocks = std::max(b->net_blocks, blocks);
    }
    const size_t B = (size_t)b->net_blocks;
    struct {
      void** p;
      size_t bytes;
    } bufs[] = {{(void**)&b->net_z, B * sc::pn::kRows * sc::pn::kPooled * 2},
                {(void**)&b->net_a.h1, B * sc::pn::kActElems1 * 2},
                {(void**)&b->net_a.h2, B * sc::pn::kActElems2 * 2},
                {(void**)&b->net_a.h3, B * sc::pn::kActElems3 * 2},
                {(void**)&b->net_logits, B * sizeof(float)}};
    for (auto& buf : bufs) {
      if (s != SC_OK) break;
      hipError_t e = hipMalloc(buf.p, buf.bytes);
      if (e != hipSuccess) {
        *buf.p = nullptr;
        s = e == hipErrorOutOfMemory ? fail(SC_ERR_OUT_OF_MEMORY, "proposal network buffers")
                                     : hip_fail(e, "proposal network buffers");
      }
    }
  }
  if (s == SC_OK && !pn) {
    hipError_t e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) s = hip_fail(e, "begin capture");
    hipEvent_t tail_done = nullptr;
    if (e == hipSuccess)
      s = enqueue_plain(b, b->capture_stream, b->tail_stream, new_event, nullptr, &tail_done, &e);
    // Join the tail chain back into the capturing stream.
    if (e == hipSuccess && s == SC_OK && tail_done)
      e = hipStreamWaitEvent(b->capture_stream, tail_done, 0);
    if (e != hipSuccess && s == SC_OK) s = hip_fail(e, "capture dependencies");
    hipGraph_t graph = nullptr;
    {
      hipError_t e2 = hipStreamEndCapture(b->capture_stream, &graph);
      if (s == SC_OK && e2 != hipSuccess) s = hip_fail(e2, "end capture");
    }
    if (s == SC_OK) {
      size_t n = 0;
      (void)hipGraphGetNodes(graph, nullptr, &n);
      b->nodes = (int)n;
      e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) s = hip_fail(e, "instantiate graph");
    }
    if (graph) (void)hipGraphDestroy(graph);
  }
  for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
  if (s != SC_OK) {
    batch_release(b);
    delete b;
    return s;
  }
  *out = b;
  return SC_OK;
}

}  // namespace

extern "C" {

sc_status sc_batch_create(const sc_instance* instances, int count, const sc_tuning* tuning,
                          int num_tuning, sc_batch** out) {
  return batch_create(instances, count, tuning, num_tuning, nullptr, 0.0f, out);
}

sc_status sc_batch_create_proposed(const sc_instance* instances, int count,
                                   const sc_tuning* tuning, int num_tuning, sc_pn* pn,
                                   float logit_threshold, sc_batch** out) {
  if (!pn) return fail(SC_ERR_INVALID_ARGUMENT, "null proposal network");
  return batch_create(instances, count, tuning, num_tuning, pn, logit_threshold, out);
}

sc_status sc_batch_run(sc_batch* b, void* stream_ptr) {
  if (!b) return fail(SC_ERR_INVALID_ARGUMENT, "null batch");
  hipStream_t stream = (hipStream_t)stream_ptr;
  if (!b->pn && b->timing) {
    const sc_status s = run_plain_timed(b, stream);
    if (s != SC_OK) return s;
  } else if (b->pn) {
    const sc_status s = enqueue_gated(b, stream);
    if (s != SC_OK) {
      // Let what was enqueued finish before the caller sees the error.
      (void)hipStreamSynchronize(b->capture_stream);
      (void)hipStreamSynchronize(b->front_stream);
      (void)hipStreamSynchronize(b->tail_stream);
      return s;
    }
  } else {
    SC_HIP(hipGraphLaunch(b->exec, stream), "batch graph launch");
  }
  SC_HIP(hipStreamSynchronize(stream), "batch");
  for (size_t i = 0; i < b->inst.size(); ++i) {
    sc_batch::Instance& bi = b->inst[i];
    bi.hits.clear();
    const sc_status s =
        device_hits(b->shapes[bi.shape], bi.r, &bi.in.meta, bi.in.beam, bi.in.coarse_channel,
                    bi.in.min_drift, b->h_cand + bi.cand_offset, b->h_stats + 4 * i, &bi.hits,
                    bi.in.num_timesteps);
    if (s != SC_OK)
      return fail(s, "instance " + std::to_string(i) + ": " + sc::last_error());
  }
  return SC_OK;
}

sc_status sc_batch_copy_hits(const sc_batch* b, int instance, sc_hit* hits, size_t capacity,
                             size_t* num_hits) {
  if (!b) return fail(SC_ERR_INVALID_ARGUMENT, "null batch");
  if (instance < 0 || instance >= (int)b->inst.size())
    return fail(SC_ERR_INVALID_ARGUMENT, "instance out of range");
  const std::vector<sc_hit>& h = b->inst[instance].hits;
  if (num_hits) *num_hits = h.size();
  if (hits && capacity) std::memcpy(hits, h.data(), std::min(capacity, h.size()) * sizeof(sc_hit));
  return SC_OK;
}

sc_status sc_batch_flagged_blocks(const sc_batch* b, int instance, int* tiles_searched,
                                  int* flagged_blocks) {
  if (!b) return fail(SC_ERR_INVALID_ARGUMENT, "null batch");
  if (!b->pn) return fail(SC_ERR_INVALID_ARGUMENT, "not a gated batch");
  if (instance < 0 || instance >= (int)b->inst.size())
    return fail(SC_ERR_INVALID_ARGUMENT, "instance out of range");
  if (tiles_searched) *tiles_searched = b->h_gate[2 * instance];
  if (flagged_blocks) *flagged_blocks = b->h_gate[2 * instance + 1];
  return SC_OK;
}

sc_status sc_batch_set_timing(sc_batch* b, int enable) {
  if (!b) return fail(SC_ERR_INVALID_ARGUMENT, "null batch");
  if (b->pn) return fail(SC_ERR_UNSUPPORTED, "kernel timing is for plain batches");
  b->timing = enable != 0;
  b->kernel_ms = 0.0;
  b->kernel_launches = 0;
  b->kernel_channels = 0;
  return SC_OK;
}

sc_status sc_batch_kernel_time(const sc_batch* b, double* total_ms, long* launches,
                               long* channels) {
  if (!b) return fail(SC_ERR_INVALID_ARGUMENT, "null batch");
  if (total_ms) *total_ms = b->kernel_ms;
  if (launches) *launches = b->kernel_launches;
  if (channels) *channels = b->kernel_channels;
  return SC_OK;
}

sc_status sc_batch_info(const sc_batch* b, int* num_shapes, int* graph_nodes) {
  if (!b) return fail(SC_ERR_INVALID_ARGUMENT, "null batch");
  if (num_shapes) {
    int n = 0;
    for (size_t k = 0; k < b->shapes.size(); k += b->alternates[k]) ++n;
    *num_shapes = n;
  }
  if (graph_nodes) *graph_nodes = b->nodes;
  return SC_OK;
}

sc_status sc_batch_destroy(sc_batch* b) {
  if (!b) return SC_OK;
  batch_release(b);
  delete b;
  return SC_OK;
}

}  // extern "C"
