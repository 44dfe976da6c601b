// Internal interface between the C-ABI host code (dd_capi.cpp) and the HIP
// kernels (dd_kernels.hip). Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/seticore_amd.h"

namespace sc {

// Drift-block range and merge window of a search (dedoppler.cu:195-198, 274):
// nmd = max_drift / |resolution * drift_timesteps|, blocks floor(-nmd) ..
// floor(nmd), window 2 * ceil(nmd * drift_timesteps).
struct DriftRange {
  int bmin, bmax, window;
  double nmd;
};
inline DriftRange drift_range_of(double resolution, int drift_timesteps, double max_drift) {
  DriftRange r;
  const double diagonal = resolution * drift_timesteps;
  r.nmd = max_drift / fabs(diagonal);
  r.bmin = (int)floor(-r.nmd);
  r.bmax = (int)floor(r.nmd);
  r.window = (int)(2 * ceil(r.nmd * drift_timesteps));
  return r;
}

// A hit as the reference's DedopplerHit constructor forms it
// (dedoppler_hit.cpp:11-23): frequency from the global channel index, beam -1
// unless it is a coherent beam of the metadata.
sc_hit make_hit(const sc_meta* meta, int index, int drift_steps, double drift_rate, float snr,
                int beam, int coarse_channel, int num_timesteps, float power);

// Default one-pass kernel shape for T_r (channels per lane, drift blocks per
// LDS window or workgroup, window mode / kernel selector), as measured best.
void default_tuning(int T, int* lanes, int* group, int* prefetch);

// Per-channel search state: the bits of an FP64 key whose value orders
// (top path sum, -code) lexicographically - the float sum widened to double
// with 0x1FFFFFFF - code in the 29 low mantissa bits, where
// code = (drift_block - min_drift_block) * T_r + path_offset. A max keeps the
// largest sum and, on equal sums, the lowest (drift block, offset): the
// reference's strict ">" in ascending order (dedoppler.cu:40,49,214).
// 0 means "no path sum > 0" (dedoppler.cu:205).
typedef unsigned long long top_key_t;

// Fused tree + top-path tracking for T_r <= 16 over all drift blocks in one
// launch. C = channels per lane, G = drift blocks per LDS window, P = 1 to
// prefetch the next window into registers. Writes column sums (unless colsum
// is null) and max-merges top keys. With tile_list (device), only the channel
// tiles tile_list[0 .. *tile_count) are searched (tiles of
// fused_tile_width(T, C) channels). With block_keys (T_r <= 16 and the staged
// kernel; fused32_kernel ignores it) the keys are block keys: (sum bits << 32) | ~(drift block -
// bmin), the offset left to launch_candidate_offsets after the window merge.
hipError_t launch_fused_tree_top(const void* x, int dtype, long stride, int T, int N,
                                 int bmin, int nblocks, int C, int G, int P,
                                 top_key_t* top_key, float* colsum, hipStream_t stream,
                                 const int* tile_list = nullptr,
                                 const int* tile_count = nullptr, bool block_keys = false);
int fused_tile_width(int T, int C);
// Up to kMaxFusedSet coarse channels of one shape (same T, N, row stride,
// sample type and drift range) searched by one launch of the T_r = 16 kernel.
constexpr int kMaxFusedSet = 8;
struct FusedSet {
  const void* x[kMaxFusedSet];
  top_key_t* key[kMaxFusedSet];
  float* colsum[kMaxFusedSet];  // may be null (column sums not written)
  // Gated set launches (launch_fused16_set with lists): per channel the
  // ascending tile list and its length (device), as in launch_fused_tree_top.
  const int* list[kMaxFusedSet];
  const int* count[kMaxFusedSet];
};
// Per-channel buffers of a gated batch run's front: one launch per step for
// all n channels (launch_front_set_stats, pn::launch_pool_set,
// pn::launch_tile_list_set), grid.y = channel.
struct FrontSet {
  const void* x[kMaxFusedSet];
  top_key_t* key[kMaxFusedSet];
  unsigned* hist[kMaxFusedSet];
  float* colsum[kMaxFusedSet];
  void* stats_scratch[kMaxFusedSet];
  float* stats_out[kMaxFusedSet];
  int* list[kMaxFusedSet];
  int* count[kMaxFusedSet];  // {tiles listed, blocks flagged}
};
// launch_reset, launch_colsum and launch_device_stats for n channels of one
// shape, each step one launch for all of them.
hipError_t launch_front_set_stats(const FrontSet& fs, int n, int dtype, long stride, int T,
                                  int N, hipStream_t stream);
// launch_device_stats for n channels (their colsum, hist, stats_scratch,
// stats_out), each pass one launch for all of them.
hipError_t launch_device_stats_set(const FrontSet& fs, int n, int N, hipStream_t stream);

// Per-channel buffers and parameters of a batch run's tail (16 time steps,
// block keys): launch_window_merge and launch_candidate_offsets for n
// channels of one width and drift-block range, one launch each (grid.y =
// channel).
struct TailSet {
  const void* x[kMaxFusedSet];
  const top_key_t* key[kMaxFusedSet];
  const float* stats[kMaxFusedSet];
  float* top_sum[kMaxFusedSet];
  int* top_blk[kMaxFusedSet];
  int* top_off[kMaxFusedSet];
  int4* cand[kMaxFusedSet];
  int window[kMaxFusedSet];
  double snr[kMaxFusedSet];
};
hipError_t launch_window_merge_set(const TailSet& ts, int n, int N, int T, int bmin,
                                   hipStream_t stream, bool skip_empty = false);
hipError_t launch_candidate_offsets_set(const TailSet& ts, int n, int dtype, long stride, int N,
                                        hipStream_t stream);

// Whether launch_fused16_set has a kernel for (T, C, G, P) (the default
// 16-time-step shape).
bool fused_set_supported(int T, int C, int G, int P);
// With lists (gated: block keys, only the listed tiles of each channel),
// the grid is sized for every tile and the idle items exit.
hipError_t launch_fused16_set(const FusedSet& set, int n, int dtype, long stride, int T, int N,
                              int bmin, int nblocks, int C, int G, int P, bool block_keys,
                              hipStream_t stream, bool lists = false);

// The T_r <= 16 kernel of launch_fused_tree_top (dd_fused16.hip), same arguments.
hipError_t launch_fused16(const void* x, int dtype, long stride, int T, int N, int bmin,
                          int nblocks, int C, int G, int P, top_key_t* top_key, float* colsum,
                          hipStream_t stream, const int* tile_list, const int* tile_count,
                          bool block_keys);

// Whether launch_fused_tree_top has a kernel for (T, C, G, P). At T = 32 the
// `prefetch` value kStaged32 selects the LDS-staged kernel of T = 64..256
// (shape (4, G, kStaged32)) instead of fused32_kernel.
constexpr int kStaged32 = 8;
bool fused_supported(int T, int C, int G, int P);
// Whether (T, C, G, P) selects the LDS-staged kernel (T = 32..256). At
// T = 128, 256, P = 1 gives each offset group its own workgroup.
bool staged_shape(int T, int C, int G, int P);

// T_r = 32..256: path sums of drift blocks [b_first, b_first + nb) folded into
// the top keys by two passes through scratch (nb * T * row_pitch floats,
// row_pitch from two_pass_row_pitch). G = drift blocks per LDS window in the
// first pass (4 or 8). Column sums are not written.
bool two_pass_supported(int T, int G);
long two_pass_row_pitch(int T, int N);
hipError_t launch_two_pass(const void* x, int dtype, long stride, int T, int N, int bmin,
                           int b_first, int nb, int G, float* scratch, long row_pitch,
                           top_key_t* top_key, hipStream_t stream);

// Level-min(T,16) sums for every time block of one drift block:
// out[(tb*Tb + p)*N + c], Tb = min(T, 16), tb < T/Tb.
hipError_t launch_tree_blocks(const void* x, int dtype, long stride, int T, int N,
                              int drift, float* out, hipStream_t stream);

// One global-memory recursion stage (taylor.cu:97-109 semantics).
hipError_t launch_one_step(const float* src, float* dst, int T, int N, int L, int drift,
                           hipStream_t stream);

// Sequential per-drift-block top update from a [T][N] path-sum buffer.
hipError_t launch_top_update(const float* sums, int T, int N, int drift, int bmin,
                             top_key_t* top_key, hipStream_t stream);

// Column sums (dedoppler.cu:63-72 order) for the paths that do not fuse them.
hipError_t launch_colsum(const void* x, int dtype, long stride, int T, int N, float* colsum,
                         hipStream_t stream);

// Zeroes the top keys and the statistics histograms before a search.
hipError_t launch_reset(top_key_t* key, long n, unsigned* hist, hipStream_t stream);

// Unpacks keys into the reference's three top arrays; channels without a
// positive path get top_sum 0 and keep their previous block/offset.
hipError_t launch_unpack_top(const top_key_t* key, int N, int T, int bmin, float* top_sum,
                             int* top_blk, int* top_off, hipStream_t stream);

// ---- device statistics -----------------------------------------------------
size_t device_stats_scratch_bytes(int N);
size_t device_stats_hist_bytes();
// From the column sums (hist: zeroed histogram scratch): median (rank N/2) and
// the trimmed mean/std over the sorted ranks [ceil(.05N), floor(.95N)] into
// stats_out = {median, std, mean}.
hipError_t launch_device_stats(const float* colsum, int N, void* scratch, unsigned* hist,
                               float* stats_out, hipStream_t stream);

// Per-window merge (dedoppler.cu:273-303) fused with unpacking the keys into
// the three top arrays. cand_out[i] = {index or -1, sum bits, drift block,
// path offset} for window i. With block keys the offsets (unpacked and in
// the candidates) are -1 until launch_candidate_offsets fills the candidates'.
// With stats_slot, cand_out[number of windows] also receives the four
// statistics (float bits), for one copy to the host.
hipError_t launch_window_merge(const top_key_t* key, int N, int T, int bmin, int window,
                               double snr_threshold, const float* stats, float* top_sum,
                               int* top_blk, int* top_off, int4* cand_out, bool block_keys,
                               hipStream_t stream, bool skip_empty = false,
                               bool stats_slot = false);
// After a block-key merge (T_r <= 16): each candidate's path offset, from the
// input of the search.
hipError_t launch_candidate_offsets(const void* x, int dtype, long stride, int T, int N,
                                    int nwin, int4* cand, int* top_off, hipStream_t stream);

// Gathers path sums for (offset, channel) pairs from a [T][N] buffer.
hipError_t launch_gather(const float* sums, int N, const int* offsets, const int* chans,
                         int n, float* out, hipStream_t stream);

// ---- normalisation and synthetic data -------------------------------------
hipError_t launch_normalize(const float* x, int T, int N, long stride, int out_dtype,
                            float clamp, void* out, double* moments /* device: mean, std, location */,
                            void* scratch, hipStream_t stream);
size_t normalize_scratch_bytes(int T, int N);

// Loader conditioning (zero rows T..T_r-1; optional DC fix on rows < T).
hipError_t launch_condition(float* data, int T, int T_r, int N, long stride, int dc_spike,
                            hipStream_t stream);

hipError_t launch_gen_waterfall(float* out, int T, int T_r, int N, uint64_t seed, float gain,
                                float taper_frac, float dc_spike, int num_tones,
                                const float* tone_start, const float* tone_slope,
                                const float* tone_amp, hipStream_t stream);

}  // namespace sc
