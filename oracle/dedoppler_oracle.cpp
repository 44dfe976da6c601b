// Dedoppler oracle: a CPU restatement of lacker/seticore's Taylor-tree search.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library links or calls this
// file. It is loaded by tests/ (as the parity checker), by
// __graft_entry__.smoke() (as the checker of the smoke run) and by bench.py's
// cpu_baseline leg (as the timed CPU restatement, "kind": "port").
//
// Every function follows the reference's arithmetic order so that float results
// are bit-identical to what the reference computes on the same bytes:
//   * the Taylor recursion (taylor.h:30-86) with its pairwise association,
//   * the global-memory tree driver (taylor.cu:116-156),
//   * column sums in sequential float order (dedoppler.cu:63-72),
//   * top-path-sum tracking with the early-return quirk (dedoppler.cu:32-55),
//   * host statistics with the same std::nth_element calls (dedoppler.cu:241-262),
//   * threshold and window merge (dedoppler.cu:264-319),
//   * hit construction (dedoppler_hit.cpp:11-23).
// Pinning (see tests/test_oracle_*.py): brute-force path enumeration, the
// integer ramp of taylor_test.cu:14, the dedoppler_test.cpp:9-33 known answer,
// the data/voyager.hits fixture (real data, affine SNR relation) and an FP64
// restatement whose statistics are computed from a full sort.
//
// Build: make -C oracle  (g++ only; no GPU code).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

extern "C" {

// Same field order and meaning as DedopplerHit (dedoppler_hit.h:11-45).
typedef struct {
  double frequency;
  int index;
  int drift_steps;
  double drift_rate;
  float snr;
  int coarse_channel;
  int beam;
  int num_timesteps;
  float power;
  float incoherent_power;
} oracle_hit;

// util.cpp:16-22
int oracle_round_up_pow2(int n) {
  int rounded = 1;
  while (rounded < n) rounded *= 2;
  return rounded;
}

// taylor.h:30-86: one recursion stage for one channel. Out-of-range
// second-half reads skip the write, leaving the target untouched.
void oracle_taylor_one_step_one_channel(const float* src, float* dst, int chan,
                                        int num_timesteps, int num_src_channels,
                                        int num_dst_channels, int path_length,
                                        int drift_block) {
  if (chan < 0 || chan >= num_src_channels || chan >= num_dst_channels) return;
  int num_time_blocks = num_timesteps / path_length;
  for (int tb = 0; tb < num_time_blocks; ++tb) {
    for (int p = path_length - 1; p >= 0; p--) {
      int half = p / 2;
      int shift = (p + 1) / 2 + drift_block * path_length / 2;
      if (chan + shift < 0 || chan + shift >= num_src_channels) continue;
      dst[(long)(tb * path_length + p) * num_dst_channels + chan] =
          src[(long)(tb * path_length + half) * num_src_channels + chan] +
          src[(long)(tb * path_length + half + path_length / 2) * num_src_channels +
              chan + shift];
    }
  }
}

// taylor.cu:116-156. Returns 1 if the result is in buf1, 2 if in buf2.
// Buffers are [T][N] floats. The input is never written.
int oracle_basic_taylor_tree(const float* input, float* buf1, float* buf2,
                             int num_timesteps, int num_channels, int drift_block) {
  const float* src = input;
  float* dst = buf1;
  int result = 0;
  for (int L = 2; L <= num_timesteps; L *= 2) {
    for (int c = 0; c < num_channels; ++c) {
      oracle_taylor_one_step_one_channel(src, dst, c, num_timesteps, num_channels,
                                         num_channels, L, drift_block);
    }
    src = dst;
    result = (dst == buf1) ? 1 : 2;
    dst = (dst == buf1) ? buf2 : buf1;
  }
  return result;
}

// dedoppler.cu:63-72: sequential float sum down each column, padding rows included.
void oracle_sum_columns(const float* input, float* sums, int num_timesteps,
                        int num_channels) {
  for (int c = 0; c < num_channels; ++c) {
    float s = 0.0f;
    for (long i = c; i < (long)num_timesteps * num_channels; i += num_channels) s += input[i];
    sums[c] = s;
  }
}

// dedoppler.cu:32-55, including the early return at the first path whose last
// channel is out of range (for negative drift blocks that skips all offsets).
void oracle_find_top_path_sums(const float* path_sums, int num_timesteps,
                               int num_channels, int drift_block, float* top_sums,
                               int* top_blocks, int* top_offsets) {
  for (int c = 0; c < num_channels; ++c) {
    for (int p = 0; p < num_timesteps; ++p) {
      int last = (num_timesteps - 1) * drift_block + p + c;
      if (last < 0 || last >= num_channels) break;
      float s = path_sums[(long)num_channels * p + c];
      if (s > top_sums[c]) {
        top_sums[c] = s;
        top_blocks[c] = drift_block;
        top_offsets[c] = p;
      }
    }
  }
}

// dedoppler.cu:195-198 (drift-block range) and dedoppler.cu:87 (resolution).
void oracle_drift_params(int num_timesteps, double foff, double tsamp, double max_drift,
                         int* min_block, int* max_block, int* window_size,
                         double* resolution) {
  int rounded = oracle_round_up_pow2(num_timesteps);
  int drift_timesteps = rounded - 1;
  double res = 1e6 * foff / (drift_timesteps * tsamp);
  double diagonal = res * drift_timesteps;
  double nmd = max_drift / std::fabs(diagonal);
  *min_block = (int)std::floor(-nmd);
  *max_block = (int)std::floor(nmd);
  *window_size = 2 * (int)std::ceil(nmd * drift_timesteps);
  *resolution = res;
}

// dedoppler.cu:241-262: median and trimmed standard deviation of the column
// sums, computed with the same three std::nth_element calls. `sums` is permuted
// in place exactly as the reference permutes its pinned host copy.
void oracle_host_stats(float* sums, int n, float* median_out, float* std_out) {
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

// dedoppler.cu:264-319 plus dedoppler_hit.cpp:11-23. Returns the number of hits
// written (at most cap; the count beyond cap is still returned).
long oracle_merge_hits(const float* top_sums, const int* top_blocks,
                       const int* top_offsets, int num_channels, int num_timesteps,
                       double foff, double tsamp, double max_drift, double min_drift,
                       double snr_threshold, float median, float std_dev, double fch1,
                       long coarse_channel_size, int num_coherent_beams, int beam,
                       int coarse_channel, oracle_hit* out, long cap) {
  int rounded = oracle_round_up_pow2(num_timesteps);
  int drift_timesteps = rounded - 1;
  double res = 1e6 * foff / (drift_timesteps * tsamp);
  double nmd = max_drift / std::fabs(res * drift_timesteps);
  float threshold = snr_threshold * std_dev + median;
  int window = 2 * (int)std::ceil(nmd * drift_timesteps);
  long count = 0;
  if (window <= 0) return -1;  // the reference would never terminate here
  for (int i = 0; (long)i * window < num_channels; ++i) {
    int cand = -1;
    float cand_sum = threshold;
    for (int j = 0; j < window; ++j) {
      int f = i * window + j;
      if (f >= num_channels) break;
      if (top_sums[f] > cand_sum) {
        cand = f;
        cand_sum = top_sums[f];
      }
    }
    if (cand < 0) continue;
    int wend = std::min(num_channels, cand + window);
    bool larger = false;
    for (int f = std::max(0, cand - window + 1); f < wend; ++f) {
      if (top_sums[f] > cand_sum) {
        larger = true;
        break;
      }
    }
    if (larger) continue;
    int drift_bins = top_blocks[cand] * drift_timesteps + top_offsets[cand];
    double drift_rate = drift_bins * res;
    float snr = (cand_sum - median) / std_dev;
    if (std::fabs(drift_rate) >= min_drift) {
      if (count < cap) {
        oracle_hit h;
        h.index = cand;
        h.drift_steps = drift_bins;
        h.drift_rate = drift_rate;
        h.snr = snr;
        h.coarse_channel = coarse_channel;
        h.beam = (beam >= 0 && beam < num_coherent_beams) ? beam : -1;
        h.num_timesteps = num_timesteps;
        h.power = cand_sum;
        h.incoherent_power = 0.0f;
        int global_index = coarse_channel * (int)coarse_channel_size + cand;
        h.frequency = fch1 + global_index * foff;
        out[count] = h;
      }
      ++count;
    }
  }
  return count;
}

// Tree + top-path tracking for every drift block over channel range
// [c_begin, c_end) of the whole buffer: the tree is computed on a slab with a
// halo wide enough for every valid path that starts in the range, which gives
// the same float results as the whole-buffer tree for those paths.
static void tree_and_top_slab(const float* input, int T, int N, int bmin, int bmax,
                              int c_begin, int c_end, float* top, int* blk, int* off) {
  int halo_left = std::max(0, -bmin) * (T - 1);
  int halo_right = (std::max(0, bmax) + 1) * (T - 1);
  int s0 = std::max(0, c_begin - halo_left);
  int s1 = std::min(N, c_end + halo_right);
  int w = s1 - s0;
  std::vector<float> slab((size_t)T * w), b1((size_t)T * w, 0.0f), b2((size_t)T * w, 0.0f);
  for (int t = 0; t < T; ++t)
    std::memcpy(&slab[(size_t)t * w], input + (size_t)t * N + s0, sizeof(float) * w);
  for (int b = bmin; b <= bmax; ++b) {
    int which = oracle_basic_taylor_tree(slab.data(), b1.data(), b2.data(), T, w, b);
    const float* sums = (which == 1) ? b1.data() : b2.data();
    for (int c = c_begin; c < c_end; ++c) {
      int lc = c - s0;
      for (int p = 0; p < T; ++p) {
        int last = (T - 1) * b + p + c;  // validity against the whole buffer
        if (last < 0 || last >= N) break;
        float s = sums[(size_t)w * p + lc];
        if (s > top[c]) {
          top[c] = s;
          blk[c] = b;
          off[c] = p;
        }
      }
    }
  }
}

// The full search of dedoppler.cu:186-320 for one coarse channel.
// `top_blocks`/`top_offsets` carry state between calls like the reference's
// device arrays (they are not reset); `top_sums` is zeroed here
// (dedoppler.cu:205). Optional outputs: colsums (as computed, before the stats
// permutation), median and std. num_threads > 1 splits the tree over channel
// slabs (same float results). Returns the hit count.
long oracle_search(const float* input, int num_timesteps, int num_channels, double foff,
                   double tsamp, double max_drift, double min_drift, double snr_threshold,
                   double fch1, long coarse_channel_size, int num_coherent_beams,
                   int beam, int coarse_channel, float* top_sums, int* top_blocks,
                   int* top_offsets, float* colsums_out, float* median_out,
                   float* std_out, int num_threads, oracle_hit* out, long cap) {
  int T = oracle_round_up_pow2(num_timesteps);
  int N = num_channels;
  int bmin, bmax, window;
  double res;
  oracle_drift_params(num_timesteps, foff, tsamp, max_drift, &bmin, &bmax, &window, &res);
  if (window <= 0) return -1;
  for (int c = 0; c < N; ++c) top_sums[c] = 0.0f;
  std::vector<float> colsums(N);
  oracle_sum_columns(input, colsums.data(), T, N);
  if (colsums_out) std::memcpy(colsums_out, colsums.data(), sizeof(float) * N);

  if (num_threads <= 1) {
    tree_and_top_slab(input, T, N, bmin, bmax, 0, N, top_sums, top_blocks, top_offsets);
  } else {
    std::vector<std::thread> pool;
    int per = (N + num_threads - 1) / num_threads;
    for (int k = 0; k < num_threads; ++k) {
      int c0 = k * per, c1 = std::min(N, c0 + per);
      if (c0 >= c1) break;
      pool.emplace_back(tree_and_top_slab, input, T, N, bmin, bmax, c0, c1, top_sums,
                        top_blocks, top_offsets);
    }
    for (auto& th : pool) th.join();
  }

  float median, std_dev;
  oracle_host_stats(colsums.data(), N, &median, &std_dev);
  if (median_out) *median_out = median;
  if (std_out) *std_out = std_dev;
  return oracle_merge_hits(top_sums, top_blocks, top_offsets, N, num_timesteps, foff,
                           tsamp, max_drift, min_drift, snr_threshold, median, std_dev,
                           fch1, coarse_channel_size, num_coherent_beams, beam,
                           coarse_channel, out, cap);
}

// ---------------------------------------------------------------------------
// Independent checks (the brute-force side of the pinning tests).

// Channel offset at time step t of path offset p over a power-of-two span L,
// from the path definition the recursion implements (taylor.h:48-54): the
// first half follows offset p/2, the second half starts (p+1)/2 further right.
int oracle_path_offset(int L, int p, int t) {
  int acc = 0;
  while (L > 1) {
    int half = L / 2;
    if (t >= half) {
      acc += (p + 1) / 2;
      t -= half;
    }
    p /= 2;
    L = half;
  }
  return acc;
}

// Brute-force path sum: explicit walk over the path's cells in double.
double oracle_path_sum_fp64(const float* input, int T, int N, int chan, int drift_block,
                            int p) {
  double s = 0.0;
  for (int t = 0; t < T; ++t) {
    int c = chan + t * drift_block + oracle_path_offset(T, p, t);
    if (c < 0 || c >= N) return NAN;
    s += (double)input[(size_t)t * N + c];
  }
  return s;
}

// FP64 statistics: exact upper median of the column sums (computed in double
// here) and the population standard deviation of the sorted ranks
// [ceil(.05N), floor(.95N)], all in double.
void oracle_stats_fp64(const double* colsums, int n, double* median_out,
                       double* mean_out, double* std_out) {
  std::vector<double> v(colsums, colsums + n);
  std::sort(v.begin(), v.end());
  int mid = n / 2;
  int first = (int)std::ceil(0.05 * n);
  int last = (int)std::floor(0.95 * n);
  long double s = 0;
  for (int i = first; i <= last; ++i) s += v[i];
  double m = (double)(s / (last + 1 - first));
  long double acc = 0;
  for (int i = first; i <= last; ++i) acc += (long double)(v[i] - m) * (v[i] - m);
  *median_out = v[mid];
  *mean_out = m;
  *std_out = std::sqrt((double)(acc / (last + 1 - first)));
}

// Column sums in double, for the FP64 statistics.
void oracle_sum_columns_fp64(const float* input, int T, int N, double* sums) {
  for (int c = 0; c < N; ++c) {
    double s = 0.0;
    for (int t = 0; t < T; ++t) s += input[(size_t)t * N + c];
    sums[c] = s;
  }
}

}  // extern "C"
