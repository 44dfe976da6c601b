// HIP kernels for the MI355X (gfx950) Taylor-tree dedoppler search.
//
// Reference semantics: taylor.h:30-118, taylor.cu:97-350, dedoppler.cu:32-72 and
// dedoppler.cu:186-320 of lacker/seticore. The dataflow here is different: the
// reference runs one tree launch plus one findTopPathSums launch per drift block,
// each streaming the whole waterfall through HBM; here one launch covers the
// whole drift range, each wavefront keeps a channel tile's tree in registers and
// reuses one LDS window across a group of drift blocks, and the max/argmax is
// folded into the same kernel. Float association of every path sum is the
// reference's pairwise recursion, so path sums are bit-identical.
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off
//        -Xclang -target-feature -Xclang -packed-fp32-ops

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "dd_internal.h"

namespace sc {

namespace {

constexpr int kWave = 64;

template <typename In>
__device__ __forceinline__ float to_f32(In v);
template <>
__device__ __forceinline__ float to_f32<float>(float v) { return v; }
template <>
__device__ __forceinline__ float to_f32<__half>(__half v) { return __half2float(v); }
template <>
__device__ __forceinline__ float to_f32<__hip_bfloat16>(__hip_bfloat16 v) {
  return __bfloat162float(v);
}

// Lane i receives lane i+1's value (DPP wave_shl:1); lane 63 receives 0
// (bound_ctrl).
__device__ __forceinline__ float from_next_lane(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x130, 0xf, 0xf, true));
}

// ---------------------------------------------------------------------------
// Top-path keys.
//
// A key is an FP64 number whose value orders (path sum, -code) lexicographically:
// the float path sum widened to double (exact; the low 29 mantissa bits are
// zero) with kCodeMask - code placed in those low bits, where
// code = (drift_block - min_drift_block) * T_r + path_offset. The max of keys
// is therefore the largest sum and, among equal sums, the lowest
// (drift block, offset) - the reference's strict ">" in ascending order
// (dedoppler.cu:40,49,214). Keys of sums <= 0 never exceed kEmptyKey, the key
// of (0, largest code), matching "only sums > 0 register" (dedoppler.cu:205).
// As unsigned 64-bit integers positive keys order the same way, so partial
// results from different workgroups merge with a 64-bit atomic max; 0 means
// "no positive path".
constexpr unsigned long long kCodeMask = 0x1FFFFFFFull;
constexpr unsigned long long kEmptyKey = kCodeMask;

__device__ __forceinline__ double make_key(float s, unsigned low_bits) {
  const unsigned long long u = (unsigned long long)__double_as_longlong((double)s) | low_bits;
  return __longlong_as_double((long long)u);
}

// ---------------------------------------------------------------------------
// Register-resident Taylor tree.
//
// v[r][j] holds, for channel j of this lane's C consecutive channels (sheared
// coordinates: the drift block is folded into the load), the level-L array
// entry r = time_block * L + path_offset (the reference's
// [time_block][offset][chan] layout, taylor.cu:31-61). Stage L combines entry
// (tb*L + q) with entry (tb*L + L/2 + q) shifted right by q or q+1 channels
// (taylor.h:53-54,80-82 with drift block 0 in sheared coordinates, as the
// reference's tiled kernel runs it, taylor.cu:234-253). Channels j+shift >= C
// come from the following lanes through chained DPP moves (the last move folds
// into the add), so a lane's outputs are exact when every input they depend
// on lies inside the wave.
template <int T, int C, int L>
__device__ __forceinline__ void tree_stage(float (&v)[T][C]) {
  float w[T][C];
#pragma unroll
  for (int tb = 0; tb < T / L; ++tb) {
#pragma unroll
    for (int q = 0; q < L / 2; ++q) {
      const int ra = tb * L + q;
      const int rb = tb * L + L / 2 + q;
      float bx[C + L / 2];
#pragma unroll
      for (int j = 0; j < C; ++j) bx[j] = v[rb][j];
#pragma unroll
      for (int k = 0; k < L / 2; ++k) {
        if (k <= q) bx[C + k] = from_next_lane(bx[k]);
      }
#pragma unroll
      for (int j = 0; j < C; ++j) {
        w[tb * L + 2 * q][j] = v[ra][j] + bx[j + q];
        w[tb * L + 2 * q + 1][j] = v[ra][j] + bx[j + q + 1];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < T; ++r)
#pragma unroll
    for (int j = 0; j < C; ++j) v[r][j] = w[r][j];
}

template <int T, int C, int L = 2>
__device__ __forceinline__ void tree_regs(float (&v)[T][C]) {
  if constexpr (L <= T) {
    tree_stage<T, C, L>(v);
    tree_regs<T, C, L * 2>(v);
  }
}

// Lanes whose outputs are exact: lane l needs input channels up to
// C*l + C-1 + T-1, which must lie inside the wave's 64*C channels.
template <int T, int C>
struct TileShape {
  static constexpr int kHaloLanes = (T - 1 + C - 1) / C;
  static constexpr int kValidLanes = kWave - kHaloLanes;
  static constexpr int kOut = kValidLanes * C;  // output channels per wave
};

// LDS window for G drift blocks starting at b0: row t holds global channels
// [c_w + t*b0, c_w + t*b0 + C*S), stored de-interleaved by C (element k at
// sub-row k % C, column k / C) so that every lane reads consecutive words
// (no bank conflicts) and every read address is a per-row base plus a
// compile-time offset.
template <int T, int C, int G>
struct Window {
  // Sub-row length, rounded so that a row is a whole number of 64-lane loads.
  static constexpr int S0 = (kWave * C + (T - 1) * (G - 1) + C - 1) / C;
  static constexpr int kStep = kWave / C;
  static constexpr int S = (S0 + kStep - 1) / kStep * kStep;
  static constexpr int kRow = C * S;
  static constexpr int kFloats = T * kRow;
};

typedef __attribute__((address_space(3))) float lds_float;

// Loads the window rows. Rows whose span lies inside [0, N) take unguarded
// loads (immediate offsets off one row pointer); rows crossing an edge clamp
// the address and substitute 0 for out-of-range samples.
template <int T, int C, int G, typename In>
__device__ __forceinline__ void load_window(lds_float* win, const In* __restrict__ x,
                                            long stride, int N, int c_w, int b0, int row0) {
  using Wn = Window<T, C, G>;
  const int lane = threadIdx.x;
  lds_float* dst0 = win + (lane % C) * Wn::S + lane / C;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int g0 = c_w + t * b0;
    const In* row = x + (long)(row0 + t) * stride;
    lds_float* dst = dst0 + t * Wn::kRow;
    if (g0 >= 0 && g0 + Wn::kRow <= N) {
      const In* src = row + g0 + lane;
#pragma unroll
      for (int k0 = 0; k0 < Wn::kRow; k0 += kWave) dst[k0 / C] = to_f32<In>(src[k0]);
    } else {
#pragma unroll
      for (int k0 = 0; k0 < Wn::kRow; k0 += kWave) {
        const int gc = g0 + k0 + lane;
        const bool ok = (unsigned)gc < (unsigned)N;
        const float val = to_f32<In>(row[ok ? gc : 0]);
        dst[k0 / C] = ok ? val : 0.0f;
      }
    }
  }
}

// Per-row LDS base addresses for this lane. The empty asm hides their common
// origin from the compiler so each read below is base + immediate offset (no
// per-read address arithmetic).
template <int T, int C, int G>
__device__ __forceinline__ void row_bases(lds_float* win, lds_float* (&rp)[T]) {
  using Wn = Window<T, C, G>;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    lds_float* p = win + t * Wn::kRow + threadIdx.x;
    asm volatile("" : "+v"(p));
    rp[t] = p;
  }
}

// Reads drift block g of the window (the sheared tile) into registers.
template <int T, int C, int G, int g>
__device__ __forceinline__ void read_block(lds_float* const (&rp)[T], float (&v)[T][C]) {
  using Wn = Window<T, C, G>;
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int k = j + t * g;
      v[t][j] = rp[t][(k % C) * Wn::S + k / C];
    }
}

// Whether every path of drift block b that starts in this wave's output tile
// is valid (dedoppler.cu:42-46), so no masking is needed.
template <int T, int C>
__device__ __forceinline__ bool block_interior(int c_w, int b, int N) {
  using Sh = TileShape<T, C>;
  return (c_w + (T - 1) * b >= 0) && (c_w + Sh::kOut - 1 + (T - 1) * b + (T - 1) < N);
}

// One drift block: tree, optional edge masking (dedoppler.cu:40-46, including
// the early return that skips every offset of a block whose first path ends
// left of channel 0), then the key max in ascending (block, offset) order.
// The interior kernel takes the blocks that need no masking, the edge kernel
// the others; together they cover every (tile, block) exactly once.
template <int T, int C, int G, bool EDGE, int g>
__device__ __forceinline__ void fused_block(lds_float* const (&rp)[T], int ng, int b0, int c_w,
                                            int N, unsigned low0, double (&best)[C]) {
  if constexpr (g < G) {
    if (g < ng) {
      const int b = b0 + g;
      if (block_interior<T, C>(c_w, b, N) != EDGE) {
        float v[T][C];
        read_block<T, C, G, g>(rp, v);
        tree_regs<T, C>(v);
        if constexpr (EDGE) {
          const int lane = threadIdx.x;
#pragma unroll
          for (int j = 0; j < C; ++j) {
            const int c = c_w + lane * C + j;
            const int last0 = c + (T - 1) * b;
            const int pmax = (c >= N || last0 < 0) ? 0 : min(T, N - last0);
#pragma unroll
            for (int p = 0; p < T; ++p)
              if (p >= pmax) v[p][j] = -INFINITY;
          }
        }
#pragma unroll
        for (int p = 0; p < T; ++p) {
          const unsigned low = low0 - (unsigned)(g * T + p);
#pragma unroll
          for (int j = 0; j < C; ++j) best[j] = fmax(best[j], make_key(v[p][j], low));
        }
      }
    }
    fused_block<T, C, G, EDGE, g + 1>(rp, ng, b0, c_w, N, low0, best);
  }
}

// Grid: one 64-lane workgroup per (channel tile, drift group). Work items are
// tile-major and dealt so that each XCD (workgroups go round-robin to the 8
// XCDs) sweeps a contiguous range of tiles: a tile's drift groups share their
// input rows in that XCD's L2. The interior kernel (EDGE=false) covers all
// tiles and also writes the column sums; the edge kernel covers the tiles
// [0, lo_tiles) and [hi_tile0, ntiles), where some drift blocks need masking.
template <int T, int C, int G, bool EDGE, typename In>
__global__ void __launch_bounds__(kWave)
    fused_tree_top_kernel(const In* __restrict__ x, long stride, int N, int bmin, int nblocks,
                          int ngroups, int lo_tiles, int hi_tile0, int total_items,
                          unsigned long long* __restrict__ top_key,
                          float* __restrict__ colsum) {
  using Sh = TileShape<T, C>;
  using Wn = Window<T, C, G>;
  __shared__ float win_storage[Wn::kFloats];
  lds_float* win = (lds_float*)win_storage;

  const int per_xcd = gridDim.x / 8;
  const int item = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (item >= total_items) return;
  int tile = item / ngroups;
  const int grp = item - tile * ngroups;
  if (EDGE && tile >= lo_tiles) tile = hi_tile0 + (tile - lo_tiles);
  const int c_w = tile * Sh::kOut;
  const int b0 = bmin + grp * G;
  const int ng = min(G, nblocks - grp * G);
  const int lane = threadIdx.x;

  if (!EDGE && b0 <= 0 && 0 < b0 + ng && lane < Sh::kValidLanes) {
    // Column sums in the reference's sequential float order (dedoppler.cu:68-71),
    // computed once per tile by the workgroup whose drift group holds block 0.
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int c = c_w + lane * C + j;
      if (c < N) {
        float s = 0.0f;
#pragma unroll
        for (int t = 0; t < T; ++t) s += to_f32<In>(x[(long)t * stride + c]);
        colsum[c] = s;
      }
    }
  }

  load_window<T, C, G, In>(win, x, stride, N, c_w, b0, 0);
  __syncthreads();

  lds_float* rp[T];
  row_bases<T, C, G>(win, rp);
  double best[C];
#pragma unroll
  for (int j = 0; j < C; ++j) best[j] = __longlong_as_double((long long)kEmptyKey);
  const unsigned low0 = (unsigned)kCodeMask - (unsigned)(grp * G * T);
  fused_block<T, C, G, EDGE, 0>(rp, ng, b0, c_w, N, low0, best);

  if (lane < Sh::kValidLanes) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int c = c_w + lane * C + j;
      const unsigned long long k = (unsigned long long)__double_as_longlong(best[j]);
      if (c < N && k > kEmptyKey) atomicMax(&top_key[c], k);
    }
  }
}

// Tree of one drift block for each T-row time block (T <= 16), written out in
// the reference's [time_block][offset][chan] layout. The shear is relative to
// each time block's first row, as in the reference's tiled kernel
// (taylor.cu:225-231).
template <int T, int C, typename In>
__global__ void __launch_bounds__(kWave)
    tree_blocks_kernel(const In* __restrict__ x, long stride, int N, int drift,
                       float* __restrict__ out) {
  using Sh = TileShape<T, C>;
  using Wn = Window<T, C, 1>;
  __shared__ float win_storage[Wn::kFloats];
  lds_float* win = (lds_float*)win_storage;
  const int c_w = blockIdx.x * Sh::kOut;
  const int tb = blockIdx.y;
  load_window<T, C, 1, In>(win, x, stride, N, c_w, drift, tb * T);
  __syncthreads();
  lds_float* rp[T];
  row_bases<T, C, 1>(win, rp);
  float v[T][C];
  read_block<T, C, 1, 0>(rp, v);
  tree_regs<T, C>(v);
  const int lane = threadIdx.x;
  if (lane < Sh::kValidLanes) {
#pragma unroll
    for (int p = 0; p < T; ++p)
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const int c = c_w + lane * C + j;
        if (c < N) out[(long)(tb * T + p) * N + c] = v[p][j];
      }
  }
}

// taylor.cu:97-109 + taylor.h:30-86: one stage in global memory.
__global__ void one_step_kernel(const float* __restrict__ src, float* __restrict__ dst, int T,
                                int N, int L, int drift) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int ntb = T / L;
  for (int tb = 0; tb < ntb; ++tb) {
    for (int p = L - 1; p >= 0; --p) {
      const int half = p / 2;
      const int shift = (p + 1) / 2 + drift * L / 2;
      if (c + shift < 0 || c + shift >= N) continue;
      dst[(long)(tb * L + p) * N + c] = src[(long)(tb * L + half) * N + c] +
                                        src[(long)(tb * L + half + L / 2) * N + c + shift];
    }
  }
}

// dedoppler.cu:32-55 for one drift block, folded into the key (launched once
// per drift block in ascending order, so a plain read-modify-write suffices).
__global__ void top_update_kernel(const float* __restrict__ sums, int T, int N, int drift,
                                  int bmin, top_key_t* __restrict__ top_key) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float best = 0.0f;
  int code = 0;
  for (int p = 0; p < T; ++p) {
    const int last = (T - 1) * drift + p + c;
    if (last < 0 || last >= N) break;
    const float s = sums[(long)N * p + c];
    if (s > best) {
      best = s;
      code = (drift - bmin) * T + p;
    }
  }
  if (best > 0.0f) {
    const top_key_t k = (top_key_t)__double_as_longlong(
        make_key(best, (unsigned)kCodeMask - (unsigned)code));
    if (k > top_key[c]) top_key[c] = k;
  }
}

template <typename In>
__global__ void colsum_kernel(const In* __restrict__ x, long stride, int T, int N,
                              float* __restrict__ colsum) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.0f;
  for (int t = 0; t < T; ++t) s += to_f32<In>(x[(long)t * stride + c]);
  colsum[c] = s;
}

__global__ void fill_u64_kernel(top_key_t* p, long n, top_key_t value) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = value;
}

__global__ void unpack_top_kernel(const top_key_t* __restrict__ key, int N, int T, int bmin,
                                  float* __restrict__ top_sum, int* __restrict__ top_blk,
                                  int* __restrict__ top_off) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const top_key_t k = key[c];
  if (k == 0) {
    top_sum[c] = 0.0f;  // block/offset keep their previous values (dedoppler.cu:205)
  } else {
    const unsigned code = (unsigned)(kCodeMask - (k & kCodeMask));
    top_sum[c] = (float)__longlong_as_double((long long)(k & ~kCodeMask));
    top_blk[c] = (int)(code / (unsigned)T) + bmin;
    top_off[c] = (int)(code % (unsigned)T);
  }
}

// ---------------------------------------------------------------------------
// Device statistics: exact order statistics by 4-pass 8-bit radix select, and
// FP64 trimmed moments reduced in a fixed order (deterministic).

struct StatsState {
  unsigned hist[3][256];
  unsigned prefix[3];
  unsigned rank[3];
  double mean_sum;  // FP64 sum over the trimmed ranks
  float mean_f;     // float(sum) / count, as the reference forms its mean
  int k_lo, k_hi;   // how many copies of the boundary values are in the set
  int pad;
};

constexpr int kStatsBlocks = 256;
constexpr int kStatsThreads = 256;

struct StatsPartials {
  double s[kStatsBlocks];
  unsigned n_in[kStatsBlocks], lt_lo[kStatsBlocks], eq_lo[kStatsBlocks], lt_hi[kStatsBlocks],
      eq_hi[kStatsBlocks];
};

__device__ __forceinline__ unsigned ord_key(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_val(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__global__ void stats_init_kernel(StatsState* st, int N) {
  const int i = threadIdx.x;
  for (int r = 0; r < 3; ++r) st->hist[r][i] = 0;
  if (i == 0) {
    const int mid = N / 2;
    const int first = (int)ceil(0.05 * N);
    const int last = (int)floor(0.95 * N);
    st->rank[0] = mid;
    st->rank[1] = first;
    st->rank[2] = last;
    for (int r = 0; r < 3; ++r) st->prefix[r] = 0;
  }
}

__global__ void __launch_bounds__(kStatsThreads)
    radix_hist_kernel(const float* __restrict__ v, int N, int shift, StatsState* st) {
  __shared__ unsigned h[3][256];
  for (int r = 0; r < 3; ++r) h[r][threadIdx.x] = 0;
  __syncthreads();
  const unsigned hi_mask = (shift >= 24) ? 0u : (0xFFFFFFFFu << (shift + 8));
  const unsigned p0 = st->prefix[0] & hi_mask, p1 = st->prefix[1] & hi_mask,
                 p2 = st->prefix[2] & hi_mask;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const unsigned k = ord_key(v[i]);
    const unsigned d = (k >> shift) & 255u;
    const unsigned hk = k & hi_mask;
    if (hk == p0) atomicAdd(&h[0][d], 1u);
    if (hk == p1) atomicAdd(&h[1][d], 1u);
    if (hk == p2) atomicAdd(&h[2][d], 1u);
  }
  __syncthreads();
  for (int r = 0; r < 3; ++r)
    if (h[r][threadIdx.x]) atomicAdd(&st->hist[r][threadIdx.x], h[r][threadIdx.x]);
}

__global__ void radix_select_kernel(StatsState* st, int shift) {
  if (threadIdx.x == 0) {
    for (int r = 0; r < 3; ++r) {
      unsigned want = st->rank[r];
      unsigned cum = 0;
      int d = 0;
      for (; d < 256; ++d) {
        const unsigned c = st->hist[r][d];
        if (cum + c > want) break;
        cum += c;
      }
      if (d > 255) d = 255;
      st->prefix[r] |= ((unsigned)d << shift);
      st->rank[r] = want - cum;
    }
  }
  __syncthreads();
  for (int r = 0; r < 3; ++r) st->hist[r][threadIdx.x] = 0;
}

template <typename T>
__device__ __forceinline__ T block_sum(T x, T* sh) {
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const T r = sh[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kStatsThreads)
    moments1_kernel(const float* __restrict__ v, int N, const StatsState* st,
                    StatsPartials* part) {
  __shared__ double shd[kStatsThreads];
  __shared__ unsigned shu[kStatsThreads];
  const float lo = ord_val(st->prefix[1]);
  const float hi = ord_val(st->prefix[2]);
  double s = 0.0;
  unsigned n_in = 0, lt_lo = 0, eq_lo = 0, lt_hi = 0, eq_hi = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const float x = v[i];
    if (x > lo && x < hi) {
      s += (double)x;
      ++n_in;
    }
    lt_lo += (x < lo);
    eq_lo += (x == lo);
    lt_hi += (x < hi);
    eq_hi += (x == hi);
  }
  s = block_sum<double>(s, shd);
  n_in = block_sum<unsigned>(n_in, shu);
  lt_lo = block_sum<unsigned>(lt_lo, shu);
  eq_lo = block_sum<unsigned>(eq_lo, shu);
  lt_hi = block_sum<unsigned>(lt_hi, shu);
  eq_hi = block_sum<unsigned>(eq_hi, shu);
  if (threadIdx.x == 0) {
    part->s[blockIdx.x] = s;
    part->n_in[blockIdx.x] = n_in;
    part->lt_lo[blockIdx.x] = lt_lo;
    part->eq_lo[blockIdx.x] = eq_lo;
    part->lt_hi[blockIdx.x] = lt_hi;
    part->eq_hi[blockIdx.x] = eq_hi;
  }
}

__global__ void moments1_final_kernel(StatsState* st, const StatsPartials* part, int N,
                                      int nblocks) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  unsigned n_in = 0, lt_lo = 0, eq_lo = 0, lt_hi = 0, eq_hi = 0;
  for (int b = 0; b < nblocks; ++b) {
    s += part->s[b];
    n_in += part->n_in[b];
    lt_lo += part->lt_lo[b];
    eq_lo += part->eq_lo[b];
    lt_hi += part->lt_hi[b];
    eq_hi += part->eq_hi[b];
  }
  (void)n_in;
  (void)eq_hi;
  const int first = (int)ceil(0.05 * N);
  const int last = (int)floor(0.95 * N);
  const int cnt = last + 1 - first;
  const float lo = ord_val(st->prefix[1]);
  const float hi = ord_val(st->prefix[2]);
  int k_lo, k_hi;
  if (lo == hi) {
    k_lo = cnt;
    k_hi = 0;
  } else {
    k_lo = min((int)(lt_lo + eq_lo), last + 1) - first;
    k_hi = last + 1 - (int)lt_hi;
  }
  s = __dadd_rn(s, __dadd_rn(__dmul_rn((double)k_lo, (double)lo), __dmul_rn((double)k_hi, (double)hi)));
  st->k_lo = k_lo;
  st->k_hi = k_hi;
  st->mean_sum = s;
  const float sum_f = (float)s;
  st->mean_f = sum_f / (float)cnt;
}

__global__ void __launch_bounds__(kStatsThreads)
    moments2_kernel(const float* __restrict__ v, int N, const StatsState* st,
                    StatsPartials* part) {
  __shared__ double shd[kStatsThreads];
  const float lo = ord_val(st->prefix[1]);
  const float hi = ord_val(st->prefix[2]);
  const double m = (double)st->mean_f;
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const float x = v[i];
    if (x > lo && x < hi) {
      const double d = __dsub_rn((double)x, m);
      acc = __dadd_rn(acc, __dmul_rn(d, d));
    }
  }
  acc = block_sum<double>(acc, shd);
  if (threadIdx.x == 0) part->s[blockIdx.x] = acc;
}

__global__ void moments2_final_kernel(const StatsState* st, const StatsPartials* part, int N,
                                      int nblocks, float* stats_out) {
  if (threadIdx.x != 0) return;
  double acc = 0.0;
  for (int b = 0; b < nblocks; ++b) acc = __dadd_rn(acc, part->s[b]);
  const double m = (double)st->mean_f;
  const float lo = ord_val(st->prefix[1]);
  const float hi = ord_val(st->prefix[2]);
  const double dl = __dsub_rn((double)lo, m), dh = __dsub_rn((double)hi, m);
  acc = __dadd_rn(acc, __dadd_rn(__dmul_rn((double)st->k_lo, __dmul_rn(dl, dl)),
                                 __dmul_rn((double)st->k_hi, __dmul_rn(dh, dh))));
  const int first = (int)ceil(0.05 * N);
  const int last = (int)floor(0.95 * N);
  const int cnt = last + 1 - first;
  stats_out[0] = ord_val(st->prefix[0]);
  stats_out[1] = (float)sqrt(acc / (double)cnt);
}

// ---------------------------------------------------------------------------
// Window merge (dedoppler.cu:273-303): one wavefront per window.
__global__ void __launch_bounds__(256)
    window_merge_kernel(const float* __restrict__ top, const int* __restrict__ blk,
                        const int* __restrict__ off, int N, int window, int nwin,
                        double snr_threshold, const float* __restrict__ stats,
                        int4* __restrict__ out) {
  const int wid = blockIdx.x * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (wid >= nwin) return;
  const float median = stats[0];
  const float sd = stats[1];
  // float path_sum_threshold = snr_threshold * std_dev + median; (double math, no FMA)
  const float thr = (float)__dadd_rn(__dmul_rn(snr_threshold, (double)sd), (double)median);
  const long lo = (long)wid * window;
  const long hi = min(lo + (long)window, (long)N);
  float bv = thr;
  int bi = -1;
  for (long f = lo + lane; f < hi; f += kWave) {
    const float v = top[f];
    if (v > bv) {
      bv = v;
      bi = (int)f;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float ov = __shfl_xor(bv, o);
    const int oi = __shfl_xor(bi, o);
    if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) {
      bv = ov;
      bi = oi;
    }
  }
  if (bi < 0) {
    if (lane == 0) out[wid] = make_int4(-1, 0, 0, 0);
    return;
  }
  const long s0 = max(0L, (long)bi - window + 1);
  const long s1 = min((long)N, (long)bi + window);
  bool larger = false;
  for (long f = s0 + lane; f < s1; f += kWave) larger |= (top[f] > bv);
  larger = __any(larger);
  if (lane == 0) {
    if (larger)
      out[wid] = make_int4(-1, 0, 0, 0);
    else
      out[wid] = make_int4(bi, __float_as_int(bv), blk[bi], off[bi]);
  }
}

__global__ void gather_kernel(const float* __restrict__ sums, int N, const int* __restrict__ offs,
                              const int* __restrict__ chans, int n, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sums[(long)offs[i] * N + chans[i]];
}

// ---------------------------------------------------------------------------
// Normalisation for half-width storage: FP64 mean and deviation over the
// T*N samples, two passes with fixed-order partial reduction.
constexpr int kNormBlocks = 512;

__global__ void __launch_bounds__(256)
    norm_sum_kernel(const float* __restrict__ x, int T, int N, long stride, double* part) {
  __shared__ double sh[256];
  double s = 0.0;
  const long total = (long)T * N;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const long t = i / N, c = i - t * N;
    s += (double)x[t * stride + c];
  }
  s = block_sum<double>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256)
    norm_var_kernel(const float* __restrict__ x, int T, int N, long stride, const double* moments,
                    double* part) {
  __shared__ double sh[256];
  const double m = moments[0];
  double s = 0.0;
  const long total = (long)T * N;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const long t = i / N, c = i - t * N;
    const double d = __dsub_rn((double)x[t * stride + c], m);
    s = __dadd_rn(s, __dmul_rn(d, d));
  }
  s = block_sum<double>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void norm_final_kernel(const double* part, int n, double count, double* moments,
                                  int which) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = __dadd_rn(s, part[i]);
  if (which == 0)
    moments[0] = s / count;
  else
    moments[1] = sqrt(s / count);
}

template <typename Out>
__device__ __forceinline__ Out from_f32(float v);
template <>
__device__ __forceinline__ __half from_f32<__half>(float v) { return __float2half_rn(v); }
template <>
__device__ __forceinline__ __hip_bfloat16 from_f32<__hip_bfloat16>(float v) {
  return __float2bfloat16(v);
}
template <>
__device__ __forceinline__ float from_f32<float>(float v) { return v; }

template <typename Out>
__global__ void norm_apply_kernel(const float* __restrict__ x, int T, int N, long stride,
                                  const double* moments, float clamp, Out* __restrict__ out) {
  const double m = moments[0];
  const double sd = moments[1] > 0.0 ? moments[1] : 1.0;
  const long total = (long)T * N;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const long t = i / N, c = i - t * N;
    float v = (float)((double)x[t * stride + c] - m) / (float)sd;
    v = fminf(fmaxf(v, -clamp), clamp);
    out[i] = from_f32<Out>(v);
  }
}

// ---------------------------------------------------------------------------
// Synthetic waterfall: exponential(1) noise from a counter-based hash.

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void gen_noise_kernel(float* __restrict__ out, int T, int T_r, int N, uint64_t seed,
                                 float gain, float taper_frac, float dc_spike) {
  const long total = (long)T_r * N;
  const int edge = (int)(taper_frac * N);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const int t = (int)(i / N);
    const int c = (int)(i - (long)t * N);
    float v = 0.0f;
    if (t < T) {
      const uint64_t z = mix64(seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull);
      const float u = (float)((z >> 40) + 1) * (1.0f / 16777216.0f);
      v = -logf(u) * gain;
      if (edge > 0) {
        const int d = min(c, N - 1 - c);
        if (d < edge) {
          const float w = 0.5f - 0.5f * cosf(3.14159265358979f * ((float)d + 0.5f) / (float)edge);
          v *= w;
        }
      }
      if (dc_spike > 0.0f && c == N / 2) v = dc_spike * gain;
    }
    out[i] = v;
  }
}

__global__ void gen_tones_kernel(float* __restrict__ out, int T, int N, int num_tones,
                                 const float* __restrict__ start, const float* __restrict__ slope,
                                 const float* __restrict__ amp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  for (int k = 0; k < num_tones; ++k) {
    const float pos = __fadd_rn(start[k], __fmul_rn(slope[k], (float)t));
    const int c = (int)floorf(__fadd_rn(pos, 0.5f));
    if (c >= 0 && c < N) out[(long)t * N + c] += amp[k];
  }
}

inline int blocks_for(long n, int threads, int cap) {
  long b = (n + threads - 1) / threads;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (int)b;
}

template <int T, int C, int G, typename In>
hipError_t launch_fused_typed(const In* x, long stride, int N, int bmin, int nblocks,
                              top_key_t* top_key, float* colsum, hipStream_t stream) {
  using Sh = TileShape<T, C>;
  const long ntiles = ((long)N + Sh::kOut - 1) / Sh::kOut;
  const long ngroups = ((long)nblocks + G - 1) / G;
  // Edge tiles: some drift block has paths that the reference skips
  // (dedoppler.cu:42-46) at the left (c + (T-1)*b < 0) or right (end >= N) end.
  const long bmax = (long)bmin + nblocks - 1;
  long lo = 0;
  if (bmin < 0) lo = std::min(ntiles, ((long)(T - 1) * (-bmin) + Sh::kOut - 1) / Sh::kOut);
  const long thr = (long)N - Sh::kOut + 1 - (long)(T - 1) * (bmax + 1);
  long hi0 = thr <= 0 ? 0 : (thr + Sh::kOut - 1) / Sh::kOut;
  hi0 = std::min(hi0, ntiles);
  if (hi0 < lo) hi0 = lo;
  const long edge_tiles = lo + (ntiles - hi0);
  const long total = ntiles * ngroups;
  if (total > (1L << 30)) return hipErrorInvalidValue;
  const int grid = (int)((total + 7) / 8 * 8);
  hipLaunchKernelGGL((fused_tree_top_kernel<T, C, G, false, In>), dim3(grid), dim3(kWave), 0,
                     stream, x, stride, N, bmin, nblocks, (int)ngroups, (int)lo, (int)hi0,
                     (int)total, top_key, colsum);
  if (edge_tiles > 0) {
    const long etotal = edge_tiles * ngroups;
    const int egrid = (int)((etotal + 7) / 8 * 8);
    hipLaunchKernelGGL((fused_tree_top_kernel<T, C, G, true, In>), dim3(egrid), dim3(kWave), 0,
                       stream, x, stride, N, bmin, nblocks, (int)ngroups, (int)lo, (int)hi0,
                       (int)etotal, top_key, colsum);
  }
  return hipGetLastError();
}

// (C, G) variants compiled for each T: C channels per lane, G drift blocks per
// LDS window.
#define SC_FUSED_VARIANTS(X) \
  X(2, 8, 8)                 \
  X(4, 8, 8)                 \
  X(8, 8, 8)                 \
  X(8, 4, 8)                 \
  X(16, 4, 2)                \
  X(16, 4, 4)                \
  X(16, 4, 8)                \
  X(16, 8, 2)                \
  X(16, 8, 4)

template <typename In>
hipError_t fused_dispatch(const In* x, long stride, int T, int N, int bmin, int nblocks, int C,
                          int G, top_key_t* top_key, float* colsum, hipStream_t stream) {
#define SC_FUSED_CASE(TT, CC, GG)                                                      \
  if (T == TT && C == CC && G == GG)                                                   \
    return launch_fused_typed<TT, CC, GG, In>(x, stride, N, bmin, nblocks, top_key, colsum, \
                                              stream);
  SC_FUSED_VARIANTS(SC_FUSED_CASE)
#undef SC_FUSED_CASE
  return hipErrorInvalidValue;
}

template <int T, typename In>
hipError_t tree_blocks_typed(const In* x, long stride, int Tall, int N, int drift, float* out,
                             hipStream_t stream) {
  constexpr int C = 4;
  using Sh = TileShape<T, C>;
  const int ntiles = (N + Sh::kOut - 1) / Sh::kOut;
  hipLaunchKernelGGL((tree_blocks_kernel<T, C, In>), dim3(ntiles, Tall / T), dim3(kWave), 0,
                     stream, x, stride, N, drift, out);
  return hipGetLastError();
}

template <typename In>
hipError_t tree_blocks_dispatch(const In* x, long stride, int T, int N, int drift, float* out,
                                hipStream_t stream) {
  switch (T) {
    case 2: return tree_blocks_typed<2, In>(x, stride, T, N, drift, out, stream);
    case 4: return tree_blocks_typed<4, In>(x, stride, T, N, drift, out, stream);
    case 8: return tree_blocks_typed<8, In>(x, stride, T, N, drift, out, stream);
    default:
      if (T >= 16 && (T % 16) == 0)
        return tree_blocks_typed<16, In>(x, stride, T, N, drift, out, stream);
      return hipErrorInvalidValue;
  }
}

}  // namespace

bool fused_supported(int T, int C, int G) {
#define SC_FUSED_SUPPORTED(TT, CC, GG) \
  if (T == TT && C == CC && G == GG) return true;
  SC_FUSED_VARIANTS(SC_FUSED_SUPPORTED)
#undef SC_FUSED_SUPPORTED
  return false;
}

hipError_t launch_fused_tree_top(const void* x, int dtype, long stride, int T, int N, int bmin,
                                 int nblocks, int C, int G, top_key_t* top_key, float* colsum,
                                 hipStream_t stream) {
  switch (dtype) {
    case 0:
      return fused_dispatch<float>((const float*)x, stride, T, N, bmin, nblocks, C, G, top_key,
                                   colsum, stream);
    case 1:
      return fused_dispatch<__half>((const __half*)x, stride, T, N, bmin, nblocks, C, G,
                                    top_key, colsum, stream);
    case 2:
      return fused_dispatch<__hip_bfloat16>((const __hip_bfloat16*)x, stride, T, N, bmin,
                                            nblocks, C, G, top_key, colsum, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_tree_blocks(const void* x, int dtype, long stride, int T, int N, int drift,
                              float* out, hipStream_t stream) {
  switch (dtype) {
    case 0: return tree_blocks_dispatch<float>((const float*)x, stride, T, N, drift, out, stream);
    case 1: return tree_blocks_dispatch<__half>((const __half*)x, stride, T, N, drift, out, stream);
    case 2:
      return tree_blocks_dispatch<__hip_bfloat16>((const __hip_bfloat16*)x, stride, T, N, drift,
                                                  out, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_one_step(const float* src, float* dst, int T, int N, int L, int drift,
                           hipStream_t stream) {
  hipLaunchKernelGGL(one_step_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, src, dst, T,
                     N, L, drift);
  return hipGetLastError();
}

hipError_t launch_top_update(const float* sums, int T, int N, int drift, int bmin,
                             top_key_t* top_key, hipStream_t stream) {
  hipLaunchKernelGGL(top_update_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, sums, T, N,
                     drift, bmin, top_key);
  return hipGetLastError();
}

hipError_t launch_colsum(const void* x, int dtype, long stride, int T, int N, float* colsum,
                         hipStream_t stream) {
  const dim3 grid((N + 255) / 256), block(256);
  switch (dtype) {
    case 0:
      hipLaunchKernelGGL(colsum_kernel<float>, grid, block, 0, stream, (const float*)x, stride, T,
                         N, colsum);
      break;
    case 1:
      hipLaunchKernelGGL(colsum_kernel<__half>, grid, block, 0, stream, (const __half*)x, stride,
                         T, N, colsum);
      break;
    case 2:
      hipLaunchKernelGGL(colsum_kernel<__hip_bfloat16>, grid, block, 0, stream,
                         (const __hip_bfloat16*)x, stride, T, N, colsum);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fill_u64(top_key_t* p, long n, top_key_t value, hipStream_t stream) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, stream, p, n,
                     value);
  return hipGetLastError();
}

hipError_t launch_unpack_top(const top_key_t* key, int N, int T, int bmin, float* top_sum,
                             int* top_blk, int* top_off, hipStream_t stream) {
  hipLaunchKernelGGL(unpack_top_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, key, N, T,
                     bmin, top_sum, top_blk, top_off);
  return hipGetLastError();
}

size_t device_stats_scratch_bytes(int /*N*/) {
  return ((sizeof(StatsState) + 255) / 256) * 256 + sizeof(StatsPartials);
}

hipError_t launch_device_stats(const float* colsum, int N, void* scratch, float* stats_out,
                               hipStream_t stream) {
  StatsState* st = (StatsState*)scratch;
  StatsPartials* part =
      (StatsPartials*)((char*)scratch + ((sizeof(StatsState) + 255) / 256) * 256);
  const int nb = blocks_for(N, kStatsThreads, kStatsBlocks);
  hipLaunchKernelGGL(stats_init_kernel, dim3(1), dim3(256), 0, stream, st, N);
  for (int shift = 24; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(kStatsThreads), 0, stream, colsum, N,
                       shift, st);
    hipLaunchKernelGGL(radix_select_kernel, dim3(1), dim3(256), 0, stream, st, shift);
  }
  hipLaunchKernelGGL(moments1_kernel, dim3(nb), dim3(kStatsThreads), 0, stream, colsum, N, st,
                     part);
  hipLaunchKernelGGL(moments1_final_kernel, dim3(1), dim3(64), 0, stream, st, part, N, nb);
  hipLaunchKernelGGL(moments2_kernel, dim3(nb), dim3(kStatsThreads), 0, stream, colsum, N, st,
                     part);
  hipLaunchKernelGGL(moments2_final_kernel, dim3(1), dim3(64), 0, stream, st, part, N, nb,
                     stats_out);
  return hipGetLastError();
}

hipError_t launch_window_merge(const float* top_sum, const int* top_blk, const int* top_off,
                               int N, int window, double snr_threshold, const float* stats,
                               int4* cand_out, hipStream_t stream) {
  const int nwin = (int)(((long)N + window - 1) / window);
  hipLaunchKernelGGL(window_merge_kernel, dim3((nwin + 3) / 4), dim3(256), 0, stream, top_sum,
                     top_blk, top_off, N, window, nwin, snr_threshold, stats, cand_out);
  return hipGetLastError();
}

hipError_t launch_gather(const float* sums, int N, const int* offsets, const int* chans, int n,
                         float* out, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, sums, N, offsets,
                     chans, n, out);
  return hipGetLastError();
}

size_t normalize_scratch_bytes(int /*T*/, int /*N*/) { return sizeof(double) * kNormBlocks; }

hipError_t launch_normalize(const float* x, int T, int N, long stride, int out_dtype, float clamp,
                            void* out, double* moments, void* scratch, hipStream_t stream) {
  double* part = (double*)scratch;
  const double count = (double)T * (double)N;
  hipLaunchKernelGGL(norm_sum_kernel, dim3(kNormBlocks), dim3(256), 0, stream, x, T, N, stride,
                     part);
  hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(64), 0, stream, part, kNormBlocks, count,
                     moments, 0);
  hipLaunchKernelGGL(norm_var_kernel, dim3(kNormBlocks), dim3(256), 0, stream, x, T, N, stride,
                     moments, part);
  hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(64), 0, stream, part, kNormBlocks, count,
                     moments, 1);
  const int nb = blocks_for((long)T * N, 256, 8192);
  switch (out_dtype) {
    case 0:
      hipLaunchKernelGGL(norm_apply_kernel<float>, dim3(nb), dim3(256), 0, stream, x, T, N,
                         stride, moments, clamp, (float*)out);
      break;
    case 1:
      hipLaunchKernelGGL(norm_apply_kernel<__half>, dim3(nb), dim3(256), 0, stream, x, T, N,
                         stride, moments, clamp, (__half*)out);
      break;
    case 2:
      hipLaunchKernelGGL(norm_apply_kernel<__hip_bfloat16>, dim3(nb), dim3(256), 0, stream, x, T,
                         N, stride, moments, clamp, (__hip_bfloat16*)out);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_gen_waterfall(float* out, int T, int T_r, int N, uint64_t seed, float gain,
                                float taper_frac, float dc_spike, int num_tones,
                                const float* tone_start, const float* tone_slope,
                                const float* tone_amp, hipStream_t stream) {
  hipLaunchKernelGGL(gen_noise_kernel, dim3(blocks_for((long)T_r * N, 256, 16384)), dim3(256), 0,
                     stream, out, T, T_r, N, seed, gain, taper_frac, dc_spike);
  if (num_tones > 0)
    hipLaunchKernelGGL(gen_tones_kernel, dim3((T + 63) / 64), dim3(64), 0, stream, out, T, N,
                       num_tones, tone_start, tone_slope, tone_amp);
  return hipGetLastError();
}

}  // namespace sc
