// Device-side building blocks shared by the dedoppler kernels
// (dd_kernels.hip, dd_fused16.hip): sample widening, the register-resident
// Taylor tree, sheared LDS windows, top-path keys and the running best.
// Internal to the library; every includer gets its own copies (anonymous
// namespace), so each translation unit can be built with its own flags.
//
// Reference semantics: taylor.h:30-118, taylor.cu:170-254, dedoppler.cu:32-55.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <math.h>
#include <stdint.h>

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
// A key packs (path sum, code) into 64 bits whose order is the lexicographic
// order of (sum, -code): the float's bit pattern, sign-extended and shifted
// left by 29, with kCodeMask - code in the low 29 bits, where
// code = (drift_block - min_drift_block) * T_r + path_offset. Keys of positive
// sums are non-negative integers below 2^60, and read as FP64 they are finite
// non-negative numbers in the same order, so the running max is one v_max_f64;
// keys of negative sums read as negative numbers. The max of keys is the
// largest sum and, among equal sums, the lowest (drift block, offset) - the
// reference's strict ">" in ascending order (dedoppler.cu:40,49,214). Keys of
// sums <= 0 never exceed kEmptyKey, the key of (0, largest code), matching
// "only sums > 0 register" (dedoppler.cu:205). Positive keys order the same way
// as unsigned integers, so partial results merge with a 64-bit atomic max;
// 0 means "no positive path". (A NaN sample would make a finite key: inputs
// must be finite, as power always is.)
constexpr unsigned long long kCodeMask = 0x1FFFFFFFull;
constexpr unsigned long long kEmptyKey = kCodeMask;

__device__ __forceinline__ double make_key(float s, unsigned low_bits) {
  const int bits = __float_as_int(s);
  const unsigned hi = (unsigned)(bits >> 3);            // v_ashrrev_i32
  const unsigned lo = ((unsigned)bits << 29) | low_bits;  // v_lshl_or_b32
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ float key_sum(unsigned long long k) {
  return __uint_as_float((unsigned)(k >> 29));
}

// Block keys (device-statistics searches with the one-pass kernels):
// (sum bits, ~block) with block = drift_block - min_drift_block, for sums > 0
// only. Their unsigned order is (sum, -block), so a 64-bit atomic max keeps
// the largest sum and, among equal sums, the lowest block; the window merge
// reads the sums, and candidate_offsets_kernel finds the offset of each
// window's candidate.
__device__ __forceinline__ unsigned long long block_key(float s, int block) {
  return ((unsigned long long)__float_as_uint(s) << 32) | (unsigned)~(unsigned)block;
}
__device__ __forceinline__ float block_key_sum(unsigned long long k) {
  return __uint_as_float((unsigned)(k >> 32));
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
// (no bank conflicts). Drift block g = C*h + g' of the window reads element
// C*lane + j + t*g = sub-row (j + t*g') % C, column lane + (j + t*g') / C + t*h:
// with g' unrolled and the t*h term folded into a per-row base that advances
// by t per h step, every read is base + a compile-time offset.
template <int T, int C, int G>
struct Window {
  static_assert(G % C == 0, "drift blocks per window must be a multiple of C");
  // Sub-row length, rounded so that a row is a whole number of 64-lane loads.
  static constexpr int S0 = (kWave * C + (T - 1) * (G - 1) + C - 1) / C;
  static constexpr int kStep = kWave / C;
  static constexpr int S = (S0 + kStep - 1) / kStep * kStep;
  static constexpr int kRow = C * S;
  static constexpr int kLoads = kRow / kWave;  // loads per lane per row
  static constexpr int kFloats = T * kRow;
};

typedef __attribute__((address_space(3))) float lds_float;

// Window rows are fetched into registers (fetch_window) and then written to
// LDS (store_window), so a window costs one memory latency and the next
// window's loads can be in flight while the current one is processed. Rows
// whose span lies inside [0, N) take unguarded loads off one row pointer; rows
// crossing an edge clamp the address and substitute 0 for out-of-range samples.
template <int T, int C, int G, typename In, int L>
__device__ __forceinline__ void fetch_window(float (&tmp)[T][L], const In* __restrict__ x,
                                             long stride, int N, int c_w, int b0, int row0) {
  using Wn = Window<T, C, G>;
  static_assert(L == Wn::kLoads, "register staging size");
  const int lane = threadIdx.x;
  // Row t starts at c_w + t*b0, monotonic in t: the first and last rows bound
  // the window, so one check covers every row and the pointer advances by
  // stride + b0 per row.
  const int g_lo = b0 >= 0 ? c_w : c_w + (T - 1) * b0;
  const int g_hi = b0 >= 0 ? c_w + (T - 1) * b0 : c_w;
  if (g_lo >= 0 && g_hi + Wn::kRow <= N) {
    const In* src = x + (long)row0 * stride + c_w + lane;
    const long step = stride + b0;
#pragma unroll
    for (int t = 0; t < T; ++t) {
#pragma unroll
      for (int i = 0; i < Wn::kLoads; ++i) tmp[t][i] = to_f32<In>(src[i * kWave]);
      src += step;
    }
    return;
  }
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int g0 = c_w + t * b0;
    const In* row = x + (long)(row0 + t) * stride;
    if (g0 >= 0 && g0 + Wn::kRow <= N) {
      const In* src = row + g0 + lane;
#pragma unroll
      for (int i = 0; i < Wn::kLoads; ++i) tmp[t][i] = to_f32<In>(src[i * kWave]);
    } else {
#pragma unroll
      for (int i = 0; i < Wn::kLoads; ++i) {
        const int gc = g0 + i * kWave + lane;
        const bool ok = (unsigned)gc < (unsigned)N;
        const float val = to_f32<In>(row[ok ? gc : 0]);
        tmp[t][i] = ok ? val : 0.0f;
      }
    }
  }
}

template <int T, int C, int G, int L>
__device__ __forceinline__ void store_window(lds_float* win, const float (&tmp)[T][L]) {
  using Wn = Window<T, C, G>;
  static_assert(L == Wn::kLoads, "register staging size");
  const int lane = threadIdx.x;
  lds_float* dst0 = win + (lane % C) * Wn::S + lane / C;
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int i = 0; i < Wn::kLoads; ++i) dst0[t * Wn::kRow + i * (kWave / C)] = tmp[t][i];
}

template <int T, int C, int G, typename In>
__device__ __forceinline__ void load_window(lds_float* win, const In* __restrict__ x,
                                            long stride, int N, int c_w, int b0, int row0) {
  float tmp[T][Window<T, C, G>::kLoads];
  fetch_window<T, C, G, In>(tmp, x, stride, N, c_w, b0, row0);
  store_window<T, C, G>(win, tmp);
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

// Reads drift block g' (of the current h step) of the window into registers.
template <int T, int C, int G, int gp>
__device__ __forceinline__ void read_block(lds_float* const (&rp)[T], float (&v)[T][C]) {
  using Wn = Window<T, C, G>;
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int k = j + t * gp;
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

// Running best of the T_r <= 16 kernel, per channel held by a lane. Full
// keys (BK false): the key max in ascending (block, offset) order, one
// v_max_f64 per path sum. Block keys (BK true): the largest sum and its block
// only - a block replaces the best when its largest sum is strictly larger
// (blocks come in ascending order), which is the reference's rule
// (dedoppler.cu:40,49,214) up to the offset within the block; the offset of a
// hit is recovered afterwards (candidate_offsets_kernel). About a third less
// VALU work per path sum.
template <int C, bool BK>
struct Best;
template <int C>
struct Best<C, false> {
  double key[C];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int j = 0; j < C; ++j) key[j] = __longlong_as_double((long long)kEmptyKey);
  }
  __device__ __forceinline__ unsigned long long merge_key(int j) const {
    const unsigned long long k = (unsigned long long)__double_as_longlong(key[j]);
    return k > kEmptyKey ? k : 0;
  }
};
template <int C>
struct Best<C, true> {
  float sum[C];
  int block[C];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int j = 0; j < C; ++j) {
      sum[j] = 0.0f;
      block[j] = 0;
    }
  }
  __device__ __forceinline__ unsigned long long merge_key(int j) const {
    return sum[j] > 0.0f ? block_key(sum[j], block[j]) : 0;
  }
};

// One drift block (block index `block` from the minimum): tree, edge masking
// when the block is not interior for this tile (dedoppler.cu:40-46, including
// the early return that skips every offset of a block whose first path ends
// left of channel 0; skipped paths become 0 and never register, as only sums
// > 0 do, dedoppler.cu:205), then the running best.
template <int T, int C, int G, bool BK, int gp>
__device__ __forceinline__ void fused_block(lds_float* const (&rp)[T], int g, int ng, int b0,
                                            int c_w, int N, int block0, Best<C, BK>& best) {
  if constexpr (gp < C) {
    if (g + gp < ng) {
      const int b = b0 + g + gp;
      float v[T][C];
      read_block<T, C, G, gp>(rp, v);
      tree_regs<T, C>(v);
      if (!block_interior<T, C>(c_w, b, N)) {
        const int lane = threadIdx.x;
#pragma unroll
        for (int j = 0; j < C; ++j) {
          const int c = c_w + lane * C + j;
          const int last0 = c + (T - 1) * b;
          const int pmax = (c >= N || last0 < 0) ? 0 : min(T, N - last0);
#pragma unroll
          for (int p = 0; p < T; ++p)
            if (p >= pmax) v[p][j] = 0.0f;
        }
      }
      const int block = block0 + g + gp;
      if constexpr (BK) {
#pragma unroll
        for (int j = 0; j < C; ++j) {
          float m = v[0][j];
#pragma unroll
          for (int p = 1; p < T; ++p) m = fmaxf(m, v[p][j]);
          if (m > best.sum[j]) {
            best.sum[j] = m;
            best.block[j] = block;
          }
        }
      } else {
        const unsigned lowg = (unsigned)kCodeMask - (unsigned)(block * T);
#pragma unroll
        for (int p = 0; p < T; ++p) {
          const unsigned low = lowg - (unsigned)p;
#pragma unroll
          for (int j = 0; j < C; ++j) best.key[j] = fmax(best.key[j], make_key(v[p][j], low));
        }
      }
    }
    fused_block<T, C, G, BK, gp + 1>(rp, g, ng, b0, c_w, N, block0, best);
  }
}

// (C, G) variants compiled for each T: C channels per lane, G drift blocks per
// LDS window.
#define SC_FUSED_VARIANTS(X) \
  X(2, 8, 16, 0)             \
  X(4, 8, 16, 0)             \
  X(8, 8, 16, 0)             \
  X(8, 4, 16, 0)             \
  X(16, 4, 4, 0)             \
  X(16, 4, 8, 0)             \
  X(16, 4, 16, 0)            \
  X(16, 4, 4, 1)             \
  X(16, 4, 8, 1)             \
  X(16, 4, 4, 2)             \
  X(16, 4, 8, 2)             \
  X(16, 4, 16, 2)            \
  X(16, 8, 8, 0)

}  // namespace

}  // namespace sc
