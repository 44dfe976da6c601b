// HIP kernels for the MI355X (gfx950) Taylor-tree dedoppler search.
//
// Reference semantics: taylor.h:30-118, taylor.cu:97-350, dedoppler.cu:32-72 and
// dedoppler.cu:186-320 of lacker/seticore. The dataflow here is different: the
// reference runs one tree launch plus one findTopPathSums launch per drift block,
// each streaming the whole waterfall through HBM; here one launch covers the
// whole drift range, each wavefront keeps a channel tile's tree in registers and
// reuses one LDS window across a group of drift blocks, and the max/argmax is
// folded into the same kernel. Float association of every path sum is the
// reference's pairwise recursion, so path sums are bit-identical. The shared
// device building blocks are in dd_device.h; the T_r <= 16 kernel is in
// dd_fused16.hip.
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off
//        -Xclang -target-feature -Xclang -packed-fp32-ops

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "dd_device.h"
#include "dd_internal.h"

namespace sc {

namespace {

// Order-preserving map of a float onto uint32 (used by the statistics).
__device__ __forceinline__ unsigned ord_key(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_val(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Level-L entry (tb, p) of drift block b's tree at start channel c, added in
// the register tree's order: entry (2tb, p/2) at c plus entry (2tb+1, p/2) at
// c + (p+1)/2 (tree_stage); level 1 is row tb at c + tb*b (the shear), 0
// outside [0, N) as in the staged windows.
template <int L, typename In>
__device__ __forceinline__ float tree_entry(const In* __restrict__ x, long stride, int N, int b,
                                            int tb, int p, int c) {
  if constexpr (L == 1) {
    const long col = (long)c + (long)tb * b;
    return (col >= 0 && col < N) ? to_f32<In>(x[(long)tb * stride + col]) : 0.0f;
  } else {
    const float lo = tree_entry<L / 2, In>(x, stride, N, b, 2 * tb, p >> 1, c);
    const float hi = tree_entry<L / 2, In>(x, stride, N, b, 2 * tb + 1, p >> 1, c + ((p + 1) >> 1));
    return lo + hi;
  }
}

// Path (b, p) at start channel c of a T-row tree (any T up to 256), in the
// tree's order: T/16 level-16 entries (tree_entry), each at the channel its
// position in the upper levels puts it (the second half of a level-L entry
// with offset q starts (q+1)/2 channels on), then summed pairwise level by
// level. The level-16 entries are formed one at a time to keep few loads in
// flight per lane.
template <int T, typename In>
__device__ float path_sum_exact(const In* __restrict__ x, long stride, int N, int b, int p,
                                int c) {
  if constexpr (T <= 16) {
    return tree_entry<T, In>(x, stride, N, b, 0, p, c);
  } else {
    constexpr int K = T / 16;
    constexpr int kLogK = K == 2 ? 1 : K == 4 ? 2 : K == 8 ? 3 : 4;
    static_assert((1 << kLogK) == K, "T is a power of two up to 256");
    float r[K];
#pragma unroll 1
    for (int n = 0; n < K; ++n) {
      int cn = c, pl = p;
      for (int lvl = kLogK - 1; lvl >= 0; --lvl) {
        if ((n >> lvl) & 1) cn += (pl + 1) >> 1;
        pl >>= 1;
      }
      // Rows 16n.. in the whole tree's shear: channel cn + 16n*b at row 16n.
      r[n] = tree_entry<16, In>(x + (long)16 * n * stride, stride, N, b, 0, pl, cn + 16 * n * b);
    }
#pragma unroll
    for (int width = K; width > 1; width >>= 1)
#pragma unroll
      for (int i = 0; i < width / 2; ++i) r[i] = r[2 * i] + r[2 * i + 1];
    return r[0];
  }
}

// Offset of each window's candidate after a block-key search: the candidate
// record holds its channel, sum and drift block; the offset is the first
// (lowest) one whose path sum in that block equals the sum (the reference
// keeps the first of equal sums, dedoppler.cu:49). Also completes the
// unpacked offset of that channel. min(T, 64) lanes per window, one offset
// each per round.
template <int T, typename In>
__device__ __forceinline__ void candidate_offsets_body(const In* __restrict__ x, long stride,
                                                       int N, int nwin, int4* __restrict__ cand,
                                                       int* __restrict__ top_off) {
  constexpr int kLanes = T < kWave ? T : kWave;
  constexpr int kPerWave = kWave / kLanes;
  const int slot = threadIdx.x / kLanes;
  const int lane = threadIdx.x % kLanes;
  const int w = blockIdx.x * kPerWave + slot;
  const unsigned long long group_mask = kLanes == kWave ? ~0ull : ((1ull << kLanes) - 1);
  int4 r = make_int4(-1, 0, 0, 0);
  int pmax = 0;
  // A candidate whose top sum is the zero the search starts from (no positive
  // path; only when the threshold is below zero) keeps the block and offset
  // the merge read from the top arrays, as the reference reports them
  // (dedoppler.cu:305-306).
  bool empty = false;
  if (w < nwin) {
    r = cand[w];
    empty = r.x >= 0 && r.y == 0;
    if (r.x >= 0 && !empty) {
      const int last0 = r.x + (T - 1) * r.z;
      pmax = last0 < 0 ? 0 : min(T, N - last0);
    }
  }
  int p_best = -1;
  for (int p0 = 0; p0 < T; p0 += kLanes) {  // uniform trip count: every lane takes the ballot
    const int p = p0 + lane;
    const bool match =
        p < pmax && path_sum_exact<T, In>(x, stride, N, r.z, p, r.x) == __int_as_float(r.y);
    const unsigned long long mine = (__ballot(match) >> (slot * kLanes)) & group_mask;
    if (p_best < 0 && mine) p_best = p0 + __ffsll((long long)mine) - 1;
  }
  if (lane == 0 && w < nwin && r.x >= 0 && !empty) {
    // -1: no offset reproduces the recorded sum (the tree kernels and this
    // re-addition disagree on an addition order); the host reports an error.
    r.w = p_best;
    cand[w] = r;
    top_off[r.x] = r.w;
  }
}

template <int T, typename In>
__global__ void __launch_bounds__(kWave)
    candidate_offsets_kernel(const In* __restrict__ x, long stride, int N, int nwin,
                             int4* __restrict__ cand, int* __restrict__ top_off) {
  candidate_offsets_body<T, In>(x, stride, N, nwin, cand, top_off);
}

// Several channels (grid.y = channel; each checks its own window count).
template <int T, typename In>
__global__ void __launch_bounds__(kWave) candidate_offsets_set_kernel(TailSet ts, long stride,
                                                                      int N) {
  const int c = blockIdx.y;
  const int nwin = (int)(((long)N + ts.window[c] - 1) / ts.window[c]);
  candidate_offsets_body<T, In>((const In*)ts.x[c], stride, N, nwin, ts.cand[c], ts.top_off[c]);
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
  using Wn = Window<T, C, C>;
  __shared__ float win_storage[Wn::kFloats];
  lds_float* win = (lds_float*)win_storage;
  const int c_w = blockIdx.x * Sh::kOut;
  const int tb = blockIdx.y;
  load_window<T, C, C, In>(win, x, stride, N, c_w, drift, tb * T);
  __syncthreads();
  lds_float* rp[T];
  row_bases<T, C, C>(win, rp);
  float v[T][C];
  read_block<T, C, C, 0>(rp, v);
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

// ---------------------------------------------------------------------------
// T_r = 32 in one pass: both 16-row time blocks' trees in registers, the top
// level through LDS.
//
// In the sheared coordinates of the whole 32-row window (row t at c + t*b),
// level 32 combines offset q of the first block's tree at channel c with
// offset q of the second block's tree at channel c + q + r (p = 2q + r): the
// reference's upper-half shift (p+1)/2 + b*16 (taylor.cu:97-109 with
// taylor.h:53-54), with b*16 absorbed by the shear. The second tree is
// computed first and parked in LDS, de-interleaved by C like the window, so
// each lane reads its shifted neighbours conflict-free; the first tree stays
// in registers. Outputs are exact for lanes whose shifted reads stay inside
// the second tree's exact lanes (56 of 64 at C = 4: 224 channels per wave).

template <int C>
struct Tile32 {
  static constexpr int kValid16 = TileShape<16, C>::kValidLanes;
  static constexpr int kValidLanes = kValid16 - 1 - 15 / C;
  static constexpr int kOut = kValidLanes * C;
  static constexpr int kHS = kWave + (C + 16 + C - 1) / C;  // LDS row length of the parked tree
};

// Rows R0 .. R0+NR-1 of drift block gp of the current h step.
template <int T, int C, int G, int gp, int R0, int NR>
__device__ __forceinline__ void read_rows(lds_float* const (&rp)[T], float (&v)[NR][C]) {
  using Wn = Window<T, C, G>;
#pragma unroll
  for (int t = 0; t < NR; ++t)
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int k = j + (R0 + t) * gp;
      v[t][j] = rp[R0 + t][(k % C) * Wn::S + k / C];
    }
}

// Orders this wave's LDS accesses (the parked tree is private to the wave):
// the workgroup-scope fences wait for the wave's outstanding LDS operations.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int C, int G, int gp>
__device__ __forceinline__ void fused32_block(lds_float* const (&rp)[32], lds_float* hp, int g,
                                              int ng, int b0, int c_w, int N, unsigned low0,
                                              double (&best)[C]) {
  if constexpr (gp < C) {
    if (g + gp < ng) {
      using Tl = Tile32<C>;
      const int b = b0 + g + gp;
      float v[16][C];
      read_rows<32, C, G, gp, 16, 16>(rp, v);
      tree_regs<16, C>(v);
      wave_lds_sync();  // the previous block's reads of the parked tree are done
#pragma unroll
      for (int q = 0; q < 16; ++q)
#pragma unroll
        for (int j = 0; j < C; ++j) hp[(q * C + j) * Tl::kHS] = v[q][j];
      read_rows<32, C, G, gp, 0, 16>(rp, v);
      tree_regs<16, C>(v);
      wave_lds_sync();
      const int lane = threadIdx.x % kWave;
      const bool interior = (c_w + 31 * b >= 0) && (c_w + Tl::kOut - 1 < N) &&
                            (c_w + Tl::kOut - 1 + 31 * b + 31 < N);
      const unsigned lowg = low0 - (unsigned)((g + gp) * 32);
#pragma unroll
      for (int q = 0; q < 16; ++q)
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          float sum[C];
#pragma unroll
          for (int j = 0; j < C; ++j) {
            const int k = j + q + r;
            sum[j] = v[q][j] + hp[(q * C + k % C) * Tl::kHS + k / C];
          }
          if (!interior) {
#pragma unroll
            for (int j = 0; j < C; ++j) {
              const int c = c_w + lane * C + j;
              const int last0 = c + 31 * b;
              const int pmax = (c >= N || last0 < 0) ? 0 : min(32, N - last0);
              if (2 * q + r >= pmax) sum[j] = 0.0f;
            }
          }
#pragma unroll
          for (int j = 0; j < C; ++j)
            best[j] = fmax(best[j], make_key(sum[j], lowg - (unsigned)(2 * q + r)));
        }
    }
    fused32_block<C, G, gp + 1>(rp, hp, g, ng, b0, c_w, N, low0, best);
  }
}

// One workgroup of W waves per (channel tile of Tile32<C>::kOut channels, group
// of G drift blocks): the waves share the group's LDS window and take turns
// over its drift blocks (wave w: h steps w, w + W, ...), each with its own
// parked tree; their key maxima merge through the atomics. The workgroup
// whose group holds block 0 writes the column sums (32 rows, sequential
// order, dedoppler.cu:68-71).
template <int C, int G, int W, typename In>
__global__ void __launch_bounds__(kWave * W)
    fused32_kernel(const In* __restrict__ x, long stride, int N, int bmin, int nblocks,
                   int ngroups, int total_items, unsigned long long* __restrict__ top_key,
                   float* __restrict__ colsum) {
  static_assert(G % (C * W) == 0, "each wave takes whole h steps");
  using Tl = Tile32<C>;
  using Wn = Window<32, C, G>;
  __shared__ float win_storage[Wn::kFloats];
  __shared__ float park[W][16 * C * Tl::kHS];
  lds_float* win = (lds_float*)win_storage;
  const int per_xcd = gridDim.x / 8;
  const int item = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (item >= total_items) return;
  const int tile = item / ngroups;
  const int grp = item - tile * ngroups;
  const int c_w = tile * Tl::kOut;
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int b0 = bmin + grp * G;
  const int ng = min(G, nblocks - grp * G);
  if (wave == 0) load_window<32, C, G, In>(win, x, stride, N, c_w, b0, 0);
  if (wave == 0 && colsum && b0 <= 0 && 0 < b0 + ng && lane < Tl::kValidLanes) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int c = c_w + lane * C + j;
      if (c < N) {
        float s = 0.0f;
#pragma unroll
        for (int t = 0; t < 32; ++t) s += to_f32<In>(x[(long)t * stride + c]);
        colsum[c] = s;
      }
    }
  }
  __syncthreads();
  lds_float* rp[32];
#pragma unroll
  for (int t = 0; t < 32; ++t) {
    lds_float* p = win + t * Wn::kRow + lane + t * wave;  // this wave's first h step
    asm volatile("" : "+v"(p));
    rp[t] = p;
  }
  lds_float* hp = (lds_float*)park[wave] + lane;
  asm volatile("" : "+v"(hp));
  double best[C];
#pragma unroll
  for (int j = 0; j < C; ++j) best[j] = __longlong_as_double((long long)kEmptyKey);
  const unsigned low0 = (unsigned)kCodeMask - (unsigned)(grp * G * 32);
  for (int g = C * wave; g < ng; g += C * W) {
    fused32_block<C, G, 0>(rp, hp, g, ng, b0, c_w, N, low0, best);
#pragma unroll
    for (int t = 1; t < 32; ++t) rp[t] += t * W;
  }
  if (lane < Tl::kValidLanes) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int c = c_w + lane * C + j;
      const unsigned long long k = (unsigned long long)__double_as_longlong(best[j]);
      if (c < N && k > kEmptyKey) atomicMax(&top_key[c], k);
    }
  }
}

// ---------------------------------------------------------------------------
// T_r = 16*M (M = 2..16): two passes through a scratch array.
//
// Unrolling the recursion (taylor.h:53-54,80-82) from level T down to level
// 16, both halves of every stage read the same offset p/2, so output offset
// p = q*M + r (q < 16, r < M) reads offset q of the level-16 tree of each of
// the M 16-row time blocks, and the shift of level 16*2^j,
// (p_j + 1)/2 with p_j = q*2^j + (r >> (log2 M - j)), is q*2^(j-1) plus the
// shift of an M-row tree at offset r >> (log2 M - j). So with
//   R_q[k][c] = level-16 tree of time block k, offset q, start channel
//               c + k*(16*b + q)      (time block k's own coordinates,
//               sheared by 16*k*b as the reference's one-step kernel does with
//               drift*L/2, taylor.cu:97-109, plus k*q),
// path (b, q*M + r, c) is offset r at channel c of the M-row tree over
// R_q[0..M-1] at drift 0, added in the same pairwise order as the reference.
// Pass 1 (level16_kernel) writes R_q for a group of drift blocks; pass 2
// (upper_tree_kernel) runs the M-row trees from registers, masks invalid
// paths (dedoppler.cu:40-46) and folds the keys.
//
// Scratch layout: row ((gi*16 + q)*M + k) of row_pitch floats holds R_q[k] of
// drift block b_first + gi, starting (k*q) & 3 floats in. Entries that no
// valid path reads are not written.

template <int C, int G, bool NT, int gp>
__device__ __forceinline__ void level16_block(lds_float* const (&rp)[16], int g, int ng, int b0,
                                              int gi0, int k, int c_w, int N, int M,
                                              long row_pitch, float* __restrict__ scratch) {
  if constexpr (gp < C) {
    if (g + gp < ng) {
      const int b = b0 + g + gp;
      float v[16][C];
      read_block<16, C, G, gp>(rp, v);
      tree_regs<16, C>(v);
      const int lane = threadIdx.x;
      const int c = c_w + lane * C;
      if (lane < TileShape<16, C>::kValidLanes && c < N) {
        float* base = scratch + (long)((gi0 + g + gp) * 16 * M + k) * row_pitch;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          // Stored (k*q) & 3 columns to the right of its place, so that every
          // lane's store is 16-byte aligned (c and 16*k*b are multiples of 4);
          // the second pass reads with that offset.
          const long col = (long)c - (long)k * (16 * b + q) + ((k * q) & 3);
          float* dst = base + (long)q * M * row_pitch;
          if (col >= 0 && col + C <= row_pitch) {
            if constexpr (C == 4) {
              typedef float f4a __attribute__((ext_vector_type(4)));
              const f4a w = {v[q][0], v[q][1], v[q][2], v[q][3]};
              if constexpr (NT)
                __builtin_nontemporal_store(w, (f4a*)(dst + col));
              else
                *(f4a*)(dst + col) = w;
            } else {
#pragma unroll
              for (int j = 0; j < C; ++j) dst[col + j] = v[q][j];
            }
          } else {
#pragma unroll
            for (int j = 0; j < C; ++j)
              if (col + j >= 0 && col + j < row_pitch) dst[col + j] = v[q][j];
          }
        }
      }
    }
    level16_block<C, G, NT, gp + 1>(rp, g, ng, b0, gi0, k, c_w, N, M, row_pitch, scratch);
  }
}

// Pass 1. One 64-lane workgroup per (channel tile, time block k, group of G
// drift blocks): the 16-row window of time block k is staged in LDS once for
// the group, as in fused_tree_top_kernel. NT: streaming (nontemporal) scratch
// stores, and loads in the second pass - measured faster at T_r = 64, slower
// at 128 and 256.
template <int C, int G, bool NT, typename In>
__global__ void __launch_bounds__(kWave)
    level16_kernel(const In* __restrict__ x, long stride, int N, int M, int b_first, int nb,
                   int ntiles, int total_items, long row_pitch, float* __restrict__ scratch) {
  using Sh = TileShape<16, C>;
  using Wn = Window<16, C, G>;
  __shared__ float win_storage[Wn::kFloats];
  lds_float* win = (lds_float*)win_storage;
  const int per_xcd = gridDim.x / 8;
  const int item = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (item >= total_items) return;
  const int tile = item % ntiles;
  const int rest = item / ntiles;
  const int k = rest % M;
  const int grp = rest / M;
  const int c_w = tile * Sh::kOut;
  const int gi0 = grp * G;
  const int b0 = b_first + gi0;
  const int ng = min(G, nb - gi0);
  load_window<16, C, G, In>(win, x, stride, N, c_w, b0, 16 * k);
  __syncthreads();
  lds_float* rp[16];
  row_bases<16, C, G>(win, rp);
  for (int g = 0; g < ng; g += C) {
    level16_block<C, G, NT, 0>(rp, g, ng, b0, gi0, k, c_w, N, M, row_pitch, scratch);
#pragma unroll
    for (int t = 1; t < 16; ++t) rp[t] += t;
  }
}

constexpr int kUpperWaves = 4;

// Pass 2. One workgroup of kUpperWaves waves per channel tile; the waves share
// the (drift block, q) pairs of the group, keep their own key maxima, and
// merge them through LDS before one atomic max per channel.
template <int M, int C>
__global__ void __launch_bounds__(kWave * kUpperWaves)
    upper_tree_kernel(const float* __restrict__ scratch, long row_pitch, int T, int N, int bmin,
                      int b_first, int nb, int total_items, top_key_t* __restrict__ top_key) {
  using Sh = TileShape<M, C>;
  __shared__ double red[kUpperWaves - 1][kWave * C];
  const int per_xcd = gridDim.x / 8;
  const int item = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (item >= total_items) return;
  const int c_w = item * Sh::kOut;
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  double best[C];
#pragma unroll
  for (int j = 0; j < C; ++j) best[j] = __longlong_as_double((long long)kEmptyKey);
  const bool tile_in = c_w + Sh::kOut - 1 < N;
  for (int pair = wave; pair < nb * 16; pair += kUpperWaves) {
    const int gi = pair / 16;
    const int q = pair % 16;
    const int b = b_first + gi;
    const float* src = scratch + (long)pair * M * row_pitch + c_w + lane * C;
    float v[M][C];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      if constexpr (C == 4) {
        typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
        const f4u* a = (const f4u*)(src + (long)k * row_pitch + ((k * q) & 3));
        f4u w;
        if constexpr (M <= 4)
          w = __builtin_nontemporal_load(a);
        else
          w = *a;
        v[k][0] = w.x;
        v[k][1] = w.y;
        v[k][2] = w.z;
        v[k][3] = w.w;
      } else {
#pragma unroll
        for (int j = 0; j < C; ++j) v[k][j] = src[(long)k * row_pitch + ((k * q) & 3) + j];
      }
    }
    tree_regs<M, C>(v);
    const bool interior = tile_in && (c_w + (T - 1) * b >= 0) &&
                          (c_w + Sh::kOut - 1 + (T - 1) * b + (T - 1) < N);
    if (!interior) {
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const int c = c_w + lane * C + j;
        const int last0 = c + (T - 1) * b;
        const int pmax = (c >= N || last0 < 0) ? 0 : min(T, N - last0);
#pragma unroll
        for (int r = 0; r < M; ++r)
          if (q * M + r >= pmax) v[r][j] = 0.0f;
      }
    }
    const unsigned low0 = (unsigned)kCodeMask - (unsigned)((b - bmin) * T + q * M);
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const unsigned low = low0 - (unsigned)r;
#pragma unroll
      for (int j = 0; j < C; ++j) best[j] = fmax(best[j], make_key(v[r][j], low));
    }
  }
  if (wave > 0) {
#pragma unroll
    for (int j = 0; j < C; ++j) red[wave - 1][lane * C + j] = best[j];
  }
  __syncthreads();
  if (wave == 0 && lane < Sh::kValidLanes) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
#pragma unroll
      for (int w = 0; w < kUpperWaves - 1; ++w) best[j] = fmax(best[j], red[w][lane * C + j]);
      const int c = c_w + lane * C + j;
      const unsigned long long key = (unsigned long long)__double_as_longlong(best[j]);
      if (c < N && key > kEmptyKey) atomicMax(&top_key[c], key);
    }
  }
}

// ---------------------------------------------------------------------------
// T_r = 64..256 in one pass: the level-16 trees staged in LDS by groups of
// offsets.
//
// With the decomposition above, a workgroup takes one channel tile and, in
// turn, each drift block of its chunk and each group (or, split, one group)
// of QG = 64/M consecutive level-16 offsets q0..q0+QG-1: phase 1 computes
// R_q[k] (q in the group, every time block k: QG*M = 64 rows) into LDS,
// phase 2 runs the M-row trees of those offsets from LDS and folds the keys.
// Nothing goes through HBM but the input. Offsets of a group differ by at most QG-1, so time block k's trees
// are needed over only Wr + (QG-1)*k channels, which one wave's 240 valid
// level-16 outputs cover for every k. Offset q of level 16 needs offset q >> 1
// of level 8, q >> 2 of level 4 and q >> 3 of level 2, so a group's partial
// tree takes at most 5 adds per output (4 for the whole tree), in exactly the
// whole tree's order.
template <int M>
struct Staged {
  static constexpr int C = 4;
  static constexpr int T = 16 * M;
  static constexpr int QG = M == 2 ? 16 : 64 / M;  // level-16 offsets per group
  static constexpr int kGroups = 16 / QG;
  static constexpr int kWaves = M == 2 ? 2 : 4;
  static constexpr int kLevel16Out = TileShape<16, C>::kOut;   // 240 channels
  static constexpr int kLevel16Lanes = TileShape<16, C>::kValidLanes;
  static constexpr int kWr = kLevel16Out - (QG - 1) * (M - 1);  // R columns valid for every k
  static constexpr int kOut = (kWr - (M - 1)) / C * C;          // output channels per tile
  // LDS row of R: columns stored de-interleaved by C (column s at sub-row
  // s % C, position s / C), so that phase 1's shifted writes and phase 2's
  // reads are both one word per lane at consecutive addresses.
  // Only positions (lanes) below kUsed feed a valid output (the tile's
  // outputs and the M-1 channels after them), so a sub-row holds kSub >= kUsed
  // positions; lanes beyond read into the next sub-row (or the pad after the
  // last row) and produce discarded outputs. 52 KiB, three workgroups per CU
  // (M = 2: 32 rows, 29 KiB).
  static constexpr int kUsed = (kOut + M - 1 + C - 1) / C;
  static constexpr int kSub = (kUsed + 3) / 4 * 4;
  static constexpr int kRow = C * kSub;
  static constexpr int kFloats = QG * M * kRow + kWave;
  // Waves per SIMD the register allocation must allow: LDS admits three
  // workgroups (12 waves) per CU. Measured: 168 VGPRs (one spilled) at M = 16
  // beat 197 at two waves; M = 4 keeps its 237 registers.
  static constexpr int kMinWaves = M <= 4 ? 2 : 3;
  static_assert(kOut <= TileShape<M, C>::kOut, "phase 2 tile");
  static_assert(kUsed <= kLevel16Lanes, "phase 1 writes every position read");
  static_assert(QG * M <= 64 && kGroups * QG == 16, "group shape");
  static_assert(3 * kWave * C * 2 <= kFloats, "key merge buffer");
};

// Stage L of the 16-row tree restricted to output offsets [LO, HI] of every
// time sub-block: entry tb*L + p = entry tb*L + p/2 + (entry tb*L + L/2 + p/2
// shifted by (p+1)/2 channels), as tree_stage computes it.
template <int C, int L, int LO, int HI>
__device__ __forceinline__ void tree_stage_range(float (&v)[16][C]) {
  float w[16][C];
#pragma unroll
  for (int tb = 0; tb < 16 / L; ++tb) {
#pragma unroll
    for (int q = LO / 2; q <= HI / 2; ++q) {
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
        if (2 * q >= LO && 2 * q <= HI) w[tb * L + 2 * q][j] = v[ra][j] + bx[j + q];
        if (2 * q + 1 >= LO && 2 * q + 1 <= HI) w[tb * L + 2 * q + 1][j] = v[ra][j] + bx[j + q + 1];
      }
    }
  }
#pragma unroll
  for (int tb = 0; tb < 16 / L; ++tb)
#pragma unroll
    for (int p = LO; p <= HI; ++p)
#pragma unroll
      for (int j = 0; j < C; ++j) v[tb * L + p][j] = w[tb * L + p][j];
}

// Level-16 offsets [Q0, Q1] (rows Q0..Q1 of v on return).
template <int C, int Q0, int Q1, int L = 2>
__device__ __forceinline__ void tree_range(float (&v)[16][C]) {
  if constexpr (L <= 16) {
    constexpr int sh = L == 2 ? 3 : L == 4 ? 2 : L == 8 ? 1 : 0;
    tree_stage_range<C, L, (Q0 >> sh), (Q1 >> sh)>(v);
    tree_range<C, Q0, Q1, L * 2>(v);
  }
}

// Phase 1 input for time block k: the 16 rows 16k.. at start channels
// base + 4*lane + j (row t shifted by t*b, the block's own coordinates).
template <int M, typename In>
__device__ __forceinline__ void staged_load(const In* __restrict__ x, long stride, int N, int k,
                                            int b, long base, float (&v)[16][Staged<M>::C]) {
  constexpr int C = Staged<M>::C;
  const int lane = threadIdx.x % kWave;
  // Row t starts at channel base + t*b, monotonic in t: rows 0 and 15 bound
  // the block, so one check covers all 16 rows and the source pointer just
  // advances by stride + b per row (no per-row address products or branches).
  const long g_lo = b >= 0 ? base : base + 15L * b;
  const long g_hi = b >= 0 ? base + 15L * b : base;
  if (g_lo >= 0 && g_hi + kWave * C <= N) {
    const In* src = x + (long)(16 * k) * stride + base + lane * C;
    const long step = stride + b;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      if constexpr (sizeof(In) == 4) {
        typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
        const f4u w = *(const f4u*)src;
        v[t][0] = w.x;
        v[t][1] = w.y;
        v[t][2] = w.z;
        v[t][3] = w.w;
      } else {
#pragma unroll
        for (int j = 0; j < C; ++j) v[t][j] = to_f32<In>(src[j]);
      }
      src += step;
    }
    return;
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const In* row = x + (long)(16 * k + t) * stride;
    const long g0 = base + (long)t * b;
    if (g0 >= 0 && g0 + kWave * C <= N) {
      const In* src = row + g0 + lane * C;
      if constexpr (sizeof(In) == 4) {
        typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
        const f4u w = *(const f4u*)src;
        v[t][0] = w.x;
        v[t][1] = w.y;
        v[t][2] = w.z;
        v[t][3] = w.w;
      } else {
#pragma unroll
        for (int j = 0; j < C; ++j) v[t][j] = to_f32<In>(src[j]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const long gc = g0 + lane * C + j;
        const bool ok = gc >= 0 && gc < N;
        const float val = to_f32<In>(row[ok ? gc : 0]);
        v[t][j] = ok ? val : 0.0f;
      }
    }
  }
}

// R_{q0+qq}[k][s] for s = 4*lane + j - k*qq into LDS row qq*M + k. With
// d = k*qq = 4m + e, element j goes to column j - e of position lane - m, or
// to column j - e + 4 of position lane - m - 1; e = ((k mod 4) * qq) mod 4 is a
// compile-time constant here (K4 = k mod 4), so every store is a fixed offset
// from one lane address and only the two position bounds are tested.
template <int M, int G_IDX, int K4>
__device__ __forceinline__ void staged_store(const float (&v)[16][Staged<M>::C], int k,
                                             lds_float* R) {
  using St = Staged<M>;
  constexpr int C = St::C;
  constexpr int Q0 = G_IDX * St::QG;
  const int lane = threadIdx.x % kWave;
#pragma unroll
  for (int qq = 0; qq < St::QG; ++qq) {
    const int e = (K4 * qq) & 3;
    const int m = (k * qq) >> 2;
    const int pos = lane - m;
    lds_float* row = R + (qq * M + k) * St::kRow + pos;
    const bool in0 = pos >= 0 && pos < St::kSub;
    const bool in1 = pos >= 1 && pos <= St::kSub;
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int jj = j - e;
      if (jj >= 0) {
        if (in0) row[jj * St::kSub] = v[Q0 + qq][j];
      } else {
        if (in1) row[(jj + C) * St::kSub - 1] = v[Q0 + qq][j];
      }
    }
  }
}

// Phase 1 for time block k: the partial tree of the loaded rows, and
// R_{q0+qq}[k] for the group's offsets into LDS.
template <int M, int G_IDX>
__device__ __forceinline__ void staged_phase1(float (&v)[16][Staged<M>::C], int k,
                                              lds_float* R) {
  using St = Staged<M>;
  constexpr int C = St::C;
  constexpr int Q0 = G_IDX * St::QG;
  const int lane = threadIdx.x % kWave;
  tree_range<C, Q0, Q0 + St::QG - 1>(v);
  if (lane < St::kLevel16Lanes) {
    switch (k & 3) {
      case 0: staged_store<M, G_IDX, 0>(v, k, R); break;
      case 1: staged_store<M, G_IDX, 1>(v, k, R); break;
      case 2: staged_store<M, G_IDX, 2>(v, k, R); break;
      default: staged_store<M, G_IDX, 3>(v, k, R); break;
    }
  }
}

// Phase 2 for one offset q = q0 + qq: R_q[0..M-1] from LDS, the M-row tree at
// drift 0, masking (dedoppler.cu:40-46) and the keys.
template <int M, bool BK>
__device__ __forceinline__ void staged_phase2(const lds_float* R, int qq, int q, int b, int c_w,
                                              int N, bool interior, int block,
                                              Best<Staged<M>::C, BK>& best) {
  using St = Staged<M>;
  constexpr int C = St::C;
  constexpr int T = St::T;
  const int lane = threadIdx.x % kWave;
  float v[M][C];
  const lds_float* src = R + qq * M * St::kRow + lane;
#pragma unroll
  for (int k = 0; k < M; ++k)
#pragma unroll
    for (int j = 0; j < C; ++j) v[k][j] = src[k * St::kRow + j * St::kSub];
  tree_regs<M, C>(v);
  if (!interior) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int c = c_w + lane * C + j;
      const int last0 = c + (T - 1) * b;
      const int pmax = (c >= N || last0 < 0) ? 0 : min(T, N - last0);
#pragma unroll
      for (int r = 0; r < M; ++r)
        if (q * M + r >= pmax) v[r][j] = 0.0f;
    }
  }
  if constexpr (BK) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
      float m = v[0][j];
#pragma unroll
      for (int r = 1; r < M; ++r) m = fmaxf(m, v[r][j]);
      if (m > best.sum[j]) {
        best.sum[j] = m;
        best.block[j] = block;
      }
    }
  } else {
    const unsigned low0 = (unsigned)kCodeMask - (unsigned)(block * T + q * M);
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const unsigned low = low0 - (unsigned)r;
#pragma unroll
      for (int j = 0; j < C; ++j) best.key[j] = fmax(best.key[j], make_key(v[r][j], low));
    }
  }
}

// One offset group of one drift block: phase 1, barrier, phase 2, barrier.
template <int M, int G_IDX, typename In, bool BK>
__device__ __forceinline__ void staged_group(const In* __restrict__ x, long stride, int N, int b,
                                             int c_w, bool interior, int block, lds_float* R,
                                             Best<Staged<M>::C, BK>& best) {
  using St = Staged<M>;
  constexpr int Q0 = G_IDX * St::QG;
  const int wave = threadIdx.x / kWave;
  // (Double-buffering the rows of the next time block measured slower: the
  // extra registers halve the waves per SIMD.)
  for (int k = wave; k < M; k += St::kWaves) {
    float v[16][St::C];
    staged_load<M, In>(x, stride, N, k, b, (long)c_w + (long)k * (16 * b + Q0), v);
    staged_phase1<M, G_IDX>(v, k, R);
  }
  __syncthreads();
  for (int qq = wave; qq < St::QG; qq += St::kWaves)
    staged_phase2<M, BK>(R, qq, Q0 + qq, b, c_w, N, interior, block, best);
  __syncthreads();
}

// Every offset group of one drift block, in order.
template <int M, int G_IDX, typename In, bool BK>
__device__ __forceinline__ void staged_groups(const In* __restrict__ x, long stride, int N, int b,
                                              int c_w, bool interior, int block, lds_float* R,
                                              Best<Staged<M>::C, BK>& best) {
  if constexpr (G_IDX < Staged<M>::kGroups) {
    staged_group<M, G_IDX, In, BK>(x, stride, N, b, c_w, interior, block, R, best);
    staged_groups<M, G_IDX + 1, In, BK>(x, stride, N, b, c_w, interior, block, R, best);
  }
}

// Offset group grp (uniform over the workgroup) of one drift block.
template <int M, int G_IDX, typename In, bool BK>
__device__ __forceinline__ void staged_group_at(int grp, const In* __restrict__ x, long stride,
                                                int N, int b, int c_w, bool interior, int block,
                                                lds_float* R, Best<Staged<M>::C, BK>& best) {
  if constexpr (G_IDX + 1 < Staged<M>::kGroups) {
    if (grp != G_IDX) {
      staged_group_at<M, G_IDX + 1, In, BK>(grp, x, stride, N, b, c_w, interior, block, R, best);
      return;
    }
  }
  staged_group<M, G_IDX, In, BK>(x, stride, N, b, c_w, interior, block, R, best);
}

// Grid: one workgroup of 4 waves per (channel tile, chunk of G consecutive
// drift blocks; G at run time), tile-major and dealt so that each XCD takes a contiguous
// range of items. The waves keep their own maxima over the chunk (full keys,
// or block keys as in fused_tree_top_kernel) and merge them through LDS
// before one atomic max per channel. With SPLIT, an item is one offset group
// of a (tile, chunk): the groups of one (tile, chunk) are consecutive items
// of one XCD, dispatched together, so the rows each group reads (mostly the
// same lines, shifted by k * q0 channels) are shared through that XCD's L2
// instead of being read once per group after other workgroups evicted them.
template <int M, typename In, bool BK, bool SPLIT>
__global__ void __launch_bounds__(kWave * Staged<M>::kWaves, Staged<M>::kMinWaves)
    staged_tree_kernel(const In* __restrict__ x, long stride, int N, int bmin, int nblocks,
                       int G, int nchunks, int total_items, top_key_t* __restrict__ top_key) {
  using St = Staged<M>;
  constexpr int C = St::C;
  constexpr int T = St::T;
  __shared__ __attribute__((aligned(16))) float r_storage[St::kFloats];
  lds_float* R = (lds_float*)r_storage;
  const int per_xcd = gridDim.x / 8;
  const int item = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (item >= total_items) return;
  const int grp = SPLIT ? item % St::kGroups : 0;
  const int tc = SPLIT ? item / St::kGroups : item;
  const int tile = tc / nchunks;
  const int chunk = tc % nchunks;
  const int c_w = tile * St::kOut;
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  Best<C, BK> best;
  best.init();
  const bool tile_in = c_w + St::kOut - 1 < N;
  const int g_end = min(nblocks, (chunk + 1) * G);
  for (int g = chunk * G; g < g_end; ++g) {
    const int b = bmin + g;
    const bool interior =
        tile_in && (c_w + (T - 1) * b >= 0) && (c_w + St::kOut - 1 + (T - 1) * b + (T - 1) < N);
    if constexpr (SPLIT)
      staged_group_at<M, 0, In, BK>(grp, x, stride, N, b, c_w, interior, g, R, best);
    else
      staged_groups<M, 0, In, BK>(x, stride, N, b, c_w, interior, g, R, best);
  }
  // The last phase 2 ended with a barrier: R is free for the merge. Keys of
  // either kind read as non-negative FP64 numbers in their unsigned order.
  double* red = (double*)r_storage;
  if (wave > 0) {
#pragma unroll
    for (int j = 0; j < C; ++j)
      red[(wave - 1) * kWave * C + lane * C + j] = __longlong_as_double((long long)best.merge_key(j));
  }
  __syncthreads();
  if (wave == 0 && lane * C < St::kOut) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
      double acc = __longlong_as_double((long long)best.merge_key(j));
#pragma unroll
      for (int w = 0; w < St::kWaves - 1; ++w) acc = fmax(acc, red[w * kWave * C + lane * C + j]);
      const int c = c_w + lane * C + j;
      const unsigned long long key = (unsigned long long)__double_as_longlong(acc);
      if (c < N && key != 0) atomicMax(&top_key[c], key);
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
__device__ __forceinline__ void colsum_body(const In* __restrict__ x, long stride, int T, int N,
                                            float* __restrict__ colsum) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.0f;
  for (int t = 0; t < T; ++t) s += to_f32<In>(x[(long)t * stride + c]);
  colsum[c] = s;
}

template <typename In>
__global__ void colsum_kernel(const In* __restrict__ x, long stride, int T, int N,
                              float* __restrict__ colsum) {
  colsum_body<In>(x, stride, T, N, colsum);
}

// Column sums of several channels (grid.y = channel).
template <typename In>
__global__ void colsum_set_kernel(FrontSet fs, long stride, int T, int N) {
  colsum_body<In>((const In*)fs.x[blockIdx.y], stride, T, N, fs.colsum[blockIdx.y]);
}

// Clears the top keys and the statistics histograms before a search (a fill
// kernel rather than a memset, so the sequence can be captured in a graph).
__device__ __forceinline__ void reset_body(top_key_t* __restrict__ key, long n,
                                           unsigned* __restrict__ hist, long nh) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    key[i] = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nh;
       i += (long)gridDim.x * blockDim.x)
    hist[i] = 0;
}

__global__ void reset_kernel(top_key_t* __restrict__ key, long n, unsigned* __restrict__ hist,
                             long nh) {
  reset_body(key, n, hist, nh);
}

__global__ void reset_set_kernel(FrontSet fs, long n, long nh) {
  reset_body(fs.key[blockIdx.y], n, fs.hist[blockIdx.y], nh);
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
    top_sum[c] = key_sum(k);
    top_blk[c] = (int)(code / (unsigned)T) + bmin;
    top_off[c] = (int)(code % (unsigned)T);
  }
}

// ---------------------------------------------------------------------------
// Device statistics (replacing the host's nth_element, dedoppler.cu:241-262):
// the exact order statistics at ranks mid = N/2, first = ceil(.05N) and
// last = floor(.95N), and the trimmed mean/deviation over the sorted ranks
// [first, last] in FP64, by a three-level radix select over the column sums'
// order keys (11 + 11 + 10 bits):
//   pass 1: histogram of the top 11 bits;
//   pass 2: every workgroup finds the level-1 bin of each rank from that
//           histogram, then histograms the next 11 bits inside those bins;
//   pass 3: likewise for the last 10 bits;
//   pass 4: every workgroup resolves the three exact values, then sums
//           (x - K) and (x - K)^2 in FP64 over the values strictly between
//           the first and last ranks' values (K = the median);
//   final:  one workgroup adds the boundary values' copies that fall inside
//           [first, last] and forms m = float(sum) / count, as the reference
//           forms its mean, and std = sqrt(sum (x - m)^2 / count) in FP64.
// Histograms are counted in LDS and flushed with one atomic per non-zero bin;
// partial sums are reduced in a fixed order, so results are deterministic.

constexpr int kStatsBlocks = 256;
constexpr int kStatsThreads = 256;
constexpr int kL1Bits = 11, kL2Bits = 11, kL3Bits = 10;
constexpr int kL1 = 1 << kL1Bits, kL2 = 1 << kL2Bits, kL3 = 1 << kL3Bits;
// Histogram words: level 1, then level 2 and level 3 for each of the 3 ranks,
// then one word counting column sums that are not finite (NaN or infinite
// samples; the search rejects such input).
constexpr int kHistWords = kL1 + 3 * kL2 + 3 * kL3;
constexpr int kNonFiniteWord = kHistWords;

struct StatsState {
  unsigned key[3];    // exact order keys of ranks mid, first, last
  unsigned below[3];  // number of values with a smaller key
  unsigned equal[3];  // number of values with the same key
  double part_s1[kStatsBlocks], part_s2[kStatsBlocks];
  unsigned part_n[kStatsBlocks];
};

size_t stats_hist_words() { return (size_t)kHistWords + 1; }

__device__ __forceinline__ int rank_first(int N) { return (int)ceil(0.05 * N); }
__device__ __forceinline__ int rank_last(int N) { return (int)floor(0.95 * N); }

// For each rank r (want[r] counted among the values matching the previous
// levels), finds the bin of `hist[r]` holding it: bin[r], the number of values
// in lower bins (added to below[r]) and the bin count. 256 threads; nbins a
// multiple of 256.
template <int NBINS>
__device__ void select_level(const unsigned* __restrict__ hist, int hist_stride,
                             const unsigned (&want)[3], unsigned (&bin)[3],
                             unsigned (&below)[3], unsigned (&count)[3]) {
  constexpr int kPer = NBINS / kStatsThreads;
  __shared__ unsigned scan[kStatsThreads];
  __shared__ unsigned found[3][3];
  const int tid = threadIdx.x;
  if (tid < 9) found[tid / 3][tid % 3] = 0;
  for (int r = 0; r < 3; ++r) {
    const unsigned* h = hist + r * hist_stride;
    unsigned v[kPer];
    unsigned local = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      v[i] = h[tid * kPer + i];
      local += v[i];
    }
    scan[tid] = local;
    __syncthreads();
    for (int off = 1; off < kStatsThreads; off <<= 1) {
      const unsigned x = tid >= off ? scan[tid - off] : 0u;
      __syncthreads();
      scan[tid] += x;
      __syncthreads();
    }
    const unsigned before = tid ? scan[tid - 1] : 0u;
    const unsigned w = want[r];
    if (w >= before && w < scan[tid]) {
      unsigned cum = before;
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        if (w >= cum && w < cum + v[i]) {
          found[r][0] = (unsigned)(tid * kPer + i);
          found[r][1] = cum;
          found[r][2] = v[i];
        }
        cum += v[i];
      }
    }
    __syncthreads();
  }
  for (int r = 0; r < 3; ++r) {
    bin[r] = found[r][0];
    below[r] = found[r][1];
    count[r] = found[r][2];
  }
  __syncthreads();
}

__device__ __forceinline__ void stats_ranks(int N, unsigned (&want)[3]) {
  want[0] = (unsigned)(N / 2);
  want[1] = (unsigned)rank_first(N);
  want[2] = (unsigned)rank_last(N);
}

// Counts `digit(key)` into an LDS histogram for every value passing `match`,
// then flushes the non-zero bins into global memory. Two 16-bit counters per
// LDS word (3 * NBINS / 2 words: 12 KB at 2048 bins), so the workgroup fits
// beside the tree kernels' 20 KB windows when one of them retires; the counts
// are flushed every 255 strides (at most 256 * 255 increments of one counter
// in between), which only happens above 16.7 M values.
template <int NBINS, typename Match, typename Digit>
__device__ void block_histogram(const float* __restrict__ v, int N, unsigned* lds,
                                unsigned* __restrict__ out, Match match, Digit digit) {
  constexpr int kWords = 3 * NBINS / 2;
  for (int i = threadIdx.x; i < kWords; i += blockDim.x) lds[i] = 0;
  __syncthreads();
  const long stride = (long)gridDim.x * blockDim.x;
  int strides = 0;
  for (long base = (long)blockIdx.x * blockDim.x; base < N; base += stride) {
    const long i = base + threadIdx.x;
    if (i < N) {
      const unsigned k = ord_key(v[i]);
#pragma unroll
      for (int r = 0; r < 3; ++r)
        if (match(r, k)) {
          const unsigned b = (unsigned)(r * NBINS) + digit(k);
          atomicAdd(&lds[b >> 1], 1u << (16 * (b & 1)));
        }
    }
    if (++strides == 255 && base + stride < N) {  // uniform over the workgroup
      __syncthreads();
      for (int w = threadIdx.x; w < kWords; w += blockDim.x) {
        const unsigned x = lds[w];
        if (x & 0xFFFFu) atomicAdd(&out[2 * w], x & 0xFFFFu);
        if (x >> 16) atomicAdd(&out[2 * w + 1], x >> 16);
        lds[w] = 0;
      }
      __syncthreads();
      strides = 0;
    }
  }
  __syncthreads();
  for (int w = threadIdx.x; w < kWords; w += blockDim.x) {
    const unsigned x = lds[w];
    if (x & 0xFFFFu) atomicAdd(&out[2 * w], x & 0xFFFFu);
    if (x >> 16) atomicAdd(&out[2 * w + 1], x >> 16);
  }
}

// The statistics passes below are written as device functions (LDS passed
// in) so that the single-channel kernels and the several-channel ones
// (grid.y = channel, a gated batch run's front) share them.
__device__ __forceinline__ void stats_pass1_body(const float* __restrict__ v, int N,
                                                 unsigned* __restrict__ hist, unsigned* h) {
  for (int i = threadIdx.x; i < kL1; i += blockDim.x) h[i] = 0;
  __syncthreads();
  unsigned bad = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const float x = v[i];
    // Exponent bits all ones: NaN or infinity (an integer test, which the
    // no-NaN compile mode cannot fold away).
    bad += (__float_as_uint(x) & 0x7F800000u) == 0x7F800000u;
    atomicAdd(&h[ord_key(x) >> (32 - kL1Bits)], 1u);
  }
  if (bad) atomicAdd(&hist[kNonFiniteWord], bad);
  __syncthreads();
  for (int i = threadIdx.x; i < kL1; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

__global__ void __launch_bounds__(kStatsThreads)
    stats_pass1_kernel(const float* __restrict__ v, int N, unsigned* __restrict__ hist) {
  __shared__ unsigned h[kL1];
  stats_pass1_body(v, N, hist, h);
}

__global__ void __launch_bounds__(kStatsThreads) stats_pass1_set_kernel(FrontSet fs, int N) {
  __shared__ unsigned h[kL1];
  stats_pass1_body(fs.colsum[blockIdx.y], N, fs.hist[blockIdx.y], h);
}

// Level-1 selection (replicated in every workgroup).
__device__ void select_l1(const unsigned* __restrict__ hist, int N, unsigned (&pre)[3],
                          unsigned (&want)[3], unsigned (&below)[3]) {
  unsigned w0[3], bin[3], bl[3], cnt[3];
  stats_ranks(N, w0);
  select_level<kL1>(hist, 0, w0, bin, bl, cnt);
  for (int r = 0; r < 3; ++r) {
    pre[r] = bin[r];
    want[r] = w0[r] - bl[r];
    below[r] = bl[r];
  }
}

__device__ void select_l2(const unsigned* __restrict__ hist, unsigned (&pre)[3],
                          unsigned (&want)[3], unsigned (&below)[3]) {
  unsigned bin[3], bl[3], cnt[3];
  select_level<kL2>(hist + kL1, kL2, want, bin, bl, cnt);
  for (int r = 0; r < 3; ++r) {
    pre[r] = (pre[r] << kL2Bits) | bin[r];
    want[r] -= bl[r];
    below[r] += bl[r];
  }
}

__device__ __forceinline__ void stats_pass2_body(const float* __restrict__ v, int N,
                                                 unsigned* __restrict__ hist, unsigned* h) {
  unsigned pre[3], want[3], below[3];
  select_l1(hist, N, pre, want, below);
  const unsigned p0 = pre[0], p1 = pre[1], p2 = pre[2];
  block_histogram<kL2>(
      v, N, h, hist + kL1,
      [=](int r, unsigned k) {
        const unsigned top = k >> (32 - kL1Bits);
        return top == (r == 0 ? p0 : (r == 1 ? p1 : p2));
      },
      [](unsigned k) { return (k >> kL3Bits) & (kL2 - 1); });
}

__global__ void __launch_bounds__(kStatsThreads)
    stats_pass2_kernel(const float* __restrict__ v, int N, unsigned* __restrict__ hist) {
  __shared__ unsigned h[3 * kL2 / 2];
  stats_pass2_body(v, N, hist, h);
}

__global__ void __launch_bounds__(kStatsThreads) stats_pass2_set_kernel(FrontSet fs, int N) {
  __shared__ unsigned h[3 * kL2 / 2];
  stats_pass2_body(fs.colsum[blockIdx.y], N, fs.hist[blockIdx.y], h);
}

__device__ __forceinline__ void stats_pass3_body(const float* __restrict__ v, int N,
                                                 unsigned* __restrict__ hist, unsigned* h) {
  unsigned pre[3], want[3], below[3];
  select_l1(hist, N, pre, want, below);
  select_l2(hist, pre, want, below);
  const unsigned p0 = pre[0], p1 = pre[1], p2 = pre[2];
  block_histogram<kL3>(
      v, N, h, hist + kL1 + 3 * kL2,
      [=](int r, unsigned k) {
        const unsigned top = k >> kL3Bits;
        return top == (r == 0 ? p0 : (r == 1 ? p1 : p2));
      },
      [](unsigned k) { return k & (kL3 - 1); });
}

__global__ void __launch_bounds__(kStatsThreads)
    stats_pass3_kernel(const float* __restrict__ v, int N, unsigned* __restrict__ hist) {
  __shared__ unsigned h[3 * kL3 / 2];
  stats_pass3_body(v, N, hist, h);
}

__global__ void __launch_bounds__(kStatsThreads) stats_pass3_set_kernel(FrontSet fs, int N) {
  __shared__ unsigned h[3 * kL3 / 2];
  stats_pass3_body(fs.colsum[blockIdx.y], N, fs.hist[blockIdx.y], h);
}

// All three levels: exact keys, counts below and equal.
__device__ void select_all(const unsigned* __restrict__ hist, int N, unsigned (&key)[3],
                           unsigned (&below)[3], unsigned (&equal)[3]) {
  unsigned pre[3], want[3];
  select_l1(hist, N, pre, want, below);
  select_l2(hist, pre, want, below);
  unsigned bin[3], bl[3], cnt[3];
  select_level<kL3>(hist + kL1 + 3 * kL2, kL3, want, bin, bl, cnt);
  for (int r = 0; r < 3; ++r) {
    key[r] = (pre[r] << kL3Bits) | bin[r];
    below[r] += bl[r];
    equal[r] = cnt[r];
  }
}

template <typename V>
__device__ __forceinline__ V block_sum(V x, V* sh) {
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const V r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void stats_pass4_body(const float* __restrict__ v, int N,
                                                 const unsigned* __restrict__ hist,
                                                 StatsState* st, double* shd, unsigned* shu) {
  unsigned key[3], below[3], equal[3];
  select_all(hist, N, key, below, equal);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int r = 0; r < 3; ++r) {
      st->key[r] = key[r];
      st->below[r] = below[r];
      st->equal[r] = equal[r];
    }
  }
  const unsigned k_lo = key[1], k_hi = key[2];
  const double K = (double)ord_val(key[0]);
  double s1 = 0.0, s2 = 0.0;
  unsigned n = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const float x = v[i];
    const unsigned k = ord_key(x);
    if (k > k_lo && k < k_hi) {
      const double d = __dsub_rn((double)x, K);
      s1 = __dadd_rn(s1, d);
      s2 = __dadd_rn(s2, __dmul_rn(d, d));
      ++n;
    }
  }
  s1 = block_sum<double>(s1, shd);
  s2 = block_sum<double>(s2, shd);
  n = block_sum<unsigned>(n, shu);
  if (threadIdx.x == 0) {
    st->part_s1[blockIdx.x] = s1;
    st->part_s2[blockIdx.x] = s2;
    st->part_n[blockIdx.x] = n;
  }
}

__global__ void __launch_bounds__(kStatsThreads)
    stats_pass4_kernel(const float* __restrict__ v, int N, const unsigned* __restrict__ hist,
                       StatsState* st) {
  __shared__ double shd[kStatsThreads];
  __shared__ unsigned shu[kStatsThreads];
  stats_pass4_body(v, N, hist, st, shd, shu);
}

__global__ void __launch_bounds__(kStatsThreads) stats_pass4_set_kernel(FrontSet fs, int N) {
  __shared__ double shd[kStatsThreads];
  __shared__ unsigned shu[kStatsThreads];
  stats_pass4_body(fs.colsum[blockIdx.y], N, fs.hist[blockIdx.y],
                   (StatsState*)fs.stats_scratch[blockIdx.y], shd, shu);
}

__device__ __forceinline__ void stats_final_body(const StatsState* st, int N,
                                                 const unsigned* __restrict__ hist,
                                                 float* __restrict__ stats_out, double* shd) {
  const int tid = threadIdx.x;
  // Fixed-order tree reduction of the per-workgroup partials.
  const double s1 = block_sum<double>(st->part_s1[tid], shd);
  const double s2 = block_sum<double>(st->part_s2[tid], shd);
  if (tid != 0) return;
  const int first = rank_first(N), last = rank_last(N);
  const long cnt = (long)last + 1 - first;
  const float median = ord_val(st->key[0]);
  const float v_lo = ord_val(st->key[1]);
  const float v_hi = ord_val(st->key[2]);
  const double K = (double)median;
  // Copies of the boundary values whose ranks fall in [first, last].
  long k_lo, k_hi;
  if (st->key[1] == st->key[2]) {
    k_lo = cnt;
    k_hi = 0;
  } else {
    k_lo = min((long)st->below[1] + (long)st->equal[1], (long)last + 1) - first;
    k_hi = (long)last + 1 - (long)st->below[2];
  }
  const double dl = __dsub_rn((double)v_lo, K), dh = __dsub_rn((double)v_hi, K);
  const double t1 =
      __dadd_rn(s1, __dadd_rn(__dmul_rn((double)k_lo, dl), __dmul_rn((double)k_hi, dh)));
  const double t2 = __dadd_rn(s2, __dadd_rn(__dmul_rn((double)k_lo, __dmul_rn(dl, dl)),
                                            __dmul_rn((double)k_hi, __dmul_rn(dh, dh))));
  // float sum = accumulate(...) (in double); float m = sum / count.
  const double total = __dadd_rn(__dmul_rn((double)cnt, K), t1);
  const float sum_f = (float)total;
  const float m = sum_f / (float)cnt;
  const double delta = __dsub_rn((double)m, K);
  // sum (x - m)^2 = sum (x-K)^2 - 2 delta sum (x-K) + cnt delta^2
  const double acc = __dadd_rn(__dsub_rn(t2, __dmul_rn(2.0 * delta, t1)),
                               __dmul_rn((double)cnt, __dmul_rn(delta, delta)));
  stats_out[0] = median;
  stats_out[1] = (float)sqrt(fmax(acc, 0.0) / (double)cnt);
  stats_out[2] = m;
  stats_out[3] = (float)hist[kNonFiniteWord];  // non-finite column sums (0 for valid input)
}

__global__ void __launch_bounds__(kStatsThreads)
    stats_final_kernel(const StatsState* st, int N, const unsigned* __restrict__ hist,
                       float* __restrict__ stats_out) {
  __shared__ double shd[kStatsThreads];
  stats_final_body(st, N, hist, stats_out, shd);
}

__global__ void __launch_bounds__(kStatsThreads) stats_final_set_kernel(FrontSet fs, int N) {
  __shared__ double shd[kStatsThreads];
  stats_final_body((const StatsState*)fs.stats_scratch[blockIdx.y], N, fs.hist[blockIdx.y],
                   fs.stats_out[blockIdx.y], shd);
}

// ---------------------------------------------------------------------------
// Window merge (dedoppler.cu:273-303), one workgroup per window, fused with the
// unpacking of the keys into the reference's three top arrays (each
// workgroup unpacks its own window; channels without a positive path keep
// their previous block and offset, dedoppler.cu:205).
__device__ __forceinline__ void window_merge_body(const top_key_t* __restrict__ key, int N, int T,
                                                  int bmin, int window, double snr_threshold,
                                                  const float* __restrict__ stats,
                                                  float* __restrict__ top_sum,
                                                  int* __restrict__ top_blk,
                                                  int* __restrict__ top_off,
                                                  int4* __restrict__ out, int flags) {
  const bool block_keys = flags & 1;
  // Gated search: a channel without a key (not searched, or no positive
  // path) neither forms nor suppresses a candidate.
  const float empty = (flags & 2) ? -INFINITY : 0.0f;
  __shared__ float sv[4];
  __shared__ int si[4];
  __shared__ int any_larger;
  const int wid = blockIdx.x;
  const int tid = threadIdx.x;
  // flags & 4: the statistics also go to out[number of windows], so that one
  // copy brings both to the host.
  if ((flags & 4) && wid == 0 && tid == 0)
    out[gridDim.x] = make_int4(__float_as_int(stats[0]), __float_as_int(stats[1]),
                               __float_as_int(stats[2]), __float_as_int(stats[3]));
  const int lane = tid % kWave, wave = tid / kWave;
  const float median = stats[0];
  const float sd = stats[1];
  // float path_sum_threshold = snr_threshold * std_dev + median; (double math, no FMA)
  const float thr = (float)__dadd_rn(__dmul_rn(snr_threshold, (double)sd), (double)median);
  const long lo = (long)wid * window;
  const long hi = min(lo + (long)window, (long)N);
  float bv = thr;
  int bi = -1;
  for (long f = lo + tid; f < hi; f += blockDim.x) {
    const top_key_t k = key[f];
    const float v = k ? (block_keys ? block_key_sum(k) : key_sum(k)) : 0.0f;
    const float cv = k ? v : empty;
    if (k && block_keys) {
      top_blk[f] = (int)~(unsigned)k + bmin;
      top_off[f] = -1;  // completed for candidates by candidate_offsets_kernel
    } else if (k) {
      const unsigned code = (unsigned)(kCodeMask - (k & kCodeMask));
      top_blk[f] = (int)(code / (unsigned)T) + bmin;
      top_off[f] = (int)(code % (unsigned)T);
    }
    top_sum[f] = v;
    if (cv > bv) {
      bv = cv;
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
  if (lane == 0) {
    sv[wave] = bv;
    si[wave] = bi;
  }
  if (tid == 0) any_larger = 0;
  __syncthreads();
  bv = sv[0];
  bi = si[0];
  for (int w = 1; w < 4; ++w) {
    const float ov = sv[w];
    const int oi = si[w];
    if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) {
      bv = ov;
      bi = oi;
    }
  }
  if (bi < 0) {
    if (tid == 0) out[wid] = make_int4(-1, 0, 0, 0);
    return;
  }
  const long s0 = max(0L, (long)bi - window + 1);
  const long s1 = min((long)N, (long)bi + window);
  int larger = 0;
  for (long f = s0 + tid; f < s1; f += blockDim.x) {
    const top_key_t k = key[f];
    larger |= (k ? (block_keys ? block_key_sum(k) : key_sum(k)) : empty) > bv;
  }
  if (larger) any_larger = 1;
  __syncthreads();
  if (tid == 0) {
    if (any_larger) {
      out[wid] = make_int4(-1, 0, 0, 0);
    } else {
      // The candidate lies in this window, so this workgroup unpacked it.
      out[wid] = make_int4(bi, __float_as_int(bv), top_blk[bi], top_off[bi]);
    }
  }
}

__global__ void __launch_bounds__(256)
    window_merge_kernel(const top_key_t* __restrict__ key, int N, int T, int bmin, int window,
                        double snr_threshold, const float* __restrict__ stats,
                        float* __restrict__ top_sum, int* __restrict__ top_blk,
                        int* __restrict__ top_off, int4* __restrict__ out, int flags) {
  window_merge_body(key, N, T, bmin, window, snr_threshold, stats, top_sum, top_blk, top_off, out,
                    flags);
}

// Several channels' merges (grid.y = channel, grid.x = the most windows of
// any of them).
__global__ void __launch_bounds__(256)
    window_merge_set_kernel(TailSet ts, int N, int T, int bmin, int flags) {
  const int c = blockIdx.y;
  const int nwin = (int)(((long)N + ts.window[c] - 1) / ts.window[c]);
  if ((int)blockIdx.x >= nwin) return;
  window_merge_body(ts.key[c], N, T, bmin, ts.window[c], ts.snr[c], ts.stats[c], ts.top_sum[c],
                    ts.top_blk[c], ts.top_off[c], ts.cand[c], flags);
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

// Location of the normalisation: the median column sum over T (stats_out[0]
// from the device statistics of the raw column sums).
__global__ void norm_location_kernel(const float* stats_out, int T, double* moments) {
  if (threadIdx.x == 0) moments[2] = (double)stats_out[0] / (double)T;
}

template <typename Out>
__global__ void norm_apply_kernel(const float* __restrict__ x, int T, int N, long stride,
                                  const double* moments, float clamp, Out* __restrict__ out) {
  const double m = moments[2];
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

// h5_reader.cpp:195-200: rows beyond the file's time steps are zero.
__global__ void zero_rows_kernel(float* __restrict__ rows, int nrows, int N, long stride) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < (long)nrows * N;
       i += (long)gridDim.x * blockDim.x)
    rows[(i / N) * stride + i % N] = 0.0f;
}

// h5_reader.cpp:202-209: the centre channel of each row becomes the mean of its
// neighbours ((a + b) / 2.0 in the reference; halving is exact, so * 0.5f).
__global__ void dc_fix_kernel(float* __restrict__ data, int T, int N, long stride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  float* row = data + (long)t * stride;
  const int mid = N / 2;
  row[mid] = (row[mid - 1] + row[mid + 1]) * 0.5f;
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

template <int C, int G, int W, typename In>
hipError_t launch_fused32_typed(const In* x, long stride, int N, int bmin, int nblocks,
                                top_key_t* top_key, float* colsum, hipStream_t stream) {
  using Tl = Tile32<C>;
  const long ntiles = ((long)N + Tl::kOut - 1) / Tl::kOut;
  const long ngroups = ((long)nblocks + G - 1) / G;
  const long total = ntiles * ngroups;
  if (total > (1L << 30)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((fused32_kernel<C, G, W, In>), dim3((int)((total + 7) / 8 * 8)),
                     dim3(kWave * W), 0, stream, x, stride, N, bmin, nblocks, (int)ngroups,
                     (int)total, top_key, colsum);
  return hipGetLastError();
}

// T=32 shapes (C, G, waves per workgroup): `prefetch` selects the waves (0: one).
template <typename In>
hipError_t fused32_dispatch(const In* x, long stride, int N, int bmin, int nblocks, int C, int G,
                            int P, top_key_t* top_key, float* colsum, hipStream_t stream) {
#define SC_F32_CASE(CC, GG, PP, WW)                                                         \
  if (C == CC && G == GG && P == PP)                                                        \
    return launch_fused32_typed<CC, GG, WW, In>(x, stride, N, bmin, nblocks, top_key, colsum, \
                                                stream);
  SC_F32_CASE(4, 4, 0, 1)
  SC_F32_CASE(2, 2, 0, 1)
  SC_F32_CASE(2, 4, 0, 1)
  SC_F32_CASE(2, 4, 2, 2)
  SC_F32_CASE(2, 8, 2, 2)
  SC_F32_CASE(2, 8, 4, 4)
  SC_F32_CASE(4, 8, 2, 2)
#undef SC_F32_CASE
  return hipErrorInvalidValue;
}

constexpr int kPass1C = 4;
constexpr int kPass2C = 4;

template <int M>
long two_pass_pitch_typed(int N) {
  using Sh = TileShape<M, kPass2C>;
  const long ntiles = ((long)N + Sh::kOut - 1) / Sh::kOut;
  return (ntiles * Sh::kOut + kWave * kPass2C + 63) / 64 * 64;
}

template <int M, int G, typename In>
hipError_t two_pass_typed(const In* x, long stride, int N, int bmin, int b_first, int nb,
                          float* scratch, long row_pitch, top_key_t* top_key,
                          hipStream_t stream) {
  constexpr int T = 16 * M;
  if (row_pitch < two_pass_pitch_typed<M>(N)) return hipErrorInvalidValue;
  using Sh1 = TileShape<16, kPass1C>;
  const long ntiles1 = ((long)N + Sh1::kOut - 1) / Sh1::kOut;
  const long ngroups = ((long)nb + G - 1) / G;
  const long total1 = ntiles1 * M * ngroups;
  using Sh2 = TileShape<M, kPass2C>;
  const long total2 = ((long)N + Sh2::kOut - 1) / Sh2::kOut;
  if (total1 > (1L << 30) || total2 > (1L << 30)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((level16_kernel<kPass1C, G, (M <= 4), In>), dim3((int)((total1 + 7) / 8 * 8)),
                     dim3(kWave), 0, stream, x, stride, N, M, b_first, nb, (int)ntiles1,
                     (int)total1, row_pitch, scratch);
  hipLaunchKernelGGL((upper_tree_kernel<M, kPass2C>), dim3((int)((total2 + 7) / 8 * 8)),
                     dim3(kWave * kUpperWaves), 0, stream, scratch, row_pitch, T, N, bmin,
                     b_first, nb, (int)total2, top_key);
  return hipGetLastError();
}

template <typename In>
hipError_t two_pass_dispatch(const In* x, long stride, int T, int N, int bmin, int b_first,
                             int nb, int G, float* scratch, long row_pitch, top_key_t* top_key,
                             hipStream_t stream) {
#define SC_TWO_PASS_CASE(MM, GG)                                                          \
  if (T == 16 * MM && G == GG)                                                            \
    return two_pass_typed<MM, GG, In>(x, stride, N, bmin, b_first, nb, scratch, row_pitch, \
                                      top_key, stream);
  SC_TWO_PASS_CASE(2, 4)
  SC_TWO_PASS_CASE(2, 8)
  SC_TWO_PASS_CASE(4, 4)
  SC_TWO_PASS_CASE(4, 8)
  SC_TWO_PASS_CASE(8, 4)
  SC_TWO_PASS_CASE(8, 8)
  SC_TWO_PASS_CASE(16, 4)
  SC_TWO_PASS_CASE(16, 8)
#undef SC_TWO_PASS_CASE
  return hipErrorInvalidValue;
}

template <int M, typename In, bool SPLIT>
hipError_t launch_staged_typed(const In* x, long stride, int N, int bmin, int nblocks, int G,
                               top_key_t* top_key, bool block_keys, hipStream_t stream) {
  using St = Staged<M>;
  const long ntiles = ((long)N + St::kOut - 1) / St::kOut;
  const long nchunks = ((long)nblocks + G - 1) / G;
  const long total = ntiles * nchunks * (SPLIT ? St::kGroups : 1);
  if (total > (1L << 30)) return hipErrorInvalidValue;
  if (block_keys)
    hipLaunchKernelGGL((staged_tree_kernel<M, In, true, SPLIT>),
                       dim3((int)((total + 7) / 8 * 8)), dim3(kWave * St::kWaves), 0, stream, x,
                       stride, N, bmin, nblocks, G, (int)nchunks, (int)total, top_key);
  else
    hipLaunchKernelGGL((staged_tree_kernel<M, In, false, SPLIT>),
                       dim3((int)((total + 7) / 8 * 8)), dim3(kWave * St::kWaves), 0, stream, x,
                       stride, N, bmin, nblocks, G, (int)nchunks, (int)total, top_key);
  return hipGetLastError();
}

// split: one workgroup per offset group (T_r = 128, 256: shape (4, G, 1)).
template <typename In>
hipError_t staged_dispatch(const In* x, long stride, int T, int N, int bmin, int nblocks, int G,
                           bool split, top_key_t* top_key, bool bk, hipStream_t stream) {
  switch (T) {
    case 32:
      return launch_staged_typed<2, In, false>(x, stride, N, bmin, nblocks, G, top_key, bk,
                                               stream);
    case 64:
      return launch_staged_typed<4, In, false>(x, stride, N, bmin, nblocks, G, top_key, bk,
                                               stream);
    case 128:
      return split ? launch_staged_typed<8, In, true>(x, stride, N, bmin, nblocks, G, top_key,
                                                      bk, stream)
                   : launch_staged_typed<8, In, false>(x, stride, N, bmin, nblocks, G, top_key,
                                                       bk, stream);
    case 256:
      return split ? launch_staged_typed<16, In, true>(x, stride, N, bmin, nblocks, G, top_key,
                                                       bk, stream)
                   : launch_staged_typed<16, In, false>(x, stride, N, bmin, nblocks, G, top_key,
                                                        bk, stream);
  }
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

bool staged_shape(int T, int C, int G, int P) {
  const bool g_ok = G == 1 || G == 2 || G == 4 || G == 8 || G == 16 || G == 32 || G == 64 ||
                    G == 128 || G == 256;
  if (T == 64) return C == 4 && P == 0 && g_ok;
  if (T == 128 || T == 256) return C == 4 && (P == 0 || P == 1) && g_ok;
  return T == 32 && C == 4 && P == kStaged32 && g_ok;
}

bool fused_supported(int T, int C, int G, int P) {
  if (T == 64 || T == 128 || T == 256 || (T == 32 && P == kStaged32))
    return staged_shape(T, C, G, P);
  if (T == 32)
    return (P == 0 && ((C == 4 && G == 4) || (C == 2 && (G == 2 || G == 4)))) ||
           (P == 2 && ((C == 2 && (G == 4 || G == 8)) || (C == 4 && G == 8))) ||
           (P == 4 && C == 2 && G == 8);
#define SC_FUSED_SUPPORTED(TT, CC, GG, PP) \
  if (T == TT && C == CC && G == GG && P == PP) return true;
  SC_FUSED_VARIANTS(SC_FUSED_SUPPORTED)
#undef SC_FUSED_SUPPORTED
  return false;
}

hipError_t launch_fused_tree_top(const void* x, int dtype, long stride, int T, int N, int bmin,
                                 int nblocks, int C, int G, int P, top_key_t* top_key,
                                 float* colsum, hipStream_t stream, const int* tile_list,
                                 const int* tile_count, bool block_keys) {
  if (T == 64 || T == 128 || T == 256 || (T == 32 && P == kStaged32)) {
    if (tile_list || !staged_shape(T, C, G, P)) return hipErrorInvalidValue;
    if (colsum) {
      const hipError_t e = launch_colsum(x, dtype, stride, T, N, colsum, stream);
      if (e != hipSuccess) return e;
    }
    switch (dtype) {
      case 0:
        return staged_dispatch<float>((const float*)x, stride, T, N, bmin, nblocks, G, P == 1,
                                      top_key, block_keys, stream);
      case 1:
        return staged_dispatch<__half>((const __half*)x, stride, T, N, bmin, nblocks, G,
                                       P == 1, top_key, block_keys, stream);
      case 2:
        return staged_dispatch<__hip_bfloat16>((const __hip_bfloat16*)x, stride, T, N, bmin,
                                               nblocks, G, P == 1, top_key, block_keys, stream);
    }
    return hipErrorInvalidValue;
  }
  if (T == 32) {
    if (tile_list) return hipErrorInvalidValue;
    switch (dtype) {
      case 0:
        return fused32_dispatch<float>((const float*)x, stride, N, bmin, nblocks, C, G, P,
                                       top_key, colsum, stream);
      case 1:
        return fused32_dispatch<__half>((const __half*)x, stride, N, bmin, nblocks, C, G, P,
                                        top_key, colsum, stream);
      case 2:
        return fused32_dispatch<__hip_bfloat16>((const __hip_bfloat16*)x, stride, N, bmin,
                                                nblocks, C, G, P, top_key, colsum, stream);
    }
    return hipErrorInvalidValue;
  }
  return launch_fused16(x, dtype, stride, T, N, bmin, nblocks, C, G, P, top_key, colsum, stream,
                        tile_list, tile_count, block_keys);
}

template <int T, typename In>
hipError_t candidate_offsets_typed(const In* x, long stride, int N, int nwin, int4* cand,
                                   int* top_off, hipStream_t stream) {
  constexpr int per_wave = kWave / (T < kWave ? T : kWave);
  hipLaunchKernelGGL((candidate_offsets_kernel<T, In>), dim3((nwin + per_wave - 1) / per_wave),
                     dim3(kWave), 0, stream, x, stride, N, nwin, cand, top_off);
  return hipGetLastError();
}

template <typename In>
hipError_t candidate_offsets_dispatch(const In* x, long stride, int T, int N, int nwin,
                                      int4* cand, int* top_off, hipStream_t stream) {
  switch (T) {
    case 2: return candidate_offsets_typed<2, In>(x, stride, N, nwin, cand, top_off, stream);
    case 4: return candidate_offsets_typed<4, In>(x, stride, N, nwin, cand, top_off, stream);
    case 8: return candidate_offsets_typed<8, In>(x, stride, N, nwin, cand, top_off, stream);
    case 16: return candidate_offsets_typed<16, In>(x, stride, N, nwin, cand, top_off, stream);
    case 32: return candidate_offsets_typed<32, In>(x, stride, N, nwin, cand, top_off, stream);
    case 64: return candidate_offsets_typed<64, In>(x, stride, N, nwin, cand, top_off, stream);
    case 128:
      return candidate_offsets_typed<128, In>(x, stride, N, nwin, cand, top_off, stream);
    case 256:
      return candidate_offsets_typed<256, In>(x, stride, N, nwin, cand, top_off, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_window_merge_set(const TailSet& ts, int n, int N, int T, int bmin,
                                   hipStream_t stream, bool skip_empty) {
  if (n < 1 || n > kMaxFusedSet) return hipErrorInvalidValue;
  int nwin = 0;
  for (int c = 0; c < n; ++c) {
    if (ts.window[c] <= 0) return hipErrorInvalidValue;
    nwin = std::max(nwin, (int)(((long)N + ts.window[c] - 1) / ts.window[c]));
  }
  hipLaunchKernelGGL(window_merge_set_kernel, dim3(nwin, n), dim3(256), 0, stream, ts, N, T,
                     bmin, 1 | (skip_empty ? 2 : 0));
  return hipGetLastError();
}

hipError_t launch_candidate_offsets_set(const TailSet& ts, int n, int dtype, long stride, int N,
                                        hipStream_t stream) {
  if (n < 1 || n > kMaxFusedSet) return hipErrorInvalidValue;
  int nwin = 0;
  for (int c = 0; c < n; ++c) {
    if (ts.window[c] <= 0) return hipErrorInvalidValue;
    nwin = std::max(nwin, (int)(((long)N + ts.window[c] - 1) / ts.window[c]));
  }
  constexpr int per_wave = kWave / 16;
  const dim3 grid((nwin + per_wave - 1) / per_wave, n);
  switch (dtype) {
    case 0:
      hipLaunchKernelGGL((candidate_offsets_set_kernel<16, float>), grid, dim3(kWave), 0, stream,
                         ts, stride, N);
      break;
    case 1:
      hipLaunchKernelGGL((candidate_offsets_set_kernel<16, __half>), grid, dim3(kWave), 0, stream,
                         ts, stride, N);
      break;
    case 2:
      hipLaunchKernelGGL((candidate_offsets_set_kernel<16, __hip_bfloat16>), grid, dim3(kWave), 0,
                         stream, ts, stride, N);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_candidate_offsets(const void* x, int dtype, long stride, int T, int N,
                                    int nwin, int4* cand, int* top_off, hipStream_t stream) {
  if (nwin <= 0) return hipSuccess;
  switch (dtype) {
    case 0:
      return candidate_offsets_dispatch<float>((const float*)x, stride, T, N, nwin, cand,
                                               top_off, stream);
    case 1:
      return candidate_offsets_dispatch<__half>((const __half*)x, stride, T, N, nwin, cand,
                                                top_off, stream);
    case 2:
      return candidate_offsets_dispatch<__hip_bfloat16>((const __hip_bfloat16*)x, stride, T, N,
                                                        nwin, cand, top_off, stream);
  }
  return hipErrorInvalidValue;
}

int fused_tile_width(int T, int C) {
  if (T == 32) return C == 4 ? Tile32<4>::kOut : C == 2 ? Tile32<2>::kOut : 0;
  if (T < 2 || C <= 0) return 0;
  return (kWave - (T - 1 + C - 1) / C) * C;
}

bool two_pass_supported(int T, int G) {
  return (T == 32 || T == 64 || T == 128 || T == 256) && (G == 4 || G == 8);
}

long two_pass_row_pitch(int T, int N) {
  switch (T) {
    case 32: return two_pass_pitch_typed<2>(N);
    case 64: return two_pass_pitch_typed<4>(N);
    case 128: return two_pass_pitch_typed<8>(N);
    case 256: return two_pass_pitch_typed<16>(N);
  }
  return 0;
}

hipError_t launch_two_pass(const void* x, int dtype, long stride, int T, int N, int bmin,
                           int b_first, int nb, int G, float* scratch, long row_pitch,
                           top_key_t* top_key, hipStream_t stream) {
  if (nb <= 0) return hipSuccess;
  switch (dtype) {
    case 0:
      return two_pass_dispatch<float>((const float*)x, stride, T, N, bmin, b_first, nb, G,
                                      scratch, row_pitch, top_key, stream);
    case 1:
      return two_pass_dispatch<__half>((const __half*)x, stride, T, N, bmin, b_first, nb, G,
                                       scratch, row_pitch, top_key, stream);
    case 2:
      return two_pass_dispatch<__hip_bfloat16>((const __hip_bfloat16*)x, stride, T, N, bmin,
                                               b_first, nb, G, scratch, row_pitch, top_key,
                                               stream);
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

hipError_t launch_reset(top_key_t* key, long n, unsigned* hist, hipStream_t stream) {
  hipLaunchKernelGGL(reset_kernel, dim3(blocks_for(n, 256, 2048)), dim3(256), 0, stream, key, n,
                     hist, (long)stats_hist_words());
  return hipGetLastError();
}

hipError_t launch_unpack_top(const top_key_t* key, int N, int T, int bmin, float* top_sum,
                             int* top_blk, int* top_off, hipStream_t stream) {
  hipLaunchKernelGGL(unpack_top_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, key, N, T,
                     bmin, top_sum, top_blk, top_off);
  return hipGetLastError();
}

size_t device_stats_scratch_bytes(int /*N*/) { return sizeof(StatsState); }
size_t device_stats_hist_bytes() { return sizeof(unsigned) * stats_hist_words(); }

hipError_t launch_device_stats(const float* colsum, int N, void* scratch, unsigned* hist,
                               float* stats_out, hipStream_t stream) {
  StatsState* st = (StatsState*)scratch;
  const dim3 grid(kStatsBlocks), block(kStatsThreads);
  hipLaunchKernelGGL(stats_pass1_kernel, grid, block, 0, stream, colsum, N, hist);
  hipLaunchKernelGGL(stats_pass2_kernel, grid, block, 0, stream, colsum, N, hist);
  hipLaunchKernelGGL(stats_pass3_kernel, grid, block, 0, stream, colsum, N, hist);
  hipLaunchKernelGGL(stats_pass4_kernel, grid, block, 0, stream, colsum, N, hist, st);
  hipLaunchKernelGGL(stats_final_kernel, dim3(1), block, 0, stream, st, N, hist, stats_out);
  return hipGetLastError();
}

hipError_t launch_front_set_stats(const FrontSet& fs, int n, int dtype, long stride, int T,
                                  int N, hipStream_t stream) {
  if (n < 1 || n > kMaxFusedSet) return hipErrorInvalidValue;
  hipLaunchKernelGGL(reset_set_kernel, dim3(blocks_for(N, 256, 2048), n), dim3(256), 0, stream,
                     fs, (long)N, (long)stats_hist_words());
  const dim3 cgrid((N + 255) / 256, n);
  switch (dtype) {
    case 0:
      hipLaunchKernelGGL(colsum_set_kernel<float>, cgrid, dim3(256), 0, stream, fs, stride, T, N);
      break;
    case 1:
      hipLaunchKernelGGL(colsum_set_kernel<__half>, cgrid, dim3(256), 0, stream, fs, stride, T,
                         N);
      break;
    case 2:
      hipLaunchKernelGGL(colsum_set_kernel<__hip_bfloat16>, cgrid, dim3(256), 0, stream, fs,
                         stride, T, N);
      break;
    default: return hipErrorInvalidValue;
  }
  const hipError_t e = hipGetLastError();
  return e != hipSuccess ? e : launch_device_stats_set(fs, n, N, stream);
}

hipError_t launch_device_stats_set(const FrontSet& fs, int n, int N, hipStream_t stream) {
  if (n < 1 || n > kMaxFusedSet) return hipErrorInvalidValue;
  const dim3 grid(kStatsBlocks, n), block(kStatsThreads);
  hipLaunchKernelGGL(stats_pass1_set_kernel, grid, block, 0, stream, fs, N);
  hipLaunchKernelGGL(stats_pass2_set_kernel, grid, block, 0, stream, fs, N);
  hipLaunchKernelGGL(stats_pass3_set_kernel, grid, block, 0, stream, fs, N);
  hipLaunchKernelGGL(stats_pass4_set_kernel, grid, block, 0, stream, fs, N);
  hipLaunchKernelGGL(stats_final_set_kernel, dim3(1, n), block, 0, stream, fs, N);
  return hipGetLastError();
}

hipError_t launch_window_merge(const top_key_t* key, int N, int T, int bmin, int window,
                               double snr_threshold, const float* stats, float* top_sum,
                               int* top_blk, int* top_off, int4* cand_out, bool block_keys,
                               hipStream_t stream, bool skip_empty, bool stats_slot) {
  const int nwin = (int)(((long)N + window - 1) / window);
  hipLaunchKernelGGL(window_merge_kernel, dim3(nwin), dim3(256), 0, stream, key, N, T, bmin,
                     window, snr_threshold, stats, top_sum, top_blk, top_off, cand_out,
                     (block_keys ? 1 : 0) | (skip_empty ? 2 : 0) | (stats_slot ? 4 : 0));
  return hipGetLastError();
}

hipError_t launch_gather(const float* sums, int N, const int* offsets, const int* chans, int n,
                         float* out, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, sums, N, offsets,
                     chans, n, out);
  return hipGetLastError();
}

namespace {
// Scratch of launch_normalize: partial sums, column sums, statistics state,
// histograms, statistics output.
constexpr size_t norm_align(size_t n) { return (n + 255) / 256 * 256; }
struct NormScratch {
  size_t part, colsum, state, hist, stats, total;
  explicit NormScratch(int N) {
    part = 0;
    colsum = norm_align(sizeof(double) * kNormBlocks);
    state = colsum + norm_align(sizeof(float) * (size_t)N);
    hist = state + norm_align(device_stats_scratch_bytes(N));
    stats = hist + norm_align(device_stats_hist_bytes());
    total = stats + norm_align(4 * sizeof(float));
  }
};
}  // namespace

size_t normalize_scratch_bytes(int /*T*/, int N) { return NormScratch(N).total; }

hipError_t launch_normalize(const float* x, int T, int N, long stride, int out_dtype, float clamp,
                            void* out, double* moments, void* scratch, hipStream_t stream) {
  const NormScratch L(N);
  char* base = (char*)scratch;
  double* part = (double*)(base + L.part);
  const double count = (double)T * (double)N;
  // Location: the median of the raw column sums over T, so the normalised
  // column sums have median ~0 whatever share of the power narrowband RFI
  // holds (the mean would put them far below zero): every positive SNR
  // threshold then lies above the zero the top path sums start from.
  float* colsum = (float*)(base + L.colsum);
  unsigned* hist = (unsigned*)(base + L.hist);
  float* stats = (float*)(base + L.stats);
  hipError_t e = launch_colsum(x, 0, stride, T, N, colsum, stream);
  if (e == hipSuccess) e = launch_reset(nullptr, 0, hist, stream);
  if (e == hipSuccess)
    e = launch_device_stats(colsum, N, base + L.state, hist, stats, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(norm_location_kernel, dim3(1), dim3(64), 0, stream, stats, T, moments);
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

hipError_t launch_condition(float* data, int T, int T_r, int N, long stride, int dc_spike,
                            hipStream_t stream) {
  if (T_r > T)
    hipLaunchKernelGGL(zero_rows_kernel, dim3(blocks_for((long)(T_r - T) * N, 256, 8192)),
                       dim3(256), 0, stream, data + (long)T * stride, T_r - T, N, stride);
  if (dc_spike && N >= 3 && T > 0)
    hipLaunchKernelGGL(dc_fix_kernel, dim3((T + 255) / 256), dim3(256), 0, stream, data, T, N,
                       stride);
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
