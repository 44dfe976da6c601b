// The fused tree + top-path kernel for T_r <= 16 (the search's hot kernel at
// 16 time steps) and its launcher.
//
// Reference semantics: taylor.cu:170-290 (tiledTaylorKernel, one launch per
// drift block) and dedoppler.cu:32-55 (findTopPathSums, one launch per drift
// block). Here one launch covers every drift block: each wavefront keeps a
// channel tile's tree in registers, reuses one LDS window across a group of
// drift blocks and folds the max/argmax into the same pass.
//
// Built in a translation unit of its own so that its instruction scheduling
// can be chosen independently of the other kernels (see build.py).

#include "dd_device.h"

namespace sc {

namespace {

// Grid: one 64-lane workgroup per work item = (channel tile, chunk of
// consecutive drift groups of G blocks). The workgroup walks its chunk window
// by window and keeps the running key max in registers across the chunk.
// MODE 0: a chunk is one window. MODE 1: the next window's rows are fetched
// into registers while the current window's blocks run out of LDS. MODE 2:
// the next window is fetched after the current one is done (no registers held
// across the compute); chunks cut the number of atomic merges per channel.
// Items are tile-major and dealt so that each XCD
// (workgroups go round-robin to the 8 XCDs) sweeps a contiguous range of
// tiles. The workgroup whose chunk holds drift block 0 also writes the tile's
// column sums in the reference's sequential order (dedoppler.cu:68-71).
template <int T, int C, int G, int MODE, typename In, bool BK>
__global__ void __launch_bounds__(kWave)
    fused_tree_top_kernel(const In* __restrict__ x, long stride, int N, int bmin, int nblocks,
                          int ngroups, int chunks, int groups_per_chunk, int total_items,
                          unsigned long long* __restrict__ top_key,
                          float* __restrict__ colsum, const int* __restrict__ tile_list,
                          const int* __restrict__ tile_count) {
  using Sh = TileShape<T, C>;
  using Wn = Window<T, C, G>;
  __shared__ float win_storage[Wn::kFloats];
  lds_float* win = (lds_float*)win_storage;

  // With a tile list (gated search) only its first *tile_count tiles are
  // searched; the active items are dealt over the XCDs the same way.
  const int active = tile_list ? min(total_items, *tile_count * chunks) : total_items;
  const int per_xcd = tile_list ? (active + 7) / 8 : gridDim.x / 8;
  if ((int)blockIdx.x / 8 >= per_xcd) return;
  const int item = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (item >= active) return;
  const int tile = tile_list ? tile_list[item / chunks] : item / chunks;
  const int chunk = item % chunks;
  const int c_w = tile * Sh::kOut;
  const int g_begin = chunk * groups_per_chunk;
  const int g_end = min(ngroups, g_begin + groups_per_chunk);
  const int lane = threadIdx.x;

  float pf[T][Wn::kLoads];
  fetch_window<T, C, G, In>(pf, x, stride, N, c_w, bmin + g_begin * G, 0);
  const int b_first = bmin + g_begin * G;
  const int b_last = min(bmin + g_end * G, bmin + nblocks) - 1;
  if (colsum && b_first <= 0 && 0 <= b_last && lane < Sh::kValidLanes) {
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
  store_window<T, C, G>(win, pf);
  __syncthreads();

  Best<C, BK> best;
  best.init();

  for (int grp = g_begin; grp < g_end; ++grp) {
    // MODE 0: every work item is one drift group (the launcher makes
    // groups_per_chunk 1), so there is no next window to fetch.
    const bool more = MODE != 0 && grp + 1 < g_end;
    if (MODE == 1 && more)
      fetch_window<T, C, G, In>(pf, x, stride, N, c_w, bmin + (grp + 1) * G, 0);
    const int b0 = bmin + grp * G;
    const int ng = min(G, nblocks - grp * G);
    lds_float* rp[T];
    row_bases<T, C, G>(win, rp);
    for (int g = 0; g < ng; g += C) {
      fused_block<T, C, G, BK, 0>(rp, g, ng, b0, c_w, N, grp * G, best);
#pragma unroll
      for (int t = 1; t < T; ++t) rp[t] += t;  // next h step: rows shift by t columns
    }
    if (more) {
      if (MODE == 2) fetch_window<T, C, G, In>(pf, x, stride, N, c_w, bmin + (grp + 1) * G, 0);
      __syncthreads();
      store_window<T, C, G>(win, pf);
      __syncthreads();
    }
  }

  if (lane < Sh::kValidLanes) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int c = c_w + lane * C + j;
      const unsigned long long k = best.merge_key(j);
      if (c < N && k != 0) atomicMax(&top_key[c], k);
    }
  }
}

// Several coarse channels of one shape in one launch (a batch's consecutive
// instances): items are instance-major, then tile-major, dealt over the XCDs
// as above. Longer chunks with the same number of items: at 16 x 2^20 x 8 the
// launcher makes one chunk per tile instead of about eight. The body below
// the item decode is fused_tree_top_kernel's, repeated as text: written as one
// shared inline function, the kernels compiled to ten more registers and the
// single-channel one ran 2-3 % slower. LIST (gated batches): each channel's
// items are its listed tiles' only (set.count[i] tiles x chunks), the active
// items of all channels dealt over the XCDs; the rest of the grid exits.
template <int T, int C, int G, int MODE, typename In, bool BK, bool LIST>
__global__ void __launch_bounds__(kWave)
    fused_tree_top_set_kernel(FusedSet set, long stride, int N, int bmin, int nblocks,
                              int ngroups, int chunks, int groups_per_chunk,
                              int items_per_channel, int total_items) {
  using Sh = TileShape<T, C>;
  using Wn = Window<T, C, G>;
  __shared__ float win_storage[Wn::kFloats];
  lds_float* win = (lds_float*)win_storage;

  int inst = 0, item = 0;
  if (LIST) {
    const int n = total_items / items_per_channel;
    int active = 0;
    for (int i = 0; i < n; ++i) active += min(items_per_channel, *set.count[i] * chunks);
    const int per_xcd = (active + 7) / 8;
    if ((int)blockIdx.x / 8 >= per_xcd) return;
    item = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (item >= active) return;
    // item < the sum of the channels' active items: the walk stops at a channel below n.
    for (;;) {
      const int a = min(items_per_channel, *set.count[inst] * chunks);
      if (item < a) break;
      item -= a;
      ++inst;
    }
  } else {
    const int per_xcd = gridDim.x / 8;
    const int item_all = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (item_all >= total_items) return;
    inst = item_all / items_per_channel;
    item = item_all - inst * items_per_channel;
  }
  const In* __restrict__ x = (const In*)set.x[inst];
  unsigned long long* __restrict__ top_key = set.key[inst];
  float* __restrict__ colsum = set.colsum[inst];
  const int tile = LIST ? set.list[inst][item / chunks] : item / chunks;
  const int chunk = item % chunks;
  const int c_w = tile * Sh::kOut;
  const int g_begin = chunk * groups_per_chunk;
  const int g_end = min(ngroups, g_begin + groups_per_chunk);
  const int lane = threadIdx.x;

  float pf[T][Wn::kLoads];
  fetch_window<T, C, G, In>(pf, x, stride, N, c_w, bmin + g_begin * G, 0);
  const int b_first = bmin + g_begin * G;
  const int b_last = min(bmin + g_end * G, bmin + nblocks) - 1;
  if (colsum && b_first <= 0 && 0 <= b_last && lane < Sh::kValidLanes) {
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
  store_window<T, C, G>(win, pf);
  __syncthreads();

  Best<C, BK> best;
  best.init();

  for (int grp = g_begin; grp < g_end; ++grp) {
    // MODE 0: every work item is one drift group (the launcher makes
    // groups_per_chunk 1), so there is no next window to fetch.
    const bool more = MODE != 0 && grp + 1 < g_end;
    if (MODE == 1 && more)
      fetch_window<T, C, G, In>(pf, x, stride, N, c_w, bmin + (grp + 1) * G, 0);
    const int b0 = bmin + grp * G;
    const int ng = min(G, nblocks - grp * G);
    lds_float* rp[T];
    row_bases<T, C, G>(win, rp);
    for (int g = 0; g < ng; g += C) {
      fused_block<T, C, G, BK, 0>(rp, g, ng, b0, c_w, N, grp * G, best);
#pragma unroll
      for (int t = 1; t < T; ++t) rp[t] += t;  // next h step: rows shift by t columns
    }
    if (more) {
      if (MODE == 2) fetch_window<T, C, G, In>(pf, x, stride, N, c_w, bmin + (grp + 1) * G, 0);
      __syncthreads();
      store_window<T, C, G>(win, pf);
      __syncthreads();
    }
  }

  if (lane < Sh::kValidLanes) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const int c = c_w + lane * C + j;
      const unsigned long long k = best.merge_key(j);
      if (c < N && k != 0) atomicMax(&top_key[c], k);
    }
  }
}


// Chunks per tile so that about `target` items exist (mode 0: one group each).
template <int MODE>
long fused_chunks(long tiles, long ngroups) {
  long chunks = ngroups;
  if (MODE != 0) {
    const long target = MODE == 1 ? 8192 : 32768;
    chunks = 1;
    while (tiles * chunks < target && chunks < ngroups) ++chunks;
  }
  const long per_chunk = (ngroups + chunks - 1) / chunks;
  return (ngroups + per_chunk - 1) / per_chunk;
}

template <int T, int C, int G, int MODE, typename In>
hipError_t launch_fused_typed(const In* x, long stride, int N, int bmin, int nblocks,
                              top_key_t* top_key, float* colsum, const int* tile_list,
                              const int* tile_count, bool block_keys, hipStream_t stream) {
  using Sh = TileShape<T, C>;
  const long ntiles = ((long)N + Sh::kOut - 1) / Sh::kOut;
  const long ngroups = ((long)nblocks + G - 1) / G;
  // MODE 1/2: chunks of drift groups per tile, as long as the chip is still
  // filled several times over (256 CUs x a few resident workgroups). Measured
  // at 16 x 2^20, 132 blocks, mode 2: 4096 / 8192 / 16384 / 32768 / 65536
  // items -> 0.64 / 0.56 / 0.54 / 0.49 / 0.49 ms.
  const long chunks = fused_chunks<MODE>(ntiles, ngroups);
  const long per_chunk = (ngroups + chunks - 1) / chunks;
  const long total = ntiles * chunks;
  if (total > (1L << 30)) return hipErrorInvalidValue;
  const int grid = (int)((total + 7) / 8 * 8);
  if (block_keys)
    hipLaunchKernelGGL((fused_tree_top_kernel<T, C, G, MODE, In, true>), dim3(grid), dim3(kWave),
                       0, stream, x, stride, N, bmin, nblocks, (int)ngroups, (int)chunks,
                       (int)per_chunk, (int)total, top_key, colsum, tile_list, tile_count);
  else
    hipLaunchKernelGGL((fused_tree_top_kernel<T, C, G, MODE, In, false>), dim3(grid),
                       dim3(kWave), 0, stream, x, stride, N, bmin, nblocks, (int)ngroups,
                       (int)chunks, (int)per_chunk, (int)total, top_key, colsum, tile_list,
                       tile_count);
  return hipGetLastError();
}

template <typename In>
hipError_t fused_dispatch(const In* x, long stride, int T, int N, int bmin, int nblocks, int C,
                          int G, int P, top_key_t* top_key, float* colsum, const int* tile_list,
                          const int* tile_count, bool block_keys, hipStream_t stream) {
#define SC_FUSED_CASE(TT, CC, GG, PP)                                                     \
  if (T == TT && C == CC && G == GG && P == PP)                                           \
    return launch_fused_typed<TT, CC, GG, PP, In>(x, stride, N, bmin, nblocks, top_key,      \
                                                       colsum, tile_list, tile_count,         \
                                                       block_keys, stream);
  SC_FUSED_VARIANTS(SC_FUSED_CASE)
#undef SC_FUSED_CASE
  return hipErrorInvalidValue;
}

template <typename In>
hipError_t launch_set_typed(const FusedSet& set, int n, long stride, int N, int bmin,
                            int nblocks, bool block_keys, bool lists, hipStream_t stream) {
  // The default 16-time-step shape (C = 4, G = 4, mode 2).
  constexpr int T = 16, C = 4, G = 4, MODE = 2;
  using Sh = TileShape<T, C>;
  const long ntiles = ((long)N + Sh::kOut - 1) / Sh::kOut;
  const long ngroups = ((long)nblocks + G - 1) / G;
  const long chunks = fused_chunks<MODE>(ntiles * n, ngroups);
  const long per_chunk = (ngroups + chunks - 1) / chunks;
  const long per_channel = ntiles * chunks;
  const long total = per_channel * n;
  if (total > (1L << 30)) return hipErrorInvalidValue;
  const int grid = (int)((total + 7) / 8 * 8);
  if (lists) {
    // Gated searches keep block keys only.
    if (!block_keys) return hipErrorInvalidValue;
    for (int i = 0; i < n; ++i)
      if (!set.list[i] || !set.count[i]) return hipErrorInvalidValue;
    hipLaunchKernelGGL((fused_tree_top_set_kernel<T, C, G, MODE, In, true, true>), dim3(grid),
                       dim3(kWave), 0, stream, set, stride, N, bmin, nblocks, (int)ngroups,
                       (int)chunks, (int)per_chunk, (int)per_channel, (int)total);
  } else if (block_keys) {
    hipLaunchKernelGGL((fused_tree_top_set_kernel<T, C, G, MODE, In, true, false>), dim3(grid),
                       dim3(kWave), 0, stream, set, stride, N, bmin, nblocks, (int)ngroups,
                       (int)chunks, (int)per_chunk, (int)per_channel, (int)total);
  } else {
    hipLaunchKernelGGL((fused_tree_top_set_kernel<T, C, G, MODE, In, false, false>),
                       dim3(grid), dim3(kWave), 0, stream, set, stride, N, bmin, nblocks,
                       (int)ngroups, (int)chunks, (int)per_chunk, (int)per_channel,
                       (int)total);
  }
  return hipGetLastError();
}

}  // namespace

bool fused_set_supported(int T, int C, int G, int P) {
  return T == 16 && C == 4 && G == 4 && P == 2;
}

hipError_t launch_fused16_set(const FusedSet& set, int n, int dtype, long stride, int T, int N,
                              int bmin, int nblocks, int C, int G, int P, bool block_keys,
                              hipStream_t stream, bool lists) {
  if (n < 1 || n > kMaxFusedSet || !fused_set_supported(T, C, G, P)) return hipErrorInvalidValue;
  switch (dtype) {
    case 0:
      return launch_set_typed<float>(set, n, stride, N, bmin, nblocks, block_keys, lists, stream);
    case 1:
      return launch_set_typed<__half>(set, n, stride, N, bmin, nblocks, block_keys, lists,
                                      stream);
    case 2:
      return launch_set_typed<__hip_bfloat16>(set, n, stride, N, bmin, nblocks, block_keys,
                                              lists, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fused16(const void* x, int dtype, long stride, int T, int N, int bmin,
                          int nblocks, int C, int G, int P, top_key_t* top_key, float* colsum,
                          hipStream_t stream, const int* tile_list, const int* tile_count,
                          bool block_keys) {
  switch (dtype) {
    case 0:
      return fused_dispatch<float>((const float*)x, stride, T, N, bmin, nblocks, C, G, P,
                                   top_key, colsum, tile_list, tile_count, block_keys, stream);
    case 1:
      return fused_dispatch<__half>((const __half*)x, stride, T, N, bmin, nblocks, C, G, P,
                                    top_key, colsum, tile_list, tile_count, block_keys, stream);
    case 2:
      return fused_dispatch<__hip_bfloat16>((const __hip_bfloat16*)x, stride, T, N, bmin,
                                            nblocks, C, G, P, top_key, colsum, tile_list,
                                            tile_count, block_keys, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace sc
