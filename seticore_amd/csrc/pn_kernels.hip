// Learned block proposal network: HIP kernels for gfx950.
//
// No counterpart in lacker/seticore. A small convolutional network scores
// 4096-channel blocks of a 16-row coarse channel from a frequency-pooled,
// noise-normalised tile; only blocks it flags are searched with the full tree
// (the plain search is the fallback and the judge of its recall).
//
// Network (23,393 parameters; weights and activations stored as BF16,
// products accumulated in FP32):
//   input  z [16][1024]           pooled power, z = 2 (sum of 4 channels -
//                                 median/4) / std (column-sum statistics),
//                                 clamped to [-4, 32]
//   conv1  1 -> 16, 3x5, stride (1,2), pad (1,2), ReLU   -> h1 [16][512][16]
//   conv2 16 -> 32, 3x5, stride (2,2), pad (1,2), ReLU   -> h2 [ 8][256][32]
//   conv3 32 -> 32, 3x5, stride (2,2), pad (1,2), ReLU   -> h3 [ 4][128][32]
//   head   1x1 32 -> 1 (FP32 weights), max over the 512 positions -> logit
// Activations are NHWC (channels innermost). The convolutions run on the
// matrix cores (v_mfma_f32_16x16x16_bf16) as implicit GEMMs: rows = output
// positions, columns = output channels, K = (tap, input channel).
//
// Training: weighted binary cross-entropy on the block logit. The max makes
// the backward pass sparse - only the receptive field of each block's argmax
// position carries gradient (3x5 positions of h2, 7x13 of h1) - so one
// workgroup per block computes that block's whole parameter gradient, and a
// second kernel sums the per-block gradients in a fixed order
// (deterministic). Adam updates FP32 master weights, then the BF16 copies are
// refreshed.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "dd_internal.h"
#include "pn_internal.h"

namespace sc {
namespace pn {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even float -> bfloat16 bits (inputs are finite).
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float((unsigned)h << 16);
}

// ---------------------------------------------------------------------------
// Pooling: z[blk][t][j] from the 4 channels 4096*blk + 4j .. +3 of row t.
template <typename In>
__device__ __forceinline__ float load_f32(const In* p);
template <>
__device__ __forceinline__ float load_f32<float>(const float* p) { return *p; }
template <>
__device__ __forceinline__ float load_f32<unsigned short>(const unsigned short* p) {
  return bf2f(*p);
}
template <>
__device__ __forceinline__ float load_f32<_Float16>(const _Float16* p) { return (float)*p; }

template <typename In>
__device__ __forceinline__ void pool_body(const In* __restrict__ x, long stride, int nblk,
                                          const float* __restrict__ stats,
                                          unsigned short* __restrict__ z) {
  const long i = blockIdx.x * 256L + threadIdx.x;  // over nblk * 16 * 1024
  if (i >= (long)nblk * kRows * kPooled) return;
  const int j = (int)(i % kPooled);
  const int t = (int)((i / kPooled) % kRows);
  const long blk = i / ((long)kPooled * kRows);
  const In* p = x + t * stride + blk * kBlockChannels + 4L * j;
  const float s = (load_f32<In>(p) + load_f32<In>(p + 1)) + (load_f32<In>(p + 2) + load_f32<In>(p + 3));
  const float median = stats[0], sd = stats[1];
  float v = 2.0f * (s - 0.25f * median) / sd;
  v = fminf(fmaxf(v, kZMin), kZMax);
  z[i] = f2bf(v);
}

template <typename In>
__global__ void __launch_bounds__(256)
    pool_kernel(const In* __restrict__ x, long stride, int nblk, const float* __restrict__ stats,
                unsigned short* __restrict__ z) {
  pool_body<In>(x, stride, nblk, stats, z);
}

// Several channels (grid.y = channel), channel c's tiles at z + c * nblk tiles.
template <typename In>
__global__ void __launch_bounds__(256)
    pool_set_kernel(FrontSet fs, long stride, int nblk, unsigned short* __restrict__ z) {
  const int c = blockIdx.y;
  pool_body<In>((const In*)fs.x[c], stride, nblk, fs.stats_out[c],
                z + (long)c * nblk * kRows * kPooled);
}

// ---------------------------------------------------------------------------
// Convolution forward on the matrix cores. One wave per (64 output positions
// along a row, output row, block). Lane layout of v_mfma_f32_16x16x16_bf16:
// A[row][k] sits in lane row + 16*(k/4), element k%4; B[k][col] in lane
// col + 16*(k/4), element k%4; D[row][col] in lane col + 16*(row/4), element
// row%4. The product is formed transposed, D[channel][position] =
// W[channel][k] * X[k][position], so each lane ends with 4 consecutive
// channels of one position: one 8-byte store per lane and channel tile.
// ROWS output rows per workgroup (blockIdx.y * ROWS ..), one after the other,
// so the weights held in registers are loaded once for all of them.
template <int CIN, int COUT, int SH, int SW, int ROWS>
__global__ void __launch_bounds__(64)
    conv_fwd_kernel(const unsigned short* __restrict__ in, int HI, int WI,
                    const unsigned short* __restrict__ wb, const float* __restrict__ bias,
                    unsigned short* __restrict__ out, int HO, int WO) {
  constexpr int NT = COUT / 16;
  constexpr int KS = CIN >= 16 ? CIN / 16 : 1;
  constexpr int TAPS = CIN >= 16 ? 15 : 1;  // CIN == 1: one K step holds all 15 taps
  constexpr int KROW = CIN >= 16 ? 15 * CIN : 16;  // weights per output channel
  const int lane = threadIdx.x;
  const int r = lane % 16, g = lane / 16;
  const int wo0 = blockIdx.x * 64;
  const long b = blockIdx.z;

  bf16x4 bw[TAPS][KS][NT];
#pragma unroll
  for (int t = 0; t < TAPS; ++t)
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const unsigned short* w = wb + (long)(n * 16 + r) * KROW +
                                  (CIN >= 16 ? t * CIN + s * 16 + 4 * g : 4 * g);
        bw[t][s][n] = *(const bf16x4*)w;
      }

  float bsv[NT][4];
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int i = 0; i < 4; ++i) bsv[n][i] = bias[n * 16 + 4 * g + i];

  for (int ho = blockIdx.y * ROWS; ho < min(HO, (int)(blockIdx.y + 1) * ROWS); ++ho) {
  f32x4 acc[4][NT];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};

  if constexpr (CIN == 1) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int pos = wo0 + 16 * m + r;
      bf16x4 a;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = 4 * g + i;
        unsigned short v = 0;
        if (k < 15) {
          const int hi = ho * SH + k / 5 - 1;
          const int wi = pos * SW + k % 5 - 2;
          if (hi >= 0 && hi < HI && wi >= 0 && wi < WI) v = in[(b * HI + hi) * WI + wi];
        }
        a[i] = (short)v;
      }
#pragma unroll
      for (int n = 0; n < NT; ++n)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bw[0][0][n], a, acc[m][n], 0, 0, 0);
    }
  } else {
    // Every operand of the row first (32-bit offsets within the block's
    // input; taps outside the input read offset 0 and are zeroed), then the
    // products in the same (position tile, tap, K step, channel tile) order.
    const unsigned short* blk_in = in + b * HI * WI * CIN;
    bf16x4 a[4][15][KS];
#pragma unroll
    for (int t = 0; t < 15; ++t) {
      const int hi = ho * SH + t / 5 - 1;
      const bool row_ok = hi >= 0 && hi < HI;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int wi = (wo0 + 16 * m + r) * SW + t % 5 - 2;
        const bool ok = row_ok && wi >= 0 && wi < WI;
        const int off = ok ? (hi * WI + wi) * CIN + 4 * g : 0;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const bf16x4 v = *(const bf16x4*)(blk_in + off + s * 16);
          a[m][t][s] = ok ? v : bf16x4{0, 0, 0, 0};
        }
      }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int t = 0; t < 15; ++t)
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
          for (int n = 0; n < NT; ++n)
            acc[m][n] =
                __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bw[t][s][n], a[m][t][s], acc[m][n], 0, 0, 0);
  }

#pragma unroll
  for (int n = 0; n < NT; ++n) {
    const int ch0 = n * 16 + 4 * g;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int pos = wo0 + 16 * m + r;
      bf16x4 v;
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = (short)f2bf(fmaxf(acc[m][n][i] + bsv[n][i], 0.0f));
      *(bf16x4*)(out + ((b * HO + ho) * WO + pos) * COUT + ch0) = v;
    }
  }
  }
}

// Head: logit[b] = max over positions of (bh + sum_c h3[pos][c] * wh[c]);
// argmax = lowest position attaining it.
__global__ void __launch_bounds__(256)
    head_kernel(const unsigned short* __restrict__ h3, const float* __restrict__ params,
                float* __restrict__ logits, int* __restrict__ argmax) {
  __shared__ float sv[256];
  __shared__ int si[256];
  const long b = blockIdx.x;
  const float* wh = params + kOffWH;
  const float bh = params[kOffBH];
  float best = -INFINITY;
  int bi = 0;
  for (int pos = threadIdx.x; pos < kH3 * kW3; pos += 256) {
    const unsigned short* h = h3 + (b * kH3 * kW3 + pos) * kC3;
    float s = bh;
#pragma unroll
    for (int c = 0; c < kC3; ++c) s += bf2f(h[c]) * wh[c];
    if (s > best) {
      best = s;
      bi = pos;
    }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const float o = sv[threadIdx.x + w];
      const int oi = si[threadIdx.x + w];
      if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = o;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    logits[b] = sv[0];
    if (argmax) argmax[b] = si[0];
  }
}

// ---------------------------------------------------------------------------
// Training.

// Weighted BCE: per block w*(max(z,0) - z*y + log1p(exp(-|z|))), w = pos_weight
// for y = 1; dz = w*(sigmoid(z) - y)/B; loss = mean (one workgroup, fixed order).
__global__ void __launch_bounds__(256)
    loss_kernel(const float* __restrict__ logits, const float* __restrict__ labels, int B,
                float pos_weight, float* __restrict__ dz, float* __restrict__ loss) {
  __shared__ double part[256];
  double acc = 0.0;
  for (int b = threadIdx.x; b < B; b += 256) {
    const float z = logits[b], y = labels[b];
    const float w = y > 0.5f ? pos_weight : 1.0f;
    const float l = fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)));
    const float p = 1.0f / (1.0f + expf(-z));
    dz[b] = w * (p - y) / (float)B;
    acc += (double)(w * l);
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) loss[0] = (float)(part[0] / B);
}

// One workgroup per block: the whole parameter gradient of that block's loss
// term into gpart[b][kNumParams] (every entry written).
__global__ void __launch_bounds__(256)
    backward_kernel(const unsigned short* __restrict__ z, const unsigned short* __restrict__ h1,
                    const unsigned short* __restrict__ h2, const unsigned short* __restrict__ h3,
                    const int* __restrict__ argmax, const float* __restrict__ dz,
                    const float* __restrict__ params, const unsigned short* __restrict__ w2b,
                    const unsigned short* __restrict__ w3b, float* __restrict__ gpart) {
  const long b = blockIdx.x;
  const int tid = threadIdx.x;
  float* g = gpart + b * (long)kNumParams;
  __shared__ float dY3[kC3];
  __shared__ float dY2[15][kC2];   // conv2 outputs at the 15 taps of conv3
  __shared__ int p2y[15], p2x[15];  // their coordinates (-1: padding)
  __shared__ float dY1[7 * 13][kC1];

  const int pos = argmax[b];
  const int y3 = pos / kW3, x3 = pos % kW3;
  const float d = dz[b];
  const unsigned short* h3p = h3 + ((b * kH3 + y3) * kW3 + x3) * kC3;

  // Head and conv3 bias.
  if (tid < kC3) {
    const float a = bf2f(h3p[tid]);
    g[kOffWH + tid] = d * a;
    const float dh = d * params[kOffWH + tid];
    const float dy = a > 0.0f ? dh : 0.0f;
    dY3[tid] = dy;
    g[kOffB3 + tid] = dy;
  }
  if (tid == 0) g[kOffBH] = d;
  if (tid < 15) {
    const int yy = 2 * y3 + tid / 5 - 1, xx = 2 * x3 + tid % 5 - 2;
    const bool ok = yy >= 0 && yy < kH2 && xx >= 0 && xx < kW2;
    p2y[tid] = ok ? yy : -1;
    p2x[tid] = ok ? xx : -1;
  }
  __syncthreads();

  // conv3 weights: dW3[co][t][ci] = dY3[co] * h2[p2(t)][ci].
  for (int e = tid; e < kC3 * 15 * kC2; e += 256) {
    const int ci = e % kC2, t = (e / kC2) % 15, co = e / (kC2 * 15);
    float v = 0.0f;
    if (p2y[t] >= 0) v = dY3[co] * bf2f(h2[((b * kH2 + p2y[t]) * kW2 + p2x[t]) * kC2 + ci]);
    g[kOffW3 + e] = v;
  }
  // Gradient at the 15 conv2 outputs, through the ReLU.
  for (int e = tid; e < 15 * kC2; e += 256) {
    const int ci = e % kC2, t = e / kC2;
    float v = 0.0f;
    if (p2y[t] >= 0) {
      const float a = bf2f(h2[((b * kH2 + p2y[t]) * kW2 + p2x[t]) * kC2 + ci]);
      if (a > 0.0f) {
        float s = 0.0f;
        for (int co = 0; co < kC3; ++co) s += dY3[co] * bf2f(w3b[(co * 15 + t) * kC2 + ci]);
        v = s;
      }
    }
    dY2[t][ci] = v;
  }
  __syncthreads();

  // conv2 bias and weights: dW2[co][t2][ci] = sum_j dY2[j][co] * h1[2*p2(j) + tap(t2)][ci].
  if (tid < kC2) {
    float s = 0.0f;
    for (int j = 0; j < 15; ++j) s += dY2[j][tid];
    g[kOffB2 + tid] = s;
  }
  for (int e = tid; e < kC2 * 15 * kC1; e += 256) {
    const int ci = e % kC1, t2 = (e / kC1) % 15, co = e / (kC1 * 15);
    float s = 0.0f;
    for (int j = 0; j < 15; ++j) {
      if (p2y[j] < 0) continue;
      const int yy = 2 * p2y[j] + t2 / 5 - 1, xx = 2 * p2x[j] + t2 % 5 - 2;
      if (yy < 0 || yy >= kH1 || xx < 0 || xx >= kW1) continue;
      s += dY2[j][co] * bf2f(h1[((b * kH1 + yy) * kW1 + xx) * kC1 + ci]);
    }
    g[kOffW2 + e] = s;
  }
  // Gradient at h1 over the 7x13 region rows 4*y3-3.., cols 4*x3-6.., through the ReLU.
  const int y1o = 4 * y3 - 3, x1o = 4 * x3 - 6;
  for (int e = tid; e < 7 * 13 * kC1; e += 256) {
    const int ci = e % kC1, cell = e / kC1;
    const int yy = y1o + cell / 13, xx = x1o + cell % 13;
    float v = 0.0f;
    if (yy >= 0 && yy < kH1 && xx >= 0 && xx < kW1 &&
        bf2f(h1[((b * kH1 + yy) * kW1 + xx) * kC1 + ci]) > 0.0f) {
      float s = 0.0f;
      for (int j = 0; j < 15; ++j) {
        if (p2y[j] < 0) continue;
        const int ty = yy - 2 * p2y[j] + 1, tx = xx - 2 * p2x[j] + 2;
        if (ty < 0 || ty >= 3 || tx < 0 || tx >= 5) continue;
        const int t2 = ty * 5 + tx;
        for (int co = 0; co < kC2; ++co) s += dY2[j][co] * bf2f(w2b[(co * 15 + t2) * kC1 + ci]);
      }
      v = s;
    }
    dY1[cell][ci] = v;
  }
  __syncthreads();

  // conv1 bias and weights: dW1[co][t] = sum_cells dY1[cell][co] * z[y + t/5 - 1][2x + t%5 - 2].
  if (tid < kC1) {
    float s = 0.0f;
    for (int cell = 0; cell < 7 * 13; ++cell) s += dY1[cell][tid];
    g[kOffB1 + tid] = s;
  }
  if (tid < kC1 * 15) {
    const int co = tid / 15, t = tid % 15;
    float s = 0.0f;
    for (int cell = 0; cell < 7 * 13; ++cell) {
      const int yy = y1o + cell / 13, xx = x1o + cell % 13;
      if (yy < 0 || yy >= kH1 || xx < 0 || xx >= kW1) continue;
      const int zy = yy + t / 5 - 1, zx = 2 * xx + t % 5 - 2;
      if (zy < 0 || zy >= kRows || zx < 0 || zx >= kPooled) continue;
      s += dY1[cell][co] * bf2f(z[(b * kRows + zy) * kPooled + zx]);
    }
    g[kOffW1 + co * 15 + t] = s;
  }
}

// grad[p] = sum over b in ascending order of gpart[b][p].
__global__ void __launch_bounds__(256)
    reduce_grad_kernel(const float* __restrict__ gpart, int B, float* __restrict__ grad) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= kNumParams) return;
  float s = 0.0f;
  for (int b = 0; b < B; ++b) s += gpart[(long)b * kNumParams + p];
  grad[p] = s;
}

__global__ void __launch_bounds__(256)
    adam_kernel(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                const float* __restrict__ grad, float lr, float beta1, float beta2, float eps,
                float c1, float c2) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= kNumParams) return;
  const float gr = grad[p];
  const float mm = beta1 * m[p] + (1.0f - beta1) * gr;
  const float vv = beta2 * v[p] + (1.0f - beta2) * gr * gr;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// BF16 copies of the convolution weights in the matrix-core layouts:
// w1b[co][16] (tap 15 = 0), w2b[co][t][ci], w3b[co][t][ci].
__global__ void __launch_bounds__(256)
    pack_kernel(const float* __restrict__ params, unsigned short* __restrict__ w1b,
                unsigned short* __restrict__ w2b, unsigned short* __restrict__ w3b) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < kC1 * 16) {
    const int co = i / 16, k = i % 16;
    w1b[i] = k < 15 ? f2bf(params[kOffW1 + co * 15 + k]) : 0;
  }
  if (i < kC2 * 15 * kC1) w2b[i] = f2bf(params[kOffW2 + i]);
  if (i < kC3 * 15 * kC2) w3b[i] = f2bf(params[kOffW3 + i]);
}

// ---------------------------------------------------------------------------
// Gating: the channel tiles of the search that overlap a flagged block, in
// ascending order (one workgroup; list[0..*count)).
__device__ __forceinline__ void tile_list_body(const float* __restrict__ logits, int nblk,
                                               float threshold, int N, int tile_width,
                                               int* __restrict__ list, int* __restrict__ count,
                                               int* __restrict__ flagged, int* sums) {
  const int ntiles = (N + tile_width - 1) / tile_width;
  const int per = (ntiles + 1023) / 1024;
  const int t0 = threadIdx.x * per;
  int n = 0;
  for (int t = t0; t < min(ntiles, t0 + per); ++t) {
    const long c0 = (long)t * tile_width, c1 = min((long)N, c0 + tile_width) - 1;
    const int b0 = (int)(c0 / kBlockChannels), b1 = (int)min((long)nblk - 1, c1 / kBlockChannels);
    bool f = false;
    for (int bb = b0; bb <= b1; ++bb) f = f || (logits[bb] > threshold);
    n += f;
  }
  sums[threadIdx.x] = n;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = (int)threadIdx.x >= off ? sums[threadIdx.x - off] : 0;
    __syncthreads();
    sums[threadIdx.x] += v;
    __syncthreads();
  }
  int w = sums[threadIdx.x] - n;
  for (int t = t0; t < min(ntiles, t0 + per); ++t) {
    const long c0 = (long)t * tile_width, c1 = min((long)N, c0 + tile_width) - 1;
    const int b0 = (int)(c0 / kBlockChannels), b1 = (int)min((long)nblk - 1, c1 / kBlockChannels);
    bool f = false;
    for (int bb = b0; bb <= b1; ++bb) f = f || (logits[bb] > threshold);
    if (f) list[w++] = t;
  }
  if (threadIdx.x == 1023) count[0] = sums[1023];
  if (flagged) {
    __syncthreads();
    if (threadIdx.x == 0) {
      int nf = 0;
      for (int bb = 0; bb < nblk; ++bb) nf += logits[bb] > threshold;
      flagged[0] = nf;
    }
  }
}

__global__ void __launch_bounds__(1024)
    tile_list_kernel(const float* __restrict__ logits, int nblk, float threshold, int N,
                     int tile_width, int* __restrict__ list, int* __restrict__ count,
                     int* __restrict__ flagged) {
  __shared__ int sums[1024];
  tile_list_body(logits, nblk, threshold, N, tile_width, list, count, flagged, sums);
}

// Several channels (one workgroup each; channel c's logits at logits + c * nblk).
__global__ void __launch_bounds__(1024)
    tile_list_set_kernel(const float* __restrict__ logits, int nblk, float threshold, int N,
                         int tile_width, FrontSet fs) {
  __shared__ int sums[1024];
  const int c = blockIdx.x;
  tile_list_body(logits + (long)c * nblk, nblk, threshold, N, tile_width, fs.list[c],
                 fs.count[c], fs.count[c] + 1, sums);
}

}  // namespace

hipError_t launch_pool(const void* x, int dtype, long stride, int nblk, const float* stats,
                       unsigned short* z, hipStream_t stream) {
  const long n = (long)nblk * kRows * kPooled;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  switch (dtype) {
    case 0:
      hipLaunchKernelGGL(pool_kernel<float>, grid, block, 0, stream, (const float*)x, stride,
                         nblk, stats, z);
      break;
    case 1:
      hipLaunchKernelGGL(pool_kernel<_Float16>, grid, block, 0, stream, (const _Float16*)x,
                         stride, nblk, stats, z);
      break;
    case 2:
      hipLaunchKernelGGL(pool_kernel<unsigned short>, grid, block, 0, stream,
                         (const unsigned short*)x, stride, nblk, stats, z);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pool_set(const FrontSet& fs, int n, int dtype, long stride, int nblk,
                           unsigned short* z, hipStream_t stream) {
  if (n < 1 || n > kMaxFusedSet) return hipErrorInvalidValue;
  const long per = (long)nblk * kRows * kPooled;
  const dim3 grid((unsigned)((per + 255) / 256), n), block(256);
  switch (dtype) {
    case 0:
      hipLaunchKernelGGL(pool_set_kernel<float>, grid, block, 0, stream, fs, stride, nblk, z);
      break;
    case 1:
      hipLaunchKernelGGL(pool_set_kernel<_Float16>, grid, block, 0, stream, fs, stride, nblk, z);
      break;
    case 2:
      hipLaunchKernelGGL(pool_set_kernel<unsigned short>, grid, block, 0, stream, fs, stride,
                         nblk, z);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_forward(const Weights& w, const unsigned short* z, int B, Acts& a,
                          float* logits, int* argmax, hipStream_t stream) {
  // Rows per workgroup: 4 / 4 / 2 (grid.y 4 / 2 / 2). Measured forward at
  // 256 / 2048 blocks: all rows of conv2 and conv3 in one workgroup and a row
  // per workgroup for conv1, 125 / 866 us; these, 108 / 865 us (the
  // workgroups load their weights a few times more, the blocks' rows spread
  // over more workgroups).
  hipLaunchKernelGGL((conv_fwd_kernel<1, kC1, 1, 2, 4>), dim3(kW1 / 64, kH1 / 4, B), dim3(64), 0,
                     stream, z, kRows, kPooled, w.w1b, w.params + kOffB1, a.h1, kH1, kW1);
  hipLaunchKernelGGL((conv_fwd_kernel<kC1, kC2, 2, 2, 4>), dim3(kW2 / 64, kH2 / 4, B), dim3(64),
                     0, stream, a.h1, kH1, kW1, w.w2b, w.params + kOffB2, a.h2, kH2, kW2);
  hipLaunchKernelGGL((conv_fwd_kernel<kC2, kC3, 2, 2, 2>), dim3(kW3 / 64, kH3 / 2, B), dim3(64),
                     0, stream, a.h2, kH2, kW2, w.w3b, w.params + kOffB3, a.h3, kH3, kW3);
  hipLaunchKernelGGL(head_kernel, dim3(B), dim3(256), 0, stream, a.h3, w.params, logits, argmax);
  return hipGetLastError();
}

hipError_t launch_gradient(const Weights& w, const unsigned short* z, const float* labels, int B,
                           float pos_weight, Acts& a, float* logits, int* argmax, float* dz,
                           float* gpart, float* grad, float* loss, hipStream_t stream) {
  hipError_t e = launch_forward(w, z, B, a, logits, argmax, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), 0, stream, logits, labels, B, pos_weight,
                     dz, loss);
  hipLaunchKernelGGL(backward_kernel, dim3(B), dim3(256), 0, stream, z, a.h1, a.h2, a.h3, argmax,
                     dz, w.params, w.w2b, w.w3b, gpart);
  hipLaunchKernelGGL(reduce_grad_kernel, dim3((kNumParams + 255) / 256), dim3(256), 0, stream,
                     gpart, B, grad);
  return hipGetLastError();
}

hipError_t launch_adam(Weights& w, float* m, float* v, const float* grad, float lr, int step,
                       hipStream_t stream) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((kNumParams + 255) / 256), dim3(256), 0, stream, w.params,
                     m, v, grad, lr, b1, b2, 1e-8f, c1, c2);
  return launch_pack(w, stream);
}

hipError_t launch_pack(Weights& w, hipStream_t stream) {
  hipLaunchKernelGGL(pack_kernel, dim3((kC3 * 15 * kC2 + 255) / 256), dim3(256), 0, stream,
                     w.params, w.w1b, w.w2b, w.w3b);
  return hipGetLastError();
}

hipError_t launch_tile_list_set(const float* logits, int nblk, float threshold, int N,
                                int tile_width, const FrontSet& fs, int n, hipStream_t stream) {
  if (n < 1 || n > kMaxFusedSet) return hipErrorInvalidValue;
  hipLaunchKernelGGL(tile_list_set_kernel, dim3(n), dim3(1024), 0, stream, logits, nblk,
                     threshold, N, tile_width, fs);
  return hipGetLastError();
}

hipError_t launch_tile_list(const float* logits, int nblk, float threshold, int N, int tile_width,
                            int* list, int* count, int* flagged, hipStream_t stream) {
  hipLaunchKernelGGL(tile_list_kernel, dim3(1), dim3(1024), 0, stream, logits, nblk, threshold,
                     N, tile_width, list, count, flagged);
  return hipGetLastError();
}

}  // namespace pn
}  // namespace sc
