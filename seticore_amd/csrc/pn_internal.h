// Internal interface of the block proposal network (pn_kernels.hip) used by
// the C-ABI host code. Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

namespace sc {
struct FrontSet;
namespace pn {

constexpr int kRows = 16;              // time steps of a tile (T_r of the search)
constexpr int kBlockChannels = 4096;   // fine channels scored together
constexpr int kPool = 4;               // fine channels per pooled column
constexpr int kPooled = kBlockChannels / kPool;  // 1024
constexpr float kZMin = -4.0f, kZMax = 32.0f;

constexpr int kC1 = 16, kH1 = 16, kW1 = 512;
constexpr int kC2 = 32, kH2 = 8, kW2 = 256;
constexpr int kC3 = 32, kH3 = 4, kW3 = 128;

// FP32 parameter vector layout.
constexpr int kOffW1 = 0;                        // [16][15]        (co, tap)
constexpr int kOffB1 = kOffW1 + kC1 * 15;        // [16]
constexpr int kOffW2 = kOffB1 + kC1;             // [32][15][16]    (co, tap, ci)
constexpr int kOffB2 = kOffW2 + kC2 * 15 * kC1;  // [32]
constexpr int kOffW3 = kOffB2 + kC2;             // [32][15][32]
constexpr int kOffB3 = kOffW3 + kC3 * 15 * kC2;  // [32]
constexpr int kOffWH = kOffB3 + kC3;             // [32]
constexpr int kOffBH = kOffWH + kC3;             // [1]
constexpr int kNumParams = kOffBH + 1;           // 23393

// Activation bytes per block (h1 + h2 + h3, BF16).
constexpr long kActElems1 = (long)kH1 * kW1 * kC1;
constexpr long kActElems2 = (long)kH2 * kW2 * kC2;
constexpr long kActElems3 = (long)kH3 * kW3 * kC3;

struct Weights {
  float* params = nullptr;          // FP32 master copy, kNumParams
  unsigned short* w1b = nullptr;    // [16][16]
  unsigned short* w2b = nullptr;    // [32][15][16]
  unsigned short* w3b = nullptr;    // [32][15][32]
};

struct Acts {
  unsigned short* h1 = nullptr;  // [B][16][512][16]
  unsigned short* h2 = nullptr;  // [B][8][256][32]
  unsigned short* h3 = nullptr;  // [B][4][128][32]
};

// z[blk][16][1024] (BF16) from a [16][N] coarse channel (dtype 0 f32, 1 f16,
// 2 bf16; N = nblk * 4096 columns read) and stats = {median, std} of its
// column sums on the device.
hipError_t launch_pool(const void* x, int dtype, long stride, int nblk, const float* stats,
                       unsigned short* z, hipStream_t stream);

// Block logits (and argmax positions when argmax != nullptr) for B tiles.
hipError_t launch_forward(const Weights& w, const unsigned short* z, int B, Acts& a,
                          float* logits, int* argmax, hipStream_t stream);

// Forward + loss + gradient: grad (kNumParams), loss (1 float, mean weighted
// BCE). gpart: B * kNumParams floats of scratch; dz: B floats.
hipError_t launch_gradient(const Weights& w, const unsigned short* z, const float* labels, int B,
                           float pos_weight, Acts& a, float* logits, int* argmax, float* dz,
                           float* gpart, float* grad, float* loss, hipStream_t stream);

// Adam step (step counts from 1) on the master weights, then launch_pack.
hipError_t launch_adam(Weights& w, float* m, float* v, const float* grad, float lr, int step,
                       hipStream_t stream);

// Refreshes the BF16 weight copies from the master weights.
hipError_t launch_pack(Weights& w, hipStream_t stream);

// Ascending list of the channel tiles (tile_width channels each, N channels)
// that overlap a block whose logit exceeds threshold; count[0] = list length,
// flagged[0] (if not null) = number of flagged blocks.
hipError_t launch_tile_list(const float* logits, int nblk, float threshold, int N,
                            int tile_width, int* list, int* count, int* flagged,
                            hipStream_t stream);

// Several channels of one shape (a gated batch run's front, FrontSet in
// dd_internal.h): launch_pool for each channel c into z + c * nblk tiles, and
// launch_tile_list for each channel from logits + c * nblk into its list and
// count (count[1] = flagged blocks); one launch each.
hipError_t launch_pool_set(const FrontSet& fs, int n, int dtype, long stride, int nblk,
                           unsigned short* z, hipStream_t stream);
hipError_t launch_tile_list_set(const float* logits, int nblk, float threshold, int N,
                                int tile_width, const FrontSet& fs, int n, hipStream_t stream);

}  // namespace pn
}  // namespace sc
