// C-ABI of the block proposal network (host side). No counterpart in
// lacker/seticore; see pn_kernels.hip for the network.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/seticore_amd.h"
#include "pn_internal.h"
#include "pn_state.h"
#include "sc_error.h"

namespace {

using namespace sc::pn;

sc_status fail(sc_status code, const std::string& msg) {
  sc::set_last_error(msg);
  return code;
}

sc_status hip_fail(hipError_t e, const char* where) {
  return fail(SC_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define PN_HIP(call, where)                            \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return hip_fail(_e, where);  \
  } while (0)

sc_status alloc(sc_pn* pn, void** p, size_t bytes, const char* what) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? SC_ERR_OUT_OF_MEMORY : SC_ERR_HIP,
                std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
  }
  pn->device_bytes += bytes;
  return SC_OK;
}

void release(sc_pn* pn) {
  (void)hipFree(pn->w.params);
  (void)hipFree(pn->w.w1b);
  (void)hipFree(pn->w.w2b);
  (void)hipFree(pn->w.w3b);
  (void)hipFree(pn->a.h1);
  (void)hipFree(pn->a.h2);
  (void)hipFree(pn->a.h3);
  (void)hipFree(pn->adam_m);
  (void)hipFree(pn->adam_v);
  (void)hipFree(pn->z);
  (void)hipFree(pn->logits);
  (void)hipFree(pn->argmax);
  (void)hipFree(pn->dz);
  (void)hipFree(pn->gpart);
  (void)hipFree(pn->grad);
  (void)hipFree(pn->loss);
}

// splitmix64
uint64_t next_u64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform(uint64_t* s) { return (next_u64(s) >> 11) * (1.0 / 9007199254740992.0); }

sc_status check_batch(const sc_pn* pn, int B) {
  if (B <= 0 || B > pn->max_batch)
    return fail(SC_ERR_INVALID_ARGUMENT, "num_blocks must be in [1, max_batch]");
  return SC_OK;
}

}  // namespace

extern "C" {

int sc_pn_num_params(void) { return kNumParams; }

sc_status sc_pn_create(int max_batch, sc_pn** out) {
  if (!out) return fail(SC_ERR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  if (max_batch <= 0 || max_batch > (1 << 16))
    return fail(SC_ERR_INVALID_ARGUMENT, "max_batch must be in [1, 65536]");
  sc_pn* pn = new sc_pn();
  pn->max_batch = max_batch;
  (void)hipGetDevice(&pn->device);
  const size_t B = max_batch;
  sc_status s = SC_OK;
  if (s == SC_OK) s = alloc(pn, (void**)&pn->w.params, kNumParams * 4, "parameters");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->w.w1b, kC1 * 16 * 2, "w1");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->w.w2b, kC2 * 15 * kC1 * 2, "w2");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->w.w3b, kC3 * 15 * kC2 * 2, "w3");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->a.h1, B * kActElems1 * 2, "h1");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->a.h2, B * kActElems2 * 2, "h2");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->a.h3, B * kActElems3 * 2, "h3");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->adam_m, kNumParams * 4, "adam m");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->adam_v, kNumParams * 4, "adam v");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->z, B * kRows * kPooled * 2, "tiles");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->logits, B * 4, "logits");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->argmax, B * 4, "argmax");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->dz, B * 4, "dz");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->gpart, B * kNumParams * 4, "gradient parts");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->grad, kNumParams * 4, "gradient");
  if (s == SC_OK) s = alloc(pn, (void**)&pn->loss, 4, "loss");
  if (s == SC_OK) {
    hipError_t e = hipMemset(pn->w.params, 0, kNumParams * 4);
    if (e == hipSuccess) e = hipMemset(pn->adam_m, 0, kNumParams * 4);
    if (e == hipSuccess) e = hipMemset(pn->adam_v, 0, kNumParams * 4);
    if (e == hipSuccess) e = launch_pack(pn->w, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) s = hip_fail(e, "initialise parameters");
  }
  if (s != SC_OK) {
    release(pn);
    delete pn;
    return s;
  }
  *out = pn;
  return SC_OK;
}

sc_status sc_pn_destroy(sc_pn* pn) {
  if (!pn) return SC_OK;
  release(pn);
  delete pn;
  return SC_OK;
}

sc_status sc_pn_memory_usage(const sc_pn* pn, size_t* bytes) {
  if (!pn || !bytes) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  *bytes = pn->device_bytes;
  return SC_OK;
}

sc_status sc_pn_set_params(sc_pn* pn, const float* params, size_t n) {
  if (!pn || !params) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  if (n != (size_t)kNumParams) return fail(SC_ERR_INVALID_ARGUMENT, "wrong parameter count");
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(params[i])) return fail(SC_ERR_INVALID_ARGUMENT, "non-finite parameter");
  PN_HIP(hipMemcpy(pn->w.params, params, n * 4, hipMemcpyHostToDevice), "upload parameters");
  PN_HIP(hipMemset(pn->adam_m, 0, n * 4), "reset adam");
  PN_HIP(hipMemset(pn->adam_v, 0, n * 4), "reset adam");
  pn->step = 0;
  PN_HIP(launch_pack(pn->w, nullptr), "pack weights");
  PN_HIP(hipDeviceSynchronize(), "pack weights");
  return SC_OK;
}

sc_status sc_pn_get_params(const sc_pn* pn, float* params, size_t n) {
  if (!pn || !params) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  if (n != (size_t)kNumParams) return fail(SC_ERR_INVALID_ARGUMENT, "wrong parameter count");
  PN_HIP(hipDeviceSynchronize(), "sync");
  PN_HIP(hipMemcpy(params, pn->w.params, n * 4, hipMemcpyDeviceToHost), "download parameters");
  return SC_OK;
}

// He-uniform weights (bound sqrt(6 / fan_in)), zero biases, from a seed.
sc_status sc_pn_init(sc_pn* pn, uint64_t seed) {
  if (!pn) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  std::vector<float> p(kNumParams, 0.0f);
  uint64_t s = seed;
  auto fill = [&](int off, int count, int fan_in) {
    const double bound = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < count; ++i) p[off + i] = (float)((2.0 * uniform(&s) - 1.0) * bound);
  };
  fill(kOffW1, kC1 * 15, 15);
  fill(kOffW2, kC2 * 15 * kC1, 15 * kC1);
  fill(kOffW3, kC3 * 15 * kC2, 15 * kC2);
  fill(kOffWH, kC3, kC3);
  return sc_pn_set_params(pn, p.data(), p.size());
}

sc_status sc_pn_pool(const void* data, int dtype, int rounded_timesteps, int num_channels,
                     long row_stride, const float* stats, void* tiles, void* stream) {
  if (!data || !stats || !tiles) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  if (dtype != SC_F32 && dtype != SC_F16 && dtype != SC_BF16)
    return fail(SC_ERR_INVALID_ARGUMENT, "unknown dtype");
  if (rounded_timesteps != kRows)
    return fail(SC_ERR_UNSUPPORTED, "the proposal network takes 16 time steps");
  if (num_channels <= 0 || num_channels % kBlockChannels != 0)
    return fail(SC_ERR_UNSUPPORTED, "num_channels must be a positive multiple of 4096");
  if (row_stride < num_channels) return fail(SC_ERR_INVALID_ARGUMENT, "row_stride < num_channels");
  PN_HIP(launch_pool(data, dtype, row_stride, num_channels / kBlockChannels, stats,
                     (unsigned short*)tiles, (hipStream_t)stream),
         "pool");
  return SC_OK;
}

sc_status sc_pn_forward(sc_pn* pn, const void* tiles, int num_blocks, float* logits,
                        void* stream) {
  if (!pn || !tiles || !logits) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  sc_status s = check_batch(pn, num_blocks);
  if (s != SC_OK) return s;
  PN_HIP(launch_forward(pn->w, (const unsigned short*)tiles, num_blocks, pn->a, logits, nullptr,
                        (hipStream_t)stream),
         "forward");
  return SC_OK;
}

sc_status sc_pn_gradient(sc_pn* pn, const void* tiles, const float* labels, int num_blocks,
                         float pos_weight, float* grad, float* loss, void* stream) {
  if (!pn || !tiles || !labels) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  sc_status s = check_batch(pn, num_blocks);
  if (s != SC_OK) return s;
  if (!(pos_weight > 0)) return fail(SC_ERR_INVALID_ARGUMENT, "pos_weight must be > 0");
  PN_HIP(launch_gradient(pn->w, (const unsigned short*)tiles, labels, num_blocks, pos_weight,
                         pn->a, pn->logits, pn->argmax, pn->dz, pn->gpart,
                         grad ? grad : pn->grad, loss ? loss : pn->loss, (hipStream_t)stream),
         "gradient");
  return SC_OK;
}

sc_status sc_pn_train_step(sc_pn* pn, const void* tiles, const float* labels, int num_blocks,
                           float lr, float pos_weight, float* loss, void* stream) {
  if (!(lr > 0)) return fail(SC_ERR_INVALID_ARGUMENT, "lr must be > 0");
  sc_status s = sc_pn_gradient(pn, tiles, labels, num_blocks, pos_weight, nullptr, loss, stream);
  if (s != SC_OK) return s;
  pn->step += 1;
  PN_HIP(launch_adam(pn->w, pn->adam_m, pn->adam_v, pn->grad, lr, pn->step, (hipStream_t)stream),
         "adam");
  return SC_OK;
}

sc_status sc_pn_activations(const sc_pn* pn, const void** h1, const void** h2, const void** h3,
                            const float** logits, const int** argmax) {
  if (!pn) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (h1) *h1 = pn->a.h1;
  if (h2) *h2 = pn->a.h2;
  if (h3) *h3 = pn->a.h3;
  if (logits) *logits = pn->logits;
  if (argmax) *argmax = pn->argmax;
  return SC_OK;
}

}  // extern "C"
