// Dedopplerer::addIncoherentPower through the C++ façade (dedoppler.cu:139-176):
// the call re-sorts the hits by drift steps (driftStepsLessThan,
// dedoppler_hit.cpp:69-71), and every hit must keep all of its own fields
// while gaining the incoherent power of its own path. The hits below start in
// an order the sort changes. Exit code 0 and "ok" on success.
#include <cmath>
#include <cstdio>
#include <vector>

#include "seticore_amd.hpp"

using namespace seticore_amd;

#define REQUIRE(c)                                                      \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "REQUIRE failed (line %d): %s\n", __LINE__, #c); \
      return 1;                                                         \
    }                                                                   \
  } while (0)

int main() {
  const int T = 16, N = 4096;
  // Small integers: every path sum is exact in float.
  FilterbankBuffer buffer(T, N);
  for (int t = 0; t < T; ++t)
    for (int c = 0; c < N; ++c) buffer.data[(size_t)t * N + c] = (float)((t * 131 + c * 7) % 23);

  const int drift_steps[] = {20, -5, 3, 22, 0, -17};
  const int n = (int)(sizeof(drift_steps) / sizeof(drift_steps[0]));
  std::vector<DedopplerHit> hits;
  for (int i = 0; i < n; ++i) {
    DedopplerHit h;
    h.frequency = 1000.0 + i;
    h.index = 500 + 300 * i;
    h.drift_steps = drift_steps[i];
    h.drift_rate = 0.25 * (i + 1);
    h.snr = 10.0f + i;
    h.coarse_channel = 7 + i;
    h.beam = i % 2 ? 1 : NO_BEAM;
    h.num_timesteps = 9 + i;
    h.power = 50.0f + i;
    h.incoherent_power = -1.0f;
    hits.push_back(h);
  }
  const std::vector<DedopplerHit> before = hits;

  Dedopplerer dedopplerer(T, N, -2.79e-6, 18.25, false);
  dedopplerer.addIncoherentPower(buffer, hits);
  REQUIRE((int)hits.size() == n);

  // The expected powers: each hit's path in the library's full tree of its
  // drift block (sc_dd_taylor_tree, [offset][channel]).
  float* tree = nullptr;
  REQUIRE(hipMalloc((void**)&tree, sizeof(float) * T * N) == hipSuccess);
  std::vector<float> host(T * N);
  bool moved = false;
  for (int i = 0; i < n; ++i) {
    if (i > 0) REQUIRE(hits[i - 1].drift_steps <= hits[i].drift_steps);
    // The same hit before the call (indices are unique).
    int j = 0;
    while (j < n && before[j].index != hits[i].index) ++j;
    REQUIRE(j < n);
    moved = moved || j != i;
    const DedopplerHit& a = before[j];
    const DedopplerHit& b = hits[i];
    REQUIRE(a.frequency == b.frequency && a.drift_steps == b.drift_steps &&
            a.drift_rate == b.drift_rate && a.snr == b.snr &&
            a.coarse_channel == b.coarse_channel && a.beam == b.beam &&
            a.num_timesteps == b.num_timesteps && a.power == b.power);
    // dedoppler.cu:156-157
    const int block = (int)std::floor((float)b.drift_steps / (T - 1));
    const int offset = b.drift_steps - block * (T - 1);
    check(sc_dd_taylor_tree(dedopplerer.handle(), buffer.device_data, SC_F32, T, N, N, block,
                            tree, nullptr));
    REQUIRE(hipMemcpy(host.data(), tree, sizeof(float) * T * N, hipMemcpyDeviceToHost) ==
            hipSuccess);
    REQUIRE(b.incoherent_power == host[(size_t)offset * N + b.index]);
    REQUIRE(b.incoherent_power > 0.0f);
  }
  REQUIRE(moved);
  (void)hipFree(tree);
  std::printf("ok %d hits\n", n);
  return 0;
}
