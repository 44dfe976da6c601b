// The reference's known-answer dedoppler test (dedoppler_test.cpp:9-33),
// run through the C++ façade, plus the hit writers on its result.
// Exit code 0 and "ok" on success.
#include <cstdio>
#include <string>
#include <vector>

#include "seticore_amd.hpp"

using namespace seticore_amd;

#define REQUIRE(c)                                              \
  do {                                                          \
    if (!(c)) {                                                 \
      std::fprintf(stderr, "REQUIRE failed: %s\n", #c);         \
      return 1;                                                 \
    }                                                           \
  } while (0)

int main(int argc, char** argv) {
  const std::string out_dir = argc > 1 ? argv[1] : ".";
  int num_timesteps = 8;
  int num_channels = 1000;
  auto buffer = makeNoisyBuffer(num_timesteps, num_channels);
  FilterbankMetadata metadata;

  // Draw a line drifting a few timesteps to the right
  buffer->set(0, 70, 1.0);
  buffer->set(1, 70, 1.0);
  buffer->set(2, 71, 1.0);
  buffer->set(3, 71, 1.0);
  buffer->set(4, 72, 1.0);
  buffer->set(5, 72, 1.0);
  buffer->set(6, 73, 1.0);
  buffer->set(7, 73, 1.0);

  Dedopplerer dedopplerer(num_timesteps, num_channels, 1.0, 1.0, false);
  std::vector<DedopplerHit> hits;
  dedopplerer.search(*buffer, metadata, NO_BEAM, 555, 0.01, 0.01, 200.0, &hits);
  REQUIRE(hits.size() == 1);
  REQUIRE(hits[0].index == 70);
  REQUIRE(hits[0].drift_steps == 3);
  REQUIRE(hits[0].beam == NO_BEAM);
  REQUIRE(hits[0].coarse_channel == 555);
  REQUIRE(dedopplerer.memoryUsage() > 0);

  metadata.coarse_channel_size = num_channels;
  metadata.num_timesteps = num_timesteps;
  metadata.source_name = "kat";
  {
    auto dat = makeHitRecorder(out_dir + "/kat.dat", metadata, 0.01);
    auto hitf = makeHitRecorder(out_dir + "/kat.hits", metadata, 0.01);
    for (const auto& h : hits) {
      dat->recordHit(h, buffer->data);
      hitf->recordHit(h, buffer->data);
    }
  }
  std::printf("ok %s\n", hits[0].toString().c_str());
  return 0;
}
