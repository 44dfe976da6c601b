// Cadence event search (find_events.cpp:16-227 of lacker/seticore): the files of
// an "ABACAD" cadence are searched coarse channel by coarse channel; a hit in
// the first "on" file becomes an event when every "on" file has exactly one
// hit where its drift predicts and no "off" file has a hit in the event's
// channel range. Events go to a .events file (event_file_writer.cpp).
//
// The matching rules are the reference's, including two of its behaviours:
// the range lookup (findHit, find_events.cpp:27-46) counts the first hit at or
// after the range start even when it lies beyond the range end, and the
// second range check spans [low - 50, high - 50] (find_events.cpp:206-207).
// Searches run on the GPU through the same entry points as sc_dd_search with
// host-exact statistics (the reference's); the inputs are .fil files.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "../../include/seticore_amd.h"
#include "sc_error.h"

namespace {

sc_status fail(sc_status code, const std::string& msg) {
  sc::set_last_error(msg);
  return code;
}

// util.cpp:69-84
bool float_eq(float a, float b) {
  while (a > 100) {
    a /= 2.0f;
    b /= 2.0f;
  }
  return std::fabs(a - b) <= 0.001f;
}

struct Found {
  int count;
  const sc_hit* hit;
};

// find_events.cpp:27-46 (inclusive range; the first hit at or after `low` is
// returned with count 1 unless a second one also starts at or before `high`).
Found find_hit(const std::map<int, const sc_hit*>& hitmap, int low, int high) {
  auto it = hitmap.lower_bound(low);
  if (it == hitmap.end()) return {0, nullptr};
  Found answer{1, it->second};
  ++it;
  if (it == hitmap.end() || it->first > high) return answer;
  return {2, nullptr};
}

// dedoppler_hit.cpp:40-44
int expected_index(const sc_hit& h, int timesteps) {
  const double per_step = (double)h.drift_steps / (double)(h.num_timesteps - 1);
  return h.index + (int)std::round(timesteps * per_step);
}

int low_index(const sc_hit& h) { return std::min(h.index, h.index + h.drift_steps); }
int high_index(const sc_hit& h) { return std::max(h.index, h.index + h.drift_steps); }

struct File {
  sc_fil* fil = nullptr;
  sc_fil_meta meta;
  sc_dd* dd = nullptr;
  float* device = nullptr;
  float* host = nullptr;  // pinned [T_r][coarse_channel_size]
  std::vector<sc_hit> hits;
};

struct Cleanup {
  std::vector<File>& files;
  sc_writer* writer = nullptr;
  ~Cleanup() {
    for (File& f : files) {
      if (f.fil) sc_fil_close(f.fil);
      if (f.dd) sc_dd_destroy(f.dd);
      if (f.device) (void)hipFree(f.device);
      if (f.host) (void)hipHostFree(f.host);
    }
    if (writer) sc_writer_close(writer);
  }
};

int round_up_pow2(long n) {
  int r = 1;
  while (r < n) r *= 2;
  return r;
}

sc_status search_file(File& f, int coarse_channel, double max_drift, double snr, bool drop_zero,
                      hipStream_t stream) {
  const int T_r = round_up_pow2(f.meta.num_timesteps);
  const long ccs = f.meta.coarse_channel_size;
  sc_status s = sc_fil_read_coarse_channel(f.fil, coarse_channel, T_r, f.host);
  if (s != SC_OK) return s;
  hipError_t e = hipMemcpyAsync(f.device, f.host, sizeof(float) * T_r * ccs,
                                hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return fail(SC_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
  sc_meta m;
  m.fch1 = f.meta.fch1;
  m.foff = f.meta.foff;
  m.coarse_channel_size = ccs;
  m.num_coherent_beams = 0;
  size_t n = 0;
  s = sc_dd_search(f.dd, f.device, SC_F32, T_r, (int)ccs, ccs, &m, -1, coarse_channel,
                   max_drift, 0.0, snr, nullptr, 0, &n, stream);
  if (s != SC_OK) return s;
  f.hits.resize(n);
  s = sc_dd_copy_hits(f.dd, f.hits.data(), n, &n);
  if (s != SC_OK) return s;
  if (drop_zero) {  // find_events.cpp:48-54
    std::vector<sc_hit> kept;
    for (const sc_hit& h : f.hits)
      if (h.drift_steps != 0) kept.push_back(h);
    f.hits.swap(kept);
  }
  return SC_OK;
}

}  // namespace

extern "C" sc_status sc_find_events(const char* const* filenames, int num_files,
                                    const char* output, double max_drift, double snr_on,
                                    double snr_off, long* num_events, void* stream_ptr) {
  if (!filenames || num_files <= 0 || !output)
    return fail(SC_ERR_INVALID_ARGUMENT, "need input files and an output file");
  hipStream_t stream = (hipStream_t)stream_ptr;
  std::vector<File> files((size_t)num_files);
  Cleanup cleanup{files};
  for (int i = 0; i < num_files; ++i) {
    if (!filenames[i]) return fail(SC_ERR_INVALID_ARGUMENT, "null file name");
    sc_status s = sc_fil_open(filenames[i], &files[i].fil);
    if (s != SC_OK) return s;
    sc_fil_metadata(files[i].fil, &files[i].meta);
  }
  // find_events.cpp:89-120: the cadence must line up, and follow ABACAD.
  const sc_fil_meta& m0 = files[0].meta;
  for (int i = 0; i < num_files; ++i) {
    const sc_fil_meta& m = files[i].meta;
    if (m.num_timesteps != m0.num_timesteps || m.coarse_channel_size != m0.coarse_channel_size ||
        m.num_coarse_channels != m0.num_coarse_channels || !float_eq(m.foff, m0.foff) ||
        !float_eq(m.tsamp, m0.tsamp) || m.has_dc_spike != m0.has_dc_spike)
      return fail(SC_ERR_INVALID_ARGUMENT,
                  std::string("file ") + filenames[i] + " does not match the first file's shape");
    const std::string name = m.source_name, first = m0.source_name;
    if (i % 2 == 0 && name != first)
      return fail(SC_ERR_INVALID_ARGUMENT, std::string("file ") + filenames[i] + " has source " +
                                               name + " when we expected source " + first);
    if (i % 2 == 1 && name == first && first != "VOYAGER-1")
      return fail(SC_ERR_INVALID_ARGUMENT, std::string("file ") + filenames[i] + " has source " +
                                               name + " but it is supposed to be an 'off'");
  }
  const int T_r = round_up_pow2(m0.num_timesteps);
  const long ccs = m0.coarse_channel_size;
  std::vector<sc_file_meta> wmeta((size_t)num_files);
  for (int i = 0; i < num_files; ++i) {
    File& f = files[i];
    sc_status s = sc_dd_create((int)f.meta.num_timesteps, (int)ccs, f.meta.foff, f.meta.tsamp,
                               f.meta.has_dc_spike, &f.dd);
    if (s != SC_OK) return s;
    if (hipMalloc((void**)&f.device, sizeof(float) * T_r * ccs) != hipSuccess ||
        hipHostMalloc((void**)&f.host, sizeof(float) * T_r * ccs, hipHostMallocDefault) !=
            hipSuccess)
      return fail(SC_ERR_OUT_OF_MEMORY, "coarse channel buffers");
    sc_file_meta& w = wmeta[i];
    w.source_name = f.meta.source_name;
    w.fch1 = f.meta.fch1;
    w.foff = f.meta.foff;
    w.tstart = f.meta.tstart;
    w.tsamp = f.meta.tsamp;
    w.src_raj = f.meta.src_raj;
    w.src_dej = f.meta.src_dej;
    w.num_timesteps = f.meta.num_timesteps;
    w.coarse_channel_size = ccs;
    w.telescope_id = f.meta.telescope_id;
  }
  sc_status s = sc_events_writer_open(output, wmeta.data(), num_files, &cleanup.writer);
  if (s != SC_OK) return s;

  long total = 0;
  std::vector<const float*> inputs((size_t)num_files);
  for (int i = 0; i < num_files; ++i) inputs[i] = files[i].host;
  for (int cc = 0; cc < m0.num_coarse_channels; ++cc) {
    // find_events.cpp:128-152: the first file, then the rest if it has hits.
    s = search_file(files[0], cc, max_drift, snr_on, true, stream);
    if (s != SC_OK) return s;
    if (files[0].hits.empty()) continue;
    for (int i = 1; i < num_files; ++i) {
      const bool on = i % 2 == 0;
      s = search_file(files[i], cc, max_drift, on ? snr_on : snr_off, on, stream);
      if (s != SC_OK) return s;
    }
    // find_events.cpp:159-165: hits keyed by start index (a later hit at the
    // same index replaces an earlier one).
    std::vector<std::map<int, const sc_hit*>> maps((size_t)num_files);
    for (int i = 0; i < num_files; ++i)
      for (const sc_hit& h : files[i].hits) maps[i][h.index] = &h;
    const double t0 = files[0].meta.tstart;
    for (const auto& kv : maps[0]) {
      const sc_hit& first = *kv.second;
      // find_events.cpp:175-196: each "on" file has exactly one hit where
      // the first hit's drift predicts it.
      bool ok = true;
      int low = INT_MAX, high = -1;
      for (int i = 0; i < num_files; i += 2) {
        const int timesteps =
            (int)std::round((files[i].meta.tstart - t0) / files[0].meta.tsamp);
        const int expected = expected_index(first, timesteps);
        const Found r = find_hit(maps[i], expected - 10, expected + 10);
        if (r.count != 1) {
          ok = false;
          break;
        }
        low = std::min(low, low_index(*r.hit));
        high = std::max(high, high_index(*r.hit));
      }
      if (!ok) continue;
      // find_events.cpp:202-217: "on" files one hit, "off" files none.
      std::vector<const sc_hit*> candidate;
      for (int i = 0; i < num_files; ++i) {
        const int ideal = i % 2 == 0 ? 1 : 0;
        const Found r = find_hit(maps[i], low - 50, high - 50);
        if (r.count != ideal) {
          ok = false;
          break;
        }
        candidate.push_back(r.hit);
      }
      if (!ok) continue;
      s = sc_events_writer_record(cleanup.writer, candidate.data(), inputs.data());
      if (s != SC_OK) return s;
      ++total;
    }
  }
  if (num_events) *num_events = total;
  s = sc_writer_close(cleanup.writer);
  cleanup.writer = nullptr;
  return s;
}
