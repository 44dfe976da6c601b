// seticore_amd C++ façade: the reference's dedoppler classes over the C ABI.
//
// Same names, argument meaning and ownership as lacker/seticore:
//   FilterbankMetadata  (filterbank_metadata.h:15-62; the fields the search and
//                        the writers read, plus isCoherentBeam)
//   FilterbankBuffer    (filterbank_buffer.h:11-48; row-major data[time][chan],
//                        owning or a view)
//   DedopplerHit        (dedoppler_hit.h:11-66)
//   Dedopplerer         (dedoppler.h:15-79: search, addIncoherentPower,
//                        memoryUsage, print_hits)
//   HitRecorder, DatFileWriter, HitFileWriter, makeHitRecorder
//                       (hit_recorder.h, dat_file_writer.h, hit_file_writer.h)
//   FilReader, findEvents (fil_reader.h, find_events.h)
// Where the reference asserts or exits, the façade throws std::runtime_error.
// Header-only; link with -lseticore_amd.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "seticore_amd.h"

namespace seticore_amd {

const int NO_BEAM = -1;  // dedoppler_hit.h:9

inline void check(sc_status s) {
  if (s != SC_OK) throw std::runtime_error(std::string("seticore_amd: ") + sc_last_error_string());
}

// util.cpp:16-22
inline int roundUpToPowerOfTwo(int n) {
  int r = 1;
  while (r < n) r *= 2;
  return r;
}

class FilterbankMetadata {
 public:
  bool has_dc_spike = false;
  std::string source_name;
  std::vector<std::string> source_names;
  std::vector<double> ras, decs;
  long num_timesteps = 0, num_channels = 0, coarse_channel_size = 0, num_coarse_channels = 0;
  double fch1 = 0, foff = 0, tstart = 0, tsamp = 0;
  double src_raj = 0, src_dej = 0;
  int telescope_id = -1;

  // filterbank_metadata.cpp:116-118
  bool isCoherentBeam(int beam) const { return 0 <= beam && beam < (int)source_names.size(); }

  sc_meta cMeta() const {
    sc_meta m;
    m.fch1 = fch1;
    m.foff = foff;
    m.coarse_channel_size = coarse_channel_size;
    m.num_coherent_beams = (int)source_names.size();
    return m;
  }

  sc_file_meta cFileMeta() const {
    sc_file_meta m;
    m.source_name = source_name.c_str();
    m.fch1 = fch1;
    m.foff = foff;
    m.tstart = tstart;
    m.tsamp = tsamp;
    m.src_raj = src_raj;
    m.src_dej = src_dej;
    m.num_timesteps = num_timesteps;
    m.coarse_channel_size = coarse_channel_size;
    m.telescope_id = telescope_id;
    return m;
  }
};

// Row-major data[time][chan] float32. The owning form is pinned host memory
// mapped into the device address space (readable and writable from the host
// like the reference's managed buffer, filterbank_buffer.cu:13-20); pass
// Location::Device for device-resident storage, or use the view constructor.
class FilterbankBuffer {
 public:
  enum class Location { HostMapped, Device };

  const int num_timesteps;
  const int num_channels;
  const bool managed;
  const int size;
  const size_t bytes;
  float* data;         // host-visible pointer (nullptr for Device buffers)
  float* device_data;  // pointer handed to the kernels

  FilterbankBuffer(int num_timesteps, int num_channels, Location where = Location::HostMapped)
      : num_timesteps(num_timesteps), num_channels(num_channels), managed(true),
        size(num_timesteps * num_channels), bytes(sizeof(float) * (size_t)size),
        data(nullptr), device_data(nullptr), location_(where) {
    if (where == Location::HostMapped) {
      if (hipHostMalloc((void**)&data, bytes, hipHostMallocMapped) != hipSuccess)
        throw std::runtime_error("FilterbankBuffer: hipHostMalloc failed");
      if (hipHostGetDevicePointer((void**)&device_data, data, 0) != hipSuccess)
        throw std::runtime_error("FilterbankBuffer: no device mapping");
    } else {
      if (hipMalloc((void**)&device_data, bytes) != hipSuccess)
        throw std::runtime_error("FilterbankBuffer: hipMalloc failed");
    }
  }

  // A view on caller-owned device-accessible memory (filterbank_buffer.cu:22-33).
  FilterbankBuffer(int num_timesteps, int num_channels, float* device_ptr)
      : num_timesteps(num_timesteps), num_channels(num_channels), managed(false),
        size(num_timesteps * num_channels), bytes(sizeof(float) * (size_t)size),
        data(nullptr), device_data(device_ptr), location_(Location::Device) {}

  ~FilterbankBuffer() {
    if (!managed) return;
    if (location_ == Location::HostMapped)
      (void)hipHostFree(data);
    else
      (void)hipFree(device_data);
  }

  FilterbankBuffer(const FilterbankBuffer&) = delete;
  FilterbankBuffer& operator=(const FilterbankBuffer&) = delete;

  void zero() {
    if (data) {
      std::memset(data, 0, bytes);
    } else {
      check(hipMemset(device_data, 0, bytes) == hipSuccess ? SC_OK : SC_ERR_HIP);
    }
  }

  void set(int time, int channel, float value) {
    if (time < 0 || time >= num_timesteps || channel < 0 || channel >= num_channels)
      throw std::out_of_range("FilterbankBuffer::set");
    const size_t index = (size_t)time * num_channels + channel;
    if (data) {
      data[index] = value;
    } else if (hipMemcpy(device_data + index, &value, sizeof(float), hipMemcpyHostToDevice) !=
               hipSuccess) {
      throw std::runtime_error("FilterbankBuffer::set");
    }
  }

  float get(int time, int channel) const {
    (void)hipDeviceSynchronize();
    const size_t index = (size_t)time * num_channels + channel;
    if (data) return data[index];
    float v = 0.0f;
    if (hipMemcpy(&v, device_data + index, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("FilterbankBuffer::get");
    return v;
  }

 private:
  Location location_;
};

// filterbank_buffer.cu:71-78
inline std::unique_ptr<FilterbankBuffer> makeNoisyBuffer(int num_timesteps, int num_channels) {
  std::unique_ptr<FilterbankBuffer> buffer(new FilterbankBuffer(num_timesteps, num_channels));
  buffer->zero();
  for (int chan = 0; chan < buffer->num_channels; ++chan)
    buffer->set(0, chan, 0.1 * chan / buffer->num_channels);
  return buffer;
}

class DedopplerHit : public sc_hit {
 public:
  DedopplerHit() : sc_hit() {}
  explicit DedopplerHit(const sc_hit& h) : sc_hit(h) {}

  // dedoppler_hit.cpp:25-30 with util.cpp:103-108 (pluralize)
  std::string toString() const {
    char buf[256];
    const int n = drift_steps;
    std::snprintf(buf, sizeof(buf),
                  "coarse channel = %d, index = %d, snr = %.5f, drift rate = %.5f (%d bin%s)",
                  coarse_channel, index, snr, drift_rate, n, (n == 1 || n == -1) ? "" : "s");
    return buf;
  }
  int lowIndex() const { return std::min(index, index + drift_steps); }
  int highIndex() const { return std::max(index, index + drift_steps); }
};

class Dedopplerer {
 public:
  const int num_timesteps;
  const int num_channels;
  const double foff;
  const double tsamp;
  const bool has_dc_spike;
  bool print_hits = false;

  Dedopplerer(int num_timesteps, int num_channels, double foff, double tsamp, bool has_dc_spike)
      : num_timesteps(num_timesteps), num_channels(num_channels), foff(foff), tsamp(tsamp),
        has_dc_spike(has_dc_spike) {
    check(sc_dd_create(num_timesteps, num_channels, foff, tsamp, has_dc_spike ? 1 : 0, &dd_));
  }
  ~Dedopplerer() { sc_dd_destroy(dd_); }
  Dedopplerer(const Dedopplerer&) = delete;
  Dedopplerer& operator=(Dedopplerer&) = delete;

  // dedoppler.cu:186-320: appends this coarse channel's hits to *output.
  void search(const FilterbankBuffer& input, const FilterbankMetadata& metadata, int beam,
              int coarse_channel, double max_drift, double min_drift, double snr_threshold,
              std::vector<DedopplerHit>* output, hipStream_t stream = nullptr) {
    const sc_meta m = metadata.cMeta();
    size_t n = 0;
    check(sc_dd_search(dd_, input.device_data, SC_F32, input.num_timesteps, input.num_channels,
                       input.num_channels, &m, beam, coarse_channel, max_drift, min_drift,
                       snr_threshold, nullptr, 0, &n, stream));
    std::vector<sc_hit> hits(n);
    check(sc_dd_copy_hits(dd_, hits.data(), n, &n));
    for (const sc_hit& h : hits) {
      DedopplerHit hit(h);
      if (print_hits) std::printf("hit: %s\n", hit.toString().c_str());
      output->push_back(hit);
    }
  }

  // dedoppler.cu:139-176
  void addIncoherentPower(const FilterbankBuffer& input, std::vector<DedopplerHit>& hits,
                          hipStream_t stream = nullptr) {
    std::vector<sc_hit> raw(hits.begin(), hits.end());
    check(sc_dd_add_incoherent_power(dd_, input.device_data, SC_F32, input.num_timesteps,
                                     input.num_channels, input.num_channels, raw.data(),
                                     raw.size(), stream));
    for (size_t i = 0; i < raw.size(); ++i) hits[i] = DedopplerHit(raw[i]);
  }

  size_t memoryUsage() const {
    size_t b = 0;
    check(sc_dd_memory_usage(dd_, &b));
    return b;
  }

  sc_dd* handle() { return dd_; }

 private:
  sc_dd* dd_ = nullptr;
};

// hit_recorder.h:10-19
class HitRecorder {
 public:
  virtual void recordHit(DedopplerHit hit, const float* input) = 0;
  virtual ~HitRecorder() {}
};

class WriterBase : public HitRecorder {
 public:
  void recordHit(DedopplerHit hit, const float* input) override {
    check(sc_writer_record_hit(w_, &hit, input));
  }
  ~WriterBase() override { sc_writer_close(w_); }

 protected:
  sc_writer* w_ = nullptr;
};

// dat_file_writer.h:15-27
class DatFileWriter : public WriterBase {
 public:
  DatFileWriter(const std::string& filename, const FilterbankMetadata& metadata,
                double max_drift) {
    const sc_file_meta m = metadata.cFileMeta();
    check(sc_dat_writer_open(filename.c_str(), &m, max_drift, &w_));
  }
};

// hit_file_writer.h:16-31
class HitFileWriter : public WriterBase {
 public:
  HitFileWriter(const std::string& filename, const FilterbankMetadata& metadata) {
    const sc_file_meta m = metadata.cFileMeta();
    std::vector<const char*> names;
    for (const auto& s : metadata.source_names) names.push_back(s.c_str());
    check(sc_hits_writer_open(filename.c_str(), &m, (int)names.size(), names.data(),
                              metadata.ras.data(), metadata.decs.data(), &w_));
  }
};

// hit_recorder.cpp:10-23 (throws where the reference exits)
inline std::unique_ptr<HitRecorder> makeHitRecorder(const std::string& filename,
                                                    const FilterbankMetadata& metadata,
                                                    double max_drift) {
  auto ends_with = [&](const char* suffix) {
    const size_t n = std::strlen(suffix);
    return filename.size() >= n && filename.compare(filename.size() - n, n, suffix) == 0;
  };
  if (ends_with(".dat"))
    return std::unique_ptr<HitRecorder>(new DatFileWriter(filename, metadata, max_drift));
  if (ends_with(".hits")) return std::unique_ptr<HitRecorder>(new HitFileWriter(filename, metadata));
  throw std::runtime_error("could not recognize the file type of " + filename);
}

// fil_reader.h: a sigproc filterbank file with the reference's inferred
// metadata; loadCoarseChannel fills a device-visible buffer with zero padding
// and the DC-spike fix (the reference leaves it unimplemented,
// fil_reader.cpp:151-154).
class FilReader : public FilterbankMetadata {
 public:
  const std::string filename;

  explicit FilReader(const std::string& filename) : filename(filename) {
    check(sc_fil_open(filename.c_str(), &f_));
    sc_fil_meta m;
    check(sc_fil_metadata(f_, &m));
    source_name = m.source_name;
    fch1 = m.fch1;
    foff = m.foff;
    tstart = m.tstart;
    tsamp = m.tsamp;
    src_raj = m.src_raj;
    src_dej = m.src_dej;
    telescope_id = m.telescope_id;
    num_timesteps = m.num_timesteps;
    num_channels = m.num_channels;
    coarse_channel_size = m.coarse_channel_size;
    num_coarse_channels = m.num_coarse_channels;
    has_dc_spike = m.has_dc_spike != 0;
  }
  ~FilReader() { sc_fil_close(f_); }
  FilReader(const FilReader&) = delete;
  FilReader& operator=(const FilReader&) = delete;

  void loadCoarseChannel(int i, FilterbankBuffer* buffer, hipStream_t stream = nullptr) const {
    if (buffer->num_channels != coarse_channel_size)
      throw std::runtime_error("FilReader::loadCoarseChannel: buffer width mismatch");
    check(sc_fil_load_coarse_channel(f_, i, buffer->num_timesteps, buffer->device_data, stream));
  }

 private:
  sc_fil* f_ = nullptr;
};

// find_events.cpp:66-227 over .fil files; returns the number of events written.
inline long findEvents(const std::vector<std::string>& input_filenames,
                       const std::string& output_filename, double max_drift, double snr_on,
                       double snr_off) {
  std::vector<const char*> names;
  for (const auto& f : input_filenames) names.push_back(f.c_str());
  long n = 0;
  check(sc_find_events(names.data(), (int)names.size(), output_filename.c_str(), max_drift,
                       snr_on, snr_off, &n, nullptr));
  return n;
}

}  // namespace seticore_amd
