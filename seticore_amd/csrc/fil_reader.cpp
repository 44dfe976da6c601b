// Sigproc filterbank (.fil) input for the search: header parsing, metadata
// inference and coarse-channel loading with the loader-side conditioning the
// search expects.
//
// Header parsing follows fil_reader.cpp:31-131 (same attribute list and
// types, RA/Dec conversion fil_reader.cpp:16-24, data size checks 112-130);
// metadata inference follows filterbank_metadata.cpp:53-93. The reference's
// FilReader::loadCoarseChannel is unimplemented (fil_reader.cpp:151-154);
// here it is implemented with the conditioning of the HDF5 loader
// (h5_reader.cpp:195-209): rows beyond the file's time steps are zero, and
// with a DC spike the centre channel of every loaded row becomes the mean of
// its two neighbours. Errors are status codes where the reference exits.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/seticore_amd.h"
#include "sc_error.h"

namespace {

sc_status fail(sc_status code, const std::string& msg) {
  sc::set_last_error(msg);
  return code;
}

constexpr int kNoTelescope = -1;  // util.h:14-18
constexpr int kGreenBank = 6;
constexpr int kAta = 9;

// fil_reader.cpp:16-24: HHMMSS.S / DDMMSS.S -> hours (degrees) with a fraction.
double sigproc_angle(double v) {
  const bool negative = v < 0;
  const double a = std::fabs(v);
  const int whole = (int)std::floor(a / 10000.0);
  const double rest = a - whole * 10000.0;
  const int minutes = (int)std::floor(rest / 100.0);
  const double seconds = rest - minutes * 100.0;
  const double out = whole + minutes / 60.0 + seconds / 3600.0;
  return negative ? -out : out;
}

class Reader {
 public:
  explicit Reader(FILE* f) : f_(f) {}
  bool ok() const { return ok_; }
  template <typename T>
  T basic() {
    T v{};
    if (std::fread(&v, sizeof(T), 1, f_) != 1) ok_ = false;
    return v;
  }
  std::string str() {
    const uint32_t n = basic<uint32_t>();
    if (!ok_ || n >= 256) {  // fil_reader.cpp:144
      ok_ = false;
      return std::string();
    }
    std::string s(n, '\0');
    if (n && std::fread(&s[0], 1, n, f_) != n) ok_ = false;
    return s;
  }

 private:
  FILE* f_;
  bool ok_ = true;
};

}  // namespace

struct sc_fil {
  FILE* file = nullptr;
  std::string path;
  long data_start = 0;
  sc_fil_meta meta;
};

extern "C" {

sc_status sc_fil_open(const char* filename, sc_fil** out) {
  if (!filename || !out) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  FILE* f = std::fopen(filename, "rb");
  if (!f) return fail(SC_ERR_IO, std::string("cannot open ") + filename);
  sc_fil* h = new sc_fil();
  h->file = f;
  h->path = filename;
  std::memset(&h->meta, 0, sizeof(h->meta));
  sc_fil_meta& m = h->meta;
  m.telescope_id = kNoTelescope;
  Reader r(f);
  auto bad = [&](sc_status code, const std::string& msg) {
    std::fclose(f);
    delete h;
    return fail(code, msg);
  };
  if (r.str() != "HEADER_START" || !r.ok())
    return bad(SC_ERR_IO, std::string(filename) + " does not start with HEADER_START");
  int nbits = 32;
  long header_timesteps = 0;
  while (true) {
    const std::string name = r.str();
    if (!r.ok()) return bad(SC_ERR_IO, "truncated header");
    if (name == "telescope_id") {
      m.telescope_id = r.basic<int>();
    } else if (name == "machine_id" || name == "data_type" || name == "barycentric" ||
               name == "pulsarcentric" || name == "nifs" || name == "nbeams" ||
               name == "ibeam") {
      r.basic<int>();
    } else if (name == "nbits") {
      nbits = r.basic<int>();
    } else if (name == "nsamples") {
      header_timesteps = r.basic<int>();
    } else if (name == "nchans") {
      m.num_channels = r.basic<int>();
    } else if (name == "rawdatafile") {
      r.str();
    } else if (name == "source_name") {
      const std::string s = r.str();
      std::snprintf(m.source_name, sizeof(m.source_name), "%s", s.c_str());
    } else if (name == "az_start" || name == "za_start" || name == "refdm" || name == "period") {
      r.basic<double>();
    } else if (name == "tstart") {
      m.tstart = r.basic<double>();
    } else if (name == "tsamp") {
      m.tsamp = r.basic<double>();
    } else if (name == "fch1") {
      m.fch1 = r.basic<double>();
    } else if (name == "foff") {
      m.foff = r.basic<double>();
    } else if (name == "src_raj") {
      m.src_raj = sigproc_angle(r.basic<double>());
    } else if (name == "src_dej") {
      m.src_dej = sigproc_angle(r.basic<double>());
    } else if (name == "HEADER_END") {
      break;
    } else {
      return bad(SC_ERR_IO, "unhandled attr name: " + name);
    }
    if (!r.ok()) return bad(SC_ERR_IO, "truncated header at " + name);
  }
  if (nbits != 32) return bad(SC_ERR_UNSUPPORTED, "only 32-bit float data is handled");
  if (m.num_channels <= 0) return bad(SC_ERR_IO, "header has no nchans");
  h->data_start = std::ftell(f);
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f) - h->data_start;
  if (bytes % 4 != 0)
    return bad(SC_ERR_IO, "indivisible amount of data is " + std::to_string(bytes) + " bytes");
  const long floats = bytes / 4;
  if (floats % m.num_channels != 0)
    return bad(SC_ERR_IO, "data does not divide into " + std::to_string(m.num_channels) +
                              " frequencies");
  const long inferred = floats / m.num_channels;
  if (header_timesteps == 0) {
    m.num_timesteps = inferred;
  } else if (header_timesteps != inferred) {
    return bad(SC_ERR_IO, "header num timesteps is " + std::to_string(header_timesteps) +
                              " but inferred num timesteps is " + std::to_string(inferred));
  } else {
    m.num_timesteps = header_timesteps;
  }
  // filterbank_metadata.cpp:53-93
  auto green_bank = [&](long nfpc) {
    if ((m.num_timesteps == 16 || m.num_timesteps == 32 || m.num_timesteps == 15) &&
        m.num_channels % nfpc == 0) {
      if (m.coarse_channel_size == 0) m.coarse_channel_size = nfpc;
      m.has_dc_spike = 1;
      return true;
    }
    return false;
  };
  if (green_bank(1048576) || green_bank(1033216) || green_bank(999424)) {
    if (m.telescope_id != kNoTelescope && m.telescope_id != kGreenBank)
      return bad(SC_ERR_IO, "Green Bank shape with another telescope id");
  } else if (m.num_channels == 50331648) {
    if (m.telescope_id != kNoTelescope && m.telescope_id != kAta)
      return bad(SC_ERR_IO, "ATA shape with another telescope id");
    if (m.coarse_channel_size == 0) m.coarse_channel_size = 262144;
    m.has_dc_spike = 0;
  } else if (m.coarse_channel_size > 0) {
    m.has_dc_spike = m.telescope_id == kGreenBank;
  } else {
    return bad(SC_ERR_UNSUPPORTED,
               "unable to infer coarse channel size for data with dimensions: " +
                   std::to_string(m.num_timesteps) + " x " + std::to_string(m.num_channels) +
                   ". please set the nfpc header.");
  }
  m.num_coarse_channels = m.num_channels / m.coarse_channel_size;
  *out = h;
  return SC_OK;
}

sc_status sc_fil_set_coarse_channel_size(sc_fil* h, long coarse_channel_size) {
  if (!h) return fail(SC_ERR_INVALID_ARGUMENT, "null handle");
  if (coarse_channel_size <= 0 || h->meta.num_channels % coarse_channel_size != 0)
    return fail(SC_ERR_INVALID_ARGUMENT, "coarse channel size must divide nchans");
  // The DC fix averages the centre channel's two neighbours (h5_reader.cpp:202-209),
  // which a coarse channel needs at least three channels to have.
  if (h->meta.has_dc_spike && coarse_channel_size < 3)
    return fail(SC_ERR_INVALID_ARGUMENT,
                "coarse channels of fewer than 3 channels cannot take the DC-spike fix");
  h->meta.coarse_channel_size = coarse_channel_size;
  h->meta.num_coarse_channels = h->meta.num_channels / coarse_channel_size;
  return SC_OK;
}

sc_status sc_fil_metadata(const sc_fil* h, sc_fil_meta* out) {
  if (!h || !out) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  *out = h->meta;
  return SC_OK;
}

// Host copy of coarse channel i: host [rounded_timesteps][coarse_channel_size]
// float32, zero rows beyond the file's time steps, DC fix when the metadata
// says so (h5_reader.cpp:195-209).
sc_status sc_fil_read_coarse_channel(const sc_fil* h, int i, int rounded_timesteps,
                                     float* host) {
  if (!h || !host) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  const sc_fil_meta& m = h->meta;
  if (i < 0 || i >= m.num_coarse_channels)
    return fail(SC_ERR_INVALID_ARGUMENT, "coarse channel out of range");
  if (rounded_timesteps < m.num_timesteps)
    return fail(SC_ERR_INVALID_ARGUMENT, "rounded_timesteps below the file's time steps");
  const long ccs = m.coarse_channel_size;
  if (m.has_dc_spike && ccs < 3)
    return fail(SC_ERR_INVALID_ARGUMENT,
                "coarse channels of fewer than 3 channels cannot take the DC-spike fix");
  for (long t = 0; t < m.num_timesteps; ++t) {
    const long off = h->data_start + 4 * (t * (long)m.num_channels + i * ccs);
    if (std::fseek(h->file, off, SEEK_SET) != 0 ||
        std::fread(host + t * ccs, 4, (size_t)ccs, h->file) != (size_t)ccs)
      return fail(SC_ERR_IO, "short read of " + h->path);
  }
  std::memset(host + m.num_timesteps * ccs, 0,
              sizeof(float) * (size_t)(rounded_timesteps - m.num_timesteps) * ccs);
  if (m.has_dc_spike) {
    const long mid = ccs / 2;
    for (long t = 0; t < m.num_timesteps; ++t) {
      float* row = host + t * ccs;
      row[mid] = (float)((row[mid - 1] + row[mid + 1]) / 2.0);
    }
  }
  return SC_OK;
}

// The same into device memory (device [rounded_timesteps][coarse_channel_size]),
// through a pinned staging buffer; returns after the copy completes.
sc_status sc_fil_load_coarse_channel(const sc_fil* h, int i, int rounded_timesteps,
                                     float* device, void* stream) {
  if (!h || !device) return fail(SC_ERR_INVALID_ARGUMENT, "null argument");
  const size_t n = (size_t)rounded_timesteps * h->meta.coarse_channel_size;
  float* staging = nullptr;
  if (hipHostMalloc((void**)&staging, n * sizeof(float), hipHostMallocDefault) != hipSuccess)
    return fail(SC_ERR_OUT_OF_MEMORY, "pinned staging buffer");
  sc_status s = sc_fil_read_coarse_channel(h, i, rounded_timesteps, staging);
  if (s == SC_OK) {
    hipError_t e = hipMemcpyAsync(device, staging, n * sizeof(float), hipMemcpyHostToDevice,
                                  (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) s = fail(SC_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
  }
  (void)hipHostFree(staging);
  return s;
}

sc_status sc_fil_close(sc_fil* h) {
  if (!h) return SC_OK;
  if (h->file) std::fclose(h->file);
  delete h;
  return SC_OK;
}

}  // extern "C"
