// Hit output writers: the turboSETI-style .dat text file
// (dat_file_writer.cpp:15-73), the .hits stream of Cap'n Proto messages
// (hit_file_writer.cpp:16-97, schema hit.capnp:1-86) and the .events stream
// of cadence events (event_file_writer.cpp:19-69, hit.capnp:91-93).
//
// The .hits encoder is written against the Cap'n Proto wire format directly
// (the capnp library is not a dependency): one unpacked message per hit, laid
// out the way capnp's MallocMessageBuilder lays it out for this schema - a
// 1024-word first segment, objects allocated in the order the reference
// builds them (root pointer, Hit, Signal, Filterbank, source name, data), and
// a second segment with a far-pointer landing pad when the filterbank slice
// does not fit in the first.

#include <errno.h>
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/seticore_amd.h"
#include "sc_error.h"

namespace {

sc_status werr(const std::string& m) {
  sc::set_last_error(m);
  return SC_ERR_IO;
}

// The sc_hit field order is DedopplerHit's; lowIndex/highIndex follow
// dedoppler_hit.cpp:33-39.
int low_index(const sc_hit& h) { return std::min(h.index, h.index + h.drift_steps); }
int high_index(const sc_hit& h) { return std::max(h.index, h.index + h.drift_steps); }

std::string fmt_fixed(double v, int width, int prec) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%*.*f", width, prec, v);
  return buf;
}

// ---- Cap'n Proto wire encoding --------------------------------------------

constexpr long kFirstSegmentWords = 1024;

struct Message {
  std::vector<std::vector<uint64_t>> segs;
  long limit;  // words segment 0 may hold
  explicit Message(long first_segment_words = kFirstSegmentWords) : limit(first_segment_words) {
    segs.reserve(2);  // references to segs[0] stay valid when segment 1 is added
    segs.emplace_back();
  }
  // Allocates `words` in segment 0 if it fits; returns the word index or -1.
  long alloc0(long words) {
    auto& s = segs[0];
    if ((long)s.size() + words > limit) return -1;
    long at = (long)s.size();
    s.resize(s.size() + words, 0);
    return at;
  }
};

uint64_t struct_ptr(long from, long to, int data_words, int ptr_words) {
  const int64_t off = to - from - 1;
  const uint32_t lo = (uint32_t)(off << 2);
  const uint32_t hi = (uint32_t)data_words | ((uint32_t)ptr_words << 16);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

uint64_t list_ptr(long from, long to, int elem_size_code, uint32_t count) {
  const int64_t off = to - from - 1;
  const uint32_t lo = (uint32_t)(off << 2) | 1u;
  const uint32_t hi = (uint32_t)elem_size_code | (count << 3);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

void put_bytes(std::vector<uint64_t>& seg, long word, int byte_off, const void* src, size_t n) {
  std::memcpy(reinterpret_cast<char*>(seg.data() + word) + byte_off, src, n);
}
template <typename V>
void put(std::vector<uint64_t>& seg, long word, int byte_off, V v) {
  put_bytes(seg, word, byte_off, &v, sizeof(v));
}

struct HitsMeta {
  std::string source_name;
  double fch1, foff, tstart, tsamp, src_raj, src_dej;
  long num_timesteps, coarse_channel_size;
  int telescope_id;
  std::vector<std::string> beam_names;
  std::vector<double> beam_ras, beam_decs;
  bool coherent(int beam) const { return beam >= 0 && beam < (int)beam_names.size(); }
};

// Signal (hit.capnp:11-46, 6 data words) for `h`, pointed to from segment-0
// word `ptr_word` (hit_file_writer.cpp:54-65).
void encode_signal(Message& msg, long ptr_word, const sc_hit& h) {
  auto& s0 = msg.segs[0];
  const long sig = msg.alloc0(6);
  s0[ptr_word] = struct_ptr(ptr_word, sig, 6, 0);
  put<double>(s0, sig, 0, h.frequency);
  put<int32_t>(s0, sig, 8, h.index);
  put<int32_t>(s0, sig, 12, h.drift_steps);
  put<double>(s0, sig, 16, h.drift_rate);
  put<float>(s0, sig, 24, h.snr);
  put<int32_t>(s0, sig, 28, h.coarse_channel);
  put<int32_t>(s0, sig, 32, h.beam);
  put<int32_t>(s0, sig, 36, h.num_timesteps);
  put<float>(s0, sig, 40, h.power);
  put<float>(s0, sig, 44, h.incoherent_power);
}

// Filterbank (hit.capnp:50-78: 9 data words + 2 pointers) for the channel
// range around [low, high] of one coarse channel (hit_file_writer.cpp:67-97),
// pointed to from segment-0 word `ptr_word`. The data goes to a second
// segment behind a far pointer when segment 0 is full.
void encode_filterbank(Message& msg, long ptr_word, const HitsMeta& m, int beam,
                       int coarse_channel, int low, int high, const float* input) {
  auto& s0 = msg.segs[0];
  const long fb = msg.alloc0(11);
  s0[ptr_word] = struct_ptr(ptr_word, fb, 9, 2);
  const std::string& name = m.coherent(beam) ? m.beam_names[beam] : m.source_name;
  const double ra = m.coherent(beam) ? m.beam_ras[beam] : m.src_raj;
  const double dec = m.coherent(beam) ? m.beam_decs[beam] : m.src_dej;
  const long extra = 40;  // EXTRA_COLUMNS, hit_file_writer.cpp:16
  const long begin = std::max((long)low - extra, 0L);
  const long end = std::min((long)high + extra, m.coarse_channel_size) - 1;
  const long nch = end - begin;
  const long coarse_offset = (long)coarse_channel * m.coarse_channel_size;
  put<double>(s0, fb, 0, m.fch1 + (coarse_offset + begin) * m.foff);
  put<double>(s0, fb, 8, m.foff);
  put<double>(s0, fb, 16, m.tstart);
  put<double>(s0, fb, 24, m.tsamp);
  put<double>(s0, fb, 32, ra);
  put<double>(s0, fb, 40, dec);
  put<int32_t>(s0, fb, 48, m.telescope_id);
  put<int32_t>(s0, fb, 52, (int32_t)m.num_timesteps);
  put<int32_t>(s0, fb, 56, (int32_t)nch);
  put<int32_t>(s0, fb, 60, coarse_channel);
  put<int32_t>(s0, fb, 64, (int32_t)begin);
  put<int32_t>(s0, fb, 68, beam);

  // sourceName: Text = byte list with a NUL terminator.
  const long name_ptr = fb + 9;
  const long text_bytes = (long)name.size() + 1;
  const long text_words = (text_bytes + 7) / 8;
  const long text = msg.alloc0(text_words);
  s0[name_ptr] = list_ptr(name_ptr, text, 2, (uint32_t)text_bytes);
  put_bytes(s0, text, 0, name.data(), name.size());

  // data: List(Float32), [t][c] over the slice.
  const long data_ptr = fb + 10;
  const long n = nch > 0 ? (long)m.num_timesteps * nch : 0;
  std::vector<float> vals((size_t)std::max(n, 0L));
  for (long t = 0; t < m.num_timesteps && nch > 0; ++t)
    for (long c = 0; c < nch; ++c)
      vals[t * nch + c] = input ? input[t * m.coarse_channel_size + begin + c] : 0.0f;
  const long data_words = (n * 4 + 7) / 8;
  const long at = msg.alloc0(data_words);
  if (at >= 0) {
    s0[data_ptr] = list_ptr(data_ptr, at, 4, (uint32_t)n);
    if (n) put_bytes(s0, at, 0, vals.data(), n * 4);
  } else {
    // New segment: landing pad (a list pointer to the words right after it),
    // then the data; the original pointer becomes a far pointer to the pad.
    msg.segs.emplace_back((size_t)(1 + data_words), 0);
    auto& s1 = msg.segs[1];
    s0[data_ptr] = (uint64_t)2u | ((uint64_t)1u << 32);  // far, offset 0, segment 1
    s1[0] = list_ptr(0, 1, 4, (uint32_t)n);
    if (n) put_bytes(s1, 1, 0, vals.data(), n * 4);
  }
}

// Stream framing: segment count - 1, segment sizes, padding to 8 bytes.
std::vector<char> frame(const Message& msg) {
  const uint32_t nseg = (uint32_t)msg.segs.size();
  std::vector<uint32_t> table;
  table.push_back(nseg - 1);
  for (auto& s : msg.segs) table.push_back((uint32_t)s.size());
  if (table.size() % 2) table.push_back(0);
  std::vector<char> out(table.size() * 4);
  std::memcpy(out.data(), table.data(), out.size());
  for (auto& s : msg.segs) {
    const char* p = reinterpret_cast<const char*>(s.data());
    out.insert(out.end(), p, p + s.size() * 8);
  }
  return out;
}

// One message for one hit (hit_file_writer.cpp:34-97): root pointer, Hit,
// Signal, Filterbank, source name, data, in the order the reference builds them.
std::vector<char> encode_hit(const HitsMeta& m, const sc_hit& h, const float* input) {
  Message msg;
  auto& s0 = msg.segs[0];
  const long root = msg.alloc0(1);
  const long hit = msg.alloc0(2);  // Hit: 0 data words, 2 pointers
  s0[root] = struct_ptr(root, hit, 0, 2);
  encode_signal(msg, hit, h);
  encode_filterbank(msg, hit + 1, m, h.beam, h.coarse_channel, low_index(h), high_index(h),
                    input);
  return frame(msg);
}

// One message for one cadence event (event_file_writer.cpp:35-69, schema
// hit.capnp:91-93): an Event with one Hit per file; every Hit has a
// Filterbank over the event's channel range in that file, and a Signal when
// the file has a hit. A single segment holds it (any segmentation is valid
// Cap'n Proto; readers do not depend on it).
std::vector<char> encode_event(const std::vector<HitsMeta>& metas, const sc_hit* const* hits,
                               const float* const* inputs) {
  const long n = (long)metas.size();
  Message msg(1L << 40);
  auto& s0 = msg.segs[0];
  const long root = msg.alloc0(1);
  const long ev = msg.alloc0(1);  // Event: 0 data words, 1 pointer
  s0[root] = struct_ptr(root, ev, 0, 1);
  // List(Hit): inline composite, tag word + n elements of 0 data / 2 pointer words.
  const long tag = msg.alloc0(1 + 2 * n);
  s0[ev] = list_ptr(ev, tag, 7, (uint32_t)(2 * n));
  s0[tag] = ((uint64_t)(uint32_t)(n << 2)) | ((uint64_t)(0u | (2u << 16)) << 32);
  int low = low_index(*hits[0]), high = high_index(*hits[0]);
  for (long i = 1; i < n; ++i) {
    if (!hits[i]) continue;
    low = std::min(low, low_index(*hits[i]));
    high = std::max(high, high_index(*hits[i]));
  }
  const int beam = hits[0]->beam, coarse_channel = hits[0]->coarse_channel;
  for (long i = 0; i < n; ++i) {
    const long elem = tag + 1 + 2 * i;
    encode_filterbank(msg, elem + 1, metas[i], beam, coarse_channel, low, high, inputs[i]);
    if (hits[i]) encode_signal(msg, elem, *hits[i]);
  }
  return frame(msg);
}

HitsMeta hits_meta(const sc_file_meta* meta) {
  HitsMeta m;
  m.source_name = meta->source_name ? meta->source_name : "";
  m.fch1 = meta->fch1;
  m.foff = meta->foff;
  m.tstart = meta->tstart;
  m.tsamp = meta->tsamp;
  m.src_raj = meta->src_raj;
  m.src_dej = meta->src_dej;
  m.num_timesteps = meta->num_timesteps;
  m.coarse_channel_size = meta->coarse_channel_size;
  m.telescope_id = meta->telescope_id;
  return m;
}

sc_status write_all(int fd, const std::vector<char>& bytes) {
  size_t done = 0;
  while (done < bytes.size()) {
    ssize_t k = ::write(fd, bytes.data() + done, bytes.size() - done);
    if (k < 0) {
      if (errno == EINTR) continue;
      return werr("write failed");
    }
    done += (size_t)k;
  }
  return SC_OK;
}

}  // namespace

struct sc_writer {
  bool dat = false;
  bool events = false;
  std::vector<HitsMeta> event_metas;  // .events: one per cadence file
  // .dat
  FILE* f = nullptr;
  int hit_count = 0;
  // .hits
  int fd = -1;
  std::string tmp_name, final_name;
  HitsMeta meta;
};

extern "C" {

sc_status sc_dat_writer_open(const char* filename, const sc_file_meta* meta, double max_drift,
                             sc_writer** out) {
  if (!filename || !meta || !out) return werr("null argument");
  FILE* f = std::fopen(filename, "w");
  if (!f) return werr(std::string("could not open ") + filename);
  std::string fn(filename);
  const size_t slash = fn.find_last_of('/');
  const std::string base = slash == std::string::npos ? fn : fn.substr(slash + 1);
  std::string h;
  h += "# -------------------------- o --------------------------\n";
  h += "# File ID: " + base + " \n";
  h += "# -------------------------- o --------------------------\n";
  h += std::string("# Source:") + (meta->source_name ? meta->source_name : "") + "\n";
  h += "# MJD: " + fmt_fixed(meta->tstart, 18, 12) + "\tRA: " + fmt_fixed(meta->src_raj, 0, 6) +
       "\tDEC: " + fmt_fixed(meta->src_dej, 0, 6) + "\n";
  h += "# DELTAT: " + fmt_fixed(meta->tsamp, 10, 6) + "\tDELTAF(Hz): " +
       fmt_fixed(meta->foff * 1000000.0, 10, 6) + "\tmax_drift_rate: " +
       fmt_fixed(max_drift, 10, 6) + "\tobs_length: " +
       fmt_fixed(meta->num_timesteps * meta->tsamp, 10, 6) + "\n";
  h += "# --------------------------\n"
       "# Top_Hit_# \t"
       "Drift_Rate \t"
       "SNR \t"
       "Uncorrected_Frequency \t"
       "Corrected_Frequency \t"
       "Index \t"
       "freq_start \t"
       "freq_end \t"
       "SEFD \t"
       "SEFD_freq \t"
       "Coarse_Channel_Number \t"
       "Full_number_of_hits \t"
       "\n"
       "# --------------------------\n";
  std::fputs(h.c_str(), f);
  std::fflush(f);
  sc_writer* w = new sc_writer();
  w->dat = true;
  w->f = f;
  w->meta.fch1 = meta->fch1;
  w->meta.foff = meta->foff;
  w->meta.coarse_channel_size = meta->coarse_channel_size;
  *out = w;
  return SC_OK;
}

sc_status sc_hits_writer_open(const char* filename, const sc_file_meta* meta, int num_beams,
                              const char* const* beam_source_names, const double* beam_ras,
                              const double* beam_decs, sc_writer** out) {
  if (!filename || !meta || !out) return werr("null argument");
  if (num_beams < 0 || (num_beams > 0 && (!beam_source_names || !beam_ras || !beam_decs)))
    return werr("bad beam metadata");
  sc_writer* w = new sc_writer();
  w->final_name = filename;
  w->tmp_name = w->final_name + ".tmp";
  // hit_file_writer.cpp:22: no O_TRUNC, like the reference.
  w->fd = open(w->tmp_name.c_str(), O_WRONLY | O_CREAT, 0664);
  if (w->fd < 0) {
    int e = errno;
    delete w;
    return werr("could not open " + std::string(filename) + ".tmp for writing. errno = " +
                std::to_string(e));
  }
  w->meta = hits_meta(meta);
  HitsMeta& m = w->meta;
  for (int b = 0; b < num_beams; ++b) {
    m.beam_names.push_back(beam_source_names[b] ? beam_source_names[b] : "");
    m.beam_ras.push_back(beam_ras[b]);
    m.beam_decs.push_back(beam_decs[b]);
  }
  *out = w;
  return SC_OK;
}

// event_file_writer.cpp:19-28: opened like the .hits writer (no O_TRUNC,
// renamed on close).
sc_status sc_events_writer_open(const char* filename, const sc_file_meta* metas, int num_files,
                                sc_writer** out) {
  if (!filename || !metas || !out || num_files <= 0) return werr("null argument");
  sc_writer* w = new sc_writer();
  w->events = true;
  w->final_name = filename;
  w->tmp_name = w->final_name + ".tmp";
  w->fd = open(w->tmp_name.c_str(), O_WRONLY | O_CREAT, 0664);
  if (w->fd < 0) {
    int e = errno;
    delete w;
    return werr("could not open " + std::string(filename) + ".tmp for writing. errno = " +
                std::to_string(e));
  }
  for (int i = 0; i < num_files; ++i) w->event_metas.push_back(hits_meta(&metas[i]));
  *out = w;
  return SC_OK;
}

sc_status sc_events_writer_record(sc_writer* w, const sc_hit* const* hits,
                                  const float* const* inputs) {
  if (!w || !hits || !inputs || !w->events) return werr("not an events writer");
  if (!hits[0]) return werr("the first file of an event must have a hit");
  for (size_t i = 0; i < w->event_metas.size(); ++i)
    if (!inputs[i]) return werr("events need every file's coarse channel data");
  return write_all(w->fd, encode_event(w->event_metas, hits, inputs));
}

sc_status sc_writer_record_hit(sc_writer* w, const sc_hit* hit, const float* input) {
  if (!w || !hit) return werr("null argument");
  if (w->events) return werr("use sc_events_writer_record for an events writer");
  if (w->dat) {
    // dat_file_writer.cpp:58-73
    ++w->hit_count;
    int global_index = (int)(hit->coarse_channel * w->meta.coarse_channel_size) + hit->index;
    double frequency = w->meta.fch1 + global_index * w->meta.foff;
    std::string row = std::to_string(w->hit_count) + "\t" + fmt_fixed(hit->drift_rate, 10, 6) +
                      "\t" + fmt_fixed(hit->snr, 10, 6) + "\t" + fmt_fixed(frequency, 14, 6) +
                      "\t" + fmt_fixed(frequency, 14, 6) + "\t" + std::to_string(hit->index) +
                      "\t" + fmt_fixed(frequency, 14, 6) + "\t" + fmt_fixed(frequency, 14, 6) +
                      "\t";
    row += "0.0\t0.000000\t" + std::to_string(hit->coarse_channel) + "\t1\t\n";
    if (std::fputs(row.c_str(), w->f) < 0) return werr("write failed");
    std::fflush(w->f);
    return SC_OK;
  }
  if (!input) return werr(".hits needs the coarse channel's input data");
  return write_all(w->fd, encode_hit(w->meta, *hit, input));
}

sc_status sc_writer_close(sc_writer* w) {
  if (!w) return SC_OK;
  sc_status s = SC_OK;
  if (w->dat) {
    if (w->f) std::fclose(w->f);
  } else {
    if (w->fd >= 0) close(w->fd);
    if (std::rename(w->tmp_name.c_str(), w->final_name.c_str()) != 0)
      s = werr("rename failed for " + w->final_name);
  }
  delete w;
  return s;
}

}  // extern "C"
