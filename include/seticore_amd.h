/*
 * seticore_amd: MI355X-native Taylor-tree dedoppler search, C ABI.
 *
 * This is the drop-in boundary for lacker/seticore's Dedopplerer
 * (dedoppler.h:15-79, dedoppler.cu:79-320). The reference exposes a C++ class;
 * every entry point below replaces one of its members, cited per function.
 * Plain pointers and sizes only: no C++ types, no torch types. All device
 * pointers are HIP device (or managed) allocations on the current device.
 *
 * Errors: every function returns an sc_status. On failure the message for the
 * calling thread is available from sc_last_error_string(). Nothing in this
 * library calls exit() or abort().
 */
#ifndef SETICORE_AMD_H
#define SETICORE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SC_ABI_VERSION 1

typedef enum {
  SC_OK = 0,
  SC_ERR_INVALID_ARGUMENT = 1, /* a precondition the reference asserts (dedoppler.cu:191-192) */
  SC_ERR_HIP = 2,              /* a HIP runtime error (the reference exits with code 2) */
  SC_ERR_OUT_OF_MEMORY = 3,
  SC_ERR_UNSUPPORTED = 4,      /* a shape this build has no kernel for */
  SC_ERR_IO = 5,               /* output file errors (writers) */
  SC_ERR_INTERNAL = 6          /* a consistency check inside the library failed */
} sc_status;

/* Storage type of the power samples handed to sc_dd_search. The reference
 * only has float32 (filterbank_buffer.h:35). The half-width forms hold
 * normalised power (see sc_normalize) and are accumulated in float32. */
typedef enum { SC_F32 = 0, SC_F16 = 1, SC_BF16 = 2 } sc_dtype;

/* Where the column-sum statistics and the window merge run.
 * SC_STATS_HOST_EXACT follows dedoppler.cu:232-319 on the host: the column
 * sums and top arrays are copied back and the median/trimmed deviation are
 * computed with the same std::nth_element calls, so SNR values are bit-equal
 * to the reference. SC_STATS_DEVICE keeps everything on the GPU: exact rank
 * median by radix select, FP64 trimmed mean/deviation over the sorted ranks,
 * and a wave-per-window merge; only the per-window candidates come back.
 * Default: SC_STATS_HOST_EXACT. */
typedef enum { SC_STATS_HOST_EXACT = 0, SC_STATS_DEVICE = 1 } sc_stats_mode;

/* The metadata fields a search reads (filterbank_metadata.h:15-62 and
 * dedoppler_hit.cpp:11-23): isCoherentBeam(beam) is
 * 0 <= beam < num_coherent_beams (filterbank_metadata.cpp:116-118). */
typedef struct {
  double fch1;               /* MHz, frequency of channel 0 of the file */
  double foff;               /* MHz per fine channel */
  long coarse_channel_size;  /* fine channels per coarse channel */
  int num_coherent_beams;    /* size of the per-beam source list */
} sc_meta;

/* One hit: the fields of DedopplerHit (dedoppler_hit.h:11-45), same meaning. */
typedef struct {
  double frequency;       /* MHz: fch1 + (coarse_channel*coarse_channel_size + index)*foff */
  int index;              /* start fine channel within the coarse channel */
  int drift_steps;        /* drift_block*(T_r-1) + path_offset */
  double drift_rate;      /* Hz/s: drift_steps * resolution */
  float snr;              /* (power - median) / std */
  int coarse_channel;
  int beam;               /* beam if coherent, else -1 (NO_BEAM) */
  int num_timesteps;      /* unrounded */
  float power;            /* the path sum */
  float incoherent_power; /* filled by sc_dd_add_incoherent_power, else 0 */
} sc_hit;

typedef struct sc_dd sc_dd;

/* Replaces Dedopplerer::Dedopplerer (dedoppler.cu:79-111). num_timesteps is
 * unrounded (> 1); the search runs over roundUpToPowerOfTwo(num_timesteps)
 * rows. Allocates all device scratch once, on the current device. */
sc_status sc_dd_create(int num_timesteps, int num_channels, double foff_mhz,
                       double tsamp_s, int has_dc_spike, sc_dd** out);

/* Replaces Dedopplerer::~Dedopplerer (dedoppler.cu:113-124). */
sc_status sc_dd_destroy(sc_dd* dd);

/* Replaces Dedopplerer::memoryUsage (dedoppler.cu:127-130): device bytes held. */
sc_status sc_dd_memory_usage(const sc_dd* dd, size_t* bytes);

/* Selects sc_stats_mode for later searches (no counterpart; default host-exact). */
sc_status sc_dd_set_stats_mode(sc_dd* dd, int mode);

/* Shape of the one-pass tree kernel used when rounded_timesteps <= 256 (the
 * default; selecting a shape undoes sc_dd_set_two_pass). At 64..256 (level-16
 * trees staged in LDS) the shape is (4, drift blocks per workgroup in
 * {1, 2, 4, 8, 16, 32, 64, 128, 256}, 0); at 128 and 256 also (4, G, 1), one workgroup per
 * group of offsets instead of one for all. At 32 the same kernel is (4, G, 8) (the default);
 * otherwise `prefetch` is the number of waves per workgroup sharing the
 * window of the register-parked kernel: (C, G, waves) in {(4,4,0), (2,2,0),
 * (2,4,0), (2,4,2), (2,8,2), (2,8,4), (4,8,2)}, 0 meaning one. At 16 and below:
 * channels held per lane (4 or 8), drift blocks per LDS window (a multiple
 * of channels_per_lane) and the window mode in `prefetch` (0: one LDS window
 * per workgroup; 1: each workgroup walks several windows, prefetching the
 * next into registers; 2: several windows, each loaded after the previous
 * one is done). Returns SC_ERR_UNSUPPORTED for shapes that are not
 * compiled. Results do not depend on the shape; only speed does. No
 * counterpart in the reference (its tile shape is fixed, taylor.cu:170-189);
 * used by the autotuner. */
sc_status sc_dd_set_tuning(sc_dd* dd, int channels_per_lane, int drift_group, int prefetch);

/* Selects the two-pass search (level-16 trees through an HBM scratch array)
 * for rounded_timesteps 32..256 instead of the one-pass kernels (the
 * reference's tiledTaylorTree<32> / twoStageTaylorTree range, taylor.cu:261-350):
 * drift blocks per LDS window in the level-16 pass (4 or 8) and drift blocks
 * per launch pair (1..64; 0 = as many as keep the scratch within 4.5 GB, at
 * most drift_group). May grow the scratch (and drops cached graphs when it
 * does). Results do not depend on the shape. No counterpart in the reference;
 * used by the autotuner. */
sc_status sc_dd_set_two_pass(sc_dd* dd, int drift_group, int blocks_per_launch);

/* Whether device-statistics searches replay a cached hipGraph of their
 * launch sequence (one graph per input pointer and search parameters, up to
 * 256; captured the second time the same search is made, so a caller that
 * passes a new buffer every time never captures). Default on. No counterpart
 * in the reference. */
sc_status sc_dd_set_graphs(sc_dd* dd, int enable);

/* Kernel timing for measurement (no counterpart in the reference, which only
 * times whole runs, taylor_benchmark.cu:22-40): when enabled, each search
 * records HIP events on its stream around the path-sum kernels (the tree and
 * top-path tracking over all drift blocks) and accumulates their elapsed time.
 * Enabling (or disabling) resets the totals. */
sc_status sc_dd_set_timing(sc_dd* dd, int enable);
sc_status sc_dd_kernel_time(const sc_dd* dd, double* total_ms, long* searches);

/* Replaces Dedopplerer::search (dedoppler.cu:186-320).
 * data: device pointer to [rounded_timesteps][row_stride] samples of `dtype`
 *       (row_stride >= num_channels, in elements; rows beyond the unrounded
 *       num_timesteps must already be zero, as the reference's loader leaves
 *       them, h5_reader.cpp:195-200).
 * The hits are kept in the handle (ascending channel order, dedoppler.cu:275).
 * Up to `capacity` of them are copied to `hits`; *num_hits receives the total,
 * which may exceed capacity (then fetch the rest with sc_dd_copy_hits).
 * stream: a hipStream_t (NULL = default stream). The call returns after the
 * search is complete, like the reference's blocking copies (dedoppler.cu:232).
 * Samples must be finite: input with NaN or infinite samples (detected through
 * the column sums) fails with SC_ERR_INVALID_ARGUMENT. */
sc_status sc_dd_search(sc_dd* dd, const void* data, int dtype, int rounded_timesteps,
                       int num_channels, long row_stride, const sc_meta* meta,
                       int beam, int coarse_channel, double max_drift,
                       double min_drift, double snr_threshold, sc_hit* hits,
                       size_t capacity, size_t* num_hits, void* stream);

/* Allocation-free search. No counterpart in the reference: it is for callers
 * that cannot allocate memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). It runs the search of
 * sc_dd_search with device statistics (SC_STATS_DEVICE) and the default
 * kernel shapes, taking the Dedopplerer's constructor arguments (num_timesteps,
 * foff, tsamp, as sc_dd_create) with the search's instead of a handle, and
 * allocates nothing: every intermediate lives in one fixed device array of
 * SC_STATIC_SCRATCH_BYTES per GPU, set up once with the library and held
 * until the process ends. Limits: 4096 <= num_channels <=
 * SC_STATIC_MAX_CHANNELS and at most 256 rounded time steps (larger inputs
 * fail with SC_ERR_INVALID_ARGUMENT or SC_ERR_UNSUPPORTED). Calls that
 * overlap on one GPU race, since they share that GPU's array. Up to `capacity`
 * hits (ascending channel order) are written to `hits`; *num_hits receives the
 * total. Returns when the search is complete. */
#define SC_STATIC_MAX_CHANNELS (1L << 20)
#define SC_STATIC_SCRATCH_BYTES 33620480L
sc_status sc_dd_search_static(int num_timesteps, int rounded_timesteps, int num_channels,
                              double foff_mhz, double tsamp_s, const void* data, int dtype,
                              long row_stride, const sc_meta* meta, int beam,
                              int coarse_channel, double max_drift, double min_drift,
                              double snr_threshold, sc_hit* hits, size_t capacity,
                              size_t* num_hits, void* stream);

/* Copies hits of the last sc_dd_search (no counterpart: the reference appends
 * to a vector). */
sc_status sc_dd_copy_hits(const sc_dd* dd, sc_hit* hits, size_t capacity,
                          size_t* num_hits);

/* Per-channel state of the last search, copied to host arrays of
 * num_channels elements each (any may be NULL): column sums (dedoppler.cu:207),
 * top path sums, their drift blocks and path offsets (dedoppler.cu:32-55),
 * and the median/std the threshold used. No counterpart (the reference keeps
 * them in private pinned arrays); used to check parity. After a
 * device-statistics search at rounded_timesteps <= 16 the path offset is
 * materialised only for each window's candidate (the channels hits can come
 * from); other channels with a positive sum read -1 there. */
sc_status sc_dd_copy_state(const sc_dd* dd, float* column_sums, float* top_sums,
                           int* top_drift_blocks, int* top_path_offsets,
                           float* median, float* std_dev);

/* Replaces Dedopplerer::addIncoherentPower (dedoppler.cu:139-176): sorts hits
 * by drift_steps (driftStepsLessThan, dedoppler_hit.cpp:69-71) and fills each
 * hit's incoherent_power with the path sum of the same path in `data`. */
sc_status sc_dd_add_incoherent_power(sc_dd* dd, const void* data, int dtype,
                                     int rounded_timesteps, int num_channels,
                                     long row_stride, sc_hit* hits, size_t num_hits,
                                     void* stream);

/* Computes the full [rounded_timesteps][num_channels] path sums of one drift
 * block, like optimizedTaylorTree (taylor.cu:333-350), into a device buffer of
 * rounded_timesteps*num_channels floats laid out [path_offset][start_channel].
 * Entries of invalid paths are unspecified (the reference leaves them stale).
 * No separate counterpart in the search; exposed for the reference's
 * differential tree test (taylor_test.cu:8-51). */
sc_status sc_dd_taylor_tree(sc_dd* dd, const void* data, int dtype,
                            int rounded_timesteps, int num_channels, long row_stride,
                            int drift_block, float* out, void* stream);

/* Per-coarse-channel normalisation for the half-width storage forms: writes
 * (x - loc) / std, clamped to +-clamp, as `out_dtype` ([T][N]), where loc is
 * the median of the T-sample column sums over T and std the FP64 standard
 * deviation of all T*N samples, both computed on the GPU. With loc at the
 * median (not the mean, which narrowband RFI can pull far above the noise)
 * the normalised column sums have median ~0, so a positive SNR threshold is a
 * positive path sum and the search's zero start for top path sums
 * (dedoppler.cu:205) never forms candidates. loc_out / std_out (may be NULL)
 * receive loc and std. No counterpart in the reference (it only stores
 * float32). */
sc_status sc_normalize(const float* data, int rounded_timesteps, int num_channels,
                       long row_stride, int out_dtype, float clamp, void* out, double* loc_out,
                       double* std_out, void* stream);

/* Synthetic waterfall generator used by the tests and the benchmark (not part
 * of the reference): exponential(1) noise from a counter-based hash of
 * (seed, t, c), times `gain`, a cosine taper over the outer `taper_frac` of the
 * channels, a DC spike of `dc_spike` times the mean at num_channels/2 (0 = off),
 * plus `num_tones` linear tones. Tone k starts at channel tone_start[k] at t=0,
 * moves tone_slope[k] channels per time step and adds tone_amp[k] to the nearest
 * channel of each of the first `num_timesteps` rows. Rows in
 * [num_timesteps, rounded_timesteps) are zero. Writes float32 [T_r][num_channels]. */
sc_status sc_gen_waterfall(float* out, int num_timesteps, int rounded_timesteps,
                           int num_channels, uint64_t seed, float gain,
                           float taper_frac, float dc_spike, int num_tones,
                           const float* tone_start, const float* tone_slope,
                           const float* tone_amp, void* stream);

/* Thread-local message of the last failing call on this thread. */
const char* sc_last_error_string(void);

/* SC_ABI_VERSION of the built library. */
int sc_abi_version(void);

/* ---- Output writers (dat_file_writer.cpp, hit_file_writer.cpp) ---------- */

/* Header fields the writers take from the file metadata
 * (filterbank_metadata.h:15-62). source_name may be NULL ("" is written). */
typedef struct {
  const char* source_name;
  double fch1, foff, tstart, tsamp, src_raj, src_dej;
  long num_timesteps;        /* unrounded */
  long coarse_channel_size;
  int telescope_id;
} sc_file_meta;

typedef struct sc_writer sc_writer;

/* Replaces DatFileWriter (dat_file_writer.cpp:15-73): turboSETI header at
 * open, one row per hit. */
sc_status sc_dat_writer_open(const char* filename, const sc_file_meta* meta,
                             double max_drift, sc_writer** out);

/* Replaces HitFileWriter (hit_file_writer.cpp:16-97): one unpacked Cap'n Proto
 * message per hit (schema hit.capnp), written to filename + ".tmp" and renamed
 * on close. `beam_source_names`/`beam_ras`/`beam_decs` are the per-beam
 * metadata (num_beams entries, may be 0/NULL). */
sc_status sc_hits_writer_open(const char* filename, const sc_file_meta* meta,
                              int num_beams, const char* const* beam_source_names,
                              const double* beam_ras, const double* beam_decs,
                              sc_writer** out);

/* Replaces HitRecorder::recordHit (hit_recorder.h:10-19). `input` is the host
 * copy of the coarse channel's [T][coarse_channel_size] float32 data (read by
 * the .hits writer for its filterbank slice; may be NULL for .dat). */
sc_status sc_writer_record_hit(sc_writer* w, const sc_hit* hit, const float* input);

/* Replaces EventFileWriter (event_file_writer.cpp:19-69): one Cap'n Proto Event
 * message (hit.capnp:91-93) per cadence event, written to filename + ".tmp" and
 * renamed on close. metas: one per cadence file. */
sc_status sc_events_writer_open(const char* filename, const sc_file_meta* metas, int num_files,
                                sc_writer** out);

/* EventFileWriter::write: hits[i] is file i's hit or NULL (hits[0] must not be
 * NULL); inputs[i] the host [T][coarse_channel_size] float32 data of file i's
 * coarse channel, for the filterbank slices over the event's channel range. */
sc_status sc_events_writer_record(sc_writer* w, const sc_hit* const* hits,
                                  const float* const* inputs);

/* Closes the file (and renames .hits.tmp / .events.tmp). */
sc_status sc_writer_close(sc_writer* w);

/* Replaces findEvents (find_events.cpp:66-227): searches the .fil files of an
 * ABACAD cadence coarse channel by coarse channel (host-exact statistics) and
 * writes every event (a hit in each "on" file where the first file's drift
 * predicts it, none in the "off" files' range) to `output`. Metadata mismatches
 * are SC_ERR_INVALID_ARGUMENT where the reference exits. */
sc_status sc_find_events(const char* const* filenames, int num_files, const char* output,
                         double max_drift, double snr_on, double snr_off, long* num_events,
                         void* stream);

/* ---- Learned block proposal ------------------------------------------------
 * No counterpart in the reference. A small BF16 convolutional network scores
 * 4096-channel blocks of a 16-row coarse channel from a pooled tile
 * (16 x 1024, 4 fine channels per column, normalised by the coarse channel's
 * column-sum median and deviation); the gated search runs the full tree only
 * on the channel tiles of blocks whose logit exceeds a threshold. The plain
 * search is its fallback and the judge of its recall. */

#define SC_PN_BLOCK_CHANNELS 4096
#define SC_PN_TILE_ROWS 16
#define SC_PN_TILE_COLUMNS 1024

typedef struct sc_pn sc_pn;

/* Number of FP32 parameters (23393): conv1 1->16, conv2 16->32, conv3 32->32
 * (3x5 kernels), 1x1 head 32->1; the layout is documented in pn_internal.h. */
int sc_pn_num_params(void);

/* Allocates weights, optimizer state and activation buffers for up to
 * max_batch blocks per call on the current device. Parameters start at 0. */
sc_status sc_pn_create(int max_batch, sc_pn** out);
sc_status sc_pn_destroy(sc_pn* pn);
sc_status sc_pn_memory_usage(const sc_pn* pn, size_t* bytes);

/* He-uniform weights and zero biases from a seed (deterministic). */
sc_status sc_pn_init(sc_pn* pn, uint64_t seed);
/* Host FP32 parameters in/out (n == sc_pn_num_params()); set resets Adam. */
sc_status sc_pn_set_params(sc_pn* pn, const float* params, size_t n);
sc_status sc_pn_get_params(const sc_pn* pn, float* params, size_t n);

/* Pooled tiles (device BF16 [N/4096][16][1024]) of a device [16][N] coarse
 * channel; stats = device {median, std} of its column sums. */
sc_status sc_pn_pool(const void* data, int dtype, int rounded_timesteps, int num_channels,
                     long row_stride, const float* stats, void* tiles, void* stream);

/* Block logits (device float[num_blocks]) of device tiles. */
sc_status sc_pn_forward(sc_pn* pn, const void* tiles, int num_blocks, float* logits,
                        void* stream);

/* Gradient (device float[sc_pn_num_params()], or the handle's own buffer if
 * NULL) of the mean weighted binary cross-entropy of the block logits against
 * device labels (0/1 floats; positives weighted by pos_weight), and the loss
 * (device float, may be NULL). */
sc_status sc_pn_gradient(sc_pn* pn, const void* tiles, const float* labels, int num_blocks,
                         float pos_weight, float* grad, float* loss, void* stream);

/* sc_pn_gradient followed by one Adam step (beta 0.9/0.999) at rate lr. */
sc_status sc_pn_train_step(sc_pn* pn, const void* tiles, const float* labels, int num_blocks,
                           float lr, float pos_weight, float* loss, void* stream);

/* Device pointers of the last forward's activations (BF16 NHWC h1 [B][16][512][16],
 * h2 [B][8][256][32], h3 [B][4][128][32]) and of the last gradient call's logits
 * and argmax positions; any output may be NULL. For tests. */
sc_status sc_pn_activations(const sc_pn* pn, const void** h1, const void** h2, const void** h3,
                            const float** logits, const int** argmax);

/* Gated form of sc_dd_search (rounded_timesteps 16, num_channels a multiple of
 * 4096): device statistics, then the network, then the full tree only on the
 * channel tiles that overlap a block with logit > logit_threshold. Hits are
 * those the plain device-statistics search reports whose start channel lies in
 * a searched tile (unsearched channels neither form nor suppress candidates).
 * flagged_blocks (may be NULL) receives the number of flagged blocks. */
sc_status sc_dd_search_proposed(sc_dd* dd, sc_pn* pn, const void* data, int dtype,
                                int rounded_timesteps, int num_channels, long row_stride,
                                const sc_meta* meta, int beam, int coarse_channel,
                                double max_drift, double min_drift, double snr_threshold,
                                float logit_threshold, sc_hit* hits, size_t capacity,
                                size_t* num_hits, int* flagged_blocks, void* stream);

/* ---- Batched search --------------------------------------------------------
 * No counterpart in the reference, which searches one coarse channel per call
 * (run_dedoppler.cpp:35-51, beamforming_pipeline.cpp:145-253). A batch of
 * coarse channels / beams with uneven time-step counts and channel widths is
 * planned once (search handles per distinct shape, each instance's launch
 * sequence captured into one hipGraph; runs of up to 8 consecutive instances
 * of one 16-step shape share one launch of the tree kernel) and then run as a
 * single graph launch. Statistics are the device statistics (SC_STATS_DEVICE). */

typedef struct {
  const void* data;          /* device [rounded_timesteps][row_stride] samples */
  int dtype;                 /* sc_dtype */
  int num_timesteps;         /* unrounded */
  int num_channels;          /* >= 4096 */
  long row_stride;
  double foff;               /* MHz */
  double tsamp;              /* s */
  sc_meta meta;
  int beam;
  int coarse_channel;
  double max_drift, min_drift, snr_threshold;
} sc_instance;

/* Kernel shape for one (rounded_timesteps, num_channels): the one-pass fields
 * apply when rounded_timesteps <= 16, or at 32 when pass_group is 0; the
 * two-pass fields otherwise, for 32..256 (see sc_dd_set_tuning /
 * sc_dd_set_two_pass). Produced by the autotuner. */
typedef struct {
  int rounded_timesteps;
  int num_channels;
  int channels_per_lane, drift_group, prefetch;
  int pass_group, pass_blocks;
} sc_tuning;

typedef struct sc_batch sc_batch;

/* Plans the batch: instance data pointers are captured, so the same buffers
 * must hold each run's input. tuning may be NULL (default shapes). */
sc_status sc_batch_create(const sc_instance* instances, int count, const sc_tuning* tuning,
                          int num_tuning, sc_batch** out);
/* Runs every instance (one graph launch on `stream`), waits, and extracts the hits. */
sc_status sc_batch_run(sc_batch* b, void* stream);
/* Hits of one instance from the last run (same contract as sc_dd_copy_hits). */
sc_status sc_batch_copy_hits(const sc_batch* b, int instance, sc_hit* hits, size_t capacity,
                             size_t* num_hits);
/* Gated batch: every instance searched as sc_dd_search_proposed does (16 time
 * steps, a multiple of 4096 channels, at most pn's max_batch blocks each).
 * Each run is direct launches on three streams (not a graph): the trees on
 * one, the column sums, statistics, network and tile lists on another, the
 * merges on the third. Runs of up to 8 consecutive instances of one shape
 * share one launch per step (their network runs over all their blocks at
 * once, in buffers the batch allocates: 448 KiB per block of the longest
 * run) and one tree launch; a run's front runs beside the previous run's
 * tree. pn must outlive the batch; its weights are read at every run.
 * sc_batch_info reports 0 graph nodes for it. */
sc_status sc_batch_create_proposed(const sc_instance* instances, int count,
                                   const sc_tuning* tuning, int num_tuning, sc_pn* pn,
                                   float logit_threshold, sc_batch** out);
/* Gated batch, last run: channel tiles searched and blocks flagged for one instance. */
sc_status sc_batch_flagged_blocks(const sc_batch* b, int instance, int* tiles_searched,
                                  int* flagged_blocks);
/* Kernel timing for measurement (plain batches; no counterpart in the
 * reference): while enabled, sc_batch_run launches the batch's sequence
 * directly instead of replaying its graph, with HIP events around each
 * path-sum launch, and accumulates their elapsed time; enabling or disabling
 * resets the totals. Runs of consecutive instances of one 16-step shape share
 * a launch (up to 8 coarse channels), so a launch may cover several. */
sc_status sc_batch_set_timing(sc_batch* b, int enable);
/* Totals since sc_batch_set_timing: kernel milliseconds, path-sum launches,
 * coarse channels those launches covered (any output may be NULL). */
sc_status sc_batch_kernel_time(const sc_batch* b, double* total_ms, long* launches,
                               long* channels);
/* Number of distinct shapes and of graph nodes. */
sc_status sc_batch_info(const sc_batch* b, int* num_shapes, int* graph_nodes);
sc_status sc_batch_destroy(sc_batch* b);

/* ---- Input: sigproc filterbank files and loader conditioning ---------------
 * Header parsing as FilReader (fil_reader.cpp:31-131), metadata inference as
 * FilterbankMetadata::inferMetadata (filterbank_metadata.cpp:53-93). Coarse
 * channel loading (unimplemented in the reference's FilReader,
 * fil_reader.cpp:151-154) applies the HDF5 loader's conditioning
 * (h5_reader.cpp:195-209): zero rows up to rounded_timesteps and, when the
 * metadata has a DC spike, the centre channel of each row replaced by the mean
 * of its neighbours. Errors are status codes where the reference exits. */

typedef struct {
  char source_name[256];
  double fch1, foff, tstart, tsamp, src_raj, src_dej; /* RA hours, Dec degrees */
  int telescope_id;                                    /* -1 if absent */
  long num_timesteps, num_channels, coarse_channel_size, num_coarse_channels;
  int has_dc_spike;
} sc_fil_meta;

typedef struct sc_fil sc_fil;

sc_status sc_fil_open(const char* filename, sc_fil** out);
sc_status sc_fil_metadata(const sc_fil* f, sc_fil_meta* out);
/* Overrides the inferred coarse channel size (the reference's nfpc header). */
sc_status sc_fil_set_coarse_channel_size(sc_fil* f, long coarse_channel_size);
/* Coarse channel i as host float32 [rounded_timesteps][coarse_channel_size]. */
sc_status sc_fil_read_coarse_channel(const sc_fil* f, int i, int rounded_timesteps,
                                     float* host);
/* The same into device memory (returns when the upload is complete). */
sc_status sc_fil_load_coarse_channel(const sc_fil* f, int i, int rounded_timesteps,
                                     float* device, void* stream);
sc_status sc_fil_close(sc_fil* f);

/* The loader conditioning on a coarse channel already in device memory
 * (float32 [rounded_timesteps][row_stride]): rows num_timesteps.. are zeroed,
 * and with has_dc_spike the centre channel (num_channels/2) of each of the
 * first num_timesteps rows becomes the mean of its neighbours
 * (h5_reader.cpp:195-209). */
sc_status sc_condition_coarse_channel(float* data, int num_timesteps, int rounded_timesteps,
                                      int num_channels, long row_stride, int has_dc_spike,
                                      void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SETICORE_AMD_H */
