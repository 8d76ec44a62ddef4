/* ssim2hip -- ssimulacra2 on AMD Instinct MI355X (gfx950), C ABI.
 *
 * Drop-in boundary for the ssimulacra2 engine of imazen/turbo-metrics
 * (crates/ssimulacra2-cuda). Each entry point below names the Rust method it
 * replaces as `file:line`; paths are relative to the turbo-metrics checkout.
 * Images are device buffers described by plain pointers and sizes; streams are
 * hipStream_t passed as an opaque pointer (NULL = the default stream). Every
 * call returns an ssim2_status; nothing aborts the process.
 *
 * Score semantics follow the reference's CPU ssimulacra2
 * (crates/ssimulacra2-cuda/examples/cpu.rs): the same f32 arithmetic, the same
 * scale count (the first scale smaller than 8 px is still scored) and the same
 * 108-weight table. Images smaller than 8x8 are rejected with
 * SSIM2_INVALID_SIZE (the documented error of cpu.rs:335-341).
 */
#ifndef SSIM2HIP_H
#define SSIM2HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SSIM2_OK = 0,
  SSIM2_INVALID_SIZE = 1,      /* width or height < 8, or beyond the supported maximum */
  SSIM2_DIM_MISMATCH = 2,      /* image size differs from the context's */
  SSIM2_NO_REFERENCE = 3,      /* cached-reference call without set_reference */
  SSIM2_OOM = 4,               /* device allocation failed or pre-flight budget exceeded */
  SSIM2_BATCH_SIZE = 5,        /* batch larger than the context was created for */
  SSIM2_ADMISSION_REFUSED = 6, /* serving: request does not fit the HBM budget */
  SSIM2_HIP_ERROR = 7,         /* a HIP runtime call failed (see ssim2_last_error) */
  SSIM2_INVALID_ARGUMENT = 8,  /* null pointer, bad pitch, bad option value */
  SSIM2_NOT_READY = 9          /* serving: no finished result to return yet */
} ssim2_status;

/* Opaque hipStream_t. NULL selects the default (null) stream. */
typedef void* ssim2_stream;

/* Packed RGB device image: 3 samples per pixel, u8 or f32, rows `pitch_bytes`
 * apart. Mirrors cudarse-npp Image<S, C<3>> (crates/cudarse/cudarse-npp/src/
 * image/mod.rs:138-147). */
typedef struct {
  void* ptr;
  size_t pitch_bytes;
  uint32_t width;
  uint32_t height;
} ssim2_image;

typedef struct ssim2_ctx ssim2_ctx;

/* YUV 4:2:0 frame in device memory, BT.709 limited range (the video input of
 * crates/turbo-metrics/src/color.rs:96-116). SSIM2_YUV420P10: three planes of
 * u16 10-bit samples (LSB-aligned), c0 = Cb, c1 = Cr, chroma planes
 * ceil(w/2) x ceil(h/2). SSIM2_P016: two planes of u16 samples in the high
 * bits (the NVDEC P016 layout the reference consumes), c0 = interleaved CbCr,
 * c1 unused. */
typedef enum { SSIM2_YUV420P10 = 0, SSIM2_P016 = 1 } ssim2_yuv_format;

typedef struct {
  const void* y;
  const void* c0;
  const void* c1;
  size_t y_pitch_bytes;
  size_t c_pitch_bytes;
  uint32_t width;
  uint32_t height;
  uint32_t format; /* ssim2_yuv_format */
  uint32_t reserved0;
} ssim2_yuv_image;

/* Context options. Zero-initialise, then set struct_size = sizeof(ssim2_options). */
typedef struct {
  uint32_t struct_size;
  uint32_t reserved0;
  /* Refuse creation when the estimated device footprint exceeds this many
   * bytes (0 = use the device's free memory, as the reference's pre-flight
   * check does, crates/ssimulacra2-cuda/src/lib.rs:88-112). */
  uint64_t memory_budget;
} ssim2_options;

/* ---- lifecycle ---------------------------------------------------------- */

/* Ssimulacra2::new (crates/ssimulacra2-cuda/src/lib.rs:80-197). Unlike the
 * reference, the context is not bound to particular input buffers: every
 * compute call names its images, so buffers may change between calls. */
ssim2_status ssim2_create(uint32_t width, uint32_t height, const ssim2_options* options,
                          ssim2_ctx** out);

/* Drop for Ssimulacra2 (lib.rs:1195-1202): waits for the context's work,
 * frees everything, never aborts. NULL is accepted. */
ssim2_status ssim2_destroy(ssim2_ctx* ctx);

/* Ssimulacra2::mem_usage (lib.rs:206-235): device bytes owned by the context. */
ssim2_status ssim2_mem_usage(const ssim2_ctx* ctx, size_t* bytes);

/* ---- scoring ------------------------------------------------------------ */

/* Ssimulacra2::compute (lib.rs:419-424): enqueue the whole pipeline on
 * `stream` for a pair of packed linear-RGB f32 images; returns at once.
 * Fetch the score with ssim2_get_score. */
ssim2_status ssim2_compute_linear_async(ssim2_ctx* ctx, const ssim2_image* ref_linear,
                                        const ssim2_image* dis_linear, ssim2_stream stream);

/* Ssimulacra2::get_score (lib.rs:426-429): waits for the last enqueued
 * computation of this context and returns its score. Idempotent (the
 * reference's version post-processes in place and must not be called twice). */
ssim2_status ssim2_get_score(ssim2_ctx* ctx, double* score);

/* Ssimulacra2::compute_sync (lib.rs:409-416). */
ssim2_status ssim2_compute_linear_sync(ssim2_ctx* ctx, const ssim2_image* ref_linear,
                                       const ssim2_image* dis_linear, ssim2_stream stream,
                                       double* score);

/* Ssimulacra2::compute_srgb_sync (lib.rs:391-404): sRGB u8 device images in.
 * When ref_linear / dis_linear are non-NULL they receive the linear f32 copy
 * (as the reference's srgb_to_linear step writes them); NULL skips that write. */
ssim2_status ssim2_compute_srgb_sync(ssim2_ctx* ctx, const ssim2_image* ref_srgb,
                                     const ssim2_image* dis_srgb, ssim2_image* ref_linear,
                                     ssim2_image* dis_linear, ssim2_stream stream, double* score);

/* Asynchronous form of ssim2_compute_srgb_sync without the linear copies:
 * u8 in, score fetched with ssim2_get_score. The reference has no separate
 * method; compute_srgb_sync (lib.rs:391-404) without its final sync is this. */
ssim2_status ssim2_compute_srgb_async(ssim2_ctx* ctx, const ssim2_image* ref_srgb,
                                      const ssim2_image* dis_srgb, ssim2_stream stream);

/* Doubles in a result record: [0] score, [1..108] the per-scale sums
 * (scale*18 + channel*6 + {ssim S1, ssim S4, artifact S1, artifact S4,
 * detail-loss S1, detail-loss S4}), [109..120] the scale widths and heights. */
#define SSIM2_RESULT_DOUBLES 121

/* No counterpart in the reference. ssim2_compute_srgb_async for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): no context, no allocation, no host
 * copy. The same kernels run on `stream` and the score equals the context
 * path's to the bit. `result` is a device buffer of SSIM2_RESULT_DOUBLES
 * doubles that receives the record when the stream reaches it.
 * The working memory is one device array per GPU, held until the process
 * ends; it fits pairs of up to about 2.2 Mpx (a 1920x1080 pair needs
 * 244 MB of the 256 MiB): larger pairs are refused with SSIM2_INVALID_SIZE.
 * Calls that overlap on one GPU share that array and race: the caller orders
 * them (one stream, or events). */
ssim2_status ssim2_compute_srgb_noalloc_async(const ssim2_image* ref_srgb, const ssim2_image* dis_srgb,
                                              double* result, ssim2_stream stream);

/* Ssimulacra2::compute_from_cpu_srgb_sync (lib.rs:329-345): packed sRGB bytes
 * in host memory (`row_bytes` apart, at least 3*width), uploaded into the
 * context's own staging images. */
ssim2_status ssim2_compute_from_host_srgb(ssim2_ctx* ctx, const uint8_t* ref, const uint8_t* dis,
                                          size_t row_bytes, ssim2_stream stream, double* score);

/* Video frames: YUV 4:2:0 converted to linear RGB inside the first kernel
 * (fused ingest), then the same exact pipeline. The reference converts with a
 * separate kernel into the buffers its graph is bound to
 * (crates/turbo-metrics/src/lib.rs:271-363, color.rs:96-116). */
ssim2_status ssim2_compute_yuv420_async(ssim2_ctx* ctx, const ssim2_yuv_image* ref,
                                        const ssim2_yuv_image* dis, ssim2_stream stream);
ssim2_status ssim2_compute_yuv420_sync(ssim2_ctx* ctx, const ssim2_yuv_image* ref,
                                       const ssim2_yuv_image* dis, ssim2_stream stream, double* score);

/* ---- cached reference ---------------------------------------------------- */

/* Ssimulacra2::set_reference_linear (lib.rs:440-541): compute and keep the
 * reference-side state (XYB pyramid, blurred ref^2 and ref) on the device. */
ssim2_status ssim2_set_reference_linear(ssim2_ctx* ctx, const ssim2_image* ref_linear,
                                        ssim2_stream stream);

/* sRGB u8 form of ssim2_set_reference_linear, covering the reference's
 * prepare_reference_from_cpu_srgb + set_reference_linear sequence
 * (lib.rs:373-388, 440-541) for device-resident bytes. */
ssim2_status ssim2_set_reference_srgb(ssim2_ctx* ctx, const ssim2_image* ref_srgb,
                                      ssim2_stream stream);

/* Ssimulacra2::compute_with_reference_linear (lib.rs:543-590): score a linear
 * distorted image against the cached reference (SSIM2_NO_REFERENCE without
 * one; the reference panics). `ref_linear` is accepted for signature parity
 * and may be NULL: the cached state is used. */
ssim2_status ssim2_compute_with_reference_linear(ssim2_ctx* ctx, const ssim2_image* ref_linear,
                                                 const ssim2_image* dis_linear,
                                                 ssim2_stream stream, double* score);

/* sRGB u8 form (the reference's compute_from_cpu_dis_only_srgb_sync,
 * lib.rs:355-371, for device-resident bytes). */
ssim2_status ssim2_compute_with_reference_srgb(ssim2_ctx* ctx, const ssim2_image* dis_srgb,
                                               ssim2_stream stream, double* score);

/* Host-byte forms of the cached-reference flow:
 * Ssimulacra2::prepare_reference_from_cpu_srgb (lib.rs:373-388) uploads the
 * reference bytes into the context's staging image and caches its state;
 * Ssimulacra2::compute_from_cpu_dis_only_srgb_sync (lib.rs:347-371) uploads
 * only the distorted bytes and scores them against the cached reference
 * (SSIM2_NO_REFERENCE without one). */
ssim2_status ssim2_prepare_reference_from_host_srgb(ssim2_ctx* ctx, const uint8_t* ref, size_t row_bytes,
                                                    ssim2_stream stream);
ssim2_status ssim2_compute_from_host_dis_srgb(ssim2_ctx* ctx, const uint8_t* dis, size_t row_bytes,
                                              ssim2_stream stream, double* score);

/* Ssimulacra2::clear_reference / has_reference (lib.rs:592-602). */
ssim2_status ssim2_clear_reference(ssim2_ctx* ctx);
ssim2_status ssim2_has_reference(const ssim2_ctx* ctx, int* has_reference);

/* ---- batched fast form ---------------------------------------------------- */
/* No counterpart in the reference's ssimulacra2-cuda (its closest relative is
 * the same-size Ssim2Batch of crates/ssim2-gpu/src/pipeline_batch.rs:120-612).
 * Many sRGB pairs of unrelated sizes are scored by three kernels captured
 * once in a hipGraph (the grids that follow a batch's size are set per
 * batch). Arithmetic: exact 9-tap FIR blur (the closed form of the
 * reference's recursive Gaussian), statistics in shifted coordinates, f32
 * maps, f64 sums; the XYB planes of every scale stored in `storage`. Its
 * scores follow the FP64 evaluation of the metric closely (|dscore| about
 * 3e-4 with f32 and 0.015 with f16 storage on natural-image-like content)
 * and are not bit-identical to the exact path. */

typedef enum {
  SSIM2_STORAGE_F32 = 0,
  SSIM2_STORAGE_F16 = 1,
  /* Refused with SSIM2_INVALID_ARGUMENT (batch, server and video options):
   * bfloat16 planes put 4K video frames 0.30-0.35 from the FP64 score,
   * seven times the 0.05 bar that f16 storage meets. */
  SSIM2_STORAGE_BF16 = 2
} ssim2_storage;

typedef struct {
  uint32_t struct_size;  /* sizeof(ssim2_batch_options) */
  uint32_t storage;      /* ssim2_storage */
  uint32_t max_pairs;    /* pairs per batch */
  uint32_t reserved0;
  uint64_t max_pixels;   /* sum of width*height over a batch */
  uint64_t memory_budget;/* refuse creation above this many device bytes (0 = free memory) */
} ssim2_batch_options;

/* One pair of packed sRGB u8 device images of equal size. */
typedef struct {
  ssim2_image ref;
  ssim2_image dis;
} ssim2_pair;

typedef struct ssim2_batch ssim2_batch;

ssim2_status ssim2_batch_create(const ssim2_batch_options* options, ssim2_batch** out);
ssim2_status ssim2_batch_destroy(ssim2_batch* batch);
ssim2_status ssim2_batch_mem_usage(const ssim2_batch* batch, size_t* bytes);
/* Enqueue one batch on `stream`. The host stages it while the previous batch
 * runs (two staging slots: the call blocks only until the batch before the
 * previous one has finished); on the device it starts after the previous
 * batch, on any stream. Scores and sums refer to the last batch enqueued. */
ssim2_status ssim2_batch_compute_async(ssim2_batch* batch, const ssim2_pair* pairs, uint32_t npairs,
                                       ssim2_stream stream);
/* Resident reference pyramids for the batched form: the batched counterpart
 * of the reference's cached reference (crates/ssimulacra2-cuda/src/lib.rs:440-860,
 * set_reference / compute_with_reference), for callers that score many
 * distorted images against one reference (an encoder's rate-distortion loop).
 * ssim2_batch_add_reference computes the reference's pyramid once (it
 * allocates, and waits on `stream`) and returns an id; pairs naming the id
 * then skip the reference side of the pyramid. A pair scores the same as the
 * pair with the reference image given directly. Removing a reference waits
 * for the batches in flight. */
typedef struct {
  uint32_t ref_id;   /* from ssim2_batch_add_reference */
  uint32_t reserved; /* 0 */
  ssim2_image dis;   /* packed sRGB u8, the reference's size */
} ssim2_ref_pair;
ssim2_status ssim2_batch_add_reference(ssim2_batch* batch, const ssim2_image* ref, ssim2_stream stream,
                                       uint32_t* ref_id);
ssim2_status ssim2_batch_remove_reference(ssim2_batch* batch, uint32_t ref_id);
ssim2_status ssim2_batch_compute_ref_async(ssim2_batch* batch, const ssim2_ref_pair* pairs, uint32_t npairs,
                                           ssim2_stream stream);
/* As ssim2_batch_compute_async for YUV 4:2:0 frame pairs (fused ingest). */
typedef struct {
  ssim2_yuv_image ref;
  ssim2_yuv_image dis;
} ssim2_yuv_pair;
ssim2_status ssim2_batch_compute_yuv_async(ssim2_batch* batch, const ssim2_yuv_pair* pairs, uint32_t npairs,
                                           ssim2_stream stream);
/* Scores of the last batch, in submission order (waits for it). */
ssim2_status ssim2_batch_get_scores(ssim2_batch* batch, double* scores, uint32_t capacity);
ssim2_status ssim2_batch_compute_sync(ssim2_batch* batch, const ssim2_pair* pairs, uint32_t npairs,
                                      ssim2_stream stream, double* scores);
/* done = 1 once the last batch has finished (non-blocking). */
ssim2_status ssim2_batch_query(ssim2_batch* batch, int* done);
/* Per-scale sums of pair `index` of the last batch (layout of ssim2_get_sums). */
ssim2_status ssim2_batch_get_sums(ssim2_batch* batch, uint32_t index, double* sums108, int* dims12);
/* Output tile shape (0: 32x16, 1: 64x16, 2: 32x32, 3: 32x24, 4: 64x32) and
 * grid size in multiples of the CU count (0: one block per tile, sized per
 * batch); a new engine starts at shape 4, grid 32 with binary16 storage and
 * shape 0, grid 64 with f32. ssim2_batch_autotune times every combination on
 * the given pairs and keeps the fastest. With binary16 storage, shape 4 runs the blur on the
 * matrix cores (binary16 operands split into hi + lo parts, f32 sums; scores
 * within about 1e-3 of the other shapes); its scaled squares stay finite for
 * shifted XYB values up to about 11, and the sRGB and YUV inputs of the batch
 * give at most 0.45. `report` (optional) receives a table. */
ssim2_status ssim2_batch_set_tiling(ssim2_batch* batch, uint32_t tile_cfg, uint32_t grid_mult);
ssim2_status ssim2_batch_get_tiling(const ssim2_batch* batch, uint32_t* tile_cfg, uint32_t* grid_mult);
ssim2_status ssim2_batch_autotune(ssim2_batch* batch, const ssim2_pair* pairs, uint32_t npairs,
                                  ssim2_stream stream, uint32_t reps, char* report, size_t report_size);

/* ---- serving --------------------------------------------------------------- */
/* No counterpart in the reference (whose callers score one frame pair at a
 * time, crates/turbo-metrics/src/lib.rs:271-363). A scheduler thread owns
 * `slots` batch engines on separate streams; requests submitted from any
 * thread wait in FIFO order and are admitted into the next idle slot while
 * their device footprint fits `hbm_budget` and the slot's capacity, so
 * batches launched while others run overlap on the device. Every result
 * carries host timestamps (seconds, steady clock) for latency accounting. A
 * request's score does not depend on which batch it joined. */

typedef struct {
  uint32_t struct_size;           /* sizeof(ssim2_server_options) */
  uint32_t storage;               /* ssim2_storage of the batch engines */
  uint32_t slots;                 /* batches in flight at most, 1..16 */
  uint32_t max_pairs_per_batch;
  uint64_t max_pixels_per_batch;  /* also the largest request accepted */
  uint64_t hbm_budget;            /* admission limit on in-flight device bytes */
  uint32_t tile_cfg;              /* as ssim2_batch_set_tiling; 0/0 = the batch engine's
                                     default shape (4 for binary16, 0 for f32), grid 8 x CUs */
  uint32_t grid_mult;
} ssim2_server_options;

typedef struct {
  uint64_t request_id;
  double score;
  double submit_time_s;
  double done_time_s;
  ssim2_status status;
  uint32_t reserved0;
} ssim2_result;

typedef struct {
  uint32_t struct_size;
  uint32_t reserved0;
  uint64_t submitted, completed, refused, batches, pixels_completed, waiting, in_flight_bytes;
  double first_submit_s, last_done_s;
} ssim2_server_stats_t;

typedef struct ssim2_server ssim2_server;

ssim2_status ssim2_server_create(const ssim2_server_options* options, ssim2_server** out);
/* Stops the scheduler after the batches in flight and frees everything. */
ssim2_status ssim2_server_destroy(ssim2_server* server);
/* Thread-safe. SSIM2_ADMISSION_REFUSED when the request can never fit the
 * server's limits; the images must stay valid until its result is polled. */
ssim2_status ssim2_server_submit(ssim2_server* server, const ssim2_pair* pair, uint64_t request_id);
/* Resident references of the server (the batched set_reference, see
 * ssim2_batch_add_reference): the pyramid is computed once and kept in HBM
 * for every slot; requests naming the id batch with any others and score as
 * the pair with the reference image would. Registration and removal are
 * thread-safe and briefly hold the scheduler; removal waits for the batches
 * in flight, and queued requests naming a removed id fail with
 * SSIM2_NO_REFERENCE. */
ssim2_status ssim2_server_add_reference(ssim2_server* server, const ssim2_image* ref, uint32_t* ref_id);
ssim2_status ssim2_server_remove_reference(ssim2_server* server, uint32_t ref_id);
ssim2_status ssim2_server_submit_ref(ssim2_server* server, uint32_t ref_id, const ssim2_image* dis,
                                     uint64_t request_id);
/* Move up to `capacity` finished results out; *n receives the count. */
ssim2_status ssim2_server_poll(ssim2_server* server, ssim2_result* out, uint32_t capacity, uint32_t* n);
/* Wait until every accepted request has finished (timeout_s <= 0: no limit). */
ssim2_status ssim2_server_drain(ssim2_server* server, double timeout_s);
ssim2_status ssim2_server_stats(ssim2_server* server, ssim2_server_stats_t* stats);

/* ---- video driver and statistics ------------------------------------------ */

/* quick_stats::full::Stats (crates/quick-stats/src/lib.rs:1-52): population and
 * sample variance, percentiles by linear interpolation between ranks. */
typedef struct {
  double min, max, mean;
  double var, sample_var, stddev, sample_stddev;
  double p1, p5, p50, p95, p99;
} ssim2_stats;

/* Stats::compute (quick-stats lib.rs:21-52) over n >= 1 values. */
ssim2_status ssim2_stats_compute(const double* values, size_t n, ssim2_stats* out);

/* One decoded frame in device memory (HwFrame, crates/turbo-metrics/src/lib.rs:128-133):
 * packed sRGB u8 or YUV 4:2:0. */
typedef enum { SSIM2_FRAME_SRGB = 0, SSIM2_FRAME_YUV420 = 1 } ssim2_frame_kind;
typedef struct {
  uint32_t kind; /* ssim2_frame_kind */
  uint32_t reserved0;
  ssim2_image srgb;
  ssim2_yuv_image yuv;
} ssim2_frame;

/* FrameSource::next_frame (lib.rs:151-160): fill *frame with the next frame and
 * return 1, return 0 at the end of the stream, or a negative value on error.
 * The frame only has to stay valid until the next call on the same source. */
typedef int (*ssim2_frame_source)(void* user, ssim2_frame* frame);

typedef enum { SSIM2_VIDEO_EXACT = 0, SSIM2_VIDEO_FAST = 1 } ssim2_video_mode;

/* Options (turbo_metrics::Options, lib.rs:44-59) plus the scoring form. */
typedef struct {
  uint32_t struct_size; /* sizeof(ssim2_video_options) */
  uint32_t every;       /* score one decoded pair in `every` (0 or 1: all) */
  uint32_t skip;        /* leading frames dropped from both sources */
  uint32_t skip_ref;    /* ... and additionally from the reference source */
  uint32_t skip_dis;    /* ... and additionally from the distorted source */
  uint32_t frames;      /* stop after this many decoded pairs (0: all) */
  uint32_t mode;        /* ssim2_video_mode: exact path pair by pair, or the batched fast form */
  uint32_t storage;     /* fast form: ssim2_storage */
  uint32_t batch;       /* fast form: pairs per launch (0: 8) */
  uint32_t reserved0;
} ssim2_video_options;

/* TurboMetrics::compute_all for ssimulacra2 (lib.rs:367-442): pulls frame pairs
 * from the two sources, applies the sampling options, scores the kept pairs and
 * returns up to `capacity` per-frame scores, their count and their statistics
 * (MetricAggregate, lib.rs:61-77). */
ssim2_status ssim2_video_compute_all(uint32_t width, uint32_t height, const ssim2_video_options* options,
                                     ssim2_frame_source ref_source, void* ref_user, ssim2_frame_source dis_source,
                                     void* dis_user, ssim2_stream stream, double* scores, size_t capacity,
                                     size_t* frame_count, ssim2_stats* stats);

/* ---- diagnostics --------------------------------------------------------- */

/* The raw per-scale sums behind the last score: sums[scale*18 + c*6 + j],
 * j = {ssim sum d, ssim sum d^4, artifact sum, artifact sum^4, detail sum,
 * detail sum^4}; dims[2*scale] = width, dims[2*scale+1] = height. Waits for
 * the last computation. Any pointer may be NULL. */
ssim2_status ssim2_get_sums(ssim2_ctx* ctx, double* sums108, int* dims12, int* nscales);

/* Copies one intermediate plane of the last computation to the host (waits
 * for it): kind 0 = XYB plane `index` (0..2 reference X, Y, B; 3..5 the
 * distorted image's), kind 1 = row-pass plane `index` = channel * 5 +
 * quantity (reference^2, distorted^2, reference*distorted, reference,
 * distorted), at scale `scale`; rows of the scale's width, `host_row_floats`
 * apart. No counterpart in the reference; for stage-by-stage parity checks. */
ssim2_status ssim2_get_plane(ssim2_ctx* ctx, uint32_t kind, uint32_t scale, uint32_t index, float* host,
                             size_t host_row_floats);

/* Per-kernel timing of the sRGB u8 pipeline: runs it `reps` times with HIP
 * events around each of its four kernels on `stream` and writes the mean
 * milliseconds per launch to ms[4] = {pyramid, row pass, column pass + maps,
 * finalize}. No counterpart in the reference (its profiling hooks are
 * cuProfilerStart/Stop, crates/cudarse/cudarse-driver/src/lib.rs:56-62). */
ssim2_status ssim2_time_kernels(ssim2_ctx* ctx, const ssim2_image* ref_srgb,
                                const ssim2_image* dis_srgb, ssim2_stream stream, int reps,
                                float* ms);

/* Diagnostic, no counterpart in the reference. The kernels replace three IEEE
 * operations by shorter sequences that must give the same results: the cube
 * root's double division (glibc's cbrtf method), the SSIM map's f32 quotient
 * and the edge map's f64 quotient (reciprocal, Newton steps, remainder
 * corrections). This evaluates both forms on the device for every float x
 * whose bit pattern lies in [lo_bits, hi_bits] and returns how many results
 * differ. which 0: cbrt(x). which 1: 1 - a / x in the SSIM map's arithmetic,
 * for 64 numerators a. which 2: the edge ratio with x as one absolute
 * difference and 64 values as the other. Synchronous. */
ssim2_status ssim2_check_arithmetic(uint32_t which, uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches);

/* ---- utilities ------------------------------------------------------------ */

/* Raw device memory (cudarse-driver async alloc, mem.rs:27-35) and a plain
 * streaming copy kernel of `bytes` (a multiple of 16), the known-size kernel
 * that byte counters are calibrated against. */
ssim2_status ssim2_device_alloc(size_t bytes, void** out);
ssim2_status ssim2_device_free(void* ptr);
ssim2_status ssim2_device_copy(void* dst, const void* src, size_t bytes, ssim2_stream stream);
ssim2_status ssim2_device_synchronize(void);

/* Kernel::srgb_to_linear (crates/ssimulacra2-cuda/src/kernel.rs:36-57):
 * packed u8 sRGB -> packed f32 linear through the reference's 256-entry table. */
ssim2_status ssim2_srgb_to_linear(const ssim2_image* src_u8, ssim2_image* dst_f32,
                                  ssim2_stream stream);

/* Kernel::downscale_by_2 (crates/ssimulacra2-cuda/src/kernel.rs:59-83; kernel
 * crates/ssimulacra2-cuda-kernel/src/downscale.rs:4-35): packed f32 RGB, 2x2
 * mean with the source clamped at its last row and column. dst must be
 * ceil(width/2) x ceil(height/2) of src. */
ssim2_status ssim2_downscale_by_2(const ssim2_image* src_f32, ssim2_image* dst_f32, ssim2_stream stream);

/* Kernel::linear_to_xyb (crates/ssimulacra2-cuda/src/kernel.rs:161-184; kernel
 * crates/ssimulacra2-cuda-kernel/src/xyb.rs:81-102): packed f32 linear RGB to
 * packed positive XYB (X, Y, B per pixel), the scoring path's arithmetic. */
ssim2_status ssim2_linear_to_xyb(const ssim2_image* src_f32, ssim2_image* dst_f32, ssim2_stream stream);

/* CudaColorspace::biplanaryuv420_to_linearrgb_16 (crates/cuda-colorspace/src/
 * lib.rs:79-123; kernel crates/cuda-colorspace-kernel/src/biplanar.rs:8-70):
 * YUV 4:2:0 BT.709 limited -> packed f32 linear RGB (the power function is
 * the device powf, where the reference uses its fast approximation). */
ssim2_status ssim2_yuv420_to_linear(const ssim2_yuv_image* src, ssim2_image* dst_f32, ssim2_stream stream);

/* Device image helpers, the C form of cudarse-npp Image::malloc /
 * copy_from_cpu / copy_to_cpu (crates/cudarse/cudarse-npp/src/image/isu.rs:18-40,
 * image/mod.rs:410-439). bytes_per_sample is 1 (u8) or 4 (f32). Upload and
 * download are synchronous. */
ssim2_status ssim2_image_alloc(uint32_t width, uint32_t height, uint32_t bytes_per_sample,
                               ssim2_image* out);
ssim2_status ssim2_image_free(ssim2_image* image);
ssim2_status ssim2_image_upload(ssim2_image* image, const void* host, size_t host_row_bytes);
ssim2_status ssim2_image_download(const ssim2_image* image, void* host, size_t host_row_bytes);

/* YUV 4:2:0 frames in device memory: the decoded surfaces the reference's
 * video loop receives from NVDEC (crates/cudarse/cudarse-video/src/dec.rs,
 * consumed at crates/turbo-metrics/src/lib.rs:271-363). Planes get 256-byte
 * aligned pitches. Upload takes tightly packed host planes: Y w*h u16; for
 * SSIM2_YUV420P10 Cb and Cr ceil(w/2)*ceil(h/2) u16; for SSIM2_P016 one
 * interleaved CbCr plane 2*ceil(w/2)*ceil(h/2) u16 (c1 ignored). */
ssim2_status ssim2_yuv_alloc(uint32_t width, uint32_t height, uint32_t format, ssim2_yuv_image* out);
ssim2_status ssim2_yuv_free(ssim2_yuv_image* image);
ssim2_status ssim2_yuv_upload(ssim2_yuv_image* image, const uint16_t* y, const uint16_t* c0, const uint16_t* c1);

/* Stream helpers (cudarse-driver CuStream::new / sync, stream.rs:46,107). */
ssim2_status ssim2_stream_create(ssim2_stream* out);
ssim2_status ssim2_stream_destroy(ssim2_stream stream);
ssim2_status ssim2_stream_synchronize(ssim2_stream stream);

/* Human-readable name of a status, and the last error message recorded on the
 * calling thread ("" if none). */
const char* ssim2_status_string(ssim2_status status);
const char* ssim2_last_error(void);

/* Library version as major*10000 + minor*100 + patch. */
int ssim2_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SSIM2HIP_H */
