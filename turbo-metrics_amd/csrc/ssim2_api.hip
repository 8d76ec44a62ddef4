// C ABI of the ssimulacra2 HIP engine (include/ssim2hip.h): contexts, memory
// planning, launch sequencing and error handling. Host code only; the kernels
// live in ssim2_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/ssim2hip.h"
#include "ssim2_internal.h"
#include "ssim2_plan.h"

using namespace ssim2;

namespace {

thread_local std::string g_last_error;

ssim2_status fail(ssim2_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

ssim2_status hip_fail(hipError_t e, const char* what) {
  return fail(SSIM2_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return hip_fail(_e, #expr);   \
  } while (0)

}  // namespace

namespace ssim2 {
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace ssim2

struct ssim2_ctx {
  int w = 0, h = 0;
  PairPlan plan;
  float* arena = nullptr;
  double* partials = nullptr;
  // The result record (kResultDoubles) in pinned, device-mapped host memory:
  // the finalisation kernel writes it there directly, so a pair needs no
  // copy command after its kernels; `result` is its device address.
  double* result = nullptr;
  double* result_host = nullptr;
  PairDesc* d_desc = nullptr;
  WorkItem* d_items = nullptr;
  unsigned char* stage = nullptr;  // two packed u8 images for host uploads
  size_t stage_pitch = 0;
  size_t bytes = 0;
  bool has_ref = false;
  bool have_result = false;
  // Every call shares the arena, partial sums, staging images and result
  // record, whichever stream it names: `done` is recorded after each call's
  // last command, and a call on another stream than `last` first waits for it.
  hipEvent_t done = nullptr;
  bool issued = false;
  hipStream_t last = nullptr;
};

namespace {

void release(ssim2_ctx* c) {
  if (c->arena) (void)hipFree(c->arena);
  if (c->partials) (void)hipFree(c->partials);
  if (c->result_host) (void)hipHostFree(c->result_host);
  if (c->d_desc) (void)hipFree(c->d_desc);
  if (c->d_items) (void)hipFree(c->d_items);
  if (c->stage) (void)hipFree(c->stage);
  if (c->done) (void)hipEventDestroy(c->done);
  c->done = nullptr;
  c->arena = nullptr;
  c->partials = c->result = c->result_host = nullptr;
  c->d_desc = nullptr;
  c->d_items = nullptr;
  c->stage = nullptr;
}

ssim2_status check_image(const ssim2_ctx* c, const ssim2_image* im, size_t bps, const char* name) {
  if (!im || !im->ptr) return fail(SSIM2_INVALID_ARGUMENT, std::string(name) + " is null");
  if ((int)im->width != c->w || (int)im->height != c->h)
    return fail(SSIM2_DIM_MISMATCH, std::string(name) + " is " + std::to_string(im->width) + "x" +
                                        std::to_string(im->height) + ", context is " +
                                        std::to_string(c->w) + "x" + std::to_string(c->h));
  if (im->pitch_bytes < size_t(im->width) * 3 * bps)
    return fail(SSIM2_INVALID_ARGUMENT, std::string(name) + " pitch smaller than a row");
  if (bps == 4 && (im->pitch_bytes % 4 != 0 || (reinterpret_cast<uintptr_t>(im->ptr) & 3)))
    return fail(SSIM2_INVALID_ARGUMENT, std::string(name) + " f32 image not 4-byte aligned");
  return SSIM2_OK;
}

ssim2_status check_yuv(const ssim2_yuv_image* im, uint32_t w, uint32_t h, const char* name) {
  if (!im || !im->y || !im->c0 || (im->format == SSIM2_YUV420P10 && !im->c1))
    return fail(SSIM2_INVALID_ARGUMENT, std::string(name) + ": null plane");
  if (im->format > SSIM2_P016) return fail(SSIM2_INVALID_ARGUMENT, std::string(name) + ": unknown YUV format");
  if (im->width != w || im->height != h)
    return fail(SSIM2_DIM_MISMATCH, std::string(name) + " is " + std::to_string(im->width) + "x" +
                                        std::to_string(im->height) + ", expected " + std::to_string(w) + "x" +
                                        std::to_string(h));
  const size_t cw = (size_t(im->width) + 1) / 2;
  if (im->y_pitch_bytes < 2 * size_t(im->width) ||
      im->c_pitch_bytes < (im->format == SSIM2_P016 ? 4 : 2) * cw || (im->y_pitch_bytes & 1) || (im->c_pitch_bytes & 1))
    return fail(SSIM2_INVALID_ARGUMENT, std::string(name) + ": pitch too small or odd");
  return SSIM2_OK;
}

void yuv_inputs(InputArgs* in, const ssim2_yuv_image* ref, const ssim2_yuv_image* dis) {
  if (ref) {
    in->ref = static_cast<const unsigned char*>(ref->y);
    in->ref_c0 = static_cast<const unsigned char*>(ref->c0);
    in->ref_c1 = static_cast<const unsigned char*>(ref->c1);
    in->ref_pitch = (long long)ref->y_pitch_bytes;
    in->ref_cpitch = (long long)ref->c_pitch_bytes;
  }
  if (dis) {
    in->dis = static_cast<const unsigned char*>(dis->y);
    in->dis_c0 = static_cast<const unsigned char*>(dis->c0);
    in->dis_c1 = static_cast<const unsigned char*>(dis->c1);
    in->dis_pitch = (long long)dis->y_pitch_bytes;
    in->dis_cpitch = (long long)dis->c_pitch_bytes;
  }
  const ssim2_yuv_image* a = ref ? ref : dis;
  in->kind = a->format == SSIM2_P016 ? kInputYuvP016 : kInputYuvP10;
  in->use = 1;
}

InputArgs make_inputs(const ssim2_image* ref, const ssim2_image* dis, int kind) {
  InputArgs in{};
  in.ref = ref ? static_cast<const unsigned char*>(ref->ptr) : nullptr;
  in.dis = dis ? static_cast<const unsigned char*>(dis->ptr) : nullptr;
  in.ref_pitch = ref ? (long long)ref->pitch_bytes : 0;
  in.dis_pitch = dis ? (long long)dis->pitch_bytes : 0;
  in.kind = kind;
  in.use = 1;
  return in;
}

// Orders stream `s` behind every command the context has issued on other
// streams (the reference's calls are ordered the same way: one graph exec,
// crates/ssimulacra2-cuda/src/lib.rs:419-424).
ssim2_status join_stream(ssim2_ctx* c, hipStream_t s) {
  if (c->issued && c->last != s) HIP_TRY(hipStreamWaitEvent(s, c->done, 0));
  c->last = s;
  return SSIM2_OK;
}

// Marks the end of the commands just issued on `s`.
ssim2_status mark_issued(ssim2_ctx* c, hipStream_t s) {
  HIP_TRY(hipEventRecord(c->done, s));
  c->issued = true;
  return SSIM2_OK;
}

// Enqueue the pipeline for the context's single pair.
ssim2_status enqueue(ssim2_ctx* c, int sides, const InputArgs& in, hipStream_t s) {
  const PairPlan& p = c->plan;
  ssim2_status st = join_stream(c, s);
  if (st != SSIM2_OK) return st;
  HIP_TRY(launch_pyramid(c->d_desc, c->d_items + p.off_pyr, p.n_pyr, sides, in, s));
  HIP_TRY(launch_rows(c->d_desc, c->d_items + p.off_rows[sides], p.n_rows[sides], sides, s));
  HIP_TRY(launch_cols(c->d_desc, c->d_items + p.off_cols, p.n_cols, sides, s));
  if (sides != kSidesRef) {
    HIP_TRY(launch_finalize(c->d_desc, 1, s));  // writes the host record (result_host)
    c->have_result = true;
  }
  return mark_issued(c, s);
}

// Every command is ordered behind the previous one, so the last recorded
// event covers the last result copy whatever stream it ran on.
ssim2_status wait_result(ssim2_ctx* c) {
  if (!c->have_result) return fail(SSIM2_INVALID_ARGUMENT, "no computation has been enqueued");
  HIP_TRY(hipEventSynchronize(c->done));
  return SSIM2_OK;
}

}  // namespace

extern "C" {

int ssim2_version(void) { return 100; }

const char* ssim2_status_string(ssim2_status st) {
  switch (st) {
    case SSIM2_OK: return "SSIM2_OK";
    case SSIM2_INVALID_SIZE: return "SSIM2_INVALID_SIZE";
    case SSIM2_DIM_MISMATCH: return "SSIM2_DIM_MISMATCH";
    case SSIM2_NO_REFERENCE: return "SSIM2_NO_REFERENCE";
    case SSIM2_OOM: return "SSIM2_OOM";
    case SSIM2_BATCH_SIZE: return "SSIM2_BATCH_SIZE";
    case SSIM2_ADMISSION_REFUSED: return "SSIM2_ADMISSION_REFUSED";
    case SSIM2_HIP_ERROR: return "SSIM2_HIP_ERROR";
    case SSIM2_INVALID_ARGUMENT: return "SSIM2_INVALID_ARGUMENT";
    case SSIM2_NOT_READY: return "SSIM2_NOT_READY";
  }
  return "SSIM2_UNKNOWN";
}

const char* ssim2_last_error(void) { return g_last_error.c_str(); }

ssim2_status ssim2_create(uint32_t width, uint32_t height, const ssim2_options* options,
                          ssim2_ctx** out) {
  if (!out) return fail(SSIM2_INVALID_ARGUMENT, "out is null");
  *out = nullptr;
  if (options && options->struct_size != sizeof(ssim2_options))
    return fail(SSIM2_INVALID_ARGUMENT, "ssim2_options.struct_size mismatch");
  if (!size_supported(width, height))
    return fail(SSIM2_INVALID_SIZE, "image size " + std::to_string(width) + "x" +
                                        std::to_string(height) + " outside [8, " +
                                        std::to_string(kMaxDim) + "] or too many pixels");
  ssim2_ctx* c = new (std::nothrow) ssim2_ctx();
  if (!c) return fail(SSIM2_OOM, "host allocation failed");
  c->w = (int)width;
  c->h = (int)height;
  plan_pair(c->w, c->h, 0, 0, &c->plan);
  const PairPlan& p = c->plan;
  // + 64 KB: the row pass prefetches up to one row past a band's last row
  const size_t arena_bytes = p.arena_elems * sizeof(float) + 65536;
  const size_t part_bytes = size_t(p.nparts) * kNumSums * sizeof(double);
  const size_t items_bytes = p.items.size() * sizeof(WorkItem);
  c->bytes = arena_bytes + part_bytes + kResultDoubles * sizeof(double) + sizeof(PairDesc) + items_bytes;

  // Pre-flight budget, as the reference refuses before allocating (lib.rs:88-112).
  size_t budget = options ? (size_t)options->memory_budget : 0;
  if (budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) budget = free_b;
  }
  if (budget && c->bytes > budget) {
    delete c;
    return fail(SSIM2_OOM, "estimated footprint " + std::to_string(c->bytes) +
                               " B exceeds the budget of " + std::to_string(budget) + " B");
  }
  hipError_t e = hipSuccess;
  if ((e = hipMalloc(&c->arena, arena_bytes)) != hipSuccess ||
      (e = hipMalloc(&c->partials, std::max<size_t>(part_bytes, 8))) != hipSuccess ||
      (e = hipHostMalloc(&c->result_host, kResultDoubles * sizeof(double),
                         hipHostMallocMapped | hipHostMallocCoherent)) != hipSuccess ||
      (e = hipHostGetDevicePointer((void**)&c->result, c->result_host, 0)) != hipSuccess ||
      (e = hipMalloc(&c->d_desc, sizeof(PairDesc))) != hipSuccess ||
      (e = hipMalloc(&c->d_items, items_bytes)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&c->done, hipEventDisableTiming)) != hipSuccess) {
    release(c);
    delete c;
    (void)hipGetLastError();
    return fail(SSIM2_OOM, std::string("device allocation failed: ") + hipGetErrorString(e));
  }
  PairDesc d = p.desc;
  d.arena = c->arena;
  d.partials = c->partials;
  d.result = c->result;
  // The padding columns right of every plane row must read as zero (the row
  // scan takes them as the inputs past the image); no kernel writes them
  // except the row pass's own outputs, which no later pass reads there.
  if ((e = hipMemset(c->arena, 0, arena_bytes)) != hipSuccess ||
      (e = hipMemcpy(c->d_desc, &d, sizeof(PairDesc), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(c->d_items, p.items.data(), items_bytes, hipMemcpyHostToDevice)) != hipSuccess) {
    release(c);
    delete c;
    return hip_fail(e, "descriptor upload");
  }
  std::memset(c->result_host, 0, kResultDoubles * sizeof(double));
  *out = c;
  return SSIM2_OK;
}

ssim2_status ssim2_destroy(ssim2_ctx* c) {
  if (!c) return SSIM2_OK;
  // only this context's own commands are waited for (Drop, lib.rs:1195-1202)
  if (c->issued) (void)hipEventSynchronize(c->done);
  release(c);
  delete c;
  return SSIM2_OK;
}

ssim2_status ssim2_mem_usage(const ssim2_ctx* c, size_t* bytes) {
  if (!c || !bytes) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  *bytes = c->bytes + (c->stage ? 2 * c->stage_pitch * size_t(c->h) : 0);
  return SSIM2_OK;
}

ssim2_status ssim2_compute_linear_async(ssim2_ctx* c, const ssim2_image* ref,
                                        const ssim2_image* dis, ssim2_stream stream) {
  if (!c) return fail(SSIM2_INVALID_ARGUMENT, "ctx is null");
  ssim2_status st;
  if ((st = check_image(c, ref, 4, "ref_linear")) != SSIM2_OK) return st;
  if ((st = check_image(c, dis, 4, "dis_linear")) != SSIM2_OK) return st;
  return enqueue(c, kSidesBoth, make_inputs(ref, dis, kInputLinearF32), (hipStream_t)stream);
}

ssim2_status ssim2_get_score(ssim2_ctx* c, double* score) {
  if (!c || !score) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_status st = wait_result(c);
  if (st != SSIM2_OK) return st;
  *score = c->result_host[0];
  return SSIM2_OK;
}

ssim2_status ssim2_compute_linear_sync(ssim2_ctx* c, const ssim2_image* ref, const ssim2_image* dis,
                                       ssim2_stream stream, double* score) {
  if (!score) return fail(SSIM2_INVALID_ARGUMENT, "score is null");
  ssim2_status st = ssim2_compute_linear_async(c, ref, dis, stream);
  if (st != SSIM2_OK) return st;
  return ssim2_get_score(c, score);
}

ssim2_status ssim2_compute_srgb_async(ssim2_ctx* c, const ssim2_image* ref, const ssim2_image* dis,
                                      ssim2_stream stream) {
  if (!c) return fail(SSIM2_INVALID_ARGUMENT, "ctx is null");
  ssim2_status st;
  if ((st = check_image(c, ref, 1, "ref_srgb")) != SSIM2_OK) return st;
  if ((st = check_image(c, dis, 1, "dis_srgb")) != SSIM2_OK) return st;
  return enqueue(c, kSidesBoth, make_inputs(ref, dis, kInputSrgbU8), (hipStream_t)stream);
}

ssim2_status ssim2_compute_srgb_sync(ssim2_ctx* c, const ssim2_image* ref, const ssim2_image* dis,
                                     ssim2_image* ref_lin, ssim2_image* dis_lin, ssim2_stream stream,
                                     double* score) {
  if (!c || !score) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_status st;
  if ((st = check_image(c, ref, 1, "ref_srgb")) != SSIM2_OK) return st;
  if ((st = check_image(c, dis, 1, "dis_srgb")) != SSIM2_OK) return st;
  if (ref_lin && (st = check_image(c, ref_lin, 4, "ref_linear")) != SSIM2_OK) return st;
  if (dis_lin && (st = check_image(c, dis_lin, 4, "dis_linear")) != SSIM2_OK) return st;
  InputArgs in = make_inputs(ref, dis, kInputSrgbU8);
  if (ref_lin) {
    in.lin_ref_out = static_cast<float*>(ref_lin->ptr);
    in.lin_ref_pitch = (long long)ref_lin->pitch_bytes;
  }
  if (dis_lin) {
    in.lin_dis_out = static_cast<float*>(dis_lin->ptr);
    in.lin_dis_pitch = (long long)dis_lin->pitch_bytes;
  }
  if ((st = enqueue(c, kSidesBoth, in, (hipStream_t)stream)) != SSIM2_OK) return st;
  return ssim2_get_score(c, score);
}

static ssim2_status ensure_stage(ssim2_ctx* c) {
  if (c->stage) return SSIM2_OK;
  c->stage_pitch = (size_t(c->w) * 3 + 255) / 256 * 256;
  hipError_t e = hipMalloc(&c->stage, 2 * c->stage_pitch * size_t(c->h));
  if (e != hipSuccess) {
    c->stage = nullptr;
    (void)hipGetLastError();
    return fail(SSIM2_OOM, "staging allocation failed");
  }
  return SSIM2_OK;
}

// Uploads `host` into staging image `which` (0 ref, 1 dis) on `s`.
static ssim2_status upload_stage(ssim2_ctx* c, int which, const uint8_t* host, size_t row_bytes, hipStream_t s,
                          ssim2_image* out) {
  const size_t rb = size_t(c->w) * 3;
  if (row_bytes < rb) return fail(SSIM2_INVALID_ARGUMENT, "row_bytes smaller than 3*width");
  ssim2_status st = ensure_stage(c);
  if (st != SSIM2_OK) return st;
  unsigned char* p = c->stage + (which ? c->stage_pitch * size_t(c->h) : 0);
  if ((st = join_stream(c, s)) != SSIM2_OK) return st;
  HIP_TRY(hipMemcpy2DAsync(p, c->stage_pitch, host, row_bytes, rb, c->h, hipMemcpyHostToDevice, s));
  *out = ssim2_image{p, c->stage_pitch, (uint32_t)c->w, (uint32_t)c->h};
  return mark_issued(c, s);
}

ssim2_status ssim2_compute_from_host_srgb(ssim2_ctx* c, const uint8_t* ref, const uint8_t* dis,
                                          size_t row_bytes, ssim2_stream stream, double* score) {
  if (!c || !ref || !dis || !score) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  hipStream_t s = (hipStream_t)stream;
  ssim2_image ri, di;
  ssim2_status st = upload_stage(c, 0, ref, row_bytes, s, &ri);
  if (st == SSIM2_OK) st = upload_stage(c, 1, dis, row_bytes, s, &di);
  if (st != SSIM2_OK) return st;
  st = enqueue(c, kSidesBoth, make_inputs(&ri, &di, kInputSrgbU8), s);
  if (st != SSIM2_OK) return st;
  return ssim2_get_score(c, score);
}

ssim2_status ssim2_prepare_reference_from_host_srgb(ssim2_ctx* c, const uint8_t* ref, size_t row_bytes,
                                                    ssim2_stream stream) {
  if (!c || !ref) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_image ri;
  ssim2_status st = upload_stage(c, 0, ref, row_bytes, (hipStream_t)stream, &ri);
  if (st != SSIM2_OK) return st;
  return ssim2_set_reference_srgb(c, &ri, stream);
}

ssim2_status ssim2_compute_from_host_dis_srgb(ssim2_ctx* c, const uint8_t* dis, size_t row_bytes,
                                              ssim2_stream stream, double* score) {
  if (!c || !dis || !score) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  int has = 0;
  ssim2_has_reference(c, &has);
  if (!has) return fail(SSIM2_NO_REFERENCE, "no cached reference: call a set/prepare_reference entry first");
  ssim2_image di;
  ssim2_status st = upload_stage(c, 1, dis, row_bytes, (hipStream_t)stream, &di);
  if (st != SSIM2_OK) return st;
  return ssim2_compute_with_reference_srgb(c, &di, stream, score);
}

ssim2_status ssim2_compute_yuv420_async(ssim2_ctx* c, const ssim2_yuv_image* ref, const ssim2_yuv_image* dis,
                                        ssim2_stream stream) {
  if (!c) return fail(SSIM2_INVALID_ARGUMENT, "ctx is null");
  ssim2_status st;
  if ((st = check_yuv(ref, c->w, c->h, "ref")) != SSIM2_OK) return st;
  if ((st = check_yuv(dis, c->w, c->h, "dis")) != SSIM2_OK) return st;
  if (ref->format != dis->format) return fail(SSIM2_INVALID_ARGUMENT, "ref and dis YUV formats differ");
  InputArgs in{};
  yuv_inputs(&in, ref, dis);
  return enqueue(c, kSidesBoth, in, (hipStream_t)stream);
}

ssim2_status ssim2_compute_yuv420_sync(ssim2_ctx* c, const ssim2_yuv_image* ref, const ssim2_yuv_image* dis,
                                       ssim2_stream stream, double* score) {
  if (!score) return fail(SSIM2_INVALID_ARGUMENT, "score is null");
  ssim2_status st = ssim2_compute_yuv420_async(c, ref, dis, stream);
  if (st != SSIM2_OK) return st;
  return ssim2_get_score(c, score);
}

// ---- cached reference

static ssim2_status set_reference(ssim2_ctx* c, const ssim2_image* ref, int kind, hipStream_t s) {
  c->has_ref = false;
  ssim2_status st = enqueue(c, kSidesRef, make_inputs(ref, nullptr, kind), s);
  if (st != SSIM2_OK) return st;
  HIP_TRY(hipStreamSynchronize(s));
  c->has_ref = true;
  return SSIM2_OK;
}

ssim2_status ssim2_set_reference_linear(ssim2_ctx* c, const ssim2_image* ref, ssim2_stream stream) {
  if (!c) return fail(SSIM2_INVALID_ARGUMENT, "ctx is null");
  ssim2_status st = check_image(c, ref, 4, "ref_linear");
  if (st != SSIM2_OK) return st;
  return set_reference(c, ref, kInputLinearF32, (hipStream_t)stream);
}

ssim2_status ssim2_set_reference_srgb(ssim2_ctx* c, const ssim2_image* ref, ssim2_stream stream) {
  if (!c) return fail(SSIM2_INVALID_ARGUMENT, "ctx is null");
  ssim2_status st = check_image(c, ref, 1, "ref_srgb");
  if (st != SSIM2_OK) return st;
  return set_reference(c, ref, kInputSrgbU8, (hipStream_t)stream);
}

static ssim2_status with_reference(ssim2_ctx* c, const ssim2_image* dis, int kind, hipStream_t s,
                                   double* score) {
  if (!c->has_ref)
    return fail(SSIM2_NO_REFERENCE, "set_reference must be called before compute_with_reference");
  ssim2_status st = enqueue(c, kSidesDis, make_inputs(nullptr, dis, kind), s);
  if (st != SSIM2_OK) return st;
  return ssim2_get_score(c, score);
}

ssim2_status ssim2_compute_with_reference_linear(ssim2_ctx* c, const ssim2_image* ref,
                                                 const ssim2_image* dis, ssim2_stream stream,
                                                 double* score) {
  if (!c || !score) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_status st;
  if (ref && (st = check_image(c, ref, 4, "ref_linear")) != SSIM2_OK) return st;
  if ((st = check_image(c, dis, 4, "dis_linear")) != SSIM2_OK) return st;
  return with_reference(c, dis, kInputLinearF32, (hipStream_t)stream, score);
}

ssim2_status ssim2_compute_with_reference_srgb(ssim2_ctx* c, const ssim2_image* dis,
                                               ssim2_stream stream, double* score) {
  if (!c || !score) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_status st = check_image(c, dis, 1, "dis_srgb");
  if (st != SSIM2_OK) return st;
  return with_reference(c, dis, kInputSrgbU8, (hipStream_t)stream, score);
}

ssim2_status ssim2_clear_reference(ssim2_ctx* c) {
  if (!c) return fail(SSIM2_INVALID_ARGUMENT, "ctx is null");
  c->has_ref = false;
  return SSIM2_OK;
}

ssim2_status ssim2_has_reference(const ssim2_ctx* c, int* has) {
  if (!c || !has) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  *has = c->has_ref ? 1 : 0;
  return SSIM2_OK;
}

ssim2_status ssim2_get_sums(ssim2_ctx* c, double* sums, int* dims, int* nscales) {
  if (!c) return fail(SSIM2_INVALID_ARGUMENT, "ctx is null");
  ssim2_status st = wait_result(c);
  if (st != SSIM2_OK) return st;
  if (sums) std::memcpy(sums, c->result_host + 1, kSumsPerPair * sizeof(double));
  if (dims)
    for (int i = 0; i < 2 * kMaxScales; ++i) dims[i] = (int)c->result_host[1 + kSumsPerPair + i];
  if (nscales) *nscales = c->plan.desc.nscales;
  return SSIM2_OK;
}

ssim2_status ssim2_get_plane(ssim2_ctx* c, uint32_t kind, uint32_t scale, uint32_t index, float* host,
                             size_t host_row_floats) {
  if (!c || !host) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  const PairDesc& d = c->plan.desc;
  if ((int)scale >= d.nscales) return fail(SSIM2_INVALID_ARGUMENT, "scale beyond the pair's scale count");
  if (kind > 1 || index >= (kind == 0 ? 6u : 3u * kNumQ)) return fail(SSIM2_INVALID_ARGUMENT, "bad plane kind or index");
  const ScaleDesc& S = d.sc[scale];
  if (host_row_floats < size_t(S.w)) return fail(SSIM2_INVALID_ARGUMENT, "host rows shorter than the plane's width");
  if (c->issued) HIP_TRY(hipEventSynchronize(c->done));
  const float* src = c->arena + (kind == 0 ? S.xyb : S.hrow) + (long long)index * S.plane;
  HIP_TRY(hipMemcpy2D(host, host_row_floats * sizeof(float), src, size_t(S.pitch) * sizeof(float),
                      size_t(S.w) * sizeof(float), S.h, hipMemcpyDeviceToHost));
  return SSIM2_OK;
}

ssim2_status ssim2_time_kernels(ssim2_ctx* c, const ssim2_image* ref, const ssim2_image* dis,
                                ssim2_stream stream, int reps, float* ms) {
  if (!c || !ms || reps <= 0) return fail(SSIM2_INVALID_ARGUMENT, "bad time_kernels arguments");
  ssim2_status st;
  if ((st = check_image(c, ref, 1, "ref_srgb")) != SSIM2_OK) return st;
  if ((st = check_image(c, dis, 1, "dis_srgb")) != SSIM2_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  const PairPlan& p = c->plan;
  const InputArgs in = make_inputs(ref, dis, kInputSrgbU8);
  if ((st = join_stream(c, s)) != SSIM2_OK) return st;
  hipEvent_t ev[5];
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  double acc[4] = {0, 0, 0, 0};
  ssim2_status out = SSIM2_OK;
  for (int r = 0; r < reps && out == SSIM2_OK; ++r) {
    hipError_t e = hipSuccess;
    if ((e = hipEventRecord(ev[0], s)) != hipSuccess ||
        (e = launch_pyramid(c->d_desc, c->d_items + p.off_pyr, p.n_pyr, kSidesBoth, in, s)) != hipSuccess ||
        (e = hipEventRecord(ev[1], s)) != hipSuccess ||
        (e = launch_rows(c->d_desc, c->d_items + p.off_rows[kSidesBoth], p.n_rows[kSidesBoth], kSidesBoth, s)) != hipSuccess ||
        (e = hipEventRecord(ev[2], s)) != hipSuccess ||
        (e = launch_cols(c->d_desc, c->d_items + p.off_cols, p.n_cols, kSidesBoth, s)) != hipSuccess ||
        (e = hipEventRecord(ev[3], s)) != hipSuccess ||
        (e = launch_finalize(c->d_desc, 1, s)) != hipSuccess ||
        (e = hipEventRecord(ev[4], s)) != hipSuccess || (e = hipEventSynchronize(ev[4])) != hipSuccess) {
      out = hip_fail(e, "time_kernels");
      break;
    }
    for (int k = 0; k < 4; ++k) {
      float t = 0.0f;
      if ((e = hipEventElapsedTime(&t, ev[k], ev[k + 1])) != hipSuccess) {
        out = hip_fail(e, "hipEventElapsedTime");
        break;
      }
      acc[k] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (out != SSIM2_OK) return out;
  for (int k = 0; k < 4; ++k) ms[k] = (float)(acc[k] / reps);
  c->have_result = true;
  return mark_issued(c, s);
}

ssim2_status ssim2_device_copy(void* dst, const void* src, size_t bytes, ssim2_stream stream) {
  if (!dst || !src || (bytes % 16)) return fail(SSIM2_INVALID_ARGUMENT, "copy needs 16-byte multiples");
  HIP_TRY(launch_copy(src, dst, (long long)bytes, (hipStream_t)stream));
  return SSIM2_OK;
}

ssim2_status ssim2_check_arithmetic(uint32_t which, uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches) {
  if (!mismatches || lo_bits > hi_bits || which > 2) return fail(SSIM2_INVALID_ARGUMENT, "bad check_arithmetic arguments");
  unsigned long long* d = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof *d));
  unsigned long long h = 0;
  hipError_t e = hipMemset(d, 0, sizeof *d);
  if (e == hipSuccess) e = launch_check_arithmetic((int)which, lo_bits, hi_bits, d, nullptr);
  if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "check_arithmetic");
  *mismatches = h;
  return SSIM2_OK;
}

ssim2_status ssim2_device_alloc(size_t bytes, void** out) {
  if (!out || !bytes) return fail(SSIM2_INVALID_ARGUMENT, "bad alloc arguments");
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return fail(SSIM2_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return SSIM2_OK;
}

ssim2_status ssim2_device_free(void* p) {
  if (p) HIP_TRY(hipFree(p));
  return SSIM2_OK;
}

ssim2_status ssim2_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return SSIM2_OK;
}

// ---- utilities

ssim2_status ssim2_srgb_to_linear(const ssim2_image* src, ssim2_image* dst, ssim2_stream stream) {
  if (!src || !dst || !src->ptr || !dst->ptr) return fail(SSIM2_INVALID_ARGUMENT, "null image");
  if (src->width != dst->width || src->height != dst->height)
    return fail(SSIM2_DIM_MISMATCH, "source and destination sizes differ");
  if (src->pitch_bytes < size_t(src->width) * 3 || dst->pitch_bytes < size_t(dst->width) * 12 ||
      dst->pitch_bytes % 4)
    return fail(SSIM2_INVALID_ARGUMENT, "pitch too small or misaligned");
  HIP_TRY(launch_srgb_to_linear(static_cast<const unsigned char*>(src->ptr), (long long)src->pitch_bytes,
                                static_cast<float*>(dst->ptr), (long long)dst->pitch_bytes,
                                (int)src->width, (int)src->height, (hipStream_t)stream));
  return SSIM2_OK;
}

static ssim2_status check_f32_image(const ssim2_image* im, const char* name) {
  if (!im || !im->ptr) return fail(SSIM2_INVALID_ARGUMENT, std::string(name) + " is null");
  if (im->width == 0 || im->height == 0) return fail(SSIM2_INVALID_SIZE, std::string(name) + " is empty");
  if (im->pitch_bytes < size_t(im->width) * 12 || im->pitch_bytes % 4 || (reinterpret_cast<uintptr_t>(im->ptr) & 3))
    return fail(SSIM2_INVALID_ARGUMENT, std::string(name) + " pitch too small or misaligned");
  return SSIM2_OK;
}

ssim2_status ssim2_downscale_by_2(const ssim2_image* src, ssim2_image* dst, ssim2_stream stream) {
  ssim2_status st;
  if ((st = check_f32_image(src, "src")) != SSIM2_OK || (st = check_f32_image(dst, "dst")) != SSIM2_OK) return st;
  if (dst->width != (src->width + 1) / 2 || dst->height != (src->height + 1) / 2)
    return fail(SSIM2_DIM_MISMATCH, "dst must be ceil(width / 2) x ceil(height / 2) of src");
  HIP_TRY(launch_downscale_by_2(static_cast<const float*>(src->ptr), (long long)src->pitch_bytes, (int)src->width,
                                (int)src->height, static_cast<float*>(dst->ptr), (long long)dst->pitch_bytes,
                                (int)dst->width, (int)dst->height, (hipStream_t)stream));
  return SSIM2_OK;
}

ssim2_status ssim2_linear_to_xyb(const ssim2_image* src, ssim2_image* dst, ssim2_stream stream) {
  ssim2_status st;
  if ((st = check_f32_image(src, "src")) != SSIM2_OK || (st = check_f32_image(dst, "dst")) != SSIM2_OK) return st;
  if (src->width != dst->width || src->height != dst->height)
    return fail(SSIM2_DIM_MISMATCH, "source and destination sizes differ");
  HIP_TRY(launch_linear_to_xyb(static_cast<const float*>(src->ptr), (long long)src->pitch_bytes,
                               static_cast<float*>(dst->ptr), (long long)dst->pitch_bytes, (int)src->width,
                               (int)src->height, (hipStream_t)stream));
  return SSIM2_OK;
}

ssim2_status ssim2_yuv420_to_linear(const ssim2_yuv_image* src, ssim2_image* dst, ssim2_stream stream) {
  if (!src || !dst || !dst->ptr) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_status st = check_yuv(src, dst->width, dst->height, "src");
  if (st != SSIM2_OK) return st;
  if (dst->pitch_bytes < size_t(dst->width) * 12 || dst->pitch_bytes % 4)
    return fail(SSIM2_INVALID_ARGUMENT, "dst pitch too small or misaligned");
  InputArgs in{};
  yuv_inputs(&in, src, nullptr);
  InputView v{in.ref, in.ref_c0, in.ref_c1, in.ref_pitch, in.ref_cpitch, in.kind};
  HIP_TRY(launch_yuv_to_linear(v, static_cast<float*>(dst->ptr), (long long)dst->pitch_bytes, (int)dst->width,
                               (int)dst->height, (hipStream_t)stream));
  return SSIM2_OK;
}

ssim2_status ssim2_image_alloc(uint32_t width, uint32_t height, uint32_t bps, ssim2_image* out) {
  if (!out || (bps != 1 && bps != 4) || width == 0 || height == 0)
    return fail(SSIM2_INVALID_ARGUMENT, "bad image_alloc arguments");
  const size_t row = size_t(width) * 3 * bps;
  const size_t pitch = (row + 255) / 256 * 256;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, pitch * height);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(SSIM2_OOM, std::string("image allocation failed: ") + hipGetErrorString(e));
  }
  out->ptr = p;
  out->pitch_bytes = pitch;
  out->width = width;
  out->height = height;
  return SSIM2_OK;
}

ssim2_status ssim2_image_free(ssim2_image* im) {
  if (!im) return SSIM2_OK;
  if (im->ptr) HIP_TRY(hipFree(im->ptr));
  im->ptr = nullptr;
  return SSIM2_OK;
}

ssim2_status ssim2_image_upload(ssim2_image* im, const void* host, size_t host_row_bytes) {
  if (!im || !im->ptr || !host) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  if (host_row_bytes < size_t(im->width) * 3)
    return fail(SSIM2_INVALID_ARGUMENT, "host_row_bytes smaller than a u8 row");
  const size_t row = std::min(host_row_bytes, im->pitch_bytes);
  HIP_TRY(hipMemcpy2D(im->ptr, im->pitch_bytes, host, host_row_bytes, row, im->height,
                      hipMemcpyHostToDevice));
  return SSIM2_OK;
}

ssim2_status ssim2_image_download(const ssim2_image* im, void* host, size_t host_row_bytes) {
  if (!im || !im->ptr || !host) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  const size_t row = std::min(host_row_bytes, im->pitch_bytes);
  HIP_TRY(hipMemcpy2D(host, host_row_bytes, im->ptr, im->pitch_bytes, row, im->height,
                      hipMemcpyDeviceToHost));
  return SSIM2_OK;
}

ssim2_status ssim2_yuv_alloc(uint32_t w, uint32_t h, uint32_t format, ssim2_yuv_image* out) {
  if (!out) return fail(SSIM2_INVALID_ARGUMENT, "out is null");
  *out = ssim2_yuv_image{};
  if (w == 0 || h == 0 || w > kMaxDim || h > kMaxDim) return fail(SSIM2_INVALID_SIZE, "bad frame size");
  if (format > SSIM2_P016) return fail(SSIM2_INVALID_ARGUMENT, "unknown YUV format");
  const size_t cw = (size_t(w) + 1) / 2, ch = (size_t(h) + 1) / 2;
  const size_t yp = (size_t(w) * 2 + 255) & ~size_t(255);
  const size_t cp = ((format == SSIM2_P016 ? 4 : 2) * cw + 255) & ~size_t(255);
  const size_t nplanes_c = format == SSIM2_P016 ? 1 : 2;
  void* p = nullptr;
  HIP_TRY(hipMalloc(&p, yp * h + nplanes_c * cp * ch));
  unsigned char* b = static_cast<unsigned char*>(p);
  out->y = b;
  out->c0 = b + yp * h;
  out->c1 = format == SSIM2_P016 ? nullptr : b + yp * h + cp * ch;
  out->y_pitch_bytes = yp;
  out->c_pitch_bytes = cp;
  out->width = w;
  out->height = h;
  out->format = format;
  return SSIM2_OK;
}

ssim2_status ssim2_yuv_free(ssim2_yuv_image* im) {
  if (!im) return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  if (im->y) HIP_TRY(hipFree(const_cast<void*>(im->y)));
  *im = ssim2_yuv_image{};
  return SSIM2_OK;
}

ssim2_status ssim2_yuv_upload(ssim2_yuv_image* im, const uint16_t* y, const uint16_t* c0, const uint16_t* c1) {
  if (!im || !im->y || !y || !c0 || (im->format == SSIM2_YUV420P10 && !c1))
    return fail(SSIM2_INVALID_ARGUMENT, "null argument");
  const size_t w = im->width, h = im->height, cw = (w + 1) / 2, ch = (h + 1) / 2;
  HIP_TRY(hipMemcpy2D(const_cast<void*>(im->y), im->y_pitch_bytes, y, w * 2, w * 2, h, hipMemcpyHostToDevice));
  const size_t crow = (im->format == SSIM2_P016 ? 4 : 2) * cw;
  HIP_TRY(hipMemcpy2D(const_cast<void*>(im->c0), im->c_pitch_bytes, c0, crow, crow, ch, hipMemcpyHostToDevice));
  if (im->format == SSIM2_YUV420P10)
    HIP_TRY(hipMemcpy2D(const_cast<void*>(im->c1), im->c_pitch_bytes, c1, crow, crow, ch, hipMemcpyHostToDevice));
  return SSIM2_OK;
}

ssim2_status ssim2_stream_create(ssim2_stream* out) {
  if (!out) return fail(SSIM2_INVALID_ARGUMENT, "out is null");
  hipStream_t s;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = s;
  return SSIM2_OK;
}

ssim2_status ssim2_stream_destroy(ssim2_stream stream) {
  if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream));
  return SSIM2_OK;
}

ssim2_status ssim2_stream_synchronize(ssim2_stream stream) {
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return SSIM2_OK;
}

}  // extern "C"
