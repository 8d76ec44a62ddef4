// The video/metric driver and its statistics: host code over the C ABI.
//
// ssim2_video_compute_all follows TurboMetrics::compute_all / compute_one
// (crates/turbo-metrics/src/lib.rs:311-442) for the ssimulacra2 metric:
// frames are pulled from two caller-supplied sources, the leading frames are
// discarded per source (skip + skip_ref, skip + skip_dis), `every` keeps one
// decoded pair in `every` (the first always), `frames` bounds the decoded
// pairs, and each kept pair is scored. Exact mode scores pair by pair with the
// exact path (as the reference does, synchronously); fast mode copies the
// frames into device slots owned by the driver (a source may reuse its
// buffers) and scores them in batches with the fast form.
//
// ssim2_stats_compute restates quick_stats::full::Stats::compute
// (crates/quick-stats/src/lib.rs:21-98): min/max of the sorted values, mean
// of the sorted values, population and sample variance about that mean over
// the values in their given order, and percentiles by linear interpolation
// between closest ranks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ssim2hip.h"
#include "ssim2_internal.h"

namespace {

ssim2_status vfail(ssim2_status st, const std::string& msg) {
  ssim2::set_last_error(msg);
  return st;
}

// quick-stats percentile_of_sorted (lib.rs:56-77)
double percentile_of_sorted(const std::vector<double>& s, double pct) {
  if (s.size() == 1) return s[0];
  if (pct == 100.0) return s.back();
  const double rank = (pct / 100.0) * (double)(s.size() - 1);
  const double lrank = std::floor(rank);
  const double d = rank - lrank;
  const size_t n = (size_t)lrank;
  return s[n] + (s[n + 1] - s[n]) * d;
}

// quick-stats compute_var (lib.rs:79-98)
double compute_var(const double* v, size_t n, double mean, bool sample) {
  if (n < 2) return 0.0;
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double x = v[i] - mean;
    acc += x * x;
  }
  return acc / (double)(sample ? n - 1 : n);
}

size_t frame_bytes(const ssim2_frame& f, size_t* plane_bytes) {
  if (f.kind == SSIM2_FRAME_SRGB) {
    plane_bytes[0] = f.srgb.pitch_bytes * f.srgb.height;
    plane_bytes[1] = plane_bytes[2] = 0;
  } else {
    const size_t ch = (f.yuv.height + 1) / 2;
    plane_bytes[0] = f.yuv.y_pitch_bytes * f.yuv.height;
    plane_bytes[1] = f.yuv.c_pitch_bytes * ch;
    plane_bytes[2] = f.yuv.format == SSIM2_YUV420P10 ? f.yuv.c_pitch_bytes * ch : 0;
  }
  return plane_bytes[0] + plane_bytes[1] + plane_bytes[2];
}

// A driver-owned copy of one frame (fast mode).
struct FrameSlot {
  void* mem = nullptr;
  size_t cap = 0;
  ssim2_frame f{};
};

ssim2_status copy_frame(FrameSlot& slot, const ssim2_frame& src, hipStream_t s) {
  size_t pb[3];
  const size_t bytes = frame_bytes(src, pb);
  if (bytes > slot.cap) {
    if (slot.mem) (void)hipFree(slot.mem);
    slot.mem = nullptr;
    slot.cap = 0;
    if (hipMalloc(&slot.mem, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return vfail(SSIM2_OOM, "frame slot allocation failed");
    }
    slot.cap = bytes;
  }
  unsigned char* d = static_cast<unsigned char*>(slot.mem);
  slot.f = src;
  if (src.kind == SSIM2_FRAME_SRGB) {
    if (hipMemcpyAsync(d, src.srgb.ptr, pb[0], hipMemcpyDeviceToDevice, s) != hipSuccess)
      return vfail(SSIM2_HIP_ERROR, "frame copy failed");
    slot.f.srgb.ptr = d;
  } else {
    const void* planes[3] = {src.yuv.y, src.yuv.c0, src.yuv.c1};
    size_t off = 0;
    void* dst[3] = {nullptr, nullptr, nullptr};
    for (int i = 0; i < 3; ++i) {
      if (!pb[i]) continue;
      dst[i] = d + off;
      if (hipMemcpyAsync(dst[i], planes[i], pb[i], hipMemcpyDeviceToDevice, s) != hipSuccess)
        return vfail(SSIM2_HIP_ERROR, "frame copy failed");
      off += pb[i];
    }
    slot.f.yuv.y = dst[0];
    slot.f.yuv.c0 = dst[1];
    slot.f.yuv.c1 = dst[2];
  }
  return SSIM2_OK;
}

ssim2_status check_frame(const ssim2_frame& f, uint32_t w, uint32_t h, const char* name) {
  const uint32_t fw = f.kind == SSIM2_FRAME_SRGB ? f.srgb.width : f.yuv.width;
  const uint32_t fh = f.kind == SSIM2_FRAME_SRGB ? f.srgb.height : f.yuv.height;
  if (f.kind > SSIM2_FRAME_YUV420) return vfail(SSIM2_INVALID_ARGUMENT, std::string(name) + ": unknown frame kind");
  if (fw != w || fh != h)
    return vfail(SSIM2_DIM_MISMATCH, std::string(name) + " frame is " + std::to_string(fw) + "x" +
                                         std::to_string(fh) + ", expected " + std::to_string(w) + "x" +
                                         std::to_string(h));
  return SSIM2_OK;
}

// next frame of a source: 1 = frame, 0 = end, error otherwise
int pull(ssim2_frame_source src, void* user, ssim2_frame* f, ssim2_status* err) {
  std::memset(f, 0, sizeof(*f));
  const int r = src(user, f);
  if (r < 0) *err = vfail(SSIM2_INVALID_ARGUMENT, "frame source reported an error");
  return r;
}

}  // namespace

extern "C" {

ssim2_status ssim2_stats_compute(const double* values, size_t n, ssim2_stats* out) {
  if (!out || (!values && n)) return vfail(SSIM2_INVALID_ARGUMENT, "null argument");
  if (n == 0) return vfail(SSIM2_INVALID_ARGUMENT, "no values");
  std::vector<double> sorted(values, values + n);
  // total order (f64::total_cmp): -NaN < -inf < ... < -0 < +0 < ... < +inf < NaN
  auto key = [](double x) {
    int64_t i;
    std::memcpy(&i, &x, 8);
    return i ^ (int64_t)((uint64_t)(i >> 63) >> 1);
  };
  std::sort(sorted.begin(), sorted.end(), [&](double a, double b) { return key(a) < key(b); });
  double sum = 0.0;
  for (double v : sorted) sum += v;
  ssim2_stats s{};
  s.min = sorted.front();
  s.max = sorted.back();
  s.mean = sum / (double)n;
  s.var = compute_var(values, n, s.mean, false);
  s.sample_var = compute_var(values, n, s.mean, true);
  s.stddev = std::sqrt(s.var);
  s.sample_stddev = std::sqrt(s.sample_var);
  s.p1 = percentile_of_sorted(sorted, 1.0);
  s.p5 = percentile_of_sorted(sorted, 5.0);
  s.p50 = percentile_of_sorted(sorted, 50.0);
  s.p95 = percentile_of_sorted(sorted, 95.0);
  s.p99 = percentile_of_sorted(sorted, 99.0);
  *out = s;
  return SSIM2_OK;
}

ssim2_status ssim2_video_compute_all(uint32_t width, uint32_t height, const ssim2_video_options* opt,
                                     ssim2_frame_source ref_src, void* ref_user, ssim2_frame_source dis_src,
                                     void* dis_user, ssim2_stream stream, double* scores, size_t capacity,
                                     size_t* frame_count, ssim2_stats* stats) {
  if (!opt || opt->struct_size < sizeof(ssim2_video_options) || !ref_src || !dis_src || !frame_count)
    return vfail(SSIM2_INVALID_ARGUMENT, "null argument or old options struct");
  if (opt->mode > SSIM2_VIDEO_FAST) return vfail(SSIM2_INVALID_ARGUMENT, "unknown mode");
  hipStream_t s = (hipStream_t)stream;
  *frame_count = 0;
  ssim2_status st = SSIM2_OK;
  std::vector<double> out;

  ssim2_ctx* ctx = nullptr;
  ssim2_batch* batch = nullptr;
  const uint32_t nb = opt->batch ? opt->batch : 8;
  std::vector<FrameSlot> slots_r, slots_d;
  if (opt->mode == SSIM2_VIDEO_EXACT) {
    if ((st = ssim2_create(width, height, nullptr, &ctx)) != SSIM2_OK) return st;
  } else {
    ssim2_batch_options bo{};
    bo.struct_size = sizeof(bo);
    bo.storage = opt->storage;
    bo.max_pairs = nb;
    bo.max_pixels = (uint64_t)nb * width * height;
    if ((st = ssim2_batch_create(&bo, &batch)) != SSIM2_OK) return st;
    slots_r.resize(nb);
    slots_d.resize(nb);
  }
  uint32_t pending = 0;
  auto flush = [&]() -> ssim2_status {
    if (!pending) return SSIM2_OK;
    const bool yuv = slots_r[0].f.kind == SSIM2_FRAME_YUV420;
    ssim2_status e;
    if (yuv) {
      std::vector<ssim2_yuv_pair> p(pending);
      for (uint32_t i = 0; i < pending; ++i) p[i] = {slots_r[i].f.yuv, slots_d[i].f.yuv};
      e = ssim2_batch_compute_yuv_async(batch, p.data(), pending, stream);
    } else {
      std::vector<ssim2_pair> p(pending);
      for (uint32_t i = 0; i < pending; ++i) p[i] = {slots_r[i].f.srgb, slots_d[i].f.srgb};
      e = ssim2_batch_compute_async(batch, p.data(), pending, stream);
    }
    if (e != SSIM2_OK) return e;
    std::vector<double> sc(pending);
    if ((e = ssim2_batch_get_scores(batch, sc.data(), pending)) != SSIM2_OK) return e;
    out.insert(out.end(), sc.begin(), sc.end());
    pending = 0;
    return SSIM2_OK;
  };

  // leading frames of each source (lib.rs:388-389)
  ssim2_frame fr, fd;
  bool ended = false;
  for (uint32_t i = 0; i < opt->skip + opt->skip_ref && !ended; ++i) ended = pull(ref_src, ref_user, &fr, &st) <= 0;
  for (uint32_t i = 0; i < opt->skip + opt->skip_dis && !ended; ++i) ended = pull(dis_src, dis_user, &fd, &st) <= 0;
  uint32_t decoded = 0;
  while (!ended && st == SSIM2_OK) {
    if (pull(ref_src, ref_user, &fr, &st) <= 0) break;
    if (pull(dis_src, dis_user, &fd, &st) <= 0) break;
    // lib.rs:396-405: keep one pair in `every`, stop after `frames` decoded pairs
    if (opt->every > 1 && decoded != 0 && decoded % opt->every != 0) {
      ++decoded;
      continue;
    }
    if (opt->frames > 0 && decoded >= opt->frames) break;
    ++decoded;
    if ((st = check_frame(fr, width, height, "ref")) != SSIM2_OK) break;
    if ((st = check_frame(fd, width, height, "dis")) != SSIM2_OK) break;
    if (fr.kind != fd.kind || (fr.kind == SSIM2_FRAME_YUV420 && fr.yuv.format != fd.yuv.format)) {
      st = vfail(SSIM2_INVALID_ARGUMENT, "ref and dis frames differ in kind or format");
      break;
    }
    if (ctx) {
      double score = 0.0;
      st = fr.kind == SSIM2_FRAME_SRGB ? ssim2_compute_srgb_async(ctx, &fr.srgb, &fd.srgb, stream)
                                       : ssim2_compute_yuv420_async(ctx, &fr.yuv, &fd.yuv, stream);
      if (st == SSIM2_OK) st = ssim2_get_score(ctx, &score);
      if (st == SSIM2_OK) out.push_back(score);
    } else {
      if (pending && slots_r[0].f.kind != fr.kind) st = flush();
      if (st == SSIM2_OK) st = copy_frame(slots_r[pending], fr, s);
      if (st == SSIM2_OK) st = copy_frame(slots_d[pending], fd, s);
      // the sources may reuse their buffers once the copies are done
      if (st == SSIM2_OK && hipStreamSynchronize(s) != hipSuccess) st = vfail(SSIM2_HIP_ERROR, "frame copy failed");
      if (st == SSIM2_OK && ++pending == nb) st = flush();
    }
  }
  if (st == SSIM2_OK && batch) st = flush();
  if (ctx) ssim2_destroy(ctx);
  if (batch) ssim2_batch_destroy(batch);
  for (auto& sl : slots_r)
    if (sl.mem) (void)hipFree(sl.mem);
  for (auto& sl : slots_d)
    if (sl.mem) (void)hipFree(sl.mem);
  if (st != SSIM2_OK) return st;
  *frame_count = out.size();
  if (scores) std::memcpy(scores, out.data(), std::min(capacity, out.size()) * sizeof(double));
  if (stats && !out.empty()) return ssim2_stats_compute(out.data(), out.size(), stats);
  return SSIM2_OK;
}

}  // extern "C"
