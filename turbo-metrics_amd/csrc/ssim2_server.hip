// Serving runtime for the fast form (include/ssim2hip.h, ssim2_server_*).
//
// Requests (pairs of device-resident sRGB images) are submitted from any
// thread. A scheduler thread keeps `slots` batch engines, each on its own
// stream: whenever a slot is idle it takes the oldest waiting requests that
// fit the slot's capacity (pairs, pixels) and the HBM budget, and launches
// them; requests that arrive while other batches run join the next batch, so
// their work overlaps on the device (continuous batching). Completion is
// detected per slot with hipEventQuery and time-stamped on the host clock.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ssim2hip.h"
#include "ssim2_batch_internal.h"
#include "ssim2_internal.h"
#include "ssim2_plan.h"

using namespace ssim2;

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

ssim2_status sfail(ssim2_status st, const std::string& m) {
  set_last_error(m);
  return st;
}

// Device bytes a request occupies while in flight in the fast form: the XYB
// planes of every scale in the storage type plus its partial-sum records.
uint64_t request_bytes(uint32_t w, uint32_t h, uint32_t storage) {
  int sw[kMaxScales], sh[kMaxScales];
  const int n = scale_dims((int)w, (int)h, sw, sh);
  const uint64_t sb = storage == SSIM2_STORAGE_F32 ? 4 : 2;
  uint64_t b = 0;
  for (int s = 0; s < n; ++s) {
    const uint64_t pitch = (uint64_t(sw[s]) + 15) / 16 * 16;
    b += 6 * pitch * sh[s] * sb + 256;
    b += 3 * (uint64_t(sw[s] + 15) / 16) * ((sh[s] + 15) / 16) * kNumSums * sizeof(double);
  }
  return b + kResultDoubles * sizeof(double);
}

struct Request {
  ssim2_pair pair;
  uint32_t ref_id = 0;  // a resident reference in place of pair.ref (0: none)
  uint64_t id;
  double t_submit;
  uint64_t px;
  uint64_t bytes;
};

struct Slot {
  ssim2_batch* batch = nullptr;
  hipStream_t stream = nullptr;
  std::vector<Request> reqs;
  std::vector<ssim2_pair> pairs;
  std::vector<uint32_t> ref_ids;
  std::vector<double> scores;
  uint64_t bytes = 0;
  bool busy = false;
  double t_launch = 0;
};

}  // namespace

struct ssim2_server {
  ssim2_server_options opt{};
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request> waiting;
  std::vector<ssim2_result> done;
  uint64_t in_flight_bytes = 0;
  std::atomic<bool> stop{false};
  std::thread worker;
  int device = 0;
  // statistics
  uint64_t n_submitted = 0, n_done = 0, n_refused = 0, n_batches = 0, px_done = 0;
  double busy_since = 0, first_submit = 0, last_done = 0;
  std::string worker_error;
  // Resident references: computed by slot 0's engine, shared with the
  // others. ref_mu serialises every batch-engine call of the worker with
  // reference registration and removal.
  std::mutex ref_mu;
  hipStream_t reg_stream = nullptr;
  struct RefInfo {
    uint32_t id;
    uint32_t w, h;
  };
  std::vector<RefInfo> refs;

  // Whether `id` names a registered reference; call with ref_mu held.
  bool has_ref(uint32_t id) const {
    for (const RefInfo& r : refs)
      if (r.id == id) return true;
    return false;
  }
  void run();
};

void ssim2_server::run() {
  (void)hipSetDevice(device);
  while (!stop.load()) {
    bool progressed = false;
    // 1. retire finished batches
    for (Slot& s : slots) {
      if (!s.busy) continue;
      int fin = 0;
      ssim2_status st;
      {
        std::lock_guard<std::mutex> rk(ref_mu);
        if (ssim2_batch_query(s.batch, &fin) != SSIM2_OK) fin = 1;
        if (!fin) continue;
        s.scores.resize(s.reqs.size());
        st = ssim2_batch_get_scores(s.batch, s.scores.data(), (uint32_t)s.scores.size());
      }
      const double t = now_s();
      std::lock_guard<std::mutex> lk(mu);
      for (size_t i = 0; i < s.reqs.size(); ++i) {
        ssim2_result r{};
        r.request_id = s.reqs[i].id;
        r.score = st == SSIM2_OK ? s.scores[i] : 0.0;
        r.status = st;
        r.submit_time_s = s.reqs[i].t_submit;
        r.done_time_s = t;
        done.push_back(r);
        px_done += s.reqs[i].px;
      }
      n_done += s.reqs.size();
      last_done = t;
      in_flight_bytes -= s.bytes;
      s.busy = false;
      s.reqs.clear();
      progressed = true;
      cv.notify_all();
    }
    // 2. launch waiting requests on idle slots
    for (Slot& s : slots) {
      if (s.busy) continue;
      {
        std::lock_guard<std::mutex> lk(mu);
        uint64_t px = 0, bytes = 0;
        while (!waiting.empty() && s.reqs.size() < opt.max_pairs_per_batch) {
          const Request& q = waiting.front();
          // FIFO admission: stop at the first request that does not fit the
          // slot's pixel capacity or the HBM budget left by the batches in
          // flight (it waits for the next slot).
          if (px + q.px > opt.max_pixels_per_batch) break;
          if (in_flight_bytes + bytes + q.bytes > opt.hbm_budget) break;
          px += q.px;
          bytes += q.bytes;
          s.reqs.push_back(q);
          waiting.pop_front();
        }
        if (s.reqs.empty()) continue;
        s.bytes = bytes;
        in_flight_bytes += bytes;
      }
      s.t_launch = now_s();
      ssim2_status st = SSIM2_OK;
      std::vector<Request> stale;
      {
        std::lock_guard<std::mutex> rk(ref_mu);
        // A resident reference removed after its requests were queued: those
        // requests fail on their own, the rest of the batch runs as it would
        // have without them. Holding ref_mu from this check to the launch
        // keeps the reference from being removed in between.
        size_t k = 0;
        for (size_t i = 0; i < s.reqs.size(); ++i) {
          if (s.reqs[i].ref_id && !has_ref(s.reqs[i].ref_id)) {
            stale.push_back(s.reqs[i]);
            continue;
          }
          s.reqs[k++] = s.reqs[i];
        }
        s.reqs.resize(k);
        s.pairs.resize(k);
        s.ref_ids.resize(k);
        for (size_t i = 0; i < k; ++i) {
          s.pairs[i] = s.reqs[i].pair;
          s.ref_ids[i] = s.reqs[i].ref_id;
        }
        if (k)
          st = batch_compute_mixed_async(s.batch, s.pairs.data(), s.ref_ids.data(), (uint32_t)k, s.stream);
      }
      if (!stale.empty()) {
        std::lock_guard<std::mutex> lk(mu);
        const double t = now_s();
        for (const Request& q : stale) {
          ssim2_result r{};
          r.request_id = q.id;
          r.status = SSIM2_NO_REFERENCE;
          r.submit_time_s = q.t_submit;
          r.done_time_s = t;
          done.push_back(r);
          s.bytes -= q.bytes;
          in_flight_bytes -= q.bytes;
        }
        n_done += stale.size();
        progressed = true;
        cv.notify_all();
      }
      if (s.reqs.empty()) continue;
      if (st != SSIM2_OK) {
        std::lock_guard<std::mutex> lk(mu);
        worker_error = ssim2_last_error();
        for (const Request& q : s.reqs) {
          ssim2_result r{};
          r.request_id = q.id;
          r.status = st;
          r.submit_time_s = q.t_submit;
          r.done_time_s = now_s();
          done.push_back(r);
        }
        n_done += s.reqs.size();
        in_flight_bytes -= s.bytes;
        s.reqs.clear();
        cv.notify_all();
        continue;
      }
      s.busy = true;
      ++n_batches;
      progressed = true;
    }
    if (!progressed) {
      std::unique_lock<std::mutex> lk(mu);
      bool any_busy = false;
      for (const Slot& s : slots) any_busy |= s.busy;
      if (any_busy)
        cv.wait_for(lk, std::chrono::microseconds(20));
      else
        cv.wait_for(lk, std::chrono::milliseconds(2), [&] { return stop.load() || !waiting.empty(); });
    }
  }
}

extern "C" {

ssim2_status ssim2_server_create(const ssim2_server_options* o, ssim2_server** out) {
  if (!o || !out) return sfail(SSIM2_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (o->struct_size != sizeof(ssim2_server_options))
    return sfail(SSIM2_INVALID_ARGUMENT, "ssim2_server_options.struct_size mismatch");
  if (o->slots == 0 || o->slots > 16 || o->max_pairs_per_batch == 0 || o->max_pixels_per_batch < 64 ||
      o->hbm_budget == 0)
    return sfail(SSIM2_INVALID_ARGUMENT, "slots in [1,16], non-zero batch limits and HBM budget required");
  ssim2_server* sv = new (std::nothrow) ssim2_server();
  if (!sv) return sfail(SSIM2_OOM, "host allocation failed");
  sv->opt = *o;
  (void)hipGetDevice(&sv->device);
  sv->slots.resize(o->slots);
  for (Slot& s : sv->slots) {
    ssim2_batch_options bo{};
    bo.struct_size = sizeof bo;
    bo.storage = o->storage;
    bo.max_pairs = o->max_pairs_per_batch;
    bo.max_pixels = o->max_pixels_per_batch;
    bo.memory_budget = 0;  // engine capacity is static; hbm_budget governs admission
    ssim2_status st = ssim2_batch_create(&bo, &s.batch);
    hipError_t e = hipSuccess;
    if (st == SSIM2_OK) e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (st != SSIM2_OK || e != hipSuccess) {
      const std::string m = st != SSIM2_OK ? std::string(ssim2_last_error()) : hipGetErrorString(e);
      for (Slot& t : sv->slots) {
        if (t.batch) ssim2_batch_destroy(t.batch);
        if (t.stream) (void)hipStreamDestroy(t.stream);
      }
      delete sv;
      return sfail(st != SSIM2_OK ? st : SSIM2_HIP_ERROR, "server slot: " + m);
    }
    // serving batches are small: a smaller grid unless the caller sets one;
    // tile shape 0 with grid 0 keeps the engine's default shape
    uint32_t cfg = o->tile_cfg, mult = 0;
    if (cfg == 0 && o->grid_mult == 0) ssim2_batch_get_tiling(s.batch, &cfg, &mult);
    ssim2_batch_set_tiling(s.batch, cfg, o->grid_mult ? o->grid_mult : 8);
  }
  if (hipStreamCreateWithFlags(&sv->reg_stream, hipStreamNonBlocking) != hipSuccess) {
    for (Slot& t : sv->slots) {
      ssim2_batch_destroy(t.batch);
      (void)hipStreamDestroy(t.stream);
    }
    delete sv;
    return sfail(SSIM2_HIP_ERROR, "server: stream creation failed");
  }
  sv->worker = std::thread([sv] { sv->run(); });
  *out = sv;
  return SSIM2_OK;
}

ssim2_status ssim2_server_add_reference(ssim2_server* sv, const ssim2_image* ref, uint32_t* ref_id) {
  if (!sv || !ref || !ref_id || !ref->ptr) return sfail(SSIM2_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> rk(sv->ref_mu);
  uint32_t id = 0;
  ssim2_status st = ssim2_batch_add_reference(sv->slots[0].batch, ref, sv->reg_stream, &id);
  if (st != SSIM2_OK) return st;
  for (size_t i = 1; i < sv->slots.size() && st == SSIM2_OK; ++i)
    st = batch_share_reference(sv->slots[i].batch, sv->slots[0].batch, id);
  if (st != SSIM2_OK) {
    const std::string m = ssim2_last_error();
    for (size_t i = sv->slots.size(); i-- > 0;) (void)ssim2_batch_remove_reference(sv->slots[i].batch, id);
    return sfail(st, "server reference: " + m);
  }
  sv->refs.push_back({id, ref->width, ref->height});
  *ref_id = id;
  return SSIM2_OK;
}

ssim2_status ssim2_server_remove_reference(ssim2_server* sv, uint32_t ref_id) {
  if (!sv) return sfail(SSIM2_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> rk(sv->ref_mu);
  size_t k = 0;
  while (k < sv->refs.size() && sv->refs[k].id != ref_id) ++k;
  if (k == sv->refs.size()) return sfail(SSIM2_NO_REFERENCE, "unknown reference id");
  // the sharing engines first (each waits for its own batches in flight),
  // then the owner, which frees the planes
  ssim2_status st = SSIM2_OK;
  for (size_t i = sv->slots.size(); i-- > 0;) {
    const ssim2_status e = ssim2_batch_remove_reference(sv->slots[i].batch, ref_id);
    if (e != SSIM2_OK && st == SSIM2_OK) st = e;
  }
  sv->refs.erase(sv->refs.begin() + (long)k);
  return st;
}

ssim2_status ssim2_server_submit(ssim2_server* sv, const ssim2_pair* pair, uint64_t request_id) {
  if (!sv || !pair || !pair->ref.ptr || !pair->dis.ptr) return sfail(SSIM2_INVALID_ARGUMENT, "null argument");
  if (pair->ref.width != pair->dis.width || pair->ref.height != pair->dis.height)
    return sfail(SSIM2_DIM_MISMATCH, "reference and distorted sizes differ");
  if (!size_supported(pair->ref.width, pair->ref.height)) return sfail(SSIM2_INVALID_SIZE, "unsupported size");
  Request q;
  q.pair = *pair;
  q.id = request_id;
  q.px = uint64_t(pair->ref.width) * pair->ref.height;
  q.bytes = request_bytes(pair->ref.width, pair->ref.height, sv->opt.storage);
  if (q.bytes > sv->opt.hbm_budget || q.px > sv->opt.max_pixels_per_batch) {
    std::lock_guard<std::mutex> lk(sv->mu);
    ++sv->n_refused;
    return sfail(SSIM2_ADMISSION_REFUSED, "request needs " + std::to_string(q.bytes) + " B / " +
                                              std::to_string(q.px) + " px, above the server limits");
  }
  q.t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->n_submitted == 0) sv->first_submit = q.t_submit;
    ++sv->n_submitted;
    sv->waiting.push_back(q);
  }
  sv->cv.notify_all();
  return SSIM2_OK;
}

ssim2_status ssim2_server_submit_ref(ssim2_server* sv, uint32_t ref_id, const ssim2_image* dis,
                                     uint64_t request_id) {
  if (!sv || !dis || !dis->ptr) return sfail(SSIM2_INVALID_ARGUMENT, "null argument");
  uint32_t w = 0, h = 0;
  {
    std::lock_guard<std::mutex> rk(sv->ref_mu);
    size_t k = 0;
    while (k < sv->refs.size() && sv->refs[k].id != ref_id) ++k;
    if (k == sv->refs.size()) return sfail(SSIM2_NO_REFERENCE, "unknown reference id");
    w = sv->refs[k].w;
    h = sv->refs[k].h;
  }
  if (dis->width != w || dis->height != h) return sfail(SSIM2_DIM_MISMATCH, "reference and distorted sizes differ");
  if (dis->pitch_bytes < size_t(w) * 3) return sfail(SSIM2_INVALID_ARGUMENT, "pitch smaller than a row");
  Request q;
  q.pair.ref = *dis;  // not read: the resident reference replaces it
  q.pair.dis = *dis;
  q.ref_id = ref_id;
  q.id = request_id;
  q.px = uint64_t(w) * h;
  q.bytes = request_bytes(w, h, sv->opt.storage);
  if (q.bytes > sv->opt.hbm_budget || q.px > sv->opt.max_pixels_per_batch) {
    std::lock_guard<std::mutex> lk(sv->mu);
    ++sv->n_refused;
    return sfail(SSIM2_ADMISSION_REFUSED, "request needs " + std::to_string(q.bytes) + " B / " +
                                              std::to_string(q.px) + " px, above the server limits");
  }
  q.t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->n_submitted == 0) sv->first_submit = q.t_submit;
    ++sv->n_submitted;
    sv->waiting.push_back(q);
  }
  sv->cv.notify_all();
  return SSIM2_OK;
}

ssim2_status ssim2_server_poll(ssim2_server* sv, ssim2_result* out, uint32_t capacity, uint32_t* n) {
  if (!sv || !n || (!out && capacity)) return sfail(SSIM2_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  const uint32_t k = (uint32_t)std::min<size_t>(capacity, sv->done.size());
  for (uint32_t i = 0; i < k; ++i) out[i] = sv->done[i];
  sv->done.erase(sv->done.begin(), sv->done.begin() + k);
  *n = k;
  return SSIM2_OK;
}

ssim2_status ssim2_server_drain(ssim2_server* sv, double timeout_s) {
  if (!sv) return sfail(SSIM2_INVALID_ARGUMENT, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  const bool ok = sv->cv.wait_for(lk, std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9),
                                  [&] { return sv->n_done >= sv->n_submitted; });
  return ok ? SSIM2_OK : sfail(SSIM2_NOT_READY, "drain timed out");
}

ssim2_status ssim2_server_stats(ssim2_server* sv, ssim2_server_stats_t* st) {
  if (!sv || !st) return sfail(SSIM2_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  std::memset(st, 0, sizeof *st);
  st->struct_size = sizeof *st;
  st->submitted = sv->n_submitted;
  st->completed = sv->n_done;
  st->refused = sv->n_refused;
  st->batches = sv->n_batches;
  st->pixels_completed = sv->px_done;
  st->waiting = sv->waiting.size();
  st->in_flight_bytes = sv->in_flight_bytes;
  st->first_submit_s = sv->first_submit;
  st->last_done_s = sv->last_done;
  return SSIM2_OK;
}

ssim2_status ssim2_server_destroy(ssim2_server* sv) {
  if (!sv) return SSIM2_OK;
  sv->stop.store(true);
  sv->cv.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  // sharing engines first: the owner (slot 0) frees the planes
  for (size_t i = sv->slots.size(); i-- > 0;) {
    Slot& s = sv->slots[i];
    if (s.batch) ssim2_batch_destroy(s.batch);
    if (s.stream) (void)hipStreamDestroy(s.stream);
  }
  if (sv->reg_stream) (void)hipStreamDestroy(sv->reg_stream);
  delete sv;
  return SSIM2_OK;
}

}  // extern "C"
