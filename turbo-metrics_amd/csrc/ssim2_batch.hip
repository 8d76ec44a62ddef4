// Batch engine of the fast form (include/ssim2hip.h, ssim2_batch_*): many
// uneven sRGB pairs per launch. Host code: planning, staging, the captured
// hipGraph of the three kernels, resident reference pyramids, and the
// tile-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/ssim2hip.h"
#include "ssim2_batch_internal.h"
#include "ssim2_internal.h"
#include "ssim2_plan.h"

using namespace ssim2;

namespace {

ssim2_status bfail(ssim2_status st, const std::string& msg) {
  set_last_error(msg);
  return st;
}

#define BHIP(expr)                                                                            \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) return bfail(SSIM2_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

size_t storage_bytes(int storage) { return storage == kStorageF32 ? 4 : 2; }

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Plan one pair of the fast form. Returns false when the size is unsupported.
struct FastPlan {
  PairDesc d{};
  int npyr = 0;
  int ntiles = 0;
  size_t arena = 0;  // bytes of XYB planes (all scales)
  int nparts = 0;
};

bool plan_fast(int w, int h, int storage, int cfg, size_t arena_off, int part_base, FastPlan* out) {
  if (!size_supported(w, h)) return false;
  FastPlan& p = *out;
  p = FastPlan{};
  PairDesc& d = p.d;
  d.w = w;
  d.h = h;
  int sw[kMaxScales], sh[kMaxScales];
  d.nscales = scale_dims(w, h, sw, sh);
  const int TX = kTileCfgW[cfg], TY = kTileCfgH[cfg];
  const size_t sb = storage_bytes(storage);
  size_t off = arena_off;
  int part = part_base, tile = 0;
  for (int s = 0; s < d.nscales; ++s) {
    ScaleDesc& S = d.sc[s];
    S.w = sw[s];
    S.h = sh[s];
    S.pitch = (sw[s] + 15) / 16 * 16;
    S.plane = (long long)S.pitch * S.h;
    S.xyb = (long long)off;  // every scale, scale 0 included
    off += align256(6 * (size_t)S.plane * sb);
    S.tiles_x = (sw[s] + TX - 1) / TX;
    S.colblocks = S.tiles_x * ((sh[s] + TY - 1) / TY);
    S.tile0 = tile;
    tile += S.colblocks;
    S.part = part;
    part += 3 * S.colblocks;
  }
  p.npyr = ((w + kPyrTile - 1) / kPyrTile) * ((h + kPyrTile - 1) / kPyrTile);
  p.ntiles = tile;
  p.arena = off - arena_off;
  p.nparts = part - part_base;
  return true;
}

}  // namespace

struct ssim2_batch {
  int storage = kStorageF32;
  int tile_cfg = 0;
  int grid_mult = 64;  // autotuned batches of a few hundred pairs pick 64..256
  uint32_t max_pairs = 0;
  uint64_t max_pixels = 0;
  size_t arena_cap = 0, part_cap = 0;
  unsigned char* arena = nullptr;
  double* partials = nullptr;
  // Two staging slots: the host stages batch k+1 while batch k runs; batch
  // k-1's slot is the one reused, so staging waits on it at most.
  struct Slot {
    double* results = nullptr;
    PairDesc* d_desc = nullptr;
    int* d_index = nullptr;  // [meta(4) | pyr prefix (max+1) | tile prefix (max+1)]
    PairDesc* h_desc = nullptr;
    int* h_index = nullptr;
    double* h_results = nullptr;
    hipEvent_t done = nullptr;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;        // kept for the grid updates of two nodes
    hipGraphNode_t pyr_node = nullptr;     // the pyramid launch in graph_src
    hipGraphNode_t tile_node = nullptr;    // the tile launch
    hipKernelNodeParams pyr_params = {}, tile_params = {};
    int graph_cfg = -1, graph_mult = -1;
    int pyr_grid = 0, tile_grid = 0;       // grids of the instantiated graph
    bool pending = false;
    int n = 0;
    int npyr = 0, ntiles = 0;              // items of the staged batch
  } slot[2];
  int cur = 0;     // slot of the last submitted batch
  int staged = 1;  // slot filled by the last stage()
  hipStream_t cap_stream = nullptr;
  int cus = 256;
  size_t bytes = 0;
  std::string tune_report;
  // Resident reference pyramids (shifted XYB of every scale, storage type),
  // each its own allocation, named by id.
  struct Ref {
    uint32_t id = 0;
    int w = 0, h = 0;
    unsigned char* planes = nullptr;
    long long off[kMaxScales] = {};  // byte offset of each scale's X, Y, B planes
    size_t bytes = 0;
    bool owned = true;               // false: another engine's planes, shared
  };
  std::vector<Ref> refs;
  uint32_t next_ref = 1;
  unsigned char* d_reg = nullptr;  // registration launch: one PairDesc + its item index
};

namespace {

int* meta_ptr(int* base) { return base; }
int* pyr_prefix(ssim2_batch*, int* base) { return base + 4; }
int* tile_prefix(ssim2_batch* b, int* base) { return base + 4 + (b->max_pairs + 1); }

void release(ssim2_batch* b) {
  for (ssim2_batch::Slot& S : b->slot) {
    if (S.graph) (void)hipGraphExecDestroy(S.graph);
    if (S.graph_src) (void)hipGraphDestroy(S.graph_src);
    if (S.done) (void)hipEventDestroy(S.done);
    if (S.results) (void)hipFree(S.results);
    if (S.d_desc) (void)hipFree(S.d_desc);
    if (S.d_index) (void)hipFree(S.d_index);
    if (S.h_desc) (void)hipHostFree(S.h_desc);
    if (S.h_index) (void)hipHostFree(S.h_index);
    if (S.h_results) (void)hipHostFree(S.h_results);
  }
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  for (ssim2_batch::Ref& r : b->refs)
    if (r.planes && r.owned) (void)hipFree(r.planes);
  b->refs.clear();
  if (b->d_reg) (void)hipFree(b->d_reg);
  if (b->arena) (void)hipFree(b->arena);
  if (b->partials) (void)hipFree(b->partials);
}

// Tile-kernel grid: mult x CUs blocks, or one block per tile for mult 0.
int tile_grid(const ssim2_batch* b, const ssim2_batch::Slot& S, int mult) {
  return mult > 0 ? b->cus * mult : (S.ntiles < 1 ? 1 : S.ntiles);
}

ssim2_status enqueue_kernels(ssim2_batch* b, const ssim2_batch::Slot& S, hipStream_t s, int cfg, int mult) {
  const int* meta = meta_ptr(S.d_index);
  BHIP(launch_fast_pyramid(b->storage, S.d_desc, pyr_prefix(b, S.d_index), meta, S.npyr, s));
  BHIP(launch_fast_tile(b->storage, cfg, S.d_desc, tile_prefix(b, S.d_index), meta, tile_grid(b, S, mult), s));
  BHIP(launch_fast_finalize(S.d_desc, meta, b->cus * 4, (int)b->max_pairs, s));
  return SSIM2_OK;
}

// Capture the three launches once per (tile shape, grid size). All sizes
// come from device memory, so the same graph runs every batch; only the
// grids sized by the batch (one block per tile) are set per batch, on their
// nodes.
ssim2_status ensure_graph(ssim2_batch* b, ssim2_batch::Slot& S) {
  if (!(S.graph && S.graph_cfg == b->tile_cfg && S.graph_mult == b->grid_mult)) {
    if (S.graph) {
      (void)hipGraphExecDestroy(S.graph);
      S.graph = nullptr;
    }
    if (S.graph_src) {
      (void)hipGraphDestroy(S.graph_src);
      S.graph_src = nullptr;
    }
    hipGraph_t g = nullptr;
    BHIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    ssim2_status st = enqueue_kernels(b, S, b->cap_stream, b->tile_cfg, b->grid_mult);
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (st != SSIM2_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return bfail(SSIM2_HIP_ERROR, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    // the pyramid's and the tile kernel's nodes
    size_t nn = 0;
    S.pyr_node = S.tile_node = nullptr;
    if ((e = hipGraphGetNodes(g, nullptr, &nn)) == hipSuccess) {
      std::vector<hipGraphNode_t> nodes(nn);
      e = hipGraphGetNodes(g, nodes.data(), &nn);
      const void* pk = fast_pyramid_kernel(b->storage);
      const void* tk = fast_tile_kernel(b->storage, b->tile_cfg);
      for (size_t i = 0; e == hipSuccess && i < nn; ++i) {
        hipGraphNodeType t;
        if ((e = hipGraphNodeGetType(nodes[i], &t)) != hipSuccess || t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams kp;
        if ((e = hipGraphKernelNodeGetParams(nodes[i], &kp)) != hipSuccess) break;
        if (kp.func == pk) {
          S.pyr_node = nodes[i];
          S.pyr_params = kp;
        } else if (kp.func == tk) {
          S.tile_node = nodes[i];
          S.tile_params = kp;
        }
      }
    }
    if (e == hipSuccess && (!S.pyr_node || !S.tile_node)) e = hipErrorNotFound;
    if (e == hipSuccess) e = hipGraphInstantiate(&S.graph, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      S.graph = nullptr;
      return bfail(SSIM2_HIP_ERROR, std::string("batch graph: ") + hipGetErrorString(e));
    }
    S.graph_src = g;
    S.graph_cfg = b->tile_cfg;
    S.graph_mult = b->grid_mult;
    S.pyr_grid = (int)S.pyr_params.gridDim.x;
    S.tile_grid = (int)S.tile_params.gridDim.x;
  }
  const int grid = S.npyr < 1 ? 1 : S.npyr;
  if (grid != S.pyr_grid) {
    hipKernelNodeParams kp = S.pyr_params;
    kp.gridDim = dim3(grid);
    BHIP(hipGraphExecKernelNodeSetParams(S.graph, S.pyr_node, &kp));
    S.pyr_grid = grid;
  }
  const int tgrid = tile_grid(b, S, b->grid_mult);
  if (tgrid != S.tile_grid) {
    hipKernelNodeParams kp = S.tile_params;
    kp.gridDim = dim3(tgrid);
    BHIP(hipGraphExecKernelNodeSetParams(S.graph, S.tile_node, &kp));
    S.tile_grid = tgrid;
  }
  return SSIM2_OK;
}

const ssim2_batch::Ref* find_ref(const ssim2_batch* b, uint32_t id) {
  for (const ssim2_batch::Ref& r : b->refs)
    if (r.id == id) return &r;
  return nullptr;
}

// Exactly one of pairs, ypairs, rpairs is non-null. With pairs, ref_ids
// (optional) names a resident reference per pair (0: the pair's own ref).
ssim2_status stage(ssim2_batch* b, const ssim2_pair* pairs, uint32_t n, int cfg,
                   const ssim2_yuv_pair* ypairs = nullptr, const ssim2_ref_pair* rpairs = nullptr,
                   const uint32_t* ref_ids = nullptr) {
  if (n == 0) return bfail(SSIM2_BATCH_SIZE, "empty batch");
  if (n > b->max_pairs)
    return bfail(SSIM2_BATCH_SIZE, std::to_string(n) + " pairs exceed the batch capacity of " +
                                       std::to_string(b->max_pairs));
  // the slot not used by the last submitted batch; its own previous batch
  // (two submissions ago) must have finished with its buffers
  const int si = b->cur ^ 1;
  ssim2_batch::Slot& SL = b->slot[si];
  if (SL.pending) BHIP(hipEventSynchronize(SL.done));
  SL.pending = false;
  b->staged = si;
  size_t arena = 0;
  int parts = 0, pyr = 0, tiles = 0;
  uint64_t px = 0;
  int* pp = pyr_prefix(b, SL.h_index);
  int* tp = tile_prefix(b, SL.h_index);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t qw, qh;
    const ssim2_batch::Ref* ref = nullptr;
    ssim2_ref_pair rq{};
    const bool by_ref = rpairs || (ref_ids && ref_ids[i]);
    if (by_ref) {
      if (rpairs) {
        rq = rpairs[i];
      } else {
        rq.ref_id = ref_ids[i];
        rq.dis = pairs[i].dis;
      }
      const ssim2_ref_pair& q = rq;
      ref = find_ref(b, q.ref_id);
      if (!ref) return bfail(SSIM2_NO_REFERENCE, "pair " + std::to_string(i) + ": unknown reference id");
      qw = q.dis.width;
      qh = q.dis.height;
      if (!q.dis.ptr) return bfail(SSIM2_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": null image");
      if (int(qw) != ref->w || int(qh) != ref->h)
        return bfail(SSIM2_DIM_MISMATCH, "pair " + std::to_string(i) + ": reference and distorted sizes differ");
      if (q.dis.pitch_bytes < size_t(qw) * 3)
        return bfail(SSIM2_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": pitch smaller than a row");
    } else if (ypairs) {
      const ssim2_yuv_pair& y = ypairs[i];
      qw = y.ref.width;
      qh = y.ref.height;
      if (y.dis.width != qw || y.dis.height != qh)
        return bfail(SSIM2_DIM_MISMATCH, "pair " + std::to_string(i) + ": reference and distorted sizes differ");
      if (y.ref.format != y.dis.format || y.ref.format > SSIM2_P016 || !y.ref.y || !y.dis.y || !y.ref.c0 ||
          !y.dis.c0 || (y.ref.format == SSIM2_YUV420P10 && (!y.ref.c1 || !y.dis.c1)) ||
          y.ref.y_pitch_bytes < 2 * size_t(qw) || y.dis.y_pitch_bytes < 2 * size_t(qw) ||
          y.ref.c_pitch_bytes < (y.ref.format == SSIM2_P016 ? 4 : 2) * ((size_t(qw) + 1) / 2) ||
          y.dis.c_pitch_bytes < (y.dis.format == SSIM2_P016 ? 4 : 2) * ((size_t(qw) + 1) / 2) ||
          ((y.ref.y_pitch_bytes | y.dis.y_pitch_bytes | y.ref.c_pitch_bytes | y.dis.c_pitch_bytes) & 1))
        return bfail(SSIM2_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": bad YUV planes");
    } else {
      const ssim2_pair& q = pairs[i];
      qw = q.ref.width;
      qh = q.ref.height;
      if (!q.ref.ptr || !q.dis.ptr) return bfail(SSIM2_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": null image");
      if (q.ref.width != q.dis.width || q.ref.height != q.dis.height)
        return bfail(SSIM2_DIM_MISMATCH, "pair " + std::to_string(i) + ": reference and distorted sizes differ");
      if (q.ref.pitch_bytes < size_t(q.ref.width) * 3 || q.dis.pitch_bytes < size_t(q.dis.width) * 3)
        return bfail(SSIM2_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": pitch smaller than a row");
    }
    FastPlan fp;
    if (!plan_fast((int)qw, (int)qh, b->storage, cfg, arena, parts, &fp))
      return bfail(SSIM2_INVALID_SIZE, "pair " + std::to_string(i) + ": size " + std::to_string(qw) + "x" +
                                           std::to_string(qh) + " unsupported");
    px += uint64_t(qw) * qh;
    PairDesc d = fp.d;
    if (by_ref) {
      // image 0 is the resident pyramid: the pyramid kernel leaves it alone
      // (its input pointers name the distorted image, never read for it)
      const ssim2_ref_pair& q = rq;
      d.in_ref = d.in_dis = static_cast<const unsigned char*>(q.dis.ptr);
      d.in_ref_pitch = d.in_dis_pitch = (long long)q.dis.pitch_bytes;
      d.in_kind = kInputSrgbU8;
      d.ref_planes = ref->planes;
      d.skip_img = 1;
      for (int sc = 0; sc < d.nscales; ++sc) d.sc[sc].cache = ref->off[sc];
    } else if (ypairs) {
      const ssim2_yuv_pair& y = ypairs[i];
      d.in_ref = static_cast<const unsigned char*>(y.ref.y);
      d.in_dis = static_cast<const unsigned char*>(y.dis.y);
      d.in_ref_pitch = (long long)y.ref.y_pitch_bytes;
      d.in_dis_pitch = (long long)y.dis.y_pitch_bytes;
      d.in_ref_c0 = static_cast<const unsigned char*>(y.ref.c0);
      d.in_ref_c1 = static_cast<const unsigned char*>(y.ref.c1);
      d.in_dis_c0 = static_cast<const unsigned char*>(y.dis.c0);
      d.in_dis_c1 = static_cast<const unsigned char*>(y.dis.c1);
      d.in_ref_cpitch = (long long)y.ref.c_pitch_bytes;
      d.in_dis_cpitch = (long long)y.dis.c_pitch_bytes;
      d.in_kind = y.ref.format == SSIM2_P016 ? kInputYuvP016 : kInputYuvP10;
    } else {
      const ssim2_pair& q = pairs[i];
      d.in_ref = static_cast<const unsigned char*>(q.ref.ptr);
      d.in_dis = static_cast<const unsigned char*>(q.dis.ptr);
      d.in_ref_pitch = (long long)q.ref.pitch_bytes;
      d.in_dis_pitch = (long long)q.dis.pitch_bytes;
      d.in_kind = kInputSrgbU8;
    }
    d.arena = reinterpret_cast<float*>(b->arena);
    d.partials = b->partials;
    d.result = SL.results + size_t(i) * kResultDoubles;
    SL.h_desc[i] = d;
    pp[i] = pyr;
    tp[i] = tiles;
    pyr += fp.npyr;
    tiles += fp.ntiles;
    arena += fp.arena;
    parts += fp.nparts;
  }
  if (px > b->max_pixels || arena > b->arena_cap || size_t(parts) > b->part_cap)
    return bfail(SSIM2_BATCH_SIZE, "batch of " + std::to_string(px) + " px exceeds the capacity of " +
                                       std::to_string(b->max_pixels) + " px");
  pp[n] = pyr;
  SL.npyr = pyr;
  SL.ntiles = tiles;
  tp[n] = tiles;
  int* meta = meta_ptr(SL.h_index);
  meta[0] = (int)n;
  meta[1] = pyr;
  meta[2] = tiles;
  meta[3] = 0;
  return SSIM2_OK;
}

ssim2_status submit(ssim2_batch* b, uint32_t n, hipStream_t s, bool use_graph, int cfg, int mult) {
  const size_t index_ints = 4 + 2 * size_t(b->max_pairs + 1);
  ssim2_batch::Slot& S = b->slot[b->staged];
  ssim2_batch::Slot& prev = b->slot[b->cur];
  // The arena and partial sums are shared by the slots: a batch starts after
  // the previous one (a no-op on the same stream, an ordering across streams).
  if (b->cur != b->staged && prev.pending) BHIP(hipStreamWaitEvent(s, prev.done, 0));
  BHIP(hipMemcpyAsync(S.d_desc, S.h_desc, n * sizeof(PairDesc), hipMemcpyHostToDevice, s));
  BHIP(hipMemcpyAsync(S.d_index, S.h_index, index_ints * sizeof(int), hipMemcpyHostToDevice, s));
  if (use_graph) {
    ssim2_status st = ensure_graph(b, S);
    if (st != SSIM2_OK) return st;
    BHIP(hipGraphLaunch(S.graph, s));
  } else {
    ssim2_status st = enqueue_kernels(b, S, s, cfg, mult);
    if (st != SSIM2_OK) return st;
  }
  BHIP(hipMemcpyAsync(S.h_results, S.results, n * kResultDoubles * sizeof(double), hipMemcpyDeviceToHost, s));
  BHIP(hipEventRecord(S.done, s));
  S.pending = true;
  S.n = (int)n;
  b->cur = b->staged;
  return SSIM2_OK;
}

}  // namespace

extern "C" {

ssim2_status ssim2_batch_create(const ssim2_batch_options* o, ssim2_batch** out) {
  if (!out || !o) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (o->struct_size != sizeof(ssim2_batch_options))
    return bfail(SSIM2_INVALID_ARGUMENT, "ssim2_batch_options.struct_size mismatch");
  if (o->storage == SSIM2_STORAGE_BF16)
    return bfail(SSIM2_INVALID_ARGUMENT,
                 "bfloat16 storage is refused: measured 0.30-0.35 from the FP64 score on 4K frames "
                 "(the bar is 0.05); use SSIM2_STORAGE_F16");
  if (o->storage > SSIM2_STORAGE_BF16) return bfail(SSIM2_INVALID_ARGUMENT, "unknown storage type");
  if (o->max_pairs == 0 || o->max_pairs > (1u << 20) || o->max_pixels < 64)
    return bfail(SSIM2_INVALID_ARGUMENT, "max_pairs must be in [1, 2^20] and max_pixels >= 64");
  ssim2_batch* b = new (std::nothrow) ssim2_batch();
  if (!b) return bfail(SSIM2_OOM, "host allocation failed");
  b->storage = (int)o->storage;
  if (b->storage == kStorageF16) {  // the matrix-core tiles, a few blocks per CU's worth of items each
    b->tile_cfg = 4;
    b->grid_mult = 32;
  }
  b->max_pairs = o->max_pairs;
  b->max_pixels = o->max_pixels;
  const size_t sb = storage_bytes(b->storage);
  // All scales hold at most ~4/3 of the scale-0 pixels (plus the odd-size
  // rounding); pitch rounding adds at most 15 columns per plane row. A batch
  // whose exact need exceeds this capacity is refused with SSIM2_BATCH_SIZE
  // (split it).
  b->arena_cap = 6 * sb * (size_t(o->max_pixels) * 3 / 2 + size_t(o->max_pairs) * 65536) +
                 size_t(256) * kMaxScales * o->max_pairs;
  b->part_cap = size_t(o->max_pixels) * 3 / 128 + size_t(o->max_pairs) * 3 * 2048;
  const size_t index_ints = 4 + 2 * size_t(o->max_pairs + 1);
  b->bytes = b->arena_cap + b->part_cap * kNumSums * sizeof(double) +
             2 * (size_t(o->max_pairs) * (kResultDoubles * sizeof(double) + sizeof(PairDesc)) +
                  index_ints * sizeof(int));
  size_t budget = (size_t)o->memory_budget;
  if (budget == 0) {
    size_t fb = 0, tb = 0;
    if (hipMemGetInfo(&fb, &tb) == hipSuccess) budget = fb;
  }
  if (budget && b->bytes > budget) {
    const std::string m = "batch footprint " + std::to_string(b->bytes) + " B exceeds the budget of " +
                          std::to_string(budget) + " B";
    delete b;
    return bfail(SSIM2_OOM, m);
  }
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) b->cus = v;
  }
  hipError_t e;
  if ((e = hipMalloc(&b->arena, b->arena_cap)) == hipSuccess &&
      (e = hipMalloc(&b->partials, b->part_cap * kNumSums * sizeof(double))) == hipSuccess)
    e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
  for (ssim2_batch::Slot& S : b->slot) {
    if (e != hipSuccess) break;
    if ((e = hipMalloc(&S.results, size_t(o->max_pairs) * kResultDoubles * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&S.d_desc, size_t(o->max_pairs) * sizeof(PairDesc))) != hipSuccess ||
        (e = hipMalloc(&S.d_index, index_ints * sizeof(int))) != hipSuccess ||
        (e = hipHostMalloc(&S.h_desc, size_t(o->max_pairs) * sizeof(PairDesc), hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc(&S.h_index, index_ints * sizeof(int), hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc(&S.h_results, size_t(o->max_pairs) * kResultDoubles * sizeof(double),
                           hipHostMallocDefault)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&S.done, hipEventDisableTiming)) != hipSuccess)
      break;
    std::memset(S.h_index, 0, index_ints * sizeof(int));
  }
  if (e != hipSuccess) {
    release(b);
    delete b;
    (void)hipGetLastError();
    return bfail(SSIM2_OOM, std::string("batch allocation failed: ") + hipGetErrorString(e));
  }
  *out = b;
  return SSIM2_OK;
}

ssim2_status ssim2_batch_destroy(ssim2_batch* b) {
  if (!b) return SSIM2_OK;
  for (ssim2_batch::Slot& S : b->slot)
    if (S.pending) (void)hipEventSynchronize(S.done);
  release(b);
  delete b;
  return SSIM2_OK;
}

ssim2_status ssim2_batch_mem_usage(const ssim2_batch* b, size_t* bytes) {
  if (!b || !bytes) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  *bytes = b->bytes;
  return SSIM2_OK;
}

ssim2_status ssim2_batch_compute_async(ssim2_batch* b, const ssim2_pair* pairs, uint32_t npairs,
                                       ssim2_stream stream) {
  if (!b || (!pairs && npairs)) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_status st = stage(b, pairs, npairs, b->tile_cfg);
  if (st != SSIM2_OK) return st;
  return submit(b, npairs, (hipStream_t)stream, true, b->tile_cfg, b->grid_mult);
}

ssim2_status ssim2_batch_add_reference(ssim2_batch* b, const ssim2_image* ref, ssim2_stream stream,
                                       uint32_t* ref_id) {
  if (!b || !ref || !ref_id || !ref->ptr) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  if (ref->pitch_bytes < size_t(ref->width) * 3) return bfail(SSIM2_INVALID_ARGUMENT, "pitch smaller than a row");
  FastPlan fp;
  if (!plan_fast((int)ref->width, (int)ref->height, b->storage, b->tile_cfg, 0, 0, &fp))
    return bfail(SSIM2_INVALID_SIZE, "size " + std::to_string(ref->width) + "x" + std::to_string(ref->height) +
                                         " unsupported");
  ssim2_batch::Ref r;
  r.w = (int)ref->width;
  r.h = (int)ref->height;
  const size_t sb = storage_bytes(b->storage);
  for (int sc = 0; sc < fp.d.nscales; ++sc) {
    r.off[sc] = (long long)r.bytes;
    r.bytes += align256(3 * size_t(fp.d.sc[sc].plane) * sb);
  }
  const size_t reg_bytes = align256(sizeof(PairDesc)) + 4 * sizeof(int) + 2 * sizeof(int);
  hipError_t e = hipSuccess;
  if (!b->d_reg) e = hipMalloc(&b->d_reg, reg_bytes);
  if (e == hipSuccess) e = hipMalloc(&r.planes, r.bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return bfail(SSIM2_OOM, std::string("reference pyramid: ") + hipGetErrorString(e));
  }
  // One pyramid launch with image 1 left alone writes image 0 (the reference)
  // into the new planes; the arena is not touched.
  PairDesc d = fp.d;
  d.in_ref = d.in_dis = static_cast<const unsigned char*>(ref->ptr);
  d.in_ref_pitch = d.in_dis_pitch = (long long)ref->pitch_bytes;
  d.in_kind = kInputSrgbU8;
  d.arena = reinterpret_cast<float*>(b->arena);
  d.ref_planes = r.planes;
  d.skip_img = 2;
  for (int sc = 0; sc < d.nscales; ++sc) d.sc[sc].cache = r.off[sc];
  const int index[6] = {1, fp.npyr, 0, 0, 0, fp.npyr};  // meta (4 ints), then prefix {0, npyr}
  unsigned char* dd = b->d_reg;
  int* di = reinterpret_cast<int*>(dd + align256(sizeof(PairDesc)));
  hipStream_t s = (hipStream_t)stream;
  // the registration buffer is reused: the previous registration's launch
  // finished (each registration waits for its own below)
  if ((e = hipMemcpy(dd, &d, sizeof(PairDesc), hipMemcpyHostToDevice)) == hipSuccess &&
      (e = hipMemcpy(di, index, sizeof(index), hipMemcpyHostToDevice)) == hipSuccess &&
      (e = launch_fast_pyramid(b->storage, reinterpret_cast<const PairDesc*>(dd), di + 4, di, fp.npyr, s)) ==
          hipSuccess)
    e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(r.planes);
    return bfail(SSIM2_HIP_ERROR, std::string("reference pyramid: ") + hipGetErrorString(e));
  }
  r.id = b->next_ref++;
  b->refs.push_back(r);
  b->bytes += r.bytes;
  *ref_id = r.id;
  return SSIM2_OK;
}

ssim2_status ssim2_batch_remove_reference(ssim2_batch* b, uint32_t ref_id) {
  if (!b) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  for (size_t i = 0; i < b->refs.size(); ++i)
    if (b->refs[i].id == ref_id) {
      for (ssim2_batch::Slot& S : b->slot)  // batches in flight may read it
        if (S.pending) BHIP(hipEventSynchronize(S.done));
      if (b->refs[i].owned) {
        BHIP(hipFree(b->refs[i].planes));
        b->bytes -= b->refs[i].bytes;
      }
      b->refs.erase(b->refs.begin() + (long)i);
      return SSIM2_OK;
    }
  return bfail(SSIM2_NO_REFERENCE, "unknown reference id");
}

ssim2_status ssim2_batch_compute_ref_async(ssim2_batch* b, const ssim2_ref_pair* pairs, uint32_t npairs,
                                           ssim2_stream stream) {
  if (!b || (!pairs && npairs)) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_status st = stage(b, nullptr, npairs, b->tile_cfg, nullptr, pairs);
  if (st != SSIM2_OK) return st;
  return submit(b, npairs, (hipStream_t)stream, true, b->tile_cfg, b->grid_mult);
}

}  // extern "C"

namespace ssim2 {

ssim2_status batch_share_reference(ssim2_batch* dst, const ssim2_batch* src, uint32_t ref_id) {
  const ssim2_batch::Ref* r = find_ref(src, ref_id);
  if (!dst || !r) return bfail(SSIM2_NO_REFERENCE, "unknown reference id");
  if (dst->storage != src->storage) return bfail(SSIM2_INVALID_ARGUMENT, "storage types differ");
  ssim2_batch::Ref c = *r;
  c.owned = false;
  dst->refs.push_back(c);
  return SSIM2_OK;
}

ssim2_status batch_compute_mixed_async(ssim2_batch* b, const ssim2_pair* pairs, const uint32_t* ref_ids,
                                       uint32_t npairs, hipStream_t stream) {
  if (!b || (!pairs && npairs)) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_status st = stage(b, pairs, npairs, b->tile_cfg, nullptr, nullptr, ref_ids);
  if (st != SSIM2_OK) return st;
  return submit(b, npairs, stream, true, b->tile_cfg, b->grid_mult);
}

}  // namespace ssim2

extern "C" {

ssim2_status ssim2_batch_compute_yuv_async(ssim2_batch* b, const ssim2_yuv_pair* pairs, uint32_t npairs,
                                           ssim2_stream stream) {
  if (!b || (!pairs && npairs)) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_status st = stage(b, nullptr, npairs, b->tile_cfg, pairs);
  if (st != SSIM2_OK) return st;
  return submit(b, npairs, (hipStream_t)stream, true, b->tile_cfg, b->grid_mult);
}

ssim2_status ssim2_batch_get_scores(ssim2_batch* b, double* scores, uint32_t capacity) {
  if (!b || !scores) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_batch::Slot& S = b->slot[b->cur];
  if (!S.pending && S.n == 0) return bfail(SSIM2_INVALID_ARGUMENT, "no batch has been submitted");
  if (capacity < (uint32_t)S.n) return bfail(SSIM2_BATCH_SIZE, "score buffer too small");
  if (S.pending) BHIP(hipEventSynchronize(S.done));
  S.pending = false;
  for (int i = 0; i < S.n; ++i) scores[i] = S.h_results[size_t(i) * kResultDoubles];
  return SSIM2_OK;
}

ssim2_status ssim2_batch_query(ssim2_batch* b, int* done) {
  if (!b || !done) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  const ssim2_batch::Slot& S = b->slot[b->cur];
  if (!S.pending) {
    *done = 1;
    return SSIM2_OK;
  }
  hipError_t e = hipEventQuery(S.done);
  if (e == hipErrorNotReady) {
    *done = 0;
    return SSIM2_OK;
  }
  BHIP(e);
  *done = 1;
  return SSIM2_OK;
}

ssim2_status ssim2_batch_get_sums(ssim2_batch* b, uint32_t index, double* sums108, int* dims12) {
  if (!b) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  ssim2_batch::Slot& S = b->slot[b->cur];
  if (index >= (uint32_t)S.n) return bfail(SSIM2_INVALID_ARGUMENT, "pair index out of range");
  if (S.pending) BHIP(hipEventSynchronize(S.done));
  S.pending = false;
  const double* r = S.h_results + size_t(index) * kResultDoubles;
  if (sums108) std::memcpy(sums108, r + 1, kSumsPerPair * sizeof(double));
  if (dims12)
    for (int i = 0; i < 2 * kMaxScales; ++i) dims12[i] = (int)r[1 + kSumsPerPair + i];
  return SSIM2_OK;
}

ssim2_status ssim2_batch_compute_sync(ssim2_batch* b, const ssim2_pair* pairs, uint32_t npairs,
                                      ssim2_stream stream, double* scores) {
  ssim2_status st = ssim2_batch_compute_async(b, pairs, npairs, stream);
  if (st != SSIM2_OK) return st;
  return ssim2_batch_get_scores(b, scores, npairs);
}

ssim2_status ssim2_batch_set_tiling(ssim2_batch* b, uint32_t tile_cfg, uint32_t grid_mult) {
  if (!b || tile_cfg >= (uint32_t)kNumTileCfgs || grid_mult > 1024)
    return bfail(SSIM2_INVALID_ARGUMENT, "tile_cfg in [0, 4], grid_mult in [0, 1024]");
  b->tile_cfg = (int)tile_cfg;
  b->grid_mult = (int)grid_mult;
  return SSIM2_OK;
}

ssim2_status ssim2_batch_get_tiling(const ssim2_batch* b, uint32_t* tile_cfg, uint32_t* grid_mult) {
  if (!b) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  if (tile_cfg) *tile_cfg = (uint32_t)b->tile_cfg;
  if (grid_mult) *grid_mult = (uint32_t)b->grid_mult;
  return SSIM2_OK;
}

ssim2_status ssim2_batch_autotune(ssim2_batch* b, const ssim2_pair* pairs, uint32_t npairs,
                                  ssim2_stream stream, uint32_t reps, char* report, size_t report_size) {
  if (!b || !pairs || npairs == 0) return bfail(SSIM2_INVALID_ARGUMENT, "null argument");
  if (reps == 0) reps = 3;
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  BHIP(hipEventCreate(&e0));
  BHIP(hipEventCreate(&e1));
  static const int kMults[] = {32, 64, 128, 256, 0};  // 0: one block per tile
  float best = 1e30f;
  int best_cfg = b->tile_cfg, best_mult = b->grid_mult;
  std::string rep = "tile     grid_x_cu  ms/batch\n";
  ssim2_status st = SSIM2_OK;
  for (int cfg = 0; cfg < kNumTileCfgs && st == SSIM2_OK; ++cfg)
    for (int mult : kMults) {
      if ((st = stage(b, pairs, npairs, cfg)) != SSIM2_OK) break;
      if ((st = submit(b, npairs, s, false, cfg, mult)) != SSIM2_OK) break;  // warm-up
      hipError_t e = hipEventRecord(e0, s);
      for (uint32_t r = 0; r < reps && e == hipSuccess; ++r)
        if (enqueue_kernels(b, b->slot[b->cur], s, cfg, mult) != SSIM2_OK) e = hipErrorUnknown;
      if (e == hipSuccess) e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        st = bfail(SSIM2_HIP_ERROR, std::string("autotune: ") + hipGetErrorString(e));
        break;
      }
      ms /= reps;
      char line[96];
      if (mult > 0)
        std::snprintf(line, sizeof line, "%2dx%-5d %-10d %.4f\n", kTileCfgW[cfg], kTileCfgH[cfg], mult, ms);
      else
        std::snprintf(line, sizeof line, "%2dx%-5d %-10s %.4f\n", kTileCfgW[cfg], kTileCfgH[cfg], "per tile", ms);
      rep += line;
      if (ms < best) {
        best = ms;
        best_cfg = cfg;
        best_mult = mult;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != SSIM2_OK) return st;
  BHIP(hipStreamSynchronize(s));
  b->tile_cfg = best_cfg;
  b->grid_mult = best_mult;
  char line[96];
  if (best_mult > 0)
    std::snprintf(line, sizeof line, "chosen %dx%d, grid %d x CUs\n", kTileCfgW[best_cfg], kTileCfgH[best_cfg],
                  best_mult);
  else
    std::snprintf(line, sizeof line, "chosen %dx%d, one block per tile\n", kTileCfgW[best_cfg], kTileCfgH[best_cfg]);
  rep += line;
  b->tune_report = rep;
  if (report && report_size) {
    std::strncpy(report, rep.c_str(), report_size - 1);
    report[report_size - 1] = 0;
  }
  return SSIM2_OK;
}

}  // extern "C"
