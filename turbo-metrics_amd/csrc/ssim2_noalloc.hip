// No-allocation form of the exact path: ssim2_compute_srgb_noalloc_async
// (include/ssim2hip.h). It is for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). It runs the kernels of ssim2_compute_srgb_async; everything they
// need beyond the caller's two input images and the caller's result record
// lives in `scratch`, a file-scope device array that each GPU holds until the
// process ends.
//
// A call is kernel launches only, all on the caller's stream: a set-up kernel
// writes the pair descriptor (passed by value) and the work list into
// `scratch` and clears the plane padding the scans read as zeros, then the
// pyramid, row, column and finalisation kernels run as in the context path.
// Nothing is copied from the host and nothing is allocated; the work list
// comes from the same functions the host planner uses (ssim2_plan.h).
//
// Every call on a GPU uses that GPU's one array, which keeps the last call's
// contents: calls that overlap on one GPU (two streams, two threads) race, and
// the caller orders them. What must read as zero (the padding columns right of
// every XYB row, the rows under every plane) is cleared by each call; the
// partial sums and the result record are overwritten, never accumulated.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/ssim2hip.h"
#include "ssim2_internal.h"
#include "ssim2_plan.h"

using namespace ssim2;

// 256 MiB per GPU (0.09 % of an MI355X's 288 GB): a 1920x1080 pair needs
// about 244 MB (21 planes per scale, no cached-reference planes).
constexpr size_t kScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
// Device address of `scratch` on each GPU (null until looked up). In host
// code `scratch` names a host-side stand-in, never the device memory.
void* g_scratch[kMaxDevices];
std::mutex g_scratch_mu;

ssim2_status nfail(ssim2_status st, const std::string& msg) {
  set_last_error(msg);
  return st;
}

// The address of `scratch` on the current device, looked up once per device
// with that device current.
ssim2_status scratch_base(unsigned char** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return nfail(SSIM2_HIP_ERROR, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return nfail(SSIM2_HIP_ERROR, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nfail(SSIM2_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch[dev]);
  return SSIM2_OK;
}

constexpr size_t align256(size_t v) { return (v + 255) & ~size_t(255); }
constexpr int kPlanesNoCache = 6 + 3 * kNumQ;  // XYB + row-pass planes, no cached-reference planes
constexpr int kSetupThreads = 256;

// Blocks [0, item_blocks) write the descriptor and the work list (pyramid
// items, row items of both sides, column items: the order launch offsets
// below assume). The other blocks clear, for every plane of every scale, the
// columns right of the image in rows 0 .. h-1 and all of rows h .. h+3: the
// row scan reads the former as the zeros past a row, the column scan the
// latter as the zeros under the image. One wave clears one plane row.
__global__ __launch_bounds__(kSetupThreads) void k_noalloc_setup(PairDesc d, PairDesc* __restrict__ d_out,
                                                                 WorkItem* __restrict__ items, int item_blocks) {
  if ((int)blockIdx.x < item_blocks) {
    const int n_pyr = pyramid_items(d), n_rows = row_items(d, kSidesBoth), n_cols = column_items(d);
    const int k = (int)blockIdx.x * kSetupThreads + (int)threadIdx.x;
    if (k == 0) *d_out = d;
    if (k < n_pyr)
      items[k] = pyramid_item(d, k, 0);
    else if (k < n_pyr + n_rows)
      items[k] = row_item(d, kSidesBoth, k - n_pyr, 0);
    else if (k < n_pyr + n_rows + n_cols)
      items[k] = column_item(d, k - n_pyr - n_rows, 0);
    return;
  }
  constexpr int kWavesPerBlock = kSetupThreads / 64;
  const int lane = (int)threadIdx.x & 63;
  const long long wave = (long long)(blockIdx.x - item_blocks) * kWavesPerBlock + (threadIdx.x >> 6);
  const long long nwaves = (long long)(gridDim.x - item_blocks) * kWavesPerBlock;
  for (int s = 0; s < d.nscales; ++s) {
    const ScaleDesc& S = d.sc[s];
    const long long prow = S.h + kPadRows, rows = (long long)kPlanesNoCache * prow;
    for (long long u = wave; u < rows; u += nwaves) {
      const long long plane = u / prow;
      const int y = (int)(u - plane * prow);
      float* row = d.arena + S.xyb + plane * S.plane + (long long)y * S.pitch;
      for (int x = (y < S.h ? S.w : 0) + lane; x < S.pitch; x += 64) row[x] = 0.0f;
    }
  }
}

ssim2_status check_srgb(const ssim2_image* im, const char* name) {
  if (!im || !im->ptr) return nfail(SSIM2_INVALID_ARGUMENT, std::string(name) + " is null");
  if (im->pitch_bytes < size_t(im->width) * 3)
    return nfail(SSIM2_INVALID_ARGUMENT, std::string(name) + " pitch smaller than a row");
  return SSIM2_OK;
}

}  // namespace

extern "C" ssim2_status ssim2_compute_srgb_noalloc_async(const ssim2_image* ref_srgb, const ssim2_image* dis_srgb,
                                                         double* result, ssim2_stream stream) {
  ssim2_status st;
  if ((st = check_srgb(ref_srgb, "ref_srgb")) != SSIM2_OK) return st;
  if ((st = check_srgb(dis_srgb, "dis_srgb")) != SSIM2_OK) return st;
  if (!result) return nfail(SSIM2_INVALID_ARGUMENT, "result is null");
  if (reinterpret_cast<uintptr_t>(result) & 7) return nfail(SSIM2_INVALID_ARGUMENT, "result not 8-byte aligned");
  const uint32_t w = ref_srgb->width, h = ref_srgb->height;
  if (dis_srgb->width != w || dis_srgb->height != h)
    return nfail(SSIM2_DIM_MISMATCH, "reference and distorted sizes differ");
  if (!size_supported(w, h))
    return nfail(SSIM2_INVALID_SIZE, "image size " + std::to_string(w) + "x" + std::to_string(h) + " outside [8, " +
                                         std::to_string(kMaxDim) + "]");
  PairDesc d;
  int nparts = 0;
  const long long elems = layout_pair((int)w, (int)h, 0, 0, kPlanesNoCache, &d, &nparts);
  const int n_pyr = pyramid_items(d), n_rows = row_items(d, kSidesBoth), n_cols = column_items(d);
  const int n_items = n_pyr + n_rows + n_cols;
  // scratch layout: descriptor | work list | partial sums | planes (+ 64 KB
  // read past the last plane by the row pass's prefetch)
  const size_t off_items = align256(sizeof(PairDesc));
  const size_t off_parts = align256(off_items + size_t(n_items) * sizeof(WorkItem));
  const size_t off_arena = align256(off_parts + size_t(nparts) * kNumSums * sizeof(double));
  const size_t need = off_arena + size_t(elems) * sizeof(float) + 65536;
  if (need > kScratchBytes)
    return nfail(SSIM2_INVALID_SIZE, "a " + std::to_string(w) + "x" + std::to_string(h) + " pair needs " +
                                         std::to_string(need) + " B of scratch; the no-allocation path holds " +
                                         std::to_string(kScratchBytes) + " B per GPU (about 2.2 Mpx)");
  unsigned char* base = nullptr;
  if ((st = scratch_base(&base)) != SSIM2_OK) return st;
  PairDesc* d_desc = reinterpret_cast<PairDesc*>(base);
  WorkItem* d_items = reinterpret_cast<WorkItem*>(base + off_items);
  d.partials = reinterpret_cast<double*>(base + off_parts);
  d.arena = reinterpret_cast<float*>(base + off_arena);
  d.result = result;
  d.in_kind = kInputSrgbU8;
  d.in_ref = static_cast<const unsigned char*>(ref_srgb->ptr);
  d.in_dis = static_cast<const unsigned char*>(dis_srgb->ptr);
  d.in_ref_pitch = (long long)ref_srgb->pitch_bytes;
  d.in_dis_pitch = (long long)dis_srgb->pitch_bytes;
  InputArgs in{};
  in.ref = d.in_ref;
  in.dis = d.in_dis;
  in.ref_pitch = d.in_ref_pitch;
  in.dis_pitch = d.in_dis_pitch;
  in.kind = kInputSrgbU8;
  in.use = 1;

  hipStream_t s = (hipStream_t)stream;
  const int item_blocks = (n_items + kSetupThreads - 1) / kSetupThreads;
  const int pad_blocks = 512;
  hipLaunchKernelGGL(k_noalloc_setup, dim3(item_blocks + pad_blocks), dim3(kSetupThreads), 0, s, d, d_desc, d_items,
                     item_blocks);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = launch_pyramid(d_desc, d_items, n_pyr, kSidesBoth, in, s);
  if (e == hipSuccess) e = launch_rows(d_desc, d_items + n_pyr, n_rows, kSidesBoth, s);
  if (e == hipSuccess) e = launch_cols(d_desc, d_items + n_pyr + n_rows, n_cols, kSidesBoth, s);
  if (e == hipSuccess) e = launch_finalize(d_desc, 1, s);
  if (e != hipSuccess) return nfail(SSIM2_HIP_ERROR, std::string("no-allocation launch: ") + hipGetErrorString(e));
  return SSIM2_OK;
}
