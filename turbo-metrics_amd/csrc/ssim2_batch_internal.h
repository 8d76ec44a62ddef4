// Batch-engine internals shared with the serving runtime (ssim2_server.hip).
// Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ssim2hip.h"

namespace ssim2 {

// Make the resident reference `ref_id` of `src` usable by `dst` (same
// storage type): `dst` reads the same planes and never frees them, so `src`
// must keep the reference until `dst` has removed it.
ssim2_status batch_share_reference(ssim2_batch* dst, const ssim2_batch* src, uint32_t ref_id);

// ssim2_batch_compute_async where ref_ids[i] != 0 replaces pair i's reference
// image with that resident reference (its `ref` field is then not read).
ssim2_status batch_compute_mixed_async(ssim2_batch* b, const ssim2_pair* pairs, const uint32_t* ref_ids,
                                       uint32_t npairs, hipStream_t stream);

}  // namespace ssim2
