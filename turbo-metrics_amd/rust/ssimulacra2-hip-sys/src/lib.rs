//! Raw bindings of `include/ssim2hip.h`, the C ABI of the MI355X ssimulacra2
//! engine. Every type and function here mirrors the header one for one; see
//! the header for the meaning of each call and the reference method it
//! replaces. The safe wrapper is the `ssimulacra2-hip` crate.
#![allow(non_camel_case_types)]

use std::ffi::{c_char, c_int, c_void};

/// `ssim2_status`: 0 is `SSIM2_OK`; `ssim2_status_string` names the others.
pub type ssim2_status = c_int;
pub const SSIM2_OK: ssim2_status = 0;
pub const SSIM2_INVALID_SIZE: ssim2_status = 1;
pub const SSIM2_DIM_MISMATCH: ssim2_status = 2;
pub const SSIM2_NO_REFERENCE: ssim2_status = 3;
pub const SSIM2_OOM: ssim2_status = 4;
pub const SSIM2_BATCH_SIZE: ssim2_status = 5;
pub const SSIM2_ADMISSION_REFUSED: ssim2_status = 6;
pub const SSIM2_HIP_ERROR: ssim2_status = 7;
pub const SSIM2_INVALID_ARGUMENT: ssim2_status = 8;
pub const SSIM2_NOT_READY: ssim2_status = 9;

/// `hipStream_t`; null selects the default stream.
pub type ssim2_stream = *mut c_void;

pub const SSIM2_RESULT_DOUBLES: usize = 121;

pub const SSIM2_YUV420P10: u32 = 0;
pub const SSIM2_P016: u32 = 1;
pub const SSIM2_STORAGE_F32: u32 = 0;
pub const SSIM2_STORAGE_F16: u32 = 1;
/// Refused by the library (measured 0.30-0.35 from FP64 on 4K frames).
pub const SSIM2_STORAGE_BF16: u32 = 2;
pub const SSIM2_FRAME_SRGB: u32 = 0;
pub const SSIM2_FRAME_YUV420: u32 = 1;
pub const SSIM2_VIDEO_EXACT: u32 = 0;
pub const SSIM2_VIDEO_FAST: u32 = 1;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ssim2_image {
    pub ptr: *mut c_void,
    pub pitch_bytes: usize,
    pub width: u32,
    pub height: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ssim2_yuv_image {
    pub y: *const c_void,
    pub c0: *const c_void,
    pub c1: *const c_void,
    pub y_pitch_bytes: usize,
    pub c_pitch_bytes: usize,
    pub width: u32,
    pub height: u32,
    pub format: u32,
    pub reserved0: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ssim2_options {
    pub struct_size: u32,
    pub reserved0: u32,
    pub memory_budget: u64,
}

#[repr(C)]
pub struct ssim2_ctx {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct ssim2_batch {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct ssim2_server {
    _opaque: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ssim2_batch_options {
    pub struct_size: u32,
    pub storage: u32,
    pub max_pairs: u32,
    pub reserved0: u32,
    pub max_pixels: u64,
    pub memory_budget: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ssim2_pair {
    pub ref_: ssim2_image,
    pub dis: ssim2_image,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ssim2_ref_pair {
    pub ref_id: u32,
    pub reserved: u32,
    pub dis: ssim2_image,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ssim2_yuv_pair {
    pub ref_: ssim2_yuv_image,
    pub dis: ssim2_yuv_image,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ssim2_server_options {
    pub struct_size: u32,
    pub storage: u32,
    pub slots: u32,
    pub max_pairs_per_batch: u32,
    pub max_pixels_per_batch: u64,
    pub hbm_budget: u64,
    pub tile_cfg: u32,
    pub grid_mult: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ssim2_result {
    pub request_id: u64,
    pub score: f64,
    pub submit_time_s: f64,
    pub done_time_s: f64,
    pub status: ssim2_status,
    pub reserved0: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ssim2_server_stats_t {
    pub struct_size: u32,
    pub reserved0: u32,
    pub submitted: u64,
    pub completed: u64,
    pub refused: u64,
    pub batches: u64,
    pub pixels_completed: u64,
    pub waiting: u64,
    pub in_flight_bytes: u64,
    pub first_submit_s: f64,
    pub last_done_s: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ssim2_stats {
    pub min: f64,
    pub max: f64,
    pub mean: f64,
    pub var: f64,
    pub sample_var: f64,
    pub stddev: f64,
    pub sample_stddev: f64,
    pub p1: f64,
    pub p5: f64,
    pub p50: f64,
    pub p95: f64,
    pub p99: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ssim2_frame {
    pub kind: u32,
    pub reserved0: u32,
    pub srgb: ssim2_image,
    pub yuv: ssim2_yuv_image,
}

/// `FrameSource::next_frame`: 1 = a frame, 0 = end of stream, < 0 = error.
pub type ssim2_frame_source = Option<unsafe extern "C" fn(user: *mut c_void, frame: *mut ssim2_frame) -> c_int>;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ssim2_video_options {
    pub struct_size: u32,
    pub every: u32,
    pub skip: u32,
    pub skip_ref: u32,
    pub skip_dis: u32,
    pub frames: u32,
    pub mode: u32,
    pub storage: u32,
    pub batch: u32,
    pub reserved0: u32,
}

extern "C" {
    // lifecycle
    pub fn ssim2_create(width: u32, height: u32, options: *const ssim2_options, out: *mut *mut ssim2_ctx)
        -> ssim2_status;
    pub fn ssim2_destroy(ctx: *mut ssim2_ctx) -> ssim2_status;
    pub fn ssim2_mem_usage(ctx: *const ssim2_ctx, bytes: *mut usize) -> ssim2_status;

    // scoring
    pub fn ssim2_compute_linear_async(ctx: *mut ssim2_ctx, ref_linear: *const ssim2_image,
                                      dis_linear: *const ssim2_image, stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_get_score(ctx: *mut ssim2_ctx, score: *mut f64) -> ssim2_status;
    pub fn ssim2_compute_linear_sync(ctx: *mut ssim2_ctx, ref_linear: *const ssim2_image,
                                     dis_linear: *const ssim2_image, stream: ssim2_stream, score: *mut f64)
        -> ssim2_status;
    pub fn ssim2_compute_srgb_sync(ctx: *mut ssim2_ctx, ref_srgb: *const ssim2_image, dis_srgb: *const ssim2_image,
                                   ref_linear: *mut ssim2_image, dis_linear: *mut ssim2_image,
                                   stream: ssim2_stream, score: *mut f64) -> ssim2_status;
    pub fn ssim2_compute_srgb_async(ctx: *mut ssim2_ctx, ref_srgb: *const ssim2_image,
                                    dis_srgb: *const ssim2_image, stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_compute_srgb_noalloc_async(ref_srgb: *const ssim2_image, dis_srgb: *const ssim2_image,
                                            result: *mut f64, stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_compute_from_host_srgb(ctx: *mut ssim2_ctx, ref_: *const u8, dis: *const u8, row_bytes: usize,
                                        stream: ssim2_stream, score: *mut f64) -> ssim2_status;
    pub fn ssim2_compute_yuv420_async(ctx: *mut ssim2_ctx, ref_: *const ssim2_yuv_image,
                                      dis: *const ssim2_yuv_image, stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_compute_yuv420_sync(ctx: *mut ssim2_ctx, ref_: *const ssim2_yuv_image,
                                     dis: *const ssim2_yuv_image, stream: ssim2_stream, score: *mut f64)
        -> ssim2_status;

    // cached reference
    pub fn ssim2_set_reference_linear(ctx: *mut ssim2_ctx, ref_linear: *const ssim2_image, stream: ssim2_stream)
        -> ssim2_status;
    pub fn ssim2_set_reference_srgb(ctx: *mut ssim2_ctx, ref_srgb: *const ssim2_image, stream: ssim2_stream)
        -> ssim2_status;
    pub fn ssim2_compute_with_reference_linear(ctx: *mut ssim2_ctx, ref_linear: *const ssim2_image,
                                               dis_linear: *const ssim2_image, stream: ssim2_stream,
                                               score: *mut f64) -> ssim2_status;
    pub fn ssim2_compute_with_reference_srgb(ctx: *mut ssim2_ctx, dis_srgb: *const ssim2_image,
                                             stream: ssim2_stream, score: *mut f64) -> ssim2_status;
    pub fn ssim2_prepare_reference_from_host_srgb(ctx: *mut ssim2_ctx, ref_: *const u8, row_bytes: usize,
                                                  stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_compute_from_host_dis_srgb(ctx: *mut ssim2_ctx, dis: *const u8, row_bytes: usize,
                                            stream: ssim2_stream, score: *mut f64) -> ssim2_status;
    pub fn ssim2_clear_reference(ctx: *mut ssim2_ctx) -> ssim2_status;
    pub fn ssim2_has_reference(ctx: *const ssim2_ctx, has: *mut c_int) -> ssim2_status;

    // batched fast form
    pub fn ssim2_batch_create(options: *const ssim2_batch_options, out: *mut *mut ssim2_batch) -> ssim2_status;
    pub fn ssim2_batch_destroy(batch: *mut ssim2_batch) -> ssim2_status;
    pub fn ssim2_batch_mem_usage(batch: *const ssim2_batch, bytes: *mut usize) -> ssim2_status;
    pub fn ssim2_batch_compute_async(batch: *mut ssim2_batch, pairs: *const ssim2_pair, npairs: u32,
                                     stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_batch_add_reference(batch: *mut ssim2_batch, ref_: *const ssim2_image, stream: ssim2_stream,
                                     ref_id: *mut u32) -> ssim2_status;
    pub fn ssim2_batch_remove_reference(batch: *mut ssim2_batch, ref_id: u32) -> ssim2_status;
    pub fn ssim2_batch_compute_ref_async(batch: *mut ssim2_batch, pairs: *const ssim2_ref_pair, npairs: u32,
                                         stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_batch_compute_yuv_async(batch: *mut ssim2_batch, pairs: *const ssim2_yuv_pair, npairs: u32,
                                         stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_batch_get_scores(batch: *mut ssim2_batch, scores: *mut f64, capacity: u32) -> ssim2_status;
    pub fn ssim2_batch_compute_sync(batch: *mut ssim2_batch, pairs: *const ssim2_pair, npairs: u32,
                                    stream: ssim2_stream, scores: *mut f64) -> ssim2_status;
    pub fn ssim2_batch_query(batch: *mut ssim2_batch, done: *mut c_int) -> ssim2_status;
    pub fn ssim2_batch_get_sums(batch: *mut ssim2_batch, index: u32, sums108: *mut f64, dims12: *mut c_int)
        -> ssim2_status;
    pub fn ssim2_batch_set_tiling(batch: *mut ssim2_batch, tile_cfg: u32, grid_mult: u32) -> ssim2_status;
    pub fn ssim2_batch_get_tiling(batch: *const ssim2_batch, tile_cfg: *mut u32, grid_mult: *mut u32)
        -> ssim2_status;
    pub fn ssim2_batch_autotune(batch: *mut ssim2_batch, pairs: *const ssim2_pair, npairs: u32,
                                stream: ssim2_stream, reps: u32, report: *mut c_char, report_size: usize)
        -> ssim2_status;

    // serving
    pub fn ssim2_server_create(options: *const ssim2_server_options, out: *mut *mut ssim2_server) -> ssim2_status;
    pub fn ssim2_server_destroy(server: *mut ssim2_server) -> ssim2_status;
    pub fn ssim2_server_submit(server: *mut ssim2_server, pair: *const ssim2_pair, request_id: u64)
        -> ssim2_status;
    pub fn ssim2_server_add_reference(server: *mut ssim2_server, ref_: *const ssim2_image, ref_id: *mut u32)
        -> ssim2_status;
    pub fn ssim2_server_remove_reference(server: *mut ssim2_server, ref_id: u32) -> ssim2_status;
    pub fn ssim2_server_submit_ref(server: *mut ssim2_server, ref_id: u32, dis: *const ssim2_image,
                                   request_id: u64) -> ssim2_status;
    pub fn ssim2_server_poll(server: *mut ssim2_server, out: *mut ssim2_result, capacity: u32, n: *mut u32)
        -> ssim2_status;
    pub fn ssim2_server_drain(server: *mut ssim2_server, timeout_s: f64) -> ssim2_status;
    pub fn ssim2_server_stats(server: *mut ssim2_server, stats: *mut ssim2_server_stats_t) -> ssim2_status;

    // video driver and statistics
    pub fn ssim2_stats_compute(values: *const f64, n: usize, out: *mut ssim2_stats) -> ssim2_status;
    pub fn ssim2_video_compute_all(width: u32, height: u32, options: *const ssim2_video_options,
                                   ref_source: ssim2_frame_source, ref_user: *mut c_void,
                                   dis_source: ssim2_frame_source, dis_user: *mut c_void, stream: ssim2_stream,
                                   scores: *mut f64, capacity: usize, frame_count: *mut usize,
                                   stats: *mut ssim2_stats) -> ssim2_status;

    // diagnostics
    pub fn ssim2_get_sums(ctx: *mut ssim2_ctx, sums108: *mut f64, dims12: *mut c_int, nscales: *mut c_int)
        -> ssim2_status;
    pub fn ssim2_time_kernels(ctx: *mut ssim2_ctx, ref_srgb: *const ssim2_image, dis_srgb: *const ssim2_image,
                              stream: ssim2_stream, reps: c_int, ms: *mut f32) -> ssim2_status;
    pub fn ssim2_get_plane(ctx: *mut ssim2_ctx, kind: u32, scale: u32, index: u32, host: *mut f32,
                           host_row_floats: usize) -> ssim2_status;
    pub fn ssim2_check_arithmetic(which: u32, lo_bits: u32, hi_bits: u32, mismatches: *mut u64) -> ssim2_status;

    // utilities
    pub fn ssim2_device_alloc(bytes: usize, out: *mut *mut c_void) -> ssim2_status;
    pub fn ssim2_device_free(ptr: *mut c_void) -> ssim2_status;
    pub fn ssim2_device_copy(dst: *mut c_void, src: *const c_void, bytes: usize, stream: ssim2_stream)
        -> ssim2_status;
    pub fn ssim2_device_synchronize() -> ssim2_status;
    pub fn ssim2_srgb_to_linear(src_u8: *const ssim2_image, dst_f32: *mut ssim2_image, stream: ssim2_stream)
        -> ssim2_status;
    pub fn ssim2_downscale_by_2(src_f32: *const ssim2_image, dst_f32: *mut ssim2_image, stream: ssim2_stream)
        -> ssim2_status;
    pub fn ssim2_linear_to_xyb(src_f32: *const ssim2_image, dst_f32: *mut ssim2_image, stream: ssim2_stream)
        -> ssim2_status;
    pub fn ssim2_yuv420_to_linear(src: *const ssim2_yuv_image, dst_f32: *mut ssim2_image, stream: ssim2_stream)
        -> ssim2_status;
    pub fn ssim2_image_alloc(width: u32, height: u32, bytes_per_sample: u32, out: *mut ssim2_image) -> ssim2_status;
    pub fn ssim2_image_free(image: *mut ssim2_image) -> ssim2_status;
    pub fn ssim2_image_upload(image: *mut ssim2_image, host: *const c_void, host_row_bytes: usize) -> ssim2_status;
    pub fn ssim2_image_download(image: *const ssim2_image, host: *mut c_void, host_row_bytes: usize)
        -> ssim2_status;
    pub fn ssim2_yuv_alloc(width: u32, height: u32, format: u32, out: *mut ssim2_yuv_image) -> ssim2_status;
    pub fn ssim2_yuv_free(image: *mut ssim2_yuv_image) -> ssim2_status;
    pub fn ssim2_yuv_upload(image: *mut ssim2_yuv_image, y: *const u16, c0: *const u16, c1: *const u16)
        -> ssim2_status;
    pub fn ssim2_stream_create(out: *mut ssim2_stream) -> ssim2_status;
    pub fn ssim2_stream_destroy(stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_stream_synchronize(stream: ssim2_stream) -> ssim2_status;
    pub fn ssim2_status_string(status: ssim2_status) -> *const c_char;
    pub fn ssim2_last_error() -> *const c_char;
    pub fn ssim2_version() -> c_int;
}
