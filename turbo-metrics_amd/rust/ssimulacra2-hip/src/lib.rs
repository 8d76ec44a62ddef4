//! `ssimulacra2_cuda::Ssimulacra2` (crates/ssimulacra2-cuda/src/lib.rs:27-602)
//! on the MI355X engine: the same type, method names and argument meaning,
//! over the C ABI of libssim2hip. Images are packed RGB in device memory
//! described by the `Img` trait (the role of cudarse-npp's `Img<S, C<3>>`);
//! streams are HIP streams. Errors are `Ssim2Error` values carrying the
//! library's status and message; nothing panics except `get_score` on a
//! context with no computation, as the reference's does.
use std::ffi::CStr;
use std::fmt;
use std::ptr;

use ssimulacra2_hip_sys as sys;

/// A packed RGB device image of samples `T` (u8 sRGB or f32 linear).
pub trait Img<T> {
    fn device_ptr(&self) -> *const T;
    fn pitch(&self) -> usize;
    fn width(&self) -> u32;
    fn height(&self) -> u32;
}

/// A writable packed RGB device image.
pub trait ImgMut<T>: Img<T> {
    fn device_ptr_mut(&mut self) -> *mut T;
}

fn image<T, I: Img<T> + ?Sized>(i: &I) -> sys::ssim2_image {
    sys::ssim2_image { ptr: i.device_ptr() as *mut _, pitch_bytes: i.pitch(), width: i.width(), height: i.height() }
}

/// Error of a library call: its status and the message of ssim2_last_error().
#[derive(Clone, Debug)]
pub struct Ssim2Error {
    pub status: sys::ssim2_status,
    pub message: String,
}

impl fmt::Display for Ssim2Error {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        let name = unsafe { CStr::from_ptr(sys::ssim2_status_string(self.status)) };
        write!(f, "{}: {}", name.to_string_lossy(), self.message)
    }
}

impl std::error::Error for Ssim2Error {}

fn check(status: sys::ssim2_status) -> Result<(), Ssim2Error> {
    if status == sys::SSIM2_OK {
        return Ok(());
    }
    let message = unsafe { CStr::from_ptr(sys::ssim2_last_error()) }.to_string_lossy().into_owned();
    Err(Ssim2Error { status, message })
}

/// A HIP stream (cudarse-driver `CuStream`), created non-blocking.
pub struct HipStream(sys::ssim2_stream);

impl HipStream {
    pub fn new() -> Result<Self, Ssim2Error> {
        let mut s = ptr::null_mut();
        check(unsafe { sys::ssim2_stream_create(&mut s) })?;
        Ok(Self(s))
    }
    /// The default (null) stream.
    pub fn default_stream() -> Self {
        Self(ptr::null_mut())
    }
    pub fn raw(&self) -> sys::ssim2_stream {
        self.0
    }
    pub fn sync(&self) -> Result<(), Ssim2Error> {
        check(unsafe { sys::ssim2_stream_synchronize(self.0) })
    }
}

impl Drop for HipStream {
    fn drop(&mut self) {
        if !self.0.is_null() {
            unsafe { sys::ssim2_stream_destroy(self.0) };
        }
    }
}

/// The stage kernels (`Ssimulacra2::kernel()`, crates/ssimulacra2-cuda/src/kernel.rs).
pub struct Kernel;

impl Kernel {
    /// kernel.rs:36-57: u8 sRGB to f32 linear through the 256-entry table.
    pub fn srgb_to_linear(&self, stream: &HipStream, src: &impl Img<u8>, dst: &mut impl ImgMut<f32>) {
        let (s, mut d) = (image(src), image(dst));
        check(unsafe { sys::ssim2_srgb_to_linear(&s, &mut d, stream.raw()) }).expect("srgb_to_linear");
    }
    /// kernel.rs:59-83: 2x2 mean, dst half the size (rounded up).
    pub fn downscale_by_2(&self, stream: &HipStream, src: &impl Img<f32>, dst: &mut impl ImgMut<f32>) {
        let (s, mut d) = (image(src), image(dst));
        check(unsafe { sys::ssim2_downscale_by_2(&s, &mut d, stream.raw()) }).expect("downscale_by_2");
    }
    /// kernel.rs:161-184: linear RGB to positive XYB.
    pub fn linear_to_xyb(&self, stream: &HipStream, src: &impl Img<f32>, dst: &mut impl ImgMut<f32>) {
        let (s, mut d) = (image(src), image(dst));
        check(unsafe { sys::ssim2_linear_to_xyb(&s, &mut d, stream.raw()) }).expect("linear_to_xyb");
    }
}

/// ssimulacra2 for one image size (lib.rs:27-77).
pub struct Ssimulacra2 {
    ctx: *mut sys::ssim2_ctx,
    ref_lin: sys::ssim2_image,
    dis_lin: sys::ssim2_image,
    kernel: Kernel,
}

unsafe impl Send for Ssimulacra2 {}

impl Ssimulacra2 {
    /// lib.rs:80-197. The two linear images are remembered for `compute` /
    /// `compute_sync`, as the reference's graph is bound to them; unlike the
    /// reference, any other image of the same size may be scored later.
    pub fn new(ref_lin: &impl Img<f32>, dis_lin: &impl Img<f32>, _stream: &HipStream) -> Result<Self, Ssim2Error> {
        if ref_lin.width() != dis_lin.width() || ref_lin.height() != dis_lin.height() {
            return Err(Ssim2Error { status: sys::SSIM2_DIM_MISMATCH, message: "image sizes differ".into() });
        }
        let opts = sys::ssim2_options {
            struct_size: std::mem::size_of::<sys::ssim2_options>() as u32,
            ..Default::default()
        };
        let mut ctx = ptr::null_mut();
        check(unsafe { sys::ssim2_create(ref_lin.width(), ref_lin.height(), &opts, &mut ctx) })?;
        Ok(Self { ctx, ref_lin: image(ref_lin), dis_lin: image(dis_lin), kernel: Kernel })
    }

    /// lib.rs:199-204.
    pub fn kernel(&self) -> &Kernel {
        &self.kernel
    }

    /// lib.rs:206-235: device bytes of the context.
    pub fn mem_usage(&self) -> usize {
        let mut b = 0;
        unsafe { sys::ssim2_mem_usage(self.ctx, &mut b) };
        b
    }

    /// lib.rs:419-424: enqueue the pipeline for the images given to `new`.
    pub fn compute(&mut self, stream: &HipStream) -> Result<(), Ssim2Error> {
        check(unsafe { sys::ssim2_compute_linear_async(self.ctx, &self.ref_lin, &self.dis_lin, stream.raw()) })
    }

    /// lib.rs:426-429: waits for the last computation and returns its score.
    pub fn get_score(&mut self) -> f64 {
        let mut s = 0.0;
        check(unsafe { sys::ssim2_get_score(self.ctx, &mut s) }).expect("get_score without a computation");
        s
    }

    /// lib.rs:409-416.
    pub fn compute_sync(&mut self, stream: &HipStream) -> Result<f64, Ssim2Error> {
        self.compute(stream)?;
        Ok(self.get_score())
    }

    /// lib.rs:391-404: u8 sRGB device images; the linear copies are written
    /// into `ref_lin` / `dis_lin` as the reference's srgb_to_linear step does.
    pub fn compute_srgb_sync(&mut self, src_ref: &impl Img<u8>, src_dis: &impl Img<u8>,
                             ref_lin: &mut impl ImgMut<f32>, dis_lin: &mut impl ImgMut<f32>,
                             stream: &HipStream) -> Result<f64, Ssim2Error> {
        let (r, d) = (image(src_ref), image(src_dis));
        let (mut rl, mut dl) = (image(ref_lin), image(dis_lin));
        let mut s = 0.0;
        check(unsafe { sys::ssim2_compute_srgb_sync(self.ctx, &r, &d, &mut rl, &mut dl, stream.raw(), &mut s) })?;
        Ok(s)
    }

    /// lib.rs:329-345: packed sRGB bytes in host memory.
    pub fn compute_from_cpu_srgb_sync(&mut self, ref_bytes: &[u8], dis_bytes: &[u8], stream: &HipStream)
        -> Result<f64, Ssim2Error> {
        let row = self.ref_lin.width as usize * 3;
        self.check_host(ref_bytes)?;
        self.check_host(dis_bytes)?;
        let mut s = 0.0;
        check(unsafe {
            sys::ssim2_compute_from_host_srgb(self.ctx, ref_bytes.as_ptr(), dis_bytes.as_ptr(), row, stream.raw(), &mut s)
        })?;
        Ok(s)
    }

    /// lib.rs:373-388: upload host sRGB bytes and cache them as the reference.
    pub fn prepare_reference_from_cpu_srgb(&mut self, ref_bytes: &[u8], stream: &HipStream) -> Result<(), Ssim2Error> {
        self.check_host(ref_bytes)?;
        let row = self.ref_lin.width as usize * 3;
        check(unsafe { sys::ssim2_prepare_reference_from_host_srgb(self.ctx, ref_bytes.as_ptr(), row, stream.raw()) })
    }

    /// lib.rs:347-371: host sRGB bytes scored against the cached reference.
    pub fn compute_from_cpu_dis_only_srgb_sync(&mut self, dis_bytes: &[u8], stream: &HipStream)
        -> Result<f64, Ssim2Error> {
        self.check_host(dis_bytes)?;
        let row = self.ref_lin.width as usize * 3;
        let mut s = 0.0;
        check(unsafe { sys::ssim2_compute_from_host_dis_srgb(self.ctx, dis_bytes.as_ptr(), row, stream.raw(), &mut s) })?;
        Ok(s)
    }

    /// lib.rs:440-541: compute and keep the reference side.
    pub fn set_reference_linear(&mut self, ref_lin: &impl Img<f32>, stream: &HipStream) -> Result<(), Ssim2Error> {
        let r = image(ref_lin);
        check(unsafe { sys::ssim2_set_reference_linear(self.ctx, &r, stream.raw()) })
    }

    /// lib.rs:543-590: score against the cached reference (the reference
    /// panics without one; this returns SSIM2_NO_REFERENCE).
    pub fn compute_with_reference_linear(&mut self, ref_lin: &impl Img<f32>, dis_lin: &impl Img<f32>,
                                         stream: &HipStream) -> Result<f64, Ssim2Error> {
        let (r, d) = (image(ref_lin), image(dis_lin));
        let mut s = 0.0;
        check(unsafe { sys::ssim2_compute_with_reference_linear(self.ctx, &r, &d, stream.raw(), &mut s) })?;
        Ok(s)
    }

    /// lib.rs:592-598.
    pub fn clear_reference(&mut self) {
        unsafe { sys::ssim2_clear_reference(self.ctx) };
    }

    /// lib.rs:600-602.
    pub fn has_reference(&self) -> bool {
        let mut h = 0;
        unsafe { sys::ssim2_has_reference(self.ctx, &mut h) };
        h != 0
    }

    fn check_host(&self, bytes: &[u8]) -> Result<(), Ssim2Error> {
        let need = self.ref_lin.width as usize * self.ref_lin.height as usize * 3;
        if bytes.len() < need {
            return Err(Ssim2Error { status: sys::SSIM2_DIM_MISMATCH,
                                    message: format!("{} host bytes, the image needs {need}", bytes.len()) });
        }
        Ok(())
    }
}

impl Drop for Ssimulacra2 {
    /// lib.rs:1195-1202: waits for the context's own work and frees it.
    fn drop(&mut self) {
        unsafe { sys::ssim2_destroy(self.ctx) };
    }
}

/// Many sRGB pairs of any sizes per launch (ssim2_batch_*; no counterpart in
/// ssimulacra2-cuda, the same-size Ssim2Batch of crates/ssim2-gpu is its
/// closest relative). Scores follow the FP64 evaluation of the metric.
pub struct Ssim2Batch {
    batch: *mut sys::ssim2_batch,
    last: usize,
}

unsafe impl Send for Ssim2Batch {}

impl Ssim2Batch {
    /// storage: sys::SSIM2_STORAGE_F32 or SSIM2_STORAGE_F16.
    pub fn new(max_pairs: u32, max_pixels: u64, storage: u32) -> Result<Self, Ssim2Error> {
        let o = sys::ssim2_batch_options {
            struct_size: std::mem::size_of::<sys::ssim2_batch_options>() as u32,
            storage,
            max_pairs,
            max_pixels,
            ..Default::default()
        };
        let mut b = ptr::null_mut();
        check(unsafe { sys::ssim2_batch_create(&o, &mut b) })?;
        Ok(Self { batch: b, last: 0 })
    }

    /// Scores of (reference, distorted) u8 sRGB image pairs, in order.
    pub fn compute<I: Img<u8>>(&mut self, pairs: &[(&I, &I)], stream: &HipStream) -> Result<Vec<f64>, Ssim2Error> {
        let v: Vec<sys::ssim2_pair> = pairs.iter().map(|(r, d)| sys::ssim2_pair { ref_: image(*r), dis: image(*d) }).collect();
        check(unsafe { sys::ssim2_batch_compute_async(self.batch, v.as_ptr(), v.len() as u32, stream.raw()) })?;
        self.last = v.len();
        let mut out = vec![0.0; self.last];
        check(unsafe { sys::ssim2_batch_get_scores(self.batch, out.as_mut_ptr(), out.len() as u32) })?;
        Ok(out)
    }
}

impl Drop for Ssim2Batch {
    fn drop(&mut self) {
        unsafe { sys::ssim2_batch_destroy(self.batch) };
    }
}
