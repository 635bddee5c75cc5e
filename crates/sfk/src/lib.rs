//! Safe layer over `sfk-sys` in the shape of splatfield's own entry points,
//! so that the render-frame path, `crates/sort` callers and `src/seg` keep
//! their call shapes when they move onto the MI355X kernels:
//!
//! * `Splats::{new, prepare_isects, finalize, render_with}` and
//!   `RenderScratch::new` (src/render.rs:143-460);
//! * `RadixScratch::{new, argsort}` and `bits_for`
//!   (crates/sort/src/sort.rs:259-381);
//! * `fractional_bits`, `Accumulators`, `Mask::from_bytes`,
//!   `Splats::{render_responsibility, accumulate_prepared}`
//!   (src/seg/mod.rs:25-210), `BetaState::{new_uniform, eig, update}` and
//!   `map_labels` (src/seg/beta.rs:214-297).
//!
//! Errors are the library's return codes with its message; the reference's
//! panics (undersized scratch, zero splats) arrive as `Err`.  Outputs that
//! alias a scratch (`DeviceSlice`) borrow it, so the borrow checker enforces
//! the reference's "valid until the next call on this scratch" rule.
use std::ffi::CStr;
use std::marker::PhantomData;
use std::mem::size_of;
use std::ptr;

use sfk_sys as sys;

/// A failed libsfk call: its SFK_E_* code and message.
#[derive(Clone, Debug)]
pub struct Error {
    pub code: i32,
    pub message: String,
}

impl std::fmt::Display for Error {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "libsfk error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for Error {}

pub type Result<T> = std::result::Result<T, Error>;

/// One device and one HIP stream: every call made through it is ordered on
/// that stream (the single-stream policy of src/lib.rs:20-27).  Calls on one
/// context must not overlap, as `&mut self` methods of the reference imply.
pub struct Context {
    raw: *mut sys::sfk_ctx,
}

impl Context {
    pub fn new(device: i32) -> Result<Self> {
        let mut raw = ptr::null_mut();
        let rc = unsafe { sys::sfk_ctx_create(device, &mut raw) };
        if rc != sys::SFK_OK {
            return Err(Error { code: rc, message: format!("no context on device {device}") });
        }
        Ok(Self { raw })
    }

    fn check(&self, rc: i32) -> Result<()> {
        if rc == sys::SFK_OK {
            return Ok(());
        }
        let message = unsafe { CStr::from_ptr(sys::sfk_last_error(self.raw)) }
            .to_string_lossy()
            .into_owned();
        Err(Error { code: rc, message })
    }

    /// Wait for everything enqueued so far.
    pub fn sync(&self) -> Result<()> {
        self.check(unsafe { sys::sfk_sync(self.raw) })
    }

    pub fn as_raw(&self) -> *mut sys::sfk_ctx {
        self.raw
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe { sys::sfk_ctx_destroy(self.raw) }
    }
}

/// Owned device buffer of `len` elements (GpuTensor,
/// crates/sort/src/tensor.rs:29-77): `write` and `read_vec` synchronize.
pub struct GpuTensor<'c, T: Copy> {
    ctx: &'c Context,
    ptr: *mut T,
    len: usize,
}

impl<'c, T: Copy> GpuTensor<'c, T> {
    pub fn empty(ctx: &'c Context, len: usize) -> Result<Self> {
        let mut p: *mut sys::c_void = ptr::null_mut();
        ctx.check(unsafe { sys::sfk_alloc(ctx.raw, len.max(1) * size_of::<T>(), &mut p) })?;
        Ok(Self { ctx, ptr: p as *mut T, len })
    }

    pub fn from_slice(ctx: &'c Context, data: &[T]) -> Result<Self> {
        let t = Self::empty(ctx, data.len())?;
        t.write(data)?;
        Ok(t)
    }

    pub fn write(&self, data: &[T]) -> Result<()> {
        assert!(data.len() <= self.len, "write past the end of a device buffer");
        let bytes = data.len() * size_of::<T>();
        self.ctx.check(unsafe {
            sys::sfk_upload(self.ctx.raw, self.ptr as *mut sys::c_void,
                            data.as_ptr() as *const sys::c_void, bytes)
        })
    }

    pub fn read_vec(&self) -> Result<Vec<T>> {
        read_device(self.ctx, self.ptr, self.len)
    }

    pub fn len(&self) -> usize {
        self.len
    }

    pub fn is_empty(&self) -> bool {
        self.len == 0
    }

    pub fn as_ptr(&self) -> *const T {
        self.ptr
    }

    pub fn as_mut_ptr(&self) -> *mut T {
        self.ptr
    }
}

impl<T: Copy> Drop for GpuTensor<'_, T> {
    fn drop(&mut self) {
        unsafe {
            sys::sfk_free(self.ctx.raw, self.ptr as *mut sys::c_void);
        }
    }
}

fn read_device<T: Copy>(ctx: &Context, src: *const T, len: usize) -> Result<Vec<T>> {
    let mut v: Vec<T> = Vec::with_capacity(len);
    ctx.check(unsafe {
        sys::sfk_download(ctx.raw, v.as_mut_ptr() as *mut sys::c_void,
                          src as *const sys::c_void, len * size_of::<T>())
    })?;
    unsafe { v.set_len(len) };
    Ok(v)
}

/// Device memory owned by a scratch object (a returned tensor that aliases
/// it, render.rs:366-370, sort.rs:278-281): valid while the borrow lasts.
pub struct DeviceSlice<'s, T: Copy> {
    ctx: &'s Context,
    ptr: *const T,
    len: usize,
    _scratch: PhantomData<&'s ()>,
}

impl<T: Copy> DeviceSlice<'_, T> {
    pub fn read_vec(&self) -> Result<Vec<T>> {
        read_device(self.ctx, self.ptr, self.len)
    }

    pub fn as_ptr(&self) -> *const T {
        self.ptr
    }

    pub fn len(&self) -> usize {
        self.len
    }

    pub fn is_empty(&self) -> bool {
        self.len == 0
    }
}

// ------------------------------------------------------------------ camera

/// Pinhole camera (src/camera.rs:4-19): per-axis fov, position, rotation
/// (x, y, z, w) from camera to world; the camera looks along local +Z.
#[derive(Clone, Copy, Debug)]
pub struct Camera {
    pub fov: [f32; 2],
    pub position: [f32; 3],
    pub rotation: [f32; 4],
}

impl Camera {
    pub const BASE_FOV: f32 = 0.8;

    /// Camera::look_at (camera.rs:57-63).
    pub fn look_at(position: [f32; 3], target: [f32; 3]) -> Self {
        let mut q = [0f32; 4];
        unsafe { sys::sfk_look_at(position.as_ptr(), target.as_ptr(), q.as_mut_ptr()) };
        Self { fov: [Self::BASE_FOV; 2], position, rotation: q }
    }

    fn raw(&self) -> sys::sfk_camera {
        sys::sfk_camera { fov: self.fov, position: self.position, rotation: self.rotation }
    }
}

// ------------------------------------------------------------------ render

/// Uploaded splats (Splats::new, render.rs:241-269): 11 field-major attribute
/// planes and 3 * sh_per_ch SH planes.
pub struct Splats<'c> {
    ctx: &'c Context,
    raw: *mut sys::sfk_splats,
    n: u32,
}

/// Per-(splat count, image size) frame buffers (RenderScratch,
/// render.rs:143-197), reused across frames.
pub struct RenderScratch<'c> {
    ctx: &'c Context,
    raw: *mut sys::sfk_render_scratch,
}

/// Sorted per-tile isect lists (TileIsects, render.rs:463-471).  They alias
/// the scratch that prepared them and are valid until its next prepare; the
/// library rejects lists prepared by another scratch.
#[derive(Clone, Copy, Debug)]
pub struct TileIsects {
    raw: sys::sfk_tile_isects,
}

/// Which buffers a finalize fills (Finalize, render.rs:220-239).
pub enum Finalize<'a, 'c> {
    Rgb,
    Total { resp: &'a GpuTensor<'c, u32>, scale: f32 },
    Evidence { mask: &'a Mask<'c>, fg: &'a GpuTensor<'c, u32>, bg: &'a GpuTensor<'c, u32>, scale: f32 },
}

impl<'c> RenderScratch<'c> {
    pub fn new(ctx: &'c Context, n: u32, img: [u32; 2]) -> Result<Self> {
        let mut raw = ptr::null_mut();
        ctx.check(unsafe { sys::sfk_render_scratch_create(ctx.raw, n, img[0], img[1], &mut raw) })?;
        Ok(Self { ctx, raw })
    }

    /// Tile width 8, 16 (the reference's TILE_WIDTH) or 32.
    pub fn set_tile_width(&mut self, tile_w: u32) -> Result<()> {
        self.ctx.check(unsafe { sys::sfk_render_scratch_set_tile_width(self.raw, tile_w) })
    }

    /// Whether the binned emission writes per-isect tile ids (default on); a
    /// training loop that only rasterizes switches them off.
    pub fn set_tile_ids(&mut self, on: bool) -> Result<()> {
        self.ctx.check(unsafe { sys::sfk_render_scratch_set_tile_ids(self.raw, on as i32) })
    }

    /// Grow the isect capacity now so the first frame renders untruncated.
    pub fn reserve_isects(&mut self, isects: u32) -> Result<()> {
        self.ctx.check(unsafe { sys::sfk_render_scratch_reserve_isects(self.raw, isects) })
    }

    fn buffers(&self) -> Result<sys::sfk_render_buffers> {
        let mut b = std::mem::MaybeUninit::<sys::sfk_render_buffers>::uninit();
        self.ctx.check(unsafe { sys::sfk_render_scratch_buffers(self.raw, b.as_mut_ptr()) })?;
        Ok(unsafe { b.assume_init() })
    }
}

impl Drop for RenderScratch<'_> {
    fn drop(&mut self) {
        unsafe { sys::sfk_render_scratch_destroy(self.raw) }
    }
}

impl<'c> Splats<'c> {
    pub fn new(ctx: &'c Context, attributes: &[f32], sh_coeffs: &[f32]) -> Result<Self> {
        let planes = sys::SFK_ATTR_PLANES as usize;
        let n = attributes.len() / planes;
        if n == 0 || attributes.len() != n * planes || sh_coeffs.len() % (3 * n) != 0 {
            return Err(Error { code: sys::SFK_E_ARG, message: "Splats::new: bad plane sizes".into() });
        }
        let per_ch = (sh_coeffs.len() / (3 * n)) as u32;
        let mut raw = ptr::null_mut();
        ctx.check(unsafe {
            sys::sfk_splats_create(ctx.raw, attributes.as_ptr(), sh_coeffs.as_ptr(), n as u32,
                                   per_ch, &mut raw)
        })?;
        Ok(Self { ctx, raw, n: n as u32 })
    }

    /// Splats from the bytes of a binary little-endian 3DGS PLY file
    /// (src/ply.rs:9-112), decoded on the device.
    pub fn from_ply(ctx: &'c Context, bytes: &[u8]) -> Result<Self> {
        let mut raw = ptr::null_mut();
        ctx.check(unsafe {
            sys::sfk_splats_from_ply(ctx.raw, bytes.as_ptr() as *const sys::c_void, bytes.len(),
                                     &mut raw)
        })?;
        let mut n = 0u32;
        unsafe {
            sys::sfk_splats_info(raw, &mut n, ptr::null_mut(), ptr::null_mut(), ptr::null_mut(),
                                 ptr::null_mut(), ptr::null_mut());
        }
        Ok(Self { ctx, raw, n })
    }

    pub fn len(&self) -> u32 {
        self.n
    }

    pub fn is_empty(&self) -> bool {
        self.n == 0
    }

    /// Splats::prepare_isects (render.rs:271-364): project, depth sort, isect
    /// emission and tile ranges, with no host synchronization.
    pub fn prepare_isects(&self, scratch: &mut RenderScratch<'c>, camera: &Camera,
                          img: [u32; 2]) -> Result<TileIsects> {
        let cam = camera.raw();
        let mut raw = std::mem::MaybeUninit::<sys::sfk_tile_isects>::uninit();
        self.ctx.check(unsafe {
            sys::sfk_prepare_isects(self.ctx.raw, self.raw, scratch.raw, &cam, img[0], img[1],
                                    raw.as_mut_ptr())
        })?;
        Ok(TileIsects { raw: unsafe { raw.assume_init() } })
    }

    /// finalize (render.rs:382-460) over isects this scratch prepared.
    pub fn finalize(&self, scratch: &mut RenderScratch<'c>, isects: &TileIsects,
                    mode: Finalize<'_, 'c>) -> Result<()> {
        let rc = unsafe {
            match mode {
                Finalize::Rgb => sys::sfk_finalize_rgb(self.ctx.raw, scratch.raw, &isects.raw),
                Finalize::Total { resp, scale } => sys::sfk_finalize_total(
                    self.ctx.raw, scratch.raw, &isects.raw, scale, resp.as_mut_ptr()),
                Finalize::Evidence { mask, fg, bg, scale } => sys::sfk_finalize_evidence(
                    self.ctx.raw, scratch.raw, &isects.raw, scale, mask.words.as_ptr(),
                    mask.words_per_row, fg.as_mut_ptr(), bg.as_mut_ptr()),
            }
        };
        self.ctx.check(rc)
    }

    /// Splats::render_with (render.rs:366-380): the RGBA8 bitmap (rows of
    /// `row_stride` words), aliasing the scratch.
    pub fn render_with<'s>(&self, scratch: &'s mut RenderScratch<'c>, camera: &Camera,
                           img: [u32; 2]) -> Result<DeviceSlice<'s, u32>> {
        let cam = camera.raw();
        let mut bitmap: *const u32 = ptr::null();
        self.ctx.check(unsafe {
            sys::sfk_render_with(self.ctx.raw, self.raw, scratch.raw, &cam, img[0], img[1],
                                 &mut bitmap, ptr::null_mut())
        })?;
        let b = scratch.buffers()?;
        Ok(DeviceSlice { ctx: self.ctx, ptr: bitmap, len: (b.h * b.row_stride) as usize,
                         _scratch: PhantomData })
    }

    /// Splats::render_responsibility (src/seg/mod.rs:146-159): total epsilon
    /// of every splat at this camera into `acc.bits`, scale tracked in acc.
    pub fn render_responsibility(&self, scratch: &mut RenderScratch<'c>,
                                 acc: &mut Accumulators<'c>, camera: &Camera,
                                 img: [u32; 2]) -> Result<()> {
        let cam = camera.raw();
        let mut scale = 0f32;
        self.ctx.check(unsafe {
            sys::sfk_render_responsibility(self.ctx.raw, self.raw, scratch.raw, &cam, img[0],
                                           img[1], acc.bits.as_mut_ptr(), &mut scale)
        })?;
        acc.scale = scale;
        Ok(())
    }

    /// accumulate_prepared (src/seg/mod.rs:161-210): over isects already
    /// prepared for this camera; no mask -> total into acc.bits, else the
    /// fg / bg split.
    pub fn accumulate_prepared(&self, scratch: &mut RenderScratch<'c>, isects: &TileIsects,
                               acc: &mut Accumulators<'c>, mask: Option<&Mask<'c>>) -> Result<()> {
        let mut scale = 0f32;
        let rc = unsafe {
            match mask {
                None => sys::sfk_accumulate_prepared(
                    self.ctx.raw, scratch.raw, &isects.raw, ptr::null(), 0,
                    acc.bits.as_mut_ptr(), ptr::null_mut(), ptr::null_mut(), &mut scale),
                Some(m) => sys::sfk_accumulate_prepared(
                    self.ctx.raw, scratch.raw, &isects.raw, m.words.as_ptr(), m.words_per_row,
                    ptr::null_mut(), acc.fg.as_mut_ptr(), acc.bg.as_mut_ptr(), &mut scale),
            }
        };
        self.ctx.check(rc)?;
        acc.scale = scale;
        Ok(())
    }
}

impl Drop for Splats<'_> {
    fn drop(&mut self) {
        unsafe { sys::sfk_splats_destroy(self.raw) }
    }
}

// -------------------------------------------------------------------- sort

/// bits_for (sort.rs:259-261).
pub fn bits_for(max_exclusive: u32) -> u32 {
    unsafe { sys::sfk_bits_for(max_exclusive) }
}

/// Reusable buffers of the stable LSD radix argsort (RadixScratch,
/// sort.rs:278-300), for u32 or u64 keys.
pub struct RadixScratch<'c> {
    ctx: &'c Context,
    raw: *mut sys::sfk_radix_scratch,
}

impl<'c> RadixScratch<'c> {
    /// RadixScratch::new: u32 keys, at most max_elems of them.
    pub fn new(ctx: &'c Context, max_elems: u32) -> Result<Self> {
        Self::with_key_bytes(ctx, max_elems, 4)
    }

    /// The same for u64 keys.
    pub fn new_u64(ctx: &'c Context, max_elems: u32) -> Result<Self> {
        Self::with_key_bytes(ctx, max_elems, 8)
    }

    fn with_key_bytes(ctx: &'c Context, max_elems: u32, key_bytes: u32) -> Result<Self> {
        let mut raw = ptr::null_mut();
        ctx.check(unsafe { sys::sfk_radix_scratch_create(ctx.raw, max_elems, key_bytes, &mut raw) })?;
        Ok(Self { ctx, raw })
    }

    /// RadixScratch::argsort (sort.rs:302-381): stable argsort of keys[0..n)
    /// on their low `bits` (keys < 2^bits), carrying `vals` (None: the
    /// identity).  Returns (sorted keys if write_keys, sorted values), both
    /// aliasing this scratch (or the inputs when nothing moves).
    pub fn argsort<'s>(&'s mut self, keys: &GpuTensor<'c, u32>, vals: Option<&GpuTensor<'c, u32>>,
                       n: u32, bits: u32, write_keys: bool)
                       -> Result<(Option<DeviceSlice<'s, u32>>, DeviceSlice<'s, u32>)> {
        let mut ok: *mut u32 = ptr::null_mut();
        let mut ov: *mut u32 = ptr::null_mut();
        let v = vals.map_or(ptr::null(), |t| t.as_ptr());
        self.ctx.check(unsafe {
            sys::sfk_argsort_u32(self.ctx.raw, self.raw, keys.as_ptr(), v, n, ptr::null(), bits,
                                 write_keys as i32, &mut ok, &mut ov)
        })?;
        let k = write_keys.then(|| DeviceSlice { ctx: self.ctx, ptr: ok as *const u32,
                                                 len: n as usize, _scratch: PhantomData });
        Ok((k, DeviceSlice { ctx: self.ctx, ptr: ov as *const u32, len: n as usize,
                             _scratch: PhantomData }))
    }

    /// u64 sorts: two passes on the top 16 bits, then every group sorted on
    /// chip when all fit (default on; a tuning knob, same results).
    pub fn set_bucket_finish(&mut self, on: bool) -> Result<()> {
        self.ctx.check(unsafe { sys::sfk_radix_set_bucket_finish(self.raw, on as i32) })
    }

    /// The same over u64 keys (a scratch made by `new_u64`).
    pub fn argsort_u64<'s>(&'s mut self, keys: &GpuTensor<'c, u64>,
                           vals: Option<&GpuTensor<'c, u32>>, n: u32, bits: u32, write_keys: bool)
                           -> Result<(Option<DeviceSlice<'s, u64>>, DeviceSlice<'s, u32>)> {
        let mut ok: *mut u64 = ptr::null_mut();
        let mut ov: *mut u32 = ptr::null_mut();
        let v = vals.map_or(ptr::null(), |t| t.as_ptr());
        self.ctx.check(unsafe {
            sys::sfk_argsort_u64(self.ctx.raw, self.raw, keys.as_ptr(), v, n, ptr::null(), bits,
                                 write_keys as i32, &mut ok, &mut ov)
        })?;
        let k = write_keys.then(|| DeviceSlice { ctx: self.ctx, ptr: ok as *const u64,
                                                 len: n as usize, _scratch: PhantomData });
        Ok((k, DeviceSlice { ctx: self.ctx, ptr: ov as *const u32, len: n as usize,
                             _scratch: PhantomData }))
    }
}

impl Drop for RadixScratch<'_> {
    fn drop(&mut self) {
        unsafe { sys::sfk_radix_scratch_destroy(self.raw) }
    }
}

// --------------------------------------------------------------------- seg

/// fractional_bits (src/seg/mod.rs:25-34): 32 - log2(next_pow2(w h)).
pub fn fractional_bits(img: [u32; 2]) -> u32 {
    unsafe { sys::sfk_fractional_bits(img[0], img[1]) }
}

/// Fixed-point epsilon accumulators (src/seg/mod.rs:45-85).
pub struct Accumulators<'c> {
    pub bits: GpuTensor<'c, u32>,
    pub fg: GpuTensor<'c, u32>,
    pub bg: GpuTensor<'c, u32>,
    pub scale: f32,
}

impl<'c> Accumulators<'c> {
    pub fn new(ctx: &'c Context, total: usize) -> Result<Self> {
        Ok(Self {
            bits: GpuTensor::empty(ctx, total)?,
            fg: GpuTensor::empty(ctx, total)?,
            bg: GpuTensor::empty(ctx, total)?,
            scale: 1.0,
        })
    }
}

/// Bit-packed 2D mask (src/seg/mod.rs:105-144).
pub struct Mask<'c> {
    words: GpuTensor<'c, u32>,
    words_per_row: u32,
    pub fg_pixels: u32,
}

impl<'c> Mask<'c> {
    /// Mask::from_bytes: one byte per pixel, row-major, nonzero = inside.
    pub fn from_bytes(ctx: &'c Context, img: [u32; 2], bytes: &[u8]) -> Result<Self> {
        if bytes.len() != (img[0] * img[1]) as usize {
            return Err(Error { code: sys::SFK_E_ARG, message: "mask: one byte per pixel".into() });
        }
        let wpr = img[0].div_ceil(32);
        let mut host = vec![0u32; (wpr * img[1]) as usize];
        let mut fg = 0u32;
        ctx.check(unsafe {
            sys::sfk_mask_pack(bytes.as_ptr(), img[0], img[1], host.as_mut_ptr(), &mut fg)
        })?;
        Ok(Self { words: GpuTensor::from_slice(ctx, &host)?, words_per_row: wpr, fg_pixels: fg })
    }
}

/// Per-splat Beta posteriors (BetaState, src/seg/beta.rs:214-287) with the
/// entropy evaluated in f64 on the device.
pub struct BetaState<'c> {
    ctx: &'c Context,
    pub a: GpuTensor<'c, f32>,
    pub b: GpuTensor<'c, f32>,
}

impl<'c> BetaState<'c> {
    /// The uniform Beta(1, 1) prior (beta.rs:227-238).
    pub fn new_uniform(ctx: &'c Context, total: usize) -> Result<Self> {
        let ones = vec![1f32; total];
        Ok(Self { ctx, a: GpuTensor::from_slice(ctx, &ones)?, b: GpuTensor::from_slice(ctx, &ones)? })
    }

    /// Analytic EIG of one responsibility map: sum over splats of
    /// H(a, b) - H(a k, b k), k = 1 + eps / (a + b) (beta.rs:240-268).
    pub fn eig(&self, acc: &Accumulators<'c>) -> Result<f64> {
        let mut out = 0f64;
        self.ctx.check(unsafe {
            sys::sfk_eig(self.ctx.raw, self.a.as_ptr(), self.b.as_ptr(), acc.bits.as_ptr(),
                         self.a.len() as u32, acc.scale, &mut out)
        })?;
        Ok(out)
    }

    /// The conjugate update a += fg / scale, b += bg / scale (beta.rs:270-286).
    pub fn update(&self, acc: &Accumulators<'c>) -> Result<()> {
        self.ctx.check(unsafe {
            sys::sfk_beta_update(self.ctx.raw, self.a.as_mut_ptr(), self.b.as_mut_ptr(),
                                 acc.fg.as_ptr(), acc.bg.as_ptr(), self.a.len() as u32, acc.scale)
        })
    }
}

/// MAP foreground labels, a > b, ties background (beta.rs:289-297).
pub fn map_labels(a: &[f32], b: &[f32]) -> Vec<bool> {
    assert_eq!(a.len(), b.len(), "posterior length mismatch");
    a.iter().zip(b).map(|(&a, &b)| a > b).collect()
}
