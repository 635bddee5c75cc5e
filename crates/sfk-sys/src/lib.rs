//! Raw bindings of libsfk, the MI355X (gfx950) render and B3-Seg evidence
//! library (include/sfk.h).  Generated by gen_sys.py from the header; do not
//! edit by hand.  Every call returns SFK_OK (0) or a negative SFK_E_* code,
//! with sfk_last_error() holding the message; see the header for the
//! contract of each entry point.
#![allow(non_camel_case_types, non_snake_case, clippy::missing_safety_doc)]
#![no_std]

pub use core::ffi::{c_char, c_int, c_void};

pub const SFK_OK: c_int = 0;
pub const SFK_E_ARG: c_int = -1;
pub const SFK_E_CAPACITY: c_int = -2;
pub const SFK_E_HIP: c_int = -3;
pub const SFK_E_NOMEM: c_int = -4;
pub const SFK_ATTR_PLANES: u32 = 11u32;
pub const SFK_PROJ_STRIDE: u32 = 12u32;
pub const SFK_EIG_NOALLOC_MAX_SPLATS: u32 = 1u32 << 26;
pub const SFK_PLY_MAX_REST: u32 = 45u32;
pub const SFK_PLY_MAX_STRIDE: u32 = 128u32;

/// Opaque handle (`sfk_ctx` of include/sfk.h).
#[repr(C)]
pub struct sfk_ctx {
    _private: [u8; 0],
}

/// Opaque handle (`sfk_radix_scratch` of include/sfk.h).
#[repr(C)]
pub struct sfk_radix_scratch {
    _private: [u8; 0],
}

/// Opaque handle (`sfk_scan_scratch` of include/sfk.h).
#[repr(C)]
pub struct sfk_scan_scratch {
    _private: [u8; 0],
}

/// Opaque handle (`sfk_splats` of include/sfk.h).
#[repr(C)]
pub struct sfk_splats {
    _private: [u8; 0],
}

/// Opaque handle (`sfk_render_scratch` of include/sfk.h).
#[repr(C)]
pub struct sfk_render_scratch {
    _private: [u8; 0],
}

/// Opaque handle (`sfk_eig_round` of include/sfk.h).
#[repr(C)]
pub struct sfk_eig_round {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct sfk_camera {
    pub fov: [f32; 2],
    pub position: [f32; 3],
    pub rotation: [f32; 4],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct sfk_view {
    pub rot0: [f32; 3],
    pub rot1: [f32; 3],
    pub rot2: [f32; 3],
    pub trans: [f32; 3],
    pub focal: [f32; 2],
    pub world_pos: [f32; 3],
    pub img: [f32; 2],
    pub tile_bounds: [f32; 2],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct sfk_tile_isects {
    pub tile_ids: *const u32,
    pub gaussian_ids: *const u32,
    pub tile_starts: *const u32,
    pub d_counters: *const u32,
    pub capacity: u32,
    pub tiles_x: u32,
    pub tiles_y: u32,
    pub tile_w: u32,
    pub row_stride: u32,
    pub w: u32,
    pub h: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct sfk_render_buffers {
    pub depth_keys: *const u32,
    pub depth_order: *const u32,
    pub projected: *const f32,
    pub tile_counts: *const u32,
    pub bbox: *const u32,
    pub offsets: *const u32,
    pub counters: *const u32,
    pub bitmap: *const u32,
    pub rgba: *const f32,
    pub final_t: *const f32,
    pub n_contrib: *const u32,
    pub n: u32,
    pub w: u32,
    pub h: u32,
    pub row_stride: u32,
    pub capacity: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct sfk_ply_layout {
    pub n: u64,
    pub data_offset: u64,
    pub stride: u32,
    pub n_rest: u32,
    pub sh_planes: u32,
    pub sh_per_ch: u32,
    pub base: [i32; 14],
    pub rest: [i32; SFK_PLY_MAX_REST as usize],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct sfk_sog_meta {
    pub n: u64,
    pub centroid_px: u64,
    pub mins: [f32; 3],
    pub maxs: [f32; 3],
    pub sh_per_ch: u32,
    pub centroid_w: u32,
    pub palette_count: u32,
}

unsafe extern "C" {
    pub fn sfk_ctx_create(device: c_int, out: *mut *mut sfk_ctx) -> c_int;
    pub fn sfk_ctx_destroy(ctx: *mut sfk_ctx);
    pub fn sfk_last_error(ctx: *const sfk_ctx) -> *const c_char;
    pub fn sfk_ctx_stream(ctx: *mut sfk_ctx) -> *mut c_void;
    pub fn sfk_sync(ctx: *mut sfk_ctx) -> c_int;
    pub fn sfk_device_name(ctx: *mut sfk_ctx, buf: *mut c_char, len: c_int) -> c_int;
    pub fn sfk_alloc(ctx: *mut sfk_ctx, bytes: usize, out: *mut *mut c_void) -> c_int;
    pub fn sfk_free(ctx: *mut sfk_ctx, ptr: *mut c_void) -> c_int;
    pub fn sfk_upload(
        ctx: *mut sfk_ctx,
        dst: *mut c_void,
        src: *const c_void,
        bytes: usize,
    ) -> c_int;
    pub fn sfk_download(
        ctx: *mut sfk_ctx,
        dst: *mut c_void,
        src: *const c_void,
        bytes: usize,
    ) -> c_int;
    pub fn sfk_fill_u32(ctx: *mut sfk_ctx, buf: *mut u32, value: u32, count: usize) -> c_int;
    pub fn sfk_copy_u32(ctx: *mut sfk_ctx, dst: *mut u32, src: *const u32, count: usize) -> c_int;
    pub fn sfk_timer_start(ctx: *mut sfk_ctx) -> c_int;
    pub fn sfk_timer_stop_ms(ctx: *mut sfk_ctx, ms: *mut f32) -> c_int;
    pub fn sfk_prof_enable(ctx: *mut sfk_ctx, on: c_int) -> c_int;
    pub fn sfk_prof_report(ctx: *mut sfk_ctx, buf: *mut c_char, len: c_int) -> c_int;
    pub fn sfk_camera_view(
        cam: *const sfk_camera,
        w: u32,
        h: u32,
        tile_w: u32,
        out: *mut sfk_view,
    ) -> c_int;
    pub fn sfk_look_at(pos: *const f32, target: *const f32, quat_xyzw: *mut f32) -> c_int;
    pub fn sfk_bits_for(max_exclusive: u32) -> u32;
    pub fn sfk_radix_scratch_create(
        ctx: *mut sfk_ctx,
        max_elems: u32,
        key_bytes: u32,
        out: *mut *mut sfk_radix_scratch,
    ) -> c_int;
    pub fn sfk_radix_scratch_destroy(s: *mut sfk_radix_scratch);
    pub fn sfk_radix_set_digit_bits(s: *mut sfk_radix_scratch, digit_bits: u32) -> c_int;
    pub fn sfk_radix_set_bucket_finish(s: *mut sfk_radix_scratch, on: c_int) -> c_int;
    pub fn sfk_argsort_u32(
        ctx: *mut sfk_ctx,
        s: *mut sfk_radix_scratch,
        keys: *const u32,
        vals: *const u32,
        n: u32,
        d_n: *const u32,
        bits: u32,
        write_keys: c_int,
        out_keys: *mut *mut u32,
        out_vals: *mut *mut u32,
    ) -> c_int;
    pub fn sfk_argsort_u64(
        ctx: *mut sfk_ctx,
        s: *mut sfk_radix_scratch,
        keys: *const u64,
        vals: *const u32,
        n: u32,
        d_n: *const u32,
        bits: u32,
        write_keys: c_int,
        out_keys: *mut *mut u64,
        out_vals: *mut *mut u32,
    ) -> c_int;
    pub fn sfk_argsort_u32_segmented(
        ctx: *mut sfk_ctx,
        s: *mut sfk_radix_scratch,
        keys: *const u32,
        vals: *const u32,
        nseg: u32,
        seg_len: u32,
        bits: u32,
        write_keys: c_int,
        out_keys: *mut *mut u32,
        out_vals: *mut *mut u32,
    ) -> c_int;
    pub fn sfk_scan_scratch_create(
        ctx: *mut sfk_ctx,
        max_elems: u32,
        out: *mut *mut sfk_scan_scratch,
    ) -> c_int;
    pub fn sfk_scan_scratch_destroy(s: *mut sfk_scan_scratch);
    pub fn sfk_exclusive_scan_gather(
        ctx: *mut sfk_ctx,
        gids: *const u32,
        counts: *const u32,
        offsets: *mut u32,
        n: u32,
        s: *mut sfk_scan_scratch,
    ) -> c_int;
    pub fn sfk_exclusive_scan(
        ctx: *mut sfk_ctx,
        buf: *mut u32,
        n: u32,
        s: *mut sfk_scan_scratch,
    ) -> c_int;
    pub fn sfk_splats_create(
        ctx: *mut sfk_ctx,
        attrs: *const f32,
        sh: *const f32,
        n: u32,
        sh_per_ch: u32,
        out: *mut *mut sfk_splats,
    ) -> c_int;
    pub fn sfk_splats_destroy(s: *mut sfk_splats);
    pub fn sfk_splats_info(
        s: *const sfk_splats,
        n: *mut u32,
        sh_per_ch: *mut u32,
        bounds_min: *mut f32,
        bounds_max: *mut f32,
        d_attrs: *mut *const f32,
        d_sh: *mut *const f32,
    ) -> c_int;
    pub fn sfk_render_scratch_create(
        ctx: *mut sfk_ctx,
        n: u32,
        w: u32,
        h: u32,
        out: *mut *mut sfk_render_scratch,
    ) -> c_int;
    pub fn sfk_render_scratch_destroy(s: *mut sfk_render_scratch);
    pub fn sfk_render_scratch_set_tile_width(s: *mut sfk_render_scratch, tile_w: u32) -> c_int;
    pub fn sfk_render_scratch_reserve_isects(s: *mut sfk_render_scratch, isects: u32) -> c_int;
    pub fn sfk_render_scratch_set_digit_bits(s: *mut sfk_render_scratch, digit_bits: u32) -> c_int;
    pub fn sfk_render_scratch_set_binned(s: *mut sfk_render_scratch, on: c_int) -> c_int;
    pub fn sfk_render_scratch_set_tile_ids(s: *mut sfk_render_scratch, on: c_int) -> c_int;
    pub fn sfk_render_scratch_set_pair_raster(s: *mut sfk_render_scratch, on: c_int) -> c_int;
    pub fn sfk_prepare_isects(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        scratch: *mut sfk_render_scratch,
        cam: *const sfk_camera,
        w: u32,
        h: u32,
        out: *mut sfk_tile_isects,
    ) -> c_int;
    pub fn sfk_prepare_isects_view(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        scratch: *mut sfk_render_scratch,
        view: *const sfk_view,
        w: u32,
        h: u32,
        out: *mut sfk_tile_isects,
    ) -> c_int;
    pub fn sfk_tile_isects_counts(
        ctx: *mut sfk_ctx,
        isects: *const sfk_tile_isects,
        counts: *mut u32,
    ) -> c_int;
    pub fn sfk_finalize_rgb(
        ctx: *mut sfk_ctx,
        scratch: *mut sfk_render_scratch,
        isects: *const sfk_tile_isects,
    ) -> c_int;
    pub fn sfk_finalize_total(
        ctx: *mut sfk_ctx,
        scratch: *mut sfk_render_scratch,
        isects: *const sfk_tile_isects,
        scale: f32,
        resp: *mut u32,
    ) -> c_int;
    pub fn sfk_finalize_evidence(
        ctx: *mut sfk_ctx,
        scratch: *mut sfk_render_scratch,
        isects: *const sfk_tile_isects,
        scale: f32,
        mask: *const u32,
        words_per_row: u32,
        fg: *mut u32,
        bg: *mut u32,
    ) -> c_int;
    pub fn sfk_render_with(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        scratch: *mut sfk_render_scratch,
        cam: *const sfk_camera,
        w: u32,
        h: u32,
        bitmap: *mut *const u32,
        rgba: *mut *const f32,
    ) -> c_int;
    pub fn sfk_render_scratch_buffers(
        s: *const sfk_render_scratch,
        out: *mut sfk_render_buffers,
    ) -> c_int;
    pub fn sfk_fractional_bits(w: u32, h: u32) -> u32;
    pub fn sfk_mask_pack(
        bytes: *const u8,
        w: u32,
        h: u32,
        words: *mut u32,
        fg_pixels: *mut u32,
    ) -> c_int;
    pub fn sfk_render_responsibility(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        scratch: *mut sfk_render_scratch,
        cam: *const sfk_camera,
        w: u32,
        h: u32,
        resp: *mut u32,
        scale_out: *mut f32,
    ) -> c_int;
    pub fn sfk_accumulate_prepared(
        ctx: *mut sfk_ctx,
        scratch: *mut sfk_render_scratch,
        isects: *const sfk_tile_isects,
        mask: *const u32,
        words_per_row: u32,
        resp: *mut u32,
        fg: *mut u32,
        bg: *mut u32,
        scale_out: *mut f32,
    ) -> c_int;
    pub fn sfk_eig(
        ctx: *mut sfk_ctx,
        a: *const f32,
        b: *const f32,
        resp: *const u32,
        n: u32,
        scale: f32,
        out: *mut f64,
    ) -> c_int;
    pub fn sfk_eig_device(
        ctx: *mut sfk_ctx,
        a: *const f32,
        b: *const f32,
        resp: *const u32,
        n: u32,
        scale: f32,
        d_out: *mut f64,
    ) -> c_int;
    pub fn sfk_eig_device_noalloc(
        ctx: *mut sfk_ctx,
        a: *const f32,
        b: *const f32,
        resp: *const u32,
        n: u32,
        scale: f32,
        d_out: *mut f64,
    ) -> c_int;
    pub fn sfk_beta_update(
        ctx: *mut sfk_ctx,
        a: *mut f32,
        b: *mut f32,
        fg: *const u32,
        bg: *const u32,
        n: u32,
        scale: f32,
    ) -> c_int;
    pub fn sfk_localize(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        a: *const f32,
        b: *const f32,
        center3: *mut f32,
        radius: *mut f32,
        found: *mut c_int,
    ) -> c_int;
    pub fn sfk_ply_parse_header(
        bytes: *const c_void,
        len: usize,
        out: *mut sfk_ply_layout,
        err: *mut c_char,
        err_len: c_int,
    ) -> c_int;
    pub fn sfk_ply_decode(
        ctx: *mut sfk_ctx,
        layout: *const sfk_ply_layout,
        rows: *const f32,
        attrs: *mut f32,
        sh: *mut f32,
    ) -> c_int;
    pub fn sfk_splats_from_ply(
        ctx: *mut sfk_ctx,
        bytes: *const c_void,
        len: usize,
        out: *mut *mut sfk_splats,
    ) -> c_int;
    pub fn sfk_sog_decode(
        ctx: *mut sfk_ctx,
        meta: *const sfk_sog_meta,
        means_lo: *const u8,
        means_hi: *const u8,
        scales: *const u8,
        quats: *const u8,
        sh0: *const u8,
        labels: *const u8,
        centroids: *const u8,
        codebooks: *const f32,
        attrs: *mut f32,
        sh: *mut f32,
        status: *mut u32,
    ) -> c_int;
    pub fn sfk_splats_create_device(
        ctx: *mut sfk_ctx,
        d_attrs: *const f32,
        d_sh: *const f32,
        n: u32,
        sh_per_ch: u32,
        out: *mut *mut sfk_splats,
    ) -> c_int;
    pub fn sfk_render_posterior(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        scratch: *mut sfk_render_scratch,
        a: *const f32,
        b: *const f32,
        n_ab: u32,
        cam: *const sfk_camera,
        w: u32,
        h: u32,
        bitmap: *mut *const u32,
        rgba: *mut *const f32,
    ) -> c_int;
    pub fn sfk_tint(
        ctx: *mut sfk_ctx,
        splats: *mut sfk_splats,
        a: *const f32,
        b: *const f32,
        n_ab: u32,
        color: *const f32,
    ) -> c_int;
    pub fn sfk_save_colors(ctx: *mut sfk_ctx, splats: *const sfk_splats, saved: *mut f32) -> c_int;
    pub fn sfk_restore_colors(
        ctx: *mut sfk_ctx,
        splats: *mut sfk_splats,
        saved: *const f32,
    ) -> c_int;
    pub fn sfk_eig_round_create(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        cams: *const sfk_camera,
        widths: *const u32,
        heights: *const u32,
        nviews: u32,
        tile_w: u32,
        digit_bits: u32,
        out: *mut *mut sfk_eig_round,
    ) -> c_int;
    pub fn sfk_eig_round_destroy(r: *mut sfk_eig_round);
    pub fn sfk_eig_round_run(
        ctx: *mut sfk_ctx,
        r: *mut sfk_eig_round,
        a: *const f32,
        b: *const f32,
        use_graph: c_int,
        eig_out: *mut f64,
        best_out: *mut u32,
    ) -> c_int;
    pub fn sfk_eig_round_set_binned(r: *mut sfk_eig_round, on: c_int) -> c_int;
    pub fn sfk_eig_round_set_pair_raster(r: *mut sfk_eig_round, on: c_int) -> c_int;
    pub fn sfk_eig_round_info(
        r: *const sfk_eig_round,
        view: u32,
        resp: *mut *const u32,
        scale: *mut f32,
        capacity: *mut u32,
        total_tiles: *mut u32,
    ) -> c_int;
    pub fn sfk_eig_round_counters(
        ctx: *mut sfk_ctx,
        r: *const sfk_eig_round,
        counters: *mut u32,
    ) -> c_int;
    pub fn sfk_eig_round_autotune(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        cams: *const sfk_camera,
        widths: *const u32,
        heights: *const u32,
        nviews: u32,
        a: *const f32,
        b: *const f32,
        reps: u32,
        best_tile_w: *mut u32,
        best_digit_bits: *mut u32,
        ms_out: *mut f32,
    ) -> c_int;
    pub fn sfk_eig_round_plain(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        scratch: *mut sfk_render_scratch,
        cams: *const sfk_camera,
        widths: *const u32,
        heights: *const u32,
        nviews: u32,
        a: *const f32,
        b: *const f32,
        resp: *mut u32,
        eig_out: *mut f64,
        best_out: *mut u32,
    ) -> c_int;
    pub fn sfk_loss_l2(
        ctx: *mut sfk_ctx,
        rgba: *const f32,
        target: *const f32,
        pixels: u32,
        d_rgba: *mut f32,
        d_loss: *mut f64,
    ) -> c_int;
    pub fn sfk_render_backward(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        scratch: *mut sfk_render_scratch,
        isects: *const sfk_tile_isects,
        d_rgba: *const f32,
        grad_attrs: *mut f32,
        grad_sh: *mut f32,
    ) -> c_int;
    pub fn sfk_render_backward_l2(
        ctx: *mut sfk_ctx,
        splats: *const sfk_splats,
        scratch: *mut sfk_render_scratch,
        isects: *const sfk_tile_isects,
        rgba: *const f32,
        target: *const f32,
        grad_attrs: *mut f32,
        grad_sh: *mut f32,
        d_loss: *mut f64,
    ) -> c_int;
    pub fn sfk_project(
        ctx: *mut sfk_ctx,
        view: *const sfk_view,
        attrs: *const f32,
        sh: *const f32,
        n: u32,
        sh_per_ch: u32,
        tile_w: u32,
        depth_keys: *mut u32,
        projected: *mut f32,
        tile_counts: *mut u32,
        bbox: *mut u32,
        counters: *mut u32,
    ) -> c_int;
    pub fn sfk_map_isects(
        ctx: *mut sfk_ctx,
        depth_order: *const u32,
        tile_counts: *const u32,
        bbox: *const u32,
        offsets: *const u32,
        n: u32,
        tiles_per_row: u32,
        max_isects: u32,
        tile_ids: *mut u32,
        gaussian_ids: *mut u32,
    ) -> c_int;
    pub fn sfk_tile_starts(
        ctx: *mut sfk_ctx,
        tile_ids: *const u32,
        n: u32,
        d_n: *const u32,
        num_tiles: u32,
        starts: *mut u32,
    ) -> c_int;
}
