"""ctypes binding of libsfk.so (include/sfk.h).

The library is loaded from this package directory; there is no fallback: a
missing or unloadable libsfk.so raises at import of any GPU entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsfk.so")

P = C.c_void_p
u32, i32, f32, f64, sz = C.c_uint32, C.c_int, C.c_float, C.c_double, C.c_size_t
PP = C.POINTER(C.c_void_p)


class Camera(C.Structure):
    _fields_ = [("fov", f32 * 2), ("position", f32 * 3), ("rotation", f32 * 4)]


class View(C.Structure):
    _fields_ = [
        ("rot0", f32 * 3), ("rot1", f32 * 3), ("rot2", f32 * 3), ("trans", f32 * 3),
        ("focal", f32 * 2), ("world_pos", f32 * 3), ("img", f32 * 2), ("tile_bounds", f32 * 2),
    ]


class TileIsects(C.Structure):
    _fields_ = [
        ("tile_ids", P), ("gaussian_ids", P), ("tile_starts", P), ("d_counters", P),
        ("capacity", u32), ("tiles_x", u32), ("tiles_y", u32), ("tile_w", u32),
        ("row_stride", u32), ("w", u32), ("h", u32),
    ]


class PlyLayout(C.Structure):
    _fields_ = [("n", C.c_uint64), ("data_offset", C.c_uint64), ("stride", u32),
                ("n_rest", u32), ("sh_planes", u32), ("sh_per_ch", u32),
                ("base", C.c_int32 * 14), ("rest", C.c_int32 * 45)]


class SogMeta(C.Structure):
    _fields_ = [("n", C.c_uint64), ("centroid_px", C.c_uint64), ("mins", f32 * 3),
                ("maxs", f32 * 3), ("sh_per_ch", u32), ("centroid_w", u32),
                ("palette_count", u32)]


class RenderBuffers(C.Structure):
    _fields_ = [
        ("depth_keys", P), ("depth_order", P), ("projected", P), ("tile_counts", P),
        ("bbox", P), ("offsets", P), ("counters", P), ("bitmap", P), ("rgba", P),
        ("final_t", P), ("n_contrib", P),
        ("n", u32), ("w", u32), ("h", u32), ("row_stride", u32), ("capacity", u32),
    ]


# name -> (restype, argtypes); every declaration of include/sfk.h.
SIGNATURES = {
    "sfk_ctx_create": (i32, [i32, PP]),
    "sfk_ctx_destroy": (None, [P]),
    "sfk_last_error": (C.c_char_p, [P]),
    "sfk_ctx_stream": (P, [P]),
    "sfk_sync": (i32, [P]),
    "sfk_device_name": (i32, [P, C.c_char_p, i32]),
    "sfk_alloc": (i32, [P, sz, PP]),
    "sfk_free": (i32, [P, P]),
    "sfk_upload": (i32, [P, P, P, sz]),
    "sfk_download": (i32, [P, P, P, sz]),
    "sfk_fill_u32": (i32, [P, P, u32, sz]),
    "sfk_copy_u32": (i32, [P, P, P, sz]),
    "sfk_timer_start": (i32, [P]),
    "sfk_timer_stop_ms": (i32, [P, C.POINTER(f32)]),
    "sfk_prof_enable": (i32, [P, i32]),
    "sfk_prof_report": (i32, [P, C.c_char_p, i32]),
    "sfk_camera_view": (i32, [C.POINTER(Camera), u32, u32, u32, C.POINTER(View)]),
    "sfk_look_at": (i32, [P, P, P]),
    "sfk_bits_for": (u32, [u32]),
    "sfk_radix_scratch_create": (i32, [P, u32, u32, PP]),
    "sfk_radix_scratch_destroy": (None, [P]),
    "sfk_radix_set_bucket_finish": (i32, [P, i32]),
    "sfk_radix_set_digit_bits": (i32, [P, u32]),
    "sfk_argsort_u32": (i32, [P, P, P, P, u32, P, u32, i32, PP, PP]),
    "sfk_argsort_u64": (i32, [P, P, P, P, u32, P, u32, i32, PP, PP]),
    "sfk_argsort_u32_segmented": (i32, [P, P, P, P, u32, u32, u32, i32, PP, PP]),
    "sfk_scan_scratch_create": (i32, [P, u32, PP]),
    "sfk_scan_scratch_destroy": (None, [P]),
    "sfk_exclusive_scan_gather": (i32, [P, P, P, P, u32, P]),
    "sfk_exclusive_scan": (i32, [P, P, u32, P]),
    "sfk_splats_create": (i32, [P, P, P, u32, u32, PP]),
    "sfk_splats_destroy": (None, [P]),
    "sfk_splats_info": (i32, [P, P, P, P, P, PP, PP]),
    "sfk_render_scratch_create": (i32, [P, u32, u32, u32, PP]),
    "sfk_render_scratch_destroy": (None, [P]),
    "sfk_render_scratch_set_tile_width": (i32, [P, u32]),
    "sfk_render_scratch_reserve_isects": (i32, [P, u32]),
    "sfk_render_scratch_set_digit_bits": (i32, [P, u32]),
    "sfk_render_scratch_set_binned": (i32, [P, i32]),
    "sfk_render_scratch_set_tile_ids": (i32, [P, i32]),
    "sfk_render_scratch_set_pair_raster": (i32, [P, i32]),
    "sfk_prepare_isects": (i32, [P, P, P, C.POINTER(Camera), u32, u32, C.POINTER(TileIsects)]),
    "sfk_prepare_isects_view": (i32, [P, P, P, C.POINTER(View), u32, u32,
                                      C.POINTER(TileIsects)]),
    "sfk_tile_isects_counts": (i32, [P, C.POINTER(TileIsects), P]),
    "sfk_finalize_rgb": (i32, [P, P, C.POINTER(TileIsects)]),
    "sfk_finalize_total": (i32, [P, P, C.POINTER(TileIsects), f32, P]),
    "sfk_finalize_evidence": (i32, [P, P, C.POINTER(TileIsects), f32, P, u32, P, P]),
    "sfk_render_with": (i32, [P, P, P, C.POINTER(Camera), u32, u32, PP, PP]),
    "sfk_render_scratch_buffers": (i32, [P, C.POINTER(RenderBuffers)]),
    "sfk_fractional_bits": (u32, [u32, u32]),
    "sfk_mask_pack": (i32, [P, u32, u32, P, P]),
    "sfk_render_responsibility": (i32, [P, P, P, C.POINTER(Camera), u32, u32, P, P]),
    "sfk_accumulate_prepared": (i32, [P, P, C.POINTER(TileIsects), P, u32, P, P, P, P]),
    "sfk_eig": (i32, [P, P, P, P, u32, f32, P]),
    "sfk_eig_device": (i32, [P, P, P, P, u32, f32, P]),
    "sfk_eig_device_noalloc": (i32, [P, P, P, P, u32, f32, P]),
    "sfk_beta_update": (i32, [P, P, P, P, P, u32, f32]),
    "sfk_eig_round_create": (i32, [P, P, P, P, P, u32, u32, u32, PP]),
    "sfk_eig_round_destroy": (None, [P]),
    "sfk_eig_round_run": (i32, [P, P, P, P, i32, P, P]),
    "sfk_eig_round_set_binned": (i32, [P, i32]),
    "sfk_eig_round_set_pair_raster": (i32, [P, i32]),
    "sfk_eig_round_info": (i32, [P, u32, PP, P, P, P]),
    "sfk_eig_round_counters": (i32, [P, P, P]),
    "sfk_eig_round_plain": (i32, [P, P, P, P, P, P, u32, P, P, P, P, P]),
    "sfk_eig_round_autotune": (i32, [P, P, P, P, P, u32, P, P, u32, P, P, P]),
    "sfk_loss_l2": (i32, [P, P, P, u32, P, P]),
    "sfk_render_backward": (i32, [P, P, P, C.POINTER(TileIsects), P, P, P]),
    "sfk_render_backward_l2": (i32, [P, P, P, C.POINTER(TileIsects), P, P, P, P, P]),
    "sfk_localize": (i32, [P, P, P, P, P, P, P]),
    "sfk_render_posterior": (i32, [P, P, P, P, P, u32, C.POINTER(Camera), u32, u32, P, P]),
    "sfk_tint": (i32, [P, P, P, P, u32, P]),
    "sfk_ply_parse_header": (i32, [P, C.c_size_t, C.POINTER(PlyLayout), C.c_char_p, i32]),
    "sfk_ply_decode": (i32, [P, C.POINTER(PlyLayout), P, P, P]),
    "sfk_splats_from_ply": (i32, [P, P, C.c_size_t, PP]),
    "sfk_sog_decode": (i32, [P, C.POINTER(SogMeta), P, P, P, P, P, P, P, P, P, P, P]),
    "sfk_splats_create_device": (i32, [P, P, P, u32, u32, PP]),
    "sfk_save_colors": (i32, [P, P, P]),
    "sfk_restore_colors": (i32, [P, P, P]),
    "sfk_project": (i32, [P, C.POINTER(View), P, P, u32, u32, u32, P, P, P, P, P]),
    "sfk_map_isects": (i32, [P, P, P, P, P, u32, u32, u32, P, P]),
    "sfk_tile_starts": (i32, [P, P, u32, P, u32, P]),
}

_lib = None


def lib():
    """The loaded libsfk (raises if it is missing: there is no CPU path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libsfk.so not built at {LIB_PATH}; run `python -m splatfield_amd.build`")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib
