"""CPU checks of the C ABI: libsfk.so loads, exports every entry point that
include/sfk.h declares, and its host-side functions (camera views, look_at,
bits_for, fractional_bits, mask packing) agree with the oracle.  No device
compute runs here."""
import ctypes as C
import os
import re

import numpy as np

from oracle import oracle as O
from splatfield_amd import _lib
from splatfield_amd import render as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "sfk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    lib = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"


def test_camera_view_matches_oracle_bitwise():
    rng = np.random.default_rng(0)
    for _ in range(20):
        pos = rng.normal(size=3).astype(np.float32) * 3
        cam = R.Camera.look_at(pos, (0.1, -0.2, 0.3), fov=(0.9, 0.7))
        for w, h, tw in [(256, 256, 16), (1920, 1080, 16), (512, 384, 8), (100, 60, 32)]:
            v = cam.view(w, h, tw)
            o = O.camera_view(cam.fov, cam.position, cam.rotation, w, h, tw)
            assert bytes(v) == bytes(o)
        q = O.look_at(pos, (0.1, -0.2, 0.3))
        np.testing.assert_array_equal(np.array(cam.rotation, np.float32), q)


def test_host_helpers_match_oracle():
    lib = _lib.lib()
    for m in [0, 1, 2, 3, 255, 256, 257, 8160, 65536, 2**32 - 1]:
        assert lib.sfk_bits_for(m) == O.bits_for(m)
    for w, h in [(70, 70), (512, 512), (1920, 1080), (1, 1), (65536, 65536)]:
        assert lib.sfk_fractional_bits(w, h) == O.fractional_bits(w, h)


def test_mask_pack_layout():
    """Mask::from_bytes (seg/mod.rs:118-143): row-major, LSB-first, 32 px/word."""
    lib = _lib.lib()
    w, h = 70, 5
    rng = np.random.default_rng(1)
    b = (rng.random(w * h) < 0.3).astype(np.uint8)
    wpr = (w + 31) // 32
    words = np.zeros(wpr * h, np.uint32)
    fg = C.c_uint32()
    assert lib.sfk_mask_pack(b.ctypes.data, w, h, words.ctypes.data, C.byref(fg)) == 0
    assert fg.value == int(b.sum())
    for y in range(h):
        for x in range(w):
            assert ((int(words[y * wpr + x // 32]) >> (x % 32)) & 1) == b[y * w + x]


def test_invalid_arguments_are_errors_not_aborts():
    lib = _lib.lib()
    v = _lib.View()
    c = R.Camera().to_c()
    assert lib.sfk_camera_view(C.byref(c), 64, 64, 12, C.byref(v)) != 0  # bad tile width
    assert lib.sfk_camera_view(C.byref(c), 0, 64, 16, C.byref(v)) != 0
    out = C.c_void_p()
    assert lib.sfk_radix_scratch_create(None, 10, 4, C.byref(out)) != 0
