"""splatfield_amd — MI355X (gfx950) implementation of the splatfield 3DGS
render path (project -> tile-key sort -> rasterize) and the B3-Seg
evidence / EIG path, behind the C ABI of include/sfk.h (libsfk.so).

The Python modules mirror the reference's entry points (render_with,
prepare_isects, finalize, RadixScratch.argsort, exclusive_scan_gather,
render_responsibility, BetaState) for tests and benchmarks; all compute runs
in libsfk's HIP kernels."""
from ._lib import LIB_PATH, lib  # noqa: F401
from .runtime import Context, GpuTensor, SfkError  # noqa: F401
