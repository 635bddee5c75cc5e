"""The rasterizers' branch-free exp (sfk_expf_neg) against the library's
full exp (the oracle's algorithm), bit for bit, and the branch-free
saturating float -> u32 conversion against its branchy definition, on the
host: a small HIP host program built with hipcc from the product header."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_expf_neg_bits_match_expf(tmp_path):
    exe = str(tmp_path / "expf_neg_check")
    src = os.path.join(HERE, "host", "expf_neg_check.hip")
    r = subprocess.run([HIPCC, "-O2", "-std=c++20", "-ffp-contract=off",
                        "--offload-arch=gfx950", src, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    # every 7th float pattern: ~3e8 inputs, a few seconds
    r = subprocess.run([exe, "7"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
