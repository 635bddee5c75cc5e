"""Scene fixtures of the reference's test suites (src/render.rs:473-502,
src/seg/tests.rs:9-110, src/seg/views.rs:98-129), as numpy arrays."""
import numpy as np

from splatfield_amd.scene import ATTR_PLANES


def sample_opaque_attributes(n, position):
    """render.rs:478-494: identity rotation, log-scale -2, opacity logit 8."""
    a = np.zeros((ATTR_PLANES, n), np.float32)
    for i in range(n):
        a[0:3, i] = np.asarray(position(i), np.float32)
    a[3] = 1.0
    a[7:10] = -2.0
    a[10] = 8.0
    return a.reshape(-1)


def row_attributes(n, z, spacing):
    """seg/tests.rs:11-15."""
    f = np.float32
    return sample_opaque_attributes(
        n, lambda i: (f((f(i) - f(n - 1) * f(0.5)) * f(spacing)), 0.0, z))


def overdraw_attributes():
    """seg/tests.rs:78-89: 40 staggered splats crowding the centre tile."""
    f = np.float32

    def pos(i):
        x = f(i % 4) - f(1.5)
        y = f(i % 3) - f(1.0)
        return (x * f(0.04), y * f(0.04), f(1.0) + f(i) * f(0.5))
    return sample_opaque_attributes(40, pos)


def jitter(i):
    """views.rs:98-107."""
    f = np.float32
    return np.array([f((i * 73) % 17) / f(17.0) - f(0.5), f((i * 151) % 13) / f(13.0) - f(0.5),
                     f((i * 201) % 11) / f(11.0) - f(0.5)], np.float32)


# seg/tests.rs:186-203 — epsilon words of the overdraw scene at 70x70,
# produced by the reference's kernel on its CI device.
GOLDEN_EPS = [411225574, 64983936, 17882252, 5200864, 773093, 333408, 108187, 61572, 12526,
              2462, 1466, 1461, 10] + [0] * 27
