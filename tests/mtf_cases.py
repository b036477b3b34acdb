"""analysis::mtf cases (src/analysis.rs:133-150 and its tests,
analysis.rs:476-494, 661-671) against an mtf callable
mtf(psf (ny, nx), dx) -> (freqs, values)."""
import math

import numpy as np


def gaussian_psf(w, n, d):
    """analysis.rs:432-442."""
    x = (np.arange(n) - (n - 1) / 2.0) * d
    return np.exp(-(x[None, :] * x[None, :] + x[:, None] * x[:, None]) / (w * w))


def gaussian_anchors_mtf(mtf, tol=1.0):
    """analysis.rs:476-494: MTF of exp(-rho^2/w^2) is exp(-pi^2 w^2 f^2)."""
    w, d = 100e-6, 10e-6
    f, v = mtf(gaussian_psf(w, 512, d), d)
    assert abs(v[0] - 1.0) < 1e-12
    worst = 0.0
    for ff, vv in zip(f[1:], v[1:]):
        if ff > 800e3:
            break
        worst = max(worst, abs(vv - math.exp(-math.pi ** 2 * w * w * ff * ff)))
    assert worst < 0.02


def flat_field_is_dc_only(mtf, tol=1.0):
    """analysis.rs:661-665."""
    f, v = mtf(np.ones((8, 8)), 1e-3)
    assert np.all(v[1:] < 1e-9)


def delta_is_flat(mtf, tol=1.0):
    """A single bright sample: every bin has |F| = |F(0)|, so MTF = 1."""
    p = np.zeros((33, 48))
    p[7, 19] = 2.5
    f, v = mtf(p, 2e-6)
    assert np.all(np.abs(v - 1.0) < 1e-12)
    assert np.allclose(f, np.arange(25) / (48 * 2e-6), rtol=0, atol=0)


def zero_field_is_safe(mtf, tol=1.0):
    """analysis.rs:661-671: a zero PSF divides by DBL_MIN, not by zero."""
    f, v = mtf(np.zeros((8, 8)), 1e-3)
    assert np.all(v == 0.0)


def matches_numpy_fft(mtf, tol=1.0):
    """Row fy = 0 of numpy's fft2 (pocketfft) on uneven sizes."""
    rng = np.random.default_rng(9)
    for ny, nx in ((37, 64), (64, 45), (128, 128), (20, 7)):
        p = rng.uniform(0.0, 1.0, (ny, nx)) ** 3
        f, v = mtf(p, 3e-6)
        F = np.fft.fft2(p)
        want = np.minimum(np.abs(F[0, : nx // 2 + 1]) / abs(F[0, 0]), 1.0)
        assert np.abs(v - want).max() < 1e-12 * tol


CASES = [gaussian_anchors_mtf, flat_field_is_dc_only, delta_is_flat, zero_field_is_safe, matches_numpy_fft]
