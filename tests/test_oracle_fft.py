"""Pins the oracle's FFT layer (restating src/propagation/fft.rs over rustfft).

Checks: numpy.fft (pocketfft) at even, odd, prime and 2^a*3^b sizes, plus the
reference's own unit tests fft.rs:159-259 (round trip, Parseval x N, delta ->
flat spectrum, circular convolution against a direct sum, plane wave in bin
+k0, fftfreq literals).  A transposed operand, a wrong sign or a missing 1/N
fails one of these.
"""
import numpy as np
import pytest


def noise(n, seed=0x2545F491):
    """xorshift64 stream of fft.rs:134-150 (deterministic complex noise)."""
    s = seed
    out = np.empty(n, np.complex128)
    m = (1 << 64) - 1
    for i in range(n):
        s ^= (s << 13) & m
        s ^= s >> 7
        s ^= (s << 17) & m
        re = s / float(m) - 0.5
        s ^= (s << 13) & m
        s ^= s >> 7
        s ^= (s << 17) & m
        im = s / float(m) - 0.5
        out[i] = complex(re, im)
    return out


@pytest.mark.parametrize("shape", [(8, 8), (7, 6), (5, 1), (1, 5), (20, 12), (13, 97), (48, 96), (3, 2),
                                   (256, 256), (384, 64), (25, 49), (11, 121)])
def test_fft2_matches_numpy(orc, shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = orc.fft2(a)
    ref = np.fft.fft2(a)
    assert np.linalg.norm(f - ref) / np.linalg.norm(ref) < 1e-14
    g = orc.ifft2(a)
    ref = np.fft.ifft2(a)
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 16, 17, 60, 127, 512, 1536, 4096])
def test_fft1_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = orc.fft1(a)
    assert np.linalg.norm(f - np.fft.fft(a)) <= 1e-14 * max(1.0, np.linalg.norm(np.fft.fft(a)))
    g = orc.fft1(a, inverse=True)
    assert np.linalg.norm(g - np.fft.ifft(a) * n) <= 1e-14 * max(1.0, np.linalg.norm(g))


def test_roundtrip_even_and_odd(orc):
    for nx, ny in [(8, 8), (6, 7), (1, 5), (5, 1)]:
        u = noise(nx * ny).reshape(ny, nx)
        back = orc.ifft2(orc.fft2(u))
        assert np.max(np.abs(back - u)) < 1e-12


def test_parseval_scales_by_n(orc):
    nx, ny = 9, 6
    u = noise(nx * ny).reshape(ny, nx)
    e_in = np.sum(np.abs(u) ** 2)
    e_out = np.sum(np.abs(orc.fft2(u)) ** 2)
    assert abs(e_out - nx * ny * e_in) < 1e-9 * max(nx * ny * e_in, 1.0)


def test_delta_gives_flat_unit_spectrum(orc):
    buf = np.zeros((4, 8), np.complex128)
    buf[0, 0] = 1.0
    assert np.all(np.abs(orc.fft2(buf) - 1.0) < 1e-12)


def test_convolution_theorem_is_exact(orc):
    nx, ny = 5, 4
    u = noise(nx * ny).reshape(ny, nx)
    h = noise(nx * ny, seed=0x9E3779B97F4A7C15).reshape(ny, nx)
    prod = orc.ifft2(orc.fft2(u) * orc.fft2(h))
    for iy in range(ny):
        for ix in range(nx):
            acc = 0j
            for jy in range(ny):
                for jx in range(nx):
                    acc += u[jy, jx] * h[(iy - jy) % ny, (ix - jx) % nx]
            assert abs(prod[iy, ix] - acc) < 1e-9 * max(abs(acc), 1.0)


def test_plane_wave_energy_sits_in_positive_bin(orc):
    nx, ny, k0 = 16, 4, 3
    n = np.arange(nx)
    buf = np.tile(np.exp(1j * 2 * np.pi * k0 * n / nx), (ny, 1))
    f = orc.fft2(buf)
    assert abs(f[0, k0] - nx * ny) < 1e-9
    assert abs(np.sum(np.abs(f)) - nx * ny) < 1e-6


def test_fftfreq_matches_numpy_semantics(orc):
    assert list(orc.fftfreq(4, 1.0)) == [0.0, 0.25, -0.5, -0.25]
    assert list(orc.fftfreq(5, 1.0)) == [0.0, 0.2, 0.4, -0.4, -0.2]
    for n, d in [(4096, 2e-6), (1023, 3e-6), (7, 0.5)]:
        assert np.array_equal(orc.fftfreq(n, d), np.array([(k if k <= (n - 1) // 2 else k - n) / (n * d)
                                                           for k in range(n)]))


def test_padded_len(orc):
    assert orc.padded_len(100, 2.0) == 200
    assert orc.padded_len(100, 1.5) == 150
    assert orc.padded_len(100, 1.0) == 100
    assert orc.padded_len(100, 0.5) == 100


def test_splitmix_matches_numpy_generator(orc):
    import workloads
    a = orc.splitmix_uniform(20261203, 1000)
    b = workloads.splitmix_uniform(20261203, 1000)
    assert np.array_equal(a, b)
    assert a.min() >= 0.0 and a.max() < 1.0
