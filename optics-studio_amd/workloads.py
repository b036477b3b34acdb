"""Synthetic inputs of the five benchmark workloads, generated deterministically.

Randomness comes from SplitMix64 -> u53 * 2^-53 uniforms in [0, 1), seeded per
workload with SEED, so every consumer (tests, bench, checker) draws
bit-identical inputs.  Geometry and physics follow the reference sources:
Gaussian beam (src/sources.rs:259-298), plane wave (src/field.rs:360-386),
Field::coord centring (src/field.rs:216-221).
"""
import math

import numpy as np

SEED = 20261203
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_uniform(seed, count, offset=0):
    """Uniforms in [0, 1): element i is draw number offset+i of the stream."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Stream:
    """Sequential consumer of one SplitMix64 stream."""

    def __init__(self, seed=SEED):
        self.seed = seed
        self.pos = 0

    def uniform(self, count):
        u = splitmix_uniform(self.seed, count, self.pos)
        self.pos += count
        return u


def coords(n, d):
    """Field::coord along one axis: (i - (n-1)/2) * d."""
    return (np.arange(n, dtype=np.float64) - (n - 1.0) / 2.0) * d


def gaussian_beam(n, d, lam, w0, amplitude=1.0, center=(0.0, 0.0)):
    """sources::gaussian at the waist (z = 0): flat phase, e^{-rho^2/w0^2}."""
    x = coords(n, d) - center[0]
    y = coords(n, d) - center[1]
    rho2 = x[None, :] * x[None, :] + y[:, None] * y[:, None]
    return (amplitude * np.exp(-rho2 / (w0 * w0))).astype(np.complex128)


def circle_mask(n, d, r):
    """EllipticalAperture::circle transmission at the origin (boundary inclusive)."""
    x = coords(n, d)
    return (x[None, :] * x[None, :] + x[:, None] * x[:, None]) <= r * r


# ---------------------------------------------------------------------------
# Workload 1: plain f64 ASM, 256^2 at 4 um, 633 nm, Gaussian w0 200 um,
# circular aperture r 300 um, z = 5 mm.
# ---------------------------------------------------------------------------
CFG1 = dict(n=256, d=4e-6, lam=633e-9, w0=200e-6, r_ap=300e-6, z=5e-3)


def workload1():
    c = CFG1
    u = gaussian_beam(c["n"], c["d"], c["lam"], c["w0"])
    u[~circle_mask(c["n"], c["d"], c["r_ap"])] = 0.0
    return u


# ---------------------------------------------------------------------------
# Workload 2: polychromatic 2048^2 at 2 um, 16 bands 450-650 nm, Gaussian
# w0 1 mm, 8-level DOE in 4x4-pixel blocks, ASM z = 20 mm then Fresnel TF
# z = 200 mm (padding 2, strict off: the strict guard fails even at 8x).
# ---------------------------------------------------------------------------
CFG2 = dict(n=2048, d=2e-6, bands=16, lam0=450e-9, lam1=650e-9, w0=1e-3, levels=8, block=4, z_asm=20e-3,
            z_fresnel=0.2)


def workload2_lambdas(bands=None):
    c = CFG2
    b = c["bands"] if bands is None else bands
    return [c["lam0"] + j * ((c["lam1"] - c["lam0"]) / (c["bands"] - 1)) for j in range(b)]


def workload2_doe(n=None, block=None, levels=None, seed=SEED):
    c = CFG2
    n = c["n"] if n is None else n
    block = c["block"] if block is None else block
    levels = c["levels"] if levels is None else levels
    nb = n // block
    u = Stream(seed).uniform(nb * nb).reshape(nb, nb)
    lv = np.floor(levels * u)
    phi_blocks = 2.0 * math.pi * lv / levels
    return np.repeat(np.repeat(phi_blocks, block, axis=0), block, axis=1)


def workload2_field(lam, n=None):
    c = CFG2
    n = c["n"] if n is None else n
    return gaussian_beam(n, c["d"], lam, c["w0"])


# ---------------------------------------------------------------------------
# Workload 3: FFT-form Gerchberg-Saxton, 1024^2, disk source r 450 px, 64
# Gaussian spots (sigma 3 px) in the central 768^2, 32 restarts x 200 its.
# ---------------------------------------------------------------------------
CFG3 = dict(n=1024, r_src=450.0, spots=64, sigma=3.0, lo=128.0, hi=896.0, restarts=32, iterations=200)


def workload3(n=None, restarts=None, spots=None, seed=SEED):
    c = CFG3
    n = c["n"] if n is None else n
    restarts = c["restarts"] if restarts is None else restarts
    spots = c["spots"] if spots is None else spots
    s = Stream(seed)
    ctr = (n - 1) / 2.0
    i = np.arange(n, dtype=np.float64)
    r2 = (i[None, :] - ctr) ** 2 + (i[:, None] - ctr) ** 2
    amp_src = (np.sqrt(r2) <= c["r_src"] * n / c["n"]).astype(np.float64)
    scale = n / c["n"]
    cxy = c["lo"] * scale + s.uniform(2 * spots).reshape(spots, 2) * (c["hi"] - c["lo"]) * scale
    a = 0.5 + 0.5 * s.uniform(spots)
    inten = np.zeros((n, n))
    sig = c["sigma"]
    for k in range(spots):
        inten += a[k] * np.exp(-((i[None, :] - cxy[k, 0]) ** 2 + (i[:, None] - cxy[k, 1]) ** 2) / (2 * sig * sig))
    t = np.sqrt(inten)
    t *= math.sqrt(np.sum(amp_src ** 2) / np.sum(t ** 2))
    phases = 2.0 * math.pi * s.uniform(restarts * n * n).reshape(restarts, n, n)
    return amp_src, t, phases


# ---------------------------------------------------------------------------
# Workload 4: adjoint inverse design, 3-plane beamline at 1024^2, 1.55 um,
# 3 um pitch, plane wave in, DOE1 -> ASM 10 mm -> DOE2 -> ASM 10 mm ->
# detector, MseToTarget against 16 Gaussian spots (sigma 4 px).
# ---------------------------------------------------------------------------
CFG4 = dict(n=1024, d=3e-6, lam=1.55e-6, z=10e-3, spots=16, sigma=4.0, lo=128.0, hi=896.0, fd_pixels=128,
            fd_step=1e-4)


def workload4(n=None, seed=SEED):
    c = CFG4
    n = c["n"] if n is None else n
    s = Stream(seed)
    phi1 = 2.0 * math.pi * s.uniform(n * n).reshape(n, n)
    phi2 = 2.0 * math.pi * s.uniform(n * n).reshape(n, n)
    scale = n / c["n"]
    cxy = c["lo"] * scale + s.uniform(2 * c["spots"]).reshape(c["spots"], 2) * (c["hi"] - c["lo"]) * scale
    i = np.arange(n, dtype=np.float64)
    t = np.zeros((n, n))
    sig = c["sigma"]
    for k in range(c["spots"]):
        t += np.exp(-((i[None, :] - cxy[k, 0]) ** 2 + (i[:, None] - cxy[k, 1]) ** 2) / (2 * sig * sig))
    u = np.ones((n, n), np.complex128)  # BandField::plane_wave(0, 0, 1, ...) is exactly 1 + 0i
    return u, phi1, phi2, t


def workload4_fd_pixels(n=None, count=None, seed=SEED + 1):
    c = CFG4
    n = c["n"] if n is None else n
    count = c["fd_pixels"] if count is None else count
    s = Stream(seed)
    p1 = np.minimum((s.uniform(count) * n * n).astype(np.int64), n * n - 1)
    p2 = np.minimum((s.uniform(count) * n * n).astype(np.int64), n * n - 1)
    return p1, p2


# ---------------------------------------------------------------------------
# Workload 5: 64 uneven Rayleigh-Sommerfeld problems (padding 2, strict off).
# ---------------------------------------------------------------------------
CFG5 = dict(problems=64, sizes=(256, 384, 512, 768, 1024, 2048), lam0=400e-9, lam1=700e-9, z0=1e-3, z1=1e-1,
            d0=2e-6, d1=8e-6, ap=0.4)


def workload5_params(problems=None, sizes=None, seed=SEED):
    c = CFG5
    problems = c["problems"] if problems is None else problems
    sizes = c["sizes"] if sizes is None else sizes
    s = Stream(seed)
    out = []
    for _ in range(problems):
        u = s.uniform(4)
        n = sizes[min(int(u[0] * len(sizes)), len(sizes) - 1)]
        lam = c["lam0"] + u[1] * (c["lam1"] - c["lam0"])
        z = math.exp(math.log(c["z0"]) + u[2] * (math.log(c["z1"]) - math.log(c["z0"])))
        d = c["d0"] + u[3] * (c["d1"] - c["d0"])
        out.append(dict(n=n, lam=lam, z=z, d=d))
    return out


def workload5_field(p, seed=SEED, index=0):
    n, d = p["n"], p["d"]
    th = 2.0 * math.pi * splitmix_uniform(seed + 7919 * (index + 1), n * n).reshape(n, n)
    x = coords(n, d)
    inside = (x[None, :] ** 2 + x[:, None] ** 2) <= (CFG5["ap"] * n * d) ** 2
    u = np.exp(1j * th)
    u[~inside] = 0.0
    return u
