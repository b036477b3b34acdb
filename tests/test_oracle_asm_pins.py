"""Pins the oracle's padded, band-limited ASM (asm.rs:438-538, band limit
asm.rs:258-265) by two routes that share nothing with its FFT:

- a direct evaluation on small padded grids: the 2-D DFT and its inverse as
  products with DFT matrices whose angles are reduced exactly ((k n) mod P),
  the transfer function e^{i (2 pi/lambda) z sqrt(s)} with the Matsushima
  limit (zero where |f| exceeds it on either axis) and the evanescent
  cut-off, embedded at offset (P - N)/2 and cropped back; even and odd
  padded sizes, band limit active and inactive;
- numpy's pocketfft ASM (np_asm of test_oracle_inverse.py) at the
  benchmark geometries (workloads 1, 2 and 4, band limit inactive there) and
  at one where the limit is active (256^2 at z = 0.2 m).

A dropped limit, a limit on the wrong axis, a wrong embedding offset or a
sign error in the phase fails the direct check at 1e-12.
"""
import math

import numpy as np
import pytest

import oc_types as T
from test_oracle_inverse import np_asm


def dft_matrix(n, sign):
    k = np.arange(n)
    e = np.outer(k, k) % n  # exact angle reduction
    return np.exp(sign * 2j * math.pi * e / n)


def direct_asm(u, dx, dy, lam, z, padding, band_limited):
    ny, nx = u.shape
    P = int(math.ceil(nx * padding)) if padding > 1 else nx
    Q = int(math.ceil(ny * padding)) if padding > 1 else ny
    ox, oy = (P - nx) // 2, (Q - ny) // 2
    buf = np.zeros((Q, P), complex)
    buf[oy:oy + ny, ox:ox + nx] = u
    F = dft_matrix(Q, -1) @ buf @ dft_matrix(P, -1).T
    fx = np.array([k / (P * dx) if k <= (P - 1) // 2 else (k - P) / (P * dx) for k in range(P)])
    fy = np.array([k / (Q * dy) if k <= (Q - 1) // 2 else (k - Q) / (Q * dy) for k in range(Q)])
    s = 1.0 - (lam * fx[None, :]) ** 2 - (lam * fy[:, None]) ** 2
    H = np.where(s >= 0, np.exp(1j * ((2 * math.pi / lam) * z) * np.sqrt(np.maximum(s, 0))), 0)
    if band_limited:
        lx = 1.0 / (lam * math.sqrt((2 * z / (P * dx)) ** 2 + 1.0))
        ly = 1.0 / (lam * math.sqrt((2 * z / (Q * dy)) ** 2 + 1.0))
        H = np.where((np.abs(fx[None, :]) > lx) | (np.abs(fy[:, None]) > ly), 0, H)
    out = dft_matrix(Q, 1) @ (F * H) @ dft_matrix(P, 1).T / (P * Q)
    return out[oy:oy + ny, ox:ox + nx], (P, Q)


def rand_field(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


CASES = [  # (ny, nx), dx, dy, lambda, z, padding, band limit active on some axis
    ((9, 12), 4e-6, 5e-6, 633e-9, 0.05, 2.0, True),
    ((9, 12), 4e-6, 5e-6, 633e-9, 2e-4, 2.0, False),
    ((10, 7), 3e-6, 3e-6, 532e-9, 0.03, 2.5, True),
    ((16, 16), 2e-6, 2e-6, 450e-9, 0.01, 2.0, True),
    ((5, 13), 6e-6, 2e-6, 1.55e-6, 0.02, 3.0, True),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0][0]}x{c[0][1]}-pad{c[5]}-z{c[4]}" for c in CASES])
def test_oracle_asm_equals_direct_dft(orc, case):
    shape, dx, dy, lam, z, pad, active = case
    u = rand_field(shape, 3 + shape[0] * shape[1])
    ref, (P, Q) = direct_asm(u, dx, dy, lam, z, pad, True)
    free, _ = direct_asm(u, dx, dy, lam, z, pad, False)
    # the case exercises the limit when it says so (a check on the check)
    assert (np.linalg.norm(ref - free) / np.linalg.norm(free) > 1e-6) == active
    cfg = T.AsmConfig(pad, 1, 0, pad, 1.0)
    st, out, rep = orc.asm_propagate(u, dx, dy, lam, z, cfg)
    assert st == 0
    assert (rep.nx_pad, rep.ny_pad) == (P, Q)
    e = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert e <= 1e-12, e
    cfg0 = T.AsmConfig(pad, 0, 0, pad, 1.0)
    st, out0, _ = orc.asm_propagate(u, dx, dy, lam, z, cfg0)
    assert st == 0
    assert np.linalg.norm(out0 - free) / np.linalg.norm(free) <= 1e-12


GEOMS = [  # n, d, lambda, z
    (256, 4e-6, 633e-9, 5e-3),     # workload 1
    (1024, 3e-6, 1.55e-6, 10e-3),  # workload 4
    (2048, 2e-6, 650e-9, 20e-3),   # workload 2, longest wavelength
    (256, 4e-6, 633e-9, 0.2),      # limit active
]


@pytest.mark.parametrize("geom", GEOMS, ids=[f"n{g[0]}-z{g[3]}" for g in GEOMS])
def test_oracle_asm_band_limited_equals_numpy(orc, geom):
    n, d, lam, z = geom
    u = rand_field((n, n), n + 1)
    ref = np_asm(u, d, lam, z, padding=2.0, band_limited=True)
    st, out, rep = orc.asm_propagate(u, d, d, lam, z, T.AsmConfig(2.0, 1, 0, 2.0, 1.0))
    assert st == 0
    assert rep.nx_pad == 2 * n
    e = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert e <= 1e-12, e
