"""Padded lengths beyond the on-chip transform (the multi-pass path) against
the f64 oracle, through the C ABI.

- Strict densification (asm.rs:124-137, 465-479) of a 2048 x 16 grid at 2 um,
  633 nm, z = 0.1 mm: the 16-sample axis fails the guard at padding 2 and 4
  and passes at 8, so both axes pad 8x and x reaches P = 16384 (256 KB per
  complex128 line; the four-step split 128 x 128): host f64 drop-in and
  complex128 device form at 1e-12, complex64 at 1e-5.
- Rayleigh-Sommerfeld at padding 12 on a 1024 x 8 grid (P_x = 12288 =
  96 x 128 in complex128; the kernel spectrum through the same passes), and
  the Fresnel transfer function at padding 16 in complex64 (P_x = 32768).
- The adjoint push_back through densified ASM steps (the gradient epilogue
  on the multi-pass path) against the oracle at 1e-10.
- A prime padded length beyond the on-chip transform has no split:
  OC_ERR_UNSUPPORTED.
"""
import math
import sys

import numpy as np
import pytest

import oc_types as T
import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def field(ny, nx, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[:ny, :nx]
    env = np.exp(-(((x - (nx - 1) / 2) / (0.3 * nx)) ** 2 + ((y - (ny - 1) / 2) / (0.3 * ny)) ** 2))
    return env * (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx)))


def test_strict_densification_to_16384(G, orc):
    u = field(16, 2048, 1)
    d, lam, z = 2e-6, 633e-9, 1e-4
    cfg = T.AsmConfig.default()
    st, ref, rref = orc.asm_propagate(u, d, d, lam, z, cfg)
    assert st == 0
    assert (rref.used_padding, rref.nx_pad, rref.ny_pad) == (8.0, 16384, 128)
    say("densified", rref.used_padding, rref.nx_pad, rref.ny_pad)
    for name, eng, tol in (("f64", G, 1e-12), ("c128", G.DevEngine("c128"), 1e-12), ("c64", G.DevEngine("c64"), 1e-5)):
        st, out, rep = eng.asm_propagate(u, d, d, lam, z, cfg)
        assert st == 0, (name, T.STATUS[st])
        assert T.report_dict(rep) == T.report_dict(rref)
        e = rel(out, ref)
        say("  ", name, e)
        assert e <= tol, (name, e)


def test_rs_and_fresnel_beyond_on_chip(G, orc):
    u = field(8, 1024, 2)
    d, lam = 3e-6, 550e-9
    st, ref, _ = orc.rs_propagate(u, d, d, lam, 2e-3, T.RsConfig(12.0, 0, 0))
    assert st == 0
    for name, eng, tol in (("f64", G, 1e-12), ("c128", G.DevEngine("c128"), 1e-12), ("c64", G.DevEngine("c64"), 1e-5)):
        st, out, rep = eng.rs_propagate(u, d, d, lam, 2e-3, T.RsConfig(12.0, 0, 0))
        assert st == 0, (name, T.STATUS[st])
        assert rep.nx_pad == 12288
        e = rel(out, ref)
        say("  rs", name, e)
        assert e <= tol, (name, e)
    u2 = field(16, 2048, 3)
    fcfg = T.FresnelConfig(16.0, 0, 0, 16.0)
    st, ref2, _ = orc.fresnel_propagate(u2, d, d, lam, 5e-3, fcfg)
    assert st == 0
    for name, eng, tol in (("c64", G.DevEngine("c64"), 1e-5), ("c128", G.DevEngine("c128"), 1e-12)):
        st, out, rep = eng.fresnel_propagate(u2, d, d, lam, 5e-3, fcfg)
        assert st == 0, (name, T.STATUS[st])
        assert rep.nx_pad == 32768
        e = rel(out, ref2)
        say("  fresnel", name, e)
        assert e <= tol, (name, e)


def test_adjoint_through_densified_steps(G, orc):
    ny, nx, d, lam, z = 16, 2048, 2e-6, 633e-9, 1e-4
    u = field(ny, nx, 4)
    p1 = 2 * math.pi * W.splitmix_uniform(41, nx * ny).reshape(ny, nx)
    dmdi = np.random.default_rng(5).standard_normal((ny, nx))
    cfg = T.AsmConfig.default()
    so = [(orc.Mask(p1, d, d), z, 1.0), (None, z, 1.0)]
    sg = [(G.Mask(p1, d, d), z, 1.0), (None, z, 1.0)]
    st, ref = orc.adjoint_push_back(so, u, d, d, lam, cfg, dmdi)
    assert st == 0
    st, got = G.adjoint_push_back(sg, u, d, d, lam, cfg, dmdi)
    assert st == 0, T.STATUS[st]
    e = rel(got[0], ref[0])
    st, _, rep = orc.asm_propagate(u, d, d, lam, z, cfg)
    assert rep.nx_pad == 16384
    say("adjoint densified", e)
    assert e <= 1e-10


def test_prime_length_beyond_on_chip_is_unsupported(G):
    import torch
    n = 10007  # prime, above the complex128 on-chip transform
    u = torch.as_tensor(field(2, n, 6)).cuda()
    st, _, _ = G.fresnel_propagate_dev(u, 2e-6, 2e-6, 550e-9, 1e-3, T.FresnelConfig(1.0, 0, 0, 1.0))
    assert st == T.CODE["OC_ERR_UNSUPPORTED"]
