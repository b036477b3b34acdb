"""A masked propagation at z = 0 and masked propagations whose guard fails:
Element::apply then the engine, as Beamline::trace does (src/beamline.rs:
118-123; PhaseMask::apply, src/element.rs:1443-1484).

- z = 0 is the engine's identity (asm.rs:439-479, fresnel.rs:430-451), so
  the output is the masked field: single calls, batches and the chained
  call, against the oracle's phasemask_apply.
- The mask's OutsideField error comes first: at z = 0 and before a failed
  strict guard.
- The engine validates the masked field: a non-finite sample outside the
  stop is zeroed by the apply and is not an error; one inside is.
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


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


N, D, LAM = 64, 4e-6, 550e-9
AP = (T.APERTURE_CIRCLE, 90e-6, 0.0)


def field(seed=3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))


def lattice():
    return 2 * math.pi * W.splitmix_uniform(17, N * N).reshape(N, N)


def masked_ref(orc, u):
    st, um = orc.phasemask_apply(u, D, D, LAM, orc.Mask(lattice(), D, D, aperture=AP, center=(8e-6, -4e-6)))
    assert st == 0
    return um


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_masked_z_zero_single_calls(G, orc, dtype):
    import torch
    dt, tol = (torch.complex128, 1e-12) if dtype == "c128" else (torch.complex64, 1e-6)
    u = field()
    ref = masked_ref(orc, u)
    ms, keep = G.dev_mask(lattice(), D, D, aperture=AP, center=(8e-6, -4e-6))
    ud = torch.as_tensor(u).to(dt).cuda()
    say("masked z = 0", dtype)
    st, out, _ = G.asm_propagate_dev(ud, D, D, LAM, 0.0, T.AsmConfig.default(), mask=ms)
    assert st == 0, T.STATUS[st]
    e = rel_l2(out.cpu().numpy().astype(np.complex128), ref)
    say("  asm", e)
    assert e <= tol
    st, out, _ = G.fresnel_propagate_dev(ud, D, D, LAM, 0.0, T.FresnelConfig.default(), mask=ms)
    assert st == 0, T.STATUS[st]
    e = rel_l2(out.cpu().numpy().astype(np.complex128), ref)
    say("  fresnel", e)
    assert e <= tol
    # in place
    ui = ud.clone()
    st, out, _ = G.asm_propagate_dev(ui, D, D, LAM, 0.0, T.AsmConfig.default(), mask=ms, out=ui)
    assert st == 0
    assert rel_l2(ui.cpu().numpy().astype(np.complex128), ref) <= tol


@pytest.mark.parametrize("double", [True, False])
def test_masked_z_zero_batch_and_chain(G, orc, double):
    import torch
    dt, tol = (torch.complex128, 1e-12) if double else (torch.complex64, 1e-5)
    u = field(5)
    um = masked_ref(orc, u)
    ms, keep = G.dev_mask(lattice(), D, D, aperture=AP, center=(8e-6, -4e-6))
    ud = torch.as_tensor(u).to(dt).cuda()
    o0, o1 = torch.empty_like(ud), torch.empty_like(ud)
    arr = (T.PropDesc * 2)()
    for k, (z, o) in enumerate(((0.0, o0), (2e-3, o1))):
        x = arr[k]
        x.geom = T.geom(N, N, D, D, LAM)
        x.engine = T.ENGINE_ASM
        x.z, x.padding, x.band_limited, x.strict, x.max_padding, x.medium = z, 2.0, 1, 0, 2.0, 1.0
        x.mask = T.C.pointer(ms)
        x.in_, x.out = ud.data_ptr(), o.data_ptr()
    say("masked z = 0 batch, double", double)
    st = G.propagate_batch_dev(arr, double=double)
    assert st == 0, T.STATUS[st]
    st, ra, _ = orc.asm_propagate(um, D, D, LAM, 2e-3, T.AsmConfig(2.0, 1, 0, 2.0, 1.0))
    assert st == 0
    e0 = rel_l2(o0.cpu().numpy().astype(np.complex128), um)
    e1 = rel_l2(o1.cpu().numpy().astype(np.complex128), ra)
    say("  identity", e0, "propagated", e1)
    assert e0 <= tol and e1 <= tol
    # chained: masked identity, then Fresnel 5 mm of the masked field
    first, second = (T.PropDesc * 1)(), (T.PropDesc * 1)()
    first[0] = arr[0]
    second[0].geom = arr[0].geom
    second[0].engine = T.ENGINE_FRESNEL
    second[0].z, second[0].padding, second[0].max_padding = 5e-3, 2.0, 2.0
    mid, out = torch.empty_like(ud), torch.empty_like(ud)
    first[0].out, second[0].out = mid.data_ptr(), out.data_ptr()
    st = G.propagate_chain_dev(first, second, double=double)
    assert st == 0, T.STATUS[st]
    st, rf, _ = orc.fresnel_propagate(um, D, D, LAM, 5e-3, T.FresnelConfig(2.0, 0, 0, 2.0))
    assert st == 0
    em = rel_l2(mid.cpu().numpy().astype(np.complex128), um)
    ef = rel_l2(out.cpu().numpy().astype(np.complex128), rf)
    say("  chain", em, ef)
    assert em <= tol and ef <= tol


def test_masked_error_precedence(G):
    import torch
    ud = torch.as_tensor(field()).cuda()
    big = (T.APERTURE_CIRCLE, 200e-6, 0.0)  # the grid's half extent is 126 um
    ms, keep = G.dev_mask(lattice(), D, D, aperture=big)
    outside = T.CODE["OC_OUTSIDE_FIELD"]
    st, _, _ = G.asm_propagate_dev(ud, D, D, LAM, 0.0, T.AsmConfig.default(), mask=ms)
    assert st == outside
    st, _, _ = G.fresnel_propagate_dev(ud, D, D, LAM, 0.0, T.FresnelConfig.default(), mask=ms)
    assert st == outside
    # a strict guard that fails at padding 1 (no densification): the apply
    # runs first
    strict1 = T.AsmConfig(1.0, 0, 1, 1.0, 1.0)
    ok_ms, keep2 = G.dev_mask(lattice(), D, D, aperture=AP)
    st, _, _ = G.asm_propagate_dev(ud, D, D, LAM, 0.5, strict1, mask=ok_ms)
    assert st == T.CODE["OC_ALIASING_RISK"]
    st, _, _ = G.asm_propagate_dev(ud, D, D, LAM, 0.5, strict1, mask=ms)
    assert st == outside


@pytest.mark.parametrize("z", [0.0, 2e-3])
def test_masked_nonfinite_outside_stop_is_zeroed(G, orc, z):
    """NaN at a sample outside the stop: the apply zeroes it, the engine sees
    a finite field.  NaN inside the stop: NonFiniteSample."""
    import torch
    u = field(9)
    ms, keep = G.dev_mask(lattice(), D, D, aperture=AP, center=(8e-6, -4e-6))
    bad_out = u.copy()
    bad_out[0, 0] = np.nan  # a corner: outside the 90 um stop
    bad_in = u.copy()
    bad_in[N // 2, N // 2] = np.nan
    cfg = T.AsmConfig(2.0, 1, 0, 2.0, 1.0)
    for dt in (torch.complex128, torch.complex64):
        st, out, _ = G.asm_propagate_dev(torch.as_tensor(bad_in).to(dt).cuda(), D, D, LAM, z, cfg, mask=ms)
        assert st == T.CODE["OC_NON_FINITE_SAMPLE"], (dt, T.STATUS[st])
        st, out, _ = G.asm_propagate_dev(torch.as_tensor(bad_out).to(dt).cuda(), D, D, LAM, z, cfg, mask=ms)
        assert st == 0, (dt, z, T.STATUS[st])
        um = masked_ref(orc, np.where(np.isnan(bad_out), 0.0, bad_out))
        ref = um
        if z != 0.0:
            st, ref, _ = orc.asm_propagate(um, D, D, LAM, z, cfg)
        e = rel_l2(out.cpu().numpy().astype(np.complex128), ref)
        say("  nan outside the stop", dt, z, e)
        assert e <= (1e-12 if dt == torch.complex128 else 1e-5)
