"""HIP path vs the f64 oracle, through the C ABI.

Bars: f64 drop-in and complex128 device forms within 1e-12 relative L2 of the
oracle; the FP32-bulk complex64 form (FP64 phase arguments and twiddles)
within 1e-5 relative L2 (the FP32 tolerance of the workload definitions).
"""
import math
import sys

import numpy as np
import pytest

import oc_types as T
import workloads as W

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
C64_TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


@pytest.fixture(scope="module")
def c64(G):
    return G.DevEngine("c64")


@pytest.fixture(scope="module")
def c128(G):
    return G.DevEngine("c128")


def rand_field(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_workload1_f64_dropin(G, c64, orc):
    c = W.CFG1
    u = W.workload1()
    say("workload1 f64", u.shape)
    st, ref, rref = orc.asm_propagate(u, c["d"], c["d"], c["lam"], c["z"], T.AsmConfig.default())
    assert st == 0
    st, out, rep = G.asm_propagate(u, c["d"], c["d"], c["lam"], c["z"], T.AsmConfig.default())
    assert st == 0, T.STATUS[st]
    assert T.report_dict(rep) == T.report_dict(rref)
    e = rel_l2(out, ref)
    say("workload1 f64 rel L2", e)
    assert e <= F64_TOL
    ir, io = np.abs(ref) ** 2, np.abs(out) ** 2
    assert rel_l2(io, ir) <= F64_TOL
    st, o64, _ = c64.asm_propagate(u, c["d"], c["d"], c["lam"], c["z"], T.AsmConfig.default())
    assert st == 0
    e64 = rel_l2(o64, ref)
    say("workload1 c64 rel L2", e64)
    assert e64 <= C64_TOL


SHAPES = [(16, 16), (8, 5), (20, 33), (48, 48), (37, 101), (200, 200), (96, 256), (384, 384), (7, 3), (1, 6)]


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{s[0]}x{s[1]}" for s in SHAPES])
def test_asm_parity_sizes(G, c64, c128, orc, shape):
    u = rand_field(shape, 1 + shape[0] * 7 + shape[1])
    d, lam, z = 3e-6, 532e-9, 1.2e-3
    cfg = T.AsmConfig(2.0, 1, 0, 8.0, 1.0)
    say("asm", shape)
    st, ref, _ = orc.asm_propagate(u, d, d * 1.25, lam, z, cfg)
    assert st == 0
    for name, eng, tol in (("f64", G, F64_TOL), ("c128", c128, F64_TOL), ("c64", c64, C64_TOL)):
        st, out, _ = eng.asm_propagate(u, d, d * 1.25, lam, z, cfg)
        assert st == 0, (name, T.STATUS[st])
        e = rel_l2(out, ref)
        say("  ", name, e)
        assert e <= tol, (name, e)


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{s[0]}x{s[1]}" for s in SHAPES])
def test_fresnel_parity_sizes(G, c64, c128, orc, shape):
    u = rand_field(shape, 11 + shape[0] * 5 + shape[1])
    d, lam, z = 4e-6, 633e-9, 3e-2
    cfg = T.FresnelConfig(2.0, 0, 0, 2.0)
    say("fresnel", shape)
    st, ref, _ = orc.fresnel_propagate(u, d, d, lam, z, cfg)
    assert st == 0
    for name, eng, tol in (("f64", G, F64_TOL), ("c128", c128, F64_TOL), ("c64", c64, C64_TOL)):
        st, out, _ = eng.fresnel_propagate(u, d, d, lam, z, cfg)
        assert st == 0, (name, T.STATUS[st])
        e = rel_l2(out, ref)
        say("  ", name, e)
        assert e <= tol, (name, e)


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{s[0]}x{s[1]}" for s in SHAPES])
def test_rs_parity_sizes(G, c64, c128, orc, shape):
    u = rand_field(shape, 21 + shape[0] * 3 + shape[1])
    d, lam, z = 2e-6, 550e-9, 4e-3
    cfg = T.RsConfig(2.0, 0, 0)
    say("rs", shape)
    st, ref, _ = orc.rs_propagate(u, d, d, lam, z, cfg)
    assert st == 0
    for name, eng, tol in (("f64", G, F64_TOL), ("c128", c128, F64_TOL), ("c64", c64, C64_TOL)):
        st, out, _ = eng.rs_propagate(u, d, d, lam, z, cfg)
        assert st == 0, (name, T.STATUS[st])
        e = rel_l2(out, ref)
        say("  ", name, e)
        assert e <= tol, (name, e)


@pytest.mark.parametrize("psi", [0.0, 0.7])
def test_masked_asm_fused_matches_apply_then_propagate(G, orc, psi):
    import torch
    n, d, lam, z = 96, 4e-6, 600e-9, 2e-3
    u = rand_field((n, n), 5)
    vals = 2 * math.pi * W.splitmix_uniform(99, 64 * 80).reshape(64, 80)
    ap = (T.APERTURE_CIRCLE, 150e-6, 0.0)
    say("masked asm psi", psi)
    m = orc.Mask(vals, 5e-6, 4e-6, orientation=psi, aperture=ap, center=(10e-6, -6e-6))
    st, um = orc.phasemask_apply(u, d, d, lam, m)
    assert st == 0
    st, ref, _ = orc.asm_propagate(um, d, d, lam, z, T.AsmConfig.default())
    assert st == 0
    # host drop-in apply
    gm = G.Mask(vals, 5e-6, 4e-6, orientation=psi, aperture=ap, center=(10e-6, -6e-6))
    st, gum = G.phasemask_apply(u, d, d, lam, gm)
    assert st == 0
    assert rel_l2(gum, um) <= 1e-15
    # fused device form, c128 and c64
    ms, keep = G.dev_mask(vals, 5e-6, 4e-6, orientation=psi, aperture=ap, center=(10e-6, -6e-6))
    for dt, tol in ((torch.complex128, F64_TOL), (torch.complex64, C64_TOL)):
        ud = torch.as_tensor(u.astype(np.complex128 if dt == torch.complex128 else np.complex64)).cuda()
        st, out, _ = G.asm_propagate_dev(ud, d, d, lam, z, T.AsmConfig.default(), mask=ms)
        assert st == 0, T.STATUS[st]
        e = rel_l2(out.cpu().numpy().astype(np.complex128), ref)
        say("  ", dt, e)
        assert e <= tol


def test_non_finite_and_order(G, c64):
    u = np.ones((16, 16), np.complex128)
    u[3, 3] = np.nan
    st, _, _ = G.asm_propagate(u, 1e-6, 1e-6, 500e-9, 1e-3, T.AsmConfig.default())
    assert st == T.CODE["OC_NON_FINITE_SAMPLE"]
    st, _, _ = c64.asm_propagate(u, 1e-6, 1e-6, 500e-9, 1e-3, T.AsmConfig.default())
    assert st == T.CODE["OC_NON_FINITE_SAMPLE"]
    # NonFiniteSample outranks AliasingRisk and the z == 0 identity
    st, _, _ = G.asm_propagate(u, 1e-6, 1e-6, 500e-9, 0.5, T.AsmConfig(1.0, 0, 1, 1.0, 1.0))
    assert st == T.CODE["OC_NON_FINITE_SAMPLE"]
    st, _, _ = G.asm_propagate(u, 1e-6, 1e-6, 500e-9, 0.0, T.AsmConfig.default())
    assert st == T.CODE["OC_NON_FINITE_SAMPLE"]
    st, _, _ = G.rs_propagate(u, 1e-6, 1e-6, 500e-9, 1e-3, T.RsConfig(2.0, 0, 0))
    assert st == T.CODE["OC_NON_FINITE_SAMPLE"]


def test_batch_uneven_rs_matches_single(G, orc):
    import torch
    ps = W.workload5_params(problems=6, sizes=(16, 24, 40, 64))
    fields = [W.workload5_field(p, index=i) for i, p in enumerate(ps)]
    ins = [torch.as_tensor(f.astype(np.complex64)).cuda() for f in fields]
    outs = [torch.empty_like(t) for t in ins]
    arr = (T.PropDesc * len(ps))()
    for i, p in enumerate(ps):
        arr[i].geom = T.geom(p["n"], p["n"], p["d"], p["d"], p["lam"])
        arr[i].z = p["z"]
        arr[i].engine = T.ENGINE_RS
        arr[i].padding = 2.0
        arr[i].strict = 0
        arr[i].in_ = ins[i].data_ptr()
        arr[i].out = outs[i].data_ptr()
    say("batch rs", [p["n"] for p in ps])
    st = G.propagate_batch_dev(arr)
    assert st == 0, T.STATUS[st]
    for i, p in enumerate(ps):
        assert arr[i].status == 0
        st, ref, _ = orc.rs_propagate(fields[i], p["d"], p["d"], p["lam"], p["z"], T.RsConfig(2.0, 0, 0))
        e = rel_l2(outs[i].cpu().numpy().astype(np.complex128), ref)
        say("  ", i, p["n"], e)
        assert e <= C64_TOL


def test_workload2_one_band_c64(G, orc):
    """Full-size 2048^2 band: DOE mask + ASM 20 mm (P=4096) + Fresnel 200 mm."""
    import torch
    c = W.CFG2
    lam = W.workload2_lambdas()[-1]
    u = W.workload2_field(lam)
    phi = W.workload2_doe()
    say("workload2 band", lam)
    m = orc.Mask(phi, c["d"], c["d"])
    st, um = orc.phasemask_apply(u, c["d"], c["d"], lam, m)
    assert st == 0
    st, ra, _ = orc.asm_propagate(um, c["d"], c["d"], lam, c["z_asm"], T.AsmConfig.default())
    assert st == 0
    fcfg = T.FresnelConfig(2.0, 0, 0, 2.0)
    st, rf, _ = orc.fresnel_propagate(ra, c["d"], c["d"], lam, c["z_fresnel"], fcfg)
    assert st == 0
    ms, keep = G.dev_mask(phi, c["d"], c["d"])
    ud = torch.as_tensor(u.astype(np.complex64)).cuda()
    st, a, _ = G.asm_propagate_dev(ud, c["d"], c["d"], lam, c["z_asm"], T.AsmConfig.default(), mask=ms)
    assert st == 0
    ea = rel_l2(a.cpu().numpy().astype(np.complex128), ra)
    st, f, _ = G.fresnel_propagate_dev(a, c["d"], c["d"], lam, c["z_fresnel"], fcfg)
    assert st == 0
    ef = rel_l2(f.cpu().numpy().astype(np.complex128), rf)
    say("  asm", ea, "fresnel", ef)
    assert ea <= C64_TOL and ef <= C64_TOL


FH_CASES = [((64, 64), 2.0, None), ((256, 256), -3.0, None), ((1024, 1024), 5.0, (2e-5, -1e-5)),
            ((37, 101), 2.5, None), ((200, 200), -2.0, (1e-5, 1e-5)), ((96, 256), 4.0, None)]


@pytest.mark.parametrize("shape,z,src", FH_CASES, ids=[f"{s[0]}x{s[1]}-z{z}" for s, z, _ in FH_CASES])
def test_fraunhofer_parity(G, c64, c128, orc, shape, z, src):
    u = rand_field(shape, 31 + shape[0] + shape[1])
    d, lam = 4e-6, 633e-9
    cfg = T.FraunhoferConfig.default(strict=False, source_pos=src)
    say("fraunhofer", shape, z)
    st, ref, rref = orc.fraunhofer_propagate(u, d, d, lam, z, cfg)
    assert st == 0
    for name, eng, tol in (("f64", G, F64_TOL), ("c128", c128, F64_TOL), ("c64", c64, C64_TOL)):
        st, out, rep = eng.fraunhofer_propagate(u, d, d, lam, z, cfg)
        assert st == 0, (name, T.STATUS[st])
        assert T.report_dict(rep) == T.report_dict(rref)
        e = rel_l2(out, ref)
        say("  ", name, e)
        assert e <= tol, (name, e)


FC_SHAPES = [(64, 64), (37, 101), (200, 200), (512, 512)]


@pytest.mark.parametrize("shape", FC_SHAPES, ids=[f"{s[0]}x{s[1]}" for s in FC_SHAPES])
def test_fresnel_convolution_parity(G, c64, c128, orc, shape):
    u = rand_field(shape, 41 + shape[0] * 3 + shape[1])
    d, lam, z = 3e-6, 600e-9, 8e-3
    cfg = T.FresnelConfig(2.0, 0, 0, 2.0)
    say("fresnel conv", shape)
    st, ref, rref = orc.fresnel_propagate_convolution(u, d, d, lam, z, cfg)
    assert st == 0
    for name, eng, tol in (("f64", G, F64_TOL), ("c128", c128, F64_TOL), ("c64", c64, C64_TOL)):
        st, out, rep = eng.fresnel_propagate_convolution(u, d, d, lam, z, cfg)
        assert st == 0, (name, T.STATUS[st])
        assert T.report_dict(rep) == T.report_dict(rref)
        e = rel_l2(out, ref)
        say("  ", name, e)
        assert e <= tol, (name, e)


@pytest.mark.parametrize("pad", [2.0, 1.5, 2.5])
def test_convolution_padding_is_linear_window(G, c64, c128, orc, pad):
    """RS and the Fresnel convolution with P >= 2N - 1 on both axes compute a
    linear convolution whose cropped window does not depend on P; such
    non-power-of-two P run at the next power of two.  Against the oracle at
    the reference's P: padding 2 and 2.5 on a 384 x 300 grid (P = 768 x 600
    and 960 x 750, run at 1024), and padding 1.5 (P < 2N - 1, wrap-around
    present: run at P itself).  The report keeps the reference's P."""
    shape = (300, 384)
    u = rand_field(shape, 77)
    d, lam = 3e-6, 600e-9
    for kind in ("rs", "fconv"):
        if kind == "rs":
            cfg = T.RsConfig(pad, 0, 0)
            st, ref, rref = orc.rs_propagate(u, d, d, lam, 5e-3, cfg)
        else:
            cfg = T.FresnelConfig(pad, 0, 0, pad)
            st, ref, rref = orc.fresnel_propagate_convolution(u, d, d, lam, 8e-3, cfg)
        assert st == 0
        for name, eng, tol in (("f64", G, F64_TOL), ("c128", c128, F64_TOL), ("c64", c64, C64_TOL)):
            if kind == "rs":
                st, out, rep = eng.rs_propagate(u, d, d, lam, 5e-3, cfg)
            else:
                st, out, rep = eng.fresnel_propagate_convolution(u, d, d, lam, 8e-3, cfg)
            assert st == 0, (kind, name, T.STATUS[st])
            assert T.report_dict(rep) == T.report_dict(rref)
            e = rel_l2(out, ref)
            say("  ", kind, pad, name, e)
            assert e <= tol, (kind, pad, name, e)
