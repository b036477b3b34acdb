"""Compile-time-specialised power-of-two kernels (padded lengths 64..8192) in
every workgroup layout, against the f64 oracle and against the runtime-plan
kernels of the same library.

Each case pins the variant on a private context, so the default context's
tuning is untouched.  Bars as in the other parity tests: 1e-12 relative L2 for
complex128, 1e-5 for the FP32-bulk complex64 form.  At P = 8192 the oracle
would need minutes, so the fast kernels are compared with the runtime-plan
kernels (themselves pinned to the oracle at smaller sizes) at the same bars,
plus an ASM(+z) then ASM(-z) round trip on a band-limited field, and RS and
the Fresnel convolution at a non-power-of-two padded length (6000) that runs
at 8192.
"""
import sys

import numpy as np
import pytest

import oc_types as T

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


def field(n, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[:n, :n]
    c = (n - 1) / 2
    env = np.exp(-(((x - c) ** 2 + (y - c) ** 2) / (0.18 * n) ** 2))
    return env * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))


def run_dev(G, ctx, engine, u, dtype, d, lam, z, cfg, mask=None):
    import torch
    t = torch.as_tensor(u.astype(np.complex64 if dtype == "c64" else np.complex128)).cuda()
    if engine == "asm":
        st, out, _ = G.asm_propagate_dev(t, d, d, lam, z, cfg, mask=mask, ctx=ctx)
    elif engine == "fresnel":
        st, out, _ = G.fresnel_propagate_dev(t, d, d, lam, z, cfg, ctx=ctx)
    else:
        st, out, _ = G.rs_propagate_dev(t, d, d, lam, z, cfg, ctx=ctx)
    assert st == 0, (engine, T.STATUS[st], ctx.last_error())
    return out.cpu().numpy().astype(np.complex128)


def variants(n):
    return [0, 1, 2, 3] if n >= 1024 else [0]


CASES = [(p, v) for p in (64, 128, 256, 512, 1024, 2048, 4096) for v in variants(p)]


@pytest.mark.parametrize("P,v", CASES, ids=[f"P{p}-v{v}" for p, v in CASES])
def test_fast_vs_oracle(G, orc, P, v):
    n = P // 2
    d, lam = 2e-6, 532e-9
    u = field(n, P + v)
    rng = np.random.default_rng(P)
    phi = rng.uniform(-np.pi, np.pi, (n, n))
    ctx = G.Context(0)
    for dbl in (False, True):
        assert ctx.set_fft_variant(P, v, double=dbl) == 0
        assert ctx.fft_variant(P, double=dbl) == v
    say("fast", P, "variant", v)
    mask, keep = G.dev_mask(phi, d, d)
    om = orc.Mask(phi, d, d)
    st, um = orc.phasemask_apply(u, d, d, lam, om)
    assert st == 0
    z_asm = 0.3 * n * d
    st, ref_a, _ = orc.asm_propagate(um, d, d, lam, z_asm, T.AsmConfig(2.0, 1, 0, 2.0, 1.0))
    assert st == 0
    fcfg = T.FresnelConfig(2.0, 0, 0, 2.0)
    z_fr = 2.0 * n * d * d / lam
    st, ref_f, _ = orc.fresnel_propagate(u, d, d, lam, z_fr, fcfg)
    assert st == 0
    rcfg = T.RsConfig(2.0, 0, 0)
    z_rs = 0.5 * n * d
    do_rs = P <= 2048
    if do_rs:
        st, ref_r, _ = orc.rs_propagate(u, d, d, lam, z_rs, rcfg)
        assert st == 0
    for dt, tol in (("c64", C64_TOL), ("c128", F64_TOL)):
        out = run_dev(G, ctx, "asm", u, dt, d, lam, z_asm, T.AsmConfig(2.0, 1, 0, 2.0, 1.0), mask=mask)
        e = rel_l2(out, ref_a)
        say("   masked asm", dt, e)
        assert e <= tol
        out = run_dev(G, ctx, "fresnel", u, dt, d, lam, z_fr, fcfg)
        e = rel_l2(out, ref_f)
        say("   fresnel", dt, e)
        assert e <= tol
        if do_rs:
            out = run_dev(G, ctx, "rs", u, dt, d, lam, z_rs, rcfg)
            e = rel_l2(out, ref_r)
            say("   rs", dt, e)
            assert e <= tol
    ctx.synchronize()


@pytest.mark.parametrize("v", [0, 1, 2, 3])
def test_fast_p8192_vs_generic(G, v):
    P, n = 8192, 4096
    d, lam = 2e-6, 633e-9
    u = field(n, 77 + v)
    fast, gen = G.Context(0), G.Context(0)
    gen.force_generic(True)
    for dbl in (False, True):
        assert fast.set_fft_variant(P, v, double=dbl) == 0
    say("P8192 variant", v)
    z = 0.2 * n * d
    acfg = T.AsmConfig(2.0, 1, 0, 2.0, 1.0)
    for dt, tol in (("c64", C64_TOL), ("c128", F64_TOL)):
        a = run_dev(G, fast, "asm", u, dt, d, lam, z, acfg)
        b = run_dev(G, gen, "asm", u, dt, d, lam, z, acfg)
        e = rel_l2(a, b)
        say("   asm fast vs generic", dt, e)
        assert e <= tol
    zf = 1.5 * n * d * d / lam
    a = run_dev(G, fast, "fresnel", u, "c128", d, lam, zf, T.FresnelConfig(2.0, 0, 0, 2.0))
    b = run_dev(G, gen, "fresnel", u, "c128", d, lam, zf, T.FresnelConfig(2.0, 0, 0, 2.0))
    e = rel_l2(a, b)
    say("   fresnel fast vs generic c128", e)
    assert e <= F64_TOL


def test_fast_asm_round_trip_p8192(G):
    """A field band-limited well inside the propagating disc and the Matsushima
    limit, with zero border, returns to itself under ASM(+z) then ASM(-z)."""
    import torch
    n, d, lam = 4096, 2e-6, 633e-9
    y, x = np.mgrid[:n, :n]
    c = (n - 1) / 2
    u = np.exp(-(((x - c) ** 2 + (y - c) ** 2) / (0.08 * n) ** 2)) * np.exp(1j * 0.002 * (x + 2 * y))
    ctx = G.Context(0)
    cfg = T.AsmConfig(2.0, 1, 0, 2.0, 1.0)
    z = 40 * d
    t = torch.as_tensor(u).cuda()
    st, a, _ = G.asm_propagate_dev(t, d, d, lam, z, cfg, ctx=ctx)
    assert st == 0
    st, b, _ = G.asm_propagate_dev(a, d, d, lam, -z, cfg, ctx=ctx)
    assert st == 0
    e = rel_l2(b.cpu().numpy(), u)
    say("round trip", e)
    assert e <= 1e-10


def test_autotune_picks_a_variant(G):
    ctx = G.Context(0)
    for P, dbl in ((2048, False), (4096, False), (1024, True)):
        v, ms = ctx.autotune(P, double=dbl)
        say("autotune", P, dbl, v, ms)
        assert v in (0, 1, 2, 3)
        assert ms[v] > 0 and ms[v] == min(m for m in ms if m > 0)
        assert ctx.fft_variant(P, double=dbl) == v
    v, ms = ctx.autotune(96)
    assert v == -1 and ms == [-1.0] * G.FFT_VARIANTS
    assert ctx.set_fft_variant(512, 1) != 0  # one layout only below 1024


def block_width(P, v):
    """Width B of the blocked intermediate of complex64 layout variant v at
    padded length P >= 1024: G = min(row-thread cap / (P/16), rows that fit
    140 KB of LDS at the row stride, 16) with caps 1024 for variants 0 and 2
    and 512 otherwise; B = min(G * (1, 1, 1, 2)[v], P/16, 16)."""
    t = P // 16
    slots = P + P // 16
    rstride = slots + (40 - slots % 32) % 32
    g = min((512 if v in (1, 3) else 1024) // t, 140 * 1024 // (rstride * 8), 16)
    return min(g * (1, 1, 1, 2)[v], t, 16)


@pytest.mark.parametrize("P", [1024, 4096])
def test_fresnel_sep_layout_pairs_are_bit_identical(G, P):
    """The two passes of the complex64 separable Fresnel take their layouts
    separately (any pair with one blocked width).  A layout only regroups
    whole transforms per workgroup, so every pair must give the same bits,
    masked or not; a pair with different widths is refused."""
    import torch
    n = P // 2
    d, lam = 2e-6, 550e-9
    u = field(n, P + 5)
    rng = np.random.default_rng(P + 6)
    phi = rng.uniform(0, 2 * np.pi, (n, n))
    mask, keep = G.dev_mask(phi, d, d)
    t = torch.as_tensor(u.astype(np.complex64)).cuda()
    z = 1.5 * n * d * d / lam
    cfg = T.FresnelConfig(2.0, 0, 0, 2.0)
    ctx = G.Context(0)
    nv = G.FFT_VARIANTS
    for v in range(nv):
        assert ctx.set_fft_variant(P, v) == 0
        assert ctx.fresnel_sep_layout(P) == (v, v)  # a forced variant drives both passes
    widths = {v: block_width(P, v) for v in range(nv)}
    say("block widths", P, widths)
    ref = None
    pairs = [(a, b) for a in range(nv) for b in range(nv) if widths[a] == widths[b]]
    assert len(pairs) > 4
    for a, b in pairs:
        assert ctx.set_fresnel_sep_layout(P, a, b) == 0
        assert ctx.fresnel_sep_layout(P) == (a, b)
        ctx.enable_timing(True)
        st, out, _ = G.fresnel_propagate_dev(t, d, d, lam, z, cfg, ctx=ctx)
        assert st == 0
        st, out_m, _ = G.fresnel_propagate_dev(t, d, d, lam, z, cfg, mask=mask, ctx=ctx)
        assert st == 0
        assert ctx.pass_timing()["sep_col"][1] == 2  # both calls ran the two-pass form
        ctx.enable_timing(False)
        got = (out.cpu().numpy(), out_m.cpu().numpy())
        say("sep pair", P, a, b)
        if ref is None:
            ref = got
        else:
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), (a, b)
    bad = [(a, b) for a in range(nv) for b in range(nv) if widths[a] != widths[b]]
    assert bad
    for a, b in bad:
        assert ctx.set_fresnel_sep_layout(P, a, b) == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    assert ctx.set_fresnel_sep_layout(P, nv, 0) == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    assert ctx.set_fresnel_sep_layout(96, 0, 0) == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    v, _ = ctx.autotune(P)
    a, b = ctx.fresnel_sep_layout(P)
    say("tuned", P, v, (a, b))
    assert (a, b) in pairs and ctx.fft_variant(P) == v
    ctx.close()


SEP_CASES = [(1024, 2.0), (1000, 2.048), (600, 3.4133), (777, 2.6357)]


@pytest.mark.parametrize("n,pad", SEP_CASES, ids=[f"n{n}-pad{p}" for n, p in SEP_CASES])
def test_fresnel_c64_two_pass_and_fallback(G, orc, n, pad):
    """complex64 Fresnel runs the two-pass separable form when the x offset
    of the embedding falls on a block boundary of the layout (n = 1024:
    offset 512, always) and the three-pass form otherwise (n = 777: odd
    offset 635, always; 1000 and 600 depend on the layout's block width);
    both forms, masked and not, in every layout, against the oracle."""
    import math
    d, lam = 3e-6, 600e-9
    P = math.ceil(n * pad)
    assert P == 2048
    u = field(n, n)
    rng = np.random.default_rng(n + 1)
    phi = rng.uniform(0, 2 * np.pi, (n // 2, n // 2))
    z = 1.2 * n * d * d / lam
    cfg = T.FresnelConfig(pad, 0, 0, pad)
    st, ref, _ = orc.fresnel_propagate(u, d, d, lam, z, cfg)
    assert st == 0
    om = orc.Mask(phi, 2 * d, 2 * d)
    st, um = orc.phasemask_apply(u, d, d, lam, om)
    assert st == 0
    st, ref_m, _ = orc.fresnel_propagate(um, d, d, lam, z, cfg)
    assert st == 0
    mask, keep = G.dev_mask(phi, 2 * d, 2 * d)
    import torch
    t = torch.as_tensor(u.astype(np.complex64)).cuda()
    for v in variants(P):
        ctx = G.Context(0)
        assert ctx.set_fft_variant(P, v) == 0
        ctx.enable_timing(True)
        st, out, _ = G.fresnel_propagate_dev(t, d, d, lam, z, cfg, ctx=ctx)
        assert st == 0
        st, out_m, _ = G.fresnel_propagate_dev(t, d, d, lam, z, cfg, mask=mask, ctx=ctx)
        assert st == 0
        timing = ctx.pass_timing()
        e = rel_l2(out.cpu().numpy().astype(np.complex128), ref)
        em = rel_l2(out_m.cpu().numpy().astype(np.complex128), ref_m)
        two = timing["sep_col"][1] > 0
        say("   variant", v, "two-pass" if two else "three-pass", e, em)
        assert e <= C64_TOL and em <= C64_TOL
        if n == 777:
            assert not two
        if n == 1024:
            assert two
        ctx.close()


@pytest.mark.parametrize("engine", [T.ENGINE_ASM, T.ENGINE_FRESNEL])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_shared_mask_batch_equals_single(G, engine, dtype):
    """Bands of one field share its phase mask: the batch evaluates the
    transmission once and multiplies by it.  That product is the same
    arithmetic as the per-problem fused mask, so every band must equal its
    single-problem result bit for bit."""
    import torch
    n, d = 512, 2e-6
    lams = [450e-9, 532e-9, 633e-9]
    rng = np.random.default_rng(5)
    phi = rng.uniform(0, 2 * np.pi, (n, n))
    mask, keep = G.dev_mask(phi, d, d, aperture=(T.APERTURE_CIRCLE, 0.45 * n * d, 0.0))
    cdt = torch.complex64 if dtype == "c64" else torch.complex128
    u = torch.as_tensor(field(n, 9)).to(cdt).cuda()
    outs = [torch.empty_like(u) for _ in lams]
    arr = (T.PropDesc * len(lams))()
    for b, lam in enumerate(lams):
        x = arr[b]
        x.geom = T.geom(n, n, d, d, lam)
        x.engine = engine
        x.z, x.padding, x.band_limited, x.strict, x.max_padding, x.medium = 2e-3, 2.0, 1, 0, 2.0, 1.0
        x.mask = T.C.pointer(mask)
        x.in_ = u.data_ptr()
        x.out = outs[b].data_ptr()
    ctx = G.Context(0)
    ctx.enable_timing(True)
    assert G.propagate_batch_dev(arr, double=dtype == "c128", ctx=ctx) == 0
    timing = ctx.pass_timing()
    assert timing["other"][1] == 1  # one transmission evaluation for the three bands
    for b, lam in enumerate(lams):
        cfg = T.AsmConfig(2.0, 1, 0, 2.0, 1.0) if engine == T.ENGINE_ASM else T.FresnelConfig(2.0, 0, 0, 2.0)
        fn = G.asm_propagate_dev if engine == T.ENGINE_ASM else G.fresnel_propagate_dev
        st, single, _ = fn(u, d, d, lam, 2e-3, cfg, mask=mask, ctx=ctx)
        assert st == 0
        say("   band", b, dtype, "max |diff|", float((outs[b] - single).abs().max()))
        assert torch.equal(outs[b], single)
    ctx.close()


@pytest.mark.parametrize("engine", [T.ENGINE_ASM, T.ENGINE_FRESNEL])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_mixed_mask_batch_equals_single(G, engine, dtype):
    """One batch with a shared mask (two bands), a mask of its own and no
    mask at all: the launch takes the mixed pre-op kernels (complex64 Fresnel:
    the two-pass form) and every problem must equal its single-problem
    result bit for bit."""
    import torch
    n, d = 512, 2e-6
    rng = np.random.default_rng(8)
    shared, k1 = G.dev_mask(rng.uniform(0, 2 * np.pi, (n, n)), d, d, aperture=(T.APERTURE_CIRCLE, 0.45 * n * d, 0.0))
    own, k2 = G.dev_mask(rng.uniform(0, 2 * np.pi, (n, n)), d, d)
    cdt = torch.complex64 if dtype == "c64" else torch.complex128
    u = torch.as_tensor(field(n, 11)).to(cdt).cuda()
    plan = [(450e-9, shared), (532e-9, shared), (600e-9, own), (633e-9, None)]
    outs = [torch.empty_like(u) for _ in plan]
    arr = (T.PropDesc * len(plan))()
    for b, (lam, m) in enumerate(plan):
        x = arr[b]
        x.geom = T.geom(n, n, d, d, lam)
        x.engine = engine
        x.z, x.padding, x.band_limited, x.strict, x.max_padding, x.medium = 2e-3, 2.0, 1, 0, 2.0, 1.0
        if m is not None:
            x.mask = T.C.pointer(m)
        x.in_ = u.data_ptr()
        x.out = outs[b].data_ptr()
    ctx = G.Context(0)
    say("mixed-mask batch", engine, dtype)
    assert G.propagate_batch_dev(arr, double=dtype == "c128", ctx=ctx) == 0
    for b, (lam, m) in enumerate(plan):
        cfg = T.AsmConfig(2.0, 1, 0, 2.0, 1.0) if engine == T.ENGINE_ASM else T.FresnelConfig(2.0, 0, 0, 2.0)
        fn = G.asm_propagate_dev if engine == T.ENGINE_ASM else G.fresnel_propagate_dev
        st, single, _ = fn(u, d, d, lam, 2e-3, cfg, mask=m, ctx=ctx)
        assert st == 0
        say("   problem", b, "max |diff|", float((outs[b] - single).abs().max()))
        assert torch.equal(outs[b], single)
    ctx.close()


@pytest.mark.parametrize("kind", ["rs", "fconv"])
def test_fast_convolution_p8192_remapped_vs_generic(G, kind):
    """RS and the Fresnel convolution at N = 3000 with padding 2: the
    reference's padded length 6000 is not a power of two, and since
    6000 >= 2N - 1 the linear convolution runs at 8192 on the
    register-resident kernels (kernel spectrum from f_rs_kgen/f_rs_kcol, the
    field's column pass in mirror-paired tile order).  Compared with the
    generic kernels at the same length in both precisions; the report keeps
    the reference's 6000."""
    import torch
    n, d, lam = 3000, 2e-6, 633e-9
    u = field(n, 91)
    fast, gen = G.Context(0), G.Context(0)
    gen.force_generic(True)
    z = 3e-3
    for dt, tol in (("c64", C64_TOL), ("c128", F64_TOL)):
        outs = []
        for ctx in (fast, gen):
            t = torch.as_tensor(u.astype(np.complex64 if dt == "c64" else np.complex128)).cuda()
            if kind == "rs":
                st, out, rep = G.rs_propagate_dev(t, d, d, lam, z, T.RsConfig(2.0, 0, 0), ctx=ctx)
            else:
                st, out, rep = G.fresnel_propagate_convolution_dev(t, d, d, lam, z, T.FresnelConfig(2.0, 0, 0, 2.0),
                                                                   ctx=ctx)
            assert st == 0, (kind, dt, T.STATUS[st], ctx.last_error())
            assert rep.nx_pad == 6000 and rep.ny_pad == 6000
            outs.append(out.cpu().numpy().astype(np.complex128))
        e = rel_l2(outs[0], outs[1])
        say("P8192 remapped", kind, dt, e)
        assert e <= tol
