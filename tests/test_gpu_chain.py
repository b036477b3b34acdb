"""oc_propagate_chain_dev_*: a propagation chained into a second one of its
output, as two gaps of Beamline::trace (src/beamline.rs:109-130).  The fused
complex64 form (masked ASM into Fresnel at a power-of-two padded length)
must give the ASM intermediate bit for bit as the ASM batch does, and the
final field within the complex64 bar of the two batch calls and of the f64
oracle; every other chain equals the two batch calls bit for bit (complex128
runs its two three-pass steps linked through their rows, the rest as two
batches).  Statuses: a failing first stage repeats its status on the second."""
import sys

import numpy as np
import pytest

import oc_types as T

pytestmark = pytest.mark.gpu

C64_TOL = 1e-5
NON_FINITE = 5  # OC_NON_FINITE_SAMPLE


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


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


def descs(engine, lams, n, d, ins, outs, z, mask=None):
    arr = (T.PropDesc * len(lams))()
    for b, lam in enumerate(lams):
        x = arr[b]
        x.geom = T.geom(n, n, d, d, lam)
        x.engine = engine
        x.z, x.padding, x.band_limited, x.strict, x.max_padding, x.medium = z, 2.0, 1, 0, 2.0, 1.0
        if mask is not None:
            x.mask = T.C.pointer(mask)
        x.in_ = ins[b].data_ptr() if ins is not None else 0
        x.out = outs[b].data_ptr() if outs is not None else 0
    return arr


def two_batches(G, ctx, u, n, d, lams, z1, z2, mask, dbl, mask2=None):
    import torch
    mid = [torch.empty_like(x) for x in u]
    out = [torch.empty_like(x) for x in u]
    assert G.propagate_batch_dev(descs(T.ENGINE_ASM, lams, n, d, u, mid, z1, mask), double=dbl, ctx=ctx) == 0
    assert G.propagate_batch_dev(descs(T.ENGINE_FRESNEL, lams, n, d, mid, out, z2, mask2), double=dbl, ctx=ctx) == 0
    return mid, out


@pytest.mark.parametrize("n", [512, 2048])
def test_fused_chain_matches_two_batches(G, n):
    import torch
    d = 2e-6
    lams = [450e-9, 550e-9, 650e-9]
    mask, keep = G.dev_mask(np.random.default_rng(n).uniform(0, 2 * np.pi, (n, n)), d, d)
    u = [torch.as_tensor(field(n, n + b)).to(torch.complex64).cuda() for b in range(len(lams))]
    z1, z2 = 20e-3 * n / 2048, 0.2 * n / 2048
    ctx = G.Context(0)
    mid_ref, out_ref = two_batches(G, ctx, u, n, d, lams, z1, z2, mask, False)
    mid = [torch.empty_like(x) for x in u]
    out = [torch.empty_like(x) for x in u]
    out2 = [torch.empty_like(x) for x in u]
    ctx.enable_timing(True)
    ctx.pass_timing(reset=True)
    first = descs(T.ENGINE_ASM, lams, n, d, u, mid, z1, mask)
    second = descs(T.ENGINE_FRESNEL, lams, n, d, None, out, z2)
    assert G.propagate_chain_dev(first, second, ctx=ctx) == 0
    tm = ctx.pass_timing()
    say("chain", n, {k: v[1] for k, v in tm.items()})
    assert tm["row_inv"][1] == 0 and tm["sep_row"][1] == 1 and tm["sep_col"][1] == 1  # the fused passes ran
    assert all(first[b].status == 0 and second[b].status == 0 for b in range(len(lams)))
    # no intermediate asked for: the same final field
    first_n = descs(T.ENGINE_ASM, lams, n, d, u, None, z1, mask)
    second_n = descs(T.ENGINE_FRESNEL, lams, n, d, None, out2, z2)
    assert G.propagate_chain_dev(first_n, second_n, ctx=ctx) == 0
    for b in range(len(lams)):
        assert torch.equal(mid[b], mid_ref[b])  # the ASM stage's arithmetic is unchanged
        assert torch.equal(out2[b], out[b])
        e = rel_l2(out[b].cpu().numpy().astype(np.complex128), out_ref[b].cpu().numpy().astype(np.complex128))
        say("   band", b, "fused vs two batches", e)
        assert e <= 1e-6
    ctx.close()


def test_fused_chain_vs_oracle(G, orc):
    import torch
    n, d, lam = 256, 4e-6, 532e-9
    u = field(n, 7)
    phi = np.random.default_rng(8).uniform(0, 2 * np.pi, (n, n))
    mask, keep = G.dev_mask(phi, d, d)
    om = orc.Mask(phi, d, d)
    st, um = orc.phasemask_apply(u, d, d, lam, om)
    assert st == 0
    z1, z2 = 4e-3, 0.05
    st, ra, _ = orc.asm_propagate(um, d, d, lam, z1, T.AsmConfig(2.0, 1, 0, 2.0, 1.0))
    assert st == 0
    st, rf, _ = orc.fresnel_propagate(ra, d, d, lam, z2, T.FresnelConfig(2.0, 0, 0, 2.0))
    assert st == 0
    t = [torch.as_tensor(u.astype(np.complex64)).cuda()]
    mid, out = [torch.empty_like(t[0])], [torch.empty_like(t[0])]
    ctx = G.Context(0)
    first = descs(T.ENGINE_ASM, [lam], n, d, t, mid, z1, mask)
    second = descs(T.ENGINE_FRESNEL, [lam], n, d, None, out, z2)
    assert G.propagate_chain_dev(first, second, ctx=ctx) == 0
    ea = rel_l2(mid[0].cpu().numpy().astype(np.complex128), ra)
    ef = rel_l2(out[0].cpu().numpy().astype(np.complex128), rf)
    say("chain vs oracle", ea, ef)
    assert ea <= C64_TOL and ef <= C64_TOL
    ctx.close()


@pytest.mark.parametrize("case", ["c128", "masked_second", "odd_size", "generic"])
def test_unfused_chains_equal_two_batches(G, case):
    import torch
    n = 777 if case == "odd_size" else 256
    d = 3e-6
    lams = [500e-9, 600e-9]
    dbl = case == "c128"
    cdt = torch.complex128 if dbl else torch.complex64
    rng = np.random.default_rng(11)
    mask, keep = G.dev_mask(rng.uniform(0, 2 * np.pi, (n, n)), d, d)
    mask2, keep2 = G.dev_mask(rng.uniform(0, 2 * np.pi, (n, n)), d, d)
    u = [torch.as_tensor(field(n, 40 + b)).to(cdt).cuda() for b in range(len(lams))]
    ctx = G.Context(0)
    if case == "generic":
        ctx.force_generic(True)
    m2 = mask2 if case == "masked_second" else None
    z1, z2 = 3e-3, 0.03
    mid_ref, out_ref = two_batches(G, ctx, u, n, d, lams, z1, z2, mask, dbl, m2)
    mid = [torch.empty_like(x) for x in u]
    out = [torch.empty_like(x) for x in u]
    ctx.enable_timing(True)
    ctx.pass_timing(reset=True)
    first = descs(T.ENGINE_ASM, lams, n, d, u, mid, z1, mask)
    second = descs(T.ENGINE_FRESNEL, lams, n, d, None, out, z2, m2)
    assert G.propagate_chain_dev(first, second, double=dbl, ctx=ctx) == 0
    tm = ctx.pass_timing()
    say("unfused chain", case, {k: v[1] for k, v in tm.items()})
    # complex128 (three-pass Fresnel): the steps run linked through their
    # rows, one forward row pass; the others as two batch calls
    assert tm["row_fwd"][1] == (1 if case == "c128" else 2 if case != "masked_second" else 1)
    for b in range(len(lams)):
        assert torch.equal(mid[b], mid_ref[b]) and torch.equal(out[b], out_ref[b]), (case, b)
    # without the intermediate: a scratch buffer carries it
    out2 = [torch.empty_like(x) for x in u]
    first = descs(T.ENGINE_ASM, lams, n, d, u, None, z1, mask)
    second = descs(T.ENGINE_FRESNEL, lams, n, d, None, out2, z2, m2)
    assert G.propagate_chain_dev(first, second, double=dbl, ctx=ctx) == 0
    for b in range(len(lams)):
        assert torch.equal(out2[b], out_ref[b])
    ctx.close()


@pytest.mark.parametrize("fused", [True, False])
def test_chain_statuses(G, fused):
    import torch
    n, d = 256, 3e-6
    lams = [500e-9, 550e-9, 600e-9]
    cdt = torch.complex64 if fused else torch.complex128
    u = [torch.as_tensor(field(n, 60 + b)).to(cdt).cuda() for b in range(len(lams))]
    u[1][7, 9] = float("nan")
    out = [torch.empty_like(x) for x in u]
    ctx = G.Context(0)
    first = descs(T.ENGINE_ASM, lams, n, d, u, None, 3e-3)
    second = descs(T.ENGINE_FRESNEL, lams, n, d, None, out, 0.03)
    first[2].z = float("nan")
    st = G.propagate_chain_dev(first, second, double=not fused, ctx=ctx)
    say("statuses", fused, st, [(first[b].status, second[b].status) for b in range(3)])
    # problem 2's host-side failure makes the whole chain unfused; problem 1's
    # NaN is found on the device
    assert st == T.CODE["OC_NON_FINITE_SAMPLE"]
    assert (first[0].status, second[0].status) == (0, 0)
    assert first[1].status == second[1].status == T.CODE["OC_NON_FINITE_SAMPLE"]
    assert first[2].status == second[2].status == T.CODE["OC_INVALID_DISTANCE"]
    # fused with a device-side failure only
    first = descs(T.ENGINE_ASM, lams, n, d, u, None, 3e-3)
    second = descs(T.ENGINE_FRESNEL, lams, n, d, None, out, 0.03)
    ctx.enable_timing(True)
    ctx.pass_timing(reset=True)
    st = G.propagate_chain_dev(first, second, double=not fused, ctx=ctx)
    assert (ctx.pass_timing()["sep_row"][1] == 1) == fused
    assert st == T.CODE["OC_NON_FINITE_SAMPLE"]
    assert [first[b].status for b in range(3)] == [0, T.CODE["OC_NON_FINITE_SAMPLE"], 0]
    assert [second[b].status for b in range(3)] == [0, T.CODE["OC_NON_FINITE_SAMPLE"], 0]
    rmid, ref = [torch.empty_like(u[0])], [torch.empty_like(u[0])]
    assert G.propagate_batch_dev(descs(T.ENGINE_ASM, lams[:1], n, d, u[:1], rmid, 3e-3), double=not fused,
                                 ctx=ctx) == 0
    assert G.propagate_batch_dev(descs(T.ENGINE_FRESNEL, lams[:1], n, d, rmid, ref, 0.03), double=not fused,
                                 ctx=ctx) == 0
    e = rel_l2(out[0].cpu().numpy().astype(np.complex128), ref[0].cpu().numpy().astype(np.complex128))
    assert e <= (1e-6 if fused else 0.0)
    # a second stage on another grid is refused before anything runs
    bad = descs(T.ENGINE_FRESNEL, lams, n, d * 2, None, out, 0.03)
    assert G.propagate_chain_dev(first, bad, double=not fused, ctx=ctx) == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    ctx.close()


def test_chain_argument_checks(G):
    """count 0 is a no-op; missing arrays with a count are refused."""
    ctx = G.Context(0)
    empty = (T.PropDesc * 1)()
    assert G.lib.oc_propagate_chain_dev_c64(ctx.h, empty, empty, 0) == 0
    assert G.lib.oc_propagate_chain_dev_c128(ctx.h, None, None, 0) == 0
    assert G.lib.oc_propagate_chain_dev_c64(ctx.h, None, empty, 1) == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    assert G.lib.oc_propagate_chain_dev_c64(ctx.h, empty, None, 1) == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    assert G.lib.oc_propagate_chain_dev_c64(None, empty, empty, 1) == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    ctx.close()


@pytest.mark.parametrize("second_mask", [False, True])
def test_linked_asm_chain_4096_equals_two_batches(G, second_mask):
    """complex64 ASM into ASM at P = 4096: the steps run linked through their
    rows (f_row_link, which keeps the second radix-16 stage's twiddles in
    LDS beside its exchange buffer); bit for bit the two batch calls."""
    import torch
    n, d = 2048, 2e-6
    lams = [480e-9, 620e-9]
    rng = np.random.default_rng(21)
    mask, keep = G.dev_mask(rng.uniform(0, 2 * np.pi, (n, n)), d, d)
    mask2, keep2 = G.dev_mask(rng.uniform(0, 2 * np.pi, (n, n)), d, d)
    m2 = mask2 if second_mask else None
    u = [torch.as_tensor(field(n, 50 + b)).to(torch.complex64).cuda() for b in range(len(lams))]
    ctx = G.Context(0)
    z1, z2 = 8e-3, 12e-3
    mid_ref = [torch.empty_like(x) for x in u]
    out_ref = [torch.empty_like(x) for x in u]
    assert G.propagate_batch_dev(descs(T.ENGINE_ASM, lams, n, d, u, mid_ref, z1, mask), ctx=ctx) == 0
    assert G.propagate_batch_dev(descs(T.ENGINE_ASM, lams, n, d, mid_ref, out_ref, z2, m2), ctx=ctx) == 0
    mid = [torch.empty_like(x) for x in u]
    out = [torch.empty_like(x) for x in u]
    ctx.enable_timing(True)
    ctx.pass_timing(reset=True)
    first = descs(T.ENGINE_ASM, lams, n, d, u, mid, z1, mask)
    second = descs(T.ENGINE_ASM, lams, n, d, None, out, z2, m2)
    assert G.propagate_chain_dev(first, second, ctx=ctx) == 0
    tm = ctx.pass_timing()
    say("linked ASM chain 4096", second_mask, {k: v[1] for k, v in tm.items()})
    assert tm["row_fwd"][1] == 1  # one forward row pass: the second runs inside the link
    for b in range(len(lams)):
        assert torch.equal(mid[b], mid_ref[b]) and torch.equal(out[b], out_ref[b]), (second_mask, b)
    ctx.close()


@pytest.mark.parametrize("case", ["fused", "linked_c128", "two_batches"])
def test_submitted_chains_equal_synchronous(G, case):
    """oc_propagate_chain_submit_dev_*: chains queued without waiting give
    the synchronous chain's outputs bit for bit; statuses after the wait.
    A non-finite input sample fails that problem's chain (both stages) at
    the wait and leaves the others OK."""
    import torch
    n = 777 if case == "two_batches" else 512
    d = 3e-6
    lams = [480e-9, 560e-9, 640e-9]
    dbl = case == "linked_c128"
    cdt = torch.complex128 if dbl else torch.complex64
    rng = np.random.default_rng(31)
    mask, keep = G.dev_mask(rng.uniform(0, 2 * np.pi, (n, n)), d, d)
    ctx = G.Context(0)
    z1, z2 = 4e-3, 0.04
    reps = 3
    ins = [[torch.as_tensor(field(n, 60 + 10 * r + b)).to(cdt).cuda() for b in range(len(lams))]
           for r in range(reps)]
    ins[1][1][n // 2, n // 3] = float("nan")
    refs, stats = [], []
    for r in range(reps):
        out = [torch.empty_like(x) for x in ins[r]]
        first = descs(T.ENGINE_ASM, lams, n, d, ins[r], None, z1, mask)
        second = descs(T.ENGINE_FRESNEL, lams, n, d, None, out, z2)
        G.propagate_chain_dev(first, second, double=dbl, ctx=ctx)
        refs.append(out)
        stats.append(([first[i].status for i in range(len(lams))], [second[i].status for i in range(len(lams))]))
    say("synchronous statuses", case, stats)
    assert stats[1][0][1] == NON_FINITE and stats[1][1][1] == NON_FINITE
    outs, keep_descs = [], []
    torch.cuda.synchronize()
    for r in range(reps):
        out = [torch.empty_like(x) for x in ins[r]]
        first = descs(T.ENGINE_ASM, lams, n, d, ins[r], None, z1, mask)
        second = descs(T.ENGINE_FRESNEL, lams, n, d, None, out, z2)
        st = G.propagate_chain_submit_dev(first, second, double=dbl, ctx=ctx, sync_torch=False)
        assert st == 0 or case == "two_batches", (case, r, st)
        outs.append(out)
        keep_descs.append((first, second))
    st = G.wait_dev(ctx=ctx)
    # a chain that runs as two batches (odd size) had its statuses at the submit
    assert st == (0 if case == "two_batches" else NON_FINITE), st
    for r in range(reps):
        first, second = keep_descs[r]
        got = ([first[i].status for i in range(len(lams))], [second[i].status for i in range(len(lams))])
        assert got == stats[r], (case, r, got, stats[r])
        for b in range(len(lams)):
            if stats[r][1][b] == 0:
                assert torch.equal(outs[r][b], refs[r][b]), (case, r, b)
    ctx.close()
