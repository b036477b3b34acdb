"""The batch schedule tuner (oc_autotune_batch_dev_*, oc_schedule_*): every
schedule it can pick computes the same propagation (against the f64 oracle:
asm.rs:439-479, rayleigh_sommerfeld.rs:384-442), and its picks survive a
save / load into a fresh context.

- Radix plans of the LDS kernels forced one by one (radix 16, 8, 4, 2 and
  3 stages in several orders) on ASM at P = 256 and Rayleigh-Sommerfeld at
  its own P = 768 (the pad pick keeps the reference's length), complex128 at
  1e-12 and complex64 at 1e-5.
- The tuner on an uneven RS batch (N = 96, 128, 192: P = 192, 256, 384, each
  either at its own length or at the next power of two): outputs against
  the oracle, search time and picks reported; the saved picks reload into a
  new context and give bit-identical outputs there.
- Malformed files and files tagged with another device.
"""
import math
import os
import sys

import numpy as np
import pytest

import oc_types as T
import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def field(n, seed):
    th = 2 * math.pi * W.splitmix_uniform(seed, n * n).reshape(n, n)
    x = np.arange(n) - (n - 1) / 2
    inside = x[None, :] ** 2 + x[:, None] ** 2 <= (0.4 * n) ** 2
    return np.where(inside, np.exp(1j * th), 0.0)


def run_one(G, torch, dt, u, engine, d, lam, z, padding, ctx):
    x = torch.as_tensor(u).to(dt).cuda().contiguous()
    y = torch.empty_like(x)
    arr = (T.PropDesc * 1)()
    arr[0].geom = T.geom(u.shape[1], u.shape[0], d, d, lam)
    arr[0].z = z
    arr[0].engine = engine
    arr[0].band_limited = 1
    arr[0].strict = 0
    arr[0].padding = padding
    arr[0].max_padding = padding
    arr[0].medium = 1.0
    arr[0].in_ = x.data_ptr()
    arr[0].out = y.data_ptr()
    st = G.propagate_batch_dev(arr, double=dt == torch.complex128, ctx=ctx)
    torch.cuda.synchronize()
    return st, y.cpu().numpy()


PLANS_256 = [[16, 16], [4, 4, 16], [8, 8, 4], [4, 4, 4, 4], [2, 8, 16], [16, 8, 2]]
PLANS_768 = [[16, 16, 3], [3, 16, 16], [8, 8, 4, 3], [3, 4, 4, 4, 4], [16, 4, 4, 3], [2, 8, 16, 3]]


@pytest.mark.parametrize("dbl", [True, False], ids=["c128", "c64"])
def test_forced_radix_plans_match_oracle(G, orc, dbl):
    import torch
    ctx = G.Context(0)
    dt, tol = (torch.complex128, 1e-12) if dbl else (torch.complex64, 1e-5)
    # ASM at P = 256 (N = 128, padding 2)
    u = field(128, 3)
    d, lam, z = 4e-6, 550e-9, 2e-3
    cfg = T.AsmConfig(2.0, 1, 0, 2.0, 1.0)
    st, ref, _ = orc.asm_propagate(u, d, d, lam, z, cfg)
    assert st == 0
    for plan in PLANS_256:
        G.schedule_clear(ctx=ctx)
        assert G.schedule_set(dict(length=256, double=dbl, kind="lds", problems=1, radix=plan), ctx=ctx) == 0
        st, out = run_one(G, torch, dt, u, T.ENGINE_ASM, d, lam, z, 2.0, ctx)
        assert st == 0
        e = rel(out, ref)
        say("asm 256", plan, e)
        assert e <= tol, (plan, e)
    # RS at its own P = 768 (N = 384): the pad pick keeps 768, the LDS plan runs it
    u = field(384, 4)
    d, lam, z = 3e-6, 600e-9, 1e-3
    st, ref, rep = orc.rs_propagate(u, d, d, lam, z, T.RsConfig(2.0, 0, 0))
    assert st == 0 and rep.nx_pad == 768
    for plan in PLANS_768:
        G.schedule_clear(ctx=ctx)
        assert G.schedule_set(dict(length=768, double=dbl, kind="pad", problems=1, run_length=768), ctx=ctx) == 0
        assert G.schedule_set(dict(length=768, double=dbl, kind="lds", problems=1, radix=plan), ctx=ctx) == 0
        st, out = run_one(G, torch, dt, u, T.ENGINE_RS, d, lam, z, 2.0, ctx)
        assert st == 0
        e = rel(out, ref)
        say("rs 768", plan, e)
        assert e <= tol, (plan, e)
    # and the power-of-two run length with each fast layout
    for v in range(4):
        G.schedule_clear(ctx=ctx)
        G.schedule_set(dict(length=768, double=dbl, kind="pad", problems=1, run_length=1024), ctx=ctx)
        G.schedule_set(dict(length=1024, double=dbl, kind="fast", problems=1, variant=v), ctx=ctx)
        st, out = run_one(G, torch, dt, u, T.ENGINE_RS, d, lam, z, 2.0, ctx)
        assert st == 0
        e = rel(out, ref)
        say("rs 768 -> 1024 layout", v, e)
        assert e <= tol, (v, e)
    ctx.close()


def uneven_batch(G, torch, sizes=(96, 128, 192, 96, 128, 192, 192)):
    ps, ins, outs = [], [], []
    for i, n in enumerate(sizes):
        p = dict(n=n, d=(3 + i % 3) * 1e-6, lam=(450 + 40 * i) * 1e-9, z=(1 + i) * 1e-3)
        ps.append(p)
        ins.append(torch.as_tensor(field(n, 10 + i)).to(torch.complex64).cuda().contiguous())
        outs.append(torch.empty_like(ins[-1]))
    arr = (T.PropDesc * len(ps))()
    for i, p in enumerate(ps):
        arr[i].geom = T.geom(p["n"], p["n"], p["d"], p["d"], p["lam"])
        arr[i].z = p["z"]
        arr[i].engine = T.ENGINE_RS
        arr[i].padding = 2.0
        arr[i].strict = 0
        arr[i].in_ = ins[i].data_ptr()
        arr[i].out = outs[i].data_ptr()
    return ps, ins, outs, arr


def test_tuner_picks_reload_and_match(G, orc, tmp_path):
    import torch
    ctx = G.Context(0)
    ps, ins, outs, arr = uneven_batch(G, torch)
    st, picks, search_ms = G.autotune_batch_dev(arr, sweeps=1, ctx=ctx)
    assert st == 0, T.STATUS[st]
    say("search ms", search_ms)
    for p in picks:
        say("  pick", p)
    assert search_ms > 0
    kinds = {(p["kind"], p["length"]) for p in picks}
    # a pad choice for each non-power-of-two reference length, a path for each run length
    assert ("pad", 192) in kinds and ("pad", 384) in kinds
    assert any(k == "fast" or k == "lds" for k, _ in kinds)
    assert all(p["ms"] > 0 for p in picks)
    # the tuned batch against the oracle
    st = G.propagate_batch_dev(arr, ctx=ctx)
    assert st == 0
    torch.cuda.synchronize()
    first = [o.cpu().numpy() for o in outs]
    for i, p in enumerate(ps):
        u = ins[i].cpu().numpy().astype(np.complex128)
        st, ref, _ = orc.rs_propagate(u, p["d"], p["d"], p["lam"], p["z"], T.RsConfig(2.0, 0, 0))
        e = rel(first[i], ref)
        assert e <= 1e-5, (i, e)
    # save, reload into a fresh context: same picks, bit-identical outputs
    path = str(tmp_path / "schedules.txt")
    assert G.schedule_save(path, ctx=ctx) == 0
    text = open(path).read()
    say(text)
    saved = G.schedules(ctx=ctx)
    ctx2 = G.Context(0)
    st, n = G.schedule_load(path, ctx=ctx2)
    assert st == 0 and n == len(saved) and n >= len(picks)
    again = G.schedules(ctx=ctx2)
    assert [dict(x, ms=0) for x in again] == [dict(x, ms=0) for x in saved]
    assert all(abs(a["ms"] - b["ms"]) <= 1e-9 * b["ms"] for a, b in zip(again, saved))
    for o in outs:
        o.zero_()
    st = G.propagate_batch_dev(arr, ctx=ctx2)
    assert st == 0
    torch.cuda.synchronize()
    for i in range(len(ps)):
        assert np.array_equal(outs[i].cpu().numpy(), first[i]), i
    ctx2.close()
    ctx.close()


def test_schedule_file_errors_and_validation(G, tmp_path):
    ctx = G.Context(0)
    bad = tmp_path / "bad.txt"
    bad.write_text("optics_gpu schedules 1\ndevice gfx950 256\nfast 1024 0 4 1.0\n")  # missing variant
    st, n = G.schedule_load(str(bad), ctx=ctx)
    assert st == T.CODE["OC_ERR_INVALID_ARGUMENT"] and n == 0
    bad.write_text("not a schedule file\n")
    assert G.schedule_load(str(bad), ctx=ctx)[0] == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    other = tmp_path / "other.txt"
    other.write_text("optics_gpu schedules 1\ndevice gfx000 1\nfast 1024 0 4 1.0 2\nlds 768 0 1 2.0 3 16 16 3\n")
    st, n = G.schedule_load(str(other), ctx=ctx)
    assert st == 0 and n == 0 and G.schedules(ctx=ctx) == []
    assert G.schedule_load(str(tmp_path / "missing.txt"), ctx=ctx)[0] == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    # plans whose product is not the length, unknown radices, bad layouts
    assert G.schedule_set(dict(length=768, kind="lds", problems=1, radix=[16, 16, 2]), ctx=ctx) != 0
    assert G.schedule_set(dict(length=121, kind="lds", problems=1, radix=[11, 11]), ctx=ctx) != 0
    assert G.schedule_set(dict(length=1024, kind="fast", problems=1, variant=7), ctx=ctx) != 0
    assert G.schedule_set(dict(length=1024, kind="pad", problems=1, run_length=512), ctx=ctx) != 0
    assert G.schedules(ctx=ctx) == []
    ctx.close()
