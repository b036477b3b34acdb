"""oc_propagate_batch_submit_dev_* / oc_context_wait: queued batches run in
submit order (a chain of dependent batches equals the synchronous calls bit
for bit), host-detected statuses come back at the submit and device-side
ones at the wait."""
import sys

import numpy as np
import pytest

import oc_types as T

pytestmark = pytest.mark.gpu


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


def descs(engine, lams, n, d, ins, outs, z, mask=None):
    arr = (T.PropDesc * len(lams))()
    for b, lam in enumerate(lams):
        x = arr[b]
        x.geom = T.geom(n, n, d, d, lam)
        x.engine = engine
        x.z, x.padding, x.band_limited, x.strict, x.max_padding, x.medium = z, 2.0, 1, 0, 2.0, 1.0
        if mask is not None:
            x.mask = T.C.pointer(mask)
        x.in_ = ins[b].data_ptr()
        x.out = outs[b].data_ptr()
    return arr


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_submitted_chain_equals_synchronous_calls(G, dtype):
    import torch
    n, d = 512, 2e-6
    lams = [450e-9, 550e-9, 650e-9]
    cdt = torch.complex64 if dtype == "c64" else torch.complex128
    dbl = dtype == "c128"
    mask, keep = G.dev_mask(np.random.default_rng(2).uniform(0, 2 * np.pi, (n, n)), d, d)
    u = [torch.as_tensor(field(n, 30 + b)).to(cdt).cuda() for b in range(len(lams))]
    a1 = [torch.empty_like(x) for x in u]
    f1 = [torch.empty_like(x) for x in u]
    a2 = [torch.empty_like(x) for x in u]
    f2 = [torch.empty_like(x) for x in u]
    ctx = G.Context(0)
    s_asm = descs(T.ENGINE_ASM, lams, n, d, u, a1, 2e-3, mask)
    s_fr = descs(T.ENGINE_FRESNEL, lams, n, d, a1, f1, 20e-3)
    assert G.propagate_batch_dev(s_asm, double=dbl, ctx=ctx) == 0
    assert G.propagate_batch_dev(s_fr, double=dbl, ctx=ctx) == 0
    q_asm = descs(T.ENGINE_ASM, lams, n, d, u, a2, 2e-3, mask)
    q_fr = descs(T.ENGINE_FRESNEL, lams, n, d, a2, f2, 20e-3)
    for rep in range(3):
        say("submit chain", dtype, rep)
        assert G.propagate_batch_submit_dev(q_asm, double=dbl, ctx=ctx) == 0
        assert G.propagate_batch_submit_dev(q_fr, double=dbl, ctx=ctx) == 0
        assert G.wait_dev(ctx) == 0
        for b in range(len(lams)):
            assert torch.equal(a2[b], a1[b]) and torch.equal(f2[b], f1[b])
    assert G.wait_dev(ctx) == 0  # nothing queued
    ctx.close()


def test_submit_statuses(G):
    import torch
    n, d = 256, 2e-6
    lams = [500e-9, 600e-9]
    good = torch.as_tensor(field(n, 40)).to(torch.complex64).cuda()
    bad = good.clone()
    bad[5, 7] = complex(float("nan"), 0.0)
    outs = [torch.empty_like(good) for _ in range(2)]
    ctx = G.Context(0)
    # host-detected: a non-finite distance is reported by the submit itself
    arr = descs(T.ENGINE_ASM, lams, n, d, [good, good], outs, 1e-3)
    arr[1].z = float("nan")
    st = G.propagate_batch_submit_dev(arr, ctx=ctx)
    assert st == T.CODE["OC_INVALID_DISTANCE"] and arr[1].status == st and arr[0].status == 0
    # device-detected: a non-finite sample is known only at the wait
    arr2 = descs(T.ENGINE_ASM, lams, n, d, [good, bad], outs, 1e-3)
    assert G.propagate_batch_submit_dev(arr2, ctx=ctx) == 0
    st = G.wait_dev(ctx)
    assert st == T.CODE["OC_INVALID_DISTANCE"]  # the first non-OK status in submit order
    assert arr[0].status == 0 and arr2[0].status == 0 and arr2[1].status == T.CODE["OC_NON_FINITE_SAMPLE"]
    ctx.close()


@pytest.mark.parametrize("shared", [False, True])
def test_threads_on_contexts(G, shared):
    """Calls from several host threads, each on its own context or all on
    one shared context (serialised by the context's lock), give the results
    of the same calls made one after another."""
    import threading
    import torch
    n, d = 256, 2e-6
    lams = [450e-9, 500e-9, 550e-9, 600e-9]
    u = [torch.as_tensor(field(n, 50 + b)).to(torch.complex64).cuda() for b in range(len(lams))]
    ref = [torch.empty_like(x) for x in u]
    got = [torch.empty_like(x) for x in u]
    seq = G.Context(0)
    for b, lam in enumerate(lams):
        assert G.propagate_batch_dev(descs(T.ENGINE_ASM, [lam], n, d, [u[b]], [ref[b]], 1e-3), ctx=seq) == 0
    torch.cuda.synchronize()
    one = G.Context(0) if shared else None
    ctxs = [one if shared else G.Context(0) for _ in lams]
    errs = []

    def work(b):
        try:
            for _ in range(5):
                arr = descs(T.ENGINE_ASM, [lams[b]], n, d, [u[b]], [got[b]], 1e-3)
                st = G.lib.oc_propagate_batch_dev_c64(ctxs[b].h, arr, len(arr))
                if st:
                    errs.append(st)
                G.lib.oc_context_synchronize(ctxs[b].h)
        except Exception as e:  # noqa: BLE001
            errs.append(repr(e))

    say("threads on", "one shared context" if shared else "own contexts")
    th = [threading.Thread(target=work, args=(b,)) for b in range(len(lams))]
    for t_ in th:
        t_.start()
    for t_ in th:
        t_.join()
    assert not errs, errs
    torch.cuda.synchronize()
    for b in range(len(lams)):
        assert torch.equal(got[b], ref[b])
