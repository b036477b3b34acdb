"""The context's descriptor cache: a transient batch whose descriptors are
byte-identical to a kept device copy launches on that copy without a new
upload; other batches take the least recently used slot.  Alternating and
repeated calls, and more distinct batches than there are slots, must give
the results of fresh calls bit for bit."""
import sys

import numpy as np
import pytest

import oc_types as T

pytestmark = pytest.mark.gpu

SLOTS = 64  # oc_context::kDescSlots


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def field(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))


def desc(n, d, lam, z, src, dst, engine=T.ENGINE_ASM):
    arr = (T.PropDesc * 1)()
    x = arr[0]
    x.geom = T.geom(n, n, d, d, lam)
    x.engine = engine
    x.z, x.padding, x.band_limited, x.strict, x.max_padding, x.medium = z, 2.0, 1, 0, 2.0, 1.0
    x.in_ = src.data_ptr()
    x.out = dst.data_ptr()
    return arr


@pytest.mark.parametrize("double", [False, True])
def test_cached_descriptors_give_fresh_results(G, double):
    import torch
    n, d, lam = 64, 4e-6, 633e-9
    cdt = torch.complex128 if double else torch.complex64
    src = torch.as_tensor(field(n, 5)).to(cdt).cuda()
    dst = torch.empty_like(src)
    zs = [1e-3 * (1 + k) for k in range(SLOTS + 8)]
    # reference outputs, each from a fresh context (empty cache)
    ref = {}
    for z in (zs[0], zs[1], zs[-1]):
        ctx = G.Context(0)
        assert G.propagate_batch_dev(desc(n, d, lam, z, src, dst), double=double, ctx=ctx) == 0
        torch.cuda.synchronize()
        ref[z] = dst.cpu().numpy().copy()
        ctx.close()
    ctx = G.Context(0)

    def run(z):
        dst.zero_()
        assert G.propagate_batch_dev(desc(n, d, lam, z, src, dst), double=double, ctx=ctx) == 0
        torch.cuda.synchronize()
        return dst.cpu().numpy()

    # alternating hits and misses
    for z in (zs[0], zs[1], zs[0], zs[1], zs[0]):
        assert np.array_equal(run(z), ref[z]), z
    say("alternating calls equal fresh calls")
    # more distinct batches than slots: the first ones are evicted and
    # uploaded again when they come back
    for z in zs:
        run(z)
    assert np.array_equal(run(zs[-1]), ref[zs[-1]])
    assert np.array_equal(run(zs[0]), ref[zs[0]])
    assert np.array_equal(run(zs[1]), ref[zs[1]])
    say("after", len(zs), "distinct batches the outputs still equal fresh calls")
    ctx.close()


def test_cached_descriptors_follow_new_buffers(G):
    """Same parameters, other buffers: the descriptors differ in their
    pointers, so each call reads its own input and writes its own output."""
    import torch
    n, d, lam, z = 64, 4e-6, 532e-9, 2e-3
    a = torch.as_tensor(field(n, 7)).to(torch.complex64).cuda()
    b = torch.as_tensor(field(n, 8)).to(torch.complex64).cuda()
    oa, ob = torch.empty_like(a), torch.empty_like(b)
    ctx = G.Context(0)
    for _ in range(3):
        oa.zero_()
        ob.zero_()
        assert G.propagate_batch_dev(desc(n, d, lam, z, a, oa), ctx=ctx) == 0
        assert G.propagate_batch_dev(desc(n, d, lam, z, b, ob), ctx=ctx) == 0
        torch.cuda.synchronize()
        ra, rb = oa.cpu().numpy(), ob.cpu().numpy()
        assert not np.array_equal(ra, rb)
    ctx2 = G.Context(0)
    ref = torch.empty_like(a)
    assert G.propagate_batch_dev(desc(n, d, lam, z, b, ref), ctx=ctx2) == 0
    torch.cuda.synchronize()
    assert np.array_equal(rb, ref.cpu().numpy())
    ctx.close()
    ctx2.close()
