"""oc_propagate_batch_noalloc_dev_c64 (the complex64 batch on the library's
per-device scratch array, no device allocation) against
oc_propagate_batch_dev_c64 through the C ABI: the same kernels and tables,
so the same bits and statuses.  Two calls in a row, the second smaller than
the first (nothing of the first call's scratch contents may leak into it),
then a batch beyond the scratch array, refused before anything runs."""
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


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def field(ny, nx, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)


def build(torch, specs, mask):
    """specs: (engine, nx, ny, d, lam, z, masked); returns descs, inputs, outputs."""
    arr = (T.PropDesc * len(specs))()
    ins, outs = [], []
    for k, (eng, nx, ny, d, lam, z, masked) in enumerate(specs):
        ins.append(torch.as_tensor(field(ny, nx, 10 + k)).cuda())
        outs.append(torch.full_like(ins[-1], complex(7.0, -7.0)))
        x = arr[k]
        x.geom = T.geom(nx, ny, d, d, lam)
        x.engine = eng
        x.z, x.padding, x.band_limited, x.strict, x.max_padding, x.medium = z, 2.0, 1, 0, 2.0, 1.0
        if masked:
            x.mask = T.C.pointer(mask)
        x.in_, x.out = ins[-1].data_ptr(), outs[-1].data_ptr()
    return arr, ins, outs


def run_both(G, torch, ctx, specs, mask):
    a, ins, outs_a = build(torch, specs, mask)
    b, _, outs_b = build(torch, specs, mask)
    for k in range(len(specs)):  # same inputs for both forms
        b[k].in_ = ins[k].data_ptr()
    st_b = G.propagate_batch_dev(b, ctx=ctx)
    st_a = G.propagate_batch_noalloc_dev(a, ctx=ctx)
    torch.cuda.synchronize()
    return st_a, st_b, a, b, outs_a, outs_b


def test_noalloc_batch_equals_batch(G):
    import torch
    ctx = G.Context(0)
    n = 512
    vals = 2 * math.pi * W.splitmix_uniform(3, n * n).reshape(n, n)
    mask, keep = G.dev_mask(vals, 2e-6, 2e-6, aperture=(T.APERTURE_CIRCLE, 400e-6, 0.0))
    first = [(T.ENGINE_ASM, n, n, 2e-6, 500e-9, 3e-3, True),          # shared transmission (two bands)
             (T.ENGINE_ASM, n, n, 2e-6, 600e-9, 3e-3, True),
             (T.ENGINE_FRESNEL, 384, 300, 3e-6, 550e-9, 0.05, False),  # runtime-plan kernels (P 768 x 600)
             (T.ENGINE_RS, 256, 256, 4e-6, 633e-9, 8e-3, False),       # kernel spectra
             (T.ENGINE_ASM, n, n, 2e-6, 700e-9, 0.0, True),            # masked identity
             (T.ENGINE_ASM, 64, 64, 2e-6, 700e-9, float("nan"), False)]  # InvalidDistance
    second = [(T.ENGINE_ASM, 128, 128, 4e-6, 532e-9, 1e-3, False)]
    for name, specs in (("first", first), ("second, smaller", second)):
        st_a, st_b, a, b, oa, ob = run_both(G, torch, ctx, specs, mask)
        say(name, "statuses", st_a, st_b, [a[k].status for k in range(len(specs))])
        assert st_a == st_b
        for k in range(len(specs)):
            assert a[k].status == b[k].status, (name, k)
            if b[k].status == 0:
                assert torch.equal(oa[k], ob[k]), (name, k)
    assert st_a == T.CODE["OC_OK"]  # the second call


def test_noalloc_refuses_beyond_scratch(G):
    import torch
    ctx = G.Context(0)
    # four 2048^2 problems at P = 4096: 4 x 64 MiB of intermediates > 256 MiB
    specs = [(T.ENGINE_ASM, 2048, 2048, 2e-6, 500e-9 + 1e-8 * k, 20e-3, False) for k in range(4)]
    a, ins, outs = build(torch, specs, None)
    st = G.propagate_batch_noalloc_dev(a, ctx=ctx)
    torch.cuda.synchronize()
    msg = ctx.last_error()
    say("refused:", st, msg)
    assert st == T.CODE["OC_ERR_INVALID_ARGUMENT"]
    assert "scratch" in msg and "MiB" in msg
    assert all(a[k].status == T.CODE["OC_ERR_INVALID_ARGUMENT"] for k in range(len(specs)))
    for o in outs:  # nothing ran
        assert bool((o == complex(7.0, -7.0)).all())
    # one of them fits and matches the allocating call
    st_a, st_b, a1, b1, oa, ob = run_both(G, torch, ctx, specs[:1], None)
    assert st_a == st_b == 0
    assert torch.equal(oa[0], ob[0])
