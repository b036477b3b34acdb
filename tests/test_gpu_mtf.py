"""analysis::mtf on the device through the C ABI: the reference's MTF cases
on the host and device forms, batched parity with the f64 oracle (1e-12
absolute on values in [0, 1]) and the Psf::new errors."""
import math
import sys

import numpy as np
import pytest

import mtf_cases as mc
import oc_types as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def host_mtf(G):
    def f(psf, dx):
        st, fr, v = G.mtf(psf, dx)
        assert st == 0, T.STATUS.get(st, st)
        return fr, v
    return f


def dev_mtf(G):
    import torch

    def f(psf, dx):
        p = torch.as_tensor(np.ascontiguousarray(psf, np.float64))[None].cuda()
        v = torch.zeros((1, psf.shape[1] // 2 + 1), dtype=torch.float64, device="cuda")
        st, fr = G.mtf_dev(p, dx, v)
        assert st == 0, T.STATUS.get(st, st)
        return fr, v[0].cpu().numpy()
    return f


@pytest.mark.parametrize("form", ["host", "dev"])
@pytest.mark.parametrize("case", mc.CASES, ids=lambda f: f.__name__)
def test_mtf_case_gpu(G, case, form):
    case(host_mtf(G) if form == "host" else dev_mtf(G))


def test_mtf_batch_matches_oracle(G, orc):
    import torch
    rng = np.random.default_rng(4)
    count, ny, nx = 24, 200, 333
    ps = rng.uniform(0.0, 1.0, (count, ny, nx)) ** 4
    pd = torch.as_tensor(ps).cuda()
    vd = torch.zeros((count, nx // 2 + 1), dtype=torch.float64, device="cuda")
    st, fr = G.mtf_dev(pd, 4e-6, vd)
    assert st == 0
    got = vd.cpu().numpy()
    for c in range(count):
        f0, v0 = orc.mtf(ps[c], 4e-6)
        assert np.array_equal(fr, f0)
        assert np.abs(got[c] - v0).max() < 1e-12


def test_mtf_errors(G):
    st, _, _ = G.mtf(np.ones((4, 4)), 0.0)
    assert st == T.CODE["OC_INVALID_GRID_PITCH"]
    st, _, _ = G.mtf(np.ones((4, 4)), float("nan"))
    assert st == T.CODE["OC_INVALID_GRID_PITCH"]
    bad = np.ones((4, 4))
    bad[1, 2] = math.inf
    st, _, _ = G.mtf(bad, 1e-3)
    assert st == T.CODE["OC_NON_FINITE_SAMPLE"]
