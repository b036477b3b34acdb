"""fd::gradient (render + MseToTarget objective) and the Adam epoch on the
HIP path through the C ABI, against the f64 oracle and the reference's FD
and Adam tests.

Bars: complex128 gradient <= 1e-9 relative (max norm) and merits <= 1e-13
against the oracle (the device reduces each merit in a different order);
FD against the adjoint gradient <= 1e-5 (complex128) and <= 2e-2
(complex64: each merit carries the FP32 rounding of the intensities,
divided by the probe separation 2h = 1.3e-3 rad); Adam against the oracle
<= 1e-14 absolute (the gradient norm is reduced in another order).
"""
import math
import sys

import numpy as np
import pytest

import fd_cases as fc
import oc_types as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def dev_fd(G, dtype):
    """fd callable with the oracle's call shape over the device form."""
    import torch

    def fd(designs, u, dx, dy, lam, cfg, target, up, dn, lo, hi):
        keep, dd = [], []
        for pair in designs:
            pp = []
            for steps in pair:
                ss = []
                for (m, z, med) in steps:
                    if m is None:
                        ss.append((None, z, med))
                        continue
                    s, t = G.dev_mask(m.values, m.dx, m.dy, aperture=m.aperture, center=m.center)
                    keep.append(t)
                    ss.append((s, z, med))
                pp.append(ss)
            dd.append(tuple(pp))
        ud = torch.as_tensor(np.ascontiguousarray(u)).to(dtype).cuda()
        td = torch.as_tensor(np.ascontiguousarray(target, np.float64)).cuda()
        return G.fd_gradient_dev(dd, ud, dx, dy, lam, cfg, td, up, dn, lo, hi)
    return fd


def dev_adam(G):
    import torch

    def step(x, g, m, v, lo, hi, t, cfg):
        ts = [torch.as_tensor(np.ascontiguousarray(a, np.float64)).cuda() for a in (x, g, m, v, lo, hi)]
        st = G.adam_step_dev(*ts, t, cfg)
        torch.cuda.synchronize()
        x[:] = ts[0].cpu().numpy()
        m[:] = ts[2].cpu().numpy()
        v[:] = ts[3].cpu().numpy()
        return st
    return step


@pytest.mark.parametrize("form", ["c128", "c64"])
@pytest.mark.parametrize("case", fc.FD_CASES, ids=lambda f: f.__name__)
def test_fd_case_gpu(G, orc, case, form):
    import torch
    say("fd case", case.__name__, form)
    fd = dev_fd(G, torch.complex128 if form == "c128" else torch.complex64)
    if case is fc.fd_matches_adjoint:
        case(fd, orc, tol=1e-5 if form == "c128" else 2e-2)
    else:
        case(fd, orc)


def test_fd_matches_oracle_c128(G, orc):
    import torch
    p1, p2, u, tgt = fc.problem()
    designs, up, dn, lo, hi = fc.probes(p1, p2)
    st, g0, m0, e0 = orc.fd_gradient(designs, u, fc.D, fc.D, fc.LAM, fc.CFG, tgt, up, dn, lo, hi)
    assert st == 0
    st, g1, m1, e1 = dev_fd(G, torch.complex128)(designs, u, fc.D, fc.D, fc.LAM, fc.CFG, tgt, up, dn, lo, hi)
    assert st == 0 and e1 == e0
    eg = np.abs(g1 - g0).max() / np.abs(g0).max()
    em = np.abs(m1 - m0).max() / np.abs(m0).max()
    say("fd c128 vs oracle: grad", eg, "merits", em)
    assert eg < 1e-9 and em < 1e-13


@pytest.mark.parametrize("case", fc.ADAM_CASES, ids=lambda f: f.__name__)
def test_adam_case_gpu(G, case):
    case(dev_adam(G))


@pytest.mark.parametrize("t,cfg", [(1, T.AdamConfig.default()), (7, T.AdamConfig.default(has_clip=0)),
                                   (3, T.AdamConfig.default(clip_norm=1e3, lr_frac=0.1))])
def test_adam_matches_oracle(G, orc, t, cfg):
    rng = np.random.default_rng(t)
    n = 200_003
    lo, hi = -np.ones(n), np.ones(n) * 2.0
    lo[::7] = hi[::7] = 0.5  # zero-width slots stay put
    x = rng.uniform(-1.0, 2.0, n)
    x[::7] = 0.5
    g = rng.standard_normal(n) * 1e-2
    m = rng.standard_normal(n) * 1e-3
    v = rng.uniform(0.0, 1e-4, n)
    ref = [a.copy() for a in (x, m, v)]
    assert orc.adam_step(ref[0], g, ref[1], ref[2], lo, hi, t, cfg) == 0
    got = [a.copy() for a in (x, m, v)]
    assert dev_adam(G)(got[0], g, got[1], got[2], lo, hi, t, cfg) == 0
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() < 1e-14
    assert np.all(got[0][::7] == 0.5)


def test_adam_errors(G):
    import torch
    z = torch.zeros(8, dtype=torch.float64, device="cuda")
    o = torch.ones(8, dtype=torch.float64, device="cuda")
    assert G.adam_step_dev(z.clone(), z, z.clone(), z.clone(), z, o, 0, T.AdamConfig.default()) == \
        T.CODE["OC_ERR_INVALID_ARGUMENT"]
    for bad in (dict(lr_frac=0.0), dict(beta1=1.0), dict(beta2=-0.1), dict(eps=-1.0), dict(clip_norm=0.0),
                dict(lr_frac=math.nan)):
        assert G.adam_step_dev(z.clone(), z, z.clone(), z.clone(), z, o, 1, T.AdamConfig.default(**bad)) == \
            T.CODE["OC_INV_RENDER"], bad
