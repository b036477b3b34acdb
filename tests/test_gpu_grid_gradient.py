"""oc_adjoint_grid_gradient_frames_dev_* (adjoint::grid_gradient,
adjoint.rs:268-340, with MseToTarget over every (band, scene) frame) against
the oracle's grid_gradient, through the C ABI.

Bars: complex128 1e-10 relative L2 per plane (the adjoint bar), merit 1e-12;
complex64 1e-4 per plane and 1e-5 on the merit (FP32 fields, FP64 seeds and
gradient sums).  The one-frame call equals the single-frame entry bit for
bit; a selector out of range gives zero planes and a NaN merit.
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


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def case(n=256, d=6e-6, bands=(1.50e-6, 1.55e-6, 1.60e-6), scenes=((0.0, 0.0), (60e-6, -30e-6)), per_band=False):
    p1 = 2 * math.pi * W.splitmix_uniform(31, n * n).reshape(n, n)
    p2 = 2 * math.pi * W.splitmix_uniform(32, n * n).reshape(n, n)
    inputs = np.zeros((len(bands), len(scenes), n, n), complex)
    for b, lam in enumerate(bands):
        for s, c in enumerate(scenes):
            inputs[b, s] = W.gaussian_beam(n, d, lam, 0.3 * n * d, center=c)
    rng = np.random.default_rng(7)
    target = np.abs(rng.standard_normal((n, n))) * 0.02
    # a band-dependent first plane when per_band (a GridThickness-like
    # carrier builds a different lattice per wavelength)
    lat1 = [p1 * (1.55e-6 / lam if per_band else 1.0) for lam in bands]
    return dict(n=n, d=d, bands=list(bands), p1=lat1, p2=p2, inputs=inputs, target=target, z=(4e-3, 3e-3))


def oracle_designs(orc, c):
    d = c["d"]
    ap = (T.APERTURE_CIRCLE, 0.4 * c["n"] * d, 0.0)
    return [[(orc.Mask(c["p1"][b], d, d, aperture=ap), c["z"][0], 1.0), (orc.Mask(c["p2"], d, d), c["z"][1], 1.0)]
            for b in range(len(c["bands"]))]


def device_designs(G, c):
    d = c["d"]
    ap = (T.APERTURE_CIRCLE, 0.4 * c["n"] * d, 0.0)
    keep, out = [], []
    m2, k2 = G.dev_mask(c["p2"], d, d)
    keep.append(k2)
    for b in range(len(c["bands"])):
        m1, k1 = G.dev_mask(c["p1"][b], d, d, aperture=ap)
        keep.append(k1)
        out.append([(m1, c["z"][0], 1.0), (m2, c["z"][1], 1.0)])
    return out, keep


@pytest.mark.parametrize("per_band", [False, True], ids=["shared", "per-band"])
@pytest.mark.parametrize("sel", [(-1, -1), (2, 1), (-1, 0)], ids=["all", "b2s1", "s0"])
def test_frames_gradient_matches_oracle(G, orc, per_band, sel):
    import torch
    c = case(per_band=per_band)
    cfg = T.AsmConfig.default()
    say("grid_gradient frames", per_band, sel)
    st, ref, planes, m_ref = orc.grid_gradient(oracle_designs(orc, c), c["inputs"], c["d"], c["d"], c["bands"], cfg,
                                               c["target"], *sel)
    assert st == 0 and planes == 2
    designs, keep = device_designs(G, c)
    td = torch.as_tensor(c["target"]).cuda()
    for dt, tol, mtol in ((torch.complex128, 1e-10, 1e-12), (torch.complex64, 1e-4, 1e-5)):
        u = torch.as_tensor(c["inputs"]).to(dt).cuda().contiguous()
        grads = torch.full((2, c["n"], c["n"]), 7.0, dtype=torch.float64, device="cuda")
        st, merit = G.adjoint_grid_gradient_frames_dev(designs, u, c["d"], c["d"], c["bands"], cfg, td, grads,
                                                       sel_band=sel[0], sel_scene=sel[1])
        assert st == 0, T.STATUS[st]
        g = grads.cpu().numpy()
        e = [rel(g[k], ref[k]) for k in range(2)]
        say("  ", dt, e, abs(merit - m_ref) / m_ref)
        assert max(e) <= tol
        assert abs(merit - m_ref) <= mtol * m_ref


def test_one_frame_equals_single_frame_entry(G):
    import torch
    c = case(n=128, bands=(1.55e-6,), scenes=((0.0, 0.0),))
    cfg = T.AsmConfig.default()
    designs, keep = device_designs(G, c)
    td = torch.as_tensor(c["target"]).cuda()
    u = torch.as_tensor(c["inputs"]).cuda().contiguous()
    g1 = torch.zeros((2, 128, 128), dtype=torch.float64, device="cuda")
    g2 = torch.zeros_like(g1)
    st, m1 = G.adjoint_grid_gradient_frames_dev(designs, u, c["d"], c["d"], c["bands"], cfg, td, g1)
    assert st == 0
    st, m2 = G.adjoint_grid_gradient_dev(designs[0], u[0, 0], c["d"], c["d"], c["bands"][0], cfg, td, g2)
    assert st == 0
    assert torch.equal(g1, g2) and m1 == m2


def test_frames_selector_out_of_range_and_errors(G):
    import torch
    c = case(n=64)
    cfg = T.AsmConfig.default()
    designs, keep = device_designs(G, c)
    td = torch.as_tensor(c["target"]).cuda()
    u = torch.as_tensor(c["inputs"]).cuda().contiguous()
    grads = torch.full((2, 64, 64), 7.0, dtype=torch.float64, device="cuda")
    st, merit = G.adjoint_grid_gradient_frames_dev(designs, u, c["d"], c["d"], c["bands"], cfg, td, grads,
                                                   sel_band=3)
    assert st == 0 and math.isnan(merit)
    assert float(grads.abs().max()) == 0.0
    # a failing forward propagation (non-finite input) is the renderer's error
    bad = u.clone()
    bad[1, 0, 32, 32] = float("nan")  # inside the first mask's stop
    st, _ = G.adjoint_grid_gradient_frames_dev(designs, bad, c["d"], c["d"], c["bands"], cfg, td, grads)
    assert st == T.CODE["OC_INV_RENDER"]
    # designs of different structure across bands
    odd = [designs[0], designs[1], [(designs[2][0][0], 0.0, 1.0), designs[2][1]]]
    st, _ = G.adjoint_grid_gradient_frames_dev(odd, u, c["d"], c["d"], c["bands"], cfg, td, grads)
    assert st == T.CODE["OC_ERR_UNSUPPORTED"]
