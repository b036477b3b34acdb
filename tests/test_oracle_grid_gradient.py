"""Pins the oracle's grid_gradient (adjoint.rs:268-340) with MseToTarget
(merit.rs:112-166) over (band, scene) frames against central finite
differences of what its seeds differentiate:

- every frame selected: obs_gradient seeds frame f with 2 (I_f - T)/n, so the
  summed planes are the gradient of J = sum_f mean((I_f - T)^2) over the
  lattice values shared by the bands' designs (a band-independent phase);
- one frame selected: the gradient of MseToTarget::evaluate itself, and the
  reported merit equals evaluate;
- a selector out of range, or a seed that is zero everywhere, gives no planes.

FD through the oracle's trace (a different code path from push_back's tape
and reverse sweep); h = 1e-5 rad, relative L2 over the probed pixels < 1e-4.
"""
import math

import numpy as np

import oc_types as T
import workloads as W

N, D = 24, 8e-6
BANDS = [520e-9, 610e-9]
SCENES = [(0.0, 0.0), (16e-6, -8e-6)]
Z1, Z2 = 1.2e-3, 0.8e-3


def setup(orc):
    p1 = 2 * math.pi * W.splitmix_uniform(21, N * N).reshape(N, N)
    p2 = 2 * math.pi * W.splitmix_uniform(22, N * N).reshape(N, N)
    inputs = np.zeros((len(BANDS), len(SCENES), N, N), complex)
    for b, lam in enumerate(BANDS):
        for s, c in enumerate(SCENES):
            inputs[b, s] = W.gaussian_beam(N, D, lam, 60e-6, center=c)
    target = np.abs(np.random.default_rng(3).standard_normal((N, N))) * 0.05
    return p1, p2, inputs, target


def designs(orc, p1, p2):
    ap = (T.APERTURE_CIRCLE, 80e-6, 0.0)
    return [[(orc.Mask(p1, D, D, aperture=ap), Z1, 1.0), (orc.Mask(p2, D, D), Z2, 1.0)] for _ in BANDS]


def intensities(orc, p1, p2, inputs, cfg):
    out = np.zeros((len(BANDS), len(SCENES), N, N))
    ds = designs(orc, p1, p2)
    for b, lam in enumerate(BANDS):
        for s in range(len(SCENES)):
            st, f = orc.trace(ds[b], inputs[b, s], D, D, lam, cfg)
            assert st == 0
            out[b, s] = f.real * f.real + f.imag * f.imag  # norm_sqr
    return out


def test_grid_gradient_all_frames_is_sum_of_frame_mse_gradients(orc):
    p1, p2, inputs, target = setup(orc)
    cfg = T.AsmConfig.default()
    st, g, planes, merit = orc.grid_gradient(designs(orc, p1, p2), inputs, D, D, BANDS, cfg, target)
    assert st == 0 and planes == 2
    inten = intensities(orc, p1, p2, inputs, cfg)
    summed = inten.sum(axis=(0, 1))
    assert abs(merit - np.mean((summed - target) ** 2)) <= 1e-13 * merit  # evaluate: frames summed first

    def J(q1, q2):
        i = intensities(orc, q1, q2, inputs, cfg)
        return sum(np.mean((i[b, s] - target) ** 2) for b in range(len(BANDS)) for s in range(len(SCENES)))

    h = 1e-5
    fd, ad = [], []
    for plane in (0, 1):
        for k in W.workload4_fd_pixels(n=N, count=6)[plane]:
            iy, ix = divmod(int(k), N)
            vals = []
            for sgn in (1, -1):
                q = [p1.copy(), p2.copy()]
                q[plane][iy, ix] += sgn * h
                vals.append(J(*q))
            fd.append((vals[0] - vals[1]) / (2 * h))
            ad.append(g[plane, iy, ix])
    fd, ad = np.array(fd), np.array(ad)
    assert np.linalg.norm(fd - ad) / np.linalg.norm(fd) < 1e-4


def test_grid_gradient_one_frame_is_gradient_of_evaluate(orc):
    p1, p2, inputs, target = setup(orc)
    cfg = T.AsmConfig.default()
    b, s = 1, 0
    st, g, planes, merit = orc.grid_gradient(designs(orc, p1, p2), inputs, D, D, BANDS, cfg, target, b, s)
    assert st == 0 and planes == 2
    inten = intensities(orc, p1, p2, inputs, cfg)
    assert abs(merit - np.mean((inten[b, s] - target) ** 2)) <= 1e-13 * merit
    # the selected frame alone is push_back of its own seed
    st, pb = orc.adjoint_push_back(designs(orc, p1, p2)[b], inputs[b, s], D, D, BANDS[b], cfg,
                                   2.0 * (inten[b, s] - target) / target.size)
    assert st == 0
    assert np.array_equal(g, pb + 0.0)
    h = 1e-5
    fd, ad = [], []
    for k in W.workload4_fd_pixels(n=N, count=6)[0]:
        iy, ix = divmod(int(k), N)
        vals = []
        for sgn in (1, -1):
            q1 = p1.copy()
            q1[iy, ix] += sgn * h
            i = intensities(orc, q1, p2, inputs, cfg)
            vals.append(np.mean((i[b, s] - target) ** 2))
        fd.append((vals[0] - vals[1]) / (2 * h))
        ad.append(g[0, iy, ix])
    fd, ad = np.array(fd), np.array(ad)
    assert np.linalg.norm(fd - ad) / np.linalg.norm(fd) < 1e-4


def test_grid_gradient_no_planes(orc):
    p1, p2, inputs, target = setup(orc)
    cfg = T.AsmConfig.default()
    st, g, planes, merit = orc.grid_gradient(designs(orc, p1, p2), inputs, D, D, BANDS, cfg, target, 5, -1)
    assert st == 0 and planes == 0 and math.isnan(merit)
    # a target equal to the only frame's intensity: the seed is zero everywhere
    inten = intensities(orc, p1, p2, inputs, cfg)
    st, g, planes, merit = orc.grid_gradient(designs(orc, p1, p2)[:1], inputs[:1, :1], D, D, BANDS[:1], cfg,
                                             inten[0, 0])
    assert st == 0 and planes == 0 and merit == 0.0
