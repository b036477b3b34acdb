"""The reference's golden fixture on the HIP path: the hypercam forward
render (crates/optics-inverse/tests/hypercam_forward.rs:89-207, fixture
crates/optics-inverse/tests/data/hypercam/{forward,thickness}.npy, copied to
tests/golden/hypercam/) through oc_render, as the Renderer the reference test
builds: a Gaussian source at each scene point, gap z_gap, the thickness DOE
(phase 2 pi (n - 1) h / lambda) then f, a thin lens then 2f, a thin lens
then f; ASM at padding 1 without band limit; x and y scene derivative
channels by central differences of step h_scene.

The thin lens is a PhaseMask sampled on the field lattice with its circular
stop: on every sample it equals ThinLens::apply (checked here against the
oracle's ThinLens at 1e-12, as test_lens_mask_equals_thinlens does for the
renderer cases).

Bars: max|obs - fixture| / max|fixture| < 1e-6 (hypercam_forward.rs:202) for
the f64 host form and the complex128 device form.  complex64: the intensity
channel within 1e-5 of its own peak, and the whole observation within 1e-4
of the fixture's peak: the derivative channels divide FP32 intensity
differences (relative rounding ~1e-7 of the peak intensity) by 2 h = 2 um,
which turns them into errors of ~5e-2 I_max per metre against derivative
peaks of ~1e3 I_max per metre.
"""
import math
import os
import sys

import numpy as np
import pytest

import oc_types as T
import render_cases as rcases

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hypercam")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def read_meta():
    out = {}
    for line in open(os.path.join(HERE, "meta.txt")):
        line = line.strip()
        if line and not line.startswith("#"):
            k, v = line.split("=", 1)
            out[k] = float(v)
    return out


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def setup_case(orc):
    meta = read_meta()
    thickness = np.load(os.path.join(HERE, "thickness.npy"))
    ref = np.load(os.path.join(HERE, "forward.npy"))
    n, p, lam, f = int(meta["n"]), meta["pitch"], meta["lambda"], meta["f_relay"]
    h = meta["h_scene"]
    ap = (T.APERTURE_CIRCLE, meta["aperture_r"], 0.0)
    lens = rcases.lens_values(n, p, lam, f)
    # the sampled lens equals ThinLens::apply on every sample
    ones = np.ones((n, n), np.complex128)
    st, a = orc.phasemask_apply(ones, p, p, lam, orc.Mask(lens, p, p, aperture=ap))
    assert st == 0
    st, b = orc.thinlens_apply(ones, p, p, lam, f, aperture=ap)
    assert st == 0
    assert np.abs(a - b).max() < 1e-12
    doe = 2.0 * math.pi * (meta["index"] - 1.0) / lam * thickness
    scenes = [(0.0, 0.0), (h, 0.0), (-h, 0.0), (0.0, h), (0.0, -h)]
    # renders per frame: base, x + h, x - h, y + h, y - h (the source closure)
    inputs = np.zeros((1, len(scenes), 5, n, n), np.complex128)
    for s, (x0, y0) in enumerate(scenes):
        for r, (ex, ey) in enumerate(((0, 0), (h, 0), (-h, 0), (0, h), (0, -h))):
            inputs[0, s, r] = orc.gaussian(meta["w0"], 1.0, (x0 + ex, y0 + ey), 0.0, lam, n, n, p, p)
    rc = T.RenderConfig.make(x=True, y=True, scene_step=h)
    cfg = T.AsmConfig(1.0, 0, 0, 1.0, 1.0)
    return dict(meta=meta, ref=ref, n=n, p=p, lam=lam, f=f, ap=ap, lens=lens, doe=doe, inputs=inputs, rc=rc,
                cfg=cfg, zg=meta["z_gap"])


def design(mk, c):
    f = c["f"]
    return [[[(None, c["zg"], 1.0), (mk(c["doe"], None), f, 1.0), (mk(c["lens"], c["ap"]), 2 * f, 1.0),
              (mk(c["lens"], c["ap"]), f, 1.0)]]]


def score(obs, ref):
    return float(np.max(np.abs(obs - ref)) / max(np.max(np.abs(ref)), 1e-30))


def test_hypercam_fixture_host_f64(G, orc):
    c = setup_case(orc)
    p = c["p"]

    def mk(v, ap):
        return G.Mask(v, p, p) if ap is None else G.Mask(v, p, p, aperture=ap)
    st, obs = G.render(design(mk, c), c["inputs"], p, p, [c["lam"]], c["cfg"], c["rc"])
    assert st == 0, T.STATUS[st]
    assert obs.shape == c["ref"].shape
    e = score(obs, c["ref"])
    say("hypercam f64 host", e)
    assert e < 1e-6


@pytest.mark.parametrize("form,tol,tol_i", [("c128", 1e-6, 1e-6), ("c64", 1e-4, 1e-5)])
def test_hypercam_fixture_device(G, orc, form, tol, tol_i):
    import torch
    c = setup_case(orc)
    p = c["p"]
    keep = []

    def mk(v, ap):
        s, k = G.dev_mask(v, p, p) if ap is None else G.dev_mask(v, p, p, aperture=ap)
        keep.append(k)
        return s
    dt = torch.complex128 if form == "c128" else torch.complex64
    u = torch.as_tensor(c["inputs"]).to(dt).cuda().contiguous()
    obs = torch.zeros(c["ref"].shape, dtype=torch.float64, device="cuda")
    st = G.render_dev(design(mk, c), u, p, p, [c["lam"]], c["cfg"], c["rc"], obs)
    assert st == 0, T.STATUS[st]
    o = obs.cpu().numpy()
    e = score(o, c["ref"])
    ei = score(o[0], c["ref"][0])
    say("hypercam", form, "all channels", e, "intensity", ei)
    assert e < tol and ei < tol_i
