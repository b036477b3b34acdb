"""Pins the oracle's ASM + PhaseMask + ThinLens chain against the reference's
golden fixture crates/optics-inverse/tests/data/hypercam/{forward,thickness}.npy
(copied to tests/golden/hypercam/), rendered the way
crates/optics-inverse/tests/hypercam_forward.rs:89-207 does: padding 1, no band
limit, scene central differences for the x/y derivative channels.  The fixture
was produced by an independent torch ASM, so agreement pins the propagation
numerics end to end.  Tolerance: max|diff|/peak < 1e-6 (hypercam_forward.rs:202).
"""
import math
import os

import numpy as np

import oc_types as T

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hypercam")


def read_meta():
    out = {}
    for line in open(os.path.join(HERE, "meta.txt")):
        line = line.strip()
        if line and not line.startswith("#"):
            k, v = line.split("=", 1)
            out[k] = float(v)
    return out


def render_frame(orc, meta, thickness, x0, y0, lam):
    n = int(meta["n"])
    p = meta["pitch"]
    cfg = T.AsmConfig(1.0, 0, 0, 1.0, 1.0)
    u = orc.gaussian(meta["w0"], 1.0, (x0, y0), 0.0, lam, n, n, p, p)
    scale = 2.0 * math.pi * (meta["index"] - 1.0) / lam
    mask = orc.Mask(scale * thickness, p, p)
    f = meta["f_relay"]
    lens_ap = (T.APERTURE_CIRCLE, meta["aperture_r"], 0.0)

    def prop(v, z):
        st, o, _ = orc.asm_propagate(v, p, p, lam, z, cfg)
        assert st == 0
        return o

    u = prop(u, meta["z_gap"])                       # design.gap(z_gap)
    st, u = orc.phasemask_apply(u, p, p, lam, mask)  # push(mask, f)
    assert st == 0
    u = prop(u, f)
    st, u = orc.thinlens_apply(u, p, p, lam, f, lens_ap)  # push(lens, 2f)
    assert st == 0
    u = prop(u, 2 * f)
    st, u = orc.thinlens_apply(u, p, p, lam, f, lens_ap)  # push(lens, f)
    assert st == 0
    u = prop(u, f)
    return u.real * u.real + u.imag * u.imag


def test_hypercam_forward_fixture(orc):
    meta = read_meta()
    thickness = np.load(os.path.join(HERE, "thickness.npy"))
    ref = np.load(os.path.join(HERE, "forward.npy"))
    n = int(meta["n"])
    h = meta["h_scene"]
    lam = meta["lambda"]
    scenes = [(0.0, 0.0), (h, 0.0), (-h, 0.0), (0.0, h), (0.0, -h)]
    obs = np.zeros((3, 1, 5, n, n))
    for si, (x0, y0) in enumerate(scenes):
        obs[0, 0, si] = render_frame(orc, meta, thickness, x0, y0, lam)
        hi = render_frame(orc, meta, thickness, x0 + h, y0, lam)
        lo = render_frame(orc, meta, thickness, x0 - h, y0, lam)
        obs[1, 0, si] = (hi - lo) / (2.0 * h)
        hi = render_frame(orc, meta, thickness, x0, y0 + h, lam)
        lo = render_frame(orc, meta, thickness, x0, y0 - h, lam)
        obs[2, 0, si] = (hi - lo) / (2.0 * h)
    scale = max(np.max(np.abs(ref)), 1e-30)
    worst = np.max(np.abs(obs - ref))
    assert worst / scale < 1e-6, f"rel {worst / scale:.3e}"
