"""Renderer::render cases (crates/optics-inverse/src/observe.rs:213-314 and
its tests, observe.rs:330-664), written once against a render callable with
the call shape of oracle.render / optics_gpu.render:

    render(designs, inputs, dx, dy, bands, asm_cfg, rc) -> (status, obs)

designs[band][variant] are step lists [(mask or None, z_after, medium)];
inputs are the source closure's fields (bands, scenes, renders, ny, nx) in
the render order base, x+, x-, y+, y-, l+, l-.  The carriers and sources
below restate the reference tests' own (relay chain, flat relief, chromatic
thin lens) as PhaseMask designs; sources come from the closed forms in
workloads.py (Gaussian at the waist, unit plane wave).
"""
import math

import numpy as np

import oc_types as T
import workloads as W

L = 500e-9
NO_BL = T.AsmConfig.default(band_limited=0, strict=0)


class MaskSpec:
    """PhaseMask::sampled arguments with the lattice kept alive; `.struct`
    is what the design tables take (host lattice)."""

    def __init__(self, values, dx, dy, aperture=(T.APERTURE_FIELD, 0.0, 0.0), center=(0.0, 0.0)):
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self.dx, self.dy, self.aperture, self.center = dx, dy, aperture, center
        ny, nx = self.values.shape
        kind, a, b = aperture
        self.struct = T.PhaseMask(self.values.ctypes.data, nx, ny, dx, dy, 0.0, T.ClearAperture(kind, 0, a, b),
                                  center[0], center[1])


def inputs_for(source, bands, scenes, rc):
    """Source closure evaluated at every render (observe.rs:258-291)."""
    frames = []
    for lam in bands:
        per_band = []
        for (x0, y0) in scenes:
            r = [source(x0, y0, lam)]
            if rc.deriv_x:
                r += [source(x0 + rc.scene_step, y0, lam), source(x0 - rc.scene_step, y0, lam)]
            if rc.deriv_y:
                r += [source(x0, y0 + rc.scene_step, lam), source(x0, y0 - rc.scene_step, lam)]
            if rc.deriv_lambda:
                r += [source(x0, y0, lam + rc.lambda_step), source(x0, y0, lam - rc.lambda_step)]
            per_band.append(np.stack(r))
        frames.append(np.stack(per_band))
    return np.stack(frames)


def designs_for(carrier, bands, rc):
    """Carrier::build at every wavelength the render needs."""
    out = []
    for lam in bands:
        v = [carrier(lam)]
        if rc.deriv_lambda:
            v += [carrier(lam + rc.lambda_step), carrier(lam - rc.lambda_step)]
        out.append(v)
    return out


def gaussian_source(w0, n=64, d=10e-6):
    return lambda x0, y0, lam: W.gaussian_beam(n, d, lam, w0, 1.0, (x0, y0))


def plane_wave_source(n=64):
    return lambda x0, y0, lam: np.ones((n, n), np.complex128)


def relay_carrier(d, n, h, grid=64, pitch=10e-6):
    """observe.rs:371-397: gap(-d), uniform relief mask, gap(+d)."""
    def build(lam):
        phase = 2.0 * math.pi * (n - 1.0) * h / lam
        m = MaskSpec(np.full((grid, grid), phase), pitch, pitch)
        return [(None, -d, 1.0), (m, 0.0, 1.0), (None, d, 1.0)]
    return build


def flat_relief(h, n, grid=32, pitch=10e-6):
    """observe.rs:399-411 via carrier.rs:338-353: a 32x32 thickness lattice,
    phase 2 pi (n - 1) / lambda * h, one element, no gap."""
    def build(lam):
        scale = 2.0 * math.pi * (n - 1.0) / lam
        m = MaskSpec(np.full((grid, grid), scale * h), pitch, pitch)
        return [(m, 0.0, 1.0)]
    return build


def lens_values(n, d, lam, f):
    """ThinLens phase -k r^2 / (2 f) on the field lattice (element.rs ThinLens)."""
    x = W.coords(n, d)
    k = 2.0 * math.pi / lam
    return -k * (x[None, :] * x[None, :] + x[:, None] * x[:, None]) / (2.0 * f)


def chromatic_lens(f0, lam0, n=128, d=10e-6, r_ap=300e-6):
    """observe.rs:506-528: f = f0 lambda0 / lambda, circular stop, then f0."""
    def build(lam):
        f = f0 * lam0 / lam
        m = MaskSpec(lens_values(n, d, lam, f), d, d, aperture=(T.APERTURE_CIRCLE, r_ap, 0.0))
        return [(m, f0, 1.0)]
    return build


def ok(st):
    assert st == 0, T.STATUS.get(st, st)


# ---------------------------------------------------------------- cases ---

def defocus_mask_defocus_chain_is_pure_piston(render, tol=1.0):
    rc = T.RenderConfig.make()
    bands = [L]
    src = gaussian_source(80e-6)
    st, obs = render(designs_for(relay_carrier(5e-3, 1.5, 2e-6), bands, rc), inputs_for(src, bands, [(0, 0)], rc),
                     10e-6, 10e-6, bands, NO_BL, rc)
    ok(st)
    i_in = np.abs(src(0, 0, L)) ** 2
    peak = i_in.max()
    assert np.abs(obs[0, 0, 0] - i_in).max() < 1e-6 * peak * tol


def plane_wave_x_y_channels_are_exactly_zero(render, tol=1.0):
    rc = T.RenderConfig.make(x=True, y=True)
    bands = [L]
    st, obs = render(designs_for(relay_carrier(5e-3, 1.5, 2e-6), bands, rc),
                     inputs_for(plane_wave_source(), bands, [(0, 0)], rc), 10e-6, 10e-6, bands,
                     T.AsmConfig.default(), rc)
    ok(st)
    assert np.all(obs[1] == 0.0) and np.all(obs[2] == 0.0)


def dx_channel_matches_translation_invariance(render, tol=1.0):
    rc = T.RenderConfig.make(x=True, scene_step=1e-6)
    bands = [L]
    st, obs = render(designs_for(lambda lam: [(None, 3e-3, 1.0)], bands, rc),
                     inputs_for(gaussian_source(80e-6), bands, [(0, 0)], rc), 10e-6, 10e-6, bands, NO_BL, rc)
    ok(st)
    i, dx = obs[0, 0, 0], obs[1, 0, 0]
    ny, nx = i.shape
    peak = i.max()
    scale = peak / 80e-6
    worst, checked = 0.0, 0
    for iy in range(1, ny - 1):
        for ix in range(2, nx - 2):
            if i[iy, ix] < 0.02 * peak:
                continue
            gx = (i[iy, ix - 2] - 8.0 * i[iy, ix - 1] + 8.0 * i[iy, ix + 1] - i[iy, ix + 2]) / (12.0 * 10e-6)
            worst = max(worst, abs(dx[iy, ix] + gx) / scale)
            checked += 1
    assert checked > 100
    assert worst < 0.02 * tol


def lambda_channel_vanishes_for_piston_only_system(render, tol=1.0):
    rc = T.RenderConfig.make(lam=True)
    bands = [450e-9, 500e-9, 550e-9]
    st, obs = render(designs_for(flat_relief(2e-6, 1.5), bands, rc),
                     inputs_for(plane_wave_source(), bands, [(0, 0)], rc), 10e-6, 10e-6, bands,
                     T.AsmConfig.default(), rc)
    ok(st)
    assert obs.shape[0] == 2
    assert np.abs(obs[1]).max() < 1e-3 * tol


def lambda_channel_detects_chromatic_focus(render, tol=1.0):
    rc = T.RenderConfig.make(lam=True, lambda_step=5e-9)
    bands = [450e-9, 500e-9, 550e-9]
    st, obs = render(designs_for(chromatic_lens(50e-3, 500e-9), bands, rc),
                     inputs_for(plane_wave_source(128), bands, [(0, 0)], rc), 10e-6, 10e-6, bands, NO_BL, rc)
    ok(st)
    peak = obs[0, 1, 0].max()
    assert np.abs(obs[1, 1, 0]).max() > 1.0e5 * peak


def observation_layout_is_contiguous_frames(render, tol=1.0):
    """observe.rs:588-608, and every (band, scene) slice equals that frame
    rendered on its own."""
    rc = T.RenderConfig.make(x=True, lam=True)
    bands = [450e-9, 550e-9]
    scenes = [(0.0, 0.0), (20e-6, 0.0)]
    car = relay_carrier(2e-3, 1.5, 1e-6)
    src = gaussian_source(60e-6)
    st, obs = render(designs_for(car, bands, rc), inputs_for(src, bands, scenes, rc), 10e-6, 10e-6, bands, NO_BL, rc)
    ok(st)
    assert obs.shape == (3, 2, 2, 64, 64)
    for b, lam in enumerate(bands):
        for s, sc in enumerate(scenes):
            st1, one = render(designs_for(car, [lam], rc), inputs_for(src, [lam], [sc], rc), 10e-6, 10e-6, [lam],
                              NO_BL, rc)
            ok(st1)
            scale = np.abs(one).max(axis=(1, 2, 3, 4), keepdims=True)[:, 0]
            assert np.abs(obs[:, b, s] - one[:, 0, 0]).max() <= 1e-13 * tol * scale.max()


def renderer_rejects_empty_axes_and_bad_steps(render, tol=1.0):
    rc = T.RenderConfig.make()
    car = flat_relief(1e-6, 1.5)
    empty = np.zeros((0, 1, 1, 64, 64), np.complex128)
    st, _ = render([], empty, 10e-6, 10e-6, [], T.AsmConfig.default(), rc)
    assert st == T.CODE["OC_INV_RENDER"]
    rc = T.RenderConfig.make(x=True, scene_step=0.0)
    bands = [L]
    st, _ = render(designs_for(car, bands, rc), inputs_for(plane_wave_source(), bands, [(0, 0)], rc), 10e-6, 10e-6,
                   bands, T.AsmConfig.default(), rc)
    assert st == T.CODE["OC_INV_RENDER"]
    rc = T.RenderConfig.make(lam=True, lambda_step=float("nan"))
    st, _ = render(designs_for(car, bands, T.RenderConfig.make()), inputs_for(plane_wave_source(), bands, [(0, 0)], rc),
                   10e-6, 10e-6, bands, T.AsmConfig.default(), rc)
    assert st == T.CODE["OC_INV_RENDER"]


CASES = [defocus_mask_defocus_chain_is_pure_piston, plane_wave_x_y_channels_are_exactly_zero,
         dx_channel_matches_translation_invariance, lambda_channel_vanishes_for_piston_only_system,
         lambda_channel_detects_chromatic_focus, observation_layout_is_contiguous_frames,
         renderer_rejects_empty_axes_and_bad_steps]
