"""The oracle's Renderer::render (observe.rs:213-314) pinned by the
reference's own renderer tests (observe.rs:413-664), restated over PhaseMask
designs in render_cases.py, and the thin-lens-as-mask design those tests
need pinned against the oracle's ThinLens."""
import math

import numpy as np
import pytest

import oc_types as T
import render_cases as rcases
import workloads as W


@pytest.mark.parametrize("case", rcases.CASES, ids=lambda f: f.__name__)
def test_render_case_oracle(orc, case):
    case(orc.render)


def test_lens_mask_equals_thinlens(orc):
    """The chromatic case's lens is a PhaseMask sampled on the field lattice:
    on every sample it equals ThinLens::apply (element.rs) with the same
    circular stop."""
    n, d = 128, 10e-6
    for lam in (450e-9, 500e-9, 550e-9):
        f = 50e-3 * 500e-9 / lam
        m = rcases.MaskSpec(rcases.lens_values(n, d, lam, f), d, d, aperture=(T.APERTURE_CIRCLE, 300e-6, 0.0))
        ones = np.ones((n, n), np.complex128)
        st, a = orc.phasemask_apply(ones, d, d, lam, m)
        assert st == 0
        st, b = orc.thinlens_apply(ones, d, d, lam, f, aperture=(T.APERTURE_CIRCLE, 300e-6, 0.0))
        assert st == 0
        assert np.abs(a - b).max() < 1e-12
        assert np.count_nonzero(a) == np.count_nonzero(W.circle_mask(n, d, 300e-6))


def test_render_frames_are_traces(orc):
    """Channel 0 of every frame is |trace|^2 of that frame's design and
    source; each FD channel is (I+ - I-)/(2 step) of the traced renders
    (observe.rs:195-210, 260-310): checked against or_trace directly on a
    two-mask design with every channel on."""
    n, d = 48, 8e-6
    rc = T.RenderConfig.make(x=True, y=True, lam=True, scene_step=2e-6, lambda_step=3e-9)
    bands = [520e-9, 610e-9]
    scenes = [(0.0, 0.0), (-16e-6, 8e-6)]
    p1 = 2 * math.pi * W.splitmix_uniform(5, n * n).reshape(n, n)

    def car(lam):
        m1 = rcases.MaskSpec(p1 * 550e-9 / lam, d, d, aperture=(T.APERTURE_CIRCLE, 150e-6, 0.0))
        m2 = rcases.MaskSpec(np.full((n, n), 1.0), d, d)
        return [(m1, 1.5e-3, 1.0), (None, -0.5e-3, 1.2), (m2, 0.0, 1.0)]

    src = rcases.gaussian_source(70e-6, n, d)
    designs = rcases.designs_for(car, bands, rc)
    inputs = rcases.inputs_for(src, bands, scenes, rc)
    cfg = T.AsmConfig.default(strict=0)
    st, obs = orc.render(designs, inputs, d, d, bands, cfg, rc)
    assert st == 0
    for b, lam in enumerate(bands):
        for s in range(len(scenes)):
            def tr(r, v, lam_r):
                st2, out = orc.trace(designs[b][v], inputs[b, s, r], d, d, lam_r, cfg)
                assert st2 == 0
                return out.real * out.real + out.imag * out.imag  # norm_sqr
            assert np.array_equal(obs[0, b, s], tr(0, 0, lam))
            assert np.array_equal(obs[1, b, s], (tr(1, 0, lam) - tr(2, 0, lam)) / (2.0 * rc.scene_step))
            assert np.array_equal(obs[2, b, s], (tr(3, 0, lam) - tr(4, 0, lam)) / (2.0 * rc.scene_step))
            want = (tr(5, 1, lam + rc.lambda_step) - tr(6, 2, lam - rc.lambda_step)) / (2.0 * rc.lambda_step)
            assert np.array_equal(obs[3, b, s], want)
