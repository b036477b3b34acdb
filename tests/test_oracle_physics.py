"""Pins the f64 oracle with the reference's own physics/convention tests
(restated in physics_cases.py) and the oracle-only analytic checks below."""
import math

import numpy as np
import pytest

import oc_types as T
import physics_cases as pc

ALL = pc.ASM_CASES + pc.FRESNEL_CASES + pc.RS_CASES + pc.FRAUNHOFER_CASES + pc.MASK_CASES


@pytest.mark.parametrize("case", ALL, ids=[c.__name__ for c in ALL])
def test_oracle_case(orc, case):
    case(orc)


def test_fresnel_closed_form_step_is_sweep_max(orc):
    """fresnel.rs:632-656: brute-force sweep of adjacent-bin chirp steps."""
    nxp, nyp, d, z, lam = 12, 8, 2e-6, 0.037, 500e-9
    fx, fy = orc.fftfreq(nxp, d), orc.fftfreq(nyp, d)
    ph = lambda ix, iy: -math.pi * lam * z * (fx[ix] ** 2 + fy[iy] ** 2)
    sweep = max(abs(ph(ix, iy) - ph((ix + 1) % nxp, iy)) for iy in range(nyp) for ix in range(nxp))
    assert abs(sweep - orc.fresnel_worst_axis_step(nxp, d, lam, z)) < 1e-9


def test_fresnel_transfer_matches_asm_low_freq(orc):
    """fresnel.rs:610-630: H_F vs H_ASM truncation O(kz s^2/8); DC exact."""
    z, lam = 0.1, 500e-9
    k = 2 * math.pi / lam
    s = 1e-4
    f = math.sqrt(s) / lam
    h_f = orc.fresnel_transfer_value(f, 0.0, lam, z)
    h_asm = complex(math.cos(k * z * math.sqrt(1 - s)), math.sin(k * z * math.sqrt(1 - s)))
    assert abs(h_f - h_asm) < 2 * k * z * s * s / 8 + 1e-12
    assert abs(orc.fresnel_transfer_value(0.0, 0.0, lam, z) - complex(math.cos(k * z), math.sin(k * z))) < 1e-12


def test_fresnel_convolution_delta_is_kernel(orc):
    """fresnel.rs:800-845: delta -> d^2/(i lam z) e^{ikz} chirp, |h| flat."""
    n, d, z, lam = 32, 1e-5, 1e-3, 500e-9
    u = np.zeros((n, n), np.complex128)
    u[0, 0] = 1
    st, out, _ = orc.fresnel_propagate_convolution(u, d, d, lam, z, T.FresnelConfig(1.0, 0, 0, 1.0))
    assert st == 0
    amp = d * d / (lam * z)
    k = 2 * math.pi / lam
    pref = complex(math.cos(k * z), math.sin(k * z)) / (1j * lam * z) * d * d
    assert abs(out[0, 0] - pref) < 1e-12
    for iy in [0, 11, 16]:
        for ix in [1, 5, 16, 31]:
            s = out[iy, ix]
            assert abs(abs(s) - amp) < 1e-12
            x = (ix if ix <= n // 2 else ix - n) * d
            y = (iy if iy <= n // 2 else iy - n) * d
            ch = math.pi * (x * x + y * y) / (lam * z)
            assert abs(s / out[0, 0] - complex(math.cos(ch), math.sin(ch))) < 1e-9


def test_fresnel_tf_equals_convolution_at_critical_sampling(orc):
    """fresnel.rs:878-913."""
    n, z, lam = 256, 2e-3, 500e-9
    d = math.sqrt(lam * z / n)
    u = pc.gaussian(n, n, d, 50 * d)
    cfg = T.FresnelConfig(1.0, 1, 0, 1.0)
    rep = orc.fresnel_sampling_report(u.shape, d, d, lam, z, cfg)
    assert rep.nyquist_ok and rep.kernel_nyquist_ok
    st, tf, _ = orc.fresnel_propagate(u, d, d, lam, z, cfg)
    assert st == 0
    st, cv, _ = orc.fresnel_propagate_convolution(u, d, d, lam, z, cfg)
    assert st == 0
    assert np.sqrt(np.sum(np.abs(tf - cv) ** 2) / np.sum(np.abs(tf) ** 2)) < 1e-6


def test_fresnel_paraxial_flag(orc):
    cfg = T.FresnelConfig.default()
    rep = orc.fresnel_sampling_report((32, 32), 1e-3, 1e-3, 500e-9, 0.1, cfg)
    assert rep.paraxial_ok and rep.max_phase_error_vs_asm < 0.01
    rep = orc.fresnel_sampling_report((32, 32), 1e-6, 1e-6, 500e-9, 0.05, cfg)
    assert not rep.paraxial_ok and rep.max_phase_error_vs_asm > 0.01


def test_rs_negative_z_kernel_is_conjugate(orc):
    """rayleigh_sommerfeld.rs:520-531: bit-exact conj."""
    f = orc.rs_impulse_response(32, 32, 10e-6, 10e-6, 500e-9, 2e-3)
    b = orc.rs_impulse_response(32, 32, 10e-6, 10e-6, 500e-9, -2e-3)
    assert np.array_equal(b, np.conj(f))


def test_rs_kernel_reduces_to_fresnel(orc):
    """rayleigh_sommerfeld.rs:486-517."""
    n, d, z, lam = 64, 10e-6, 5.0, 500e-9
    k = 2 * math.pi / lam
    worst = 0.0
    for iy in [0, 16, 31, 32, 48, 63]:
        for ix in [0, 16, 31, 32, 48, 63]:
            x = (ix if ix <= n // 2 else ix - n) * d
            y = (iy if iy <= n // 2 else iy - n) * d
            hf = complex(math.cos(k * z), math.sin(k * z)) / (1j * lam * z) * np.exp(
                1j * math.pi * (x * x + y * y) / (lam * z))
            worst = max(worst, abs(orc.rs_kernel_value(x, y, lam, z) / hf - 1.0))
    assert worst < 1e-6


def test_rs_kernel_step_brackets_sweep(orc):
    """rayleigh_sommerfeld.rs:594-632."""
    n, d, z, lam = 64, 2e-6, 3e-4, 500e-9
    k = 2 * math.pi / lam
    c = lambda i: (i if i <= n // 2 else i - n) * d

    def ph(x, y):
        r = math.sqrt(x * x + y * y + z * z)
        return k * r - math.atan(k * r)

    sweep = max(abs(ph(c(ix), c(iy)) - ph(c((ix + 1) % n), c(iy))) for iy in range(n) for ix in range(n))
    closed = orc.rs_worst_kernel_axis_step(n, d, lam, z)
    assert sweep <= closed + 1e-12 and closed < 2 * sweep + 1e-12


def test_sampled_phase_bilinear_lookup(orc):
    """element.rs:3300-3341."""
    vals = np.array([(k % 3) + 2.0 * (k // 3) for k in range(9)]).reshape(3, 3)
    m = orc.Mask(vals, 1e-3, 1e-3)
    assert abs(orc.sampled_phase_at(m, 0.0, 0.0) - 3.0) < 1e-12
    assert abs(orc.sampled_phase_at(m, 1e-3, 0.0) - 4.0) < 1e-12
    assert abs(orc.sampled_phase_at(m, -1e-3, 1e-3) - 4.0) < 1e-12
    assert abs(orc.sampled_phase_at(m, -0.5e-3, -0.5e-3) - 1.5) < 1e-12
    assert orc.sampled_phase_at(m, 2e-3, 0.0) == 0.0
    assert orc.sampled_phase_at(m, 0.0, -2e-3) == 0.0
    ramp = np.array([2.0 * ((k % 3) - 1.0) + 3.0 * ((k // 3) - 1.0) for k in range(9)]).reshape(3, 3)
    mr = orc.Mask(ramp, 1e-3, 1e-3)
    for px, py in [(0.3e-3, -0.7e-3), (-0.1e-3, 0.9e-3), (0.0, 0.0)]:
        assert abs(orc.sampled_phase_at(mr, px, py) - (2 * px / 1e-3 + 3 * py / 1e-3)) < 1e-9


def test_phasemask_validation_order(orc):
    assert orc.phasemask_validate(orc.Mask(np.zeros((2, 2)), 1e-3, 1e-3)) == 0
    assert orc.phasemask_validate(orc.Mask(np.zeros((2, 2)), 0.0, 1e-3)) == T.CODE["OC_INVALID_GRID_PITCH"]
    bad = np.zeros((2, 2))
    bad[1, 1] = np.inf
    assert orc.phasemask_validate(orc.Mask(bad, 1e-3, 1e-3)) == T.CODE["OC_INVALID_PHASE"]
    assert orc.phasemask_validate(orc.Mask(np.zeros((2, 2)), 1e-3, 1e-3, orientation=float("nan"))) == \
        T.CODE["OC_INVALID_ORIENTATION"]
    assert orc.phasemask_validate(orc.Mask(np.zeros((2, 2)), 1e-3, 1e-3, aperture=(T.APERTURE_CIRCLE, -1.0, 0))) \
        == T.CODE["OC_INVALID_APERTURE"]


def test_asm_band_limit_report(orc):
    """asm.rs:564-575: band limit uses the padded width and shrinks with z."""
    g = (100, 100)
    r0 = orc.asm_sampling_report(g, 10e-6, 10e-6, 500e-9, 0.0, T.AsmConfig.default())
    assert abs(r0.band_limit_fx - 1.0 / 500e-9) < 1e-6
    r1 = orc.asm_sampling_report(g, 10e-6, 10e-6, 500e-9, 0.05, T.AsmConfig.default())
    assert r1.band_limit_fx < r0.band_limit_fx and r1.nx_pad == 200
    r2 = orc.asm_sampling_report(g, 10e-6, 10e-6, 500e-9, 1.0, T.AsmConfig.default(band_limited=0))
    assert abs(r2.band_limit_fx - 1.0 / 500e-9) < 1e-9


def test_asm_medium_scales_band_limit_and_guard(orc):
    """asm.rs:706-740."""
    base = T.AsmConfig(2.0, 0, 0, 8.0, 1.0)
    dense = T.AsmConfig(2.0, 0, 0, 8.0, 1.5)
    r1 = orc.asm_sampling_report((32, 32), 1e-6, 1e-6, 500e-9, 0.5, base)
    r2 = orc.asm_sampling_report((32, 32), 1e-6, 1e-6, 500e-9, 0.5, dense)
    assert abs(r2.band_limit_fx - 1.5 * r1.band_limit_fx) < 1e-9
    ratio = r2.max_phase_step_x / r1.max_phase_step_x
    assert 0.55 <= ratio < 1.0 and r2.medium == 1.5


def test_workload_guards(orc):
    """The workload geometries land where the survey's probe says they do."""
    import workloads as W
    c1 = W.CFG1
    r = orc.asm_sampling_report((c1["n"], c1["n"]), c1["d"], c1["d"], c1["lam"], c1["z"], T.AsmConfig.default())
    assert r.nyquist_ok and r.nx_pad == 512 and abs(r.max_phase_step_x - 1.24) < 0.01
    for lam in (450e-9, 650e-9):
        r = orc.asm_sampling_report((2048, 2048), 2e-6, 2e-6, lam, 20e-3, T.AsmConfig.default())
        assert r.nyquist_ok and r.nx_pad == 4096
    r = orc.asm_sampling_report((1024, 1024), 3e-6, 3e-6, 1.55e-6, 10e-3, T.AsmConfig.default())
    assert r.nyquist_ok and r.nx_pad == 2048


@pytest.mark.parametrize("shape,z,src", [((7, 6), 2.0, None), ((8, 8), -1.5, None), ((9, 5), 3.0, (3e-5, -2e-5))])
def test_fraunhofer_oracle_equals_direct_sum(orc, shape, z, src):
    """The far field as the textbook sum (fraunhofer.rs:1-17):
    U(X, Y) = e^{ikz}/(i lam z) e^{i pi ((X-xs)^2 + (Y-ys)^2)/(lam z)}
              dx dy sum u(xi, eta) e^{-i 2 pi (fx xi + fy eta)},
    f = (X - xs)/(lam z), X = (i - c) dX, dX = lam |z| / (n dx); brute force
    over every input and output sample."""
    import math
    ny, nx = shape
    d, lam = 1e-5, 500e-9
    rng = np.random.default_rng(nx * 10 + ny)
    u = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    cfg = T.FraunhoferConfig.default(strict=False, source_pos=src)
    st, out, rep = orc.fraunhofer_propagate(u, d, d, lam, z, cfg)
    assert st == 0
    xs, ys = src if src else (0.0, 0.0)
    dxo, dyo = lam * abs(z) / (nx * d), lam * abs(z) / (ny * d)
    xi = (np.arange(nx) - (nx - 1) / 2) * d
    eta = (np.arange(ny) - (ny - 1) / 2) * d
    X = (np.arange(nx) - (nx - 1) / 2) * dxo
    Y = (np.arange(ny) - (ny - 1) / 2) * dyo
    k = 2 * math.pi / lam
    pref = np.exp(1j * k * z) / (1j * lam * z)
    want = np.empty(shape, complex)
    for iy in range(ny):
        for ix in range(nx):
            fx, fy = (X[ix] - xs) / (lam * z), (Y[iy] - ys) / (lam * z)
            ft = d * d * np.sum(u * np.exp(-2j * math.pi * (fx * xi[None, :] + fy * eta[:, None])))
            chirp = np.exp(1j * math.pi * ((X[ix] - xs) ** 2 + (Y[iy] - ys) ** 2) / (lam * z))
            want[iy, ix] = pref * chirp * ft
    e = np.linalg.norm(out - want) / np.linalg.norm(want)
    assert e < 1e-11, e
