"""GS and adjoint on the HIP path vs the f64 oracle, through the C ABI.

Bars: f64 (complex128) forms match the oracle to 1e-9 relative on merit
traces/gradients over a few iterations (GS is a nonlinear iteration: rounding
differences of ~1e-16 grow slowly) and 1e-10 on adjoint gradients; the
FP32-bulk FFT-form GS matches the amplitude-weighted field to 1e-4 after a
few iterations; the adjoint gradient agrees with central finite differences
to 1e-4.
"""
import math
import sys

import numpy as np
import pytest

import oc_types as T
import workloads as W

pytestmark = pytest.mark.gpu

ASM_PLAIN = T.AsmConfig(2.0, 0, 0, 8.0, 1.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def tophat(n, d, target_r):
    x = (np.arange(n) - (n - 1) / 2) * d
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    u = np.exp(-r2 / (150e-6) ** 2).astype(complex)
    disk = r2 <= target_r ** 2
    amp = math.sqrt(np.sum(np.abs(u) ** 2) / disk.sum())
    return u, np.where(disk, amp, 0.0)


@pytest.mark.parametrize("variant,beta", [(0, 1.0), (1, 1.5), (2, 0.8)])
def test_gs_run_matches_oracle(G, orc, variant, beta):
    n, d, lam, z = 48, 8e-6, 500e-9, 4e-3
    u, target = tophat(n, d, 100e-6)
    init = -math.pi + 2 * math.pi * W.splitmix_uniform(5, n * n).reshape(n, n)
    say("gs_run variant", variant)
    kw = dict(max_iterations=8, patience=3, merit_tolerance=1e-4, init_phase=init, variant=variant, beta=beta,
              aperture=(T.APERTURE_CIRCLE, 180e-6, 0.0))
    st, ref = orc.gs_run(u, d, d, lam, z, target, target ** 2, ASM_PLAIN, **kw)
    assert st == 0
    st, got = G.gs_run(u, d, d, lam, z, target, target ** 2, ASM_PLAIN, **kw)
    assert st == 0, T.STATUS[st]
    assert got["iterations"] == ref["iterations"] and got["converged"] == ref["converged"]
    assert got["best_iteration"] == ref["best_iteration"]
    for a, b in zip(got["history"], ref["history"]):
        assert a[0] == b[0]
        for x, y in zip(a[1:], b[1:]):
            assert abs(x - y) <= 1e-9 * max(abs(y), 1e-30), (a, b)
    w = ref["illumination_amplitude"]
    assert np.allclose(got["illumination_amplitude"], w, rtol=1e-12, atol=1e-15)
    dp = np.abs(np.exp(1j * got["phase"]) - np.exp(1j * ref["phase"])) * w
    assert dp.sum() / w.sum() < 1e-9
    ti = ref["target_intensity"]
    assert np.linalg.norm(got["target_intensity"] - ti) / np.linalg.norm(ti) < 1e-9


def test_gs_run_validation(G):
    n = 8
    u = np.ones((n, n), complex)
    t = np.ones((n, n))
    bad = t.copy()
    bad[2, 2] = -1
    st, _ = G.gs_run(u, 1e-5, 1e-5, 5e-7, 1e-3, bad, t, ASM_PLAIN)
    assert st == T.CODE["OC_INV_OUT_OF_BOUNDS"]
    bad[2, 2] = np.nan
    st, _ = G.gs_run(u, 1e-5, 1e-5, 5e-7, 1e-3, bad, t, ASM_PLAIN)
    assert st == T.CODE["OC_INV_NOT_FINITE"]
    st, _ = G.gs_run(u, 1e-5, 1e-5, 5e-7, -1e-3, t, t, ASM_PLAIN)
    assert st == T.CODE["OC_INV_RENDER"]
    st, _ = G.gs_run(u, 1e-5, 1e-5, 5e-7, 1e-3, t, t, ASM_PLAIN, variant=1, beta=2.5)
    assert st == T.CODE["OC_INV_RENDER"]


@pytest.mark.parametrize("chunk", [0, 3])
def test_gs_fft_matches_oracle(G, orc, chunk):
    import torch
    n, R, its = 64, 3, 5
    amp, tgt, ph = W.workload3(n=n, restarts=R, spots=6)
    tgt = np.fft.ifftshift(tgt)
    say("gs_fft chunk", chunk)
    ref, err_ref = orc.gs_fft(amp, tgt, ph, its)
    # complex128 form
    p64 = torch.as_tensor(ph.copy()).cuda()
    st, err = G.gs_fft_run_dev(torch.as_tensor(amp).cuda(), torch.as_tensor(tgt).cuda(), p64, its, chunk)
    assert st == 0, T.STATUS[st]
    got = p64.cpu().numpy()
    for r in range(R):
        w = amp
        d = (np.abs(np.exp(1j * got[r]) - np.exp(1j * ref[r])) * w).sum() / w.sum()
        say("  c128 restart", r, d, err[r], err_ref[r])
        assert d < 1e-9
        assert abs(err[r] - err_ref[r]) < 1e-9
    # FP32-bulk form
    p32 = torch.as_tensor(ph.astype(np.float32)).cuda()
    st, err32 = G.gs_fft_run_dev(torch.as_tensor(amp.astype(np.float32)).cuda(),
                                 torch.as_tensor(tgt.astype(np.float32)).cuda(), p32, its, chunk)
    assert st == 0
    got32 = p32.cpu().numpy().astype(np.float64)
    for r in range(R):
        f1 = amp * np.exp(1j * got32[r])
        f0 = amp * np.exp(1j * ref[r])
        e = np.linalg.norm(f1 - f0) / np.linalg.norm(f0)
        say("  c64 restart", r, e, err32[r])
        assert e < 1e-4
        assert abs(err32[r] - err_ref[r]) < 1e-4


def test_gs_fft_graph_equals_eager(G):
    import torch
    n, R, its = 128, 4, 6
    amp, tgt, ph = W.workload3(n=n, restarts=R, spots=8)
    tgt = np.fft.ifftshift(tgt)
    a = torch.as_tensor(amp.astype(np.float32)).cuda()
    t = torch.as_tensor(tgt.astype(np.float32)).cuda()
    p1 = torch.as_tensor(ph.astype(np.float32)).cuda()
    p2 = p1.clone()
    st, e1 = G.gs_fft_run_dev(a, t, p1, its, 0)
    assert st == 0
    st, e2 = G.gs_fft_run_dev(a, t, p2, its, 2)  # three replays of a 2-iteration graph
    assert st == 0
    assert torch.equal(p1, p2)
    # the error sums are wave-reduced f64 atomics: order-dependent in the last bits
    assert np.allclose(e1, e2, rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("dbl", [False, True])
def test_gs_fft_fused_rows_equal_three_pass(G, dbl):
    """Power-of-two grids run each iteration as the column pass plus one row
    pass that fuses the inverse rows, the phase update and the next forward
    rows; the generic kernels run the three passes apart.  Same iteration,
    different FFT kernels: the phases agree to rounding."""
    import torch
    n, R, its = 256, 2, 7
    amp, tgt, ph = W.workload3(n=n, restarts=R, spots=8)
    tgt = np.fft.ifftshift(tgt)
    dt = np.float64 if dbl else np.float32
    a = torch.as_tensor(amp.astype(dt)).cuda()
    t = torch.as_tensor(tgt.astype(dt)).cuda()
    fused = torch.as_tensor(ph.astype(dt)).cuda()
    split = fused.clone()
    say("gs_fft fused vs three-pass, double", dbl)
    st, e1 = G.gs_fft_run_dev(a, t, fused, its, 3)
    assert st == 0
    gen = G.Context(0)
    gen.force_generic(True)
    st, e2 = G.gs_fft_run_dev(a, t, split, its, 3, ctx=gen)
    assert st == 0
    p1 = fused.cpu().numpy().astype(np.float64)
    p2 = split.cpu().numpy().astype(np.float64)
    tol = 1e-9 if dbl else 1e-4
    for r in range(R):
        f1 = amp * np.exp(1j * p1[r])
        f2 = amp * np.exp(1j * p2[r])
        e = np.linalg.norm(f1 - f2) / np.linalg.norm(f2)
        say("  restart", r, e, e1[r], e2[r])
        assert e < tol
        assert abs(e1[r] - e2[r]) < tol


def design(G, orc, n, d, z, seeds=(11, 12)):
    p1 = 2 * math.pi * W.splitmix_uniform(seeds[0], n * n).reshape(n, n)
    p2 = 2 * math.pi * W.splitmix_uniform(seeds[1], n * n).reshape(n, n)
    return p1, p2


def test_adjoint_push_back_matches_oracle(G, orc):
    n, d, z, lam = 40, 6e-6, 0.8e-3, 633e-9
    p1, p2 = design(G, orc, n, d, z)
    u = W.gaussian_beam(n, d, lam, 80e-6) * np.exp(0.2j)
    rng = np.random.default_rng(1)
    dmdi = rng.standard_normal((n, n))
    cfg = T.AsmConfig.default()
    ap = (T.APERTURE_CIRCLE, 100e-6, 0.0)
    # element 2 has no propagation after it (the z == 0 element kernel); a
    # last element after it keeps its gradient away from zero (a detector
    # right behind a phase-only element with a real dmdi has dphi == 0)
    steps_o = [(None, z / 2, 1.0), (orc.Mask(p1, d, d, aperture=ap), z, 1.0), (orc.Mask(p2, d, d), z, 1.3),
               (orc.Mask(p1, d, d, orientation=0.4), 0.0, 1.0), (orc.Mask(p2, d, d, orientation=-0.3), z / 3, 1.0)]
    steps_g = [(None, z / 2, 1.0), (G.Mask(p1, d, d, aperture=ap), z, 1.0), (G.Mask(p2, d, d), z, 1.3),
               (G.Mask(p1, d, d, orientation=0.4), 0.0, 1.0), (G.Mask(p2, d, d, orientation=-0.3), z / 3, 1.0)]
    say("push_back")
    st, ref = orc.adjoint_push_back(steps_o, u, d, d, lam, cfg, dmdi)
    assert st == 0
    st, got = G.adjoint_push_back(steps_g, u, d, d, lam, cfg, dmdi)
    assert st == 0, T.STATUS[st]
    for e in range(4):
        rel = np.linalg.norm(got[e] - ref[e]) / np.linalg.norm(ref[e])
        say("  element", e, rel)
        assert rel < 1e-10


def test_grid_gradient_dev_matches_fd_and_oracle(G, orc):
    """Workload-4 geometry at 128^2: plane wave, DOE1 -> 10 mm -> DOE2 -> 10 mm."""
    import torch
    n = 128
    c = W.CFG4
    d, lam, z = c["d"], c["lam"], c["z"]
    u, p1, p2, tgt = W.workload4(n=n)
    cfg = T.AsmConfig.default()
    m1, k1 = G.dev_mask(p1, d, d)
    m2, k2 = G.dev_mask(p2, d, d)
    ud = torch.as_tensor(u).cuda()
    td = torch.as_tensor(tgt).cuda()
    grads = torch.zeros((2, n, n), dtype=torch.float64, device="cuda")
    say("grid_gradient 128")
    st, merit = G.adjoint_grid_gradient_dev([(m1, z, 1.0), (m2, z, 1.0)], ud, d, d, lam, cfg, td, grads)
    assert st == 0, T.STATUS[st]
    g = grads.cpu().numpy()
    # oracle: trace + MseToTarget seed + push_back
    so = [(orc.Mask(p1, d, d), z, 1.0), (orc.Mask(p2, d, d), z, 1.0)]
    st, det = orc.trace(so, u, d, d, lam, cfg)
    inten = det.real ** 2 + det.imag ** 2
    assert abs(merit - np.mean((inten - tgt) ** 2)) < 1e-10 * merit
    st, ref = orc.adjoint_push_back(so, u, d, d, lam, cfg, 2.0 * (inten - tgt) / inten.size)
    for e in range(2):
        assert np.linalg.norm(g[e] - ref[e]) / np.linalg.norm(ref[e]) < 1e-10
    # central finite differences, h = 1e-4 rad, on seeded pixels of both planes
    h = c["fd_step"]
    px1, px2 = W.workload4_fd_pixels(n=n, count=8)
    fd, ad = [], []
    for plane, pix in ((0, px1), (1, px2)):
        vals = (p1 if plane == 0 else p2).copy()
        for k in pix:
            iy, ix = divmod(int(k), n)
            mp = []
            for sgn in (1, -1):
                vv = vals.copy()
                vv[iy, ix] += sgn * h
                mm, kk = G.dev_mask(vv, d, d)
                st_ = [(mm, z, 1.0), (m2, z, 1.0)] if plane == 0 else [(m1, z, 1.0), (mm, z, 1.0)]
                st, mval = G.adjoint_grid_gradient_dev(st_, ud, d, d, lam, cfg, td, None)
                assert st == 0
                mp.append(mval)
            fd.append((mp[0] - mp[1]) / (2 * h))
            ad.append(g[plane, iy, ix])
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    say("  fd rel", rel)
    assert rel < 1e-4


def test_adjoint_asm_entry(G, orc):
    """oc_adjoint_asm (adjoint.rs:90-92): propagate by -z with the same
    config, bit for bit the drop-in at -z; the inner-product identity
    <w, A(z) d> = <A(z)^dagger w, d> (adjoint.rs:406-427) to 1e-12 with the
    band limit on and off; a failing propagation maps to InverseError::Adjoint."""
    n, d, lam, z = 48, 4e-6, 633e-9, 3e-3
    rng = np.random.default_rng(4)
    dd = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    w = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    for bl in (1, 0):
        cfg = T.AsmConfig(2.0, bl, 0, 2.0, 1.0)
        st, back, rep = G.adjoint_asm(w, d, d, lam, z, cfg)
        assert st == 0, T.STATUS[st]
        st, ref, rref = G.asm_propagate(w, d, d, lam, -z, cfg)
        assert st == 0
        assert np.array_equal(back, ref)
        assert T.report_dict(rep) == T.report_dict(rref)
        st, fwd, _ = G.asm_propagate(dd, d, d, lam, z, cfg)
        assert st == 0
        lhs = np.vdot(w, fwd)
        rhs = np.vdot(back, dd)
        say("inner product, band limit", bl, abs(lhs - rhs) / abs(lhs))
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
        st, o_ref, _ = orc.asm_propagate(w, d, d, lam, -z, cfg)
        assert np.linalg.norm(back - o_ref) / np.linalg.norm(o_ref) <= 1e-12
    st, _, _ = G.adjoint_asm(w, d, d, lam, float("nan"), T.AsmConfig.default())
    assert st == T.CODE["OC_INV_ADJOINT"]


def test_adjoint_transmission_entry(G, orc):
    """oc_adjoint_transmission (adjoint.rs:102-108): apply(1), bit for bit the
    host phase-mask apply of a unit field and e^{i phi} inside the stop (0
    outside) against the oracle to 1e-12; a stop beyond the grid is
    InverseError::Adjoint."""
    n, d, lam = 40, 5e-6, 550e-9
    vals = 2 * math.pi * W.splitmix_uniform(8, n * n).reshape(n, n)
    ap = (T.APERTURE_CIRCLE, 80e-6, 0.0)
    gm = G.Mask(vals, d, d, orientation=0.3, aperture=ap, center=(5e-6, 0.0))
    st, t = G.adjoint_transmission(gm, (n, n), d, d, lam)
    assert st == 0, T.STATUS[st]
    st, t2 = G.phasemask_apply(np.ones((n, n), complex), d, d, lam, gm)
    assert st == 0
    assert np.array_equal(t, t2)
    st, tr = orc.phasemask_apply(np.ones((n, n), complex), d, d, lam,
                                 orc.Mask(vals, d, d, orientation=0.3, aperture=ap, center=(5e-6, 0.0)))
    assert st == 0
    assert np.abs(t - tr).max() <= 1e-12
    assert np.count_nonzero(t) == np.count_nonzero(tr) > 0
    big = G.Mask(vals, d, d, aperture=(T.APERTURE_CIRCLE, 150e-6, 0.0))
    st, _ = G.adjoint_transmission(big, (n, n), d, d, lam)
    assert st == T.CODE["OC_INV_ADJOINT"]
