"""Pins the oracle's Gerchberg-Saxton loop and adjoint push_back.

- GS (gs.rs:347-597): the reference's top-hat acceptance test
  (gs.rs:857-947: monotone best-so-far, merit < 0.5 x first, interior CV <
  0.35) with GsInit::Given phases (GsInit::Random's rand 0.8 stream is not
  restated), and two iterations against an independent numpy restatement of
  the loop (numpy.fft ASM, projections, wrap/encode).
- FFT-form GS: two iterations against numpy; error decreases.
- adjoint (adjoint.rs:498-544): the coarse-grid central finite-difference
  acceptance (< 1e-4 global relative error), plus a 2-mask design.
"""
import math

import numpy as np

import oc_types as T
import workloads as W


def np_asm(u, d, lam, z, padding=2.0, band_limited=False):
    """numpy ASM (pocketfft) restating asm.rs:481-535 for the pins below."""
    ny, nx = u.shape
    P = int(math.ceil(nx * padding)) if padding > 1 else nx
    Q = int(math.ceil(ny * padding)) if padding > 1 else ny
    ox, oy = (P - nx) // 2, (Q - ny) // 2
    buf = np.zeros((Q, P), complex)
    buf[oy:oy + ny, ox:ox + nx] = u
    F = np.fft.fft2(buf)
    fx = np.fft.fftfreq(P, d)
    fy = np.fft.fftfreq(Q, d)
    s = 1.0 - (lam * fx[None, :]) ** 2 - (lam * fy[:, None]) ** 2
    H = np.where(s >= 0, np.exp(1j * (2 * math.pi / lam) * z * np.sqrt(np.maximum(s, 0))), 0)
    if band_limited:
        lx = 1 / (lam * math.sqrt((2 * z / (P * d)) ** 2 + 1))
        ly = 1 / (lam * math.sqrt((2 * z / (Q * d)) ** 2 + 1))
        H = np.where((np.abs(fx[None, :]) > lx) | (np.abs(fy[:, None]) > ly), 0, H)
    return np.fft.ifft2(F * H)[oy:oy + ny, ox:ox + nx]


def wrap_pi(v):
    return np.mod(v + math.pi, 2 * math.pi) - math.pi


def tophat_problem(n=128, d=8e-6, target_r=250e-6):
    x = (np.arange(n) - (n - 1) / 2) * d
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    u = np.exp(-r2 / (150e-6) ** 2).astype(complex)
    power = np.sum(np.abs(u) ** 2)
    disk = r2 <= target_r ** 2
    amp = math.sqrt(power / disk.sum())
    target = np.where(disk, amp, 0.0)
    return u, target, r2


ASM_PLAIN = T.AsmConfig(2.0, 0, 0, 8.0, 1.0)


def test_gs_tophat_acceptance(orc):
    n, d, lam = 128, 8e-6, 500e-9
    u, target, r2 = tophat_problem(n, d)
    init = -math.pi + 2 * math.pi * W.splitmix_uniform(7, n * n).reshape(n, n)
    st, sol = orc.gs_run(u, d, d, lam, 8e-3, target, target ** 2, ASM_PLAIN, max_iterations=120, patience=10,
                         merit_tolerance=1e-4, init_phase=init)
    assert st == 0
    bm = [h[2] for h in sol["history"]]
    assert all(b1 <= b0 + 1e-15 for b0, b1 in zip(bm, bm[1:]))
    assert sol["merit"] < 0.5 * sol["history"][0][1]
    inner = r2 <= (0.7 * 250e-6) ** 2
    v = sol["target_intensity"][inner]
    assert v.std() / v.mean() < 0.35
    assert np.all(np.isfinite(sol["phase_error"])) and np.all(np.isfinite(sol["target_amplitude_error"]))


def test_gs_two_iterations_match_numpy(orc):
    n, d, lam, z = 24, 10e-6, 500e-9, 2e-3
    u, target, _ = tophat_problem(n, d, target_r=60e-6)
    init = -math.pi + 2 * math.pi * W.splitmix_uniform(3, n * n).reshape(n, n)
    mt = target ** 2 * 0.9
    st, sol = orc.gs_run(u, d, d, lam, z, target, mt, ASM_PLAIN, max_iterations=2, patience=10,
                         merit_tolerance=1e-4, init_phase=init)
    assert st == 0
    illum = np.abs(np_asm(u, d, lam, z / 2))
    phase = np.mod(init + math.pi, 2 * math.pi) - math.pi
    merits = []
    for _ in range(2):
        ut = np_asm(illum * np.exp(1j * phase), d, lam, z / 2)
        at = np.abs(ut)
        merits.append(np.mean((at ** 2 - mt) ** 2))
        proj = np.where(at > 0, ut * (target / np.where(at > 0, at, 1)), target)
        v = np_asm(proj, d, lam, -z / 2)
        dphi = wrap_pi(np.angle(v) - phase)
        phase = np.mod(phase + dphi + math.pi, 2 * math.pi) - math.pi
    assert abs(sol["history"][0][1] - merits[0]) <= 1e-9 * merits[0]
    assert abs(sol["history"][1][1] - merits[1]) <= 1e-9 * merits[1]
    assert np.allclose(sol["illumination_amplitude"], illum, rtol=1e-9, atol=1e-12)
    best = 0 if merits[0] <= merits[1] else 1
    assert sol["best_iteration"] == best


def test_gs_fft_two_iterations_match_numpy(orc):
    n = 32
    amp, tgt, ph = W.workload3(n=n, restarts=2, spots=4)
    tgt = np.fft.ifftshift(tgt)
    out, err = orc.gs_fft(amp, tgt, ph, 2)
    for r in range(2):
        p = ph[r].copy()
        for _ in range(2):
            U = np.fft.fft2(amp * np.exp(1j * p))
            m = np.abs(U)
            U = np.where(m > 0, U * (tgt / np.where(m > 0, m, 1)), tgt)
            V = np.fft.ifft2(U)
            p = np.mod(p + wrap_pi(np.angle(V) - p), 2 * math.pi)
        diff = np.abs(np.exp(1j * p) - np.exp(1j * out[r])) * amp
        assert diff.max() < 1e-9
        inten = np.abs(np.fft.fft2(amp * np.exp(1j * p))) ** 2
        t2 = tgt ** 2
        e = np.linalg.norm(inten / inten.sum() - t2 / t2.sum()) / np.linalg.norm(t2 / t2.sum())
        assert abs(e - err[r]) < 1e-9


def test_gs_fft_error_decreases(orc):
    amp, tgt, ph = W.workload3(n=64, restarts=1, spots=8)
    tgt = np.fft.ifftshift(tgt)
    _, e1 = orc.gs_fft(amp, tgt, ph, 1)
    _, e20 = orc.gs_fft(amp, tgt, ph, 20)
    assert e20[0] < 0.8 * e1[0]


def gaussian_intensity(n, d, w):
    x = (np.arange(n) - (n - 1) / 2) * d
    return np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (w * w))


def grid_gradient(orc, steps_fn, values, u, d, lam, cfg, target):
    """grid_gradient (adjoint.rs:268-340) for one band/scene over the oracle."""
    steps = steps_fn(values)
    st, det = orc.trace(steps, u, d, d, lam, cfg)
    assert st == 0
    inten = det.real * det.real + det.imag * det.imag
    dmdi = 2.0 * (inten - target) / inten.size
    st, g = orc.adjoint_push_back(steps, u, d, d, lam, cfg, dmdi)
    assert st == 0
    return g, float(np.mean((inten - target) ** 2))


def merit_of(orc, steps, u, d, lam, cfg, target):
    st, det = orc.trace(steps, u, d, d, lam, cfg)
    assert st == 0
    inten = det.real * det.real + det.imag * det.imag
    return float(np.mean((inten - target) ** 2))


def test_adjoint_matches_fd_coarse_grid(orc):
    """adjoint.rs:498-544: 6x6 DOE relay, values sin(1.3 k)*1.1."""
    n, d, z, lam = 6, 30e-6, 1.2e-3, 500e-9
    values = (np.sin(np.arange(n * n) * 1.3) * 1.1).reshape(n, n)
    target = gaussian_intensity(n, d, 60e-6)
    u = np.ones((n, n), complex)

    def steps_fn(v):
        return [(None, z, 1.0), (orc.Mask(v, d, d), z, 1.0)]

    g, _ = grid_gradient(orc, steps_fn, values, u, d, lam, ASM_PLAIN, target)
    h = 1e-6
    fd = np.zeros(n * n)
    for k in range(n * n):
        vp = values.copy().ravel()
        vm = values.copy().ravel()
        vp[k] += h
        vm[k] -= h
        fd[k] = (merit_of(orc, steps_fn(vp.reshape(n, n)), u, d, lam, ASM_PLAIN, target)
                 - merit_of(orc, steps_fn(vm.reshape(n, n)), u, d, lam, ASM_PLAIN, target)) / (2 * h)
    ga = g[0].ravel()
    rel = np.linalg.norm(fd - ga) / np.linalg.norm(fd)
    assert rel < 1e-4, rel
    big = np.abs(fd) > 1e-3 * np.linalg.norm(fd)
    assert np.all(np.sign(fd[big]) == np.sign(ga[big]))


def test_adjoint_two_masks_fd(orc):
    """Two phase masks with a circular stop, band-limited default config."""
    n, d, z, lam = 16, 8e-6, 0.6e-3, 633e-9
    p1 = 2 * math.pi * W.splitmix_uniform(11, n * n).reshape(n, n)
    p2 = 2 * math.pi * W.splitmix_uniform(12, n * n).reshape(n, n)
    target = gaussian_intensity(n, d, 30e-6)
    u = np.ones((n, n), complex)
    cfg = T.AsmConfig.default()

    def steps_fn(a, b):
        return [(orc.Mask(a, d, d, aperture=(T.APERTURE_CIRCLE, 56e-6, 0.0)), z, 1.0), (orc.Mask(b, d, d), z, 1.0)]

    st, det = orc.trace(steps_fn(p1, p2), u, d, d, lam, cfg)
    assert st == 0
    inten = np.abs(det) ** 2
    dmdi = 2.0 * (inten - target) / inten.size
    st, g = orc.adjoint_push_back(steps_fn(p1, p2), u, d, d, lam, cfg, dmdi)
    assert st == 0
    rng = np.random.default_rng(0)
    h = 1e-6
    for plane in (0, 1):
        idx = rng.choice(n * n, 12, replace=False)
        fd = []
        for k in idx:
            a, b = p1.copy().ravel(), p2.copy().ravel()
            arr = a if plane == 0 else b
            arr[k] += h
            mp = merit_of(orc, steps_fn(a.reshape(n, n), b.reshape(n, n)), u, d, lam, cfg, target)
            arr[k] -= 2 * h
            mm = merit_of(orc, steps_fn(a.reshape(n, n), b.reshape(n, n)), u, d, lam, cfg, target)
            fd.append((mp - mm) / (2 * h))
        fd = np.array(fd)
        ga = g[plane].ravel()[idx]
        assert np.linalg.norm(fd - ga) / np.linalg.norm(fd) < 1e-4
