"""fd::gradient and Adam cases (crates/optics-inverse/src/fd.rs:139-244 and
its tests, fd.rs:260-480).

Two families:
- the render objective (trace a PhaseMask design with ASM, MseToTarget on
  the intensity), written against an fd callable with the call shape of
  oracle.fd_gradient; pinned against the adjoint gradient, which is
  independent of any FD machinery (adjoint.rs:140-246);
- the reference's Adam tests, whose objectives are closures (a quadratic
  bowl), run through fd::adam restated here with the Adam epoch update
  taken from an adam_step callable (oracle.adam_step or the device form).
"""
import math

import numpy as np

import oc_types as T
import render_cases as rcases
import workloads as W

# ------------------------------------------------ render objective ------

N, D, LAM, Z = 32, 8e-6, 633e-9, 0.8e-3
CFG = T.AsmConfig.default(strict=0)


def problem():
    """Two phase masks on the field lattice, a Gaussian source, a 6-spot
    intensity target (the adjoint test's design at 32^2)."""
    p1 = 2 * math.pi * W.splitmix_uniform(21, N * N).reshape(N, N)
    p2 = 2 * math.pi * W.splitmix_uniform(22, N * N).reshape(N, N)
    u = W.gaussian_beam(N, D, LAM, 70e-6)
    tgt = np.zeros((N, N))
    for k in range(6):
        tgt[8 + 3 * k, 10 + 2 * k] = 1e-3
    return p1, p2, u, tgt


def design(p1, p2):
    return [(rcases.MaskSpec(p1, D, D), Z, 1.0), (rcases.MaskSpec(p2, D, D), Z, 1.0)]


SLOTS = [(0, 5, 7), (0, 16, 16), (0, 20, 3), (1, 9, 12), (1, 15, 15), (1, 28, 30)]


def probes(p1, p2, x_lo=0.0, x_hi=2 * math.pi, slots=SLOTS, steps=None, at_bound=None):
    """Probe designs for pixel-phase slots, fd_step = 1e-4 of the span
    (params.rs:62-71), probes clamped into the bounds (fd.rs:163-166)."""
    lat = [p1, p2]
    designs, up, dn, lo, hi = [], [], [], [], []
    for j, (e, iy, ix) in enumerate(slots):
        x = lat[e][iy, ix] if at_bound is None or j not in at_bound else at_bound[j]
        h = (x_hi - x_lo) * 1e-4 if steps is None else steps[j]
        u_, d_ = min(max(x + h, x_lo), x_hi), min(max(x - h, x_lo), x_hi)
        pair = []
        for val in (u_, d_):
            q = [lat[0].copy(), lat[1].copy()]
            q[e][iy, ix] = val
            pair.append(design(q[0], q[1]))
        designs.append(tuple(pair))
        up.append(u_)
        dn.append(d_)
        lo.append(x_lo)
        hi.append(x_hi)
    return designs, up, dn, lo, hi


def adjoint_reference(orc, p1, p2, u, tgt):
    """dM/d(lattice value) of MseToTarget through the oracle's adjoint."""
    st, out = orc.trace(design(p1, p2), u, D, D, LAM, CFG)
    assert st == 0
    inten = out.real * out.real + out.imag * out.imag
    dmdi = 2.0 * (inten - tgt) / inten.size
    st, g = orc.adjoint_push_back(design(p1, p2), u, D, D, LAM, CFG, dmdi)
    assert st == 0
    return g


def fd_matches_adjoint(fd, orc, tol=1e-5):
    p1, p2, u, tgt = problem()
    designs, up, dn, lo, hi = probes(p1, p2)
    st, g, mer, ev = fd(designs, u, D, D, LAM, CFG, tgt, up, dn, lo, hi)
    assert st == 0, T.STATUS.get(st, st)
    assert ev == 2 * len(SLOTS)
    ga = adjoint_reference(orc, p1, p2, u, tgt)
    want = np.array([ga[e, iy, ix] for (e, iy, ix) in SLOTS])
    err = np.abs(g - want).max() / np.abs(want).max()
    assert err < tol, (g, want)
    return g, mer


def zero_width_slot_is_ignored(fd, orc, tol=1.0):
    """fd.rs:321-333: a slot with lo == hi is not probed and gets 0."""
    p1, p2, u, tgt = problem()
    designs, up, dn, lo, hi = probes(p1, p2, slots=SLOTS[:2])
    lo[0] = hi[0] = p1[5, 7]
    st, g, mer, ev = fd(designs, u, D, D, LAM, CFG, tgt, up, dn, lo, hi)
    assert st == 0
    assert ev == 2 and g[0] == 0.0 and g[1] != 0.0
    assert mer[0, 0] == 0.0 and mer[0, 1] == 0.0


def probe_clamped_at_bound_is_one_sided(fd, orc, tol=1.0):
    """fd.rs:309-319: at the upper bound the +h probe clamps and the quotient
    is the one-sided slope over the clamped separation; against the adjoint
    gradient at the bound it is off by O(h)."""
    p1, p2, u, tgt = problem()
    slots = [(0, 16, 16)]
    q1 = p1.copy()
    q1[16, 16] = 2 * math.pi
    h = 0.01
    designs, up, dn, lo, hi = probes(q1, p2, slots=slots, steps=[h])
    assert up[0] == 2 * math.pi and abs(dn[0] - (2 * math.pi - h)) < 1e-15
    st, g, mer, ev = fd(designs, u, D, D, LAM, CFG, tgt, up, dn, lo, hi)
    assert st == 0 and ev == 2
    assert g[0] == (mer[0, 0] - mer[0, 1]) / (up[0] - dn[0])
    ga = adjoint_reference(orc, q1, p2, u, tgt)[0, 16, 16]
    assert abs(g[0] - ga) < 0.05 * abs(ga)


def collapsed_separation_is_an_error(fd, orc, tol=1.0):
    p1, p2, u, tgt = problem()
    designs, up, dn, lo, hi = probes(p1, p2, slots=SLOTS[:1])
    dn[0] = up[0]
    st, g, mer, ev = fd(designs, u, D, D, LAM, CFG, tgt, up, dn, lo, hi)
    assert st == T.CODE["OC_INV_RENDER"]


FD_CASES = [fd_matches_adjoint, zero_width_slot_is_ignored, probe_clamped_at_bound_is_one_sided,
            collapsed_separation_is_an_error]

# ------------------------------------------------------------ Adam ------


def bowl(target):
    """fd.rs:269-281: sum (x_i - t_i)^2 + x0 x1."""
    def f(v):
        s = 0.0
        for x, t in zip(v, target):
            s += (x - t) * (x - t)
        return s + v[0] * v[1]
    return f


def fd_closure(obj, x, lo, hi, steps):
    """fd::gradient (fd.rs:139-184) for a closure objective."""
    g = np.zeros(len(x))
    ev = 0
    for i in range(len(x)):
        if lo[i] == hi[i]:
            continue
        up = x.copy()
        up[i] = min(max(x[i] + steps[i], lo[i]), hi[i])
        dn = x.copy()
        dn[i] = min(max(x[i] - steps[i], lo[i]), hi[i])
        fu, fd = obj(up), obj(dn)
        ev += 2
        g[i] = (fu - fd) / (up[i] - dn[i])
    return g, ev


def adam_solve(step, obj, x0, lo, hi, epochs, cfg, fd_steps=None):
    """fd::adam (fd.rs:193-257) with a fixed start; the epoch update is
    step(x, g, m, v, lo, hi, t, cfg)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    steps = fd_steps if fd_steps is not None else [max((h - l) * 1e-4, 0.0) or 1e-9 for l, h in zip(lo, hi)]
    x = np.clip(np.asarray(x0, np.float64).copy(), lo, hi)
    val = obj(x)
    best, best_val, hist, evals = x.copy(), val, [val], 1
    m = np.zeros(len(x))
    v = np.zeros(len(x))
    seen = [x.copy()]
    for t in range(1, epochs + 1):
        g, ev = fd_closure(lambda y: (seen.append(y.copy()), obj(y))[1], x, lo, hi, steps)
        evals += ev
        assert step(x, g, m, v, lo, hi, t, cfg) == 0
        val = obj(x)
        seen.append(x.copy())
        evals += 1
        hist.append(val)
        if val < best_val:
            best_val, best = val, x.copy()
    return best, best_val, hist, evals, seen


def adam_reduces_a_quadratic_bowl(step, tol=1.0):
    """fd.rs:341-395."""
    target = [0.3, -0.2, 0.5]
    best, val, hist, evals, _ = adam_solve(step, bowl(target), [0.0, 0.0, 0.0], [-1.0] * 3, [1.0] * 3, 200,
                                           T.AdamConfig.default())
    assert hist[0] > 0.3
    star = [8.0 / 15.0, -7.0 / 15.0, 0.5]
    f_min = sum((s - t) ** 2 for s, t in zip(star, target)) + star[0] * star[1]
    assert val < f_min + 0.01
    for x, s in zip(best, star):
        assert abs(x - s) < 0.08
    assert evals == 1 + 200 * (2 * 3 + 1)


def adam_clips_and_stays_in_bounds(step, tol=1.0):
    """fd.rs:398-441."""
    obj = lambda v: (v[0] - 0.2) * (v[0] - 0.2) * 1e6
    best, val, hist, evals, seen = adam_solve(step, obj, [0.9], [0.0], [1.0], 100,
                                              T.AdamConfig.default(lr_frac=0.5, clip_norm=1.0))
    for v in seen:
        assert 0.0 <= v[0] <= 1.0
    assert val < 1e3
    assert 0.0 < best[0] < 1.0


def adam_first_epoch_closed_form(step, tol=1.0):
    """With m = v = 0 at t = 1: m_hat = g s, v_hat = (g s)^2, so each
    coordinate moves by lr (hi - lo) g s / (|g s| + eps) before the clamp
    (s = the clip scale)."""
    rng = np.random.default_rng(3)
    n = 64
    g = rng.standard_normal(n)
    lo, hi = -np.ones(n), 2.0 * np.ones(n)
    for cfg in (T.AdamConfig.default(), T.AdamConfig.default(has_clip=0), T.AdamConfig.default(clip_norm=100.0)):
        x = rng.uniform(-0.5, 1.5, n)
        x0 = x.copy()
        m, v = np.zeros(n), np.zeros(n)
        assert step(x, g, m, v, lo, hi, 1, cfg) == 0
        norm = math.sqrt(float(np.sum(g * g)))
        s = cfg.clip_norm / norm if (cfg.has_clip and norm > cfg.clip_norm) else 1.0
        gs = g * s
        want = np.clip(x0 - cfg.lr_frac * (hi - lo) * gs / (np.abs(gs) + cfg.eps), lo, hi)
        assert np.abs(x - want).max() < 1e-14
        assert np.abs(m - (1.0 - cfg.beta1) * gs).max() < 1e-15
        assert np.abs(v - (1.0 - cfg.beta2) * gs * gs).max() < 1e-15


def adam_deterministic(step, tol=1.0):
    """fd.rs:444-480 (fixed start): two runs agree exactly; f(0, 0) = 0.4."""
    a = adam_solve(step, bowl([0.2, -0.6]), [0.0, 0.0], [-1.0] * 2, [1.0] * 2, 50, T.AdamConfig.default())
    b = adam_solve(step, bowl([0.2, -0.6]), [0.0, 0.0], [-1.0] * 2, [1.0] * 2, 50, T.AdamConfig.default())
    assert np.array_equal(a[0], b[0]) and a[2] == b[2]
    assert abs(a[2][0] - 0.4) < 1e-12


ADAM_CASES = [adam_reduces_a_quadratic_bowl, adam_clips_and_stays_in_bounds, adam_first_epoch_closed_form,
              adam_deterministic]
