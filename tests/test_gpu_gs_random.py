"""GsInit::Random on the HIP path (crates/optics-inverse/src/gs.rs:449-458).

The starting phases are drawn on the host by the library's own restatement
of rand 0.8's StdRng; the oracle draws them with an independent one.
- oc_gs_run with a seed equals the oracle's gs::run with the same seed
  (history to 1e-9 relative over 10 iterations, the same best iteration),
  which holds only if both draw the same starting phases.
- The reference's three random-start tests, on the HIP path: the top-hat
  shaper with seed 7 (gs.rs:858-947: monotone best-so-far, merit below half
  the first, interior CV < 0.35), DVR with seed 11 (gs.rs:950-1002) and FPR
  with seed 3 (gs.rs:1005-1049: the patience stop fires).
- oc_gs_run_batch with a seed per restart equals one oc_gs_run per seed,
  bit for bit.
"""
import math
import sys

import numpy as np
import pytest

import oc_types as T
import workloads as W

pytestmark = pytest.mark.gpu

ASM_PLAIN = T.AsmConfig(2.0, 0, 0, 8.0, 1.0)
LAM = 500e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


def tophat(n, d, w0, target_r, scale_to_power=True):
    x = (np.arange(n) - (n - 1) / 2) * d
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    u = W.gaussian_beam(n, d, LAM, w0).astype(complex)
    disk = r2 <= target_r ** 2
    amp = math.sqrt(np.sum(np.abs(u) ** 2) / disk.sum()) if scale_to_power else 1.0
    return u, np.where(disk, amp, 0.0), r2


@pytest.mark.parametrize("variant,beta,seed", [(0, 1.0, 7), (1, 0.7, 11), (2, 0.8, 3)])
def test_random_start_matches_oracle(G, orc, variant, beta, seed):
    n, d, z = 48, 8e-6, 4e-3
    u, target, _ = tophat(n, d, 150e-6, 100e-6)
    kw = dict(max_iterations=10, patience=10, merit_tolerance=1e-4, variant=variant, beta=beta, init_seed=seed)
    say("gs random start vs oracle", variant, seed)
    st, ref = orc.gs_run(u, d, d, LAM, z, target, target ** 2, ASM_PLAIN, **kw)
    assert st == 0
    st, got = G.gs_run(u, d, d, LAM, z, target, target ** 2, ASM_PLAIN, **kw)
    assert st == 0, T.STATUS[st]
    assert got["iterations"] == ref["iterations"] and got["best_iteration"] == ref["best_iteration"]
    for a, b in zip(got["history"], ref["history"]):
        assert a[0] == b[0]
        for x, y in zip(a[1:], b[1:]):
            assert abs(x - y) <= 1e-9 * max(abs(y), 1e-30), (a, b)
    # a different seed starts elsewhere
    st, other = G.gs_run(u, d, d, LAM, z, target, target ** 2, ASM_PLAIN, **dict(kw, init_seed=seed + 1))
    assert st == 0 and other["history"][0][1] != got["history"][0][1]


def test_tophat_shaper_random_start(G):
    n, d = 128, 8e-6
    u, target, r2 = tophat(n, d, 150e-6, 250e-6)
    say("gs top-hat shaper, seed 7")
    st, sol = G.gs_run(u, d, d, LAM, 8e-3, target, target ** 2, ASM_PLAIN, max_iterations=120, patience=10,
                       merit_tolerance=1e-4, variant=0, init_seed=7)
    assert st == 0, T.STATUS[st]
    bm = [h[2] for h in sol["history"]]
    assert all(b1 <= b0 + 1e-15 for b0, b1 in zip(bm, bm[1:]))
    assert sol["merit"] < 0.5 * sol["history"][0][1]
    v = sol["target_intensity"][r2 <= (0.7 * 250e-6) ** 2]
    assert v.std() / v.mean() < 0.35
    assert all(math.isfinite(h[4]) and math.isfinite(h[5]) for h in sol["history"])
    assert np.all(sol["phase"] >= -math.pi) and np.all(sol["phase"] <= math.pi)


def test_dvr_random_start(G):
    n, d = 96, 8e-6
    u, target, _ = tophat(n, d, 140e-6, 220e-6)
    say("gs DVR, seed 11")
    st, sol = G.gs_run(u, d, d, LAM, 8e-3, target, target ** 2, ASM_PLAIN, max_iterations=150, patience=10,
                       merit_tolerance=1e-4, variant=1, beta=0.7, init_seed=11)
    assert st == 0, T.STATUS[st]
    bm = [h[2] for h in sol["history"]]
    assert all(b1 <= b0 + 1e-15 for b0, b1 in zip(bm, bm[1:]))
    assert sol["merit"] < 0.5 * sol["history"][0][1]


def test_fpr_random_start_stops(G):
    n, d = 64, 10e-6
    u, target, _ = tophat(n, d, 120e-6, 180e-6, scale_to_power=False)
    say("gs FPR, seed 3")
    st, sol = G.gs_run(u, d, d, LAM, 6e-3, target, target ** 2, ASM_PLAIN, max_iterations=120, patience=5,
                       merit_tolerance=1e-3, variant=2, beta=0.8, init_seed=3)
    assert st == 0, T.STATUS[st]
    assert sol["converged"] and sol["iterations"] <= 120


def test_batch_random_starts_equal_single_calls(G):
    n, d, z = 48, 8e-6, 4e-3
    u, target, _ = tophat(n, d, 150e-6, 100e-6)
    base = dict(u_in=u, dx=d, dy=d, lam=LAM, z_total=z, target_amplitude=target, merit_target=target ** 2,
                asm_cfg=ASM_PLAIN, max_iterations=12, patience=4, merit_tolerance=1e-4)
    seeds = [5, 6, 7]
    say("gs batch of random starts", seeds)
    st, res = G.gs_run_batch([dict(base, init_seed=s) for s in seeds], graph_chunk=4)
    assert st == 0, T.STATUS[st]
    for s, r in zip(seeds, res):
        kw = {k: v for k, v in base.items() if k not in ("u_in", "dx", "dy", "lam", "z_total", "target_amplitude",
                                                        "merit_target", "asm_cfg")}
        st, one = G.gs_run(u, d, d, LAM, z, target, target ** 2, ASM_PLAIN, init_seed=s, **kw)
        assert st == 0
        assert r["iterations"] == one["iterations"] and r["best_iteration"] == one["best_iteration"]
        assert r["history"] == one["history"]
        assert np.array_equal(r["phase"], one["phase"])


@pytest.mark.parametrize("lo,hi", [(-0.5, 2.0), (0.0, 10.0)], ids=["span2.5", "span-clamped"])
def test_random_start_bounds_match_oracle(G, orc, lo, hi):
    """Slot bounds other than [-pi, pi]: the span is (hi - lo) clamped to
    [0, 2 pi] (gs.rs:451) and the start is encoded into the bounds."""
    n, d, z = 40, 8e-6, 4e-3
    u, target, _ = tophat(n, d, 150e-6, 100e-6)
    kw = dict(max_iterations=6, patience=10, merit_tolerance=1e-4, lo=lo, hi=hi, init_seed=21)
    say("gs random start with bounds", lo, hi)
    st, ref = orc.gs_run(u, d, d, LAM, z, target, target ** 2, ASM_PLAIN, **kw)
    assert st == 0
    st, got = G.gs_run(u, d, d, LAM, z, target, target ** 2, ASM_PLAIN, **kw)
    assert st == 0, T.STATUS[st]
    assert got["iterations"] == ref["iterations"] and got["best_iteration"] == ref["best_iteration"]
    for a, b in zip(got["history"], ref["history"]):
        for x, y in zip(a[1:], b[1:]):
            assert abs(x - y) <= 1e-9 * max(abs(y), 1e-30), (a, b)
    assert np.all(got["phase"] >= lo) and np.all(got["phase"] <= hi)
