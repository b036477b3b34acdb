"""oc_gs_run_batch (gs::run for several restarts at once, graph-captured
chunks, device-side bookkeeping) against oc_gs_run per restart and the
oracle's gs::run (crates/optics-inverse/src/gs.rs:347-597).

The restarts share a grid but differ in initial phase, variant, merit
tolerance and iteration cap: one converges at iteration 4, one at 9 and one
runs to its cap of 9, so chunks of 3 and 5 iterations put the stops both
inside a chunk (rolled back and replayed alone) and on a chunk boundary.  Each must equal
oc_gs_run bit for bit (history, phase, target intensity, error maps: the GS
sums are reduced in a fixed order, so neither batching nor chunking changes
a bit) and the oracle to the GS bar of 1e-9.
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


def problems(n=48, d=8e-6, lam=500e-9, z=4e-3):
    x = (np.arange(n) - (n - 1) / 2) * d
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    u = np.exp(-r2 / (150e-6) ** 2).astype(complex)
    disk = r2 <= (100e-6) ** 2
    target = np.where(disk, math.sqrt(np.sum(np.abs(u) ** 2) / disk.sum()), 0.0)
    out = []
    for k, (tol, cap, variant, beta) in enumerate(((0.5, 13, 0, 1.0), (0.2, 13, 1, 1.5), (0.0, 9, 2, 0.8))):
        init = -math.pi + 2 * math.pi * W.splitmix_uniform(50 + k, n * n).reshape(n, n)
        out.append(dict(u_in=u, dx=d, dy=d, lam=lam, z_total=z, target_amplitude=target, merit_target=target ** 2,
                        asm_cfg=ASM_PLAIN, max_iterations=cap, patience=2, merit_tolerance=tol, init_phase=init,
                        variant=variant, beta=beta, aperture=(T.APERTURE_CIRCLE, 180e-6, 0.0)))
    return out


def call_single(mod, p):
    kw = {k: v for k, v in p.items() if k not in ("u_in", "dx", "dy", "lam", "z_total", "target_amplitude",
                                                   "merit_target", "asm_cfg")}
    return mod.gs_run(p["u_in"], p["dx"], p["dy"], p["lam"], p["z_total"], p["target_amplitude"], p["merit_target"],
                      p["asm_cfg"], **kw)


@pytest.mark.parametrize("chunk", [0, 3, 5])
def test_gs_batch_equals_single_runs(G, orc, chunk):
    ps = problems()
    say("gs batch chunk", chunk)
    st, got = G.gs_run_batch(ps, graph_chunk=chunk)
    assert st == 0, T.STATUS[st]
    stops = []
    for k, p in enumerate(ps):
        st, one = call_single(G, p)
        assert st == 0
        g = got[k]
        say("  restart", k, "iterations", g["iterations"], "converged", g["converged"], "best", g["best_iteration"])
        stops.append((g["iterations"], g["converged"]))
        for key in ("iterations", "converged", "best_iteration", "merit", "amplitude_mse", "history"):
            assert g[key] == one[key], (k, key)
        for key in ("phase", "target_intensity", "target_amplitude_error", "phase_error", "illumination_amplitude"):
            assert np.array_equal(g[key], one[key]), (k, key)
        st, ref = call_single(orc, p)
        assert st == 0
        assert g["iterations"] == ref["iterations"] and g["converged"] == ref["converged"]
        assert g["best_iteration"] == ref["best_iteration"]
        for a, b in zip(g["history"], ref["history"]):
            assert a[0] == b[0]
            for x, y in zip(a[1:], b[1:]):
                assert abs(x - y) <= 1e-9 * max(abs(y), 1e-30), (k, a, b)
    # restart 0 converges at iteration 4 (inside a chunk of 3 or 5), restart 1
    # at 9 (a chunk boundary for 3, inside one for 5), restart 2 at its cap
    assert stops == [(4, True), (9, True), (9, False)]


def test_gs_batch_validation(G):
    ps = problems()
    ps[1] = dict(ps[1], target_amplitude=-ps[1]["target_amplitude"] - 1.0)
    st, _ = G.gs_run_batch(ps, graph_chunk=3)
    assert st == T.CODE["OC_INV_OUT_OF_BOUNDS"]
    ps = problems()
    small = problems(n=32)[0]
    st, _ = G.gs_run_batch([ps[0], small], graph_chunk=3)
    assert st == T.CODE["OC_ERR_INVALID_ARGUMENT"]
