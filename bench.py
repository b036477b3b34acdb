#!/usr/bin/env python3
"""Throughput benchmark of the propagation hot path on one MI355X.

Workload (the polychromatic BandField case): 16 bands, lambda 450-650 nm,
2048^2 at 2 um pitch, Gaussian beam w0 = 1 mm, 8-level DOE phase mask in 4x4
pixel blocks, then ASM z = 20 mm (padding 2 -> P = 4096, band-limited,
strict) and Fresnel transfer-function z = 200 mm (padding 2, strict off).
One step = one pass over the whole batch in the FP32-bulk form (complex64
field, FP64 phase arguments and twiddles): 16 masked ASM propagations and 16
Fresnel propagations of their results, as one chained call
(oc_propagate_chain_dev_c64: every problem of a pass in one persistent
launch; the ASM's inverse row pass and the Fresnel's row pass fused, the ASM
field kept on chip).  value = propagations/s (32 per step), inputs resident
in HBM.  The same steps queued without waiting
(oc_propagate_chain_submit_dev_c64, one oc_context_wait), the chain with the
ASM field written out, and the two stages as two batch calls are reported
beside it.

Also reported: per-pass kernel times from HIP events on the library stream
(roofline of the dominant kernel), the complex128 form, the f64 host drop-in
(PCIe included), the FP32 form's relative L2 error against the f64 oracle,
and the oracle itself timed on the host (cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "optics-studio_amd"))

METRIC = "propagations/s @2048² FP32-vs-f64 (rel L2≤1e-5), GS iter/s, adjoint grad/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--bands", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the f64 and host drop-in legs")
    ap.add_argument("--only", choices=["gs", "gs_asm", "frames", "adjoint", "rs", "render", "fd"],
                    help="run only that secondary workload and print its record (for profiling)")
    ap.add_argument("--autotune", action="store_true",
                    help="time every workgroup layout at P = 4096 first and run the fastest (default: the layouts "
                         "measured fastest on MI355X, so that every run, profiled or not, runs the same kernels)")
    ap.add_argument("--layout", type=str, default=None,
                    help="V or V,C,R: run layout variant V (and the separable Fresnel's column/row layouts C,R)")
    ap.add_argument("--schedule-cache", type=str, default=None,
                    help="the RS batch loads its schedules from this file when it exists, else tunes and saves "
                         "them there (default: tune, keep nothing)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 1 GiB (8 launches) for PMC counter calibration")
    return ap.parse_args()


def copy_calibration():
    """A plain copy kernel of known size: under rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE its per-dispatch counters give the KiB each counter reports
    per byte actually moved."""
    import torch
    nbytes = 1 << 30
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(8):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                      "launches": 8}), flush=True)


def bench_gs(G, W, torch, np):
    """Workload 3: FFT-form GS at 1024^2, complex64, 32 restarts x 200
    iterations in one batched launch sequence, hipGraph-captured in chunks."""
    c = W.CFG3
    amp, tgt, ph = W.workload3()
    tgt = np.fft.ifftshift(tgt)  # Fourier plane in DFT order
    a = torch.as_tensor(amp.astype(np.float32)).cuda()
    t = torch.as_tensor(tgt.astype(np.float32)).cuda()
    p0 = torch.as_tensor(ph.astype(np.float32)).cuda()
    chunk = 20
    # the default layout at 1024 (variant 1) is the fastest for GS: 103k
    # restart-iterations/s against 98k, 95k and 86k for variants 3, 2, 0
    variant = G.context().fft_variant(c["n"])
    st, _ = G.gs_fft_run_dev(a, t, p0.clone(), chunk, chunk)  # warm-up (plans, graph)
    if st:
        raise RuntimeError(f"gs warm-up failed: {st}")
    p = p0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, err = G.gs_fft_run_dev(a, t, p, c["iterations"], chunk)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if st:
        raise RuntimeError(f"gs failed: {st}")
    its = c["restarts"] * c["iterations"]
    r = {"workload": f"GS FFT form {c['n']}^2 complex64, {c['restarts']} restarts x {c['iterations']} iterations, "
                     f"graph chunk {chunk}",
         "iterations_per_s": its / (t1 - t0), "seconds": t1 - t0, "fft_variant": variant,
         "final_error_mean": float(np.mean(err)), "final_error_best": float(np.min(err))}
    log(f"gs: {r['iterations_per_s']:.0f} restart-iterations/s, error mean {r['final_error_mean']:.4f} "
        f"best {r['final_error_best']:.4f}")
    return r


def bench_gs_asm(G, W, T, np):
    """The ASM form of GS (gs::run with the MseToTarget merit, complex128):
    512^2 at 8 um, 633 nm, z = 20 mm (each half padded 2x to 1024^2), a
    Gaussian source and a flat-top disk target, 8 restarts x 200 iterations
    run to their cap.  Timed as one oc_gs_run_batch call (graph chunks of 20)
    and as 8 oc_gs_run calls; both calls include the host uploads, the
    illumination propagation and the outputs' read-back."""
    n, d, lam, z = 512, 8e-6, 633e-9, 20e-3
    restarts, iters, chunk = 8, 200, 20
    u = W.gaussian_beam(n, d, lam, 1.2e-3)
    x = W.coords(n, d)
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    disk = r2 <= (1.0e-3) ** 2
    target = np.where(disk, np.sqrt(np.sum(np.abs(u) ** 2) / disk.sum()), 0.0)
    cfg = T.AsmConfig(2.0, 0, 0, 8.0, 1.0)
    probs = []
    for k in range(restarts):
        init = -np.pi + 2 * np.pi * W.splitmix_uniform(900 + k, n * n).reshape(n, n)
        probs.append(dict(u_in=u, dx=d, dy=d, lam=lam, z_total=z, target_amplitude=target,
                          merit_target=target ** 2, asm_cfg=cfg, max_iterations=iters, patience=5,
                          merit_tolerance=0.0, init_phase=init, aperture=(T.APERTURE_CIRCLE, 1.8e-3, 0.0)))
    st, _ = G.gs_run_batch(probs[:2], graph_chunk=chunk)  # warm-up (plans, tables, graph)
    if st:
        raise RuntimeError(f"gs asm warm-up failed: {st}")
    t0 = time.perf_counter()
    st, sols = G.gs_run_batch(probs, graph_chunk=chunk)
    t1 = time.perf_counter()
    if st:
        raise RuntimeError(f"gs asm batch failed: {st}")
    single = [k for k in ("u_in", "dx", "dy", "lam", "z_total", "target_amplitude", "merit_target", "asm_cfg")]
    s0 = time.perf_counter()
    for p in probs:
        st, _ = G.gs_run(*[p[k] for k in single], **{k: v for k, v in p.items() if k not in single})
        if st:
            raise RuntimeError(f"gs asm single failed: {st}")
    s1 = time.perf_counter()
    its = restarts * iters
    r = {"workload": f"GS ASM form (gs::run) {n}^2 complex128, P = {2 * n}, {restarts} restarts x {iters} "
                     f"iterations, graph chunk {chunk}",
         "iterations_per_s": its / (t1 - t0), "seconds": t1 - t0,
         "single_calls_iterations_per_s": its / (s1 - s0),
         "merit_best": min(s["merit"] for s in sols), "merit_mean": float(np.mean([s["merit"] for s in sols]))}
    log(f"gs asm: {r['iterations_per_s']:.0f} restart-iterations/s batched, "
        f"{r['single_calls_iterations_per_s']:.0f}/s as single calls")
    return r


def bench_adjoint(G, W, T, torch, np, cpu=True):
    """Workload 4: adjoint gradient of the intensity MSE w.r.t. both DOEs of
    DOE1 -> ASM 10 mm -> DOE2 -> ASM 10 mm at 1024^2 (f64, per-plane
    checkpoints), against central finite differences at h = 1e-4 rad on 256
    seeded pixels (also central at 1e-3 and 2e-3 rad, Richardson-
    extrapolated) and, with the CPU leg, against the f64 oracle adjoint."""
    c = W.CFG4
    n, d, lam, z = c["n"], c["d"], c["lam"], c["z"]
    u, p1, p2, tgt = W.workload4()
    cfg = T.AsmConfig.default()
    m1, k1 = G.dev_mask(p1, d, d)
    m2, k2 = G.dev_mask(p2, d, d)
    ud = torch.as_tensor(u).cuda()
    td = torch.as_tensor(tgt).cuda()
    grads = torch.zeros((2, n, n), dtype=torch.float64, device="cuda")
    steps = [(m1, z, 1.0), (m2, z, 1.0)]
    st, merit = G.adjoint_grid_gradient_dev(steps, ud, d, d, lam, cfg, td, grads)
    if st:
        raise RuntimeError(f"adjoint failed: {st}")
    # the timed calls go straight to the C ABI with arguments marshalled
    # once (what a compiled caller does); the Python wrapper's per-call
    # marshalling and torch synchronisation are timed beside it
    call = G.adjoint_grid_gradient_dev_prepared(steps, ud, d, d, lam, cfg, td, grads)
    k = 50
    for _ in range(5):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        st_, _m = call()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if st_:
        raise RuntimeError(f"adjoint failed: {st_}")
    w0 = time.perf_counter()
    for _ in range(k):
        G.adjoint_grid_gradient_dev(steps, ud, d, d, lam, cfg, td, grads)
    torch.cuda.synchronize()
    wrapped_s = (time.perf_counter() - w0) / k
    g = grads.cpu().numpy()
    # finite differences on 2 x 128 seeded pixels through the batched FD
    # gradient (fixed-order merit reductions): central at h0 = 1e-4 rad, and
    # central at h = 10 h0 and 2h combined by Richardson extrapolation
    # ((4 g_h - g_2h) / 3 cancels the h^2 truncation term)
    h0 = c["fd_step"]
    h = 10 * h0
    px1, px2 = W.workload4_fd_pixels()
    fd, ad, fd_h, fd_0 = [], [], [], []
    f0 = time.perf_counter()
    for plane, pix in ((0, px1), (1, px2)):
        base = torch.as_tensor(p1 if plane == 0 else p2).cuda()
        for c0_ in range(0, len(pix), 32):
            chunk = pix[c0_:c0_ + 32]
            gs_ = []
            for hh in (h0, h, 2 * h):
                keep, designs, up, dn = [], [], [], []
                for kk in chunk:
                    iy, ix = divmod(int(kk), n)
                    pair = []
                    for sgn in (1.0, -1.0):
                        vv = base.clone()
                        vv[iy, ix] += sgn * hh
                        keep.append(vv)
                        mm = G.dev_mask_from_tensor(vv, d, d)
                        pair.append([(mm, z, 1.0), (m2, z, 1.0)] if plane == 0 else [(m1, z, 1.0), (mm, z, 1.0)])
                    designs.append(tuple(pair))
                    x0 = float((p1 if plane == 0 else p2)[iy, ix])
                    up.append(x0 + hh)
                    dn.append(x0 - hh)
                st, gg, _, _ = G.fd_gradient_dev(designs, ud, d, d, lam, cfg, td, up, dn, [-1e3] * len(chunk),
                                                 [1e3] * len(chunk))
                if st:
                    raise RuntimeError(f"fd check failed: {st}")
                gs_.append(gg)
                del keep
            fd_0.extend(gs_[0])
            fd.extend((4.0 * gs_[1] - gs_[2]) / 3.0)
            fd_h.extend(gs_[1])
            for kk in chunk:
                iy, ix = divmod(int(kk), n)
                ad.append(g[plane, iy, ix])
    f1 = time.perf_counter()
    fd, ad, fd_0 = np.array(fd), np.array(ad), np.array(fd_0)
    # the same gradient in the complex64 form (one band, one scene of the
    # multi-frame entry point): FP32 fields, FP64 phases and sums
    u64 = torch.as_tensor(u[None, None]).to(torch.complex64).cuda().contiguous()
    g64 = torch.zeros_like(grads)
    call64 = G.adjoint_grid_gradient_frames_dev_prepared([steps], u64, d, d, [lam], cfg, td, g64)
    st64, merit64 = call64()
    if st64:
        raise RuntimeError(f"complex64 adjoint failed: {st64}")
    for _ in range(5):
        call64()
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    for _ in range(k):
        call64()
    torch.cuda.synchronize()
    s1 = time.perf_counter()
    h64 = g64.cpu().numpy()
    ad64 = np.array([h64[plane, int(kk) // n, int(kk) % n] for plane, pix in ((0, px1), (1, px2)) for kk in pix])
    r = {"workload": f"adjoint grid gradient {n}^2 f64, 2 DOEs + 2 ASM {z * 1e3:.0f} mm, MSE to a 16-spot target",
         "gradients_per_s": k / (t1 - t0), "ms_per_gradient": 1e3 * (t1 - t0) / k, "merit": merit,
         "call": "C ABI, arguments marshalled once",
         "python_wrapper_gradients_per_s": 1.0 / wrapped_s,
         "c64": {"gradients_per_s": k / (s1 - s0), "ms_per_gradient": 1e3 * (s1 - s0) / k, "merit": merit64,
                 "rel_l2_vs_c128_gradient": float(np.linalg.norm(h64 - g) / np.linalg.norm(g)),
                 "fd_rel_l2": float(np.linalg.norm(fd_0 - ad64) / np.linalg.norm(fd_0))},
         "fd_step_rad": h0, "fd_pixels": int(fd.size),
         "fd_rel_l2": float(np.linalg.norm(fd_0 - ad) / np.linalg.norm(fd_0)),
         "fd_max_abs_over_max": float(np.max(np.abs(fd_0 - ad)) / np.max(np.abs(fd_0))),
         "fd_richardson_1e-3_2e-3": {
             "rel_l2": float(np.linalg.norm(fd - ad) / np.linalg.norm(fd)),
             "max_rel": float(np.max(np.abs(fd - ad) / np.maximum(np.abs(fd), 1e-30))),
             "central_1e-3_rel_l2": float(np.linalg.norm(np.array(fd_h) - ad) / np.linalg.norm(fd))},
         "fd_seconds": f1 - f0}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle
        so = [(oracle.Mask(p1, d, d), z, 1.0), (oracle.Mask(p2, d, d), z, 1.0)]
        c0 = time.perf_counter()
        st, det = oracle.trace(so, u, d, d, lam, cfg)
        inten = det.real ** 2 + det.imag ** 2
        st2, ref = oracle.adjoint_push_back(so, u, d, d, lam, cfg, 2.0 * (inten - tgt) / inten.size)
        c1 = time.perf_counter()
        r["vs_f64_oracle_rel_l2"] = [float(np.linalg.norm(g[e] - ref[e]) / np.linalg.norm(ref[e])) for e in range(2)]
        r["cpu_baseline"] = {"gradients_per_s": 1.0 / (c1 - c0), "cores": 1, "kind": "port",
                             "sample": f"one gradient (trace + push_back) with the f64 oracle, {c1 - c0:.1f} s"}
    log(f"adjoint: {r['gradients_per_s']:.1f} gradients/s (C ABI; {r['python_wrapper_gradients_per_s']:.1f}/s through the Python wrapper), FD rel L2 {r['fd_rel_l2']:.2e}; "
        f"complex64 {r['c64']['gradients_per_s']:.1f}/s, rel L2 {r['c64']['rel_l2_vs_c128_gradient']:.1e} vs complex128, FD {r['c64']['fd_rel_l2']:.1e}"
        + (f", vs oracle {r['vs_f64_oracle_rel_l2']}" if cpu else ""))
    return r


def bench_frames(G, W, T, torch, np):
    """The multi-frame adjoint gradient (adjoint::grid_gradient with
    MseToTarget summed over every (band, scene) frame): workload 4's design
    (DOE1 -> ASM 10 mm -> DOE2 -> ASM 10 mm at 1024^2, P = 2048) under 3
    bands x 2 scenes (Gaussian beams, centred and shifted), complex128 and
    complex64 fields.  One gradient = 6 forward traces + 6 reverse sweeps,
    24 propagations; roofline of the column pass from the library's own
    per-pass timing (2 N P B algorithmic bytes per propagation)."""
    c = W.CFG4
    n, d, z = c["n"], c["d"], c["z"]
    u, p1, p2, tgt = W.workload4()
    bands = [1.50e-6, 1.55e-6, 1.60e-6]
    scenes = [(0.0, 0.0), (40 * d, -25 * d)]
    cfg = T.AsmConfig.default()
    m1, k1 = G.dev_mask(p1, d, d)
    m2, k2 = G.dev_mask(p2, d, d)
    designs = [[(m1, z, 1.0), (m2, z, 1.0)] for _ in bands]
    inp = np.zeros((len(bands), len(scenes), n, n), complex)
    for b, lam in enumerate(bands):
        for s_, cc in enumerate(scenes):
            inp[b, s_] = W.gaussian_beam(n, d, lam, 0.3 * n * d, center=cc)
    td = torch.as_tensor(tgt).cuda()
    grads = torch.zeros((2, n, n), dtype=torch.float64, device="cuda")
    ctx = G.context()
    P, frames = 2 * n, len(bands) * len(scenes)
    props = 4 * frames
    out = {"workload": f"adjoint grid gradient over {len(bands)} bands x {len(scenes)} scenes, {n}^2, 2 DOEs + "
                       f"2 ASM {z * 1e3:.0f} mm (P = {P}), MSE to a 16-spot target, per-element sums over frames",
           "propagations_per_gradient": props}
    for name, dt, B in (("c128", torch.complex128, 16), ("c64", torch.complex64, 8)):
        ud = torch.as_tensor(inp).to(dt).cuda().contiguous()
        st, merit = G.adjoint_grid_gradient_frames_dev(designs, ud, d, d, bands, cfg, td, grads, ctx=ctx)
        if st:
            raise RuntimeError(f"frames gradient failed: {st}")
        k = 5
        torch.cuda.synchronize()
        ctx.pass_timing(reset=True)
        ctx.set_timing_passes(["col"])
        ctx.enable_timing(True)
        t0 = time.perf_counter()
        for _ in range(k):
            G.adjoint_grid_gradient_frames_dev(designs, ud, d, d, bands, cfg, td, grads, ctx=ctx)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.enable_timing(False)
        ctx.set_timing_passes(None)
        col_ms, col_launches = ctx.pass_timing(reset=True)["col"]
        alg = k * props * 2 * n * P * B  # column-pass bytes over the timed calls
        achieved = alg / (col_ms * 1e-3) / 1e9 if col_ms > 0 else None
        out[name] = {"gradients_per_s": k / (t1 - t0), "ms_per_gradient": 1e3 * (t1 - t0) / k, "merit": merit,
                     "roofline": {"bound": "hbm", "kernel": "col", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
                                  "traffic": None, "col_ms_per_gradient": col_ms / k,
                                  "col_launches_per_gradient": col_launches / k}}
        log(f"frames gradient {name}: {out[name]['gradients_per_s']:.1f} gradients/s, column pass "
            f"{col_ms / k:.3f} ms per gradient at {achieved:.0f} GB/s")
        del ud
    return out


def bench_rs(G, W, T, torch, np, schedule_cache=None):
    """Workload 5: 64 uneven Rayleigh-Sommerfeld problems (sizes 256..2048,
    padding 2) in one batched call, complex64."""
    ps = W.workload5_params()
    ins = [torch.as_tensor(W.workload5_field(p, index=i).astype(np.complex64)).cuda() for i, p in enumerate(ps)]
    outs = [torch.empty_like(x) for x in ins]
    arr = (T.PropDesc * len(ps))()
    for i, p in enumerate(ps):
        arr[i].geom = T.geom(p["n"], p["n"], p["d"], p["d"], p["lam"])
        arr[i].z = p["z"]
        arr[i].engine = T.ENGINE_RS
        arr[i].padding = 2.0
        arr[i].strict = 0
        arr[i].in_ = ins[i].data_ptr()
        arr[i].out = outs[i].data_ptr()
    ctx = G.Context(0)  # a context of its own: the tuner's picks stay in it

    def timed(k=5):
        st = G.propagate_batch_dev(arr, ctx=ctx)
        if st:
            raise RuntimeError(f"rs batch failed: {st}")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            G.propagate_batch_dev(arr, ctx=ctx)
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / k

    ms_default = timed()
    # the batch schedule tuner on this batch: pad length, layout or radix plan per length
    cache = schedule_cache
    loaded = 0
    if cache and os.path.exists(cache):
        st, loaded = G.schedule_load(cache, ctx=ctx)
        if st:
            raise RuntimeError(f"schedule load failed: {st}")
        picks, search_ms = G.schedules(ctx=ctx), 0.0
    else:
        st, picks, search_ms = G.autotune_batch_dev(arr, sweeps=1, ctx=ctx)
        if st:
            raise RuntimeError(f"rs tuning failed: {st}")
        if cache:
            G.schedule_save(cache, ctx=ctx)
    ms = timed()
    ctx.close()
    px = sum(p["n"] ** 2 for p in ps)
    r = {"workload": "64 uneven RS problems (256..2048, padding 2), complex64, one call, tuned schedules",
         "problems_per_s": len(ps) / (ms * 1e-3), "ms_per_batch": ms, "ms_per_batch_default_schedule": ms_default,
         "tuner": {"search_ms": search_ms, "loaded_from_cache": loaded, "cache": cache, "picks": picks},
         "field_samples": px, "sizes": sorted({p["n"] for p in ps})}
    log(f"rs batch: {ms:.2f} ms per 64 problems tuned ({ms_default:.2f} default), "
        f"search {search_ms:.0f} ms, {len(picks)} picks")
    return r


def render_design(G, W, n, d, bands):
    """Two DOEs (height profiles: phase scales as 550 nm / lambda) and two
    ASM gaps, built for each band: designs[band][0] = step list with
    device-resident lattices (the keep list holds them)."""
    import numpy as np
    h1 = 2 * np.pi * W.splitmix_uniform(101, n * n).reshape(n, n)
    h2 = 2 * np.pi * W.splitmix_uniform(102, n * n).reshape(n, n)
    keep, designs = [], []
    for lam in bands:
        m1, k1 = G.dev_mask(h1 * 550e-9 / lam, d, d)
        m2, k2 = G.dev_mask(h2 * 550e-9 / lam, d, d)
        keep += [k1, k2]
        designs.append([[(m1, 4e-3, 1.0), (m2, 6e-3, 1.0)]])
    return designs, keep, (h1, h2)


def bench_render(G, W, T, torch, np, cpu=True):
    """Renderer::render: 8 bands (450-650 nm) x 4 scenes with the x and y FD
    channels (5 renders per frame, 160 renders), 512^2 at 4 um, two DOEs
    and two ASM gaps (P = 1024), complex64, one call."""
    n, d = 512, 4e-6
    bands = list(np.linspace(450e-9, 650e-9, 8))
    scenes = [(0.0, 0.0), (40e-6, 0.0), (0.0, -40e-6), (-60e-6, 60e-6)]
    rc = T.RenderConfig.make(x=True, y=True, scene_step=1e-6)
    designs, keep, hs = render_design(G, W, n, d, bands)
    src = lambda x0, y0, lam: W.gaussian_beam(n, d, lam, 400e-6, 1.0, (x0, y0))
    ins = []
    for lam in bands:
        for (x0, y0) in scenes:
            h = rc.scene_step
            ins.append(np.stack([src(x0, y0, lam), src(x0 + h, y0, lam), src(x0 - h, y0, lam),
                                 src(x0, y0 + h, lam), src(x0, y0 - h, lam)]))
    u = torch.as_tensor(np.stack(ins).reshape(len(bands), len(scenes), 5, n, n).astype(np.complex64)).cuda()
    obs = torch.zeros((3, len(bands), len(scenes), n, n), dtype=torch.float64, device="cuda")
    cfg = T.AsmConfig.default(strict=0)
    st = G.render_dev(designs, u, d, d, bands, cfg, rc, obs)
    if st:
        raise RuntimeError(f"render failed: {st}")
    k = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        G.render_dev(designs, u, d, d, bands, cfg, rc, obs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nr = len(bands) * len(scenes) * 5
    r = {"workload": "Renderer::render 8 bands x 4 scenes, x and y FD channels (160 renders), 512^2 at 4 um, "
                     "2 DOEs + 2 ASM gaps (P = 1024), complex64, one call",
         "renders_per_s": nr * k / (t1 - t0), "frames_per_s": len(bands) * len(scenes) * k / (t1 - t0),
         "ms_per_call": 1e3 * (t1 - t0) / k}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as orc
        m1 = orc.Mask(hs[0] * 550e-9 / bands[0], d, d)
        m2 = orc.Mask(hs[1] * 550e-9 / bands[0], d, d)
        steps = [(m1, 4e-3, 1.0), (m2, 6e-3, 1.0)]
        one = np.ascontiguousarray(u[0:1, 0:1].cpu().numpy().astype(np.complex128))
        c0 = time.perf_counter()
        st, _ = orc.render([[steps]], one, d, d, bands[:1], cfg, rc)
        c1 = time.perf_counter()
        r["cpu_baseline"] = {"renders_per_s": 5 / (c1 - c0), "cores": 1, "kind": "port",
                             "sample": f"one frame (5 renders) with the f64 oracle, single thread, {c1 - c0:.1f} s"}
    log(f"render: {r['renders_per_s']:.0f} renders/s ({r['ms_per_call']:.2f} ms per 160-render call)")
    return r


def bench_fd(G, W, T, torch, np):
    """fd::gradient with the render + MseToTarget objective: 64 pixel-phase
    slots of a two-DOE design (128 probe renders), 256^2 at 4 um, P = 512,
    complex128, one call."""
    n, d, lam = 256, 4e-6, 633e-9
    p1 = 2 * np.pi * W.splitmix_uniform(111, n * n).reshape(n, n)
    p2 = 2 * np.pi * W.splitmix_uniform(112, n * n).reshape(n, n)
    u = torch.as_tensor(W.gaussian_beam(n, d, lam, 300e-6)).cuda()
    tgt = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    tgt[n // 2, n // 2] = 1.0
    rng = np.random.default_rng(7)
    pix = rng.choice(n * n, 64, replace=False)
    h = 2 * np.pi * 1e-4
    o1, k1 = G.dev_mask(p1, d, d)
    o2, k2 = G.dev_mask(p2, d, d)
    keep, designs, up, dn = [k1, k2], [], [], []
    for k_, q in enumerate(pix):
        e = k_ % 2  # probes alternate between the two DOEs
        pair = []
        for sgn in (1.0, -1.0):
            lat = [p1, p2][e].copy().reshape(-1)
            lat[q] = min(max(lat[q] + sgn * h, 0.0), 2 * np.pi)
            (up if sgn > 0 else dn).append(lat[q])
            m, t_ = G.dev_mask(lat.reshape(n, n), d, d)
            keep.append(t_)
            pair.append([(m, 2e-3, 1.0), (o2, 2e-3, 1.0)] if e == 0 else [(o1, 2e-3, 1.0), (m, 2e-3, 1.0)])
        designs.append(tuple(pair))
    lo, hi = [0.0] * 64, [2 * np.pi] * 64
    cfg = T.AsmConfig.default(strict=0)
    st, g, _, ev = G.fd_gradient_dev(designs, u, d, d, lam, cfg, tgt, up, dn, lo, hi)
    if st:
        raise RuntimeError(f"fd failed: {st}")
    k = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        G.fd_gradient_dev(designs, u, d, d, lam, cfg, tgt, up, dn, lo, hi)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    r = {"workload": "fd::gradient, render + MseToTarget objective, 64 pixel slots (128 probe renders) of a "
                     "2-DOE + 2-gap design, 256^2 at 4 um (P = 512), complex128, one call",
         "probe_renders_per_s": ev * k / (t1 - t0), "ms_per_gradient": 1e3 * (t1 - t0) / k}
    log(f"fd: {r['probe_renders_per_s']:.0f} probe renders/s ({r['ms_per_gradient']:.2f} ms per gradient)")
    return r


def asm_only_rate(G, torch, arr, nb, k, dbl, ctx):
    """Masked ASM propagations/s of one batch call over the bands."""
    st = G.propagate_batch_dev(arr, double=dbl, ctx=ctx)
    if st:
        raise RuntimeError(f"ASM batch failed: {st}")
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    for _ in range(k):
        G.propagate_batch_dev(arr, double=dbl, ctx=ctx)
    torch.cuda.synchronize()
    return nb * k / (time.perf_counter() - s0)


TRAFFIC_FILE = "profiles/traffic_col_c64_workload2.txt"


def read_traffic(variant):
    """HBM bytes per launch of the column pass for the layout variant in use,
    from the committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE,
    copy-kernel calibrated), if any."""
    p = os.path.join(REPO, TRAFFIC_FILE)
    if not os.path.exists(p):
        return None
    key = f"traffic_bytes_per_launch_v{variant}="
    for line in open(p):
        if line.startswith(key):
            return float(line.split("=", 1)[1])
    return None


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the workloads split into "
              "independent bands/problems: one process per GPU would run this on its share).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import oc_types as T
    import optics_gpu as G
    import workloads as W

    if args.only:
        runners = {"gs": lambda: bench_gs(G, W, torch, np),
                   "gs_asm": lambda: bench_gs_asm(G, W, T, np),
                   "frames": lambda: bench_frames(G, W, T, torch, np),
                   "adjoint": lambda: bench_adjoint(G, W, T, torch, np, cpu=not args.no_cpu),
                   "rs": lambda: bench_rs(G, W, T, torch, np, args.schedule_cache),
                   "render": lambda: bench_render(G, W, T, torch, np, cpu=not args.no_cpu),
                   "fd": lambda: bench_fd(G, W, T, torch, np)}
        print(json.dumps({"only": args.only, "record": runners[args.only]()}))
        return

    c = W.CFG2
    n, d = c["n"], c["d"]
    lams = W.workload2_lambdas()[: args.bands]
    nb = len(lams)
    ctx = G.context()
    phi = W.workload2_doe()
    mask, keep = G.dev_mask(phi, d, d)
    field = W.workload2_field(lams[0])  # the waist field is wavelength-independent
    u_in = torch.as_tensor(field.astype(np.complex64)).cuda().unsqueeze(0).repeat(nb, 1, 1).contiguous()
    u_asm = torch.empty_like(u_in)
    u_fr = torch.empty_like(u_in)

    def descs(engine, src, dst, dbl=False):
        arr = (T.PropDesc * nb)()
        for b, lam in enumerate(lams):
            x = arr[b]
            x.geom = T.geom(n, n, d, d, lam)
            x.engine = engine
            if engine == T.ENGINE_ASM:
                x.z, x.padding, x.band_limited, x.strict, x.max_padding, x.medium = c["z_asm"], 2.0, 1, 1, 8.0, 1.0
                x.mask = T.C.pointer(mask)
            else:
                x.z, x.padding, x.strict, x.max_padding = c["z_fresnel"], 2.0, 0, 2.0
            x.in_ = src[b].data_ptr() if src is not None else 0
            x.out = dst[b].data_ptr() if dst is not None else 0
        return arr

    d_asm = descs(T.ENGINE_ASM, u_in, u_asm)
    d_fr = descs(T.ENGINE_FRESNEL, u_asm, u_fr)
    # the step: one chained call, mask + ASM then Fresnel per band; the ASM
    # field is kept on chip (first stage's out = NULL)
    ch_first = descs(T.ENGINE_ASM, u_in, None)
    ch_second = descs(T.ENGINE_FRESNEL, None, u_fr)

    def chain_step(first, second):
        st = G.propagate_chain_dev(first, second, ctx=ctx)
        if st:
            raise RuntimeError(f"chain failed: {T.STATUS.get(st, st)} {ctx.last_error()}")

    def chain_submit(first, second):
        # queued without waiting: the next step's host work overlaps this
        # step's kernels; the statuses of every queued step come at the wait
        st = G.propagate_chain_submit_dev(first, second, ctx=ctx, sync_torch=False)
        if st:
            raise RuntimeError(f"chain failed: {T.STATUS.get(st, st)} {ctx.last_error()}")

    def chain_wait():
        st = G.wait_dev(ctx)
        if st:
            raise RuntimeError(f"queued chains failed: {T.STATUS.get(st, st)} {ctx.last_error()}")

    def step(da, df, dbl=False):
        # both batches queued, one wait: the Fresnel batch reads the ASM
        # batch's outputs and runs after it on the context stream
        for what, ds in (("ASM", da), ("Fresnel", df)):
            st = G.propagate_batch_submit_dev(ds, double=dbl, ctx=ctx)
            if st:
                raise RuntimeError(f"{what} batch failed: {T.STATUS.get(st, st)} {ctx.last_error()}")
        st = G.wait_dev(ctx)
        if st:
            raise RuntimeError(f"step failed: {T.STATUS.get(st, st)} {ctx.last_error()}")

    # workgroup layouts for the padded length (untimed)
    if args.layout:
        lv = [int(x) for x in args.layout.split(",")]
        if ctx.set_fft_variant(2 * n, lv[0]) != 0 or (len(lv) == 3 and ctx.set_fresnel_sep_layout(2 * n, lv[1], lv[2])):
            raise SystemExit(f"--layout {args.layout}: no such layout at P = {2 * n}")
        variant, variant_ms = lv[0], None
    elif args.autotune:
        variant, variant_ms = ctx.autotune(2 * n, double=False)
    else:
        variant, variant_ms = ctx.fft_variant(2 * n), None
    sep_layout = ctx.fresnel_sep_layout(2 * n)
    log(f"layout P={2 * n} c64: variant {variant}, autotune ms per variant {variant_ms}; "
        f"two-pass Fresnel layouts (columns, rows) {sep_layout}")
    log(f"workload: {nb} bands x 2048^2, c64; warmup {args.warmup}")
    for _ in range(args.warmup):
        chain_step(ch_first, ch_second)
    torch.cuda.synchronize()
    # the timed region brackets the column pass (the dominant kernel) with
    # HIP events on the library's stream; the other passes run unbracketed
    ctx.pass_timing(reset=True)
    ctx.set_timing_passes(["col"])
    ctx.enable_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        chain_step(ch_first, ch_second)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.enable_timing(False)
    timing_col = ctx.pass_timing(reset=True)
    # the same steps queued with one wait (oc_propagate_chain_submit_dev_c64)
    torch.cuda.synchronize()
    y0 = time.perf_counter()
    for _ in range(args.steps):
        chain_submit(ch_first, ch_second)
    chain_wait()
    torch.cuda.synchronize()
    queued_ms_step = 1e3 * (time.perf_counter() - y0) / args.steps
    # per-pass breakdown: a few more steps with every pass bracketed (untimed)
    ctx.set_timing_passes(None)
    ctx.enable_timing(True)
    nbreak = max(5, args.steps // 5)
    for _ in range(nbreak):
        chain_step(ch_first, ch_second)
    torch.cuda.synchronize()
    ctx.enable_timing(False)
    timing = ctx.pass_timing(reset=True)
    elapsed = t1 - t0
    props = 2 * nb * args.steps
    value = props / elapsed
    ms_step = 1e3 * elapsed / args.steps
    log(f"c64: {ms_step:.3f} ms/step, {value:.1f} propagations/s ({queued_ms_step:.3f} ms/step with the steps "
        f"queued); passes {timing}")

    # Algorithmic bytes per launch (minimal-sweep model, complex64 = 8 B),
    # per band, 16 bands per launch.  The chained step: row_fwd reads N^2
    # and writes N P; col reads and writes N P; the fused row pass (timing
    # slot sep_row: the ASM's inverse rows and the Fresnel's rows) reads N P
    # and writes N^2; the Fresnel's column pass (slot sep_col) reads and
    # writes N^2.
    B = 8
    P = 2 * n
    alg = {"row_fwd": nb * (n * n + n * P) * B, "col": nb * (2 * n * P) * B, "row_inv": nb * (n * P + n * n) * B,
           "sep_col": nb * (2 * n * n) * B, "sep_row": nb * (n * P + n * n) * B}
    alg = {k: v for k, v in alg.items() if timing[k][1]}
    dom = max(alg, key=lambda k: timing[k][0])
    src = timing_col if dom == "col" else timing  # the timed region's events for the column pass
    avg_ms = src[dom][0] / max(src[dom][1], 1)
    achieved = alg[dom] / (avg_ms * 1e-3) / 1e9
    traffic = read_traffic(variant) if dom == "col" else None
    passes = {k: {"avg_ms": timing[k][0] / max(timing[k][1], 1), "launches": timing[k][1],
                  "alg_GBps": alg[k] / (timing[k][0] / max(timing[k][1], 1) * 1e-3) / 1e9}
              for k in alg if timing[k][1]}
    # reference point for the HBM figures: a plain 1 GiB device copy
    # (read + write) on the same GPU, timed with events
    ca = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    cb = torch.empty_like(ca)
    for _ in range(3):
        cb.copy_(ca)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        cb.copy_(ca)
    ev1.record()
    torch.cuda.synchronize()
    copy_gbps = 2 * (1 << 30) / (ev0.elapsed_time(ev1) / 10 * 1e-3) / 1e9
    del ca, cb
    # + the shared DOE: k_mask_trans reads the f64 lattice and writes the
    # complex64 transmission once, and row_fwd reads it once (L2 across bands)
    step_bytes = sum(alg[k] * timing[k][1] for k in alg) / nbreak + 3 * n * n * 8
    extra = {"passes": passes, "fft_variant": variant, "autotune_ms": variant_ms,
             "queued_steps_ms_per_step": queued_ms_step,
             "queued_steps_propagations_per_s": 2 * nb / (queued_ms_step * 1e-3),
             "fresnel_sep_layout": list(sep_layout), "step_alg_GB": step_bytes / 1e9,
             "device_copy_GBps": copy_gbps,
             "step_alg_GBps": step_bytes / (elapsed / args.steps) / 1e9}

    if not args.no_extra:
        # the same chain with the ASM field written out, and the two stages
        # as two queued batch calls (untimed for `value`)
        k = max(5, args.steps // 5)
        for name, fn in (("chain_with_intermediate", lambda: chain_step(d_asm, ch_second)),
                         ("two_batch_calls", lambda: step(d_asm, d_fr))):
            fn()
            torch.cuda.synchronize()
            s0 = time.perf_counter()
            for _ in range(k):
                fn()
            torch.cuda.synchronize()
            extra[name + "_propagations_per_s"] = 2 * nb * k / (time.perf_counter() - s0)
        log(f"chain with the ASM field written: {extra['chain_with_intermediate_propagations_per_s']:.0f}/s; "
            f"two batch calls: {extra['two_batch_calls_propagations_per_s']:.0f}/s")
        # ASM alone: the 16 masked ASM propagations of a step as one batch call
        extra["asm_only_c64_propagations_per_s"] = asm_only_rate(G, torch, d_asm, nb, k, False, ctx)

    # parity of this run's last band output against the f64 oracle (and the
    # oracle timed on the host as the CPU baseline)
    cpu = None
    if not args.no_cpu:
        chain_step(d_asm, ch_second)  # both stages' fields of the chained call
        torch.cuda.synchronize()
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle
        lam = lams[-1]
        u = W.workload2_field(lam)
        om = oracle.Mask(phi, d, d)
        t_c0 = time.perf_counter()
        st, um = oracle.phasemask_apply(u, d, d, lam, om)
        st, ra, _ = oracle.asm_propagate(um, d, d, lam, c["z_asm"], T.AsmConfig.default())
        st2, rf, _ = oracle.fresnel_propagate(ra, d, d, lam, c["z_fresnel"], T.FresnelConfig(2.0, 0, 0, 2.0))
        t_c1 = time.perf_counter()
        if st or st2:
            raise RuntimeError("oracle failed")
        ga = u_asm[nb - 1].cpu().numpy().astype(np.complex128)
        gf = u_fr[nb - 1].cpu().numpy().astype(np.complex128)
        e_asm = float(np.linalg.norm(ga - ra) / np.linalg.norm(ra))
        e_fr = float(np.linalg.norm(gf - rf) / np.linalg.norm(rf))
        extra["rel_l2_c64_vs_f64_oracle"] = {"asm": e_asm, "fresnel": e_fr, "band_nm": lam * 1e9}
        cpu = {"value": 2.0 / (t_c1 - t_c0), "unit": "propagations/s", "cores": 1, "kind": "port",
               "sample": f"1 of {nb} bands ({lam * 1e9:.1f} nm): DOE mask + ASM 20 mm + Fresnel 200 mm at 2048^2 "
                         f"(P=4096), f64 oracle, single thread, {t_c1 - t_c0:.1f} s"}
        log(f"cpu oracle: {t_c1 - t_c0:.1f} s for 2 propagations; rel L2 asm {e_asm:.2e} fresnel {e_fr:.2e}")
        # all host cores: one band per thread (the oracle keeps no shared
        # state, and ctypes releases the GIL inside each call)
        from concurrent.futures import ThreadPoolExecutor
        ncores = min(16, os.cpu_count() or 1, nb)

        def one_band(lam_b):
            ub = W.workload2_field(lam_b)
            s1, ubm = oracle.phasemask_apply(ub, d, d, lam_b, om)
            s2, ub_a, _ = oracle.asm_propagate(ubm, d, d, lam_b, c["z_asm"], T.AsmConfig.default())
            s3, _, _ = oracle.fresnel_propagate(ub_a, d, d, lam_b, c["z_fresnel"], T.FresnelConfig(2.0, 0, 0, 2.0))
            return s1 or s2 or s3

        m0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=ncores) as ex:
            sts = list(ex.map(one_band, lams[:ncores]))
        m1 = time.perf_counter()
        if any(sts):
            raise RuntimeError("oracle failed")
        cpu["all_cores"] = {"value": 2.0 * ncores / (m1 - m0), "unit": "propagations/s", "cores": ncores,
                            "sample": f"{ncores} bands, one per thread, same chain, {m1 - m0:.1f} s"}
        log(f"cpu oracle on {ncores} threads: {m1 - m0:.1f} s for {2 * ncores} propagations")

    if not args.no_extra:
        # complex128 device form of the same batch
        v_in = u_in.to(torch.complex128)
        v_asm = torch.empty_like(v_in)
        v_fr = torch.empty_like(v_in)
        e_asm_d = descs(T.ENGINE_ASM, v_in, None)
        e_fr_d = descs(T.ENGINE_FRESNEL, None, v_fr)

        def c128_step():
            st = G.propagate_chain_dev(e_asm_d, e_fr_d, double=True, ctx=ctx)
            if st:
                raise RuntimeError(f"c128 chain failed: {T.STATUS.get(st, st)} {ctx.last_error()}")

        c128_step()
        torch.cuda.synchronize()
        k = max(3, args.steps // 10)
        s0 = time.perf_counter()
        for _ in range(k):
            c128_step()
        torch.cuda.synchronize()
        s1 = time.perf_counter()
        # the same chained call in complex128 (its three-pass steps linked through their rows)
        extra["c128_device_propagations_per_s"] = 2 * nb * k / (s1 - s0)
        # ASM alone in complex128 on the device, and the FP32-bulk form's ratio over it
        e_asm_full = descs(T.ENGINE_ASM, v_in, v_asm)
        extra["asm_only_c128_device_propagations_per_s"] = asm_only_rate(G, torch, e_asm_full, nb, k, True, ctx)
        extra["asm_only_c64_over_c128_device"] = (extra["asm_only_c64_propagations_per_s"] /
                                                  extra["asm_only_c128_device_propagations_per_s"])
        log(f"ASM only: c64 {extra['asm_only_c64_propagations_per_s']:.0f}/s, c128 device "
            f"{extra['asm_only_c128_device_propagations_per_s']:.0f}/s, ratio {extra['asm_only_c64_over_c128_device']:.2f}")
        del v_in, v_asm, v_fr
        # f64 host drop-in (oc_asm_propagate: host memory in and out, PCIe included)
        um = (W.workload2_field(lams[0]) * np.exp(1j * phi)).astype(np.complex128)
        G.asm_propagate(um, d, d, lams[0], c["z_asm"], T.AsmConfig.default(), ctx=ctx)
        h0 = time.perf_counter()
        for _ in range(3):
            G.asm_propagate(um, d, d, lams[0], c["z_asm"], T.AsmConfig.default(), ctx=ctx)
        h1 = time.perf_counter()
        extra["f64_host_dropin_propagations_per_s"] = 3 / (h1 - h0)
        log(f"c128 device {extra['c128_device_propagations_per_s']:.1f}/s, f64 host drop-in "
            f"{extra['f64_host_dropin_propagations_per_s']:.2f}/s")
        del u_in, u_asm, u_fr
        extra["gs"] = bench_gs(G, W, torch, np)
        extra["gs_asm"] = bench_gs_asm(G, W, T, np)
        extra["adjoint"] = bench_adjoint(G, W, T, torch, np, cpu=not args.no_cpu)
        extra["adjoint_frames"] = bench_frames(G, W, T, torch, np)
        extra["rs_batch"] = bench_rs(G, W, T, torch, np, args.schedule_cache)
        extra["render"] = bench_render(G, W, T, torch, np, cpu=not args.no_cpu)
        extra["fd"] = bench_fd(G, W, T, torch, np)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "propagations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"polychromatic BandField {nb} bands 450-650 nm, 2048^2 at 2 um, 8-level DOE mask "
                               "+ ASM 20 mm (P=4096) + Fresnel TF 200 mm, FP32 bulk / FP64 phase",
                   "bands": nb, "grid": n, "padded": P, "propagations_per_step": 2 * nb,
                   "call": "chained (ASM field not written)"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": (f"{TRAFFIC_FILE} (rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel and "
                                        f"layout, copy-calibrated; committed, not measured in this run)"
                                        if traffic is not None else None),
                     "alg_bytes_per_launch": alg[dom], "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
