"""The benchmark workloads' own hot kernels at the sizes they run, against
the f64 oracle, through the C ABI.

- Workload 5 (uneven Rayleigh-Sommerfeld, rayleigh_sommerfeld.rs:384-442):
  one N = 2048 problem (P = 4096) in complex64 and complex128, and one draw
  of each grid size 256/384/512/768/1024/2048 through the batched call,
  each problem against or_rs_propagate.
- Workload 3 (FFT-form GS, 1024^2): 2 restarts x 5 iterations of the fused,
  graph-captured path against the oracle's FFT-form GS, and the full 32 x
  200 run's reported final error against a recomputation (numpy f64) from
  the phases it returns.
- Workload 4 (adjoint, adjoint.rs:140-340): the 1024^2 (P = 2048) grid
  gradient against the oracle's trace + MseToTarget seed + push_back, and
  against central finite differences at h = 1e-4 rad on 32 seeded pixels.
- Workload 2 (polychromatic DOE + ASM + Fresnel at 2048^2): one band in
  complex128 through the chained call, and the 16-band complex64 chained
  call (the timed step) on its first and last band.

Bars: complex128 1e-12 relative L2, complex64 1e-5 (GS: the nonlinear
iteration's bars of test_gpu_inverse.py), adjoint 1e-10 against the oracle
and 1e-4 against FD.
"""
import math
import sys

import numpy as np
import pytest

import oc_types as T
import workloads as W

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
C64_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def G():
    import optics_gpu
    return optics_gpu


_ORACLE = {}


def memo(key, fn):
    """Oracle results shared by the tests of this module (each takes seconds
    at these sizes)."""
    if key not in _ORACLE:
        _ORACLE[key] = fn()
    return _ORACLE[key]


# --------------------------------------------------------------- workload 5

def w5_draws():
    """The first draw of each grid size among the workload's 64, with its
    index (the field's seed offset)."""
    ps = W.workload5_params()
    first = {}
    for i, p in enumerate(ps):
        first.setdefault(p["n"], (i, p))
    return [first[n] for n in sorted(first)]


def w5_oracle(orc, i, p):
    def run():
        u = W.workload5_field(p, index=i)
        st, ref, _ = orc.rs_propagate(u, p["d"], p["d"], p["lam"], p["z"], T.RsConfig(2.0, 0, 0))
        assert st == 0
        return u, ref
    return memo(("rs", i), run)


def test_rs_p4096_both_forms(G, orc):
    """RS at P = 4096: the kernel row spectra (half rows) and the column pass
    with the kernel-spectrum product, complex64 and complex128."""
    i, p = [x for x in w5_draws() if x[1]["n"] == 2048][0]
    say("rs 2048 draw", i, p)
    u, ref = w5_oracle(orc, i, p)
    for name, dt, tol in (("c64", "c64", C64_TOL), ("c128", "c128", F64_TOL)):
        eng = G.DevEngine(dt)
        st, out, rep = eng.rs_propagate(u, p["d"], p["d"], p["lam"], p["z"], T.RsConfig(2.0, 0, 0))
        assert st == 0, (name, T.STATUS[st])
        assert (rep.nx_pad, rep.ny_pad) == (4096, 4096)
        e = rel_l2(out, ref)
        say("  ", name, e)
        assert e <= tol, (name, e)


@pytest.mark.parametrize("double", [False, True], ids=["c64", "c128"])
def test_workload5_sizes_in_one_batch(G, orc, double):
    """One batched call over a draw of every grid size of workload 5 (padded
    512/768/1024/1536/2048/4096, the non-power-of-two ones run at the next
    power of two), each problem against the oracle at the reference's P."""
    import torch
    draws = w5_draws()
    assert sorted(p["n"] for _, p in draws) == [256, 384, 512, 768, 1024, 2048]
    dt = torch.complex128 if double else torch.complex64
    ins, outs = [], []
    arr = (T.PropDesc * len(draws))()
    for k, (i, p) in enumerate(draws):
        u, _ = w5_oracle(orc, i, p)
        ins.append(torch.as_tensor(u).to(dt).cuda())
        outs.append(torch.empty_like(ins[-1]))
        arr[k].geom = T.geom(p["n"], p["n"], p["d"], p["d"], p["lam"])
        arr[k].z = p["z"]
        arr[k].engine = T.ENGINE_RS
        arr[k].padding = 2.0
        arr[k].strict = 0
        arr[k].in_ = ins[-1].data_ptr()
        arr[k].out = outs[-1].data_ptr()
    say("rs batch sizes", [p["n"] for _, p in draws], "double", double)
    st = G.propagate_batch_dev(arr, double=double)
    assert st == 0, T.STATUS[st]
    tol = F64_TOL if double else C64_TOL
    for k, (i, p) in enumerate(draws):
        assert arr[k].status == 0
        _, ref = w5_oracle(orc, i, p)
        e = rel_l2(outs[k].cpu().numpy().astype(np.complex128), ref)
        say("  ", p["n"], e)
        assert e <= tol, (p["n"], e)


# --------------------------------------------------------------- workload 3

def gs_error(amp, t_dft, phases):
    """||I/sum I - T^2/sum T^2|| / ||T^2/sum T^2|| per restart, numpy f64."""
    t2 = t_dft ** 2
    t2 = t2 / t2.sum()
    out = []
    for ph in phases:
        inten = np.abs(np.fft.fft2(amp * np.exp(1j * ph.astype(np.float64)))) ** 2
        out.append(np.linalg.norm(inten / inten.sum() - t2) / np.linalg.norm(t2))
    return np.array(out)


def test_workload3_gs_1024_vs_oracle(G, orc):
    """2 restarts x 5 iterations at 1024^2 on the fused row pass and column
    pass (default layout), graph-captured in chunks of 2."""
    import torch
    n, R, its = 1024, 2, 5
    amp, tgt, ph = W.workload3(n=n, restarts=R)
    tgt = np.fft.ifftshift(tgt)
    say("gs 1024 vs oracle")
    ref, err_ref = orc.gs_fft(amp, tgt, ph, its)
    p64 = torch.as_tensor(ph.copy()).cuda()
    st, err = G.gs_fft_run_dev(torch.as_tensor(amp).cuda(), torch.as_tensor(tgt).cuda(), p64, its, 2)
    assert st == 0, T.STATUS[st]
    got = p64.cpu().numpy()
    for r in range(R):
        dpw = (np.abs(np.exp(1j * got[r]) - np.exp(1j * ref[r])) * amp).sum() / amp.sum()
        say("  c128", r, dpw, err[r], err_ref[r])
        assert dpw < 1e-9
        assert abs(err[r] - err_ref[r]) < 1e-9
    p32 = torch.as_tensor(ph.astype(np.float32)).cuda()
    st, err32 = G.gs_fft_run_dev(torch.as_tensor(amp.astype(np.float32)).cuda(),
                                 torch.as_tensor(tgt.astype(np.float32)).cuda(), p32, its, 2)
    assert st == 0, T.STATUS[st]
    got32 = p32.cpu().numpy().astype(np.float64)
    for r in range(R):
        e = rel_l2(amp * np.exp(1j * got32[r]), amp * np.exp(1j * ref[r]))
        say("  c64", r, e, err32[r])
        assert e < 1e-4
        assert abs(err32[r] - err_ref[r]) < 1e-4


def test_workload3_full_run_reports_its_error(G):
    """The full workload (32 restarts x 200 iterations, complex64, graph
    chunk 20): the reported final error of every restart equals the error
    recomputed in f64 from the returned phases, and GS lowered it."""
    import torch
    c = W.CFG3
    amp, tgt, ph = W.workload3()
    tgt = np.fft.ifftshift(tgt)
    p = torch.as_tensor(ph.astype(np.float32)).cuda()
    say("gs full 32 x 200")
    st, err = G.gs_fft_run_dev(torch.as_tensor(amp.astype(np.float32)).cuda(),
                               torch.as_tensor(tgt.astype(np.float32)).cuda(), p, c["iterations"], 20)
    assert st == 0, T.STATUS[st]
    got = p.cpu().numpy()
    assert got.shape == (c["restarts"], c["n"], c["n"]) and np.all(np.isfinite(got))
    recomputed = gs_error(amp, tgt, got)
    start = gs_error(amp, tgt, ph[:2].astype(np.float32))
    say("  reported", err[:4], "recomputed", recomputed[:4], "start", start)
    assert np.all(np.abs(err - recomputed) <= 1e-4 * recomputed + 1e-6)
    assert np.all(err[:2] < 0.5 * start)


# --------------------------------------------------------------- workload 4

def test_workload4_gradient_1024(G, orc):
    """The 1024^2 grid gradient (P = 2048, f64) against the oracle's adjoint
    (1e-10) and against central FD at h = 1e-4 rad on 16 + 16 seeded pixels
    of the two planes, evaluated as one batch of 64 probe renders (1e-4)."""
    import torch
    c = W.CFG4
    n, d, lam, z = c["n"], c["d"], c["lam"], c["z"]
    u, p1, p2, tgt = W.workload4()
    cfg = T.AsmConfig.default()
    m1, k1 = G.dev_mask(p1, d, d)
    m2, k2 = G.dev_mask(p2, d, d)
    ud = torch.as_tensor(u).cuda()
    td = torch.as_tensor(tgt).cuda()
    grads = torch.zeros((2, n, n), dtype=torch.float64, device="cuda")
    say("adjoint 1024")
    st, merit = G.adjoint_grid_gradient_dev([(m1, z, 1.0), (m2, z, 1.0)], ud, d, d, lam, cfg, td, grads)
    assert st == 0, T.STATUS[st]
    g = grads.cpu().numpy()
    so = [(orc.Mask(p1, d, d), z, 1.0), (orc.Mask(p2, d, d), z, 1.0)]
    st, det = orc.trace(so, u, d, d, lam, cfg)
    assert st == 0
    inten = det.real ** 2 + det.imag ** 2
    m_ref = np.mean((inten - tgt) ** 2)
    assert abs(merit - m_ref) <= 1e-12 * m_ref
    st, ref = orc.adjoint_push_back(so, u, d, d, lam, cfg, 2.0 * (inten - tgt) / inten.size)
    assert st == 0
    for e in range(2):
        r = rel_l2(g[e], ref[e])
        say("  plane", e, "vs oracle", r)
        assert r < 1e-10
    h = c["fd_step"]
    assert h == 1e-4
    px1, px2 = W.workload4_fd_pixels(count=16)
    keep, designs, up, dn, ad = [], [], [], [], []
    for plane, pix in ((0, px1), (1, px2)):
        base = torch.as_tensor(p1 if plane == 0 else p2).cuda()
        for kk in pix:
            iy, ix = divmod(int(kk), n)
            pair = []
            for sgn in (1.0, -1.0):
                vv = base.clone()
                vv[iy, ix] += sgn * h
                keep.append(vv)
                mm = G.dev_mask_from_tensor(vv, d, d)
                pair.append([(mm, z, 1.0), (m2, z, 1.0)] if plane == 0 else [(m1, z, 1.0), (mm, z, 1.0)])
            designs.append(tuple(pair))
            x0 = float((p1 if plane == 0 else p2)[iy, ix])
            up.append(x0 + h)
            dn.append(x0 - h)
            ad.append(g[plane, iy, ix])
    st, fd, _, ev = G.fd_gradient_dev(designs, ud, d, d, lam, cfg, td, up, dn, [-1e3] * len(up), [1e3] * len(up))
    assert st == 0, T.STATUS[st]
    assert ev == 2 * len(up) == 64
    ad = np.array(ad)
    r = float(np.linalg.norm(fd - ad) / np.linalg.norm(fd))
    say("  fd rel L2", r, "pixels", len(ad))
    assert r <= 1e-4


# --------------------------------------------------------------- workload 2

def w2_oracle(orc, band):
    c = W.CFG2
    lam = W.workload2_lambdas()[band]

    def run():
        u = W.workload2_field(lam)
        m = orc.Mask(W.workload2_doe(), c["d"], c["d"])
        st, um = orc.phasemask_apply(u, c["d"], c["d"], lam, m)
        assert st == 0
        st, ra, _ = orc.asm_propagate(um, c["d"], c["d"], lam, c["z_asm"], T.AsmConfig.default())
        assert st == 0
        st, rf, _ = orc.fresnel_propagate(ra, c["d"], c["d"], lam, c["z_fresnel"], T.FresnelConfig(2.0, 0, 0, 2.0))
        assert st == 0
        return ra, rf
    return memo(("w2", band), run)


def w2_descs(lams, ins, mids, outs, mask):
    c = W.CFG2
    n, d = c["n"], c["d"]
    a, f = (T.PropDesc * len(lams))(), (T.PropDesc * len(lams))()
    for b, lam in enumerate(lams):
        for x, eng, z, pad in ((a[b], T.ENGINE_ASM, c["z_asm"], 2.0), (f[b], T.ENGINE_FRESNEL, c["z_fresnel"], 2.0)):
            x.geom = T.geom(n, n, d, d, lam)
            x.engine = eng
            x.z, x.padding, x.max_padding, x.medium = z, pad, 8.0 if eng == T.ENGINE_ASM else 2.0, 1.0
            x.band_limited, x.strict = (1, 1) if eng == T.ENGINE_ASM else (0, 0)
        a[b].mask = T.C.pointer(mask)
        a[b].in_ = ins[b].data_ptr()
        a[b].out = mids[b].data_ptr() if mids is not None else 0
        f[b].out = outs[b].data_ptr()
    return a, f


def test_workload2_band_c128_chain(G, orc):
    """One band of workload 2 in the f64 plain form at full size: mask + ASM
    20 mm (P = 4096) then Fresnel 200 mm through the complex128 chained call
    (its two three-pass steps linked through their rows)."""
    import torch
    c = W.CFG2
    band = 0
    lam = W.workload2_lambdas()[band]
    say("workload2 c128 band", band)
    ra, rf = w2_oracle(orc, band)
    mask, keep = G.dev_mask(W.workload2_doe(), c["d"], c["d"])
    u = torch.as_tensor(W.workload2_field(lam)).cuda()
    mid, out = torch.empty_like(u), torch.empty_like(u)
    a, f = w2_descs([lam], [u], [mid], [out], mask)
    st = G.propagate_chain_dev(a, f, double=True)
    assert st == 0, T.STATUS[st]
    ea = rel_l2(mid.cpu().numpy(), ra)
    ef = rel_l2(out.cpu().numpy(), rf)
    say("  asm", ea, "fresnel", ef)
    assert ea <= F64_TOL and ef <= F64_TOL


def test_workload2_sixteen_band_chain_c64(G, orc):
    """The timed step of workload 2: all 16 bands through the complex64
    chained call without the ASM field (fused ASM inverse rows + Fresnel
    rows, then the Fresnel columns); its first and last bands against the
    oracle."""
    import torch
    c = W.CFG2
    lams = W.workload2_lambdas()
    mask, keep = G.dev_mask(W.workload2_doe(), c["d"], c["d"])
    ins = [torch.as_tensor(W.workload2_field(lam).astype(np.complex64)).cuda() for lam in lams]
    outs = [torch.empty_like(x) for x in ins]
    a, f = w2_descs(lams, ins, None, outs, mask)
    say("workload2 16-band chain c64")
    st = G.propagate_chain_dev(a, f)
    assert st == 0, T.STATUS[st]
    for band in (0, len(lams) - 1):
        _, rf = w2_oracle(orc, band)
        e = rel_l2(outs[band].cpu().numpy().astype(np.complex128), rf)
        say("  band", band, e)
        assert e <= C64_TOL, (band, e)
