// Gerchberg-Saxton (drop-in and FFT form) and the adjoint gradient.
#include <climits>

#include "runtime.hpp"
#include "std_rng.hpp"

using namespace ocr;

namespace {


inline double rem_euclid_h(double x, double m) {
    const double r = std::fmod(x, m);
    return r < 0.0 ? r + std::fabs(m) : r;
}
inline double encode_bounds_h(double phi, double lo, double hi) {  // gs.rs:115-121
    if (hi - lo >= 2.0 * och::PI) return lo + rem_euclid_h(phi - lo, 2.0 * och::PI);
    return phi < lo ? lo : (phi > hi ? hi : phi);
}
inline bool contains_h(const oc_clear_aperture& ap, double x, double y) {  // element.rs:476-486
    switch (ap.kind) {
        case OC_APERTURE_FIELD: return true;
        case OC_APERTURE_CIRCLE: return x * x + y * y <= ap.a * ap.a;
        case OC_APERTURE_ELLIPSE: {
            const double a2 = ap.a * ap.a, b2 = ap.b * ap.b;
            return b2 * x * x + a2 * y * y <= a2 * b2;
        }
        default: return std::fabs(x) <= ap.a && std::fabs(y) <= ap.b;
    }
}

// Runs the already prepared persistent batch `b` and reads `nflags` flags.
int run_persistent(oc_context* ctx, Batch& b, int* dflags, int nflags, unsigned passes, std::vector<int>* flags_out) {
    HIP_TRY(ctx, ocl::launch_fill_int(dflags, 0, nflags, ctx->stream));
    int st = launch_batch(ctx, b, passes);
    if (st) return st;
    if (flags_out) {
        flags_out->assign(nflags, 0);
        HIP_TRY(ctx, hipMemcpyAsync(flags_out->data(), dflags, sizeof(int) * nflags, hipMemcpyDeviceToHost,
                                    ctx->stream));
    }
    return OC_OK;
}

int gs_fft_impl(oc_context* ctx, bool dbl, uint64_t n, uint64_t R, const void* A, const void* Tamp, void* phases,
                uint64_t iterations, uint64_t chunk, double* err_out) {
    if (!ctx || !A || !Tamp || !phases || n == 0 || R == 0) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    const size_t esz_r = dbl ? 8 : 4, esz = 2 * esz_r;
    const size_t nn = (size_t)n * n;
    Scoped mem(ctx);
    // layout (same plan on both axes)
    FftPlanDev pl;
    int st = make_plan(ctx, (int)n, dbl, &pl);
    if (st) return st;
    Layout L = choose_layout(pl, pl, esz, n, n, false);
    if (!L.ok) return fail(ctx, OC_ERR_UNSUPPORTED, "grid side exceeds the on-chip transform size");
    void *tblk, *sums, *flags;
    if ((st = mem.alloc(nn * esz_r, &tblk))) return st;
    if ((st = mem.alloc(sizeof(double) * 4 * R, &sums))) return st;
    if ((st = mem.alloc(sizeof(int) * R, &flags))) return st;
    oc_band_geom g{n, n, 1.0, 1.0, 1.0};
    std::vector<Prob> it(R), me(R);
    for (uint64_t r = 0; r < R; ++r) {
        Prob& p = it[r];
        p.g = g;
        p.engine = ENG_GS_PROJECT;
        p.mask = nullptr;
        p.in = nullptr;
        p.out = nullptr;
        p.ops.pre = PRE_POLAR;
        p.ops.amp = A;
        p.ops.phase = static_cast<unsigned char*>(phases) + r * nn * esz_r;
        p.ops.open = nullptr;
        p.ops.post = POST_GS_PHASE;
        p.ops.phase_io = static_cast<unsigned char*>(phases) + r * nn * esz_r;
        p.ops.beta_p = 1.0;
        p.ops.lo = 0.0;
        p.ops.hi = 2.0 * och::PI;
        p.ops.tamp_blk = tblk;
        p.ops.sums = nullptr;
        prepare(p);
        me[r] = p;
        me[r].engine = ENG_MEASURE;
        me[r].ops.post = POST_NONE;
        me[r].ops.sums = static_cast<double*>(sums) + 4 * r;
    }
    Batch bi, bm;
    st = build_batch(ctx, dbl, it, static_cast<int*>(flags), bi, true);
    if (st) { free_batch(ctx, bi); return st; }
    st = build_batch(ctx, dbl, me, static_cast<int*>(flags), bm, true);
    if (st) { free_batch(ctx, bi); free_batch(ctx, bm); return st; }
    auto cleanup = [&]() {
        free_batch(ctx, bi);
        free_batch(ctx, bm);
    };
    if (bi.pd.size() != R || bm.pd.size() != R) { cleanup(); return it[0].status ? it[0].status : OC_ERR_UNSUPPORTED; }
    // the target amplitude in the column pass's blocked order, with the block
    // width of the kernels the batches were built for
    if (bi.pd[0].B != bm.pd[0].B ||
        ocl::launch_block_reorder(dbl, Tamp, tblk, (int)n, (int)n, bi.pd[0].B, ctx->stream) != hipSuccess) {
        cleanup();
        return fail(ctx, OC_ERR_DEVICE, "target reorder failed");
    }
    const unsigned passes3 = (1u << 0) | (1u << 1) | (1u << 2);
    uint64_t done = 0;
    if (bi.fast && iterations > 0) {
        // power-of-two grids: row_fwd once, then per iteration the column
        // pass and the fused row pass (inverse rows + phase update + next
        // forward rows); the last iteration ends with the plain inverse
        // rows.  Graph-captured in chunks of [col, fused row].
        // beta = 1 over [0, 2 pi) with no per-iteration outputs: amplitude
        // replacement in the fused pass, phases from the last inverse pass
        const bool amp_only = it[0].ops.beta_p == 1.0 && it[0].ops.hi - it[0].ops.lo >= 2.0 * och::PI &&
                              !it[0].ops.sums && !it[0].ops.dphi_out && !it[0].out;
        const unsigned cf = (1u << 1) | (1u << (amp_only ? PASS_GS_AMP : PASS_GS_ROW));
        st = launch_batch(ctx, bi, 1u << 0);
        if (st) { cleanup(); return st; }
        const uint64_t body = iterations - 1;  // iterations with a fused row pass
        if (chunk > 0 && body >= chunk) {
            const bool was = ctx->timing;
            ctx->timing = false;
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
            if (e != hipSuccess) { ctx->timing = was; cleanup(); return fail(ctx, OC_ERR_DEVICE, "begin capture"); }
            for (uint64_t k = 0; k < chunk && st == OC_OK; ++k) st = launch_batch(ctx, bi, cf);
            e = hipStreamEndCapture(ctx->stream, &graph);
            ctx->timing = was;
            if (st || e != hipSuccess) { cleanup(); return st ? st : fail(ctx, OC_ERR_DEVICE, "end capture"); }
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) { (void)hipGraphDestroy(graph); cleanup(); return fail(ctx, OC_ERR_DEVICE, "instantiate"); }
            while (body - done >= chunk) {
                e = timed(ctx, 4, [&] { return hipGraphLaunch(exec, ctx->stream); });
                if (e != hipSuccess) break;
                done += chunk;
            }
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) { cleanup(); return fail(ctx, OC_ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e)); }
        }
        for (; done < body; ++done) {
            st = launch_batch(ctx, bi, cf);
            if (st) { cleanup(); return st; }
        }
        st = launch_batch(ctx, bi, (1u << 1) | (1u << 2));
        if (st) { cleanup(); return st; }
        done = iterations;
    }
    if (chunk > 0 && iterations - done >= chunk) {
        // capture `chunk` iterations once, replay
        const bool was = ctx->timing;
        ctx->timing = false;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) { ctx->timing = was; cleanup(); return fail(ctx, OC_ERR_DEVICE, "begin capture"); }
        for (uint64_t k = 0; k < chunk && st == OC_OK; ++k) st = launch_batch(ctx, bi, passes3);
        e = hipStreamEndCapture(ctx->stream, &graph);
        ctx->timing = was;
        if (st || e != hipSuccess) { cleanup(); return st ? st : fail(ctx, OC_ERR_DEVICE, "end capture"); }
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) { (void)hipGraphDestroy(graph); cleanup(); return fail(ctx, OC_ERR_DEVICE, "instantiate"); }
        while (iterations - done >= chunk) {
            e = timed(ctx, 4, [&] { return hipGraphLaunch(exec, ctx->stream); });
            if (e != hipSuccess) break;
            done += chunk;
        }
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) { cleanup(); return fail(ctx, OC_ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e)); }
    }
    for (; done < iterations; ++done) {
        st = launch_batch(ctx, bi, passes3);
        if (st) { cleanup(); return st; }
    }
    // final error: spectrum statistics of the final phases
    HIP_TRY(ctx, hipMemsetAsync(sums, 0, sizeof(double) * 4 * R, ctx->stream));
    std::vector<int> fl;
    st = run_persistent(ctx, bm, static_cast<int*>(flags), (int)R, (1u << 0) | (1u << 1), &fl);
    if (st) { cleanup(); return st; }
    std::vector<double> hs(4 * R);
    std::vector<unsigned char> th(nn * esz_r);
    HIP_TRY(ctx, hipMemcpyAsync(hs.data(), sums, sizeof(double) * 4 * R, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(th.data(), Tamp, nn * esz_r, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timing(ctx);
    cleanup();
    double st1 = 0.0, st2 = 0.0;
    for (size_t k = 0; k < nn; ++k) {
        const double t = dbl ? reinterpret_cast<const double*>(th.data())[k] : reinterpret_cast<const float*>(th.data())[k];
        st1 += t * t;
        st2 += (t * t) * (t * t);
    }
    if (err_out) {
        for (uint64_t r = 0; r < R; ++r) {
            const double S = hs[4 * r], Q = hs[4 * r + 1], X = hs[4 * r + 2];
            const double den = st2 / (st1 * st1);
            const double num = Q / (S * S) - 2.0 * X / (S * st1) + den;
            err_out[r] = std::sqrt(std::fmax(num, 0.0) / den);
        }
    }
    for (int f : fl)
        if (f) return OC_NON_FINITE_SAMPLE;
    return OC_OK;
}

}  // namespace

extern "C" {

int oc_gs_fft_run_dev_c64(oc_context* ctx, uint64_t n, uint64_t restarts, const float* source_amp,
                          const float* target_amp, float* phases, uint64_t iterations, uint64_t graph_chunk,
                          double* err_out) {
    OC_CONTEXT_LOCK(ctx);
    return gs_fft_impl(ctx, false, n, restarts, source_amp, target_amp, phases, iterations, graph_chunk, err_out);
}
int oc_gs_fft_run_dev_c128(oc_context* ctx, uint64_t n, uint64_t restarts, const double* source_amp,
                           const double* target_amp, double* phases, uint64_t iterations, uint64_t graph_chunk,
                           double* err_out) {
    OC_CONTEXT_LOCK(ctx);
    return gs_fft_impl(ctx, true, n, restarts, source_amp, target_amp, phases, iterations, graph_chunk, err_out);
}

}  // extern "C"

namespace {

// gs.rs:378-418: the problem checks of gs::run.
int gs_validate(const oc_gs_problem* pb) {
    const size_t n = pb->geom.nx * pb->geom.ny;
    if (n == 0 || !pb->input || !pb->target_amplitude || !pb->merit_target) return OC_ERR_INVALID_ARGUMENT;
    for (size_t k = 0; k < n; ++k) {
        const double t = pb->target_amplitude[k];
        if (!std::isfinite(t)) return OC_INV_NOT_FINITE;
        if (t < 0.0) return OC_INV_OUT_OF_BOUNDS;
    }
    if (!(std::isfinite(pb->z_total) && pb->z_total > 0.0)) return OC_INV_RENDER;
    if (pb->variant == 1 && !(std::isfinite(pb->beta) && pb->beta > 0.0 && pb->beta < 2.0)) return OC_INV_RENDER;
    if (pb->variant == 2 && !(std::isfinite(pb->beta) && pb->beta > 0.0 && pb->beta <= 1.0)) return OC_INV_RENDER;
    if (pb->max_iterations == 0) return OC_INV_RENDER;
    if (pb->init_kind < 0 || pb->init_kind > 2) return OC_ERR_INVALID_ARGUMENT;
    if (pb->init_kind == 0 && !pb->carrier_values) return OC_ERR_INVALID_ARGUMENT;
    if (pb->init_kind == 2 && !pb->init_phase) return OC_ERR_INVALID_ARGUMENT;
    return OC_OK;
}

Prob gs_half(const oc_gs_problem* pb, double z, const void* in, void* out) {
    const oc_asm_config& c = pb->asm_cfg;
    return make_prob(&pb->geom, z, ENG_ASM, c.band_limited, c.strict, c.padding, c.max_padding, c.medium, nullptr, in,
                     out);
}

// gs::run (gs.rs:347-597) for R problems (restarts) on one grid size, run
// together: one persistent launch per pass over all restarts, the merit,
// best-so-far, history and best-phase copy on the device (k_gs_after_fwd,
// k_gs_keep_best, k_gs_after_bwd), the GS sums reduced in a fixed order.
// `chunk` > 1 queues that many iterations at once (captured in a hipGraph
// and replayed) and reads the histories back once per chunk; the patience
// stop is applied to them on the host, iteration by iteration, and a
// restart that stopped inside a chunk is rolled back to the chunk's start
// (phase, best phase and state were saved) and replayed alone up to its
// stop, so every output is that of the iteration-by-iteration loop.
int gs_batch_impl(oc_context* ctx, const oc_gs_problem* pbs, oc_gs_solution* sols, size_t R, uint64_t chunk) {
    if (!ctx || !pbs || !sols || R == 0) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    const size_t nx = pbs[0].geom.nx, ny = pbs[0].geom.ny, n = nx * ny;
    uint64_t max_it = 0;
    for (size_t r = 0; r < R; ++r) {
        if (pbs[r].geom.nx != nx || pbs[r].geom.ny != ny)
            return fail(ctx, OC_ERR_INVALID_ARGUMENT, "the restarts of a batch share one grid size");
        const int v = gs_validate(&pbs[r]);
        if (v) return v;
        max_it = std::max<uint64_t>(max_it, pbs[r].max_iterations);
    }
    Scoped mem(ctx);
    const size_t rn = R * n;
    void *d_in, *d_arr, *d_illum, *d_phase, *d_best, *d_open, *d_tamp, *d_mt, *d_proj, *d_err, *d_dphi, *d_snap,
        *d_state, *d_snap_state, *d_hist, *d_flags, *d_parts_f, *d_parts_b, *d_nparts;
    // partials: at most one row per tile and 16 waves per workgroup
    const size_t pstride = ny * 16 * 4;
    int st;
    if ((st = mem.alloc(rn * 16, &d_in)) || (st = mem.alloc(rn * 16, &d_arr)) || (st = mem.alloc(rn * 8, &d_illum)) ||
        (st = mem.alloc(rn * 8, &d_phase)) || (st = mem.alloc(rn * 8, &d_best)) || (st = mem.alloc(rn, &d_open)) ||
        (st = mem.alloc(rn * 8, &d_tamp)) || (st = mem.alloc(rn * 8, &d_mt)) || (st = mem.alloc(rn * 16, &d_proj)) ||
        (st = mem.alloc(rn * 8, &d_err)) || (st = mem.alloc(rn * 8, &d_dphi)) || (st = mem.alloc(2 * rn * 8, &d_snap)) ||
        (st = mem.alloc(R * sizeof(GsState), &d_state)) || (st = mem.alloc(R * sizeof(GsState), &d_snap_state)) ||
        (st = mem.alloc(R * max_it * sizeof(GsRec), &d_hist)) || (st = mem.alloc(2 * R * sizeof(int), &d_flags)) ||
        (st = mem.alloc(R * pstride * 8, &d_parts_f)) || (st = mem.alloc(R * pstride * 8, &d_parts_b)) ||
        (st = mem.alloc(2 * R * sizeof(int), &d_nparts)))
        return st;
    auto at = [n](void* base, size_t r, size_t esz) { return static_cast<unsigned char*>(base) + r * n * esz; };
    GsState* state = static_cast<GsState*>(d_state);
    GsRec* hist = static_cast<GsRec*>(d_hist);
    int* flags = static_cast<int*>(d_flags);
    // inputs, targets, aperture indicators (gs.rs:427-433), initial phases (gs.rs:447-475)
    std::vector<unsigned char> open(n);
    std::vector<double> phase(n);
    for (size_t r = 0; r < R; ++r) {
        const oc_gs_problem* pb = &pbs[r];
        const oc_band_geom& g = pb->geom;
        HIP_TRY(ctx, hipMemcpy(at(d_in, r, 16), pb->input, n * 16, hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipMemcpy(at(d_tamp, r, 8), pb->target_amplitude, n * 8, hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipMemcpy(at(d_mt, r, 8), pb->merit_target, n * 8, hipMemcpyHostToDevice));
        for (size_t iy = 0; iy < ny; ++iy)
            for (size_t ix = 0; ix < nx; ++ix) {
                const double x = ((double)ix - ((double)nx - 1.0) / 2.0) * g.dx;
                const double y = ((double)iy - ((double)ny - 1.0) / 2.0) * g.dy;
                open[iy * nx + ix] = contains_h(pb->aperture, x - pb->cx, y - pb->cy) ? 1 : 0;
            }
        HIP_TRY(ctx, hipMemcpy(at(d_open, r, 1), open.data(), n, hipMemcpyHostToDevice));
        if (pb->init_kind == 1) {  // GsInit::Random, drawn on the host as the reference does
            och::gs_random_phase(pb->seed, pb->lo, pb->hi, n, phase.data());
        } else {
            const double* src = pb->init_kind == 0 ? pb->carrier_values : pb->init_phase;
            std::copy(src, src + n, phase.begin());
        }
        for (size_t k = 0; k < n; ++k) phase[k] = encode_bounds_h(phase[k], pb->lo, pb->hi);
        HIP_TRY(ctx, hipMemcpy(at(d_phase, r, 8), phase.data(), n * 8, hipMemcpyHostToDevice));
    }
    // illumination arriving at the DOE plane (gs.rs:436-444), every restart at once
    std::vector<std::vector<double>> illum(R, std::vector<double>(n));
    {
        std::vector<Prob> v;
        for (size_t r = 0; r < R; ++r) v.push_back(gs_half(&pbs[r], pbs[r].z_total / 2.0, at(d_in, r, 16), at(d_arr, r, 16)));
        if (run_batch(ctx, true, v) != OC_OK) return OC_INV_RENDER;
        std::vector<double> arr(2 * n);
        for (size_t r = 0; r < R; ++r) {
            HIP_TRY(ctx, hipMemcpy(arr.data(), at(d_arr, r, 16), n * 16, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < n; ++k) illum[r][k] = std::hypot(arr[2 * k], arr[2 * k + 1]);
            HIP_TRY(ctx, hipMemcpy(at(d_illum, r, 8), illum[r].data(), n * 8, hipMemcpyHostToDevice));
        }
    }
    // the two half-transforms of every restart, projections fused
    std::vector<Prob> fw, bw;
    for (size_t r = 0; r < R; ++r) {
        const oc_gs_problem* pb = &pbs[r];
        Prob f = gs_half(pb, pb->z_total / 2.0, nullptr, at(d_proj, r, 16));
        f.ops.pre = PRE_POLAR;
        f.ops.amp = at(d_illum, r, 8);
        f.ops.phase = at(d_phase, r, 8);
        f.ops.open = at(d_open, r, 1);
        f.ops.post = POST_GS_TARGET;
        f.ops.tamp = reinterpret_cast<const double*>(at(d_tamp, r, 8));
        f.ops.merit_target = reinterpret_cast<const double*>(at(d_mt, r, 8));
        f.ops.beta_f = pb->variant == 2 ? pb->beta : 1.0;
        f.ops.err_out = reinterpret_cast<double*>(at(d_err, r, 8));
        f.ops.parts = static_cast<double*>(d_parts_f) + r * pstride;
        Prob b = gs_half(pb, -pb->z_total / 2.0, at(d_proj, r, 16), nullptr);
        b.ops.post = POST_GS_PHASE;
        b.ops.phase_io = at(d_phase, r, 8);
        b.ops.beta_p = pb->variant == 1 ? pb->beta : 1.0;
        b.ops.lo = pb->lo;
        b.ops.hi = pb->hi;
        b.ops.dphi_out = reinterpret_cast<double*>(at(d_dphi, r, 8));
        b.ops.parts = static_cast<double*>(d_parts_b) + r * pstride;
        prepare(f);
        prepare(b);
        if (f.status || b.status || f.check_only || b.check_only || !fits_on_chip(f, true)) return OC_INV_RENDER;
        fw.push_back(f);
        bw.push_back(b);
    }
    Batch bf, bb;
    st = build_batch(ctx, true, fw, flags, bf, true);
    if (!st) st = build_batch(ctx, true, bw, flags + R, bb, true);
    struct Cleanup {
        oc_context* c;
        Batch *a, *b;
        ~Cleanup() {
            free_batch(c, *a);
            free_batch(c, *b);
        }
    } cleanup{ctx, &bf, &bb};
    if (st || bf.pd.size() != R || bb.pd.size() != R) return st ? st : OC_INV_RENDER;
    // fixed-order partial counts: row tiles x waves of the last pass
    std::vector<int> np(2 * R);
    for (size_t r = 0; r < R; ++r) {
        np[r] = bf.pd[r].tiles_row * ((bf.pc[2].block + 63) / 64);
        np[R + r] = bb.pd[r].tiles_row * ((bb.pc[2].block + 63) / 64);
        if ((size_t)np[r] * 4 > pstride || (size_t)np[R + r] * 4 > pstride) return fail(ctx, OC_ERR_DEVICE, "GS partials");
    }
    HIP_TRY(ctx, hipMemcpy(d_nparts, np.data(), np.size() * sizeof(int), hipMemcpyHostToDevice));
    const int* np_f = static_cast<const int*>(d_nparts);
    const int* np_b = np_f + R;
    HIP_TRY(ctx, hipMemsetAsync(d_state, 0, R * sizeof(GsState), ctx->stream));
    HIP_TRY(ctx, ocl::launch_fill_int(flags, 0, (int)(2 * R), ctx->stream));
    double* phase_d = static_cast<double*>(d_phase);
    double* best_d = static_cast<double*>(d_best);
    // one iteration of the restarts of batches (fwd, bwd) whose state,
    // partial counts and history start at restart r0 (`cnt` restarts)
    auto iterate = [&](Batch& f, Batch& b, size_t r0, int cnt) -> int {
        int s2 = launch_batch(ctx, f);
        if (s2) return s2;
        HIP_TRY(ctx, ocl::launch_gs_after_fwd(state + r0, static_cast<double*>(d_parts_f) + r0 * pstride, np_f + r0,
                                              pstride, n, hist + r0 * max_it, max_it, cnt, ctx->stream));
        HIP_TRY(ctx, ocl::launch_gs_keep_best(state + r0, phase_d + r0 * n, best_d + r0 * n, n, cnt, ctx->stream));
        if ((s2 = launch_batch(ctx, b))) return s2;
        HIP_TRY(ctx, ocl::launch_gs_after_bwd(state + r0, static_cast<double*>(d_parts_b) + r0 * pstride, np_b + r0,
                                              pstride, n, hist + r0 * max_it, max_it, cnt, ctx->stream));
        return OC_OK;
    };
    hipGraphExec_t exec = nullptr;
    struct GraphGuard {
        hipGraphExec_t* e;
        ~GraphGuard() {
            if (*e) (void)hipGraphExecDestroy(*e);
        }
    } graph_guard{&exec};
    if (chunk > 1 && max_it >= chunk) {
        const bool was = ctx->timing;
        ctx->timing = false;
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) {
            ctx->timing = was;
            return fail(ctx, OC_ERR_DEVICE, "begin capture");
        }
        for (uint64_t k = 0; k < chunk && st == OC_OK; ++k) st = iterate(bf, bb, 0, (int)R);
        e = hipStreamEndCapture(ctx->stream, &graph);
        ctx->timing = was;
        if (st || e != hipSuccess) return st ? st : fail(ctx, OC_ERR_DEVICE, "end capture");
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(ctx, OC_ERR_DEVICE, "graph instantiate");
    }
    std::vector<char> done(R, 0), conv(R, 0);
    std::vector<uint64_t> stop_at(R, 0);
    std::vector<GsRec> h(R * max_it);
    std::vector<GsState> hs(R);
    size_t live = R;
    // copies a finished restart's outputs (its buffers hold the state of its
    // last iteration)
    auto finalize = [&](size_t r, uint64_t last) -> int {
        const oc_gs_problem* pb = &pbs[r];
        oc_gs_solution* sol = &sols[r];
        HIP_TRY(ctx, hipMemcpyAsync(h.data() + r * max_it, hist + r * max_it, (last + 1) * sizeof(GsRec),
                                    hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(&hs[r], state + r, sizeof(GsState), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(sol->phase, at(d_best, r, 8), n * 8, hipMemcpyDeviceToHost, ctx->stream));
        if (sol->target_amplitude_error)
            HIP_TRY(ctx, hipMemcpyAsync(sol->target_amplitude_error, at(d_err, r, 8), n * 8, hipMemcpyDeviceToHost,
                                        ctx->stream));
        if (sol->phase_error)
            HIP_TRY(ctx, hipMemcpyAsync(sol->phase_error, at(d_dphi, r, 8), n * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        const GsRec* hr = h.data() + r * max_it;
        if (sol->history) std::memcpy(sol->history, hr, (last + 1) * sizeof(oc_gs_record));
        if (sol->illumination_amplitude) std::memcpy(sol->illumination_amplitude, illum[r].data(), n * 8);
        sol->merit = hs[r].best_merit;
        sol->best_iteration = hs[r].best_it;
        sol->amplitude_mse = hr[hs[r].best_it].amplitude_mse;
        sol->converged = conv[r];
        sol->iterations = last + 1;
        if (sol->target_intensity) {  // |U_t|^2 of the best phase (the render of its iteration)
            std::vector<Prob> v{gs_half(pb, pb->z_total / 2.0, nullptr, at(d_arr, r, 16))};
            v[0].ops.pre = PRE_POLAR;
            v[0].ops.amp = at(d_illum, r, 8);
            v[0].ops.phase = at(d_best, r, 8);
            v[0].ops.open = at(d_open, r, 1);
            if (run_batch(ctx, true, v) != OC_OK) return OC_INV_RENDER;
            std::vector<double> arr(2 * n);
            HIP_TRY(ctx, hipMemcpy(arr.data(), at(d_arr, r, 16), n * 16, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < n; ++k) {
                const double a = std::hypot(arr[2 * k], arr[2 * k + 1]);
                sol->target_intensity[k] = a * a;
            }
        }
        done[r] = 1;
        --live;
        return OC_OK;
    };
    uint64_t it0 = 0;
    while (live) {
        const uint64_t c = (exec && it0 + chunk <= max_it) ? chunk : 1;
        if (c > 1) {  // the chunk's starting point, for restarts that stop inside it
            HIP_TRY(ctx, hipMemcpyAsync(d_snap, d_phase, rn * 8, hipMemcpyDeviceToDevice, ctx->stream));
            HIP_TRY(ctx, hipMemcpyAsync(static_cast<unsigned char*>(d_snap) + rn * 8, d_best, rn * 8,
                                        hipMemcpyDeviceToDevice, ctx->stream));
            HIP_TRY(ctx, hipMemcpyAsync(d_snap_state, d_state, R * sizeof(GsState), hipMemcpyDeviceToDevice,
                                        ctx->stream));
            const hipError_t e = timed(ctx, 4, [&] { return hipGraphLaunch(exec, ctx->stream); });
            if (e != hipSuccess) return fail(ctx, OC_ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e));
        } else if ((st = iterate(bf, bb, 0, (int)R))) {
            return st;
        }
        std::vector<int> hflags(2 * R);
        for (size_t r = 0; r < R; ++r)
            if (!done[r])
                HIP_TRY(ctx, hipMemcpyAsync(h.data() + r * max_it + it0, hist + r * max_it + it0, c * sizeof(GsRec),
                                            hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(hflags.data(), flags, 2 * R * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        collect_timing(ctx);
        // a non-finite sample in a propagation (gs.rs:488, 554); restarts that
        // already stopped keep running in the batched launches, their flags
        // are not theirs to raise any more
        for (size_t r = 0; r < R; ++r)
            if (!done[r] && (hflags[r] || hflags[R + r])) return OC_INV_RENDER;
        for (size_t r = 0; r < R; ++r) {
            if (done[r]) continue;
            const oc_gs_problem* pb = &pbs[r];
            const GsRec* hr = h.data() + r * max_it;
            // the loop's exit test, iteration by iteration (gs.rs:482, 570-577)
            uint64_t last = 0;
            bool stop = false;
            for (uint64_t it = it0; it < it0 + c && !stop; ++it) {
                if (it + 1 > pb->patience) {
                    const double earlier = hr[it - pb->patience].best_merit;
                    const double scale = std::fmax(std::fabs(earlier), 1e-30);
                    if ((earlier - hr[it].best_merit) / scale < pb->merit_tolerance) {
                        conv[r] = 1;
                        stop = true;
                    }
                }
                if (it + 1 == pb->max_iterations) stop = true;
                last = it;
            }
            if (!stop) continue;
            if (last + 1 < it0 + c) {
                // roll this restart back to the chunk's start and run it alone to its stop
                HIP_TRY(ctx, hipMemcpyAsync(at(d_phase, r, 8), static_cast<unsigned char*>(d_snap) + r * n * 8, n * 8,
                                            hipMemcpyDeviceToDevice, ctx->stream));
                HIP_TRY(ctx, hipMemcpyAsync(at(d_best, r, 8), static_cast<unsigned char*>(d_snap) + (rn + r * n) * 8,
                                            n * 8, hipMemcpyDeviceToDevice, ctx->stream));
                HIP_TRY(ctx, hipMemcpyAsync(state + r, static_cast<GsState*>(d_snap_state) + r, sizeof(GsState),
                                            hipMemcpyDeviceToDevice, ctx->stream));
                std::vector<int> one{(int)r};
                Batch sf, sb;
                st = build_batch(ctx, true, fw, flags, sf, false, &one, "gsf");
                if (!st) st = build_batch(ctx, true, bw, flags + R, sb, false, &one, "gsb");
                for (uint64_t it = it0; !st && it <= last; ++it) st = iterate(sf, sb, r, 1);
                if (st) return st;
            }
            if ((st = finalize(r, last))) return st;
        }
        it0 += c;
    }
    return OC_OK;
}

}  // namespace

extern "C" {

// gs.rs:347-597
int oc_gs_run(oc_context* ctx, const oc_gs_problem* pb, oc_gs_solution* sol) {
    OC_CONTEXT_LOCK(ctx);
    return gs_batch_impl(ctx, pb, sol, 1, 0);
}

int oc_gs_run_batch(oc_context* ctx, const oc_gs_problem* problems, oc_gs_solution* solutions, size_t count,
                    uint64_t graph_chunk) {
    OC_CONTEXT_LOCK(ctx);
    return gs_batch_impl(ctx, problems, solutions, count, graph_chunk);
}

int oc_adjoint_asm(oc_context* ctx, const oc_band_geom* g, const double* in, double z, const oc_asm_config* cfg,
                   double* out, oc_asm_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    int st = oc_asm_propagate(ctx, g, in, -z, cfg, out, rep);
    return st == OC_OK ? OC_OK : OC_INV_ADJOINT;
}

int oc_adjoint_transmission(oc_context* ctx, const oc_band_geom* g, const oc_phase_mask* m, double* t_out) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !m || !t_out) return OC_ERR_INVALID_ARGUMENT;
    const size_t n = g->nx * g->ny;
    std::vector<double> ones(2 * n, 0.0);
    for (size_t k = 0; k < n; ++k) ones[2 * k] = 1.0;
    int st = oc_phasemask_apply(ctx, g, ones.data(), m, t_out);
    return st == OC_OK ? OC_OK : OC_INV_ADJOINT;
}

}  // extern "C"

namespace {

// The frames an adjoint run traces and pushes back: frame f has wavelength
// lam[f], design steps step(f, s) (every frame's design has frame 0's
// structure: element presence and zero gaps) and input field input(f)
// (device).  Frames are (band, scene) pairs, f = b * ns + s.
struct AdjFrames {
    size_t F = 1, nb = 1, ns = 1;
    std::vector<double> lam;
    std::function<const oc_design_step&(size_t, size_t)> step;
    std::function<const void*(size_t)> input;
    int sel_band = -1, sel_scene = -1;  // MseToTarget selector (-1 = all)
};

// Forward tape + reverse sweep on device (adjoint.rs:140-246) for every
// frame at once: each propagation step is one batch over the frames.  Mask
// values are device pointers.  `dmdi` (device, f64, one frame) seeds
// w = dmdi * U_det; when null, the MseToTarget seeds of the selected frames
// (merit.rs:143-165) are formed from `target` and the merit of their summed
// intensity (merit.rs:124-137) is returned in *merit.  grads: n_elements * n
// (device); several frames are pushed back into per-frame planes that are
// then added in frame order (adjoint.rs:310-332).  May be null for a forward
// render + merit only.  fwd_err: the status of a failing forward trace
// (InverseError::Render for the renderer's trace, ::Adjoint for push_back's
// tape).
int adjoint_impl(oc_context* ctx, bool dbl, const oc_band_geom& g, const AdjFrames& fr, size_t nsteps,
                 const oc_asm_config& cfg, const double* dmdi, const double* target, double* grads, double* merit,
                 int fwd_err) {
    const size_t n = g.nx * g.ny, F = fr.F, esz = dbl ? 16 : 8, fb = n * esz;
    size_t ne = 0;
    for (size_t s = 0; s < nsteps; ++s) ne += fr.step(0, s).has_element ? 1 : 0;
    for (size_t f = 1; f < F; ++f)
        for (size_t s = 0; s < nsteps; ++s) {
            const oc_design_step &d = fr.step(f, s), &d0 = fr.step(0, s);
            if (d.has_element != d0.has_element || (d.z_after == 0.0) != (d0.z_after == 0.0))
                return fail(ctx, OC_ERR_UNSUPPORTED, "the bands' designs differ in structure");
        }
    // the context's cached buffers (a gradient per call in an inverse loop
    // must not pay an allocation and a free)
    void *cur, *nxt, *ckpt = nullptr, *sums, *flags, *slab = nullptr, *fsbuf;
    // the flags, then the merit sums, in one buffer, so that one copy brings
    // back the flags and the merit
    const size_t flag_bytes = (sizeof(int) * (2 * nsteps + 1) * F + 63) / 64 * 64;
    int st;
    if ((st = ensure_buf(ctx, "adj_cur", F * fb, &cur)) || (st = ensure_buf(ctx, "adj_nxt", F * fb, &nxt)) ||
        (st = ensure_buf(ctx, "adj_flags_sums", flag_bytes + sizeof(double) * (8 + ocd::kSeedParts), &fsbuf)))
        return st;
    flags = fsbuf;
    sums = static_cast<unsigned char*>(fsbuf) + flag_bytes;
    if (ne && (st = ensure_buf(ctx, "adj_ckpt", ne * F * fb, &ckpt))) return st;
    if (ne && grads && F > 1 && (st = ensure_buf(ctx, "adj_slab", sizeof(double) * F * ne * n, &slab))) return st;
    // A field set: frame f's field at base + f * fb, or the inputs (base null).
    auto at = [&](const unsigned char* base, size_t f) -> const void* { return base ? base + f * fb : fr.input(f); };
    auto geom_f = [&](size_t f) {
        oc_band_geom gi = g;
        gi.lambda = fr.lam[f];
        return gi;
    };
    // Every step is queued without waiting; the finiteness flags of all
    // propagations (zeroed once here) and the merit are read back once at
    // the end.
    std::vector<std::vector<Prob>> queued;
    queued.reserve(2 * nsteps);
    int* dflags = static_cast<int*>(flags);
    HIP_TRY(ctx, ocl::launch_fill_int(dflags, 0, (int)((2 * nsteps + 1) * F), ctx->stream));
    auto propagate = [&](std::vector<Prob> ps, int err) -> int {
        queued.push_back(std::move(ps));
        int s2 = launch_problems(ctx, dbl, queued.back(), dflags + (queued.size() - 1) * F, false);
        if (s2 == OC_ERR_DEVICE || s2 == OC_ERR_OUT_OF_MEMORY) return s2;
        for (const Prob& p : queued.back())
            if (s2 != OC_OK || p.status != OC_OK) return fail(ctx, err, "propagation failed");
        return OC_OK;
    };
    // forward, the mask fused into the propagation.  The fields arriving at
    // element e are its checkpoints: the inputs themselves for a first-step
    // element, else the previous step writes them straight into slot e.
    // Other intermediates alternate between two scratch buffers.
    unsigned char* const sa = static_cast<unsigned char*>(cur);
    unsigned char* const sb = static_cast<unsigned char*>(nxt);
    std::vector<const unsigned char*> ck_at(ne, nullptr);
    const unsigned char* src = nullptr;  // the inputs
    size_t e = 0;
    // consecutive forward propagations are launched together, linked
    // through their rows when they qualify (each checkpoint still written)
    std::vector<std::vector<Prob>> run;
    auto flush = [&]() -> int {
        if (run.empty()) return OC_OK;
        std::vector<int> off;
        for (size_t k = 0; k < run.size(); ++k) off.push_back((int)((queued.size() + k) * F));
        int sl = run.size() > 1 ? launch_linked(ctx, dbl, run, dflags, off, queued, true) : 1;
        if (sl < 0) return -sl;
        if (sl > 0)
            for (auto& r : run)
                if (int s2 = propagate(r, fwd_err)) return s2;
        for (size_t k = queued.size() - run.size(); k < queued.size(); ++k)
            for (const Prob& p : queued[k])
                if (p.status != OC_OK) return fail(ctx, fwd_err, "forward propagation failed");
        run.clear();
        return OC_OK;
    };
    for (size_t s = 0; s < nsteps; ++s) {
        const oc_design_step& d0 = fr.step(0, s);
        if (d0.has_element) ck_at[e++] = src;
        const bool next_elem = s + 1 < nsteps && fr.step(0, s + 1).has_element;
        unsigned char* dst = next_elem ? static_cast<unsigned char*>(ckpt) + e * F * fb : (src == sa ? sb : sa);
        if (d0.z_after != 0.0) {
            std::vector<Prob> ps;
            for (size_t f = 0; f < F; ++f) {
                const oc_design_step& d = fr.step(f, s);
                const oc_band_geom gi = geom_f(f);
                ps.push_back(make_prob(&gi, d.z_after, ENG_ASM, cfg.band_limited, cfg.strict, cfg.padding,
                                       cfg.max_padding, d.medium, d.has_element ? &d.mask : nullptr, at(src, f),
                                       dst + f * fb));
            }
            run.push_back(std::move(ps));
            src = dst;
            continue;
        }
        if ((st = flush())) return st;
        if (d0.has_element) {
            for (size_t f = 0; f < F; ++f) {
                const oc_design_step& d = fr.step(f, s);
                if (mask_footprint(g, d.mask)) return fail(ctx, fwd_err, "phase mask does not fit the grid");
                MaskDev md;
                if ((st = mask_tables(ctx, g, d.mask, d.mask.values, &md))) return st;
                HIP_TRY(ctx, dbl ? ocl::launch_mask_apply_c128(at(src, f), dst + f * fb, md, (int)g.nx, (int)g.ny,
                                                               nullptr, ctx->stream)
                                 : ocl::launch_mask_apply_c64(at(src, f), dst + f * fb, md, (int)g.nx, (int)g.ny,
                                                              nullptr, ctx->stream));
            }
            src = dst;
        } else if (next_elem) {  // an empty step before an element: its field is that element's checkpoint
            for (size_t f = 0; f < F; ++f)
                HIP_TRY(ctx, hipMemcpyAsync(dst + f * fb, at(src, f), fb, hipMemcpyDeviceToDevice, ctx->stream));
            src = dst;
        }
    }
    if ((st = flush())) return st;
    // the detector fields, contiguous in a scratch buffer
    const unsigned char* det = src;
    if (!det) {
        for (size_t f = 0; f < F; ++f)
            HIP_TRY(ctx, hipMemcpyAsync(sa + f * fb, fr.input(f), fb, hipMemcpyDeviceToDevice, ctx->stream));
        det = sa;
    }
    unsigned char* co = det == sa ? sb : sa;
    // seed at the detector
    if (dmdi) {
        HIP_TRY(ctx, hipMemsetAsync(sums, 0, 64, ctx->stream));
        HIP_TRY(ctx, ocl::launch_seed(dbl, det, dmdi, co, n, ctx->stream));
    } else if (F == 1 && (fr.sel_band == -1 || fr.sel_band == 0) && (fr.sel_scene == -1 || fr.sel_scene == 0)) {
        HIP_TRY(ctx, ocl::launch_seed_mse(dbl, det, target, co, n, static_cast<double*>(sums), ctx->stream));
    } else {
        HIP_TRY(ctx, ocl::launch_seed_frames(dbl, det, target, co, n, (int)fr.nb, (int)fr.ns, fr.sel_band,
                                             fr.sel_scene, static_cast<double*>(sums), ctx->stream));
    }
    // the flags and the merit sums come back through the context's pinned
    // read-back buffer: one asynchronous copy of the flags and the first
    // four sums (adjacent in the device buffer), and one wait
    auto finish = [&]() -> int {
        const size_t nf = queued.size() * F;
        void* rb = nullptr;
        if (int s2 = ensure_readback(ctx, flag_bytes + 32, &rb)) return s2;
        const int* hf = static_cast<const int*>(rb);
        const double* hs = reinterpret_cast<const double*>(static_cast<unsigned char*>(rb) + flag_bytes);
        HIP_TRY(ctx, hipMemcpyAsync(rb, flags, flag_bytes + 32, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        collect_timing(ctx);
        if (merit) *merit = hs[0] / (double)n;
        for (size_t i = 0; i < nf; ++i)
            if (hf[i]) return fail(ctx, fwd_err, "non-finite samples in a propagation");
        return OC_OK;
    };
    if (!grads) return finish();
    // every element plane (per frame with a slab) is written exactly once by
    // the reverse sweep, as 0 + its term (the reference's zero-initialised
    // sum of one term), so the planes need no clearing first
    double* const gbase = slab ? static_cast<double*>(slab) : grads;
    auto gplane = [&](size_t f, size_t el) { return gbase + ((slab ? f * ne : 0) + el) * n; };
    // reverse sweep
    unsigned char* c0 = co;
    unsigned char* c1 = co == sa ? sb : sa;
    e = ne;
    for (size_t s = nsteps; s-- > 0;) {
        const oc_design_step& d0 = fr.step(0, s);
        if (d0.has_element) --e;
        if (d0.z_after != 0.0) {
            std::vector<Prob> ps;
            for (size_t f = 0; f < F; ++f) {
                const oc_design_step& d = fr.step(f, s);
                const oc_band_geom gi = geom_f(f);
                Prob pb = make_prob(&gi, -d.z_after, ENG_ASM, cfg.band_limited, cfg.strict, cfg.padding,
                                    cfg.max_padding, d.medium, nullptr, c0 + f * fb, c1 + f * fb);
                if (d.has_element) {
                    pb.mask = &d.mask;  // mask tables for the epilogue (no pre-op)
                    pb.ops.pre = PRE_NONE;
                    pb.ops.post = POST_ADJ_GRAD;
                    pb.ops.ckpt = at(ck_at[e], f);
                    pb.ops.grad = gplane(f, e);
                }
                ps.push_back(pb);
            }
            if (int s2 = propagate(std::move(ps), OC_INV_ADJOINT)) return s2;
            std::swap(c0, c1);
        } else if (d0.has_element) {
            for (size_t f = 0; f < F; ++f) {
                const oc_design_step& d = fr.step(f, s);
                MaskDev md;
                if ((st = mask_tables(ctx, g, d.mask, d.mask.values, &md))) return st;
                HIP_TRY(ctx, ocl::launch_adj_elem(dbl, c0 + f * fb, at(ck_at[e], f), gplane(f, e), md, (int)g.nx,
                                                  (int)g.ny, ctx->stream));
            }
        }
    }
    if (slab && ne) HIP_TRY(ctx, ocl::launch_frame_sum(static_cast<double*>(slab), grads, ne * n, (int)F, ctx->stream));
    return finish();
}

// One frame (band and scene) of geometry g.
AdjFrames one_frame(const oc_band_geom& g, const oc_design_step* steps, const void* input) {
    AdjFrames fr;
    fr.lam = {g.lambda};
    fr.step = [steps](size_t, size_t s) -> const oc_design_step& { return steps[s]; };
    fr.input = [input](size_t) { return input; };
    return fr;
}

int grid_gradient_frames(oc_context* ctx, bool dbl, const oc_band_geom* g, const double* bands, size_t nb,
                         size_t ns, const oc_design_step* designs, size_t nsteps, const oc_asm_config* cfg,
                         const void* inputs, const double* target, int32_t sel_band, int32_t sel_scene,
                         double* dphi_out, double* merit_out) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg || !target || (!designs && nsteps) || (nb && ns && (!bands || !inputs)))
        return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    // observe.rs:214-218 (Renderer::render), then the frame length
    if (nb == 0 || ns == 0) return fail(ctx, OC_INV_RENDER, "renderer needs at least one band and one scene");
    const size_t n = g->nx * g->ny;
    if (n == 0) return fail(ctx, OC_INV_RENDER, "empty grid");
    const size_t esz = dbl ? 16 : 8;
    AdjFrames fr;
    fr.F = nb * ns;
    fr.nb = nb;
    fr.ns = ns;
    for (size_t f = 0; f < fr.F; ++f) fr.lam.push_back(bands[f / ns]);
    fr.step = [designs, nsteps, ns](size_t f, size_t s) -> const oc_design_step& {
        return designs[(f / ns) * nsteps + s];
    };
    fr.input = [inputs, n, esz](size_t f) -> const void* {
        return static_cast<const unsigned char*>(inputs) + f * n * esz;
    };
    // a selector that picks no frame: obs_gradient seeds nothing, so no
    // plane (adjoint.rs:297-304) and all-zero gradients; evaluate would fail
    const bool in_range = sel_band >= -1 && sel_scene >= -1 && (sel_band < 0 || (size_t)sel_band < nb) &&
                          (sel_scene < 0 || (size_t)sel_scene < ns);
    fr.sel_band = in_range ? sel_band : INT_MAX;  // matches no frame
    fr.sel_scene = in_range ? sel_scene : INT_MAX;
    double m = 0.0;
    int st = adjoint_impl(ctx, dbl, *g, fr, nsteps, *cfg, nullptr, target, dphi_out, &m, OC_INV_RENDER);
    if (merit_out) *merit_out = in_range ? m : std::nan("");
    return st;
}

}  // namespace

extern "C" {

int oc_adjoint_push_back(oc_context* ctx, const oc_design_step* steps, size_t nsteps, const oc_band_geom* g,
                         const double* input, const oc_asm_config* cfg, const double* dmdi, double* dphi_out) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg || !input || !dmdi || (!steps && nsteps)) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    const size_t n = g->nx * g->ny;
    if (n == 0) return OC_INV_ADJOINT;
    size_t ne = 0;
    for (size_t s = 0; s < nsteps; ++s) ne += steps[s].has_element ? 1 : 0;
    Scoped mem(ctx);
    void *din, *ddm, *dgr = nullptr;
    int st;
    if ((st = mem.alloc(n * 16, &din)) || (st = mem.alloc(n * 8, &ddm))) return st;
    if (ne && (st = mem.alloc(ne * n * 8, &dgr))) return st;
    HIP_TRY(ctx, hipMemcpy(din, input, n * 16, hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(ddm, dmdi, n * 8, hipMemcpyHostToDevice));
    std::vector<oc_design_step> ds(steps, steps + nsteps);
    for (auto& s : ds) {
        if (!s.has_element) continue;
        int vs = oc_phasemask_validate(&s.mask);
        if (vs) return OC_INV_ADJOINT;
        void* mv;
        const size_t mn = s.mask.nx * s.mask.ny;
        if ((st = mem.alloc(mn * 8, &mv))) return st;
        HIP_TRY(ctx, hipMemcpy(mv, s.mask.values, mn * 8, hipMemcpyHostToDevice));
        s.mask.values = static_cast<const double*>(mv);
    }
    st = adjoint_impl(ctx, true, *g, one_frame(*g, ds.data(), din), nsteps, *cfg, static_cast<double*>(ddm), nullptr,
                      static_cast<double*>(dgr), nullptr, OC_INV_ADJOINT);
    if (st) return st;
    if (ne) HIP_TRY(ctx, hipMemcpy(dphi_out, dgr, ne * n * 8, hipMemcpyDeviceToHost));
    return OC_OK;
}

int oc_adjoint_grid_gradient_dev_c128(oc_context* ctx, const oc_design_step* steps, size_t nsteps,
                                      const oc_band_geom* g, const double* input, const oc_asm_config* cfg,
                                      const double* target, double* dphi_out, double* merit_out) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg || !input || !target || (!steps && nsteps)) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    if (g->nx * g->ny == 0) return OC_INV_ADJOINT;
    double m = 0.0;
    int st = adjoint_impl(ctx, true, *g, one_frame(*g, steps, input), nsteps, *cfg, nullptr, target, dphi_out, &m,
                          OC_INV_RENDER);
    if (merit_out) *merit_out = m;
    return st;
}

int oc_adjoint_grid_gradient_frames_dev_c128(oc_context* ctx, const oc_band_geom* geom, const double* bands,
                                             size_t n_bands, size_t n_scenes, const oc_design_step* designs,
                                             size_t nsteps, const oc_asm_config* cfg, const void* inputs,
                                             const double* target, int32_t sel_band, int32_t sel_scene,
                                             double* dphi_out, double* merit_out) {
    return grid_gradient_frames(ctx, true, geom, bands, n_bands, n_scenes, designs, nsteps, cfg, inputs, target,
                                sel_band, sel_scene, dphi_out, merit_out);
}

int oc_adjoint_grid_gradient_frames_dev_c64(oc_context* ctx, const oc_band_geom* geom, const double* bands,
                                            size_t n_bands, size_t n_scenes, const oc_design_step* designs,
                                            size_t nsteps, const oc_asm_config* cfg, const void* inputs,
                                            const double* target, int32_t sel_band, int32_t sel_scene,
                                            double* dphi_out, double* merit_out) {
    return grid_gradient_frames(ctx, false, geom, bands, n_bands, n_scenes, designs, nsteps, cfg, inputs, target,
                                sel_band, sel_scene, dphi_out, merit_out);
}

}  // extern "C"
