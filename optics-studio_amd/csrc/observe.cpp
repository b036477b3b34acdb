// Renderer::render, the batched FD gradient and the Adam epoch.
#include "runtime.hpp"

using namespace ocr;

namespace {

// observe.rs:214-231, before any buffer is read.
int render_check(oc_context* ctx, const oc_band_geom& g, size_t nb, size_t ns, const oc_render_config& rc) {
    if (nb == 0 || ns == 0) return fail(ctx, OC_INV_RENDER, "renderer needs at least one band and one scene");
    if ((rc.deriv_x || rc.deriv_y) && (!std::isfinite(rc.scene_step) || rc.scene_step <= 0.0))
        return fail(ctx, OC_INV_RENDER, "scene_step must be positive and finite");
    if (rc.deriv_lambda && (!std::isfinite(rc.lambda_step) || rc.lambda_step <= 0.0))
        return fail(ctx, OC_INV_RENDER, "lambda_step must be positive and finite");
    if (g.nx * g.ny == 0) return fail(ctx, OC_INV_RENDER, "empty grid");
    return OC_OK;
}

// Beamline::trace (beamline.rs:109-130) of many renders at once.  Render
// i starts from input(i), has wavelength lam[i] and design steps step(i, s);
// every design has the structure of step(0, .) (element presence and zero
// gaps; distances and media may differ), so each propagation step is one
// batch over all renders.  A mask with z_after = 0 is carried into the next
// propagation's pre-op (apply, then propagate: the same as apply; identity;
// apply nothing; propagate), or applied on its own where none follows.  The
// final fields land in a contiguous [nrend][n] buffer (*fields).  Nothing
// waits: the propagations are left in `queued` with their flags at dflags
// (nrend per propagation step) for finish_problems.
struct TraceSet {
    size_t nrend;
    const double* lam;
    std::function<const oc_design_step&(size_t, size_t)> step;
    std::function<const void*(size_t)> input;
};

int trace_batch(oc_context* ctx, bool dbl, const oc_band_geom& g, const TraceSet& ts, size_t nsteps,
                const oc_asm_config& cfg, int* dflags, std::vector<std::vector<Prob>>& queued,
                const unsigned char** fields) {
    const size_t n = g.nx * g.ny, nrend = ts.nrend, esz = dbl ? 16 : 8;
    for (size_t i = 0; i < nrend; ++i)
        for (size_t s = 0; s < nsteps; ++s) {
            const oc_design_step& d = ts.step(i, s);
            const oc_design_step& d0 = ts.step(0, s);
            if (d.has_element != d0.has_element || (d.z_after == 0.0) != (d0.z_after == 0.0))
                return fail(ctx, OC_ERR_UNSUPPORTED, "designs differ in structure across renders");
            if (d.has_element && mask_footprint(g, d.mask) != OC_OK)
                return fail(ctx, OC_INV_RENDER, "phase mask does not fit the grid");
        }
    void *bufa, *bufb;
    int st;
    if ((st = ensure_buf(ctx, "trace_a", nrend * n * esz, &bufa)) ||
        (st = ensure_buf(ctx, "trace_b", nrend * n * esz, &bufb)))
        return st;
    const unsigned char* cur = nullptr;  // nullptr: the inputs
    unsigned char* nxt = static_cast<unsigned char*>(bufa);
    auto src = [&](size_t i) -> const void* { return cur ? cur + i * n * esz : ts.input(i); };
    auto advance = [&]() {
        cur = nxt;
        nxt = nxt == static_cast<unsigned char*>(bufa) ? static_cast<unsigned char*>(bufb)
                                                       : static_cast<unsigned char*>(bufa);
    };
    queued.clear();
    queued.reserve(nsteps);
    int nflag = 0;
    long pend = -1;  // step whose mask is pending
    // consecutive propagations not yet launched (a run): launched linked
    // when they qualify, else one batch each
    std::vector<std::vector<Prob>> run;
    std::vector<int> run_off;
    auto launch_run = [&]() -> int {
        int s2 = run.size() > 1 ? launch_linked(ctx, dbl, run, dflags, run_off, queued, false) : 1;
        if (s2 < 0) return -s2;  // a device error
        if (s2 > 0) {            // one batch per step
            for (size_t k = 0; k < run.size(); ++k) {
                queued.push_back(std::move(run[k]));
                s2 = launch_problems(ctx, dbl, queued.back(), dflags + run_off[k]);
                if (s2) return s2;
                for (const Prob& p : queued.back())
                    if (p.status != OC_OK)
                        return fail(ctx, OC_INV_RENDER, std::string("propagation: ") + oc_status_name(p.status));
            }
        }
        run.clear();
        run_off.clear();
        return OC_OK;
    };
    auto flush = [&]() -> int {
        if (pend < 0) return OC_OK;
        if (int s3 = launch_run()) return s3;
        for (size_t i = 0; i < nrend; ++i) {
            const oc_design_step& d = ts.step(i, (size_t)pend);
            MaskDev md;
            int s2 = mask_tables(ctx, g, d.mask, d.mask.values, &md);
            if (s2) return s2;
            const hipError_t e = dbl ? ocl::launch_mask_apply_c128(src(i), nxt + i * n * esz, md, (int)g.nx,
                                                                   (int)g.ny, nullptr, ctx->stream)
                                     : ocl::launch_mask_apply_c64(src(i), nxt + i * n * esz, md, (int)g.nx,
                                                                  (int)g.ny, nullptr, ctx->stream);
            if (e != hipSuccess) return fail(ctx, OC_ERR_DEVICE, std::string("mask apply: ") + hipGetErrorString(e));
        }
        advance();
        pend = -1;
        return OC_OK;
    };
    for (size_t s = 0; s < nsteps; ++s) {
        const oc_design_step& d0 = ts.step(0, s);
        if (d0.has_element) {
            if ((st = flush())) return st;
            pend = (long)s;
        }
        if (d0.z_after == 0.0) continue;
        std::vector<Prob> probs;
        probs.reserve(nrend);
        for (size_t i = 0; i < nrend; ++i) {
            const oc_design_step& d = ts.step(i, s);
            const oc_design_step* dm = pend >= 0 ? &ts.step(i, (size_t)pend) : nullptr;
            oc_band_geom gi = g;
            gi.lambda = ts.lam[i];
            probs.push_back(make_prob(&gi, d.z_after, ENG_ASM, cfg.band_limited, cfg.strict, cfg.padding,
                                      cfg.max_padding, d.medium, dm ? &dm->mask : nullptr, src(i),
                                      nxt + i * n * esz));
        }
        run.push_back(std::move(probs));
        run_off.push_back(nflag);
        nflag += (int)nrend;
        pend = -1;
        advance();
    }
    if ((st = launch_run())) return st;
    if ((st = flush())) return st;
    if (!cur) {  // nothing applied: the fields are the inputs
        for (size_t i = 0; i < nrend; ++i)
            HIP_TRY(ctx, ocl::launch_copy(dbl, ts.input(i), nxt + i * n * esz, n, ctx->stream));
        advance();
    }
    *fields = cur;
    return OC_OK;
}

// Waits for the traced propagations and maps a failed one to OC_INV_RENDER.
int finish_traces(oc_context* ctx, std::vector<std::vector<Prob>>& queued, const int* dflags) {
    int off = 0;
    for (auto& q : queued) {
        const int st = finish_problems(ctx, q, dflags + off);
        off += (int)q.size();
        if (st) return fail(ctx, OC_INV_RENDER, std::string("propagation: ") + oc_status_name(st));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OC_OK;
}

// Renderer::render (observe.rs:213-314) over PhaseMask designs: every
// render of every (band, scene) through trace_batch, then one kernel
// writes the intensity and FD channels.
int render_impl(oc_context* ctx, bool dbl, const oc_band_geom& g, const double* bands, size_t nb, size_t ns,
                const oc_render_config& rc, const oc_design_step* designs, size_t nsteps, const oc_asm_config& cfg,
                const void* inputs, double* obs) {
    int st = render_check(ctx, g, nb, ns, rc);
    if (st) return st;
    const size_t n = g.nx * g.ny;
    const int nd = (rc.deriv_x ? 1 : 0) + (rc.deriv_y ? 1 : 0) + (rc.deriv_lambda ? 1 : 0);
    const int nr = 1 + 2 * nd;               // renders per frame
    const int nv = rc.deriv_lambda ? 3 : 1;  // designs per band
    const size_t frames = nb * ns, nrend = frames * nr;
    const size_t esz = dbl ? 16 : 8;
    // per render: band, design variant, wavelength
    std::vector<int> rb(nrend), rv(nrend);
    std::vector<double> rl(nrend);
    for (size_t f = 0; f < frames; ++f) {
        const size_t b = f / ns;
        for (int r = 0; r < nr; ++r) {
            const size_t i = f * nr + r;
            rb[i] = (int)b;
            int v = 0;
            if (rc.deriv_lambda && r >= nr - 2) v = r == nr - 2 ? 1 : 2;
            rv[i] = v;
            rl[i] = v == 0 ? bands[b] : (v == 1 ? bands[b] + rc.lambda_step : bands[b] - rc.lambda_step);
        }
    }
    TraceSet ts;
    ts.nrend = nrend;
    ts.lam = rl.data();
    ts.step = [&](size_t i, size_t s) -> const oc_design_step& {
        return designs[((size_t)rb[i] * nv + rv[i]) * nsteps + s];
    };
    ts.input = [&](size_t i) -> const void* { return static_cast<const unsigned char*>(inputs) + i * n * esz; };
    void* fl;
    if ((st = ensure_buf(ctx, "trace_flags", sizeof(int) * (nrend * (nsteps + 1) + 1), &fl))) return st;
    int* dflags = static_cast<int*>(fl);
    int* oflag = dflags + nrend * (nsteps + 1);
    HIP_TRY(ctx, ocl::launch_fill_int(oflag, 0, 1, ctx->stream));
    std::vector<std::vector<Prob>> queued;
    const unsigned char* fields = nullptr;
    if ((st = trace_batch(ctx, dbl, g, ts, nsteps, cfg, dflags, queued, &fields))) return st;
    HIP_TRY(ctx, ocl::launch_render_obs(dbl, fields, obs, n, (int)frames, nr, rc.deriv_x, rc.deriv_y, rc.deriv_lambda,
                                        rc.scene_step, rc.lambda_step, oflag, ctx->stream));
    int of = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&of, oflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if ((st = finish_traces(ctx, queued, dflags))) return st;
    if (of) return fail(ctx, OC_INV_RENDER, "non-finite intensity");
    return OC_OK;
}

// fd::gradient (fd.rs:139-184) with the render + MseToTarget objective:
// the live slots' probes traced as one batch, merits reduced per probe.
int fd_impl(oc_context* ctx, bool dbl, const oc_band_geom& g, const oc_design_step* designs, size_t nsteps, size_t dim,
            const double* up, const double* dn, const double* lo, const double* hi, const oc_asm_config& cfg,
            const void* input, const double* target, double* grad, double* merits, size_t* evals) {
    const size_t n = g.nx * g.ny;
    if (n == 0) return fail(ctx, OC_INV_RENDER, "empty grid");
    std::vector<size_t> live;
    for (size_t i = 0; i < dim; ++i) {
        grad[i] = 0.0;
        if (merits) merits[2 * i] = merits[2 * i + 1] = 0.0;
        if (lo[i] == hi[i]) continue;  // degenerate slot: gradient 0
        const double d = up[i] - dn[i];
        if (d <= 0.0 || std::isnan(d)) return fail(ctx, OC_INV_RENDER, "fd step collapses to zero separation");
        live.push_back(i);
    }
    if (evals) *evals = 2 * live.size();
    if (live.empty()) return OC_OK;
    const size_t nrend = 2 * live.size();
    std::vector<double> lam(nrend, g.lambda);
    TraceSet ts;
    ts.nrend = nrend;
    ts.lam = lam.data();
    ts.step = [&](size_t i, size_t s) -> const oc_design_step& {
        return designs[(live[i / 2] * 2 + (i % 2)) * nsteps + s];
    };
    ts.input = [&](size_t) -> const void* { return input; };
    void *fl, *part, *dm;
    int st;
    if ((st = ensure_buf(ctx, "trace_flags", sizeof(int) * (nrend * (nsteps + 1) + 1), &fl)) ||
        (st = ensure_buf(ctx, "fd_part", sizeof(double) * 64 * nrend, &part)) ||
        (st = ensure_buf(ctx, "fd_merit", sizeof(double) * nrend, &dm)))
        return st;
    int* dflags = static_cast<int*>(fl);
    std::vector<std::vector<Prob>> queued;
    const unsigned char* fields = nullptr;
    if ((st = trace_batch(ctx, dbl, g, ts, nsteps, cfg, dflags, queued, &fields))) return st;
    HIP_TRY(ctx, ocl::launch_mse(dbl, fields, target, n, (int)nrend, static_cast<double*>(part),
                                 static_cast<double*>(dm), ctx->stream));
    std::vector<double> hm(nrend);
    HIP_TRY(ctx, hipMemcpyAsync(hm.data(), dm, sizeof(double) * nrend, hipMemcpyDeviceToHost, ctx->stream));
    if ((st = finish_traces(ctx, queued, dflags))) return st;
    for (size_t j = 0; j < live.size(); ++j) {
        const size_t i = live[j];
        const double fu = hm[2 * j], fd = hm[2 * j + 1];
        if (merits) {
            merits[2 * i] = fu;
            merits[2 * i + 1] = fd;
        }
        if (!std::isfinite(fu) || !std::isfinite(fd)) return fail(ctx, OC_INV_RENDER, "non-finite merit");
        const double gi = (fu - fd) / (up[i] - dn[i]);
        if (!std::isfinite(gi)) return fail(ctx, OC_INV_RENDER, "non-finite gradient component");
        grad[i] = gi;
    }
    return OC_OK;
}

}  // namespace

namespace {

// analysis::mtf (analysis.rs:133-150) over `count` device-resident PSFs.
int mtf_impl(oc_context* ctx, const double* psfs, size_t count, size_t nx, size_t ny, double dx, double* values,
             double* freqs) {
    if (nx == 0 || ny == 0) return fail(ctx, OC_EMPTY_GRID, "empty PSF grid");
    if (!std::isfinite(dx) || dx <= 0.0) return fail(ctx, OC_INVALID_GRID_PITCH, "PSF pitch must be positive");
    const size_t half = nx / 2;
    if (freqs)
        for (size_t k = 0; k <= half; ++k) freqs[k] = (double)k / ((double)nx * dx);
    if (count == 0) return OC_OK;
    // e^{-2 pi i j / nx}, j = 0..nx-1, from long double
    std::vector<double> tw(2 * nx);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (size_t j = 0; j < nx; ++j) {
        const long double a = two_pi * (long double)j / (long double)nx;
        tw[2 * j] = (double)cosl(a);
        tw[2 * j + 1] = (double)-sinl(a);
    }
    void *dtw, *sums, *fl;
    int st;
    if ((st = ensure_buf(ctx, "mtf_tw", 16 * nx, &dtw)) || (st = ensure_buf(ctx, "mtf_sums", 8 * nx * count, &sums)) ||
        (st = ensure_buf(ctx, "mtf_flag", sizeof(int), &fl)))
        return st;
    HIP_TRY(ctx, hipMemcpyAsync(dtw, tw.data(), 16 * nx, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, ocl::launch_fill_int(static_cast<int*>(fl), 0, 1, ctx->stream));
    HIP_TRY(ctx, ocl::launch_mtf(psfs, nx, ny, count, static_cast<double*>(sums), dtw, values, static_cast<int*>(fl),
                                 ctx->stream));
    int hf = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&hf, fl, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (hf) return fail(ctx, OC_NON_FINITE_SAMPLE, "non-finite PSF sample");
    return OC_OK;
}

}  // namespace

extern "C" {

int oc_fd_gradient_dev_c128(oc_context* ctx, const oc_band_geom* g, const oc_design_step* designs, size_t nsteps,
                            size_t dim, const double* up, const double* dn, const double* lo, const double* hi,
                            const oc_asm_config* cfg, const void* input, const double* target, double* grad_out,
                            double* merits_out, size_t* evaluations_out) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg || (!designs && nsteps && dim) || (dim && (!up || !dn || !lo || !hi || !grad_out)) ||
        !input || !target)
        return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    return fd_impl(ctx, true, *g, designs, nsteps, dim, up, dn, lo, hi, *cfg, input, target, grad_out, merits_out,
                   evaluations_out);
}

int oc_fd_gradient_dev_c64(oc_context* ctx, const oc_band_geom* g, const oc_design_step* designs, size_t nsteps,
                           size_t dim, const double* up, const double* dn, const double* lo, const double* hi,
                           const oc_asm_config* cfg, const void* input, const double* target, double* grad_out,
                           double* merits_out, size_t* evaluations_out) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg || (!designs && nsteps && dim) || (dim && (!up || !dn || !lo || !hi || !grad_out)) ||
        !input || !target)
        return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    return fd_impl(ctx, false, *g, designs, nsteps, dim, up, dn, lo, hi, *cfg, input, target, grad_out, merits_out,
                   evaluations_out);
}

int oc_adam_step_dev(oc_context* ctx, size_t dim, double* x, const double* g, double* m, double* v, const double* lo,
                     const double* hi, uint64_t t, const oc_adam_config* cfg) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !cfg || t == 0 || (dim && (!x || !g || !m || !v || !lo || !hi))) return OC_ERR_INVALID_ARGUMENT;
    // AdamConfig::validate (fd.rs:95-118)
    if (!std::isfinite(cfg->lr_frac) || cfg->lr_frac <= 0.0) return fail(ctx, OC_INV_RENDER, "invalid Adam lr_frac");
    for (double b : {cfg->beta1, cfg->beta2})
        if (!std::isfinite(b) || !(b >= 0.0 && b < 1.0)) return fail(ctx, OC_INV_RENDER, "invalid Adam beta");
    if (!std::isfinite(cfg->eps) || cfg->eps < 0.0) return fail(ctx, OC_INV_RENDER, "invalid Adam eps");
    if (cfg->has_clip && (!std::isfinite(cfg->clip_norm) || cfg->clip_norm <= 0.0))
        return fail(ctx, OC_INV_RENDER, "invalid Adam clip_norm");
    if (dim == 0) return OC_OK;
    (void)hipSetDevice(ctx->device);
    void* part;
    int st = ensure_buf(ctx, "adam_part", sizeof(double) * 72, &part);
    if (st) return st;
    double* scale = static_cast<double*>(part) + 64;
    const double bc1 = 1.0 - std::pow(cfg->beta1, (double)t);
    const double bc2 = 1.0 - std::pow(cfg->beta2, (double)t);
    HIP_TRY(ctx, ocl::launch_adam(x, g, m, v, lo, hi, dim, static_cast<double*>(part), scale, cfg->has_clip,
                                  cfg->clip_norm, cfg->lr_frac, cfg->beta1, cfg->beta2, cfg->eps, bc1, bc2,
                                  ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OC_OK;
}

static size_t render_counts(const oc_render_config* rc, size_t nb, size_t ns, size_t* nv, size_t* nch) {
    const size_t nd = (rc->deriv_x ? 1 : 0) + (rc->deriv_y ? 1 : 0) + (rc->deriv_lambda ? 1 : 0);
    *nv = rc->deriv_lambda ? 3 : 1;
    *nch = 1 + nd;
    return nb * ns * (1 + 2 * nd);
}

int oc_render_dev_c128(oc_context* ctx, const oc_band_geom* g, const double* bands, size_t nb, size_t ns,
                       const oc_render_config* rc, const oc_design_step* designs, size_t nsteps,
                       const oc_asm_config* cfg, const void* inputs, double* obs) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !rc || !cfg) return OC_ERR_INVALID_ARGUMENT;
    int st = render_check(ctx, *g, nb, ns, *rc);
    if (st) return st;
    if (!bands || (!designs && nsteps) || !inputs || !obs) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    return render_impl(ctx, true, *g, bands, nb, ns, *rc, designs, nsteps, *cfg, inputs, obs);
}

int oc_render_dev_c64(oc_context* ctx, const oc_band_geom* g, const double* bands, size_t nb, size_t ns,
                      const oc_render_config* rc, const oc_design_step* designs, size_t nsteps,
                      const oc_asm_config* cfg, const void* inputs, double* obs) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !rc || !cfg) return OC_ERR_INVALID_ARGUMENT;
    int st = render_check(ctx, *g, nb, ns, *rc);
    if (st) return st;
    if (!bands || (!designs && nsteps) || !inputs || !obs) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    return render_impl(ctx, false, *g, bands, nb, ns, *rc, designs, nsteps, *cfg, inputs, obs);
}

int oc_render(oc_context* ctx, const oc_band_geom* g, const double* bands, size_t nb, size_t ns,
              const oc_render_config* rc, const oc_design_step* designs, size_t nsteps, const oc_asm_config* cfg,
              const double* inputs, double* obs) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !rc || !cfg) return OC_ERR_INVALID_ARGUMENT;
    int st = render_check(ctx, *g, nb, ns, *rc);
    if (st) return st;
    if (!bands || (!designs && nsteps) || !inputs || !obs) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    size_t nv, nch;
    const size_t nrend = render_counts(rc, nb, ns, &nv, &nch);
    const size_t n = g->nx * g->ny;
    Scoped mem(ctx);
    void *din, *dobs;
    if ((st = mem.alloc(nrend * n * 16, &din)) || (st = mem.alloc(nch * nb * ns * n * 8, &dobs))) return st;
    HIP_TRY(ctx, hipMemcpy(din, inputs, nrend * n * 16, hipMemcpyHostToDevice));
    std::vector<oc_design_step> ds(designs, designs + nb * nv * nsteps);
    std::map<const double*, const double*> up;
    for (auto& d : ds) {
        if (!d.has_element) continue;
        if (oc_phasemask_validate(&d.mask)) return fail(ctx, OC_INV_RENDER, "invalid phase mask");
        auto it = up.find(d.mask.values);
        if (it == up.end()) {
            void* mv;
            const size_t mn = d.mask.nx * d.mask.ny;
            if ((st = mem.alloc(mn * 8, &mv))) return st;
            HIP_TRY(ctx, hipMemcpy(mv, d.mask.values, mn * 8, hipMemcpyHostToDevice));
            it = up.emplace(d.mask.values, static_cast<const double*>(mv)).first;
        }
        d.mask.values = it->second;
    }
    st = render_impl(ctx, true, *g, bands, nb, ns, *rc, ds.data(), nsteps, *cfg, din, static_cast<double*>(dobs));
    if (st) return st;
    HIP_TRY(ctx, hipMemcpy(obs, dobs, nch * nb * ns * n * 8, hipMemcpyDeviceToHost));
    return OC_OK;
}

int oc_mtf_dev(oc_context* ctx, const double* psfs, size_t count, size_t nx, size_t ny, double dx, double* values,
               double* freqs) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || (count && (!psfs || !values))) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    return mtf_impl(ctx, psfs, count, nx, ny, dx, values, freqs);
}

int oc_mtf(oc_context* ctx, const double* psfs, size_t count, size_t nx, size_t ny, double dx, double* values,
           double* freqs) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || (count && (!psfs || !values))) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (nx == 0 || ny == 0 || count == 0) return mtf_impl(ctx, psfs, count, nx, ny, dx, values, freqs);
    Scoped mem(ctx);
    void *dp, *dv;
    int st;
    if ((st = mem.alloc(8 * nx * ny * count, &dp)) || (st = mem.alloc(8 * (nx / 2 + 1) * count, &dv))) return st;
    HIP_TRY(ctx, hipMemcpy(dp, psfs, 8 * nx * ny * count, hipMemcpyHostToDevice));
    st = mtf_impl(ctx, static_cast<const double*>(dp), count, nx, ny, dx, static_cast<double*>(dv), freqs);
    if (st) return st;
    HIP_TRY(ctx, hipMemcpy(values, dv, 8 * (nx / 2 + 1) * count, hipMemcpyDeviceToHost));
    return OC_OK;
}

}  // extern "C"
