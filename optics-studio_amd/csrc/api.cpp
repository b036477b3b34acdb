// C entry points of the propagation engines, phase masks, batches,
// contexts and the layout autotuner (include/optics_gpu.h).
#include "runtime.hpp"

using namespace ocr;

extern "C" {

const char* oc_version(void) { return "optics-gpu 0.1.0 (gfx950)"; }

const char* oc_status_name(int s) {
    switch (s) {
        case OC_OK: return "OC_OK";
        case OC_INVALID_DISTANCE: return "OC_INVALID_DISTANCE";
        case OC_INVALID_WAVELENGTH: return "OC_INVALID_WAVELENGTH";
        case OC_INVALID_MEDIUM: return "OC_INVALID_MEDIUM";
        case OC_EMPTY_GRID: return "OC_EMPTY_GRID";
        case OC_NON_FINITE_SAMPLE: return "OC_NON_FINITE_SAMPLE";
        case OC_ALIASING_RISK: return "OC_ALIASING_RISK";
        case OC_REGIME_VIOLATION: return "OC_REGIME_VIOLATION";
        case OC_INVALID_APERTURE: return "OC_INVALID_APERTURE";
        case OC_OUTSIDE_FIELD: return "OC_OUTSIDE_FIELD";
        case OC_INVALID_AXIS: return "OC_INVALID_AXIS";
        case OC_INVALID_PHASE: return "OC_INVALID_PHASE";
        case OC_INVALID_ORIENTATION: return "OC_INVALID_ORIENTATION";
        case OC_INVALID_GRID: return "OC_INVALID_GRID";
        case OC_INVALID_GRID_PITCH: return "OC_INVALID_GRID_PITCH";
        case OC_INV_NOT_FINITE: return "OC_INV_NOT_FINITE";
        case OC_INV_OUT_OF_BOUNDS: return "OC_INV_OUT_OF_BOUNDS";
        case OC_INV_WRONG_DIMENSION: return "OC_INV_WRONG_DIMENSION";
        case OC_INV_CARRIER: return "OC_INV_CARRIER";
        case OC_INV_RENDER: return "OC_INV_RENDER";
        case OC_INV_ADJOINT: return "OC_INV_ADJOINT";
        case OC_ERR_INVALID_ARGUMENT: return "OC_ERR_INVALID_ARGUMENT";
        case OC_ERR_UNSUPPORTED: return "OC_ERR_UNSUPPORTED";
        case OC_ERR_DEVICE: return "OC_ERR_DEVICE";
        case OC_ERR_OUT_OF_MEMORY: return "OC_ERR_OUT_OF_MEMORY";
        default: return "OC_UNKNOWN";
    }
}

int oc_context_create(int device, oc_context** out) {
    if (!out) return OC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return OC_ERR_DEVICE;
    if (device < 0 || device >= ndev) return OC_ERR_INVALID_ARGUMENT;
    std::unique_ptr<oc_context> c(new oc_context());
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) return OC_ERR_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) return OC_ERR_DEVICE;
    c->stream = c->own;
    if (ocl::set_lds_limits() != hipSuccess || ocl::big_set_lds_limits() != hipSuccess) return OC_ERR_DEVICE;
    if (ocl::fast_set_lds_limits_c64() != hipSuccess || ocl::fast_set_lds_limits_c128() != hipSuccess)
        return OC_ERR_DEVICE;
    if (noalloc_attach(c.get()) != OC_OK) {
        (void)hipStreamDestroy(c->own);
        return OC_ERR_DEVICE;
    }
    *out = c.release();
    return OC_OK;
}

void oc_context_destroy(oc_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    for (auto& kv : ctx->twiddles) (void)hipFree(kv.second);
    for (auto& kv : ctx->tables) (void)hipFree(kv.second.dev);
    collect_timing(ctx);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    for (int k = 0; k < oc_context::kPinSlots; ++k) {
        if (ctx->pinned_done[k]) (void)hipEventSynchronize(ctx->pinned_done[k]);
        if (ctx->pinned[k]) (void)hipHostFree(ctx->pinned[k]);
        if (ctx->pinned_done[k]) (void)hipEventDestroy(ctx->pinned_done[k]);
    }
    if (ctx->readback) (void)hipHostFree(ctx->readback);
    if (ctx->desc_pool) (void)hipFree(ctx->desc_pool);
    release_queued(ctx);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
}

int oc_context_set_stream(oc_context* ctx, void* stream) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own;
    return OC_OK;
}

void* oc_context_stream(oc_context* ctx) { return ctx ? ctx->stream : nullptr; }

int oc_context_synchronize(oc_context* ctx) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OC_OK;
}

int oc_context_enable_timing(oc_context* ctx, int enable) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    ctx->timing = enable != 0;
    return OC_OK;
}

int oc_context_set_timing_mask(oc_context* ctx, uint32_t mask) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    ctx->timing_mask = mask;
    return OC_OK;
}

int oc_context_pass_timing(oc_context* ctx, double* ms, uint64_t* launches, int reset) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timing(ctx);
    for (int i = 0; i < OC_NUM_PASS_KINDS; ++i) {
        if (ms) ms[i] = ctx->pass_ms[i];
        if (launches) launches[i] = ctx->pass_n[i];
        if (reset) {
            ctx->pass_ms[i] = 0.0;
            ctx->pass_n[i] = 0;
        }
    }
    return OC_OK;
}

const char* oc_last_error(oc_context* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int oc_asm_sampling_report(const oc_band_geom* g, double z, const oc_asm_config* cfg, oc_asm_report* rep) {
    if (!g || !cfg || !rep) return OC_ERR_INVALID_ARGUMENT;
    *rep = och::asm_analyze(*g, z, *cfg, cfg->padding);
    return OC_OK;
}

int oc_fresnel_sampling_report(const oc_band_geom* g, double z, const oc_fresnel_config* cfg,
                               oc_fresnel_report* rep) {
    if (!g || !cfg || !rep) return OC_ERR_INVALID_ARGUMENT;
    *rep = och::fresnel_analyze(*g, z, *cfg, cfg->padding);
    return OC_OK;
}

int oc_rs_sampling_report(const oc_band_geom* g, double z, const oc_rs_config* cfg, oc_rs_report* rep) {
    if (!g || !cfg || !rep) return OC_ERR_INVALID_ARGUMENT;
    *rep = och::rs_analyze(*g, z, cfg->padding);
    return OC_OK;
}

static void fill_asm_report(const Prob& p, oc_asm_report* rep) {
    if (rep) *rep = p.asm_rep;
}

int oc_asm_propagate(oc_context* ctx, const oc_band_geom* g, const double* in, double z, const oc_asm_config* cfg,
                     double* out, oc_asm_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg || (!in && g->nx * g->ny) || (!out && g->nx * g->ny)) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    Prob done;
    Prob p = make_prob(g, z, ENG_ASM, cfg->band_limited, cfg->strict, cfg->padding, cfg->max_padding, cfg->medium,
                       nullptr, nullptr, nullptr);
    int st = run_host(ctx, p, in, out, nullptr, &done);
    fill_asm_report(done, rep);
    return st;
}

static int dev_asm(oc_context* ctx, bool dbl, const oc_band_geom* g, const void* in, double z,
                   const oc_asm_config* cfg, const oc_phase_mask* mask, void* out, oc_asm_report* rep) {
    if (!ctx || !g || !cfg) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    std::vector<Prob> v{make_prob(g, z, ENG_ASM, cfg->band_limited, cfg->strict, cfg->padding, cfg->max_padding,
                                  cfg->medium, mask, in, out)};
    int st = run_batch(ctx, dbl, v);
    fill_asm_report(v[0], rep);
    return st;
}

int oc_asm_propagate_dev_c128(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                              const oc_asm_config* cfg, const oc_phase_mask* mask, double* out, oc_asm_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_asm(ctx, true, g, in, z, cfg, mask, out, rep);
}
int oc_asm_propagate_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                             const oc_asm_config* cfg, const oc_phase_mask* mask, float* out, oc_asm_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_asm(ctx, false, g, in, z, cfg, mask, out, rep);
}

int oc_fresnel_propagate(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                         const oc_fresnel_config* cfg, double* out, oc_fresnel_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    Prob done;
    Prob p = make_prob(g, z, ENG_FRESNEL, 0, cfg->strict, cfg->padding, cfg->max_padding, 1.0, nullptr, nullptr,
                       nullptr);
    int st = run_host(ctx, p, in, out, nullptr, &done);
    if (rep) *rep = done.fr_rep;
    return st;
}

static int dev_fresnel(oc_context* ctx, bool dbl, const oc_band_geom* g, const void* in, double z,
                       const oc_fresnel_config* cfg, const oc_phase_mask* mask, void* out, oc_fresnel_report* rep) {
    if (!ctx || !g || !cfg) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    std::vector<Prob> v{make_prob(g, z, ENG_FRESNEL, 0, cfg->strict, cfg->padding, cfg->max_padding, 1.0, mask, in,
                                  out)};
    int st = run_batch(ctx, dbl, v);
    if (rep) *rep = v[0].fr_rep;
    return st;
}

int oc_fresnel_propagate_dev_c128(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                                  const oc_fresnel_config* cfg, const oc_phase_mask* mask, double* out,
                                  oc_fresnel_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_fresnel(ctx, true, g, in, z, cfg, mask, out, rep);
}
int oc_fresnel_propagate_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                                 const oc_fresnel_config* cfg, const oc_phase_mask* mask, float* out,
                                 oc_fresnel_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_fresnel(ctx, false, g, in, z, cfg, mask, out, rep);
}

int oc_fresnel_propagate_convolution(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                                     const oc_fresnel_config* cfg, double* out, oc_fresnel_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    Prob done;
    Prob p = make_prob(g, z, ENG_FRESNEL_CONV, 0, cfg->strict, cfg->padding, cfg->max_padding, 1.0, nullptr, nullptr,
                       nullptr);
    int st = run_host(ctx, p, in, out, nullptr, &done);
    if (rep) *rep = done.fr_rep;
    return st;
}

static int dev_fresnel_conv(oc_context* ctx, bool dbl, const oc_band_geom* g, const void* in, double z,
                            const oc_fresnel_config* cfg, void* out, oc_fresnel_report* rep) {
    if (!ctx || !g || !cfg) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    std::vector<Prob> v{
        make_prob(g, z, ENG_FRESNEL_CONV, 0, cfg->strict, cfg->padding, cfg->max_padding, 1.0, nullptr, in, out)};
    int st = run_batch(ctx, dbl, v);
    if (rep) *rep = v[0].fr_rep;
    return st;
}

int oc_fresnel_propagate_convolution_dev_c128(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                                              const oc_fresnel_config* cfg, double* out, oc_fresnel_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_fresnel_conv(ctx, true, g, in, z, cfg, out, rep);
}
int oc_fresnel_propagate_convolution_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                                             const oc_fresnel_config* cfg, float* out, oc_fresnel_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_fresnel_conv(ctx, false, g, in, z, cfg, out, rep);
}

int oc_fraunhofer_regime_report(const oc_band_geom* g, double z, const oc_fraunhofer_config* cfg,
                                oc_fraunhofer_report* rep) {
    if (!g || !cfg || !rep) return OC_ERR_INVALID_ARGUMENT;
    *rep = och::fraunhofer_report(*g, z, *cfg);
    return OC_OK;
}

int oc_fraunhofer_propagate(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                            const oc_fraunhofer_config* cfg, double* out, oc_fraunhofer_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    Prob done;
    Prob p = make_prob(g, z, ENG_FRAUNHOFER, 0, cfg->strict, 1.0, 1.0, 1.0, nullptr, nullptr, nullptr);
    p.fh_cfg = *cfg;
    int st = run_host(ctx, p, in, out, nullptr, &done);
    if (rep) *rep = done.fh_rep;
    return st;
}

static int dev_fraunhofer(oc_context* ctx, bool dbl, const oc_band_geom* g, const void* in, double z,
                          const oc_fraunhofer_config* cfg, void* out, oc_fraunhofer_report* rep) {
    if (!ctx || !g || !cfg) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    std::vector<Prob> v{make_prob(g, z, ENG_FRAUNHOFER, 0, cfg->strict, 1.0, 1.0, 1.0, nullptr, in, out)};
    v[0].fh_cfg = *cfg;
    int st = run_batch(ctx, dbl, v);
    if (rep) *rep = v[0].fh_rep;
    return st;
}

int oc_fraunhofer_propagate_dev_c128(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                                     const oc_fraunhofer_config* cfg, double* out, oc_fraunhofer_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_fraunhofer(ctx, true, g, in, z, cfg, out, rep);
}
int oc_fraunhofer_propagate_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                                    const oc_fraunhofer_config* cfg, float* out, oc_fraunhofer_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_fraunhofer(ctx, false, g, in, z, cfg, out, rep);
}

int oc_rs_propagate(oc_context* ctx, const oc_band_geom* g, const double* in, double z, const oc_rs_config* cfg,
                    double* out, oc_rs_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !cfg) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    Prob done;
    Prob p = make_prob(g, z, ENG_RS, 0, cfg->strict, cfg->padding, cfg->padding, 1.0, nullptr, nullptr, nullptr);
    int st = run_host(ctx, p, in, out, nullptr, &done);
    if (rep) *rep = done.rs_rep;
    return st;
}

static int dev_rs(oc_context* ctx, bool dbl, const oc_band_geom* g, const void* in, double z,
                  const oc_rs_config* cfg, void* out, oc_rs_report* rep) {
    if (!ctx || !g || !cfg) return OC_ERR_INVALID_ARGUMENT;
    if (rep) std::memset(rep, 0, sizeof *rep);
    (void)hipSetDevice(ctx->device);
    std::vector<Prob> v{make_prob(g, z, ENG_RS, 0, cfg->strict, cfg->padding, cfg->padding, 1.0, nullptr, in, out)};
    int st = run_batch(ctx, dbl, v);
    if (rep) *rep = v[0].rs_rep;
    return st;
}

int oc_rs_propagate_dev_c128(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                             const oc_rs_config* cfg, double* out, oc_rs_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_rs(ctx, true, g, in, z, cfg, out, rep);
}
int oc_rs_propagate_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                            const oc_rs_config* cfg, float* out, oc_rs_report* rep) {
    OC_CONTEXT_LOCK(ctx);
    return dev_rs(ctx, false, g, in, z, cfg, out, rep);
}

static int batch_probs(const oc_prop_desc* d, size_t count, std::vector<Prob>& v) {
    return probs_from_descs(d, count, v);
}

static int batch(oc_context* ctx, bool dbl, oc_prop_desc* d, size_t count) {
    if (!ctx || (!d && count)) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    std::vector<Prob> v;
    int st = batch_probs(d, count, v);
    if (st) return st;
    st = run_batch(ctx, dbl, v);
    for (size_t i = 0; i < count; ++i) d[i].status = v[i].status;
    return st;
}

// Submit / wait: batches queued on the context stream with their flags in
// buffers of their own, resolved together by oc_context_wait.
static int submit(oc_context* ctx, bool dbl, oc_prop_desc* d, size_t count) {
    if (!ctx || (!d && count)) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    QueuedSet& qs = queued_set(ctx);
    Queued q;
    int st = batch_probs(d, count, q.probs);
    if (st) return st;
    void* f;
    st = ensure_buf(ctx, "queued_flags_" + std::to_string(qs.batches.size()), sizeof(int) * std::max<size_t>(count, 1),
                    &f);
    if (st) return st;
    q.dflags = static_cast<int*>(f);
    q.descs = d;
    qs.batches.push_back(std::move(q));
    Queued& b = qs.batches.back();
    st = launch_problems(ctx, dbl, b.probs, b.dflags);
    if (st) return st;  // a device error: the queued batches are resolved by the next wait
    int first = OC_OK;
    for (size_t i = 0; i < count; ++i) {
        d[i].status = b.probs[i].status;  // host-detected statuses now, the rest at the wait
        if (first == OC_OK && b.probs[i].status != OC_OK) first = b.probs[i].status;
    }
    return first;
}

// A propagation chained into a second one of its output (two gaps of
// Beamline::trace, src/beamline.rs:109-130, for many fields at once).  When
// every problem is a complex64 ASM at a power-of-two padded length followed
// by a Fresnel on the same padded grid, the ASM's inverse row pass and the
// Fresnel's row pass run as one kernel and the Fresnel's column pass reads
// the blocked intermediate (f_row_chain, f_sep_col_fin).  Three-pass chains
// (complex128, or complex64 ASM into ASM) at one power-of-two length run
// linked through their rows (f_row_link); all others as two batches.
// async: the submit form.  A fused or linked chain is queued with its flags
// in a buffer of its own and resolved by oc_context_wait; a chain that runs
// as two batches waits between them, so its statuses are final on return.
static int chain(oc_context* ctx, bool dbl, oc_prop_desc* a, oc_prop_desc* f, size_t count, bool async = false) {
    if (!ctx || ((!a || !f) && count)) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (count == 0) return OC_OK;
    for (size_t i = 0; i < count; ++i) {
        const oc_band_geom &ga = a[i].geom, &gf = f[i].geom;
        if (ga.nx != gf.nx || ga.ny != gf.ny || ga.dx != gf.dx || ga.dy != gf.dy || ga.lambda != gf.lambda)
            return fail(ctx, OC_ERR_INVALID_ARGUMENT, "the second stage's grid must be the first stage's");
    }
    std::vector<Prob> pa, pf;
    int st = batch_probs(a, count, pa);
    if (!st) st = batch_probs(f, count, pf);
    if (st) return st;
    const int np = (int)count;
    evict_tables(ctx);
    // fusable? (prepared copies; the unfused path prepares its own)
    std::vector<Prob> qa = pa, qf = pf;
    bool fuse = !dbl && !ctx->force_generic;
    for (int i = 0; i < np && fuse; ++i) {
        Prob &x = qa[i], &y = qf[i];
        prepare(x);
        prepare(y);
        if (x.status == OC_OK && x.mask) x.status = mask_footprint(x.g, *x.mask);
        fuse = x.status == OC_OK && y.status == OC_OK && x.launch && y.launch && !x.check_only && !y.check_only &&
               x.engine == ENG_ASM && y.engine == ENG_FRESNEL && !y.mask && x.nxp == x.nyp && y.nxp == x.nxp &&
               y.nyp == x.nyp && x.nxp == qa[0].nxp && ocl::fast_variants((int)x.nxp) > 0;
    }
    void* fl;
    QueuedSet* qs = async ? &queued_set(ctx) : nullptr;
    st = ensure_buf(ctx, async ? "queued_flags_" + std::to_string(qs->batches.size()) : std::string("chain_flags"),
                    sizeof(int) * 2 * count, &fl);
    if (st) return st;
    int* dflags = static_cast<int*>(fl);
    // the submit form: queue the one-launch chain (all = first stages, then
    // second stages), host-detected statuses now, the rest at the wait
    auto enqueue = [&](std::vector<Prob>& all) -> int {
        Queued q;
        q.probs = all;
        q.dflags = dflags;
        q.descs = a;
        q.second = f;
        qs->batches.push_back(std::move(q));
        int first = OC_OK;
        for (int i = 0; i < np; ++i) {
            const int sa = all[i].status;
            const int sf = sa != OC_OK ? sa : all[np + i].status;
            a[i].status = sa;
            f[i].status = sf;
            if (first == OC_OK && sa != OC_OK) first = sa;
            if (first == OC_OK && sf != OC_OK) first = sf;
        }
        return first;
    };
    // per-problem statuses of a one-launch chain (all = first stages, then
    // second stages, flags at dflags): a failed first stage fails the chain
    auto resolve = [&](std::vector<Prob>& all) -> int {
        const int fs = finish_problems(ctx, all, dflags);
        if (fs == OC_ERR_DEVICE || fs == OC_ERR_OUT_OF_MEMORY) return fs;  // flags never read back
        int first = OC_OK;
        for (int i = 0; i < np; ++i) {
            const int sa = all[i].status;
            const int sf = sa != OC_OK ? sa : all[np + i].status;
            a[i].status = sa;
            f[i].status = sf;
            if (first == OC_OK && sa != OC_OK) first = sa;
            if (first == OC_OK && sf != OC_OK) first = sf;
        }
        return first;
    };
    if (fuse) {
        for (int i = 0; i < np; ++i) qf[i].in = qa[i].out;  // unused by the fused passes
        HIP_TRY(ctx, ocl::launch_fill_int(dflags, 0, 2 * np, ctx->stream));
        Batch ba, bf;
        // the Fresnel batch takes the separable form's layouts (rows for the
        // fused row pass, columns for its column pass); both must block the
        // intermediates with the ASM batch's width
        st = build_batch(ctx, false, qa, dflags, ba, false, nullptr, "chA");
        if (!st) st = build_batch(ctx, false, qf, dflags + np, bf, false, nullptr, "chF");
        if (!st && ba.fast && bf.fast && !ba.pd.empty() && !bf.pd.empty() && ba.pd[0].B != bf.pd[0].B) {
            bf = Batch();  // the tuned pair blocks differently: both Fresnel passes take the ASM's layout
            st = build_batch(ctx, false, qf, dflags + np, bf, false, nullptr, "chF", ba.fast_v);
        }
        if (st) return st;
        fuse = ba.fast && !ba.sep && !ba.rs && ba.col_kind == ENG_ASM && ba.post_kind == POST_NONE &&
               (int)ba.pd.size() == np && bf.fast && bf.sep && bf.fast_n == ba.fast_n && (int)bf.pd.size() == np &&
               ba.pd[0].B == bf.pd[0].B;
        if (fuse) {
            st = launch_batch(ctx, ba, (1u << 0) | (1u << 1));
            if (st) return st;
            const ProbDev* da = static_cast<const ProbDev*>(ba.dprobs);
            const ProbDev* df = static_cast<const ProbDev*>(bf.dprobs);
            const int rows = bf.pc[2].total, cols = bf.pc[1].total;
            const int hw = (ba.half && bf.half) ? 2 : 0;  // the kernels for N = P/2
            hipError_t e = timed(ctx, PASS_T_SEP_ROW, [&] {
                return ocl::fast_launch_chain_c64(hw + 0, bf.fast_n, bf.fast_v_row, da, df, np, rows,
                                                  std::min(rows, ctx->num_cu * 8), ctx->stream);
            });
            if (e == hipSuccess)
                e = timed(ctx, PASS_T_SEP_COL, [&] {
                    return ocl::fast_launch_chain_c64(hw + 1, bf.fast_n, bf.fast_v, nullptr, df, np, cols,
                                                      std::min(cols, ctx->num_cu * 8), ctx->stream);
                });
            if (e != hipSuccess) return fail(ctx, OC_ERR_DEVICE, std::string("chain launch: ") + hipGetErrorString(e));
            std::vector<Prob> all = qa;
            all.insert(all.end(), qf.begin(), qf.end());
            return async ? enqueue(all) : resolve(all);
        }
        // the batches were only built (and any shared transmission computed): fall through
    }
    // three-pass transfer-function engines at one power-of-two length: the
    // two steps linked through their rows (f_row_link), the intermediate
    // written where the caller asked for it
    if (!ctx->force_generic) {
        std::vector<std::vector<Prob>> run{pa, pf}, done;
        const int sl = launch_linked(ctx, dbl, run, dflags, std::vector<int>{0, np}, done, true);
        if (sl < 0) return -sl;
        if (sl == 0) {
            std::vector<Prob> all = done[0];
            all.insert(all.end(), done[1].begin(), done[1].end());
            return async ? enqueue(all) : resolve(all);
        }
    }
    // two batches: the first stage into the caller's buffer or a scratch one
    size_t need = 0;
    const size_t esz = dbl ? 16 : 8;
    for (int i = 0; i < np; ++i)
        if (!pa[i].out) need += pa[i].g.nx * pa[i].g.ny * esz;
    void* scratch = nullptr;
    if (need) {
        st = ensure_buf(ctx, "chain_mid", need, &scratch);
        if (st) return st;
    }
    size_t off = 0;
    for (int i = 0; i < np; ++i) {
        if (!pa[i].out) {
            pa[i].out = static_cast<unsigned char*>(scratch) + off;
            off += pa[i].g.nx * pa[i].g.ny * esz;
        }
        pf[i].in = pa[i].out;
    }
    st = launch_problems(ctx, dbl, pa, dflags);
    if (st) return st;
    st = finish_problems(ctx, pa, dflags);
    if (st == OC_ERR_DEVICE || st == OC_ERR_OUT_OF_MEMORY) return st;
    std::vector<Prob> rest;
    std::vector<int> where;
    for (int i = 0; i < np; ++i)
        if (pa[i].status == OC_OK) {
            rest.push_back(pf[i]);
            where.push_back(i);
        } else {
            pf[i].status = pa[i].status;
        }
    if (!rest.empty()) {
        st = launch_problems(ctx, dbl, rest, dflags + np);
        if (st) return st;
        st = finish_problems(ctx, rest, dflags + np);
        if (st == OC_ERR_DEVICE || st == OC_ERR_OUT_OF_MEMORY) return st;
        for (size_t k = 0; k < rest.size(); ++k) pf[where[k]].status = rest[k].status;
    }
    int first = OC_OK;
    for (int i = 0; i < np; ++i) {
        a[i].status = pa[i].status;
        f[i].status = pf[i].status;
        if (first == OC_OK && pa[i].status != OC_OK) first = pa[i].status;
        if (first == OC_OK && pf[i].status != OC_OK) first = pf[i].status;
    }
    return first;
}

int oc_propagate_chain_dev_c64(oc_context* ctx, oc_prop_desc* first, oc_prop_desc* second, size_t count) {
    OC_CONTEXT_LOCK(ctx);
    return chain(ctx, false, first, second, count);
}
int oc_propagate_chain_dev_c128(oc_context* ctx, oc_prop_desc* first, oc_prop_desc* second, size_t count) {
    OC_CONTEXT_LOCK(ctx);
    return chain(ctx, true, first, second, count);
}
int oc_propagate_chain_submit_dev_c64(oc_context* ctx, oc_prop_desc* first, oc_prop_desc* second, size_t count) {
    OC_CONTEXT_LOCK(ctx);
    return chain(ctx, false, first, second, count, true);
}
int oc_propagate_chain_submit_dev_c128(oc_context* ctx, oc_prop_desc* first, oc_prop_desc* second, size_t count) {
    OC_CONTEXT_LOCK(ctx);
    return chain(ctx, true, first, second, count, true);
}

int oc_propagate_batch_dev_c64(oc_context* ctx, oc_prop_desc* d, size_t count) {
    OC_CONTEXT_LOCK(ctx);
    return batch(ctx, false, d, count);
}
int oc_propagate_batch_dev_c128(oc_context* ctx, oc_prop_desc* d, size_t count) {
    OC_CONTEXT_LOCK(ctx);
    return batch(ctx, true, d, count);
}
int oc_propagate_batch_submit_dev_c64(oc_context* ctx, oc_prop_desc* d, size_t count) {
    OC_CONTEXT_LOCK(ctx);
    return submit(ctx, false, d, count);
}
int oc_propagate_batch_submit_dev_c128(oc_context* ctx, oc_prop_desc* d, size_t count) {
    OC_CONTEXT_LOCK(ctx);
    return submit(ctx, true, d, count);
}

int oc_context_wait(oc_context* ctx) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (!ctx->queued) return OC_OK;
    QueuedSet& qs = queued_set(ctx);
    int first = OC_OK;
    for (Queued& q : qs.batches) {
        const int st = finish_problems(ctx, q.probs, q.dflags);
        if (q.second) {  // a chain: a failed first stage fails the chain
            const size_t np = q.probs.size() / 2;
            for (size_t i = 0; i < np; ++i) {
                const int sa = q.probs[i].status;
                const int sf = sa != OC_OK ? sa : q.probs[np + i].status;
                q.descs[i].status = sa;
                q.second[i].status = sf;
                if (first == OC_OK && sa != OC_OK) first = sa;
                if (first == OC_OK && sf != OC_OK) first = sf;
            }
            if (first == OC_OK && st != OC_OK) first = st;
            continue;
        }
        for (size_t i = 0; i < q.probs.size(); ++i) q.descs[i].status = q.probs[i].status;
        if (first == OC_OK && st != OC_OK) first = st;
    }
    qs.batches.clear();
    return first;
}

int oc_phasemask_validate(const oc_phase_mask* m) {
    if (!m) return OC_ERR_INVALID_ARGUMENT;
    if (!std::isfinite(och::wrap_angle(m->orientation))) return OC_INVALID_ORIENTATION;
    if (m->nx == 0 || m->ny == 0 || !m->values) return OC_INVALID_GRID;
    if (!std::isfinite(m->dx) || m->dx <= 0.0 || !std::isfinite(m->dy) || m->dy <= 0.0) return OC_INVALID_GRID_PITCH;
    for (size_t k = 0; k < m->nx * m->ny; ++k)
        if (!std::isfinite(m->values[k])) return OC_INVALID_PHASE;
    auto fp = [&](double hx, double hy) {
        return (!std::isfinite(hx) || hx <= 0.0 || !std::isfinite(hy) || hy <= 0.0 || !std::isfinite(m->cx) ||
                !std::isfinite(m->cy))
                   ? OC_INVALID_APERTURE
                   : OC_OK;
    };
    switch (m->aperture.kind) {
        case OC_APERTURE_FIELD: return (std::isfinite(m->cx) && std::isfinite(m->cy)) ? OC_OK : OC_INVALID_AXIS;
        case OC_APERTURE_CIRCLE: return fp(m->aperture.a, m->aperture.a);
        case OC_APERTURE_ELLIPSE:
        case OC_APERTURE_RECTANGLE: return fp(m->aperture.a, m->aperture.b);
        default: return OC_ERR_INVALID_ARGUMENT;
    }
}

static int mask_dev(oc_context* ctx, bool dbl, const oc_band_geom* g, const void* in, const oc_phase_mask* m,
                    void* out) {
    if (!ctx || !g || !m) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    int st = mask_footprint(*g, *m);
    if (st) return st;
    MaskDev md;
    st = mask_tables(ctx, *g, *m, m->values, &md);
    if (st) return st;
    void* f;
    st = ensure_buf(ctx, "flags", sizeof(int), &f);
    if (st) return st;
    HIP_TRY(ctx, ocl::launch_fill_int(static_cast<int*>(f), 0, 1, ctx->stream));
    hipError_t e = dbl ? ocl::launch_mask_apply_c128(in, out, md, (int)g->nx, (int)g->ny, nullptr, ctx->stream)
                       : ocl::launch_mask_apply_c64(in, out, md, (int)g->nx, (int)g->ny, nullptr, ctx->stream);
    if (e != hipSuccess) return fail(ctx, OC_ERR_DEVICE, hipGetErrorString(e));
    return OC_OK;
}

int oc_phasemask_apply_dev_c128(oc_context* ctx, const oc_band_geom* g, const double* in, const oc_phase_mask* m,
                                double* out) {
    OC_CONTEXT_LOCK(ctx);
    return mask_dev(ctx, true, g, in, m, out);
}
int oc_phasemask_apply_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, const oc_phase_mask* m,
                               float* out) {
    OC_CONTEXT_LOCK(ctx);
    return mask_dev(ctx, false, g, in, m, out);
}

int oc_phasemask_apply(oc_context* ctx, const oc_band_geom* g, const double* in, const oc_phase_mask* m, double* out) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !g || !m) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    int st = mask_footprint(*g, *m);
    if (st) return st;
    // PhaseMask::sampled validated its lattice at construction; a non-finite
    // lattice value would surface as InvalidPhase at apply time.
    for (size_t k = 0; k < m->nx * m->ny; ++k)
        if (!std::isfinite(m->values[k])) return OC_INVALID_PHASE;
    const size_t n = g->nx * g->ny;
    if (n == 0) return OC_OK;
    void *din, *dout, *dmask;
    st = ensure_buf(ctx, "host_in", n * 16, &din);
    if (st) return st;
    st = ensure_buf(ctx, "host_out", n * 16, &dout);
    if (st) return st;
    st = ensure_buf(ctx, "host_mask", std::max<size_t>(m->nx * m->ny, 1) * 8, &dmask);
    if (st) return st;
    HIP_TRY(ctx, hipMemcpyAsync(din, in, n * 16, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dmask, m->values, m->nx * m->ny * 8, hipMemcpyHostToDevice, ctx->stream));
    oc_phase_mask dm = *m;
    dm.values = static_cast<const double*>(dmask);
    st = mask_dev(ctx, true, g, din, &dm, dout);
    if (st) return st;
    HIP_TRY(ctx, hipMemcpyAsync(out, dout, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OC_OK;
}

}  // extern "C"

extern "C" {

int oc_context_set_fft_variant(oc_context* ctx, uint64_t n, int dbl, int variant) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || variant < 0 || variant >= std::max(1, ocl::fast_variants((int)n))) return OC_ERR_INVALID_ARGUMENT;
    ctx->fft_variant[std::make_pair((int)n, dbl ? 1 : 0)] = variant;
    if (!dbl) ctx->sep_variant.erase((int)n);  // both Fresnel passes follow the forced layout
    return OC_OK;
}

int oc_context_fft_variant(oc_context* ctx, uint64_t n, int dbl) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || ocl::fast_variants((int)n) == 0) return -1;
    return ctx->variant_for((int)n, dbl != 0);
}

int oc_context_set_fresnel_sep_layout(oc_context* ctx, uint64_t n, int v_col, int v_row) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    int cG, cB, cC, rG, rB, rC;
    const int nv = ocl::fast_variants((int)n);
    if (nv == 0 || v_col < 0 || v_row < 0 || v_col >= nv || v_row >= nv ||
        !ocl::fast_supported(false, (int)n, v_col, &cG, &cB, &cC) ||
        !ocl::fast_supported(false, (int)n, v_row, &rG, &rB, &rC))
        return fail(ctx, OC_ERR_INVALID_ARGUMENT, "no such layout at this length");
    if (cB != rB) return fail(ctx, OC_ERR_INVALID_ARGUMENT, "the two layouts block the intermediate differently");
    ctx->sep_variant[(int)n] = std::make_pair(v_col, v_row);
    return OC_OK;
}

int oc_context_fresnel_sep_layout(oc_context* ctx, uint64_t n, int* v_col, int* v_row) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !v_col || !v_row) return OC_ERR_INVALID_ARGUMENT;
    if (ocl::fast_variants((int)n) == 0) {
        *v_col = *v_row = -1;
        return OC_OK;
    }
    const std::pair<int, int> sv = ctx->sep_variants_for((int)n);
    *v_col = sv.first;
    *v_row = sv.second;
    return OC_OK;
}

int oc_context_force_generic(oc_context* ctx, int enable) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    ctx->force_generic = enable != 0;
    return OC_OK;
}

// Times every layout variant of the specialised kernels at padded length n on
// a synthetic round of masked ASM then Fresnel propagations (up to 16 fields
// of n/2 x n/2, padding 2) and keeps the fastest; for complex64 it then times
// the Fresnel round alone over the layout pairs of its two separable passes.
int oc_autotune_fft(oc_context* ctx, uint64_t n, int dbl, int* variant_out, double* ms_out) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    const int nv = ocl::fast_variants((int)n);
    if (ms_out)
        for (int v = 0; v < OC_FFT_VARIANTS; ++v) ms_out[v] = -1.0;
    if (nv == 0) {
        if (variant_out) *variant_out = -1;
        return OC_OK;
    }
    const size_t half = n / 2, esz = dbl ? 16 : 8;
    const size_t per = half * half;
    const size_t nb = std::max<size_t>(1, std::min<size_t>(16, ((size_t)1 << 27) / per));
    Scoped mem(ctx);
    void *in, *out, *mv;
    int st;
    if ((st = mem.alloc(nb * per * esz, &in)) || (st = mem.alloc(nb * per * esz, &out)) ||
        (st = mem.alloc(per * 8, &mv)))
        return st;
    HIP_TRY(ctx, hipMemsetAsync(in, 0, nb * per * esz, ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(mv, 0, per * 8, ctx->stream));
    oc_phase_mask m;
    std::memset(&m, 0, sizeof m);
    m.values = static_cast<const double*>(mv);
    m.nx = m.ny = half;
    m.dx = m.dy = 2e-6;
    oc_band_geom g{half, half, 2e-6, 2e-6, 550e-9};
    std::vector<Prob> probs, probs_f;
    for (size_t b = 0; b < nb; ++b) {
        unsigned char* bi = static_cast<unsigned char*>(in) + b * per * esz;
        unsigned char* bo = static_cast<unsigned char*>(out) + b * per * esz;
        probs.push_back(make_prob(&g, 5e-3, ENG_ASM, 1, 0, 2.0, 2.0, 1.0, &m, bi, bo));
        probs_f.push_back(make_prob(&g, 0.2, ENG_FRESNEL, 0, 0, 2.0, 2.0, 1.0, nullptr, bo, bi));
    }
    const auto key = std::make_pair((int)n, dbl ? 1 : 0);
    const bool had = ctx->fft_variant.count(key) != 0;
    const int prev = had ? ctx->fft_variant[key] : 0;
    const bool had_sep = !dbl && ctx->sep_variant.count((int)n) != 0;
    const std::pair<int, int> prev_sep = had_sep ? ctx->sep_variant[(int)n] : std::make_pair(0, 0);
    if (!dbl) ctx->sep_variant.erase((int)n);
    const bool timing = ctx->timing;
    ctx->timing = false;
    // best of three timed rounds (after one untimed) of the ASM batch and/or
    // the Fresnel batch
    auto round_ms = [&](bool with_asm, double* ms) {
        *ms = 1e300;
        for (int rep = 0; rep < 4; ++rep) {
            std::vector<Prob> pv = probs, pf = probs_f;
            hipEvent_t a = pool_event(ctx), e = pool_event(ctx);
            (void)hipEventRecord(a, ctx->stream);
            int s = with_asm ? run_batch(ctx, dbl, pv) : OC_OK;
            if (!s) s = run_batch(ctx, dbl, pf);
            (void)hipEventRecord(e, ctx->stream);
            (void)hipEventSynchronize(e);
            float f = 0.f;
            (void)hipEventElapsedTime(&f, a, e);
            ctx->event_pool.push_back(a);
            ctx->event_pool.push_back(e);
            if (s) return s;
            if (rep > 0) *ms = std::min(*ms, (double)f);
        }
        return (int)OC_OK;
    };
    int best = 0;
    double best_ms = 1e300;
    for (int v = 0; v < nv; ++v) {
        ctx->fft_variant[key] = v;
        double ms;
        if ((st = round_ms(true, &ms))) break;
        if (ms_out) ms_out[v] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    // complex64: then the two separable Fresnel passes' layouts, over every
    // pair that shares the blocked width
    if (!st && !dbl) {
        ctx->fft_variant[key] = best;
        std::pair<int, int> best_sep(best, best);
        double best_sep_ms = 1e300;
        for (int vc = 0; vc < nv && !st; ++vc)
            for (int vr = 0; vr < nv && !st; ++vr) {
                int cG, cB, cC, rG, rB, rC;
                if (!ocl::fast_supported(false, (int)n, vc, &cG, &cB, &cC) ||
                    !ocl::fast_supported(false, (int)n, vr, &rG, &rB, &rC) || cB != rB)
                    continue;
                ctx->sep_variant[(int)n] = std::make_pair(vc, vr);
                double ms;
                if ((st = round_ms(false, &ms))) break;
                if (ms < best_sep_ms) {
                    best_sep_ms = ms;
                    best_sep = std::make_pair(vc, vr);
                }
            }
        if (!st) ctx->sep_variant[(int)n] = best_sep;
    }
    ctx->timing = timing;
    if (st) {
        if (had) ctx->fft_variant[key] = prev;
        else ctx->fft_variant.erase(key);
        if (had_sep) ctx->sep_variant[(int)n] = prev_sep;
        else if (!dbl) ctx->sep_variant.erase((int)n);
        return st;
    }
    ctx->fft_variant[key] = best;
    if (variant_out) *variant_out = best;
    return OC_OK;
}

}  // extern "C"
