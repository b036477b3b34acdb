// Runtime of the propagation library: context caches (buffers, twiddles,
// axis and mask tables), plans, host guards and the batch builder.
//
// Every public propagation call funnels into run_batch(): host-side validation
// in the reference's order, O(1) sampling guards, plan/table lookup (cached
// per context), then the persistent pass launches of prop_kernels.hip over all
// problems of the call at once.
#include "runtime.hpp"

namespace ocr {


int fail(oc_context* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

hipEvent_t pool_event(oc_context* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}


// Folds completed brackets into the per-pass totals (stream synchronised).
void collect_timing(oc_context* ctx) {
    for (auto& p : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ctx->pass_ms[p.pass] += ms;
            ctx->pass_n[p.pass] += 1;
        }
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
}


int arena_alloc(oc_context* ctx, size_t bytes, void** out) {
    const size_t a = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    if (!ctx->scratch || ctx->arena_used > ctx->scratch_cap || a > ctx->scratch_cap - ctx->arena_used)
        return fail(ctx, OC_ERR_INVALID_ARGUMENT,
                    "the no-allocation scratch (" + std::to_string(ctx->scratch_cap >> 20) + " MiB) is exhausted");
    *out = static_cast<unsigned char*>(ctx->scratch) + ctx->arena_used;
    ctx->arena_used += a;
    return OC_OK;
}

int arena_upload(oc_context* ctx, const void* host, size_t bytes, const void** out) {
    void* d;
    int st = arena_alloc(ctx, bytes, &d);
    if (st) return st;
    ctx->arena_keep.emplace_back(static_cast<const unsigned char*>(host),
                                 static_cast<const unsigned char*>(host) + bytes);
    HIP_TRY(ctx, hipMemcpyAsync(d, ctx->arena_keep.back().data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    *out = d;
    return OC_OK;
}

const void* find_table(oc_context* ctx, const std::string& key) {
    if (ctx->arena) {
        auto it = ctx->arena_tables.find(key);
        return it == ctx->arena_tables.end() ? nullptr : it->second;
    }
    auto it = ctx->tables.find(key);
    return it == ctx->tables.end() ? nullptr : it->second.dev;
}

int store_table(oc_context* ctx, const std::string& key, const void* host, size_t bytes, const void** out) {
    if (ctx->arena) {
        int st = arena_upload(ctx, host, bytes, out);
        if (st) return st;
        ctx->arena_tables[key] = *out;
        return OC_OK;
    }
    TableEntry te;
    HIP_TRY(ctx, hipMalloc(&te.dev, std::max<size_t>(bytes, 1)));
    HIP_TRY(ctx, hipMemcpy(te.dev, host, bytes, hipMemcpyHostToDevice));
    te.bytes = bytes;
    ctx->tables.emplace(key, te);
    *out = te.dev;
    return OC_OK;
}

int ensure_buf(oc_context* ctx, const std::string& name, size_t bytes, void** out) {
    if (ctx->arena) return arena_alloc(ctx, bytes, out);
    DevBuf& b = ctx->bufs[name];
    if (b.cap < bytes) {
        if (b.p) {
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            HIP_TRY(ctx, hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
        }
        size_t cap = std::max<size_t>(bytes, 256);
        HIP_TRY(ctx, hipMalloc(&b.p, cap));
        b.cap = cap;
    }
    *out = b.p;
    return OC_OK;
}

int ensure_readback(oc_context* ctx, size_t bytes, void** out) {
    if (ctx->readback_cap < bytes) {
        if (ctx->readback) HIP_TRY(ctx, hipHostFree(ctx->readback));
        ctx->readback = nullptr;
        ctx->readback_cap = 0;
        const size_t cap = std::max(bytes, (size_t)1 << 16);
        HIP_TRY(ctx, hipHostMalloc(&ctx->readback, cap, hipHostMallocDefault));
        ctx->readback_cap = cap;
    }
    *out = ctx->readback;
    return OC_OK;
}

int ensure_pinned(oc_context* ctx, int slot, size_t bytes) {
    if (ctx->pinned_cap[slot] >= bytes) return OC_OK;
    if (!ctx->pinned[slot] && bytes <= (size_t)(1 << 18)) {
        // first use: every empty slot at once (pinned allocation is slow, and
        // the ring would otherwise allocate on the first kPinSlots uploads)
        for (int k = 0; k < oc_context::kPinSlots; ++k)
            if (!ctx->pinned[k]) {
                HIP_TRY(ctx, hipHostMalloc(&ctx->pinned[k], (size_t)1 << 18, hipHostMallocDefault));
                ctx->pinned_cap[k] = (size_t)1 << 18;
            }
        return OC_OK;
    }
    if (ctx->pinned[slot]) {
        HIP_TRY(ctx, hipHostFree(ctx->pinned[slot]));
        ctx->pinned[slot] = nullptr;
        ctx->pinned_cap[slot] = 0;
    }
    size_t cap = std::max<size_t>(bytes, (size_t)1 << 18);
    HIP_TRY(ctx, hipHostMalloc(&ctx->pinned[slot], cap, hipHostMallocDefault));
    ctx->pinned_cap[slot] = cap;
    return OC_OK;
}

// exp(-2 pi i k/n), k < n, evaluated in extended precision then rounded.
int get_twiddles(oc_context* ctx, int n, bool dbl, const void** out) {
    auto key = std::make_pair(n, dbl ? 1 : 0);
    const std::string akey = fmtkey("tw", {(double)n, dbl ? 1.0 : 0.0});
    if (ctx->arena) {
        if (const void* t = find_table(ctx, akey)) {
            *out = t;
            return OC_OK;
        }
    } else {
        auto it = ctx->twiddles.find(key);
        if (it != ctx->twiddles.end()) {
            *out = it->second;
            return OC_OK;
        }
    }
    const long double two_pi = 6.283185307179586476925286766559005768L;
    size_t esz = dbl ? 16 : 8;
    std::vector<unsigned char> host(esz * (size_t)n);
    for (int k = 0; k < n; ++k) {
        long double a = -two_pi * (long double)k / (long double)n;
        long double c = cosl(a), s = sinl(a);
        if (dbl) {
            double v[2] = {(double)c, (double)s};
            std::memcpy(host.data() + esz * k, v, esz);
        } else {
            float v[2] = {(float)c, (float)s};
            std::memcpy(host.data() + esz * k, v, esz);
        }
    }
    if (ctx->arena) return store_table(ctx, akey, host.data(), host.size(), out);
    void* d = nullptr;
    HIP_TRY(ctx, hipMalloc(&d, host.size()));
    HIP_TRY(ctx, hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice));
    ctx->twiddles[key] = d;
    *out = d;
    return OC_OK;
}

// Radix plan: 16/8/4/2 for the power-of-two part, then 3, 5, 7; any other
// prime factor makes a generic (ping-pong) plan.
bool make_radices(int n, std::vector<int>& rad, bool& generic) {
    rad.clear();
    generic = false;
    int m = n;
    int p2 = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++p2;
    }
    while (p2 >= 4) {
        rad.push_back(16);
        p2 -= 4;
    }
    if (p2 == 3) rad.push_back(8);
    if (p2 == 2) rad.push_back(4);
    if (p2 == 1) rad.push_back(2);
    for (int p : {3, 5, 7}) {
        while (m % p == 0) {
            rad.push_back(p);
            m /= p;
        }
    }
    for (int p = 11; m > 1 && p * p <= m; p += 2) {
        while (m % p == 0) {
            rad.push_back(p);
            m /= p;
            generic = true;
        }
    }
    if (m > 1) {
        rad.push_back(m);
        generic = true;
    }
    if ((int)rad.size() > kMaxStages) return false;
    return true;
}

bool valid_radices(int n, const int* rad, int nrad) {
    if (n <= 0 || nrad <= 0 || nrad > kMaxStages) return false;
    long long prod = 1;
    for (int i = 0; i < nrad; ++i) {
        const int r = rad[i];
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 7 && r != 8 && r != 16) return false;
        prod *= r;
        if (prod > n) return false;
    }
    return prod == n;
}

int make_plan(oc_context* ctx, int n, bool dbl, FftPlanDev* pl, const int* radices, int nrad) {
    std::vector<int> rad;
    bool generic = false;
    if (radices && nrad > 0) {
        if (!valid_radices(n, radices, nrad)) return fail(ctx, OC_ERR_INVALID_ARGUMENT, "radix plan does not fit n");
        rad.assign(radices, radices + nrad);
    } else if (!make_radices(n, rad, generic)) {
        return fail(ctx, OC_ERR_UNSUPPORTED, "too many FFT stages");
    }
    std::memset(pl, 0, sizeof *pl);
    pl->n = n;
    pl->nstages = (int)rad.size();
    pl->generic = generic ? 1 : 0;
    for (size_t i = 0; i < rad.size(); ++i) pl->radix[i] = rad[i];
    pl->threads = generic ? std::min(256, std::max(1, n)) : std::max(1, (n + 15) / 16);
    return get_twiddles(ctx, n, dbl, &pl->tw);
}

size_t plan_lds_elems(const FftPlanDev& pl) {
    return (size_t)lds_stride(pl.n) * (pl.generic ? 2 : 1);
}


// Cache key: the tag and the exact bit patterns of the values (no text
// formatting: these keys are built for every problem of every call).
std::string fmtkey(const char* tag, std::initializer_list<double> vals) {
    std::string s(tag);
    s.reserve(s.size() + 1 + 8 * vals.size());
    s += '|';
    for (double v : vals) s.append(reinterpret_cast<const char*>(&v), sizeof v);
    return s;
}

// Builds (or reuses) the per-axis engine tables of one problem.
// Per-axis factors of the Fresnel transfer function for complex64 launches
// (see EngineDev::hx): phases k z - c fx^2 and -c fy^2 in f64, sin/cos in
// long double, rounded to complex64.
int fresnel_factors(oc_context* ctx, const Prob& p, EngineDev* e) {
    const size_t nxp = p.nxp, nyp = p.nyp;
    std::string key = fmtkey("frh", {(double)nxp, (double)nyp, p.g.dx, p.g.dy, p.g.lambda, p.z});
    const void* dev = find_table(ctx, key);
    if (!dev) {
        std::vector<float> h(2 * (nxp + nyp));
        for (size_t k = 0; k < nxp + nyp; ++k) {
            const bool isx = k < nxp;
            const double f = isx ? och::fftfreq_at(k, nxp, p.g.dx) : och::fftfreq_at(k - nxp, nyp, p.g.dy);
            const double ph = isx ? e->kz - e->c * (f * f) : -(e->c * (f * f));
            h[2 * k] = (float)std::cos((long double)ph);
            h[2 * k + 1] = (float)std::sin((long double)ph);
        }
        int st = store_table(ctx, key, h.data(), h.size() * sizeof(float), &dev);
        if (st) return st;
    }
    e->hx = dev;
    e->hy = static_cast<const float*>(dev) + 2 * nxp;
    return OC_OK;
}

int engine_tables(oc_context* ctx, const Prob& p, bool dbl, EngineDev* e) {
    const size_t nxp = p.nxp, nyp = p.nyp;
    std::memset(e, 0, sizeof *e);
    e->kind = p.engine;
    if (p.engine == ENG_FRAUNHOFER) {  // fraunhofer.rs:360-428
        const oc_fraunhofer_report& r = p.fh_rep;
        e->lambda = p.g.lambda;
        e->z = p.z;
        e->sgn = p.z > 0.0 ? 1.0 : -1.0;
        e->dx = p.g.dx;
        e->dy = p.g.dy;
        e->fcx = ((double)p.g.nx - 1.0) / 2.0;
        e->fcy = ((double)p.g.ny - 1.0) / 2.0;
        e->fs = r.source_x / (p.g.lambda * p.z);
        e->gs = r.source_y / (p.g.lambda * p.z);
        e->dxo = r.dx_out;
        e->dyo = r.dy_out;
        e->xs = r.source_x;
        e->ys = r.source_y;
        // e^{ikz} / (i lambda z), complex division written out
        const double k = 2.0 * och::PI / p.g.lambda;
        const double c = std::cos(k * p.z), s = std::sin(k * p.z);
        const double dre = 0.0 * p.g.lambda * p.z, dim = 1.0 * p.g.lambda * p.z;
        const double nsq = dre * dre + dim * dim;
        e->quad_re = (c * dre + s * dim) / nsq;
        e->quad_im = (s * dre - c * dim) / nsq;
        return OC_OK;
    }
    if (p.engine == ENG_FRESNEL_CONV) {  // fresnel.rs:319-321
        e->lambda = p.g.lambda;
        e->z = p.z;
        e->dx = p.g.dx;
        e->dy = p.g.dy;
        const double k = 2.0 * och::PI / p.g.lambda;
        const double c = std::cos(k * p.z), s = std::sin(k * p.z);
        const double dre = 0.0 * p.g.lambda * p.z, dim = 1.0 * p.g.lambda * p.z;
        const double nsq = dre * dre + dim * dim;
        const double qre = (c * dre + s * dim) / nsq, qim = (s * dre - c * dim) / nsq;
        e->pref_re = qre * (p.g.dx * p.g.dy);
        e->pref_im = qim * (p.g.dx * p.g.dy);
        return OC_OK;
    }
    if (p.engine == ENG_RS) {
        e->lambda = p.g.lambda;
        e->z = p.z;
        e->k = 2.0 * och::PI / p.g.lambda;
        e->sgn = p.z > 0.0 ? 1.0 : -1.0;
        e->dx = p.g.dx;
        e->dy = p.g.dy;
        return OC_OK;
    }
    double lim_x = 0, lim_y = 0;
    double lam_n = p.g.lambda;
    if (p.engine == ENG_ASM) {
        lam_n = p.g.lambda / p.medium;
        const double k = 2.0 * och::PI / lam_n;
        e->kz = k * p.z;
        lim_x = p.asm_rep.band_limit_fx;
        lim_y = p.asm_rep.band_limit_fy;
    } else {
        const double k = 2.0 * och::PI / p.g.lambda;
        e->kz = k * p.z;
        e->c = och::PI * p.g.lambda * p.z;
    }
    std::string key = fmtkey(p.engine == ENG_ASM ? "asm" : "fr",
                             {(double)nxp, (double)nyp, p.g.dx, p.g.dy, lam_n, (double)p.band_limited, lim_x, lim_y});
    const void* dev = find_table(ctx, key);
    if (!dev) {
        // ASM: (lam_n f)^2, or +inf where the band limit masks the frequency
        // (then 1 - ax - ay = -inf < 0 and the sample is zeroed like an
        // evanescent one).  Fresnel: f^2.
        std::vector<double> h(nxp + nyp);
        for (size_t k = 0; k < nxp + nyp; ++k) {
            const bool isx = k < nxp;
            const double f = isx ? och::fftfreq_at(k, nxp, p.g.dx) : och::fftfreq_at(k - nxp, nyp, p.g.dy);
            if (p.engine == ENG_ASM) {
                const double a = lam_n * f;
                const bool masked = p.band_limited && std::fabs(f) > (isx ? lim_x : lim_y);
                h[k] = masked ? HUGE_VAL : a * a;
            } else {
                h[k] = f * f;
            }
        }
        int st = store_table(ctx, key, h.data(), h.size() * sizeof(double), &dev);
        if (st) return st;
    }
    e->ax = static_cast<const double*>(dev);
    e->ay = e->ax + nxp;
    if (p.engine == ENG_FRESNEL && !dbl) return fresnel_factors(ctx, p, e);
    return OC_OK;
}

// PhaseMask::apply footprint check (src/element.rs:1446-1458).
int mask_footprint(const oc_band_geom& g, const oc_phase_mask& m) {
    const double psi = och::wrap_angle(m.orientation);
    double fx, fy;
    switch (m.aperture.kind) {
        case OC_APERTURE_FIELD: return OC_OK;
        case OC_APERTURE_CIRCLE: fx = fy = m.aperture.a; break;
        default: {
            const double c = std::fabs(std::cos(psi)), s = std::fabs(std::sin(psi));
            fx = m.aperture.a * c + m.aperture.b * s;
            fy = m.aperture.a * s + m.aperture.b * c;
        }
    }
    const double ghx = ((double)g.nx - 1.0) / 2.0 * g.dx;
    const double ghy = ((double)g.ny - 1.0) / 2.0 * g.dy;
    if (std::fabs(m.cx) + fx > ghx || std::fabs(m.cy) + fy > ghy) return OC_OUTSIDE_FIELD;
    return OC_OK;
}

// Mask descriptor; `values_dev` is the device copy of the lattice.
int mask_tables(oc_context* ctx, const oc_band_geom& g, const oc_phase_mask& m, const double* values_dev,
                MaskDev* md) {
    std::memset(md, 0, sizeof *md);
    const double psi = och::wrap_angle(m.orientation);
    md->ap_kind = m.aperture.kind;
    md->ap_a = m.aperture.a;
    md->ap_b = m.aperture.b;
    md->values = values_dev;
    md->mnx = (int)m.nx;
    md->mny = (int)m.ny;
    md->mdx = m.dx;
    md->mdy = m.dy;
    md->cpsi = std::cos(psi);
    md->spsi = std::sin(psi);
    md->cx = m.cx;
    md->cy = m.cy;
    md->fnx = (int)g.nx;
    md->fny = (int)g.ny;
    md->fdx = g.dx;
    md->fdy = g.dy;
    if (psi != 0.0) {
        md->mode = MASK_GENERAL;
        return OC_OK;
    }
    md->mode = MASK_SEPARABLE;
    const size_t nx = g.nx, ny = g.ny;
    std::string key = fmtkey("mask", {(double)nx, (double)ny, g.dx, g.dy, (double)m.nx, (double)m.ny, m.dx, m.dy,
                                      m.cx, m.cy});
    const void* dev = find_table(ctx, key);
    if (!dev) {
        std::vector<MaskAxis> h(nx + ny);
        auto fill = [](MaskAxis* out, size_t n, double d, size_t mn, double md_, double c) {
            for (size_t i = 0; i < n; ++i) {
                const double x = ((double)i - ((double)n - 1.0) / 2.0) * d;
                MaskAxis a{};
                a.r = x - c;
                const double f = a.r / md_ + ((double)mn - 1.0) / 2.0;
                const bool inside = !(f < 0.0 || f > (double)mn - 1.0);
                a.in = inside ? 1 : 0;
                if (inside) {
                    const size_t k = (size_t)std::floor(f);
                    a.i0 = (int)k;
                    a.i1 = (int)std::min(k + 1, mn - 1);
                    a.t = f - (double)k;
                }
                out[i] = a;
            }
        };
        fill(h.data(), nx, g.dx, m.nx, m.dx, m.cx);
        fill(h.data() + nx, ny, g.dy, m.ny, m.dy, m.cy);
        int st = store_table(ctx, key, h.data(), h.size() * sizeof(MaskAxis), &dev);
        if (st) return st;
    }
    md->sx = static_cast<const MaskAxis*>(dev);
    md->sy = md->sx + nx;
    return OC_OK;
}

int pick_pow2_divisor(size_t n, int cap) {
    int b = 1;
    while (b * 2 <= cap && n % (size_t)(b * 2) == 0) b *= 2;
    return b;
}

// Host validation + guards for one problem (reference order per engine).
void prepare(Prob& p) {
    const oc_band_geom& g = p.g;
    if (p.engine == ENG_NONE || p.engine == ENG_GS_PROJECT || p.engine == ENG_MEASURE) {
        if (g.nx == 0 || g.ny == 0) { p.status = OC_EMPTY_GRID; return; }
        p.nxp = g.nx;
        p.nyp = g.ny;
        p.launch = true;
        return;
    }
    if (p.engine == ENG_FRAUNHOFER) {  // fraunhofer.rs:269-287, 351-358
        if (!std::isfinite(p.z) || p.z == 0.0) { p.status = OC_INVALID_DISTANCE; return; }
        if (!std::isfinite(g.lambda) || g.lambda <= 0.0) { p.status = OC_INVALID_WAVELENGTH; return; }
        if (g.nx == 0 || g.ny == 0) { p.status = OC_EMPTY_GRID; return; }
        p.fh_rep = och::fraunhofer_report(g, p.z, p.fh_cfg);
        if (p.fh_cfg.strict && !p.fh_rep.regime_ok) {
            p.check_only = true;  // the finiteness check keeps its precedence
            p.pending = OC_REGIME_VIOLATION;
            return;
        }
        p.nxp = g.nx;  // unpadded
        p.nyp = g.ny;
        p.ops.pre = PRE_FRAUN;
        p.launch = true;
        return;
    }
    if (p.engine == ENG_RS) {
        if (!std::isfinite(p.z) || p.z == 0.0) { p.status = OC_INVALID_DISTANCE; return; }
    } else if (!std::isfinite(p.z)) { p.status = OC_INVALID_DISTANCE; return; }
    if (!std::isfinite(g.lambda) || g.lambda <= 0.0) { p.status = OC_INVALID_WAVELENGTH; return; }
    if (p.engine == ENG_ASM && (!std::isfinite(p.medium) || p.medium <= 0.0)) { p.status = OC_INVALID_MEDIUM; return; }
    if (g.nx == 0 || g.ny == 0) { p.status = OC_EMPTY_GRID; return; }
    if (p.engine != ENG_RS && p.z == 0.0) {  // identity after the finiteness check
        p.check_only = true;
        p.pending = OC_OK;
        return;
    }
    if (p.engine == ENG_ASM) {
        oc_asm_config c{p.padding, p.band_limited, p.strict, p.max_padding, p.medium};
        oc_asm_report r = och::asm_analyze(g, p.z, c, c.padding);
        if (c.strict && !r.nyquist_ok) {
            double pad = c.padding;
            while (!r.nyquist_ok && pad < c.max_padding) {
                pad = std::fmin(pad * 2.0, c.max_padding);
                r = och::asm_analyze(g, p.z, c, pad);
            }
        }
        p.asm_rep = r;
        if (c.strict && !r.nyquist_ok) { p.check_only = true; p.pending = OC_ALIASING_RISK; return; }
        p.nxp = r.nx_pad;
        p.nyp = r.ny_pad;
    } else if (p.engine == ENG_FRESNEL_CONV) {  // fresnel.rs:531-553: no densification
        oc_fresnel_config c{p.padding, p.strict, 0, p.max_padding};
        oc_fresnel_report r = och::fresnel_analyze(g, p.z, c, c.padding);
        p.fr_rep = r;
        if (c.strict && !r.kernel_nyquist_ok) { p.check_only = true; p.pending = OC_ALIASING_RISK; return; }
        p.nxp = r.nx_pad;
        p.nyp = r.ny_pad;
    } else if (p.engine == ENG_FRESNEL) {
        oc_fresnel_config c{p.padding, p.strict, 0, p.max_padding};
        oc_fresnel_report r = och::fresnel_analyze(g, p.z, c, c.padding);
        if (c.strict && !r.nyquist_ok) {
            double pad = c.padding;
            while (!r.nyquist_ok && pad < c.max_padding) {
                pad = std::fmin(pad * 2.0, c.max_padding);
                r = och::fresnel_analyze(g, p.z, c, pad);
            }
        }
        p.fr_rep = r;
        if (c.strict && !r.nyquist_ok) { p.check_only = true; p.pending = OC_ALIASING_RISK; return; }
        p.nxp = r.nx_pad;
        p.nyp = r.ny_pad;
    } else {
        oc_rs_report r = och::rs_analyze(g, p.z, p.padding);
        p.rs_rep = r;
        if (p.strict && !r.kernel_nyquist_ok) { p.check_only = true; p.pending = OC_ALIASING_RISK; return; }
        p.nxp = r.nx_pad;
        p.nyp = r.ny_pad;
    }
    p.ref_nxp = p.nxp;
    p.ref_nyp = p.nyp;
    if (conv_engine(p.engine) && p.nxp + 1 >= 2 * g.nx && p.nyp + 1 >= 2 * g.ny) {
        // Sampled-kernel convolution with P >= 2N - 1 on both axes: every
        // cropped output sample is sum_m u[m] h((n - m) d) over offsets
        // |n - m| < N, which the DFT-order sampling places at the same
        // coordinates for any such P, and no term wraps around.  The result
        // is the same linear convolution for every padded length >= 2N - 1,
        // so a non-power-of-two P runs at the next power of two (the
        // register-resident kernels) up to 8192, square.  The report keeps
        // the reference's padded size.
        size_t q = 64;
        while (q < std::max(p.nxp, p.nyp)) q <<= 1;
        if (q <= 8192 && (q != p.nxp || q != p.nyp)) p.nxp = p.nyp = q;
    }
    p.launch = true;
}



Layout choose_layout(const FftPlanDev& prow, const FftPlanDev& pcol, size_t esz, size_t ny, size_t nyp, bool rs) {
    Layout L;
    L.row_bytes = plan_lds_elems(prow) * esz;
    L.col_bytes = plan_lds_elems(pcol) * esz;
    const int tr = rs ? 2 : 1;
    if (L.row_bytes > kLdsBudget || L.col_bytes * tr > kLdsBudget) return L;
    int cnt = (int)std::min<size_t>(kLdsBudget / L.row_bytes, (size_t)std::max(1, 1024 / prow.threads));
    cnt = std::max(1, std::min(cnt, 16));
    cnt = std::min<int>(cnt, (int)std::max<size_t>(nyp, ny));
    int bcap = (int)std::min<size_t>(kLdsBudget / (L.col_bytes * tr), (size_t)std::max(1, 1024 / (pcol.threads * tr)));
    bcap = std::max(1, std::min(bcap, 16));
    L.cnt = cnt;
    L.B = pick_pow2_divisor((size_t)prow.n, bcap);
    L.ccnt = L.B;
    L.ok = true;
    return L;
}


void free_batch(oc_context* ctx, Batch& b) {
    if (!b.owned.empty()) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : b.owned) (void)hipFree(p);
    b.owned.clear();
    b.pd.clear();
    b.idx.clear();
    b.dprobs = nullptr;
}

QueuedSet& queued_set(oc_context* ctx) {
    if (!ctx->queued) ctx->queued = new QueuedSet();
    return *static_cast<QueuedSet*>(ctx->queued);
}
void release_queued(oc_context* ctx) {
    delete static_cast<QueuedSet*>(ctx->queued);
    ctx->queued = nullptr;
}

int batch_alloc(oc_context* ctx, Batch& b, bool persistent, const std::string& name, size_t bytes, void** out) {
    if (!persistent || ctx->arena) return ensure_buf(ctx, name, bytes, out);
    HIP_TRY(ctx, hipMalloc(out, std::max<size_t>(bytes, 256)));
    b.owned.push_back(*out);
    return OC_OK;
}

// Descriptors for every prepared problem with launch == true.  A persistent
// batch owns its intermediates and descriptor array (reusable across calls,
// capturable in a graph); a transient one uses the context's shared buffers.
// `subset` (optional) restricts the batch to those problem indices; `tag`
// names the transient buffers so that several batches can be in flight.
// Hash of a descriptor set for the descriptor cache (a match is confirmed
// by a full compare, so only speed and spread matter): FNV-1a's multiply
// over 8-byte words in four independent lanes, folded at the end.  A
// byte-at-a-time FNV-1a took ~10 us for the 16-problem batches of a step.
static uint64_t desc_hash(const unsigned char* p, size_t bytes) {
    constexpr uint64_t kPrime = 1099511628211ull;
    uint64_t h[4] = {1469598103934665603ull, 0x9e3779b97f4a7c15ull, 0xc2b2ae3d27d4eb4full, 0x165667b19e3779f9ull};
    const size_t words = bytes / 8;
    size_t i = 0;
    for (; i + 4 <= words; i += 4)
        for (int l = 0; l < 4; ++l) {
            uint64_t w;
            std::memcpy(&w, p + 8 * (i + l), 8);
            h[l] = (h[l] ^ w) * kPrime;
        }
    for (; i < words; ++i) {
        uint64_t w;
        std::memcpy(&w, p + 8 * i, 8);
        h[0] = (h[0] ^ w) * kPrime;
    }
    for (size_t j = 8 * words; j < bytes; ++j) h[1] = (h[1] ^ p[j]) * kPrime;
    uint64_t r = bytes;
    for (int l = 0; l < 4; ++l) r = (r ^ h[l] ^ (h[l] >> 29)) * kPrime;
    return r;
}

int build_batch(oc_context* ctx, bool dbl, std::vector<Prob>& probs, int* dflags, Batch& b, bool persistent,
                const std::vector<int>* subset, const std::string& tag, int force_v) {
    const size_t esz = dbl ? 16 : 8;
    b.dbl = dbl;
    size_t work_elems = 0, kbuf_elems = 0;
    std::vector<size_t> woff, koff;
    std::vector<int> cand;
    std::vector<int> all;
    if (!subset) {
        for (int i = 0; i < (int)probs.size(); ++i) all.push_back(i);
        subset = &all;
    }
    for (int i : *subset) {
        Prob& p = probs[i];
        if (p.status != OC_OK || !p.launch) continue;
        if (p.nxp > (1u << 20) || p.nyp > (1u << 20) || p.g.nx > (1u << 20) || p.g.ny > (1u << 20)) {
            p.status = fail(ctx, OC_ERR_UNSUPPORTED, "grid too large");
            continue;
        }
        woff.push_back(work_elems);
        work_elems += p.g.ny * p.nxp;
        koff.push_back(kbuf_elems);
        if (conv_engine(p.engine)) kbuf_elems += p.nyp * p.nxp;
        cand.push_back(i);
    }
    // power-of-two fast path: one shared square padded size, uniform RS-ness
    int fG = 0, fB = 0, fC = 0;
    if (!cand.empty() && !ctx->force_generic) {
        const size_t n0 = probs[cand[0]].nxp;
        bool ok = n0 >= 32 && n0 <= 8192 && (n0 & (n0 - 1)) == 0;
        const bool rs0 = conv_engine(probs[cand[0]].engine);
        for (int i : cand)
            ok = ok && probs[i].nxp == n0 && probs[i].nyp == n0 && conv_engine(probs[i].engine) == rs0 &&
                 probs[i].path != 0;
        // a scheduled layout (every problem of a batch length carries the same pick)
        const int sv0 = probs[cand[0]].path == 1 ? probs[cand[0]].variant : -1;
        const int v = force_v >= 0 ? force_v : (sv0 >= 0 ? sv0 : ctx->variant_for((int)n0, dbl));
        if (ok && ocl::fast_supported(dbl, (int)n0, v, &fG, &fB, &fC)) {
            b.fast = true;
            b.fast_n = (int)n0;
            b.fast_v = v;
            b.rs = rs0;
            bool half = true;
            for (int i : cand) half = half && probs[i].g.nx * 2 == n0 && probs[i].g.ny * 2 == n0;
            b.half = half;
        }
    }
    // complex64 Fresnel batches without post-ops factor per axis: two passes,
    // each with its own layout (the pair shares the blocked width B, and the
    // x embedding offset must fall on a block boundary)
    if (b.fast && !dbl && !b.rs) {
        const std::pair<int, int> sv =
            force_v >= 0 ? std::make_pair(force_v, force_v)
                         : (probs[cand[0]].path == 1 ? std::make_pair(b.fast_v, b.fast_v) : ctx->sep_variants_for(b.fast_n));
        int cG, cB, cC, rG, rB, rC;
        bool sep = ocl::fast_supported(false, b.fast_n, sv.first, &cG, &cB, &cC) &&
                   ocl::fast_supported(false, b.fast_n, sv.second, &rG, &rB, &rC) && cB == rB;
        for (int i : cand) {
            const Prob& p = probs[i];
            const size_t ox = (p.nxp - p.g.nx) / 2;
            sep = sep && p.engine == ENG_FRESNEL && (p.ops.pre == PRE_NONE || p.ops.pre == PRE_MASK) &&
                  p.ops.post == POST_NONE && ox % (size_t)cB == 0;
        }
        b.sep = sep;
        if (sep) {  // rows per workgroup from the row pass, columns from the column pass
            b.fast_v = sv.first;
            b.fast_v_row = sv.second;
            fG = rG;
            fB = cB;
            fC = cC;
        }
    }
    void *work = nullptr, *kbuf = nullptr;
    if (!cand.empty()) {
        int st = batch_alloc(ctx, b, persistent, "work" + tag, work_elems * esz, &work);
        if (st) return st;
        if (kbuf_elems) {
            st = batch_alloc(ctx, b, persistent, "kbuf" + tag, kbuf_elems * esz, &kbuf);
            if (st) return st;
        }
    }
    // Phase masks shared by several problems (the bands of one field): the
    // transmission does not depend on the wavelength, so it is evaluated
    // once into a buffer and the first pass multiplies by it.
    std::map<std::string, int> mask_uses;
    auto mask_key = [](const Prob& p) {
        const oc_phase_mask& m = *p.mask;
        return fmtkey("tm", {(double)(uintptr_t)m.values, (double)m.nx, (double)m.ny, m.dx, m.dy, m.orientation,
                             (double)m.aperture.kind, m.aperture.a, m.aperture.b, m.cx, m.cy, (double)p.g.nx,
                             (double)p.g.ny, p.g.dx, p.g.dy});
    };
    for (int i : cand)
        if (probs[i].mask && probs[i].ops.pre == PRE_MASK) mask_uses[mask_key(probs[i])] += 1;
    std::map<std::string, void*> trans;
    int base_row = 0, base_col = 0, base_kgen = 0;
    for (size_t a = 0; a < cand.size(); ++a) {
        Prob& p = probs[cand[a]];
        ProbDev d;
        std::memset(&d, 0, sizeof d);
        d.nx = (int)p.g.nx;
        d.ny = (int)p.g.ny;
        d.px = (int)p.nxp;
        d.py = (int)p.nyp;
        d.ox = (int)((p.nxp - p.g.nx) / 2);
        d.oy = (int)((p.nyp - p.g.ny) / 2);
        // a scheduled radix plan is for the square run length
        const bool sr = !b.fast && p.path == 0 && p.nrad > 0 && p.nxp == p.nyp;
        int st = make_plan(ctx, d.px, dbl, &d.prow, sr ? p.rad : nullptr, sr ? p.nrad : 0);
        if (st) return st;
        st = make_plan(ctx, d.py, dbl, &d.pcol, sr ? p.rad : nullptr, sr ? p.nrad : 0);
        if (st) return st;
        const bool rs = conv_engine(p.engine);
        Layout L;
        if (b.fast) {
            const size_t slots = (size_t)b.fast_n + b.fast_n / 16;
            L.cnt = fG;
            L.B = fB;
            L.ccnt = fC;
            L.row_bytes = slots * esz;
            L.col_bytes = slots * esz;
            L.ok = true;
        } else {
            L = choose_layout(d.prow, d.pcol, esz, p.g.ny, p.nyp, rs);
        }
        if (!L.ok) {
            p.status = fail(ctx, OC_ERR_UNSUPPORTED, "padded length exceeds the on-chip transform size");
            continue;
        }
        const int tr = (rs && !b.fast) ? 2 : 1;
        d.B = L.B;
        d.rows_cnt = L.cnt;
        d.tiles_row = (d.ny + L.cnt - 1) / L.cnt;
        d.tiles_col = b.sep ? (d.nx + L.ccnt - 1) / L.ccnt : d.px / L.ccnt;
        // the power-of-two kernels transform only the distinct kernel rows 0..P/2
        d.tiles_kgen = rs ? ((b.fast ? d.py / 2 + 1 : d.py) + L.cnt - 1) / L.cnt : 0;
        d.base_row = base_row;
        d.base_col = base_col;
        d.base_kgen = base_kgen;
        base_row += d.tiles_row;
        base_col += d.tiles_col;
        base_kgen += d.tiles_kgen;
        d.in = p.in;
        d.out = p.out;
        d.work = static_cast<unsigned char*>(work) + woff[a] * esz;
        d.kbuf = rs ? static_cast<unsigned char*>(kbuf) + koff[a] * esz : nullptr;
        d.scale = 1.0 / (double)(p.nxp * p.nyp);
        d.flag = dflags + cand[a];
        d.ops = p.ops;
        if (p.engine == ENG_ASM || p.engine == ENG_FRESNEL || conv_engine(p.engine) || p.engine == ENG_FRAUNHOFER) {
            st = engine_tables(ctx, p, dbl, &d.eng);
            if (st) return st;
        } else {
            d.eng.kind = p.engine;
        }
        if (p.mask) {
            st = mask_tables(ctx, p.g, *p.mask, p.mask->values, &d.mask);
            if (st) return st;
            if (d.ops.pre == PRE_MASK) {
                const std::string key = mask_key(p);
                if (mask_uses[key] > 1) {
                    auto it = trans.find(key);
                    if (it == trans.end()) {
                        void* tb = nullptr;
                        const std::string name = "trans" + tag + "_" + std::to_string(trans.size());
                        st = batch_alloc(ctx, b, persistent, name.c_str(), p.g.nx * p.g.ny * esz, &tb);
                        if (st) return st;
                        HIP_TRY(ctx, timed(ctx, 5, [&] {
                                    return ocl::launch_mask_trans(dbl, tb, d.mask, (int)p.g.nx, (int)p.g.ny,
                                                                  ctx->stream);
                                }));
                        it = trans.emplace(key, tb).first;
                    }
                    d.ops.pre = PRE_TMUL;
                    d.ops.trans = it->second;
                }
            }
        }
        auto rnd = [](int t) { return ((t + 63) / 64) * 64; };
        const int brow = b.fast ? (b.fast_n / 16) * L.cnt : rnd(d.prow.threads * L.cnt);
        const int bcol = b.fast ? (b.fast_n / 16) * L.ccnt : rnd(d.pcol.threads * d.B * tr);
        b.pc[0].block = std::max(b.pc[0].block, brow);
        b.pc[0].lds = std::max(b.pc[0].lds, L.row_bytes * L.cnt);
        b.pc[1].block = std::max(b.pc[1].block, bcol);
        b.pc[1].lds = std::max(b.pc[1].lds, L.col_bytes * L.ccnt * tr);
        if (rs) {
            b.pc[3].block = std::max(b.pc[3].block, brow);
            b.pc[3].lds = std::max(b.pc[3].lds, L.row_bytes * L.cnt);
        }
        const int qk =
            (d.ops.post == POST_NONE || d.ops.post == POST_GS_PHASE || d.ops.post == POST_ADJ_GRAD) ? d.ops.post : -1;
        b.post_kind = b.post_kind == -2 ? qk : (b.post_kind == qk ? qk : -1);
        const int ek = (p.engine == ENG_ASM || p.engine == ENG_FRESNEL || p.engine == ENG_GS_PROJECT) ? p.engine : -1;
        b.col_kind = b.col_kind == -2 ? ek : (b.col_kind == ek ? ek : -1);
        const int pk =
            (d.ops.pre == PRE_NONE || d.ops.pre == PRE_MASK || d.ops.pre == PRE_TMUL || d.ops.pre == PRE_POLAR)
                ? d.ops.pre
                : -1;
        // mixes of none / mask / shared transmission keep a specialised
        // kernel that branches per problem
        auto masklike = [](int k) { return k == PRE_NONE || k == PRE_MASK || k == PRE_TMUL || k == PRE_MASKS; };
        if (b.pre_kind == -2 || b.pre_kind == pk) b.pre_kind = pk;
        else if (masklike(b.pre_kind) && masklike(pk)) b.pre_kind = PRE_MASKS;
        else b.pre_kind = -1;
        b.pd.push_back(d);
        b.idx.push_back(cand[a]);
    }
    if (b.pd.empty()) return OC_OK;
    b.pc[0].total = base_row;
    b.pc[1].total = base_col;
    b.pc[3].total = base_kgen;
    b.pc[2] = b.pc[0];
    if (b.sep) b.pc[0].total = 0;  // pass 1: columns, pass 2: rows
    if (b.fast && !b.sep && b.pre_kind == PRE_TMUL) {
        bool even = true;
        for (const ProbDev& d : b.pd) even = even && d.tiles_row == b.pd[0].tiles_row;
        for (ProbDev& d : b.pd) d.ilv = even ? 1 : 0;
    }
    bool fraun = true;
    for (const ProbDev& d : b.pd) fraun = fraun && d.eng.kind == ENG_FRAUNHOFER;
    if (fraun) b.pc[2].total = 0;  // the column pass writes the far field
    const size_t bytes = b.pd.size() * sizeof(ProbDev);
    const bool cached = !ctx->arena && !persistent && bytes <= oc_context::kDescSlotBytes;
    int st = OC_OK;
    if (cached) {
        if (!ctx->desc_pool)
            HIP_TRY(ctx, hipMalloc(&ctx->desc_pool, oc_context::kDescSlots * oc_context::kDescSlotBytes));
        const unsigned char* hb = reinterpret_cast<const unsigned char*>(b.pd.data());
        const uint64_t h = desc_hash(hb, bytes);
        int hit = -1, lru = 0;
        for (int k = 0; k < oc_context::kDescSlots; ++k) {
            if (ctx->desc_hash[k] == h && ctx->desc_host[k].size() == bytes &&
                std::memcmp(ctx->desc_host[k].data(), hb, bytes) == 0) {
                hit = k;
                break;
            }
            if (ctx->desc_used[k] < ctx->desc_used[lru]) lru = k;
        }
        const int k = hit >= 0 ? hit : lru;
        ctx->desc_used[k] = ++ctx->desc_clock;
        b.dprobs = static_cast<unsigned char*>(ctx->desc_pool) + (size_t)k * oc_context::kDescSlotBytes;
        if (hit >= 0) return OC_OK;
        ctx->desc_hash[k] = h;
        ctx->desc_host[k].assign(hb, hb + bytes);
    } else if (!ctx->arena) {
        st = batch_alloc(ctx, b, persistent, "probs" + tag, bytes, &b.dprobs);
    }
    if (st) return st;
    if (ctx->arena) {
        const void* d = nullptr;
        st = arena_upload(ctx, b.pd.data(), bytes, &d);
        if (st) return st;
        b.dprobs = const_cast<void*>(d);
    } else if (persistent) {
        HIP_TRY(ctx, hipMemcpy(b.dprobs, b.pd.data(), bytes, hipMemcpyHostToDevice));
    } else {
        // the slot is free once its previous copy ran (not the kernels
        // queued after that copy)
        const int slot = ctx->pin_next;
        ctx->pin_next = (slot + 1) % oc_context::kPinSlots;
        if (ctx->pinned_done[slot]) HIP_TRY(ctx, hipEventSynchronize(ctx->pinned_done[slot]));
        else HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->pinned_done[slot], hipEventDisableTiming));
        st = ensure_pinned(ctx, slot, bytes);
        if (st) return st;
        std::memcpy(ctx->pinned[slot], b.pd.data(), bytes);
        HIP_TRY(ctx, hipMemcpyAsync(b.dprobs, ctx->pinned[slot], bytes, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->pinned_done[slot], ctx->stream));
    }
    return OC_OK;
}


// Enqueues the passes of `b` selected by `mask` (bit i = pass i), in the
// order kernel generation, row forward, column, row inverse.
int launch_batch(oc_context* ctx, Batch& b, unsigned mask) {
    if (b.pd.empty()) return OC_OK;
    const int nprob = (int)b.pd.size();
    const ProbDev* dprobs = static_cast<const ProbDev*>(b.dprobs);
    const int order[6] = {3, 0, 1, 2, PASS_GS_ROW, PASS_GS_AMP};
    for (int k = 0; k < 6; ++k) {
        const int pass = order[k];
        if (!(mask & (1u << pass))) continue;
        const bool fused = pass == PASS_GS_ROW || pass == PASS_GS_AMP;
        if (fused && !b.fast) return fail(ctx, OC_ERR_UNSUPPORTED, "fused GS row pass needs N = 2^k");
        const PassCfg& c = b.pc[fused ? 0 : pass];
        if (c.total <= 0) continue;
        const int grid = std::min(c.total, ctx->num_cu * 8);
        const int tpass = b.sep ? (pass == 1 ? PASS_T_SEP_COL : PASS_T_SEP_ROW) : (fused ? 4 : pass);
        hipError_t e = timed(ctx, tpass, [&] {
            if (b.fast) {
                const int v = (b.sep && pass == 2) ? b.fast_v_row : b.fast_v;
                const ocd::FastSel sel{pass,       b.fast_n,   v,             b.rs ? 1 : 0, b.post_kind,
                                       b.col_kind, b.pre_kind, b.sep ? 1 : 0, b.half ? 1 : 0};
                return b.dbl ? ocl::fast_launch_c128(sel, dprobs, nprob, c.total, grid, ctx->stream)
                             : ocl::fast_launch_c64(sel, dprobs, nprob, c.total, grid, ctx->stream);
            }
            return b.dbl ? ocl::launch_pass_c128(pass, dprobs, nprob, c.total, c.block, c.lds, grid, ctx->stream)
                         : ocl::launch_pass_c64(pass, dprobs, nprob, c.total, c.block, c.lds, grid, ctx->stream);
        });
        if (e != hipSuccess) return fail(ctx, OC_ERR_DEVICE, std::string("pass launch: ") + hipGetErrorString(e));
    }
    return OC_OK;
}

void evict_tables(oc_context* ctx) {
    if (!ctx->arena && ctx->tables.size() > 1024) {  // between calls only: tables in flight stay valid
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& kv : ctx->tables) (void)hipFree(kv.second.dev);
        ctx->tables.clear();
    }
}

// A length the generic LDS transform holds on chip: a plan within the
// stage limit and the LDS budget (ping-pong plans twice over).
static bool line_fits(int n, size_t esz) {
    std::vector<int> rad;
    bool generic;
    if (n <= 0 || !make_radices(n, rad, generic)) return false;
    return (size_t)lds_stride(n) * (generic ? 2 : 1) * esz <= kLdsBudget;
}

bool fits_on_chip(const Prob& p, bool dbl) {
    const size_t esz = dbl ? 16 : 8;
    std::vector<int> rad;
    bool gy;
    if (!line_fits((int)p.nxp, esz) || !make_radices((int)p.nyp, rad, gy)) return false;
    return (size_t)lds_stride((int)p.nyp) * (gy ? 2 : 1) * esz * (conv_engine(p.engine) ? 2 : 1) <= kLdsBudget;
}

// n = a * (n / a) with both factors on chip and a as close to sqrt(n) as
// possible; 0 when n has no such split (a prime beyond the on-chip size).
static int split_len(int n, size_t esz) {
    int best = 0;
    double bd = 1e300;
    for (int a = 2; (long long)a * a <= n; ++a) {
        if (n % a) continue;
        const int b = n / a;
        if (!line_fits(a, esz) || !line_fits(b, esz)) continue;
        const double d = std::fabs(std::log((double)b / (double)a));
        if (d < bd) {
            bd = d;
            best = a;
        }
    }
    return best;
}

// The 1-D transforms of every line of `cur` along axis x (rows) or y
// (columns) of the px x py buffer, direction dir; a length beyond the
// on-chip transform runs as four steps through `tmp`, and the result then
// sits in tmp (the two are swapped).
static int big_axis(oc_context* ctx, bool dbl, void*& cur, void*& tmp, long long px, long long py, bool along_x,
                    int dir) {
    const size_t esz = dbl ? 16 : 8;
    const int L = (int)(along_x ? px : py);
    const long long lines = along_x ? py : px;  // transforms along the axis
    // one LDS pass of `count` lines of length n
    auto pass = [&](const void* in, void* out, int n, long long count, long long inner, LineView vin, LineView vout,
                    int tw_len, int tw_outer) -> int {
        FftPlanDev pl;
        int st = make_plan(ctx, n, dbl, &pl);
        if (st) return st;
        const void* twl = nullptr;
        if (tw_len > 0 && (st = get_twiddles(ctx, tw_len, dbl, &twl))) return st;
        const size_t line_bytes = (size_t)lds_stride(n) * (pl.generic ? 2 : 1) * esz;
        int cnt = (int)std::min<size_t>(kLdsBudget / line_bytes, (size_t)std::max(1, 1024 / pl.threads));
        cnt = std::max(1, std::min(cnt, 16));
        LineSpec s{n, cnt, count, inner, vin, vout, tw_len, tw_outer};
        HIP_TRY(ctx, timed(ctx, 5, [&] {
                    return ocl::big_line(dbl, dir, in, out, s, pl, twl, line_bytes * cnt, ctx->num_cu * 8,
                                         ctx->stream);
                }));
        return OC_OK;
    };
    if (line_fits(L, esz)) {
        // rows: line r at r px; columns: line c at c, element stride px
        const LineView v = along_x ? LineView{px, 0, 1} : LineView{0, 1, px};
        return pass(cur, cur, L, lines, along_x ? 1 : px, v, v, 0, 0);
    }
    const int L1 = split_len(L, esz);
    if (!L1) return fail(ctx, OC_ERR_UNSUPPORTED, "a padded length with a prime factor beyond the on-chip transform");
    const int L2 = L / L1;
    int st;
    if (along_x) {
        // step 1: row r, n2 < L2: elements n1 L2 + n2 (in place), twiddle W_L^{n2 k1}
        const LineView a{px, 1, L2};
        if ((st = pass(cur, cur, L1, py * L2, L2, a, a, L, 0))) return st;
        // step 2: row r, k1 < L1: elements k1 L2 + n2 -> k1 + L1 k2
        if ((st = pass(cur, tmp, L2, py * L1, L1, LineView{px, L2, 1}, LineView{px, 1, L1}, 0, 0))) return st;
    } else {
        // columns, lines numbered (n2 or k1) * px + c so neighbouring lines are neighbouring columns
        const LineView a{px, 1, (long long)L2 * px};
        if ((st = pass(cur, cur, L1, px * L2, px, a, a, L, 1))) return st;
        if ((st = pass(cur, tmp, L2, px * L1, px, LineView{(long long)L2 * px, 1, px},
                       LineView{px, 1, (long long)L1 * px}, 0, 0)))
            return st;
    }
    std::swap(cur, tmp);
    return OC_OK;
}

int launch_big(oc_context* ctx, bool dbl, Prob& p, int* flag) {
    if (p.engine != ENG_ASM && p.engine != ENG_FRESNEL && !conv_engine(p.engine)) {
        p.status = fail(ctx, OC_ERR_UNSUPPORTED, "padded length exceeds the on-chip transform size");
        return OC_OK;
    }
    const size_t esz = dbl ? 16 : 8;
    const long long px = (long long)p.nxp, py = (long long)p.nyp;
    if ((!line_fits((int)px, esz) && !split_len((int)px, esz)) || (!line_fits((int)py, esz) && !split_len((int)py, esz))) {
        p.status = fail(ctx, OC_ERR_UNSUPPORTED, "a padded length with a prime factor beyond the on-chip transform");
        return OC_OK;
    }
    const size_t bytes = (size_t)px * (size_t)py * esz;
    const bool conv = conv_engine(p.engine);
    void *a, *b, *k = nullptr;
    int st;
    if ((st = ensure_buf(ctx, "big_a", bytes, &a)) || (st = ensure_buf(ctx, "big_b", bytes, &b))) return st;
    if (conv && (st = ensure_buf(ctx, "big_k", bytes, &k))) return st;
    ProbDev d;
    std::memset(&d, 0, sizeof d);
    d.nx = (int)p.g.nx;
    d.ny = (int)p.g.ny;
    d.px = (int)px;
    d.py = (int)py;
    d.ox = (int)((p.nxp - p.g.nx) / 2);
    d.oy = (int)((p.nyp - p.g.ny) / 2);
    d.in = p.in;
    d.out = p.out;
    d.scale = 1.0 / (double)(p.nxp * p.nyp);
    d.flag = flag;
    d.ops = p.ops;
    if ((st = engine_tables(ctx, p, dbl, &d.eng))) return st;
    if (p.mask && (st = mask_tables(ctx, p.g, *p.mask, p.mask->values, &d.mask))) return st;
    void* kspec = nullptr;
    void *fld = a, *tmp = b;
    if (conv) {  // the kernel's spectrum first, in k or a; the field then takes the other
        void *kc = k, *kt = a;
        HIP_TRY(ctx, ocl::big_kernel(dbl, d, kc, ctx->stream));
        if ((st = big_axis(ctx, dbl, kc, kt, px, py, true, -1)) || (st = big_axis(ctx, dbl, kc, kt, px, py, false, -1)))
            return st;
        kspec = kc;
        fld = kt;
    }
    HIP_TRY(ctx, ocl::big_embed(dbl, d, fld, ctx->stream));
    if ((st = big_axis(ctx, dbl, fld, tmp, px, py, true, -1)) || (st = big_axis(ctx, dbl, fld, tmp, px, py, false, -1)))
        return st;
    HIP_TRY(ctx, ocl::big_transfer(dbl, d, fld, kspec, ctx->stream));
    if ((st = big_axis(ctx, dbl, fld, tmp, px, py, false, 1)) || (st = big_axis(ctx, dbl, fld, tmp, px, py, true, 1)))
        return st;
    HIP_TRY(ctx, ocl::big_crop(dbl, d, fld, ctx->stream));
    return OC_OK;
}

// Queues all problems on the context stream (host-side validation, the
// standalone finiteness checks, then one batch per padded-size group) with
// their finiteness flags in dflags[0..n).  No synchronisation.
int launch_problems(oc_context* ctx, bool dbl, std::vector<Prob>& probs, int* dflags, bool zero_flags) {
    evict_tables(ctx);
    // A masked problem is Element::apply then the engine (beamline.rs:
    // 118-123): the mask's footprint error comes first, also for the z == 0
    // identity and a failed guard.
    for (auto& p : probs) {
        prepare(p);
        if (p.status == OC_OK && p.mask) {
            int st = mask_footprint(p.g, *p.mask);
            if (st) p.status = st;
        }
    }
    apply_schedules(ctx, dbl, probs);
    const int np = (int)probs.size();
    if (zero_flags) HIP_TRY(ctx, ocl::launch_fill_int(dflags, 0, np, ctx->stream));
    for (int i = 0; i < np; ++i) {
        Prob& p = probs[i];
        if (p.status != OC_OK || !p.check_only) continue;
        const size_t n = p.g.nx * p.g.ny;
        if (p.mask) {
            // the engine validates the masked field; the identity returns it
            // (a failed guard writes nothing)
            MaskDev md;
            int st = mask_tables(ctx, p.g, *p.mask, p.mask->values, &md);
            if (st) return st;
            void* o = p.pending == OC_OK ? p.out : nullptr;
            HIP_TRY(ctx, dbl ? ocl::launch_mask_apply_c128(p.in, o, md, (int)p.g.nx, (int)p.g.ny, dflags + i,
                                                           ctx->stream)
                             : ocl::launch_mask_apply_c64(p.in, o, md, (int)p.g.nx, (int)p.g.ny, dflags + i,
                                                          ctx->stream));
            continue;
        }
        HIP_TRY(ctx, ocl::launch_check_finite(dbl, p.in, n, dflags + i, ctx->stream));
        if (p.pending == OC_OK && p.out != p.in) HIP_TRY(ctx, ocl::launch_copy(dbl, p.in, p.out, n, ctx->stream));
    }
    // One batch per padded size that has specialised kernels (uneven grids
    // then each get the power-of-two kernels), one for all other sizes that
    // fit on chip; the rest run one by one through the multi-pass path.
    std::map<std::pair<size_t, size_t>, std::vector<int>> groups;
    std::vector<int> big;
    for (int i = 0; i < np; ++i) {
        const Prob& p = probs[i];
        if (p.status != OC_OK || !p.launch) continue;
        const bool fast =
            !ctx->force_generic && p.nxp == p.nyp && ocl::fast_variants((int)p.nxp) > 0 && p.path != 0;
        if (!fast && p.engine != ENG_FRAUNHOFER && !fits_on_chip(p, dbl)) {
            big.push_back(i);
            continue;
        }
        // Fraunhofer problems (forward transforms only) never share a batch
        // with the others: their batch has no inverse row pass
        const size_t fh = p.engine == ENG_FRAUNHOFER ? ((size_t)1 << 40) : 0;
        groups[fast ? std::make_pair(p.nxp, p.nyp + fh) : std::make_pair((size_t)0, fh)].push_back(i);
    }
    int gi = 0;
    for (auto& kv : groups) {
        Batch b;
        int st = build_batch(ctx, dbl, probs, dflags, b, false, &kv.second, groups.size() > 1 ? std::to_string(gi) : "");
        if (st) return st;
        st = launch_batch(ctx, b);
        if (st) return st;
        ++gi;
    }
    for (int i : big) {
        if (probs[i].nxp > (1u << 20) || probs[i].nyp > (1u << 20)) {
            probs[i].status = fail(ctx, OC_ERR_UNSUPPORTED, "grid too large");
            continue;
        }
        int st = launch_big(ctx, dbl, probs[i], dflags + i);
        if (st) return st;
    }
    return OC_OK;
}

// Consecutive propagations (every problem of step k reads step k-1's
// output, the next pre-op at most a mask or shared transmission; transfer
// function engines, three-pass form) on one power-of-two padded length:
// step 0's forward row and column passes,
// then per later step f_row_link (step k-1's inverse rows, the pre-op, step
// k's forward rows) and step k's column pass, then the last step's inverse
// rows.  The fields between the steps are written only with keep_mid, to
// each problem's own output pointer where it is set.  Returns 0 when
// launched (the steps' problems appended to `queued`, flags at dflags +
// off[k]), 1 when the run does not qualify (nothing launched), or minus an
// error status.
int launch_linked(oc_context* ctx, bool dbl, std::vector<std::vector<Prob>>& run, int* dflags,
                  const std::vector<int>& off, std::vector<std::vector<Prob>>& queued, bool keep_mid) {
    const size_t K = run.size();
    std::vector<std::vector<Prob>> q = run;  // prepared copies
    size_t P = 0;
    for (size_t k = 0; k < K; ++k)
        for (Prob& p : q[k]) {
            prepare(p);
            if (p.status == OC_OK && p.mask) p.status = mask_footprint(p.g, *p.mask);
            if (p.status != OC_OK || !p.launch || p.check_only || (p.engine != ENG_ASM && p.engine != ENG_FRESNEL) ||
                p.nxp != p.nyp || ocl::fast_variants((int)p.nxp) == 0 || (P && p.nxp != P))
                return 1;
            P = p.nxp;
            if (k + 1 < K && !keep_mid) p.out = nullptr;  // stays on chip
        }
    if (ctx->force_generic) return 1;
    std::vector<Batch> b(K);
    for (size_t k = 0; k < K; ++k) {
        int st = build_batch(ctx, dbl, q[k], dflags + off[k], b[k], false, nullptr, "ln" + std::to_string(k));
        if (st) return -st;
        const Batch& x = b[k];
        const bool pre_ok = k == 0 || x.pre_kind == PRE_NONE || x.pre_kind == PRE_MASK || x.pre_kind == PRE_TMUL ||
                            x.pre_kind == PRE_MASKS;
        if (!x.fast || x.sep || x.rs || x.post_kind != POST_NONE || (x.col_kind != ENG_ASM && x.col_kind != ENG_FRESNEL) ||
            !pre_ok ||
            x.pd.size() != q[k].size() || x.fast_v != b[0].fast_v || x.pd[0].B != b[0].pd[0].B)
            return 1;
    }
    const hipError_t ef = ocl::launch_fill_int(dflags + off[0], 0, (int)(K * q[0].size()), ctx->stream);
    if (ef != hipSuccess) return -fail(ctx, OC_ERR_DEVICE, std::string("flags: ") + hipGetErrorString(ef));
    int st = launch_batch(ctx, b[0], (1u << 0) | (1u << 1));
    for (size_t k = 1; k < K && !st; ++k) {
        const int np = (int)b[k].pd.size(), rows = b[k].pc[0].total;
        const ProbDev* da = static_cast<const ProbDev*>(b[k - 1].dprobs);
        const ProbDev* db = static_cast<const ProbDev*>(b[k].dprobs);
        hipError_t e = timed(ctx, 2, [&] {
            const int grid = std::min(rows, ctx->num_cu * 8);
            return dbl ? ocl::fast_launch_link_c128(b[k].pre_kind, b[k].fast_n, b[k].fast_v, da, db, np, rows, grid,
                                                    ctx->stream)
                       : ocl::fast_launch_link_c64(b[k].pre_kind, b[k].fast_n, b[k].fast_v, da, db, np, rows, grid,
                                                   ctx->stream);
        });
        if (e != hipSuccess) return -fail(ctx, OC_ERR_DEVICE, std::string("link launch: ") + hipGetErrorString(e));
        st = launch_batch(ctx, b[k], 1u << 1);
    }
    if (!st) st = launch_batch(ctx, b[K - 1], 1u << 2);
    if (st) return -st;
    for (size_t k = 0; k < K; ++k) queued.push_back(std::move(q[k]));
    run.clear();
    return 0;
}

// Waits for the queued work and turns the flags of launch_problems into
// per-problem statuses; returns the first non-OK one.
int finish_problems(oc_context* ctx, std::vector<Prob>& probs, const int* dflags) {
    const int np = (int)probs.size();
    // Through the context's pinned buffer: a copy into pageable memory is
    // staged and waits by itself, and the stream wait after it cost another
    // ~18 us per call.  The no-allocation form (arena) keeps the pageable
    // copy, since the pinned buffer would be allocated on first use.
    std::vector<int> pageable;
    void* rb = nullptr;
    if (ctx->arena) {
        pageable.assign(std::max(np, 1), 0);
        rb = pageable.data();
    } else {
        const int st = ensure_readback(ctx, sizeof(int) * std::max(np, 1), &rb);
        if (st) return st;
    }
    const int* hflags = static_cast<const int*>(rb);
    if (np) HIP_TRY(ctx, hipMemcpyAsync(rb, dflags, sizeof(int) * np, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timing(ctx);
    int first = OC_OK;
    for (int i = 0; i < np; ++i) {
        Prob& p = probs[i];
        if (p.status == OC_OK) {
            if (p.check_only) p.status = (hflags[i] & 1) ? OC_NON_FINITE_SAMPLE : p.pending;
            else if (hflags[i]) p.status = OC_NON_FINITE_SAMPLE;
        }
        if (first == OC_OK && p.status != OC_OK) first = p.status;
    }
    return first;
}

// Runs all problems.  Device pointers throughout.  Synchronises once at the
// end to read the per-problem finiteness flags.
int run_batch(oc_context* ctx, bool dbl, std::vector<Prob>& probs) {
    void* f;
    int st = ensure_buf(ctx, "flags", sizeof(int) * std::max<size_t>(probs.size(), 1), &f);
    if (st) return st;
    int* dflags = static_cast<int*>(f);
    st = launch_problems(ctx, dbl, probs, dflags);
    if (st) return st;
    return finish_problems(ctx, probs, dflags);
}

Prob make_prob(const oc_band_geom* g, double z, int engine, int band_limited, int strict, double padding,
               double max_padding, double medium, const oc_phase_mask* mask, const void* in, void* out) {
    Prob p;
    p.g = *g;
    p.z = z;
    p.engine = engine;
    p.band_limited = band_limited;
    p.strict = strict;
    p.padding = padding;
    p.max_padding = max_padding;
    p.medium = medium;
    p.mask = mask;
    p.in = in;
    p.out = out;
    if (mask) p.ops.pre = PRE_MASK;
    return p;
}

bool geom_ok(const oc_band_geom* g) { return g != nullptr; }

int probs_from_descs(const oc_prop_desc* d, size_t count, std::vector<Prob>& v) {
    v.reserve(count);
    for (size_t i = 0; i < count; ++i) {
        const oc_prop_desc& x = d[i];
        if ((x.engine < OC_ENGINE_ASM || x.engine > OC_ENGINE_RS) && x.engine != OC_ENGINE_FRESNEL_CONV)
            return OC_ERR_INVALID_ARGUMENT;
        if ((x.engine == OC_ENGINE_RS || x.engine == OC_ENGINE_FRESNEL_CONV) && x.mask) return OC_ERR_INVALID_ARGUMENT;
        v.push_back(make_prob(&x.geom, x.z, x.engine, x.band_limited, x.strict, x.padding, x.max_padding,
                              x.engine == OC_ENGINE_ASM ? x.medium : 1.0, x.mask, x.in, x.out));
    }
    return OC_OK;
}

// Upper bound of what launch_problems allocates for `probs` (unprepared):
// every buffer, table and descriptor rounded up to 256 bytes, one shared
// transmission per distinct mask and padded size (a batch is built per
// padded size), tables counted per problem rather than per key.
size_t arena_need(const std::vector<Prob>& probs, bool dbl) {
    auto a = [](size_t x) { return (std::max<size_t>(x, 1) + 255) & ~(size_t)255; };
    const size_t esz = dbl ? 16 : 8;
    size_t need = a(sizeof(int) * std::max<size_t>(probs.size(), 1)) + 4096;
    std::map<std::string, int> trans;
    for (Prob p : probs) {
        prepare(p);
        if (p.status == OC_OK && p.mask) p.status = mask_footprint(p.g, *p.mask);
        if (p.status != OC_OK) continue;
        const size_t nx = p.g.nx, ny = p.g.ny;
        if (p.mask) {
            const oc_phase_mask& m = *p.mask;
            const std::string k =
                fmtkey("tm", {(double)(uintptr_t)m.values, (double)m.nx, (double)m.ny, m.dx, m.dy, m.orientation,
                              (double)m.aperture.kind, m.aperture.a, m.aperture.b, m.cx, m.cy, (double)nx, (double)ny,
                              p.g.dx, p.g.dy, (double)p.nxp, (double)p.nyp});
            need += a((nx + ny) * sizeof(MaskAxis));
            if (trans[k]++ == 0) need += a(nx * ny * esz);
        }
        if (p.check_only || !p.launch) continue;
        const size_t nxp = p.nxp, nyp = p.nyp;
        if (!fits_on_chip(p, dbl)) {  // the multi-pass path: two or three padded buffers, six plans
            need += a(nxp * nyp * esz) * (conv_engine(p.engine) ? 3 : 2) + 6 * a(std::max(nxp, nyp) * esz);
            need += a((nxp + nyp) * 8) * 2 + 256;
            continue;
        }
        need += a(ny * nxp * esz);                           // intermediate
        if (conv_engine(p.engine)) need += a(nyp * nxp * esz);  // kernel spectra
        need += a(nxp * esz) + a(nyp * esz);                 // twiddles
        need += a((nxp + nyp) * 8) + a((nxp + nyp) * 8);     // axis tables, Fresnel factors
        need += a(sizeof(ProbDev)) + 256;                    // descriptor
    }
    return need;
}

// Host-memory wrapper: H2D, run, D2H.
int run_host(oc_context* ctx, Prob p, const double* in, double* out, const oc_phase_mask* host_mask, Prob* done) {
    const size_t n = p.g.nx * p.g.ny;
    void *din = nullptr, *dout = nullptr, *dmask = nullptr;
    int st;
    if (n) {
        st = ensure_buf(ctx, "host_in", n * 16, &din);
        if (st) return st;
        st = ensure_buf(ctx, "host_out", n * 16, &dout);
        if (st) return st;
        HIP_TRY(ctx, hipMemcpyAsync(din, in, n * 16, hipMemcpyHostToDevice, ctx->stream));
    }
    oc_phase_mask dm;
    if (host_mask) {
        dm = *host_mask;
        const size_t mn = host_mask->nx * host_mask->ny;
        st = ensure_buf(ctx, "host_mask", std::max<size_t>(mn, 1) * 8, &dmask);
        if (st) return st;
        HIP_TRY(ctx, hipMemcpyAsync(dmask, host_mask->values, mn * 8, hipMemcpyHostToDevice, ctx->stream));
        dm.values = static_cast<const double*>(dmask);
        p.mask = &dm;
    }
    p.in = din;
    p.out = dout;
    std::vector<Prob> v{p};
    st = run_batch(ctx, true, v);
    *done = v[0];
    if (st == OC_OK && n) {
        HIP_TRY(ctx, hipMemcpyAsync(out, dout, n * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    return st;
}

}  // namespace ocr
