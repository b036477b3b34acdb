// C ABI of the propagation library: context, plans, guards and the
// entry points declared in include/optics_gpu.h.
//
// Every public propagation call funnels into run_batch(): host-side validation
// in the reference's order, O(1) sampling guards, plan/table lookup (cached
// per context), then the persistent pass launches of prop_kernels.hip over all
// problems of the call at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/optics_gpu.h"
#include "host_guards.hpp"
#include "prop_kernels.hpp"

namespace ocl {
hipError_t launch_pass_c64(int pass, const ocd::ProbDev* d, int np, int total, int block, size_t lds, int grid,
                           hipStream_t st);
hipError_t launch_pass_c128(int pass, const ocd::ProbDev* d, int np, int total, int block, size_t lds, int grid,
                            hipStream_t st);
hipError_t set_lds_limits();
hipError_t launch_mask_apply_c64(const void* in, void* out, const ocd::MaskDev& m, int nx, int ny, int* flag,
                                 hipStream_t st);
hipError_t launch_mask_trans(bool dbl, void* out, const ocd::MaskDev& m, int nx, int ny, hipStream_t st);
hipError_t launch_mask_apply_c128(const void* in, void* out, const ocd::MaskDev& m, int nx, int ny, int* flag,
                                  hipStream_t st);
hipError_t launch_check_finite(bool dbl, const void* in, size_t n, int* flag, hipStream_t st);
hipError_t launch_copy(bool dbl, const void* in, void* out, size_t n, hipStream_t st);
hipError_t launch_fill_int(int* p, int v, int n, hipStream_t st);
hipError_t fast_launch_c64(const ocd::FastSel& s, const ocd::ProbDev* d, int np, int total, int grid, hipStream_t st);
hipError_t fast_launch_c128(const ocd::FastSel& s, const ocd::ProbDev* d, int np, int total, int grid, hipStream_t st);
void fast_layout_c64(int n, int v, int* G, int* B, int* CB);
void fast_layout_c128(int n, int v, int* G, int* B, int* CB);
hipError_t fast_set_lds_limits_c64();
hipError_t fast_set_lds_limits_c128();

// Number of layout variants of the power-of-two kernels at length n (0: none).
int fast_variants(int n) {
    switch (n) {
        case 64: case 128: case 256: case 512: return 1;
        case 1024: case 2048: case 4096: case 8192: return 4;
        default: return 0;
    }
}

// true when n has a specialised kernel; fills its workgroup layout.
bool fast_supported(bool dbl, int n, int v, int* G, int* B, int* CB) {
    if (fast_variants(n) == 0) return false;
    if (dbl) fast_layout_c128(n, v, G, B, CB);
    else fast_layout_c64(n, v, G, B, CB);
    return *G > 0;
}
hipError_t launch_seed(bool dbl, const void* u, const double* dmdi, void* out, size_t n, hipStream_t st);
hipError_t launch_adj_elem(bool dbl, void* co, const void* ckpt, double* grad, const ocd::MaskDev& m, int nx, int ny,
                           hipStream_t st);
hipError_t launch_block_reorder(bool dbl, const void* src, void* dst, int nx, int ny, int B, hipStream_t st);
hipError_t launch_seed_mse(bool dbl, const void* u, const double* target, void* out, size_t n, double* sums,
                           hipStream_t st);
hipError_t launch_render_obs(bool dbl, const void* u, double* obs, size_t n, int frames, int nr, int cx, int cy,
                             int cl, double hs, double hl, int* flag, hipStream_t st);
hipError_t launch_mse(bool dbl, const void* u, const double* target, size_t n, int nr, double* part, double* merit,
                      hipStream_t st);
hipError_t launch_adam(double* x, const double* g, double* m, double* v, const double* lo, const double* hi, size_t n,
                       double* part, double* scale, int has_clip, double cn, double lr, double b1, double b2,
                       double eps, double bc1, double bc2, hipStream_t st);
}  // namespace ocl

using namespace ocd;

namespace {

constexpr size_t kLdsBudget = 140 * 1024;  // per workgroup (gfx950: 160 KB per CU)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct TableEntry {
    void* dev = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct oc_context {
    int device = 0;
    int num_cu = 256;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, DevBuf> bufs;
    std::map<std::pair<int, int>, void*> twiddles;  // (n, double?) -> table
    std::map<std::string, TableEntry> tables;       // per-parameter axis tables
    size_t tables_bytes = 0;
    void* pinned = nullptr;  // host staging for descriptors / flags
    size_t pinned_cap = 0;
    hipEvent_t pinned_done = nullptr;  // recorded after the last copy out of `pinned`
    // per-pass timing
    bool timing = false;
    struct Pending {
        int pass;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double pass_ms[OC_NUM_PASS_KINDS] = {0};
    uint64_t pass_n[OC_NUM_PASS_KINDS] = {0};
    bool force_generic = false;  // route every batch through the runtime-plan kernels
    void* queued = nullptr;      // batches submitted without waiting (QueuedSet, created on first use)
    std::map<std::pair<int, int>, int> fft_variant;  // (n, double?) -> fast-kernel layout variant
    // The autotuned layout if there is one, else the layout that timed
    // fastest for a masked-ASM + Fresnel round on MI355X (oc_autotune_fft
    // over 16 problems, both precisions; 1024/2048 complex128 are within 1%
    // across layouts).
    int variant_for(int n, bool dbl) const {
        auto it = fft_variant.find(std::make_pair(n, dbl ? 1 : 0));
        if (it != fft_variant.end()) return it->second;
        switch (n) {
            case 1024: case 2048: return 1;
            case 4096: case 8192: return dbl ? 3 : 2;
            default: return 0;
        }
    }
};

namespace {

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

// Records a timing bracket around `launch` when timing is enabled.
template <typename F>
hipError_t timed(oc_context* ctx, int pass, F&& launch) {
    if (!ctx->timing) return launch();
    hipEvent_t a = pool_event(ctx), b = pool_event(ctx);
    if (!a || !b) return launch();
    (void)hipEventRecord(a, ctx->stream);
    hipError_t e = launch();
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({pass, a, b});
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

#define HIP_TRY(ctx, expr)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail(ctx, e_ == hipErrorOutOfMemory ? OC_ERR_OUT_OF_MEMORY : OC_ERR_DEVICE,       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                         \
    } while (0)

int ensure_buf(oc_context* ctx, const std::string& name, size_t bytes, void** out) {
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

int ensure_pinned(oc_context* ctx, size_t bytes) {
    if (ctx->pinned_cap >= bytes) return OC_OK;
    if (ctx->pinned) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        HIP_TRY(ctx, hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_cap = 0;
    }
    size_t cap = std::max<size_t>(bytes, 1 << 16);
    HIP_TRY(ctx, hipHostMalloc(&ctx->pinned, cap, hipHostMallocDefault));
    ctx->pinned_cap = cap;
    return OC_OK;
}

// exp(-2 pi i k/n), k < n, evaluated in extended precision then rounded.
int get_twiddles(oc_context* ctx, int n, bool dbl, const void** out) {
    auto key = std::make_pair(n, dbl ? 1 : 0);
    auto it = ctx->twiddles.find(key);
    if (it != ctx->twiddles.end()) {
        *out = it->second;
        return OC_OK;
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

int make_plan(oc_context* ctx, int n, bool dbl, FftPlanDev* pl) {
    std::vector<int> rad;
    bool generic;
    if (!make_radices(n, rad, generic)) return fail(ctx, OC_ERR_UNSUPPORTED, "too many FFT stages");
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

// Internal description of one propagation problem.
struct Prob {
    oc_band_geom g;
    double z;
    int engine;
    int band_limited, strict;
    double padding, max_padding, medium;
    const oc_phase_mask* mask;
    const void* in;
    void* out;
    int status = OC_OK;
    bool launch = false;     // goes through the FFT passes
    bool check_only = false; // finiteness check first (z == 0 copy / guard failure)
    int pending = OC_OK;     // status to report if the input is finite
    size_t nxp = 0, nyp = 0;
    oc_asm_report asm_rep{};
    oc_fresnel_report fr_rep{};
    oc_rs_report rs_rep{};
    oc_fraunhofer_config fh_cfg{};
    oc_fraunhofer_report fh_rep{};
    OpsDev ops{};
};

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
    auto it = ctx->tables.find(key);
    if (it == ctx->tables.end()) {
        std::vector<float> h(2 * (nxp + nyp));
        for (size_t k = 0; k < nxp + nyp; ++k) {
            const bool isx = k < nxp;
            const double f = isx ? och::fftfreq_at(k, nxp, p.g.dx) : och::fftfreq_at(k - nxp, nyp, p.g.dy);
            const double ph = isx ? e->kz - e->c * (f * f) : -(e->c * (f * f));
            h[2 * k] = (float)std::cos((long double)ph);
            h[2 * k + 1] = (float)std::sin((long double)ph);
        }
        TableEntry te;
        const size_t bytes = h.size() * sizeof(float);
        HIP_TRY(ctx, hipMalloc(&te.dev, bytes));
        HIP_TRY(ctx, hipMemcpy(te.dev, h.data(), bytes, hipMemcpyHostToDevice));
        te.bytes = bytes;
        it = ctx->tables.emplace(key, te).first;
    }
    e->hx = it->second.dev;
    e->hy = static_cast<const float*>(it->second.dev) + 2 * nxp;
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
    auto it = ctx->tables.find(key);
    const size_t bytes = (nxp + nyp) * 8;
    if (it == ctx->tables.end()) {
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
        TableEntry te;
        HIP_TRY(ctx, hipMalloc(&te.dev, bytes));
        HIP_TRY(ctx, hipMemcpy(te.dev, h.data(), bytes, hipMemcpyHostToDevice));
        te.bytes = bytes;
        it = ctx->tables.emplace(key, te).first;
    }
    e->ax = static_cast<const double*>(it->second.dev);
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
    auto it = ctx->tables.find(key);
    if (it == ctx->tables.end()) {
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
        TableEntry te;
        const size_t bytes = h.size() * sizeof(MaskAxis);
        HIP_TRY(ctx, hipMalloc(&te.dev, bytes));
        HIP_TRY(ctx, hipMemcpy(te.dev, h.data(), bytes, hipMemcpyHostToDevice));
        te.bytes = bytes;
        it = ctx->tables.emplace(key, te).first;
    }
    md->sx = static_cast<const MaskAxis*>(it->second.dev);
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
    p.launch = true;
}

struct PassCfg {
    int block = 64;
    size_t lds = 0;
    int total = 0;
};

// Workgroup shape of one problem: transforms per row-pass workgroup (cnt) and
// the column block width B of the blocked intermediate.
struct Layout {
    int cnt = 1, B = 1;
    int ccnt = 1;  // columns per column-pass workgroup (divides B)
    size_t row_bytes = 0, col_bytes = 0;
    bool ok = false;
};

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

struct Batch {
    bool dbl = false;
    bool fast = false;  // power-of-two kernels (all problems share P_x = P_y = fast_n)
    int fast_n = 0;
    int fast_v = 0;
    int post_kind = -2;      // post-op shared by every problem (-1: mixed)
    int col_kind = -2;       // engine shared by every problem (-1: mixed)
    int pre_kind = -2;       // pre-op shared by every problem (-1: mixed)
    bool sep = false;        // separable complex64 Fresnel: two passes
    bool rs = false;
    std::vector<ProbDev> pd;
    std::vector<int> idx;  // index into the problem vector of each descriptor
    void* dprobs = nullptr;
    PassCfg pc[4];
    std::vector<void*> owned;  // allocations of a persistent batch
};

void free_batch(oc_context* ctx, Batch& b) {
    if (!b.owned.empty()) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : b.owned) (void)hipFree(p);
    b.owned.clear();
    b.pd.clear();
    b.idx.clear();
    b.dprobs = nullptr;
}

// Batches of oc_propagate_batch_submit_dev_* not yet waited for.
struct Queued {
    std::vector<Prob> probs;
    int* dflags = nullptr;
    oc_prop_desc* descs = nullptr;
};
struct QueuedSet {
    std::vector<Queued> batches;
};
QueuedSet& queued_set(oc_context* ctx) {
    if (!ctx->queued) ctx->queued = new QueuedSet();
    return *static_cast<QueuedSet*>(ctx->queued);
}
void release_queued(oc_context* ctx) {
    delete static_cast<QueuedSet*>(ctx->queued);
    ctx->queued = nullptr;
}

int batch_alloc(oc_context* ctx, Batch& b, bool persistent, const std::string& name, size_t bytes, void** out) {
    if (!persistent) return ensure_buf(ctx, name, bytes, out);
    HIP_TRY(ctx, hipMalloc(out, std::max<size_t>(bytes, 256)));
    b.owned.push_back(*out);
    return OC_OK;
}

// Descriptors for every prepared problem with launch == true.  A persistent
// batch owns its intermediates and descriptor array (reusable across calls,
// capturable in a graph); a transient one uses the context's shared buffers.
// `subset` (optional) restricts the batch to those problem indices; `tag`
// names the transient buffers so that several batches can be in flight.
int build_batch(oc_context* ctx, bool dbl, std::vector<Prob>& probs, int* dflags, Batch& b, bool persistent,
                const std::vector<int>* subset = nullptr, const std::string& tag = "") {
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
            ok = ok && probs[i].nxp == n0 && probs[i].nyp == n0 && conv_engine(probs[i].engine) == rs0;
        const int v = ctx->variant_for((int)n0, dbl);
        if (ok && ocl::fast_supported(dbl, (int)n0, v, &fG, &fB, &fC)) {
            b.fast = true;
            b.fast_n = (int)n0;
            b.fast_v = v;
            b.rs = rs0;
        }
    }
    // complex64 Fresnel batches without post-ops factor per axis: two passes
    // (the x embedding offset must fall on a block boundary of the layout)
    if (b.fast && !dbl && !b.rs) {
        bool sep = true;
        for (int i : cand) {
            const Prob& p = probs[i];
            const size_t ox = (p.nxp - p.g.nx) / 2;
            sep = sep && p.engine == ENG_FRESNEL && (p.ops.pre == PRE_NONE || p.ops.pre == PRE_MASK) &&
                  p.ops.post == POST_NONE && ox % (size_t)fB == 0;
        }
        b.sep = sep;
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
        int st = make_plan(ctx, d.px, dbl, &d.prow);
        if (st) return st;
        st = make_plan(ctx, d.py, dbl, &d.pcol);
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
        d.tiles_kgen = rs ? (d.py + L.cnt - 1) / L.cnt : 0;
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
    int st = batch_alloc(ctx, b, persistent, "probs" + tag, bytes, &b.dprobs);
    if (st) return st;
    if (persistent) {
        HIP_TRY(ctx, hipMemcpy(b.dprobs, b.pd.data(), bytes, hipMemcpyHostToDevice));
    } else {
        st = ensure_pinned(ctx, bytes);
        if (st) return st;
        // the staging buffer is free once the previous copy out of it ran
        // (not the kernels queued after that copy)
        if (ctx->pinned_done) HIP_TRY(ctx, hipEventSynchronize(ctx->pinned_done));
        else HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->pinned_done, hipEventDisableTiming));
        std::memcpy(ctx->pinned, b.pd.data(), bytes);
        HIP_TRY(ctx, hipMemcpyAsync(b.dprobs, ctx->pinned, bytes, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->pinned_done, ctx->stream));
    }
    return OC_OK;
}

constexpr unsigned kAllPasses = 0xF;
// pass 4 (power-of-two batches with a GS phase post-op): one iteration's
// inverse row pass fused with the next one's forward row pass
constexpr int PASS_GS_ROW = 4;
// pass 5: the same with beta = 1 and 2 pi wide bounds (amplitude replacement;
// the phases are left to the last iteration's inverse pass)
constexpr int PASS_GS_AMP = 5;
// timing slots of the separable Fresnel passes (see OC_NUM_PASS_KINDS)
constexpr int PASS_T_SEP_COL = 6, PASS_T_SEP_ROW = 7;

// Enqueues the passes of `b` selected by `mask` (bit i = pass i), in the
// order kernel generation, row forward, column, row inverse.
int launch_batch(oc_context* ctx, Batch& b, unsigned mask = kAllPasses) {
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
                const ocd::FastSel sel{pass,       b.fast_n,   b.fast_v,     b.rs ? 1 : 0, b.post_kind,
                                       b.col_kind, b.pre_kind, b.sep ? 1 : 0};
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
    if (ctx->tables.size() > 1024) {  // between calls only: tables in flight stay valid
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& kv : ctx->tables) (void)hipFree(kv.second.dev);
        ctx->tables.clear();
    }
}

// Queues all problems on the context stream (host-side validation, the
// standalone finiteness checks, then one batch per padded-size group) with
// their finiteness flags in dflags[0..n).  No synchronisation.
int launch_problems(oc_context* ctx, bool dbl, std::vector<Prob>& probs, int* dflags) {
    evict_tables(ctx);
    for (auto& p : probs) {
        prepare(p);
        if (p.status == OC_OK && p.mask && !p.check_only) {
            int st = mask_footprint(p.g, *p.mask);
            if (st) p.status = st;
        }
    }
    const int np = (int)probs.size();
    HIP_TRY(ctx, ocl::launch_fill_int(dflags, 0, np, ctx->stream));
    for (int i = 0; i < np; ++i) {
        Prob& p = probs[i];
        if (p.status != OC_OK || !p.check_only) continue;
        const size_t n = p.g.nx * p.g.ny;
        HIP_TRY(ctx, ocl::launch_check_finite(dbl, p.in, n, dflags + i, ctx->stream));
        if (p.pending == OC_OK && p.out != p.in) HIP_TRY(ctx, ocl::launch_copy(dbl, p.in, p.out, n, ctx->stream));
    }
    // One batch per padded size that has specialised kernels (uneven grids
    // then each get the power-of-two kernels), one for all other sizes.
    std::map<std::pair<size_t, size_t>, std::vector<int>> groups;
    for (int i = 0; i < np; ++i) {
        const Prob& p = probs[i];
        if (p.status != OC_OK || !p.launch) continue;
        const bool fast = !ctx->force_generic && p.nxp == p.nyp && ocl::fast_variants((int)p.nxp) > 0;
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
    return OC_OK;
}

// Waits for the queued work and turns the flags of launch_problems into
// per-problem statuses; returns the first non-OK one.
int finish_problems(oc_context* ctx, std::vector<Prob>& probs, const int* dflags) {
    const int np = (int)probs.size();
    std::vector<int> hflags(std::max(np, 1), 0);
    if (np) HIP_TRY(ctx, hipMemcpyAsync(hflags.data(), dflags, sizeof(int) * np, hipMemcpyDeviceToHost, ctx->stream));
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

}  // namespace

/* ====================================================================== */
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
    if (ocl::set_lds_limits() != hipSuccess) return OC_ERR_DEVICE;
    if (ocl::fast_set_lds_limits_c64() != hipSuccess || ocl::fast_set_lds_limits_c128() != hipSuccess)
        return OC_ERR_DEVICE;
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
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->pinned_done) (void)hipEventDestroy(ctx->pinned_done);
    release_queued(ctx);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
}

int oc_context_set_stream(oc_context* ctx, void* stream) {
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own;
    return OC_OK;
}

void* oc_context_stream(oc_context* ctx) { return ctx ? ctx->stream : nullptr; }

int oc_context_synchronize(oc_context* ctx) {
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OC_OK;
}

int oc_context_enable_timing(oc_context* ctx, int enable) {
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    ctx->timing = enable != 0;
    return OC_OK;
}

int oc_context_pass_timing(oc_context* ctx, double* ms, uint64_t* launches, int reset) {
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
    return dev_asm(ctx, true, g, in, z, cfg, mask, out, rep);
}
int oc_asm_propagate_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                             const oc_asm_config* cfg, const oc_phase_mask* mask, float* out, oc_asm_report* rep) {
    return dev_asm(ctx, false, g, in, z, cfg, mask, out, rep);
}

int oc_fresnel_propagate(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                         const oc_fresnel_config* cfg, double* out, oc_fresnel_report* rep) {
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
    return dev_fresnel(ctx, true, g, in, z, cfg, mask, out, rep);
}
int oc_fresnel_propagate_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                                 const oc_fresnel_config* cfg, const oc_phase_mask* mask, float* out,
                                 oc_fresnel_report* rep) {
    return dev_fresnel(ctx, false, g, in, z, cfg, mask, out, rep);
}

int oc_fresnel_propagate_convolution(oc_context* ctx, const oc_band_geom* g, const double* in, double z,
                                     const oc_fresnel_config* cfg, double* out, oc_fresnel_report* rep) {
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
    return dev_fresnel_conv(ctx, true, g, in, z, cfg, out, rep);
}
int oc_fresnel_propagate_convolution_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                                             const oc_fresnel_config* cfg, float* out, oc_fresnel_report* rep) {
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
    return dev_fraunhofer(ctx, true, g, in, z, cfg, out, rep);
}
int oc_fraunhofer_propagate_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                                    const oc_fraunhofer_config* cfg, float* out, oc_fraunhofer_report* rep) {
    return dev_fraunhofer(ctx, false, g, in, z, cfg, out, rep);
}

int oc_rs_propagate(oc_context* ctx, const oc_band_geom* g, const double* in, double z, const oc_rs_config* cfg,
                    double* out, oc_rs_report* rep) {
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
    return dev_rs(ctx, true, g, in, z, cfg, out, rep);
}
int oc_rs_propagate_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, double z,
                            const oc_rs_config* cfg, float* out, oc_rs_report* rep) {
    return dev_rs(ctx, false, g, in, z, cfg, out, rep);
}

static int batch_probs(const oc_prop_desc* d, size_t count, std::vector<Prob>& v) {
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

int oc_propagate_batch_dev_c64(oc_context* ctx, oc_prop_desc* d, size_t count) { return batch(ctx, false, d, count); }
int oc_propagate_batch_dev_c128(oc_context* ctx, oc_prop_desc* d, size_t count) { return batch(ctx, true, d, count); }
int oc_propagate_batch_submit_dev_c64(oc_context* ctx, oc_prop_desc* d, size_t count) {
    return submit(ctx, false, d, count);
}
int oc_propagate_batch_submit_dev_c128(oc_context* ctx, oc_prop_desc* d, size_t count) {
    return submit(ctx, true, d, count);
}

int oc_context_wait(oc_context* ctx) {
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (!ctx->queued) return OC_OK;
    QueuedSet& qs = queued_set(ctx);
    int first = OC_OK;
    for (Queued& q : qs.batches) {
        const int st = finish_problems(ctx, q.probs, q.dflags);
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
    return mask_dev(ctx, true, g, in, m, out);
}
int oc_phasemask_apply_dev_c64(oc_context* ctx, const oc_band_geom* g, const float* in, const oc_phase_mask* m,
                               float* out) {
    return mask_dev(ctx, false, g, in, m, out);
}

int oc_phasemask_apply(oc_context* ctx, const oc_band_geom* g, const double* in, const oc_phase_mask* m, double* out) {
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

/* ====================================================================== */
/* Gerchberg-Saxton and adjoint                                            */
/* ====================================================================== */
namespace {

// Device allocation released at scope exit (after a stream sync).
struct Scoped {
    oc_context* ctx;
    std::vector<void*> ptrs;
    explicit Scoped(oc_context* c) : ctx(c) {}
    ~Scoped() {
        if (!ptrs.empty()) (void)hipStreamSynchronize(ctx->stream);
        for (void* p : ptrs) (void)hipFree(p);
    }
    int alloc(size_t bytes, void** out) {
        hipError_t e = hipMalloc(out, std::max<size_t>(bytes, 256));
        if (e != hipSuccess) return fail(ctx, e == hipErrorOutOfMemory ? OC_ERR_OUT_OF_MEMORY : OC_ERR_DEVICE,
                                         std::string("hipMalloc: ") + hipGetErrorString(e));
        ptrs.push_back(*out);
        return OC_OK;
    }
};

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
    return gs_fft_impl(ctx, false, n, restarts, source_amp, target_amp, phases, iterations, graph_chunk, err_out);
}
int oc_gs_fft_run_dev_c128(oc_context* ctx, uint64_t n, uint64_t restarts, const double* source_amp,
                           const double* target_amp, double* phases, uint64_t iterations, uint64_t graph_chunk,
                           double* err_out) {
    return gs_fft_impl(ctx, true, n, restarts, source_amp, target_amp, phases, iterations, graph_chunk, err_out);
}

// gs.rs:347-597
int oc_gs_run(oc_context* ctx, const oc_gs_problem* pb, oc_gs_solution* sol) {
    if (!ctx || !pb || !sol) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    const oc_band_geom g = pb->geom;
    const size_t nx = g.nx, ny = g.ny, n = nx * ny;
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
    if (pb->init_kind != 0 && pb->init_kind != 2) return OC_ERR_UNSUPPORTED;
    if (pb->init_kind == 0 && !pb->carrier_values) return OC_ERR_INVALID_ARGUMENT;
    if (pb->init_kind == 2 && !pb->init_phase) return OC_ERR_INVALID_ARGUMENT;
    const double half_z = pb->z_total / 2.0;
    const double lo = pb->lo, hi = pb->hi;
    const double beta_f = pb->variant == 2 ? pb->beta : 1.0;
    const double beta_p = pb->variant == 1 ? pb->beta : 1.0;
    Scoped mem(ctx);
    void *d_in, *d_arr, *d_illum, *d_phase, *d_best, *d_open, *d_tamp, *d_mt, *d_proj, *d_err, *d_dphi, *d_sums,
        *d_flags;
    int st;
    if ((st = mem.alloc(n * 16, &d_in)) || (st = mem.alloc(n * 16, &d_arr)) || (st = mem.alloc(n * 8, &d_illum)) ||
        (st = mem.alloc(n * 8, &d_phase)) || (st = mem.alloc(n * 8, &d_best)) || (st = mem.alloc(n, &d_open)) ||
        (st = mem.alloc(n * 8, &d_tamp)) || (st = mem.alloc(n * 8, &d_mt)) || (st = mem.alloc(n * 16, &d_proj)) ||
        (st = mem.alloc(n * 8, &d_err)) || (st = mem.alloc(n * 8, &d_dphi)) || (st = mem.alloc(64, &d_sums)) ||
        (st = mem.alloc(64, &d_flags)))
        return st;
    HIP_TRY(ctx, hipMemcpy(d_in, pb->input, n * 16, hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(d_tamp, pb->target_amplitude, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(d_mt, pb->merit_target, n * 8, hipMemcpyHostToDevice));
    // aperture indicator (gs.rs:427-433)
    std::vector<unsigned char> open(n);
    for (size_t iy = 0; iy < ny; ++iy)
        for (size_t ix = 0; ix < nx; ++ix) {
            const double x = ((double)ix - ((double)nx - 1.0) / 2.0) * g.dx;
            const double y = ((double)iy - ((double)ny - 1.0) / 2.0) * g.dy;
            open[iy * nx + ix] = contains_h(pb->aperture, x - pb->cx, y - pb->cy) ? 1 : 0;
        }
    HIP_TRY(ctx, hipMemcpy(d_open, open.data(), n, hipMemcpyHostToDevice));
    // illumination arriving at the DOE plane (gs.rs:436-444)
    const oc_asm_config& c = pb->asm_cfg;
    {
        std::vector<Prob> v{make_prob(&g, half_z, ENG_ASM, c.band_limited, c.strict, c.padding, c.max_padding,
                                      c.medium, nullptr, d_in, d_arr)};
        if (run_batch(ctx, true, v) != OC_OK) return OC_INV_RENDER;
    }
    std::vector<double> arr(2 * n), illum(n);
    HIP_TRY(ctx, hipMemcpy(arr.data(), d_arr, n * 16, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; ++k) illum[k] = std::hypot(arr[2 * k], arr[2 * k + 1]);
    HIP_TRY(ctx, hipMemcpy(d_illum, illum.data(), n * 8, hipMemcpyHostToDevice));
    // initial phase (gs.rs:447-475)
    std::vector<double> phase(n);
    const double* src = pb->init_kind == 0 ? pb->carrier_values : pb->init_phase;
    for (size_t k = 0; k < n; ++k) phase[k] = encode_bounds_h(src[k], lo, hi);
    HIP_TRY(ctx, hipMemcpy(d_phase, phase.data(), n * 8, hipMemcpyHostToDevice));
    // fused forward / backward half-transforms
    std::vector<Prob> fw{make_prob(&g, half_z, ENG_ASM, c.band_limited, c.strict, c.padding, c.max_padding, c.medium,
                                   nullptr, nullptr, d_proj)};
    fw[0].ops.pre = PRE_POLAR;
    fw[0].ops.amp = d_illum;
    fw[0].ops.phase = d_phase;
    fw[0].ops.open = static_cast<const unsigned char*>(d_open);
    fw[0].ops.post = POST_GS_TARGET;
    fw[0].ops.tamp = static_cast<const double*>(d_tamp);
    fw[0].ops.merit_target = static_cast<const double*>(d_mt);
    fw[0].ops.beta_f = beta_f;
    fw[0].ops.err_out = static_cast<double*>(d_err);
    fw[0].ops.sums = static_cast<double*>(d_sums);
    std::vector<Prob> bw{make_prob(&g, -half_z, ENG_ASM, c.band_limited, c.strict, c.padding, c.max_padding, c.medium,
                                   nullptr, d_proj, nullptr)};
    bw[0].ops.post = POST_GS_PHASE;
    bw[0].ops.phase_io = d_phase;
    bw[0].ops.beta_p = beta_p;
    bw[0].ops.lo = lo;
    bw[0].ops.hi = hi;
    bw[0].ops.dphi_out = static_cast<double*>(d_dphi);
    bw[0].ops.sums = static_cast<double*>(d_sums);
    prepare(fw[0]);
    prepare(bw[0]);
    if (fw[0].status || bw[0].status || fw[0].check_only || bw[0].check_only) return OC_INV_RENDER;
    Batch bf, bb;
    st = build_batch(ctx, true, fw, static_cast<int*>(d_flags), bf, true);
    if (!st) st = build_batch(ctx, true, bw, static_cast<int*>(d_flags) + 1, bb, true);
    if (st || bf.pd.empty() || bb.pd.empty()) {
        free_batch(ctx, bf);
        free_batch(ctx, bb);
        return st ? st : OC_INV_RENDER;
    }
    bool have_best = false;
    double best_merit = 0.0;
    uint64_t best_it = 0;
    std::vector<oc_gs_record> hist;
    bool converged = false;
    double hsum[4];
    int hfl[2];
    for (uint64_t it = 0; it < pb->max_iterations; ++it) {
        HIP_TRY(ctx, hipMemsetAsync(d_sums, 0, 64, ctx->stream));
        HIP_TRY(ctx, ocl::launch_fill_int(static_cast<int*>(d_flags), 0, 2, ctx->stream));
        if ((st = launch_batch(ctx, bf))) break;
        HIP_TRY(ctx, hipMemcpyAsync(hsum, d_sums, 32, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(hfl, d_flags, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (hfl[0]) { st = OC_INV_RENDER; break; }
        const double amp_mse = hsum[0] / (double)n;
        const double merit = hsum[1] / (double)n;
        const bool improved = !have_best || merit < best_merit;
        if (improved) {
            have_best = true;
            best_merit = merit;
            best_it = it;
            HIP_TRY(ctx, hipMemcpyAsync(d_best, d_phase, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        }
        hist.push_back({it, merit, best_merit, amp_mse, std::sqrt(amp_mse), 0.0});
        if ((st = launch_batch(ctx, bb))) break;
        HIP_TRY(ctx, hipMemcpyAsync(hsum, d_sums, 32, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(hfl, d_flags, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (hfl[1]) { st = OC_INV_RENDER; break; }
        hist.back().phase_update_rms = std::sqrt(hsum[2]) / std::sqrt((double)n);
        if (it + 1 > pb->patience) {
            const double earlier = hist[hist.size() - 1 - pb->patience].best_merit;
            const double scale = std::fmax(std::fabs(earlier), 1e-30);
            if ((earlier - best_merit) / scale < pb->merit_tolerance) {
                converged = true;
                break;
            }
        }
    }
    collect_timing(ctx);
    free_batch(ctx, bf);
    free_batch(ctx, bb);
    if (st) return st;
    // outputs
    HIP_TRY(ctx, hipMemcpy(sol->phase, d_best, n * 8, hipMemcpyDeviceToHost));
    if (sol->target_amplitude_error)
        HIP_TRY(ctx, hipMemcpy(sol->target_amplitude_error, d_err, n * 8, hipMemcpyDeviceToHost));
    if (sol->phase_error) HIP_TRY(ctx, hipMemcpy(sol->phase_error, d_dphi, n * 8, hipMemcpyDeviceToHost));
    if (sol->illumination_amplitude) std::memcpy(sol->illumination_amplitude, illum.data(), n * 8);
    if (sol->history) std::memcpy(sol->history, hist.data(), hist.size() * sizeof(oc_gs_record));
    sol->merit = best_merit;
    sol->best_iteration = best_it;
    sol->amplitude_mse = hist[best_it].amplitude_mse;
    sol->converged = converged ? 1 : 0;
    sol->iterations = hist.size();
    if (sol->target_intensity) {  // |U_t|^2 of the best phase (same render as its iteration)
        std::vector<Prob> v{make_prob(&g, half_z, ENG_ASM, c.band_limited, c.strict, c.padding, c.max_padding,
                                      c.medium, nullptr, nullptr, d_proj)};
        v[0].ops.pre = PRE_POLAR;
        v[0].ops.amp = d_illum;
        v[0].ops.phase = d_best;
        v[0].ops.open = static_cast<const unsigned char*>(d_open);
        if (run_batch(ctx, true, v) != OC_OK) return OC_INV_RENDER;
        HIP_TRY(ctx, hipMemcpy(arr.data(), d_proj, n * 16, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) {
            const double a = std::hypot(arr[2 * k], arr[2 * k + 1]);
            sol->target_intensity[k] = a * a;
        }
    }
    return OC_OK;
}

int oc_adjoint_asm(oc_context* ctx, const oc_band_geom* g, const double* in, double z, const oc_asm_config* cfg,
                   double* out, oc_asm_report* rep) {
    int st = oc_asm_propagate(ctx, g, in, -z, cfg, out, rep);
    return st == OC_OK ? OC_OK : OC_INV_ADJOINT;
}

int oc_adjoint_transmission(oc_context* ctx, const oc_band_geom* g, const oc_phase_mask* m, double* t_out) {
    if (!ctx || !g || !m || !t_out) return OC_ERR_INVALID_ARGUMENT;
    const size_t n = g->nx * g->ny;
    std::vector<double> ones(2 * n, 0.0);
    for (size_t k = 0; k < n; ++k) ones[2 * k] = 1.0;
    int st = oc_phasemask_apply(ctx, g, ones.data(), m, t_out);
    return st == OC_OK ? OC_OK : OC_INV_ADJOINT;
}

}  // extern "C"

namespace {

// Forward tape + reverse sweep on device (adjoint.rs:140-246).  Mask values
// are device pointers.  `dmdi` (device, f64) seeds w = dmdi * U_det; when
// null, the MseToTarget seed 2(|U|^2 - target)/n is formed from `target` and
// the merit is returned in *merit.  grads: n_elements * n (device, zeroed
// here); may be null for a forward render + merit only.
int adjoint_impl(oc_context* ctx, const oc_design_step* steps, size_t nsteps, const oc_band_geom& g, const void* d_input,
                 const oc_asm_config& cfg, const double* dmdi, const double* target, double* grads, double* merit) {
    const size_t n = g.nx * g.ny;
    size_t ne = 0;
    for (size_t s = 0; s < nsteps; ++s) ne += steps[s].has_element ? 1 : 0;
    Scoped mem(ctx);
    void *cur, *nxt, *ckpt = nullptr, *sums, *flags;
    int st;
    if ((st = mem.alloc(n * 16, &cur)) || (st = mem.alloc(n * 16, &nxt)) || (st = mem.alloc(64, &sums)) ||
        (st = mem.alloc(sizeof(int) * (2 * nsteps + 1), &flags)))
        return st;
    if (ne && (st = mem.alloc(ne * n * 16, &ckpt))) return st;
    // Every step is queued without waiting; the finiteness flags of all
    // propagations and the merit are read back once at the end.
    std::vector<std::vector<Prob>> queued;
    queued.reserve(2 * nsteps);
    int* dflags = static_cast<int*>(flags);
    auto propagate = [&](Prob p) -> int {
        queued.push_back({p});
        int s2 = launch_problems(ctx, true, queued.back(), dflags + (queued.size() - 1));
        if (s2 == OC_OK) s2 = queued.back()[0].status;
        return s2;
    };
    HIP_TRY(ctx, hipMemcpyAsync(cur, d_input, n * 16, hipMemcpyDeviceToDevice, ctx->stream));
    // forward: store each element's incoming field, fuse the mask into the propagation
    size_t e = 0;
    for (size_t s = 0; s < nsteps; ++s) {
        const oc_design_step& stp = steps[s];
        if (stp.has_element) {
            HIP_TRY(ctx, hipMemcpyAsync(static_cast<unsigned char*>(ckpt) + e * n * 16, cur, n * 16,
                                        hipMemcpyDeviceToDevice, ctx->stream));
            ++e;
        }
        if (stp.z_after != 0.0) {
            if (propagate(make_prob(&g, stp.z_after, ENG_ASM, cfg.band_limited, cfg.strict, cfg.padding,
                                    cfg.max_padding, stp.medium, stp.has_element ? &stp.mask : nullptr, cur,
                                    nxt)) != OC_OK)
                return fail(ctx, OC_INV_ADJOINT, "forward propagation failed");
            std::swap(cur, nxt);
        } else if (stp.has_element) {
            const int fs = mask_footprint(g, stp.mask);
            if (fs) return OC_INV_ADJOINT;
            MaskDev md;
            if ((st = mask_tables(ctx, g, stp.mask, stp.mask.values, &md))) return st;
            HIP_TRY(ctx, ocl::launch_mask_apply_c128(cur, nxt, md, (int)g.nx, (int)g.ny, nullptr, ctx->stream));
            std::swap(cur, nxt);
        }
    }
    // seed at the detector
    HIP_TRY(ctx, hipMemsetAsync(sums, 0, 64, ctx->stream));
    if (dmdi) HIP_TRY(ctx, ocl::launch_seed(true, cur, dmdi, nxt, n, ctx->stream));
    else HIP_TRY(ctx, ocl::launch_seed_mse(true, cur, target, nxt, n, static_cast<double*>(sums), ctx->stream));
    std::swap(cur, nxt);
    auto finish = [&]() -> int {
        double hs[4];
        std::vector<int> hf(queued.size() + 1, 0);
        HIP_TRY(ctx, hipMemcpyAsync(hs, sums, 32, hipMemcpyDeviceToHost, ctx->stream));
        if (!queued.empty())
            HIP_TRY(ctx, hipMemcpyAsync(hf.data(), dflags, sizeof(int) * queued.size(), hipMemcpyDeviceToHost,
                                        ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        collect_timing(ctx);
        if (merit) *merit = hs[0] / (double)n;
        for (size_t i = 0; i < queued.size(); ++i)
            if (hf[i]) return fail(ctx, OC_INV_ADJOINT, "non-finite samples in a propagation");
        return OC_OK;
    };
    if (!grads) return finish();
    HIP_TRY(ctx, hipMemsetAsync(grads, 0, ne * n * 8, ctx->stream));
    // reverse sweep
    e = ne;
    for (size_t s = nsteps; s-- > 0;) {
        const oc_design_step& stp = steps[s];
        if (stp.has_element) --e;
        void* ck = stp.has_element ? static_cast<unsigned char*>(ckpt) + e * n * 16 : nullptr;
        double* gr = stp.has_element ? grads + e * n : nullptr;
        if (stp.z_after != 0.0) {
            Prob pb = make_prob(&g, -stp.z_after, ENG_ASM, cfg.band_limited, cfg.strict, cfg.padding, cfg.max_padding,
                                stp.medium, nullptr, cur, nxt);
            if (stp.has_element) {
                pb.mask = &stp.mask;  // mask tables for the epilogue (no pre-op)
                pb.ops.pre = PRE_NONE;
                pb.ops.post = POST_ADJ_GRAD;
                pb.ops.ckpt = ck;
                pb.ops.grad = gr;
            }
            if (propagate(pb) != OC_OK) return fail(ctx, OC_INV_ADJOINT, "adjoint propagation failed");
            std::swap(cur, nxt);
        } else if (stp.has_element) {
            MaskDev md;
            if ((st = mask_tables(ctx, g, stp.mask, stp.mask.values, &md))) return st;
            HIP_TRY(ctx, ocl::launch_adj_elem(true, cur, ck, gr, md, (int)g.nx, (int)g.ny, ctx->stream));
        }
    }
    return finish();
}

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
    auto flush = [&]() -> int {
        if (pend < 0) return OC_OK;
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
        queued.push_back(std::move(probs));
        st = launch_problems(ctx, dbl, queued.back(), dflags + nflag);
        if (st) return st;
        for (const Prob& p : queued.back())
            if (p.status != OC_OK)
                return fail(ctx, OC_INV_RENDER, std::string("propagation: ") + oc_status_name(p.status));
        nflag += (int)nrend;
        pend = -1;
        advance();
    }
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

extern "C" {

int oc_fd_gradient_dev_c128(oc_context* ctx, const oc_band_geom* g, const oc_design_step* designs, size_t nsteps,
                            size_t dim, const double* up, const double* dn, const double* lo, const double* hi,
                            const oc_asm_config* cfg, const void* input, const double* target, double* grad_out,
                            double* merits_out, size_t* evaluations_out) {
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

int oc_adjoint_push_back(oc_context* ctx, const oc_design_step* steps, size_t nsteps, const oc_band_geom* g,
                         const double* input, const oc_asm_config* cfg, const double* dmdi, double* dphi_out) {
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
    st = adjoint_impl(ctx, ds.data(), nsteps, *g, din, *cfg, static_cast<double*>(ddm), nullptr,
                      static_cast<double*>(dgr), nullptr);
    if (st) return st;
    if (ne) HIP_TRY(ctx, hipMemcpy(dphi_out, dgr, ne * n * 8, hipMemcpyDeviceToHost));
    return OC_OK;
}

int oc_adjoint_grid_gradient_dev_c128(oc_context* ctx, const oc_design_step* steps, size_t nsteps,
                                      const oc_band_geom* g, const double* input, const oc_asm_config* cfg,
                                      const double* target, double* dphi_out, double* merit_out) {
    if (!ctx || !g || !cfg || !input || !target || (!steps && nsteps)) return OC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    evict_tables(ctx);
    if (g->nx * g->ny == 0) return OC_INV_ADJOINT;
    double m = 0.0;
    int st = adjoint_impl(ctx, steps, nsteps, *g, input, *cfg, nullptr, target, dphi_out, &m);
    if (merit_out) *merit_out = m;
    return st;
}

}  // extern "C"

extern "C" {

int oc_context_set_fft_variant(oc_context* ctx, uint64_t n, int dbl, int variant) {
    if (!ctx || variant < 0 || variant >= std::max(1, ocl::fast_variants((int)n))) return OC_ERR_INVALID_ARGUMENT;
    ctx->fft_variant[std::make_pair((int)n, dbl ? 1 : 0)] = variant;
    return OC_OK;
}

int oc_context_fft_variant(oc_context* ctx, uint64_t n, int dbl) {
    if (!ctx || ocl::fast_variants((int)n) == 0) return -1;
    return ctx->variant_for((int)n, dbl != 0);
}

int oc_context_force_generic(oc_context* ctx, int enable) {
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    ctx->force_generic = enable != 0;
    return OC_OK;
}

// Times every layout variant of the specialised kernels at padded length n on
// a synthetic round of masked ASM then Fresnel propagations (up to 16 fields
// of n/2 x n/2, padding 2) and keeps the fastest.
int oc_autotune_fft(oc_context* ctx, uint64_t n, int dbl, int* variant_out, double* ms_out) {
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
    const bool timing = ctx->timing;
    ctx->timing = false;
    int best = 0;
    double best_ms = 1e300;
    for (int v = 0; v < nv; ++v) {
        ctx->fft_variant[key] = v;
        double ms = 1e300;
        for (int rep = 0; rep < 4; ++rep) {
            std::vector<Prob> pv = probs, pf = probs_f;
            hipEvent_t a = pool_event(ctx), e = pool_event(ctx);
            (void)hipEventRecord(a, ctx->stream);
            st = run_batch(ctx, dbl, pv);
            if (!st) st = run_batch(ctx, dbl, pf);
            (void)hipEventRecord(e, ctx->stream);
            (void)hipEventSynchronize(e);
            float f = 0.f;
            (void)hipEventElapsedTime(&f, a, e);
            ctx->event_pool.push_back(a);
            ctx->event_pool.push_back(e);
            if (st) break;
            if (rep > 0) ms = std::min(ms, (double)f);
        }
        if (st) break;
        if (ms_out) ms_out[v] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    ctx->timing = timing;
    if (st) {
        if (had) ctx->fft_variant[key] = prev;
        else ctx->fft_variant.erase(key);
        return st;
    }
    ctx->fft_variant[key] = best;
    if (variant_out) *variant_out = best;
    return OC_OK;
}

}  // extern "C"
