/*
 * ORACLE — test infrastructure, not product code (see oracle.h).
 *
 * f64 CPU restatement of the reference's propagation hot path, following the
 * Rust source expression by expression (same operation order, no FMA
 * contraction: built with -ffp-contract=off) so that it reproduces the
 * reference's f64 results up to the FFT library's own rounding.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   fft2/ifft2/fftfreq ........ numpy.fft (pocketfft) vectors at odd, even and
 *                               prime sizes; reference fft.rs unit tests
 *                               (round trip, Parseval, delta, convolution,
 *                               plane-wave bin, fftfreq literals).
 *   asm propagate/report ...... hypercam golden fixture (ASM + PhaseMask +
 *                               ThinLens chain, padding 1); reference asm.rs
 *                               and tests/asm.rs closed forms (z=0 identity,
 *                               e^{ikz}, e^{inkt}, round trip, strict guard,
 *                               energy); padded, band-limited ASM against a
 *                               direct DFT-matrix evaluation on small even
 *                               and odd padded grids (limit active and not)
 *                               and against numpy's pocketfft at the
 *                               benchmark geometries.
 *   fresnel TF / convolution .. exact bin phase, round trip, closed-form step
 *                               vs sweep, TF == convolution at critical
 *                               sampling, delta -> sampled kernel.
 *   rs propagate .............. delta -> sampled kernel (1e-12), -z kernel is
 *                               conj (bit-exact), paraxial kernel == Fresnel,
 *                               guard brackets the sweep, direct O(n^4) RS
 *                               quadrature on a small grid, RS ~ ASM.
 *   phasemask apply ........... identity mask exact, bilinear lattice exact on
 *                               linear profiles and zero beyond the span,
 *                               transmission probe == e^{i phi} in a circle.
 *   gs run .................... convention pins (encode_bounds literals,
 *                               monotone best-so-far, single-iteration step
 *                               against a numpy restatement, top-hat
 *                               convergence).  GsInit::Random: the
 *                               ChaCha block function is pinned by the RFC
 *                               7539 ChaCha20 known answers; the seeding
 *                               (PCG32), the 12-round stream order and the
 *                               Uniform<f64> mapping have no vector here
 *                               (rand is not vendored): parity unpinned for
 *                               the stream, checked only for its range,
 *                               determinism and moments.
 *   gs fft form ............... added form (no reference counterpart); pinned
 *                               by a numpy restatement of one iteration and by
 *                               monotone error decrease.
 *   adjoint push_back ......... inner-product identity <1e-12 (band limit on
 *                               and off), central finite differences <1e-4.
 */
#include "oracle.h"

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

const double PI = 3.14159265358979323846264338327950288;  // f64::consts::PI
const double TAU = 2.0 * PI;

struct cx {
    double re, im;
};

// num-complex 0.4 Mul: re = a.re*b.re - a.im*b.im; im = a.re*b.im + a.im*b.re
inline cx cmul(cx a, cx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline cx cadd(cx a, cx b) { return {a.re + b.re, a.im + b.im}; }
inline cx csub(cx a, cx b) { return {a.re - b.re, a.im - b.im}; }
// Complex * f64 (num-complex impl Mul<T>): componentwise
inline cx crmul(cx a, double s) { return {a.re * s, a.im * s}; }
// (theta * Complex64::I).exp() == from_polar(1.0, theta) == (cos, sin)
inline cx cexpi(double th) { return {std::cos(th), std::sin(th)}; }
inline bool cfinite(cx a) { return std::isfinite(a.re) && std::isfinite(a.im); }
inline double cnorm(cx a) { return std::hypot(a.re, a.im); }
inline double carg(cx a) { return std::atan2(a.im, a.re); }
// Rust f64::rem_euclid
inline double rem_euclid(double x, double m) {
    double r = std::fmod(x, m);
    return r < 0.0 ? r + std::fabs(m) : r;
}
// Rust f64::max / min (NaN-ignoring like fmax/fmin)
inline double fmaxr(double a, double b) { return std::fmax(a, b); }
inline double fminr(double a, double b) { return std::fmin(a, b); }

/* ---------------------------------------------------------------------- */
/* FFT: mixed-radix decimation-in-time, unnormalised, e^{-2 pi i kn/N}     */
/* forward (src/propagation/fft.rs:1-27).  The reference's numerics live   */
/* in rustfft 6.4.1 (not vendored); this is an independent restatement of  */
/* the same DFT, pinned against numpy.fft.                                 */
/* ---------------------------------------------------------------------- */
struct Plan {
    size_t n = 0;
    bool inverse = false;
    std::vector<cx> tw;       // tw[i] = exp(s * 2 pi i * i / n), s = -1 fwd
    std::vector<size_t> fac;  // pairs (radix p, remaining length m)
};

void factorize(size_t n, std::vector<size_t>& fac) {
    size_t p = 4;
    while (n > 1) {
        while (n % p) {
            if (p == 4) p = 2;
            else if (p == 2) p = 3;
            else p += 2;
            if (p * p > n) p = n;
        }
        n /= p;
        fac.push_back(p);
        fac.push_back(n);
    }
}

Plan make_plan(size_t n, bool inverse) {
    Plan pl;
    pl.n = n;
    pl.inverse = inverse;
    pl.tw.resize(n);
    const double sgn = inverse ? 1.0 : -1.0;
    for (size_t i = 0; i < n; ++i) {
        double ang = sgn * (TAU * (double)i / (double)n);
        pl.tw[i] = {std::cos(ang), std::sin(ang)};
    }
    if (n > 1) factorize(n, pl.fac);
    return pl;
}

void bfly2(cx* out, size_t fstride, const Plan& pl, size_t m) {
    for (size_t k = 0; k < m; ++k) {
        cx t = cmul(out[m + k], pl.tw[k * fstride]);
        out[m + k] = csub(out[k], t);
        out[k] = cadd(out[k], t);
    }
}

void bfly4(cx* out, size_t fstride, const Plan& pl, size_t m) {
    for (size_t k = 0; k < m; ++k) {
        cx s0 = cmul(out[k + m], pl.tw[k * fstride]);
        cx s1 = cmul(out[k + 2 * m], pl.tw[2 * k * fstride]);
        cx s2 = cmul(out[k + 3 * m], pl.tw[3 * k * fstride]);
        cx s5 = csub(out[k], s1);
        cx a02 = cadd(out[k], s1);
        cx s3 = cadd(s0, s2);
        cx s4 = csub(s0, s2);
        out[k + 2 * m] = csub(a02, s3);
        out[k] = cadd(a02, s3);
        if (!pl.inverse) {
            out[k + m] = {s5.re + s4.im, s5.im - s4.re};
            out[k + 3 * m] = {s5.re - s4.im, s5.im + s4.re};
        } else {
            out[k + m] = {s5.re - s4.im, s5.im + s4.re};
            out[k + 3 * m] = {s5.re + s4.im, s5.im - s4.re};
        }
    }
}

void bfly_generic(cx* out, size_t fstride, const Plan& pl, size_t m, size_t p) {
    std::vector<cx> scratch(p);
    const size_t n = pl.n;
    for (size_t u = 0; u < m; ++u) {
        for (size_t q = 0; q < p; ++q) scratch[q] = out[u + q * m];
        for (size_t q1 = 0; q1 < p; ++q1) {
            size_t k = u + q1 * m;
            size_t twidx = 0;
            cx acc = scratch[0];
            for (size_t q = 1; q < p; ++q) {
                twidx += fstride * k;
                twidx %= n;
                acc = cadd(acc, cmul(scratch[q], pl.tw[twidx]));
            }
            out[k] = acc;
        }
    }
}

void work(cx* out, const cx* in, size_t fstride, const size_t* fac, const Plan& pl) {
    const size_t p = fac[0], m = fac[1];
    if (m == 1) {
        for (size_t i = 0; i < p; ++i) out[i] = in[i * fstride];
    } else {
        for (size_t i = 0; i < p; ++i)
            work(out + i * m, in + i * fstride, fstride * p, fac + 2, pl);
    }
    switch (p) {
        case 2: bfly2(out, fstride, pl, m); break;
        case 4: bfly4(out, fstride, pl, m); break;
        default: bfly_generic(out, fstride, pl, m, p); break;
    }
}

// In-place 1-D transform of a contiguous vector (tmp is scratch of size n).
void fft1(cx* data, std::vector<cx>& tmp, const Plan& pl) {
    if (pl.n <= 1) return;
    std::copy(data, data + pl.n, tmp.begin());
    work(data, tmp.data(), 1, pl.fac.data(), pl);
}

// fft.rs transform2 (:73-108): rows in place, columns gather/transform/
// scatter, then the optional scale sweep.
void transform2(cx* buf, size_t nx, size_t ny, bool inverse, bool scale) {
    Plan px = make_plan(nx, inverse);  // a fresh plan per call (fft.rs:41,53)
    Plan py = make_plan(ny, inverse);
    std::vector<cx> tmp(std::max(nx, ny));
    for (size_t iy = 0; iy < ny; ++iy) fft1(buf + iy * nx, tmp, px);
    std::vector<cx> col(ny);
    for (size_t ix = 0; ix < nx; ++ix) {
        for (size_t r = 0; r < ny; ++r) col[r] = buf[r * nx + ix];
        fft1(col.data(), tmp, py);
        for (size_t r = 0; r < ny; ++r) buf[r * nx + ix] = col[r];
    }
    if (scale) {
        // Complex64::from(1/(nx*ny)) multiplied as a complex number
        const cx s = {1.0 / (double)(nx * ny), 0.0};
        for (size_t i = 0; i < nx * ny; ++i) buf[i] = cmul(buf[i], s);
    }
}

void fft2(cx* buf, size_t nx, size_t ny) { transform2(buf, nx, ny, false, false); }
void ifft2(cx* buf, size_t nx, size_t ny) { transform2(buf, nx, ny, true, true); }

// fft.rs:115-126
std::vector<double> fftfreq(size_t n, double d) {
    std::vector<double> f(n);
    for (size_t k = 0; k < n; ++k) {
        double kk = (k <= (n - 1) / 2) ? (double)k : (double)k - (double)n;
        f[k] = kk / ((double)n * d);
    }
    return f;
}

// propagation/mod.rs:58-64
size_t padded_len(size_t n, double padding) {
    if (padding <= 1.0) return n;
    return (size_t)std::ceil((double)n * padding);
}

bool all_finite(const cx* s, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!cfinite(s[i])) return false;
    return true;
}

// Embed the centred window (asm.rs:481-498), shared by all engines.
std::vector<cx> embed(const cx* in, size_t nx, size_t ny, size_t nxp, size_t nyp) {
    std::vector<cx> buf(nxp * nyp, cx{0.0, 0.0});
    const size_t ox = (nxp - nx) / 2, oy = (nyp - ny) / 2;
    for (size_t iy = 0; iy < ny; ++iy)
        for (size_t ix = 0; ix < nx; ++ix) buf[(iy + oy) * nxp + (ix + ox)] = in[iy * nx + ix];
    return buf;
}

// Extract the centred window (asm.rs:529-534) + Field::new finiteness.
bool extract(const std::vector<cx>& buf, size_t nx, size_t ny, size_t nxp, size_t nyp, cx* out) {
    const size_t ox = (nxp - nx) / 2, oy = (nyp - ny) / 2;
    for (size_t iy = 0; iy < ny; ++iy)
        for (size_t ix = 0; ix < nx; ++ix) out[iy * nx + ix] = buf[(iy + oy) * nxp + (ix + ox)];
    return all_finite(out, nx * ny);
}

/* ---------------------------------------------------------------------- */
/* ASM (src/propagation/asm.rs)                                            */
/* ---------------------------------------------------------------------- */

// asm.rs:258-265
double axis_band_limit(double lambda, double z, double grid_width, bool band_limited, double medium) {
    double lam_n = lambda / medium;
    if (band_limited) {
        double t = 2.0 * z / grid_width;
        return 1.0 / (lam_n * std::sqrt(t * t + 1.0));
    }
    return 1.0 / lam_n;
}

// asm.rs:273-281; returns false for the evanescent (None) case.
bool asm_transfer_value(double fx, double fy, double lambda, double z, double medium, cx* out) {
    double lam_n = lambda / medium;
    double a = lam_n * fx, b = lam_n * fy;
    double s = 1.0 - a * a - b * b;
    if (s < 0.0) return false;
    double k = 2.0 * PI / lam_n;
    *out = cexpi(k * z * std::sqrt(s));
    return true;
}

struct AsmPlan {
    double lambda, medium, z;
    size_t nx_pad, ny_pad;
    double dx, dy, lim_x, lim_y;
    bool band_limited;
};

double asm_step_along(double lam_n, double kz, size_t n_pad, double d, double f_other,
                      double lim_self, double lim_other, bool band_limited) {
    double df = 1.0 / ((double)n_pad * d);
    double nyq = (double)(n_pad / 2) * df;
    double inv = 1.0 / lam_n;
    double f_circle = std::sqrt(fmaxr(inv * inv - f_other * f_other, 0.0));
    double f_edge = fminr(f_circle, nyq);
    if (band_limited) {
        if (std::fabs(f_other) > lim_other) return 0.0;
        f_edge = fminr(f_edge, lim_self);
    }
    double hi = fmaxr(std::floor(f_edge / df), 0.0) * df;
    double lo = fmaxr(hi - df, 0.0);
    auto root = [&](double f) {
        return std::sqrt(fmaxr(1.0 - lam_n * lam_n * (f * f + f_other * f_other), 0.0));
    };
    return kz * std::fabs(root(hi) - root(lo));
}

// asm.rs:307-368
void worst_phase_steps(const AsmPlan& p, double* sx, double* sy) {
    double lam_n = p.lambda / p.medium;
    double kz = (2.0 * PI / lam_n) * std::fabs(p.z);
    double step_x = asm_step_along(lam_n, kz, p.nx_pad, p.dx, 0.0, p.lim_x, p.lim_y, p.band_limited);
    double step_y = asm_step_along(lam_n, kz, p.ny_pad, p.dy, 0.0, p.lim_y, p.lim_x, p.band_limited);
    if (p.band_limited) {
        step_x = fmaxr(step_x, asm_step_along(lam_n, kz, p.nx_pad, p.dx, p.lim_y, p.lim_x, p.lim_y, true));
        step_y = fmaxr(step_y, asm_step_along(lam_n, kz, p.ny_pad, p.dy, p.lim_x, p.lim_y, p.lim_x, true));
    }
    *sx = step_x;
    *sy = step_y;
}

// asm.rs:373-412
oc_asm_report asm_analyze(const oc_band_geom* g, double z, const oc_asm_config* cfg, double padding) {
    size_t nx_pad = padded_len(g->nx, padding), ny_pad = padded_len(g->ny, padding);
    double lim_x = axis_band_limit(g->lambda, z, (double)nx_pad * g->dx, cfg->band_limited != 0, cfg->medium);
    double lim_y = axis_band_limit(g->lambda, z, (double)ny_pad * g->dy, cfg->band_limited != 0, cfg->medium);
    AsmPlan p{g->lambda, cfg->medium, z, nx_pad, ny_pad, g->dx, g->dy, lim_x, lim_y, cfg->band_limited != 0};
    double sx, sy;
    worst_phase_steps(p, &sx, &sy);
    oc_asm_report r;
    std::memset(&r, 0, sizeof r);
    r.lambda = g->lambda;
    r.medium = cfg->medium;
    r.z = z;
    r.nx = g->nx;
    r.ny = g->ny;
    r.dx = g->dx;
    r.dy = g->dy;
    r.requested_padding = cfg->padding;
    r.used_padding = padding;
    r.nx_pad = nx_pad;
    r.ny_pad = ny_pad;
    r.band_limit_fx = lim_x;
    r.band_limit_fy = lim_y;
    r.max_phase_step_x = sx;
    r.max_phase_step_y = sy;
    r.nyquist_ok = (sx < PI && sy < PI) ? 1 : 0;
    return r;
}

/* ---------------------------------------------------------------------- */
/* Fresnel TF (src/propagation/fresnel.rs)                                 */
/* ---------------------------------------------------------------------- */

// fresnel.rs:258-262
cx fresnel_transfer_value(double fx, double fy, double lambda, double z) {
    double k = 2.0 * PI / lambda;
    double phase = k * z - PI * lambda * z * (fx * fx + fy * fy);
    return cexpi(phase);
}

// fresnel.rs:267-273
double fresnel_worst_axis_step(size_t n_pad, double d, double lambda, double z) {
    double df = 1.0 / ((double)n_pad * d);
    double f_nyq = (double)(n_pad / 2) * df;
    double hi = fmaxr(std::floor(f_nyq / df), 0.0) * df;
    double lo = fmaxr(hi - df, 0.0);
    return PI * lambda * std::fabs(z) * (hi * hi - lo * lo);
}

// fresnel.rs:280-286
double fresnel_worst_paraxial_error(double dx, double dy, double lambda, double z) {
    double fx = 1.0 / (2.0 * dx);
    double fy = 1.0 / (2.0 * dy);
    double s = (lambda * fx) * (lambda * fx) + (lambda * fy) * (lambda * fy);
    double k = 2.0 * PI / lambda;
    return k * std::fabs(z) * (1.0 - 0.5 * s - std::sqrt(fmaxr(1.0 - s, 0.0)));
}

// fresnel.rs:294-298
double fresnel_worst_kernel_axis_step(size_t n_pad, double d, double lambda, double z) {
    double hi = (double)(n_pad / 2) * d;
    double lo = hi - d;
    return PI * (hi * hi - lo * lo) / (lambda * std::fabs(z));
}

// fresnel.rs:342-373
oc_fresnel_report fresnel_analyze(const oc_band_geom* g, double z, const oc_fresnel_config* cfg, double padding) {
    size_t nx_pad = padded_len(g->nx, padding), ny_pad = padded_len(g->ny, padding);
    oc_fresnel_report r;
    std::memset(&r, 0, sizeof r);
    double sx = fresnel_worst_axis_step(nx_pad, g->dx, g->lambda, z);
    double sy = fresnel_worst_axis_step(ny_pad, g->dy, g->lambda, z);
    double kx = fresnel_worst_kernel_axis_step(nx_pad, g->dx, g->lambda, z);
    double ky = fresnel_worst_kernel_axis_step(ny_pad, g->dy, g->lambda, z);
    double err = fresnel_worst_paraxial_error(g->dx, g->dy, g->lambda, z);
    r.lambda = g->lambda;
    r.z = z;
    r.nx = g->nx;
    r.ny = g->ny;
    r.dx = g->dx;
    r.dy = g->dy;
    r.requested_padding = cfg->padding;
    r.used_padding = padding;
    r.nx_pad = nx_pad;
    r.ny_pad = ny_pad;
    r.max_phase_step_x = sx;
    r.max_phase_step_y = sy;
    r.nyquist_ok = (sx < PI && sy < PI) ? 1 : 0;
    r.max_kernel_step_x = kx;
    r.max_kernel_step_y = ky;
    r.kernel_nyquist_ok = (kx < PI && ky < PI) ? 1 : 0;
    r.max_phase_error_vs_asm = err;
    r.paraxial_ok = err < 0.01 ? 1 : 0;
    return r;
}

// fresnel.rs:406-425
int fresnel_validate(const oc_band_geom* g, const cx* s, double z) {
    if (!std::isfinite(z)) return OC_INVALID_DISTANCE;
    if (!std::isfinite(g->lambda) || g->lambda <= 0.0) return OC_INVALID_WAVELENGTH;
    if (g->nx == 0 || g->ny == 0) return OC_EMPTY_GRID;
    if (!all_finite(s, g->nx * g->ny)) return OC_NON_FINITE_SAMPLE;
    return OC_OK;
}

/* ---------------------------------------------------------------------- */
/* Rayleigh-Sommerfeld (src/propagation/rayleigh_sommerfeld.rs)            */
/* ---------------------------------------------------------------------- */

// rayleigh_sommerfeld.rs:228-235
cx rs_kernel_value(double x, double y, double lambda, double z) {
    double k = 2.0 * PI / lambda;
    double s = z > 0.0 ? 1.0 : -1.0;
    double r = std::sqrt(x * x + y * y + z * z);
    double amp = std::fabs(z) / (2.0 * PI * r * r);
    cx obliq = {1.0 / r, -s * k};
    cx a = {amp * obliq.re, amp * obliq.im};  // f64 * Complex
    return cmul(a, cexpi(s * k * r));
}

// rayleigh_sommerfeld.rs:240-249
double rs_worst_kernel_axis_step(size_t n_pad, double d, double lambda, double z) {
    double k = 2.0 * PI / lambda;
    double x_hi = (double)(n_pad / 2) * d;
    double x_lo = x_hi - d;
    double r_hi = std::sqrt(x_hi * x_hi + z * z);
    double r_lo = std::sqrt(x_lo * x_lo + z * z);
    double geom = k * (r_hi - r_lo);
    double obliq = k * d / (1.0 + k * k * z * z);
    return geom + obliq;
}

// rayleigh_sommerfeld.rs:254-269
double rs_worst_paraxial_excess(size_t nx_pad, size_t ny_pad, double dx, double dy, double lambda, double z) {
    double k = 2.0 * PI / lambda;
    double x = (double)(nx_pad / 2) * dx;
    double y = (double)(ny_pad / 2) * dy;
    double az = std::fabs(z);
    double r = std::sqrt(x * x + y * y + az * az);
    double excess = k * (r - az - (x * x + y * y) / (2.0 * az));
    return fmaxr(excess, 0.0);
}

// DFT-order coordinate (rayleigh_sommerfeld.rs:285-291, fresnel.rs:320-326)
inline double dft_coord(size_t i, size_t n, double d) {
    return (i <= n / 2) ? (double)i * d : ((double)i - (double)n) * d;
}

// rayleigh_sommerfeld.rs:277-301
std::vector<cx> rs_impulse_response(size_t nxp, size_t nyp, double dx, double dy, double lambda, double z) {
    std::vector<cx> h(nxp * nyp);
    for (size_t iy = 0; iy < nyp; ++iy) {
        double y = dft_coord(iy, nyp, dy);
        for (size_t ix = 0; ix < nxp; ++ix) {
            double x = dft_coord(ix, nxp, dx);
            h[iy * nxp + ix] = crmul(rs_kernel_value(x, y, lambda, z), dx * dy);
        }
    }
    return h;
}

// rayleigh_sommerfeld.rs:304-329
oc_rs_report rs_analyze(const oc_band_geom* g, double z, double padding) {
    size_t nx_pad = padded_len(g->nx, padding), ny_pad = padded_len(g->ny, padding);
    oc_rs_report r;
    std::memset(&r, 0, sizeof r);
    double kx = rs_worst_kernel_axis_step(nx_pad, g->dx, g->lambda, z);
    double ky = rs_worst_kernel_axis_step(ny_pad, g->dy, g->lambda, z);
    double err = rs_worst_paraxial_excess(nx_pad, ny_pad, g->dx, g->dy, g->lambda, z);
    r.lambda = g->lambda;
    r.z = z;
    r.nx = g->nx;
    r.ny = g->ny;
    r.dx = g->dx;
    r.dy = g->dy;
    r.padding = padding;
    r.nx_pad = nx_pad;
    r.ny_pad = ny_pad;
    r.max_kernel_step_x = kx;
    r.max_kernel_step_y = ky;
    r.kernel_nyquist_ok = (kx < PI && ky < PI) ? 1 : 0;
    r.max_phase_error_vs_fresnel = err;
    r.paraxial_ok = err < 0.01 ? 1 : 0;
    return r;
}

/* ---------------------------------------------------------------------- */
/* Elements (src/element.rs)                                               */
/* ---------------------------------------------------------------------- */

// element.rs:557-560
inline void to_element_frame(double x, double y, double psi, double* fx, double* fy) {
    double c = std::cos(psi), s = std::sin(psi);
    *fx = c * x + s * y;
    *fy = -s * x + c * y;
}

// element.rs:566-569
inline void rotated_box(double hx, double hy, double psi, double* bx, double* by) {
    double c = std::fabs(std::cos(psi)), s = std::fabs(std::sin(psi));
    *bx = hx * c + hy * s;
    *by = hx * s + hy * c;
}

// element.rs:476-486
bool aperture_contains(const oc_clear_aperture* ap, double x, double y) {
    switch (ap->kind) {
        case OC_APERTURE_FIELD: return true;
        case OC_APERTURE_CIRCLE: return x * x + y * y <= ap->a * ap->a;
        case OC_APERTURE_ELLIPSE: {
            double a2 = ap->a * ap->a, b2 = ap->b * ap->b;
            return b2 * x * x + a2 * y * y <= a2 * b2;
        }
        default: return std::fabs(x) <= ap->a && std::fabs(y) <= ap->b;
    }
}

// element.rs:584-600
int validate_footprint(double hx, double hy, double cx_, double cy_) {
    if (!std::isfinite(hx) || hx <= 0.0 || !std::isfinite(hy) || hy <= 0.0 || !std::isfinite(cx_) ||
        !std::isfinite(cy_))
        return OC_INVALID_APERTURE;
    return OC_OK;
}

// element.rs:415-433
int aperture_validate(const oc_clear_aperture* ap, double cx_, double cy_) {
    switch (ap->kind) {
        case OC_APERTURE_FIELD:
            if (!std::isfinite(cx_) || !std::isfinite(cy_)) return OC_INVALID_AXIS;
            return OC_OK;
        case OC_APERTURE_CIRCLE: return validate_footprint(ap->a, ap->a, cx_, cy_);
        case OC_APERTURE_ELLIPSE:
        case OC_APERTURE_RECTANGLE: return validate_footprint(ap->a, ap->b, cx_, cy_);
        default: return OC_ERR_INVALID_ARGUMENT;
    }
}

// element.rs:543-548 footprint_at; returns false for the field-wide stop.
bool aperture_footprint_at(const oc_clear_aperture* ap, double psi, double* fx, double* fy) {
    switch (ap->kind) {
        case OC_APERTURE_FIELD: return false;
        case OC_APERTURE_CIRCLE: *fx = ap->a; *fy = ap->a; return true;
        default: rotated_box(ap->a, ap->b, psi, fx, fy); return true;
    }
}

// element.rs:217-221 Field::coord
inline double coord(size_t i, size_t n, double d) { return ((double)i - ((double)n - 1.0) / 2.0) * d; }

// element.rs:1189-1207 SampledPhase::at
double sampled_at(const oc_phase_mask* m, double x, double y) {
    double fx = x / m->dx + ((double)m->nx - 1.0) / 2.0;
    double fy = y / m->dy + ((double)m->ny - 1.0) / 2.0;
    double last_x = (double)m->nx - 1.0, last_y = (double)m->ny - 1.0;
    if (fx < 0.0 || fx > last_x || fy < 0.0 || fy > last_y) return 0.0;
    size_t i0 = (size_t)std::floor(fx), j0 = (size_t)std::floor(fy);
    size_t i1 = std::min(i0 + 1, (size_t)m->nx - 1), j1 = std::min(j0 + 1, (size_t)m->ny - 1);
    double tx = fx - (double)i0, ty = fy - (double)j0;
    auto v = [&](size_t i, size_t j) { return m->values[j * m->nx + i]; };
    double a = v(i0, j0) + (v(i1, j0) - v(i0, j0)) * tx;
    double b = v(i0, j1) + (v(i1, j1) - v(i0, j1)) * tx;
    return a + (b - a) * ty;
}

inline double wrap_angle(double psi) { return rem_euclid(psi, 2.0 * PI); }  // common.rs:30-32

// element.rs:1443-1484 (PhaseMask::apply); grid-half extents element.rs:549-553
int phasemask_apply(const oc_band_geom* g, const cx* in, const oc_phase_mask* m, cx* out) {
    const double psi = wrap_angle(m->orientation);
    double fpx, fpy;
    if (aperture_footprint_at(&m->aperture, psi, &fpx, &fpy)) {
        double ghx = ((double)g->nx - 1.0) / 2.0 * g->dx;
        double ghy = ((double)g->ny - 1.0) / 2.0 * g->dy;
        if (std::fabs(m->cx) + fpx > ghx || std::fabs(m->cy) + fpy > ghy) return OC_OUTSIDE_FIELD;
    }
    const size_t nx = g->nx, ny = g->ny;
    std::vector<cx> res(nx * ny);
    for (size_t iy = 0; iy < ny; ++iy) {
        for (size_t ix = 0; ix < nx; ++ix) {
            double x = coord(ix, nx, g->dx), y = coord(iy, ny, g->dy);
            double xr = x - m->cx, yr = y - m->cy;
            double fx, fy;
            to_element_frame(xr, yr, psi, &fx, &fy);
            cx s = in[iy * nx + ix];
            if (!aperture_contains(&m->aperture, fx, fy)) {
                res[iy * nx + ix] = {0.0, 0.0};
                continue;
            }
            double phi = sampled_at(m, fx, fy);
            if (!std::isfinite(phi)) return OC_INVALID_PHASE;
            res[iy * nx + ix] = cmul(s, cexpi(phi));
        }
    }
    std::copy(res.begin(), res.end(), out);
    return OC_OK;
}

/* ---------------------------------------------------------------------- */
/* ASM propagate (asm.rs:438-538)                                          */
/* ---------------------------------------------------------------------- */
int asm_propagate(const oc_band_geom* g, const cx* in, double z, const oc_asm_config* cfg, cx* out,
                  oc_asm_report* rep) {
    if (rep) std::memset(rep, 0, sizeof *rep);
    if (!std::isfinite(z)) return OC_INVALID_DISTANCE;
    const double lambda = g->lambda;
    if (!std::isfinite(lambda) || lambda <= 0.0) return OC_INVALID_WAVELENGTH;
    if (!std::isfinite(cfg->medium) || cfg->medium <= 0.0) return OC_INVALID_MEDIUM;
    const size_t nx = g->nx, ny = g->ny;
    if (nx == 0 || ny == 0) return OC_EMPTY_GRID;
    if (!all_finite(in, nx * ny)) return OC_NON_FINITE_SAMPLE;
    if (z == 0.0) {
        if (out != in) std::copy(in, in + nx * ny, out);
        return OC_OK;
    }
    oc_asm_report r = asm_analyze(g, z, cfg, cfg->padding);
    if (cfg->strict && !r.nyquist_ok) {
        double padding = cfg->padding;
        while (!r.nyquist_ok && padding < cfg->max_padding) {
            padding = fminr(padding * 2.0, cfg->max_padding);
            r = asm_analyze(g, z, cfg, padding);
        }
        if (!r.nyquist_ok) {
            if (rep) *rep = r;
            return OC_ALIASING_RISK;
        }
    }
    if (rep) *rep = r;
    const size_t nxp = r.nx_pad, nyp = r.ny_pad;
    std::vector<cx> buf = embed(in, nx, ny, nxp, nyp);
    fft2(buf.data(), nxp, nyp);
    std::vector<double> fxa = fftfreq(nxp, g->dx), fya = fftfreq(nyp, g->dy);
    for (size_t iy = 0; iy < nyp; ++iy) {
        const double fy = fya[iy];
        for (size_t ix = 0; ix < nxp; ++ix) {
            const double fx = fxa[ix];
            cx& h = buf[iy * nxp + ix];
            bool masked = cfg->band_limited && (std::fabs(fx) > r.band_limit_fx || std::fabs(fy) > r.band_limit_fy);
            if (masked) {
                h = {0.0, 0.0};
            } else {
                cx t;
                if (asm_transfer_value(fx, fy, lambda, z, cfg->medium, &t)) h = cmul(h, t);
                else h = {0.0, 0.0};
            }
        }
    }
    ifft2(buf.data(), nxp, nyp);
    if (!extract(buf, nx, ny, nxp, nyp, out)) return OC_NON_FINITE_SAMPLE;
    return OC_OK;
}

}  // namespace

/* ====================================================================== */
/* C ABI                                                                   */
/* ====================================================================== */
extern "C" {

void or_fft2(double* buf, uint64_t nx, uint64_t ny) { fft2(reinterpret_cast<cx*>(buf), nx, ny); }
void or_ifft2(double* buf, uint64_t nx, uint64_t ny) { ifft2(reinterpret_cast<cx*>(buf), nx, ny); }
void or_fft1(double* buf, uint64_t n, int inverse) {
    Plan p = make_plan(n, inverse != 0);
    std::vector<cx> tmp(n);
    fft1(reinterpret_cast<cx*>(buf), tmp, p);
}
void or_fftfreq(uint64_t n, double d, double* out) {
    std::vector<double> f = fftfreq(n, d);
    std::copy(f.begin(), f.end(), out);
}
uint64_t or_padded_len(uint64_t n, double padding) { return padded_len(n, padding); }

int or_asm_sampling_report(const oc_band_geom* g, double z, const oc_asm_config* cfg, oc_asm_report* rep) {
    *rep = asm_analyze(g, z, cfg, cfg->padding);
    return OC_OK;
}

int or_asm_propagate(const oc_band_geom* g, const double* in, double z, const oc_asm_config* cfg, double* out,
                     oc_asm_report* rep) {
    return asm_propagate(g, reinterpret_cast<const cx*>(in), z, cfg, reinterpret_cast<cx*>(out), rep);
}

int or_fresnel_sampling_report(const oc_band_geom* g, double z, const oc_fresnel_config* cfg,
                               oc_fresnel_report* rep) {
    *rep = fresnel_analyze(g, z, cfg, cfg->padding);
    return OC_OK;
}

// fresnel.rs:430-509
int or_fresnel_propagate(const oc_band_geom* g, const double* in_, double z, const oc_fresnel_config* cfg,
                         double* out_, oc_fresnel_report* rep) {
    const cx* in = reinterpret_cast<const cx*>(in_);
    cx* out = reinterpret_cast<cx*>(out_);
    if (rep) std::memset(rep, 0, sizeof *rep);
    int st = fresnel_validate(g, in, z);
    if (st) return st;
    const size_t nx = g->nx, ny = g->ny;
    if (z == 0.0) {
        if (out != in) std::copy(in, in + nx * ny, out);
        return OC_OK;
    }
    oc_fresnel_report r = fresnel_analyze(g, z, cfg, cfg->padding);
    if (cfg->strict && !r.nyquist_ok) {
        double padding = cfg->padding;
        while (!r.nyquist_ok && padding < cfg->max_padding) {
            padding = fminr(padding * 2.0, cfg->max_padding);
            r = fresnel_analyze(g, z, cfg, padding);
        }
        if (!r.nyquist_ok) {
            if (rep) *rep = r;
            return OC_ALIASING_RISK;
        }
    }
    if (rep) *rep = r;
    const size_t nxp = r.nx_pad, nyp = r.ny_pad;
    std::vector<cx> buf = embed(in, nx, ny, nxp, nyp);
    fft2(buf.data(), nxp, nyp);
    std::vector<double> fxa = fftfreq(nxp, g->dx), fya = fftfreq(nyp, g->dy);
    for (size_t iy = 0; iy < nyp; ++iy)
        for (size_t ix = 0; ix < nxp; ++ix)
            buf[iy * nxp + ix] = cmul(buf[iy * nxp + ix], fresnel_transfer_value(fxa[ix], fya[iy], g->lambda, z));
    ifft2(buf.data(), nxp, nyp);
    if (!extract(buf, nx, ny, nxp, nyp, out)) return OC_NON_FINITE_SAMPLE;
    return OC_OK;
}

// fresnel.rs:531-590 (convolution form)
int or_fresnel_propagate_convolution(const oc_band_geom* g, const double* in_, double z,
                                     const oc_fresnel_config* cfg, double* out_, oc_fresnel_report* rep) {
    const cx* in = reinterpret_cast<const cx*>(in_);
    cx* out = reinterpret_cast<cx*>(out_);
    if (rep) std::memset(rep, 0, sizeof *rep);
    int st = fresnel_validate(g, in, z);
    if (st) return st;
    const size_t nx = g->nx, ny = g->ny;
    if (z == 0.0) {
        if (out != in) std::copy(in, in + nx * ny, out);
        return OC_OK;
    }
    oc_fresnel_report r = fresnel_analyze(g, z, cfg, cfg->padding);
    if (rep) *rep = r;
    if (cfg->strict && !r.kernel_nyquist_ok) return OC_ALIASING_RISK;
    const size_t nxp = r.nx_pad, nyp = r.ny_pad;
    const double lambda = g->lambda, dx = g->dx, dy = g->dy;
    std::vector<cx> buf = embed(in, nx, ny, nxp, nyp);
    fft2(buf.data(), nxp, nyp);
    // impulse_response (fresnel.rs:307-333)
    double k = 2.0 * PI / lambda;
    cx e = cexpi(k * z);
    cx den = {0.0 * lambda * z, 1.0 * lambda * z};  // I * lambda * z
    double ns = den.re * den.re + den.im * den.im;
    cx pref = {(e.re * den.re + e.im * den.im) / ns, (e.im * den.re - e.re * den.im) / ns};
    pref = crmul(pref, dx * dy);
    std::vector<cx> h(nxp * nyp);
    for (size_t iy = 0; iy < nyp; ++iy) {
        double y = dft_coord(iy, nyp, dy);
        for (size_t ix = 0; ix < nxp; ++ix) {
            double x = dft_coord(ix, nxp, dx);
            h[iy * nxp + ix] = cmul(pref, cexpi(PI * (x * x + y * y) / (lambda * z)));
        }
    }
    fft2(h.data(), nxp, nyp);
    for (size_t i = 0; i < nxp * nyp; ++i) buf[i] = cmul(buf[i], h[i]);
    ifft2(buf.data(), nxp, nyp);
    if (!extract(buf, nx, ny, nxp, nyp, out)) return OC_NON_FINITE_SAMPLE;
    return OC_OK;
}

void or_fresnel_transfer_value(double fx, double fy, double lambda, double z, double* out2) {
    cx v = fresnel_transfer_value(fx, fy, lambda, z);
    out2[0] = v.re;
    out2[1] = v.im;
}
double or_fresnel_worst_axis_step(uint64_t n_pad, double d, double lambda, double z) {
    return fresnel_worst_axis_step(n_pad, d, lambda, z);
}

// fraunhofer.rs:289-334
int or_fraunhofer_regime_report(const oc_band_geom* g, double z, const oc_fraunhofer_config* cfg,
                                oc_fraunhofer_report* rep) {
    const size_t nx = g->nx, ny = g->ny;
    const double dx = g->dx, dy = g->dy, lambda = g->lambda, az = std::fabs(z);
    const double xs = cfg->has_source_pos ? cfg->source_x : 0.0;
    const double ys = cfg->has_source_pos ? cfg->source_y : 0.0;
    const double hx = ((double)nx - 1.0) / 2.0 * dx;
    const double hy = ((double)ny - 1.0) / 2.0 * dy;
    const bool ap = cfg->has_aperture_extent && std::isfinite(cfg->aperture_extent) && cfg->aperture_extent > 0.0;
    const double rx = ap ? cfg->aperture_extent : hx + std::fabs(xs);
    const double ry = ap ? cfg->aperture_extent : hy + std::fabs(ys);
    std::memset(rep, 0, sizeof *rep);
    rep->lambda = lambda;
    rep->z = z;
    rep->nx = nx;
    rep->ny = ny;
    rep->dx = dx;
    rep->dy = dy;
    rep->dx_out = lambda * az / ((double)nx * dx);
    rep->dy_out = lambda * az / ((double)ny * dy);
    rep->n_f_x = rx * rx / (lambda * az);
    rep->n_f_y = ry * ry / (lambda * az);
    const double n_f = fmaxr(rep->n_f_x, rep->n_f_y);
    rep->source_x = xs;
    rep->source_y = ys;
    rep->extent_used = fmaxr(rx, ry);
    rep->extent_is_aperture = ap ? 1 : 0;
    rep->quad_phase_error_rad = PI * n_f;
    rep->max_n_f = cfg->max_n_f;
    rep->regime_ok = n_f <= cfg->max_n_f ? 1 : 0;
    return OC_OK;
}

// fraunhofer.rs:351-443: offset DFT with the pre-phase (and the axis ramp),
// unpadded fft2, then the bin map (mirrored for z < 0), post-phase, chirp
// about the axis, e^{ikz}/(i lambda z) and dx dy.
int or_fraunhofer_propagate(const oc_band_geom* g, const double* in_, double z, const oc_fraunhofer_config* cfg,
                            double* out_, oc_fraunhofer_report* rep) {
    const cx* in = reinterpret_cast<const cx*>(in_);
    cx* out = reinterpret_cast<cx*>(out_);
    if (rep) std::memset(rep, 0, sizeof *rep);
    if (!std::isfinite(z) || z == 0.0) return OC_INVALID_DISTANCE;
    const double lambda = g->lambda;
    if (!std::isfinite(lambda) || lambda <= 0.0) return OC_INVALID_WAVELENGTH;
    const size_t nx = g->nx, ny = g->ny;
    if (nx == 0 || ny == 0) return OC_EMPTY_GRID;
    for (size_t i = 0; i < nx * ny; ++i)
        if (!cfinite(in[i])) return OC_NON_FINITE_SAMPLE;
    oc_fraunhofer_report r;
    or_fraunhofer_regime_report(g, z, cfg, &r);
    if (rep) *rep = r;
    if (cfg->strict && !r.regime_ok) return OC_REGIME_VIOLATION;
    const double dx = g->dx, dy = g->dy, dxo = r.dx_out, dyo = r.dy_out;
    const double xs = r.source_x, ys = r.source_y;
    const double k = 2.0 * PI / lambda;
    const double f_s = xs / (lambda * z), g_s = ys / (lambda * z);
    const double cxo = ((double)nx - 1.0) / 2.0, cyo = ((double)ny - 1.0) / 2.0;
    std::vector<cx> buf(nx * ny);
    for (size_t iy = 0; iy < ny; ++iy) {
        const double eta = ((double)iy - cyo) * dy;
        const cx pre_y = cexpi(2.0 * PI * cyo * (double)iy / (double)ny);
        const cx ramp_y = cexpi(2.0 * PI * g_s * eta);
        for (size_t ix = 0; ix < nx; ++ix) {
            const double xi = ((double)ix - cxo) * dx;
            const cx pre = cmul(cmul(pre_y, ramp_y), cexpi(2.0 * PI * (cxo * (double)ix / (double)nx + f_s * xi)));
            buf[iy * nx + ix] = cmul(in[iy * nx + ix], pre);
        }
    }
    fft2(buf.data(), nx, ny);
    const double sgn = z > 0.0 ? 1.0 : -1.0;
    // Complex::exp(i k z) / (Complex::I * lambda * z) with num-complex Div
    const cx e = cexpi(k * z);
    const cx den = {0.0 * lambda * z, 1.0 * lambda * z};
    const double nsq = den.re * den.re + den.im * den.im;
    const cx quad = {(e.re * den.re + e.im * den.im) / nsq, (e.im * den.re - e.re * den.im) / nsq};
    for (size_t iy = 0; iy < ny; ++iy) {
        const double dyp = ((double)iy - cyo) * dyo - ys;
        const size_t my = (size_t)std::round(sgn * ((double)iy - cyo) + cyo);
        const cx post_y = cexpi(2.0 * PI * cyo * ((double)my - cyo) / (double)ny);
        for (size_t ix = 0; ix < nx; ++ix) {
            const double dxp = ((double)ix - cxo) * dxo - xs;
            const size_t mx = (size_t)std::round(sgn * ((double)ix - cxo) + cxo);
            const cx post = cmul(post_y, cexpi(2.0 * PI * cxo * ((double)mx - cxo) / (double)nx));
            const cx chirp = cexpi(PI * (dxp * dxp + dyp * dyp) / (lambda * z));
            out[iy * nx + ix] = crmul(cmul(cmul(cmul(quad, chirp), post), buf[my * nx + mx]), dx * dy);
        }
    }
    for (size_t i = 0; i < nx * ny; ++i)
        if (!cfinite(out[i])) return OC_NON_FINITE_SAMPLE;
    return OC_OK;
}

int or_rs_sampling_report(const oc_band_geom* g, double z, const oc_rs_config* cfg, oc_rs_report* rep) {
    *rep = rs_analyze(g, z, cfg->padding);
    return OC_OK;
}

// rayleigh_sommerfeld.rs:345-442
int or_rs_propagate(const oc_band_geom* g, const double* in_, double z, const oc_rs_config* cfg, double* out_,
                    oc_rs_report* rep) {
    const cx* in = reinterpret_cast<const cx*>(in_);
    cx* out = reinterpret_cast<cx*>(out_);
    if (rep) std::memset(rep, 0, sizeof *rep);
    if (!std::isfinite(z) || z == 0.0) return OC_INVALID_DISTANCE;
    if (!std::isfinite(g->lambda) || g->lambda <= 0.0) return OC_INVALID_WAVELENGTH;
    const size_t nx = g->nx, ny = g->ny;
    if (nx == 0 || ny == 0) return OC_EMPTY_GRID;
    if (!all_finite(in, nx * ny)) return OC_NON_FINITE_SAMPLE;
    oc_rs_report r = rs_analyze(g, z, cfg->padding);
    if (rep) *rep = r;
    if (cfg->strict && !r.kernel_nyquist_ok) return OC_ALIASING_RISK;
    const size_t nxp = r.nx_pad, nyp = r.ny_pad;
    std::vector<cx> buf = embed(in, nx, ny, nxp, nyp);
    fft2(buf.data(), nxp, nyp);
    std::vector<cx> h = rs_impulse_response(nxp, nyp, g->dx, g->dy, g->lambda, z);
    fft2(h.data(), nxp, nyp);
    for (size_t i = 0; i < nxp * nyp; ++i) buf[i] = cmul(buf[i], h[i]);
    ifft2(buf.data(), nxp, nyp);
    if (!extract(buf, nx, ny, nxp, nyp, out)) return OC_NON_FINITE_SAMPLE;
    return OC_OK;
}

void or_rs_kernel_value(double x, double y, double lambda, double z, double* out2) {
    cx v = rs_kernel_value(x, y, lambda, z);
    out2[0] = v.re;
    out2[1] = v.im;
}
void or_rs_impulse_response(uint64_t nx_pad, uint64_t ny_pad, double dx, double dy, double lambda, double z,
                            double* out) {
    std::vector<cx> h = rs_impulse_response(nx_pad, ny_pad, dx, dy, lambda, z);
    std::memcpy(out, h.data(), h.size() * sizeof(cx));
}
double or_rs_worst_kernel_axis_step(uint64_t n_pad, double d, double lambda, double z) {
    return rs_worst_kernel_axis_step(n_pad, d, lambda, z);
}

// element.rs:1321-1368 (PhaseMask::sampled validation order)
int or_phasemask_validate(const oc_phase_mask* m) {
    if (!std::isfinite(wrap_angle(m->orientation))) return OC_INVALID_ORIENTATION;
    if (m->nx == 0 || m->ny == 0 || m->values == nullptr) return OC_INVALID_GRID;
    if (!std::isfinite(m->dx) || m->dx <= 0.0 || !std::isfinite(m->dy) || m->dy <= 0.0)
        return OC_INVALID_GRID_PITCH;
    for (size_t k = 0; k < m->nx * m->ny; ++k)
        if (!std::isfinite(m->values[k])) return OC_INVALID_PHASE;
    return aperture_validate(&m->aperture, m->cx, m->cy);
}

int or_phasemask_apply(const oc_band_geom* g, const double* in, const oc_phase_mask* m, double* out) {
    return phasemask_apply(g, reinterpret_cast<const cx*>(in), m, reinterpret_cast<cx*>(out));
}

double or_sampled_phase_at(const oc_phase_mask* m, double x, double y) { return sampled_at(m, x, y); }

// element.rs:754-790 (EllipticalAperture::apply)
int or_elliptical_aperture_apply(const oc_band_geom* g, const double* in_, double a, double b, double psi_in,
                                 double cx_, double cy_, double* out_) {
    const cx* in = reinterpret_cast<const cx*>(in_);
    cx* out = reinterpret_cast<cx*>(out_);
    const double psi = wrap_angle(psi_in);
    double bx = a, by = b;
    if (a != b) rotated_box(a, b, psi, &bx, &by);
    double ghx = ((double)g->nx - 1.0) / 2.0 * g->dx;
    double ghy = ((double)g->ny - 1.0) / 2.0 * g->dy;
    if (std::fabs(cx_) + bx > ghx || std::fabs(cy_) + by > ghy) return OC_OUTSIDE_FIELD;
    const double a2 = a * a, b2 = b * b, limit = a2 * b2;
    for (size_t iy = 0; iy < g->ny; ++iy)
        for (size_t ix = 0; ix < g->nx; ++ix) {
            double x = coord(ix, g->nx, g->dx), y = coord(iy, g->ny, g->dy);
            double fx, fy;
            to_element_frame(x - cx_, y - cy_, psi, &fx, &fy);
            size_t k = iy * g->nx + ix;
            out[k] = (b2 * fx * fx + a2 * fy * fy > limit) ? cx{0.0, 0.0} : in[k];
        }
    return OC_OK;
}

// element.rs:1115-1145 (ThinLens::apply, fixed focal length)
int or_thinlens_apply(const oc_band_geom* g, const double* in_, double f, const oc_clear_aperture* ap,
                      double psi_in, double cx_, double cy_, double* out_) {
    const cx* in = reinterpret_cast<const cx*>(in_);
    cx* out = reinterpret_cast<cx*>(out_);
    if (!std::isfinite(f) || f == 0.0) return OC_ERR_INVALID_ARGUMENT;
    const double psi = wrap_angle(psi_in);
    double fpx, fpy;
    if (aperture_footprint_at(ap, psi, &fpx, &fpy)) {
        double ghx = ((double)g->nx - 1.0) / 2.0 * g->dx;
        double ghy = ((double)g->ny - 1.0) / 2.0 * g->dy;
        if (std::fabs(cx_) + fpx > ghx || std::fabs(cy_) + fpy > ghy) return OC_OUTSIDE_FIELD;
    }
    const double k = 2.0 * PI / g->lambda;
    for (size_t iy = 0; iy < g->ny; ++iy)
        for (size_t ix = 0; ix < g->nx; ++ix) {
            double x = coord(ix, g->nx, g->dx), y = coord(iy, g->ny, g->dy);
            double xr = x - cx_, yr = y - cy_;
            size_t kk = iy * g->nx + ix;
            double fx, fy;
            to_element_frame(xr, yr, psi, &fx, &fy);
            if (!aperture_contains(ap, fx, fy)) {
                out[kk] = {0.0, 0.0};
                continue;
            }
            double phase = -k * (xr * xr + yr * yr) / (2.0 * f);
            out[kk] = cmul(in[kk], cexpi(phase));
        }
    return OC_OK;
}

// sources.rs:259-298
int or_gaussian(double w0, double amplitude, double cx_, double cy_, double z, double lambda, uint64_t nx,
                uint64_t ny, double dx, double dy, double* out_) {
    cx* out = reinterpret_cast<cx*>(out_);
    if (!std::isfinite(w0) || w0 <= 0.0 || !std::isfinite(amplitude) || !std::isfinite(cx_) || !std::isfinite(cy_))
        return OC_ERR_INVALID_ARGUMENT;
    if (!std::isfinite(z)) return OC_INVALID_DISTANCE;
    if (!std::isfinite(lambda) || lambda <= 0.0) return OC_INVALID_WAVELENGTH;
    double z_r = PI * w0 * w0 / lambda;
    double t = z / z_r;
    double w_z = w0 * std::sqrt(1.0 + t * t);
    double k = 2.0 * PI / lambda;
    double inv_r = z / (z * z + z_r * z_r);
    double carrier = k * z - std::atan(z / z_r);
    double amp = amplitude * w0 / w_z;
    for (size_t iy = 0; iy < ny; ++iy) {
        double y = ((double)iy - ((double)ny - 1.0) / 2.0) * dy - cy_;
        for (size_t ix = 0; ix < nx; ++ix) {
            double x = ((double)ix - ((double)nx - 1.0) / 2.0) * dx - cx_;
            double rho2 = x * x + y * y;
            double envelope = amp * std::exp(-rho2 / (w_z * w_z));
            double phase = carrier + 0.5 * k * rho2 * inv_r;
            cx e = cexpi(phase);
            out[iy * nx + ix] = {envelope * e.re, envelope * e.im};
        }
    }
    return all_finite(out, nx * ny) ? OC_OK : OC_NON_FINITE_SAMPLE;
}

// field.rs:360-386
int or_plane_wave(double theta_x, double theta_y, double amplitude, uint64_t nx, uint64_t ny, double dx, double dy,
                  double lambda, double* out_) {
    cx* out = reinterpret_cast<cx*>(out_);
    if (!std::isfinite(lambda) || lambda <= 0.0) return OC_INVALID_WAVELENGTH;
    double k = 2.0 * PI / lambda;
    double alpha = std::sin(theta_x), beta = std::sin(theta_y);
    for (size_t iy = 0; iy < ny; ++iy) {
        double y = ((double)iy - ((double)ny - 1.0) / 2.0) * dy;
        for (size_t ix = 0; ix < nx; ++ix) {
            double x = ((double)ix - ((double)nx - 1.0) / 2.0) * dx;
            cx e = cexpi(k * (alpha * x + beta * y));
            out[iy * nx + ix] = {amplitude * e.re, amplitude * e.im};
        }
    }
    return all_finite(out, nx * ny) ? OC_OK : OC_NON_FINITE_SAMPLE;
}

void or_splitmix_uniform(uint64_t seed, uint64_t count, double* out) {
    uint64_t state = seed;
    for (uint64_t i = 0; i < count; ++i) {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        out[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
}

/* rand 0.8 StdRng (not vendored with the reference; Cargo.lock pins rand
 * 0.8.8, rand_chacha 0.3.1, rand_core 0.6.4).  Restated from the published
 * algorithms:
 *  - ChaCha block (Bernstein; RFC 7539 section 2.3 with the original 64-bit
 *    counter / 64-bit nonce split that rand_chacha uses): state = constants,
 *    key words, counter lo/hi, stream lo/hi; double rounds of column then
 *    diagonal quarter rounds; output = state after the rounds + input state;
 *  - seed_from_u64 (rand_core 0.6): 8 PCG32 XSH-RR outputs, state advanced
 *    first (mul 6364136223846793005, inc 11634580027462260723), stored
 *    little-endian as the 32-byte key;
 *  - ChaCha12Rng: blocks 0, 1, 2, ... of stream 0, 6 double rounds, words in
 *    order; next_u64 = word[2j] | word[2j+1] << 32;
 *  - Uniform::new(low, high) for f64: scale = high - low, decreased one ulp
 *    at a time while scale * (1 - 2^-52) + low >= high; sample = (f64 with
 *    exponent 0 and mantissa u64 >> 12) - 1, times scale, plus low. */
static inline uint32_t chacha_rotl(uint32_t v, unsigned c) { return (v << c) | (v >> (32u - c)); }

void or_chacha_block(const uint32_t* key, uint64_t counter, uint64_t stream, int double_rounds, uint32_t* out) {
    uint32_t st[16];
    st[0] = 0x61707865u;  // "expa"
    st[1] = 0x3320646eu;  // "nd 3"
    st[2] = 0x79622d32u;  // "2-by"
    st[3] = 0x6b206574u;  // "te k"
    for (int i = 0; i < 8; ++i) st[4 + i] = key[i];
    st[12] = (uint32_t)(counter & 0xffffffffu);
    st[13] = (uint32_t)(counter >> 32);
    st[14] = (uint32_t)(stream & 0xffffffffu);
    st[15] = (uint32_t)(stream >> 32);
    uint32_t w[16];
    for (int i = 0; i < 16; ++i) w[i] = st[i];
    static const int QR[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13},  {2, 6, 10, 14}, {3, 7, 11, 15},
                                 {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13},  {3, 4, 9, 14}};
    for (int r = 0; r < double_rounds; ++r)
        for (int q = 0; q < 8; ++q) {
            uint32_t &a = w[QR[q][0]], &b = w[QR[q][1]], &c = w[QR[q][2]], &d = w[QR[q][3]];
            a += b; d ^= a; d = chacha_rotl(d, 16);
            c += d; b ^= c; b = chacha_rotl(b, 12);
            a += b; d ^= a; d = chacha_rotl(d, 8);
            c += d; b ^= c; b = chacha_rotl(b, 7);
        }
    for (int i = 0; i < 16; ++i) out[i] = w[i] + st[i];
}

namespace {
struct OrStdRng {
    uint32_t key[8];
    uint32_t block[16];
    uint64_t next_block = 0;
    unsigned used = 16;
    explicit OrStdRng(uint64_t s) {
        const uint64_t mul = 6364136223846793005ull, inc = 11634580027462260723ull;
        for (int i = 0; i < 8; ++i) {
            s = s * mul + inc;
            uint32_t xorshifted = (uint32_t)(((s >> 18) ^ s) >> 27);
            unsigned rot = (unsigned)(s >> 59);
            key[i] = rot ? (xorshifted >> rot) | (xorshifted << (32u - rot)) : xorshifted;
        }
    }
    uint64_t u64() {
        if (used == 16) {
            or_chacha_block(key, next_block++, 0, 6, block);
            used = 0;
        }
        uint64_t v = (uint64_t)block[used] | ((uint64_t)block[used + 1] << 32);
        used += 2;
        return v;
    }
};
}  // namespace

void or_std_rng_u64(uint64_t seed, uint64_t count, uint64_t* out) {
    OrStdRng g(seed);
    for (uint64_t i = 0; i < count; ++i) out[i] = g.u64();
}

void or_gs_random_phase(uint64_t seed, double lo, double hi, uint64_t n, double* out) {
    const double TAU = 6.283185307179586;
    double span = hi - lo;  // (hi - lo).clamp(0.0, TAU); NaN stays NaN
    if (span < 0.0) span = 0.0;
    if (span > TAU) span = TAU;
    if (!(span > 0.0)) {
        for (uint64_t k = 0; k < n; ++k) out[k] = lo;
        return;
    }
    // Uniform::new(0.0, span)
    const double low = 0.0, high = span;
    double max_rand;
    {
        uint64_t b = (~0ull >> 12) | (1023ull << 52);
        std::memcpy(&max_rand, &b, 8);
        max_rand -= 1.0;
    }
    double scale = high - low;
    while (scale * max_rand + low >= high) {
        uint64_t b;
        std::memcpy(&b, &scale, 8);
        b -= 1;
        std::memcpy(&scale, &b, 8);
    }
    OrStdRng g(seed);
    for (uint64_t k = 0; k < n; ++k) {
        uint64_t b = (g.u64() >> 12) | (1023ull << 52);
        double v12;
        std::memcpy(&v12, &b, 8);
        double v01 = v12 - 1.0;
        out[k] = lo + (v01 * scale + low);
    }
}

/* ---------------------------------------------------------------------- */
/* Gerchberg-Saxton (crates/optics-inverse/src/gs.rs)                      */
/* ---------------------------------------------------------------------- */
static inline double wrap_pi(double v) { return rem_euclid(v + PI, TAU) - PI; }  // gs.rs:108-110
static inline double encode_bounds(double phi, double lo, double hi) {       // gs.rs:115-121
    if (hi - lo >= TAU) return lo + rem_euclid(phi - lo, TAU);
    return phi < lo ? lo : (phi > hi ? hi : phi);
}

// gs.rs:325-338
static int propagate_samples(const std::vector<cx>& s, const oc_band_geom* g, double z, const oc_asm_config* cfg,
                             std::vector<cx>& out) {
    out.resize(s.size());
    int st = asm_propagate(g, s.data(), z, cfg, out.data(), nullptr);
    return st ? OC_INV_RENDER : OC_OK;
}

// gs.rs:347-597 (merit = MseToTarget, merit.rs:112-141)
int or_gs_run(const or_gs_problem* p, double* phase_out, double* merit_out, uint64_t* best_iteration,
              double* amplitude_mse_out, int32_t* converged_out, uint64_t* iterations_out, or_gs_record* history_out,
              double* target_intensity_out, double* target_amplitude_error_out, double* phase_error_out,
              double* illumination_out) {
    const size_t nx = p->geom.nx, ny = p->geom.ny, n = nx * ny;
    const double lo = p->lo, hi = p->hi;
    for (size_t k = 0; k < n; ++k) {
        double t = p->target_amplitude[k];
        if (!std::isfinite(t)) return OC_INV_NOT_FINITE;
        if (t < 0.0) return OC_INV_OUT_OF_BOUNDS;
    }
    if (!(std::isfinite(p->z_total) && p->z_total > 0.0)) return OC_INV_RENDER;
    if (p->variant == 1 && !(std::isfinite(p->beta) && p->beta > 0.0 && p->beta < 2.0)) return OC_INV_RENDER;
    if (p->variant == 2 && !(std::isfinite(p->beta) && p->beta > 0.0 && p->beta <= 1.0)) return OC_INV_RENDER;
    if (p->max_iterations == 0) return OC_INV_RENDER;
    const double half_z = p->z_total / 2.0;
    std::vector<char> open(n);
    for (size_t iy = 0; iy < ny; ++iy)
        for (size_t ix = 0; ix < nx; ++ix) {
            double x = coord(ix, nx, p->geom.dx), y = coord(iy, ny, p->geom.dy);
            open[iy * nx + ix] = aperture_contains(&p->aperture, x - p->cx, y - p->cy);
        }
    std::vector<cx> input(reinterpret_cast<const cx*>(p->input), reinterpret_cast<const cx*>(p->input) + n);
    std::vector<cx> arriving;
    if (propagate_samples(input, &p->geom, half_z, &p->asm_cfg, arriving)) return OC_INV_RENDER;
    std::vector<double> illum(n);
    for (size_t k = 0; k < n; ++k) illum[k] = cnorm(arriving[k]);
    std::vector<double> phase(n);
    if (p->init_kind == 0) {
        for (size_t k = 0; k < n; ++k) phase[k] = p->carrier_values[k];
    } else if (p->init_kind == 2) {
        for (size_t k = 0; k < n; ++k) phase[k] = p->init_phase[k];
    } else if (p->init_kind == 1) {  // gs.rs:449-458
        or_gs_random_phase(p->seed, lo, hi, n, phase.data());
    } else {
        return OC_ERR_UNSUPPORTED;
    }
    for (size_t k = 0; k < n; ++k) phase[k] = encode_bounds(phase[k], lo, hi);
    const double beta_f = p->variant == 2 ? p->beta : 1.0;
    const double beta_p = p->variant == 1 ? p->beta : 1.0;

    bool have_best = false;
    double best_merit_v = 0.0;
    size_t best_it = 0;
    std::vector<double> best_phase, best_int;
    std::vector<or_gs_record> history;
    std::vector<double> last_target_error(n, 0.0), last_phase_error(n, 0.0);
    bool converged = false;
    std::vector<cx> mask(n), u_t, u_proj(n), v;
    std::vector<double> amp_t(n), intensity(n), err(n), dphi(n);
    for (size_t it = 0; it < p->max_iterations; ++it) {
        for (size_t k = 0; k < n; ++k) {
            if (open[k]) {
                cx e = {std::cos(phase[k]), std::sin(phase[k])};
                mask[k] = {illum[k] * e.re, illum[k] * e.im};  // from_polar
            } else {
                mask[k] = {0.0, 0.0};
            }
        }
        if (propagate_samples(mask, &p->geom, half_z, &p->asm_cfg, u_t)) return OC_INV_RENDER;
        for (size_t k = 0; k < n; ++k) {
            amp_t[k] = cnorm(u_t[k]);
            intensity[k] = amp_t[k] * amp_t[k];
            err[k] = amp_t[k] - p->target_amplitude[k];
        }
        double se = 0.0;
        for (size_t k = 0; k < n; ++k) se += err[k] * err[k];
        double amp_mse = se / (double)n;
        double mm = 0.0;
        for (size_t k = 0; k < n; ++k) {
            double d = intensity[k] - p->merit_target[k];
            mm += d * d;
        }
        double merit = mm / (double)n;
        bool improved = !have_best || merit < best_merit_v;
        if (improved) {
            have_best = true;
            best_merit_v = merit;
            best_it = it;
            best_phase = phase;
            best_int = intensity;
        }
        or_gs_record rec{it, merit, best_merit_v, amp_mse, std::sqrt(amp_mse), 0.0};
        history.push_back(rec);
        last_target_error = err;
        for (size_t k = 0; k < n; ++k) {
            double a = (1.0 - beta_f) * amp_t[k] + beta_f * p->target_amplitude[k];
            double m = cnorm(u_t[k]);
            u_proj[k] = m > 0.0 ? crmul(u_t[k], a / m) : cx{a, 0.0};
        }
        if (propagate_samples(u_proj, &p->geom, -half_z, &p->asm_cfg, v)) return OC_INV_RENDER;
        for (size_t k = 0; k < n; ++k) dphi[k] = wrap_pi(carg(v[k]) - phase[k]);
        for (size_t k = 0; k < n; ++k) phase[k] = encode_bounds(phase[k] + beta_p * dphi[k], lo, hi);
        last_phase_error = dphi;
        double sd = 0.0;
        for (size_t k = 0; k < n; ++k) sd += dphi[k] * dphi[k];
        history.back().phase_update_rms = std::sqrt(sd) / std::sqrt((double)n);
        if (it + 1 > p->patience) {
            double earlier = history[history.size() - 1 - p->patience].best_merit;
            double scale = fmaxr(std::fabs(earlier), 1e-30);
            if ((earlier - best_merit_v) / scale < p->merit_tolerance) {
                converged = true;
                break;
            }
        }
    }
    std::copy(best_phase.begin(), best_phase.end(), phase_out);
    *merit_out = best_merit_v;
    *best_iteration = best_it;
    *amplitude_mse_out = history[best_it].amplitude_mse;
    *converged_out = converged ? 1 : 0;
    *iterations_out = history.size();
    if (history_out) std::copy(history.begin(), history.end(), history_out);
    std::copy(best_int.begin(), best_int.end(), target_intensity_out);
    std::copy(last_target_error.begin(), last_target_error.end(), target_amplitude_error_out);
    std::copy(last_phase_error.begin(), last_phase_error.end(), phase_error_out);
    std::copy(illum.begin(), illum.end(), illumination_out);
    return OC_OK;
}

void or_gs_fft(uint64_t n, const double* A, const double* T, double* phases, uint64_t restarts,
               uint64_t iterations, double* err_out) {
    const size_t nn = n * n;
    std::vector<cx> buf(nn);
    for (uint64_t r = 0; r < restarts; ++r) {
        double* ph = phases + r * nn;
        for (uint64_t it = 0; it < iterations; ++it) {
            for (size_t k = 0; k < nn; ++k) {
                cx e = {std::cos(ph[k]), std::sin(ph[k])};
                buf[k] = {A[k] * e.re, A[k] * e.im};
            }
            fft2(buf.data(), n, n);
            for (size_t k = 0; k < nn; ++k) {
                double m = cnorm(buf[k]);
                double a = T[k];
                buf[k] = m > 0.0 ? crmul(buf[k], a / m) : cx{a, 0.0};
            }
            ifft2(buf.data(), n, n);
            for (size_t k = 0; k < nn; ++k) {
                double d = wrap_pi(carg(buf[k]) - ph[k]);
                ph[k] = encode_bounds(ph[k] + d, 0.0, TAU);
            }
        }
        for (size_t k = 0; k < nn; ++k) {
            cx e = {std::cos(ph[k]), std::sin(ph[k])};
            buf[k] = {A[k] * e.re, A[k] * e.im};
        }
        fft2(buf.data(), n, n);
        double si = 0.0, st = 0.0;
        for (size_t k = 0; k < nn; ++k) {
            si += buf[k].re * buf[k].re + buf[k].im * buf[k].im;
            st += T[k] * T[k];
        }
        double num = 0.0, den = 0.0;
        for (size_t k = 0; k < nn; ++k) {
            double i = (buf[k].re * buf[k].re + buf[k].im * buf[k].im) / si;
            double t = T[k] * T[k] / st;
            num += (i - t) * (i - t);
            den += t * t;
        }
        err_out[r] = std::sqrt(num) / std::sqrt(den);
    }
}

/* ---------------------------------------------------------------------- */
/* Adjoint (crates/optics-inverse/src/adjoint.rs)                          */
/* ---------------------------------------------------------------------- */

// beamline.rs:109-130 with the renderer's ASM closure (observe.rs:195-210)
int or_trace(const or_design_step* steps, uint64_t nsteps, const oc_band_geom* g, const double* input,
             const oc_asm_config* asm_cfg, double* out) {
    const size_t n = g->nx * g->ny;
    std::vector<cx> cur(reinterpret_cast<const cx*>(input), reinterpret_cast<const cx*>(input) + n), tmp(n);
    for (uint64_t s = 0; s < nsteps; ++s) {
        if (steps[s].has_element) {
            int st = phasemask_apply(g, cur.data(), &steps[s].mask, tmp.data());
            if (st) return st;
            cur.swap(tmp);
        }
        if (steps[s].z_after != 0.0) {
            oc_asm_config c = *asm_cfg;
            c.medium = steps[s].medium;
            int st = asm_propagate(g, cur.data(), steps[s].z_after, &c, tmp.data(), nullptr);
            if (st) return st;
            cur.swap(tmp);
        }
    }
    std::memcpy(out, cur.data(), n * sizeof(cx));
    return OC_OK;
}

// fd.rs:139-184; the objective: observe.rs:195-210 then merit.rs:124-137
int or_fd_gradient(const oc_band_geom* g, const or_design_step* designs, uint64_t nsteps, uint64_t dim,
                   const double* up, const double* dn, const double* lo, const double* hi,
                   const oc_asm_config* asm_cfg, const double* input, const double* target, double* grad,
                   double* merits, uint64_t* evaluations) {
    const size_t n = g->nx * g->ny;
    std::vector<double> fld(2 * n);
    auto objective = [&](const or_design_step* steps, double* out) -> int {
        if (or_trace(steps, nsteps, g, input, asm_cfg, fld.data())) return OC_INV_RENDER;
        double s = 0.0;
        for (size_t k = 0; k < n; ++k) {
            const double a = fld[2 * k] * fld[2 * k] + fld[2 * k + 1] * fld[2 * k + 1];
            s += (a - target[k]) * (a - target[k]);
        }
        *out = s / (double)n;
        return OC_OK;
    };
    uint64_t ev = 0;
    for (uint64_t i = 0; i < dim; ++i) {
        grad[i] = 0.0;
        if (merits) merits[2 * i] = merits[2 * i + 1] = 0.0;
    }
    for (uint64_t i = 0; i < dim; ++i) {
        if (lo[i] == hi[i]) continue;
        double fu, fd;
        if (objective(designs + (2 * i) * nsteps, &fu) || objective(designs + (2 * i + 1) * nsteps, &fd))
            return OC_INV_RENDER;
        ev += 2;
        if (merits) {
            merits[2 * i] = fu;
            merits[2 * i + 1] = fd;
        }
        const double d = up[i] - dn[i];
        if (d <= 0.0 || std::isnan(d)) return OC_INV_RENDER;
        const double gi = (fu - fd) / d;
        if (!std::isfinite(gi)) return OC_INV_RENDER;
        grad[i] = gi;
    }
    if (evaluations) *evaluations = ev;
    return OC_OK;
}

// fd.rs:210-244 (one epoch, gradient given), solve.rs:74-78
int or_adam_step(uint64_t dim, double* x, const double* g, double* m, double* v, const double* lo, const double* hi,
                 uint64_t t, const oc_adam_config* cfg) {
    double norm = 0.0;
    for (uint64_t i = 0; i < dim; ++i) norm += g[i] * g[i];
    norm = std::sqrt(norm);
    const double scale = (cfg->has_clip && norm > cfg->clip_norm && norm > 0.0) ? cfg->clip_norm / norm : 1.0;
    const double bc1 = 1.0 - std::pow(cfg->beta1, (double)t);
    const double bc2 = 1.0 - std::pow(cfg->beta2, (double)t);
    for (uint64_t i = 0; i < dim; ++i) {
        const double gi = g[i] * scale;
        m[i] = cfg->beta1 * m[i] + (1.0 - cfg->beta1) * gi;
        v[i] = cfg->beta2 * v[i] + (1.0 - cfg->beta2) * gi * gi;
        const double mh = m[i] / bc1;
        const double vh = v[i] / bc2;
        const double denom = std::sqrt(vh) + cfg->eps;
        if (denom > 0.0) x[i] -= cfg->lr_frac * (hi[i] - lo[i]) * mh / denom;
    }
    for (uint64_t i = 0; i < dim; ++i) {
        double xi = x[i];
        if (xi < lo[i]) xi = lo[i];
        if (xi > hi[i]) xi = hi[i];
        x[i] = xi;
    }
    return OC_OK;
}

// analysis.rs:133-150
void or_mtf(const double* psf, uint64_t nx, uint64_t ny, double dx, double* freqs, double* values) {
    std::vector<double> buf(2 * nx * ny, 0.0);
    for (uint64_t k = 0; k < nx * ny; ++k) buf[2 * k] = psf[k];
    or_fft2(buf.data(), nx, ny);
    const double dc = std::max(std::hypot(buf[0], buf[1]), std::numeric_limits<double>::min());
    const uint64_t half = nx / 2;
    for (uint64_t k = 0; k <= half; ++k) {
        freqs[k] = (double)k / ((double)nx * dx);
        values[k] = std::min(std::hypot(buf[2 * k], buf[2 * k + 1]) / dc, 1.0);
    }
}

// observe.rs:213-314
int or_render(const oc_band_geom* g, const double* bands, uint64_t nb, uint64_t ns, const oc_render_config* rc,
              const or_design_step* designs, uint64_t nsteps, const oc_asm_config* asm_cfg, const double* inputs,
              double* obs) {
    if (nb == 0 || ns == 0) return OC_INV_RENDER;
    if ((rc->deriv_x || rc->deriv_y) && (!std::isfinite(rc->scene_step) || rc->scene_step <= 0.0))
        return OC_INV_RENDER;
    if (rc->deriv_lambda && (!std::isfinite(rc->lambda_step) || rc->lambda_step <= 0.0)) return OC_INV_RENDER;
    const size_t n = g->nx * g->ny;
    const int nd = (rc->deriv_x ? 1 : 0) + (rc->deriv_y ? 1 : 0) + (rc->deriv_lambda ? 1 : 0);
    const int nr = 1 + 2 * nd, nv = rc->deriv_lambda ? 3 : 1;
    const size_t frames = nb * ns;
    // one render_frame: the design at this wavelength traced from the source
    std::vector<double> fld(2 * n);
    auto frame = [&](size_t b, size_t s, int r, int v, double lambda, std::vector<double>& inten) -> int {
        oc_band_geom gi = *g;
        gi.lambda = lambda;
        const double* in = inputs + ((b * ns + s) * nr + r) * 2 * n;
        if (or_trace(designs + (b * nv + v) * nsteps, nsteps, &gi, in, asm_cfg, fld.data())) return OC_INV_RENDER;
        inten.resize(n);
        for (size_t k = 0; k < n; ++k) inten[k] = fld[2 * k] * fld[2 * k] + fld[2 * k + 1] * fld[2 * k + 1];
        return OC_OK;
    };
    auto slot = [&](int c, size_t f) { return obs + ((size_t)c * frames + f) * n; };
    std::vector<double> base, hi, lo;
    for (size_t b = 0; b < nb; ++b) {
        for (size_t s = 0; s < ns; ++s) {
            const size_t f = b * ns + s;
            if (frame(b, s, 0, 0, bands[b], base)) return OC_INV_RENDER;
            for (double v : base)
                if (!std::isfinite(v)) return OC_INV_RENDER;  // check_grid, observe.rs:296-303
            std::copy(base.begin(), base.end(), slot(0, f));
            int c = 1, r = 1;
            auto fd = [&](int v_hi, int v_lo, double l_hi, double l_lo, double step) -> int {
                if (frame(b, s, r, v_hi, l_hi, hi) || frame(b, s, r + 1, v_lo, l_lo, lo)) return OC_INV_RENDER;
                double* d = slot(c, f);
                for (size_t k = 0; k < n; ++k) d[k] = (hi[k] - lo[k]) / (2.0 * step);
                ++c;
                r += 2;
                return OC_OK;
            };
            if (rc->deriv_x && fd(0, 0, bands[b], bands[b], rc->scene_step)) return OC_INV_RENDER;
            if (rc->deriv_y && fd(0, 0, bands[b], bands[b], rc->scene_step)) return OC_INV_RENDER;
            if (rc->deriv_lambda && fd(1, 2, bands[b] + rc->lambda_step, bands[b] - rc->lambda_step, rc->lambda_step))
                return OC_INV_RENDER;
        }
    }
    return OC_OK;
}

// adjoint.rs:268-340 with MseToTarget as the merit (merit.rs:112-166):
// render every (band, scene) frame (observe.rs:195-210, the band's design
// on its input), seed each frame with obs_gradient (merit.rs:143-165:
// 2 (I_f - T)/n for a frame the selector picks, 0 otherwise), return no
// planes when the whole seed is zero (adjoint.rs:297-304), else push every
// frame back through its band's design and add the planes frame by frame,
// band-major (adjoint.rs:310-332: the first frame's plane, then += the
// next).  merit_out = MseToTarget::evaluate (merit.rs:124-137: the
// selected frames' intensities summed in frame order, then the mean
// squared difference), NaN for a selector out of range (evaluate's error).
int or_grid_gradient(const oc_band_geom* g, const double* bands, uint64_t nb, uint64_t ns,
                     const or_design_step* designs, uint64_t nsteps, const oc_asm_config* asm_cfg,
                     const double* inputs, const double* target, int32_t sel_band, int32_t sel_scene,
                     double* dphi_out, uint64_t* n_planes, double* merit_out) {
    const size_t n = g->nx * g->ny;
    if (nb == 0 || ns == 0 || n == 0) return OC_INV_RENDER;
    size_t ne = 0;
    for (uint64_t s = 0; s < nsteps; ++s) ne += designs[s].has_element ? 1 : 0;
    const size_t frames = nb * ns;
    std::vector<double> obs(frames * n);
    std::vector<double> fld(2 * n);
    for (uint64_t b = 0; b < nb; ++b) {
        oc_band_geom gb = *g;
        gb.lambda = bands[b];
        for (uint64_t sc = 0; sc < ns; ++sc) {
            const double* in = inputs + 2 * n * (b * ns + sc);
            if (or_trace(designs + b * nsteps, nsteps, &gb, in, asm_cfg, fld.data())) return OC_INV_RENDER;
            double* o = obs.data() + (b * ns + sc) * n;
            for (size_t k = 0; k < n; ++k) o[k] = fld[2 * k] * fld[2 * k] + fld[2 * k + 1] * fld[2 * k + 1];
        }
    }
    auto picked = [&](uint64_t b, uint64_t sc) {
        return (sel_band < 0 || (uint64_t)sel_band == b) && (sel_scene < 0 || (uint64_t)sel_scene == sc);
    };
    if (merit_out) {
        const bool in_range = (sel_band < 0 || (uint64_t)sel_band < nb) && (sel_scene < 0 || (uint64_t)sel_scene < ns);
        if (!in_range) {
            *merit_out = std::numeric_limits<double>::quiet_NaN();
        } else {
            std::vector<double> acc(n, 0.0);
            for (uint64_t b = 0; b < nb; ++b)
                for (uint64_t sc = 0; sc < ns; ++sc)
                    if (picked(b, sc))
                        for (size_t k = 0; k < n; ++k) acc[k] += obs[(b * ns + sc) * n + k];
            double m = 0.0;
            for (size_t k = 0; k < n; ++k) m += (acc[k] - target[k]) * (acc[k] - target[k]);
            *merit_out = m / (double)n;
        }
    }
    std::vector<double> seed(frames * n, 0.0);
    bool any = false;
    for (uint64_t b = 0; b < nb; ++b)
        for (uint64_t sc = 0; sc < ns; ++sc) {
            if (!picked(b, sc)) continue;
            for (size_t k = 0; k < n; ++k) {
                const double v = 2.0 * (obs[(b * ns + sc) * n + k] - target[k]) / (double)n;
                seed[(b * ns + sc) * n + k] = v;
                any = any || v != 0.0;
            }
        }
    *n_planes = 0;
    if (!any) return OC_OK;
    std::vector<double> plane(ne * n);
    for (uint64_t b = 0; b < nb; ++b) {
        oc_band_geom gb = *g;
        gb.lambda = bands[b];
        for (uint64_t sc = 0; sc < ns; ++sc) {
            const size_t f = b * ns + sc;
            int st = or_adjoint_push_back(designs + b * nsteps, nsteps, &gb, inputs + 2 * n * f, asm_cfg,
                                          seed.data() + f * n, plane.data());
            if (st) return st;
            if (f == 0) std::memcpy(dphi_out, plane.data(), ne * n * sizeof(double));
            else
                for (size_t k = 0; k < ne * n; ++k) dphi_out[k] += plane[k];
        }
    }
    *n_planes = ne;
    return OC_OK;
}

// adjoint.rs:140-246
int or_adjoint_push_back(const or_design_step* steps, uint64_t nsteps, const oc_band_geom* g, const double* input,
                         const oc_asm_config* asm_cfg, const double* dmdi, double* dphi_out) {
    const size_t n = g->nx * g->ny;
    struct Tape {
        bool element;
        size_t eidx;
        std::vector<cx> t, v_after;
        double z, medium;
    };
    std::vector<Tape> tape;
    std::vector<cx> cur(reinterpret_cast<const cx*>(input), reinterpret_cast<const cx*>(input) + n), tmp(n);
    std::vector<cx> ones(n, cx{1.0, 0.0});
    size_t ne = 0;
    for (uint64_t s = 0; s < nsteps; ++s) {
        Tape tp;
        tp.element = steps[s].has_element != 0;
        tp.z = steps[s].z_after;
        tp.medium = steps[s].medium;
        if (tp.element) {
            tp.eidx = ne++;
            tp.t.resize(n);
            if (phasemask_apply(g, ones.data(), &steps[s].mask, tp.t.data())) return OC_INV_ADJOINT;
            if (phasemask_apply(g, cur.data(), &steps[s].mask, tmp.data())) return OC_INV_ADJOINT;
            tp.v_after = tmp;
            cur.swap(tmp);
        }
        if (tp.z != 0.0) {
            oc_asm_config c = *asm_cfg;
            c.medium = tp.medium;
            if (asm_propagate(g, cur.data(), tp.z, &c, tmp.data(), nullptr)) return OC_INV_ADJOINT;
            cur.swap(tmp);
        }
        tape.push_back(std::move(tp));
    }
    std::vector<cx> co(n);
    for (size_t k = 0; k < n; ++k) co[k] = {dmdi[k] * cur[k].re, dmdi[k] * cur[k].im};
    std::fill(dphi_out, dphi_out + ne * n, 0.0);
    for (size_t s = tape.size(); s-- > 0;) {
        const Tape& tp = tape[s];
        if (tp.z != 0.0) {
            oc_asm_config c = *asm_cfg;
            c.medium = tp.medium;
            if (asm_propagate(g, co.data(), -tp.z, &c, tmp.data(), nullptr)) return OC_INV_ADJOINT;
            co.swap(tmp);
        }
        if (tp.element) {
            double* gp = dphi_out + tp.eidx * n;
            for (size_t p = 0; p < n; ++p) {
                cx mc = {co[p].re, -co[p].im};
                gp[p] += -2.0 * cmul(mc, tp.v_after[p]).im;
            }
            for (size_t p = 0; p < n; ++p) {
                cx tc = {tp.t[p].re, -tp.t[p].im};
                cx c = co[p];
                co[p] = {c.re * tc.re - c.im * tc.im, c.im * tc.re + c.re * tc.im};  // MulAssign
            }
        }
    }
    return OC_OK;
}

}  // extern "C"
