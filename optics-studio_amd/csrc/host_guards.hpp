// Host-side sampling guards and reports: O(1) closed forms evaluated on the
// CPU before any launch, exactly as the reference engines evaluate them.
#pragma once

#include <cmath>
#include <cstring>

#include "../../include/optics_gpu.h"

namespace och {

constexpr double PI = 3.14159265358979323846264338327950288;

inline size_t padded_len(size_t n, double padding) {  // propagation/mod.rs:58-64
    if (padding <= 1.0) return n;
    return (size_t)std::ceil((double)n * padding);
}

// asm.rs:258-265
inline double asm_axis_band_limit(double lambda, double z, double width, bool band_limited, double medium) {
    const double lam_n = lambda / medium;
    if (band_limited) {
        const double t = 2.0 * z / width;
        return 1.0 / (lam_n * std::sqrt(t * t + 1.0));
    }
    return 1.0 / lam_n;
}

// One axis of asm.rs:307-368 worst_phase_steps.
inline double asm_step_along(double lam_n, double kz, size_t n_pad, double d, double f_other, double lim_self,
                             double lim_other, bool band_limited) {
    const double df = 1.0 / ((double)n_pad * d);
    const double nyq = (double)(n_pad / 2) * df;
    const double inv = 1.0 / lam_n;
    const double f_circle = std::sqrt(std::fmax(inv * inv - f_other * f_other, 0.0));
    double f_edge = std::fmin(f_circle, nyq);
    if (band_limited) {
        if (std::fabs(f_other) > lim_other) return 0.0;
        f_edge = std::fmin(f_edge, lim_self);
    }
    const double hi = std::fmax(std::floor(f_edge / df), 0.0) * df;
    const double lo = std::fmax(hi - df, 0.0);
    auto root = [&](double f) {
        return std::sqrt(std::fmax(1.0 - lam_n * lam_n * (f * f + f_other * f_other), 0.0));
    };
    return kz * std::fabs(root(hi) - root(lo));
}

// asm.rs:373-412
inline oc_asm_report asm_analyze(const oc_band_geom& g, double z, const oc_asm_config& c, double padding) {
    oc_asm_report r;
    std::memset(&r, 0, sizeof r);
    const size_t nxp = padded_len(g.nx, padding), nyp = padded_len(g.ny, padding);
    const bool bl = c.band_limited != 0;
    const double lim_x = asm_axis_band_limit(g.lambda, z, (double)nxp * g.dx, bl, c.medium);
    const double lim_y = asm_axis_band_limit(g.lambda, z, (double)nyp * g.dy, bl, c.medium);
    const double lam_n = g.lambda / c.medium;
    const double kz = (2.0 * PI / lam_n) * std::fabs(z);
    double sx = asm_step_along(lam_n, kz, nxp, g.dx, 0.0, lim_x, lim_y, bl);
    double sy = asm_step_along(lam_n, kz, nyp, g.dy, 0.0, lim_y, lim_x, bl);
    if (bl) {
        sx = std::fmax(sx, asm_step_along(lam_n, kz, nxp, g.dx, lim_y, lim_x, lim_y, true));
        sy = std::fmax(sy, asm_step_along(lam_n, kz, nyp, g.dy, lim_x, lim_y, lim_x, true));
    }
    r.lambda = g.lambda;
    r.medium = c.medium;
    r.z = z;
    r.nx = g.nx;
    r.ny = g.ny;
    r.dx = g.dx;
    r.dy = g.dy;
    r.requested_padding = c.padding;
    r.used_padding = padding;
    r.nx_pad = nxp;
    r.ny_pad = nyp;
    r.band_limit_fx = lim_x;
    r.band_limit_fy = lim_y;
    r.max_phase_step_x = sx;
    r.max_phase_step_y = sy;
    r.nyquist_ok = (sx < PI && sy < PI) ? 1 : 0;
    return r;
}

// fresnel.rs:267-298
inline double fresnel_axis_step(size_t n_pad, double d, double lambda, double z) {
    const double df = 1.0 / ((double)n_pad * d);
    const double f_nyq = (double)(n_pad / 2) * df;
    const double hi = std::fmax(std::floor(f_nyq / df), 0.0) * df;
    const double lo = std::fmax(hi - df, 0.0);
    return PI * lambda * std::fabs(z) * (hi * hi - lo * lo);
}
inline double fresnel_paraxial_error(double dx, double dy, double lambda, double z) {
    const double fx = 1.0 / (2.0 * dx), fy = 1.0 / (2.0 * dy);
    const double s = (lambda * fx) * (lambda * fx) + (lambda * fy) * (lambda * fy);
    const double k = 2.0 * PI / lambda;
    return k * std::fabs(z) * (1.0 - 0.5 * s - std::sqrt(std::fmax(1.0 - s, 0.0)));
}
inline double fresnel_kernel_step(size_t n_pad, double d, double lambda, double z) {
    const double hi = (double)(n_pad / 2) * d;
    const double lo = hi - d;
    return PI * (hi * hi - lo * lo) / (lambda * std::fabs(z));
}

// fresnel.rs:342-373
inline oc_fresnel_report fresnel_analyze(const oc_band_geom& g, double z, const oc_fresnel_config& c,
                                         double padding) {
    oc_fresnel_report r;
    std::memset(&r, 0, sizeof r);
    const size_t nxp = padded_len(g.nx, padding), nyp = padded_len(g.ny, padding);
    r.lambda = g.lambda;
    r.z = z;
    r.nx = g.nx;
    r.ny = g.ny;
    r.dx = g.dx;
    r.dy = g.dy;
    r.requested_padding = c.padding;
    r.used_padding = padding;
    r.nx_pad = nxp;
    r.ny_pad = nyp;
    r.max_phase_step_x = fresnel_axis_step(nxp, g.dx, g.lambda, z);
    r.max_phase_step_y = fresnel_axis_step(nyp, g.dy, g.lambda, z);
    r.nyquist_ok = (r.max_phase_step_x < PI && r.max_phase_step_y < PI) ? 1 : 0;
    r.max_kernel_step_x = fresnel_kernel_step(nxp, g.dx, g.lambda, z);
    r.max_kernel_step_y = fresnel_kernel_step(nyp, g.dy, g.lambda, z);
    r.kernel_nyquist_ok = (r.max_kernel_step_x < PI && r.max_kernel_step_y < PI) ? 1 : 0;
    r.max_phase_error_vs_asm = fresnel_paraxial_error(g.dx, g.dy, g.lambda, z);
    r.paraxial_ok = r.max_phase_error_vs_asm < 0.01 ? 1 : 0;
    return r;
}

// rayleigh_sommerfeld.rs:240-269
inline double rs_kernel_step(size_t n_pad, double d, double lambda, double z) {
    const double k = 2.0 * PI / lambda;
    const double x_hi = (double)(n_pad / 2) * d;
    const double x_lo = x_hi - d;
    const double r_hi = std::sqrt(x_hi * x_hi + z * z);
    const double r_lo = std::sqrt(x_lo * x_lo + z * z);
    return k * (r_hi - r_lo) + k * d / (1.0 + k * k * z * z);
}
inline double rs_paraxial_excess(size_t nxp, size_t nyp, double dx, double dy, double lambda, double z) {
    const double k = 2.0 * PI / lambda;
    const double x = (double)(nxp / 2) * dx, y = (double)(nyp / 2) * dy;
    const double az = std::fabs(z);
    const double r = std::sqrt(x * x + y * y + az * az);
    return std::fmax(k * (r - az - (x * x + y * y) / (2.0 * az)), 0.0);
}

// rayleigh_sommerfeld.rs:304-329
inline oc_rs_report rs_analyze(const oc_band_geom& g, double z, double padding) {
    oc_rs_report r;
    std::memset(&r, 0, sizeof r);
    const size_t nxp = padded_len(g.nx, padding), nyp = padded_len(g.ny, padding);
    r.lambda = g.lambda;
    r.z = z;
    r.nx = g.nx;
    r.ny = g.ny;
    r.dx = g.dx;
    r.dy = g.dy;
    r.padding = padding;
    r.nx_pad = nxp;
    r.ny_pad = nyp;
    r.max_kernel_step_x = rs_kernel_step(nxp, g.dx, g.lambda, z);
    r.max_kernel_step_y = rs_kernel_step(nyp, g.dy, g.lambda, z);
    r.kernel_nyquist_ok = (r.max_kernel_step_x < PI && r.max_kernel_step_y < PI) ? 1 : 0;
    r.max_phase_error_vs_fresnel = rs_paraxial_excess(nxp, nyp, g.dx, g.dy, g.lambda, z);
    r.paraxial_ok = r.max_phase_error_vs_fresnel < 0.01 ? 1 : 0;
    return r;
}

// fft.rs:115-126
inline double fftfreq_at(size_t k, size_t n, double d) {
    const double kk = (k <= (n - 1) / 2) ? (double)k : (double)k - (double)n;
    return kk / ((double)n * d);
}

inline double wrap_angle(double psi) {  // common.rs:30-32 (rem_euclid into [0, 2 pi))
    const double m = 2.0 * PI;
    const double r = std::fmod(psi, m);
    return r < 0.0 ? r + m : r;
}

// fraunhofer::regime_report (src/propagation/fraunhofer.rs:289-334): Fresnel
// numbers from the declared aperture half-extent or, by default, the window
// extent measured from the optical axis; the far-field pitch lambda|z|/(n d).
inline oc_fraunhofer_report fraunhofer_report(const oc_band_geom& g, double z, const oc_fraunhofer_config& c) {
    oc_fraunhofer_report r;
    std::memset(&r, 0, sizeof r);
    const double az = std::fabs(z);
    const double xs = c.has_source_pos ? c.source_x : 0.0, ys = c.has_source_pos ? c.source_y : 0.0;
    const double hx = ((double)g.nx - 1.0) / 2.0 * g.dx, hy = ((double)g.ny - 1.0) / 2.0 * g.dy;
    const bool ap = c.has_aperture_extent && std::isfinite(c.aperture_extent) && c.aperture_extent > 0.0;
    const double rx = ap ? c.aperture_extent : hx + std::fabs(xs);
    const double ry = ap ? c.aperture_extent : hy + std::fabs(ys);
    r.lambda = g.lambda;
    r.z = z;
    r.nx = g.nx;
    r.ny = g.ny;
    r.dx = g.dx;
    r.dy = g.dy;
    r.dx_out = g.lambda * az / ((double)g.nx * g.dx);
    r.dy_out = g.lambda * az / ((double)g.ny * g.dy);
    r.n_f_x = rx * rx / (g.lambda * az);
    r.n_f_y = ry * ry / (g.lambda * az);
    const double nf = std::fmax(r.n_f_x, r.n_f_y);
    r.source_x = xs;
    r.source_y = ys;
    r.extent_used = std::fmax(rx, ry);
    r.extent_is_aperture = ap ? 1 : 0;
    r.quad_phase_error_rad = PI * nf;
    r.max_n_f = c.max_n_f;
    r.regime_ok = nf <= c.max_n_f ? 1 : 0;
    return r;
}

}  // namespace och
