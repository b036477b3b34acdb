// Device-side element-wise operations shared by the generic and the
// power-of-two pass kernels: phase-mask transmission, transfer functions,
// the RS kernel sample, fused pre/post ops, accumulators and tile lookup.
#pragma once

#include <hip/hip_runtime.h>

#include "prop_kernels.hpp"

namespace ocd {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kTwoPiHi = 6.28318530717958623199592693708837032;   // nearest double to 2 pi
constexpr double kTwoPiLo = 2.44929359829470635445213186455000e-16;  // 2 pi - kTwoPiHi
constexpr double kInvTwoPi = 0.159154943091895335768883763372514362;

// sin/cos of a phase argument.  Argument reduction modulo pi/2 is done in
// FP64 by Cody-Waite with a three-part pi/2 (the product n*C1 and n*C2 are
// exact for |n| < 2^31, so phases up to ~3e9 rad reduce to ~1e-16 absolute);
// the reduced |x| <= pi/4 goes through Taylor polynomials of degree 17/18
// in FP64 (truncation < 5e-17) for the complex128 path.  The quadrant
// n mod 4 then rotates (sin, cos).
constexpr double kPio2_1 = 1.57079632673412561416626;  // pi/2 rounded to 33 bits
constexpr double kPio2_2 = 6.077100506303965976595549e-11;  // next 33 bits
constexpr double kPio2_3 = 2.022266248795950631541144e-21;  // remainder (53 bits)
constexpr double kInvPio2 = 6.36619772367581382433e-01;

__device__ __forceinline__ double reduce_pio2(double th, int* quad) {
    const double n = rint(th * kInvPio2);
    double r = fma(-n, kPio2_1, th);
    r = fma(-n, kPio2_2, r);
    r = fma(-n, kPio2_3, r);
    *quad = (int)n & 3;  // |n| < 2^31 (see above)
    return r;
}

// (cos, sin) of x + q pi/2 from (sin x, cos x), with selects (no branches):
// q = 0: (c, s); 1: (-s, c); 2: (-c, -s); 3: (s, -c).
template <typename R>
__device__ __forceinline__ cpx<R> quadrant(R s, R c, int q) {
    const R a = (q & 1) ? s : c;
    const R b = (q & 1) ? c : s;
    return {((q + 1) & 2) ? -a : a, (q & 2) ? -b : b};
}

template <typename R>
__device__ __forceinline__ cpx<R> unit_phase(double th);

// complex128 path: (cos th, sin th) to ~1 ulp.
template <>
__device__ __forceinline__ cpx<double> unit_phase<double>(double th) {
    int q;
    const double x = reduce_pio2(th, &q);
    const double x2 = x * x;
    // sin x = x + x^3 P(x^2),  P(u) = -1/3! + u/5! - u^2/7! + ... + u^7/17!
    double ps = 2.81145725434552076319e-15;         // 1/17!
    ps = fma(ps, x2, -7.64716373181981647590e-13);  // -1/15!
    ps = fma(ps, x2, 1.60590438368216145994e-10);   // 1/13!
    ps = fma(ps, x2, -2.50521083854417187751e-08);  // -1/11!
    ps = fma(ps, x2, 2.75573192239858906526e-06);   // 1/9!
    ps = fma(ps, x2, -1.98412698412698412698e-04);  // -1/7!
    ps = fma(ps, x2, 8.33333333333333333333e-03);   // 1/5!
    ps = fma(ps, x2, -1.66666666666666666667e-01);  // -1/3!
    const double s = fma(x * x2, ps, x);
    // cos x = 1 + x^2 Q(x^2),  Q(u) = -1/2 + u/4! - u^2/6! + ... - u^8/18!
    double pc = -1.56192069685862264622e-16;        // -1/18!
    pc = fma(pc, x2, 4.77947733238738529744e-14);   // 1/16!
    pc = fma(pc, x2, -1.14707455977297247139e-11);  // -1/14!
    pc = fma(pc, x2, 2.08767569878680989792e-09);   // 1/12!
    pc = fma(pc, x2, -2.75573192239858906526e-07);  // -1/10!
    pc = fma(pc, x2, 2.48015873015873015873e-05);   // 1/8!
    pc = fma(pc, x2, -1.38888888888888888889e-03);  // -1/6!
    pc = fma(pc, x2, 4.16666666666666666667e-02);   // 1/4!
    pc = fma(pc, x2, -0.5);
    const double c = fma(x2, pc, 1.0);
    return quadrant<double>(s, c, q);
}

// FP32-bulk path: the phase in turns, reduced in FP64 to f = th/(2 pi) - n,
// n = rint(th/(2 pi)), with one rounding (fma), so |f| <= 1/2 carries an
// error of |th| 2^-53 / (2 pi) turns (below 1e-9 rad for |th| < 1e7); f is
// rounded to FP32 (at most 1e-7 rad) and goes to the hardware sine and
// cosine, which take turns (v_sin_f32 / v_cos_f32; measured on gfx950 over
// 2^22 points of [-1/2, 1/2): at most 1.26e-7 absolute error, rms 3.5e-8).
template <>
__device__ __forceinline__ cpx<float> unit_phase<float>(double th) {
    const double n = rint(th * kInvTwoPi);
    const float f = (float)fma(th, kInvTwoPi, -n);
    return {__builtin_amdgcn_cosf(f), __builtin_amdgcn_sinf(f)};
}

template <typename R>
__device__ __forceinline__ bool finite_c(cpx<R> v) {
    return isfinite(v.re) && isfinite(v.im);
}

/* ------------------------------------------------------------------ */
/* Phase-mask transmission (src/element.rs:1189-1207, 1443-1484).       */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ bool ap_contains(const MaskDev& m, double x, double y) {
    switch (m.ap_kind) {
        case 0: return true;
        case 1: return x * x + y * y <= m.ap_a * m.ap_a;
        case 2: {
            const double a2 = m.ap_a * m.ap_a, b2 = m.ap_b * m.ap_b;
            return b2 * x * x + a2 * y * y <= a2 * b2;
        }
        default: return fabs(x) <= m.ap_a && fabs(y) <= m.ap_b;
    }
}

__device__ __forceinline__ double bilinear(const double* __restrict__ v, int mnx, int i0, int i1, int j0, int j1,
                                           double tx, double ty) {
    const gmem<const double>* g = gp(v);
    const double v00 = g[(size_t)j0 * mnx + i0], v10 = g[(size_t)j0 * mnx + i1];
    const double v01 = g[(size_t)j1 * mnx + i0], v11 = g[(size_t)j1 * mnx + i1];
    const double a = v00 + (v10 - v00) * tx;
    const double b = v01 + (v11 - v01) * tx;
    return a + (b - a) * ty;
}

// Returns false when the sample is outside the stop (transmission 0);
// otherwise *phi is the lattice phase (0 beyond the lattice span).
__device__ __forceinline__ bool mask_phase(const MaskDev& m, int ix, int iy, double* phi) {
    if (m.mode == MASK_SEPARABLE) {
        // lattice reads stay in bounds (indices 0 outside the span), so
        // they are issued unconditionally and selected afterwards
        MaskAxis X, Y;
        X = at32(m.sx, ix);
        Y = at32(m.sy, iy);
        const double v = bilinear(m.values, m.mnx, X.i0, X.i1, Y.i0, Y.i1, X.t, Y.t);
        *phi = (X.in & Y.in) ? v : 0.0;
        return ap_contains(m, X.r, Y.r);
    }
    const double x = ((double)ix - ((double)m.fnx - 1.0) / 2.0) * m.fdx;
    const double y = ((double)iy - ((double)m.fny - 1.0) / 2.0) * m.fdy;
    const double xr = x - m.cx, yr = y - m.cy;
    const double fx = m.cpsi * xr + m.spsi * yr;
    const double fy = -m.spsi * xr + m.cpsi * yr;
    if (!ap_contains(m, fx, fy)) return false;
    const double gx = fx / m.mdx + ((double)m.mnx - 1.0) / 2.0;
    const double gy = fy / m.mdy + ((double)m.mny - 1.0) / 2.0;
    if (gx < 0.0 || gx > (double)m.mnx - 1.0 || gy < 0.0 || gy > (double)m.mny - 1.0) {
        *phi = 0.0;
        return true;
    }
    const int i0 = (int)floor(gx), j0 = (int)floor(gy);
    const int i1 = min(i0 + 1, m.mnx - 1), j1 = min(j0 + 1, m.mny - 1);
    *phi = bilinear(m.values, m.mnx, i0, i1, j0, j1, gx - (double)i0, gy - (double)j0);
    return true;
}

template <typename R>
__device__ __forceinline__ cpx<R> apply_mask(const MaskDev& m, int ix, int iy, cpx<R> s) {
    double phi;
    if (!mask_phase(m, ix, iy, &phi)) return {R(0), R(0)};
    return cmul(s, unit_phase<R>(phi));
}

/* ------------------------------------------------------------------ */
/* Transfer functions (asm.rs:273-281, fresnel.rs:258-262).             */
/* ------------------------------------------------------------------ */
// sqrt of an ASM radicand for the phase k z sqrt(s).  complex128: IEEE sqrt
// (the phase argument rounds as the reference's).  complex64: hardware
// reciprocal square root and one Newton step (relative error ~1e-14, a phase
// error below 1e-8 rad at |k z| < 1e6), clamped so that s = 0 gives ~0.
template <typename R>
__device__ __forceinline__ double asm_sqrt(double s) {
    if constexpr (sizeof(R) == 8) {
        return sqrt(s);
    } else {
        const double x = fmax(s, 1e-300);
        const double r = __builtin_amdgcn_rsq(x);
        const double y = x * r;
        return fma(fma(-y, y, x), 0.5 * r, y);
    }
}

// ASM transfer for one sample with its axis table entries (asm.rs:273-281):
// zero when band-limited (an entry is +inf) or evanescent, else
// h exp(i kz sqrt(s)).  Written without branches: every lane computes the
// phase and a select keeps it.
template <typename R>
__device__ __forceinline__ cpx<R> asm_transfer(double kz, double axk, double ayk, cpx<R> h) {
    const double s = (1.0 - axk) - ayk;
    const bool dead = s < 0.0;
    const cpx<R> t = cmul(h, unit_phase<R>(kz * asm_sqrt<R>(dead ? 0.0 : s)));
    return dead ? cpx<R>{R(0), R(0)} : t;
}

// The ASM transfer value itself (asm.rs:273-281): zero when band-limited or
// evanescent, else exp(i kz sqrt(s)); cmul(h, asm_tf(...)) equals
// asm_transfer(..., h) wherever the value is not zero.
template <typename R>
__device__ __forceinline__ cpx<R> asm_tf(double kz, double axk, double ayk) {
    const double s = (1.0 - axk) - ayk;
    const bool dead = s < 0.0;
    const cpx<R> u = unit_phase<R>(kz * asm_sqrt<R>(dead ? 0.0 : s));
    return dead ? cpx<R>{R(0), R(0)} : u;
}

// EK: the engine when known at compile time (ENG_ASM / ENG_FRESNEL), else -1
// and the descriptor's kind decides.
template <typename R, int EK = -1>
__device__ __forceinline__ cpx<R> transfer(const EngineDev& e, int kx, int ky, cpx<R> h) {
    const gmem<const double>* ax = gp(e.ax);
    const gmem<const double>* ay = gp(e.ay);
    if (EK == ENG_ASM || (EK < 0 && e.kind == ENG_ASM))
        return asm_transfer<R>(e.kz, ax[kx], ay[ky], h);
    // Fresnel TF: phase = k z - (pi lambda z) (fx^2 + fy^2)
    if constexpr (sizeof(R) == 4) {
        // complex64: the factored form (the host always provides hx, hy)
        cpx<R> a, b;
        a = at32(static_cast<const cpx<R>*>(e.hx), kx);
        b = at32(static_cast<const cpx<R>*>(e.hy), ky);
        return cmul(h, cmul(a, b));
    } else {
        return cmul(h, unit_phase<R>(e.kz - e.c * (ax[kx] + ay[ky])));
    }
}

// rayleigh_sommerfeld.rs:228-235 sampled at DFT-order coordinates, times dx dy.
template <typename R>
__device__ __forceinline__ cpx<R> rs_sample(const EngineDev& e, int ix, int iy, int px, int py) {
    const double x = (ix <= px / 2) ? (double)ix * e.dx : ((double)ix - (double)px) * e.dx;
    const double y = (iy <= py / 2) ? (double)iy * e.dy : ((double)iy - (double)py) * e.dy;
    const double r = sqrt(x * x + y * y + e.z * e.z);
    const double amp = fabs(e.z) / (2.0 * kPi * r * r);
    const double ore = 1.0 / r, oim = -e.sgn * e.k;
    const double are = amp * ore, aim = amp * oim;
    const double th = e.sgn * e.k * r;
    const double dxdy = e.dx * e.dy;
    if constexpr (sizeof(R) == 8) {
        const cpx<double> u = unit_phase<double>(th);
        const double vre = are * u.re - aim * u.im, vim = are * u.im + aim * u.re;
        return {vre * dxdy, vim * dxdy};
    } else {
        const cpx<float> u = unit_phase<float>(th);
        const double vre = are * (double)u.re - aim * (double)u.im;
        const double vim = are * (double)u.im + aim * (double)u.re;
        return {(float)(vre * dxdy), (float)(vim * dxdy)};
    }
}

// complex64 RS sample with the FP64 reciprocal square root refined by one
// Newton step instead of the IEEE sqrt and two divides: r = rho2 q, 1/r = q,
// 1/r^2 = q^2 with q = 1/sqrt(rho2) to ~1e-14 relative, so the phase k r
// (< ~2e6 rad) keeps an error below 1e-7 rad before the FP32 evaluation.
__device__ __forceinline__ cpx<float> rs_sample_fast(const EngineDev& e, int ix, int iy, int px, int py) {
    const double x = (ix <= px / 2) ? (double)ix * e.dx : ((double)ix - (double)px) * e.dx;
    const double y = (iy <= py / 2) ? (double)iy * e.dy : ((double)iy - (double)py) * e.dy;
    const double rho2 = x * x + y * y + e.z * e.z;
    double q = __builtin_amdgcn_rsq(rho2);
    q = fma(q * 0.5, fma(-rho2 * q, q, 1.0), q);
    const double r = rho2 * q;
    const double amp = fabs(e.z) * (q * q) * (0.5 / kPi);
    const double are = amp * q, aim = amp * (-e.sgn * e.k);
    const cpx<float> u = unit_phase<float>(e.sgn * e.k * r);
    const double dxdy = e.dx * e.dy;
    return {(float)((are * (double)u.re - aim * (double)u.im) * dxdy),
            (float)((are * (double)u.im + aim * (double)u.re) * dxdy)};
}

// Paraxial impulse response of the Fresnel convolution form in DFT order
// (fresnel.rs:312-333): pref exp(i pi (x^2 + y^2)/(lambda z)).
template <typename R>
__device__ __forceinline__ cpx<R> fconv_sample(const EngineDev& e, int ix, int iy, int px, int py) {
    const double x = (ix <= px / 2) ? (double)ix * e.dx : ((double)ix - (double)px) * e.dx;
    const double y = (iy <= py / 2) ? (double)iy * e.dy : ((double)iy - (double)py) * e.dy;
    const double th = kPi * (x * x + y * y) / (e.lambda * e.z);
    if constexpr (sizeof(R) == 8) {
        return cmul(cpx<double>{e.pref_re, e.pref_im}, unit_phase<double>(th));
    } else {
        return cmul(cpx<float>{(float)e.pref_re, (float)e.pref_im}, unit_phase<float>(th));
    }
}

// Kernel sample of a convolution engine.
template <typename R>
__device__ __forceinline__ cpx<R> conv_sample(const EngineDev& e, int ix, int iy, int px, int py) {
    if (e.kind == ENG_FRESNEL_CONV) return fconv_sample<R>(e, ix, iy, px, py);
    if constexpr (sizeof(R) == 4) return rs_sample_fast(e, ix, iy, px, py);
    else return rs_sample<R>(e, ix, iy, px, py);
}

/* ------------------------------------------------------------------ */
/* Fraunhofer offset DFT (fraunhofer.rs:384-437).                       */
/* ------------------------------------------------------------------ */
// Pre-phase of input sample (ix, iy): e^{i 2 pi c_y iy/n_y} e^{i 2 pi g_s eta}
// e^{i 2 pi (c_x ix/n_x + f_s xi)}.  complex128: the three factors and their
// products as the reference forms them; complex64: one phase, summed in f64.
template <typename R>
__device__ __forceinline__ cpx<R> fraun_pre(const EngineDev& e, int ix, int iy, int nx, int ny) {
    const double eta = ((double)iy - e.fcy) * e.dy;
    const double xi = ((double)ix - e.fcx) * e.dx;
    const double ty = 2.0 * kPi * e.fcy * (double)iy / (double)ny;
    const double tr = 2.0 * kPi * e.gs * eta;
    const double tx = 2.0 * kPi * (e.fcx * (double)ix / (double)nx + e.fs * xi);
    if constexpr (sizeof(R) == 8) {
        return cmul(cmul(unit_phase<double>(ty), unit_phase<double>(tr)), unit_phase<double>(tx));
    } else {
        return unit_phase<float>(ty + tr + tx);
    }
}

// Far-field sample for raw bin (mx, my) with value b: its output position
// (ix, iy) (identity for z > 0, mirrored for z < 0) and the value
// e^{ikz}/(i lambda z) chirp(X, Y) post(mx, my) b dx dy.
template <typename R>
__device__ __forceinline__ cpx<R> fraun_post(const EngineDev& e, int mx, int my, int nx, int ny, cpx<R> b, int* ix,
                                             int* iy) {
    *ix = e.sgn > 0.0 ? mx : nx - 1 - mx;
    *iy = e.sgn > 0.0 ? my : ny - 1 - my;
    const double dxp = ((double)*ix - e.fcx) * e.dxo - e.xs;
    const double dyp = ((double)*iy - e.fcy) * e.dyo - e.ys;
    const double py = 2.0 * kPi * e.fcy * ((double)my - e.fcy) / (double)ny;
    const double px = 2.0 * kPi * e.fcx * ((double)mx - e.fcx) / (double)nx;
    const double pc = kPi * (dxp * dxp + dyp * dyp) / (e.lambda * e.z);
    const double dxdy = e.dx * e.dy;
    if constexpr (sizeof(R) == 8) {
        const cpx<double> q = {e.quad_re, e.quad_im};
        const cpx<double> post = cmul(unit_phase<double>(py), unit_phase<double>(px));
        const cpx<double> v = cmul(cmul(cmul(q, unit_phase<double>(pc)), post), b);
        return {v.re * dxdy, v.im * dxdy};
    } else {
        const cpx<float> q = {(float)(e.quad_re * dxdy), (float)(e.quad_im * dxdy)};
        return cmul(cmul(q, unit_phase<float>(pc + py + px)), b);
    }
}

/* ------------------------------------------------------------------ */
/* Fused element-wise ops.                                              */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ double rem_euclid(double x, double m) {
    const double r = fmod(x, m);
    return r < 0.0 ? r + fabs(m) : r;
}
__device__ __forceinline__ double wrap_pi(double v) {  // gs.rs:108-110
    return rem_euclid(v + kPi, 2.0 * kPi) - kPi;
}
__device__ __forceinline__ double encode_bounds(double phi, double lo, double hi) {  // gs.rs:115-121
    if (hi - lo >= 2.0 * kPi) return lo + rem_euclid(phi - lo, 2.0 * kPi);
    return phi < lo ? lo : (phi > hi ? hi : phi);
}

// FP32 forms of wrap_pi / encode_bounds for the complex64 GS update.
// rem_euclid(x, 2 pi) takes the quotient from a product by 1/(2 pi): where
// that quotient's floor is one off (x within an ulp of a multiple of 2 pi)
// the range guard moves the remainder back into [0, 2 pi).
__device__ __forceinline__ float rem_2pi_f(float x) {
    const float two_pi = 6.28318530717959f, inv_two_pi = 0.159154943091895f;
    const float r = x - two_pi * floorf(x * inv_two_pi);
    return r >= two_pi ? r - two_pi : (r < 0.0f ? r + two_pi : r);
}
__device__ __forceinline__ float wrap_pi_f(float v) {
    const float pi = 3.14159265358979f;
    return rem_2pi_f(v + pi) - pi;
}
__device__ __forceinline__ float encode_bounds_f(float phi, float lo, float hi) {
    const float two_pi = 6.28318530717959f;
    if (hi - lo >= two_pi) return lo + rem_2pi_f(phi - lo);
    return phi < lo ? lo : (phi > hi ? hi : phi);
}

// atan2 in FP32 for the complex64 phase updates.  t = min(|x|,|y|) /
// max(|x|,|y|) in [0, 1] (hardware reciprocal plus one Newton step), then
// atan(t) = t p(t^2) with p the degree-8 polynomial fitted here to atan(t)/t
// on [0, 1] by iteratively reweighted least squares toward the minimax fit
// (relative error 1.6e-8 in exact arithmetic, <= 9e-8 absolute in FP32
// Horner), then the octant: pi/2 - r where |y| > |x|, pi - r where x has its
// sign bit set, and y's sign.  (0, 0) gives +-0 or +-pi as IEEE atan2 does;
// non-finite inputs are the caller's to flag.
__device__ __forceinline__ float atan2_f(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    float rc = __builtin_amdgcn_rcpf(mx);
    rc = fmaf(rc, fmaf(-mx, rc, 1.0f), rc);
    const float t = mx > 0.0f ? mn * rc : 0.0f;
    const float s = t * t;
    float p = 2.850835325e-03f;
    p = fmaf(p, s, -1.607220454e-02f);
    p = fmaf(p, s, 4.269698496e-02f);
    p = fmaf(p, s, -7.504727103e-02f);
    p = fmaf(p, s, 1.064112236e-01f);
    p = fmaf(p, s, -1.420368861e-01f);
    p = fmaf(p, s, 1.999262442e-01f);
    p = fmaf(p, s, -3.333307356e-01f);
    p = fmaf(p, s, 9.999999848e-01f);
    float r = t * p;
    r = ay > ax ? 1.57079632679490f - r : r;
    r = __builtin_signbit(x) ? 3.14159265358979f - r : r;
    return copysignf(r, y);
}

// Sums a[0..3] over the workgroup's waves into sums[0..3] (one atomic per wave
// and slot).  Every thread of the block must call it.
__device__ __forceinline__ void accumulate4(double a0, double a1, double a2, double a3, double* sums) {
    double v[4] = {a0, a1, a2, a3};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_down(v[i], off, 64);
    }
    if ((threadIdx.x & 63) == 0 && sums) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (v[i] != 0.0) atomicAdd(sums + i, v[i]);
    }
}

// accumulate4 into a problem's fixed-order partials when it has them:
// wave w of the problem's row tile `slot` writes its four sums to
// parts[(slot * waves + w) * 4 + i].  Every thread of the block must call it.
__device__ __forceinline__ void accumulate4_tile(double a0, double a1, double a2, double a3, double* sums,
                                                 double* parts, int slot) {
    if (!parts) {
        accumulate4(a0, a1, a2, a3, sums);
        return;
    }
    double v[4] = {a0, a1, a2, a3};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_down(v[i], off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        const int waves = ((int)blockDim.x + 63) / 64, w = (int)threadIdx.x / 64;
#pragma unroll
        for (int i = 0; i < 4; ++i) parts[((size_t)slot * waves + w) * 4 + i] = v[i];
    }
}

// First-pass load of input sample (ix, iy) with the problem's pre-op.
// PK: the pre-op when known at compile time (PRE_NONE / PRE_MASK /
// PRE_TMUL), else -1 and the descriptor's op decides.
template <typename R, int PK = -1>
__device__ __forceinline__ cpx<R> load_pre(const ProbDev& P, int ix, int iy, bool& bad, double& acc) {
    const size_t k = (size_t)iy * P.nx + ix;
    const OpsDev& o = P.ops;
    const int pre = PK >= 0 ? PK : o.pre;
    if (pre == PRE_POLAR) {  // Complex::from_polar(A, phi) masked by the stop (gs.rs:484-492)
        if (o.open && !gp(o.open)[k]) return {R(0), R(0)};
        const R a = gp(static_cast<const R*>(o.amp))[k];
        const R ph = gp(static_cast<const R*>(o.phase))[k];
        const cpx<R> e = unit_phase<R>((double)ph);
        return {a * e.re, a * e.im};
    }
    cpx<R> v = gp(static_cast<const cpx<R>*>(P.in))[k];
    // a mask is Element::apply before the engine: the engine validates the
    // masked field, in which samples outside the stop are zero
    if (pre != PRE_MASK && pre != PRE_TMUL) bad |= !finite_c(v);
    if (pre == PRE_MASK) {
        v = apply_mask(P.mask, ix, iy, v);
        bad |= !finite_c(v);
    } else if (pre == PRE_FRAUN) {
        v = cmul(v, fraun_pre<R>(P.eng, ix, iy, P.nx, P.ny));
    } else if (pre == PRE_TMUL) {
        cpx<R> t;
        t = gp(static_cast<const cpx<R>*>(o.trans))[k];
        v = tmul(v, t);
        bad |= !finite_c(v);
    } else if (pre == PRE_SEED_MSE) {  // w = (dM/dI) U, dM/dI = 2 (I - T)/n (merit.rs:143-165)
        const double re = v.re, im = v.im;
        const double d = (re * re + im * im) - gp(o.seed_target)[k];
        acc += d * d;
        const double g = 2.0 * d / o.n_frame;
        v = {(R)(g * re), (R)(g * im)};
    }
    return v;
}

// Last-pass store of output sample (ix, iy) with the problem's post-op.
template <typename R>
__device__ __forceinline__ void store_post(const ProbDev& P, int ix, int iy, cpx<R> o, double& acc0, double& acc1,
                                           double& acc2) {
    const size_t k = (size_t)iy * P.nx + ix;
    cpx<R>* out = static_cast<cpx<R>*>(P.out);
    const OpsDev& op = P.ops;
    switch (op.post) {
        case POST_GS_TARGET: {  // gs.rs:494-551
            const double re = o.re, im = o.im;
            const double a_t = hypot(re, im);
            const double inten = a_t * a_t;
            const double e = a_t - op.tamp[k];
            acc0 += e * e;
            const double dm = inten - op.merit_target[k];
            acc1 += dm * dm;
            if (op.err_out) op.err_out[k] = e;
            const double a = (1.0 - op.beta_f) * a_t + op.beta_f * op.tamp[k];
            const double m = hypot(re, im);
            cpx<R> r;
            if (m > 0.0) r = {(R)(re * (a / m)), (R)(im * (a / m))};
            else r = {(R)a, R(0)};
            out[k] = r;
            return;
        }
        case POST_GS_PHASE: {  // gs.rs:557-567
            R* ph = static_cast<R*>(op.phase_io);
            if constexpr (sizeof(R) == 4) {
                // complex64 form: the same update in FP32
                const float p = ph[k];
                const float dphi = wrap_pi_f(atan2_f(o.im, o.re) - p);
                ph[k] = encode_bounds_f(p + (float)op.beta_p * dphi, (float)op.lo, (float)op.hi);
                acc2 += (double)dphi * dphi;
                if (op.dphi_out) op.dphi_out[k] = dphi;
            } else {
                const double p = ph[k];
                const double dphi = wrap_pi(atan2((double)o.im, (double)o.re) - p);
                ph[k] = (R)encode_bounds(p + op.beta_p * dphi, op.lo, op.hi);
                acc2 += dphi * dphi;
                if (op.dphi_out) op.dphi_out[k] = dphi;
            }
            if (out) out[k] = o;
            return;
        }
        case POST_ADJ_GRAD: {  // adjoint.rs:222-236
            const cpx<R> t = apply_mask(P.mask, ix, iy, cpx<R>{R(1), R(0)});
            const cpx<R> v = apply_mask(P.mask, ix, iy, static_cast<const cpx<R>*>(op.ckpt)[k]);
            const double mre = o.re, mim = o.im;
            // -2 Im(conj(mu) V), conj(mu) = (mre, -mim)
            const double gim = mre * (double)v.im + (-mim) * (double)v.re;
            op.grad[k] = 0.0 + -2.0 * gim;  // the plane's one term (no clearing pass)
            // co *= conj(t): re = mu.re*t.re - mu.im*(-t.im); im = mu.im*t.re + mu.re*(-t.im)
            const R tr = t.re, ti = -t.im;
            cpx<R> c;
            c.re = o.re * tr - o.im * ti;
            c.im = o.im * tr + o.re * ti;
            out[k] = c;
            return;
        }
        default:
            out[k] = o;
    }
}

/* ------------------------------------------------------------------ */
/* Tile lookup.                                                         */
/* ------------------------------------------------------------------ */

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on
// XCD b % 8), each XCD behind its own L2.  Renumbering them so that the
// workgroups of one XCD take consecutive tiles keeps neighbouring rows --
// which share cache lines of the blocked intermediate -- in one L2.
__device__ __forceinline__ int xcd_block() {
    const int g = (int)gridDim.x, b = (int)blockIdx.x;
    if (g % 8) return b;
    return (b % 8) * (g / 8) + b / 8;
}
enum PassKind : int { PASS_ROW = 0, PASS_COL = 1, PASS_KGEN = 2 };

template <int PASS>
__device__ __forceinline__ int find_problem(const ProbDev* __restrict__ probs, int nprob, int tile) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int b = PASS == PASS_ROW ? probs[mid].base_row : (PASS == PASS_COL ? probs[mid].base_col
                                                                                 : probs[mid].base_kgen);
        if (b <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace ocd
