// Device-side FFT building blocks for gfx950 (CDNA4, 64-wide wavefronts).
//
// A transform of length n lives in LDS; T threads cooperate on it.  Stages are
// self-sorting Stockham decimation-in-time steps: stage s with radix R and
// Ns = product of the earlier radices reads the R inputs of butterfly j at
// j + r*(n/R), twiddles them by w_n^{(j mod Ns)*r*n/(Ns*R)}, runs an R-point
// DFT in registers and writes output r at (j - j mod Ns)*R + j mod Ns + r*Ns.
// Fast stages are in place (all inputs staged in registers, barrier, then
// stores); a generic ping-pong stage handles prime factors above 7.
//
// Sign convention: forward kernel e^{-2 pi i kn/N}, unnormalised; inverse uses
// the conjugate twiddles (src/propagation/fft.rs:1-27).
#pragma once

#include <hip/hip_runtime.h>

namespace ocd {

// Aligned to its full size so that LDS and global accesses are single
// 64/128-bit operations.
template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R re, im;
};
using c64 = cpx<float>;
using c128 = cpx<double>;

// Global-memory view of a generic pointer.  Every buffer the kernels address
// through a descriptor lives in device (global) memory; accessing it through
// this view emits global_* instead of flat_* instructions, which do not tie
// up the LDS counter.
// (The host pass of a .hip file only type-checks device code; it sees the
// plain type.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gmem = __attribute__((address_space(1))) T;
#else
template <typename T>
using gmem = T;
#endif
template <typename T>
__device__ __forceinline__ gmem<T>* gp(T* p) {
    return (gmem<T>*)p;
}

// Element idx of a device array addressed with a 32-bit byte offset from a
// wave-uniform base (one offset register per access instead of a 64-bit
// address pair).  For arrays below 4 GiB.
template <typename T>
__device__ __forceinline__ gmem<T>& at32(T* base, unsigned idx) {
    const gmem<const char>* b = (const gmem<const char>*)gp(base);
    return *(gmem<T>*)(b + idx * (unsigned)sizeof(T));
}

// Fused multiply-add.  The transforms and field products use it explicitly
// (the files are built without automatic contraction so that f64 phase
// arguments round exactly as written).
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
    return {fmadd(a.re, b.re, -(a.im * b.im)), fmadd(a.re, b.im, a.im * b.re)};
}
// Product with a phase-mask transmission t: t is e^{i phi} (|t| = 1) inside
// the stop and exactly 0 outside, where the apply sets the sample to 0
// whatever it held (a non-finite sample there included).
template <typename R>
__device__ __forceinline__ cpx<R> tmul(cpx<R> v, cpx<R> t) {
    const bool stop = t.re == R(0) && t.im == R(0);
    const cpx<R> o = cmul(v, t);
    return {stop ? R(0) : o.re, stop ? R(0) : o.im};
}
template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename R>
__device__ __forceinline__ cpx<R> csub(cpx<R> a, cpx<R> b) {
    return {a.re - b.re, a.im - b.im};
}
// multiply by -i*DIR... : rot<DIR>(a) = a * (DIR * i)   (DIR = -1 forward)
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> rot(cpx<R> a) {
    // a * (i*DIR): (re, im) * (0, DIR) = (-im*DIR, re*DIR)
    if (DIR < 0) return {a.im, -a.re};
    return {-a.im, a.re};
}

// LDS index with a one-element pad every 32 elements (spreads strided accesses
// of the early stages over the banks).
__device__ __forceinline__ int lidx(int i) { return i + (i >> 5); }
__host__ __device__ inline int lds_stride(int n) { return n + (n >> 5) + 1; }

/* ------------------------------------------------------------------ */
/* Small DFTs in registers.                                             */
/* ------------------------------------------------------------------ */
template <int DIR, typename R>
__device__ __forceinline__ void dft2(cpx<R>& a, cpx<R>& b) {
    cpx<R> t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft4(cpx<R>& x0, cpx<R>& x1, cpx<R>& x2, cpx<R>& x3) {
    cpx<R> a02 = cadd(x0, x2), s02 = csub(x0, x2);
    cpx<R> a13 = cadd(x1, x3), s13 = csub(x1, x3);
    x0 = cadd(a02, a13);
    x2 = csub(a02, a13);
    cpx<R> r = rot<DIR>(s13);  // s13 * (DIR i)
    x1 = cadd(s02, r);
    x3 = csub(s02, r);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft3(cpx<R>* v) {
    const R c = R(-0.5);
    const R s = R(0.86602540378443864676372317075294);  // sin(2 pi/3)
    cpx<R> b = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    cpx<R> t = {v[0].re + c * b.re, v[0].im + c * b.im};
    v[0] = cadd(v[0], b);
    // forward: y1 = t - i s d, y2 = t + i s d
    cpx<R> sd = {s * d.re, s * d.im};
    cpx<R> r = rot<DIR>(sd);  // DIR * i * s d
    v[1] = cadd(t, r);
    v[2] = csub(t, r);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft5(cpx<R>* v) {
    const R c1 = R(0.30901699437494742410229341718282);   // cos(2pi/5)
    const R c2 = R(-0.80901699437494742410229341718282);  // cos(4pi/5)
    const R s1 = R(0.95105651629515357211643933337938);   // sin(2pi/5)
    const R s2 = R(0.58778525229247312916870595463907);   // sin(4pi/5)
    cpx<R> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    cpx<R> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    cpx<R> x0 = v[0];
    v[0] = {x0.re + a1.re + a2.re, x0.im + a1.im + a2.im};
    cpx<R> t1 = {x0.re + c1 * a1.re + c2 * a2.re, x0.im + c1 * a1.im + c2 * a2.im};
    cpx<R> t2 = {x0.re + c2 * a1.re + c1 * a2.re, x0.im + c2 * a1.im + c1 * a2.im};
    cpx<R> u1 = {s1 * b1.re + s2 * b2.re, s1 * b1.im + s2 * b2.im};
    cpx<R> u2 = {s2 * b1.re - s1 * b2.re, s2 * b1.im - s1 * b2.im};
    cpx<R> r1 = rot<DIR>(u1), r2 = rot<DIR>(u2);
    v[1] = cadd(t1, r1);
    v[4] = csub(t1, r1);
    v[2] = cadd(t2, r2);
    v[3] = csub(t2, r2);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft7(cpx<R>* v) {
    const R c[3] = {R(0.62348980185873353052500488400424), R(-0.22252093395631440428890256449680),
                    R(-0.90096886790241912623610231950745)};
    const R s[3] = {R(0.78183148246802980870844452667406), R(0.97492791218182360701813168299393),
                    R(0.43388373911755812047576833284836)};
    cpx<R> a[3], b[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        a[m] = cadd(v[1 + m], v[6 - m]);
        b[m] = csub(v[1 + m], v[6 - m]);
    }
    cpx<R> x0 = v[0];
    v[0] = {x0.re + a[0].re + a[1].re + a[2].re, x0.im + a[0].im + a[1].im + a[2].im};
#pragma unroll
    for (int q = 1; q <= 3; ++q) {
        cpx<R> t = x0, u = {R(0), R(0)};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int e = ((m + 1) * q) % 7;  // cos(2 pi e/7), sin(2 pi e/7)
            const int ei = e <= 3 ? e : 7 - e;
            const R ce = c[ei - 1];
            const R se = e <= 3 ? s[ei - 1] : -s[ei - 1];
            t.re += ce * a[m].re;
            t.im += ce * a[m].im;
            u.re += se * b[m].re;
            u.im += se * b[m].im;
        }
        cpx<R> r = rot<DIR>(u);
        v[q] = cadd(t, r);
        v[7 - q] = csub(t, r);
    }
}

// Radix 8 as 2 x 4: X[q + 4s] = DFT2_r( W8^{rq} * DFT4_m(v[r + 2m]) ).
// W8^1 = h (1 + D i) and W8^3 = h (-1 + D i) (h = cos(pi/4), D = DIR): the
// products are formed unscaled and h is applied by the FMAs of the last
// layer.
template <int DIR, typename R>
__device__ __forceinline__ void dft8(cpx<R>* v) {
    constexpr R D = DIR < 0 ? R(-1) : R(1);
    const R h = R(0.70710678118654752440084436210485);
    dft4<DIR>(v[0], v[2], v[4], v[6]);
    dft4<DIR>(v[1], v[3], v[5], v[7]);
    const cpx<R> t1 = {v[3].re - D * v[3].im, v[3].im + D * v[3].re};     // v3 (1 + D i)
    const cpx<R> t3 = {-v[7].re - D * v[7].im, D * v[7].re - v[7].im};    // v7 (-1 + D i)
    const cpx<R> r2 = rot<DIR>(v[5]);
    cpx<R> o[8];
    o[0] = cadd(v[0], v[1]);
    o[4] = csub(v[0], v[1]);
    o[1] = {fmadd(h, t1.re, v[2].re), fmadd(h, t1.im, v[2].im)};
    o[5] = {fmadd(-h, t1.re, v[2].re), fmadd(-h, t1.im, v[2].im)};
    o[2] = cadd(v[4], r2);
    o[6] = csub(v[4], r2);
    o[3] = {fmadd(h, t3.re, v[6].re), fmadd(h, t3.im, v[6].im)};
    o[7] = {fmadd(-h, t3.re, v[6].re), fmadd(-h, t3.im, v[6].im)};
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = o[i];
}

// The last layer of a row q of dft16p: X0 = a02 + s u, X2 = a02 - s u,
// X1 = s02 + s (D i) w, X3 = s02 - s (D i) w, the scale s by FMA.
template <int DIR, typename R>
__device__ __forceinline__ void dft16_row_out(cpx<R>* o, cpx<R> a02, cpx<R> s02, cpx<R> u, cpx<R> w, R s) {
    const cpx<R> rw = rot<DIR>(w);
    o[0] = {fmadd(s, u.re, a02.re), fmadd(s, u.im, a02.im)};
    o[1] = {fmadd(s, rw.re, s02.re), fmadd(s, rw.im, s02.im)};
    o[2] = {fmadd(-s, u.re, a02.re), fmadd(-s, u.im, a02.im)};
    o[3] = {fmadd(-s, rw.re, s02.re), fmadd(-s, rw.im, s02.im)};
}

// Radix 16 as 4 x 4: X[q + 4s] = DFT4_r( W16^{rq} * DFT4_m(v[r + 4m]) ).
// The twiddles of the second layer are written as a real scale times a
// factor with a unit part (c = cos(pi/8), t = tan(pi/8), h = cos(pi/4),
// D = DIR): W16^1 = c (1 + D i t), W16^3 = c (t + D i), W16^9 = -W16^1,
// W16^2 = h (1 + D i), W16^4 = D i, W16^6 = h (-1 + D i).  A row's two
// entries of one scale are combined unscaled and the scale is applied by the
// FMAs that form the row's outputs: 144 real operations per radix-16
// butterfly instead of 160, the same transform to rounding.
template <int DIR, typename R>
__device__ __forceinline__ void dft16p(cpx<R>* v) {
    constexpr R D = DIR < 0 ? R(-1) : R(1);
    const R c1 = R(0.92387953251128675612818318939679);  // cos(pi/8)
    const R tn = R(0.41421356237309504880168872420970);  // tan(pi/8)
    const R h = R(0.70710678118654752440084436210485);
#pragma unroll
    for (int r = 0; r < 4; ++r) dft4<DIR>(v[r], v[r + 4], v[r + 8], v[r + 12]);
    // A[r][q] now sits at v[r + 4q]; row q takes A[r][q] W16^{rq}, r = 0..3
    dft4<DIR>(v[0], v[1], v[2], v[3]);
    {  // q = 1: W^1 (scale c), W^2 (h), W^3 (c)
        const cpx<R> y1 = {fmadd(-D * tn, v[5].im, v[5].re), fmadd(D * tn, v[5].re, v[5].im)};
        const cpx<R> y3 = {fmadd(tn, v[7].re, -D * v[7].im), fmadd(tn, v[7].im, D * v[7].re)};
        const cpx<R> t2 = {v[6].re - D * v[6].im, v[6].im + D * v[6].re};
        const cpx<R> a02 = {fmadd(h, t2.re, v[4].re), fmadd(h, t2.im, v[4].im)};
        const cpx<R> s02 = {fmadd(-h, t2.re, v[4].re), fmadd(-h, t2.im, v[4].im)};
        dft16_row_out<DIR>(v + 4, a02, s02, cadd(y1, y3), csub(y1, y3), c1);
    }
    {  // q = 2: W^2 (h), W^4, W^6 (h)
        const cpx<R> t1 = {v[9].re - D * v[9].im, v[9].im + D * v[9].re};
        const cpx<R> t3 = {-v[11].re - D * v[11].im, D * v[11].re - v[11].im};
        const cpx<R> x2 = rot<DIR>(v[10]);
        const cpx<R> a02 = cadd(v[8], x2), s02 = csub(v[8], x2);
        dft16_row_out<DIR>(v + 8, a02, s02, cadd(t1, t3), csub(t1, t3), h);
    }
    {  // q = 3: W^3 (c), W^6 (h), W^9 (c)
        const cpx<R> y1 = {fmadd(tn, v[13].re, -D * v[13].im), fmadd(tn, v[13].im, D * v[13].re)};
        const cpx<R> y3 = {fmadd(D * tn, v[15].im, -v[15].re), fmadd(-D * tn, v[15].re, -v[15].im)};
        const cpx<R> t2 = {-v[14].re - D * v[14].im, D * v[14].re - v[14].im};
        const cpx<R> a02 = {fmadd(h, t2.re, v[12].re), fmadd(h, t2.im, v[12].im)};
        const cpx<R> s02 = {fmadd(-h, t2.re, v[12].re), fmadd(-h, t2.im, v[12].im)};
        dft16_row_out<DIR>(v + 12, a02, s02, cadd(y1, y3), csub(y1, y3), c1);
    }
    // X[q + 4s] now sits in v[4q + s]
}

// In-register transposed-output radix 16: X[q + 4s] is left in v[4q + s]
// (see dft16p_pos).  dft16 below restores natural order.
template <int DIR, typename R>
__device__ __forceinline__ void dft16(cpx<R>* v) {
    dft16p<DIR>(v);
    cpx<R> t;
    // transpose the 4x4 index grid in place
    t = v[1]; v[1] = v[4]; v[4] = t;
    t = v[2]; v[2] = v[8]; v[8] = t;
    t = v[3]; v[3] = v[12]; v[12] = t;
    t = v[6]; v[6] = v[9]; v[9] = t;
    t = v[7]; v[7] = v[13]; v[13] = t;
    t = v[11]; v[11] = v[14]; v[14] = t;
}

template <int RAD, int DIR, typename R>
__device__ __forceinline__ void dft(cpx<R>* v) {
    if constexpr (RAD == 2) dft2<DIR>(v[0], v[1]);
    else if constexpr (RAD == 3) dft3<DIR>(v);
    else if constexpr (RAD == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
    else if constexpr (RAD == 5) dft5<DIR>(v);
    else if constexpr (RAD == 7) dft7<DIR>(v);
    else if constexpr (RAD == 8) dft8<DIR>(v);
    else if constexpr (RAD == 16) dft16<DIR>(v);
}

/* ------------------------------------------------------------------ */
/* Stages.                                                              */
/* ------------------------------------------------------------------ */
constexpr int kMaxStages = 12;

struct FftPlanDev {
    int n;
    int nstages;
    int generic;  // 1: ping-pong generic stages (prime factors > 7)
    int threads;  // T: threads per transform
    int radix[kMaxStages];
    const void* tw;  // n twiddles exp(-2 pi i k/n), element type = transform type
};

template <int DIR, typename R>
__device__ __forceinline__ cpx<R> tw_at(const cpx<R>* __restrict__ tw, int idx) {
    cpx<R> w = gp(tw)[idx];
    if (DIR > 0) w.im = -w.im;
    return w;
}

// In-place register-staged radix-RAD stage on one transform stored at `buf`
// (padded LDS indexing).  t in [0, T) for working threads; idle threads pass
// t >= n and still reach both barriers.
template <int RAD, int DIR, typename R>
__device__ __forceinline__ void stage_inplace(cpx<R>* __restrict__ buf, int n, int Ns,
                                              const cpx<R>* __restrict__ tw, int t, int T) {
    constexpr int NB = (16 + RAD - 1) / RAD;
    const int nb = n / RAD;
    const int tws = n / (Ns * RAD);
    cpx<R> v[NB][RAD];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = t + b * T;
        if (j < nb) {
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[b][r] = buf[lidx(j + r * nb)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = t + b * T;
        if (j < nb) {
            const int k = j % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < RAD; ++r) v[b][r] = cmul(v[b][r], tw_at<DIR>(tw, k * r * tws));
            }
            dft<RAD, DIR>(v[b]);
            const int base = (j - k) * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; ++r) buf[lidx(base + r * Ns)] = v[b][r];
        }
    }
    __syncthreads();
}

// Generic ping-pong stage for any radix RAD (runtime): O(RAD^2) per butterfly.
template <int DIR, typename R>
__device__ void stage_generic(const cpx<R>* __restrict__ src, cpx<R>* __restrict__ dst, int n, int Ns, int RAD,
                              const cpx<R>* __restrict__ tw, int t, int T) {
    const int nb = n / RAD;
    const int tws = n / (Ns * RAD);
    const int wstep = n / RAD;  // w_RAD = w_n^{n/RAD}
    for (int j = t; j < nb; j += T) {
        const int k = j % Ns;
        const int base = (j - k) * RAD + k;
        for (int q = 0; q < RAD; ++q) {
            cpx<R> acc = {R(0), R(0)};
            for (int r = 0; r < RAD; ++r) {
                cpx<R> x = src[lidx(j + r * nb)];
                if (Ns > 1 && r > 0) x = cmul(x, tw_at<DIR>(tw, k * r * tws));
                const int e = (r * q) % RAD;
                if (e) x = cmul(x, tw_at<DIR>(tw, e * wstep));
                acc = cadd(acc, x);
            }
            dst[lidx(base + q * Ns)] = acc;
        }
    }
    __syncthreads();
}

// Runs all stages of `pl` (in `reverse` radix order when asked) on the
// transform at `buf`; for generic plans `alt` is the ping-pong partner and the
// result is copied back into `buf`.
template <int DIR, typename R>
__device__ __forceinline__ void run_fft(cpx<R>* __restrict__ buf, cpx<R>* __restrict__ alt, const FftPlanDev& pl,
                                        int t, bool reverse) {
    const cpx<R>* tw = static_cast<const cpx<R>*>(pl.tw);
    const int n = pl.n, T = pl.threads;
    int Ns = 1;
    if (!pl.generic) {
        for (int s = 0; s < pl.nstages; ++s) {
            const int rad = pl.radix[reverse ? pl.nstages - 1 - s : s];
            switch (rad) {
                case 16: stage_inplace<16, DIR>(buf, n, Ns, tw, t, T); break;
                case 8: stage_inplace<8, DIR>(buf, n, Ns, tw, t, T); break;
                case 4: stage_inplace<4, DIR>(buf, n, Ns, tw, t, T); break;
                case 2: stage_inplace<2, DIR>(buf, n, Ns, tw, t, T); break;
                case 3: stage_inplace<3, DIR>(buf, n, Ns, tw, t, T); break;
                case 5: stage_inplace<5, DIR>(buf, n, Ns, tw, t, T); break;
                case 7: stage_inplace<7, DIR>(buf, n, Ns, tw, t, T); break;
                default: break;
            }
            Ns *= rad;
        }
        return;
    }
    cpx<R>* src = buf;
    cpx<R>* dst = alt;
    for (int s = 0; s < pl.nstages; ++s) {
        const int rad = pl.radix[reverse ? pl.nstages - 1 - s : s];
        stage_generic<DIR>(src, dst, n, Ns, rad, tw, t, T);
        Ns *= rad;
        cpx<R>* tmp = src;
        src = dst;
        dst = tmp;
    }
    if (src != buf) {
        for (int i = t; i < n; i += T) buf[lidx(i)] = src[lidx(i)];
        __syncthreads();
    }
}

}  // namespace ocd
