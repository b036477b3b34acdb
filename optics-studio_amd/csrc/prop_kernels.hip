// Propagation passes for gfx950.
//
// A transfer-function propagation of an N_y x N_x field on a P_y x P_x padded
// grid (src/propagation/asm.rs:481-535) runs as three passes over HBM:
//   row_fwd : the N_y non-zero rows: embed (+ fused phase mask + finiteness
//             flag), length-P_x FFT, store to the blocked intermediate.
//   col     : blocks of B columns: zero-pad to P_y, FFT, multiply by the
//             transfer function (or the RS kernel spectrum), inverse FFT, keep
//             the N_y rows of the physical window, store back in place.
//   row_inv : inverse length-P_x FFT, crop the N_x window, scale 1/(P_x P_y).
// Rows outside the window are never materialised.  Rayleigh-Sommerfeld adds a
// kernel-generation pass (sample h in DFT order, row FFT) whose column FFT is
// fused into `col`.
//
// All kernels are persistent: a launch covers the tiles of every problem of a
// batch (uneven sizes allowed); each workgroup loops over tiles.
#include <hip/hip_runtime.h>

#include "prop_kernels.hpp"

#include "pass_ops.hpp"

namespace ocd {

__device__ __forceinline__ int fft_lds_stride(const FftPlanDev& pl) {
    return lds_stride(pl.n) * (pl.generic ? 2 : 1);
}

/* ------------------------------------------------------------------ */
/* Pass 1: row forward FFT with embed + mask.                           */
/* ------------------------------------------------------------------ */
template <typename R>
__global__ void k_row_fwd(const ProbDev* __restrict__ probs, int nprob, int total) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x, bdim = blockDim.x;
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_ROW>(probs, nprob, tile)];
        const int iy0 = (tile - P.base_row) * P.rows_cnt;
        const int cnt = P.rows_cnt, px = P.px, nx = P.nx, ny = P.ny, ox = P.ox;
        const int stride = fft_lds_stride(P.prow);
        bool bad = false;
        double acc = 0.0;
        for (int i = tid; i < cnt * px; i += bdim) {
            const int u = i / px, x = i - u * px, iy = iy0 + u;
            cpx<R> v = {R(0), R(0)};
            if (iy < ny && x >= ox && x < ox + nx) v = load_pre<R>(P, x - ox, iy, bad, acc);
            lds[u * stride + lidx(x)] = v;
        }
        if (bad) atomicOr(P.flag, 1);
        if (P.ops.pre == PRE_SEED_MSE) accumulate4(acc, 0.0, 0.0, 0.0, P.ops.sums ? P.ops.sums + 3 : nullptr);
        __syncthreads();
        const int T = P.prow.threads;
        const int u = tid / T;
        const int uu = u < cnt ? u : 0;
        const int t = u < cnt ? tid - u * T : px;
        run_fft<-1>(lds + uu * stride, lds + uu * stride + lds_stride(px), P.prow, t, false);
        // store to blocked [px/B][ny][B]
        const int B = P.B;
        cpx<R>* __restrict__ work = static_cast<cpx<R>*>(P.work);
        for (int i = tid; i < cnt * px; i += bdim) {
            const int b = i % B, rest = i / B;
            const int u2 = rest % cnt, q = rest / cnt;
            const int iy = iy0 + u2;
            if (iy < ny) work[((size_t)q * ny + iy) * B + b] = lds[u2 * stride + lidx(q * B + b)];
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------ */
/* Pass 2: column FFT -> transfer function -> inverse column FFT.       */
/* ------------------------------------------------------------------ */
template <typename R>
__global__ void k_col(const ProbDev* __restrict__ probs, int nprob, int total) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x, bdim = blockDim.x;
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_COL>(probs, nprob, tile)];
        const int q = tile - P.base_col;
        const int B = P.B, py = P.py, ny = P.ny, oy = P.oy;
        const int stride = fft_lds_stride(P.pcol);
        const bool rs = conv_engine(P.eng.kind);
        const int ntr = rs ? 2 * B : B;  // RS: B field columns + B kernel columns
        cpx<R>* __restrict__ work = static_cast<cpx<R>*>(P.work);
        for (int i = tid; i < B * py; i += bdim) {
            const int b = i % B, y = i / B;
            cpx<R> v = {R(0), R(0)};
            if (y >= oy && y < oy + ny) v = work[((size_t)q * ny + (y - oy)) * B + b];
            lds[b * stride + lidx(y)] = v;
        }
        if (rs) {
            const cpx<R>* __restrict__ kb = static_cast<const cpx<R>*>(P.kbuf);
            for (int i = tid; i < B * py; i += bdim) {
                const int b = i % B, y = i / B;
                lds[(B + b) * stride + lidx(y)] = kb[((size_t)q * py + y) * B + b];
            }
        }
        __syncthreads();
        const int T = P.pcol.threads;
        const int u = tid / T;
        const int uu = u < ntr ? u : 0;
        const int t = u < ntr ? tid - u * T : py;
        run_fft<-1>(lds + uu * stride, lds + uu * stride + lds_stride(py), P.pcol, t, false);
        const int ek = P.eng.kind;
        if (ek == ENG_FRAUNHOFER) {  // forward only: each bin to its far-field sample
            __syncthreads();
            cpx<R>* out = static_cast<cpx<R>*>(P.out);
            bool bad = false;
            for (int i = tid; i < B * py; i += bdim) {
                const int b = i % B, my = i / B;
                int ox_, oy_;
                const cpx<R> o = fraun_post<R>(P.eng, q * B + b, my, P.nx, P.ny, lds[b * stride + lidx(my)], &ox_,
                                               &oy_);
                bad |= !finite_c(o);
                out[(size_t)oy_ * P.nx + ox_] = o;
            }
            if (bad) atomicOr(P.flag, 2);
            __syncthreads();
            continue;
        }
        if (ek == ENG_MEASURE) {  // |F|^2 statistics of the spectrum; no inverse, no store
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            if (u < B) {
                const R* tb = static_cast<const R*>(P.ops.tamp_blk);
                const cpx<R>* col = lds + u * stride;
                for (int ky = t; ky < py; ky += T) {
                    const cpx<R> v = col[lidx(ky)];
                    const double in2 = (double)v.re * v.re + (double)v.im * v.im;
                    const double ta = tb[((size_t)q * py + ky) * B + u];
                    s0 += in2;
                    s1 += in2 * in2;
                    s2 += in2 * (ta * ta);
                }
            }
            accumulate4(s0, s1, s2, 0.0, P.ops.sums);
            __syncthreads();
            continue;
        }
        if (u < B) {
            const int kx = q * B + u;
            cpx<R>* col = lds + u * stride;
            if (rs) {
                const cpx<R>* kc = lds + (B + u) * stride;
                for (int ky = t; ky < py; ky += T) col[lidx(ky)] = cmul(col[lidx(ky)], kc[lidx(ky)]);
            } else if (ek == ENG_GS_PROJECT) {  // Fourier-plane projection (gs.rs:539-551 with beta = 1)
                const R* tb = static_cast<const R*>(P.ops.tamp_blk);
                for (int ky = t; ky < py; ky += T) {
                    const cpx<R> v = col[lidx(ky)];
                    const R a = tb[((size_t)q * py + ky) * B + u];
                    const R m = hypot(v.re, v.im);
                    col[lidx(ky)] = m > R(0) ? cpx<R>{v.re * (a / m), v.im * (a / m)} : cpx<R>{a, R(0)};
                }
            } else if (ek != ENG_NONE) {
                for (int ky = t; ky < py; ky += T) col[lidx(ky)] = transfer<R>(P.eng, kx, ky, col[lidx(ky)]);
            }
        }
        __syncthreads();
        const int ti = u < B ? t : py;
        const int ui = u < B ? u : 0;
        run_fft<+1>(lds + ui * stride, lds + ui * stride + lds_stride(py), P.pcol, ti, true);
        for (int i = tid; i < B * ny; i += bdim) {
            const int b = i % B, yy = i / B;
            work[((size_t)q * ny + yy) * B + b] = lds[b * stride + lidx(yy + oy)];
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------ */
/* Pass 3: row inverse FFT, crop, scale.                                */
/* ------------------------------------------------------------------ */
template <typename R>
__global__ void k_row_inv(const ProbDev* __restrict__ probs, int nprob, int total) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x, bdim = blockDim.x;
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_ROW>(probs, nprob, tile)];
        const int iy0 = (tile - P.base_row) * P.rows_cnt;
        const int cnt = P.rows_cnt, px = P.px, nx = P.nx, ny = P.ny, ox = P.ox, B = P.B;
        const int stride = fft_lds_stride(P.prow);
        const cpx<R>* __restrict__ work = static_cast<const cpx<R>*>(P.work);
        for (int i = tid; i < cnt * px; i += bdim) {
            const int b = i % B, rest = i / B;
            const int u2 = rest % cnt, q = rest / cnt;
            const int iy = iy0 + u2;
            cpx<R> v = {R(0), R(0)};
            if (iy < ny) v = work[((size_t)q * ny + iy) * B + b];
            lds[u2 * stride + lidx(q * B + b)] = v;
        }
        __syncthreads();
        const int T = P.prow.threads;
        const int u = tid / T;
        const int uu = u < cnt ? u : 0;
        const int t = u < cnt ? tid - u * T : px;
        run_fft<+1>(lds + uu * stride, lds + uu * stride + lds_stride(px), P.prow, t, true);
        const R s = (R)P.scale;
        bool bad = false;
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        for (int i = tid; i < cnt * nx; i += bdim) {
            const int u2 = i / nx, x = i - u2 * nx, iy = iy0 + u2;
            if (iy < ny) {
                const cpx<R> v = lds[u2 * stride + lidx(x + ox)];
                const cpx<R> o = {v.re * s, v.im * s};
                bad |= !finite_c(o);
                store_post<R>(P, x, iy, o, acc0, acc1, acc2);
            }
        }
        if (bad) atomicOr(P.flag, 2);
        if (P.ops.post == POST_GS_TARGET || P.ops.post == POST_GS_PHASE)
            accumulate4_tile(acc0, acc1, acc2, 0.0, P.ops.sums, P.ops.parts, tile - P.base_row);
        __syncthreads();
    }
}

/* ------------------------------------------------------------------ */
/* RS kernel generation + row FFT (rayleigh_sommerfeld.rs:277-301, 424).*/
/* ------------------------------------------------------------------ */
template <typename R>
__global__ void k_rs_kgen(const ProbDev* __restrict__ probs, int nprob, int total) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x, bdim = blockDim.x;
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_KGEN>(probs, nprob, tile)];
        const int y0 = (tile - P.base_kgen) * P.rows_cnt;
        const int cnt = P.rows_cnt, px = P.px, py = P.py, B = P.B;
        const int stride = fft_lds_stride(P.prow);
        for (int i = tid; i < cnt * px; i += bdim) {
            const int u = i / px, x = i - u * px, y = y0 + u;
            cpx<R> v = {R(0), R(0)};
            if (y < py) v = conv_sample<R>(P.eng, x, y, px, py);
            lds[u * stride + lidx(x)] = v;
        }
        __syncthreads();
        const int T = P.prow.threads;
        const int u = tid / T;
        const int uu = u < cnt ? u : 0;
        const int t = u < cnt ? tid - u * T : px;
        run_fft<-1>(lds + uu * stride, lds + uu * stride + lds_stride(px), P.prow, t, false);
        cpx<R>* __restrict__ kb = static_cast<cpx<R>*>(P.kbuf);
        for (int i = tid; i < cnt * px; i += bdim) {
            const int b = i % B, rest = i / B;
            const int u2 = rest % cnt, q = rest / cnt;
            const int y = y0 + u2;
            if (y < py) kb[((size_t)q * py + y) * B + b] = lds[u2 * stride + lidx(q * B + b)];
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------ */
/* Elementwise helpers.                                                 */
/* ------------------------------------------------------------------ */
template <typename R>
__global__ void k_mask_apply(const cpx<R>* in, cpx<R>* out, MaskDev m, int nx, int ny, int* flag) {
    // `flag` (optional) receives bit 0 for a non-finite sample of the masked
    // field (what a propagation after the apply validates: samples outside
    // the stop are zero); out == nullptr only checks.
    const size_t n = (size_t)nx * ny;
    bool bad = false;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int iy = (int)(i / nx), ix = (int)(i - (size_t)iy * nx);
        const cpx<R> o = apply_mask(m, ix, iy, in[i]);
        bad |= !finite_c(o);
        if (out) out[i] = o;
    }
    if (bad && flag) atomicOr(flag, 1);
}

// Transmission of a phase mask over the field grid (the mask applied to 1).
template <typename R>
__global__ void k_mask_trans(cpx<R>* __restrict__ out, MaskDev m, int nx, int ny) {
    const size_t n = (size_t)nx * ny;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int iy = (int)(i / nx), ix = (int)(i - (size_t)iy * nx);
        out[i] = apply_mask(m, ix, iy, cpx<R>{R(1), R(0)});
    }
}

template <typename R>
__global__ void k_check_finite(const cpx<R>* __restrict__ in, size_t n, int* flag) {
    bool bad = false;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bad |= !finite_c(in[i]);
    if (bad) atomicOr(flag, 1);
}

// co = dmdi * u  (adjoint.rs:200-201)
template <typename R>
__global__ void k_seed(const cpx<R>* __restrict__ u, const double* __restrict__ dmdi, cpx<R>* __restrict__ out,
                       size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const cpx<R> v = u[i];
        const double g = dmdi[i];
        out[i] = {(R)(g * (double)v.re), (R)(g * (double)v.im)};
    }
}

// Element gradient at a plane with no propagation after it (z_after == 0):
// grad = 0 + -2 Im(conj(co) V) (the plane's one term: the reverse sweep
// writes each element plane once), co *= conj(t)   (adjoint.rs:222-236)
template <typename R>
__global__ void k_adj_elem(cpx<R>* __restrict__ co, const cpx<R>* __restrict__ ckpt, double* __restrict__ grad,
                           MaskDev m, int nx, int ny) {
    const size_t n = (size_t)nx * ny;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int iy = (int)(i / nx), ix = (int)(i - (size_t)iy * nx);
        const cpx<R> t = apply_mask(m, ix, iy, cpx<R>{R(1), R(0)});
        const cpx<R> v = apply_mask(m, ix, iy, ckpt[i]);
        const cpx<R> o = co[i];
        const double gim = (double)o.re * (double)v.im + (-(double)o.im) * (double)v.re;
        grad[i] = 0.0 + -2.0 * gim;
        const R tr = t.re, ti = -t.im;
        cpx<R> c;
        c.re = o.re * tr - o.im * ti;
        c.im = o.im * tr + o.re * ti;
        co[i] = c;
    }
}

// dst[((q*ny) + y)*B + b] = src[y*nx + q*B + b]: a row-major real plane in the
// blocked layout of the column pass.
template <typename T>
__global__ void k_block_reorder(const T* __restrict__ src, T* __restrict__ dst, int nx, int ny, int B) {
    const size_t n = (size_t)nx * ny;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / nx), x = (int)(i - (size_t)y * nx);
        const int q = x / B, b = x - q * B;
        dst[((size_t)q * ny + y) * B + b] = src[i];
    }
}

// out = (2 (|u|^2 - target)/n) u and sums[0] += (|u|^2 - target)^2
// (merit.rs:123-165 with adjoint.rs:200-201)
// Observation frames of Renderer::render (observe.rs:258-310): per frame
// f = band * n_scenes + scene and pixel k, channel 0 = |U_base|^2 and each
// enabled derivative (I+ - I-) / (2 step) from the frame's renders
// [base, x+, x-, y+, y-, l+, l-] (enabled ones, in that order).  Flags a
// non-finite base intensity (observe.rs:296-303).
template <typename R>
__global__ void k_render_obs(const cpx<R>* __restrict__ u, double* __restrict__ obs, size_t n, int frames, int nr,
                             int cx, int cy, int cl, double hs, double hl, int* flag) {
    const size_t total = n * (size_t)frames;
    bool bad = false;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t f = i / n, k = i - f * n;
        const cpx<R>* uf = u + f * (size_t)nr * n + k;
        auto inten = [&](int r) {
            const cpx<R> v = uf[(size_t)r * n];
            return (double)v.re * (double)v.re + (double)v.im * (double)v.im;
        };
        const double i0 = inten(0);
        bad |= !isfinite(i0);
        obs[i] = i0;
        int c = 1, r = 1;
        if (cx) {
            obs[((size_t)c * frames + f) * n + k] = (inten(r) - inten(r + 1)) / (2.0 * hs);
            ++c;
            r += 2;
        }
        if (cy) {
            obs[((size_t)c * frames + f) * n + k] = (inten(r) - inten(r + 1)) / (2.0 * hs);
            ++c;
            r += 2;
        }
        if (cl) obs[((size_t)c * frames + f) * n + k] = (inten(r) - inten(r + 1)) / (2.0 * hl);
    }
    if (bad) atomicOr(flag, 1);
}

// Per-render MseToTarget sums (merit.rs:124-137) in a fixed order:
// workgroup (chunk c, render r) reduces its strided share of render r's
// samples through LDS into part[r][c]; k_mse_final adds the chunks in order
// and divides by n.  The result does not depend on the launch.
constexpr int kRedChunks = 64;
template <typename R>
__global__ __launch_bounds__(256) void k_mse_part(const cpx<R>* __restrict__ u, const double* __restrict__ t,
                                                  size_t n, int r0, double* __restrict__ part) {
    const int r = r0 + (int)blockIdx.y, c = (int)blockIdx.x;
    const cpx<R>* ur = u + (size_t)r * n;
    double acc = 0.0;
    for (size_t k = (size_t)c * 256 + threadIdx.x; k < n; k += (size_t)kRedChunks * 256) {
        const double re = ur[k].re, im = ur[k].im;
        const double d = (re * re + im * im) - t[k];
        acc += d * d;
    }
    __shared__ double sm[256];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)r * kRedChunks + c] = sm[0];
}

__global__ void k_mse_final(const double* __restrict__ part, int nr, double nf, double* __restrict__ merit) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nr; r += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < kRedChunks; ++c) s += part[(size_t)r * kRedChunks + c];
        merit[r] = s / nf;
    }
}

// Adam (fd.rs:223-244).  k_sq_part: fixed-order partial sums of g^2;
// k_clip_scale: norm = sqrt(sum), scale = cn / norm when clipping applies
// (else 1), written to scale[0]; k_adam: the per-coordinate update and the
// clamp into [lo, hi] (solve.rs:74-78).
__global__ __launch_bounds__(256) void k_sq_part(const double* __restrict__ g, size_t n, double* __restrict__ part) {
    const int c = (int)blockIdx.x;
    double acc = 0.0;
    for (size_t k = (size_t)c * 256 + threadIdx.x; k < n; k += (size_t)kRedChunks * 256) acc += g[k] * g[k];
    __shared__ double sm[256];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[c] = sm[0];
}

__global__ void k_clip_scale(const double* __restrict__ part, int has_clip, double cn, double* __restrict__ scale) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int c = 0; c < kRedChunks; ++c) s += part[c];
    const double norm = sqrt(s);
    scale[0] = (has_clip && norm > cn && norm > 0.0) ? cn / norm : 1.0;
}

__global__ void k_adam(double* __restrict__ x, const double* __restrict__ g, double* __restrict__ m,
                       double* __restrict__ v, const double* __restrict__ lo, const double* __restrict__ hi, size_t n,
                       const double* __restrict__ scale, double lr, double b1, double b2, double eps, double bc1,
                       double bc2) {
    const double sc = scale[0];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double gi = g[i] * sc;
        const double mi = b1 * m[i] + (1.0 - b1) * gi;
        const double vi = b2 * v[i] + (1.0 - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const double mh = mi / bc1;
        const double vh = vi / bc2;
        const double denom = sqrt(vh) + eps;
        const double l = lo[i], h = hi[i];
        double xi = x[i];
        if (denom > 0.0) xi -= lr * (h - l) * mh / denom;
        xi = xi < l ? l : (xi > h ? h : xi);  // f64::clamp: NaN stays NaN
        x[i] = xi;
    }
}

// analysis::mtf (analysis.rs:133-150).  k_colsum: the column sums of each
// PSF (f64, in order over y), flagging non-finite samples; k_mtf_dft: the
// 1-D DFT of those sums at bins 0..nx/2, e^{-2 pi i k x / nx} from a table
// indexed by (k x) mod nx, then |X_k| / max(|X_0|, DBL_MIN) capped at 1.
__global__ void k_colsum(const double* __restrict__ psf, size_t nx, size_t ny, size_t count,
                         double* __restrict__ sums, int* flag) {
    const size_t total = nx * count;
    bool bad = false;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t c = i / nx, x = i - c * nx;
        const double* p = psf + c * nx * ny + x;
        double s = 0.0;
        for (size_t y = 0; y < ny; ++y) {
            const double v = p[y * nx];
            bad |= !isfinite(v);
            s += v;
        }
        sums[i] = s;
    }
    if (bad) atomicOr(flag, 1);
}

__global__ void k_mtf_dft(const double* __restrict__ sums, const double2* __restrict__ tw, size_t nx, size_t count,
                          double* __restrict__ values) {
    const size_t nb = nx / 2 + 1, total = nb * count;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t c = i / nb, k = i - c * nb;
        const double* p = sums + c * nx;
        double re = 0.0, im = 0.0, d0 = 0.0;
        size_t e = 0;  // (k x) mod nx
        for (size_t x = 0; x < nx; ++x) {
            const double2 w = tw[e];
            re = fma(p[x], w.x, re);
            im = fma(p[x], w.y, im);
            d0 += p[x];
            e += k;
            if (e >= nx) e -= nx;
        }
        const double dc = fmax(fabs(d0), 2.2250738585072014e-308);
        values[i] = fmin(hypot(re, im) / dc, 1.0);
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_seed_mse(const cpx<R>* __restrict__ u, const double* __restrict__ target,
                                                  cpx<R>* __restrict__ out, size_t n, double nf,
                                                  double* __restrict__ part) {
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double re = u[i].re, im = u[i].im;
        const double d = (re * re + im * im) - target[i];
        acc += d * d;
        const double g = 2.0 * d / nf;
        out[i] = {(R)(g * re), (R)(g * im)};
    }
    __shared__ double sm[256];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// MseToTarget over the (band, scene) frames of a render, [frame][n] fields
// with frame f = b ns + s (merit.rs:112-166): a selected frame's intensity
// joins the per-pixel sum Selector::frame takes (frames in order) and gets
// its own seed 2 (I_f - T)/n times its field (obs_gradient); an unselected
// frame gets a zero seed.  part[workgroup] = the workgroup's share of
// sum_k (frame_k - T_k)^2.
template <typename R>
__global__ __launch_bounds__(256) void k_seed_frames(const cpx<R>* __restrict__ u, const double* __restrict__ target,
                                                     cpx<R>* __restrict__ out, size_t n, int nb, int ns, int sel_b,
                                                     int sel_s, double nf, double* __restrict__ part) {
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double t = target[i];
        double frame = 0.0;
        for (int b = 0; b < nb; ++b)
            for (int s = 0; s < ns; ++s) {
                const size_t f = (size_t)b * ns + s;
                const cpx<R> v = u[f * n + i];
                cpx<R> o = {R(0), R(0)};
                if ((sel_b < 0 || sel_b == b) && (sel_s < 0 || sel_s == s)) {
                    const double re = v.re, im = v.im;
                    const double inten = re * re + im * im;
                    frame += inten;
                    const double g = 2.0 * (inten - t) / nf;
                    o = {(R)(g * re), (R)(g * im)};
                }
                out[f * n + i] = o;
            }
        const double d = frame - t;
        acc += d * d;
    }
    __shared__ double sm[256];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// out[i] = slab[0][i] + slab[1][i] + ... in frame order (the per-frame
// gradient planes added as adjoint.rs:310-332 does).
__global__ void k_frame_sum(const double* __restrict__ slab, double* __restrict__ out, size_t m, int frames) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
        double s = slab[i];
        for (int f = 1; f < frames; ++f) s += slab[(size_t)f * m + i];
        out[i] = s;
    }
}

// The device-side bookkeeping of gs::run (gs.rs:482-578) for restart
// r = blockIdx.x, between the passes of an iteration, so that iterations can
// be queued (and graph-captured) without reading anything back.  sum_gs_parts
// adds a restart's fixed-order partials (accumulate4_tile) in index order.
__device__ void sum_gs_parts(const double* __restrict__ parts, int nparts, double* out) {
    __shared__ double sm[4][256];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < nparts; j += 256) {
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] += parts[(size_t)j * 4 + i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) sm[i][threadIdx.x] = a[i];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int i = 0; i < 4; ++i) sm[i][threadIdx.x] += sm[i][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) out[i] = sm[i][0];
    }
}

// After the forward half (gs.rs:494-532): amplitude MSE and merit, the
// best-so-far, the history record; st.improved tells k_gs_keep_best to
// keep this iteration's phase.
__global__ __launch_bounds__(256) void k_gs_after_fwd(GsState* __restrict__ st, const double* __restrict__ parts,
                                                      const int* __restrict__ nparts, size_t pstride, double nf,
                                                      GsRec* __restrict__ hist, size_t hstride) {
    const int r = blockIdx.x;
    __shared__ double s[4];
    sum_gs_parts(parts + (size_t)r * pstride, nparts[r], s);
    if (threadIdx.x == 0) {
        GsState& S = st[r];
        const unsigned long long it = S.it;
        const double amp_mse = s[0] / nf;
        const double merit = s[1] / nf;
        const int improved = !S.have_best || merit < S.best_merit;
        if (improved) {
            S.have_best = 1;
            S.best_merit = merit;
            S.best_it = it;
        }
        S.improved = improved;
        GsRec h;
        h.iteration = it;
        h.merit = merit;
        h.best_merit = S.best_merit;
        h.amplitude_mse = amp_mse;
        h.target_error_rms = sqrt(amp_mse);
        h.phase_update_rms = 0.0;
        hist[(size_t)r * hstride + it] = h;
    }
}

// best = phase where the iteration improved (the phase of its forward half).
__global__ void k_gs_keep_best(const GsState* __restrict__ st, const double* __restrict__ phase,
                               double* __restrict__ best, size_t n, int restarts) {
    const size_t total = n * (size_t)restarts;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
        if (st[i / n].improved) best[i] = phase[i];
}

// After the backward half (gs.rs:557-567): the phase-update RMS of the
// record, then the next iteration.
__global__ __launch_bounds__(256) void k_gs_after_bwd(GsState* __restrict__ st, const double* __restrict__ parts,
                                                      const int* __restrict__ nparts, size_t pstride, double nf,
                                                      GsRec* __restrict__ hist, size_t hstride) {
    const int r = blockIdx.x;
    __shared__ double s[4];
    sum_gs_parts(parts + (size_t)r * pstride, nparts[r], s);
    if (threadIdx.x == 0) {
        GsState& S = st[r];
        hist[(size_t)r * hstride + S.it].phase_update_rms = sqrt(s[2]) / sqrt(nf);
        S.it += 1;
    }
}

// sums[0] = the np workgroup partials added in a fixed order (thread t adds
// t, t + 256, ... in order, then an LDS tree).
__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int np,
                                                   double* __restrict__ sums) {
    double s = 0.0;
    for (int c = threadIdx.x; c < np; c += 256) s += part[c];
    __shared__ double sm[256];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[0] = sm[0];
}

__global__ void k_fill_int(int* __restrict__ p, int v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

template <typename R>
__global__ void k_copy(const cpx<R>* __restrict__ in, cpx<R>* __restrict__ out, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

}  // namespace ocd

/* ---------------------------------------------------------------------- */
/* Launchers (host).                                                       */
/* ---------------------------------------------------------------------- */
namespace ocl {
using namespace ocd;

template <typename R>
static hipError_t launch_pass(int pass, const ProbDev* dprobs, int nprob, int total, int block, size_t lds_bytes,
                              int grid, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    dim3 g(grid), b(block);
    switch (pass) {
        case 0: hipLaunchKernelGGL(k_row_fwd<R>, g, b, lds_bytes, st, dprobs, nprob, total); break;
        case 1: hipLaunchKernelGGL(k_col<R>, g, b, lds_bytes, st, dprobs, nprob, total); break;
        case 2: hipLaunchKernelGGL(k_row_inv<R>, g, b, lds_bytes, st, dprobs, nprob, total); break;
        case 3: hipLaunchKernelGGL(k_rs_kgen<R>, g, b, lds_bytes, st, dprobs, nprob, total); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pass_c64(int pass, const ProbDev* d, int np, int total, int block, size_t lds, int grid,
                           hipStream_t st) {
    return launch_pass<float>(pass, d, np, total, block, lds, grid, st);
}
hipError_t launch_pass_c128(int pass, const ProbDev* d, int np, int total, int block, size_t lds, int grid,
                            hipStream_t st) {
    return launch_pass<double>(pass, d, np, total, block, lds, grid, st);
}

hipError_t set_lds_limits() {
    const int maxl = 160 * 1024;
    hipError_t e = hipSuccess;
    e = hipFuncSetAttribute((const void*)k_row_fwd<float>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl);
    if (e) return e;
    e = hipFuncSetAttribute((const void*)k_row_fwd<double>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl);
    if (e) return e;
    e = hipFuncSetAttribute((const void*)k_col<float>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl);
    if (e) return e;
    e = hipFuncSetAttribute((const void*)k_col<double>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl);
    if (e) return e;
    e = hipFuncSetAttribute((const void*)k_row_inv<float>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl);
    if (e) return e;
    e = hipFuncSetAttribute((const void*)k_row_inv<double>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl);
    if (e) return e;
    e = hipFuncSetAttribute((const void*)k_rs_kgen<float>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl);
    if (e) return e;
    return hipFuncSetAttribute((const void*)k_rs_kgen<double>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl);
}

hipError_t launch_mask_apply_c64(const void* in, void* out, const MaskDev& m, int nx, int ny, int* flag,
                                 hipStream_t st) {
    const size_t n = (size_t)nx * ny;
    int grid = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_mask_apply<float>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<float>*>(in),
                       static_cast<cpx<float>*>(out), m, nx, ny, flag);
    return hipGetLastError();
}
hipError_t launch_mask_trans(bool dbl, void* out, const MaskDev& m, int nx, int ny, hipStream_t st) {
    const size_t n = (size_t)nx * ny;
    int grid = (int)std::min<size_t>((n + 255) / 256, 65536);
    if (grid < 1) grid = 1;
    if (dbl)
        hipLaunchKernelGGL(k_mask_trans<double>, dim3(grid), dim3(256), 0, st, static_cast<cpx<double>*>(out), m, nx,
                           ny);
    else
        hipLaunchKernelGGL(k_mask_trans<float>, dim3(grid), dim3(256), 0, st, static_cast<cpx<float>*>(out), m, nx,
                           ny);
    return hipGetLastError();
}

hipError_t launch_mask_apply_c128(const void* in, void* out, const MaskDev& m, int nx, int ny, int* flag,
                                  hipStream_t st) {
    const size_t n = (size_t)nx * ny;
    int grid = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_mask_apply<double>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<double>*>(in),
                       static_cast<cpx<double>*>(out), m, nx, ny, flag);
    return hipGetLastError();
}
hipError_t launch_check_finite(bool dbl, const void* in, size_t n, int* flag, hipStream_t st) {
    int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    if (grid < 1) grid = 1;
    if (dbl)
        hipLaunchKernelGGL(k_check_finite<double>, dim3(grid), dim3(256), 0, st,
                           static_cast<const cpx<double>*>(in), n, flag);
    else
        hipLaunchKernelGGL(k_check_finite<float>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<float>*>(in),
                           n, flag);
    return hipGetLastError();
}
hipError_t launch_copy(bool dbl, const void* in, void* out, size_t n, hipStream_t st) {
    int grid = (int)std::min<size_t>((n + 255) / 256, 16384);
    if (grid < 1) grid = 1;
    if (dbl)
        hipLaunchKernelGGL(k_copy<double>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<double>*>(in),
                           static_cast<cpx<double>*>(out), n);
    else
        hipLaunchKernelGGL(k_copy<float>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<float>*>(in),
                           static_cast<cpx<float>*>(out), n);
    return hipGetLastError();
}

hipError_t launch_seed(bool dbl, const void* u, const double* dmdi, void* out, size_t n, hipStream_t st) {
    int grid = (int)std::min<size_t>((n + 255) / 256, 16384);
    if (grid < 1) grid = 1;
    if (dbl)
        hipLaunchKernelGGL(k_seed<double>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<double>*>(u), dmdi,
                           static_cast<cpx<double>*>(out), n);
    else
        hipLaunchKernelGGL(k_seed<float>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<float>*>(u), dmdi,
                           static_cast<cpx<float>*>(out), n);
    return hipGetLastError();
}

hipError_t launch_adj_elem(bool dbl, void* co, const void* ckpt, double* grad, const MaskDev& m, int nx, int ny,
                           hipStream_t st) {
    const size_t n = (size_t)nx * ny;
    int grid = (int)std::min<size_t>((n + 255) / 256, 16384);
    if (grid < 1) grid = 1;
    if (dbl)
        hipLaunchKernelGGL(k_adj_elem<double>, dim3(grid), dim3(256), 0, st, static_cast<cpx<double>*>(co),
                           static_cast<const cpx<double>*>(ckpt), grad, m, nx, ny);
    else
        hipLaunchKernelGGL(k_adj_elem<float>, dim3(grid), dim3(256), 0, st, static_cast<cpx<float>*>(co),
                           static_cast<const cpx<float>*>(ckpt), grad, m, nx, ny);
    return hipGetLastError();
}

hipError_t launch_block_reorder(bool dbl, const void* src, void* dst, int nx, int ny, int B, hipStream_t st) {
    const size_t n = (size_t)nx * ny;
    int grid = (int)std::min<size_t>((n + 255) / 256, 16384);
    if (grid < 1) grid = 1;
    if (dbl)
        hipLaunchKernelGGL(k_block_reorder<double>, dim3(grid), dim3(256), 0, st, static_cast<const double*>(src),
                           static_cast<double*>(dst), nx, ny, B);
    else
        hipLaunchKernelGGL(k_block_reorder<float>, dim3(grid), dim3(256), 0, st, static_cast<const float*>(src),
                           static_cast<float*>(dst), nx, ny, B);
    return hipGetLastError();
}

hipError_t launch_seed_mse(bool dbl, const void* u, const double* target, void* out, size_t n, double* sums,
                           hipStream_t st) {
    // a fixed grid (for a given n) of workgroup partials at sums + 8, added
    // in order into sums[0]: the merit does not depend on scheduling
    int grid = (int)std::min<size_t>((n + 255) / 256, kSeedParts);
    if (grid < 1) grid = 1;
    if (dbl)
        hipLaunchKernelGGL(k_seed_mse<double>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<double>*>(u),
                           target, static_cast<cpx<double>*>(out), n, (double)n, sums + 8);
    else
        hipLaunchKernelGGL(k_seed_mse<float>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<float>*>(u),
                           target, static_cast<cpx<float>*>(out), n, (double)n, sums + 8);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, sums + 8, grid, sums);
    return hipGetLastError();
}

hipError_t launch_seed_frames(bool dbl, const void* u, const double* target, void* out, size_t n, int nb, int ns,
                              int sel_b, int sel_s, double* sums, hipStream_t st) {
    int grid = (int)std::min<size_t>((n + 255) / 256, kSeedParts);
    if (grid < 1) grid = 1;
    if (dbl)
        hipLaunchKernelGGL(k_seed_frames<double>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<double>*>(u),
                           target, static_cast<cpx<double>*>(out), n, nb, ns, sel_b, sel_s, (double)n, sums + 8);
    else
        hipLaunchKernelGGL(k_seed_frames<float>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<float>*>(u),
                           target, static_cast<cpx<float>*>(out), n, nb, ns, sel_b, sel_s, (double)n, sums + 8);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, sums + 8, grid, sums);
    return hipGetLastError();
}

hipError_t launch_gs_after_fwd(GsState* st, const double* parts, const int* nparts, size_t pstride, size_t n,
                               GsRec* hist, size_t hstride, int restarts, hipStream_t stream) {
    hipLaunchKernelGGL(k_gs_after_fwd, dim3(restarts), dim3(256), 0, stream, st, parts, nparts, pstride, (double)n,
                       hist, hstride);
    return hipGetLastError();
}
hipError_t launch_gs_after_bwd(GsState* st, const double* parts, const int* nparts, size_t pstride, size_t n,
                               GsRec* hist, size_t hstride, int restarts, hipStream_t stream) {
    hipLaunchKernelGGL(k_gs_after_bwd, dim3(restarts), dim3(256), 0, stream, st, parts, nparts, pstride, (double)n,
                       hist, hstride);
    return hipGetLastError();
}
hipError_t launch_gs_keep_best(const GsState* st, const double* phase, double* best, size_t n, int restarts,
                               hipStream_t stream) {
    const int grid = (int)std::max<size_t>(1, std::min<size_t>((n * restarts + 255) / 256, 16384));
    hipLaunchKernelGGL(k_gs_keep_best, dim3(grid), dim3(256), 0, stream, st, phase, best, n, restarts);
    return hipGetLastError();
}

hipError_t launch_frame_sum(const double* slab, double* out, size_t m, int frames, hipStream_t st) {
    int grid = (int)std::min<size_t>((m + 255) / 256, 16384);
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_frame_sum, dim3(grid), dim3(256), 0, st, slab, out, m, frames);
    return hipGetLastError();
}

hipError_t launch_render_obs(bool dbl, const void* u, double* obs, size_t n, int frames, int nr, int cx, int cy,
                             int cl, double hs, double hl, int* flag, hipStream_t st) {
    int grid = (int)std::min<size_t>((n * (size_t)frames + 255) / 256, 16384);
    if (grid < 1) grid = 1;
    if (dbl)
        hipLaunchKernelGGL(k_render_obs<double>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<double>*>(u), obs,
                           n, frames, nr, cx, cy, cl, hs, hl, flag);
    else
        hipLaunchKernelGGL(k_render_obs<float>, dim3(grid), dim3(256), 0, st, static_cast<const cpx<float>*>(u), obs,
                           n, frames, nr, cx, cy, cl, hs, hl, flag);
    return hipGetLastError();
}

hipError_t launch_mse(bool dbl, const void* u, const double* target, size_t n, int nr, double* part, double* merit,
                      hipStream_t st) {
    for (int r0 = 0; r0 < nr; r0 += 65535) {
        const dim3 grid(kRedChunks, (unsigned)std::min(65535, nr - r0));
        if (dbl)
            hipLaunchKernelGGL(k_mse_part<double>, grid, dim3(256), 0, st, static_cast<const cpx<double>*>(u), target,
                               n, r0, part);
        else
            hipLaunchKernelGGL(k_mse_part<float>, grid, dim3(256), 0, st, static_cast<const cpx<float>*>(u), target,
                               n, r0, part);
    }
    hipLaunchKernelGGL(k_mse_final, dim3((nr + 255) / 256), dim3(256), 0, st, part, nr, (double)n, merit);
    return hipGetLastError();
}

hipError_t launch_adam(double* x, const double* g, double* m, double* v, const double* lo, const double* hi, size_t n,
                       double* part, double* scale, int has_clip, double cn, double lr, double b1, double b2,
                       double eps, double bc1, double bc2, hipStream_t st) {
    hipLaunchKernelGGL(k_sq_part, dim3(kRedChunks), dim3(256), 0, st, g, n, part);
    hipLaunchKernelGGL(k_clip_scale, dim3(1), dim3(64), 0, st, part, has_clip, cn, scale);
    int grid = (int)std::min<size_t>((n + 255) / 256, 16384);
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_adam, dim3(grid), dim3(256), 0, st, x, g, m, v, lo, hi, n, scale, lr, b1, b2, eps, bc1,
                       bc2);
    return hipGetLastError();
}

hipError_t launch_mtf(const double* psf, size_t nx, size_t ny, size_t count, double* sums, const void* tw,
                      double* values, int* flag, hipStream_t st) {
    int grid = (int)std::min<size_t>((nx * count + 255) / 256, 16384);
    hipLaunchKernelGGL(k_colsum, dim3(grid), dim3(256), 0, st, psf, nx, ny, count, sums, flag);
    const size_t nb = (nx / 2 + 1) * count;
    grid = (int)std::min<size_t>((nb + 63) / 64, 16384);
    hipLaunchKernelGGL(k_mtf_dft, dim3(grid), dim3(64), 0, st, sums, static_cast<const double2*>(tw), nx, count,
                       values);
    return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int v, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

}  // namespace ocl
