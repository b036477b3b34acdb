// Propagations whose padded transforms do not fit on chip (strict-mode
// densification to max_padding: e.g. a 2048-sample axis padded 8x to 16384,
// 256 KB per complex128 line against 160 KB of LDS per CU).
//
// The padded field lives in HBM as a natural-layout P_y x P_x buffer and
// every stage is a kernel over it:
//   k_big_embed     zero-pad the field (with the first pass's pre-op: phase
//                   mask, shared transmission, GS source) and flag non-finite
//                   samples of the (masked) field;
//   k_big_line      batched 1-D transforms along either axis: each line is
//                   gathered into LDS (lines fastest, so neighbouring lines
//                   are neighbouring addresses), transformed by the generic
//                   Stockham stages, optionally twiddled, and scattered;
//   k_big_kernel    the sampled RS / Fresnel-convolution kernel;
//   k_big_transfer  the transfer function or the kernel spectrum;
//   k_big_crop      crop, scale 1/(P_x P_y) and the last pass's post-op.
// A line longer than the on-chip transform runs as four steps, L = L1 L2:
// L2 strided length-L1 transforms with the twiddle W_L^{n2 k1}, then L1
// contiguous length-L2 transforms scattered at stride L1 into a second
// buffer.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "prop_kernels.hpp"

#include "pass_ops.hpp"

namespace ocd {

__device__ __forceinline__ long long line_at(const LineView& v, long long inner, long long l, int e) {
    return (l / inner) * v.outer + (l % inner) * v.inner_step + (long long)e * v.elem;
}

template <typename R, int DIR>
__global__ void k_big_line(const cpx<R>* in, cpx<R>* out, LineSpec s, FftPlanDev pl, const cpx<R>* twl) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x, bdim = blockDim.x;
    const int n = s.n, cnt = s.cnt;
    const int stride = lds_stride(n) * (pl.generic ? 2 : 1);
    const long long groups = (s.count + cnt - 1) / cnt;
    const int T = pl.threads;
    const int u = tid / T;
    const int uu = u < cnt ? u : 0;
    const int t = u < cnt ? tid - u * T : n;  // idle threads still reach the barriers
    for (long long gi = blockIdx.x; gi < groups; gi += gridDim.x) {
        const long long l0 = gi * cnt;
        for (int i = tid; i < cnt * n; i += bdim) {
            const int w = i % cnt, e = i / cnt;
            const long long l = l0 + w;
            cpx<R> v = {R(0), R(0)};
            if (l < s.count) v = in[line_at(s.in, s.inner, l, e)];
            lds[w * stride + lidx(e)] = v;
        }
        __syncthreads();
        run_fft<DIR>(lds + uu * stride, lds + uu * stride + lds_stride(n), pl, t, false);
        for (int i = tid; i < cnt * n; i += bdim) {
            const int w = i % cnt, e = i / cnt;
            const long long l = l0 + w;
            if (l >= s.count) continue;
            cpx<R> v = lds[w * stride + lidx(e)];
            if (s.tw_len > 0) {
                const long long m = s.tw_outer ? l / s.inner : l % s.inner;
                v = cmul(v, tw_at<DIR>(twl, (int)((m * e) % s.tw_len)));
            }
            out[line_at(s.out, s.inner, l, e)] = v;
        }
        __syncthreads();
    }
}

template <typename R>
__global__ void k_big_embed(ProbDev P, cpx<R>* __restrict__ a) {
    const size_t px = P.px, total = px * (size_t)P.py;
    bool bad = false;
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / px), x = (int)(i - (size_t)y * px);
        cpx<R> v = {R(0), R(0)};
        if (y >= P.oy && y < P.oy + P.ny && x >= P.ox && x < P.ox + P.nx)
            v = load_pre<R>(P, x - P.ox, y - P.oy, bad, acc);
        a[i] = v;
    }
    if (bad) atomicOr(P.flag, 1);
    if (P.ops.pre == PRE_SEED_MSE) accumulate4(acc, 0.0, 0.0, 0.0, P.ops.sums ? P.ops.sums + 3 : nullptr);
}

template <typename R>
__global__ void k_big_kernel(ProbDev P, cpx<R>* __restrict__ k) {
    const size_t px = P.px, total = px * (size_t)P.py;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / px), x = (int)(i - (size_t)y * px);
        k[i] = conv_sample<R>(P.eng, x, y, P.px, P.py);
    }
}

template <typename R>
__global__ void k_big_transfer(ProbDev P, cpx<R>* __restrict__ a, const cpx<R>* __restrict__ kspec) {
    const size_t px = P.px, total = px * (size_t)P.py;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int ky = (int)(i / px), kx = (int)(i - (size_t)ky * px);
        a[i] = kspec ? cmul(a[i], kspec[i]) : transfer<R>(P.eng, kx, ky, a[i]);
    }
}

template <typename R>
__global__ void k_big_crop(ProbDev P, const cpx<R>* __restrict__ a) {
    const size_t nx = P.nx, total = nx * (size_t)P.ny;
    const R s = (R)P.scale;
    bool bad = false;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int iy = (int)(i / nx), ix = (int)(i - (size_t)iy * nx);
        const cpx<R> v = a[(size_t)(iy + P.oy) * P.px + (ix + P.ox)];
        const cpx<R> o = {v.re * s, v.im * s};
        bad |= !finite_c(o);
        store_post<R>(P, ix, iy, o, acc0, acc1, acc2);
    }
    if (bad) atomicOr(P.flag, 2);
    if (P.ops.post == POST_GS_TARGET || P.ops.post == POST_GS_PHASE) accumulate4(acc0, acc1, acc2, 0.0, P.ops.sums);
}

}  // namespace ocd

using namespace ocd;

namespace ocl {

static int elem_grid(size_t total) {
    return (int)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 16384));
}

hipError_t big_set_lds_limits() {
    const int maxl = 160 * 1024;
    hipError_t e;
    if ((e = hipFuncSetAttribute((const void*)k_big_line<float, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl)))
        return e;
    if ((e = hipFuncSetAttribute((const void*)k_big_line<float, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl)))
        return e;
    if ((e = hipFuncSetAttribute((const void*)k_big_line<double, -1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 maxl)))
        return e;
    return hipFuncSetAttribute((const void*)k_big_line<double, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, maxl);
}

// Lines of `s` (s.cnt lines per workgroup, pl.threads threads each) with
// lds bytes of LDS per workgroup.
hipError_t big_line(bool dbl, int dir, const void* in, void* out, const LineSpec& s, const FftPlanDev& pl,
                    const void* twl, size_t lds, int max_grid, hipStream_t st) {
    const long long groups = (s.count + s.cnt - 1) / s.cnt;
    const int grid = (int)std::max<long long>(1, std::min<long long>(groups, max_grid));
    const int block = ((s.cnt * pl.threads + 63) / 64) * 64;
    if (dbl) {
        auto k = dir < 0 ? k_big_line<double, -1> : k_big_line<double, 1>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, st, static_cast<const cpx<double>*>(in),
                           static_cast<cpx<double>*>(out), s, pl, static_cast<const cpx<double>*>(twl));
    } else {
        auto k = dir < 0 ? k_big_line<float, -1> : k_big_line<float, 1>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, st, static_cast<const cpx<float>*>(in),
                           static_cast<cpx<float>*>(out), s, pl, static_cast<const cpx<float>*>(twl));
    }
    return hipGetLastError();
}

hipError_t big_embed(bool dbl, const ProbDev& P, void* a, hipStream_t st) {
    const int g = elem_grid((size_t)P.px * P.py);
    if (dbl) hipLaunchKernelGGL(k_big_embed<double>, dim3(g), dim3(256), 0, st, P, static_cast<cpx<double>*>(a));
    else hipLaunchKernelGGL(k_big_embed<float>, dim3(g), dim3(256), 0, st, P, static_cast<cpx<float>*>(a));
    return hipGetLastError();
}

hipError_t big_kernel(bool dbl, const ProbDev& P, void* k, hipStream_t st) {
    const int g = elem_grid((size_t)P.px * P.py);
    if (dbl) hipLaunchKernelGGL(k_big_kernel<double>, dim3(g), dim3(256), 0, st, P, static_cast<cpx<double>*>(k));
    else hipLaunchKernelGGL(k_big_kernel<float>, dim3(g), dim3(256), 0, st, P, static_cast<cpx<float>*>(k));
    return hipGetLastError();
}

hipError_t big_transfer(bool dbl, const ProbDev& P, void* a, const void* kspec, hipStream_t st) {
    const int g = elem_grid((size_t)P.px * P.py);
    if (dbl)
        hipLaunchKernelGGL(k_big_transfer<double>, dim3(g), dim3(256), 0, st, P, static_cast<cpx<double>*>(a),
                           static_cast<const cpx<double>*>(kspec));
    else
        hipLaunchKernelGGL(k_big_transfer<float>, dim3(g), dim3(256), 0, st, P, static_cast<cpx<float>*>(a),
                           static_cast<const cpx<float>*>(kspec));
    return hipGetLastError();
}

hipError_t big_crop(bool dbl, const ProbDev& P, const void* a, hipStream_t st) {
    const int g = elem_grid((size_t)P.nx * P.ny);
    if (dbl) hipLaunchKernelGGL(k_big_crop<double>, dim3(g), dim3(256), 0, st, P, static_cast<const cpx<double>*>(a));
    else hipLaunchKernelGGL(k_big_crop<float>, dim3(g), dim3(256), 0, st, P, static_cast<const cpx<float>*>(a));
    return hipGetLastError();
}

}  // namespace ocl
