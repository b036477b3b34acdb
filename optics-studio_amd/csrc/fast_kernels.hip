// Power-of-two pass kernels (P_x = P_y = N, 64 <= N <= 8192) on the
// register-resident FFT of fft_fast.hpp.  Same passes, fused ops and blocked
// intermediate as prop_kernels.hip; every global access is coalesced and no
// transform data goes through LDS except the inter-stage exchanges.
//
// Thread maps:
//   row passes : G rows per workgroup, thread (tb, g, tq) with t = tb + B*tq,
//                so a wave touches B*G contiguous elements of the blocked
//                intermediate per access.
//   column pass: B columns per workgroup, thread (g, t) with g fastest, so a
//                wave reads/writes 64 contiguous elements of the slab.
//
// Kernels are specialised at compile time on what a launch holds uniformly
// (the pre-op of the row pass, the engine of the column pass, whether the
// inverse row pass has a post-op), so a plain propagation does not pay
// registers for the GS and adjoint epilogues.  When a launch mixes kinds the
// general kernels dispatch per problem.
//
// The file is compiled once per element type: OC_FAST_DOUBLE selects
// complex128, otherwise complex64.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fft_fast.hpp"
#include "pass_ops.hpp"

namespace ocd {

// Workgroup layout of variant V at length N:
//   G  rows per row-pass workgroup (G*T threads),
//   B  width of the blocked intermediate [P/B][ny][B] (a row workgroup
//      writes runs of G*B contiguous elements; a column workgroup reads
//      slabs of B*ny),
//   CB columns per column-pass workgroup (CB*T threads, CB | B).
// Thread caps per variant (rows / cols) and the run width:
//   V0: 1024 / 1024, B = G    -- fewest, widest workgroups
//   V1:  512 /  512, B = G    -- more registers per thread (complex128:
//                                 half the LDS, two workgroups per CU)
//   V2: 1024 /  512, B = G    -- wide row passes, two column workgroups/CU
//   V3:  512 /  512, B = 2G   -- narrow row passes with V0's run width
// each within the 140 KB LDS budget and at most 16 rows/columns.  The
// autotuner picks per size; N < 1024 has V0 only.
template <typename R, int N, int V = 0>
struct FastLayout {
    static constexpr int T = N / 16;
    static constexpr int ESZ = (int)sizeof(cpx<R>);
    static constexpr int BYTES = Pow2Plan<N>::RSTRIDE * ESZ;  // per transform (the row layout's stride)
    // LDS budget per workgroup: 140 KB; complex128 variant 1 takes 70 KB so
    // that two workgroups share a CU
    static constexpr int LDSB = (V == 1 && ESZ == 16) ? 70 * 1024 : 140 * 1024;
    static constexpr int LG = LDSB / BYTES > 1 ? LDSB / BYTES : 1;
    static constexpr int ROWCAP = (V == 1 || V == 3) ? 512 : 1024;
    static constexpr int COLCAP = V == 0 ? 1024 : 512;
    static constexpr int BM = V == 3 ? 2 : 1;
    static constexpr int min3(int a, int b, int c) { return a < b ? (a < c ? a : c) : (b < c ? b : c); }
    static constexpr int G = min3(ROWCAP / T, LG, 16);
    static constexpr int B = min3(G * BM, T, 16);
    static constexpr int CB = min3(B, COLCAP / T, LG);
    static_assert(G >= 1 && B >= 1 && CB >= 1 && (T % B) == 0 && (B % CB) == 0, "layout");
};

// The transmission or phase-mask pre-op of a row pass on the row's 16
// samples in registers (x = t + m T in the padded row, iy the field row):
// PRE_TMUL multiplies by the shared transmission, PRE_MASK applies the
// problem's mask, PRE_MASKS either per problem; others leave v alone.
template <typename R, int N, int PK>
__device__ __forceinline__ void row_pre(const ProbDev& P, cpx<R> (&v)[16], int iy, bool rowok, int t) {
    constexpr int T = N / 16;
    const int ox = P.ox, nx = P.nx;
    const unsigned r0 = (unsigned)iy * nx - ox;
    if (PK == PRE_TMUL || (PK == PRE_MASKS && P.ops.pre == PRE_TMUL)) {
        const cpx<R>* tr = static_cast<const cpx<R>*>(P.ops.trans);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int x = t + m * T;
            if (rowok && x >= ox && x < ox + nx) {
                cpx<R> w;
                w = at32(tr, r0 + x);
                v[m] = tmul(v[m], w);
            }
        }
    }
    if (PK == PRE_MASK || (PK == PRE_MASKS && P.ops.pre == PRE_MASK)) {
        const MaskDev& mk = P.mask;
        if (mk.mode == MASK_SEPARABLE) {
            MaskAxis Y;
            Y = at32(mk.sy, rowok ? iy : 0);
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int x = t + m * T;
                if (rowok && x >= ox && x < ox + nx) {
                    MaskAxis X;
                    X = at32(mk.sx, x - ox);
                    const double lv = bilinear(mk.values, mk.mnx, X.i0, X.i1, Y.i0, Y.i1, X.t, Y.t);
                    const double phi = (X.in & Y.in) ? lv : 0.0;
                    const cpx<R> o = cmul(v[m], unit_phase<R>(phi));
                    v[m] = ap_contains(mk, X.r, Y.r) ? o : cpx<R>{R(0), R(0)};
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int x = t + m * T;
                if (rowok && x >= ox && x < ox + nx) v[m] = apply_mask(mk, x - ox, iy, v[m]);
            }
        }
    }
}

// PK: pre-op (PRE_NONE, PRE_MASK, PRE_TMUL, PRE_POLAR, PRE_MASKS for a mix of the
// first three, or -1 for the descriptor's).
template <typename R, int N, int V, int PK>
__global__ __launch_bounds__((FastLayout<R, N, V>::G * (N / 16))) void f_row_fwd(const ProbDev* __restrict__ probs,
                                                                         int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, G = LY::G, Q = LY::B;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int tb = tid % Q, g = (tid / Q) % G, tq = tid / (Q * G);
    const int t = tb + Q * tq;
    TwKeep<R, N, (N == 1024 || N == 2048 || N == 4096) && sizeof(R) == 4, N == 4096 && sizeof(R) == 4 && (PK == PRE_NONE || PK == PRE_TMUL)> tk;
    tk.template init<G * T>(lds + G * RowStride<N, G>::value, static_cast<const cpx<R>*>(probs[0].prow.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        int pi, lt;
        if (PK == PRE_TMUL && probs[0].ilv) {
            pi = tile % nprob;
            lt = tile / nprob;
        } else {
            pi = find_problem<PASS_ROW>(probs, nprob, tile);
            lt = tile - probs[pi].base_row;
        }
        const ProbDev& P = probs[pi];
        const int iy = lt * G + g;
        const bool rowok = iy < P.ny;
        const int ox = P.ox, nx = P.nx;
        bool bad = false;
        double acc = 0.0;
        cpx<R> v[16];
        if constexpr (PK == PRE_NONE || PK == PRE_MASK || PK == PRE_TMUL || PK == PRE_MASKS) {
            // the row's 16 samples as one batch of loads, then the checks,
            // then the mask with its mode decided once for the row
            const cpx<R>* in = static_cast<const cpx<R>*>(P.in);
            const unsigned r0 = (unsigned)iy * nx - ox;
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int x = t + m * T;
                v[m] = {R(0), R(0)};
                if (rowok && x >= ox && x < ox + nx) v[m] = at32(in, r0 + x);
            }
            // the engine validates the masked field (samples outside the
            // stop are zero after the apply)
            row_pre<R, N, PK>(P, v, iy, rowok, t);
#pragma unroll
            for (int m = 0; m < 16; ++m) bad |= !finite_c(v[m]);
        } else if constexpr (PK == PRE_POLAR) {
            // GS source plane: amplitude and phase of the 16 samples as one
            // batch of loads, then from_polar (masked by the stop)
            const R* amp = static_cast<const R*>(P.ops.amp);
            const R* phs = static_cast<const R*>(P.ops.phase);
            const unsigned char* open = P.ops.open;
            const unsigned r0 = (unsigned)iy * nx - ox;
            R a[16], ph[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int x = t + m * T;
                a[m] = R(0);
                ph[m] = R(0);
                if (rowok && x >= ox && x < ox + nx) {
                    a[m] = at32(amp, r0 + x);
                    ph[m] = at32(phs, r0 + x);
                    if (open && !at32(open, r0 + x)) a[m] = R(0);
                }
            }
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const cpx<R> e = unit_phase<R>((double)ph[m]);
                v[m] = {a[m] * e.re, a[m] * e.im};
            }
        } else {
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int x = t + m * T;
                v[m] = {R(0), R(0)};
                if (rowok && x >= ox && x < ox + nx) v[m] = load_pre<R, PK>(P, x - ox, iy, bad, acc);
            }
        }
        if (bad) atomicOr(P.flag, 1);
        if (PK < 0 && P.ops.pre == PRE_SEED_MSE)
            accumulate4(acc, 0.0, 0.0, 0.0, P.ops.sums ? P.ops.sums + 3 : nullptr);
        fft_regs<R, N, G, -1, true>(v, lds, g, t, static_cast<const cpx<R>*>(P.prow.tw), tk);
        if (rowok) {
            cpx<R>* work = static_cast<cpx<R>*>(P.work);
            const unsigned ny = P.ny, b0 = ((unsigned)tq * ny + iy) * Q + tb;
#pragma unroll
            for (int m = 0; m < 16; ++m) at32(work, b0 + (unsigned)(m * (T / Q)) * ny * Q) = v[m];
        }
    }
}

// The forward row pass of a batch whose problems share one transmission
// (PRE_TMUL) and are all N = P/2: each thread's window is samples
// m = 4..11, and those eight inputs are loaded one tile ahead, so that the
// next tile's loads are in flight during this tile's transform.  Per tile:
// the transmission loads first, then the next tile's inputs, then the
// multiply (waiting on the transmission only), transform and store.
template <typename R, int N, int V>
__global__ __launch_bounds__((FastLayout<R, N, V>::G * (N / 16))) void f_row_fwd_pf(
    const ProbDev* __restrict__ probs, int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, G = LY::G, Q = LY::B;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int tb = tid % Q, g = (tid / Q) % G, tq = tid / (Q * G);
    const int t = tb + Q * tq;
    // the table (not the hoist: 138 VGPRs or a spill) at N = 4096
    // (row pass 0.456 -> 0.435 ms on the 16-band 2048^2 batch)
    TwKeep<R, N, (N == 1024 || N == 2048 || N == 4096) && sizeof(R) == 4> tk;
    tk.template init<G * T>(lds + G * RowStride<N, G>::value, static_cast<const cpx<R>*>(probs[0].prow.tw), t);
    const bool ilv = probs[0].ilv != 0;
    auto locate = [&](int tile, int& pi, int& iy) {
        int lt;
        if (ilv) {
            pi = tile % nprob;
            lt = tile / nprob;
        } else {
            pi = find_problem<PASS_ROW>(probs, nprob, tile);
            lt = tile - probs[pi].base_row;
        }
        iy = lt * G + g;
    };
    // window sample m + 4 of row iy of problem pi (x = t + (m + 4) T - ox)
    auto load8 = [&](int pi, int iy, cpx<R> (&w)[8]) {
        const ProbDev& P = probs[pi];
        const bool ok = iy < P.ny;
        const cpx<R>* in = static_cast<const cpx<R>*>(P.in);
        const unsigned r0 = (unsigned)iy * P.nx - P.ox;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            w[m] = {R(0), R(0)};
            if (ok) w[m] = at32(in, r0 + (unsigned)(t + (m + 4) * T));
        }
    };
    cpx<R> nxt[8];
    int tile = xcd_block(), pi = 0, iy = 0;
    if (tile < total) {
        locate(tile, pi, iy);
        load8(pi, iy, nxt);
    }
    for (; tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[pi];
        const int cy = iy;
        const bool rowok = cy < P.ny;
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = (m >= 4 && m < 12) ? nxt[m - 4] : cpx<R>{R(0), R(0)};
        cpx<R> tr[8];
        const cpx<R>* trs = static_cast<const cpx<R>*>(P.ops.trans);
        const unsigned r0 = (unsigned)cy * P.nx - P.ox;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            tr[m] = {R(0), R(0)};
            if (rowok) tr[m] = at32(trs, r0 + (unsigned)(t + (m + 4) * T));
        }
        const ProbDev* Pc = &P;
        if (tile + (int)gridDim.x < total) {
            locate(tile + gridDim.x, pi, iy);
            load8(pi, iy, nxt);
        }
#pragma unroll
        for (int m = 4; m < 12; ++m) v[m] = tmul(v[m], tr[m - 4]);
        bool bad = false;
#pragma unroll
        for (int m = 4; m < 12; ++m) bad |= !finite_c(v[m]);
        if (bad) atomicOr(Pc->flag, 1);
        fft_regs<R, N, G, -1, true>(v, lds, g, t, static_cast<const cpx<R>*>(Pc->prow.tw), tk);
        if (rowok) {
            cpx<R>* work = static_cast<cpx<R>*>(Pc->work);
            const unsigned ny = Pc->ny, b0 = ((unsigned)tq * ny + cy) * Q + tb;
#pragma unroll
            for (int m = 0; m < 16; ++m) at32(work, b0 + (unsigned)(m * (T / Q)) * ny * Q) = v[m];
        }
    }
}

// EK: engine (ENG_ASM, ENG_FRESNEL, ENG_GS_PROJECT, or -1 for the descriptor's); RS: the
// Rayleigh-Sommerfeld product with the kernel spectrum instead.
// HW: every problem is N = P/2 on both axes, so the window is exactly
// samples m = 4..11 of each thread's sixteen; the zeros outside it are
// compile-time constants, and the compiler drops the forward transform's
// first-stage work on them and the inverse's last-stage work on outputs never
// stored.
template <typename R, int N, int V, bool RS, int EK, bool HW = false>
__global__ __launch_bounds__((FastLayout<R, N, V>::CB * (N / 16))) void f_col(const ProbDev* __restrict__ probs,
                                                                      int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, B = LY::B, CB = LY::CB;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int g = tid % CB, t = tid / CB;
    // the half-window complex64 ASM form at N = 4096 in the 512-thread
    // layouts (V2, V3) has the registers to keep its twiddles (table and
    // hoist) within 128 VGPRs; that took the pass from 0.87 to 0.76 ms on
    // MI355X.  The half-window convolution form (RS) has them in V1-V3.
    // Elsewhere at 4096 the kept factors pushed the kernel past 128 VGPRs.
    constexpr bool KEEP = HW && sizeof(R) == 4 && N == 4096 &&
                          ((!RS && EK == ENG_ASM && (V == 2 || V == 3)) || (RS && V != 0));
    // complex64 at 1024 and 2048 (two radix-16 stages: the table only)
    // keeps it in every form; the table replaces the factor products, so
    // these kernels use 2-10 fewer VGPRs with it (the GS projection pass at
    // 1024: 6% fewer VALU instructions per tile).
    constexpr bool KEEP_TAB = KEEP || (sizeof(R) == 4 && (N == 1024 || N == 2048));
    TwKeep<R, N, KEEP_TAB, KEEP> tk;
    tk.template init<CB * T>(lds + CB * Pow2Plan<N>::SLOTS, static_cast<const cpx<R>*>(probs[0].pcol.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_COL>(probs, nprob, tile)];
        int lt = tile - P.base_col;
        if (RS && !(P.tiles_col & 1)) {
            // mirror-paired order: local tiles 2i and 2i + 1 take column
            // blocks i and Q - 1 - i, which read the same kernel-spectrum
            // columns min(kx, N - kx) (all but one), and run together on one
            // XCD, so the pair's second read comes from that XCD's L2
            lt = (lt & 1) ? P.tiles_col - 1 - (lt >> 1) : (lt >> 1);
        }
        const int kx = lt * CB + g;                 // this thread's column
        const int q = kx / B, bq = kx % B;          // its block and place in it
        const int ny = P.ny, oy = P.oy;
        const cpx<R>* tw = static_cast<const cpx<R>*>(P.pcol.tw);
        cpx<R>* work = static_cast<cpx<R>*>(P.work);
        // element (q, y - oy, bq) of the slab: b0 + y*B
        const unsigned b0 = ((unsigned)q * ny - oy) * B + bq;
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int y = t + m * T;
            v[m] = {R(0), R(0)};
            if (HW ? (m >= 4 && m < 12) : (y >= oy && y < oy + ny)) v[m] = at32(work, b0 + (unsigned)y * B);
        }
        fft_regs<R, N, CB, -1>(v, lds, g, t, tw, tk);
        if constexpr (RS) {
            // the kernel's 2-D spectrum from f_rs_kcol: W(kx, ky) is entry
            // (min(kx, N - kx), min(ky, N - ky)) of the stored quarter
            constexpr int KR = N / 2 + 1;
            const cpx<R>* wb = static_cast<const cpx<R>*>(P.kbuf) + (size_t)N * KR;
            const int kh = min(kx, N - kx);
            const unsigned w0 = (unsigned)(kh / B) * KR * B + kh % B;
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int y = t + m * T;
                v[m] = cmul(v[m], at32(wb, w0 + (unsigned)min(y, N - y) * B));
            }
        } else {
            const int ek = EK >= 0 ? EK : P.eng.kind;
            if (ek == ENG_FRAUNHOFER) {
                // forward only: every bin goes to its far-field sample
                cpx<R>* out = static_cast<cpx<R>*>(P.out);
                const int nxo = P.nx, nyo = P.ny;
#pragma unroll
                for (int m = 0; m < 16; ++m) {
                    int ox_, oy_;
                    const cpx<R> o = fraun_post<R>(P.eng, kx, t + m * T, nxo, nyo, v[m], &ox_, &oy_);
                    if (!finite_c(o)) atomicOr(P.flag, 2);
                    at32(out, (unsigned)oy_ * nxo + ox_) = o;
                }
                continue;
            } else if (ek == ENG_MEASURE) {
                const R* tb = static_cast<const R*>(P.ops.tamp_blk);
                const unsigned k0 = ((unsigned)q * N + t) * B + bq;
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int m = 0; m < 16; ++m) {
                    const double in2 = (double)v[m].re * v[m].re + (double)v[m].im * v[m].im;
                    const double ta = at32(tb, k0 + (unsigned)(m * T * B));
                    s0 += in2;
                    s1 += in2 * in2;
                    s2 += in2 * (ta * ta);
                }
                accumulate4(s0, s1, s2, 0.0, P.ops.sums);
                continue;
            } else if (ek == ENG_GS_PROJECT) {
                const R* tb = static_cast<const R*>(P.ops.tamp_blk);
                const unsigned k0 = ((unsigned)q * N + t) * B + bq;
#pragma unroll
                for (int m = 0; m < 16; ++m) {
                    const R a = at32(tb, k0 + (unsigned)(m * T * B));
                    if constexpr (sizeof(R) == 4) {
                        // complex64: a U / |U| through the hardware rsqrt
                        const float r2 = fmadd(v[m].re, v[m].re, v[m].im * v[m].im);
                        const float sc = a * __builtin_amdgcn_rsqf(r2);
                        v[m] = r2 > 0.0f ? cpx<R>{v[m].re * sc, v[m].im * sc} : cpx<R>{a, R(0)};
                    } else {
                        const R mg = hypot(v[m].re, v[m].im);
                        v[m] = mg > R(0) ? cpx<R>{v[m].re * (a / mg), v[m].im * (a / mg)} : cpx<R>{a, R(0)};
                    }
                }
            } else if ((EK == ENG_ASM || (EK == ENG_FRESNEL && sizeof(R) == 8)) && (sizeof(R) == 8 || N >= 4096)) {
                // The ASM transfer and the complex128 Fresnel transfer depend
                // on ky only through the axis table entry ay[ky], and
                // ay[ky] = ay[N - ky] exactly (fftfreq negates exactly), so
                // H(kx, ky) = H(kx, N - ky) bit for bit.  Each thread
                // evaluates its rows t + mT for m < 8 and takes rows m >= 8
                // from thread T - t, since N - (t + mT) = (T - t) + (15 - m) T.
                // The values pass through LDS, free between the two
                // transforms, as hs[j][T - t][g]; slot T holds thread 0's own
                // rows (16 - m) T, row N/2 among them.  (The complex64
                // Fresnel multiplies two table factors instead.)  Measured
                // on MI355X it pays in complex128 and from N = 4096 in
                // complex64 (column pass 0.91 -> 0.87 ms at 4096); the
                // shorter complex64 columns run faster without the extra
                // exchange (0.45 against 0.58 ms at 2048).
                const EngineDev& e = P.eng;
                const double axk = at32(e.ax, kx);
                auto tf = [&](int ky) -> cpx<R> {
                    if constexpr (EK == ENG_ASM) return asm_tf<R>(e.kz, axk, at32(e.ay, ky));
                    else return unit_phase<R>(e.kz - e.c * (axk + at32(e.ay, ky)));
                };
                cpx<R> h[8];
#pragma unroll
                for (int m = 0; m < 8; ++m) h[m] = tf(t + m * T);
                cpx<R>* hs = lds;  // [8][T + 1][CB] within the exchange buffer
#pragma unroll
                for (int j = 0; j < 8; ++j) hs[(j * (T + 1) + t) * CB + g] = h[j];
                if (t == 0) {
                    const cpx<R> h8 = tf(8 * T);
#pragma unroll
                    for (int j = 0; j < 8; ++j) hs[(j * (T + 1) + T) * CB + g] = j < 7 ? h[j + 1] : h8;
                }
                __syncthreads();
#pragma unroll
                for (int m = 0; m < 8; ++m) v[m] = cmul(v[m], h[m]);
#pragma unroll
                for (int m = 8; m < 16; ++m) v[m] = cmul(v[m], hs[((15 - m) * (T + 1) + (T - t)) * CB + g]);
                __syncthreads();  // before the inverse transform's exchanges reuse the buffer
            } else if (EK == ENG_ASM) {
                // the column's table entries once, the 16 rows' as a batch
                const EngineDev& e = P.eng;
                const double axk = at32(e.ax, kx);
#pragma unroll
                for (int m = 0; m < 16; ++m) v[m] = asm_transfer<R>(e.kz, axk, at32(e.ay, t + m * T), v[m]);
            } else if (ek != ENG_NONE) {
#pragma unroll
                for (int m = 0; m < 16; ++m) v[m] = transfer<R, EK>(P.eng, kx, t + m * T, v[m]);
            }
        }
        fft_regs<R, N, CB, +1>(v, lds, g, t, tw, tk);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int y = t + m * T;
            if (HW ? (m >= 4 && m < 12) : (y >= oy && y < oy + ny)) at32(work, b0 + (unsigned)y * B) = v[m];
        }
    }
}

// QK: the post-op every problem of the launch has (POST_NONE: store the
// cropped, scaled field; POST_GS_PHASE: the GS phase update), or -1 for the
// descriptor's.
template <typename R, int N, int V, int QK>
__global__ __launch_bounds__((FastLayout<R, N, V>::G * (N / 16))) void f_row_inv(const ProbDev* __restrict__ probs,
                                                                         int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, G = LY::G, Q = LY::B;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int tb = tid % Q, g = (tid / Q) % G, tq = tid / (Q * G);
    const int t = tb + Q * tq;
    TwKeep<R, N, (N == 1024 || N == 2048 || N == 4096) && sizeof(R) == 4, N == 4096 && sizeof(R) == 4 && QK == POST_NONE> tk;
    tk.template init<G * T>(lds + G * RowStride<N, G>::value, static_cast<const cpx<R>*>(probs[0].prow.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_ROW>(probs, nprob, tile)];
        const int iy = (tile - P.base_row) * G + g;
        const bool rowok = iy < P.ny;
        const int ny = P.ny, nx = P.nx, ox = P.ox;
        const cpx<R>* work = static_cast<const cpx<R>*>(P.work);
        const unsigned b0 = ((unsigned)tq * ny + iy) * Q + tb;
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            v[m] = {R(0), R(0)};
            if (rowok) v[m] = at32(work, b0 + (unsigned)(m * (T / Q)) * ny * Q);
        }
        fft_regs<R, N, G, +1, true>(v, lds, g, t, static_cast<const cpx<R>*>(P.prow.tw), tk);
        const R s = (R)P.scale;
        bool bad = false;
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        if (rowok) {
            cpx<R>* out = static_cast<cpx<R>*>(P.out);
            const unsigned o0 = (unsigned)iy * nx;
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int x = t + m * T - ox;
                if (x >= 0 && x < nx) {
                    const cpx<R> o = {v[m].re * s, v[m].im * s};
                    bad |= !finite_c(o);
                    if constexpr (QK == POST_NONE) {
                        at32(out, o0 + x) = o;
                    } else if constexpr (QK == POST_GS_PHASE) {
                        R* ph = static_cast<R*>(P.ops.phase_io);
                        const unsigned k = o0 + x;
                        if constexpr (sizeof(R) == 4) {
                            const float p = at32(ph, k);
                            const float dphi = wrap_pi_f(atan2_f(o.im, o.re) - p);
                            at32(ph, k) =
                                encode_bounds_f(p + (float)P.ops.beta_p * dphi, (float)P.ops.lo, (float)P.ops.hi);
                            acc2 += (double)dphi * dphi;
                            if (P.ops.dphi_out) at32(P.ops.dphi_out, k) = dphi;
                        } else {
                            const double p = at32(ph, k);
                            const double dphi = wrap_pi(atan2((double)o.im, (double)o.re) - p);
                            at32(ph, k) = (R)encode_bounds(p + P.ops.beta_p * dphi, P.ops.lo, P.ops.hi);
                            acc2 += dphi * dphi;
                            if (P.ops.dphi_out) at32(P.ops.dphi_out, k) = dphi;
                        }
                        if (P.out) at32(out, k) = o;
                    } else if constexpr (QK == POST_ADJ_GRAD) {
                        // adjoint.rs:222-236 with t and V from the element's mask
                        const unsigned k = o0 + x;
                        const cpx<R> tt = apply_mask(P.mask, x, iy, cpx<R>{R(1), R(0)});
                        cpx<R> ck;
                        ck = at32(static_cast<const cpx<R>*>(P.ops.ckpt), k);
                        const cpx<R> vv = cmul(ck, tt);
                        const double gim = (double)o.re * (double)vv.im + (-(double)o.im) * (double)vv.re;
                        at32(P.ops.grad, k) = 0.0 + -2.0 * gim;  // the plane's one term (no clearing pass)
                        at32(out, k) = cmul(o, cpx<R>{tt.re, -tt.im});
                    } else {
                        store_post<R>(P, x, iy, o, acc0, acc1, acc2);
                    }
                }
            }
        }
        if (bad) atomicOr(P.flag, 2);
        if (QK == POST_GS_PHASE || (QK < 0 && (P.ops.post == POST_GS_TARGET || P.ops.post == POST_GS_PHASE)))
            accumulate4_tile(acc0, acc1, acc2, 0.0, P.ops.sums, P.ops.parts, tile - P.base_row);
    }
}

// One GS iteration's inverse row pass fused with the next iteration's
// forward row pass (gs.rs:484-492, 557-567): row IFFT, scale, then
//  FULL:  phase update phi <- encode(phi + beta wrap(arg V - phi)) and the
//         new source field A e^{i phi} (masked by the stop); the phase
//         written is the one read back, as the separate passes would;
//  !FULL: beta = 1 with bounds spanning 2 pi, where the updated phase is
//         arg V mod 2 pi and A e^{i phi} = A V / |V| (V = 0: arg is 0 or pi
//         by the sign bit of Re V): the source field without the phase,
//         which the last iteration's inverse pass writes;
// then the row FFT, stored back in place.
template <typename R, int N, int V, bool FULL>
__global__ __launch_bounds__((FastLayout<R, N, V>::G * (N / 16))) void f_row_gs(const ProbDev* __restrict__ probs,
                                                                        int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, G = LY::G, Q = LY::B;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int tb = tid % Q, g = (tid / Q) % G, tq = tid / (Q * G);
    const int t = tb + Q * tq;
    // complex64 at 1024 keeps the second stage's factors in the LDS table
    // (103 VGPRs; the FFT-form GS loop at 1024^2 went from 103k to 106k
    // restart-iterations/s on MI355X)
    TwKeep<R, N, sizeof(R) == 4 && N == 1024> tk;
    tk.template init<G * T>(lds + G * RowStride<N, G>::value, static_cast<const cpx<R>*>(probs[0].prow.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_ROW>(probs, nprob, tile)];
        const int iy = (tile - P.base_row) * G + g;
        const bool rowok = iy < P.ny;
        const int ny = P.ny, nx = P.nx, ox = P.ox;
        cpx<R>* work = static_cast<cpx<R>*>(P.work);
        const unsigned b0 = ((unsigned)tq * ny + iy) * Q + tb;
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            v[m] = {R(0), R(0)};
            if (rowok) v[m] = at32(work, b0 + (unsigned)(m * (T / Q)) * ny * Q);
        }
        const R* amp = static_cast<const R*>(P.ops.amp);
        const unsigned char* open = P.ops.open;
        const unsigned o0 = (unsigned)iy * nx;
        // !FULL: the source amplitudes of the tile's samples are loaded here,
        // behind the row, so that the inverse transform covers their latency
        // (loaded after it, they waited behind its barriers: 0.161 ms per
        // pass at 1024^2 x 32 restarts on MI355X before the change)
        R av[16];
        if constexpr (!FULL) {
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int x = t + m * T - ox;
                av[m] = R(0);
                if (rowok && x >= 0 && x < nx) {
                    const unsigned k = o0 + x;
                    R a = at32(amp, k);
                    if (open && !at32(open, k)) a = R(0);
                    av[m] = a;
                }
            }
        }
        fft_regs<R, N, G, +1, true>(v, lds, g, t, static_cast<const cpx<R>*>(P.prow.tw), tk);
        const R s = (R)P.scale;
        bool bad = false;
        double acc2 = 0.0;
        R* ph = static_cast<R*>(P.ops.phase_io);
        cpx<R>* out = static_cast<cpx<R>*>(P.out);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int x = t + m * T - ox;
            if (rowok && x >= 0 && x < nx && !FULL) {
                const cpx<R> o = {v[m].re * s, v[m].im * s};
                bad |= !finite_c(o);
                const R a = av[m];
                if constexpr (sizeof(R) == 4) {
                    const float r2 = fmadd(o.re, o.re, o.im * o.im);
                    const float sc = a * __builtin_amdgcn_rsqf(r2);
                    v[m] = r2 > 0.0f ? cpx<R>{o.re * sc, o.im * sc}
                                     : cpx<R>{__builtin_signbit(o.re) ? -a : a, R(0)};
                } else {
                    const double mg = hypot(o.re, o.im);
                    v[m] = mg > 0.0 ? cpx<R>{o.re * (a / mg), o.im * (a / mg)}
                                    : cpx<R>{__builtin_signbit(o.re) ? -a : a, R(0)};
                }
            } else if (rowok && x >= 0 && x < nx) {
                const cpx<R> o = {v[m].re * s, v[m].im * s};
                bad |= !finite_c(o);
                const unsigned k = o0 + x;
                R pn;
                if constexpr (sizeof(R) == 4) {
                    const float p = at32(ph, k);
                    const float dphi = wrap_pi_f(atan2_f(o.im, o.re) - p);
                    pn = encode_bounds_f(p + (float)P.ops.beta_p * dphi, (float)P.ops.lo, (float)P.ops.hi);
                    acc2 += (double)dphi * dphi;
                    if (P.ops.dphi_out) at32(P.ops.dphi_out, k) = dphi;
                } else {
                    const double p = at32(ph, k);
                    const double dphi = wrap_pi(atan2((double)o.im, (double)o.re) - p);
                    pn = (R)encode_bounds(p + P.ops.beta_p * dphi, P.ops.lo, P.ops.hi);
                    acc2 += dphi * dphi;
                    if (P.ops.dphi_out) at32(P.ops.dphi_out, k) = dphi;
                }
                at32(ph, k) = pn;
                if (out) at32(out, k) = o;
                // the next iteration's source sample (the phase just written
                // is the one its forward pass reads)
                R a = at32(amp, k);
                if (open && !at32(open, k)) a = R(0);
                const cpx<R> e = unit_phase<R>((double)pn);
                v[m] = {a * e.re, a * e.im};
            } else {
                v[m] = {R(0), R(0)};
            }
        }
        if (bad) atomicOr(P.flag, 2);
        if (FULL) accumulate4(0.0, 0.0, acc2, 0.0, P.ops.sums);
        fft_regs<R, N, G, -1, true>(v, lds, g, t, static_cast<const cpx<R>*>(P.prow.tw), tk);
        if (rowok) {
#pragma unroll
            for (int m = 0; m < 16; ++m) at32(work, b0 + (unsigned)(m * (T / Q)) * ny * Q) = v[m];
        }
    }
}

// Row spectra of the sampled convolution kernel (RS or Fresnel
// convolution).  The kernel depends on y only through y^2 and the DFT-order
// coordinates of rows y and N - y are negatives of each other exactly, so
// those rows are equal: only rows 0..N/2 are sampled and transformed, and
// stored blocked [N/B][N/2 + 1][B]; the column pass reads row y from row
// min(y, N - y).
template <typename R, int N, int V>
__global__ __launch_bounds__((FastLayout<R, N, V>::G * (N / 16))) void f_rs_kgen(const ProbDev* __restrict__ probs,
                                                                         int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, G = LY::G, Q = LY::B;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int tb = tid % Q, g = (tid / Q) % G, tq = tid / (Q * G);
    const int t = tb + Q * tq;
    TwKeep<R, N, (N == 1024 || N == 2048 || N == 4096) && sizeof(R) == 4> tk;
    tk.template init<G * T>(lds + G * RowStride<N, G>::value, static_cast<const cpx<R>*>(probs[0].prow.tw), t);
    constexpr int KR = N / 2 + 1;  // distinct kernel rows
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_KGEN>(probs, nprob, tile)];
        const int y = (tile - P.base_kgen) * G + g;
        const bool rowok = y < KR;
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m)
            v[m] = rowok ? conv_sample<R>(P.eng, t + m * T, y, N, N) : cpx<R>{R(0), R(0)};
        fft_regs<R, N, G, -1, true>(v, lds, g, t, static_cast<const cpx<R>*>(P.prow.tw), tk);
        if (rowok) {
            // the kernel is even in x too, so only kx = t + mT <= N/2 (m < 8,
            // and m = 8 at t = 0) is needed
            cpx<R>* kb = static_cast<cpx<R>*>(P.kbuf);
            const unsigned k0 = ((unsigned)tq * KR + y) * Q + tb;
#pragma unroll
            for (int m = 0; m < 9; ++m)
                if (m < 8 || t == 0) at32(kb, k0 + (unsigned)(m * (T / Q)) * KR * Q) = v[m];
        }
    }
}

// Column transforms of the kernel's row spectra: the 2-D spectrum of the
// sampled kernel, W(kx, ky).  The kernel depends on x and y only through x^2
// and y^2, and the DFT-order coordinates of index i and N - i are exact
// negatives, so W(kx, ky) = W(N - kx, ky) = W(kx, N - ky).  This pass
// computes the distinct columns kx = 0..N/2 (each a transform over all N
// rows, row y read from the stored row min(y, N - y)) and keeps rows
// ky = 0..N/2.  W sits after the row spectra in the problem's kernel buffer,
// blocked [ceil(KR/B)][KR][B] with KR = N/2 + 1.  The field's column pass
// then reads W instead of transforming the kernel for every column: half
// the kernel transforms, and one transform fewer held in its registers.
template <typename R, int N, int V>
__global__ __launch_bounds__((FastLayout<R, N, V>::CB * (N / 16))) void f_rs_kcol(const ProbDev* __restrict__ probs,
                                                                          int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, B = LY::B, CB = LY::CB;
    constexpr int KR = N / 2 + 1;
    constexpr int TPP = (KR + CB - 1) / CB;  // tiles per problem
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int g = tid % CB, t = tid / CB;
    // complex64 keeps the table from 1024 and the third stage's factors at
    // 4096 (90-110 VGPRs)
    TwKeep<R, N, sizeof(R) == 4 && N >= 1024 && N <= 4096, sizeof(R) == 4 && N == 4096> tk;
    tk.template init<CB * T>(lds + CB * Pow2Plan<N>::SLOTS, static_cast<const cpx<R>*>(probs[0].pcol.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[tile / TPP];
        const int kx = (tile % TPP) * CB + g;  // 0..KR-1 where colok
        const bool colok = kx < KR;
        const cpx<R>* kb = static_cast<const cpx<R>*>(P.kbuf);
        const unsigned k0 = (unsigned)(kx / B) * KR * B + kx % B;
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int y = t + m * T;
            v[m] = {R(0), R(0)};
            if (colok) v[m] = at32(kb, k0 + (unsigned)min(y, N - y) * B);
        }
        fft_regs<R, N, CB, -1>(v, lds, g, t, static_cast<const cpx<R>*>(P.pcol.tw), tk);
        if (colok) {
            cpx<R>* wb = static_cast<cpx<R>*>(P.kbuf) + (size_t)N * KR;
            const unsigned w0 = (unsigned)(kx / B) * KR * B + kx % B;
#pragma unroll
            for (int m = 0; m < 9; ++m)
                if (m < 8 || t == 0) at32(wb, w0 + (unsigned)(t + m * T) * B) = v[m];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Separable complex64 Fresnel: two passes instead of three.            */
/* ------------------------------------------------------------------ */
// The Fresnel transfer function factors per axis (H = hx[kx] hy[ky]), and
// embedding, transforms and cropping act on each axis separately, so
//   out = crop_y IFFT_y hy FFT_y embed_y  .  in  .  (same along x)^T.
// Pass A applies the y factor column by column and keeps the ny cropped
// rows; pass B applies the x factor row by row and writes the output rows.
// 4N transforms of length P and 4 N^2 elements of traffic per band, against
// 6N and 2N^2 + 4NP for the three-pass form.  The intermediate is the
// cropped N_y x N_x field, blocked [ceil(N_x/B)][N_y][B].

// Pass A: CB input columns per tile, read from the natural-layout input
// (pre-op PK: none, phase mask, shared transmission or PRE_MASKS for a mix),
// written to the blocked intermediate.
template <typename R, int N, int V, int PK>
__global__ __launch_bounds__((FastLayout<R, N, V>::CB * (N / 16))) void f_sep_col(const ProbDev* __restrict__ probs,
                                                                          int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, B = LY::B, CB = LY::CB;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int g = tid % CB, t = tid / CB;
    TwKeep<R, N, (N == 1024 || N == 2048 || N == 4096) && sizeof(R) == 4, N == 4096 && sizeof(R) == 4 && PK == PRE_NONE> tk;
    tk.template init<CB * T>(lds + CB * Pow2Plan<N>::SLOTS, static_cast<const cpx<R>*>(probs[0].pcol.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_COL>(probs, nprob, tile)];
        const int kx = (tile - P.base_col) * CB + g;  // input column
        const int nx = P.nx, ny = P.ny, oy = P.oy;
        const bool colok = kx < nx;
        const cpx<R>* in = static_cast<const cpx<R>*>(P.in);
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int y = t + m * T;
            v[m] = {R(0), R(0)};
            if (colok && y >= oy && y < oy + ny) v[m] = at32(in, (unsigned)(y - oy) * nx + kx);
        }
        if (PK == PRE_MASK || (PK == PRE_MASKS && P.ops.pre == PRE_MASK)) {
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int y = t + m * T;
                if (colok && y >= oy && y < oy + ny) v[m] = apply_mask(P.mask, kx, y - oy, v[m]);
            }
        }
        if (PK == PRE_TMUL || (PK == PRE_MASKS && P.ops.pre == PRE_TMUL)) {
            const cpx<R>* tr = static_cast<const cpx<R>*>(P.ops.trans);
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int y = t + m * T;
                if (colok && y >= oy && y < oy + ny) {
                    cpx<R> w;
                    w = at32(tr, (unsigned)(y - oy) * nx + kx);
                    v[m] = tmul(v[m], w);
                }
            }
        }
        // the engine validates the masked field
        bool bad = false;
#pragma unroll
        for (int m = 0; m < 16; ++m) bad |= !finite_c(v[m]);
        if (bad) atomicOr(P.flag, 1);
        const cpx<R>* tw = static_cast<const cpx<R>*>(P.pcol.tw);
        fft_regs<R, N, CB, -1>(v, lds, g, t, tw, tk);
        const cpx<R>* hy = static_cast<const cpx<R>*>(P.eng.hy);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            cpx<R> h;
            h = at32(hy, t + m * T);
            v[m] = cmul(v[m], h);
        }
        fft_regs<R, N, CB, +1>(v, lds, g, t, tw, tk);
        if (colok) {
            cpx<R>* work = static_cast<cpx<R>*>(P.work);
            const unsigned b0 = ((unsigned)(kx / B) * ny - oy) * B + kx % B;
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int y = t + m * T;
                if (y >= oy && y < oy + ny) at32(work, b0 + (unsigned)y * B) = v[m];
            }
        }
    }
}

// Pass B: G rows per tile from the blocked intermediate (ox is a multiple
// of B), x factor, crop, 1/(P_x P_y), natural-layout output.
template <typename R, int N, int V>
__global__ __launch_bounds__((FastLayout<R, N, V>::G * (N / 16))) void f_sep_row(const ProbDev* __restrict__ probs,
                                                                         int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, G = LY::G, Q = LY::B;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int tb = tid % Q, g = (tid / Q) % G, tq = tid / (Q * G);
    const int t = tb + Q * tq;
    TwKeep<R, N, (N == 1024 || N == 2048 || N == 4096) && sizeof(R) == 4, N == 4096 && sizeof(R) == 4> tk;
    tk.template init<G * T>(lds + G * RowStride<N, G>::value, static_cast<const cpx<R>*>(probs[0].prow.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_ROW>(probs, nprob, tile)];
        const int iy = (tile - P.base_row) * G + g;
        const bool rowok = iy < P.ny;
        const int ny = P.ny, nx = P.nx, ox = P.ox;
        // sample x = t + m T is column ix = x - ox: block (x - ox)/Q, lane tb
        const cpx<R>* work = static_cast<const cpx<R>*>(P.work);
        const unsigned b0 = ((unsigned)tq * ny + iy) * Q + tb - (unsigned)(ox / Q) * ny * Q;
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int x = t + m * T;
            v[m] = {R(0), R(0)};
            if (rowok && x >= ox && x < ox + nx) v[m] = at32(work, b0 + (unsigned)(m * (T / Q)) * ny * Q);
        }
        const cpx<R>* tw = static_cast<const cpx<R>*>(P.prow.tw);
        fft_regs<R, N, G, -1, true>(v, lds, g, t, tw, tk);
        const cpx<R>* hx = static_cast<const cpx<R>*>(P.eng.hx);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            cpx<R> h;
            h = at32(hx, t + m * T);
            v[m] = cmul(v[m], h);
        }
        fft_regs<R, N, G, +1, true>(v, lds, g, t, tw, tk);
        const R s = (R)P.scale;
        bool bad = false;
        if (rowok) {
            cpx<R>* out = static_cast<cpx<R>*>(P.out);
            const unsigned o0 = (unsigned)iy * nx;
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int x = t + m * T - ox;
                if (x >= 0 && x < nx) {
                    const cpx<R> o = {v[m].re * s, v[m].im * s};
                    bad |= !finite_c(o);
                    at32(out, o0 + x) = o;
                }
            }
        }
        if (bad) atomicOr(P.flag, 2);
    }
}

/* ------------------------------------------------------------------ */
/* Two consecutive propagations of a trace, linked through their rows.  */
/* ------------------------------------------------------------------ */
// Propagation k's inverse row pass and propagation k+1's forward row pass
// on the same grid and padded length: per row, inverse row transform, crop
// and scale (propagation k's output, written only where A.out is set), the
// next step's pre-op PK (mask or shared transmission, from B), re-embed at
// the same offset, forward row transform, blocked store into B's
// intermediate for B's column pass.  Row tiles follow B (both batches
// share the layout).
template <typename R, int N, int V, int PK>
__global__ __launch_bounds__((FastLayout<R, N, V>::G * (N / 16))) void f_row_link(const ProbDev* __restrict__ A,
                                                                          const ProbDev* __restrict__ Bp,
                                                                          int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, G = LY::G, Q = LY::B;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int tb = tid % Q, g = (tid / Q) % G, tq = tid / (Q * G);
    const int t = tb + Q * tq;
    // (at 2048 the 1024-thread layouts V0/V2 spill more with the table)
    TwKeep<R, N, (N == 1024 || N == 4096 || (N == 2048 && (V == 1 || V == 3))) && sizeof(R) == 4,
           N == 4096 && sizeof(R) == 4 && PK == PRE_NONE>
        tk;
    tk.template init<G * T>(lds + G * RowStride<N, G>::value, static_cast<const cpx<R>*>(A[0].prow.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const int pi = find_problem<PASS_ROW>(Bp, nprob, tile);
        const ProbDev& P = A[pi];
        const ProbDev& Nx = Bp[pi];
        const int iy = (tile - Nx.base_row) * G + g;
        const bool rowok = iy < P.ny;
        const int ny = P.ny, nx = P.nx, ox = P.ox;
        const unsigned b0 = ((unsigned)tq * ny + iy) * Q + tb;
        cpx<R> v[16];
        const cpx<R>* work = static_cast<const cpx<R>*>(P.work);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            v[m] = {R(0), R(0)};
            if (rowok) v[m] = at32(work, b0 + (unsigned)(m * (T / Q)) * ny * Q);
        }
        const cpx<R>* tw = static_cast<const cpx<R>*>(P.prow.tw);
        fft_regs<R, N, G, +1, true>(v, lds, g, t, tw, tk);
        const R s = (R)P.scale;
        cpx<R>* mid = static_cast<cpx<R>*>(P.out);
        bool bad = false;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int x = t + m * T - ox;
            const bool in = rowok && x >= 0 && x < nx;
            const cpx<R> o = {v[m].re * s, v[m].im * s};
            v[m] = in ? o : cpx<R>{R(0), R(0)};
            bad |= in && !finite_c(o);
            if (in && mid) at32(mid, (unsigned)iy * nx + x) = o;
        }
        if (bad) {
            atomicOr(P.flag, 2);
            atomicOr(Nx.flag, 1);
        }
        row_pre<R, N, PK>(Nx, v, iy, rowok, t);
        fft_regs<R, N, G, -1, true>(v, lds, g, t, tw, tk);
        if (rowok) {
            cpx<R>* nw = static_cast<cpx<R>*>(Nx.work);
#pragma unroll
            for (int m = 0; m < 16; ++m) at32(nw, b0 + (unsigned)(m * (T / Q)) * ny * Q) = v[m];
        }
    }
}

/* ------------------------------------------------------------------ */
/* A propagation chained into a complex64 separable Fresnel.            */
/* ------------------------------------------------------------------ */
// The first propagation's inverse row pass and the Fresnel's row pass touch
// the same rows, and the Fresnel embeds the first result at the same offset
// (same grid, same padded length), so one row kernel does both: inverse row
// transform, crop and scale (the first propagation's output, written only
// when the caller asks for it), forward row transform, x factor, inverse,
// crop, blocked store for the Fresnel's column pass.  That pass then runs
// on contiguous slabs and writes the natural-layout output.  A and F are
// the two stages' descriptors, problem by problem; row tiles follow F (the
// Fresnel batch's row layout), and A's intermediate must be blocked with
// this layout's width B.
// HW: every problem is N = P/2 on both axes, so the window is exactly rows
// (here: samples) m = 4..11 of each thread's sixteen; the zeros outside it
// are compile-time constants, and the compiler drops the forward
// transform's first-stage work on them and the inverse's last-stage work on
// outputs never stored (VALU 2458 -> 2260 per tile at N = 4096).  The
// column pass has the same property but ran 3% slower specialised, so it
// keeps the runtime window.
template <typename R, int N, int V, bool HW = false>
__global__ __launch_bounds__((FastLayout<R, N, V>::G * (N / 16))) void f_row_chain(const ProbDev* __restrict__ A,
                                                                           const ProbDev* __restrict__ F,
                                                                           int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, G = LY::G, Q = LY::B;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int tb = tid % Q, g = (tid / Q) % G, tq = tid / (Q * G);
    const int t = tb + Q * tq;
    // table and hoist in the half-window form at N = 4096 (117 VGPRs; the
    // pass went from 0.72 to 0.62 ms on the 16-band 2048^2 chain)
    TwKeep<R, N, HW && N == 4096, HW && N == 4096> tk;
    tk.template init<G * T>(lds + G * RowStride<N, G>::value, static_cast<const cpx<R>*>(A[0].prow.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const int pi = find_problem<PASS_ROW>(F, nprob, tile);  // F's row tiles (this kernel's G)
        const ProbDev& P = A[pi];
        const ProbDev& Fp = F[pi];
        const int iy = (tile - Fp.base_row) * G + g;
        const bool rowok = iy < P.ny;
        const int ny = P.ny, nx = P.nx, ox = P.ox;
        const cpx<R>* work = static_cast<const cpx<R>*>(P.work);
        const unsigned b0 = ((unsigned)tq * ny + iy) * Q + tb;
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            v[m] = {R(0), R(0)};
            if (rowok) v[m] = at32(work, b0 + (unsigned)(m * (T / Q)) * ny * Q);
        }
        const cpx<R>* tw = static_cast<const cpx<R>*>(P.prow.tw);
        fft_regs<R, N, G, +1, true>(v, lds, g, t, tw, tk);
        // the first propagation's output, re-embedded where the Fresnel puts it
        const R s = (R)P.scale;
        cpx<R>* mid = static_cast<cpx<R>*>(P.out);
        bool bad = false;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int x = t + m * T - ox;
            const bool in = HW ? ((m >= 4 && m < 12) && rowok) : (rowok && x >= 0 && x < nx);
            const cpx<R> o = {v[m].re * s, v[m].im * s};
            v[m] = in ? o : cpx<R>{R(0), R(0)};
            bad |= in && !finite_c(o);
            if (in && mid) at32(mid, (unsigned)iy * nx + x) = o;
        }
        if (bad) {
            atomicOr(P.flag, 2);
            atomicOr(Fp.flag, 1);
        }
        fft_regs<R, N, G, -1, true>(v, lds, g, t, tw, tk);
        const cpx<R>* hx = static_cast<const cpx<R>*>(Fp.eng.hx);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            cpx<R> h;
            h = at32(hx, t + m * T);
            v[m] = cmul(v[m], h);
        }
        fft_regs<R, N, G, +1, true>(v, lds, g, t, tw, tk);
        // column c = x - ox goes to block c / Q, lane c % Q = tb (ox % Q == 0)
        if (rowok) {
            cpx<R>* fw = static_cast<cpx<R>*>(Fp.work);
            const unsigned w0 = ((unsigned)tq * ny + iy) * Q + tb - (unsigned)(ox / Q) * ny * Q;
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int x = t + m * T;
                if (HW ? (m >= 4 && m < 12) : (x >= ox && x < ox + nx))
                    at32(fw, w0 + (unsigned)(m * (T / Q)) * ny * Q) = v[m];
            }
        }
    }
}

// The chained Fresnel's column pass: CB columns per tile from the blocked
// intermediate, y factor, crop, 1/(P_x P_y), natural-layout output.
template <typename R, int N, int V, bool HW = false>
__global__ __launch_bounds__((FastLayout<R, N, V>::CB * (N / 16))) void f_sep_col_fin(
    const ProbDev* __restrict__ probs, int nprob, int total) {
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16, B = LY::B, CB = LY::CB;
    extern __shared__ unsigned char smem_raw[];
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int g = tid % CB, t = tid / CB;
    // (keeping the twiddles here measured no faster: 0.51-0.52 ms either way)
    TwKeep<R, N> tk;
    tk.template init<CB * T>(lds + CB * Pow2Plan<N>::SLOTS, static_cast<const cpx<R>*>(probs[0].pcol.tw), t);
    for (int tile = xcd_block(); tile < total; tile += gridDim.x) {
        const ProbDev& P = probs[find_problem<PASS_COL>(probs, nprob, tile)];
        const int kx = (tile - P.base_col) * CB + g;
        const int nx = P.nx, ny = P.ny, oy = P.oy;
        const bool colok = kx < nx;
        const cpx<R>* work = static_cast<const cpx<R>*>(P.work);
        // element y of column kx: b0 + y * B
        const unsigned b0 = ((unsigned)(kx / B) * ny - oy) * B + kx % B;
        cpx<R> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int y = t + m * T;
            v[m] = {R(0), R(0)};
            if ((HW ? (m >= 4 && m < 12) : (y >= oy && y < oy + ny)) && colok) v[m] = at32(work, b0 + (unsigned)y * B);
        }
        const cpx<R>* tw = static_cast<const cpx<R>*>(P.pcol.tw);
        fft_regs<R, N, CB, -1>(v, lds, g, t, tw, tk);
        const cpx<R>* hy = static_cast<const cpx<R>*>(P.eng.hy);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            cpx<R> h;
            h = at32(hy, t + m * T);
            v[m] = cmul(v[m], h);
        }
        fft_regs<R, N, CB, +1>(v, lds, g, t, tw, tk);
        const R s = (R)P.scale;
        bool bad = false;
        if (colok) {
            cpx<R>* out = static_cast<cpx<R>*>(P.out);
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int y = t + m * T;
                if (HW ? (m >= 4 && m < 12) : (y >= oy && y < oy + ny)) {
                    const cpx<R> o = {v[m].re * s, v[m].im * s};
                    bad |= !finite_c(o);
                    at32(out, (unsigned)(y - oy) * nx + kx) = o;
                }
            }
        }
        if (bad) atomicOr(P.flag, 2);
    }
}

}  // namespace ocd

namespace ocl {
using namespace ocd;

#ifdef OC_FAST_DOUBLE
using Real = double;
#define OC_FAST_FN(name) name##_c128
#else
using Real = float;
#define OC_FAST_FN(name) name##_c64
#endif

// Whether the half-window ASM column pass stays within its register budget:
// in complex64 the 1024-thread layout V0 spills at N = 2048 and 8192 with it
// (56 and 24 bytes) where the runtime-window kernel does not, so those run
// the runtime window.
template <typename R, int N, int V>
constexpr bool col_half_fits() {
    return !(sizeof(R) == 4 && V == 0 && (N == 2048 || N == 8192));
}

template <int N, int V>
static hipError_t launch_nv(const FastSel& s, const ProbDev* d, int np, int total, int grid, hipStream_t st) {
    using R = Real;
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16;
    // exchange buffer, then the kept twiddle table
    const size_t row_lds = (size_t)LY::G * RowStride<N, LY::G>::value * sizeof(cpx<R>) + tw_keep_bytes<R, N>(true);
    const size_t col_lds = (size_t)LY::CB * Pow2Plan<N>::SLOTS * sizeof(cpx<R>) + tw_keep_bytes<R, N>(true);
    const dim3 brow(LY::G * T), bcol(LY::CB * T), g(grid);
    if (s.sep) {  // separable complex64 Fresnel (pass 1: columns, pass 2: rows)
        if constexpr (sizeof(R) == 4) {
            if (s.pass == 1) {
                if (s.pre_kind == PRE_MASK)
                    hipLaunchKernelGGL((f_sep_col<R, N, V, PRE_MASK>), g, bcol, col_lds, st, d, np, total);
                else if (s.pre_kind == PRE_TMUL)
                    hipLaunchKernelGGL((f_sep_col<R, N, V, PRE_TMUL>), g, bcol, col_lds, st, d, np, total);
                else if (s.pre_kind == PRE_MASKS)
                    hipLaunchKernelGGL((f_sep_col<R, N, V, PRE_MASKS>), g, bcol, col_lds, st, d, np, total);
                else if (s.pre_kind == PRE_NONE)
                    hipLaunchKernelGGL((f_sep_col<R, N, V, PRE_NONE>), g, bcol, col_lds, st, d, np, total);
                else
                    return hipErrorInvalidValue;  // the two-pass form has no other pre-op
                return hipGetLastError();
            }
            if (s.pass == 2) {
                hipLaunchKernelGGL((f_sep_row<R, N, V>), g, brow, row_lds, st, d, np, total);
                return hipGetLastError();
            }
        }
        return hipErrorInvalidValue;
    }
    switch (s.pass) {
        case 0:
            if (s.pre_kind == PRE_NONE)
                hipLaunchKernelGGL((f_row_fwd<R, N, V, PRE_NONE>), g, brow, row_lds, st, d, np, total);
            else if (s.pre_kind == PRE_MASK)
                hipLaunchKernelGGL((f_row_fwd<R, N, V, PRE_MASK>), g, brow, row_lds, st, d, np, total);
            else if (s.pre_kind == PRE_TMUL && s.half)
                hipLaunchKernelGGL((f_row_fwd_pf<R, N, V>), g, brow, row_lds, st, d, np, total);
            else if (s.pre_kind == PRE_TMUL)
                hipLaunchKernelGGL((f_row_fwd<R, N, V, PRE_TMUL>), g, brow, row_lds, st, d, np, total);
            else if (s.pre_kind == PRE_POLAR)
                hipLaunchKernelGGL((f_row_fwd<R, N, V, PRE_POLAR>), g, brow, row_lds, st, d, np, total);
            else if (s.pre_kind == PRE_MASKS)
                hipLaunchKernelGGL((f_row_fwd<R, N, V, PRE_MASKS>), g, brow, row_lds, st, d, np, total);
            else
                hipLaunchKernelGGL((f_row_fwd<R, N, V, -1>), g, brow, row_lds, st, d, np, total);
            break;
        case 1:
            if (s.rs && s.half)
                hipLaunchKernelGGL((f_col<R, N, V, true, -1, true>), g, bcol, col_lds, st, d, np, total);
            else if (s.rs)
                hipLaunchKernelGGL((f_col<R, N, V, true, -1>), g, bcol, col_lds, st, d, np, total);
            else if (s.col_kind == ENG_ASM && s.half && col_half_fits<R, N, V>())
                hipLaunchKernelGGL((f_col<R, N, V, false, ENG_ASM, true>), g, bcol, col_lds, st, d, np, total);
            else if (s.col_kind == ENG_ASM)
                hipLaunchKernelGGL((f_col<R, N, V, false, ENG_ASM>), g, bcol, col_lds, st, d, np, total);
            else if (s.col_kind == ENG_FRESNEL)
                hipLaunchKernelGGL((f_col<R, N, V, false, ENG_FRESNEL>), g, bcol, col_lds, st, d, np, total);
            else if (s.col_kind == ENG_GS_PROJECT)
                hipLaunchKernelGGL((f_col<R, N, V, false, ENG_GS_PROJECT>), g, bcol, col_lds, st, d, np, total);
            else
                hipLaunchKernelGGL((f_col<R, N, V, false, -1>), g, bcol, col_lds, st, d, np, total);
            break;
        case 2:
            if (s.post_kind == POST_NONE)
                hipLaunchKernelGGL((f_row_inv<R, N, V, POST_NONE>), g, brow, row_lds, st, d, np, total);
            else if (s.post_kind == POST_GS_PHASE)
                hipLaunchKernelGGL((f_row_inv<R, N, V, POST_GS_PHASE>), g, brow, row_lds, st, d, np, total);
            else if (s.post_kind == POST_ADJ_GRAD)
                hipLaunchKernelGGL((f_row_inv<R, N, V, POST_ADJ_GRAD>), g, brow, row_lds, st, d, np, total);
            else
                hipLaunchKernelGGL((f_row_inv<R, N, V, -1>), g, brow, row_lds, st, d, np, total);
            break;
        case 3: {  // the kernel's row spectra (total = row tiles), then its distinct columns
            hipLaunchKernelGGL((f_rs_kgen<R, N, V>), g, brow, row_lds, st, d, np, total);
            // persistent: at most the grid the caller chose when it capped
            // the row tiles, else one workgroup per tile up to 2048
            const int kt = np * ((N / 2 + 1 + LY::CB - 1) / LY::CB);
            const int kg = grid < total ? std::min(kt, grid) : std::min(kt, std::max(grid, 2048));
            hipLaunchKernelGGL((f_rs_kcol<R, N, V>), dim3(kg), bcol, col_lds, st, d, np, kt);
            break;
        }
        case 4: hipLaunchKernelGGL((f_row_gs<R, N, V, true>), g, brow, row_lds, st, d, np, total); break;
        case 5: hipLaunchKernelGGL((f_row_gs<R, N, V, false>), g, brow, row_lds, st, d, np, total); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// which = 0: f_row_chain over (A, F); 1: f_sep_col_fin over F; 2, 3: the same
// for batches whose problems are all N = P/2.  complex64.
template <int N, int V>
static hipError_t launch_chain_nv(int which, const ProbDev* a, const ProbDev* f, int np, int total, int grid,
                                  hipStream_t st) {
    using R = Real;
    if constexpr (sizeof(R) == 4) {
        using LY = FastLayout<R, N, V>;
        constexpr int T = N / 16;
        const dim3 g(grid);
        const size_t rlds = (size_t)LY::G * RowStride<N, LY::G>::value * sizeof(cpx<R>) + tw_keep_bytes<R, N>(true);
        const size_t clds = (size_t)LY::CB * Pow2Plan<N>::SLOTS * sizeof(cpx<R>) + tw_keep_bytes<R, N>(true);
        switch (which) {  // + 2: every problem is N = P/2
            case 0: hipLaunchKernelGGL((f_row_chain<R, N, V>), g, dim3(LY::G * T), rlds, st, a, f, np, total); break;
            case 1: hipLaunchKernelGGL((f_sep_col_fin<R, N, V>), g, dim3(LY::CB * T), clds, st, f, np, total); break;
            case 2:
                hipLaunchKernelGGL((f_row_chain<R, N, V, true>), g, dim3(LY::G * T), rlds, st, a, f, np, total);
                break;
            case 3:
                hipLaunchKernelGGL((f_sep_col_fin<R, N, V, true>), g, dim3(LY::CB * T), clds, st, f, np, total);
                break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    (void)a; (void)f; (void)np; (void)total; (void)grid; (void)st; (void)which;
    return hipErrorInvalidValue;
}

// f_row_link over (a, b) with b's pre-op pk (PRE_NONE, PRE_MASK, PRE_TMUL,
// PRE_MASKS); both precisions.
template <int N, int V>
static hipError_t launch_link_nv(int pk, const ProbDev* a, const ProbDev* b, int np, int total, int grid,
                                 hipStream_t st) {
    using R = Real;
    using LY = FastLayout<R, N, V>;
    constexpr int T = N / 16;
    const size_t lds = (size_t)LY::G * RowStride<N, LY::G>::value * sizeof(cpx<R>) + tw_keep_bytes<R, N>(true);
    const dim3 g(grid), blk(LY::G * T);
    switch (pk) {
        case PRE_NONE: hipLaunchKernelGGL((f_row_link<R, N, V, PRE_NONE>), g, blk, lds, st, a, b, np, total); break;
        case PRE_MASK: hipLaunchKernelGGL((f_row_link<R, N, V, PRE_MASK>), g, blk, lds, st, a, b, np, total); break;
        case PRE_TMUL: hipLaunchKernelGGL((f_row_link<R, N, V, PRE_TMUL>), g, blk, lds, st, a, b, np, total); break;
        case PRE_MASKS: hipLaunchKernelGGL((f_row_link<R, N, V, PRE_MASKS>), g, blk, lds, st, a, b, np, total); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static hipError_t allow_lds(const void* kernel) {
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

template <int N, int V>
static hipError_t attr_nv() {
    using R = Real;
    const void* k[] = {(const void*)f_row_fwd<R, N, V, PRE_NONE>,
                       (const void*)f_row_fwd<R, N, V, PRE_MASK>,
                       (const void*)f_row_fwd<R, N, V, PRE_TMUL>,
                       (const void*)f_row_fwd_pf<R, N, V>,
                       (const void*)f_row_fwd<R, N, V, PRE_POLAR>,
                       (const void*)f_row_fwd<R, N, V, PRE_MASKS>,
                       (const void*)f_row_fwd<R, N, V, -1>,
                       (const void*)f_col<R, N, V, true, -1>,
                       (const void*)f_col<R, N, V, true, -1, true>,
                       (const void*)f_col<R, N, V, false, ENG_ASM>,
                       (const void*)f_col<R, N, V, false, ENG_ASM, true>,
                       (const void*)f_col<R, N, V, false, ENG_FRESNEL>,
                       (const void*)f_col<R, N, V, false, ENG_GS_PROJECT>,
                       (const void*)f_col<R, N, V, false, -1>,
                       (const void*)f_row_inv<R, N, V, POST_NONE>,
                       (const void*)f_row_inv<R, N, V, POST_GS_PHASE>,
                       (const void*)f_row_inv<R, N, V, POST_ADJ_GRAD>,
                       (const void*)f_row_inv<R, N, V, -1>,
                       (const void*)f_rs_kgen<R, N, V>,
                       (const void*)f_rs_kcol<R, N, V>,
                       (const void*)f_row_gs<R, N, V, true>,
                       (const void*)f_row_gs<R, N, V, false>,
                       (const void*)f_row_link<R, N, V, PRE_NONE>,
                       (const void*)f_row_link<R, N, V, PRE_MASK>,
                       (const void*)f_row_link<R, N, V, PRE_TMUL>,
                       (const void*)f_row_link<R, N, V, PRE_MASKS>};
    for (const void* f : k) {
        const hipError_t e = allow_lds(f);
        if (e != hipSuccess) return e;
    }
    if constexpr (sizeof(R) == 4) {
        const void* ks[] = {(const void*)f_sep_col<R, N, V, PRE_NONE>, (const void*)f_sep_col<R, N, V, PRE_MASK>,
                            (const void*)f_sep_col<R, N, V, PRE_TMUL>, (const void*)f_sep_col<R, N, V, PRE_MASKS>,
                            (const void*)f_sep_row<R, N, V>,         (const void*)f_row_chain<R, N, V>,
                            (const void*)f_sep_col_fin<R, N, V>,     (const void*)f_row_chain<R, N, V, true>,
                            (const void*)f_sep_col_fin<R, N, V, true>};
        for (const void* f : ks) {
            const hipError_t e = allow_lds(f);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// Sizes with one layout (64..512) and with two (the 512-thread cap only
// changes the layout when N/16 * 16 > 512, i.e. N >= 1024).
#ifndef OC_FAST_SIZES_1
#define OC_FAST_SIZES_1(X) X(64) X(128) X(256) X(512)
#define OC_FAST_SIZES_2(X) X(1024) X(2048) X(4096) X(8192)
#endif

hipError_t OC_FAST_FN(fast_launch)(const FastSel& s, const ProbDev* d, int np, int total, int grid, hipStream_t st) {
    switch (s.n) {
#define OC_C1(NN) \
    case NN: return launch_nv<NN, 0>(s, d, np, total, grid, st);
        OC_FAST_SIZES_1(OC_C1)
#undef OC_C1
#define OC_C2(NN)                                                  \
    case NN:                                                       \
        switch (s.v) {                                             \
            case 1: return launch_nv<NN, 1>(s, d, np, total, grid, st); \
            case 2: return launch_nv<NN, 2>(s, d, np, total, grid, st); \
            case 3: return launch_nv<NN, 3>(s, d, np, total, grid, st); \
            default: return launch_nv<NN, 0>(s, d, np, total, grid, st); \
        }
        OC_FAST_SIZES_2(OC_C2)
#undef OC_C2
        default: return hipErrorInvalidValue;
    }
}

hipError_t OC_FAST_FN(fast_launch_link)(int pk, int n, int v, const ProbDev* a, const ProbDev* b, int np, int total,
                                        int grid, hipStream_t st) {
    switch (n) {
#define OC_K1(NN) \
    case NN: return launch_link_nv<NN, 0>(pk, a, b, np, total, grid, st);
        OC_FAST_SIZES_1(OC_K1)
#undef OC_K1
#define OC_K2(NN)                                                              \
    case NN:                                                                   \
        switch (v) {                                                           \
            case 1: return launch_link_nv<NN, 1>(pk, a, b, np, total, grid, st); \
            case 2: return launch_link_nv<NN, 2>(pk, a, b, np, total, grid, st); \
            case 3: return launch_link_nv<NN, 3>(pk, a, b, np, total, grid, st); \
            default: return launch_link_nv<NN, 0>(pk, a, b, np, total, grid, st); \
        }
        OC_FAST_SIZES_2(OC_K2)
#undef OC_K2
        default: return hipErrorInvalidValue;
    }
}

// The chained propagation's fused row pass (which = 0, descriptors a and f)
// and final Fresnel column pass (which = 1, descriptors f); complex64 only.
hipError_t OC_FAST_FN(fast_launch_chain)(int which, int n, int v, const ProbDev* a, const ProbDev* f, int np,
                                         int total, int grid, hipStream_t st) {
    switch (n) {
#define OC_H1(NN) \
    case NN: return launch_chain_nv<NN, 0>(which, a, f, np, total, grid, st);
        OC_FAST_SIZES_1(OC_H1)
#undef OC_H1
#define OC_H2(NN)                                                                  \
    case NN:                                                                       \
        switch (v) {                                                               \
            case 1: return launch_chain_nv<NN, 1>(which, a, f, np, total, grid, st); \
            case 2: return launch_chain_nv<NN, 2>(which, a, f, np, total, grid, st); \
            case 3: return launch_chain_nv<NN, 3>(which, a, f, np, total, grid, st); \
            default: return launch_chain_nv<NN, 0>(which, a, f, np, total, grid, st); \
        }
        OC_FAST_SIZES_2(OC_H2)
#undef OC_H2
        default: return hipErrorInvalidValue;
    }
}

template <int N, int V>
static void layout_nv(int* G, int* B, int* CB) {
    using LY = FastLayout<Real, N, V>;
    *G = LY::G;
    *B = LY::B;
    *CB = LY::CB;
}

// Workgroup layout of length n, variant v (all 0 when there is no kernel).
void OC_FAST_FN(fast_layout)(int n, int v, int* G, int* B, int* CB) {
    *G = *B = *CB = 0;
    switch (n) {
#define OC_L1(NN) \
    case NN: layout_nv<NN, 0>(G, B, CB); return;
        OC_FAST_SIZES_1(OC_L1)
#undef OC_L1
#define OC_L2(NN)                                      \
    case NN:                                           \
        switch (v) {                                   \
            case 1: layout_nv<NN, 1>(G, B, CB); return; \
            case 2: layout_nv<NN, 2>(G, B, CB); return; \
            case 3: layout_nv<NN, 3>(G, B, CB); return; \
            default: layout_nv<NN, 0>(G, B, CB); return; \
        }
        OC_FAST_SIZES_2(OC_L2)
#undef OC_L2
        default: return;
    }
}

hipError_t OC_FAST_FN(fast_set_lds_limits)() {
    hipError_t e;
#define OC_A1(NN) \
    if ((e = attr_nv<NN, 0>())) return e;
    OC_FAST_SIZES_1(OC_A1)
    OC_FAST_SIZES_2(OC_A1)
#undef OC_A1
#define OC_A2(NN)                                                                        \
    if ((e = attr_nv<NN, 1>()) || (e = attr_nv<NN, 2>()) || (e = attr_nv<NN, 3>())) return e;
    OC_FAST_SIZES_2(OC_A2)
#undef OC_A2
    return hipSuccess;
}

}  // namespace ocl
