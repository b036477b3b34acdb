// Register-resident power-of-two FFTs for gfx950 with compile-time plans.
//
// A group of G transforms of length N runs on G*N/16 threads; thread (g, t)
// keeps 16 elements in registers.  On entry it holds x[t + m*T] (T = N/16,
// m < 16) and on exit the transform X[t + m*T]: natural order on both sides,
// so callers load and store straight from/to global memory and a forward
// transform can be followed by an element-wise op and an inverse transform
// without any exchange in between.  Stages are radix 16 (Stockham, input
// pattern t + r*T) with a final radix-L stage (L = N / 16^k in {2, 4, 8}) that
// works on the thread's own registers; consecutive stages exchange through
// LDS laid out as slot(idx)*G + g with one pad element per 16.
#pragma once

#include "fft_device.hpp"

namespace ocd {

template <int N>
struct Pow2Plan {
    static constexpr int log2n() {
        int l = 0;
        while ((1 << l) < N) ++l;
        return l;
    }
    static constexpr int LOG = log2n();
    static constexpr int NS16 = LOG / 4;
    static constexpr int L = 1 << (LOG % 4);
    static constexpr int T = N / 16;
    static constexpr int NSTAGES = NS16 + (L > 1 ? 1 : 0);
    static constexpr int SLOTS = N + N / 16;  // LDS elements per transform
    // Row-major exchange stride (transform g at g * stride) for workgroups of
    // four or more transforms: SLOTS rounded up to 8 mod 32 elements, which
    // puts consecutive transforms 64 bytes (complex64) apart in the bank
    // space.  RowStride picks the stride per workgroup size.
    static constexpr int RSTRIDE = SLOTS + (40 - SLOTS % 32) % 32;
    static_assert((1 << LOG) == N && N >= 32, "power of two >= 32");
};

// The row layout's stride for a workgroup of G transforms (the row kernels'
// lanes run (tb, g, tq), so a wave holds a few slots of each transform).
// Chosen by measurement on the complex64 4096 row passes over the residues
// 4..28 mod 32: 8 mod 32 for G >= 4 (0.443-0.448 ms against 0.455-0.458 for the
// others) and 16 mod 32 for G = 2 (0.607 against 0.611-0.64 ms).  The LDS
// bank-conflict counter does not track these times: the 16 mod 32 choice
// for G = 2 counts twice the conflict cycles of 8 mod 32, and the 4 mod 32
// choice for G = 4 half those of 8 mod 32.
template <int N, int G>
struct RowStride {
    static constexpr int S = Pow2Plan<N>::SLOTS;
    static constexpr int value = G >= 4 ? Pow2Plan<N>::RSTRIDE : S + (48 - S % 32) % 32;
};

__device__ __forceinline__ int fslot(int idx) { return idx + (idx >> 4); }

// Stage-output scatter to LDS for radix-16 stage with stride Ns (a power of
// 16, a constant once the stage loop is unrolled), then the next stage's
// gather x[t + m*T].  Both sides are written as one per-thread base plus
// compile-time steps: with Ns = 1, slot(16 t + r) = 17 t + r; with Ns a
// multiple of 16, slot(b + r Ns) = slot(b) + r (Ns + Ns/16); likewise
// slot(t + m T) = slot(t) + m (T + T/16) when 16 | T.  The accesses then use
// immediate offsets from a single address register instead of sixteen.
// RM: transform g's slots are contiguous at g * RowStride (row kernels, whose
// lanes run along a transform first); else interleaved, slot * G + g
// (column kernels, whose lanes run across transforms first).
template <typename R, int N, int G, bool RM = false>
__device__ __forceinline__ void fft_exchange16(cpx<R> (&v)[16], cpx<R>* __restrict__ lds, int g, int t, int Ns) {
    constexpr int T = N / 16;
    constexpr int SG = RM ? 1 : G;                              // element step of one slot
    cpx<R>* __restrict__ own = lds + (RM ? g * RowStride<N, G>::value : g);  // this transform's slot 0
    const int k = t & (Ns - 1);
    const int base = (t - k) * 16 + k;
    const int wstep = Ns == 1 ? 1 : Ns + Ns / 16;
    cpx<R>* __restrict__ wp = own + (Ns == 1 ? 17 * t : fslot(base)) * SG;
#pragma unroll
    for (int r = 0; r < 16; ++r) wp[r * wstep * SG] = v[r];
    __syncthreads();
    if constexpr (T % 16 == 0) {
        const cpx<R>* __restrict__ rp = own + fslot(t) * SG;
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = rp[m * (T + T / 16) * SG];
    } else {
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = own[fslot(t + m * T) * SG];
    }
    __syncthreads();
}

// w^r for r = 1..15 (w[0] = 1) from the table entries w, w^2, w^4, w^8 with
// at most three extra products each.  Every radix-16 twiddle stage forms its
// factors this way, kept or not, so all kernels round them alike.
template <typename R>
__device__ __forceinline__ void tw_powers(cpx<R> w1, cpx<R> w2, cpx<R> w4, cpx<R> w8, cpx<R> (&w)[16]) {
    w[0] = {R(1), R(0)};
    w[1] = w1;
    w[2] = w2;
    w[3] = cmul(w1, w2);
    w[4] = w4;
    w[5] = cmul(w1, w4);
    w[6] = cmul(w2, w4);
    w[7] = cmul(w[3], w4);
    w[8] = w8;
#pragma unroll
    for (int r = 9; r < 16; ++r) w[r] = cmul(w[r - 8], w8);
}

// Twiddles a kernel keeps for all of its tiles (every transform of one
// launch has the same length N, so the same factors):
//  - TAB: the second radix-16 stage (Ns = 16) multiplies input r of
//    butterfly k = t mod 16 by w_k^r, w_k = W_N^{k N/256}; those 16 x 16
//    factors sit in LDS as [r][k] (the lanes of a wave read consecutive k:
//    no bank conflicts) and are read instead of formed per transform;
//  - HOIST: the third radix-16 stage (Ns = 256, N = 4096 and 8192) has k =
//    t mod 256, fixed for the thread, so its fifteen factors are formed once,
//    before the tile loop.
// Both hold the products of tw_powers, so a kernel gives the same bits with
// or without them; each kernel enables what its registers allow.  The
// inverse transform conjugates at use (the conjugate of a product of
// conjugates is that product, exactly).  `init` is called by every thread of
// the block before its first transform.
template <typename R, int N, bool TAB_OK = false, bool HOIST_OK = false>
struct TwKeep {
    static constexpr bool TAB = TAB_OK && Pow2Plan<N>::NS16 >= 2;
    static constexpr bool HOIST = HOIST_OK && Pow2Plan<N>::NS16 >= 3;
    static constexpr int TAB_ELEMS = TAB ? 256 : 0;
    const cpx<R>* tab = nullptr;
    cpx<R> w3[16];
    // `slot`: TAB_ELEMS elements of LDS beside the exchange buffer; `tw`: the
    // length-N table exp(-2 pi i e/N); t: the thread's index in its
    // transform; NT: the block size (a constant: with blockDim.x read inside
    // the fill loop the compiler took later uniform values, gridDim.x among
    // them, for divergent ones and moved the tile loop's descriptor reads to
    // vector loads)
    template <int NT>
    __device__ __forceinline__ void init(cpx<R>* slot, const cpx<R>* __restrict__ tw, int t) {
        if constexpr (TAB) {
            for (int i = (int)threadIdx.x; i < 256; i += NT) {
                const int r = i >> 4, e = (i & 15) * (N / 256);
                cpx<R> w[16];
                tw_powers(tw_at<-1>(tw, e), tw_at<-1>(tw, 2 * e), tw_at<-1>(tw, 4 * e), tw_at<-1>(tw, 8 * e), w);
                cpx<R> x = w[0];
#pragma unroll
                for (int q = 1; q < 16; ++q) x = q == r ? w[q] : x;
                slot[i] = x;
            }
            tab = slot;
            __syncthreads();
        }
        if constexpr (HOIST) {
            const int e = (t & 255) * (N / 4096);
            tw_powers(tw_at<-1>(tw, e), tw_at<-1>(tw, 2 * e), tw_at<-1>(tw, 4 * e), tw_at<-1>(tw, 8 * e), w3);
        }
    }
};
// LDS bytes a kernel with table `tab` adds for it (after the exchange buffer).
template <typename R, int N>
constexpr size_t tw_keep_bytes(bool tab) {
    return tab && Pow2Plan<N>::NS16 >= 2 ? (size_t)256 * sizeof(cpx<R>) : 0;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft_small(cpx<R>* w, int L) {
    if (L == 2) dft2<DIR>(w[0], w[1]);
    else if (L == 4) dft4<DIR>(w[0], w[1], w[2], w[3]);
    else dft8<DIR>(w);
}

// G transforms of length N in place in registers (see header).  `active`
// threads participate in the arithmetic; every thread of the block must call
// this (barriers).
template <typename R, int N, int G, int DIR, bool RM = false, bool KT = false, bool KH = false>
__device__ __forceinline__ void fft_regs(cpx<R> (&v)[16], cpx<R>* __restrict__ lds, int g, int t,
                                         const cpx<R>* __restrict__ tw, const TwKeep<R, N, KT, KH>& tk) {
    using PL = Pow2Plan<N>;
    using TK = TwKeep<R, N, KT, KH>;
    constexpr int T = PL::T;
#pragma unroll
    for (int s = 0; s < PL::NS16; ++s) {
        const int Ns = 1 << (4 * s);
        if (s == 1 && TK::TAB) {
            const cpx<R>* __restrict__ tb = tk.tab + (t & 15);
#pragma unroll
            for (int r = 1; r < 16; ++r) {
                cpx<R> w = tb[r * 16];
                if (DIR > 0) w.im = -w.im;
                v[r] = cmul(v[r], w);
            }
        } else if (s == 2 && TK::HOIST) {
#pragma unroll
            for (int r = 1; r < 16; ++r) {
                cpx<R> w = tk.w3[r];
                if (DIR > 0) w.im = -w.im;
                v[r] = cmul(v[r], w);
            }
        } else if (s > 0) {
            // the products of tw_powers, interleaved with their use
            const int k = t & (Ns - 1);
            const int e = k * (N / (Ns * 16));
            const cpx<R> w1 = tw_at<DIR>(tw, e), w2 = tw_at<DIR>(tw, 2 * e);
            const cpx<R> w4 = tw_at<DIR>(tw, 4 * e), w8 = tw_at<DIR>(tw, 8 * e);
            const cpx<R> w3 = cmul(w1, w2), w5 = cmul(w1, w4), w6 = cmul(w2, w4), w7 = cmul(w3, w4);
            v[1] = cmul(v[1], w1);
            v[2] = cmul(v[2], w2);
            v[3] = cmul(v[3], w3);
            v[4] = cmul(v[4], w4);
            v[5] = cmul(v[5], w5);
            v[6] = cmul(v[6], w6);
            v[7] = cmul(v[7], w7);
            v[8] = cmul(v[8], w8);
            v[9] = cmul(v[9], cmul(w1, w8));
            v[10] = cmul(v[10], cmul(w2, w8));
            v[11] = cmul(v[11], cmul(w3, w8));
            v[12] = cmul(v[12], cmul(w4, w8));
            v[13] = cmul(v[13], cmul(w5, w8));
            v[14] = cmul(v[14], cmul(w6, w8));
            v[15] = cmul(v[15], cmul(w7, w8));
        }
        dft16<DIR>(v);
        if (s + 1 < PL::NSTAGES) fft_exchange16<R, N, G, RM>(v, lds, g, t, Ns);
    }
    if constexpr (PL::L > 1) {
        constexpr int L = PL::L;
        constexpr int NB = 16 / L;  // butterflies per thread
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * T;
            cpx<R> w[L];
#pragma unroll
            for (int r = 0; r < L; ++r) w[r] = v[b + r * NB];
            const cpx<R> u1 = tw_at<DIR>(tw, j);
            w[1] = cmul(w[1], u1);
            if constexpr (L >= 4) {
                const cpx<R> u2 = tw_at<DIR>(tw, 2 * j);
                w[2] = cmul(w[2], u2);
                w[3] = cmul(w[3], cmul(u1, u2));
                if constexpr (L == 8) {
                    const cpx<R> u4 = tw_at<DIR>(tw, 4 * j);
                    const cpx<R> u3 = cmul(u1, u2);
                    w[4] = cmul(w[4], u4);
                    w[5] = cmul(w[5], cmul(u1, u4));
                    w[6] = cmul(w[6], cmul(u2, u4));
                    w[7] = cmul(w[7], cmul(u3, u4));
                }
            }
            dft_small<DIR>(w, L);
#pragma unroll
            for (int r = 0; r < L; ++r) v[b + r * NB] = w[r];
        }
    }
}

}  // namespace ocd
