// Shared host/device descriptors of the propagation passes.
#pragma once

#include <cstdint>

#include "fft_device.hpp"

namespace ocd {

enum EngineKind : int {
    ENG_ASM = 0,
    ENG_FRESNEL = 1,
    ENG_RS = 2,
    ENG_NONE = 3,        // plain 2-D FFT round trip (identity)
    ENG_GS_PROJECT = 4,  // Fourier-plane amplitude projection (FFT-form GS)
    ENG_MEASURE = 5,     // forward only: accumulate intensity statistics
    ENG_FRAUNHOFER = 6,  // forward only: far-field bin map and prefactor, written to the output
    ENG_FRESNEL_CONV = 7 // Fresnel convolution form: the RS pipeline with the paraxial kernel
};
// Engines that multiply by the spectrum of a sampled kernel (kernel
// generation pass + product in the column pass).
__host__ __device__ inline bool conv_engine(int kind) { return kind == ENG_RS || kind == ENG_FRESNEL_CONV; }
// Element-wise op fused into the first pass's loads.
// PRE_TMUL multiplies by a precomputed transmission (a phase mask shared by
// several problems of a launch, evaluated once: it does not depend on the
// wavelength).
// PRE_FRAUN multiplies by the Fraunhofer offset-DFT pre-phase and axis ramp.
enum PreOp : int { PRE_NONE = 0, PRE_MASK = 1, PRE_POLAR = 2, PRE_SEED_MSE = 3, PRE_TMUL = 4, PRE_FRAUN = 5 };
// Kernel selector (never a descriptor's op): a launch whose problems mix
// PRE_NONE, PRE_MASK and PRE_TMUL; each tile takes its problem's branch.
constexpr int PRE_MASKS = 6;
// Workgroup partials of the MseToTarget seed's merit sum (launch_seed_mse
// keeps them after the first 8 doubles of its sums buffer).
constexpr int kSeedParts = 1024;
// Element-wise op fused into the last pass's (cropped) stores.
enum PostOp : int { POST_NONE = 0, POST_GS_TARGET = 1, POST_GS_PHASE = 2, POST_ADJ_GRAD = 3 };
enum MaskMode : int { MASK_NONE = 0, MASK_SEPARABLE = 1, MASK_GENERAL = 2 };

// PhaseMask::apply ingredients (src/element.rs:1443-1484).  The separable form
// is used when psi == 0: element-frame coordinates are then exactly the
// centre-relative field coordinates, so lattice indices/weights are per axis.
// One field sample's position on one axis of an unrotated phase mask.
struct MaskAxis {
    double t;    // fractional lattice coordinate f - i0
    double r;    // field coordinate relative to the mask centre (stop test)
    int i0, i1;  // bracketing lattice indices (0, 0 outside the span)
    int in;      // 1 when the coordinate lies within the lattice span
    int pad;
};

struct MaskDev {
    int mode;
    int ap_kind;
    double ap_a, ap_b;
    // general form
    const double* values;  // lattice, row-major mny x mnx
    int mnx, mny;
    double mdx, mdy;
    double cpsi, spsi;
    double cx, cy;
    int fnx, fny;
    double fdx, fdy;
    // separable form, per field axis
    const MaskAxis* sx;
    const MaskAxis* sy;
};

struct EngineDev {
    int kind;
    int _pad;
    double kz;  // ASM: (2 pi/lam_n)*z ; Fresnel: (2 pi/lambda)*z
    double c;   // Fresnel: (pi*lambda)*z
    const double* ax;  // ASM: (lam_n*fx)^2, +inf where band-limited ; Fresnel: fx*fx
    const double* ay;
    // Fresnel, complex64 launches: the transfer function factored per axis,
    // H = hx[kx] hy[ky] with hx = exp(i (k z - c fx^2)), hy = exp(-i c fy^2)
    // (phases in f64 on the host, factors rounded to complex64).
    const void* hx;
    const void* hy;
    // Rayleigh-Sommerfeld kernel sampling
    double lambda, z, k, sgn, dx, dy;
    // Fraunhofer (fraunhofer.rs:376-437): origin indices c = (n - 1)/2, axis
    // ramp frequencies, far-field pitch and axis, pi/(lambda z), and the
    // prefactor e^{ikz}/(i lambda z)
    double fcx, fcy, fs, gs, dxo, dyo, xs, ys, chirp, quad_re, quad_im;
    // Fresnel convolution kernel (fresnel.rs:312-333): pref = e^{ikz}/(i lambda z) dx dy
    double pref_re, pref_im;
};

// What a launch of the power-of-two kernels holds uniformly, selecting the
// compile-time specialisation: pass (0 row_fwd, 1 col, 2 row_inv, 3 rs_kgen),
// length and layout variant, and the kinds every problem shares (-1: mixed).
struct FastSel {
    int pass;
    int n, v;
    int rs;          // Rayleigh-Sommerfeld batch
    int post_kind;   // POST_NONE / POST_GS_PHASE when uniform, else -1
    int col_kind;    // ENG_ASM / ENG_FRESNEL when uniform, else -1
    int pre_kind;    // PRE_NONE / PRE_MASK / PRE_TMUL / PRE_POLAR when uniform, else -1
    int sep;         // separable complex64 Fresnel (two passes)
    int half;        // every problem is N = P/2 on both axes (window = samples m = 4..11 of each thread)
};

// Parameters of the fused element-wise ops (gs.rs:484-564, adjoint.rs:200-236,
// merit.rs:143-165).  Real arrays have the transform's element type R unless
// noted; `sums` are per-problem f64 accumulators.
struct OpsDev {
    int pre, post;
    // PRE_POLAR: v = open[k] ? from_polar(amp[k], phase[k]) : 0
    const void* amp;
    const void* phase;
    const unsigned char* open;
    // PRE_TMUL: v = u * trans[k]   (element type R, field layout)
    const void* trans;
    // PRE_SEED_MSE: v = (2 (|u|^2 - target)/n) u   (target f64, n = frame size)
    const double* seed_target;
    double n_frame;  // frame size n as f64 (divisor)
    // POST_GS_TARGET: |U| vs target amplitude, relaxed projection
    const double* tamp;          // target amplitude (f64)
    const double* merit_target;  // MseToTarget target intensity (f64)
    double beta_f;
    double* err_out;  // |U| - T (f64, may be null)
    // POST_GS_PHASE: dphi = wrap_pi(arg V - phi); phi <- encode(phi + beta_p dphi)
    void* phase_io;  // element type R
    double beta_p, lo, hi;
    double* dphi_out;  // f64, may be null
    // POST_ADJ_GRAD: mu = output; V = ckpt * t; grad = 0 + -2 Im(conj(mu) V) (the element plane's one
    // term); out = conj(t) mu
    const void* ckpt;
    double* grad;
    // ENG_GS_PROJECT / ENG_MEASURE: Fourier-plane target amplitude, blocked
    // like the intermediate ([P_x/B][P_y][B], element type R)
    const void* tamp_blk;
    double* sums;  // accumulators (4 per problem)
    // when set, the last pass's GS sums go here in a fixed order instead of
    // into `sums`: wave w of row tile t (counted from the problem's first)
    // writes its four sums to parts[(t * waves + w) * 4 + i]
    double* parts;
};

struct ProbDev {
    int nx, ny, px, py, ox, oy;
    int rows_cnt;  // transforms per workgroup in the row passes
    int B;         // column block width of the intermediate layout
    int base_row, base_col, base_kgen;  // tile prefix offsets per pass
    int tiles_row, tiles_col, tiles_kgen;
    // set on every problem of a batch whose row pass reads one shared
    // transmission and whose problems have equal row tile counts: row tile
    // k of problem p is launch tile k*nprob + p, so that the problems' same
    // rows run together and share the transmission rows in L2
    int ilv;
    FftPlanDev prow, pcol;
    const void* in;
    void* out;
    void* work;  // N_y x P_x intermediate, blocked [P_x/B][N_y][B]
                 // (separable Fresnel: N_y x N_x, blocked [ceil(N_x/B)][N_y][B])
    void* kbuf;  // RS kernel row spectra, blocked [P_x/B][P_y][B]
    double scale;  // 1/(P_x P_y)
    MaskDev mask;
    EngineDev eng;
    OpsDev ops;
    int* flag;  // bit 0: non-finite input sample; bit 1: non-finite output
};

// Per-restart state of the device-side gs::run loop (gs.rs:482-578): the
// next iteration, the best-so-far and whether the last iteration improved.
struct GsState {
    unsigned long long it;
    unsigned long long best_it;
    double best_merit;
    int have_best;
    int improved;
};
// GsRecord (gs.rs:254-268), the layout of oc_gs_record.
struct GsRec {
    unsigned long long iteration;
    double merit, best_merit, amplitude_mse, target_error_rms, phase_update_rms;
};

// One batch of 1-D transforms over a buffer (the multi-pass path for padded
// lengths beyond the on-chip transform).  Line l starts at
// (l / inner) * outer + (l % inner) * inner_step and its element e sits e *
// elem further on (all in elements), in the input and output views.
struct LineView {
    long long outer, inner_step, elem;
};
struct LineSpec {
    int n;             // line length
    int cnt;           // lines per workgroup iteration
    long long count;   // lines
    long long inner;   // lines per outer index
    LineView in, out;
    int tw_len;        // > 0: output e times W_{tw_len}^{m e}, m = l / inner (tw_outer) or l % inner
    int tw_outer;
};

}  // namespace ocd
