/*
 * optics_gpu.h — C ABI of the MI355X (gfx950) propagation library.
 *
 * Each entry point replaces one public Rust function of the reference
 * (optics-core / optics-inverse); the comment above it names the file:line it
 * stands in for.  Conventions shared by every call:
 *
 *  - A complex buffer is interleaved (re, im): `double*` for complex128 and
 *    `float*` for complex64, row-major with sample (ix, iy) at index
 *    iy*nx + ix, exactly the layout of `Vec<Complex64>` in `Field`
 *    (src/field.rs:99-105).
 *  - Entry points without a `_dev` suffix take and return HOST memory, like
 *    the Rust API (a `&BandField` in, a new owned `BandField` out).  `_dev`
 *    entry points take device pointers and enqueue on the context's stream
 *    without synchronising.
 *  - Every call returns an `oc_status`.  Validation order and the z == 0
 *    identity follow the reference engines (src/propagation/asm.rs:439-479).
 *    Sampling reports are filled on every call that computes one, including
 *    the OC_ALIASING_RISK failure (the Rust error boxes the same report).
 *  - No entry point aborts or panics; device failures return OC_ERR_DEVICE
 *    and the message is available from oc_last_error().
 *
 * Threading: a context owns one HIP stream, a scratch arena and plan caches.
 * Every entry point holds its context's lock for the length of the call, so
 * calls on one context from several threads are serialised by the library
 * (in an unspecified order: a caller that needs an order between them, or
 * between queued batches, orders its own calls); separate contexts run
 * concurrently from separate threads.
 */
#ifndef OPTICS_GPU_H
#define OPTICS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes.  One code per variant of the reference error enums.         */
/* ------------------------------------------------------------------------ */
typedef enum oc_status {
    OC_OK = 0,
    /* PropagationError / FresnelError / RayleighSommerfeldError
       (src/propagation/asm.rs:218-250, fresnel.rs:225-254,
       rayleigh_sommerfeld.rs:197-223) */
    OC_INVALID_DISTANCE = 1,
    OC_INVALID_WAVELENGTH = 2,
    OC_INVALID_MEDIUM = 3,
    OC_EMPTY_GRID = 4,
    OC_NON_FINITE_SAMPLE = 5,
    OC_ALIASING_RISK = 6,
    /* FraunhoferError::RegimeViolation (src/propagation/fraunhofer.rs:246-252) */
    OC_REGIME_VIOLATION = 7,
    /* OpticsError (src/element.rs:60-170) */
    OC_INVALID_APERTURE = 20,
    OC_OUTSIDE_FIELD = 21,
    OC_INVALID_AXIS = 22,
    OC_INVALID_PHASE = 23,
    OC_INVALID_ORIENTATION = 24,
    OC_INVALID_GRID = 25,
    OC_INVALID_GRID_PITCH = 26,
    /* InverseError (crates/optics-inverse/src/error.rs:11-60) */
    OC_INV_NOT_FINITE = 40,
    OC_INV_OUT_OF_BOUNDS = 41,
    OC_INV_WRONG_DIMENSION = 42,
    OC_INV_CARRIER = 43,
    OC_INV_RENDER = 44,
    OC_INV_ADJOINT = 45,
    /* Library-level failures (no reference counterpart). */
    OC_ERR_INVALID_ARGUMENT = 100,
    OC_ERR_UNSUPPORTED = 101,
    OC_ERR_DEVICE = 102,
    OC_ERR_OUT_OF_MEMORY = 103
} oc_status;

/* ------------------------------------------------------------------------ */
/* Data and configuration PODs.                                              */
/* ------------------------------------------------------------------------ */

/* Geometry of a BandField: grid dims, pitch (m) and wavelength (m)
   (src/field.rs:99-105, 287-290). */
typedef struct oc_band_geom {
    uint64_t nx, ny;
    double dx, dy;
    double lambda;
} oc_band_geom;

/* AsmConfig (src/propagation/asm.rs:100-137). Default: {2, 1, 1, 8, 1}. */
typedef struct oc_asm_config {
    double padding;
    int32_t band_limited;
    int32_t strict;
    double max_padding;
    double medium;
} oc_asm_config;

/* SamplingReport (src/propagation/asm.rs:144-179). */
typedef struct oc_asm_report {
    double lambda, medium, z;
    uint64_t nx, ny;
    double dx, dy;
    double requested_padding, used_padding;
    uint64_t nx_pad, ny_pad;
    double band_limit_fx, band_limit_fy;
    double max_phase_step_x, max_phase_step_y;
    int32_t nyquist_ok;
    int32_t _pad0;
} oc_asm_report;

/* FresnelConfig (src/propagation/fresnel.rs:99-123). Default: {2, 1, 8}. */
typedef struct oc_fresnel_config {
    double padding;
    int32_t strict;
    int32_t _pad0;
    double max_padding;
} oc_fresnel_config;

/* FresnelSamplingReport (src/propagation/fresnel.rs:130-217). */
typedef struct oc_fresnel_report {
    double lambda, z;
    uint64_t nx, ny;
    double dx, dy;
    double requested_padding, used_padding;
    uint64_t nx_pad, ny_pad;
    double max_phase_step_x, max_phase_step_y;
    int32_t nyquist_ok;
    int32_t kernel_nyquist_ok;
    double max_kernel_step_x, max_kernel_step_y;
    double max_phase_error_vs_asm;
    int32_t paraxial_ok;
    int32_t _pad0;
} oc_fresnel_report;

/* RayleighSommerfeldConfig (src/propagation/rayleigh_sommerfeld.rs:95-116). */
typedef struct oc_rs_config {
    double padding;
    int32_t strict;
    int32_t _pad0;
} oc_rs_config;

/* FraunhoferConfig (src/propagation/fraunhofer.rs:117-170): the Option
   fields are a flag plus value. */
typedef struct oc_fraunhofer_config {
    int32_t strict;
    int32_t has_source_pos;       /* 0: axis at the grid origin */
    double max_n_f;
    double source_x, source_y;    /* optical axis (m), when has_source_pos */
    int32_t has_aperture_extent;  /* 0: conservative window extent */
    int32_t _pad0;
    double aperture_extent;       /* declared source half-extent (m) */
} oc_fraunhofer_config;

/* FraunhoferReport (src/propagation/fraunhofer.rs:178-222).  The output
   grid has the input's sample counts and pitch dx_out x dy_out. */
typedef struct oc_fraunhofer_report {
    double lambda, z;
    uint64_t nx, ny;
    double dx, dy;
    double dx_out, dy_out;
    double n_f_x, n_f_y;
    double source_x, source_y;
    double extent_used;
    int32_t extent_is_aperture;
    int32_t regime_ok;
    double quad_phase_error_rad;
    double max_n_f;
} oc_fraunhofer_report;

/* RayleighSommerfeldSamplingReport (rayleigh_sommerfeld.rs:123-189). */
typedef struct oc_rs_report {
    double lambda, z;
    uint64_t nx, ny;
    double dx, dy;
    double padding;
    uint64_t nx_pad, ny_pad;
    double max_kernel_step_x, max_kernel_step_y;
    int32_t kernel_nyquist_ok;
    int32_t paraxial_ok;
    double max_phase_error_vs_fresnel;
} oc_rs_report;

/* ClearAperture (src/element.rs:384-406). */
typedef enum oc_aperture_kind {
    OC_APERTURE_FIELD = 0,
    OC_APERTURE_CIRCLE = 1,    /* a = radius */
    OC_APERTURE_ELLIPSE = 2,   /* a, b = semi-axes */
    OC_APERTURE_RECTANGLE = 3  /* a, b = half-widths */
} oc_aperture_kind;

typedef struct oc_clear_aperture {
    int32_t kind;
    int32_t _pad0;
    double a, b;
} oc_clear_aperture;

/* PhaseMask::sampled (src/element.rs:1321-1368): lattice values (radians,
   row-major nx*ny), lattice pitch, orientation psi (radians; wrapped into
   [0, 2pi) like common::Angle), clear aperture and centre (m).
   `values` is a host pointer for host entry points and a device pointer for
   `_dev` entry points. */
typedef struct oc_phase_mask {
    const double* values;
    uint64_t nx, ny;
    double dx, dy;
    double orientation;
    oc_clear_aperture aperture;
    double cx, cy;
} oc_phase_mask;

/* ------------------------------------------------------------------------ */
/* Context.                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct oc_context oc_context;

/* Creates a context on HIP device `device` with its own stream. */
int oc_context_create(int device, oc_context** out);
void oc_context_destroy(oc_context* ctx);
/* Replaces the context stream (a hipStream_t passed as void*; NULL = the
   context's own stream).  The caller keeps ownership of a supplied stream. */
int oc_context_set_stream(oc_context* ctx, void* stream);
void* oc_context_stream(oc_context* ctx);
int oc_context_synchronize(oc_context* ctx);
/* Message of the last failure on this context (never NULL). */
const char* oc_last_error(oc_context* ctx);
/* Static name of a status code. */
const char* oc_status_name(int status);
/* Library version string. */
const char* oc_version(void);

/* Per-pass kernel timing (no reference counterpart; measurement support).
   When enabled, every pass launch is bracketed by HIP events recorded on the
   context stream; oc_context_pass_timing returns the accumulated milliseconds
   and launch counts per pass: [0] row forward, [1] column, [2] row inverse,
   [3] RS kernel generation, [4] GS fused passes, [5] other kernels,
   [6] column and [7] row pass of the two-pass (separable) complex64
   Fresnel form. */
#define OC_NUM_PASS_KINDS 8
int oc_context_enable_timing(oc_context* ctx, int enable);
/* Restricts the timing brackets to the pass kinds whose bit is set in mask
   (bit k = slot k above; all by default), so that a caller timing one pass
   records two events per launch of that pass only. */
int oc_context_set_timing_mask(oc_context* ctx, uint32_t mask);
int oc_context_pass_timing(oc_context* ctx, double* ms, uint64_t* launches,
                           int reset);

/* Autotuner (no reference counterpart).  Power-of-two padded lengths
   64..8192 have compile-time-specialised kernels in up to OC_FFT_VARIANTS
   workgroup layouts (rows per row-pass workgroup, width of the blocked
   intermediate, columns per column-pass workgroup; 1024- or 512-thread
   caps).  oc_autotune_fft times a masked-ASM plus Fresnel propagation round
   of every variant at length n on a synthetic batch and keeps the fastest for
   this context; ms_out (optional, OC_FFT_VARIANTS entries) receives the
   per-variant times (negative for variants that do not exist).  oc_context_set_fft_variant forces a choice;
   oc_context_fft_variant reports the current one (-1: no specialised kernel);
   untuned lengths use the layout measured fastest on MI355X for that length
   and precision.
   oc_context_force_generic(ctx, 1) routes every batch through the
   runtime-plan kernels instead. */
#define OC_FFT_VARIANTS 4
int oc_autotune_fft(oc_context* ctx, uint64_t n, int double_precision,
                    int* variant_out, double* ms_out);
int oc_context_set_fft_variant(oc_context* ctx, uint64_t n,
                               int double_precision, int variant);
int oc_context_fft_variant(oc_context* ctx, uint64_t n, int double_precision);
/* The complex64 Fresnel transfer function factors per axis, so its batches
   run two passes (columns, then rows) whose layouts are chosen separately:
   any two variants that block the intermediate with the same width.
   oc_autotune_fft (complex64) times every such pair after the variants and
   keeps the fastest; oc_context_set_fft_variant makes both passes follow
   the forced variant; untuned, 4096 uses (0, 3), the pair measured fastest
   on MI355X, and other lengths the propagation layout for both.  Setting a
   pair whose widths differ returns OC_ERR_INVALID_ARGUMENT.  No reference
   counterpart. */
int oc_context_set_fresnel_sep_layout(oc_context* ctx, uint64_t n, int v_col,
                                      int v_row);
int oc_context_fresnel_sep_layout(oc_context* ctx, uint64_t n, int* v_col,
                                  int* v_row);
int oc_context_force_generic(oc_context* ctx, int enable);

/* Batch schedule tuner (no reference counterpart; the reference runs
   rustfft's plan for every length).  A propagation batch
   (oc_propagate_*_dev_* on square padded grids) makes two kinds of choice,
   each kept per (padded length, precision, batch size class) where the
   class is the number of problems of that length in the call rounded up to
   a power of two:
     OC_SCHED_PAD   a convolution-engine length (Rayleigh-Sommerfeld,
                    Fresnel convolution) runs at `run_length`: its own
                    length, or the next power of two when that is at most
                    8192 (with P >= 2N - 1 on both axes every cropped
                    output sample is the same linear convolution for any
                    such P, so the reports keep the reference's P);
     OC_SCHED_FAST  a power-of-two run length uses the register-resident
                    kernels in layout `variant`;
     OC_SCHED_LDS   a run length uses the LDS kernels with the radix plan
                    radix[0..nradix) (stages in order; radices 2, 3, 4, 5,
                    7, 8, 16, product = the length).
   oc_autotune_batch_dev_* times the caller's batch (it runs it several
   times: outputs and statuses are those of the last run) under every
   candidate of every choice, one choice at a time with the others fixed
   (coordinate descent, `sweeps` rounds, 1 if 0), keeping the fastest; the
   picks stay in the context and, up to max_picks, are copied to picks
   (npicks receives their count); search_ms receives the wall time spent.
   Lengths the search did not see use the defaults (power-of-two lengths:
   next power of two for convolutions, the layout of oc_context_fft_variant;
   others: the radix plan 16.., 8/4/2, 3, 5, 7).  A pick found for another
   size class applies when its own class has none (the nearest class).
   oc_schedule_save writes the context's picks as text tagged with the
   device (name and compute units); oc_schedule_load reads such a file and
   installs the entries tagged with this context's device (loaded receives
   their count), OC_ERR_INVALID_ARGUMENT for a malformed file.
   oc_schedule_set installs one entry (ms ignored), oc_schedule_get copies
   entry i (OC_ERR_INVALID_ARGUMENT past the end), oc_schedule_clear drops
   all.  The tuner's picks affect performance only: every schedule computes
   the same propagation to rounding. */
#define OC_SCHED_PAD 0
#define OC_SCHED_FAST 1
#define OC_SCHED_LDS 2
typedef struct oc_schedule {
    uint64_t length;          /* padded length (square grids) */
    int32_t double_precision; /* 0: complex64, 1: complex128 */
    int32_t kind;             /* OC_SCHED_* */
    uint64_t problems;        /* batch size class (a power of two) */
    uint64_t run_length;      /* OC_SCHED_PAD */
    int32_t variant;          /* OC_SCHED_FAST */
    int32_t nradix;           /* OC_SCHED_LDS */
    int32_t radix[12];
    double ms;                /* batch time with this pick when tuned */
} oc_schedule;
struct oc_prop_desc; /* the batch descriptor, defined with the batch calls */
int oc_autotune_batch_dev_c64(oc_context* ctx, struct oc_prop_desc* descs,
                              size_t count, int sweeps, oc_schedule* picks,
                              size_t max_picks, size_t* npicks,
                              double* search_ms);
int oc_autotune_batch_dev_c128(oc_context* ctx, struct oc_prop_desc* descs,
                               size_t count, int sweeps, oc_schedule* picks,
                               size_t max_picks, size_t* npicks,
                               double* search_ms);
int oc_schedule_set(oc_context* ctx, const oc_schedule* s);
int oc_schedule_get(oc_context* ctx, size_t i, oc_schedule* out);
size_t oc_schedule_count(oc_context* ctx);
int oc_schedule_clear(oc_context* ctx);
int oc_schedule_save(oc_context* ctx, const char* path);
int oc_schedule_load(oc_context* ctx, const char* path, size_t* loaded);

/* ------------------------------------------------------------------------ */
/* Angular spectrum (src/propagation/asm.rs).                                */
/* ------------------------------------------------------------------------ */

/* Replaces asm::sampling_report (src/propagation/asm.rs:418-420). Host only. */
int oc_asm_sampling_report(const oc_band_geom* geom, double z,
                           const oc_asm_config* cfg, oc_asm_report* report);

/* Replaces asm::propagate (src/propagation/asm.rs:438-538): f64 drop-in, host
   buffers in and out (2*nx*ny doubles each; `in` and `out` may alias). */
int oc_asm_propagate(oc_context* ctx, const oc_band_geom* geom,
                     const double* in, double z, const oc_asm_config* cfg,
                     double* out, oc_asm_report* report);

/* Device-resident asm::propagate: complex128 (`double*`) and the FP32-bulk
   complex64 form (`float*`, FP64 phase arguments and twiddles).  `in` and
   `out` are device pointers and may alias.  A non-NULL `mask` (device
   `values`) is applied before propagating, fused into the first pass: it is
   Element::apply for PhaseMask (src/element.rs:1443-1484) followed by
   asm::propagate, as Beamline::trace does (src/beamline.rs:109-130). */
int oc_asm_propagate_dev_c128(oc_context* ctx, const oc_band_geom* geom,
                              const double* in, double z,
                              const oc_asm_config* cfg,
                              const oc_phase_mask* mask, double* out,
                              oc_asm_report* report);
int oc_asm_propagate_dev_c64(oc_context* ctx, const oc_band_geom* geom,
                             const float* in, double z,
                             const oc_asm_config* cfg,
                             const oc_phase_mask* mask, float* out,
                             oc_asm_report* report);

/* ------------------------------------------------------------------------ */
/* Fresnel transfer-function form (src/propagation/fresnel.rs).              */
/* ------------------------------------------------------------------------ */

/* Replaces fresnel::sampling_report (src/propagation/fresnel.rs:382-384). */
int oc_fresnel_sampling_report(const oc_band_geom* geom, double z,
                               const oc_fresnel_config* cfg,
                               oc_fresnel_report* report);
/* Replaces fresnel::propagate (src/propagation/fresnel.rs:430-509). */
int oc_fresnel_propagate(oc_context* ctx, const oc_band_geom* geom,
                         const double* in, double z,
                         const oc_fresnel_config* cfg, double* out,
                         oc_fresnel_report* report);
int oc_fresnel_propagate_dev_c128(oc_context* ctx, const oc_band_geom* geom,
                                  const double* in, double z,
                                  const oc_fresnel_config* cfg,
                                  const oc_phase_mask* mask, double* out,
                                  oc_fresnel_report* report);
int oc_fresnel_propagate_dev_c64(oc_context* ctx, const oc_band_geom* geom,
                                 const float* in, double z,
                                 const oc_fresnel_config* cfg,
                                 const oc_phase_mask* mask, float* out,
                                 oc_fresnel_report* report);

/* Replaces fresnel::propagate_convolution (src/propagation/fresnel.rs:531-590):
   the sampled paraxial impulse response's spectrum multiplies the field's;
   strict mode checks the kernel sampling without densifying. */
int oc_fresnel_propagate_convolution(oc_context* ctx, const oc_band_geom* geom,
                                     const double* in, double z,
                                     const oc_fresnel_config* cfg, double* out,
                                     oc_fresnel_report* report);
int oc_fresnel_propagate_convolution_dev_c128(oc_context* ctx,
                                              const oc_band_geom* geom,
                                              const double* in, double z,
                                              const oc_fresnel_config* cfg,
                                              double* out,
                                              oc_fresnel_report* report);
int oc_fresnel_propagate_convolution_dev_c64(oc_context* ctx,
                                             const oc_band_geom* geom,
                                             const float* in, double z,
                                             const oc_fresnel_config* cfg,
                                             float* out,
                                             oc_fresnel_report* report);

/* ------------------------------------------------------------------------ */
/* Fraunhofer far field (src/propagation/fraunhofer.rs).                     */
/* ------------------------------------------------------------------------ */

/* Replaces fraunhofer::regime_report (src/propagation/fraunhofer.rs:289-334). */
int oc_fraunhofer_regime_report(const oc_band_geom* geom, double z,
                                const oc_fraunhofer_config* cfg,
                                oc_fraunhofer_report* report);
/* Replaces fraunhofer::propagate (src/propagation/fraunhofer.rs:351-443):
   the output field has the input's dims on the far-field grid
   (report->dx_out, report->dy_out). */
int oc_fraunhofer_propagate(oc_context* ctx, const oc_band_geom* geom,
                            const double* in, double z,
                            const oc_fraunhofer_config* cfg, double* out,
                            oc_fraunhofer_report* report);
int oc_fraunhofer_propagate_dev_c128(oc_context* ctx, const oc_band_geom* geom,
                                     const double* in, double z,
                                     const oc_fraunhofer_config* cfg,
                                     double* out,
                                     oc_fraunhofer_report* report);
int oc_fraunhofer_propagate_dev_c64(oc_context* ctx, const oc_band_geom* geom,
                                    const float* in, double z,
                                    const oc_fraunhofer_config* cfg,
                                    float* out, oc_fraunhofer_report* report);

/* ------------------------------------------------------------------------ */
/* Rayleigh-Sommerfeld first kind (src/propagation/rayleigh_sommerfeld.rs).  */
/* ------------------------------------------------------------------------ */

/* Replaces rayleigh_sommerfeld::sampling_report (rayleigh_sommerfeld.rs:333). */
int oc_rs_sampling_report(const oc_band_geom* geom, double z,
                          const oc_rs_config* cfg, oc_rs_report* report);
/* Replaces rayleigh_sommerfeld::propagate (rayleigh_sommerfeld.rs:384-442). */
int oc_rs_propagate(oc_context* ctx, const oc_band_geom* geom,
                    const double* in, double z, const oc_rs_config* cfg,
                    double* out, oc_rs_report* report);
int oc_rs_propagate_dev_c128(oc_context* ctx, const oc_band_geom* geom,
                             const double* in, double z,
                             const oc_rs_config* cfg, double* out,
                             oc_rs_report* report);
int oc_rs_propagate_dev_c64(oc_context* ctx, const oc_band_geom* geom,
                            const float* in, double z,
                            const oc_rs_config* cfg, float* out,
                            oc_rs_report* report);

/* ------------------------------------------------------------------------ */
/* Batched propagation: many (grid, wavelength, z) problems per launch.      */
/* ------------------------------------------------------------------------ */

typedef enum oc_engine {
    OC_ENGINE_ASM = 0,
    OC_ENGINE_FRESNEL = 1,
    OC_ENGINE_RS = 2,
    OC_ENGINE_FRESNEL_CONV = 7  /* fresnel::propagate_convolution */
} oc_engine;

/* One problem of a batch.  `in`/`out` are device pointers (complex64 for the
   c64 batch, complex128 for the c128 batch); `mask` may be NULL (ASM and
   Fresnel only).  `padding` and `strict`/`band_limited`/`max_padding`/
   `medium` follow the engine's config; fields an engine does not have are
   ignored. `status` receives the per-problem status. */
typedef struct oc_prop_desc {
    oc_band_geom geom;
    double z;
    int32_t engine;
    int32_t band_limited;
    int32_t strict;
    int32_t status;
    double padding;
    double max_padding;
    double medium;
    const oc_phase_mask* mask;
    const void* in;
    void* out;
} oc_prop_desc;

/* Runs every problem of `descs` with persistent launches (one per pass for
   the whole batch).  Problems with a failing guard are skipped with their
   status set; the call returns the first non-OK status, or OC_OK. */
int oc_propagate_batch_dev_c64(oc_context* ctx, oc_prop_desc* descs,
                               size_t count);
int oc_propagate_batch_dev_c128(oc_context* ctx, oc_prop_desc* descs,
                                size_t count);

/* Two propagations in a row, problem by problem: second[i] propagates the
   output of first[i] (two gaps of Beamline::trace, src/beamline.rs:109-130,
   which calls asm::propagate then fresnel::propagate for a
   mask-ASM-Fresnel line).  second[i].geom must equal first[i].geom and
   second[i].in is ignored; first[i].out receives the intermediate field, or
   may be NULL when the caller does not need it.  Statuses as two batch calls
   would give them, except that a problem whose first stage fails is not
   propagated further: second[i].status repeats first[i].status and
   second[i].out is unspecified.  complex64 chains of masked or plain ASM
   into Fresnel at one power-of-two padded length run fused (the ASM's
   inverse row pass and the Fresnel's row pass in one kernel, the Fresnel's
   column pass on the blocked intermediate); three-pass chains (complex128,
   or complex64 ASM into ASM) at one power-of-two length run linked through
   their rows (one kernel for the first inverse and the second forward row
   pass); all others run as two batches.
   No single reference counterpart: it replaces the pair of calls. */
int oc_propagate_chain_dev_c64(oc_context* ctx, oc_prop_desc* first,
                               oc_prop_desc* second, size_t count);
int oc_propagate_chain_dev_c128(oc_context* ctx, oc_prop_desc* first,
                                oc_prop_desc* second, size_t count);

/* The chain without waiting (no counterpart in the reference: the
   pipelining form for serving loops and throughput runs).  A chain that
   runs fused or linked is queued on the context stream with its
   finiteness flags in a buffer of its own: host-side checks fill the
   statuses now, and oc_context_wait fills the device-side ones (a failed
   first stage fails the chain) into first[] and second[], which must stay
   valid until then.  A chain that runs as two batches waits between them,
   so its statuses are final when the call returns.  Queued work runs in
   submit order.  Returns the first host-detected status or OC_OK. */
int oc_propagate_chain_submit_dev_c64(oc_context* ctx, oc_prop_desc* first,
                                      oc_prop_desc* second, size_t count);
int oc_propagate_chain_submit_dev_c128(oc_context* ctx, oc_prop_desc* first,
                                       oc_prop_desc* second, size_t count);

/* oc_propagate_batch_dev_c64 without device allocation.  No counterpart in
   the reference: it is for callers that cannot allocate (code under stream
   capture, tight serving loops, harnesses that forbid hipMalloc).  Every
   buffer, table and descriptor of the call comes from one 256 MiB scratch
   array per device that the library reserves when it loads and holds until
   the process ends; the call allocates no device or pinned memory.  A batch
   whose buffers would exceed it (e.g. more than three 2048^2 problems at
   P = 4096) is refused with OC_ERR_INVALID_ARGUMENT before anything runs
   (every status set to it; oc_last_error says how much it needs).  Calls on
   one device take turns on the array (each waits for its own kernels before
   the next starts), and the array keeps the last call's contents.  Same
   kernels and results as oc_propagate_batch_dev_c64; returns after the
   batch has run. */
int oc_propagate_batch_noalloc_dev_c64(oc_context* ctx, oc_prop_desc* descs,
                                       size_t count);

/* The same without waiting (no counterpart in the reference: a pipelining
   form for serving loops).  Host-side checks fill descs[i].status now and
   the batch is queued on the context stream; oc_context_wait resolves the
   device-side statuses of every queued batch, writes them into their descs
   (which must stay valid until then) and returns the first non-OK status
   in submit order.  Queued batches run in submit order, so a later batch
   may read an earlier one's outputs.  The submit returns the first
   host-detected status (its other problems still run) or OC_OK. */
int oc_propagate_batch_submit_dev_c64(oc_context* ctx, oc_prop_desc* descs,
                                      size_t count);
int oc_propagate_batch_submit_dev_c128(oc_context* ctx, oc_prop_desc* descs,
                                       size_t count);
int oc_context_wait(oc_context* ctx);

/* ------------------------------------------------------------------------ */
/* Phase mask (src/element.rs).                                              */
/* ------------------------------------------------------------------------ */

/* Replaces PhaseMask::sampled validation (src/element.rs:1321-1368). Host. */
int oc_phasemask_validate(const oc_phase_mask* mask);
/* Replaces Element::apply for PhaseMask (src/element.rs:1443-1484): host f64
   buffers in and out (may alias).  On OC_INVALID_PHASE nothing is written. */
int oc_phasemask_apply(oc_context* ctx, const oc_band_geom* geom,
                       const double* in, const oc_phase_mask* mask,
                       double* out);
int oc_phasemask_apply_dev_c128(oc_context* ctx, const oc_band_geom* geom,
                                const double* in, const oc_phase_mask* mask,
                                double* out);
int oc_phasemask_apply_dev_c64(oc_context* ctx, const oc_band_geom* geom,
                               const float* in, const oc_phase_mask* mask,
                               float* out);

/* ------------------------------------------------------------------------ */
/* Gerchberg-Saxton (crates/optics-inverse/src/gs.rs).                       */
/* ------------------------------------------------------------------------ */

/* GsProblem + GsConfig (gs.rs:125-305) for a GridPhase carrier with the
   MseToTarget merit (merit.rs:112-141) as the merit seam.  Host pointers.
   init_kind: 0 = GsInit::Zero (carrier_values), 1 = GsInit::Random{seed}
   (rand 0.8 StdRng::seed_from_u64(seed), Uniform::new(0, span), drawn on
   the host: gs.rs:449-458), 2 = GsInit::Given (init_phase).
   variant: 0 = Classic, 1 = Dvr{beta}, 2 = Fpr{beta}. */
typedef struct oc_gs_problem {
    oc_band_geom geom;
    const double* input;            /* complex128 input field, host */
    double z_total;
    const double* target_amplitude; /* nx*ny */
    const double* merit_target;     /* nx*ny, MseToTarget target intensity */
    oc_asm_config asm_cfg;
    oc_clear_aperture aperture;     /* carrier stop */
    double cx, cy;                  /* carrier centre */
    double lo, hi;                  /* carrier slot bounds */
    const double* carrier_values;   /* nx*ny (GsInit::Zero start) */
    uint64_t max_iterations, patience;
    double merit_tolerance;
    int32_t variant;
    int32_t init_kind;
    double beta;
    const double* init_phase;       /* nx*ny when init_kind == 2 */
    uint64_t seed;                  /* init_kind == 1 */
} oc_gs_problem;

/* GsRecord (gs.rs:254-268). */
typedef struct oc_gs_record {
    uint64_t iteration;
    double merit, best_merit, amplitude_mse, target_error_rms,
        phase_update_rms;
} oc_gs_record;

/* GsSolution (gs.rs:271-305); every array is caller-allocated host memory
   of nx*ny doubles, history holds max_iterations records. */
typedef struct oc_gs_solution {
    double* phase;
    double merit;
    uint64_t best_iteration;
    double amplitude_mse;
    int32_t converged;
    int32_t _pad0;
    uint64_t iterations;
    oc_gs_record* history;
    double* target_intensity;
    double* target_amplitude_error;
    double* phase_error;
    double* illumination_amplitude;
} oc_gs_solution;

/* Replaces gs::run (gs.rs:347-597): two ASM half-transforms per iteration
   with the amplitude/phase projections, merit statistics and best-so-far
   tracking fused into the propagation passes. */
int oc_gs_run(oc_context* ctx, const oc_gs_problem* problem,
              oc_gs_solution* solution);

/* gs::run for `count` problems (restarts: typically one problem with
   different initial phases) on one grid size, together.  No single
   reference counterpart: it replaces `count` calls of gs::run
   (gs.rs:347-597) and returns the same solutions.  Every pass runs once for
   all restarts (persistent launches); the merit, best-so-far, history and
   best-phase copy are kept on the device and the GS sums are reduced in a
   fixed order, so a restart's history does not depend on the batch or the
   chunking.  graph_chunk > 1 queues that many iterations as one captured
   hipGraph and reads the histories back once per chunk; the patience stop is
   applied per restart and iteration, and a restart that stops inside a chunk
   is rolled back to the chunk's start and replayed alone up to its stop.
   graph_chunk <= 1 reads back every iteration.  A restart that has finished
   keeps running with the others until all have (its outputs were taken at
   its stop).  Errors: the first failing problem's validation status
   (nothing runs), OC_ERR_INVALID_ARGUMENT for unequal grid sizes,
   OC_INV_RENDER for a failing propagation. */
int oc_gs_run_batch(oc_context* ctx, const oc_gs_problem* problems,
                    oc_gs_solution* solutions, size_t count,
                    uint64_t graph_chunk);

/* Gerchberg-Saxton in the Fourier (FFT) form, `restarts` independent runs
   batched in one persistent launch per pass (no reference counterpart: the
   reference loop uses ASM half-transforms).  Projection semantics of
   gs.rs:484-564 (Classic), unpadded fft2 forward and ifft2 back, phase
   bounds [0, 2 pi).  Device pointers: source_amp and target_amp are n*n
   real (target_amp in DFT bin order), phases is restarts*n*n, updated in
   place.  `graph_chunk` > 0 captures that many iterations in a hipGraph and
   replays it.  err_out (host, restarts) receives
   ||I/sum I - T^2/sum T^2|| / ||T^2/sum T^2|| of the final phases. */
int oc_gs_fft_run_dev_c64(oc_context* ctx, uint64_t n, uint64_t restarts,
                          const float* source_amp, const float* target_amp,
                          float* phases, uint64_t iterations,
                          uint64_t graph_chunk, double* err_out);
int oc_gs_fft_run_dev_c128(oc_context* ctx, uint64_t n, uint64_t restarts,
                           const double* source_amp, const double* target_amp,
                           double* phases, uint64_t iterations,
                           uint64_t graph_chunk, double* err_out);

/* ------------------------------------------------------------------------ */
/* Adjoint gradients (crates/optics-inverse/src/adjoint.rs).                 */
/* ------------------------------------------------------------------------ */

/* One Design step (carrier.rs:35-43) whose element, if any, is a PhaseMask:
   apply the mask, then propagate z_after through `medium` (beamline.rs:
   109-130).  Element index = order of appearance. */
typedef struct oc_design_step {
    int32_t has_element;
    int32_t _pad0;
    oc_phase_mask mask;
    double z_after;
    double medium;
} oc_design_step;

/* Replaces adjoint::asm_adjoint (adjoint.rs:90-92): propagate by -z. Host. */
int oc_adjoint_asm(oc_context* ctx, const oc_band_geom* geom, const double* in,
                   double z, const oc_asm_config* cfg, double* out,
                   oc_asm_report* report);

/* Replaces adjoint::transmission (adjoint.rs:102-108): t = apply(1). Host. */
int oc_adjoint_transmission(oc_context* ctx, const oc_band_geom* geom,
                            const oc_phase_mask* mask, double* t_out);

/* Replaces adjoint::push_back (adjoint.rs:140-246): forward tape, seed
   w = dmdi * U_det, reverse sweep.  The tape keeps only each element's
   incoming field (per-plane checkpoint); t and V are recomputed from the
   mask lattice inside the fused gradient epilogue.  Host buffers; dphi_out
   holds one nx*ny gradient per element. */
int oc_adjoint_push_back(oc_context* ctx, const oc_design_step* steps,
                         size_t nsteps, const oc_band_geom* geom,
                         const double* input, const oc_asm_config* cfg,
                         const double* dmdi, double* dphi_out);

/* Replaces adjoint::grid_gradient (adjoint.rs:268-340) for one band and one
   scene with MseToTarget (merit.rs:112-166): render, merit, seed and
   push_back in one forward and one reverse sweep.  Device pointers (mask
   values included).  dphi_out (device, n_elements*nx*ny) may be NULL: then
   only the forward render and merit run.  merit_out (host) may be NULL.
   OC_INV_RENDER for a failing forward trace, OC_INV_ADJOINT for a failing
   adjoint propagation. */
int oc_adjoint_grid_gradient_dev_c128(oc_context* ctx,
                                      const oc_design_step* steps,
                                      size_t nsteps, const oc_band_geom* geom,
                                      const double* input,
                                      const oc_asm_config* cfg,
                                      const double* target, double* dphi_out,
                                      double* merit_out);

/* Replaces adjoint::grid_gradient (adjoint.rs:268-340) with MseToTarget
   (merit.rs:112-166) over every (band, scene) frame of a render: the forward
   trace of all frames as one batch per beamline step (per-plane
   checkpoints), the seeds, one reverse sweep for all frames, and the
   per-frame planes added in frame order, band-major (adjoint.rs:310-332).
   As the reference, each selected frame is seeded with obs_gradient's
   2 (I_f - T)/n (merit.rs:143-165), so with several frames selected the
   gradient is that of the sum of the frames' own MSEs; an unselected frame
   contributes nothing, and a selector that picks no frame (out of range)
   gives all-zero gradients (the reference returns no planes).
   designs: [band][nsteps] (every band's design has the same structure;
   mask values on the device); inputs: [band][scene] fields (device,
   complex128 / complex64); target: nx*ny f64 (device); sel_band,
   sel_scene: the selector (-1 = all); dphi_out: n_elements*nx*ny f64
   (device) or NULL for the render and merit only; merit_out (host, may be
   NULL): MseToTarget::evaluate of the rendered observation (the selected
   frames' intensities summed), NaN for a selector out of range.  Errors:
   OC_INV_RENDER for an empty band or scene list or a failing forward trace,
   OC_INV_ADJOINT for a failing adjoint propagation, OC_ERR_UNSUPPORTED when
   the bands' designs differ in structure. */
int oc_adjoint_grid_gradient_frames_dev_c128(
    oc_context* ctx, const oc_band_geom* geom, const double* bands,
    size_t n_bands, size_t n_scenes, const oc_design_step* designs,
    size_t nsteps, const oc_asm_config* cfg, const void* inputs,
    const double* target, int32_t sel_band, int32_t sel_scene,
    double* dphi_out, double* merit_out);
int oc_adjoint_grid_gradient_frames_dev_c64(
    oc_context* ctx, const oc_band_geom* geom, const double* bands,
    size_t n_bands, size_t n_scenes, const oc_design_step* designs,
    size_t nsteps, const oc_asm_config* cfg, const void* inputs,
    const double* target, int32_t sel_band, int32_t sel_scene,
    double* dphi_out, double* merit_out);

/* Renderer::render (crates/optics-inverse/src/observe.rs:213-314) for
   designs whose elements are PhaseMasks.  The caller keeps the two closures
   of the Rust renderer: it builds the design for every wavelength the
   render needs (Carrier::build) and evaluates the source for every render
   (the Source closure), then this call runs every frame and FD render as
   one batch per beamline step and writes the Observation data.

   Renders per (band, scene), in this order: base; (x0 + h, x0 - h) if
   deriv_x; (y0 + h, y0 - h) if deriv_y; (lambda + dl, lambda - dl) if
   deriv_lambda (h = scene_step, dl = lambda_step).
   inputs: [band][scene][render] fields of geom's grid (complex, row-major).
   designs: [band][variant][nsteps] with variant 0 = the band's wavelength
   and, when deriv_lambda, 1 = lambda + dl, 2 = lambda - dl.
   obs: [channel][band][scene][ny][nx] f64, channel 0 = |U|^2, then the
   enabled derivatives in x, y, lambda order, each (I+ - I-) / (2 step).
   geom->lambda is not used: every render's wavelength comes from bands.
   Errors: OC_INV_RENDER for an empty band or scene list, a non-positive or
   non-finite step of an enabled derivative, a failed propagation or mask,
   or a non-finite base intensity (observe.rs:214-231, 296-303). */
typedef struct oc_render_config {
    int32_t deriv_x, deriv_y, deriv_lambda, _pad0;
    double scene_step;  /* m */
    double lambda_step; /* m */
} oc_render_config;

/* Host form: inputs (complex128) and obs in host memory. */
int oc_render(oc_context* ctx, const oc_band_geom* geom, const double* bands,
              size_t n_bands, size_t n_scenes, const oc_render_config* rc,
              const oc_design_step* designs, size_t nsteps,
              const oc_asm_config* cfg, const double* inputs, double* obs);
/* Device forms: inputs (complex128 / complex64), mask values and obs on
   the device. */
int oc_render_dev_c128(oc_context* ctx, const oc_band_geom* geom,
                       const double* bands, size_t n_bands, size_t n_scenes,
                       const oc_render_config* rc,
                       const oc_design_step* designs, size_t nsteps,
                       const oc_asm_config* cfg, const void* inputs,
                       double* obs);
int oc_render_dev_c64(oc_context* ctx, const oc_band_geom* geom,
                      const double* bands, size_t n_bands, size_t n_scenes,
                      const oc_render_config* rc,
                      const oc_design_step* designs, size_t nsteps,
                      const oc_asm_config* cfg, const void* inputs,
                      double* obs);

/* fd::gradient (crates/optics-inverse/src/fd.rs:139-184) for the objective
   "trace one band and scene through a PhaseMask design with ASM and score
   the intensity with MseToTarget" (observe.rs:195-210, merit.rs:124-137).
   The caller keeps Params: it validates x, clamps the probe points into
   [lo, hi] and builds their designs (Carrier::build).
   designs: [dim][2][nsteps], slot i's up probe then its down probe.
   up, dn, lo, hi: host arrays of dim probe values and bounds; a slot with
   lo == hi is skipped (gradient 0, designs unread, no evaluation).
   All probes run as one batch per beamline step from the one input field
   (device); each probe's merit is reduced on the device in a fixed order.
   g_i = (M(up) - M(dn)) / (up_i - dn_i).  grad_out[dim] (host);
   merits_out[2*dim] (host, may be NULL; 0 for skipped slots);
   evaluations_out (host, may be NULL) = 2 x live slots.
   Errors: OC_INV_RENDER for a failed render, a non-finite merit or
   gradient, or a live slot with up_i - dn_i <= 0. */
int oc_fd_gradient_dev_c128(oc_context* ctx, const oc_band_geom* geom,
                            const oc_design_step* designs, size_t nsteps,
                            size_t dim, const double* up, const double* dn,
                            const double* lo, const double* hi,
                            const oc_asm_config* cfg, const void* input,
                            const double* target, double* grad_out,
                            double* merits_out, size_t* evaluations_out);
int oc_fd_gradient_dev_c64(oc_context* ctx, const oc_band_geom* geom,
                           const oc_design_step* designs, size_t nsteps,
                           size_t dim, const double* up, const double* dn,
                           const double* lo, const double* hi,
                           const oc_asm_config* cfg, const void* input,
                           const double* target, double* grad_out,
                           double* merits_out, size_t* evaluations_out);

/* AdamConfig (fd.rs:62-118); has_clip = 0 is clip_norm = None. */
typedef struct oc_adam_config {
    double lr_frac, beta1, beta2, eps;
    int32_t has_clip, _pad0;
    double clip_norm;
} oc_adam_config;

/* One Adam epoch of fd::adam (fd.rs:223-244) on the device: gradient-norm
   clip, moment updates with bias corrections 1 - beta^t, the step
   lr_frac (hi - lo) m_hat / (sqrt(v_hat) + eps), then the clamp into
   [lo, hi] (solve.rs:74-78).  x, m, v updated in place; g, lo, hi read;
   all device f64 arrays of dim.  t >= 1 is the epoch.  The norm is
   reduced in a fixed order.  Errors: OC_INV_RENDER for an invalid config
   (fd.rs:95-118), OC_ERR_INVALID_ARGUMENT for t = 0. */
int oc_adam_step_dev(oc_context* ctx, size_t dim, double* x, const double* g,
                     double* m, double* v, const double* lo, const double* hi,
                     uint64_t t, const oc_adam_config* cfg);

/* analysis::mtf (src/analysis.rs:133-150) for a batch of PSFs: the fy = 0
   row of the 2-D DFT of each intensity, |F(k, 0)| / max(|F(0, 0)|,
   DBL_MIN) capped at 1, for k = 0..nx/2, at frequencies k / (nx dx).  That
   row is the 1-D DFT of the column sums, which is what the device computes
   (the sums in f64 over y, then a direct DFT with an exact twiddle index).
   psfs: [count][ny][nx] f64; values: [count][nx/2 + 1]; freqs (may be
   NULL): [nx/2 + 1], host.  Errors: OC_EMPTY_GRID, OC_INVALID_GRID_PITCH
   (Psf::new, field.rs), OC_NON_FINITE_SAMPLE for a non-finite sample. */
int oc_mtf(oc_context* ctx, const double* psfs, size_t count, size_t nx,
           size_t ny, double dx, double* values, double* freqs);
int oc_mtf_dev(oc_context* ctx, const double* psfs, size_t count, size_t nx,
               size_t ny, double dx, double* values, double* freqs);

#ifdef __cplusplus
}
#endif

#endif /* OPTICS_GPU_H */
