/*
 * ORACLE — test infrastructure, not product code.
 *
 * A plain C++ f64 CPU restatement of the reference's propagation path, used
 * only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker (never as the thing measured or shipped).  Each function cites
 * the reference file:line it follows.  The reference is Rust and cannot be
 * built in this image (no rustc/cargo, crates not vendored), so this
 * restatement is the oracle; it is pinned by the reference's own closed-form
 * unit tests, its hypercam golden fixture and numpy.fft (tests/test_oracle*.py).
 *
 * Parity status per function is listed in the header of oracle.cpp.
 *
 * The POD config/report/mask structs are the ones of include/optics_gpu.h so
 * that the Python checker can pass the same structs to both libraries.
 */
#ifndef OPTICS_ORACLE_H
#define OPTICS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/optics_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* fft.rs */
void or_fft2(double* buf, uint64_t nx, uint64_t ny);
void or_ifft2(double* buf, uint64_t nx, uint64_t ny);
void or_fft1(double* buf, uint64_t n, int inverse); /* unnormalised 1-D */
void or_fftfreq(uint64_t n, double d, double* out);
uint64_t or_padded_len(uint64_t n, double padding);

/* asm.rs */
int or_asm_sampling_report(const oc_band_geom* g, double z,
                           const oc_asm_config* cfg, oc_asm_report* rep);
int or_asm_propagate(const oc_band_geom* g, const double* in, double z,
                     const oc_asm_config* cfg, double* out,
                     oc_asm_report* rep);

/* fresnel.rs */
int or_fresnel_sampling_report(const oc_band_geom* g, double z,
                               const oc_fresnel_config* cfg,
                               oc_fresnel_report* rep);
int or_fresnel_propagate(const oc_band_geom* g, const double* in, double z,
                         const oc_fresnel_config* cfg, double* out,
                         oc_fresnel_report* rep);
int or_fresnel_propagate_convolution(const oc_band_geom* g, const double* in,
                                     double z, const oc_fresnel_config* cfg,
                                     double* out, oc_fresnel_report* rep);
void or_fresnel_transfer_value(double fx, double fy, double lambda, double z,
                               double* out2);
double or_fresnel_worst_axis_step(uint64_t n_pad, double d, double lambda,
                                  double z);

/* rayleigh_sommerfeld.rs */
int or_rs_sampling_report(const oc_band_geom* g, double z,
                          const oc_rs_config* cfg, oc_rs_report* rep);
int or_fraunhofer_regime_report(const oc_band_geom* g, double z, const oc_fraunhofer_config* cfg,
                                oc_fraunhofer_report* rep);
int or_fraunhofer_propagate(const oc_band_geom* g, const double* in, double z, const oc_fraunhofer_config* cfg,
                            double* out, oc_fraunhofer_report* rep);
int or_rs_propagate(const oc_band_geom* g, const double* in, double z,
                    const oc_rs_config* cfg, double* out, oc_rs_report* rep);
void or_rs_kernel_value(double x, double y, double lambda, double z,
                        double* out2);
void or_rs_impulse_response(uint64_t nx_pad, uint64_t ny_pad, double dx,
                            double dy, double lambda, double z, double* out);
double or_rs_worst_kernel_axis_step(uint64_t n_pad, double d, double lambda,
                                    double z);

/* element.rs */
int or_phasemask_validate(const oc_phase_mask* m);
int or_phasemask_apply(const oc_band_geom* g, const double* in,
                       const oc_phase_mask* m, double* out);
double or_sampled_phase_at(const oc_phase_mask* m, double x, double y);
int or_elliptical_aperture_apply(const oc_band_geom* g, const double* in,
                                 double a, double b, double psi, double cx,
                                 double cy, double* out);
int or_thinlens_apply(const oc_band_geom* g, const double* in, double f,
                      const oc_clear_aperture* ap, double psi, double cx,
                      double cy, double* out);

/* sources.rs / field.rs */
int or_gaussian(double w0, double amplitude, double cx, double cy, double z,
                double lambda, uint64_t nx, uint64_t ny, double dx, double dy,
                double* out);
int or_plane_wave(double theta_x, double theta_y, double amplitude,
                  uint64_t nx, uint64_t ny, double dx, double dy,
                  double lambda, double* out);

/* Deterministic input generation shared with the GPU tests: SplitMix64 ->
   u53 * 2^-53 uniforms in [0, 1). */
void or_splitmix_uniform(uint64_t seed, uint64_t count, double* out);

/* GsInit::Random's generator (gs.rs:449-458): rand 0.8 StdRng =
   rand_chacha 0.3 ChaCha12Rng seeded by rand_core 0.6 seed_from_u64, and
   Uniform<f64>.  or_chacha_block is the ChaCha block function with a 64-bit
   counter and 64-bit stream id and `double_rounds` double rounds (10 for
   ChaCha20, 6 for the ChaCha12 of StdRng), for known-answer checks;
   or_std_rng_u64 draws `count` next_u64 values; or_gs_random_phase fills
   the starting phases for slot bounds lo, hi. */
void or_chacha_block(const uint32_t* key, uint64_t counter, uint64_t stream,
                     int double_rounds, uint32_t* out);
void or_std_rng_u64(uint64_t seed, uint64_t count, uint64_t* out);
void or_gs_random_phase(uint64_t seed, double lo, double hi, uint64_t n,
                        double* out);

/* optics-inverse gs.rs (ASM half-transforms; merit = MseToTarget on the
   intensity frame with `merit_target`). */
typedef struct or_gs_problem {
    oc_band_geom geom;            /* input band geometry */
    const double* input;          /* complex input field, 2*n doubles */
    double z_total;
    const double* target_amplitude; /* n */
    const double* merit_target;   /* n (MseToTarget target intensity) */
    oc_asm_config asm_cfg;
    /* carrier (GridPhase): aperture + centre + slot bounds + current values */
    oc_clear_aperture aperture;
    double cx, cy;
    double lo, hi;
    const double* carrier_values; /* n (GsInit::Zero start) */
    /* GsConfig */
    uint64_t max_iterations, patience;
    double merit_tolerance;
    int32_t variant;              /* 0 Classic, 1 Dvr, 2 Fpr */
    int32_t init_kind;            /* 0 Zero, 1 Random{seed}, 2 Given */
    double beta;
    const double* init_phase;     /* n when init_kind == 2 */
    uint64_t seed;                /* init_kind == 1 */
} or_gs_problem;

typedef struct or_gs_record {
    uint64_t iteration;
    double merit, best_merit, amplitude_mse, target_error_rms,
        phase_update_rms;
} or_gs_record;

int or_gs_run(const or_gs_problem* p, double* phase_out, double* merit,
              uint64_t* best_iteration, double* amplitude_mse,
              int32_t* converged, uint64_t* iterations,
              or_gs_record* history, double* target_intensity,
              double* target_amplitude_error, double* phase_error,
              double* illumination_amplitude);

/* Gerchberg-Saxton with unpadded fft2 forward / ifft2 back (the FFT form of
   the loop, projection semantics of gs.rs:484-564, Classic variant, bounds
   [0, 2pi)).  phases: restarts * n*n, updated in place. err_out: per restart
   ||I/sum I - T^2/sum T^2|| / ||T^2/sum T^2|| of the final phase. */
void or_gs_fft(uint64_t n, const double* source_amp, const double* target_amp,
               double* phases, uint64_t restarts, uint64_t iterations,
               double* err_out);

/* optics-inverse adjoint.rs */
typedef struct or_design_step {
    int32_t has_element;
    int32_t _pad0;
    oc_phase_mask mask;
    double z_after;
    double medium;
} or_design_step;

/* push_back (adjoint.rs:140-246) for a design whose elements are phase masks,
   element index = order of appearance.  dphi_out: n_elements * n. */
int or_adjoint_push_back(const or_design_step* steps, uint64_t nsteps,
                         const oc_band_geom* g, const double* input,
                         const oc_asm_config* asm_cfg, const double* dmdi,
                         double* dphi_out);
/* grid_gradient (adjoint.rs:268-340) with MseToTarget (merit.rs:112-166)
   over every (band, scene) frame: designs [band][nsteps], inputs
   [band][scene] fields, selector band/scene (-1 = all).  dphi_out:
   n_elements * n, the per-frame planes added band-major; *n_planes = 0
   (dphi_out unwritten) when the seed is zero everywhere.  merit_out (may be
   NULL): MseToTarget::evaluate of the observation, NaN for a selector out of
   range. */
int or_grid_gradient(const oc_band_geom* g, const double* bands, uint64_t nb,
                     uint64_t ns, const or_design_step* designs,
                     uint64_t nsteps, const oc_asm_config* asm_cfg,
                     const double* inputs, const double* target,
                     int32_t sel_band, int32_t sel_scene, double* dphi_out,
                     uint64_t* n_planes, double* merit_out);
/* Beamline::trace with ASM (beamline.rs:109-130, observe.rs:195-210):
   output field (complex) of the design. */
int or_trace(const or_design_step* steps, uint64_t nsteps,
             const oc_band_geom* g, const double* input,
             const oc_asm_config* asm_cfg, double* out);
/* Renderer::render (observe.rs:213-314): renders in the reference's order
   (per band, per scene: base, x+/x-, y+/y-, lambda+/lambda-), each a trace
   of the design built for its wavelength; inputs [band][scene][render],
   designs [band][variant][nsteps], obs [channel][band][scene][ny][nx]
   (the layout of oc_render).  Returns OC_INV_RENDER on any failure. */
/* fd::gradient (fd.rs:139-184) with the objective "trace the design,
   MseToTarget on the intensity" (merit.rs:124-137); designs [dim][2][nsteps],
   the oc_fd_gradient_dev layout.  merits (2*dim) may be NULL. */
int or_fd_gradient(const oc_band_geom* g, const or_design_step* designs,
                   uint64_t nsteps, uint64_t dim, const double* up,
                   const double* dn, const double* lo, const double* hi,
                   const oc_asm_config* asm_cfg, const double* input,
                   const double* target, double* grad, double* merits,
                   uint64_t* evaluations);
/* One epoch of the Adam update (fd.rs:210-244) given the gradient g, then
   the clamp (solve.rs:74-78); x, m, v updated in place. */
int or_adam_step(uint64_t dim, double* x, const double* g, double* m,
                 double* v, const double* lo, const double* hi, uint64_t t,
                 const oc_adam_config* cfg);
/* analysis::mtf (analysis.rs:133-150): fft2 of the intensity, row fy = 0,
   |F(k,0)| / max(|F(0,0)|, DBL_MIN) capped at 1 for k = 0..nx/2. */
void or_mtf(const double* psf, uint64_t nx, uint64_t ny, double dx,
            double* freqs, double* values);
int or_render(const oc_band_geom* g, const double* bands, uint64_t nb,
              uint64_t ns, const oc_render_config* rc,
              const or_design_step* designs, uint64_t nsteps,
              const oc_asm_config* asm_cfg, const double* inputs,
              double* obs);

#ifdef __cplusplus
}
#endif

#endif
