/* oracle.h - TEST INFRASTRUCTURE.  C ABI of the CPU oracle (a serial restatement
 * of the reference Julia pixel-track stage).  Loaded by tests/ and by the
 * cpu_baseline leg of bench.py only; the product never links it. */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H
#include <stdint.h>
#include "../include/pixeltrack.h"
#ifdef __cplusplus
extern "C" {
#endif

void oracle_params_default(pt_params *p);

/* make_tuples restated (ca_hit_ntuplet_generator.jl:65-97) on host memory.
 * frames: rot[9*module]; stop_after: 0 = full stage, 1 = after doublets,
 * 2 = after connect, 3 = after fishbone, 4 = after ntuplets/multiplicity
 * (no fit, no classification), 5 = after the fit (no classification).  out->state/covariance must be zeroed by the
 * caller (TrackSOA() does the same). */
int oracle_make_tuples(const pt_params *p, const pt_hits_view *hits, const float *frames, double b_field,
                       pt_track_soa *out, pt_counters *cnt, pt_ca_debug *dbg, int stop_after);

/* Plants a mistake in fish_bone (gpu_fishbone.jl:43-56) for the tests that
 * show the fishbone pins catch it: no `break` after killing the i-th cell, no
 * inner-module comparison.  (0, 0) restores the restatement.  Tests only. */
void oracle_plant_fishbone(int no_break, int no_module_check);

/* fill_many_from_vector (hist_to_container.jl:265-270) for the 10-layer phi histogram */
int oracle_build_phi_hist(const int16_t *iphi, uint32_t n_hits, const uint32_t *layer_start, uint32_t *off,
                          uint16_t *bins);

/* load_tracks (gpu_vertex_finder.jl:51-82) on one host TrackSOA: n = kept
 * rows, index_track (1-based row), zt, ezt2, ptt2 as in the WorkSpace. */
int oracle_load_tracks(const pt_track_soa *t, float pt_min, uint32_t *n, uint16_t *index_track, float *zt,
                       float *ezt2, float *ptt2);

/* FP64 broken-line fit of pre-gathered tracks; same input layout as
 * pt_fit_bl_device_async.  Double outputs: out64[t*23 + k] = state[0..4],
 * cov15[5..19], pt, eta, chi2.  flags[t] = oracle fit flags. */
int oracle_fit_bl(uint32_t n_tracks, const float *hits, const float *ge, const uint8_t *nhit, double b_field,
                  double *out64, uint32_t *flags, int n_threads);

/* Plants a mistake in mult_scatt (broken_line.jl:46-79) for the tests that
 * show the scattering pins catch it: the result times `scale`, log10 in place
 * of the natural log, no 20 GeV cap.  (1, 0, 0) restores the restatement.
 * Process-wide; tests only. */
void oracle_plant_ms(double scale, int log10_for_ln, int drop_cap);

/* One track with every intermediate (for the known-answer checks).
 * hits: 3xN column-major doubles (x,y,z per hit), ge: 6xN column-major floats.
 * Returns the number of doubles written to out (see tests/oracle_layout.py). */
int oracle_bl_detail(int N, const double *hits, const float *ge, double B, double *out, int cap);

/* Count the arithmetic of one fit of N hits: counts[0]=add/sub, [1]=mul,
 * [2]=div, [3]=sqrt, [4]=transcendental calls, [5]=comparisons. */
int oracle_fit_opcount(int N, const double *hits, const float *ge, double B, uint64_t *counts);

/* Forward-mode derivative of L = chi2_line + chi2_circle (+ ||params - truth||^2, phi difference modulo 2 pi,
 * if truth != NULL) with respect to (x, y, z, sigma_rphi, sigma_z) of each of
 * the N hits, through local variances rotated by the module frames rot[9*k].
 * grad[5*k + c]; value[0] = L. */
int oracle_fit_grad(int N, const double *xyz, const double *sig, const float *rot, double B, const double *truth,
                    double *value, double *grad);

/* oracle_fit_grad over a fit-only batch in the layout of
 * pt_fit_bl_loss_grad_async (local sigmas, module ids, truth may be NULL),
 * on n_threads host threads: loss[t], grad[(c*5+k)*n+t]. */
int oracle_fit_grad_many(uint32_t n, const float *hits, const float *sig, const int16_t *det, const uint8_t *nhit,
                         const float *rot, double b_field, const float *truth, double *loss, double *grad,
                         int n_threads);

/* Riemann helix fit (riemann_fit.jl restated, see oracle_riemann.hpp for the
 * statements it corrects) of pre-gathered tracks; same input layout and
 * out64 layout as oracle_fit_bl (state, cov15, pt, eta, chi2 in the TrackSOA
 * convention).  5-hit tracks are fitted on their first 4 hits. */
int oracle_fit_riemann(uint32_t n_tracks, const float *hits, const float *ge, const uint8_t *nhit, double b_field,
                       double *out64, int n_threads);

/* oracle_fit_riemann with the hit positions multiplied by `scale` in double
 * after the gather: scale = 1 +- 2^-52 measures how much a one-ulp change of
 * the input moves each output (the conditioning of the method per track). */
int oracle_fit_riemann_scaled(uint32_t n_tracks, const float *hits, const float *ge, const uint8_t *nhit,
                              double b_field, double scale, double *out64, int n_threads);

/* One Riemann fit with every intermediate (tests/oracle_layout.py unpacks it).
 * hits: 3xN column-major doubles, ge: 6xN column-major floats, N = 3, 4 or 5. */
int oracle_riemann_detail(int N, const double *hits, const float *ge, double B, double *out, int cap);

/* Plants one mistake in the Riemann restatement (bit mask of
 * oracle::riemann::kPlant*); 0 restores it.  Process-wide; tests only. */
void oracle_plant_riemann(uint32_t mask);

/* The symmetric 3x3 eigen-solver of the Riemann restatement: A row-major,
 * eigenvalues w, eigenvectors in the columns of Q (row-major). */
int oracle_sym_eigen3(const double *A, double *w, double *Q);

#ifdef __cplusplus
}
#endif
#endif
