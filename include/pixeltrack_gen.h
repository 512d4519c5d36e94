/*
 * pixeltrack_gen.h - seeded synthetic input for the pixel-track stage.
 *
 * Stand-in for the upstream producers that the stage consumes: it emits the
 * hit SoA with the field meanings of the rechit producer (gpuPixelRecHits.jl:284-331:
 * local variances, rg = sqrt(xg^2+yg^2) in f32, iphi = unsafe_atan2s(yg,xg,7)
 * from approx_atan2.jl:164-184, 0-based module ids laid out as
 * phase1PixelTopology.jl:25-37), the layer starts (PixelRecHits.jl:20-27) and
 * the 10x128 phi histogram filled like fill_many_from_vector
 * (hist_to_container.jl:217-270, descending hit id inside a bin).
 * Geometry: 4 barrel layers (r = 2.9, 6.8, 10.9, 16.0 cm, |z| <= 26.6 cm) and
 * 3 disks per side (|z| = 29.1, 39.6, 51.6 cm, 4.5 <= r <= 16 cm); prompt
 * helices in a uniform solenoid field; Gaussian smearing in the module frame.
 */
#ifndef PIXELTRACK_GEN_H
#define PIXELTRACK_GEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define PTG_PT_LOGUNIFORM 0  /* pT log-uniform in [pt_min, pt_max]                */
#define PTG_PT_POWERLAW   1  /* dN/dpT ~ pT^-pt_index above pt_min                 */

typedef struct ptg_config {
    uint32_t n_helices;
    uint32_t n_vertices;     /* helices are assigned uniformly to the vertices  */
    double   noise_frac;     /* extra uniform hits, fraction of the track hits  */
    int32_t  pt_mode;
    double   pt_min, pt_max, pt_index;
    double   eta_max;
    double   z_sigma;        /* vertex z spread (cm)                            */
    double   b_tesla;        /* field used to bend the helices                  */
    double   sigma_rphi;     /* smear along local x (cm)                        */
    double   sigma_z;        /* smear along local y: z (barrel) or r (disks)    */
    uint64_t seed;
} ptg_config;

typedef struct ptg_event {
    uint32_t  n_hits;
    float    *xg, *yg, *zg, *rg, *xerr, *yerr;
    int16_t  *iphi, *ysize, *det_ind;
    uint32_t  layer_start[11];
    uint32_t  hist_off[1281];
    uint16_t *hist_bins;
    int32_t  *truth_track;   /* per hit: helix index, -1 for noise               */
    uint32_t  n_tracks;
    float    *trk_pt, *trk_phi, *trk_eta, *trk_z0;
    int8_t   *trk_q;
    uint32_t  n_filler;      /* hits added so that no layer is empty             */
} ptg_event;

void ptg_config_default(ptg_config *c);
int  ptg_generate_event(const ptg_config *c, ptg_event **out);
void ptg_free_event(ptg_event *e);

/* Module frames for all 1856 modules: rot[9*m + k] = R11..R33 (rows = local x,
 * local y, normal in global coordinates), as DetParams.frame.rot. */
int  ptg_frames(float *rot, uint32_t n_modules);

/* iphi[i] = unsafe_atan2s(y[i], x[i], 7) (approx_atan2.jl:164-184), for hits
 * that callers place themselves. */
int  ptg_iphi(uint32_t n, const float *x, const float *y, int16_t *iphi);

/* Phi histogram of a hit list (fill_many_from_vector semantics). */
int  ptg_build_phi_hist(const int16_t *iphi, const uint32_t *layer_start, uint32_t *off, uint16_t *bins);

/* Fit-only tracks in the pt_fit_bl layout (hits[(c*5+k)*n+t], ge[(c*5+k)*n+t]).
 * len_mix[0..2] = fractions of 3/4/5-hit tracks; per-hit sigma grows linearly
 * from (srphi0, sz0) on the first hit to (srphi1, sz1) on the fifth.
 * Also returns the local sigmas (sig[(c*5+k)*n+t], c = rphi, z) and the
 * module of each hit (det[k*n+t]) for gradient workloads. */
typedef struct ptg_fit_config {
    uint32_t n_tracks;
    double   len_mix[3];
    double   pt_min, pt_max;      /* log-uniform */
    double   eta_max, z_sigma, b_tesla;
    double   srphi0, srphi1, sz0, sz1;
    uint64_t seed;
    /* Multiple scattering, 0 = none.  Otherwise the track is propagated hit
     * to hit and deflected by Gaussian kinks whose variance is ms_scale times
     * the Highland angle (PDG, beta = 1, unit charge) of the material the
     * broken-line fit assumes: a uniform medium with 1/X0 = 0.06/16 per cm of
     * path and a geometry factor 0.7 (broken_line.jl:46-79).  The material of
     * each path segment between two hits acts at both ends of the segment, in
     * both planes that contain the direction: an interior hit takes both
     * adjacent segments (the fit's VarBeta, broken_line.jl:200-206), the first
     * hit the segment to the second hit (the term the fit adds to the slope
     * and phi variances, broken_line.jl:477-479,619-622).  Truth stays the
     * vertex helix. */
    double   ms_scale;
} ptg_fit_config;
int ptg_generate_fit_tracks(const ptg_fit_config *c, float *hits, float *ge, uint8_t *nhit, float *sig,
                            int16_t *det, float *truth5);

#ifdef __cplusplus
}
#endif
#endif
