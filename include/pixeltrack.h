/*
 * pixeltrack.h - C ABI of the MI355X (gfx950) pixel-track stage.
 *
 * The library replaces the pixel-track stage of the serial Julia Patatrack
 * port: CA doublets -> cell connection -> fishbone -> ntuplets -> multiplicity
 * -> broken-line fit -> track classification.  Every entry point takes plain
 * pointers and sizes; index conventions are the reference's:
 *   - stored hit ids are 1-based (gpu_pixel_doublets.jl:214, hist_to_container.jl:242)
 *   - module (detector) indices are 0-based (gpuPixelRecHits.jl:287)
 *   - hits_layer_start holds 0-based hit offsets (PixelRecHits.jl:20-27)
 *   - the tuple container off[] holds 1-based start positions (hist_to_container.jl:141,159)
 *   - count-filled containers hold exclusive offsets (hist_to_container.jl:201-212)
 *
 * Status: every function returns PT_OK (0) or a negative PT_ERR_* code;
 * pt_last_error() gives a message.  Capacity overflows are NOT errors: they
 * truncate exactly like the reference and set bits in pt_counters.overflow
 * (one exception: list-column exhaustion, PT_INT_LIST_SHARE below).
 * Inputs on which the reference would stop with an AssertionError /
 * InexactError / DomainError are reported in pt_counters.ref_assert.
 */
#ifndef PIXELTRACK_H
#define PIXELTRACK_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- capacities (ca_constants.jl:10-29, PixelTrackHeterogeneous.jl:55-60,
 *      gpu_clustering_constants.jl:14, ca_hit_ntuplet_generator.jl:12) ---- */
#define PT_MAX_HITS            49152u  /* MAX_NUMBER_OF_HITS                     */
#define PT_MAX_DOUBLETS        458752u /* maxNumberOfDoublets                    */
#define PT_MAX_CELLS_PER_HIT   128u    /* OuterHitOfCell capacity                */
#define PT_MAX_NEIGHBORS       64u     /* CellNeighbors capacity                 */
#define PT_MAX_CELL_TRACKS     64u     /* CellTracks capacity                    */
#define PT_MAX_ACTIVE_LISTS    65536u  /* MAX_NUM_OF_ACTIVE_DOUBLETS (columns)   */
#define PT_MAX_TUPLES          32768u  /* TrackSOA stride / HitContainer bins    */
#define PT_MAX_TUPLE_HITS      163840u /* HitContainer capacity                  */
#define PT_MAX_FITS            24576u  /* MAX_NUM_OF_CONCURRENT_FITS per length  */
#define PT_N_LAYERS            10u
#define PT_N_PHI_BINS          128u
#define PT_HIST_NOFF           1281u   /* 10 layers x 128 bins + 1               */
#define PT_N_MODULES           1856u   /* phase1PixelTopology.jl:23              */
#define PT_N_LAYER_PAIRS       19u

/* ---- status codes ---- */
#define PT_OK                  0
#define PT_ERR_ARG            (-1)
#define PT_ERR_HIP            (-2)
#define PT_ERR_UNSUPPORTED    (-3)
#define PT_ERR_CAPACITY       (-4)

/* ---- pt_counters.overflow bits (silent truncations in the reference) ---- */
#define PT_OVF_DOUBLETS        0x001u /* > max_num_of_doublets cells             */
#define PT_OVF_OUTER_HIT_CELLS 0x002u /* a hit is the outer hit of > 128 cells   */
#define PT_OVF_NEIGHBORS       0x004u /* a cell has > 64 outer neighbours        */
#define PT_OVF_NEIGHBOR_LISTS  0x008u /* > 65535 neighbour lists (extend!)       */
#define PT_OVF_CELL_TRACKS     0x010u /* a cell is on > 64 tuples                */
#define PT_OVF_TRACK_LISTS     0x020u /* > 65535 cell-track lists (extend!)      */
#define PT_OVF_TUPLES          0x040u /* > 32768 tuples                          */
#define PT_OVF_FITS            0x080u /* a length bucket holds > 24576 tuples    */
#define PT_OVF_HIT_TO_TUPLE    0x100u /* hit->tuple association full             */

/* ---- pt_counters.ref_assert bits (the reference would throw here) ---- */
#define PT_REF_TOO_MANY_HITS   0x001u /* n_hits > 49152 (ca_hit_ntuplet_generator_kernels.jl:173) */
#define PT_REF_HIST_LAYER      0x002u /* histogram bin holds a hit of another layer (gpu_pixel_doublets.jl:318-319) */
#define PT_REF_SIZE_CUT_INF    0x004u /* Int(Inf/NaN) in z_size_cut (gpu_pixel_doublets.jl:87) */
#define PT_REF_DEPTH           0x008u /* > 4 doublets in a chain (gpu_ca_cell.jl:234) */
#define PT_REF_NO_CELLS        0x010u /* zero doublets (ca_hit_ntuplet_generator_kernels_impl.jl:145) */
#define PT_REF_FIT_NAN         0x020u /* NaN fast fit (broken_line_fit_on_gpu.jl:89-92) */
#define PT_REF_FIT_CHARGE      0x040u /* q*d > 0 (broken_line.jl:444) */
#define PT_REF_SQRT_DOMAIN     0x080u /* sqrt of a negative number in the fit (Julia DomainError) */
#define PT_REF_BAD_MODULE      0x100u /* B1 hit with module >= 96 (gpu_pixel_doublets.jl:231) */
#define PT_REF_MULT            0x200u /* tuple with >= 8 hits (ca_hit_ntuplet_generator_kernels_impl.jl:189) */

/* internal_error bits: a limit of this library, not of the reference, was hit
 * and the event's results are not valid.  No counterpart in the reference. */
#define PT_INT_LOOKBACK        0x001u /* a prefix-scan wait exceeded its spin bound */
#define PT_INT_NBR_POOL        0x002u /* neighbour pool (lists longer than 4) full   */
#define PT_INT_OVF_LIST        0x004u /* internal overflow list full                 */
#define PT_INT_ROUNDS          0x008u /* duplicate-remover rounds exceeded the shared cells */
#define PT_INT_LIST_SHARE      0x010u /* > 65535 cells on tuples: the reference's extend! would hand
                                         the later cells its last tuple-list column, shared
                                         (simple_matrix.jl:13-19), which is not reproduced (the shared
                                         neighbour column is); or the shared neighbour column's
                                         candidates exceed this library's sort buffer */

/* Quality enum (track_quality.jl:129-136) */
#define PT_QUALITY_BAD         0
#define PT_QUALITY_DUP         1
#define PT_QUALITY_LOOSE       2

/* Algorithm parameters: field-for-field mirror of Params / QualityCuts
 * (ca_hit_ntuplet_generator_kernels.jl:31-87), defaults from
 * ca_hit_ntuplet_generator.jl:36-58 via pt_params_default(). */
typedef struct pt_params {
    uint32_t min_hits_per_ntuplet;             /* 3      */
    uint32_t max_num_of_doublets;              /* 458752 */
    uint8_t  use_riemann_fit;                  /* 0 (no effect, as in the reference: make_tuples
                                                  always runs the broken-line fit)                 */
    uint8_t  fit_5_as_4;                       /* 1 (0 is unsupported: the reference throws)       */
    uint8_t  include_jumping_forward_doublets; /* 1      */
    uint8_t  early_fish_bone;                  /* 1      */
    uint8_t  late_fish_bone;                   /* 0 (1 is unsupported: the reference throws)       */
    uint8_t  ideal_conditions;                 /* 1      */
    uint8_t  do_stats;                         /* 0 (1 is unsupported: the reference throws)       */
    uint8_t  do_cluster_cut;                   /* 1      */
    uint8_t  do_z0_cut;                        /* 1      */
    uint8_t  do_pt_cut;                        /* 1      */
    uint8_t  pad_[2];
    float    pt_min;                           /* 0.9f   */
    float    ca_theta_cut_barrel;              /* 0.002f */
    float    ca_theta_cut_forward;             /* 0.003f */
    float    hard_curv_cut;                    /* 0.0328407224959f */
    float    dca_cut_inner_triplet;            /* 0.15f  */
    float    dca_cut_outer_triplet;            /* 0.25f  */
    /* QualityCuts */
    float    chi2_coeff[4];                    /* 0.68177776, 0.74609577, -0.08035491, 0.00315399 */
    float    chi2_max_pt;                      /* 10     */
    float    chi2_scale;                       /* 30     */
    float    triplet_max_tip, triplet_min_pt, triplet_max_zip;          /* 0.3, 0.5, 12 */
    float    quadruplet_max_tip, quadruplet_min_pt, quadruplet_max_zip; /* 0.5, 0.3, 12 */
} pt_params;

/* Hits of one event: TrackingRecHit2DSOAView (TrackingRecHit2DSOAView.jl:35-54)
 * plus the phi histogram `Hist` (TrackingRecHit2DHeterogeneous.jl:28) and the
 * layer starts.  Arrays have n_hits entries (hist_off has 1281).
 * hist_off and hist_bins both NULL: the stage builds the histogram on the
 * device from iphi and layer_start first, as makeHits does on the host
 * (PixelRecHits.jl:46-48), into the context's event scratch. */
typedef struct pt_hits_view {
    const float    *xg, *yg, *zg, *rg;   /* global position (cm), rg = sqrt(xg^2+yg^2) in f32 */
    const float    *xerr, *yerr;         /* LOCAL variances (gpuPixelRecHits.jl:303-306)      */
    const int16_t  *iphi;                /* unsafe_atan2s(yg, xg, 7)                          */
    const int16_t  *ysize;               /* cluster size y                                    */
    const int16_t  *det_ind;             /* 0-based module index                              */
    const uint32_t *hist_off;            /* [1281] exclusive bin offsets, or NULL (see below) */
    const uint16_t *hist_bins;           /* [n_hits] 1-based hit ids, descending in a bin     */
    uint32_t        n_hits;
    uint32_t        layer_start[11];     /* host copy of hits_layer_start (0-based)           */
} pt_hits_view;

/* Output: TrackSOA = TrackSOAT{32768,163840} (PixelTrackHeterogeneous.jl:10-33,59).
 * state and covariance are column-major (one column of 32768 per component,
 * TrajectoryStateSOA.jl:8-19).  Fields the reference leaves uninitialised
 * (quality of absent tuples, pt/eta/chi2 of unfitted tuples) are not written by
 * the stage; the host-buffer entry points copy the stored tuples' rows of
 * chi2, eta and pt back whole, so there the unfitted ones hold whatever the
 * library's device buffers held.
 * hit_bins and det_bins are defined up to hit_off[32768] - 1 (the container);
 * past it they hold whatever an earlier event left there, as the reference's
 * reused buffers do. */
typedef struct pt_track_soa {
    int32_t  *quality;     /* [32768]                                   */
    float    *chi2;        /* [32768]                                   */
    float    *eta;         /* [32768]                                   */
    float    *pt;          /* [32768]                                   */
    float    *state;       /* [5][32768]  (phi, tip, 1/pt signed, cot(theta), zip) */
    float    *covariance;  /* [15][32768]                               */
    uint32_t *hit_off;     /* [32769] 1-based starts                    */
    uint16_t *hit_bins;    /* [163840] 1-based hit ids                  */
    uint32_t *det_off;     /* [32769]                                   */
    uint16_t *det_bins;    /* [163840] module ids                       */
} pt_track_soa;

/* Per-event counters.  Sizes feed the algorithmic byte model. */
typedef struct pt_counters {
    uint32_t n_hits;
    uint32_t n_inner_slots;     /* sum over layer pairs of inner-layer sizes       */
    uint32_t n_cells;           /* doublets kept                                   */
    uint32_t n_cells_found;     /* doublets passing the cuts before the cap        */
    uint32_t n_connections;     /* neighbour-list entries stored                   */
    uint32_t n_connect_pairs;   /* (cell, neighbour) pairs evaluated in connect    */
    uint32_t n_killed;          /* cells killed by the fishbone                    */
    uint32_t n_tuples;          /* tuples found (may exceed 32768)                 */
    uint32_t n_tuple_hits;      /* hits over all found tuples                      */
    uint32_t n_by_len[8];       /* multiplicity bucket sizes, index = hit count    */
    uint32_t n_fitted;          /* tracks fitted                                   */
    uint32_t n_loose;           /* tracks with quality loose after classification  */
    uint32_t overflow;          /* PT_OVF_* bits                                   */
    uint32_t ref_assert;        /* PT_REF_* bits                                   */
    uint32_t internal_error;    /* PT_INT_* bits: this library failed, results invalid */
} pt_counters;

/* Internal CA state of one event, for the stage-level parity checks.  Arrays
 * are caller-allocated with the given capacities; cell ids are 1-based.
 * nbr_off/trk_off have n_cells+1 entries, ohc_off has n_hits+1 entries. */
typedef struct pt_ca_debug {
    uint32_t  cap_cells, cap_nbr, cap_ohc, cap_trk;
    uint32_t  n_cells;          /* out */
    uint16_t *cell_inner;       /* [cap_cells] 1-based inner hit id  */
    uint16_t *cell_outer;       /* [cap_cells] 1-based outer hit id  */
    uint8_t  *cell_pair;        /* [cap_cells] 0-based layer pair    */
    int32_t  *cell_doublet_id;  /* [cap_cells] 1-based, -1 = killed  */
    uint8_t  *cell_used;        /* [cap_cells] the_used after connect */
    uint32_t *nbr_off;          /* [cap_cells+1] outer neighbours, CSR */
    uint32_t *nbr;              /* [cap_nbr]                          */
    uint32_t *ohc_off;          /* [n_hits+1] is_outer_hit_of_cell    */
    uint32_t *ohc;              /* [cap_ohc]                          */
    uint32_t *trk_off;          /* [cap_cells+1] cell -> tuple ids    */
    uint16_t *trk;              /* [cap_trk]                          */
} pt_ca_debug;

typedef struct pt_ctx pt_ctx;

/* ---------------------------------------------------------------------------
 * Context (one per Julia task / stream, mirrors the TaskLocalValue scratch of
 * ca_hit_ntuplet_generator.jl:9).  max_events = events in flight per batch.
 * ------------------------------------------------------------------------- */
void        pt_params_default(pt_params *p);
int         pt_ctx_create(pt_ctx **out, const pt_params *p, int device, uint32_t max_events);
int         pt_ctx_destroy(pt_ctx *ctx);
const char *pt_last_error(void);
int         pt_device_count(int *n);

/* Module frames (CondFormats pixelCPEforGPU.jl:54-73 DetParams.frame.rot):
 * rot[9*m + k] = (R11,R12,R13,R21,R22,R23,R31,R32,R33) of module m, host memory. */
int pt_ctx_set_frames(pt_ctx *ctx, const float *rot, uint32_t n_modules);

/* Drop-in for make_tuples (ca_hit_ntuplet_generator.jl:65-97 / CAHitNtuplet.produce,
 * ca_hit_ntuplet.jl:10-16): host hits in, host TrackSOA out, synchronous.
 * b_field is the `bf` argument (1/field in GeV, 0.0114256972711507).
 * out_zeroed != 0 tells the library that out->state/covariance are already
 * zero (Julia's TrackSOA() constructor zero-fills them); otherwise the
 * library zero-fills them like the constructor does. */
int pt_make_tuples(pt_ctx *ctx, const pt_hits_view *hits, double b_field,
                   pt_track_soa *out, int out_zeroed, pt_counters *cnt);

/* Same, for n events whose hit arrays and TrackSOAs are already in device
 * memory.  Enqueued on the context's stream; returns before completion.
 * Call pt_ctx_sync, then pt_get_counters. */
int pt_make_tuples_device_async(pt_ctx *ctx, uint32_t n_events, const pt_hits_view *hits,
                                double b_field, const pt_track_soa *out, int out_zeroed);

/* make_tuples for host tasks that share one batcher instead of holding a
 * context each (CAHitNtuplet.produce, ca_hit_ntuplet.jl:10-16, which the
 * reference's tasks call with per-task scratch, ca_hit_ntuplet_generator.jl:9).
 * pt_batcher_make_tuples has the contract of pt_make_tuples and may be called
 * from any number of threads at once: calls that arrive while the GPU runs a
 * batch form the next one (up to max_batch events, 1..64, with the same
 * b_field), which runs as one launch sequence with one copy each way.
 * Up to three batches run on the GPU side by side.  pt_batcher_destroy must
 * not overlap the calls.  Holds four contexts of max_batch event slots and,
 * per slot, about 1.6 MB of pinned input staging and 4.1 MB of pinned
 * readback, four times. */
typedef struct pt_batcher pt_batcher;
int pt_batcher_create(pt_batcher **out, const pt_params *p, int device, uint32_t max_batch, const float *rot,
                      uint32_t n_modules);
int pt_batcher_destroy(pt_batcher *b);
int pt_batcher_make_tuples(pt_batcher *b, const pt_hits_view *hits, double b_field, pt_track_soa *out,
                           int out_zeroed, pt_counters *cnt);
int pt_ctx_sync(pt_ctx *ctx);
int pt_get_counters(pt_ctx *ctx, uint32_t event, pt_counters *cnt);

/* The phi histogram of n events on the device: fill_many_from_vector
 * (hist_to_container.jl:265-270) as called by makeHits (PixelRecHits.jl:46-48)
 * over hits[i].iphi (device memory) with hits[i].layer_start as the offsets,
 * hits below layer_start[10] (the searchsortedfirst layer lookup, including
 * its filing of the first hit after an empty layer into that layer).
 * hist_off[i] (device, 1281 entries) gets the exclusive bin starts and
 * hist_bins[i] (device, layer_start[10] entries) the 1-based hit ids in
 * descending order inside each bin.  Only iphi, n_hits and layer_start of
 * hits[i] are read.  Enqueued on the context's stream. */
int pt_fill_phi_histogram_async(pt_ctx *ctx, uint32_t n_events, const pt_hits_view *hits,
                                uint32_t *const *hist_off, uint16_t *const *hist_bins);

/* Input of the pixel vertex finder for one event: the WorkSpace fields
 * (gpu_vertex_finder.jl:9-32) that load_tracks fills.  Device arrays of
 * PT_MAX_TUPLES (= MAX_TRACKS, z_vertex_soa.jl:2) entries. */
typedef struct pt_vertex_ws {
    uint32_t *n_tracks;     /* [1] selected tracks                     */
    uint16_t *index_track;  /* 1-based TrackSOA row of each            */
    float    *zt;           /* zip (state column 5)                    */
    float    *ezt2;         /* covariance column 15                    */
    float    *ptt2;         /* pt^2                                    */
} pt_vertex_ws;

/* load_tracks (gpu_vertex_finder.jl:51-82), the track selection of
 * gpuVertexFinder.make (:208-211; pt_min = 0.5 in PixelVertexProducerCUDA,
 * pixel_vertex_producer_cuda.jl:19), over n device TrackSOAs: rows in order
 * up to the first row without hits; kept when n_hits >= 4, quality == loose
 * and !(pt < pt_min).  Enqueued on the context's stream (after a batch of
 * pt_make_tuples_device_async, in stream order). */
int pt_load_tracks_async(pt_ctx *ctx, uint32_t n_events, const pt_track_soa *tracks, float pt_min,
                         const pt_vertex_ws *ws);

/* setHitsLayerStart (PixelRecHits.jl:20-27): layer_start[L] =
 * hits_module_start[first module of layer L] for the phase-1 layer table
 * (phase1PixelTopology.jl:24-36); hits_module_start has 1857 entries. Host only. */
void pt_set_hits_layer_start(const uint32_t *hits_module_start, uint32_t layer_start[11]);

/* Internal CA state of event `event` of the last batch (device -> host copy). */
int pt_get_ca_debug(pt_ctx *ctx, uint32_t event, pt_ca_debug *dbg);

/* Events of a batch are split into n_streams groups launched on separate
 * streams so their kernels overlap (default 4; 1..8).  No reference counterpart:
 * the reference runs events on host threads (EventProcessor.jl:43). */
int pt_ctx_set_streams(pt_ctx *ctx, int n_streams);

/* Precision of the stage's fit: PT_FIT_FP64 (default, the reference's) or
 * PT_FIT_MIXED (see pt_fit_bl_device_async).  Set before pt_pipeline_init. */
int pt_ctx_set_fit_precision(pt_ctx *ctx, int precision);

/* ---------------------------------------------------------------------------
 * Events in flight under captured hipGraphs (no reference counterpart: the
 * reference runs one event per host thread, EventProcessor.jl:43).  Group g
 * (0 <= g < n_groups <= 8) owns the context's event slots
 * [g*group_size, (g+1)*group_size) and a stream of its own; the stage's launch
 * sequence for one group (grids sized for the capacity limits) is captured
 * once into a hipGraph and replayed for every submission.  Needs
 * max_events >= group_size * n_groups and the module frames.
 * pt_pipeline_submit enqueues exactly group_size device-resident events and
 * returns; it first waits for the group's previous submission.
 * pt_pipeline_wait blocks until the group is done; *ms = GPU time of the
 * replay (HIP events on the group's stream).
 * ------------------------------------------------------------------------- */
int pt_pipeline_init(pt_ctx *ctx, uint32_t group_size, uint32_t n_groups, double b_field, int out_zeroed);
int pt_pipeline_submit(pt_ctx *ctx, uint32_t group, const pt_hits_view *hits, const pt_track_soa *out);
int pt_pipeline_wait(pt_ctx *ctx, uint32_t group, float *ms);
int pt_pipeline_get_counters(pt_ctx *ctx, uint32_t group, uint32_t event, pt_counters *cnt);

/* Per-kernel timing of the last batch, milliseconds, from HIP events recorded
 * on the context stream after every kernel when enabled (enable=1 before the
 * batch; a timed batch runs on one stream).  *n_stages: capacity of ms/names
 * on input, entries on output. */
#define PT_N_STAGES 48
int pt_ctx_enable_timing(pt_ctx *ctx, int enable);
int pt_get_stage_times(pt_ctx *ctx, float *ms, const char **names, int *n_stages);

/* Tile sizes of the CA kernels whose wavefronts take work by ticket (no
 * reference counterpart).  Every setting gives the same results; only the
 * speed differs, and the best one depends on the hit density (cells per hit
 * grow about fourfold from PU50 to PU200).  Defaults: 24, 192, 4, 3, 16, 1. */
typedef struct pt_ca_tuning {
    uint32_t connect_hits;   /* hits per connect ticket: 4, 8, 16, 24 or 32             */
    uint32_t connect_stage;  /* cell columns a connect wavefront stages in LDS: 128, 192,
                                256 or 384                                               */
    uint32_t count_rounds;   /* doublet count: rounds of 64 candidate loads in flight:
                                1, 2 or 4                                                */
    uint32_t fill_rounds;    /* doublet fill: rounds of 64 cells per chunk: 1, 2 or 3    */
    uint32_t fish_hits;      /* hits per fishbone ticket: 8, 16 or 32                    */
    uint32_t classify_split; /* classification as five kernels, several blocks per event
                                for the per-tuple phases (1), or one block per event (0) */
} pt_ca_tuning;
/* PT_ERR_ARG for a value not listed above.  Used by later batches and by
 * pipelines initialised later. */
int pt_ca_set_tuning(pt_ctx *ctx, const pt_ca_tuning *t);
int pt_ca_get_tuning(pt_ctx *ctx, pt_ca_tuning *t);
/* Tunes the six settings on the given batch (device-resident events as for
 * pt_make_tuples_device_async; the outputs are overwritten), one kernel at a
 * time with the others at their current setting: each candidate runs the
 * whole stage three times and is scored by its kernels' own duration (the
 * lowest of the last two runs).  Keeps the fastest in the context, returns it
 * in *best and the stage's time per batch with it in *ms.  Synchronous. */
int pt_ca_autotune(pt_ctx *ctx, uint32_t n_events, const pt_hits_view *hits, double b_field,
                   const pt_track_soa *out, pt_ca_tuning *best, float *ms);

/* ---------------------------------------------------------------------------
 * Broken-line fit on pre-gathered tracks (BrokenLineFitOnGPU kernelBLFastFit +
 * kernelBLFit, broken_line_fit_on_gpu.jl:24-163), for fit-only workloads.
 * Layout, all device pointers, track-minor (coalesced):
 *   hits[(c*5 + k)*n + t]   c = x,y,z          (float, cm)
 *   ge  [(c*5 + k)*n + t]   c = xx,xy,yy,xz,yz,zz global covariance (float)
 *   nhit[t]                 3, 4 or 5 (5 is fitted on its first 4 hits, fit5as4)
 * Outputs (device): state[k*n + t] (5), cov[k*n + t] (15), pt/eta/chi2[t].
 * ------------------------------------------------------------------------- */
#define PT_FIT_FP64   0   /* plain double-precision fit, the reference's method        */
#define PT_FIT_MIXED  1   /* float weights, scattering terms and covariance propagation;
                             double fast fit, arc lengths, band solve and translations   */
int pt_fit_bl_device_async(pt_ctx *ctx, uint32_t n_tracks, const float *hits, const float *ge,
                           const uint8_t *nhit, double b_field, int precision,
                           float *state, float *cov, float *pt, float *eta, float *chi2);

/* Schedules of a fit-only batch of uneven lengths (no reference counterpart:
 * the reference loops over one length at a time, broken_line_fit_on_gpu.jl:166-193).
 *   PT_SCHED_PLAIN       one thread per track in input order (mixed lengths
 *                        in one wavefront run both bodies)
 *   PT_SCHED_PER_LENGTH  a partition kernel, then one launch per length
 *   PT_SCHED_PERSISTENT  one launch; wavefronts take tiles of tracks by
 *                        ticket and fit each tile's lengths separately
 * pt_fit_bl_device_async uses PT_SCHED_PERSISTENT with the context's tuning. */
#define PT_SCHED_PLAIN       0
#define PT_SCHED_PER_LENGTH  1
#define PT_SCHED_PERSISTENT  2
typedef struct pt_fit_tuning {
    uint32_t block;   /* threads per block: 64, 128, 192 or 256             */
    uint32_t tile;    /* tracks per ticket, a multiple of 64 (persistent)   */
    uint32_t blocks;  /* blocks of the persistent launch                    */
} pt_fit_tuning;
int pt_fit_bl_sched_async(pt_ctx *ctx, uint32_t n_tracks, const float *hits, const float *ge,
                          const uint8_t *nhit, double b_field, int precision, int schedule,
                          float *state, float *cov, float *pt, float *eta, float *chi2);
/* Same as pt_fit_bl_sched_async, allocating nothing: no reference counterpart;
 * for callers that cannot allocate device memory during the call (code under
 * hipGraph capture, tight serving loops).  pt_fit_bl_sched_async grows its
 * per-length lists with hipMalloc on demand; this variant takes its counters
 * and lists from a fixed device array of the library (one per GPU, about
 * 2 MiB, kept until the process ends), so n_tracks is at most
 * PT_NOALLOC_MAX_TRACKS (PT_ERR_CAPACITY above).  All calls on one GPU share
 * that array: calls that overlap on one GPU (other streams, other contexts,
 * other threads) race and must be serialised by the caller. */
#define PT_NOALLOC_MAX_TRACKS 262144u
int pt_fit_bl_sched_noalloc_async(pt_ctx *ctx, uint32_t n_tracks, const float *hits, const float *ge,
                                  const uint8_t *nhit, double b_field, int precision, int schedule,
                                  float *state, float *cov, float *pt, float *eta, float *chi2);
/* Tuning of the persistent schedule per precision (PT_FIT_FP64, PT_FIT_MIXED,
 * PT_FIT_ADJOINT or PT_FIT_RIEMANN; default: 256 threads, 512-track tiles,
 * 4 blocks per CU). */
int pt_fit_set_tuning(pt_ctx *ctx, int precision, const pt_fit_tuning *t);
/* Schedule used by pt_fit_bl_device_async for PT_FIT_FP64 / PT_FIT_MIXED
 * (default PT_SCHED_PERSISTENT) and by the adjoint entry points for
 * PT_FIT_ADJOINT (default PT_SCHED_PLAIN: the fastest for batches of one
 * length; PT_SCHED_PERSISTENT is faster for mixed lengths).  No reference
 * counterpart. */
int pt_fit_set_schedule(pt_ctx *ctx, int precision, int schedule);
int pt_fit_get_tuning(pt_ctx *ctx, int precision, pt_fit_tuning *t);
/* Times every candidate launch shape of the persistent schedule on the given
 * batch (device pointers; outputs are overwritten), keeps the fastest in the
 * context and returns it.  *ms: its time per launch.  precision: PT_FIT_FP64,
 * PT_FIT_MIXED or PT_FIT_RIEMANN (the Riemann fit's slot). */
int pt_fit_autotune(pt_ctx *ctx, uint32_t n_tracks, const float *hits, const float *ge, const uint8_t *nhit,
                    double b_field, int precision, float *state, float *cov, float *pt, float *eta,
                    float *chi2, pt_fit_tuning *best, float *ms);

/* ---------------------------------------------------------------------------
 * Riemann helix fit of pre-gathered tracks (launch_riemann_kernels!,
 * helix_fit_on_gpu.jl:94-96, which only prints; the method is helix_fit of
 * riemann_fit.jl:989-1023: fast_fit, circle_fit on the paraboloid with its
 * error propagation, line_fit in the rotated s-z plane), FP64.  The
 * reference's make_tuples never runs it (use_riemann_fit is not read,
 * ca_hit_ntuplet_generator.jl:83-93), so the stage keeps the broken-line fit
 * and this is a fit-only entry point.  Same layout as pt_fit_bl_device_async;
 * outputs in the TrackSOA convention (perigee phi, tip, q/pT, cot(theta), zip
 * through fromCircleToPerigee and copyFromCircle!, fit_utils.jl:115-141,
 * TrajectoryStateSOA.jl:20-44); chi2 = (chi2_circle + chi2_line) / (2N - 5);
 * 5-hit tracks are fitted on their first 4 hits.  PT_FIT_RIEMANN is its slot
 * in pt_fit_set_tuning / pt_fit_set_schedule / pt_fit_autotune (default:
 * persistent schedule).
 * ------------------------------------------------------------------------- */
#define PT_FIT_RIEMANN 3
int pt_fit_riemann_device_async(pt_ctx *ctx, uint32_t n_tracks, const float *hits, const float *ge,
                                const uint8_t *nhit, double b_field, float *state, float *cov, float *pt, float *eta,
                                float *chi2);
/* The same under a given schedule (PT_SCHED_*; no reference counterpart). */
int pt_fit_riemann_sched_async(pt_ctx *ctx, uint32_t n_tracks, const float *hits, const float *ge,
                               const uint8_t *nhit, double b_field, int schedule, float *state, float *cov, float *pt,
                               float *eta, float *chi2);

/* ---------------------------------------------------------------------------
 * Adjoint of the mixed-precision broken-line fit (no reference counterpart:
 * the reference fit has no derivatives).  Tracks as for the fit-only entry
 * points, but hit errors are given as local sigmas and module ids, as the
 * rechit producer stores them (gpuPixelRecHits.jl:303-306), and rotated to
 * the global frame with the context's module frames (pt_ctx_set_frames):
 *   sig[(c*5 + k)*n + t]  c = sigma_rphi, sigma_z (cm)
 *   det[k*n + t]          0-based module id
 * Gradients: grad[(c*5 + k)*n + t], c = x, y, z, sigma_rphi, sigma_z.  The
 * forward fit is recomputed inside the backward kernel (nothing is stored
 * between the two).  Hits a track does not use get zero gradient.
 * ------------------------------------------------------------------------- */
#define PT_FIT_ADJOINT 2   /* tuning slot of the adjoint kernels (pt_fit_set_tuning) */
/* ge[(c*5 + k)*n + t] from sig and det: toGlobal of SOARotation.jl:139-147. */
int pt_hit_errors_to_global_async(pt_ctx *ctx, uint32_t n_tracks, const float *sig, const int16_t *det, float *ge);
/* L = chi2_line + chi2_circle + |state - truth|^2 per track (phi difference
 * modulo 2 pi; truth[k*n + t] in the state convention) and dL/d(hits).
 * state, chi2 (forward results) may be NULL. */
int pt_fit_bl_loss_grad_async(pt_ctx *ctx, uint32_t n_tracks, const float *hits, const float *sig,
                              const int16_t *det, const uint8_t *nhit, const float *truth, double b_field,
                              float *state, float *chi2, float *loss, float *grad);
/* Pullback (vector-Jacobian product) of (state[5], chi2) -> hits: the
 * backward half of a ChainRulesCore rrule around the mixed fit.  dstate[k*n + t],
 * dchi2[t] are the output cotangents. */
int pt_fit_bl_vjp_async(pt_ctx *ctx, uint32_t n_tracks, const float *hits, const float *sig, const int16_t *det,
                        const uint8_t *nhit, double b_field, const float *dstate, const float *dchi2, float *grad);

/* Device memory helpers (so hosts without a device allocator can stage data). */
int pt_malloc(pt_ctx *ctx, void **ptr, size_t bytes);
int pt_free(pt_ctx *ctx, void *ptr);
int pt_memcpy_h2d(pt_ctx *ctx, void *dst, const void *src, size_t bytes);
int pt_memcpy_d2h(pt_ctx *ctx, void *dst, const void *src, size_t bytes);
int pt_memset(pt_ctx *ctx, void *dst, int value, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* PIXELTRACK_H */
