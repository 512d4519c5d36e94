// pt_device.hpp - device-side layout shared by the pixel-track kernels.
//
// One "event slot" owns a fixed-capacity scratch region in HBM sized for the
// reference limits (49,152 hits, 458,752 doublets, 32,768 tuples), so a batch
// of events runs with no allocation.  Kernels are batched: blockIdx.y selects
// the event, blockIdx.x strides over the event's items.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pixeltrack.h"

namespace pt {

// Device pointers in the descriptors are typed as global-memory pointers in
// the device pass, so every access compiles to global_* instructions (a flat
// access would also count against the LDS wait counter and stall LDS waits
// behind memory loads); the host pass sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define PT_G __attribute__((address_space(1)))
#else
#define PT_G
#endif

constexpr uint32_t kMaxHits = PT_MAX_HITS;
constexpr uint32_t kMaxCells = PT_MAX_DOUBLETS;
constexpr uint32_t kMaxSlots = 6u * PT_MAX_HITS;   // layer 0 is the inner layer of 6 pairs
constexpr uint32_t kMaxTuples = PT_MAX_TUPLES;
constexpr uint32_t kNbrPool = 1024u * 1024u;   // neighbour entries beyond the 4 kept inline, per event
constexpr uint32_t kOhcStride = PT_MAX_CELLS_PER_HIT;  // outer-hit list capacity per hit
// Entry r (< kOhcStride) of outer hit h's cell list, in a buffer of kOhcSize entries.
constexpr size_t kOhcSize = (size_t)kMaxHits * kOhcStride;
__host__ __device__ __forceinline__ size_t ohc_index(uint32_t h, uint32_t r) { return (size_t)h * kOhcStride + r; }
constexpr uint32_t kOhcOvfCap = kMaxCells / (PT_MAX_CELLS_PER_HIT + 1) + 1;  // hits that can exceed it
constexpr uint32_t kTrkOvfCap = 4u * PT_MAX_TUPLES / (PT_MAX_CELL_TRACKS + 1) + 8u;
constexpr uint32_t kSlotTile = 1024;              // slots per prefix-sum tile of the doublet scan
constexpr uint32_t kDblAgg = (kMaxSlots + kSlotTile - 1) / kSlotTile;
constexpr uint32_t kDfsTile = 1024;               // cells per prefix-sum tile of the ntuplet scan
constexpr uint32_t kDfsAgg = (kMaxCells + kDfsTile - 1) / kDfsTile;
constexpr uint32_t kTupBuf = 2u * PT_MAX_TUPLES;   // tuples kept between the ntuplet count and emit
constexpr uint32_t kDcCap = 2u * PT_MAX_TUPLES;    // cells on >= 2 stored tuples (<= 4 cells per tuple / 2)

// Layer pairs (gpu_pixel_doublets.jl:20-28) and per-pair cuts (:29-56)
struct PairTables {
    int8_t inner[19], outer[19];
    int8_t by_inner[10][6];   // pairs whose inner layer is L, in pair order
    int8_t n_by_inner[10];
    int8_t pair_of[10][10];   // layer pair of (inner layer, outer layer), -1 if none
    int16_t phi_cut[19];
    float min_z[19], max_z[19], max_r[19];  // integral values: exact in float
};

// Scratch of one event slot (device pointers into the context arena).
struct Scratch {
    PT_G uint32_t *slot_cnt;     // [kMaxSlots+1]  passing candidates per inner slot, then the
                            //                slot's first cell id (total at [n_slots])
    PT_G uint64_t *slot_mask;    // [kMaxSlots]    passing candidates among the slot's first 64
    PT_G uint64_t *slot_mask_ext;  // [3*kMaxSlots] candidates 64..255, for slots with more than 64
    PT_G uint4 *slot_pos;        // [kMaxSlots]    candidate positions (a0, na, b0) | n | pair << 16
    PT_G uint32_t *dbl_agg;      // [kDblAgg]      passing candidates per tile of kSlotTile slots
    PT_G uint32_t *cell_io;      // [kMaxCells]    inner | outer << 16 (1-based hit ids)
    PT_G uint8_t *cell_flags;    // [2*kMaxCells]  byte 2c+1 of the pair array: kUsed (connect), kKilled (fishbone)
    PT_G uint8_t *cell_pair;     // [kMaxCells]    layer pair
    PT_G uint8_t *cell_nn;       // [2*kMaxCells]  byte 2c: stored outer neighbours (<= 64); cell_flags = cell_nn + 1,
                            //                so one 16-bit load reads both (index 2c)
    PT_G float4 *cell_oxyzr;     // [kMaxCells]    (x, y, z, r) of the outer hit
    PT_G uint4 *nbr4;            // [kMaxCells]    first 4 neighbours (0-based cell ids, ascending)
    PT_G uint32_t *nbr_ext;      // [kMaxCells]    pool offset of neighbours 5..64
    PT_G uint32_t *nbr_pool;     // [kNbrPool]
    PT_G uint32_t *ext_list;     // [kMaxCells]    cells with more than 4 neighbours
    PT_G uint2 *trk;             // [kMaxCells]    (tuples through the cell, largest hit count among them)
    PT_G uint32_t *start_list;   // [kMaxCells]    ntuplet start candidates (any order)
    PT_G uint64_t *start_bits;   // [kMaxCells/64] the same as a bit set
    PT_G uint64_t *tup;          // [kMaxCells]    start cell: (tuples << 32 | hits), then the prefix
    PT_G uint64_t *dfs_agg;      // [kDfsAgg]      the same summed per tile of kDfsTile cells
    PT_G uint2 *tup_where;       // [kMaxCells]    per start-list entry: (buffer offset, tuples) or ~0
    PT_G uint4 *tup_buf;         // [kTupBuf]      cells of the tuples kept by the count pass
    PT_G uint32_t *ohc_cursor;   // [kMaxHits]     cells whose outer hit is h (all, not capped)
    PT_G uint32_t *fish_cnt;     // [kMaxHits]     cells ending on h that connect used
    PT_G uint32_t *fish_hits;    // [kMaxHits]     hits with at least two of them (fishbone work)
    PT_G uint2 *ohc_list;        // [kOhcSize] (cell, inner hit) 0-based, unordered unless
                            //                the hit overflowed (then the first 128 cells, ascending)
    PT_G uint32_t *ohc_ovf;      // [kOhcOvfCap]   hits with more than kOhcStride cells
    PT_G uint32_t *trk_ovf;      // [kTrkOvfCap]   cells whose tuple list needs the exact path
    PT_G uint32_t *tuple_cells;  // [kMaxTuples*4] cells of each stored tuple
    PT_G uint32_t *tm_off;       // [16]           tuple multiplicity (9 used)
    PT_G uint16_t *tm_bins;      // [kMaxTuples]
    PT_G uint64_t *h2t_win;      // [kMaxHits]     min over the loose tuples on the hit of
                            //                (15 - hits) << 60 | |tip| bits << 15 | (32767 - tuple id)
    PT_G uint32_t *misc;         // [16]           internal scalars (see Misc)
    PT_G float4 *hxyzr;          // [kMaxHits]     packed (x, y, z, r) per hit
    PT_G uint4 *hdbl;            // [kMaxHits]     packed (z, r, iphi | det << 16, ysize) per hit
    PT_G uint4 *binrec;          // [kMaxHits]     per histogram position: (z, r, iphi | det << 16,
                            //                ysize | hit << 16) of the hit stored there
    PT_G float2 *binxy;          // [kMaxHits]     per histogram position: (x, y)
    // classification: cells on >= 2 stored tuples ("shared cells") and their tuple lists
    PT_G uint32_t *dc_slot;      // [kMaxCells]    cell -> shared-cell slot (valid when trk.x >= 2)
    PT_G uint32_t *dc_cells;     // [kDcCap]       slot -> cell
    PT_G uint32_t *dc_off;       // [kDcCap]       slot -> start of its list in dc_ent
    PT_G uint32_t *dc_cur;       // [kDcCap]       slot -> list length
    PT_G uint32_t *dc_done;      // [kDcCap]       round in which the duplicate remover resolved it
    PT_G uint4 *dc_pred;         // [kDcCap]       up to 4 earlier shared cells on its tuples (slots)
    PT_G uint32_t *dc_ent;       // [4*kMaxTuples] tuple ids (0-based)
    // phi histogram built on the device when the caller gives none
    PT_G uint32_t *hist_off;     // [1281]
    PT_G uint16_t *hist_bins;    // [kMaxHits]
};

constexpr uint8_t kUsed = 1, kKilled = 2;

// One contiguous piece of a device array copied into a readback block
// (launch_pack_segments): n16 16-bit words from src to dst + dst_off.
struct PackSeg {
    const uint16_t *src;
    uint64_t dst_off;
    uint32_t n16, pad;
};

enum Misc : int {
    kMiscStart = 0,       // entries of start_list
    kMiscExt = 1,         // entries of ext_list
    kMiscTkZero = 2,      // next chunk of state/covariance to clear (doublet count)
    kMiscOhcOvf = 3,      // entries of ohc_ovf
    kMiscTrkOvf = 4,      // entries of trk_ovf
    kMiscPool = 5,        // neighbour pool entries used
    kMiscNbrLists = 6,    // cells with at least one neighbour
    kMiscTrkLists = 7,    // cells with at least one tuple
    kMiscTkCount = 8,     // next tile (ticket) of the doublet count
    kMiscTkFill = 9,      // ... of the doublet fill
    kMiscTkConn = 10,     // ... of connect
    kMiscTkFish = 11,     // ... of the fishbone
    kMiscTupBuf = 12,     // entries of tup_buf
    kMiscDc = 13,         // shared cells (classification)
    kMiscH2tEnt = 14,     // hit -> loose-tuple entries (classification)
    kMiscFish = 15,       // entries of fish_hits
};

// Per-event descriptor: inputs, outputs and scratch.  Built on the host,
// copied to the device once per batch.
struct EvDesc {
    PT_G const float *xg, *yg, *zg, *rg, *xerr, *yerr;
    PT_G const int16_t *iphi, *ysize, *det;
    PT_G const uint32_t *hist_off;
    PT_G const uint16_t *hist_bins;
    uint32_t n_hits;
    uint32_t n_pairs;         // layer pairs in use (19, 15 or 13)
    uint32_t n_slots;         // sum of inner-layer sizes over the pairs
    uint32_t layer_start[11];
    uint32_t slot_cum[20];    // inclusive cumulative inner-layer sizes per pair
    // when set, the phi histogram is built here from iphi and layer_start
    // first (hist_off / hist_bins then point to the same arrays)
    PT_G uint32_t *hist_off_build;
    PT_G uint16_t *hist_bins_build;
    // vertex-finder input (load_tracks) written from the TrackSOA outputs
    PT_G uint32_t *ws_n;
    PT_G uint16_t *ws_index;
    PT_G float *ws_zt, *ws_ezt2, *ws_ptt2;
    // outputs (TrackSOA)
    PT_G int32_t *quality;
    PT_G float *chi2, *eta, *pt, *state, *cov;
    PT_G uint32_t *hit_off;
    PT_G uint16_t *hit_bins;
    PT_G uint32_t *det_off;
    PT_G uint16_t *det_bins;
    PT_G pt_counters *cnt;         // device counters of this event
    Scratch s;
};

struct Config {
    pt_params p;
    uint32_t xcd;             // XCD of the batch's first event (EvGrid::xcd)
    double B;                 // Float64(Float32(b_field))
    PT_G const float *frames;      // [1856*9]
    PairTables pairs;
    int fit_precision;        // PT_FIT_FP64 (the reference's) or PT_FIT_MIXED
    // tile sizes of the ticketed CA kernels (pt_ca_tuning; every setting gives
    // the same results)
    uint32_t conn_hits = 24, conn_stage = 192, count_rounds = 4, fill_rounds = 3, fish_hits = 16;
    uint32_t classify_split = 1;  // classification: five kernels (1) or one block per event (0)
};

// Fit-only batch (pre-gathered tracks, track-minor layout of pixeltrack.h).
struct FlatArgs {
    PT_G const float *hits, *ge;
    PT_G const uint8_t *nhit;
    uint32_t n;
    double B;
    PT_G float *state, *cov, *pt, *eta, *chi2;
};

// Adjoint of the mixed fit over a fit-only batch (hit errors as local sigmas
// sig[(c*5+k)*n+t], c = rphi, z, and module ids det[k*n+t]).
struct AdjArgs {
    PT_G const float *hits, *sig;
    PT_G const int16_t *det;
    PT_G const uint8_t *nhit;
    uint32_t n;
    double B;
    PT_G const float *frames;  // [1856*9]
    PT_G const float *truth;   // [5][n]  loss mode
    PT_G const float *dstate;  // [5][n]  pullback mode
    PT_G const float *dchi2;   // [n]     pullback mode
    PT_G float *state, *chi2;  // forward outputs (may be null)
    PT_G float *loss;          // [n]     loss mode
    PT_G float *grad;          // [25][n] grad[(c*5+k)*n+t], c = x, y, z, sigma_rphi, sigma_z
};

// Launch shape of a fit-only batch: schedule (PT_SCHED_*), threads per block,
// tracks per ticket (persistent) and blocks (persistent); device scratch for
// the per-length lists (2 n ids) and the counters / ticket (4 words).
struct FitLaunch {
    int schedule;
    uint32_t block, tile, blocks;
    PT_G uint32_t *lists;
    PT_G uint32_t *counts;
};

// Host-side kernel timing marks (HIP events recorded on the launch stream).
struct Marks {
    static constexpr int kMax = 48;
    hipEvent_t ev[kMax];
    const char *name[kMax];
    int n = 0;
    void record(const char *nm, hipStream_t st) {
        if (n < kMax) {
            name[n] = nm;
            (void)hipEventRecord(ev[n], st);
            ++n;
        }
    }
};

// hist_to_container.jl:81-86 : ((t >> 9) & 127) + 1 for Int16 t
__device__ __forceinline__ int hist_bin(int16_t t) { return (int)(((uint16_t)t) >> 9) + 1; }

__device__ __forceinline__ int layer_of(const EvDesc &e, uint32_t h0) {
    int L = 0;
#pragma unroll
    for (int k = 1; k < 10; ++k) L += (h0 >= e.layer_start[k]) ? 1 : 0;
    return L;
}

// ---------------------------------------------------------------- wave / block scan
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// exclusive scan across the block; lds needs blockDim/64 + 1 words; returns the
// exclusive prefix of v and the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *lds, uint32_t &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        uint32_t w = (lane < nw) ? lds[lane] : 0u;
        uint32_t wi = wave_incl_scan(w);
        if (lane < nw) lds[lane] = wi - w;
        if (lane == nw - 1) lds[nw] = wi;
    }
    __syncthreads();
    uint32_t r = lds[wid] + inc - v;
    total = lds[nw];
    __syncthreads();
    return r;
}

// Batched launch shape.  A kernel over a batch runs a 1-D grid of
// 8 * per_event * ceil(n_events / 8) blocks.  Workgroups are dispatched
// round-robin over the 8 XCDs in linear order, so block L runs on XCD L % 8;
// event e gets the blocks L = e % 8 + 8 * (per_event * (e / 8) + b),
// b < per_event, all on one XCD, and its working set (hit records, cells,
// lists) stays in that XCD's L2 instead of being pulled into all eight.
// xcd rotates the events over the XCDs: event e runs on XCD (e + xcd) % 8, so
// contexts that each run one event at a time spread over the chip.
struct EvGrid {
    uint32_t n_events, per_event;
    uint32_t xcd = 0;
};

__device__ __forceinline__ bool ev_block(const EvGrid &g, uint32_t &ev, uint32_t &b) {
    const uint32_t L = blockIdx.x;
    const uint32_t s = L >> 3;
    const uint32_t q = s / g.per_event;
    ev = ((L - g.xcd) & 7u) + 8u * q;
    b = s - q * g.per_event;
    return ev < g.n_events;
}

inline dim3 ev_grid_dim(const EvGrid &g) { return dim3(8u * g.per_event * ((g.n_events + 7u) / 8u)); }

// Blocks per event for a grid-stride kernel over `items` items per event:
// enough to cover the items, at most what keeps about 2048 blocks in flight.
inline EvGrid ev_grid(uint32_t items, uint32_t n_events, uint32_t xcd, uint32_t bs = 256) {
    uint32_t per = (items + bs - 1) / bs;
    const uint32_t target = (2048u + n_events - 1) / n_events;
    if (per > target) per = target;
    if (per < 1) per = 1;
    return EvGrid{n_events, per, xcd};
}

// Grid-stride loop over the items of this block's event; blocks beyond the
// last event return.
#define PT_EVENT_LOOP(e_, i_, n_)                                                                   \
    uint32_t pt_ev_, pt_b_;                                                                         \
    if (!ev_block(grid, pt_ev_, pt_b_)) return;                                                     \
    const EvDesc &e_ = evs[pt_ev_];                                                                 \
    for (uint32_t i_ = pt_b_ * blockDim.x + threadIdx.x; i_ < (n_); i_ += grid.per_event * blockDim.x)

}  // namespace pt
