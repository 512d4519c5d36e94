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

constexpr uint32_t kMaxHits = PT_MAX_HITS;
constexpr uint32_t kMaxCells = PT_MAX_DOUBLETS;
constexpr uint32_t kMaxSlots = 6u * PT_MAX_HITS;   // layer 0 is the inner layer of 6 pairs
constexpr uint32_t kMaxTuples = PT_MAX_TUPLES;
constexpr uint32_t kNbrCap = 2u * 1024u * 1024u;    // neighbour entries per event
constexpr uint32_t kTrkCap = 4u * PT_MAX_TUPLES;    // cell->tuple entries (<= 4 cells per tuple)
constexpr uint32_t kScanBlock = 1024;
constexpr uint32_t kScanItems = 8;

// Layer pairs (gpu_pixel_doublets.jl:20-28) and per-pair cuts (:29-56)
struct PairTables {
    int8_t inner[19], outer[19];
    int16_t phi_cut[19];
    float min_z[19], max_z[19], max_r[19];  // integral values: exact in float
};

// Scratch of one event slot (device pointers into the context arena).
struct Scratch {
    uint32_t *slot_cnt;     // [kMaxSlots+1]  doublets per inner slot -> exclusive offsets
    uint16_t *cell_inner;   // [kMaxCells]    1-based hit ids
    uint16_t *cell_outer;   // [kMaxCells]
    uint8_t *cell_pair;     // [kMaxCells]    0-based layer pair
    uint8_t *cell_used;     // [kMaxCells]
    uint8_t *cell_killed;   // [kMaxCells]
    uint32_t *ohc_cnt;      // [kMaxHits+1]   cells per outer hit -> exclusive offsets
    uint32_t *ohc_cursor;   // [kMaxHits]
    uint32_t *ohc_list;     // [kMaxCells]    0-based cell ids, ascending per hit
    uint32_t *nn;           // [kMaxCells+1]  stored neighbours per cell -> offsets
    uint32_t *nbr;          // [kNbrCap]      0-based cell ids, ascending
    uint32_t *tcnt;         // [kMaxCells+1]  tuples per start cell -> offsets
    uint32_t *hcnt;         // [kMaxCells+1]  tuple hits per start cell -> offsets
    uint32_t *tuple_cells;  // [kMaxTuples*4] cells of each stored tuple
    uint32_t *trk_cnt;      // [kMaxCells+1]  tuples per cell -> offsets
    uint32_t *trk_cursor;   // [kMaxCells]
    uint16_t *trk_list;     // [kTrkCap]      1-based tuple ids, ascending per cell
    uint32_t *tm_off;       // [16]           tuple multiplicity (9 used)
    uint16_t *tm_bins;      // [kMaxTuples]
    uint32_t *h2t_off;      // [kMaxHits+2]   hit -> loose tuples (classification)
    uint16_t *h2t_bins;     // [4*PT_MAX_TUPLES... 196608]
    uint32_t *h2t_cursor;   // [kMaxHits]
    uint32_t *misc;         // [16]           internal scalars (see Misc)
};

enum Misc : int {
    kMiscTupleOverflowHits = 0,
};

// Per-event descriptor: inputs, outputs and scratch.  Built on the host,
// copied to the device once per batch.
struct EvDesc {
    const float *xg, *yg, *zg, *rg, *xerr, *yerr;
    const int16_t *iphi, *ysize, *det;
    const uint32_t *hist_off;
    const uint16_t *hist_bins;
    uint32_t n_hits;
    uint32_t n_pairs;         // layer pairs in use (19, 15 or 13)
    uint32_t n_slots;         // sum of inner-layer sizes over the pairs
    uint32_t layer_start[11];
    uint32_t slot_cum[20];    // inclusive cumulative inner-layer sizes per pair
    // outputs (TrackSOA)
    int32_t *quality;
    float *chi2, *eta, *pt, *state, *cov;
    uint32_t *hit_off;
    uint16_t *hit_bins;
    uint32_t *det_off;
    uint16_t *det_bins;
    pt_counters *cnt;         // device counters of this event
    Scratch s;
};

struct Config {
    pt_params p;
    double B;                 // Float64(Float32(b_field))
    const float *frames;      // [1856*9]
    PairTables pairs;
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

// Grid-stride helpers for the batched launch shape (x: items, y: event).
#define PT_EVENT_LOOP(e_, i_, n_)                                                               \
    const EvDesc &e_ = evs[blockIdx.y];                                                         \
    for (uint32_t i_ = blockIdx.x * blockDim.x + threadIdx.x; i_ < (n_); i_ += gridDim.x * blockDim.x)

}  // namespace pt
