// pt_hist.hip - the phi histogram of an event's hits on gfx950.
//
// fill_many_from_vector (hist_to_container.jl:217-270) over iphi, with
// hits_layer_start as the 11 histogram offsets (makeHits, PixelRecHits.jl:47),
// for Hist = HisToContainer{Int16,128,49152,16,UInt16,10}
// (TrackingRecHit2DHeterogeneous.jl:28):
//   * layer of hit i (0-based): searchsortedfirst over the offsets
//     (count_from_vector, :217-229).  With n_lt offsets below i and `eq` if
//     one equals i, the layer is n_lt when eq, else n_lt - 1; so the first hit
//     after an empty layer is filed into that empty layer, as in the reference;
//   * bin: ((iphi >> 9) & 127) + 128 * layer (bin(), :81-86, 0-based here);
//   * off: after finalize! and the decrementing fill! (:186-205) the 1281
//     offsets are the exclusive bin starts, off[1280] = number of hits;
//   * bins: 1-based hit ids; inside a bin in descending id order, since the
//     serial fill visits the hits in ascending order and fills each bin from
//     its end.
//
// One block of 1024 threads (16 waves) per event; wave w owns a contiguous
// range of hits.  Pass 1 counts hits per (wave, bin) in LDS.  A per-bin pass
// turns those counts into the number of same-bin hits in the higher waves
// (larger ids) and the bin totals into offsets (block scan).  Pass 2 walks
// each wave's range from the top, 64 hits at a time; a hit's rank among the
// same-bin hits of larger id in its step comes from a lane match on the 11
// bin bits (ballots).  Every position is computed, none taken by an atomic,
// so the result is the serial one.
#include <hip/hip_runtime.h>

#include "pt_device.hpp"

namespace pt {

namespace {
constexpr uint32_t kHistWaves = 16;
constexpr uint32_t kHistBins = PT_N_LAYERS * PT_N_PHI_BINS;  // 1280 (the catch-all bin stays empty)

__device__ __forceinline__ uint32_t hist_bin_of(const EvDesc &e, uint32_t i) {
    uint32_t n_lt = 0;
    bool eq = false;
#pragma unroll
    for (int k = 0; k < 11; ++k) {
        n_lt += (e.layer_start[k] < i) ? 1u : 0u;
        eq |= e.layer_start[k] == i;
    }
    const uint32_t layer = eq ? n_lt : n_lt - 1u;
    return ((uint32_t)(uint16_t)e.iphi[i] >> 9) + PT_N_PHI_BINS * layer;
}

__device__ __forceinline__ uint64_t lanes_below_me() { return (1ull << (threadIdx.x & 63u)) - 1ull; }

// lanes of the wave (among `valid`) whose bin equals this lane's
__device__ __forceinline__ uint64_t bin_match(uint32_t b, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 11; ++bit) {
        const bool set = (b >> bit) & 1u;
        const uint64_t bal = __ballot(valid && set);
        m &= set ? bal : ~bal;
    }
    return m;
}
}  // namespace

__global__ void __launch_bounds__(1024) k_phi_hist(const EvDesc *__restrict__ evs, uint32_t n_events,
                                                    uint32_t xcd) {
    __shared__ uint16_t cnt[kHistWaves][kHistBins];
    __shared__ uint32_t off[kHistBins];
    __shared__ uint32_t scan_lds[1024 / 64 + 1];
    // block xcd + e takes event e, so the event's histogram is built on the
    // XCD its stage kernels run on
    if (blockIdx.x < xcd) return;
    const uint32_t ev = blockIdx.x - xcd;
    if (ev >= n_events) return;
    const EvDesc &e = evs[ev];
    if (!e.hist_bins_build) return;
    const uint32_t n = e.layer_start[10];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t per_wave = ((n + kHistWaves - 1) / kHistWaves + 63u) & ~63u;
    const uint32_t a = min(n, w * per_wave), b = min(n, a + per_wave);
    for (uint32_t k = tid; k < kHistWaves * kHistBins / 2; k += blockDim.x) reinterpret_cast<uint32_t *>(cnt)[k] = 0u;
    __syncthreads();
    // pass 1: counts per (wave, bin); one writer per bin and step (the lowest matching lane)
    for (uint32_t base = a; base < b; base += 64u) {
        const uint32_t i = base + lane;
        const bool valid = i < b;
        const uint32_t bn = valid ? hist_bin_of(e, i) : 0u;
        const uint64_t m = bin_match(bn, valid);
        if (valid && (m & lanes_below_me()) == 0) cnt[w][bn] = (uint16_t)(cnt[w][bn] + __popcll(m));
    }
    __syncthreads();
    // per bin: hits of the higher waves in front of each wave's; bin totals
    uint32_t tsum = 0;
    for (uint32_t bn = tid; bn < kHistBins; bn += blockDim.x) {
        uint32_t run = 0;
        for (int ww = (int)kHistWaves - 1; ww >= 0; --ww) {
            const uint32_t c = cnt[ww][bn];
            cnt[ww][bn] = (uint16_t)run;
            run += c;
        }
        off[bn] = run;
    }
    __syncthreads();
    // exclusive offsets: thread t scans bins [2t, 2t + 2)
    {
        const uint32_t b0 = 2u * tid;
        if (b0 < kHistBins) tsum = off[b0] + off[b0 + 1];
        uint32_t total;
        const uint32_t ex = block_excl_scan(tsum, scan_lds, total);
        if (b0 < kHistBins) {
            const uint32_t c0 = off[b0];
            off[b0] = ex;
            off[b0 + 1] = ex + c0;
            e.hist_off_build[b0] = ex;
            e.hist_off_build[b0 + 1] = ex + c0;
        }
        if (tid == 0) e.hist_off_build[kHistBins] = total;
    }
    __syncthreads();
    // pass 2: positions, each wave from the top of its range down
    const uint64_t above = (lane == 63u) ? 0ull : (~0ull << (lane + 1u));
    for (uint32_t top = b; top > a; top -= min(64u, top - a)) {
        const uint32_t step = min(64u, top - a);
        const uint32_t i = top - step + lane;
        const bool valid = lane < step;
        const uint32_t bn = valid ? hist_bin_of(e, i) : 0u;
        const uint64_t m = bin_match(bn, valid);
        if (valid) {
            const uint32_t c = cnt[w][bn];
            e.hist_bins_build[off[bn] + c + (uint32_t)__popcll(m & above)] = (uint16_t)(i + 1u);
            if ((m & lanes_below_me()) == 0) cnt[w][bn] = (uint16_t)(c + __popcll(m));
        }
    }
}

void launch_phi_hist(const EvDesc *d_evs, uint32_t n_events, uint32_t xcd, hipStream_t st) {
    if (n_events) k_phi_hist<<<n_events + xcd, 1024, 0, st>>>(d_evs, n_events, xcd);
}

}  // namespace pt
