// pt_load_tracks.hip - the vertex finder's track selection on gfx950:
// load_tracks (gpu_vertex_finder.jl:51-82), the first step of
// gpuVertexFinder.make (:208-211), which reads the TrackSOA this library
// writes.  For TrackSOA rows in order it stops at the first row with no hits
// (n_hits_track = size of the det_indices bin), skips rows with fewer than 4
// hits, rows whose quality is not loose and rows with pt < pt_min (a NaN pt is
// kept, as the comparison is false), and appends the row (1-based), zip
// (state column 5), covariance column 15 and pt^2 to the WorkSpace arrays.
//
// One block of 1024 threads per event; thread t owns rows [32t, 32t + 32).
// The first empty row is an LDS minimum; the kept rows are counted per thread,
// scanned over the block and written in row order, so the order is the serial
// one.
#include <hip/hip_runtime.h>

#include "pt_device.hpp"

namespace pt {

__global__ void __launch_bounds__(1024) k_load_tracks(const EvDesc *__restrict__ evs, uint32_t n_events, float pt_min) {
    constexpr uint32_t kRows = kMaxTuples / 1024u;
    __shared__ uint32_t s_first;
    __shared__ uint32_t scan_lds[1024 / 64 + 1];
    const uint32_t ev = blockIdx.x;
    if (ev >= n_events) return;
    const EvDesc &e = evs[ev];
    const uint32_t tid = threadIdx.x, r0 = tid * kRows;
    if (tid == 0) s_first = kMaxTuples;
    __syncthreads();
    uint32_t first = kMaxTuples;
    for (uint32_t k = 0; k < kRows; ++k)
        if (e.det_off[r0 + k + 1] == e.det_off[r0 + k]) {
            first = r0 + k;
            break;
        }
    if (first < kMaxTuples) atomicMin(&s_first, first);
    __syncthreads();
    const uint32_t r1 = min(r0 + kRows, s_first);
    auto keep = [&](uint32_t r) {
        return e.det_off[r + 1] - e.det_off[r] >= 4u && e.quality[r] == PT_QUALITY_LOOSE && !(e.pt[r] < pt_min);
    };
    uint32_t n = 0;
    for (uint32_t r = r0; r < r1; ++r) n += keep(r) ? 1u : 0u;
    uint32_t total;
    uint32_t pos = block_excl_scan(n, scan_lds, total);
    for (uint32_t r = r0; r < r1 && n; ++r) {
        if (!keep(r)) continue;
        const float pt = e.pt[r];
        e.ws_index[pos] = (uint16_t)(r + 1u);
        e.ws_zt[pos] = e.state[4u * kMaxTuples + r];
        e.ws_ezt2[pos] = e.cov[14u * kMaxTuples + r];
        e.ws_ptt2[pos] = pt * pt;
        ++pos;
    }
    if (tid == 0) *e.ws_n = total;
}

void launch_load_tracks(const EvDesc *d_evs, uint32_t n_events, float pt_min, hipStream_t st) {
    if (n_events) k_load_tracks<<<n_events, 1024, 0, st>>>(d_evs, n_events, pt_min);
}

}  // namespace pt
