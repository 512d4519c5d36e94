// pt_classify.hip - classify_tuples of make_tuples on gfx950
// (ca_hit_ntuplet_generator_kernels.jl:222-231):
//   kernel_classify_tracks          (ca_hit_ntuplet_generator_kernels_impl.jl:213-247)
//                                   runs in the fit epilogue (pt_fit.hip): only fitted
//                                   tuples can pass its cuts
//   kernel_fast_duplicate_remover   (:250-279)                 phases 1-3 of k_classify_event
//   kernel_count/fill_hit_in_tracks (:281-311) + finalize!    phase 4
//   kernel_triplet_cleaner          (:313-362)                 phase 5
//
// Order in the reference loops and how it is kept here:
//   * the fast duplicate remover walks the cells in id order and reads the
//     qualities the earlier cells wrote, so it is serial: a shared cell (one
//     carrying >= 2 stored tuples) is resolved in a round once every earlier
//     shared cell on one of its tuples (its predecessors) was resolved in an
//     earlier round, which gives the serial result exactly (two cells resolved
//     in one round share no tuple).  The smallest unresolved cell is always
//     ready, so the rounds end; on PU50 events there are about four;
//   * the triplet cleaner only ever writes dup, and its tests (!= bad) do not
//     change while it runs (every tuple on a hit's list was loose), so its
//     decision for a tuple depends only on, per hit of the tuple: the largest
//     hit count of the loose tuples on the hit and the one with the smallest
//     |tip| (ties: the first of the hit's list, which holds the tuples in
//     descending id order by fill_direct, i.e. the largest id).  Both come
//     from one 64-bit atomic minimum per hit; the decision is per tuple.  (The
//     reference skips hits with a single loose tuple; for those the largest
//     count is the tuple's own and it is the smallest |tip|, so the test
//     below passes them too.)
#include <hip/hip_runtime.h>

#include "pt_device.hpp"

namespace pt {

namespace {

// the reference's loops over tuples stop at the first empty one; with the
// container full, the last stored tuple reads as empty after finalize
__device__ __forceinline__ uint32_t n_classified(const EvDesc &e) {
    const uint32_t m = e.cnt->n_tuples;
    return m > kMaxTuples ? kMaxTuples - 1u : m;
}
__device__ __forceinline__ uint32_t n_stored(const EvDesc &e) {
    const uint32_t m = e.cnt->n_tuples;
    return m > kMaxTuples ? kMaxTuples : m;
}
__device__ __forceinline__ uint32_t n_shared(const EvDesc &e) { return min(e.s.misc[kMiscDc], kDcCap); }
__device__ __forceinline__ float abs_tip(const EvDesc &e, uint32_t t) { return fabsf(e.state[1 * kMaxTuples + t]); }

// dc_pred marker: more than four predecessors, check them from the lists
constexpr uint32_t kPredMany = 0xFFFFFFFEu;
// HitToTuple = OneToManyAssoc{UInt16, 49152, 4 * 49152}: its entry capacity
constexpr uint32_t kH2tCap = 4u * 49152u;
// round stamps of the first kDoneLds shared cells live in LDS
constexpr uint32_t kDoneLds = 8192u;

}  // namespace

// classify_tuples after the fit (classification itself is in the fit
// epilogue) runs in five phases; every phase reads what the previous one wrote
// for the same event.  Two launch shapes (tuned per context): separate
// kernels, where phases 1, 2, 4 and 5, independent per tuple or per shared
// cell, take several blocks per event (item i of the event goes to the
// block-strided loop of that event's blocks), phase 3's first kParRounds rounds
// are a launch each (the launch boundary is the round's barrier) and its
// remaining rounds, which need block barriers, run in one block of 1024
// threads per event; or one block of 1024 threads per event for all five
// phases with barriers between them, which saves launches where events are
// small and many.

// The phases are chains of dependent loads (tuple -> cells -> cell fields ->
// lists); each is written so that the loads of one step are issued together
// (all cells of a tuple, all hits of a tuple, up to four list entries at a
// time), which keeps the number of memory round trips per item small.

// the cells of stored tuple t that carry >= 2 tuples and have a shared-cell
// slot (cells past the first empty one are absent)
__device__ __forceinline__ void shared_cells_of(const EvDesc &e, uint32_t t, uint32_t (&cs)[4], uint32_t (&sl)[4],
                                                bool (&v)[4]) {
    const uint4 tc = reinterpret_cast<const uint4 *>(e.s.tuple_cells)[t];
    cs[0] = tc.x, cs[1] = tc.y, cs[2] = tc.z, cs[3] = tc.w;
    bool run = true;
    uint32_t nt[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        run = run && cs[q] != ~0u;
        v[q] = run;
        const uint32_t c = run ? cs[q] : 0u;
        nt[q] = e.s.trk[c].x;
        sl[q] = e.s.dc_slot[c];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = v[q] && nt[q] >= 2u && sl[q] != ~0u;
}

// Phase 1: every stored tuple appends its id to the lists of its shared cells
// (list offsets come from the multiplicity scan, cursors start at zero).
__device__ __forceinline__ void cl_lists(const EvDesc &e, uint32_t i0, uint32_t step) {
    for (uint32_t t = i0; t < n_stored(e); t += step) {
        uint32_t cs[4], sl[4], pos[4], off[4];
        bool v[4];
        shared_cells_of(e, t, cs, sl, v);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (v[q]) {
                pos[q] = atomicAdd(&e.s.dc_cur[sl[q]], 1u);
                off[q] = e.s.dc_off[sl[q]];
            }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (v[q]) e.s.dc_ent[off[q] + pos[q]] = t;
    }
}

// Phase 2, per shared cell: a list longer than the CellTracks capacity keeps
// its 64 smallest tuple ids in ascending order in front (the reference's list
// is the first 64 pushed, in tuple id order); then the predecessors: the
// shared cells with a smaller id on the tuples of the (capped) list.
__device__ __forceinline__ void cl_prepare(const EvDesc &e, uint32_t i0, uint32_t step) {
    for (uint32_t s = i0; s < n_shared(e); s += step) {
        uint32_t *L = e.s.dc_ent + e.s.dc_off[s];
        const uint32_t n_all = e.s.dc_cur[s];
        const uint32_t c = e.s.dc_cells[s];
        if (n_all > PT_MAX_CELL_TRACKS) {
            constexpr uint32_t K = PT_MAX_CELL_TRACKS;
            for (uint32_t a = 1; a < n_all; ++a) {
                const uint32_t v = L[a];
                uint32_t b = a;
                if (a >= K) {
                    if (v > L[K - 1u]) continue;
                    L[a] = L[K - 1u];  // the displaced id goes to the unsorted tail
                    b = K - 1u;
                }
                for (; b > 0 && L[b - 1] > v; --b) L[b] = L[b - 1];
                L[b] = v;
            }
        }
        const uint32_t n = min(n_all, PT_MAX_CELL_TRACKS);
        uint32_t pr[4] = {~0u, ~0u, ~0u, ~0u};
        uint32_t np = 0;
        bool many = false;
        // once `many` is set the outcome is fixed, so the tuples are taken two at
        // a time: their cells' counts and slots are loaded together
        for (uint32_t k0 = 0; k0 < n && !many; k0 += 2u) {
            uint32_t cs[8], nt[8], sq[8];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const bool in = k0 + (uint32_t)j < n;
                const uint4 tc = in ? reinterpret_cast<const uint4 *>(e.s.tuple_cells)[L[k0 + (uint32_t)j]]
                                    : make_uint4(~0u, ~0u, ~0u, ~0u);
                cs[4 * j] = tc.x, cs[4 * j + 1] = tc.y, cs[4 * j + 2] = tc.z, cs[4 * j + 3] = tc.w;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint32_t cq = (cs[q] == ~0u || cs[q] >= c) ? 0u : cs[q];
                nt[q] = e.s.trk[cq].x;
                sq[q] = e.s.dc_slot[cq];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (cs[q] == ~0u || cs[q] >= c || nt[q] < 2u) continue;
                const uint32_t x = sq[q];
                if (x == ~0u || x == pr[0] || x == pr[1] || x == pr[2] || x == pr[3]) continue;
                if (np < 4u) {
                    pr[np] = x;
                    ++np;
                } else {
                    many = true;
                }
            }
        }
        e.s.dc_pred[s] = many ? make_uint4(kPredMany, ~0u, ~0u, ~0u) : make_uint4(pr[0], pr[1], pr[2], pr[3]);
    }
}

// The loop body of kernel_fast_duplicate_remover for one shared cell, with
// list L of n tuples whose first four ids and |tip| are given (tips do not
// change while it runs; qualities do, so they are read now): the loose tuple
// with the smallest |tip| stays (ties: the smallest id, the first in the
// list), every other tuple that is not bad becomes dup.
__device__ __forceinline__ void cl_resolve(const EvDesc &e, const uint32_t *L, uint32_t n, const uint32_t (&t4)[4],
                                           const float (&tip4)[4]) {
    int8_t q4[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) q4[j] = e.quality[t4[j] == ~0u ? 0u : t4[j]];
    float mc = 10000.0f;
    uint32_t im = ~0u;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (t4[j] != ~0u && q4[j] == PT_QUALITY_LOOSE && (tip4[j] < mc || (tip4[j] == mc && im != ~0u && t4[j] < im))) {
            mc = tip4[j];
            im = t4[j];
        }
    for (uint32_t k0 = 4; k0 < n; k0 += 4u) {  // lists longer than four (rare)
        uint32_t t[4];
        float sc[4];
        int8_t ql[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) t[j] = (k0 + (uint32_t)j < n) ? L[k0 + (uint32_t)j] : ~0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t tt = t[j] == ~0u ? 0u : t[j];
            sc[j] = abs_tip(e, tt);
            ql[j] = e.quality[tt];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (t[j] != ~0u && ql[j] == PT_QUALITY_LOOSE && (sc[j] < mc || (sc[j] == mc && im != ~0u && t[j] < im))) {
                mc = sc[j];
                im = t[j];
            }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (t4[j] != ~0u && t4[j] != im && q4[j] != PT_QUALITY_BAD) e.quality[t4[j]] = PT_QUALITY_DUP;
    for (uint32_t k0 = 4; k0 < n; k0 += 4u) {
        uint32_t t[4];
        int8_t ql[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) t[j] = (k0 + (uint32_t)j < n) ? L[k0 + (uint32_t)j] : ~0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) ql[j] = e.quality[t[j] == ~0u ? 0u : t[j]];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (t[j] != ~0u && t[j] != im && ql[j] != PT_QUALITY_BAD) e.quality[t[j]] = PT_QUALITY_DUP;
    }
}

// the first four tuple ids of a list and their |tip|
__device__ __forceinline__ void cl_head(const EvDesc &e, const uint32_t *L, uint32_t n, uint32_t (&t4)[4],
                                        float (&tip4)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) t4[j] = ((uint32_t)j < n) ? L[j] : ~0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) tip4[j] = abs_tip(e, t4[j] == ~0u ? 0u : t4[j]);
}

// Phase 3: kernel_fast_duplicate_remover in rounds; the round that resolved a
// shared cell is its stamp (0: not yet), in LDS for the first kDoneLds cells.
// A thread's first kRoundCache cells keep their list head, tips and
// predecessors in registers across the rounds.
// The count of cells still pending after round r is s_pending[r % 3]: thread
// 0 clears the counter of round r + 1 during round r, when every wave has
// passed the barrier that follows its read of that counter in round r - 2, so
// no wave can read a counter that was cleared under it.
// The rounds before first_round ran as kernels of their own (cl_round_par):
// their stamps are read from global memory here.
constexpr uint32_t kRoundCache = 2;
__device__ __forceinline__ void cl_rounds(const EvDesc &e, uint32_t *s_done, uint32_t *s_pending,
                                          uint32_t first_round) {
    const uint32_t nd = n_shared(e);
    const uint32_t tid = threadIdx.x, nthr = blockDim.x;
    for (uint32_t s = tid; s < min(nd, kDoneLds); s += nthr) s_done[s] = e.s.dc_done[s];
    if (tid == 0) s_pending[first_round % 3u] = 0u;
    // the stamps start cleared by the multiplicity kernel
    auto done_of = [&](uint32_t s) -> uint32_t { return s < kDoneLds ? s_done[s] : e.s.dc_done[s]; };
    // predecessors ready: each resolved in an earlier round
    auto ready_of = [&](uint32_t s, const uint4 &pred, const uint32_t *L, uint32_t n, uint32_t round) {
        bool ready = true;
        if (pred.x != kPredMany) {
            const uint32_t ps[4] = {pred.x, pred.y, pred.z, pred.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (ps[q] == ~0u) continue;
                const uint32_t d = done_of(ps[q]);
                if (d == 0u || d == round) ready = false;
            }
        } else {
            const uint32_t c = e.s.dc_cells[s];
            for (uint32_t k = 0; k < n && ready; ++k) {
                uint32_t cs[4], sl[4];
                bool v[4];
                shared_cells_of(e, L[k], cs, sl, v);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (!v[q] || cs[q] >= c) continue;
                    const uint32_t d = done_of(sl[q]);
                    if (d == 0u || d == round) ready = false;
                }
            }
        }
        return ready;
    };
    uint32_t c_n[kRoundCache], c_t[kRoundCache][4];
    const uint32_t *c_L[kRoundCache];
    uint4 c_pred[kRoundCache];
    float c_tip[kRoundCache][4];
#pragma unroll
    for (uint32_t k = 0; k < kRoundCache; ++k) {
        const uint32_t s = tid + k * nthr;
        const bool in = s < nd;
        c_L[k] = e.s.dc_ent + (in ? e.s.dc_off[s] : 0u);
        c_n[k] = in ? min(e.s.dc_cur[s], PT_MAX_CELL_TRACKS) : 0u;
        c_pred[k] = in ? e.s.dc_pred[s] : make_uint4(~0u, ~0u, ~0u, ~0u);
    }
#pragma unroll
    for (uint32_t k = 0; k < kRoundCache; ++k) cl_head(e, c_L[k], c_n[k], c_t[k], c_tip[k]);
    __syncthreads();
    for (uint32_t round = first_round;; ++round) {
        if (tid == 0) s_pending[(round + 1u) % 3u] = 0u;
        uint32_t pending = 0;
#pragma unroll
        for (uint32_t k = 0; k < kRoundCache; ++k) {
            const uint32_t s = tid + k * nthr;
            if (s >= nd || done_of(s)) continue;
            if (!ready_of(s, c_pred[k], c_L[k], c_n[k], round)) {
                ++pending;
                continue;
            }
            cl_resolve(e, c_L[k], c_n[k], c_t[k], c_tip[k]);
            if (s < kDoneLds)
                s_done[s] = round;
            else
                e.s.dc_done[s] = round;
        }
        for (uint32_t s = tid + kRoundCache * nthr; s < nd; s += nthr) {
            if (done_of(s)) continue;
            const uint32_t *L = e.s.dc_ent + e.s.dc_off[s];
            const uint32_t n = min(e.s.dc_cur[s], PT_MAX_CELL_TRACKS);
            if (!ready_of(s, e.s.dc_pred[s], L, n, round)) {
                ++pending;
                continue;
            }
            uint32_t t4[4];
            float tip4[4];
            cl_head(e, L, n, t4, tip4);
            cl_resolve(e, L, n, t4, tip4);
            if (s < kDoneLds)
                s_done[s] = round;
            else
                e.s.dc_done[s] = round;
        }
        if (pending) atomicAdd(&s_pending[round % 3u], pending);
        __syncthreads();
        if (s_pending[round % 3u] == 0u) break;
        if (round - first_round + 1u >= nd) {  // every round resolves at least one cell: not reached
            if (tid == 0) atomicOr(&e.cnt->internal_error, PT_INT_ROUNDS);
            break;
        }
    }
}

// One round of phase 3 in a kernel of its own, several blocks per event: the
// shared cells whose predecessors were all stamped by earlier launches are
// resolved and stamped `round`.  A predecessor stamped `round` in this launch
// is not ready (the rule of the rounds above), so the cells resolved together
// share no tuple and the result is the serial one.
constexpr uint32_t kParRounds = 3;
__device__ __forceinline__ void cl_round_par(const EvDesc &e, uint32_t round, uint32_t i0, uint32_t step) {
    auto done_before = [&](uint32_t p) {
        const uint32_t d = e.s.dc_done[p];
        return d != 0u && d != round;
    };
    for (uint32_t s = i0; s < n_shared(e); s += step) {
        if (e.s.dc_done[s]) continue;
        const uint4 pred = e.s.dc_pred[s];
        const uint32_t *L = e.s.dc_ent + e.s.dc_off[s];
        const uint32_t n = min(e.s.dc_cur[s], PT_MAX_CELL_TRACKS);
        bool ready = true;
        if (pred.x != kPredMany) {
            const uint32_t ps[4] = {pred.x, pred.y, pred.z, pred.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (ps[q] != ~0u && !done_before(ps[q])) ready = false;
        } else {
            const uint32_t c = e.s.dc_cells[s];
            for (uint32_t k = 0; k < n && ready; ++k) {
                uint32_t cs[4], sl[4];
                bool v[4];
                shared_cells_of(e, L[k], cs, sl, v);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (v[q] && cs[q] < c && !done_before(sl[q])) ready = false;
            }
        }
        if (!ready) continue;
        uint32_t t4[4];
        float tip4[4];
        cl_head(e, L, n, t4, tip4);
        cl_resolve(e, L, n, t4, tip4);
        e.s.dc_done[s] = round;
    }
}

// hits of a tuple looked at together (longer tuples take further rounds)
constexpr uint32_t kHitsAtOnce = 6u;

// Phase 4: per loose tuple and hit, the hit's loose-tuple count, their largest
// hit count and the (|tip|, largest id) minimum; the association's entries
// are counted against its capacity.  The per-hit fields start from k_reset.
__device__ __forceinline__ void cl_h2t_reduce(const EvDesc &e, uint32_t i0, uint32_t step) {
    for (uint32_t t = i0; t < n_classified(e); t += step) {
        const bool loose = e.quality[t] == PT_QUALITY_LOOSE;
        const uint32_t b0 = e.hit_off[t] - 1u, b1 = e.hit_off[t + 1] - 1u;
        const float tip = abs_tip(e, t);
        const uint32_t nh = loose ? b1 - b0 : 0u;
        if (loose) {
            // smallest key: the largest hit count, then the smallest |tip|, then the
            // largest id (t < 32768 = kMaxTuples, |tip| bits < 2^31)
            const uint64_t key = ((uint64_t)(15u - nh) << 60) | ((uint64_t)__float_as_uint(tip) << 15) |
                                 (uint64_t)(0x7FFFu - t);
            for (uint32_t p0 = b0; p0 < b1; p0 += kHitsAtOnce) {
                uint32_t h[kHitsAtOnce];
#pragma unroll
                for (uint32_t j = 0; j < kHitsAtOnce; ++j) h[j] = (p0 + j < b1) ? e.hit_bins[p0 + j] - 1u : ~0u;
#pragma unroll
                for (uint32_t j = 0; j < kHitsAtOnce; ++j)
                    if (h[j] != ~0u)
                        atomicMin(reinterpret_cast<unsigned long long *>(&e.s.h2t_win[h[j]]),
                                  (unsigned long long)key);
            }
        }
        // entries of the wave: tuple sizes are below 8, sum them bit by bit
        uint32_t sum = 0;
#pragma unroll
        for (int b = 0; b < 3; ++b) sum += (uint32_t)__popcll(__ballot((nh >> b) & 1u)) << b;
        if ((threadIdx.x & 63u) == 0 && sum) atomicAdd(&e.s.misc[kMiscH2tEnt], sum);
    }
}

// Phase 5: kernel_triplet_cleaner per loose tuple: on any hit shared with
// another loose tuple it becomes dup when shorter than the longest there, or,
// when they are all triplets, when it is not the one with the smallest |tip|.
// The survivors are the loose count.
__device__ __forceinline__ void cl_h2t_decide(const EvDesc &e, uint32_t i0, uint32_t step, bool first) {
    for (uint32_t t = i0; t < n_classified(e); t += step) {
        bool keep = false;
        const bool loose = e.quality[t] == PT_QUALITY_LOOSE;
        const uint32_t b0 = e.hit_off[t] - 1u, b1 = e.hit_off[t + 1] - 1u;
        if (loose) {
            const uint32_t nh = b1 - b0;
            bool dup = false;
            for (uint32_t p0 = b0; p0 < b1; p0 += kHitsAtOnce) {
                uint32_t h[kHitsAtOnce];
                uint64_t kw[kHitsAtOnce];
#pragma unroll
                for (uint32_t j = 0; j < kHitsAtOnce; ++j) h[j] = (p0 + j < b1) ? e.hit_bins[p0 + j] - 1u : ~0u;
#pragma unroll
                for (uint32_t j = 0; j < kHitsAtOnce; ++j) {
                    const uint32_t hh = h[j] == ~0u ? 0u : h[j];
                    kw[j] = e.s.h2t_win[hh];
                }
#pragma unroll
                for (uint32_t j = 0; j < kHitsAtOnce; ++j) {
                    const uint32_t mx = 15u - (uint32_t)(kw[j] >> 60);
                    const uint32_t wt = 0x7FFFu - (uint32_t)(kw[j] & 0x7FFFu);
                    if (h[j] != ~0u && (nh != mx || (mx <= 3u && wt != t))) dup = true;
                }
            }
            if (dup)
                e.quality[t] = PT_QUALITY_DUP;
            else
                keep = true;
        }
        const uint64_t m = __ballot(keep);
        if ((threadIdx.x & 63u) == 0 && m) atomicAdd(&e.cnt->n_loose, (uint32_t)__popcll(m));
    }
    // the entry count was summed by atomics (in this kernel, in the one-block
    // shape): read it from L2, where an earlier plain load of misc may have left
    // a stale line in this CU's cache
    if (first && threadIdx.x == 0 && atomicAdd(&e.s.misc[kMiscH2tEnt], 0u) > kH2tCap)
        atomicOr(&e.cnt->overflow, PT_OVF_HIT_TO_TUPLE);
}

// the block-strided loop of a multi-block phase: first item and stride
#define CL_PHASE(e_, i0_, step_)                                                                    \
    uint32_t ev_, blk_;                                                                             \
    if (!ev_block(grid, ev_, blk_)) return;                                                         \
    const EvDesc &e_ = evs[ev_];                                                                    \
    const uint32_t i0_ = blk_ * blockDim.x + threadIdx.x, step_ = grid.per_event * blockDim.x

__global__ void __launch_bounds__(256) k_cl_lists(const EvDesc *__restrict__ evs, EvGrid grid) {
    CL_PHASE(e, i0, step);
    cl_lists(e, i0, step);
}

__global__ void __launch_bounds__(256) k_cl_prepare(const EvDesc *__restrict__ evs, EvGrid grid) {
    CL_PHASE(e, i0, step);
    cl_prepare(e, i0, step);
}

__global__ void __launch_bounds__(256) k_cl_round_par(const EvDesc *__restrict__ evs, EvGrid grid, uint32_t round) {
    CL_PHASE(e, i0, step);
    cl_round_par(e, round, i0, step);
}

// the rounds after the parallel ones (few cells are left for them)
__global__ void __launch_bounds__(1024) k_cl_rounds(const EvDesc *__restrict__ evs, EvGrid grid) {
    __shared__ uint32_t s_done[kDoneLds];
    __shared__ uint32_t s_pending[3];
    uint32_t ev, blk;
    if (!ev_block(grid, ev, blk)) return;
    cl_rounds(evs[ev], s_done, s_pending, kParRounds + 1u);
}

__global__ void __launch_bounds__(256) k_cl_h2t_reduce(const EvDesc *__restrict__ evs, EvGrid grid) {
    CL_PHASE(e, i0, step);
    cl_h2t_reduce(e, i0, step);
}

__global__ void __launch_bounds__(256) k_cl_h2t_decide(const EvDesc *__restrict__ evs, EvGrid grid) {
    CL_PHASE(e, i0, step);
    cl_h2t_decide(e, i0, step, blk_ == 0);
}

__global__ void __launch_bounds__(1024) k_classify_event(const EvDesc *__restrict__ evs, EvGrid grid) {
    __shared__ uint32_t s_done[kDoneLds];
    __shared__ uint32_t s_pending[3];
    uint32_t ev, blk;
    if (!ev_block(grid, ev, blk)) return;
    const EvDesc &e = evs[ev];
    cl_lists(e, threadIdx.x, blockDim.x);
    __syncthreads();
    cl_prepare(e, threadIdx.x, blockDim.x);
    __syncthreads();
    cl_rounds(e, s_done, s_pending, 1u);
    __syncthreads();
    cl_h2t_reduce(e, threadIdx.x, blockDim.x);
    __syncthreads();
    cl_h2t_decide(e, threadIdx.x, blockDim.x, true);
}

void launch_classify(const EvDesc *d_evs, uint32_t n_events, const Config &cfg, hipStream_t st, Marks *tm) {
    const EvGrid one{n_events, 1u, cfg.xcd};
    if (!cfg.classify_split) {
        k_classify_event<<<ev_grid_dim(one), 1024, 0, st>>>(d_evs, one);
        if (tm) tm->record("k_classify_event", st);
        return;
    }
    const EvGrid tup = ev_grid(kMaxTuples, n_events, cfg.xcd);
    const EvGrid shr = ev_grid(kDcCap, n_events, cfg.xcd);
    auto mark = [&](const char *name) {
        if (tm) tm->record(name, st);
    };
    k_cl_lists<<<ev_grid_dim(tup), 256, 0, st>>>(d_evs, tup);
    mark("k_cl_lists");
    k_cl_prepare<<<ev_grid_dim(shr), 256, 0, st>>>(d_evs, shr);
    mark("k_cl_prepare");
    for (uint32_t r = 1; r <= kParRounds; ++r) k_cl_round_par<<<ev_grid_dim(shr), 256, 0, st>>>(d_evs, shr, r);
    mark("k_cl_round_par");
    k_cl_rounds<<<ev_grid_dim(one), 1024, 0, st>>>(d_evs, one);
    mark("k_cl_rounds");
    k_cl_h2t_reduce<<<ev_grid_dim(tup), 256, 0, st>>>(d_evs, tup);
    mark("k_cl_h2t_reduce");
    k_cl_h2t_decide<<<ev_grid_dim(tup), 256, 0, st>>>(d_evs, tup);
    mark("k_cl_h2t_decide");
}

}  // namespace pt
