// pt_classify.hip - classify_tuples of make_tuples on gfx950
// (ca_hit_ntuplet_generator_kernels.jl:222-231):
//   kernel_classify_tracks          (ca_hit_ntuplet_generator_kernels_impl.jl:213-247)
//   kernel_fast_duplicate_remover   (:250-279)
//   kernel_count/fill_hit_in_tracks (:281-311) + finalize!
//   kernel_triplet_cleaner          (:313-362)
//
// Order in the reference loops and how it is kept here:
//   * classification is per tuple, independent;
//   * the fast duplicate remover walks the cells in id order and reads the
//     qualities the earlier cells wrote, so it is serial: a cell is resolved
//     in a round once every earlier cell sharing one of its tuples has been
//     resolved in an earlier round, which gives the serial result exactly
//     (two cells resolved in one round share no tuple);
//   * the triplet cleaner only ever writes dup and its tests (!= bad) do not
//     change while it runs (every tuple in a hit's list was loose), so each
//     hit is independent; ties in |tip| go to the first tuple of the hit's
//     list, which holds the tuples in descending id order (fill_direct).
#include <hip/hip_runtime.h>

#include "pt_device.hpp"

namespace pt {

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
__device__ __forceinline__ uint32_t tuple_size(const EvDesc &e, uint32_t t) { return e.hit_off[t + 1] - e.hit_off[t]; }

// K-C1: kernel_classify_tracks per tuple; per used cell on >= 2 tuples a slot
// of the duplicate-remover work list (cells sharing tuples), with room for
// its tuple list.
__global__ void __launch_bounds__(256) k_classify(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    PT_EVENT_LOOP(e, i, max(n_classified(e), e.cnt->n_cells)) {
        if (i < n_classified(e)) {
            const uint32_t t = i;
            const uint32_t nh = tuple_size(e, t);
            if (e.quality[t] != PT_QUALITY_DUP && nh >= 3) {
                bool is_nan = false;
#pragma unroll
                for (int k = 0; k < 5; ++k) is_nan |= isnan(e.state[k * kMaxTuples + t]);
                if (!is_nan) {
                    const float ptv = e.pt[t];
                    const float pt = isnan(ptv) ? ptv : (cfg.p.chi2_max_pt < ptv ? cfg.p.chi2_max_pt : ptv);
                    const float *cf = cfg.p.chi2_coeff;
                    const float chi2_cut = cfg.p.chi2_scale * (cf[0] + pt * (cf[1] + pt * (cf[2] + pt * cf[3])));
                    if (!(3.0f * e.chi2[t] >= chi2_cut)) {
                        const bool quad = nh > 3;
                        const float max_tip = quad ? cfg.p.quadruplet_max_tip : cfg.p.triplet_max_tip;
                        const float min_pt = quad ? cfg.p.quadruplet_min_pt : cfg.p.triplet_min_pt;
                        const float max_zip = quad ? cfg.p.quadruplet_max_zip : cfg.p.triplet_max_zip;
                        const float tip = e.state[1 * kMaxTuples + t], zip = e.state[4 * kMaxTuples + t];
                        if (fabsf(tip) < max_tip && ptv > min_pt && fabsf(zip) < max_zip)
                            e.quality[t] = PT_QUALITY_LOOSE;
                    }
                }
            }
        }
        if (i < e.cnt->n_cells) {
            const uint32_t c = i;
            uint32_t slot = ~0u;
            if (e.s.cell_flags[c] & kUsed) {
                const uint32_t nt = e.s.trk[c].x;
                if (nt >= 2u) {
                    slot = atomicAdd(&e.s.misc[kMiscDc], 1u);
                    if (slot < kDcCap) {
                        e.s.dc_cells[slot] = c;
                        e.s.dc_off[slot] = atomicAdd(&e.s.misc[kMiscDcEnt], nt);
                        e.s.dc_cur[slot] = 0u;
                        e.s.dc_done[slot] = 0u;
                    } else {
                        atomicOr(&e.cnt->internal_error, PT_INT_OVF_LIST);
                        slot = ~0u;
                    }
                }
            }
            e.s.dc_slot[c] = slot;
        }
    }
}

// K-C2: every stored tuple appends its id to the lists of its shared cells.
__global__ void __launch_bounds__(256) k_classify_lists(const EvDesc *__restrict__ evs, EvGrid grid) {
    PT_EVENT_LOOP(e, t, n_stored(e)) {
        const uint4 tc = reinterpret_cast<const uint4 *>(e.s.tuple_cells)[t];
        const uint32_t cs[4] = {tc.x, tc.y, tc.z, tc.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (cs[q] == ~0u) break;
            const uint32_t slot = e.s.dc_slot[cs[q]];
            if (slot == ~0u) continue;
            const uint32_t pos = atomicAdd(&e.s.dc_cur[slot], 1u);
            e.s.dc_ent[e.s.dc_off[slot] + pos] = t;
        }
    }
}

// K-C3, one block per event: the fast duplicate remover, then the hit ->
// loose-tuple association and the triplet cleaner, then the loose count.
__global__ void __launch_bounds__(1024) k_classify_tail(const EvDesc *__restrict__ evs, EvGrid grid) {
    __shared__ uint32_t lds[1024 / 64 + 1];
    __shared__ uint32_t s_pending;
    uint32_t ev, blk;
    if (!ev_block(grid, ev, blk)) return;
    const EvDesc &e = evs[ev];
    const uint32_t nd = min(e.s.misc[kMiscDc], kDcCap);
    const uint32_t tid = threadIdx.x, nthr = blockDim.x;
    // each shared cell's list in ascending tuple id (the order add_track pushed them)
    for (uint32_t s = tid; s < nd; s += nthr) {
        uint32_t *L = e.s.dc_ent + e.s.dc_off[s];
        const uint32_t n = e.s.dc_cur[s];
        for (uint32_t a = 1; a < n; ++a) {
            const uint32_t v = L[a];
            uint32_t b = a;
            for (; b > 0 && L[b - 1] > v; --b) L[b] = L[b - 1];
            L[b] = v;
        }
    }
    __syncthreads();
    // rounds: a cell is ready when every earlier shared cell on one of its
    // tuples was resolved in an earlier round; dc_done = round of resolution
    for (uint32_t round = 1;; ++round) {
        if (tid == 0) s_pending = 0u;
        __syncthreads();
        uint32_t pending = 0;
        for (uint32_t s = tid; s < nd; s += nthr) {
            if (e.s.dc_done[s]) continue;
            const uint32_t c = e.s.dc_cells[s];
            const uint32_t *L = e.s.dc_ent + e.s.dc_off[s];
            const uint32_t n = min(e.s.dc_cur[s], PT_MAX_CELL_TRACKS);
            bool ready = true;
            for (uint32_t k = 0; k < n && ready; ++k) {
                const uint4 tc = reinterpret_cast<const uint4 *>(e.s.tuple_cells)[L[k]];
                const uint32_t cs[4] = {tc.x, tc.y, tc.z, tc.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t cq = cs[q];
                    if (cq == ~0u || cq >= c) continue;
                    const uint32_t sq = e.s.dc_slot[cq];
                    if (sq == ~0u) continue;
                    const uint32_t d = e.s.dc_done[sq];
                    if (d == 0u || d == round) ready = false;
                }
            }
            if (!ready) {
                ++pending;
                continue;
            }
            // kernel_fast_duplicate_remover body for cell c
            float mc = 10000.0f;
            uint32_t im = ~0u;
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t t = L[k];
                const float sc = fabsf(e.state[1 * kMaxTuples + t]);
                if (e.quality[t] == PT_QUALITY_LOOSE && sc < mc) {
                    mc = sc;
                    im = t;
                }
            }
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t t = L[k];
                if (e.quality[t] != PT_QUALITY_BAD && t != im) e.quality[t] = PT_QUALITY_DUP;
            }
            e.s.dc_done[s] = round;
        }
        if (pending) atomicAdd(&s_pending, pending);
        __syncthreads();
        if (s_pending == 0u) break;
    }
    __syncthreads();
    // kernel_count_hit_in_tracks: loose tuples per hit (counts zeroed by k_reset)
    const uint32_t ncl = n_classified(e);
    for (uint32_t t = tid; t < ncl; t += nthr) {
        if (e.quality[t] != PT_QUALITY_LOOSE) continue;
        for (uint32_t p = e.hit_off[t] - 1u; p + 1u < e.hit_off[t + 1]; ++p)
            atomicAdd(&e.s.h2t_cursor[e.hit_bins[p] - 1u], 1u);
    }
    __syncthreads();
    // finalize!: exclusive offsets over the hits, cursors back to zero
    const uint32_t nh = e.n_hits;
    const uint32_t chunk = (nh + nthr - 1) / nthr;
    const uint32_t h0 = min(nh, tid * chunk), h1 = min(nh, h0 + chunk);
    uint32_t sum = 0;
    for (uint32_t h = h0; h < h1; ++h) sum += e.s.h2t_cursor[h];
    uint32_t total;
    uint32_t run = block_excl_scan(sum, lds, total);
    for (uint32_t h = h0; h < h1; ++h) {
        const uint32_t cnt = e.s.h2t_cursor[h];
        e.s.h2t_off[h] = run;
        run += cnt;
        e.s.h2t_cursor[h] = 0u;
    }
    if (tid == 0) e.s.h2t_off[nh] = total;
    __syncthreads();
    // kernel_fill_hit_in_tracks (list order does not matter: ties are broken by id below)
    for (uint32_t t = tid; t < ncl; t += nthr) {
        if (e.quality[t] != PT_QUALITY_LOOSE) continue;
        for (uint32_t p = e.hit_off[t] - 1u; p + 1u < e.hit_off[t + 1]; ++p) {
            const uint32_t h = e.hit_bins[p] - 1u;
            const uint32_t pos = atomicAdd(&e.s.h2t_cursor[h], 1u);
            e.s.h2t_bins[e.s.h2t_off[h] + pos] = (uint16_t)t;
        }
    }
    __syncthreads();
    // kernel_triplet_cleaner per hit: the decisions read only sizes and tips
    for (uint32_t h = tid; h < nh; h += nthr) {
        const uint32_t b0 = e.s.h2t_off[h], b1 = e.s.h2t_off[h + 1];
        if (b1 - b0 < 2u) continue;
        uint32_t max_nh = 0;
        for (uint32_t p = b0; p < b1; ++p) max_nh = max(max_nh, tuple_size(e, e.s.h2t_bins[p]));
        float mc = 10000.0f;
        uint32_t im = ~0u;
        for (uint32_t p = b0; p < b1; ++p) {
            const uint32_t t = e.s.h2t_bins[p];
            if (tuple_size(e, t) != max_nh) e.quality[t] = PT_QUALITY_DUP;
            // first of the descending-id list on ties = the largest id
            const float sc = fabsf(e.state[1 * kMaxTuples + t]);
            if (sc < mc || (sc == mc && im != ~0u && t > im)) {
                mc = sc;
                im = t;
            }
        }
        if (max_nh > 3u) continue;
        for (uint32_t p = b0; p < b1; ++p) {
            const uint32_t t = e.s.h2t_bins[p];
            if (t != im) e.quality[t] = PT_QUALITY_DUP;
        }
    }
    __syncthreads();
    uint32_t loose = 0;
    for (uint32_t t = tid; t < n_stored(e); t += nthr) loose += (e.quality[t] == PT_QUALITY_LOOSE) ? 1u : 0u;
    uint32_t tot_loose;
    block_excl_scan(loose, lds, tot_loose);
    if (tid == 0) e.cnt->n_loose = tot_loose;
}

void launch_classify(const EvDesc *d_evs, uint32_t n_events, const Config &cfg, hipStream_t st, Marks *tm) {
    uint32_t per = (kMaxTuples + 255u) / 256u;
    const uint32_t target = (2048u + n_events - 1) / n_events;
    if (per > target) per = target;
    if (per < 1) per = 1;
    const EvGrid g{n_events, per};
    k_classify<<<ev_grid_dim(g), 256, 0, st>>>(d_evs, g, cfg);
    if (tm) tm->record("k_classify", st);
    k_classify_lists<<<ev_grid_dim(g), 256, 0, st>>>(d_evs, g);
    if (tm) tm->record("k_classify_lists", st);
    const EvGrid one{n_events, 1u};
    k_classify_tail<<<ev_grid_dim(one), 1024, 0, st>>>(d_evs, one);
    if (tm) tm->record("k_classify_tail", st);
}

}  // namespace pt
