// pt_ca.hip - CA stage of the pixel-track path for gfx950: doublets, outer-hit
// lists, cell connection, fishbone, ntuplet search, early duplicate removal,
// tuple multiplicity and hit->module indices.
//
// Every stage reproduces the serial creation order of the reference (so ids,
// caps and list contents are identical) without serialising the work:
//   * doublets: count per inner slot -> exclusive scan -> fill at the scanned
//     offset; cell id = serial doublet_id - 1 (gpu_pixel_doublets.jl:187-355)
//   * neighbour lists are built from the outer cell's side ("o-centric"): the
//     candidates of an outer cell are the cells starting on its outer hit,
//     which the doublet scan already lays out in ascending id order per layer
//     pair, so each list comes out ascending and capped at 64 exactly as the
//     serial push! order gives (ca_hit_ntuplet_generator_kernels_impl.jl:45-114)
//   * ntuplets: per start cell count -> scan -> emit, so tuple ids equal the
//     serial DFS emission order (gpu_ca_cell.jl:232-270, impl.jl:116-132).
#include <hip/hip_runtime.h>

#include "pt_device.hpp"

namespace pt {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ void wave_add_atomic(uint32_t *addr, uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(addr, v);
}
__device__ __forceinline__ void wave_or_atomic(uint32_t *addr, uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v |= __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicOr(addr, v);
}

// ------------------------------------------------------------------ doublet cuts
// gpu_pixel_doublets.jl:61-65 : Float64 dphi and r2t4, Float32 products
__device__ __forceinline__ bool pt_cut(float ro, int16_t idphi, float ri) {
    const double min_radius = 0.5 * 87.78;
    const double r2t4 = 4. * min_radius * min_radius;
    const double dphi = (double)idphi * ((2.0 * 3.141592653589793) / 65536.0);
    const float rr = ri * ro;
    const float d = ro - ri;
    const float d2 = d * d;
    return dphi * dphi * (r2t4 - (double)rr) > (double)d2;
}

// Walk the phi window of inner slot j (gpu_pixel_doublets.jl:187-355) and call
// on_pass(outer_hit_1based) for every candidate passing all cuts, in the
// serial order: bins kl..kh-1 with wrap, stored (descending-id) order inside a bin.
template <class F>
__device__ __forceinline__ void walk_slot(const EvDesc &e, const Config &cfg, uint32_t j, uint32_t &ref_bits,
                                          F &&on_pass) {
    const pt_params &P = cfg.p;
    int p = 0;
    while (j >= e.slot_cum[p]) ++p;
    const int inner = cfg.pairs.inner[p], outer = cfg.pairs.outer[p];
    const uint32_t i0 = ((p == 0) ? j : j - e.slot_cum[p - 1]) + e.layer_start[inner];
    const int16_t mi = e.det[i0];
    if (mi > 2000) return;
    const float me_z = e.zg[i0];
    if (me_z < cfg.pairs.min_z[p] || me_z > cfg.pairs.max_z[p]) return;
    int me_s = -1;
    if (P.do_cluster_cut) {
        if (inner == 0 && !(mi < 96)) ref_bits |= PT_REF_BAD_MODULE;
        const bool is_outer_ladder = P.ideal_conditions ? true : (((mi / 8) % 2) == 0);
        me_s = (inner > 0 || is_outer_ladder) ? (int)e.ysize[i0] : -1;
        if (inner == 0 && outer > 3 && me_s > 0 && me_s < 36) return;
        if (inner == 1 && outer > 3 && me_s > 0 && me_s < 28) return;
    }
    const int16_t me_p = e.iphi[i0];
    const float me_r = e.rg[i0];
    const int16_t cut = cfg.pairs.phi_cut[p];
    const float max_r = cfg.pairs.max_r[p];
    const float dz_dr_fact = (float)(8 * 0.0285 / 0.015);
    const int dy = (inner == 0) ? 28 : 20;
    const uint32_t lo = e.layer_start[outer], hi = e.layer_start[outer + 1];
    int kl = hist_bin((int16_t)(me_p - cut));
    int kh = hist_bin((int16_t)(me_p + cut)) + 1;
    if (kh == 129) kh = 1;
    const int h_off = 128 * outer;
    for (int cb = kl; cb != kh; cb = (cb == 128) ? 1 : cb + 1) {
        const uint32_t pb = e.hist_off[cb + h_off - 1];
        uint32_t pe = e.hist_off[cb + h_off];
        if (pe > e.n_hits) {  // malformed histogram: never read past the hit arrays
            ref_bits |= PT_REF_HIST_LAYER;
            pe = e.n_hits;
        }
        for (uint32_t pos = pb; pos < pe; ++pos) {
            const uint32_t oi = e.hist_bins[pos];  // 1-based
            if (!(oi > lo && oi <= hi)) {
                ref_bits |= PT_REF_HIST_LAYER;
                if (oi == 0 || oi > e.n_hits) continue;
            }
            const uint32_t o0 = oi - 1;
            if (e.det[o0] > 2000) continue;
            const float ro = e.rg[o0];
            const float zo = e.zg[o0];
            if (P.do_z0_cut) {
                // gpu_pixel_doublets.jl:69-74
                const float dr = ro - me_r;
                if (dr > max_r || dr < 0.0f || fabsf(me_z * ro - me_r * zo) > 12.0f * dr) continue;
            }
            const int16_t mo_p = e.iphi[o0];
            int16_t i_dphi = (int16_t)(mo_p - me_p);
            i_dphi = (int16_t)(i_dphi < 0 ? -i_dphi : i_dphi);  // Int16 abs wraps at -32768
            if (i_dphi > cut) continue;
            if (P.do_cluster_cut) {
                // gpu_pixel_doublets.jl:79-88
                if (outer < 4) {
                    const int so = e.ysize[o0];
                    if (me_s > 0 && so > 0 && abs(so - me_s) > dy) continue;
                } else if (inner < 4 && me_s > 0) {
                    const float q = fabsf((me_z - zo) / (me_r - ro)) * dz_dr_fact;
                    const double t = trunc((double)q + 0.5);
                    if (!isfinite(t) || fabs(t) > 9.0e15) {
                        ref_bits |= PT_REF_SIZE_CUT_INF;
                        continue;
                    }
                    if (llabs((long long)me_s - (long long)t) > 20) continue;
                }
            }
            if (P.do_pt_cut && pt_cut(ro, i_dphi, me_r)) continue;
            on_pass(oi, i0);
        }
    }
}

// ------------------------------------------------------------------ K0: reset per event
// Clears the counters, the per-hit list heads and (unless the caller's
// TrackSOA is already zero, as TrackSOA() makes it) state and covariance.
__global__ void k_reset(const EvDesc *__restrict__ evs, int zero_state) {
    constexpr uint32_t kHead = 32u;
    PT_EVENT_LOOP(e, i, kHead + e.n_hits + 1u + (zero_state ? 20u * kMaxTuples : 0u)) {
        uint32_t k = i;
        if (k < kHead) {
            if (k < sizeof(pt_counters) / 4) reinterpret_cast<uint32_t *>(e.cnt)[k] = 0u;
            if (k < 16u) e.s.misc[k] = 0u;
            if (k == 0) e.cnt->n_hits = e.n_hits;
            if (k == 1) e.cnt->n_inner_slots = e.n_slots;
            continue;
        }
        k -= kHead;
        if (k < e.n_hits + 1u) {
            e.s.ohc_cnt[k] = 0u;
            if (k < e.n_hits) e.s.ohc_cursor[k] = 0u;
            continue;
        }
        k -= e.n_hits + 1u;
        if (k < 5u * kMaxTuples)
            e.state[k] = 0.0f;
        else
            e.cov[k - 5u * kMaxTuples] = 0.0f;
    }
}

// ------------------------------------------------------------------ K1: doublet count
__global__ void __launch_bounds__(256) k_doublet_count(const EvDesc *__restrict__ evs, Config cfg) {
    uint32_t ref = 0;
    PT_EVENT_LOOP(e, j, e.n_slots) {
        uint32_t c = 0;
        walk_slot(e, cfg, j, ref, [&](uint32_t, uint32_t) { ++c; });
        e.s.slot_cnt[j] = c;
    }
    wave_or_atomic(&evs[blockIdx.y].cnt->ref_assert, ref);
}

// ------------------------------------------------------------------ K3: doublet fill
__global__ void __launch_bounds__(256) k_doublet_fill(const EvDesc *__restrict__ evs, Config cfg) {
    uint32_t ref = 0;
    PT_EVENT_LOOP(e, j, e.n_slots) {
        const uint32_t n_cells = e.cnt->n_cells;
        uint32_t base = e.s.slot_cnt[j];
        const uint32_t end = e.s.slot_cnt[j + 1];
        if (base >= n_cells || base == end) continue;
        int p = 0;
        while (j >= e.slot_cum[p]) ++p;
        walk_slot(e, cfg, j, ref, [&](uint32_t oi, uint32_t i0) {
            if (base < n_cells) {
                e.s.cell_inner[base] = (uint16_t)(i0 + 1);
                e.s.cell_outer[base] = (uint16_t)oi;
                e.s.cell_pair[base] = (uint8_t)p;
                e.s.cell_used[base] = 0;
                e.s.cell_killed[base] = 0;
                atomicAdd(&e.s.ohc_cnt[oi - 1], 1u);
            }
            ++base;
        });
    }
    (void)ref;
}

// ------------------------------------------------------------------ scans
// Exclusive scan in place of arr[0..n), total to arr[n].  One block per event,
// tiles of kScanBlock*kScanItems staged through LDS for coalesced access.
enum ScanWhat : int { kScanSlots = 0, kScanOhc = 1, kScanNn = 2, kScanTcnt = 3, kScanHcnt = 4, kScanTrk = 5 };

template <int W>
__global__ void __launch_bounds__(kScanBlock) k_scan(const EvDesc *__restrict__ evs, uint32_t max_doublets) {
    __shared__ uint32_t tile[kScanBlock * kScanItems];
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    const EvDesc &e = evs[blockIdx.y];
    uint32_t *arr;
    uint32_t n;
    if (W == kScanSlots) {
        arr = e.s.slot_cnt;
        n = e.n_slots;
    } else if (W == kScanOhc) {
        arr = e.s.ohc_cnt;
        n = e.n_hits;
    } else {
        n = e.cnt->n_cells;
        arr = (W == kScanNn) ? e.s.nn : (W == kScanTcnt) ? e.s.tcnt : (W == kScanHcnt) ? e.s.hcnt : e.s.trk_cnt;
    }
    const uint32_t T = kScanBlock * kScanItems;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += T) {
        for (uint32_t k = threadIdx.x; k < T; k += kScanBlock) tile[k] = (base + k < n) ? arr[base + k] : 0u;
        __syncthreads();
        uint32_t sum = 0;
        const uint32_t t0 = threadIdx.x * kScanItems;
#pragma unroll
        for (uint32_t q = 0; q < kScanItems; ++q) sum += tile[t0 + q];
        uint32_t total;
        uint32_t run = block_excl_scan(sum, lds, total) + carry;
#pragma unroll
        for (uint32_t q = 0; q < kScanItems; ++q) {
            const uint32_t v = tile[t0 + q];
            tile[t0 + q] = run;
            run += v;
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < T; k += kScanBlock)
            if (base + k < n) arr[base + k] = tile[k];
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        arr[n] = carry;
        if (W == kScanSlots) {
            e.cnt->n_cells_found = carry;
            e.cnt->n_cells = carry < max_doublets ? carry : max_doublets;
            if (carry > max_doublets) e.cnt->overflow |= PT_OVF_DOUBLETS;
        } else if (W == kScanNn) {
            e.cnt->n_connections = carry;
        } else if (W == kScanTcnt) {
            e.cnt->n_tuples = carry;
            if (carry > kMaxTuples) e.cnt->overflow |= PT_OVF_TUPLES;
        } else if (W == kScanHcnt) {
            e.cnt->n_tuple_hits = carry;
        }
    }
}

// ------------------------------------------------------------------ outer-hit lists
// is_outer_hit_of_cell (OuterHitOfCellVector(128, n_hits)): ascending cell ids per
// outer hit; consumers use the first 128 entries like push! keeps them.
__global__ void __launch_bounds__(256) k_ohc_fill(const EvDesc *__restrict__ evs) {
    PT_EVENT_LOOP(e, c, e.cnt->n_cells) {
        const uint32_t h0 = e.s.cell_outer[c] - 1u;
        const uint32_t pos = atomicAdd(&e.s.ohc_cursor[h0], 1u);
        e.s.ohc_list[e.s.ohc_cnt[h0] + pos] = c;
    }
}
__global__ void __launch_bounds__(256) k_ohc_sort(const EvDesc *__restrict__ evs) {
    uint32_t ovf = 0;
    PT_EVENT_LOOP(e, h, e.n_hits) {
        const uint32_t a = e.s.ohc_cnt[h], b = e.s.ohc_cnt[h + 1];
        uint32_t *L = e.s.ohc_list;
        for (uint32_t x = a + 1; x < b; ++x) {
            const uint32_t v = L[x];
            uint32_t y = x;
            while (y > a && L[y - 1] > v) {
                L[y] = L[y - 1];
                --y;
            }
            L[y] = v;
        }
        if (b - a > PT_MAX_CELLS_PER_HIT) ovf |= PT_OVF_OUTER_HIT_CELLS;
    }
    wave_or_atomic(&evs[blockIdx.y].cnt->overflow, ovf);
}

// ------------------------------------------------------------------ connect
// gpu_ca_cell.jl:130-136
__device__ __forceinline__ bool are_aligned(float r1, float z1, float ri, float zi, float ro, float zo, float pt_min,
                                            float theta_cut) {
    const float radius_diff = fabsf(r1 - ro);
    const float d13 = radius_diff * radius_diff + (z1 - zo) * (z1 - zo);
    const float p_min = pt_min * sqrtf(d13);
    const float t = fabsf(z1 * (ri - ro) + zi * (ro - r1) + zo * (r1 - ri));
    return t * p_min <= theta_cut * d13 * radius_diff;
}
// circle_eq.jl:24-44,89-93 and gpu_ca_cell.jl:138-154
__device__ __forceinline__ bool dca_cut(float x1, float y1, float x2, float y2, float x3, float y3, float region,
                                        float max_curv) {
    const bool no_flip = fabsf(x3 - x1) < fabsf(y3 - y1);
    const float x1p = no_flip ? x1 - x2 : y1 - y2;
    const float y1p = no_flip ? y1 - y2 : x1 - x2;
    const float d12 = x1p * x1p + y1p * y1p;
    const float x3p = no_flip ? x3 - x2 : y3 - y2;
    const float y3p = no_flip ? y3 - y2 : x3 - x2;
    const float d32 = x3p * x3p + y3p * y3p;
    const float num = x1p * y3p - y1p * x3p;
    const float det = d12 * y3p - d32 * y1p;
    const float st2 = (d12 * x3p - d32 * x1p);
    const float seq = det * det + st2 * st2;
    float al2 = 1.0f / sqrtf(seq);
    const float be2 = -st2 * al2;
    const float ct = 2.0f * num * al2;
    al2 *= det;
    const float m_c = no_flip ? ct : -ct;
    const float m_alpha = no_flip ? al2 : -be2;
    const float m_beta = no_flip ? be2 : -al2;
    if (m_c > max_curv) return false;
    const float xx = m_c * x2 + m_alpha;
    const float yy = m_c * y2 + m_beta;
    const float dca0 = sqrtf(xx * xx + yy * yy) - 1.0f;
    return fabsf(dca0) < region * fabsf(m_c);
}

// Candidates of outer cell o: the cells whose inner hit is o's outer hit, in
// ascending cell id, provided o is among the first 128 cells of that hit's
// outer list (the only ones kernel_connect visits).
template <class F>
__device__ __forceinline__ void for_each_candidate(const EvDesc &e, const Config &cfg, uint32_t o, uint32_t n_cells,
                                                   uint32_t &n_eval, F &&f) {
    const uint32_t h0 = e.s.cell_outer[o] - 1u;
    const uint32_t la = e.s.ohc_cnt[h0], lb = e.s.ohc_cnt[h0 + 1];
    if (lb - la > PT_MAX_CELLS_PER_HIT && o > e.s.ohc_list[la + PT_MAX_CELLS_PER_HIT - 1]) return;
    const int L = layer_of(e, h0);
    const uint32_t i1 = e.s.cell_inner[o] - 1u;
    const float r1 = e.rg[i1], z1 = e.zg[i1], x1 = e.xg[i1], y1 = e.yg[i1];
    const float region = (e.det[i1] < 96) ? cfg.p.dca_cut_inner_triplet : cfg.p.dca_cut_outer_triplet;
    const float ri = e.rg[h0], zi = e.zg[h0], x2 = e.xg[h0], y2 = e.yg[h0];
    const float theta = (e.det[h0] < 1184) ? cfg.p.ca_theta_cut_barrel : cfg.p.ca_theta_cut_forward;
    const uint32_t lpos = h0 - e.layer_start[L];
    for (uint32_t p = 0; p < e.n_pairs; ++p) {
        if (cfg.pairs.inner[p] != L) continue;
        const uint32_t j = ((p == 0) ? 0u : e.slot_cum[p - 1]) + lpos;
        uint32_t a = e.s.slot_cnt[j], b = e.s.slot_cnt[j + 1];
        if (b > n_cells) b = n_cells;
        for (uint32_t t = a; t < b; ++t) {
            const uint32_t h3 = e.s.cell_outer[t] - 1u;
            const float ro = e.rg[h3], zo = e.zg[h3], x3 = e.xg[h3], y3 = e.yg[h3];
            ++n_eval;
            const bool ok = are_aligned(r1, z1, ri, zi, ro, zo, cfg.p.pt_min, theta) &&
                            dca_cut(x1, y1, x2, y2, x3, y3, region, cfg.p.hard_curv_cut);
            if (ok && !f(t)) return;
        }
    }
}

__global__ void __launch_bounds__(256) k_connect_count(const EvDesc *__restrict__ evs, Config cfg) {
    uint32_t n_eval = 0, ovf = 0;
    PT_EVENT_LOOP(e, o, e.cnt->n_cells) {
        uint32_t npass = 0;
        for_each_candidate(e, cfg, o, e.cnt->n_cells, n_eval, [&](uint32_t t) {
            e.s.cell_used[t] = 1;
            ++npass;
            return true;
        });
        if (npass) e.s.cell_used[o] = 1;
        if (npass > PT_MAX_NEIGHBORS) ovf |= PT_OVF_NEIGHBORS;
        e.s.nn[o] = npass < PT_MAX_NEIGHBORS ? npass : PT_MAX_NEIGHBORS;
    }
    const EvDesc &e0 = evs[blockIdx.y];
    wave_add_atomic(&e0.cnt->n_connect_pairs, n_eval);
    wave_or_atomic(&e0.cnt->overflow, ovf);
}

__global__ void __launch_bounds__(256) k_connect_fill(const EvDesc *__restrict__ evs, Config cfg) {
    uint32_t dummy = 0, ovf = 0;
    PT_EVENT_LOOP(e, o, e.cnt->n_cells) {
        const uint32_t a = e.s.nn[o], b = e.s.nn[o + 1];
        if (a == b) continue;
        if (b > kNbrCap) {
            ovf |= 1u;
            continue;
        }
        uint32_t w = a;
        for_each_candidate(e, cfg, o, e.cnt->n_cells, dummy, [&](uint32_t t) {
            e.s.nbr[w++] = t;
            return w < b;
        });
    }
    if (ovf) atomicOr(&evs[blockIdx.y].s.misc[1], 1u);
}

// ------------------------------------------------------------------ fishbone (gpu_fishbone.jl:4-62)
__global__ void __launch_bounds__(256) k_fishbone(const EvDesc *__restrict__ evs) {
    uint32_t killed = 0;
    PT_EVENT_LOOP(e, h, e.n_hits) {
        const uint32_t la = e.s.ohc_cnt[h], lb = e.s.ohc_cnt[h + 1];
        uint32_t m = lb - la;
        if (m < 2) continue;
        if (m > PT_MAX_CELLS_PER_HIT) m = PT_MAX_CELLS_PER_HIT;
        const float xo = e.xg[h], yo = e.yg[h], zo = e.zg[h];
        const uint32_t *L = e.s.ohc_list + la;
        for (uint32_t a = 0; a + 1 < m; ++a) {
            const uint32_t ca = L[a];
            if (!e.s.cell_used[ca]) continue;
            const uint32_t ia = e.s.cell_inner[ca] - 1u;
            const float dxa = e.xg[ia] - xo, dya = e.yg[ia] - yo, dza = e.zg[ia] - zo;
            const float na = dxa * dxa + dya * dya + dza * dza;
            const int16_t da = e.det[ia];
            for (uint32_t b = a + 1; b < m; ++b) {
                const uint32_t cb = L[b];
                if (!e.s.cell_used[cb]) continue;
                const uint32_t ib = e.s.cell_inner[cb] - 1u;
                const float dxb = e.xg[ib] - xo, dyb = e.yg[ib] - yo, dzb = e.zg[ib] - zo;
                const float nb = dxb * dxb + dyb * dyb + dzb * dzb;
                const float cos12 = dxa * dxb + dya * dyb + dza * dzb;
                if (da != e.det[ib] && cos12 * cos12 >= 0.99999f * na * nb) {
                    if (na > nb) {
                        if (!e.s.cell_killed[ca]) ++killed;
                        e.s.cell_killed[ca] = 1;
                        break;
                    } else {
                        if (!e.s.cell_killed[cb]) ++killed;
                        e.s.cell_killed[cb] = 1;
                    }
                }
            }
        }
    }
    (void)killed;
}

// ------------------------------------------------------------------ ntuplets (gpu_ca_cell.jl:232-270)
// Iterative DFS with the serial visiting order; leaf(cells, n_cells) is called
// at every emission point of the recursive reference.
template <class F>
__device__ __forceinline__ void dfs(const EvDesc &e, uint32_t s, uint32_t min_doublets, uint32_t &ref_bits,
                                    F &&leaf) {
    uint32_t stk[4], pos[4];
    bool child[4];
    int d = 0;
    stk[0] = s;
    pos[0] = 0;
    child[0] = false;
    const uint32_t *nn = e.s.nn;
    while (d >= 0) {
        const uint32_t c = stk[d];
        const uint32_t a = nn[c], cnt = nn[c + 1] - a;
        bool pushed = false;
        while (pos[d] < cnt) {
            const uint32_t n = e.s.nbr[a + pos[d]];
            pos[d]++;
            if (n >= e.cnt->n_cells || e.s.cell_killed[n]) continue;
            child[d] = true;
            if (d + 1 >= 4) {  // the reference asserts: at most 4 doublets in a chain
                ref_bits |= PT_REF_DEPTH;
                continue;
            }
            ++d;
            stk[d] = n;
            pos[d] = 0;
            child[d] = false;
            pushed = true;
            break;
        }
        if (pushed) continue;
        if (!child[d] && (uint32_t)(d + 1) >= min_doublets) leaf(stk, (uint32_t)(d + 1));
        --d;
    }
}

__device__ __forceinline__ bool is_start(const EvDesc &e, const Config &cfg, uint32_t c) {
    if (e.s.cell_killed[c]) return false;
    const int p = e.s.cell_pair[c];
    return (cfg.p.min_hits_per_ntuplet > 3) ? (p < 3) : (p < 8 || p > 12);
}

__global__ void __launch_bounds__(256) k_dfs_count(const EvDesc *__restrict__ evs, Config cfg) {
    uint32_t ref = 0, killed = 0;
    const uint32_t min_d = cfg.p.min_hits_per_ntuplet - 1;
    PT_EVENT_LOOP(e, c, e.cnt->n_cells) {
        killed += e.s.cell_killed[c];
        uint32_t nt = 0, nh = 0;
        if (is_start(e, cfg, c)) {
            dfs(e, c, min_d, ref, [&](const uint32_t *, uint32_t len) {
                ++nt;
                nh += len + 1;
            });
        }
        e.s.tcnt[c] = nt;
        e.s.hcnt[c] = nh;
        e.s.trk_cnt[c] = 0;
        e.s.trk_cursor[c] = 0;
    }
    wave_or_atomic(&evs[blockIdx.y].cnt->ref_assert, ref);
    wave_add_atomic(&evs[blockIdx.y].cnt->n_killed, killed);
}

__global__ void __launch_bounds__(256) k_dfs_emit(const EvDesc *__restrict__ evs, Config cfg) {
    uint32_t ref = 0;
    const uint32_t min_d = cfg.p.min_hits_per_ntuplet - 1;
    PT_EVENT_LOOP(e, c, e.cnt->n_cells) {
        uint32_t it = e.s.tcnt[c];  // 0-based id of the first tuple of this start cell
        if (it == e.s.tcnt[c + 1]) continue;
        uint32_t hp = e.s.hcnt[c];
        dfs(e, c, min_d, ref, [&](const uint32_t *stk, uint32_t len) {
            ++it;  // 1-based tuple id (bulk_fill)
            if (it <= kMaxTuples) {
                e.hit_off[it - 1] = hp + 1u;
                for (uint32_t q = 0; q < len; ++q) e.hit_bins[hp + q] = e.s.cell_inner[stk[q]];
                e.hit_bins[hp + len] = e.s.cell_outer[stk[len - 1]];
                e.quality[it - 1] = PT_QUALITY_BAD;
                uint32_t *tc = e.s.tuple_cells + 4u * (it - 1u);
                for (uint32_t q = 0; q < 4; ++q) tc[q] = (q < len) ? stk[q] : 0xFFFFFFFFu;
                for (uint32_t q = 0; q < len; ++q) atomicAdd(&e.s.trk_cnt[stk[q]], 1u);
            }
            hp += len + 1;
        });
    }
    (void)ref;
}

// bulk_finalize_fill (hist_to_container.jl:150-161)
__global__ void __launch_bounds__(256) k_finalize_tuples(const EvDesc *__restrict__ evs) {
    PT_EVENT_LOOP(e, i, kMaxTuples + 1u) {
        const uint32_t m = e.cnt->n_tuples, n = e.cnt->n_tuple_hits;
        if (m > kMaxTuples) {
            if (i == kMaxTuples) e.hit_off[kMaxTuples] = e.hit_off[kMaxTuples - 1];
        } else if (i >= m) {
            e.hit_off[i] = n + 1u;
        }
    }
}

// cell -> tuple lists (CellTracksVector), ascending tuple id
__global__ void __launch_bounds__(256) k_trk_fill(const EvDesc *__restrict__ evs) {
    PT_EVENT_LOOP(e, t, (e.cnt->n_tuples < kMaxTuples ? e.cnt->n_tuples : kMaxTuples)) {
        const uint32_t *tc = e.s.tuple_cells + 4u * t;
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t c = tc[q];
            if (c == 0xFFFFFFFFu) break;
            const uint32_t pos = atomicAdd(&e.s.trk_cursor[c], 1u);
            const uint32_t w = e.s.trk_cnt[c] + pos;
            if (w < kTrkCap) e.s.trk_list[w] = (uint16_t)(t + 1u);
        }
    }
}
__global__ void __launch_bounds__(256) k_trk_sort_dup(const EvDesc *__restrict__ evs) {
    uint32_t ovf = 0;
    PT_EVENT_LOOP(e, c, e.cnt->n_cells) {
        const uint32_t a = e.s.trk_cnt[c], b = e.s.trk_cnt[c + 1];
        if (b - a < 2) continue;
        uint16_t *L = e.s.trk_list;
        for (uint32_t x = a + 1; x < b; ++x) {
            const uint16_t v = L[x];
            uint32_t y = x;
            while (y > a && L[y - 1] > v) {
                L[y] = L[y - 1];
                --y;
            }
            L[y] = v;
        }
        uint32_t m = b - a;
        if (m > PT_MAX_CELL_TRACKS) {
            ovf |= PT_OVF_CELL_TRACKS;
            m = PT_MAX_CELL_TRACKS;
        }
        // kernel_early_duplicate_remover (ca_hit_ntuplet_generator_kernels_impl.jl:143-166)
        uint32_t maxh = 0;
        for (uint32_t k = 0; k < m; ++k) {
            const uint32_t it = L[a + k];
            const uint32_t sz = e.hit_off[it] - e.hit_off[it - 1];
            maxh = sz > maxh ? sz : maxh;
        }
        for (uint32_t k = 0; k < m; ++k) {
            const uint32_t it = L[a + k];
            if (e.hit_off[it] - e.hit_off[it - 1] != maxh) e.quality[it - 1] = PT_QUALITY_DUP;
        }
    }
    wave_or_atomic(&evs[blockIdx.y].cnt->overflow, ovf);
}

// ------------------------------------------------------------------ multiplicity
// kernel_countMultiplicity / finalize! / kernel_fillMultiplicity
// (ca_hit_ntuplet_generator_kernels_impl.jl:174-211): bucket = hit count, each
// bucket lists tuple ids in DESCENDING order (fill_direct decrements).
__global__ void __launch_bounds__(1024) k_multiplicity(const EvDesc *__restrict__ evs) {
    __shared__ uint32_t lds[5][1024 / 64 + 1];
    __shared__ uint32_t tot[5];
    const EvDesc &e = evs[blockIdx.y];
    const uint32_t m = e.cnt->n_tuples < kMaxTuples ? e.cnt->n_tuples : kMaxTuples;
    const uint32_t chunk = (m + blockDim.x - 1) / blockDim.x;
    const uint32_t t0 = threadIdx.x * chunk, t1 = min(m, t0 + chunk);
    uint32_t c[5] = {0, 0, 0, 0, 0};
    uint32_t ref = 0;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t nh = e.hit_off[t + 1] - e.hit_off[t];
        if (nh < 3 || e.quality[t] == PT_QUALITY_DUP) continue;
        if (nh >= 8) {
            ref |= PT_REF_MULT;
            continue;
        }
        c[nh - 3]++;
    }
    uint32_t ex[5];
    for (int b = 0; b < 5; ++b) {
        uint32_t total;
        ex[b] = block_excl_scan(c[b], lds[b], total);
        if (threadIdx.x == 0) tot[b] = total;
    }
    __syncthreads();
    // off[] after finalize!: exclusive start of bucket nh at tm_off[nh-1]
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = 0; b < 9; ++b) {
            e.s.tm_off[b] = run;
            if (b >= 2 && b < 7) run += tot[b - 2];
        }
        for (int b = 3; b < 8; ++b) e.cnt->n_by_len[b] = tot[b - 3];
        for (int b = 3; b <= 5; ++b)
            if (tot[b - 3] > PT_MAX_FITS) e.cnt->overflow |= PT_OVF_FITS;
    }
    __syncthreads();
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t nh = e.hit_off[t + 1] - e.hit_off[t];
        if (nh < 3 || nh >= 8 || e.quality[t] == PT_QUALITY_DUP) continue;
        const int b = nh - 3;
        const uint32_t rank = ex[b]++;
        e.s.tm_bins[e.s.tm_off[nh - 1] + (tot[b] - 1u - rank)] = (uint16_t)(t + 1u);
    }
    if (ref) atomicOr(&e.cnt->ref_assert, ref);
}

// ------------------------------------------------------------------ hit -> module indices
// kernel_fill_hit_indices (ca_hit_ntuplet_generator_kernels_impl.jl:25-38)
__global__ void __launch_bounds__(256) k_fill_det(const EvDesc *__restrict__ evs) {
    PT_EVENT_LOOP(e, i, kMaxTuples + 1u + PT_MAX_TUPLE_HITS) {
        if (i <= kMaxTuples) {
            e.det_off[i] = e.hit_off[i];
        } else {
            const uint32_t k = i - (kMaxTuples + 1u);
            const uint32_t sz = e.hit_off[kMaxTuples] - 1u;
            if (k < sz) e.det_bins[k] = (uint16_t)e.det[e.hit_bins[k] - 1u];
        }
    }
}

// ------------------------------------------------------------------ launchers
static inline dim3 grid_for(uint32_t items, uint32_t n_events, uint32_t bs = 256) {
    uint32_t per = (items + bs - 1) / bs;
    const uint32_t target = (2048u + n_events - 1) / n_events;
    if (per > target) per = target;
    if (per < 1) per = 1;
    return dim3(per, n_events, 1);
}

// Launch sequence of the CA stage.  When `tm` is given, an event is recorded
// after every kernel so the host can read per-kernel durations.
void launch_ca(const EvDesc *d_evs, uint32_t n_events, const Config &cfg, uint32_t max_hits, uint32_t max_slots,
               int zero_state, hipStream_t st, Marks *tm) {
    const uint32_t est_cells = max_slots * 4u + 1024u;
    auto mark = [&](const char *name) {
        if (tm) tm->record(name, st);
    };
    mark("start");
    k_reset<<<grid_for(max_hits + 33u + (zero_state ? 20u * kMaxTuples : 0u), n_events), 256, 0, st>>>(d_evs,
                                                                                                          zero_state);
    mark("k_reset");
    k_doublet_count<<<grid_for(max_slots, n_events), 256, 0, st>>>(d_evs, cfg);
    mark("k_doublet_count");
    k_scan<kScanSlots><<<dim3(1, n_events), kScanBlock, 0, st>>>(d_evs, cfg.p.max_num_of_doublets);
    mark("k_scan_slots");
    k_doublet_fill<<<grid_for(max_slots, n_events), 256, 0, st>>>(d_evs, cfg);
    mark("k_doublet_fill");
    k_scan<kScanOhc><<<dim3(1, n_events), kScanBlock, 0, st>>>(d_evs, 0);
    mark("k_scan_ohc");
    k_ohc_fill<<<grid_for(est_cells, n_events), 256, 0, st>>>(d_evs);
    mark("k_ohc_fill");
    k_ohc_sort<<<grid_for(max_hits, n_events), 256, 0, st>>>(d_evs);
    mark("k_ohc_sort");
    k_connect_count<<<grid_for(est_cells, n_events), 256, 0, st>>>(d_evs, cfg);
    mark("k_connect_count");
    k_scan<kScanNn><<<dim3(1, n_events), kScanBlock, 0, st>>>(d_evs, 0);
    mark("k_scan_nn");
    k_connect_fill<<<grid_for(est_cells, n_events), 256, 0, st>>>(d_evs, cfg);
    mark("k_connect_fill");
    if (cfg.p.early_fish_bone) k_fishbone<<<grid_for(max_hits, n_events), 256, 0, st>>>(d_evs);
    mark("k_fishbone");
    k_dfs_count<<<grid_for(est_cells, n_events), 256, 0, st>>>(d_evs, cfg);
    mark("k_dfs_count");
    k_scan<kScanTcnt><<<dim3(1, n_events), kScanBlock, 0, st>>>(d_evs, 0);
    mark("k_scan_tcnt");
    k_scan<kScanHcnt><<<dim3(1, n_events), kScanBlock, 0, st>>>(d_evs, 0);
    mark("k_scan_hcnt");
    k_dfs_emit<<<grid_for(est_cells, n_events), 256, 0, st>>>(d_evs, cfg);
    mark("k_dfs_emit");
    k_finalize_tuples<<<grid_for(kMaxTuples + 1u, n_events), 256, 0, st>>>(d_evs);
    mark("k_finalize_tuples");
    k_scan<kScanTrk><<<dim3(1, n_events), kScanBlock, 0, st>>>(d_evs, 0);
    mark("k_scan_trk");
    k_trk_fill<<<grid_for(kMaxTuples, n_events), 256, 0, st>>>(d_evs);
    mark("k_trk_fill");
    k_trk_sort_dup<<<grid_for(est_cells, n_events), 256, 0, st>>>(d_evs);
    mark("k_trk_sort_dup");
    k_multiplicity<<<dim3(1, n_events), 1024, 0, st>>>(d_evs);
    mark("k_multiplicity");
    k_fill_det<<<grid_for(kMaxTuples + 1u + PT_MAX_TUPLE_HITS, n_events), 256, 0, st>>>(d_evs);
    mark("k_fill_det");
}

}  // namespace pt
