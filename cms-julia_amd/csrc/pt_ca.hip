// pt_ca.hip - CA stage of the pixel-track path for gfx950: doublets, outer-hit
// lists, cell connection, fishbone, ntuplet search, early duplicate removal,
// tuple multiplicity and hit->module indices.
//
// Every stage reproduces the serial creation order of the reference, so ids,
// caps and list contents are identical, without serialising the work:
//   * doublets (gpu_pixel_doublets.jl:187-355): count the passing candidates of
//     each inner slot, then give slot j the cell ids after those of slots < j
//     (cell id = serial doublet id - 1) and write the cells;
//   * kernel_connect (ca_hit_ntuplet_generator_kernels_impl.jl:45-114) is laid
//     out per hit h: every (cell ending on h) x (cell starting on h) pair is one
//     lane's work, and a cell's neighbours come out in ascending id order, the
//     serial push! order, capped at 64;
//   * fishbone (gpu_fishbone.jl:4-62) likewise per hit, one lane per cell pair;
//   * ntuplets (gpu_ca_cell.jl:232-270, impl.jl:116-132): count the tuples of
//     each start cell, scan in cell order, emit, so tuple ids equal the serial
//     DFS emission order.
// The kernels are latency-bound gathers: work is spread so that every lane of
// a wavefront is busy and several independent loads are in flight per lane,
// and prefix sums use per-tile totals instead of waiting on other tiles.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "pt_device.hpp"

namespace pt {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ void wave_add_atomic(uint32_t *addr, uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (lane_id() == 0 && v) atomicAdd(addr, v);
}
__device__ __forceinline__ void wave_or_atomic(uint32_t *addr, uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v |= __shfl_xor(v, d, 64);
    if (lane_id() == 0 && v) atomicOr(addr, v);
}
// orders the wavefront's own LDS accesses (one wavefront, no block barrier)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d, 64));
    return v;
}
// bits [s0, s1) of a 64-bit mask, 0 <= s0 <= s1 <= 64
__device__ __forceinline__ uint64_t range_mask(uint32_t s0, uint32_t s1) {
    const uint64_t hi = (s1 >= 64u) ? ~0ull : ((1ull << s1) - 1ull);
    const uint64_t lo = (s0 >= 64u) ? ~0ull : ((1ull << s0) - 1ull);
    return hi & ~lo;
}
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

// One ticket per wavefront from a per-event counter.
__device__ __forceinline__ uint32_t take_ticket(uint32_t *ctr) {
    uint32_t t = 0;
    if (lane_id() == 0) t = atomicAdd(ctr, 1u);
    return __shfl(t, 0, 64);
}

// owner of item g when lane l owns items [incl_l - n_l, incl_l): the first lane
// whose inclusive prefix exceeds g
__device__ __forceinline__ int owner_of(uint32_t incl, uint32_t g) {
    int lo = 0, hi = 63;
#pragma unroll
    for (int step = 0; step < 6; ++step) {
        const int mid = (lo + hi) >> 1;
        if ((uint32_t)__shfl(incl, mid, 64) > g)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// owner_of for N items at once, the searches in lockstep so that their
// cross-lane reads are in flight together
template <int N>
__device__ __forceinline__ void owners_of(uint32_t incl, const uint32_t (&g)[N], int (&o)[N]) {
    int lo[N], hi[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
        lo[u] = 0;
        hi[u] = 63;
    }
#pragma unroll
    for (int step = 0; step < 6; ++step) {
        int mid[N];
        uint32_t v[N];
#pragma unroll
        for (int u = 0; u < N; ++u) {
            mid[u] = (lo[u] + hi[u]) >> 1;
            v[u] = (uint32_t)__shfl(incl, mid[u], 64);
        }
#pragma unroll
        for (int u = 0; u < N; ++u) {
            if (v[u] > g[u])
                hi[u] = mid[u];
            else
                lo[u] = mid[u] + 1;
        }
    }
#pragma unroll
    for (int u = 0; u < N; ++u) o[u] = lo[u];
}

// position of the r-th set bit (r < popcount(m))
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t r) {
    uint32_t pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint32_t lo = (uint32_t)__popcll(m & ((1ull << w) - 1ull));
        if (r >= lo) {
            r -= lo;
            m >>= w;
            pos += (uint32_t)w;
        }
    }
    return pos;
}

// position of the r-th set bit (0-based) of the 192-bit mask (w1 lowest)
__device__ __forceinline__ uint32_t nth_set_bit3(uint64_t w1, uint64_t w2, uint64_t w3, uint32_t r) {
    const uint32_t c1 = (uint32_t)__popcll(w1), c2 = (uint32_t)__popcll(w2);
    if (r < c1) return nth_set_bit(w1, r);
    r -= c1;
    if (r < c2) return 64u + nth_set_bit(w2, r);
    return 128u + nth_set_bit(w3, r - c2);
}

// Wavefront owners hold lane-contiguous items; after a round of 64 items
// starting at `base`, fold the pass bits of this lane's own items (lanes
// [s0, s1), item k at lane k - (excl - base)) into its count and 64-bit mask.
__device__ __forceinline__ void fold_round(uint64_t pass, uint32_t excl, uint32_t incl, uint32_t base, uint32_t &cnt,
                                           uint64_t &msk) {
    const int64_t d = (int64_t)excl - (int64_t)base;
    const int64_t d1 = (int64_t)incl - (int64_t)base;
    const uint32_t s0 = (uint32_t)(d < 0 ? 0 : (d > 64 ? 64 : d));
    const uint32_t s1 = (uint32_t)(d1 < 0 ? 0 : (d1 > 64 ? 64 : d1));
    if (s1 > s0) {
        const uint64_t seg = pass & range_mask(s0, s1);
        cnt += (uint32_t)__popcll(seg);
        msk |= (d >= 0) ? (seg >> d) : ((-d < 64) ? (seg << (-d)) : 0ull);
    }
}

// Block-wide (256 threads) sums and exclusive scans; lds holds 4 words.
template <class T>
__device__ __forceinline__ T block_sum(T v, T *lds) {
    const uint32_t wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (lane_id() == 0) lds[wid] = v;
    __syncthreads();
    const T r = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return r;
}
template <class T>
__device__ __forceinline__ T block_excl_scan4(T v, T *lds, T &total) {
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
    T inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T o = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += o;
    }
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    T pre = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
        if (k < wid) pre += lds[k];
    total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return pre + inc - v;
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

// Inner-hit part of the doublet search for inner slot j (gpu_pixel_doublets.jl:187-308):
// the inner-hit cuts and the candidate range(s): bins kl..kh-1 of the outer
// layer are contiguous histogram positions (two ranges when the window wraps
// past bin 128), visited by the serial loop in position order.
struct SlotCtx {
    float me_z, me_r;
    int32_t me_s;
    uint32_t pk;          // me_p (16 b) | pair << 16
    uint32_t a0, na, b0;  // candidates: positions [a0, a0+na) then [b0, ...)
    // the pair's constants the candidate cuts use, so that the candidate loop
    // reads them with the slot instead of indexing the pair tables per lane
    uint32_t lo, hi;      // outer-layer hit ids (lo, hi], 1-based
    float max_r;
    uint32_t cut;         // phi cut (16 b) | inner layer << 16 | outer layer << 24
};

__device__ __forceinline__ int slot_pair(const SlotCtx &s) { return (int)(s.pk >> 16); }
__device__ __forceinline__ int16_t slot_me_p(const SlotCtx &s) { return (int16_t)(s.pk & 0xFFFFu); }

// The layer-pair constants of the block's event in LDS: a lane indexes them
// by its own pair, which from the kernel arguments or the event descriptor
// would be a chain of memory reads per lane.
struct PairLds {
    uint32_t slot_cum[20];                  // as EvDesc::slot_cum
    uint32_t ls_inner[19], lo[19], hi[19];  // inner layer's first hit; outer layer's hits (lo, hi]
    int32_t inner[19], outer[19], cut[19];
    float min_z[19], max_z[19], max_r[19];
    int32_t by_inner[10][6], n_by_inner[10];
};
// every thread of the block calls it (it ends with a barrier)
__device__ __forceinline__ void load_pairs(PairLds &P, const EvDesc &e, const Config &cfg) {
    const uint32_t t = threadIdx.x;
    if (t < 20u) P.slot_cum[t] = e.slot_cum[t];
    if (t < 19u) {
        const int in = cfg.pairs.inner[t], out = cfg.pairs.outer[t];
        P.inner[t] = in;
        P.outer[t] = out;
        P.cut[t] = cfg.pairs.phi_cut[t];
        P.min_z[t] = cfg.pairs.min_z[t];
        P.max_z[t] = cfg.pairs.max_z[t];
        P.max_r[t] = cfg.pairs.max_r[t];
        P.ls_inner[t] = e.layer_start[in];
        P.lo[t] = e.layer_start[out];
        P.hi[t] = e.layer_start[out + 1];
    }
    if (t < 60u) P.by_inner[t / 6u][t % 6u] = cfg.pairs.by_inner[t / 6u][t % 6u];
    if (t < 10u) P.n_by_inner[t] = cfg.pairs.n_by_inner[t];
    __syncthreads();
}

// the pair of inner slot j: slot_cum is non-decreasing (unused pairs hold
// 0xFFFFFFFF), so the first pair with j < slot_cum is the count of the others
__device__ __forceinline__ int pair_of_slot(const EvDesc &e, uint32_t j) {
    int p = 0;
#pragma unroll
    for (int k = 0; k < 20; ++k) p += (j >= e.slot_cum[k]) ? 1 : 0;
    return p;
}

// Inner-slot setup (gpu_pixel_doublets.jl:262-316) in three steps around its
// two memory reads, so that the reads of several slots can be in flight
// together: the inner-hit record gives the cuts and the phi window, the
// histogram offsets of the window give the candidate ranges.
struct SlotPrep {
    uint32_t hoff, kl, kh;  // window bins (kl == kh: none); hoff = 128 * outer layer
    bool ok;
};
__device__ __forceinline__ SlotPrep slot_prep(const EvDesc &e, const Config &cfg, const PairLds &T, int p, uint4 q,
                                              uint32_t &ref_bits, SlotCtx &s) {
    const pt_params &P = cfg.p;
    const int inner = T.inner[p], outer = T.outer[p];
    SlotPrep r{0u, 0u, 0u, false};
    const int16_t mi = (int16_t)(q.z >> 16);
    s.pk = (uint32_t)p << 16;
    if (mi > 2000) return r;
    s.me_z = __uint_as_float(q.x);
    if (s.me_z < T.min_z[p] || s.me_z > T.max_z[p]) return r;
    int me_s = -1;
    if (P.do_cluster_cut) {
        if (inner == 0 && !(mi < 96)) ref_bits |= PT_REF_BAD_MODULE;
        const bool is_outer_ladder = P.ideal_conditions ? true : (((mi / 8) % 2) == 0);
        me_s = (inner > 0 || is_outer_ladder) ? (int)(int16_t)q.w : -1;
        if (inner == 0 && outer > 3 && me_s > 0 && me_s < 36) return r;
        if (inner == 1 && outer > 3 && me_s > 0 && me_s < 28) return r;
    }
    s.me_s = me_s;
    const int16_t me_p = (int16_t)(q.z & 0xFFFFu);
    s.me_r = __uint_as_float(q.y);
    s.pk = (uint32_t)(uint16_t)me_p | ((uint32_t)p << 16);
    const int16_t cut = (int16_t)T.cut[p];
    s.lo = T.lo[p];
    s.hi = T.hi[p];
    s.max_r = T.max_r[p];
    s.cut = (uint32_t)(uint16_t)cut | ((uint32_t)inner << 16) | ((uint32_t)outer << 24);
    const int kl = hist_bin((int16_t)(me_p - cut));
    int kh = hist_bin((int16_t)(me_p + cut)) + 1;
    if (kh == 129) kh = 1;
    if (kl == kh) return r;  // while (current_bin != kh) visits no bin
    r.hoff = 128u * (uint32_t)outer;
    r.kl = (uint32_t)kl;
    r.kh = (uint32_t)kh;
    r.ok = true;
    return r;
}
// histogram offsets of the window (Julia off[k + h_off]): kl..kh-1 or, when it
// wraps, kl..128 and 1..kh-1
__device__ __forceinline__ void slot_offsets(const EvDesc &e, const SlotPrep &r, uint32_t (&o)[4]) {
    const uint32_t ho = r.hoff;
    o[0] = e.hist_off[r.kl + ho - 1u];
    if (r.kl < r.kh) {
        o[1] = e.hist_off[r.kh + ho - 1u];
        o[2] = o[3] = 0u;
    } else {
        o[1] = e.hist_off[128u + ho];
        o[2] = e.hist_off[ho];
        o[3] = e.hist_off[r.kh + ho - 1u];
    }
}
// candidate ranges from the offsets, clamped to the hit count; returns the
// number of candidates
__device__ __forceinline__ uint32_t slot_window(const EvDesc &e, const SlotPrep &r, const uint32_t (&o)[4],
                                                uint32_t &ref_bits, SlotCtx &s) {
    auto clampo = [&](uint32_t v) {
        if (v > e.n_hits) {
            ref_bits |= PT_REF_HIST_LAYER;
            v = e.n_hits;
        }
        return v;
    };
    s.a0 = clampo(o[0]);
    if (r.kl < r.kh) {
        const uint32_t a1 = clampo(o[1]);
        s.na = a1 > s.a0 ? a1 - s.a0 : 0u;
        s.b0 = 0;
        return s.na;
    }
    const uint32_t a1 = clampo(o[1]);  // off(129)
    s.na = a1 > s.a0 ? a1 - s.a0 : 0u;
    s.b0 = clampo(o[2]);
    const uint32_t b1 = clampo(o[3]);
    return s.na + (b1 > s.b0 ? b1 - s.b0 : 0u);
}

// returns the number of candidates (0 when the inner hit fails its cuts)
__device__ __forceinline__ uint32_t slot_setup(const EvDesc &e, const Config &cfg, const PairLds &T, uint32_t j,
                                               uint32_t &ref_bits, SlotCtx &s) {
    const int p = pair_of_slot(e, j);
    const uint32_t i0 = ((p == 0) ? j : j - T.slot_cum[p - 1]) + T.ls_inner[p];
    const SlotPrep r = slot_prep(e, cfg, T, p, e.s.hdbl[i0], ref_bits, s);
    if (!r.ok) return 0;
    uint32_t o[4];
    slot_offsets(e, r, o);
    return slot_window(e, r, o, ref_bits, s);
}

__device__ __forceinline__ uint32_t cand_pos(uint32_t a0, uint32_t na, uint32_t b0, uint32_t k) {
    return (k < na) ? a0 + k : b0 + (k - na);
}
__device__ __forceinline__ uint32_t cand_pos(const SlotCtx &s, uint32_t k) { return cand_pos(s.a0, s.na, s.b0, k); }

// Candidate cuts (gpu_pixel_doublets.jl:317-340) on the hit record q stored at
// a histogram position (q.w >> 16 = 1-based hit id).  The cuts are combined
// into one flag instead of returning at the first failure, so the wavefront's
// lanes (candidates of different slots and layer pairs) do not split into
// branches at every cut; a reference assertion is flagged only where the
// serial loop would reach it (all earlier cuts passed).
__device__ __forceinline__ bool cand_pass(const EvDesc &e, const Config &cfg, const SlotCtx &s, uint4 q,
                                          uint32_t &ref_bits) {
    const pt_params &P = cfg.p;
    const int inner = (int)((s.cut >> 16) & 0xFFu), outer = (int)(s.cut >> 24);
    const uint32_t oi = q.w >> 16;
    const bool in_layer = oi > s.lo && oi <= s.hi;
    ref_bits |= in_layer ? 0u : PT_REF_HIST_LAYER;
    // (bitwise & and |: every operand is cheap and safe to evaluate, so no branches)
    bool ok = in_layer | ((oi != 0) & (oi <= e.n_hits));
    ok &= (int16_t)(q.z >> 16) <= 2000;
    const float zo = __uint_as_float(q.x), ro = __uint_as_float(q.y);
    if (P.do_z0_cut) {
        // gpu_pixel_doublets.jl:69-74
        const float dr = ro - s.me_r;
        ok &= !((dr > s.max_r) | (dr < 0.0f) | (fabsf(s.me_z * ro - s.me_r * zo) > 12.0f * dr));
    }
    const int16_t cut = (int16_t)(s.cut & 0xFFFFu);
    int16_t i_dphi = (int16_t)((int16_t)(q.z & 0xFFFFu) - slot_me_p(s));
    i_dphi = (int16_t)(i_dphi < 0 ? -i_dphi : i_dphi);  // Int16 abs wraps at -32768
    ok &= !(i_dphi > cut);
    if (P.do_cluster_cut) {
        // gpu_pixel_doublets.jl:79-88
        const int me_s = s.me_s;
        if (outer < 4) {
            const int so = (int16_t)(q.w & 0xFFFFu);
            ok &= !((me_s > 0) & (so > 0) & (abs(so - me_s) > ((inner == 0) ? 28 : 20)));
        } else if (inner < 4 && me_s > 0 && ok) {
            const float qq = fabsf((s.me_z - zo) / (s.me_r - ro)) * (float)(8 * 0.0285 / 0.015);
            const double t = trunc((double)qq + 0.5);
            const bool bad = !isfinite(t) || fabs(t) > 9.0e15;
            ref_bits |= bad ? PT_REF_SIZE_CUT_INF : 0u;
            ok = !bad && !(llabs((long long)me_s - (long long)(bad ? 0.0 : t)) > 20);
        }
    }
    if (P.do_pt_cut) ok &= !pt_cut(ro, i_dphi, s.me_r);
    return ok;
}

// state and covariance can be cleared with 16-byte stores
__device__ __forceinline__ bool zero_vec(const EvDesc &e) {
    return ((reinterpret_cast<uintptr_t>(e.state) | reinterpret_cast<uintptr_t>(e.cov)) & 15u) == 0u;
}
// Clearing state and covariance (20 floats per tuple row, 2.5 MB per event) is
// pure store bandwidth; the doublet count's wavefronts take it on by ticket
// once its tiles are gone, so it fills that latency-bound kernel's tail
// instead of holding every CU in a pass of its own.
constexpr uint32_t kZeroChunk = 512;  // float4 per ticket
__device__ __forceinline__ void zero_chunks(const EvDesc &e) {
    constexpr uint32_t n4 = 5u * kMaxTuples;  // float4: state, then covariance
    constexpr uint32_t n_state4 = 5u * kMaxTuples / 4u;
    // non-temporal stores: the cleared lines are not kept in the caches the CA
    // kernels work from
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v zv = {0.f, 0.f, 0.f, 0.f};
    for (;;) {
        const uint32_t c = take_ticket(&e.s.misc[kMiscTkZero]);
        if (c * kZeroChunk >= n4) break;
#pragma unroll
        for (uint32_t v = 0; v < kZeroChunk / 64u; ++v) {
            const uint32_t k = c * kZeroChunk + 64u * v + (threadIdx.x & 63u);
            if (k < n_state4)
                __builtin_nontemporal_store(zv, &reinterpret_cast<PT_G f4v *>(e.state)[k]);
            else if (k < n4)
                __builtin_nontemporal_store(zv, &reinterpret_cast<PT_G f4v *>(e.cov)[k - n_state4]);
        }
    }
}

// ------------------------------------------------------------------ K0: reset per event
// Clears the counters, list cursors and per-tile totals; packs the per-hit
// fields the CA reads into 16-byte records, once in hit order and once in
// histogram order (so a slot's candidates are contiguous records); and
// (unless the caller's TrackSOA is already zero, as TrackSOA() makes it)
// zeroes state and covariance.
__global__ void __launch_bounds__(256) k_reset(const EvDesc *__restrict__ evs, EvGrid grid, int zero_state) {
    constexpr uint32_t kHead = 32u;
    // state and covariance are cleared here only when they are not both 16-byte
    // aligned; otherwise the doublet count clears them (zero_chunks).  The
    // per-cell bits and tile sums are cleared by the doublet fill, which knows
    // the cell count.
    PT_EVENT_LOOP(e, i, kHead + e.n_hits + (e.n_slots + kSlotTile - 1) / kSlotTile +
                            ((zero_state && !zero_vec(e)) ? 20u * kMaxTuples : 0u)) {
        uint32_t k = i;
        if (k < kHead) {
            if (k < sizeof(pt_counters) / 4) reinterpret_cast<uint32_t *>(e.cnt)[k] = 0u;
            if (k < 16u) e.s.misc[k] = 0u;
            if (k == 0) e.cnt->n_hits = e.n_hits;
            if (k == 1) e.cnt->n_inner_slots = e.n_slots;
            continue;
        }
        k -= kHead;
        if (k < e.n_hits) {
            e.s.ohc_cursor[k] = 0u;
            e.s.fish_cnt[k] = 0u;
            e.s.h2t_win[k] = ~0ull;
            const float z = e.zg[k], r = e.rg[k];
            e.s.hxyzr[k] = make_float4(e.xg[k], e.yg[k], z, r);
            e.s.hdbl[k] = make_uint4(__float_as_uint(z), __float_as_uint(r),
                                     (uint32_t)(uint16_t)e.iphi[k] | ((uint32_t)(uint16_t)e.det[k] << 16),
                                     (uint32_t)(uint16_t)e.ysize[k]);
            const uint32_t h = e.hist_bins[k];  // 1-based hit at histogram position k
            uint4 rec = make_uint4(0u, 0u, 0x7FFFu << 16, h << 16);
            float2 xy = make_float2(0.f, 0.f);
            if (h >= 1u && h <= e.n_hits) {
                const uint32_t h0 = h - 1u;
                rec = make_uint4(__float_as_uint(e.zg[h0]), __float_as_uint(e.rg[h0]),
                                 (uint32_t)(uint16_t)e.iphi[h0] | ((uint32_t)(uint16_t)e.det[h0] << 16),
                                 (uint32_t)(uint16_t)e.ysize[h0] | (h << 16));
                xy = make_float2(e.xg[h0], e.yg[h0]);
            }
            e.s.binrec[k] = rec;
            e.s.binxy[k] = xy;
            continue;
        }
        k -= e.n_hits;
        const uint32_t n_dbl_agg = (e.n_slots + kSlotTile - 1) / kSlotTile;
        if (k < n_dbl_agg) {
            e.s.dbl_agg[k] = 0u;
            continue;
        }
        k -= n_dbl_agg;
        if (k < 5u * kMaxTuples)
            e.state[k] = 0.0f;
        else
            e.cov[k - 5u * kMaxTuples] = 0.0f;
    }
}

// ------------------------------------------------------------------ K1: doublet count
// A wavefront takes a tile of 64 consecutive inner slots.  Their candidates
// (about 16 per slot, lane-contiguous in slot order) are spread over all 64
// lanes, U rounds of 64 loads in flight at a time (U = 1, 2 or 4, tuned).  Each item finds its
// slot through an owner map in LDS (a slot's lane at its first item, carried
// forward by a running maximum), and each round's pass ballot is stored as a
// 64-bit word; afterwards every lane reads its own slot's bits back: the
// count and, for the first 256 candidates, the bit mask (the first 64 in
// slot_mask, the next 192 in slot_mask_ext for slots with more; at PU200 a
// third of the slots have more than 64 candidates).  Windows of
// kCountWin items bound the map.  Per tile of kSlotTile slots the counts are
// also summed.
constexpr uint32_t kCountWin = 1024;
static_assert(kCountWin == 64u * 16u, "window = 16 words per lane");

struct alignas(16) SlotItem {
    SlotCtx s;
    uint32_t excl;  // the slot's first item in the tile
};

template <int U>
__global__ void __launch_bounds__(256, 5) k_doublet_count(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg,
                                                         int zero_state) {
    static_assert(kCountWin % (64u * U) == 0, "whole rounds per window");
    __shared__ SlotItem ctx[4][64];
    __shared__ uint4 omap[4][kCountWin / 4];
    __shared__ uint64_t pbits[4][kCountWin / 64 + 1];
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    __shared__ PairLds PT;
    load_pairs(PT, e, cfg);
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t n_slots = e.n_slots;
    const uint32_t n_tiles = (n_slots + 63u) / 64u;
    const uint32_t *own = reinterpret_cast<const uint32_t *>(&omap[w][0]);
    uint32_t ref = 0;
    for (;;) {
        const uint32_t t = take_ticket(&e.s.misc[kMiscTkCount]);
        if (t >= n_tiles) break;
        const uint32_t j0 = t * 64u, j = j0 + lane;
        wave_sync();  // the previous tile's reads of ctx are done
        SlotCtx s = {};
        const uint32_t n = (j < n_slots) ? slot_setup(e, cfg, PT, j, ref, s) : 0u;
        const uint32_t incl = wave_incl_scan(n);
        const uint32_t excl = incl - n;
        const uint32_t total = __shfl(incl, 63, 64);
        ctx[w][lane] = SlotItem{s, excl};
        uint32_t cnt = 0;
        uint64_t msk = 0, mx1 = 0, mx2 = 0, mx3 = 0;  // pass bits of candidates 0..255
        for (uint32_t w0 = 0; w0 < total; w0 += kCountWin) {
            const uint32_t wn = min(total - w0, kCountWin);
            // owner map of items [w0, w0 + wn): lane + 1 where a slot's items start
            // (or continue from the previous window), zero elsewhere; then the
            // running maximum gives every item its slot's lane + 1
#pragma unroll
            for (int v = 0; v < 4; ++v) omap[w][4u * lane + v] = make_uint4(0u, 0u, 0u, 0u);
            wave_sync();
            if (n > 0 && incl > w0 && excl < w0 + wn)
                reinterpret_cast<uint32_t *>(&omap[w][0])[excl > w0 ? excl - w0 : 0u] = lane + 1u;
            wave_sync();
            uint32_t m[16];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint4 x = omap[w][4u * lane + v];
                m[4 * v] = x.x;
                m[4 * v + 1] = x.y;
                m[4 * v + 2] = x.z;
                m[4 * v + 3] = x.w;
            }
#pragma unroll
            for (int i = 1; i < 16; ++i) m[i] = max(m[i], m[i - 1]);
            uint32_t run = m[15];
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up(run, d, 64);
                if (lane >= (uint32_t)d) run = max(run, y);
            }
            uint32_t carry = (uint32_t)__shfl_up(run, 1, 64);
            if (lane == 0) carry = 0u;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                omap[w][4u * lane + v] = make_uint4(max(m[4 * v], carry), max(m[4 * v + 1], carry),
                                                    max(m[4 * v + 2], carry), max(m[4 * v + 3], carry));
            wave_sync();
            for (uint32_t base = 0; base < wn; base += 64u * U) {
                SlotItem c[U];
                uint4 q[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t g = base + 64u * u + lane;
                    const bool act = g < wn;
                    const uint32_t o = act ? own[g] - 1u : 0u;
                    c[u] = ctx[w][o];
                    q[u] = make_uint4(0u, 0u, 0u, 0u);
                    if (act) q[u] = e.s.binrec[cand_pos(c[u].s, w0 + g - c[u].excl)];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    // every lane evaluates the cuts (no branch); only real items count
                    const bool valid = base + 64u * u + lane < wn;
                    uint32_t r = 0;
                    const bool pass = cand_pass(e, cfg, c[u].s, q[u], r);
                    ref |= valid ? r : 0u;
                    const bool ok = valid & pass;
                    const uint64_t bal = __ballot(ok);
                    if (lane == 0) pbits[w][(base >> 6) + u] = bal;
                }
            }
            wave_sync();
            // this lane's items in the window
            const uint32_t s0 = max(excl, w0), s1 = min(incl, w0 + wn);
            for (uint32_t x = s0; x < s1; x += 64u) {
                const uint32_t r = x - w0, sh = r & 63u, len = min(64u, s1 - x);
                const uint64_t lo = pbits[w][r >> 6], hi = pbits[w][(r >> 6) + 1u];
                uint64_t bits = sh ? ((lo >> sh) | (hi << (64u - sh))) : lo;
                if (len < 64u) bits &= (1ull << len) - 1ull;
                cnt += (uint32_t)__popcll(bits);
                const uint32_t k0 = x - excl;
                if (k0 < 256u) {
                    const uint32_t wi = k0 >> 6, s2 = k0 & 63u;
                    const uint64_t lo_part = bits << s2, hi_part = s2 ? (bits >> (64u - s2)) : 0ull;
                    msk |= (wi == 0u) ? lo_part : 0ull;
                    mx1 |= (wi == 1u) ? lo_part : ((wi == 0u) ? hi_part : 0ull);
                    mx2 |= (wi == 2u) ? lo_part : ((wi == 1u) ? hi_part : 0ull);
                    mx3 |= (wi == 3u) ? lo_part : ((wi == 2u) ? hi_part : 0ull);
                }
            }
            wave_sync();  // reads of omap and pbits done before the next window
        }
        if (j < n_slots) {
            e.s.slot_cnt[j] = cnt;
            e.s.slot_mask[j] = msk;
            if (n > 64u) {
                e.s.slot_mask_ext[3u * j] = mx1;
                e.s.slot_mask_ext[3u * j + 1u] = mx2;
                e.s.slot_mask_ext[3u * j + 2u] = mx3;
            }
            e.s.slot_pos[j] = make_uint4(s.a0, s.na, s.b0, n | (s.pk & 0xFFFF0000u));
        }
        const uint32_t sum = wave_sum(cnt);
        if (lane == 0 && sum) atomicAdd(&e.s.dbl_agg[j0 / kSlotTile], sum);
    }
    if (zero_state && zero_vec(e)) zero_chunks(e);
    wave_or_atomic(&e.cnt->ref_assert, ref);
}

// ------------------------------------------------------------------ K2: doublet ids
// Exclusive scan of the slot counts in slot order, in place: block per tile of
// kSlotTile slots (4 per thread), the tile's prefix summed from the preceding
// tiles' totals.  Slot j's first cell id lands in slot_cnt[j], the total in
// slot_cnt[n_slots].
__global__ void __launch_bounds__(256) k_doublet_scan(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    __shared__ uint32_t lsum[4];
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    const uint32_t n_slots = e.n_slots;
    const uint32_t n_tiles = (n_slots + kSlotTile - 1) / kSlotTile;
    const uint32_t cap = cfg.p.max_num_of_doublets;
    for (uint32_t T = b; T < n_tiles; T += grid.per_event) {
        uint32_t part = 0;
        for (uint32_t k = threadIdx.x; k < T; k += 256u) part += e.s.dbl_agg[k];
        const uint32_t j0 = T * kSlotTile + threadIdx.x * 4u;
        uint32_t v[4], mine = 0;
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            v[u] = (j0 + u < n_slots) ? e.s.slot_cnt[j0 + u] : 0u;
            mine += v[u];
        }
        const uint32_t pre = block_sum(part, lsum);
        uint32_t tot;
        uint32_t run = pre + block_excl_scan4(mine, lsum, tot);
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            if (j0 + u < n_slots) e.s.slot_cnt[j0 + u] = run;
            run += v[u];
        }
        if (T == n_tiles - 1 && threadIdx.x == 0) {
            const uint32_t all = pre + tot;
            e.s.slot_cnt[n_slots] = all;
            e.cnt->n_cells_found = all;
            e.cnt->n_cells = all < cap ? all : cap;
            if (all > cap) atomicOr(&e.cnt->overflow, PT_OVF_DOUBLETS);
        }
    }
}

// ------------------------------------------------------------------ K3: doublet fill
// A wavefront takes a tile of 64 slots by ticket; the passing candidates of the
// tile are spread over the lanes and each lane writes one cell: the (inner,
// outer) record, the pair, the outer hit's coordinates, cleared flags and
// neighbour count, and an entry in the outer hit's list.  Doublets beyond
// max_num_of_doublets are counted but not stored, like the serial loop.
//
// The cells of a tile share outer hits (neighbouring inner hits see the same
// phi window; about 2.5 cells per distinct outer hit in a chunk of 192 cells
// at PU50), so the list cursors are advanced once per distinct outer hit of a
// chunk: the chunk's outer hits are counted in an LDS hash table, one global
// atomic per table entry reserves that many list entries, and the cells take
// consecutive entries of the reservation.  That cuts the cursor atomics, and
// the entries of one outer hit are written by one instruction to adjacent
// addresses.
// U rounds of 64 cells per chunk (U = 1, 2 or 3, tuned); the table holds more
// than 64 * U distinct keys, so an insert always ends
constexpr uint32_t kFillHash = 256;
constexpr uint32_t kFillEmpty = 0xFFFFFFFFu;
static_assert((kFillHash & (kFillHash - 1u)) == 0u, "hash table size");

template <int U> struct FillHash {
    static_assert(kFillHash > 64u * U, "hash table size");
    uint32_t key[kFillHash], cnt[kFillHash], base[kFillHash];
    uint32_t pre[kFillHash];   // the slot's first entry in the chunk's entries grouped by slot
    uint2 ent[64u * U];        // those entries: (cell, slot | inner hit << 8)
};

// table slot of key k (inserted if absent); the table never fills, so the
// linear probe ends
template <class HT> __device__ __forceinline__ uint32_t fill_hash_slot(HT &H, uint32_t k) {
    uint32_t s = (k * 2654435761u) >> 24;  // 8 bits
    for (;;) {
        const uint32_t old = atomicCAS(&H.key[s], kFillEmpty, k);
        if (old == kFillEmpty || old == k) return s;
        s = (s + 1u) & (kFillHash - 1u);
    }
}

template <int U>
__global__ void __launch_bounds__(256) k_doublet_fill(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    __shared__ uint32_t lid[4][64];
    __shared__ unsigned long long lmsk[4][64];  // a slot's first 64 pass bits (the rest stay in slot_mask_ext)
    __shared__ uint4 lpos[4][64];
    __shared__ FillHash<U> hash_all[4];
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    __shared__ PairLds PT;
    load_pairs(PT, e, cfg);
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    FillHash<U> &H = hash_all[w];
#pragma unroll
    for (uint32_t v = 0; v < kFillHash / 64u; ++v) {
        H.key[64u * v + lane] = kFillEmpty;
        H.cnt[64u * v + lane] = 0u;
    }
    const uint32_t n_slots = e.n_slots;
    const uint32_t n_tiles = (n_slots + 63u) / 64u;
    const uint32_t cap = cfg.p.max_num_of_doublets;
    uint32_t ref = 0, ovf = 0, err = 0;
    {
        // the start-cell bits of connect and the tile sums of the ntuplet count
        // start from zero over the stored cells
        const uint32_t nc = e.cnt->n_cells;
        const uint32_t nw = (nc + 63u) / 64u, nt = (nc + kDfsTile - 1u) / kDfsTile;
        for (uint32_t k = b * blockDim.x + threadIdx.x; k < nw + nt; k += grid.per_event * blockDim.x) {
            if (k < nw)
                e.s.start_bits[k] = 0ull;
            else
                e.s.dfs_agg[k - nw] = 0ull;
        }
    }
    for (;;) {
        const uint32_t t = take_ticket(&e.s.misc[kMiscTkFill]);
        if (t >= n_tiles) break;
        const uint32_t j = t * 64u + lane;
        uint32_t id = 0, nm = 0;  // nm: passing candidates among the slot's first 256
        uint64_t msk = 0;
        uint4 sp = make_uint4(0u, 0u, 0u, 0u);
        if (j < n_slots) {
            id = e.s.slot_cnt[j];
            msk = e.s.slot_mask[j];
            sp = e.s.slot_pos[j];
            nm = (uint32_t)__popcll(msk);
            if ((sp.w & 0xFFFFu) > 64u)
                nm += (uint32_t)(__popcll(e.s.slot_mask_ext[3u * j]) + __popcll(e.s.slot_mask_ext[3u * j + 1u]) +
                                 __popcll(e.s.slot_mask_ext[3u * j + 2u]));
        }
        wave_sync();  // the previous tile's reads of the LDS arrays are done
        if (__shfl(id, 0, 64) >= cap) continue;  // nothing of this tile is stored
        lid[w][lane] = id;
        lmsk[w][lane] = msk;
        lpos[w][lane] = sp;
        wave_sync();
        // writes cell cid from its outer-hit record q, xy at histogram position; returns
        // the outer hit (0-based) and inner hit (1-based) for the list entry
        auto write_cell = [&](uint32_t js, int p, uint4 q, float2 xy, uint32_t cid, uint32_t &oh, uint32_t &i1) {
            i1 = ((p == 0) ? js : js - PT.slot_cum[p - 1]) + PT.ls_inner[p] + 1u;
            oh = (q.w >> 16) - 1u;
            e.s.cell_io[cid] = i1 | ((oh + 1u) << 16);
            e.s.cell_pair[cid] = (uint8_t)p;
            *reinterpret_cast<PT_G uint16_t *>(e.s.cell_nn + 2u * cid) = 0;  // neighbours and flags
            e.s.cell_oxyzr[cid] = make_float4(xy.x, xy.y, __uint_as_float(q.x), __uint_as_float(q.y));
        };
        auto list_add = [&](uint32_t oh, uint32_t i1, uint32_t cid, uint32_t r) {
            if (r < kOhcStride) {
                e.s.ohc_list[ohc_index(oh, r)] = make_uint2(cid, i1 - 1u);
            } else if (r == kOhcStride) {
                ovf |= PT_OVF_OUTER_HIT_CELLS;
                const uint32_t k = atomicAdd(&e.s.misc[kMiscOhcOvf], 1u);
                if (k < kOhcOvfCap)
                    e.s.ohc_ovf[k] = oh;
                else
                    err |= PT_INT_OVF_LIST;
            }
        };
        auto emit = [&](uint32_t js, int p, uint32_t pos, uint32_t cid) {
            uint32_t oh, i1;
            write_cell(js, p, e.s.binrec[pos], e.s.binxy[pos], cid, oh, i1);
            list_add(oh, i1, cid, atomicAdd(&e.s.ohc_cursor[oh], 1u));
        };

        const uint32_t fincl = wave_incl_scan(nm);
        const uint32_t fexcl = fincl - nm;
        const uint32_t ftotal = __shfl(fincl, 63, 64);
        const uint32_t jw = t * 64u;
        for (uint32_t base = 0; base < ftotal; base += 64u * U) {
            // a chunk of U rounds: loads; cell writes and the outer hits into the
            // table; one cursor reservation per table entry; list entries
            bool act[U];
            uint32_t cid[U], pos[U], js[U], oh[U], i1[U], hs[U];
            int pp[U];
            uint4 q[U];
            float2 xy[U];
            uint32_t gg[U];
            int ow[U];
#pragma unroll
            for (int u = 0; u < U; ++u) gg[u] = base + 64u * u + lane;
            owners_of(fincl, gg, ow);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t g = gg[u];
                const int o = ow[u];
                const uint32_t rk = g - (uint32_t)__shfl(fexcl, o, 64);
                cid[u] = lid[w][o] + rk;
                act[u] = g < ftotal && cid[u] < cap;
                const uint4 so = lpos[w][o];
                pp[u] = (int)(so.w >> 16);
                js[u] = jw + (uint32_t)o;
                pos[u] = 0u;
                if (act[u]) {
                    const uint64_t m0 = lmsk[w][o];
                    const uint32_t c0 = (uint32_t)__popcll(m0);
                    uint32_t k;
                    if (rk < c0) {
                        k = nth_set_bit(m0, rk);
                    } else {  // past the first 64 candidates (PU200-like densities)
                        const PT_G uint64_t *x = e.s.slot_mask_ext + 3u * (size_t)js[u];
                        k = 64u + nth_set_bit3(x[0], x[1], x[2], rk - c0);
                    }
                    pos[u] = cand_pos(so.x, so.y, so.z, k);
                }
                if (act[u]) {
                    q[u] = e.s.binrec[pos[u]];
                    xy[u] = e.s.binxy[pos[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (act[u]) {
                    write_cell(js[u], pp[u], q[u], xy[u], cid[u], oh[u], i1[u]);
                    hs[u] = fill_hash_slot(H, oh[u]);
                    atomicAdd(&H.cnt[hs[u]], 1u);
                }
            wave_sync();
            // per table slot (lane l owns slots 4l..4l+3): the reservation, and the
            // slot's first position when the chunk's entries are grouped by slot
            static_assert(kFillHash == 256u, "four table slots per lane");
            const uint4 k4 = reinterpret_cast<const uint4 *>(H.key)[lane];
            const uint4 c4 = reinterpret_cast<const uint4 *>(H.cnt)[lane];
            const uint32_t kk[4] = {k4.x, k4.y, k4.z, k4.w}, cc[4] = {c4.x, c4.y, c4.z, c4.w};
            uint32_t bb[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (kk[v] != kFillEmpty) bb[v] = atomicAdd(&e.s.ohc_cursor[kk[v]], cc[v]);
            const uint32_t lsum = cc[0] + cc[1] + cc[2] + cc[3];
            const uint32_t lincl = wave_incl_scan(lsum);
            const uint32_t n_ent = (uint32_t)__shfl(lincl, 63, 64);
            const uint32_t p0 = lincl - lsum;
            reinterpret_cast<uint4 *>(H.base)[lane] = make_uint4(bb[0], bb[1], bb[2], bb[3]);
            reinterpret_cast<uint4 *>(H.pre)[lane] =
                make_uint4(p0, p0 + cc[0], p0 + cc[0] + cc[1], p0 + cc[0] + cc[1] + cc[2]);
            wave_sync();
            // each entry takes a rank in its slot (the decrements bring the counts back
            // to zero) and its place in the grouped order
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (act[u]) {
                    const uint32_t r = atomicSub(&H.cnt[hs[u]], 1u) - 1u;
                    H.ent[H.pre[hs[u]] + r] = make_uint2(cid[u], hs[u] | ((i1[u] - 1u) << 8));
                }
            wave_sync();
            // list entries in the grouped order: lanes with the same outer hit write
            // adjacent entries of its list
            for (uint32_t g0 = 0; g0 < n_ent; g0 += 64u) {
                const uint32_t g = g0 + lane;
                if (g < n_ent) {
                    const uint2 en = H.ent[g];
                    const uint32_t sl = en.y & 0xFFu;
                    list_add(H.key[sl], (en.y >> 8) + 1u, en.x, H.base[sl] + (g - H.pre[sl]));
                }
            }
            wave_sync();
#pragma unroll
            for (uint32_t v = 0; v < kFillHash / 64u; ++v) H.key[64u * v + lane] = kFillEmpty;
            wave_sync();
        }
        if (j < n_slots && (sp.w & 0xFFFFu) > 256u) {  // candidates beyond a slot's first 256: serial, rare
            SlotCtx s;
            const uint32_t n = slot_setup(e, cfg, PT, j, ref, s);
            uint32_t cid = id + nm;
            for (uint32_t k = 256; k < n && cid < cap; ++k) {
                const uint32_t pos = cand_pos(s, k);
                if (cand_pass(e, cfg, s, e.s.binrec[pos], ref)) emit(j, slot_pair(s), pos, cid++);
            }
        }
        wave_sync();
    }
    (void)ref;
    wave_or_atomic(&e.cnt->overflow, ovf);
    wave_or_atomic(&e.cnt->internal_error, err);
}

// Outer-hit lists that overflowed kOhcStride: rebuild them as the first 128
// cells in id order (the entries push! keeps).  Cell ids run pair by pair, so
// only the id ranges of the layer pairs that end on the hit's layer are
// scanned, in pair order, 2048 cells a round: each thread loads eight cells
// 256 apart, and a match's rank in id order is the matches of the earlier
// sub-rounds, of the earlier wavefronts in its sub-round (per-wavefront
// ballot counts through LDS) and of the lanes below it.  One block per event;
// exits at once when no list overflowed.
__global__ void __launch_bounds__(256) k_ohc_rebuild(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    constexpr uint32_t U = 8;
    __shared__ uint32_t wcnt[U][4];
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    const uint32_t n_ovf = e.s.misc[kMiscOhcOvf];
    if (n_ovf == 0u) return;  // block-uniform
    __shared__ PairLds PT;
    load_pairs(PT, e, cfg);
    const uint32_t w = threadIdx.x >> 6;
    uint32_t err = 0;
    if (n_ovf > kOhcOvfCap) err |= PT_INT_OVF_LIST;
    const uint32_t n = min(n_ovf, kOhcOvfCap);
    const uint32_t n_cells = e.cnt->n_cells, n_slots = e.n_slots;
    for (uint32_t q = 0; q < n; ++q) {
        const uint32_t h = e.s.ohc_ovf[q];
        const int L = layer_of(e, h);
        uint32_t r = 0;
        for (int p = 0; p < 19 && r < kOhcStride; ++p) {
            if (PT.outer[p] != L) continue;
            const uint32_t s0 = (p == 0) ? 0u : PT.slot_cum[p - 1], s1 = PT.slot_cum[p];
            if (s1 == 0xFFFFFFFFu || s0 >= n_slots) break;  // pairs not in use
            const uint32_t c0 = min(e.s.slot_cnt[s0], n_cells), c1 = min(e.s.slot_cnt[min(s1, n_slots)], n_cells);
            for (uint32_t base = c0; base < c1 && r < kOhcStride; base += 256u * U) {
                uint32_t io[U];
                uint64_t bal[U];
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t c = base + 256u * u + threadIdx.x;
                    io[u] = (c < c1) ? e.s.cell_io[c] : 0u;
                }
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t c = base + 256u * u + threadIdx.x;
                    bal[u] = __ballot(c < c1 && (io[u] >> 16) == h + 1u);
                    if (lane_id() == 0) wcnt[u][w] = (uint32_t)__popcll(bal[u]);
                }
                __syncthreads();
                uint32_t run = r;  // matches before sub-round u
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t n0 = wcnt[u][0], n1 = wcnt[u][1], n2 = wcnt[u][2], n3 = wcnt[u][3];
                    const uint32_t before = (w > 0 ? n0 : 0u) + (w > 1 ? n1 : 0u) + (w > 2 ? n2 : 0u);
                    const uint32_t rank = run + before + (uint32_t)__popcll(bal[u] & lanes_below());
                    if (((bal[u] >> lane_id()) & 1u) && rank < kOhcStride)
                        e.s.ohc_list[ohc_index(h, rank)] =
                            make_uint2(base + 256u * u + threadIdx.x, (io[u] & 0xFFFFu) - 1u);
                    run += n0 + n1 + n2 + n3;
                }
                r = run;
                __syncthreads();  // wcnt is rewritten by the next round
            }
        }
    }
    if (err && threadIdx.x == 0) atomicOr(&e.cnt->internal_error, err);
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
// The same test without the square root where that is safe: with every
// quantity >= 0, t * (pt_min * sqrt(d13)) <= rhs squares to
// (t * pt_min)^2 * d13 <= rhs^2.  Each side of the float expression above is
// within (1 +- 2^-24)^3 of its real value and the squares here within
// (1 +- 2^-24)^4 of theirs, so a squared comparison that clears a relative
// margin of 16 * 2^-24 (rhs^2 normal) decides it exactly as the float
// expression would.  Returns 1 (aligned), -1 (not aligned) or 0: inside the
// margin, or out of range, where are_aligned decides.
__device__ __forceinline__ int aligned_sq(float r1, float z1, float ri, float zi, float ro, float zo, float pt_min,
                                          float theta_cut) {
    const float radius_diff = fabsf(r1 - ro);
    const float d13 = radius_diff * radius_diff + (z1 - zo) * (z1 - zo);
    const float t = fabsf(z1 * (ri - ro) + zi * (ro - r1) + zo * (r1 - ri));
    const float rhs = theta_cut * d13 * radius_diff;
    const float a = t * pt_min;
    const float lhs2 = (a * a) * d13, rhs2 = rhs * rhs;
    const bool normal = rhs2 >= 0x1p-100f && rhs2 <= 0x1p100f;
    if (normal && lhs2 * (1.0f + 0x1p-20f) < rhs2) return 1;
    if (normal && lhs2 * (1.0f - 0x1p-20f) > rhs2) return -1;
    return 0;
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

struct ConnCtx {
    float x1, y1, z1, r1, x2, y2, zi, ri, region, theta;
};

__device__ __forceinline__ bool conn_pass(const ConnCtx &c, const Config &cfg, float4 h) {
    return are_aligned(c.r1, c.z1, c.ri, c.zi, h.w, h.z, cfg.p.pt_min, c.theta) &&
           dca_cut(c.x1, c.y1, c.x2, c.y2, h.x, h.y, c.region, cfg.p.hard_curv_cut);
}

__device__ __forceinline__ float theta_cut_of(const EvDesc &e, const Config &cfg, uint32_t h0) {
    return ((int16_t)(e.s.hdbl[h0].z >> 16) < 1184) ? cfg.p.ca_theta_cut_barrel : cfg.p.ca_theta_cut_forward;
}
__device__ __forceinline__ float region_of(const Config &cfg, uint32_t det_inner) {
    return ((int16_t)det_inner < 96) ? cfg.p.dca_cut_inner_triplet : cfg.p.dca_cut_outer_triplet;
}

// candidate cells of the pair by_inner[L][q] starting on hit h (lpos = h's index in L)
__device__ __forceinline__ void pair_range(const EvDesc &e, const PairLds &T, int L, uint32_t lpos, int q,
                                           uint32_t n_cells, uint32_t &a, uint32_t &b) {
    const int p = T.by_inner[L][q];
    const uint32_t j = ((p == 0) ? 0u : T.slot_cum[p - 1]) + lpos;
    a = e.s.slot_cnt[j];
    b = e.s.slot_cnt[j + 1];
    if (b > n_cells) b = n_cells;
    if (a > b) a = b;
}

// ntuplet start cells: pair < 8 || pair > 12 (pair < 3 when min_hits > 3)
// (ca_hit_ntuplet_generator_kernels_impl.jl:116-132)
__device__ __forceinline__ bool start_pair(const Config &cfg, int p) {
    return (cfg.p.min_hits_per_ntuplet > 3) ? (p < 3) : (p < 8 || p > 12);
}

// Per hit h, the rows are the cells ending on h that kernel_connect visits (the
// first 128 of h's list) and the columns the cells starting on h, in id order
// (one contiguous id range per layer pair).  A wavefront takes H hits
// per ticket and sorts them by column count, descending (so that rows of
// similar length share a round).  Consecutive sorted hits form groups whose
// columns fit the LDS staging area; a hit with more columns is a group of its
// own and is staged in chunks.  A group's columns - the outer hit and its
// coordinates - are loaded once into LDS, in id order per hit (a column's cell
// id is its hit's range offset plus its index, recomputed when it passes);
// then every lane takes one row and walks its hit's staged columns kConnCols
// at a time, keeping the passing ones in order - the serial push! order of the
// row cell's neighbour list.  The alignment cut runs in every lane; the dca
// cut (about 2% of the pairs get there) runs in place when no lane has more
// than one survivor in the round, else on the survivors of the whole round,
// compacted through LDS.  Each mark of a used cell is counted per
// outer hit for the fishbone's work list.
// H hits per ticket (4, 8, 16, 24 or 32) and ST columns staged at a time (128,
// 192, 256 or 384) are tuned: denser events favour fewer hits per ticket.
constexpr int kConnCols = 4;

template <uint32_t H, uint32_t ST> struct ConnLds {
    static_assert(H <= 64u && ST % 64u == 0u, "one hit per lane, whole staging rounds");
    float hx[H], hy[H], hz[H], hr[H], theta[H];
    uint32_t off[6][H], cum[6][H];  // hit lane's column c of range q: cell off[q] + c
    uint32_t coh[ST];               // staged columns: outer hit, its (z, r) and (x, y)
    float2 czr[ST], cxy[ST];
    // the row batch's rows (per lane), and the pairs waiting for the dca cut:
    // row lane | staged column << 8
    float rx1[64], ry1[64], rreg[64];
    uint32_t rhl[64], q[64];
};

template <uint32_t kConnHits, uint32_t kConnStage>
__global__ void __launch_bounds__(256, 5) k_connect(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    __shared__ ConnLds<kConnHits, kConnStage> lds_all[4];
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    __shared__ PairLds PT;
    load_pairs(PT, e, cfg);
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    ConnLds<kConnHits, kConnStage> &S = lds_all[w];
    const uint32_t n_cells = e.cnt->n_cells, n_hits = e.n_hits;
    uint32_t n_eval = 0, n_conn = 0, n_lists = 0, ovf = 0;
    for (;;) {
        const uint32_t h0 = take_ticket(&e.s.misc[kMiscTkConn]) * kConnHits;
        if (h0 >= n_hits) break;
        wave_sync();  // the previous tile's reads of S are done
        const uint32_t h = h0 + lane;
        uint32_t R = 0, C = 0;
        if (lane < kConnHits && h < n_hits) {
            R = min(e.s.ohc_cursor[h], kOhcStride);
            if (R) {
                const int L = layer_of(e, h);
                const int np = PT.n_by_inner[L];
                const uint32_t lpos = h - e.layer_start[L];
                uint32_t ra[6], rb[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    ra[q] = rb[q] = 0;
                    if (q < np) pair_range(e, PT, L, lpos, q, n_cells, ra[q], rb[q]);
                }
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    S.off[q][lane] = ra[q] - C;
                    C += rb[q] - ra[q];
                    S.cum[q][lane] = C;
                }
                if (C) {
                    const float4 p = e.s.hxyzr[h];
                    S.hx[lane] = p.x;
                    S.hy[lane] = p.y;
                    S.hz[lane] = p.z;
                    S.hr[lane] = p.w;
                    S.theta[lane] = theta_cut_of(e, cfg, h);
                }
            }
            if (!C) R = 0;
        }
        // sort the hits by column count, descending (bitonic over the wavefront)
        uint32_t key = R ? ((C << 8) | (63u - lane)) : 0u;  // ties: lower lane first
#pragma unroll
        for (uint32_t k = 2; k <= 64; k <<= 1) {
#pragma unroll
            for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
                const uint32_t other = (uint32_t)__shfl_xor(key, (int)jj, 64);
                const bool up = ((lane & k) == 0);  // descending blocks where (lane & k) == 0
                const bool lower = (lane & jj) == 0;
                const bool take_max = (up == lower);
                key = take_max ? max(key, other) : min(key, other);
            }
        }
        const uint32_t hl = 63u - (key & 0xFFu);  // hit lane of the i-th sorted hit
        const uint32_t Rsrc = (uint32_t)__shfl(R, (int)hl, 64);
        const uint32_t Rs = key ? Rsrc : 0u;
        const uint32_t Cs = key >> 8;  // columns of the i-th sorted hit (0: no rows)
        const uint32_t cincl = wave_incl_scan(Cs);
        wave_sync();
        for (uint32_t first = 0; first < 64u;) {
            if ((uint32_t)__shfl(Cs, (int)first, 64) == 0u) break;  // the rest have no rows
            // the group: sorted hits first..last whose columns fit the staging area
            const uint32_t cbefore = first ? (uint32_t)__shfl(cincl, (int)first - 1, 64) : 0u;
            const uint64_t fits = __ballot(lane >= first && Cs != 0u && cincl - cbefore <= kConnStage);
            const uint32_t last = fits ? 63u - (uint32_t)__clzll(fits) : first;
            const uint32_t T = (uint32_t)__shfl(cincl, (int)last, 64) - cbefore;
            const bool in_group = lane >= first && lane <= last;
            const uint32_t gincl = in_group ? cincl - cbefore : (lane < first ? 0u : T);
            const uint32_t cb = gincl - (in_group ? Cs : 0u);  // group column of the hit's first column
            // stage group columns [s0, s0 + n) at slots 0..n-1
            auto stage = [&](uint32_t s0, uint32_t n) {
                wave_sync();  // earlier reads of the staging area are done
                for (uint32_t g0 = 0; g0 < n; g0 += 64u) {
                    const uint32_t g = g0 + lane, gc = s0 + g;
                    const int oi = owner_of(gincl, gc);
                    const uint32_t ohl = (uint32_t)__shfl(hl, oi, 64);
                    const uint32_t c = gc - (uint32_t)__shfl(cb, oi, 64);  // column of the hit
                    if (g < n) {
                        uint32_t base = S.off[0][ohl];
#pragma unroll
                        for (int q = 0; q < 5; ++q) base = (c >= S.cum[q][ohl]) ? S.off[q + 1][ohl] : base;
                        const uint32_t t = base + c;
                        S.coh[g] = (e.s.cell_io[t] >> 16) - 1u;
                        const float4 o = e.s.cell_oxyzr[t];
                        S.czr[g] = make_float2(o.z, o.w);
                        S.cxy[g] = make_float2(o.x, o.y);
                    }
                }
                wave_sync();
            };
            const bool one_chunk = T <= kConnStage;
            if (one_chunk) stage(0u, T);
            const uint32_t rv = in_group ? Rs : 0u;
            const uint32_t rincl = wave_incl_scan(rv);
            const uint32_t rows = __shfl(rincl, 63, 64);
            for (uint32_t r0 = 0; r0 < rows; r0 += 64u) {
                const uint32_t rr = r0 + lane;
                const bool act = rr < rows;
                const int oi = owner_of(rincl, rr);
                const uint32_t ohl = (uint32_t)__shfl(hl, oi, 64);
                const uint32_t rbase = (uint32_t)__shfl(rincl, oi, 64) - (uint32_t)__shfl(rv, oi, 64);
                // (cross-lane reads outside conditionals: a lane outside the batch still
                // supplies its value)
                const uint32_t cols_o = (uint32_t)__shfl(Cs, oi, 64);
                const uint32_t cfirst = (uint32_t)__shfl(cb, oi, 64);
                const uint32_t cols = act ? cols_o : 0u;
                uint32_t cell = 0;
                float x1 = 0.f, y1 = 0.f, z1 = 0.f, r1 = 0.f, reg = 0.f, zi = 0.f, ri = 0.f, th = 0.f;
                if (act) {
                    const uint2 en = e.s.ohc_list[ohc_index(h0 + ohl, rr - rbase)];
                    cell = en.x;
                    const float4 p1 = e.s.hxyzr[en.y];
                    x1 = p1.x, y1 = p1.y, z1 = p1.z, r1 = p1.w;
                    reg = region_of(cfg, e.s.hdbl[en.y].z >> 16);
                    zi = S.hz[ohl], ri = S.hr[ohl], th = S.theta[ohl];
                    n_eval += cols;
                }
                S.rx1[lane] = x1;
                S.ry1[lane] = y1;
                S.rreg[lane] = reg;
                S.rhl[lane] = ohl;
                uint32_t cnt = 0;
                uint4 nb = make_uint4(~0u, ~0u, ~0u, ~0u);
                for (uint32_t cc0 = 0; cc0 < T; cc0 += kConnStage) {
                    if (!one_chunk) stage(cc0, min(T - cc0, kConnStage));
                    // this row's columns in the chunk, as staging slots [lo, lo + span)
                    const uint32_t a = max(cfirst, cc0), z = min(cfirst + cols, cc0 + kConnStage);
                    const uint32_t span = (act && z > a) ? z - a : 0u;
                    const uint32_t lo = a - cc0;
                    const uint32_t smax = wave_max(span);
                    for (uint32_t j0 = 0; j0 < smax; j0 += kConnCols) {
                        uint32_t sl[kConnCols];
                        float2 o3[kConnCols];  // (z, r) of the column's outer hit
                        bool al[kConnCols];
                        uint32_t nq = 0, qi[kConnCols];
                        // the round's reads first, then the cuts, then the ballots
#pragma unroll
                        for (int k = 0; k < kConnCols; ++k) {
                            sl[k] = (j0 + (uint32_t)k < span) ? lo + j0 + (uint32_t)k : 0u;
                            o3[k] = S.czr[sl[k]];
                        }
                        // the alignment cut squared; the rare pairs it leaves open (a
                        // wavefront-uniform branch) take the reference's expression
                        uint32_t open = 0;
#pragma unroll
                        for (int k = 0; k < kConnCols; ++k) {
                            const int d = aligned_sq(r1, z1, ri, zi, o3[k].y, o3[k].x, cfg.p.pt_min, th);
                            const bool in = j0 + (uint32_t)k < span;
                            al[k] = in & (d > 0);
                            open |= (in && d == 0) ? 1u << k : 0u;
                        }
                        if (__ballot(open != 0u)) {
#pragma unroll
                            for (int k = 0; k < kConnCols; ++k)
                                if ((open >> k) & 1u)
                                    al[k] = are_aligned(r1, z1, ri, zi, o3[k].y, o3[k].x, cfg.p.pt_min, th);
                        }
                        uint32_t am4 = 0;  // this lane's aligned columns of the round, bit k
#pragma unroll
                        for (int k = 0; k < kConnCols; ++k) am4 |= al[k] ? 1u << k : 0u;
                        uint32_t pm = 0;  // this lane's passing columns of the round, bit k
                        const bool many = __ballot((am4 & (am4 - 1u)) != 0u) != 0ull;
                        if (!many && __ballot(am4 != 0u)) {
                            // at most one aligned column per lane (the common case): the dca
                            // cut in place, on the same values the compacted path reads
                            if (am4) {
                                const int k1 = __builtin_ctz(am4);
                                uint32_t s1 = sl[0];
#pragma unroll
                                for (int k = 1; k < kConnCols; ++k) s1 = (k1 == k) ? sl[k] : s1;
                                const float2 o = S.cxy[s1];
                                if (dca_cut(x1, y1, S.hx[ohl], S.hy[ohl], o.x, o.y, reg, cfg.p.hard_curv_cut))
                                    pm = am4;
                            }
                        }
                        if (many) {
#pragma unroll
                            for (int k = 0; k < kConnCols; ++k) {
                                const uint64_t am = __ballot(al[k]);
                                qi[k] = nq + (uint32_t)__popcll(am & lanes_below());
                                nq += (uint32_t)__popcll(am);
                            }
                        }
                        // otherwise the survivors in (k, lane) order, the dca cut 64 at a time
                        for (uint32_t qb = 0; qb < nq; qb += 64u) {
#pragma unroll
                            for (int k = 0; k < kConnCols; ++k)
                                if (al[k] && qi[k] >= qb && qi[k] < qb + 64u) S.q[qi[k] - qb] = lane | (sl[k] << 8);
                            wave_sync();
                            bool ok = false;
                            if (qb + lane < nq) {
                                const uint32_t v = S.q[lane], rl = v & 63u;
                                const uint32_t hh = S.rhl[rl];
                                const float2 o = S.cxy[v >> 8];
                                ok = dca_cut(S.rx1[rl], S.ry1[rl], S.hx[hh], S.hy[hh], o.x, o.y, S.rreg[rl],
                                             cfg.p.hard_curv_cut);
                            }
                            const uint64_t okm = __ballot(ok);
#pragma unroll
                            for (int k = 0; k < kConnCols; ++k)
                                if (al[k] && qi[k] >= qb && qi[k] < qb + 64u && ((okm >> (qi[k] - qb)) & 1ull))
                                    pm |= 1u << k;
                            wave_sync();
                        }
#pragma unroll
                        for (int k = 0; k < kConnCols; ++k)
                            if ((pm >> k) & 1u) {
                                // the column's cell id again from the hit's ranges
                                const uint32_t c = sl[k] + cc0 - cfirst;
                                uint32_t base = S.off[0][ohl];
#pragma unroll
                                for (int q = 0; q < 5; ++q) base = (c >= S.cum[q][ohl]) ? S.off[q + 1][ohl] : base;
                                const uint32_t t = base + c;
                                e.s.cell_flags[2u * t] = kUsed;
                                // marks per outer hit (a cell can be counted more than once):
                                // the fishbone only looks at hits with two or more
                                atomicAdd(&e.s.fish_cnt[S.coh[sl[k]]], 1u);
                                e.s.trk[t] = make_uint2(0u, 0u);
                                if (cnt == 0) nb.x = t;
                                if (cnt == 1) nb.y = t;
                                if (cnt == 2) nb.z = t;
                                if (cnt == 3) nb.w = t;
                                ++cnt;
                            }
                    }
                }
                bool st = false;
                if (cnt) {
                    const uint32_t nn = cnt < PT_MAX_NEIGHBORS ? cnt : PT_MAX_NEIGHBORS;
                    e.s.cell_flags[2u * cell] = kUsed;
                    atomicAdd(&e.s.fish_cnt[h0 + ohl], 1u);
                    e.s.trk[cell] = make_uint2(0u, 0u);
                    e.s.cell_nn[2u * cell] = (uint8_t)nn;
                    e.s.nbr4[cell] = nb;
                    n_conn += nn;
                    ++n_lists;
                    if (cnt > PT_MAX_NEIGHBORS) ovf |= PT_OVF_NEIGHBORS;
                    if (nn > 4u) e.s.ext_list[atomicAdd(&e.s.misc[kMiscExt], 1u)] = cell;
                    st = start_pair(cfg, e.s.cell_pair[cell]);
                }
                const uint64_t sb = __ballot(st);
                if (sb) {
                    uint32_t sbase = 0;
                    if (lane == 0) sbase = atomicAdd(&e.s.misc[kMiscStart], (uint32_t)__popcll(sb));
                    sbase = __shfl(sbase, 0, 64);
                    if (st) {
                        e.s.start_list[sbase + (uint32_t)__popcll(sb & lanes_below())] = cell;
                        atomicOr(reinterpret_cast<unsigned long long *>(&e.s.start_bits[cell >> 6]),
                                 1ull << (cell & 63u));
                    }
                }
                wave_sync();  // this batch's reads of the row arrays are done
            }
            first = last + 1u;
        }
    }
    wave_add_atomic(&e.cnt->n_connect_pairs, n_eval);
    wave_add_atomic(&e.cnt->n_connections, n_conn);
    wave_add_atomic(&e.s.misc[kMiscNbrLists], n_lists);
    wave_or_atomic(&e.cnt->overflow, ovf);
}

// Cells with more than 4 neighbours (rare): walk the candidates again in id
// order and put neighbours 5..64 in the per-event pool.  Items i0, i0+stride, ...
__device__ __forceinline__ void connect_ext_items(const EvDesc &e, const Config &cfg, const PairLds &T, uint32_t i0,
                                                  uint32_t stride,
                                                  uint32_t &err) {
    const uint32_t n_ext = e.s.misc[kMiscExt];
    for (uint32_t i = i0; i < n_ext; i += stride) {
        const uint32_t o = e.s.ext_list[i];
        const uint32_t nkeep = e.s.cell_nn[2u * o];
        const uint32_t base = atomicAdd(&e.s.misc[kMiscPool], nkeep - 4u);
        if (base + (nkeep - 4u) > kNbrPool) {
            err |= PT_INT_NBR_POOL;
            e.s.cell_nn[2u * o] = 4;
            continue;
        }
        e.s.nbr_ext[o] = base;
        const uint32_t n_cells = e.cnt->n_cells;
        const uint32_t io = e.s.cell_io[o];
        const uint32_t h0 = (io >> 16) - 1u, i1 = (io & 0xFFFFu) - 1u;
        const int L = layer_of(e, h0);
        const int np = T.n_by_inner[L];
        const uint32_t lpos = h0 - e.layer_start[L];
        const float4 p1 = e.s.hxyzr[i1], p2 = e.s.hxyzr[h0];
        const ConnCtx c{p1.x, p1.y, p1.z, p1.w, p2.x, p2.y, p2.z, p2.w, region_of(cfg, e.s.hdbl[i1].z >> 16),
                        theta_cut_of(e, cfg, h0)};
        uint32_t k = 0;
        for (int q = 0; q < np && k < nkeep; ++q) {
            uint32_t a, bq;
            pair_range(e, T, L, lpos, q, n_cells, a, bq);
            for (uint32_t t = a; t < bq && k < nkeep; ++t) {
                if (!conn_pass(c, cfg, e.s.cell_oxyzr[t])) continue;
                if (k >= 4u) e.s.nbr_pool[base + k - 4u] = t;
                ++k;
            }
        }
    }
}

// Used when the fishbone (which otherwise does this work first) is off.
__global__ void __launch_bounds__(256) k_connect_ext(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    uint32_t err = 0;
    __shared__ PairLds PT;
    load_pairs(PT, e, cfg);
    connect_ext_items(e, cfg, PT, b * blockDim.x + threadIdx.x, grid.per_event * blockDim.x, err);
    wave_or_atomic(&e.cnt->internal_error, err);
}

// The fishbone's work list: the hits with two or more marks from connect (a
// hit with fewer has at most one used cell ending on it, so nothing to
// compare).  The fishbone decides every hit on its own (a cell is on one
// hit's list), so the list is in no particular order: several blocks per
// event, one hit per thread per round, one atomic per wavefront to reserve
// its entries.
__global__ void __launch_bounds__(256) k_fish_list(const EvDesc *__restrict__ evs, EvGrid grid) {
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    const uint32_t n = e.n_hits;
    for (uint32_t h0 = b * blockDim.x; h0 < n; h0 += grid.per_event * blockDim.x) {
        const uint32_t h = h0 + threadIdx.x;
        const bool take = h < n && e.s.fish_cnt[h] >= 2u;
        const uint64_t m = __ballot(take);
        uint32_t pos = 0;
        if (lane_id() == 0 && m) pos = atomicAdd(&e.s.misc[kMiscFish], (uint32_t)__popcll(m));
        pos = __shfl(pos, 0, 64) + (uint32_t)__popcll(m & lanes_below());
        if (take) e.s.fish_hits[pos] = h;
    }
}

// ------------------------------------------------------------------ fishbone (gpu_fishbone.jl:4-62)
// The serial loop, per hit, over the ascending cell list L (first 128):
//   for ic, for jc > ic, both used: if collinear and on different modules:
//     if |ic| > |jc| (inner-hit distance): kill ic and stop; else kill jc.
// Kills never feed back into the loop, so it is evaluated per cell pair: with
// first(ic) = smallest qualifying jc with |jc| < |ic|, ic dies if first(ic)
// exists, and jc dies if some ic qualifies with it and jc < first(ic).
// Cells that connect did not use take no part (about a tenth of the cells are
// used at PU50), so only the used entries of each list are kept, in order.
// A wavefront takes kHitTile hits per ticket (8, 16 or 32, tuned), packs them into groups of at
// most kFishRows list entries, keeps the used ones (rows) in LDS and spreads
// every unordered row pair of a hit over the lanes; a second step decides
// each row.
constexpr uint32_t kFishRows = 128;

struct FishLds {
    uint32_t id[kFishRows], mod[kFishRows], first[kFishRows];
    uint32_t upre[kFishRows];  // used entries of the group before entry rr
    float dx[kFishRows], dy[kFishRows], dz[kFishRows], n[kFishRows];
    unsigned long long qt[2][kFishRows];  // rows (index in the hit) that qualify with this one as jc
};

// pair k of the b-major enumeration (0,1), (0,2), (1,2), (0,3), ...: (k - b(b-1)/2, b);
// the hardware square root only seeds the integer result, which is then corrected
__device__ __forceinline__ uint32_t tri_b(uint32_t k) {
    uint32_t bb = (uint32_t)((1.0f + __builtin_amdgcn_sqrtf(8.0f * (float)k + 1.0f)) * 0.5f);
    if (bb * (bb - 1u) / 2u > k) --bb;
    if ((bb + 1u) * bb / 2u <= k) ++bb;
    return bb;
}

template <uint32_t kHitTile>
__global__ void __launch_bounds__(256) k_fishbone(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    static_assert(kHitTile <= 64u, "one hit per lane");
    __shared__ FishLds lds_all[4];
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    __shared__ PairLds PT;
    load_pairs(PT, e, cfg);
    {  // the rare neighbour lists longer than 4 first (independent of the fishbone)
        uint32_t err = 0;
        connect_ext_items(e, cfg, PT, b * blockDim.x + threadIdx.x, grid.per_event * blockDim.x, err);
        wave_or_atomic(&e.cnt->internal_error, err);
    }
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    FishLds &S = lds_all[w];
    // the hits on which connect used at least two cells (each listed once)
    const uint32_t n_fh = min(e.s.misc[kMiscFish], e.n_hits);
    uint32_t killed = 0;
    for (;;) {
        const uint32_t f0 = take_ticket(&e.s.misc[kMiscTkFish]) * kHitTile;
        if (f0 >= n_fh) break;
        const bool mine = lane < kHitTile && f0 + lane < n_fh;
        const uint32_t h = mine ? e.s.fish_hits[f0 + lane] : 0u;
        const uint32_t cur = mine ? e.s.ohc_cursor[h] : 0u;
        const uint32_t R = (cur >= 2u) ? (cur < kOhcStride ? cur : kOhcStride) : 0u;
        for (uint32_t first = 0; first < kHitTile;) {
            const uint32_t v = (lane >= first) ? R : 0u;
            const uint32_t rincl = wave_incl_scan(v);
            const uint64_t fits = __ballot(lane >= first && lane < kHitTile && rincl <= kFishRows);
            const uint32_t last = 63u - (uint32_t)__clzll(fits);
            const uint32_t rows = __shfl(rincl, (int)last, 64);
            if (rows) {
                const uint32_t rexcl = rincl - v;
                wave_sync();  // the previous group's reads of S are done
                // the group's list entries; the used ones become rows, in order
                uint32_t n_used = 0;
#pragma unroll
                for (uint32_t u = 0; u < kFishRows / 64u; ++u) {
                    const uint32_t rr = 64u * u + lane;
                    const int o = owner_of(rincl, rr);
                    const uint32_t s0 = (uint32_t)__shfl(rexcl, o, 64);
                    const uint32_t ho = (uint32_t)__shfl(h, o, 64);  // the row's hit
                    uint2 en = make_uint2(0u, 0u);
                    bool used = false;
                    if (rr < rows) {
                        en = e.s.ohc_list[ohc_index(ho, rr - s0)];
                        used = (e.s.cell_flags[2u * en.x] & kUsed) != 0;
                    }
                    const uint64_t um = __ballot(used);
                    const uint32_t pos = n_used + (uint32_t)__popcll(um & lanes_below());
                    S.upre[rr] = pos;
                    if (used) {
                        const float4 po = e.s.hxyzr[ho];
                        const float4 pi = e.s.hxyzr[en.y];
                        const float dx = pi.x - po.x, dy = pi.y - po.y, dz = pi.z - po.z;
                        S.id[pos] = en.x;
                        S.mod[pos] = e.s.hdbl[en.y].z >> 16;
                        S.dx[pos] = dx;
                        S.dy[pos] = dy;
                        S.dz[pos] = dz;
                        S.n[pos] = dx * dx + dy * dy + dz * dz;
                        S.first[pos] = ~0u;
                        S.qt[0][pos] = 0ull;
                        S.qt[1][pos] = 0ull;
                    }
                    n_used += (uint32_t)__popcll(um);
                }
                wave_sync();
                // per hit of the group: its first row and its number of rows
                uint32_t us0 = 0, uR = 0;
                if (lane >= first && lane <= last && v) {
                    us0 = S.upre[rexcl];
                    uR = (rincl < kFishRows ? S.upre[rincl] : n_used) - us0;
                }
                // every unordered pair of rows of the same hit
                const uint32_t items = uR * (uR - (uR ? 1u : 0u)) / 2u;
                const uint32_t iincl = wave_incl_scan(items);
                const uint32_t iexcl = iincl - items;
                const uint32_t total = __shfl(iincl, 63, 64);
                for (uint32_t base = 0; base < total; base += 64u) {
                    const uint32_t g = base + lane;
                    const int o = owner_of(iincl, g);
                    const uint32_t ox = (uint32_t)__shfl(iexcl, o, 64), s0 = (uint32_t)__shfl(us0, o, 64);
                    if (g < total) {
                        const uint32_t k = g - ox;
                        const uint32_t bb = tri_b(k), aa = k - bb * (bb - 1u) / 2u;
                        const uint32_t ra = s0 + aa, rb = s0 + bb;
                        const uint32_t ida = S.id[ra], idb = S.id[rb];
                        const bool a_ic = ida < idb;
                        const uint32_t ri = a_ic ? ra : rb, rj = a_ic ? rb : ra;  // ic, jc rows
                        if (S.mod[ri] != S.mod[rj]) {
                            const float ni = S.n[ri], nj = S.n[rj];
                            const float cos12 = S.dx[ri] * S.dx[rj] + S.dy[ri] * S.dy[rj] + S.dz[ri] * S.dz[rj];
                            if (cos12 * cos12 >= 0.99999f * ni * nj) {
                                const uint32_t il = ri - s0;
                                atomicOr(&S.qt[il >> 6][rj], 1ull << (il & 63u));
                                if (ni > nj) atomicMin(&S.first[ri], a_ic ? idb : ida);
                            }
                        }
                    }
                }
                wave_sync();
                // decide each row
                const uint32_t uincl = wave_incl_scan(uR);
                for (uint32_t r0 = 0; r0 < n_used; r0 += 64u) {
                    const uint32_t rr = r0 + lane;
                    const int o = owner_of(uincl, rr);
                    const uint32_t s0 = (uint32_t)__shfl(us0, o, 64);
                    if (rr < n_used) {
                        const uint32_t myid = S.id[rr];
                        bool kill = S.first[rr] != ~0u;
                        // the rows ic that qualify with this row as jc (rare): killed if jc < first(ic)
#pragma unroll
                        for (int hw = 0; hw < 2; ++hw) {
                            unsigned long long m = S.qt[hw][rr];
                            while (m && !kill) {
                                const uint32_t x = s0 + 64u * hw + (uint32_t)__builtin_ctzll(m);
                                m &= m - 1ull;
                                if (myid < S.first[x]) kill = true;
                            }
                        }
                        if (kill) {
                            e.s.cell_flags[2u * myid] = kUsed | kKilled;
                            ++killed;
                        }
                    }
                }
            }
            first = last + 1u;
        }
    }
    wave_add_atomic(&e.cnt->n_killed, killed);
}

// ------------------------------------------------------------------ neighbour-list columns
// The reference keeps neighbour lists in CellNeighborsVector(64, 65536),
// column 1 the empty sentinel, and hands a column to a cell at its first
// connection (add_outer_neighbor, gpu_ca_cell.jl:164-177).  extend! returns
// the last column again once all are taken (simple_matrix.jl:13-19): the
// 65,535th cell to get a list and every later one share column 65,536, which
// holds the first 64 of their connections in connect's order.  Connect walks
// the `this` cells in id order and each outer-hit list in ascending cell
// order, so a cell o asks for its column at key (first this, o) and the
// connections (this, o) reach the shared column in (this, o) order.  Here the
// lists are kept per cell; when more than 65,535 cells have one, this kernel
// (one block per event, after the lists are complete) finds the cells past the
// 65,534th key and gives them all the shared list.
__device__ __forceinline__ uint32_t comp4_(uint4 v, uint32_t i) {
    return (i == 0) ? v.x : (i == 1) ? v.y : (i == 2) ? v.z : v.w;
}
constexpr uint32_t kOwnCols = PT_MAX_ACTIVE_LISTS - 2u;  // columns 2..65535 stay private
constexpr uint32_t kShareCand = 4096;                    // candidate connections sorted in LDS
constexpr uint32_t kShareHist = 8192;

__device__ __forceinline__ uint64_t owner_key(const EvDesc &e, uint32_t c) {
    return ((uint64_t)e.s.nbr4[c].x << 19) | c;
}
__device__ __forceinline__ uint32_t nbr_of(const EvDesc &e, uint32_t c, uint32_t i) {
    return (i < 4u) ? comp4_(e.s.nbr4[c], i) : e.s.nbr_pool[e.s.nbr_ext[c] + i - 4u];
}

__global__ void __launch_bounds__(1024) k_nbr_share(const EvDesc *__restrict__ evs, EvGrid grid) {
    __shared__ union {
        uint32_t hist[kShareHist];
        uint64_t cand[kShareCand];
    } L;
    __shared__ uint32_t scan_lds[1024 / 64 + 1];
    __shared__ uint32_t n_cand, sum_nn;
    __shared__ uint64_t sel;
    uint32_t ev, blk;
    if (!ev_block(grid, ev, blk)) return;
    const EvDesc &e = evs[ev];
    if (e.s.misc[kMiscNbrLists] <= kOwnCols + 1u) return;  // every list keeps its own column
    const uint32_t n_cells = e.cnt->n_cells, tid = threadIdx.x;
    // 1. the key of rank kOwnCols among the owners' keys (38 bits), by radix
    //    selection over 13 + 13 + 12 bits
    uint64_t prefix = 0;
    uint32_t rank = kOwnCols;
    const int bits[3] = {12, 13, 13};
    int shift = 38;
    for (int pass = 0; pass < 3; ++pass) {
        const int nb = bits[pass];
        shift -= nb;
        for (uint32_t i = tid; i < kShareHist; i += blockDim.x) L.hist[i] = 0u;
        __syncthreads();
        for (uint32_t c = tid; c < n_cells; c += blockDim.x) {
            if (!e.s.cell_nn[2u * c]) continue;
            const uint64_t k = owner_key(e, c);
            if ((k >> (shift + nb)) == prefix) atomicAdd(&L.hist[(uint32_t)(k >> shift) & ((1u << nb) - 1u)], 1u);
        }
        __syncthreads();
        // the bucket holding `rank`: 8 bins per thread, block scan of the sums
        uint32_t loc = 0;
        for (uint32_t j = 0; j < 8u; ++j) loc += L.hist[8u * tid + j];
        uint32_t total;
        const uint32_t before = block_excl_scan(loc, scan_lds, total);
        if (rank >= before && rank < before + loc) {
            uint32_t run = before;
            for (uint32_t j = 0; j < 8u; ++j) {
                const uint32_t h = L.hist[8u * tid + j];
                if (rank < run + h) {
                    sel = ((uint64_t)(8u * tid + j));
                    n_cand = rank - run;
                    break;
                }
                run += h;
            }
        }
        __syncthreads();
        prefix = (prefix << nb) | sel;
        rank = n_cand;
        __syncthreads();
    }
    const uint64_t kstar = prefix;
    // 2. the shared set S = owners with key >= kstar; the 64th smallest first
    //    connection among S bounds the connections that can make the column
    if (tid == 0) {
        n_cand = 0u;
        sum_nn = 0u;
    }
    for (uint32_t i = tid; i < kShareHist; i += blockDim.x) L.hist[i] = 0u;
    __syncthreads();
    // histogram of the first connection's top 13 bits (of 19) over S
    uint32_t my_nn = 0;
    for (uint32_t c = tid; c < n_cells; c += blockDim.x) {
        const uint32_t nn = e.s.cell_nn[2u * c];
        if (!nn || owner_key(e, c) < kstar) continue;
        my_nn += nn;
        atomicAdd(&L.hist[e.s.nbr4[c].x >> 6], 1u);
    }
    atomicAdd(&sum_nn, my_nn);
    __syncthreads();
    uint32_t loc = 0;
    for (uint32_t j = 0; j < 8u; ++j) loc += L.hist[8u * tid + j];
    uint32_t total;
    const uint32_t before = block_excl_scan(loc, scan_lds, total);
    const uint32_t want = min(total, PT_MAX_NEIGHBORS) - 1u;  // rank of the 64th (or last) first connection
    if (want >= before && want < before + loc) {
        uint32_t run = before;
        for (uint32_t j = 0; j < 8u; ++j) {
            const uint32_t h = L.hist[8u * tid + j];
            if (want < run + h) {
                sel = ((uint64_t)(8u * tid + j + 1u)) << 6;  // connections below this `this` bound qualify
                break;
            }
            run += h;
        }
    }
    __syncthreads();
    // fewer than 64 cells share: all their connections can make the column
    const uint32_t tbound = (total < PT_MAX_NEIGHBORS) ? ~0u : (uint32_t)sel;
    __syncthreads();
    // 3. candidates: every connection (this, o) of S with this < tbound
    bool full = false;
    for (uint32_t c = tid; c < n_cells; c += blockDim.x) {
        const uint32_t nn = e.s.cell_nn[2u * c];
        if (!nn || owner_key(e, c) < kstar) continue;
        for (uint32_t i = 0; i < nn; ++i) {
            const uint32_t t = nbr_of(e, c, i);
            if (t >= tbound) break;  // ascending
            const uint32_t slot = atomicAdd(&n_cand, 1u);
            if (slot < kShareCand) L.cand[slot] = ((uint64_t)t << 19) | c;
            else full = true;
        }
    }
    __syncthreads();
    const uint32_t nc = min(n_cand, kShareCand);
    const bool overflow_cand = __syncthreads_or(full);
    if (overflow_cand) {  // more candidates than the sort holds: not reproduced
        if (tid == 0) atomicOr(&e.cnt->internal_error, PT_INT_LIST_SHARE);
        return;
    }
    // 4. sort the candidates (bitonic, padded with ~0) and keep the first 64
    uint32_t np = 1;
    while (np < nc) np <<= 1;
    for (uint32_t i = nc + tid; i < np; i += blockDim.x) L.cand[i] = ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= np; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < np; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = L.cand[i], b = L.cand[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        L.cand[i] = b;
                        L.cand[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    const uint32_t len = min(nc, PT_MAX_NEIGHBORS);
    // 5. the shared list: entries 5..len in the pool, the first four inline,
    //    the same for every cell of S
    __shared__ uint32_t pool_base;
    if (tid == 0) {
        pool_base = (len > 4u) ? atomicAdd(&e.s.misc[kMiscPool], len - 4u) : 0u;
        uint32_t ovf = 0, err = 0;
        if (len > 4u && pool_base + (len - 4u) > kNbrPool) err |= PT_INT_NBR_POOL;
        if (sum_nn > PT_MAX_NEIGHBORS) ovf |= PT_OVF_NEIGHBORS;  // pushes dropped from the shared column
        if (ovf) atomicOr(&e.cnt->overflow, ovf);
        if (err) atomicOr(&e.cnt->internal_error, err);
        // each list counted once: the shared column replaces the lists of S
        atomicSub(&e.cnt->n_connections, sum_nn - len);
    }
    __syncthreads();
    if (len > 4u && pool_base + (len - 4u) > kNbrPool) return;
    for (uint32_t i = 4u + tid; i < len; i += blockDim.x) e.s.nbr_pool[pool_base + i - 4u] = (uint32_t)(L.cand[i] >> 19);
    uint4 first4 = make_uint4(~0u, ~0u, ~0u, ~0u);
    if (len > 0) first4.x = (uint32_t)(L.cand[0] >> 19);
    if (len > 1) first4.y = (uint32_t)(L.cand[1] >> 19);
    if (len > 2) first4.z = (uint32_t)(L.cand[2] >> 19);
    if (len > 3) first4.w = (uint32_t)(L.cand[3] >> 19);
    for (uint32_t c = tid; c < n_cells; c += blockDim.x) {
        if (!e.s.cell_nn[2u * c] || owner_key(e, c) < kstar) continue;
        // (reads of nbr4 for the key above happen before this write for c only)
        e.s.nbr_ext[c] = pool_base;
        e.s.cell_nn[2u * c] = (uint8_t)len;
        e.s.nbr4[c] = first4;
    }
}

// ------------------------------------------------------------------ ntuplets (gpu_ca_cell.jl:232-270)
// DFS with the serial visiting order, one template level per depth (the
// reference asserts at most 4 doublets per chain, gpu_ca_cell.jl:234), so the
// path lives in registers.  At each level the data of all (up to 4 inline)
// neighbours is loaded at once - id, hits, neighbour count and list, killed
// flag - so a level costs one memory round trip.  leaf(cells, hits, n) is
// called at every emission point of the recursive reference.
struct DNode {
    uint32_t id, io, nn;
    uint4 nb;
    bool killed;
};

__device__ __forceinline__ DNode load_node(const EvDesc &e, uint32_t c) {
    DNode n;
    n.id = c;
    n.io = e.s.cell_io[c];
    const uint32_t fn = *reinterpret_cast<const PT_G uint16_t *>(e.s.cell_nn + 2u * c);  // one load for both
    n.nn = fn & 0xFFu;
    n.killed = ((fn >> 8) & kKilled) != 0;
    n.nb = e.s.nbr4[c];
    return n;
}
__device__ __forceinline__ uint32_t comp4(uint4 v, uint32_t i) {
    return (i == 0) ? v.x : (i == 1) ? v.y : (i == 2) ? v.z : v.w;
}
__device__ __forceinline__ DNode pick4(uint32_t i, const DNode &a, const DNode &b, const DNode &c, const DNode &d) {
    DNode r;
    r.id = (i == 0) ? a.id : (i == 1) ? b.id : (i == 2) ? c.id : d.id;
    r.io = (i == 0) ? a.io : (i == 1) ? b.io : (i == 2) ? c.io : d.io;
    r.nn = (i == 0) ? a.nn : (i == 1) ? b.nn : (i == 2) ? c.nn : d.nn;
    r.killed = (i == 0) ? a.killed : (i == 1) ? b.killed : (i == 2) ? c.killed : d.killed;
    r.nb.x = (i == 0) ? a.nb.x : (i == 1) ? b.nb.x : (i == 2) ? c.nb.x : d.nb.x;
    r.nb.y = (i == 0) ? a.nb.y : (i == 1) ? b.nb.y : (i == 2) ? c.nb.y : d.nb.y;
    r.nb.z = (i == 0) ? a.nb.z : (i == 1) ? b.nb.z : (i == 2) ? c.nb.z : d.nb.z;
    r.nb.w = (i == 0) ? a.nb.w : (i == 1) ? b.nb.w : (i == 2) ? c.nb.w : d.nb.w;
    return r;
}

template <int D, class F>
__device__ __forceinline__ void dfs_level(const EvDesc &e, const DNode &node, uint32_t (&cells)[4], uint32_t (&ios)[4],
                                          uint32_t min_doublets, uint32_t &ref_bits, F &leaf) {
    cells[D] = node.id;
    ios[D] = node.io;
    const uint32_t nn = node.nn;
    bool child = false;
    if (nn) {
        if constexpr (D + 1 < 4) {
            DNode c0 = {}, c1 = {}, c2 = {}, c3 = {};
            c0 = load_node(e, node.nb.x);
            if (nn > 1) c1 = load_node(e, node.nb.y);
            if (nn > 2) c2 = load_node(e, node.nb.z);
            if (nn > 3) c3 = load_node(e, node.nb.w);
            // one loop over all neighbours (5..64 from the pool, rare), so that each
            // depth is instantiated once and the kernel's code stays small
            const uint32_t ext = (nn > 4u) ? e.s.nbr_ext[node.id] : 0u;
            for (uint32_t i = 0; i < nn; ++i) {
                const DNode c = (i < 4u) ? pick4(i, c0, c1, c2, c3) : load_node(e, e.s.nbr_pool[ext + i - 4u]);
                if (c.killed) continue;
                child = true;
                // @assert(inner hit of the neighbour != own inner hit) (gpu_ca_cell.jl:247):
                // only a shared neighbour column can break it
                if (((c.io ^ node.io) & 0xFFFFu) == 0u) ref_bits |= PT_REF_DEPTH;
                dfs_level<D + 1>(e, c, cells, ios, min_doublets, ref_bits, leaf);
            }
        } else {
            // a live fifth doublet: the reference's @assert fails
            const uint32_t ext = (nn > 4u) ? e.s.nbr_ext[node.id] : 0u;
            for (uint32_t i = 0; i < nn; ++i) {
                const uint32_t n = (i < 4u) ? comp4(node.nb, i) : e.s.nbr_pool[ext + i - 4u];
                if (e.s.cell_flags[2u * n] & kKilled) continue;
                child = true;
                ref_bits |= PT_REF_DEPTH;
            }
        }
    }
    if (!child && (uint32_t)(D + 1) >= min_doublets) leaf(cells, ios, (uint32_t)(D + 1));
}

template <class F>
__device__ __forceinline__ void dfs(const EvDesc &e, const DNode &root, uint32_t min_doublets, uint32_t &ref_bits,
                                    F &&leaf) {
    uint32_t cells[4] = {0u, 0u, 0u, 0u}, ios[4] = {0u, 0u, 0u, 0u};
    dfs_level<0>(e, root, cells, ios, min_doublets, ref_bits, leaf);
}

// Start candidates (cells with neighbours on a start pair) that the fishbone
// left alive: one lane per candidate runs the DFS once, counts the tuples and
// tuple hits (summed per tile of kDfsTile cells for the scan), and keeps up to
// kTupKeep tuples' cells in a per-event buffer; a candidate with more tuples
// runs the DFS again when emitting.
constexpr uint32_t kTupKeep = 4;

__global__ void __launch_bounds__(256) k_dfs_count(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    uint32_t ref = 0, err = 0;
    const uint32_t min_d = cfg.p.min_hits_per_ntuplet - 1;
    PT_EVENT_LOOP(e, i, e.s.misc[kMiscStart]) {
        const uint32_t c = e.s.start_list[i];
        const DNode root = load_node(e, c);
        uint64_t v = 0;
        uint2 tb = make_uint2(~0u, 0u);
        if (!root.killed) {
            uint32_t nt = 0, nh = 0;
            uint4 keep[kTupKeep];
#pragma unroll
            for (uint32_t k = 0; k < kTupKeep; ++k) keep[k] = make_uint4(~0u, ~0u, ~0u, ~0u);
            dfs(e, root, min_d, ref, [&](const uint32_t *cells, const uint32_t *, uint32_t len) {
                const uint4 t4 = make_uint4(cells[0], cells[1], len > 2 ? cells[2] : ~0u, len > 3 ? cells[3] : ~0u);
#pragma unroll
                for (uint32_t k = 0; k < kTupKeep; ++k)
                    if (nt == k) keep[k] = t4;
                ++nt;
                nh += len + 1;
            });
            v = ((uint64_t)nt << 32) | nh;
            tb.y = nt;
            // buffer space for the lanes that keep their tuples: one atomic for the
            // wavefront's active lanes, each lane's offset from the bits of the
            // requests of the lanes below it (requests are 1..kTupKeep)
            const uint32_t req = (nt && nt <= kTupKeep) ? nt : 0u;
            uint32_t below = 0, total = 0;
#pragma unroll
            for (int bit = 0; bit < 3; ++bit) {
                const uint64_t m = __ballot((req >> bit) & 1u);
                below += (uint32_t)__popcll(m & lanes_below()) << bit;
                total += (uint32_t)__popcll(m) << bit;
            }
            const uint64_t want = __ballot(req != 0u);
            if (want) {
                const int leader = __ffsll((long long)want) - 1;
                uint32_t base = 0;
                if ((int)lane_id() == leader) base = atomicAdd(&e.s.misc[kMiscTupBuf], total);
                base = __shfl(base, leader, 64);
                const uint32_t off = base + below;
                if (req && off + nt <= kTupBuf) {
                    tb = make_uint2(off, nt);
#pragma unroll
                    for (uint32_t k = 0; k < kTupKeep; ++k)
                        if (k < nt) e.s.tup_buf[off + k] = keep[k];
                }
            }
        }
        e.s.tup[c] = v;
        e.s.tup_where[i] = tb;
        if (v) atomicAdd(reinterpret_cast<unsigned long long *>(&e.s.dfs_agg[c / kDfsTile]), (unsigned long long)v);
    }
    wave_or_atomic(&e.cnt->ref_assert, ref);
    (void)err;
}

// Exclusive scan, in cell order, of the start cells' (tuples, hits): block per
// tile of kDfsTile cells (4 per thread), the tile's prefix summed from the
// preceding tiles' totals.
__global__ void __launch_bounds__(256) k_dfs_scan(const EvDesc *__restrict__ evs, EvGrid grid) {
    __shared__ uint64_t lsum[4];
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    const uint32_t n_cells = e.cnt->n_cells;
    const uint32_t n_tiles = (n_cells + kDfsTile - 1) / kDfsTile;
    for (uint32_t T = b; T < n_tiles; T += grid.per_event) {
        uint64_t part = 0;
        for (uint32_t k = threadIdx.x; k < T; k += 256u) part += e.s.dfs_agg[k];
        const uint32_t c0 = T * kDfsTile + threadIdx.x * 4u;
        const uint64_t bits = (c0 < n_cells) ? (e.s.start_bits[c0 >> 6] >> (c0 & 63u)) : 0ull;
        uint64_t v[4];
        uint64_t mine = 0;
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            v[u] = (((bits >> u) & 1ull) && c0 + u < n_cells) ? e.s.tup[c0 + u] : 0ull;
            mine += v[u];
        }
        const uint64_t pre = block_sum(part, lsum);
        uint64_t tot;
        uint64_t run = pre + block_excl_scan4(mine, lsum, tot);
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            if (v[u]) e.s.tup[c0 + u] = run;
            run += v[u];
        }
        if (T == n_tiles - 1 && threadIdx.x == 0) {
            const uint64_t all = pre + tot;
            e.cnt->n_tuples = (uint32_t)(all >> 32);
            e.cnt->n_tuple_hits = (uint32_t)all;
            if ((uint32_t)(all >> 32) > kMaxTuples) atomicOr(&e.cnt->overflow, PT_OVF_TUPLES);
        }
    }
}

// Emit the tuples of each live start candidate at its scanned position:
// container offsets and hit ids (bulk_fill), the tuple's cells, quality bad,
// and per cell the tuple count and largest tuple size for the early duplicate
// remover.  The first 32768 tuples are stored, like the capped container.
// Tuples kept by the count pass are copied; the others come from a second DFS.
__global__ void __launch_bounds__(256) k_dfs_emit(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    uint32_t ref = 0, err = 0, ovf = 0, lists = 0;
    const uint32_t min_d = cfg.p.min_hits_per_ntuplet - 1;
    PT_EVENT_LOOP(e, i, e.s.misc[kMiscStart]) {
        const uint32_t c = e.s.start_list[i];
        const uint2 tb = e.s.tup_where[i];  // (kept-tuple offset or ~0, tuples)
        if (!tb.y) continue;               // killed, or no tuple
        const uint64_t pre = e.s.tup[c];
        uint32_t it = (uint32_t)(pre >> 32);  // 0-based id of this cell's first tuple
        uint32_t hp = (uint32_t)pre;
        // per cell of a stored tuple: the tuple count before this one (its
        // atomic issued with the others of the tuple) decides the lists
        // (sl_pre: the shared-cell slot when reserved beforehand, else ~0u)
        auto after_count = [&](uint32_t cq, uint32_t before, uint32_t sl_pre) {
            if (before == 0u) ++lists;
            if (before == 1u) {  // the cell now carries two tuples: a shared cell
                const uint32_t sl = (sl_pre != ~0u) ? sl_pre : atomicAdd(&e.s.misc[kMiscDc], 1u);
                if (sl < kDcCap) {
                    e.s.dc_cells[sl] = cq;
                    e.s.dc_slot[cq] = sl;
                } else {
                    e.s.dc_slot[cq] = ~0u;
                    err |= PT_INT_OVF_LIST;
                }
            }
            if (before == PT_MAX_CELL_TRACKS) {  // list capped: exact path
                ovf |= PT_OVF_CELL_TRACKS;
                const uint32_t k = atomicAdd(&e.s.misc[kMiscTrkOvf], 1u);
                if (k < kTrkOvfCap)
                    e.s.trk_ovf[k] = cq;
                else
                    err |= PT_INT_OVF_LIST;
            }
        };
        // the tuple's container entries (bulk_fill: 1-based tuple id it), cells and quality
        auto store = [&](const uint32_t *cells, const uint32_t *ios, uint32_t len) {
            e.hit_off[it - 1] = hp + 1u;
            uint4 tc = make_uint4(~0u, ~0u, ~0u, ~0u);
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                if (q < len) {
                    e.hit_bins[hp + q] = (uint16_t)(ios[q] & 0xFFFFu);
                    if (q + 1 == len) e.hit_bins[hp + len] = (uint16_t)(ios[q] >> 16);
                    if (q == 0) tc.x = cells[q];
                    if (q == 1) tc.y = cells[q];
                    if (q == 2) tc.z = cells[q];
                    if (q == 3) tc.w = cells[q];
                }
            }
            reinterpret_cast<PT_G uint4 *>(e.s.tuple_cells)[it - 1] = tc;
            e.quality[it - 1] = PT_QUALITY_BAD;
        };
        auto emit = [&](const uint32_t *cells, const uint32_t *ios, uint32_t len) {
            ++it;  // 1-based tuple id (bulk_fill)
            if (it <= kMaxTuples) {
                uint32_t before[4];
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q)
                    if (q < len) {
                        before[q] = atomicAdd(&e.s.trk[cells[q]].x, 1u);
                        atomicMax(&e.s.trk[cells[q]].y, len + 1u);
                    }
                store(cells, ios, len);
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q)
                    if (q < len) after_count(cells[q], before[q], ~0u);
            }
            hp += len + 1;
        };
        if (tb.x != ~0u) {
            // the kept tuples: cells, then their hits, all loads in flight together;
            // then the counters of all their cells
            uint4 t4[kTupKeep];
            uint32_t io[kTupKeep][4], cc[kTupKeep][4], len[kTupKeep], before[kTupKeep][4];
#pragma unroll
            for (uint32_t k = 0; k < kTupKeep; ++k) t4[k] = (k < tb.y) ? e.s.tup_buf[tb.x + k] : make_uint4(~0u, ~0u, ~0u, ~0u);
#pragma unroll
            for (uint32_t k = 0; k < kTupKeep; ++k) {
                cc[k][0] = t4[k].x, cc[k][1] = t4[k].y, cc[k][2] = t4[k].z, cc[k][3] = t4[k].w;
                len[k] = (cc[k][2] == ~0u) ? 2u : (cc[k][3] == ~0u) ? 3u : 4u;
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) io[k][q] = (cc[k][q] != ~0u) ? e.s.cell_io[cc[k][q]] : 0u;
            }
#pragma unroll
            for (uint32_t k = 0; k < kTupKeep; ++k)
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q)
                    if (k < tb.y && it + k < kMaxTuples && q < len[k]) {
                        before[k][q] = atomicAdd(&e.s.trk[cc[k][q]].x, 1u);
                        atomicMax(&e.s.trk[cc[k][q]].y, len[k] + 1u);
                    }
            // shared-cell slots for the cells that now carry two tuples: one atomic
            // for the wavefront's lanes on this path (at most 16 per lane)
            uint32_t n1 = 0;
#pragma unroll
            for (uint32_t k = 0; k < kTupKeep; ++k)
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q)
                    if (k < tb.y && it + k < kMaxTuples && q < len[k] && before[k][q] == 1u) ++n1;
            uint32_t below = 0, total = 0;
#pragma unroll
            for (int bit = 0; bit < 5; ++bit) {
                const uint64_t m = __ballot((n1 >> bit) & 1u);
                below += (uint32_t)__popcll(m & lanes_below()) << bit;
                total += (uint32_t)__popcll(m) << bit;
            }
            const uint64_t want = __ballot(n1 != 0u);
            uint32_t sl_next = 0;
            if (want) {
                const int leader = __ffsll((long long)want) - 1;
                uint32_t base = 0;
                if ((int)lane_id() == leader) base = atomicAdd(&e.s.misc[kMiscDc], total);
                sl_next = (uint32_t)__shfl(base, leader, 64) + below;
            }
#pragma unroll
            for (uint32_t k = 0; k < kTupKeep; ++k) {
                if (k >= tb.y) break;
                ++it;
                if (it <= kMaxTuples) {
                    store(cc[k], io[k], len[k]);
#pragma unroll
                    for (uint32_t q = 0; q < 4; ++q)
                        if (q < len[k]) {
                            const bool two = before[k][q] == 1u;
                            after_count(cc[k][q], before[k][q], two ? sl_next : ~0u);
                            sl_next += two ? 1u : 0u;
                        }
                }
                hp += len[k] + 1;
            }
        } else {
            const DNode root = load_node(e, c);
            dfs(e, root, min_d, ref, emit);
        }
    }
    wave_or_atomic(&e.cnt->overflow, ovf);
    wave_or_atomic(&e.cnt->internal_error, err);
    wave_add_atomic(&e.s.misc[kMiscTrkLists], lists);
    (void)ref;
}

// ------------------------------------------------------------------ after the ntuplets
// Per stored tuple: kernel_early_duplicate_remover
// (ca_hit_ntuplet_generator_kernels_impl.jl:143-166) - a tuple is dup if one of
// its cells carries a longer tuple.  Cells whose list was capped at 64, or
// that hold the last stored tuple when the container overflowed (its size
// reads as 0 after finalize), are left to the exact path in k_multiplicity.
// Per container offset: bulk_finalize_fill (hist_to_container.jl:150-161) and
// the module-index copy kernel_fill_hit_indices
// (ca_hit_ntuplet_generator_kernels_impl.jl:25-38).
__global__ void __launch_bounds__(256) k_tuples_post(const EvDesc *__restrict__ evs, EvGrid grid) {
    PT_EVENT_LOOP(e, i, kMaxTuples + 1u + min(e.cnt->n_tuple_hits, PT_MAX_TUPLE_HITS)) {
        const uint32_t m = e.cnt->n_tuples, nhits = e.cnt->n_tuple_hits;
        const uint32_t mc = m < kMaxTuples ? m : kMaxTuples;
        if (i <= kMaxTuples) {
            if (i < mc) {
                const uint4 tc = reinterpret_cast<const uint4 *>(e.s.tuple_cells)[i];
                uint4 last = make_uint4(~0u, ~0u, ~0u, ~0u);
                if (m > kMaxTuples) last = reinterpret_cast<const uint4 *>(e.s.tuple_cells)[kMaxTuples - 1];
                const uint32_t cs[4] = {tc.x, tc.y, tc.z, tc.w};
                const uint32_t len = (tc.y == ~0u) ? 1u : (tc.z == ~0u) ? 2u : (tc.w == ~0u) ? 3u : 4u;
                bool dup = false;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t cq = cs[q];
                    if (cq == ~0u) break;
                    if (cq == last.x || cq == last.y || cq == last.z || cq == last.w) continue;
                    const uint2 tr = e.s.trk[cq];
                    if (tr.x < 2u || tr.x > PT_MAX_CELL_TRACKS) continue;
                    if (tr.y != len + 1u) dup = true;
                }
                if (dup) e.quality[i] = PT_QUALITY_DUP;
            }
            uint32_t off;
            if (m > kMaxTuples) {
                off = (i == kMaxTuples) ? e.hit_off[kMaxTuples - 1] : e.hit_off[i];
                if (i == kMaxTuples) e.hit_off[kMaxTuples] = off;
            } else if (i >= m) {
                off = nhits + 1u;
                e.hit_off[i] = off;
            } else {
                off = e.hit_off[i];
            }
            e.det_off[i] = off;
        } else {
            const uint32_t k = i - (kMaxTuples + 1u);
            const uint32_t end = (m > kMaxTuples) ? e.hit_off[kMaxTuples - 1] : nhits + 1u;
            if (k + 1u < end) e.det_bins[k] = (uint16_t)e.det[e.hit_bins[k] - 1u];
        }
    }
}

// ------------------------------------------------------------------ multiplicity
// Exact early-duplicate pass for the cells set aside above: the cell's list is
// its first 64 tuples in id order; tuples shorter than the longest are dup.
__device__ void trk_exact(const EvDesc &e, uint32_t cell, uint32_t mc) {
    const uint4 *tcs = reinterpret_cast<const uint4 *>(e.s.tuple_cells);
    auto size_of = [&](uint32_t t) { return e.hit_off[t + 1] - e.hit_off[t]; };
    const uint32_t cnt = e.s.trk[cell].x;  // stored tuples on the cell
    if (cnt < 2u) return;
    if (cnt <= PT_MAX_CELL_TRACKS) {
        // the list is every stored tuple on the cell, so the order of the scan
        // does not matter: from the last stored tuple down (the cells of the
        // container's last tuple come here), stopping once all cnt are found
        uint32_t maxh = 0;
        for (int pass = 0; pass < 2; ++pass) {
            uint32_t r = 0;
            for (uint32_t t1 = mc; t1 > 0u && r < cnt; t1 = (t1 > 64u) ? t1 - 64u : 0u) {
                const uint32_t t = t1 - 1u - lane_id();
                bool has = false;
                if (lane_id() < t1) {
                    const uint4 tc = tcs[t];
                    has = tc.x == cell || tc.y == cell || tc.z == cell || tc.w == cell;
                }
                if (pass == 0)
                    maxh = max(maxh, wave_max(has ? size_of(t) : 0u));
                else if (has && size_of(t) != maxh)
                    e.quality[t] = PT_QUALITY_DUP;
                r += (uint32_t)__popcll(__ballot(has));
            }
        }
        return;
    }
    // a capped list: its first PT_MAX_CELL_TRACKS tuples in id order
    uint32_t r = 0, maxh = 0;
    for (uint32_t t0 = 0; t0 < mc && r < PT_MAX_CELL_TRACKS; t0 += 64u) {
        const uint32_t t = t0 + lane_id();
        bool has = false;
        if (t < mc) {
            const uint4 tc = tcs[t];
            has = tc.x == cell || tc.y == cell || tc.z == cell || tc.w == cell;
        }
        const uint64_t bal = __ballot(has);
        const bool in = has && r + (uint32_t)__popcll(bal & lanes_below()) < PT_MAX_CELL_TRACKS;
        maxh = max(maxh, wave_max(in ? size_of(t) : 0u));
        r += (uint32_t)__popcll(bal);
    }
    if (r < 2u) return;
    r = 0;
    for (uint32_t t0 = 0; t0 < mc && r < PT_MAX_CELL_TRACKS; t0 += 64u) {
        const uint32_t t = t0 + lane_id();
        bool has = false;
        if (t < mc) {
            const uint4 tc = tcs[t];
            has = tc.x == cell || tc.y == cell || tc.z == cell || tc.w == cell;
        }
        const uint64_t bal = __ballot(has);
        const bool in = has && r + (uint32_t)__popcll(bal & lanes_below()) < PT_MAX_CELL_TRACKS;
        if (in && size_of(t) != maxh) e.quality[t] = PT_QUALITY_DUP;
        r += (uint32_t)__popcll(bal);
    }
}

// kernel_countMultiplicity / finalize! / kernel_fillMultiplicity
// (ca_hit_ntuplet_generator_kernels_impl.jl:174-211): bucket = hit count, each
// bucket lists tuple ids in DESCENDING order (fill_direct decrements).  Two
// blocks per event run its independent parts side by side: block 0 the
// offsets of the shared cells' tuple lists, block 1 the exact early-duplicate
// pass, the event's flags and the buckets.
constexpr uint32_t kMulU = 8, kMulChunk = 1024u * kMulU;  // shared cells per chunk of the list offsets
constexpr uint8_t kMulNone = 5, kMulLong = 6;
__global__ void __launch_bounds__(1024) k_multiplicity(const EvDesc *__restrict__ evs, EvGrid grid) {
    __shared__ uint32_t lds[5][1024 / 64 + 1];
    __shared__ uint32_t tot[5];
    __shared__ uint32_t s_dcn[kMulChunk];
    __shared__ uint8_t s_code[kMaxTuples];
    uint32_t ev, blk;
    if (!ev_block(grid, ev, blk)) return;
    const EvDesc &e = evs[ev];
    const uint32_t m = e.cnt->n_tuples < kMaxTuples ? e.cnt->n_tuples : kMaxTuples;
    if (blk == 0) {
        // offsets of the shared cells' tuple lists (classification), in slot
        // order, kMulChunk cells at a time: their tuple counts are gathered with
        // coalesced loads into LDS, then each thread scans a contiguous run
        uint32_t base = 0;
        const uint32_t nd = min(e.s.misc[kMiscDc], kDcCap);
        for (uint32_t d0 = 0; d0 < nd; d0 += kMulChunk) {
            const uint32_t nc = min(nd - d0, kMulChunk);
            uint32_t cell[kMulU], v[kMulU];
#pragma unroll
            for (uint32_t u = 0; u < kMulU; ++u) {
                const uint32_t k = threadIdx.x + 1024u * u;
                cell[u] = (k < nc) ? e.s.dc_cells[d0 + k] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < kMulU; ++u) v[u] = (threadIdx.x + 1024u * u < nc) ? e.s.trk[cell[u]].x : 0u;
#pragma unroll
            for (uint32_t u = 0; u < kMulU; ++u) {
                const uint32_t k = threadIdx.x + 1024u * u;
                if (k < nc) {
                    s_dcn[k] = v[u];
                    e.s.dc_cur[d0 + k] = 0u;
                    e.s.dc_done[d0 + k] = 0u;
                }
            }
            __syncthreads();
            const uint32_t k0 = threadIdx.x * kMulU;
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t u = 0; u < kMulU; ++u) sum += (k0 + u < nc) ? s_dcn[k0 + u] : 0u;
            uint32_t total;
            uint32_t run = base + block_excl_scan(sum, lds[0], total);
#pragma unroll
            for (uint32_t u = 0; u < kMulU; ++u)
                if (k0 + u < nc) {
                    e.s.dc_off[d0 + k0 + u] = run;
                    run += s_dcn[k0 + u];
                }
            base += total;
            __syncthreads();  // s_dcn is refilled by the next chunk
        }
        return;
    }
    // block 1: the exact early-duplicate pass, the event's flags, the buckets
    {
        const uint32_t n_ovf = min(e.s.misc[kMiscTrkOvf], kTrkOvfCap);
        const uint32_t extra = (e.cnt->n_tuples > kMaxTuples) ? 4u : 0u;
        const uint32_t wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
        for (uint32_t k = wv; k < n_ovf + extra; k += nw) {
            uint32_t cell;
            if (k < n_ovf) {
                cell = e.s.trk_ovf[k];
            } else {
                cell = e.s.tuple_cells[4u * (kMaxTuples - 1u) + (k - n_ovf)];
                if (cell == ~0u) continue;
            }
            trk_exact(e, cell, m);
        }
        if (threadIdx.x == 0) {
            if (e.cnt->n_cells == 0) atomicOr(&e.cnt->ref_assert, PT_REF_NO_CELLS);
            uint32_t o = 0;
            if (e.s.misc[kMiscNbrLists] > PT_MAX_ACTIVE_LISTS - 1u) o |= PT_OVF_NEIGHBOR_LISTS;
            if (e.s.misc[kMiscTrkLists] > PT_MAX_ACTIVE_LISTS - 1u) o |= PT_OVF_TRACK_LISTS;
            if (o) atomicOr(&e.cnt->overflow, o);
            // tuple lists: the reference's later cells would share its last
            // column (extend!, simple_matrix.jl:13-19); kept per cell here, so
            // the event's results are not the reference's
            if (o & PT_OVF_TRACK_LISTS) atomicOr(&e.cnt->internal_error, PT_INT_LIST_SHARE);
        }
        __syncthreads();
    }
    // per tuple, coalesced: its bucket (hit count 3..7 -> 0..4), kMulNone when
    // it takes no part (fewer than 3 hits or dup), kMulLong for 8 or more hits
    for (uint32_t t0 = 0; t0 < m; t0 += 8192u) {
        uint32_t a[8], b[8];
        int32_t q[8];
#pragma unroll
        for (uint32_t u = 0; u < 8u; ++u) {
            const uint32_t t = t0 + threadIdx.x + 1024u * u;
            a[u] = (t < m) ? e.hit_off[t] : 0u;
            b[u] = (t < m) ? e.hit_off[t + 1] : 0u;
            q[u] = (t < m) ? e.quality[t] : 0;
        }
#pragma unroll
        for (uint32_t u = 0; u < 8u; ++u) {
            const uint32_t t = t0 + threadIdx.x + 1024u * u;
            const uint32_t nh = b[u] - a[u];
            if (t < m)
                s_code[t] = (nh < 3u || q[u] == PT_QUALITY_DUP) ? kMulNone : (nh >= 8u) ? kMulLong : (uint8_t)(nh - 3u);
        }
    }
    __syncthreads();
    // each thread a contiguous run of tuples, so ranks in a bucket follow id order
    const uint32_t chunk = (m + blockDim.x - 1) / blockDim.x;
    const uint32_t t0 = threadIdx.x * chunk, t1 = min(m, t0 + chunk);
    uint32_t c[5] = {0, 0, 0, 0, 0};
    uint32_t ref = 0;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t k = s_code[t];
        if (k == kMulLong) ref |= PT_REF_MULT;
#pragma unroll
        for (uint32_t bk = 0; bk < 5u; ++bk) c[bk] += (k == bk) ? 1u : 0u;
    }
    uint32_t ex[5];
    for (int bk = 0; bk < 5; ++bk) {
        uint32_t total;
        ex[bk] = block_excl_scan(c[bk], lds[bk], total);
        if (threadIdx.x == 0) tot[bk] = total;
    }
    __syncthreads();
    // off[] after finalize!: exclusive start of bucket nh at tm_off[nh-1]
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int bk = 0; bk < 9; ++bk) {
            e.s.tm_off[bk] = run;
            if (bk >= 2 && bk < 7) run += tot[bk - 2];
        }
        for (int bk = 3; bk < 8; ++bk) e.cnt->n_by_len[bk] = tot[bk - 3];
        for (int bk = 3; bk <= 5; ++bk)
            if (tot[bk - 3] > PT_MAX_FITS) atomicOr(&e.cnt->overflow, PT_OVF_FITS);
    }
    // last position of each bucket (tm_off[bk + 2] + tot[bk] - 1), in registers:
    // a bucket's ranks fill it from the end
    uint32_t last[5];
    {
        uint32_t run = 0;
#pragma unroll
        for (int bk = 0; bk < 5; ++bk) {
            run += tot[bk];
            last[bk] = run - 1u;
        }
    }
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t bk = s_code[t];
        if (bk >= 5u) continue;
        uint32_t pos = 0;
#pragma unroll
        for (uint32_t j = 0; j < 5u; ++j)
            if (bk == j) pos = last[j] - ex[j]++;
        e.s.tm_bins[pos] = (uint16_t)(t + 1u);
    }
    if (ref) atomicOr(&e.cnt->ref_assert, ref);
}

// ------------------------------------------------------------------ launchers
// Blocks per event for a kernel whose wavefronts take work by ticket: exactly
// the blocks that are resident at once on the events' XCDs (occupancy from
// the runtime), so every wavefront starts at once and none waits for a slot.
static std::mutex g_occ_mu;
static std::map<std::pair<const void *, int>, std::pair<int, int>> g_occ;  // (kernel, device) -> (blocks/CU, CUs)
template <class K>
static EvGrid ev_grid_resident(K kernel, uint32_t n_events, uint32_t xcd, int bs = 256) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    int per_cu = 0, cus = 0;
    {
        std::lock_guard<std::mutex> lk(g_occ_mu);
        const auto key = std::make_pair((const void *)kernel, dev);
        auto it = g_occ.find(key);
        if (it == g_occ.end()) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, bs, 0) != hipSuccess || per_cu < 1)
                per_cu = 1;
            hipDeviceProp_t prop;
            cus = (hipGetDeviceProperties(&prop, dev) == hipSuccess) ? prop.multiProcessorCount : 256;
            g_occ[key] = std::make_pair(per_cu, cus);
        } else {
            per_cu = it->second.first;
            cus = it->second.second;
        }
    }
    const uint32_t ev_per_xcd = (n_events + 7u) / 8u;
    uint32_t per = (uint32_t)(cus / 8) * (uint32_t)per_cu / ev_per_xcd;
    if (per < 1) per = 1;
    return EvGrid{n_events, per, xcd};
}

// Launch sequence of the CA stage.  When `tm` is given, an event is recorded
// after every kernel so the host can read per-kernel durations (the caller
// records the first one before the stage).
void launch_ca(const EvDesc *d_evs, uint32_t n_events, const Config &cfg, uint32_t max_hits, uint32_t max_slots,
               int zero_state, hipStream_t st, Marks *tm) {
    const uint32_t est_cells = max_slots * 4u + 1024u;
    auto mark = [&](const char *name) {
        if (tm) tm->record(name, st);
    };
    const uint32_t xcd = cfg.xcd;
    const EvGrid one{n_events, 1u, xcd};
    const EvGrid g_reset = ev_grid(max_hits + max_slots / kSlotTile + 64u +
                                       (zero_state ? 20u * kMaxTuples : 0u),
                                   n_events, xcd);
    k_reset<<<ev_grid_dim(g_reset), 256, 0, st>>>(d_evs, g_reset, zero_state);
    mark("k_reset");
    // the ticketed kernels: exactly the resident blocks, in the tuned instantiation
    auto resident = [&](auto kernel, auto... args) {
        const EvGrid g = ev_grid_resident(kernel, n_events, xcd);
        hipLaunchKernelGGL(kernel, ev_grid_dim(g), dim3(256), 0, st, d_evs, g, args...);
    };
    switch (cfg.count_rounds) {
        case 1: resident(k_doublet_count<1>, cfg, zero_state); break;
        case 2: resident(k_doublet_count<2>, cfg, zero_state); break;
        default: resident(k_doublet_count<4>, cfg, zero_state); break;
    }
    mark("k_doublet_count");
    const EvGrid g_dscan = ev_grid(max_slots / 4u, n_events, xcd);
    k_doublet_scan<<<ev_grid_dim(g_dscan), 256, 0, st>>>(d_evs, g_dscan, cfg);
    mark("k_doublet_scan");
    switch (cfg.fill_rounds) {
        case 1: resident(k_doublet_fill<1>, cfg); break;
        case 2: resident(k_doublet_fill<2>, cfg); break;
        default: resident(k_doublet_fill<3>, cfg); break;
    }
    mark("k_doublet_fill");
    k_ohc_rebuild<<<ev_grid_dim(one), 256, 0, st>>>(d_evs, one, cfg);
    mark("k_ohc_rebuild");
    switch (cfg.conn_hits * 1000u + cfg.conn_stage) {
        case 4128: resident(k_connect<4, 128>, cfg); break;
        case 4192: resident(k_connect<4, 192>, cfg); break;
        case 4256: resident(k_connect<4, 256>, cfg); break;
        case 4384: resident(k_connect<4, 384>, cfg); break;
        case 8128: resident(k_connect<8, 128>, cfg); break;
        case 8192: resident(k_connect<8, 192>, cfg); break;
        case 8256: resident(k_connect<8, 256>, cfg); break;
        case 8384: resident(k_connect<8, 384>, cfg); break;
        case 16128: resident(k_connect<16, 128>, cfg); break;
        case 16192: resident(k_connect<16, 192>, cfg); break;
        case 16256: resident(k_connect<16, 256>, cfg); break;
        case 16384: resident(k_connect<16, 384>, cfg); break;
        case 24128: resident(k_connect<24, 128>, cfg); break;
        case 24256: resident(k_connect<24, 256>, cfg); break;
        case 24384: resident(k_connect<24, 384>, cfg); break;
        case 32128: resident(k_connect<32, 128>, cfg); break;
        case 32192: resident(k_connect<32, 192>, cfg); break;
        case 32256: resident(k_connect<32, 256>, cfg); break;
        case 32384: resident(k_connect<32, 384>, cfg); break;
        default: resident(k_connect<24, 192>, cfg); break;
    }
    mark("k_connect");
    if (cfg.p.early_fish_bone) {
        const EvGrid g_fl = ev_grid(max_hits, n_events, xcd);
        k_fish_list<<<ev_grid_dim(g_fl), 256, 0, st>>>(d_evs, g_fl);
        mark("k_fish_list");
        switch (cfg.fish_hits) {
            case 8: resident(k_fishbone<8>, cfg); break;
            case 32: resident(k_fishbone<32>, cfg); break;
            default: resident(k_fishbone<16>, cfg); break;
        }
    } else {
        k_connect_ext<<<ev_grid_dim(one), 256, 0, st>>>(d_evs, one, cfg);
    }
    mark("k_fishbone");
    k_nbr_share<<<ev_grid_dim(one), 1024, 0, st>>>(d_evs, one);
    mark("k_nbr_share");
    const EvGrid g_start = ev_grid(est_cells / 8u, n_events, xcd);
    k_dfs_count<<<ev_grid_dim(g_start), 256, 0, st>>>(d_evs, g_start, cfg);
    mark("k_dfs_count");
    const EvGrid g_scan = ev_grid(est_cells / 4u, n_events, xcd);
    k_dfs_scan<<<ev_grid_dim(g_scan), 256, 0, st>>>(d_evs, g_scan);
    mark("k_dfs_scan");
    k_dfs_emit<<<ev_grid_dim(g_start), 256, 0, st>>>(d_evs, g_start, cfg);
    mark("k_dfs_emit");
    const EvGrid g_post = ev_grid(kMaxTuples + 1u + PT_MAX_TUPLE_HITS / 4u, n_events, xcd);
    k_tuples_post<<<ev_grid_dim(g_post), 256, 0, st>>>(d_evs, g_post);
    mark("k_tuples_post");
    const EvGrid two{n_events, 2u, xcd};
    k_multiplicity<<<ev_grid_dim(two), 1024, 0, st>>>(d_evs, two);
    mark("k_multiplicity");
}

}  // namespace pt
