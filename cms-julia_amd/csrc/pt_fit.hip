// pt_fit.hip - broken-line fit kernels for gfx950.
//
// k_fit_tuples: the fit stage of make_tuples (launchBrokenLineKernelsOnCPU,
//   broken_line_fit_on_gpu.jl:166-203): for each length bucket (3 -> N=3,
//   4 -> N=4, 5 -> N=4 on the first four hits, fit5as4) fit the first 24,576
//   entries of the tuple-multiplicity bucket, i.e. the highest tuple ids, and
//   classify each fitted tuple (the first step of classify_tuples).
//   Gather per hit: global position, module frame, local variances rotated to
//   the global covariance in float (SOARotation.jl:139-147).
// k_fit_flat: fit of pre-gathered tracks (fit-only workloads), broken-line or
//   Riemann (pt_riemann.hpp; the counterpart of launch_riemann_kernels!,
//   helix_fit_on_gpu.jl:94-96).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "pt_bl_adjoint.hpp"
#include "pt_bl_fit.hpp"
#include "pt_device.hpp"
#include "pt_riemann.hpp"

namespace pt {

__device__ __forceinline__ void to_global_err(const float *__restrict__ r, float cxx, float cyy, float (&g)[6]) {
    const float cxy = 0.0f;
    const float xx = r[0], xy = r[1], xz = r[2], yx = r[3], yy = r[4], yz = r[5];
    g[0] = xx * (xx * cxx + yx * cxy) + yx * (xx * cxy + yx * cyy);
    g[1] = xx * (xy * cxx + yy * cxy) + yx * (xy * cxy + yy * cyy);
    g[2] = xy * (xy * cxx + yy * cxy) + yy * (xy * cxy + yy * cyy);
    g[3] = xx * (xz * cxx + yz * cxy) + yx * (xz * cxy + yz * cyy);
    g[4] = xy * (xz * cxx + yz * cxy) + yy * (xz * cxy + yz * cyy);
    g[5] = xz * (xz * cxx + yz * cxy) + yz * (xz * cxy + yz * cyy);
}

__device__ __forceinline__ uint32_t fit_flag_bits(uint32_t f) {
    uint32_t r = 0;
    if (f & 1u) r |= PT_REF_SQRT_DOMAIN;
    if (f & 2u) r |= PT_REF_FIT_CHARGE;
    if (f & 4u) r |= PT_REF_FIT_NAN;
    return r;
}

// The fits of one length bucket, items idx0, idx0 + stride, ...
template <class P, int N>
__device__ __forceinline__ void fit_bucket(const EvDesc &e, const Config &cfg, int bucket, uint32_t idx0,
                                           uint32_t stride, uint32_t &ref, uint32_t &nfit) {
    const uint32_t b0 = e.s.tm_off[bucket - 1], b1 = e.s.tm_off[bucket];
    uint32_t n = b1 - b0;
    if (n > PT_MAX_FITS) n = PT_MAX_FITS;
    for (uint32_t idx = idx0; idx < n; idx += stride) {
        const uint32_t tkid = e.s.tm_bins[b0 + idx];  // 1-based
        const uint32_t start = e.hit_off[tkid - 1] - 1u;
        float hx[N], hy[N], hz[N];
        float ge[6][N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const uint32_t h = e.hit_bins[start + k] - 1u;
            int det = e.det[h];
            if (det < 0 || det >= (int)PT_N_MODULES) {  // the reference's frame lookup would throw
                ref |= PT_REF_BAD_MODULE;
                det = 0;
            }
            float g[6];
            to_global_err(cfg.frames + 9 * det, e.xerr[h], e.yerr[h], g);
#pragma unroll
            for (int c = 0; c < 6; ++c) ge[c][k] = g[c];
            hx[k] = e.xg[h];
            hy[k] = e.yg[h];
            hz[k] = e.zg[h];
        }
        FitOut<typename P::Prop> o;
        BLFit<P, N> f;
        f.run(hx, hy, hz, ge, cfg.B, o);
        ref |= fit_flag_bits(o.flags);
        const uint32_t row = tkid - 1u;
        float st[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            st[k] = (float)o.state[k];
            e.state[k * kMaxTuples + row] = st[k];
        }
#pragma unroll
        for (int k = 0; k < 15; ++k) e.cov[k * kMaxTuples + row] = (float)o.cov[k];
        const float ptv = (float)o.pt, chi2 = (float)o.chi2;
        e.pt[row] = ptv;
        e.eta[row] = (float)o.eta;
        e.chi2[row] = chi2;
        // kernel_classify_tracks (ca_hit_ntuplet_generator_kernels_impl.jl:213-247)
        // on the stored float values: the tuple is not dup (the buckets hold none)
        // and has `bucket` hits
        bool is_nan = false;
#pragma unroll
        for (int k = 0; k < 5; ++k) is_nan |= isnan(st[k]);
        if (!is_nan) {
            const float pt = isnan(ptv) ? ptv : (cfg.p.chi2_max_pt < ptv ? cfg.p.chi2_max_pt : ptv);
            const float *cf = cfg.p.chi2_coeff;
            const float chi2_cut = cfg.p.chi2_scale * (cf[0] + pt * (cf[1] + pt * (cf[2] + pt * cf[3])));
            if (!(3.0f * chi2 >= chi2_cut)) {
                const bool quad = bucket > 3;
                const float max_tip = quad ? cfg.p.quadruplet_max_tip : cfg.p.triplet_max_tip;
                const float min_pt = quad ? cfg.p.quadruplet_min_pt : cfg.p.triplet_min_pt;
                const float max_zip = quad ? cfg.p.quadruplet_max_zip : cfg.p.triplet_max_zip;
                if (fabsf(st[1]) < max_tip && ptv > min_pt && fabsf(st[4]) < max_zip) e.quality[row] = PT_QUALITY_LOOSE;
            }
        }
        ++nfit;
    }
}

// All three length buckets in one launch (3 -> N=3, 4 -> N=4, 5 -> N=4 on the
// first four hits when fit_5_as_4): the event's blocks are split in three.
// A block is one wavefront: at one wave per SIMD (the fit's registers) a
// larger block would hold a whole CU while only some of its waves have fits,
// and a bucket fills only the first of its blocks.
constexpr uint32_t kFitBlock = 64;
template <class P>
__global__ void __launch_bounds__(kFitBlock) k_fit_tuples(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    const uint32_t per = grid.per_event / 3u;
    const uint32_t bk = b / per, bb = b - bk * per;
    uint32_t ref = 0, nfit = 0;
    if (bk == 0)
        fit_bucket<P, 3>(e, cfg, 3, bb * blockDim.x + threadIdx.x, per * blockDim.x, ref, nfit);
    else if (bk == 1)
        fit_bucket<P, 4>(e, cfg, 4, bb * blockDim.x + threadIdx.x, per * blockDim.x, ref, nfit);
    else if (bk == 2 && cfg.p.fit_5_as_4)
        fit_bucket<P, 4>(e, cfg, 5, bb * blockDim.x + threadIdx.x, per * blockDim.x, ref, nfit);
    uint32_t v = nfit;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    uint32_t r = ref;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) r |= __shfl_xor(r, d, 64);
    if ((threadIdx.x & 63) == 0) {
        if (v) atomicAdd(&e.cnt->n_fitted, v);
        if (r) atomicOr(&e.cnt->ref_assert, r);
    }
}

// ------------------------------------------------------------------ fit-only batches
// Pre-gathered tracks, track-minor: hits[(c*5+k)*n+t], ge[(c*5+k)*n+t];
// nhit 3 -> N=3, 4 or 5 -> N=4 (fit5as4); any other count gives NaN outputs.
// A "job" is the per-track work of a batch kernel: cls(t) gives the length
// class (0: 3 hits, 1: 4 or 5 hits, 2: invalid), one<N>(t) the work,
// invalid(t) the output for an invalid track.  The schedules below run any job.

__device__ __forceinline__ int fit_class(uint8_t nh) { return nh == 3 ? 0 : ((nh == 4 || nh == 5) ? 1 : 2); }

template <class P> struct FlatJob {
    static constexpr bool kOutOfLine = false;
    FlatArgs a;
    __device__ __forceinline__ uint32_t size() const { return a.n; }
    __device__ __forceinline__ int cls(uint32_t t) const { return fit_class(a.nhit[t]); }
    template <int N> __device__ __forceinline__ void one(uint32_t t) const {
        const size_t n = a.n;
        float hx[N], hy[N], hz[N];
        float g[6][N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            hx[k] = a.hits[(size_t)(0 * 5 + k) * n + t];
            hy[k] = a.hits[(size_t)(1 * 5 + k) * n + t];
            hz[k] = a.hits[(size_t)(2 * 5 + k) * n + t];
#pragma unroll
            for (int c = 0; c < 6; ++c) g[c][k] = a.ge[(size_t)(c * 5 + k) * n + t];
        }
        FitOut<typename P::Prop> o;
        BLFit<P, N> f;
        f.run(hx, hy, hz, g, a.B, o);
#pragma unroll
        for (int k = 0; k < 5; ++k) a.state[(size_t)k * n + t] = (float)o.state[k];
#pragma unroll
        for (int k = 0; k < 15; ++k) a.cov[(size_t)k * n + t] = (float)o.cov[k];
        a.pt[t] = (float)o.pt;
        a.eta[t] = (float)o.eta;
        a.chi2[t] = (float)o.chi2;
    }
    __device__ __forceinline__ void invalid(uint32_t t) const {
        const float nan = __builtin_nanf("");
#pragma unroll
        for (int k = 0; k < 5; ++k) a.state[(size_t)k * a.n + t] = nan;
        a.pt[t] = nan;
        a.eta[t] = nan;
        a.chi2[t] = nan;
    }
};

// Riemann helix fit of a pre-gathered track (FP64), same input and output
// layout as FlatJob.
struct RiemannJob {
    static constexpr bool kOutOfLine = false;
    FlatArgs a;
    __device__ __forceinline__ uint32_t size() const { return a.n; }
    __device__ __forceinline__ int cls(uint32_t t) const { return fit_class(a.nhit[t]); }
    template <int N> __device__ __forceinline__ void one(uint32_t t) const {
        const size_t n = a.n;
        float x[N], y[N], z[N], g[6][N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            x[k] = a.hits[(size_t)(0 * 5 + k) * n + t];
            y[k] = a.hits[(size_t)(1 * 5 + k) * n + t];
            z[k] = a.hits[(size_t)(2 * 5 + k) * n + t];
#pragma unroll
            for (int c = 0; c < 6; ++c) g[c][k] = a.ge[(size_t)(c * 5 + k) * n + t];
        }
        riemann::RiemannOut<N> o;
        riemann::riemann_fit<N>(x, y, z, g, a.B, o);
#pragma unroll
        for (int k = 0; k < 5; ++k) a.state[(size_t)k * n + t] = (float)o.state[k];
#pragma unroll
        for (int k = 0; k < 15; ++k) a.cov[(size_t)k * n + t] = (float)o.cov15[k];
        a.pt[t] = (float)o.pt;
        a.eta[t] = (float)o.eta;
        a.chi2[t] = (float)o.chi2;
    }
    __device__ __forceinline__ void invalid(uint32_t t) const {
        const float nan = __builtin_nanf("");
#pragma unroll
        for (int k = 0; k < 5; ++k) a.state[(size_t)k * a.n + t] = nan;
        a.pt[t] = nan;
        a.eta[t] = nan;
        a.chi2[t] = nan;
    }
};

// Adjoint of the mixed fit per track: hit errors come as local sigmas and
// module ids (global covariance from the module frame in float, as the CA
// path does), the forward fit is recomputed and the reverse sweep gives
// dL/d(x, y, z, sigma_rphi, sigma_z) of each hit:
//   LOSS = true:  L = chi2_line + chi2_circle + |state - truth|^2 (phi
//                 difference taken modulo 2 pi), written to loss[t]
//   LOSS = false: the pullback of the cotangents dstate[5][n], dchi2[n] of
//                 (state, chi2) (chi2 = (chi2_line + chi2_circle) / (2N - 5))
// Hits a 3-hit track does not have, and the fifth hit of a 5-hit track
// (fitted on its first four), get zero gradient.
template <bool LOSS> struct AdjJob {
    // the persistent schedule calls one<N> through a non-inlined function
    // (see run_one)
    static constexpr bool kOutOfLine = true;
    AdjArgs a;
    __device__ __forceinline__ uint32_t size() const { return a.n; }
    __device__ __forceinline__ int cls(uint32_t t) const { return fit_class(a.nhit[t]); }
    template <int N> __device__ __forceinline__ void one(uint32_t t) const {
        const size_t n = a.n;
        float hx[N], hy[N], hz[N], sr[N], sz[N];
        float g[6][N];
        int det[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            hx[k] = a.hits[(size_t)(0 * 5 + k) * n + t];
            hy[k] = a.hits[(size_t)(1 * 5 + k) * n + t];
            hz[k] = a.hits[(size_t)(2 * 5 + k) * n + t];
            sr[k] = a.sig[(size_t)(0 * 5 + k) * n + t];
            sz[k] = a.sig[(size_t)(1 * 5 + k) * n + t];
            int d = a.det[(size_t)k * n + t];
            det[k] = (d < 0 || d >= (int)PT_N_MODULES) ? 0 : d;
            float gl[6];
            to_global_err(a.frames + 9 * det[k], sr[k] * sr[k], sz[k] * sz[k], gl);
#pragma unroll
            for (int c = 0; c < 6; ++c) g[c][k] = gl[c];
        }
        FitOut<float> o;
        BLFit<PolicyMixed, N> f;
        f.run(hx, hy, hz, g, a.B, o);
        double bstate[5], bcl, bcc;
        if (LOSS) {
            double L = (double)o.chi2_line + (double)o.chi2_circle;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                double dlt = (double)o.state[k] - (double)a.truth[(size_t)k * n + t];
                if (k == 0) dlt -= 2.0 * kPiD * rint(dlt / (2.0 * kPiD));
                L += dlt * dlt;
                bstate[k] = 2.0 * dlt;
            }
            bcl = bcc = 1.0;
            a.loss[t] = (float)L;
        } else {
#pragma unroll
            for (int k = 0; k < 5; ++k) bstate[k] = a.dstate[(size_t)k * n + t];
            bcl = bcc = (double)a.dchi2[t] / (double)(2 * N - 5);
        }
        double gx[N], gy[N], gz[N], gge[6][N];
        bl_fit_vjp<PolicyMixed, N>(f, hx, hy, hz, g, a.B, bstate, bcl, bcc, gx, gy, gz, gge);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            double gsr = 0.0, gsz = 0.0, gxk = 0.0, gyk = 0.0, gzk = 0.0;
            if (k < N) {
                double bg[6], bcxx, bcyy;
#pragma unroll
                for (int c = 0; c < 6; ++c) bg[c] = gge[c][k];
                local_err_vjp(a.frames + 9 * det[k], bg, bcxx, bcyy);
                gsr = 2.0 * (double)sr[k] * bcxx;
                gsz = 2.0 * (double)sz[k] * bcyy;
                gxk = gx[k];
                gyk = gy[k];
                gzk = gz[k];
            }
            a.grad[(size_t)(0 * 5 + k) * n + t] = (float)gxk;
            a.grad[(size_t)(1 * 5 + k) * n + t] = (float)gyk;
            a.grad[(size_t)(2 * 5 + k) * n + t] = (float)gzk;
            a.grad[(size_t)(3 * 5 + k) * n + t] = (float)gsr;
            a.grad[(size_t)(4 * 5 + k) * n + t] = (float)gsz;
        }
        if (a.state) {
#pragma unroll
            for (int k = 0; k < 5; ++k) a.state[(size_t)k * n + t] = o.state[k];
        }
        if (a.chi2) a.chi2[t] = o.chi2;
    }
    __device__ __forceinline__ void invalid(uint32_t t) const {
        const float nan = __builtin_nanf("");
        for (int c = 0; c < 25; ++c) a.grad[(size_t)c * a.n + t] = nan;
        if (LOSS) a.loss[t] = nan;
        if (a.state)
            for (int k = 0; k < 5; ++k) a.state[(size_t)k * a.n + t] = nan;
        if (a.chi2) a.chi2[t] = nan;
    }
};

// One thread per track in input order: lanes of one wavefront take both
// lengths, so a mixed wavefront runs both bodies.
template <class J> __global__ void __launch_bounds__(256) k_fit_plain(J j) {
    const uint32_t n = j.size();
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int c = j.cls(t);
        if (c == 0)
            j.template one<3>(t);
        else if (c == 1)
            j.template one<4>(t);
        else
            j.invalid(t);
    }
}

// Per-length schedule, step 1: split the track ids by length class into
// lists[0..n) (3 hits) and lists[n..2n) (4 or 5 hits); counts[2] must start
// at 0.  Each wavefront appends its ids in input order at a position taken
// once per wavefront, so neighbouring ids stay together.
template <class J> __global__ void __launch_bounds__(256) k_fit_partition(J j, uint32_t *lists, uint32_t *counts) {
    const uint32_t n = j.size();
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < n; base += stride) {
        const uint32_t t = base + lane;
        const int c = (t < n) ? j.cls(t) : 3;
        if (c == 2) j.invalid(t);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint64_t m = __ballot(c == k);
            if (!m) continue;
            uint32_t pos = 0;
            if (lane == 0) pos = atomicAdd(&counts[k], (uint32_t)__popcll(m));
            pos = __shfl(pos, 0, 64);
            if (c == k) lists[(size_t)k * n + pos + __popcll(m & below)] = t;
        }
    }
}

// Per-length schedule, step 2: the tracks of one length class.
template <class J, int N> __global__ void __launch_bounds__(256) k_fit_list(J j, const uint32_t *list,
                                                                            const uint32_t *count) {
    const uint32_t m = *count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        j.template one<N>(list[i]);
}

// The per-track body of a job inside the persistent schedule.  The adjoint's
// body is called out of line: inlined into k_fit_persistent it gave x/y hit
// gradients off by up to 16-26% on about a fifth of the tracks (forward state
// bit-equal, at -O2 and -O3, on 262,144-track batches), while the same code is
// exact inlined into k_fit_plain, exact called out of line, and clean on the
// host under ASan, UBSan and MSan with pattern and zero auto-initialisation.
template <class J, int N> __device__ __attribute__((noinline)) void one_out_of_line(const J &j, uint32_t t) {
    j.template one<N>(t);
}
template <class J, int N> __device__ __forceinline__ void run_one(const J &j, uint32_t t) {
    if constexpr (J::kOutOfLine)
        one_out_of_line<J, N>(j, t);
    else
        j.template one<N>(t);
}

// One persistent launch over uneven lengths: every wavefront takes tiles of
// `tile` consecutive tracks by ticket, sorts the tile's ids into one LDS list
// per length class (ballot compaction, input order kept) and then runs each
// list 64 tracks at a time, so all active lanes of a pass run the same body.
// *ticket must start at 0; dynamic LDS = 8 * tile bytes per wavefront.
template <class J> __global__ void __launch_bounds__(256) k_fit_persistent(J j, uint32_t tile, uint32_t *ticket) {
    extern __shared__ uint32_t fit_lds[];
    const uint32_t n = j.size();
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t *l3 = fit_lds + (size_t)wid * 2u * tile;
    uint32_t *l4 = l3 + tile;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    const uint32_t n_tiles = (n + tile - 1u) / tile;
    for (;;) {
        uint32_t tk = 0;
        if (lane == 0) tk = atomicAdd(ticket, 1u);
        tk = __shfl(tk, 0, 64);
        if (tk >= n_tiles) break;
        const uint32_t t0 = tk * tile, t1 = min(t0 + tile, n);
        uint32_t c3 = 0, c4 = 0;
        for (uint32_t base = t0; base < t1; base += 64u) {
            const uint32_t t = base + lane;
            const int c = (t < t1) ? j.cls(t) : 3;
            if (c == 2) j.invalid(t);
            const uint64_t m3 = __ballot(c == 0), m4 = __ballot(c == 1);
            if (c == 0) l3[c3 + __popcll(m3 & below)] = t;
            if (c == 1) l4[c4 + __popcll(m4 & below)] = t;
            c3 += (uint32_t)__popcll(m3);
            c4 += (uint32_t)__popcll(m4);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (uint32_t p = 0; p < c3; p += 64u)
            if (p + lane < c3) run_one<J, 3>(j, l3[p + lane]);
        for (uint32_t p = 0; p < c4; p += 64u)
            if (p + lane < c4) run_one<J, 4>(j, l4[p + lane]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------------ launchers
void launch_fit_tuples(const EvDesc *d_evs, uint32_t n_events, const Config &cfg, hipStream_t st, Marks *tm) {
    // per bucket: enough wavefronts for the fit cap, at most about 4096 in all
    // (fewer idle ones: 48.9 us per 100 PU50 events against 50.4 at 8192 and
    // 55.9 at 16384)
    uint32_t per = (PT_MAX_FITS + kFitBlock - 1) / kFitBlock;
    const uint32_t target = (4096u + n_events - 1) / n_events / 3u;
    if (per > target) per = target;
    if (per < 1) per = 1;
    const EvGrid grid{n_events, 3u * per, cfg.xcd};
    if (cfg.fit_precision == PT_FIT_MIXED)
        k_fit_tuples<PolicyMixed><<<ev_grid_dim(grid), kFitBlock, 0, st>>>(d_evs, grid, cfg);
    else
        k_fit_tuples<PolicyFP64><<<ev_grid_dim(grid), kFitBlock, 0, st>>>(d_evs, grid, cfg);
    if (tm) tm->record("k_fit_tuples", st);
}

__global__ void k_fill_bytes(uint8_t *p, uint8_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
void launch_fill_bytes(void *dst, uint8_t value, size_t bytes, hipStream_t st) {
    const size_t blocks = std::min<size_t>((bytes + 255) / 256, 65535);
    k_fill_bytes<<<(unsigned)blocks, 256, 0, st>>>((uint8_t *)dst, value, bytes);
}

// Readback gather: block per segment (the defined rows of the events of a
// batch, packed for one device-to-host copy).  Whole 32-bit words when the
// segment's ends allow, the odd 16-bit word last.
__global__ void __launch_bounds__(256) k_pack_segments(const PackSeg *__restrict__ segs, char *dst) {
    const PackSeg sg = segs[blockIdx.x];
    uint16_t *d = reinterpret_cast<uint16_t *>(dst + sg.dst_off);
    const bool words = ((reinterpret_cast<uintptr_t>(sg.src) | reinterpret_cast<uintptr_t>(d)) & 3u) == 0u;
    const uint32_t n32 = words ? sg.n16 / 2u : 0u;
    for (uint32_t i = threadIdx.x; i < n32; i += blockDim.x)
        reinterpret_cast<uint32_t *>(d)[i] = reinterpret_cast<const uint32_t *>(sg.src)[i];
    for (uint32_t i = 2u * n32 + threadIdx.x; i < sg.n16; i += blockDim.x) d[i] = sg.src[i];
}
void launch_pack_segments(const PackSeg *d_segs, uint32_t n_segs, char *d_dst, hipStream_t st) {
    if (n_segs) k_pack_segments<<<n_segs, 256, 0, st>>>(d_segs, d_dst);
}

template <class J>
static void launch_job(const J &j, uint32_t n, const FitLaunch &L, hipStream_t st) {
    if (n == 0) return;
    if (L.schedule == PT_SCHED_PLAIN) {
        uint32_t blocks = (n + L.block - 1) / L.block;
        if (blocks > 65535u * 4u) blocks = 65535u * 4u;
        k_fit_plain<J><<<blocks, L.block, 0, st>>>(j);
    } else if (L.schedule == PT_SCHED_PER_LENGTH) {
        launch_fill_bytes(L.counts, 0, 2 * sizeof(uint32_t), st);
        uint32_t blocks = (n + 255u) / 256u;
        if (blocks > 4096u) blocks = 4096u;
        k_fit_partition<J><<<blocks, 256, 0, st>>>(j, L.lists, L.counts);
        uint32_t fb = (n + L.block - 1) / L.block;
        if (fb > 65535u * 4u) fb = 65535u * 4u;
        k_fit_list<J, 3><<<fb, L.block, 0, st>>>(j, L.lists, L.counts);
        k_fit_list<J, 4><<<fb, L.block, 0, st>>>(j, L.lists + n, L.counts + 1);
    } else {
        launch_fill_bytes(L.counts, 0, sizeof(uint32_t), st);
        const uint32_t waves = L.block / 64u;
        const size_t lds = (size_t)waves * 8u * L.tile;
        uint32_t blocks = L.blocks;
        const uint32_t need = (n + L.tile * waves - 1) / (L.tile * waves);
        if (blocks > need) blocks = need;
        if (blocks < 1) blocks = 1;
        k_fit_persistent<J><<<blocks, L.block, lds, st>>>(j, L.tile, L.counts);
    }
}

void launch_fit_flat(const FlatArgs &a, int precision, const FitLaunch &L, hipStream_t st) {
    if (precision == PT_FIT_RIEMANN)
        launch_job(RiemannJob{a}, a.n, L, st);
    else if (precision == PT_FIT_MIXED)
        launch_job(FlatJob<PolicyMixed>{a}, a.n, L, st);
    else
        launch_job(FlatJob<PolicyFP64>{a}, a.n, L, st);
}

void launch_fit_adjoint(const AdjArgs &a, int loss_mode, const FitLaunch &L, hipStream_t st) {
    if (loss_mode)
        launch_job(AdjJob<true>{a}, a.n, L, st);
    else
        launch_job(AdjJob<false>{a}, a.n, L, st);
}

// ------------------------------------------------------------------ fixed scratch
// Scratch of the fit-only entry point that allocates nothing
// (pt_fit_bl_sched_noalloc_async): kNoAllocCounters words of counters /
// tickets, then the per-length id lists (2 per track).  One copy per device,
// held until the process ends; every call on a device shares it.
constexpr uint32_t kNoAllocCounters = 64u;  // 256 bytes, keeps the lists 256-byte aligned
static __device__ __align__(256) uint32_t scratch[kNoAllocCounters + 2u * PT_NOALLOC_MAX_TRACKS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on `device` (the current device): looked up
// once per device with hipGetSymbolAddress and kept.  The array's name is a
// host-side stand-in and is never passed to a kernel.
int noalloc_scratch(int device, uint32_t **counts, uint32_t **lists) {
    if (device < 0 || device >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -2;
        g_scratch_ptr[device] = (uint32_t *)p;
    }
    if (counts) *counts = g_scratch_ptr[device];
    if (lists) *lists = g_scratch_ptr[device] + kNoAllocCounters;
    return 0;
}

// Global covariances of pre-gathered hits from local sigmas and module ids
// (SOARotation.jl:139-147, float): ge[(c*5+k)*n+t].
__global__ void __launch_bounds__(256) k_local_to_global(uint32_t n, const float *__restrict__ sig,
                                                         const int16_t *__restrict__ det,
                                                         const float *__restrict__ frames, float *__restrict__ ge) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 5u * n; i += gridDim.x * blockDim.x) {
        const uint32_t k = i / n, t = i - k * n;
        const float sr = sig[(size_t)(0 * 5 + k) * n + t], sz = sig[(size_t)(1 * 5 + k) * n + t];
        int d = det[(size_t)k * n + t];
        if (d < 0 || d >= (int)PT_N_MODULES) d = 0;
        float gl[6];
        to_global_err(frames + 9 * d, sr * sr, sz * sz, gl);
#pragma unroll
        for (int c = 0; c < 6; ++c) ge[(size_t)(c * 5 + k) * n + t] = gl[c];
    }
}

void launch_local_to_global(uint32_t n, const float *sig, const int16_t *det, const float *frames, float *ge,
                            hipStream_t st) {
    if (n == 0) return;
    uint32_t blocks = (5u * n + 255u) / 256u;
    if (blocks > 65535u) blocks = 65535u;
    k_local_to_global<<<blocks, 256, 0, st>>>(n, sig, det, frames, ge);
}

}  // namespace pt
