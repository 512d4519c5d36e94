// pt_fit.hip - broken-line fit kernels for gfx950.
//
// k_fit_tuples: the fit stage of make_tuples (launchBrokenLineKernelsOnCPU,
//   broken_line_fit_on_gpu.jl:166-203): for each length bucket (3 -> N=3,
//   4 -> N=4, 5 -> N=4 on the first four hits, fit5as4) fit the first 24,576
//   entries of the tuple-multiplicity bucket, i.e. the highest tuple ids.
//   Gather per hit: global position, module frame, local variances rotated to
//   the global covariance in float (SOARotation.jl:139-147).
// k_fit_flat: fit of pre-gathered tracks (fit-only workloads).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pt_bl_fit.hpp"
#include "pt_device.hpp"

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
template <int N>
__device__ __forceinline__ void fit_bucket(const EvDesc &e, const Config &cfg, int bucket, uint32_t idx0,
                                           uint32_t stride, uint32_t &ref, uint32_t &nfit) {
    const uint32_t b0 = e.s.tm_off[bucket - 1], b1 = e.s.tm_off[bucket];
    uint32_t n = b1 - b0;
    if (n > PT_MAX_FITS) n = PT_MAX_FITS;
    for (uint32_t idx = idx0; idx < n; idx += stride) {
        const uint32_t tkid = e.s.tm_bins[b0 + idx];  // 1-based
        const uint32_t start = e.hit_off[tkid - 1] - 1u;
        double hx[N], hy[N], hz[N];
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
        FitOut<double> o;
        bl_fit_fp64<N>(hx, hy, hz, ge, cfg.B, o);
        ref |= fit_flag_bits(o.flags);
        const uint32_t row = tkid - 1u;
#pragma unroll
        for (int k = 0; k < 5; ++k) e.state[k * kMaxTuples + row] = (float)o.state[k];
#pragma unroll
        for (int k = 0; k < 15; ++k) e.cov[k * kMaxTuples + row] = (float)o.cov[k];
        e.pt[row] = (float)o.pt;
        e.eta[row] = (float)o.eta;
        e.chi2[row] = (float)o.chi2;
        ++nfit;
    }
}

// All three length buckets in one launch (3 -> N=3, 4 -> N=4, 5 -> N=4 on the
// first four hits when fit_5_as_4): the event's blocks are split in three.
__global__ void __launch_bounds__(256) k_fit_tuples(const EvDesc *__restrict__ evs, EvGrid grid, Config cfg) {
    uint32_t ev, b;
    if (!ev_block(grid, ev, b)) return;
    const EvDesc &e = evs[ev];
    const uint32_t per = grid.per_event / 3u;
    const uint32_t bk = b / per, bb = b - bk * per;
    uint32_t ref = 0, nfit = 0;
    if (bk == 0)
        fit_bucket<3>(e, cfg, 3, bb * blockDim.x + threadIdx.x, per * blockDim.x, ref, nfit);
    else if (bk == 1)
        fit_bucket<4>(e, cfg, 4, bb * blockDim.x + threadIdx.x, per * blockDim.x, ref, nfit);
    else if (bk == 2 && cfg.p.fit_5_as_4)
        fit_bucket<4>(e, cfg, 5, bb * blockDim.x + threadIdx.x, per * blockDim.x, ref, nfit);
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

// Pre-gathered tracks: hits[(c*5+k)*n+t], ge[(c*5+k)*n+t]; nhit 3 -> N=3, 4/5 -> N=4.
template <int N>
__device__ __forceinline__ void fit_flat_one(uint32_t n, uint32_t t, const float *__restrict__ hits,
                                             const float *__restrict__ ge, double B, float *__restrict__ state,
                                             float *__restrict__ cov, float *__restrict__ pt, float *__restrict__ eta,
                                             float *__restrict__ chi2) {
    double hx[N], hy[N], hz[N];
    float g[6][N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        hx[k] = hits[(size_t)(0 * 5 + k) * n + t];
        hy[k] = hits[(size_t)(1 * 5 + k) * n + t];
        hz[k] = hits[(size_t)(2 * 5 + k) * n + t];
#pragma unroll
        for (int c = 0; c < 6; ++c) g[c][k] = ge[(size_t)(c * 5 + k) * n + t];
    }
    FitOut<double> o;
    bl_fit_fp64<N>(hx, hy, hz, g, B, o);
#pragma unroll
    for (int k = 0; k < 5; ++k) state[(size_t)k * n + t] = (float)o.state[k];
#pragma unroll
    for (int k = 0; k < 15; ++k) cov[(size_t)k * n + t] = (float)o.cov[k];
    pt[t] = (float)o.pt;
    eta[t] = (float)o.eta;
    chi2[t] = (float)o.chi2;
}

__global__ void __launch_bounds__(256) k_fit_flat_fp64(uint32_t n, const float *__restrict__ hits,
                                                       const float *__restrict__ ge, const uint8_t *__restrict__ nhit,
                                                       double B, float *state, float *cov, float *pt, float *eta,
                                                       float *chi2) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        if (nhit[t] == 3)
            fit_flat_one<3>(n, t, hits, ge, B, state, cov, pt, eta, chi2);
        else
            fit_flat_one<4>(n, t, hits, ge, B, state, cov, pt, eta, chi2);
    }
}

// ------------------------------------------------------------------ launchers
void launch_fit_tuples(const EvDesc *d_evs, uint32_t n_events, const Config &cfg, hipStream_t st, Marks *tm) {
    uint32_t per = (PT_MAX_FITS + 255) / 256;
    const uint32_t target = (2048u + n_events - 1) / n_events / 3u;
    if (per > target) per = target;
    if (per < 1) per = 1;
    const EvGrid grid{n_events, 3u * per};
    k_fit_tuples<<<ev_grid_dim(grid), 256, 0, st>>>(d_evs, grid, cfg);
    if (tm) tm->record("k_fit_tuples", st);
}

__global__ void k_fill_bytes(uint8_t *p, uint8_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
void launch_fill_bytes(void *dst, uint8_t value, size_t bytes, hipStream_t st) {
    const size_t blocks = std::min<size_t>((bytes + 255) / 256, 65535);
    k_fill_bytes<<<(unsigned)blocks, 256, 0, st>>>((uint8_t *)dst, value, bytes);
}

void launch_fit_flat(uint32_t n, const float *hits, const float *ge, const uint8_t *nhit, double B, int precision,
                     float *state, float *cov, float *pt, float *eta, float *chi2, hipStream_t st) {
    (void)precision;
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 65535u * 4u) blocks = 65535u * 4u;
    if (blocks == 0) return;
    k_fit_flat_fp64<<<blocks, 256, 0, st>>>(n, hits, ge, nhit, B, state, cov, pt, eta, chi2);
}

}  // namespace pt
