// oracle_fit.cpp - TEST INFRASTRUCTURE (the oracle).  C entry points of the
// broken-line fit restatement (oracle_fit.hpp): batch FP64 fit of pre-gathered
// tracks, per-track intermediates for the known-answer checks, an op-counting
// scalar for the flop model and a forward-mode dual number for gradient checks.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"
#include "oracle_fit.hpp"

namespace {

using oracle::TrackOut;

// ------------------------------------------------------------------ gathering
template <int N>
void gather(const float *hits, const float *ge, uint32_t n, uint32_t t, double (&h)[3][N], double (&g)[6][N]) {
    for (int k = 0; k < N; ++k) {
        for (int c = 0; c < 3; ++c) h[c][k] = hits[(size_t)(c * 5 + k) * n + t];
        for (int c = 0; c < 6; ++c) g[c][k] = ge[(size_t)(c * 5 + k) * n + t];
    }
}

template <int N> void fit_one(const float *hits, const float *ge, uint32_t n, uint32_t t, double B, double *o, uint32_t *fl) {
    double h[3][N], g[6][N];
    gather<N>(hits, ge, n, t, h, g);
    TrackOut<double> r;
    oracle::fit_track<double, N>(h, g, B, r);
    for (int k = 0; k < 5; ++k) o[k] = r.state[k];
    for (int k = 0; k < 15; ++k) o[5 + k] = r.cov15[k];
    o[20] = r.pt;
    o[21] = r.eta;
    o[22] = r.chi2;
    *fl = r.flags;
}

// ------------------------------------------------------------------ op counting
struct OpCounts {
    uint64_t add = 0, mul = 0, div = 0, sqrt = 0, trans = 0, cmp = 0;
};
thread_local OpCounts g_ops;

struct Counted {
    double v;
    Counted() : v(0) {}
    Counted(double x) : v(x) {}
};
inline Counted operator+(Counted a, Counted b) { ++g_ops.add; return a.v + b.v; }
inline Counted operator-(Counted a, Counted b) { ++g_ops.add; return a.v - b.v; }
inline Counted operator*(Counted a, Counted b) { ++g_ops.mul; return a.v * b.v; }
inline Counted operator/(Counted a, Counted b) { ++g_ops.div; return a.v / b.v; }
inline Counted operator-(Counted a) { return -a.v; }
inline Counted operator+(double a, Counted b) { ++g_ops.add; return a + b.v; }
inline Counted operator-(double a, Counted b) { ++g_ops.add; return a - b.v; }
inline Counted operator*(double a, Counted b) { ++g_ops.mul; return a * b.v; }
inline Counted operator/(double a, Counted b) { ++g_ops.div; return a / b.v; }
inline Counted operator+(Counted a, double b) { ++g_ops.add; return a.v + b; }
inline Counted operator-(Counted a, double b) { ++g_ops.add; return a.v - b; }
inline Counted operator*(Counted a, double b) { ++g_ops.mul; return a.v * b; }
inline Counted operator/(Counted a, double b) { ++g_ops.div; return a.v / b; }
inline bool operator<(Counted a, double b) { ++g_ops.cmp; return a.v < b; }
inline bool operator>(Counted a, double b) { ++g_ops.cmp; return a.v > b; }
inline bool operator<=(Counted a, double b) { ++g_ops.cmp; return a.v <= b; }
inline bool operator==(Counted a, double b) { ++g_ops.cmp; return a.v == b; }
inline bool operator<(Counted a, Counted b) { ++g_ops.cmp; return a.v < b.v; }
inline Counted sqrt(Counted a) { ++g_ops.sqrt; return std::sqrt(a.v); }
inline Counted abs(Counted a) { return std::fabs(a.v); }
inline Counted atan(Counted a) { ++g_ops.trans; return std::atan(a.v); }
inline Counted log(Counted a) { ++g_ops.trans; return std::log(a.v); }
inline Counted sin(Counted a) { ++g_ops.trans; return std::sin(a.v); }
inline Counted cos(Counted a) { ++g_ops.trans; return std::cos(a.v); }
inline Counted asinh(Counted a) { ++g_ops.trans; return std::asinh(a.v); }
inline bool isnan(Counted a) { return std::isnan(a.v); }

// ------------------------------------------------------------------ forward-mode dual
constexpr int kMaxPartials = 25;  // 5 inputs x 5 hits
struct Dual {
    double v;
    double d[kMaxPartials];
    Dual() : v(0) { std::memset(d, 0, sizeof(d)); }
    Dual(double x) : v(x) { std::memset(d, 0, sizeof(d)); }
};
inline Dual make_unary(const Dual &a, double val, double der) {
    Dual r(val);
    for (int i = 0; i < kMaxPartials; ++i) r.d[i] = der * a.d[i];
    return r;
}
inline Dual operator+(const Dual &a, const Dual &b) {
    Dual r(a.v + b.v);
    for (int i = 0; i < kMaxPartials; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
inline Dual operator-(const Dual &a, const Dual &b) {
    Dual r(a.v - b.v);
    for (int i = 0; i < kMaxPartials; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
inline Dual operator*(const Dual &a, const Dual &b) {
    Dual r(a.v * b.v);
    for (int i = 0; i < kMaxPartials; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
inline Dual operator/(const Dual &a, const Dual &b) {
    Dual r(a.v / b.v);
    const double ib = 1.0 / b.v;
    for (int i = 0; i < kMaxPartials; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
inline Dual operator-(const Dual &a) { return make_unary(a, -a.v, -1.0); }
inline Dual operator+(double a, const Dual &b) { return Dual(a) + b; }
inline Dual operator-(double a, const Dual &b) { return Dual(a) - b; }
inline Dual operator*(double a, const Dual &b) { return make_unary(b, a * b.v, a); }
inline Dual operator/(double a, const Dual &b) { return Dual(a) / b; }
inline Dual operator+(const Dual &a, double b) { return a + Dual(b); }
inline Dual operator-(const Dual &a, double b) { return a - Dual(b); }
inline Dual operator*(const Dual &a, double b) { return make_unary(a, a.v * b, b); }
inline Dual operator/(const Dual &a, double b) { return make_unary(a, a.v / b, 1.0 / b); }
inline bool operator<(const Dual &a, double b) { return a.v < b; }
inline bool operator>(const Dual &a, double b) { return a.v > b; }
inline bool operator<=(const Dual &a, double b) { return a.v <= b; }
inline bool operator==(const Dual &a, double b) { return a.v == b; }
inline bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
inline Dual sqrt(const Dual &a) {
    double s = std::sqrt(a.v);
    return make_unary(a, s, 0.5 / s);
}
inline Dual abs(const Dual &a) { return a.v < 0 ? -a : a; }
inline Dual atan(const Dual &a) { return make_unary(a, std::atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
inline Dual log(const Dual &a) { return make_unary(a, std::log(a.v), 1.0 / a.v); }
inline Dual sin(const Dual &a) { return make_unary(a, std::sin(a.v), std::cos(a.v)); }
inline Dual cos(const Dual &a) { return make_unary(a, std::cos(a.v), -std::sin(a.v)); }
inline Dual asinh(const Dual &a) { return make_unary(a, std::asinh(a.v), 1.0 / std::sqrt(1.0 + a.v * a.v)); }
inline bool isnan(const Dual &a) { return std::isnan(a.v); }

// global covariance from local variances (SOARotation.jl:139-147) in T
template <class T> void to_global_err_t(const float *r, T cxx, T cyy, T *gl) {
    const double xx = r[0], xy = r[1], xz = r[2], yx = r[3], yy = r[4], yz = r[5];
    const T cxy(0.0);
    gl[0] = xx * (xx * cxx + yx * cxy) + yx * (xx * cxy + yx * cyy);
    gl[1] = xx * (xy * cxx + yy * cxy) + yx * (xy * cxy + yy * cyy);
    gl[2] = xy * (xy * cxx + yy * cxy) + yy * (xy * cxy + yy * cyy);
    gl[3] = xx * (xz * cxx + yz * cxy) + yx * (xz * cxy + yz * cyy);
    gl[4] = xy * (xz * cxx + yz * cxy) + yy * (xz * cxy + yz * cyy);
    gl[5] = xz * (xz * cxx + yz * cxy) + yz * (xz * cxy + yz * cyy);
}

template <int N>
int grad_impl(const double *xyz, const double *sig, const float *rot, double B, const double *truth, double *value,
              double *grad) {
    Dual h[3][N];
    Dual g[6][N];
    for (int k = 0; k < N; ++k) {
        for (int c = 0; c < 3; ++c) {
            h[c][k] = Dual(xyz[3 * k + c]);
            h[c][k].d[5 * k + c] = 1.0;
        }
        Dual srphi(sig[2 * k + 0]);
        srphi.d[5 * k + 3] = 1.0;
        Dual sz(sig[2 * k + 1]);
        sz.d[5 * k + 4] = 1.0;
        Dual gl[6];
        to_global_err_t<Dual>(rot + 9 * k, srphi * srphi, sz * sz, gl);
        for (int c = 0; c < 6; ++c) g[c][k] = gl[c];
    }
    TrackOut<Dual> o;
    oracle::fit_track<Dual, N>(h, g, B, o);
    Dual L = o.line_chi2 + o.circle_chi2;
    if (truth) {
        for (int k = 0; k < 5; ++k) {
            Dual dlt = o.state[k] - truth[k];
            if (k == 0) dlt = dlt - 2.0 * oracle::kPi * std::rint(dlt.v / (2.0 * oracle::kPi));  // phi modulo 2 pi
            L = L + dlt * dlt;
        }
    }
    value[0] = L.v;
    for (int i = 0; i < 5 * N; ++i) grad[i] = L.d[i];
    return PT_OK;
}

template <int N> int detail_impl(const double *hits, const float *ge, double B, double *out, int cap) {
    double h[3][N], g[6][N];
    for (int k = 0; k < N; ++k) {
        for (int c = 0; c < 3; ++c) h[c][k] = hits[3 * k + c];
        for (int c = 0; c < 6; ++c) g[c][k] = ge[6 * k + c];
    }
    oracle::LineDetail<double, N> ld;
    oracle::CircleDetail<double, N> cd;
    oracle::Prepared<double, N> prep;
    TrackOut<double> o;
    oracle::fit_track<double, N>(h, g, B, o, &ld, &cd, &prep);
    std::vector<double> v;
    for (int k = 0; k < 4; ++k) v.push_back(o.fast[k]);
    v.push_back((double)prep.q);
    for (int k = 0; k < N; ++k) v.push_back(prep.radii[0][k]);
    for (int k = 0; k < N; ++k) v.push_back(prep.radii[1][k]);
    for (int k = 0; k < N; ++k) v.push_back(prep.s[k]);
    for (int k = 0; k < N; ++k) v.push_back(prep.S[k]);
    for (int k = 0; k < N; ++k) v.push_back(prep.Z[k]);
    for (int k = 0; k < N; ++k) v.push_back(prep.VarBeta[k]);
    for (int k = 0; k < N; ++k) v.push_back(ld.w[k]);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) v.push_back(ld.C[i][j]);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) v.push_back(ld.I[i][j]);
    for (int k = 0; k < N; ++k) v.push_back(ld.u[k]);
    for (int k = 0; k < N; ++k) v.push_back(cd.w[k]);
    for (int k = 0; k < N; ++k) v.push_back(cd.VarBeta[k]);
    for (int k = 0; k < N; ++k) v.push_back(cd.Z[k]);
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) v.push_back(cd.C[i][j]);
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) v.push_back(cd.I[i][j]);
    for (int k = 0; k <= N; ++k) v.push_back(cd.u[k]);
    for (int k = 0; k < 5; ++k) v.push_back(o.state[k]);
    for (int k = 0; k < 15; ++k) v.push_back(o.cov15[k]);
    v.push_back(o.pt);
    v.push_back(o.eta);
    v.push_back(o.chi2);
    v.push_back(o.line_chi2);
    v.push_back(o.circle_chi2);
    v.push_back((double)o.q);
    v.push_back((double)o.flags);
    if ((int)v.size() > cap) return PT_ERR_ARG;
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    return (int)v.size();
}

template <int N> void opcount_impl(const double *hits, const float *ge, double B, uint64_t *counts) {
    Counted h[3][N], g[6][N];
    for (int k = 0; k < N; ++k) {
        for (int c = 0; c < 3; ++c) h[c][k] = Counted(hits[3 * k + c]);
        for (int c = 0; c < 6; ++c) g[c][k] = Counted((double)ge[6 * k + c]);
    }
    g_ops = OpCounts();
    TrackOut<Counted> o;
    oracle::fit_track<Counted, N>(h, g, B, o);
    counts[0] = g_ops.add;
    counts[1] = g_ops.mul;
    counts[2] = g_ops.div;
    counts[3] = g_ops.sqrt;
    counts[4] = g_ops.trans;
    counts[5] = g_ops.cmp;
}

}  // namespace

extern "C" {

int oracle_fit_bl(uint32_t n, const float *hits, const float *ge, const uint8_t *nhit, double b_field, double *out64,
                  uint32_t *flags, int n_threads) {
    const double B = (double)(float)b_field;
    auto work = [&](uint32_t t0, uint32_t t1) {
        for (uint32_t t = t0; t < t1; ++t) {
            const int nh = nhit[t];
            if (nh == 3)
                fit_one<3>(hits, ge, n, t, B, out64 + 23 * (size_t)t, flags + t);
            else
                fit_one<4>(hits, ge, n, t, B, out64 + 23 * (size_t)t, flags + t);  // 4, and 5 fitted as 4
        }
    };
    if (n_threads <= 1) {
        work(0, n);
    } else {
        std::vector<std::thread> th;
        const uint32_t chunk = (n + n_threads - 1) / n_threads;
        for (int i = 0; i < n_threads; ++i) {
            uint32_t a = i * chunk, b = std::min<uint32_t>(n, a + chunk);
            if (a >= b) break;
            th.emplace_back(work, a, b);
        }
        for (auto &x : th) x.join();
    }
    return PT_OK;
}

void oracle_plant_ms(double scale, int log10_for_ln, int drop_cap) {
    oracle::ms_plant().scale = scale;
    oracle::ms_plant().log10 = log10_for_ln != 0;
    oracle::ms_plant().no_cap = drop_cap != 0;
}

int oracle_bl_detail(int N, const double *hits, const float *ge, double B, double *out, int cap) {
    if (N == 3) return detail_impl<3>(hits, ge, B, out, cap);
    if (N == 4) return detail_impl<4>(hits, ge, B, out, cap);
    return PT_ERR_ARG;
}

int oracle_fit_opcount(int N, const double *hits, const float *ge, double B, uint64_t *counts) {
    if (N == 3) {
        opcount_impl<3>(hits, ge, B, counts);
        return PT_OK;
    }
    if (N == 4) {
        opcount_impl<4>(hits, ge, B, counts);
        return PT_OK;
    }
    return PT_ERR_ARG;
}

int oracle_fit_grad(int N, const double *xyz, const double *sig, const float *rot, double B, const double *truth,
                    double *value, double *grad) {
    if (N == 3) return grad_impl<3>(xyz, sig, rot, B, truth, value, grad);
    if (N == 4) return grad_impl<4>(xyz, sig, rot, B, truth, value, grad);
    return PT_ERR_ARG;
}

// Forward-mode loss gradients of a fit-only batch (layout of
// pt_fit_bl_loss_grad_async): hits[(c*5+k)*n+t], sig[(c*5+k)*n+t] (local
// sigmas), det[k*n+t], nhit[t] (5 fitted as 4), truth[k*n+t] (may be NULL),
// rot[9*m] module frames; loss[t] and grad[(c*5+k)*n+t], c = x, y, z,
// sigma_rphi, sigma_z.  Used as the CPU baseline of the adjoint workload.
int oracle_fit_grad_many(uint32_t n, const float *hits, const float *sig, const int16_t *det, const uint8_t *nhit,
                         const float *rot, double b_field, const float *truth, double *loss, double *grad,
                         int n_threads) {
    const double B = (double)(float)b_field;
    auto work = [&](uint32_t t0, uint32_t t1) {
        for (uint32_t t = t0; t < t1; ++t) {
            const int N = nhit[t] == 3 ? 3 : 4;
            double xyz[15], sg[10], tr[5], g[20] = {0}, v = 0.0;
            float R[36];
            for (int k = 0; k < N; ++k) {
                for (int c = 0; c < 3; ++c) xyz[3 * k + c] = hits[(size_t)(c * 5 + k) * n + t];
                for (int c = 0; c < 2; ++c) sg[2 * k + c] = sig[(size_t)(c * 5 + k) * n + t];
                int d = det[(size_t)k * n + t];
                if (d < 0 || d >= (int)PT_N_MODULES) d = 0;
                for (int c = 0; c < 9; ++c) R[9 * k + c] = rot[9 * d + c];
            }
            if (truth)
                for (int k = 0; k < 5; ++k) tr[k] = truth[(size_t)k * n + t];
            if (N == 3)
                grad_impl<3>(xyz, sg, R, B, truth ? tr : nullptr, &v, g);
            else
                grad_impl<4>(xyz, sg, R, B, truth ? tr : nullptr, &v, g);
            loss[t] = v;
            for (int k = 0; k < 5; ++k)
                for (int c = 0; c < 5; ++c) grad[(size_t)(c * 5 + k) * n + t] = (k < N) ? g[5 * k + c] : 0.0;
        }
    };
    if (n_threads <= 1) {
        work(0, n);
    } else {
        std::vector<std::thread> th;
        const uint32_t chunk = (n + n_threads - 1) / n_threads;
        for (int i = 0; i < n_threads; ++i) {
            uint32_t a = i * chunk, b = std::min<uint32_t>(n, a + chunk);
            if (a >= b) break;
            th.emplace_back(work, a, b);
        }
        for (auto &x : th) x.join();
    }
    return PT_OK;
}

}  // extern "C"
