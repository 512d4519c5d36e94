// oracle_riemann.cpp - TEST INFRASTRUCTURE (the oracle).  C entry points of the
// Riemann helix fit restatement (oracle_riemann.hpp): batch fit of pre-gathered
// tracks in the TrackSOA convention, and one track with every intermediate.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "oracle.h"
#include "oracle_riemann.hpp"

namespace {

using oracle::riemann::RiemannTrack;

template <int N> void gather(const float *hits, const float *ge, uint32_t n, uint32_t t, double (&h)[3][N], double (&g)[6][N]) {
    for (int k = 0; k < N; ++k) {
        for (int c = 0; c < 3; ++c) h[c][k] = hits[(size_t)(c * 5 + k) * n + t];
        for (int c = 0; c < 6; ++c) g[c][k] = ge[(size_t)(c * 5 + k) * n + t];
    }
}

template <int N>
void fit_one(const float *hits, const float *ge, uint32_t n, uint32_t t, double B, double scale, double *o) {
    double h[3][N], g[6][N];
    gather<N>(hits, ge, n, t, h, g);
    if (scale != 1.0)
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < N; ++k) h[c][k] = h[c][k] * scale;
    RiemannTrack<N> r;
    oracle::riemann::helix_fit<N>(h, g, B, r);
    for (int k = 0; k < 5; ++k) o[k] = r.state[k];
    for (int k = 0; k < 15; ++k) o[5 + k] = r.cov15[k];
    o[20] = r.pt;
    o[21] = r.eta;
    o[22] = r.chi2;
}

template <int N> int detail(const double *hits, const float *ge, double B, double *out, int cap) {
    double h[3][N], g[6][N];
    for (int k = 0; k < N; ++k) {
        for (int c = 0; c < 3; ++c) h[c][k] = hits[3 * k + c];
        for (int c = 0; c < 6; ++c) g[c][k] = ge[6 * k + c];
    }
    RiemannTrack<N> r;
    oracle::riemann::helix_fit<N>(h, g, B, r);
    std::vector<double> v;
    auto put = [&](const double *p, int k) { v.insert(v.end(), p, p + k); };
    const auto &c = r.circle;
    put(r.fast, 4);
    put(c.par, 3);
    put(&c.cov[0][0], 9);
    v.push_back(c.q);
    v.push_back(c.chi2);
    put(r.line.par, 2);
    put(&r.line.cov[0][0], 4);
    v.push_back(r.line.chi2);
    put(r.helix.par, 5);
    put(&r.helix.cov[0][0], 25);
    put(r.state, 5);
    put(r.cov15, 15);
    v.push_back(r.pt);
    v.push_back(r.eta);
    v.push_back(r.chi2);
    put(c.weight, N);
    put(&c.G[0][0], N * N);
    put(&c.p3D[0][0], 3 * N);
    put(&c.A[0][0], 9);
    put(c.v, 3);
    v.push_back(c.c);
    put(c.r0, 3);
    v.push_back(c.s);
    put(c.h_, 2);
    v.push_back(c.lambda);
    put(&c.E[0][0], 36);
    put(&c.J2[0][0], 18);
    put(&c.Cvc[0][0], 16);
    v.push_back(c.var_scale);
    put(&c.V[0][0], 4 * N * N);
    v.push_back(r.line.theta);
    put(&r.line.jx[0][0], 5 * N);
    put(&r.line.pr[0][0], 2 * N);
    put(&r.line.cov_ms[0][0], N * N);
    put(r.line.sol, 2);
    if ((int)v.size() > cap) return -1;
    std::copy(v.begin(), v.end(), out);
    return (int)v.size();
}

}  // namespace

extern "C" {

int oracle_fit_riemann(uint32_t n, const float *hits, const float *ge, const uint8_t *nhit, double b_field,
                       double *out64, int n_threads) {
    return oracle_fit_riemann_scaled(n, hits, ge, nhit, b_field, 1.0, out64, n_threads);
}

int oracle_fit_riemann_scaled(uint32_t n, const float *hits, const float *ge, const uint8_t *nhit, double b_field,
                              double scale, double *out64, int n_threads) {
    const double B = (double)(float)b_field;
    auto work = [&](uint32_t t0, uint32_t t1) {
        for (uint32_t t = t0; t < t1; ++t) {
            if (nhit[t] == 3)
                fit_one<3>(hits, ge, n, t, B, scale, out64 + 23 * (size_t)t);
            else
                fit_one<4>(hits, ge, n, t, B, scale, out64 + 23 * (size_t)t);  // 4, and 5 fitted as 4
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

int oracle_riemann_detail(int N, const double *hits, const float *ge, double B, double *out, int cap) {
    switch (N) {
    case 3: return detail<3>(hits, ge, B, out, cap);
    case 4: return detail<4>(hits, ge, B, out, cap);
    case 5: return detail<5>(hits, ge, B, out, cap);
    default: return -1;
    }
}

void oracle_plant_riemann(uint32_t mask) { oracle::riemann::plant() = mask; }

int oracle_sym_eigen3(const double *A, double *w, double *Q) {
    double a[3][3], ww[3], q[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = A[3 * i + j];
    oracle::riemann::sym_eigen3(a, ww, q);
    for (int i = 0; i < 3; ++i) {
        w[i] = ww[i];
        for (int j = 0; j < 3; ++j) Q[3 * i + j] = q[i][j];
    }
    return 0;
}

}  // extern "C"
