// oracle_riemann.hpp - TEST INFRASTRUCTURE (the oracle), never linked into the product.
//
// Serial restatement of the Riemann helix fit of
//   src/julia-serial/plugin-PixelTriplets/riemann_fit.jl
//     :31-38    compute_rad_len_uniform_material
//     :59-97    scatter_cov_line
//     :112-150  scatter_cov_rad
//     :158-180  cov_radtocart
//     :226-258  cov_carttorad_prefit
//     :264-266  weight_circle
//     :275-277  Charge
//     :292-298  min_eigen3D
//     :338-364  atan2 (x < 0 -> atan(y/x) + pi), Float64 and Float32 methods
//     :378-418  fast_fit
//     :445-823  circle_fit (with error propagation)
//     :841-955  line_fit
//     :989-1023 helix_fit
//   fit_utils.jl:23-45 loadCovariance2D, :89-113 par_uvrtopak, :115-141 fromCircleToPerigee
//   cholesky_inversion.jl:8-19 invert22 (invert33/44 from oracle_fit.hpp)
//   TrajectoryStateSOA.jl:20-44 copyFromCircle! (the TrackSOA output)
//
// The reference never runs this fit: make_tuples always calls the broken-line
// fit (ca_hit_ntuplet_generator.jl:83-93), launch_riemann_kernels! only prints
// (helix_fit_on_gpu.jl:94-96), and riemann_fit.jl is not included by the
// module (src/Patatrack.jl).  Several of its statements cannot run or read
// values that are never computed.  The restatement follows the algorithm the
// file documents (the Riemann-Chernov circle fit on the paraboloid, the s-z
// line fit) and changes exactly these statements, marked R1..R9 below:
//   R1 :513-514   the plane offset c is read from a freshly zeroed 1x1 matrix;
//                 here c = -v.r0, the offset of the fitted plane
//   R2 :549-793   later temporaries overwrite that scalar c (Julia has no block
//                 scope); J3 uses the plane offset
//   R3 :799-809   J3 is allocated undef and four entries are never set; zero here
//   R4 :581       t01 = p3D[1,:]' .* p3D[2,:] gives x_j*y_i, the line's own
//                 comment (p3D.row(0)' * p3D.row(1)) says x_i*y_j; the comment is followed
//   R5 :756-761   the sign fix negates the whole J2 when one column needs it;
//                 here only column a
//   R6 :814       (par*par') * Jq * V * Jq' multiplies 3x3 by 1x2N
//                 (DimensionMismatch); Jq*V*Jq' is the scalar variance of the scale
//   R7 :67-68     scatter_cov_line reads fast_fit[2], fast_fit[3] (Y0, R) for
//                 (R, tan theta), the 0-based indices of the original; scatter_cov_rad
//                 (:114-115) reads the 1-based ones, as here
//   R8 :928-952   hcat of a row and a vector; sol(2,1) call syntax; line_par
//                 undefined: A = [1...1; s'], sol[2], results stored
//   R9 :996,1005  vecnorm and par_urvtopak are undefined (par_uvrtopak meant);
//                 fit_utils.jl:97 J4[1,1] uses X0 where d(phi)/dX0 needs Y0;
//                 fit_utils.jl:120-122 atan2 is undefined there (the two-argument
//                 arctangent meant) and fromCircleToPerigee's third parameter is
//                 R*B with B undefined while its Jacobian (:135) is that of q/R
// Everything else is restated as written, including the Float32 arc lengths of
// scatter_cov_rad (:131-134), the custom atan2 quadrants and the second-order
// terms of C[3][3] (:591-592).
#pragma once
#include <cmath>
#include <cstdint>

#include "oracle_fit.hpp"

namespace oracle {
namespace riemann {

constexpr double kXX0Inv = 0.06 / 16.0;  // :33, one radiation length = 16 cm / 0.06
constexpr double kEps = 0.0001;          // :8, relative step of the numerical Jacobian J2

// Planted mistakes (tests only): each reverts one correction to the statement
// as the reference file writes it, or drops a term, to show a pin catches it.
enum : uint32_t {
    kPlantZeroC = 1u,        // R1 as written: c = 0
    kPlantJ3OverwrittenC = 2u,  // R2 as written: J3 uses Cvc[4,4]
    kPlantT01AsWritten = 4u,    // R4 as written
    kPlantLineMsIndices = 8u,   // R7 as written: p_t from Y0, p_2 from R
    kPlantNoScaleVar = 16u,     // R6 term dropped
    kPlantNoMs = 32u,           // scatter_cov_rad returns zero
};
inline uint32_t &plant() {
    static uint32_t p = 0;
    return p;
}

template <int N> struct CircleOut {
    double par[3];      // (X0, Y0, R)
    double cov[3][3];
    int q;
    double chi2;
    // intermediates for the known-answer checks
    double weight[N];
    double G[N][N];       // inverse radial covariance, normalised by its sum
    double p3D[3][N];     // centred, scaled points on the paraboloid
    double A[3][3];
    double v[3];          // plane normal (v[2] > 0)
    double c;             // plane offset
    double r0[3];
    double s;             // scale
    double h_[2];         // centre of the unweighted mean
    double lambda;        // smallest eigenvalue of A
    double E[6][6];
    double J2[3][6];
    double Cvc[4][4];
    double var_scale;     // Jq V Jq'
    double V[2 * N][2 * N];  // hit covariance plus scattering, Cartesian (x block, y block)
};

template <int N> struct LineOut {
    double par[2];  // (cot theta, Zip)
    double cov[2][2];
    double chi2;
    // intermediates for the known-answer checks
    double theta;
    double jx[N][5];      // d s_i / d(X0, Y0, R, x_i, y_i)
    double pr[2][N];      // points in the rotated s-z frame
    double cov_ms[N][N];  // covariance of the rotated z (hits, circle, scattering)
    double sol[2];
};

struct HelixOut {
    double par[5];  // (phi, Tip, pT, cot theta, Zip)
    double cov[5][5];
    int q;
    double chi2_circle, chi2_line;
};

// riemann_fit.jl:352-364, the Float32 method: atan(y/x) in Float32, + Float32(pi)
inline float atan2_f32(float y, float x) {
    if (x > 0.0f) return std::atan(y / x);
    if (x < 0.0f) return std::atan(y / x) + float(kPi);
    if (y > 0.0f && x == 0.0f) return float(kPi / 2.0);
    if (y < 0.0f && x == 0.0f) return float(-kPi / 2.0);
    return std::numeric_limits<float>::quiet_NaN();
}

inline double sqr(double x) { return x * x; }

// riemann_fit.jl:31-38
template <int N> inline void rad_len_uniform(const double (&len)[N], double (&rl)[N]) {
    rl[0] = len[0] * kXX0Inv;
    for (int j = 1; j < N; ++j) rl[j] = std::abs(len[j] - len[j - 1]) * kXX0Inv;
}

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (Rutishauser's
// rotation form); columns of Q are the eigenvectors.  The reference calls
// LAPACK through eigen(Symmetric(A)) (:293); any accurate solver gives the
// same eigenpair up to its sign, which circle_fit fixes (:502-507).
inline void sym_eigen3(const double (&Ain)[3][3], double (&w)[3], double (&Q)[3][3]) {
    double a[3][3];  // Symmetric(A): the upper triangle
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            a[i][j] = i <= j ? Ain[i][j] : Ain[j][i];
            Q[i][j] = i == j ? 1.0 : 0.0;
        }
    static const int P[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 100; ++sweep) {
        if (a[0][1] == 0.0 && a[0][2] == 0.0 && a[1][2] == 0.0) break;
        for (const auto &pq : P) {
            const int p = pq[0], q = pq[1], r = 3 - p - q;
            const double apq = a[p][q];
            if (apq == 0.0) continue;
            const double g = 100.0 * std::abs(apq);
            if (std::abs(a[p][p]) + g == std::abs(a[p][p]) && std::abs(a[q][q]) + g == std::abs(a[q][q])) {
                a[p][q] = a[q][p] = 0.0;
                continue;
            }
            const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            double t;
            if (std::abs(theta) > 1e150)
                t = 0.5 / theta;
            else
                t = (theta >= 0.0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
            const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c, tau = s / (1.0 + c);
            const double h = t * apq;
            a[p][p] -= h;
            a[q][q] += h;
            a[p][q] = a[q][p] = 0.0;
            const double arp = a[r][p], arq = a[r][q];
            a[r][p] = a[p][r] = arp - s * (arq + arp * tau);
            a[r][q] = a[q][r] = arq + s * (arp - arq * tau);
            for (int k = 0; k < 3; ++k) {
                const double qp = Q[k][p], qq = Q[k][q];
                Q[k][p] = qp - s * (qq + qp * tau);
                Q[k][q] = qq + s * (qp - qq * tau);
            }
        }
    }
    for (int i = 0; i < 3; ++i) w[i] = a[i][i];
}

// riemann_fit.jl:292-298: the eigenvector of the smallest eigenvalue, negated,
// and that eigenvalue
inline void min_eigen3D(const double (&A)[3][3], double (&vec)[3], double &lambda) {
    double w[3], Q[3][3];
    sym_eigen3(A, w, Q);
    int m = 0;
    for (int i = 1; i < 3; ++i)
        if (w[i] < w[m]) m = i;
    lambda = w[m];
    for (int k = 0; k < 3; ++k) vec[k] = -Q[k][m];
}

// cholesky_inversion.jl:8-19 (+ symmetrize22!)
inline void invert22(const double (&s)[2][2], double (&d)[2][2]) {
    const double luc0 = 1.0 / s[0][0];
    const double luc1 = s[1][0] * s[1][0] * luc0;
    const double luc2 = 1.0 / (s[1][1] - luc1);
    const double li21 = luc1 * luc0 * luc2;
    d[0][0] = li21 + luc0;
    d[1][0] = -s[1][0] * luc0 * luc2;
    d[1][1] = luc2;
    d[0][1] = d[1][0];
}

// riemann_fit.jl:378-418 ; hits[c][k], c = x,y,z ; res = (X0, Y0, R, tan theta)
template <int N> inline void fast_fit(const double (&hits)[3][N], double (&res)[4]) {
    constexpr int mid = (N + 2) / 2;  // ceil(n/2) with n = N + 1 (:380,384), 1-based
    const double b0 = hits[0][mid - 1] - hits[0][0], b1 = hits[1][mid - 1] - hits[1][0];
    const double c0 = hits[0][N - 1] - hits[0][0], c1 = hits[1][N - 1] - hits[1][0];
    const double b2 = b0 * b0 + b1 * b1;
    const double c2 = c0 * c0 + c1 * c1;
    const bool flip = std::abs(b0) < std::abs(b1);
    const double bx = flip ? b1 : b0, by = flip ? b0 : b1;
    const double cx = flip ? c1 : c0, cy = flip ? c0 : c1;
    const double div = 2 * (cx * by - bx * cy);
    const double Y0 = (cx * b2 - bx * c2) / div;
    const double X0 = (0.5 * b2 - Y0 * by) / bx;
    res[0] = hits[0][0] + (flip ? Y0 : X0);
    res[1] = hits[1][0] + (flip ? X0 : Y0);
    res[2] = std::sqrt((X0 * X0) + (Y0 * Y0));
    const double d0 = hits[0][0] - res[0], d1 = hits[1][0] - res[1];
    const double e0 = hits[0][N - 1] - res[0], e1 = hits[1][N - 1] - res[1];
    const double dr = res[2] * atan2_ref(d0 * e1 - d1 * e0, d0 * e0 + d1 * e1);
    const double dz = hits[2][N - 1] - hits[2][0];
    res[3] = dr / dz;
}

// riemann_fit.jl:112-150
template <int N>
inline void scatter_cov_rad(const double (&p2D)[2][N], const double (&ff)[4], const double (&rad)[N], double B,
                            double (&out)[N][N]) {
    const double p_t = jl_min(20.0, ff[2] * B);
    const double p_2 = p_t * p_t * (1.0 + 1.0 / (ff[3] * ff[3]));
    double theta = std::atan(ff[3]);
    theta = theta < 0.0 ? theta + kPi : theta;
    double s_values[N], len[N], rl[N], sig2[N];
    const double o0 = ff[0], o1 = ff[1];
    for (int i = 0; i < N; ++i) {
        const double p0 = p2D[0][i] - o0, p1 = p2D[1][i] - o1;
        const float cross = float((-o0) * p1 - (-o1) * p0);
        const float dot = float((-o0) * p0 + (-o1) * p1);
        const float a = atan2_f32(cross, dot);
        s_values[i] = std::abs(double(a) * ff[2]);
    }
    const double fac = std::sqrt(1.0 + 1.0 / (ff[3] * ff[3]));
    for (int i = 0; i < N; ++i) len[i] = s_values[i] * fac;
    rad_len_uniform<N>(len, rl);
    const double den = p_2 * sqr(std::sin(theta));
    for (int i = 0; i < N; ++i) {
        const double t = 1.0 + 0.038 * std::log(rl[i]);
        sig2[i] = (t * t) * rl[i];
        sig2[i] = sig2[i] * 0.000225 / den;
    }
    for (int k = 0; k < N; ++k)
        for (int l = 0; l < N; ++l) out[k][l] = 0.0;
    for (int k = 0; k < N; ++k)
        for (int l = k; l < N; ++l) {
            for (int i = 0; i <= (k < l ? k : l); ++i) out[k][l] += (rad[k] - rad[i]) * (rad[l] - rad[i]) * sig2[i];
            out[l][k] = out[k][l];
        }
    if (plant() & kPlantNoMs)
        for (int k = 0; k < N; ++k)
            for (int l = 0; l < N; ++l) out[k][l] = 0.0;
}

// riemann_fit.jl:158-180
template <int N>
inline void cov_radtocart(const double (&p2D)[2][N], const double (&cr)[N][N], const double (&rad)[N],
                          double (&cc)[2 * N][2 * N]) {
    double ri[N];
    for (int i = 0; i < N; ++i) ri[i] = 1.0 / rad[i];
    for (int i = 0; i < 2 * N; ++i)
        for (int j = 0; j < 2 * N; ++j) cc[i][j] = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = i; j < N; ++j) {
            cc[i][j] = cr[i][j] * p2D[1][i] * ri[i] * p2D[1][j] * ri[j];
            cc[i + N][j + N] = cr[i][j] * p2D[0][i] * ri[i] * p2D[0][j] * ri[j];
            cc[i][j + N] = -cr[i][j] * p2D[1][i] * ri[i] * p2D[0][j] * ri[j];
            cc[i + N][j] = -cr[i][j] * p2D[0][i] * ri[i] * p2D[1][j] * ri[j];
            cc[j][i] = cc[i][j];
            cc[j + N][i + N] = cc[i + N][j + N];
            cc[j + N][i] = cc[i][j + N];
            cc[j][i + N] = cc[i + N][j];
        }
}

// riemann_fit.jl:226-258 (diagonal only)
template <int N>
inline void cov_carttorad_prefit(const double (&p2D)[2][N], const double (&V)[2 * N][2 * N], const double (&ff)[4],
                                 const double (&rad)[N], double (&cr)[N][N]) {
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) cr[i][j] = 0.0;
    for (int i = 0; i < N; ++i) {
        if (rad[i] < 0.0001) {
            cr[i][i] = V[i][i];
        } else {
            const double a0 = p2D[0][i], a1 = p2D[1][i];
            const double b0 = a0 - ff[0], b1 = a1 - ff[1];
            const double x2 = a0 * b0 + a1 * b1;
            const double y2 = a0 * b1 - a1 * b0;
            const double tan_c = -y2 / x2;
            const double tan_c2 = tan_c * tan_c;
            cr[i][i] = 1.0 / (1.0 + tan_c2) * (V[i][i] + V[i + N][i + N] * tan_c2 + 2 * V[i][i + N] * tan_c);
        }
    }
}

// riemann_fit.jl:445-823 ; hits2D[c][k] ; V = loadCovariance2D of the hits (fit_utils.jl:23-45)
template <int N>
inline void circle_fit(const double (&hits2D)[2][N], const double (&Vin)[2 * N][2 * N], const double (&ff)[4],
                       const double (&rad)[N], double B, CircleOut<N> &o) {
    const uint32_t pl = plant();
    double V[2 * N][2 * N];
    double cov_rad[N][N], ms[N][N], cc[2 * N][2 * N];
    cov_carttorad_prefit<N>(hits2D, Vin, ff, rad, cov_rad);
    scatter_cov_rad<N>(hits2D, ff, rad, B, ms);
    cov_radtocart<N>(hits2D, ms, rad, cc);
    for (int i = 0; i < 2 * N; ++i)
        for (int j = 0; j < 2 * N; ++j) V[i][j] = Vin[i][j] + cc[i][j];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) cov_rad[i][j] = cov_rad[i][j] + ms[i][j];

    double G[N][N];
    invert(cov_rad, G);
    double renorm = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) renorm += G[i][j];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) G[i][j] = G[i][j] / renorm;
    double w[N];
    for (int j = 0; j < N; ++j) {
        w[j] = 0.0;
        for (int i = 0; i < N; ++i) w[j] += G[i][j];
    }

    double h_[2] = {0.0, 0.0};
    for (int i = 0; i < N; ++i) {
        h_[0] += hits2D[0][i];
        h_[1] += hits2D[1][i];
    }
    h_[0] = h_[0] / N;
    h_[1] = h_[1] / N;
    double p3D[3][N], mc[2 * N];
    for (int i = 0; i < N; ++i) {
        p3D[0][i] = hits2D[0][i] - h_[0];
        p3D[1][i] = hits2D[1][i] - h_[1];
        mc[i] = p3D[0][i];
        mc[i + N] = p3D[1][i];
    }
    double nrm = 0.0;
    for (int i = 0; i < 2 * N; ++i) nrm += mc[i] * mc[i];
    nrm = std::sqrt(nrm);
    const double q = nrm * nrm;  // norm(mc)^2 (:484)
    const double s = std::sqrt(N / q);
    for (int i = 0; i < N; ++i) {
        p3D[0][i] = p3D[0][i] * s;
        p3D[1][i] = p3D[1][i] * s;
        p3D[2][i] = p3D[0][i] * p3D[0][i] + p3D[1][i] * p3D[1][i];
    }

    double r0[3], X[3][N], A[3][3];
    for (int k = 0; k < 3; ++k) {
        r0[k] = 0.0;
        for (int j = 0; j < N; ++j) r0[k] += p3D[k][j] * w[j];
    }
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < N; ++j) X[k][j] = p3D[k][j] - r0[k];
    {
        double XG[3][N];
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < N; ++j) {
                XG[k][j] = 0.0;
                for (int i = 0; i < N; ++i) XG[k][j] += X[k][i] * G[i][j];
            }
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) {
                A[k][l] = 0.0;
                for (int j = 0; j < N; ++j) A[k][l] += XG[k][j] * X[l][j];
            }
    }

    double v[3], lambda;
    min_eigen3D(A, v, lambda);
    if (!(v[2] > 0.0))
        for (int k = 0; k < 3; ++k) v[k] = -v[k];

    // R1: the plane offset
    double c = -(v[0] * r0[0] + v[1] * r0[1] + v[2] * r0[2]);
    if (pl & kPlantZeroC) c = 0.0;
    const double h = std::sqrt(1 - sqr(v[2]) - 4 * c * v[2]);
    const double v2x2_inv = 1 / (2 * v[2]);
    const double s_inv = 1 / s;
    const double par_uvr[3] = {-v[0] * v2x2_inv, -v[1] * v2x2_inv, h * v2x2_inv};
    o.par[0] = par_uvr[0] * s_inv + h_[0];
    o.par[1] = par_uvr[1] * s_inv + h_[1];
    o.par[2] = par_uvr[2] * s_inv;
    // Charge (:275-277)
    o.q = (hits2D[0][1] - hits2D[0][0]) * (o.par[1] - hits2D[1][0]) -
                      (hits2D[1][1] - hits2D[1][0]) * (o.par[0] - hits2D[0][0]) >
                  0
              ? -1
              : 1;
    o.chi2 = std::abs(lambda) * renorm * 1 / sqr(2 * v[2] * par_uvr[2] * s);

    // ---------------------------------------------------------------- errors
    double cmc = 0.0, normV2 = 0.0;
    for (int a = 0; a < 2 * N; ++a)
        for (int b = 0; b < 2 * N; ++b) {
            cmc += mc[a] * V[a][b] * mc[b];
            normV2 += V[a][b] * V[a][b];
        }
    normV2 = sqr(std::sqrt(normV2));
    const double kap = sqr(sqr(s)) * 1.0 / (4.0 * q * N) * (2.0 * normV2 + 4.0 * cmc);
    double Vcs[2 * N][2 * N];
    for (int a = 0; a < 2 * N; ++a)
        for (int b = 0; b < 2 * N; ++b) Vcs[a][b] = sqr(s) * V[a][b] + kap * (mc[a] * mc[b]);

    // blocks C[i][j] (i <= j) of the covariance of the paraboloid points (:557-603)
    double C[3][3][N][N];
    const double *px = p3D[0], *py = p3D[1];
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            const int lo = a < b ? a : b, hi = a < b ? b : a;
            const double c00 = Vcs[lo][hi], c11 = Vcs[N + lo][N + hi];
            const double c01 = Vcs[a][N + b], c10 = Vcs[b][N + a];
            C[0][0][a][b] = c00;
            C[1][1][a][b] = c11;
            C[0][1][a][b] = c01;
            C[0][2][a][b] = 2 * ((c00 * px[b]) + (c01 * py[b]));
            C[1][2][a][b] = 2 * (c10 * px[b] + c11 * py[b]);
        }
    for (int a = 0; a < N; ++a)
        for (int b = a; b < N; ++b) {
            const double c00 = Vcs[a][b], c11 = Vcs[N + a][N + b];
            const double c01 = Vcs[a][N + b], c10 = Vcs[b][N + a];
            const double t00 = px[b] * px[a], t11 = py[b] * py[a];
            double t01 = px[a] * py[b], t10 = px[b] * py[a];
            if (pl & kPlantT01AsWritten) {
                t01 = px[b] * py[a];
                t10 = px[a] * py[b];
            }
            const double tmp = 2.0 * (c00 * c00 + c00 * c01 + c11 * c10 + c11 * c11) +
                               4.0 * (c00 * t00 + c01 * t01 + c10 * t10 + c11 * t11);
            C[2][2][a][b] = tmp;
            C[2][2][b][a] = tmp;
        }

    // C0 (:611-622)
    double C0[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
            double t = 0.0;
            for (int a = 0; a < N; ++a) {
                double u = 0.0;
                for (int b = 0; b < N; ++b) u += C[i][j][a][b] * w[b];
                t += w[a] * u;
            }
            C0[i][j] = t;
            C0[j][i] = t;
        }

    // H, s_v, D (:624-660)
    double H[N][N], s_v[N][3];
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) H[a][b] = (a == b ? 1.0 : 0.0) - w[b];
    for (int a = 0; a < N; ++a)
        for (int k = 0; k < 3; ++k) {
            s_v[a][k] = 0.0;
            for (int b = 0; b < N; ++b) s_v[a][k] += H[a][b] * p3D[k][b];
        }
    double D[3][3][N][N];
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
            double HC[N][N];
            for (int a = 0; a < N; ++a)
                for (int b = 0; b < N; ++b) {
                    HC[a][b] = 0.0;
                    for (int k = 0; k < N; ++k) HC[a][b] += H[a][k] * C[i][j][k][b];
                }
            for (int a = 0; a < N; ++a)
                for (int b = 0; b < N; ++b) {
                    double t = 0.0;
                    for (int k = 0; k < N; ++k) t += HC[a][k] * H[b][k];
                    D[i][j][a][b] = t * (w[a] * w[b]);
                }
            if (i != j)
                for (int a = 0; a < N; ++a)
                    for (int b = 0; b < N; ++b) D[j][i][b][a] = D[i][j][a][b];
        }

    // E (:662-742)
    static const int nu[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    double E[6][6];
    auto Dmv = [&](int x, int y, int col, double (&out)[N], double f) {
        for (int a = 0; a < N; ++a) {
            double t = 0.0;
            for (int b = 0; b < N; ++b) t += D[x][y][a][b] * s_v[b][col];
            out[a] = f * t;
        }
    };
    auto dotc = [&](int col, const double (&u)[N]) {
        double t = 0.0;
        for (int a = 0; a < N; ++a) t += s_v[a][col] * u[a];
        return t;
    };
    for (int a = 0; a < 6; ++a) {
        const int i = nu[a][0], j = nu[a][1];
        for (int b = a; b < 6; ++b) {
            const int k = nu[b][0], l = nu[b][1];
            double t0[N], t1[N], u[N];
            if (l == k) {
                Dmv(j, l, l, t0, 2.0);
                if (i == j)
                    for (int z = 0; z < N; ++z) t1[z] = t0[z];
                else
                    Dmv(i, l, l, t1, 2.0);
            } else {
                Dmv(j, l, k, t0, 1.0);
                Dmv(j, k, l, u, 1.0);
                for (int z = 0; z < N; ++z) t0[z] = t0[z] + u[z];
                if (i == j) {
                    for (int z = 0; z < N; ++z) t1[z] = t0[z];
                } else {
                    Dmv(i, l, k, t1, 1.0);
                    Dmv(i, k, l, u, 1.0);
                    for (int z = 0; z < N; ++z) t1[z] = t1[z] + u[z];
                }
            }
            double e;
            if (i == j) {
                double sum[N];
                for (int z = 0; z < N; ++z) sum[z] = t0[z] + t1[z];
                e = dotc(i, sum);
            } else {
                e = dotc(i, t0) + dotc(j, t1);
            }
            E[a][b] = e;
            E[b][a] = e;
        }
    }

    // J2: numerical Jacobian of the eigenvector (:746-775), R5
    double J2[3][6];
    for (int a = 0; a < 6; ++a) {
        const int i = nu[a][0], j = nu[a][1];
        double Ap[3][3];
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) Ap[x][y] = A[x][y];
        const double delta = std::abs(A[i][j] * kEps);
        Ap[i][j] = A[i][j] + delta;
        Ap[j][i] = Ap[i][j];
        double vec[3], lam;
        min_eigen3D(Ap, vec, lam);
        if (!(vec[2] > 0.0))
            for (int k = 0; k < 3; ++k) vec[k] = -vec[k];
        for (int k = 0; k < 3; ++k) J2[k][a] = (vec[k] * 1.0 - v[k]) / delta;
    }

    // Cvc (:779-795)
    double t0m[3][3];
    {
        double JE[3][6];
        for (int x = 0; x < 3; ++x)
            for (int b = 0; b < 6; ++b) {
                JE[x][b] = 0.0;
                for (int a = 0; a < 6; ++a) JE[x][b] += J2[x][a] * E[a][b];
            }
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) {
                t0m[x][y] = 0.0;
                for (int b = 0; b < 6; ++b) t0m[x][y] += JE[x][b] * J2[y][b];
            }
    }
    double Cvc[4][4];
    for (int x = 0; x < 3; ++x) {
        double t = 0.0;
        for (int y = 0; y < 3; ++y) {
            Cvc[x][y] = t0m[x][y];
            t += (-t0m[x][y]) * r0[y];
        }
        Cvc[x][3] = t;
        Cvc[3][x] = t;
    }
    {
        double cm1 = 0.0, cm3 = 0.0, hs = 0.0;
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) {
                cm1 += v[x] * C0[x][y] * v[y];
                hs += C0[x][y] * t0m[x][y];
                cm3 += r0[x] * t0m[x][y] * r0[y];
            }
        Cvc[3][3] = cm1 + hs + cm3;
    }

    // J3 (:797-809), R2 + R3
    const double cJ = (pl & kPlantJ3OverwrittenC) ? Cvc[3][3] : c;
    const double t = 1 / h;
    double J3[3][4] = {{-v2x2_inv, 0.0, v[0] * sqr(v2x2_inv) * 2.0, 0.0},
                       {0.0, -v2x2_inv, v[1] * sqr(v2x2_inv) * 2.0, 0.0},
                       {v[0] * v2x2_inv * t, v[1] * v2x2_inv * t,
                        -h * sqr(v2x2_inv) * 2.0 - (2.0 * cJ + v[2]) * v2x2_inv * t, -t}};

    // Jq, R6 (:811-814)
    double Jq[2 * N];
    for (int a = 0; a < 2 * N; ++a) Jq[a] = mc[a] * s * 1 / N;
    double varq = 0.0;
    for (int a = 0; a < 2 * N; ++a) {
        double u = 0.0;
        for (int b = 0; b < 2 * N; ++b) u += Jq[b] * V[b][a];
        varq += u * Jq[a];
    }
    if (pl & kPlantNoScaleVar) varq = 0.0;
    {
        double JC[3][4];
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 4; ++y) {
                JC[x][y] = 0.0;
                for (int z = 0; z < 4; ++z) JC[x][y] += J3[x][z] * Cvc[z][y];
            }
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) {
                double u = 0.0;
                for (int z = 0; z < 4; ++z) u += JC[x][z] * J3[y][z];
                o.cov[x][y] = u * sqr(s_inv) + (par_uvr[x] * par_uvr[y]) * varq;
            }
    }

    for (int i = 0; i < N; ++i) {
        o.weight[i] = w[i];
        for (int j = 0; j < N; ++j) o.G[i][j] = G[i][j];
        for (int k = 0; k < 3; ++k) o.p3D[k][i] = p3D[k][i];
    }
    for (int x = 0; x < 3; ++x) {
        o.v[x] = v[x];
        o.r0[x] = r0[x];
        for (int y = 0; y < 3; ++y) o.A[x][y] = A[x][y];
        for (int y = 0; y < 6; ++y) o.J2[x][y] = J2[x][y];
    }
    for (int x = 0; x < 6; ++x)
        for (int y = 0; y < 6; ++y) o.E[x][y] = E[x][y];
    for (int x = 0; x < 4; ++x)
        for (int y = 0; y < 4; ++y) o.Cvc[x][y] = Cvc[x][y];
    o.c = c;
    o.s = s;
    o.h_[0] = h_[0];
    o.h_[1] = h_[1];
    o.lambda = lambda;
    o.var_scale = varq;
    for (int a = 0; a < 2 * N; ++a)
        for (int b = 0; b < 2 * N; ++b) o.V[a][b] = V[a][b];
}

// riemann_fit.jl:59-97, R7
template <int N>
inline void scatter_cov_line(const double (&cov_sz)[N][2][2], const double (&ff)[4], const double (&s_arcs)[N],
                             const double (&z)[N], double B, double (&ret)[N][N]) {
    const bool as_written = plant() & kPlantLineMsIndices;
    const double p_t = jl_min(20.0, (as_written ? ff[1] : ff[2]) * B);
    const double tq = as_written ? ff[2] : ff[3];
    const double p_2 = p_t * p_t * (1.0 + 1.0 / (tq * tq));
    double S[N], rl[N], sig2[N];
    for (int i = 0; i < N; ++i) S[i] = std::sqrt(s_arcs[i] * s_arcs[i] + z[i] * z[i]);
    rad_len_uniform<N>(S, rl);
    for (int i = 0; i < N; ++i) {
        const double t = 1.0 + 0.038 * std::log(rl[i]);
        sig2[i] = 0.000225 / p_2 * (t * t) * rl[i];
    }
    for (int k = 0; k < N; ++k)
        for (int l = 0; l < N; ++l) ret[k][l] = k == l ? cov_sz[k][1][1] : 0.0;
    for (int k = 0; k < N; ++k)
        for (int l = k; l < N; ++l) {
            for (int i = 0; i <= (k < l ? k : l); ++i)
                ret[k][l] += std::abs(S[k] - S[i]) * std::abs(S[l] - S[i]) * sig2[i];
            ret[l][k] = ret[k][l];
        }
}

// riemann_fit.jl:841-955 ; ge[c][k] = (xx, xy, yy, xz, yz, zz), R8
template <int N>
inline void line_fit(const double (&hits)[3][N], const double (&ge)[6][N], const CircleOut<N> &circle,
                     const double (&ff)[4], double B, LineOut<N> &o) {
    const double q = circle.q;
    double theta = -q * std::atan(ff[3]);
    theta = theta < 0 ? theta + kPi : theta;
    const double sn = std::sin(theta), cs = std::cos(theta);
    const double rot[2][2] = {{sn, cs}, {-cs, sn}};
    const double o0 = circle.par[0], o1 = circle.par[1], R = circle.par[2];
    double p2D[2][N], cov_sz[N][2][2];
    double Cov[6][6] = {};
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) Cov[x][y] = circle.cov[x][y];
    for (int i = 0; i < N; ++i) {
        const double p0 = hits[0][i] - o0, p1 = hits[1][i] - o1;
        const double cross = (-o0) * p1 - (-o1) * p0;
        const double dot = (-o0) * p0 + (-o1) * p1;
        const double atan2_ = -q * atan2_ref(cross, dot);
        p2D[0][i] = atan2_ * R;
        const double temp0 = -q * R * 1 / (sqr(dot) + sqr(cross));
        const double d_X0 = -temp0 * ((p1 + o1) * dot - (p0 - o0) * cross);
        const double d_Y0 = temp0 * ((p0 + o0) * dot - (o1 - p1) * cross);
        const double d_R = atan2_;
        const double d_x = temp0 * (o1 * dot + o0 * cross);
        const double d_y = temp0 * (-o0 * dot + o1 * cross);
        const double Jx[2][6] = {{d_X0, d_Y0, d_R, d_x, d_y, 0.0}, {0.0, 0.0, 0.0, 0.0, 0.0, 1.0}};
        for (int k = 0; k < 5; ++k) o.jx[i][k] = Jx[0][k];
        Cov[3][3] = ge[0][i];
        Cov[4][4] = ge[2][i];
        Cov[5][5] = ge[5][i];
        Cov[3][4] = Cov[4][3] = ge[1][i];
        Cov[3][5] = Cov[5][3] = ge[3][i];
        Cov[4][5] = Cov[5][4] = ge[4][i];
        double JC[2][6], tmp[2][2], rt[2][2];
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 6; ++y) {
                JC[x][y] = 0.0;
                for (int z = 0; z < 6; ++z) JC[x][y] += Jx[x][z] * Cov[z][y];
            }
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y) {
                tmp[x][y] = 0.0;
                for (int z = 0; z < 6; ++z) tmp[x][y] += JC[x][z] * Jx[y][z];
            }
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y) rt[x][y] = rot[x][0] * tmp[0][y] + rot[x][1] * tmp[1][y];
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y) cov_sz[i][x][y] = rt[x][0] * rot[y][0] + rt[x][1] * rot[y][1];
    }
    for (int i = 0; i < N; ++i) p2D[1][i] = hits[2][i];

    double cov_ms[N][N];
    scatter_cov_line<N>(cov_sz, ff, p2D[0], p2D[1], B, cov_ms);

    double pr[2][N];
    for (int i = 0; i < N; ++i) {
        pr[0][i] = rot[0][0] * p2D[0][i] + rot[0][1] * p2D[1][i];
        pr[1][i] = rot[1][0] * p2D[0][i] + rot[1][1] * p2D[1][i];
    }
    double A[2][N];
    for (int i = 0; i < N; ++i) {
        A[0][i] = 1.0;
        A[1][i] = pr[0][i];
    }
    double Vy[N][N];
    invert(cov_ms, Vy);
    double AV[2][N];
    for (int x = 0; x < 2; ++x)
        for (int j = 0; j < N; ++j) {
            AV[x][j] = 0.0;
            for (int i = 0; i < N; ++i) AV[x][j] += A[x][i] * Vy[i][j];
        }
    double Cp[2][2], Cpi[2][2];
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) {
            Cp[x][y] = 0.0;
            for (int j = 0; j < N; ++j) Cp[x][y] += AV[x][j] * A[y][j];
        }
    invert22(Cp, Cpi);
    double sol[2];
    for (int x = 0; x < 2; ++x) {
        double u = 0.0;
        for (int j = 0; j < N; ++j) {
            const double cav = Cpi[x][0] * AV[0][j] + Cpi[x][1] * AV[1][j];
            u += cav * pr[1][j];
        }
        sol[x] = u;
    }
    const double cf = 1 / (sn - sol[1] * cs);
    const double J[2][2] = {{0.0, cf * cf}, {cf, sol[0] * cs * cf * cf}};
    o.par[0] = cf * (sol[1] * sn + cs);
    o.par[1] = cf * sol[0];
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) {
            double u = 0.0;
            for (int z = 0; z < 2; ++z) {
                const double jc = J[x][0] * Cpi[0][z] + J[x][1] * Cpi[1][z];
                u += jc * J[y][z];
            }
            o.cov[x][y] = u;
        }
    double res[N];
    for (int i = 0; i < N; ++i) res[i] = pr[1][i] - (A[0][i] * sol[0] + A[1][i] * sol[1]);
    double chi2 = 0.0;
    for (int i = 0; i < N; ++i) {
        double u = 0.0;
        for (int j = 0; j < N; ++j) u += res[j] * Vy[j][i];
        chi2 += u * res[i];
    }
    o.chi2 = chi2;
    o.theta = theta;
    for (int i = 0; i < N; ++i) {
        o.pr[0][i] = pr[0][i];
        o.pr[1][i] = pr[1][i];
        for (int j = 0; j < N; ++j) o.cov_ms[i][j] = cov_ms[i][j];
    }
    o.sol[0] = sol[0];
    o.sol[1] = sol[1];
}

// fit_utils.jl:89-113, R9 ; circle par (X0, Y0, R) -> (phi, Tip, pT)
template <int N> inline void par_uvrtopak(const CircleOut<N> &c, double B, double (&par)[3], double (&cov)[3][3]) {
    const double X0 = c.par[0], Y0 = c.par[1], R = c.par[2], q = c.q;
    const double temp0 = X0 * X0 + Y0 * Y0;
    const double temp1 = std::sqrt(temp0);
    par[0] = std::atan2(q * X0, -q * Y0);
    par[1] = q * (temp1 - R);
    par[2] = R * B;
    const double temp2 = sqr(X0) * 1.0 / temp0;
    const double temp3 = 1.0 / temp1 * q;
    const double J4[3][3] = {{-Y0 * temp2 / (X0 * X0), temp2 / X0, 0.0}, {X0 * temp3, Y0 * temp3, -q}, {0.0, 0.0, B}};
    sandwich(J4, c.cov, cov);
}

// fit_utils.jl:115-141, R9 ; circle par (X0, Y0, R) -> (phi, Tip, q/R)
template <int N> inline void circle_to_perigee(const CircleOut<N> &c, double (&par)[3], double (&cov)[3][3]) {
    const double X0 = c.par[0], Y0 = c.par[1], R = c.par[2], q = c.q;
    const double temp0 = X0 * X0 + Y0 * Y0;
    const double temp1 = std::sqrt(temp0);
    par[0] = std::atan2(q * X0, -q * Y0);
    par[1] = q * (temp1 - R);
    par[2] = q / R;
    const double temp2 = sqr(X0) * 1.0 / temp0;
    const double temp3 = 1.0 / temp1 * q;
    const double J4[3][3] = {{-Y0 * temp2 / (X0 * X0), temp2 / X0, 0.0},
                             {X0 * temp3, Y0 * temp3, -q},
                             {0.0, 0.0, -q / (R * R)}};
    sandwich(J4, c.cov, cov);
}

template <int N> struct RiemannTrack {
    double fast[4];
    CircleOut<N> circle;  // (X0, Y0, R) form, as line_fit receives it
    LineOut<N> line;
    HelixOut helix;       // helix_fit's result (par_uvrtopak form)
    // TrackSOA convention (copyFromCircle!, TrajectoryStateSOA.jl:20-44)
    double state[5];
    double cov15[15];
    double pt, eta, chi2;
};

// riemann_fit.jl:989-1023 + the TrackSOA output of the perigee form ; hits[c][k], ge[c][k]
template <int N> inline void helix_fit(const double (&hits)[3][N], const double (&ge)[6][N], double B, RiemannTrack<N> &o) {
    double rad[N];
    for (int i = 0; i < N; ++i) rad[i] = std::sqrt(hits[0][i] * hits[0][i] + hits[1][i] * hits[1][i]);
    fast_fit<N>(hits, o.fast);
    double hc[2 * N][2 * N];
    for (int i = 0; i < 2 * N; ++i)
        for (int j = 0; j < 2 * N; ++j) hc[i][j] = 0.0;
    for (int i = 0; i < N; ++i) {  // loadCovariance2D (fit_utils.jl:23-45)
        hc[i][i] = ge[0][i];
        hc[i + N][i + N] = ge[2][i];
        hc[i][i + N] = hc[i + N][i] = ge[1][i];
    }
    double h2[2][N];
    for (int i = 0; i < N; ++i) {
        h2[0][i] = hits[0][i];
        h2[1][i] = hits[1][i];
    }
    circle_fit<N>(h2, hc, o.fast, rad, B, o.circle);
    line_fit<N>(hits, ge, o.circle, o.fast, B, o.line);

    double pp[3], pc[3][3];
    par_uvrtopak<N>(o.circle, B, pp, pc);
    for (int x = 0; x < 5; ++x)
        for (int y = 0; y < 5; ++y) o.helix.cov[x][y] = 0.0;
    for (int x = 0; x < 3; ++x) {
        o.helix.par[x] = pp[x];
        for (int y = 0; y < 3; ++y) o.helix.cov[x][y] = pc[x][y];
    }
    for (int x = 0; x < 2; ++x) {
        o.helix.par[3 + x] = o.line.par[x];
        for (int y = 0; y < 2; ++y) o.helix.cov[3 + x][3 + y] = o.line.cov[x][y];
    }
    o.helix.q = o.circle.q;
    o.helix.chi2_circle = o.circle.chi2;
    o.helix.chi2_line = o.line.chi2;

    double kp[3], kc[3][3];
    circle_to_perigee<N>(o.circle, kp, kc);
    const double b = 1.0 / B;
    o.state[0] = kp[0];
    o.state[1] = kp[1];
    o.state[2] = kp[2] * b;
    o.state[3] = o.line.par[0];
    o.state[4] = o.line.par[1];
    o.cov15[0] = kc[0][0];
    o.cov15[1] = kc[0][1];
    o.cov15[2] = b * kc[0][2];
    o.cov15[3] = 0.0;
    o.cov15[4] = 0.0;
    o.cov15[5] = kc[1][1];
    o.cov15[6] = b * kc[1][2];
    o.cov15[7] = 0.0;
    o.cov15[8] = 0.0;
    o.cov15[9] = (b * b) * kc[2][2];
    o.cov15[10] = 0.0;
    o.cov15[11] = 0.0;
    o.cov15[12] = o.line.cov[0][0];
    o.cov15[13] = o.line.cov[0][1];
    o.cov15[14] = o.line.cov[1][1];
    o.pt = B / std::abs(kp[2]);
    o.eta = std::asinh(o.line.par[0]);
    o.chi2 = (o.circle.chi2 + o.line.chi2) / double(2 * N - 5);
}

}  // namespace riemann
}  // namespace oracle
