// pt_riemann.hpp - Riemann helix fit, device code (one track per lane, FP64).
//
// Method of the reference's Riemann fit (src/julia-serial/plugin-PixelTriplets/
// riemann_fit.jl: fast_fit :378-418, circle_fit :445-823, line_fit :841-955,
// helix_fit :989-1023; fit_utils.jl:115-141 for the perigee form and
// TrajectoryStateSOA.jl:20-44 for the TrackSOA output): the hits are lifted
// onto the paraboloid z = x^2 + y^2 and a plane is fitted to them (weighted
// by the inverse radial covariance, multiple scattering included); the plane
// gives the circle; its covariance comes from the covariance of the lifted
// points through the eigenvector Jacobian (a forward difference with a
// relative step of 1e-4, as the reference computes it); the line fit is a
// weighted least-squares line in the s-z plane rotated by the polar angle.
// Where riemann_fit.jl cannot run or reads values that are never computed,
// the fit follows the method it documents:
//   * the plane offset is c = -v.r0 (the file reads a freshly zeroed matrix,
//     :513-514) and the circle Jacobian J3 uses it (:549-809 overwrite the
//     scalar); J3's unset entries are zero;
//   * t01 is x_i*y_j as the line's comment says (:581);
//   * the eigenvector sign fix applies to one column of J2 (:756-761);
//   * (par*par')*(Jq*V*Jq') is a matrix times the scalar variance of the
//     scale (:814);
//   * scatter_cov_line takes R and tan(theta) from fast_fit (:67-68 index it
//     0-based);
//   * the line fit's design matrix is [1...1; s'] and its solution is stored
//     (:928-952); the perigee form is (phi, Tip, q/R) (fit_utils.jl:115-141).
//
// GPU formulation (same method, different evaluation): nothing of size
// (2N)^2 or 3x3 blocks of N x N is stored.  The hit covariance V (hits plus
// the scattering term) is formed entry by entry from per-hit values; the
// covariance of the plane-fit scatter (E, :662-742) is accumulated directly as
// bilinear forms y_i' C_jl y_k with y_k = w .* (p3D[k,:] - r0[k]), which equal
// the reference's s_v' * ((H C H') .* (w w')) * s_v because H' y = y when the
// weights sum to one; the perturbed eigen-solves of J2 start from the
// eigenbasis of the unperturbed matrix (two Jacobi sweeps reach convergence
// from there).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_bl_fit.hpp"

namespace pt {
namespace riemann {

constexpr double kXX0Inv = 0.06 / 16.0;  // one radiation length = 16 cm / 0.06 (riemann_fit.jl:33)
constexpr double kEps = 0.0001;          // relative step of the eigenvector Jacobian (riemann_fit.jl:8)

PT_HD double sqr(double x) { return x * x; }

// One Jacobi rotation of the symmetric 3x3 a (upper and lower kept equal) in
// the (p, q) plane, accumulated into the eigenvector columns of Q.
PT_HD void jacobi_rot(double (&a)[3][3], double (&Q)[3][3], int p, int q) {
    const double apq = a[p][q];
    if (apq == 0.0) return;
    const double g = 100.0 * fabs(apq);
    if (fabs(a[p][p]) + g == fabs(a[p][p]) && fabs(a[q][q]) + g == fabs(a[q][q])) {
        a[p][q] = a[q][p] = 0.0;
        return;
    }
    const int r = 3 - p - q;
    const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
    double t;
    if (fabs(theta) > 1e150)
        t = 0.5 / theta;
    else
        t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c, tau = s / (1.0 + c);
    const double h = t * apq;
    a[p][p] -= h;
    a[q][q] += h;
    a[p][q] = a[q][p] = 0.0;
    const double arp = a[r][p], arq = a[r][q];
    a[r][p] = a[p][r] = arp - s * (arq + arp * tau);
    a[r][q] = a[q][r] = arq + s * (arp - arq * tau);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double qp = Q[k][p], qq = Q[k][q];
        Q[k][p] = qp - s * (qq + qp * tau);
        Q[k][q] = qq + s * (qp - qq * tau);
    }
}

// Cyclic Jacobi to convergence (at most `sweeps`); a is overwritten by the
// eigenvalues on its diagonal.
PT_HD void jacobi3(double (&a)[3][3], double (&Q)[3][3], int sweeps) {
    for (int sw = 0; sw < sweeps; ++sw) {
        if (a[0][1] == 0.0 && a[0][2] == 0.0 && a[1][2] == 0.0) break;
        jacobi_rot(a, Q, 0, 1);
        jacobi_rot(a, Q, 0, 2);
        jacobi_rot(a, Q, 1, 2);
    }
}

PT_HD int argmin3(double a, double b, double c) {
    int m = 0;
    double v = a;
    if (b < v) {
        m = 1;
        v = b;
    }
    if (c < v) m = 2;
    return m;
}

// riemann_fit.jl:378-418 ; (X0, Y0, R, tan theta)
template <int N> PT_HD void fast_fit(const double (&x)[N], const double (&y)[N], const double (&z)[N], double (&ff)[4]) {
    constexpr int mid = (N + 2) / 2 - 1;
    const double b0 = x[mid] - x[0], b1 = y[mid] - y[0];
    const double c0 = x[N - 1] - x[0], c1 = y[N - 1] - y[0];
    const double b2 = b0 * b0 + b1 * b1;
    const double c2 = c0 * c0 + c1 * c1;
    const bool flip = fabs(b0) < fabs(b1);
    const double bx = flip ? b1 : b0, by = flip ? b0 : b1;
    const double cx = flip ? c1 : c0, cy = flip ? c0 : c1;
    const double div = 2 * (cx * by - bx * cy);
    const double Y0 = (cx * b2 - bx * c2) / div;
    const double X0 = (0.5 * b2 - Y0 * by) / bx;
    ff[0] = x[0] + (flip ? Y0 : X0);
    ff[1] = y[0] + (flip ? X0 : Y0);
    ff[2] = sqrt((X0 * X0) + (Y0 * Y0));
    const double d0 = x[0] - ff[0], d1 = y[0] - ff[1];
    const double e0 = x[N - 1] - ff[0], e1 = y[N - 1] - ff[1];
    ff[3] = ff[2] * atan2_bl(d0 * e1 - d1 * e0, d0 * e0 + d1 * e1) / (z[N - 1] - z[0]);
}

// riemann_fit.jl:31-38 in place
template <int N> PT_HD void rad_len_uniform(double (&len)[N]) {
    double prev = len[0];
    len[0] = len[0] * kXX0Inv;
#pragma unroll
    for (int j = 1; j < N; ++j) {
        const double cur = len[j];
        len[j] = fabs(cur - prev) * kXX0Inv;
        prev = cur;
    }
}

template <int N> struct RiemannOut {
    double state[5];
    double cov15[15];
    double pt, eta, chi2;
};

// The line fit (riemann_fit.jl:841-955), the perigee form (fit_utils.jl:115-141)
// and the TrackSOA output (TrajectoryStateSOA.jl:20-44), given the circle
// (X0, Y0, R, q) and its covariance.  Called out of line: inlined into the
// circle fit, the per-length N = 4 kernel of the fit-only schedules gave line
// results (cot(theta), zip, their covariance, chi2) different from the same
// code in the plain and persistent kernels and from the oracle, with the
// circle bit-identical.
template <int N>
__host__ __device__ __attribute__((noinline)) void riemann_line_and_output(
    const float (&x)[N], const float (&y)[N], const float (&z)[N], const float (&ge)[6][N], const double (&ff)[4],
    double X0, double Y0, double Rc, int q, const double (&ccov)[3][3], double chi2_circle, double B,
    RiemannOut<N> &out) {
    auto X = [&](int i) -> double { return (double)x[i]; };
    auto Y = [&](int i) -> double { return (double)y[i]; };
    auto Z = [&](int i) -> double { return (double)z[i]; };
    auto GE = [&](int c, int i) -> double { return (double)ge[c][i]; };
    // ---------------------------------------------------------------- line fit (:841-955)
    double theta = -q * atan(ff[3]);
    theta = theta < 0 ? theta + kPiD : theta;
    const double sn = sin(theta), cs = cos(theta);
    double sa[N], czz[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double p0 = X(i) - X0, p1 = Y(i) - Y0;
        const double cross = (-X0) * p1 - (-Y0) * p0;
        const double dot = (-X0) * p0 + (-Y0) * p1;
        const double at = -q * atan2_bl(cross, dot);
        sa[i] = at * Rc;
        const double temp0 = -q * Rc * 1 / (sqr(dot) + sqr(cross));
        const double jc[3] = {-temp0 * ((p1 + Y0) * dot - (p0 - X0) * cross),
                              temp0 * ((p0 + X0) * dot - (Y0 - p1) * cross), at};
        const double dx = temp0 * (Y0 * dot + X0 * cross);
        const double dy = temp0 * (-X0 * dot + Y0 * cross);
        // Jx Cov Jx' with Cov = diag(circle cov, hit cov), then the z'z' entry of rot (.) rot'
        double t00 = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n2 = 0; n2 < 3; ++n2) t00 += jc[m] * ccov[m][n2] * jc[n2];
        t00 += dx * (GE(0, i) * dx + GE(1, i) * dy) + dy * (GE(1, i) * dx + GE(2, i) * dy);
        const double t01 = dx * GE(3, i) + dy * GE(4, i);
        const double t11 = GE(5, i);
        czz[i] = cs * cs * t00 - 2.0 * cs * sn * t01 + sn * sn * t11;
    }
    double cm[N][N];
    {
        const double p_t = jl_min(20.0, ff[2] * B);
        const double p_2 = p_t * p_t * (1.0 + 1.0 / (ff[3] * ff[3]));
        double S[N], rl[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            S[i] = sqrt(sa[i] * sa[i] + Z(i) * Z(i));
            rl[i] = S[i];
        }
        rad_len_uniform<N>(rl);
        double sig2[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double tt = 1.0 + 0.038 * log(rl[i]);
            sig2[i] = 0.000225 / p_2 * (tt * tt) * rl[i];
        }
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int l = k; l < N; ++l) {
                double acc = k == l ? czz[k] : 0.0;
#pragma unroll
                for (int i = 0; i <= (k < l ? k : l); ++i) acc += fabs(S[k] - S[i]) * fabs(S[l] - S[i]) * sig2[i];
                cm[k][l] = acc;
                cm[l][k] = acc;
            }
    }
    double xr[N], yr[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        xr[i] = sn * sa[i] + cs * Z(i);
        yr[i] = -cs * sa[i] + sn * Z(i);
    }
    double Vy[N][N];
    invert_n(cm, Vy);
    double av0[N], av1[N];  // rows of A * Vy
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            a0 += Vy[i][j];
            a1 += xr[i] * Vy[i][j];
        }
        av0[j] = a0;
        av1[j] = a1;
    }
    double cp00 = 0.0, cp01 = 0.0, cp11 = 0.0, r0s = 0.0, r1s = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        cp00 += av0[j];
        cp01 += av0[j] * xr[j];
        cp11 += av1[j] * xr[j];
        r0s += av0[j] * yr[j];
        r1s += av1[j] * yr[j];
    }
    // invert22 (cholesky_inversion.jl:8-19)
    const double luc0 = 1.0 / cp00;
    const double luc1 = cp01 * cp01 * luc0;
    const double luc2 = 1.0 / (cp11 - luc1);
    const double ci00 = luc1 * luc0 * luc2 + luc0, ci01 = -cp01 * luc0 * luc2, ci11 = luc2;
    const double sol0 = ci00 * r0s + ci01 * r1s, sol1 = ci01 * r0s + ci11 * r1s;
    const double cf = 1 / (sn - sol1 * cs);
    const double Jl[2][2] = {{0.0, cf * cf}, {cf, sol0 * cs * cf * cf}};
    const double m_ = cf * (sol1 * sn + cs);
    const double q_ = cf * sol0;
    const double Ci[2][2] = {{ci00, ci01}, {ci01, ci11}};
    double lcov[2][2];
#pragma unroll
    for (int xk = 0; xk < 2; ++xk)
#pragma unroll
        for (int yk = 0; yk < 2; ++yk) {
            double acc = 0.0;
#pragma unroll
            for (int zk = 0; zk < 2; ++zk) acc += (Jl[xk][0] * Ci[0][zk] + Jl[xk][1] * Ci[1][zk]) * Jl[yk][zk];
            lcov[xk][yk] = acc;
        }
    double chi2_line = 0.0;
    {
        double res[N];
#pragma unroll
        for (int i = 0; i < N; ++i) res[i] = yr[i] - (sol0 + xr[i] * sol1);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) acc += res[j] * Vy[j][i];
            chi2_line += acc * res[i];
        }
    }

    // ---------------------------------------------------------------- perigee and TrackSOA
    const double temp0 = X0 * X0 + Y0 * Y0;
    const double temp1 = sqrt(temp0);
    const double kp[3] = {atan2(q * X0, -q * Y0), q * (temp1 - Rc), q / Rc};
    const double temp2 = sqr(X0) * 1.0 / temp0;
    const double temp3 = 1.0 / temp1 * q;
    const double J4[3][3] = {{-Y0 * temp2 / (X0 * X0), temp2 / X0, 0.0},
                             {X0 * temp3, Y0 * temp3, (double)-q},
                             {0.0, 0.0, -q / (Rc * Rc)}};
    double kc[3][3];
#pragma unroll
    for (int xk = 0; xk < 3; ++xk)
#pragma unroll
        for (int yk = 0; yk < 3; ++yk) {
            double acc = 0.0;
#pragma unroll
            for (int zk = 0; zk < 3; ++zk) {
                double jc = 0.0;
#pragma unroll
                for (int m = 0; m < 3; ++m) jc += J4[xk][m] * ccov[m][zk];
                acc += jc * J4[yk][zk];
            }
            kc[xk][yk] = acc;
        }
    const double bi = 1.0 / B;
    out.state[0] = kp[0];
    out.state[1] = kp[1];
    out.state[2] = kp[2] * bi;
    out.state[3] = m_;
    out.state[4] = q_;
    out.cov15[0] = kc[0][0];
    out.cov15[1] = kc[0][1];
    out.cov15[2] = bi * kc[0][2];
    out.cov15[3] = 0.0;
    out.cov15[4] = 0.0;
    out.cov15[5] = kc[1][1];
    out.cov15[6] = bi * kc[1][2];
    out.cov15[7] = 0.0;
    out.cov15[8] = 0.0;
    out.cov15[9] = (bi * bi) * kc[2][2];
    out.cov15[10] = 0.0;
    out.cov15[11] = 0.0;
    out.cov15[12] = lcov[0][0];
    out.cov15[13] = lcov[0][1];
    out.cov15[14] = lcov[1][1];
    out.pt = B / fabs(kp[2]);
    out.eta = asinh(m_);
    out.chi2 = (chi2_circle + chi2_line) / double(2 * N - 5);
}

template <int N>
PT_HD void riemann_fit(const float (&x)[N], const float (&y)[N], const float (&z)[N], const float (&ge)[6][N],
                       double B, RiemannOut<N> &out) {
    // the inputs stay float (they arrive as float) and are widened where used
    auto X = [&](int i) -> double { return (double)x[i]; };
    auto Y = [&](int i) -> double { return (double)y[i]; };
    auto Z = [&](int i) -> double { return (double)z[i]; };
    auto GE = [&](int c, int i) -> double { return (double)ge[c][i]; };
    double ff[4];
    {
        double xd[N], yd[N], zd[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            xd[i] = X(i);
            yd[i] = Y(i);
            zd[i] = Z(i);
        }
        fast_fit<N>(xd, yd, zd, ff);
    }
    double rad[N];
#pragma unroll
    for (int i = 0; i < N; ++i) rad[i] = sqrt(X(i) * X(i) + Y(i) * Y(i));

    // ---------------------------------------------------------------- circle weights
    // scattering in radial coordinates, scatter_cov_rad (:112-150)
    double ms[N][N];
    {
        const double p_t = jl_min(20.0, ff[2] * B);
        const double p_2 = p_t * p_t * (1.0 + 1.0 / (ff[3] * ff[3]));
        double theta = atan(ff[3]);
        theta = theta < 0.0 ? theta + kPiD : theta;
        const double fac = sqrt(1.0 + 1.0 / (ff[3] * ff[3]));
        double len[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double p0 = X(i) - ff[0], p1 = Y(i) - ff[1];
            const float cr = (float)((-ff[0]) * p1 - (-ff[1]) * p0);
            const float dt = (float)((-ff[0]) * p0 + (-ff[1]) * p1);
            len[i] = fabs((double)atan2_bl(cr, dt) * ff[2]) * fac;
        }
        rad_len_uniform<N>(len);
        const double den = p_2 * sqr(sin(theta));
        double sig2[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double t = 1.0 + 0.038 * log(len[i]);
            sig2[i] = (t * t) * len[i] * 0.000225 / den;
        }
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int l = k; l < N; ++l) {
                double acc = 0.0;
#pragma unroll
                for (int i = 0; i <= (k < l ? k : l); ++i) acc += (rad[k] - rad[i]) * (rad[l] - rad[i]) * sig2[i];
                ms[k][l] = acc;
                ms[l][k] = acc;
            }
    }
    // V(a, b): hits (loadCovariance2D, fit_utils.jl:23-45) plus the Cartesian
    // scattering term of cov_radtocart (:158-180), formed per entry with the
    // same products; indices 0..N-1 are x, N..2N-1 are y.
    auto Vab = [&](int a, int b) -> double {
        const int i = a % N, j = b % N;
        const bool ay = a >= N, by = b >= N;
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const double crv = ms[lo][hi];
        double e;
        if (!ay && !by)
            e = crv * Y(lo) * (1.0 / rad[lo]) * Y(hi) * (1.0 / rad[hi]);
        else if (ay && by)
            e = crv * X(lo) * (1.0 / rad[lo]) * X(hi) * (1.0 / rad[hi]);
        else {
            // x index ix, y index iy: cc[ix][iy+N] = -cr * y_ix / r_ix * x_iy / r_iy when ix <= iy,
            // cc[iy+N][ix] = -cr * x_iy / r_iy * y_ix / r_ix when iy < ix (mirrored)
            const int ix = ay ? j : i, iy = ay ? i : j;
            if (ix <= iy)
                e = -crv * Y(ix) * (1.0 / rad[ix]) * X(iy) * (1.0 / rad[iy]);
            else
                e = -crv * X(iy) * (1.0 / rad[iy]) * Y(ix) * (1.0 / rad[ix]);
        }
        double hcv = 0.0;
        if (i == j) hcv = (!ay && !by) ? GE(0, i) : ((ay && by) ? GE(2, i) : GE(1, i));
        return hcv + e;
    };

    // radial covariance: cov_carttorad_prefit (:226-258) + scattering
    double cr[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double d;
        if (rad[i] < 0.0001) {
            d = GE(0, i);
        } else {
            const double b0 = X(i) - ff[0], b1 = Y(i) - ff[1];
            const double x2 = X(i) * b0 + Y(i) * b1;
            const double y2 = X(i) * b1 - Y(i) * b0;
            const double tc = -y2 / x2;
            const double tc2 = tc * tc;
            d = 1.0 / (1.0 + tc2) * (GE(0, i) + GE(2, i) * tc2 + 2 * GE(1, i) * tc);
        }
#pragma unroll
        for (int j = 0; j < N; ++j) cr[i][j] = (i == j ? d : 0.0) + ms[i][j];
    }
    double G[N][N];
    invert_n(cr, G);
    double renorm = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) renorm += G[i][j];
    double w[N];
#pragma unroll
    for (int j = 0; j < N; ++j) w[j] = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            G[i][j] = G[i][j] / renorm;
            w[j] += G[i][j];
        }

    // ---------------------------------------------------------------- plane fit
    double hx = 0.0, hy = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        hx += X(i);
        hy += Y(i);
    }
    hx = hx / N;
    hy = hy / N;
    double mcx[N], mcy[N];
    double qn = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        mcx[i] = X(i) - hx;
        mcy[i] = Y(i) - hy;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) qn += mcx[i] * mcx[i];
#pragma unroll
    for (int i = 0; i < N; ++i) qn += mcy[i] * mcy[i];
    qn = sqrt(qn);
    const double q2 = qn * qn;
    const double s = sqrt(N / q2);
    double P[3][N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        P[0][i] = mcx[i] * s;
        P[1][i] = mcy[i] * s;
        P[2][i] = P[0][i] * P[0][i] + P[1][i] * P[1][i];
    }
    double r0[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) acc += P[k][j] * w[j];
        r0[k] = acc;
    }
    double Xc[3][N];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < N; ++j) Xc[k][j] = P[k][j] - r0[k];
    double A[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = k; l < 3; ++l) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                double xg = 0.0;
#pragma unroll
                for (int i = 0; i < N; ++i) xg += Xc[k][i] * G[i][j];
                acc += xg * Xc[l][j];
            }
            A[k][l] = acc;
            A[l][k] = acc;
        }
    double Q[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    double lam[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) lam[k][l] = A[k][l];
    jacobi3(lam, Q, 60);
    const int mi = argmin3(lam[0][0], lam[1][1], lam[2][2]);
    const double lambda = mi == 0 ? lam[0][0] : (mi == 1 ? lam[1][1] : lam[2][2]);
    double v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = -(mi == 0 ? Q[k][0] : (mi == 1 ? Q[k][1] : Q[k][2]));
    if (!(v[2] > 0.0))
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = -v[k];

    const double c = -(v[0] * r0[0] + v[1] * r0[1] + v[2] * r0[2]);
    const double h = sqrt(1 - sqr(v[2]) - 4 * c * v[2]);
    const double v2x2_inv = 1 / (2 * v[2]);
    const double s_inv = 1 / s;
    const double pu[3] = {-v[0] * v2x2_inv, -v[1] * v2x2_inv, h * v2x2_inv};
    const double X0 = pu[0] * s_inv + hx, Y0 = pu[1] * s_inv + hy, Rc = pu[2] * s_inv;
    const int q = (X(1) - X(0)) * (Y0 - Y(0)) - (Y(1) - Y(0)) * (X0 - X(0)) > 0 ? -1 : 1;
    const double chi2_circle = fabs(lambda) * renorm * 1 / sqr(2 * v[2] * pu[2] * s);

    // ---------------------------------------------------------------- circle errors
    double cmc = 0.0, normV2 = 0.0;
#pragma unroll
    for (int a = 0; a < 2 * N; ++a)
#pragma unroll
        for (int b = 0; b < 2 * N; ++b) {
            const double vab = Vab(a, b);
            const double ma = a < N ? mcx[a] : mcy[a - N], mb = b < N ? mcx[b] : mcy[b - N];
            cmc += ma * vab * mb;
            normV2 += vab * vab;
        }
    normV2 = sqr(sqrt(normV2));
    const double kap = sqr(sqr(s)) * 1.0 / (4.0 * q2 * N) * (2.0 * normV2 + 4.0 * cmc);
    const double s2 = sqr(s);
    auto Vcs = [&](int a, int b) -> double {
        const double ma = a < N ? mcx[a] : mcy[a - N], mb = b < N ? mcx[b] : mcy[b - N];
        return s2 * Vab(a, b) + kap * (ma * mb);
    };

    // y_k = w .* Xc[k,:]
    double yv[3][N];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int a = 0; a < N; ++a) yv[k][a] = w[a] * Xc[k][a];

    // C0 = w' C_pr w and E (:611-742) from the bilinear forms
    // M_pr[x][z] = y_x' C_pr y_z of the six upper blocks C_pr (p <= r) of the
    // covariance of the lifted points (:557-603), one block at a time; the
    // lower blocks are the transposes (C_rp = C_pr'), so F(x; r, p; z) is
    // M_pr[z][x].  Term list of E[(i,j)][(k,l)] (:694-741):
    //   k == l, i == j:  4 F(i; j,l; l)
    //   k == l, i != j:  2 F(i; j,l; l) + 2 F(j; i,l; l)
    //   k != l, i == j:  2 F(i; j,l; k) + 2 F(i; j,k; l)
    //   k != l, i != j:  F(i; j,l; k) + F(i; j,k; l) + F(j; i,l; k) + F(j; i,k; l)
    double C0[3][3];
    double E[6][6] = {};
    constexpr int nu[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    // one block per iteration, not unrolled: the entries of V are recomputed
    // per block instead of all being kept live across the six
#pragma unroll 1
    for (int pb = 0; pb < 6; ++pb) {
        const int p = pb < 3 ? 0 : (pb < 5 ? 1 : 2), r = pb < 3 ? pb : (pb < 5 ? pb - 2 : 2);
        double M[3][3] = {};
        double c0 = 0.0;
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
            for (int b = 0; b < N; ++b) {
                const int lo = a < b ? a : b, hi = a < b ? b : a;
                double cv;
                if (p == 0 && r == 0) {
                    cv = Vcs(lo, hi);
                } else if (p == 1 && r == 1) {
                    cv = Vcs(N + lo, N + hi);
                } else if (p == 0 && r == 1) {
                    cv = Vcs(a, N + b);
                } else if (p == 0 && r == 2) {
                    cv = 2 * ((Vcs(lo, hi) * P[0][b]) + (Vcs(a, N + b) * P[1][b]));
                } else if (p == 1 && r == 2) {
                    cv = 2 * (Vcs(b, N + a) * P[0][b] + Vcs(N + lo, N + hi) * P[1][b]);
                } else {
                    // C22 = Symmetric(tmp, :U): tmp at (lo, hi), t01 = x_i y_j (:581)
                    const double d00 = Vcs(lo, hi), d11 = Vcs(N + lo, N + hi);
                    const double d01 = Vcs(lo, N + hi), d10 = Vcs(hi, N + lo);
                    const double plx = P[0][lo], phx = P[0][hi], ply = P[1][lo], phy = P[1][hi];
                    cv = 2.0 * (d00 * d00 + d00 * d01 + d11 * d10 + d11 * d11) +
                         4.0 * (d00 * (phx * plx) + d01 * (plx * phy) + d10 * (phx * ply) + d11 * (phy * ply));
                }
                c0 += cv * (w[a] * w[b]);
#pragma unroll
                for (int xk = 0; xk < 3; ++xk) {
                    const double yc = yv[xk][a] * cv;
#pragma unroll
                    for (int zk = 0; zk < 3; ++zk) M[xk][zk] += yc * yv[zk][b];
                }
            }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if ((i == p && j == r) || (i == r && j == p)) C0[i][j] = c0;
        // add this block's terms to E
#pragma unroll
        for (int A_ = 0; A_ < 6; ++A_) {
            const int i = nu[A_][0], j = nu[A_][1];
#pragma unroll
            for (int B_ = A_; B_ < 6; ++B_) {
                const int k = nu[B_][0], l = nu[B_][1];
                double e = 0.0;
                // weights of 0 or 1, not a choice between two addresses, so M
                // stays in registers with the block index known only at run time
                auto F = [&](double coef, int xk, int pp, int rr, int zk) {
                    const double m1 = (pp == p && rr == r) ? 1.0 : 0.0;
                    const double m2 = (pp != rr && pp == r && rr == p) ? 1.0 : 0.0;
                    e += coef * (m1 * M[xk][zk] + m2 * M[zk][xk]);
                };
                if (k == l) {
                    if (i == j) {
                        F(4.0, i, j, l, l);
                    } else {
                        F(2.0, i, j, l, l);
                        F(2.0, j, i, l, l);
                    }
                } else {
                    if (i == j) {
                        F(2.0, i, j, l, k);
                        F(2.0, i, j, k, l);
                    } else {
                        F(1.0, i, j, l, k);
                        F(1.0, i, j, k, l);
                        F(1.0, j, i, l, k);
                        F(1.0, j, i, k, l);
                    }
                }
                E[A_][B_] += e;
            }
        }
    }

    // J2: forward differences of the eigenvector (:746-775), each perturbed
    // matrix solved in the eigenbasis Q of A
    double J2[3][6];
#pragma unroll
    for (int a_ = 0; a_ < 6; ++a_) {
        const int i = nu[a_][0], j = nu[a_][1];
        const double delta = fabs(A[i][j] * kEps);
        double bm[3][3];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n2 = 0; n2 < 3; ++n2) {
                const double d = (i == j) ? Q[i][m] * Q[i][n2] : (Q[i][m] * Q[j][n2] + Q[j][m] * Q[i][n2]);
                bm[m][n2] = (m == n2 ? lam[m][m] : 0.0) + delta * d;
            }
        double U[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
        jacobi3(bm, U, 8);
        const int m2 = argmin3(bm[0][0], bm[1][1], bm[2][2]);
        double vec[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double u0 = m2 == 0 ? U[0][0] : (m2 == 1 ? U[0][1] : U[0][2]);
            const double u1 = m2 == 0 ? U[1][0] : (m2 == 1 ? U[1][1] : U[1][2]);
            const double u2 = m2 == 0 ? U[2][0] : (m2 == 1 ? U[2][1] : U[2][2]);
            vec[k] = -(Q[k][0] * u0 + Q[k][1] * u1 + Q[k][2] * u2);
        }
        if (!(vec[2] > 0.0))
#pragma unroll
            for (int k = 0; k < 3; ++k) vec[k] = -vec[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) J2[k][a_] = (vec[k] * 1.0 - v[k]) / delta;
    }

    // Cvc (:779-795)
    double t0m[3][3];
    {
        double JE[3][6];
#pragma unroll
        for (int xk = 0; xk < 3; ++xk)
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < 6; ++a) acc += J2[xk][a] * (a <= b ? E[a][b] : E[b][a]);
                JE[xk][b] = acc;
            }
#pragma unroll
        for (int xk = 0; xk < 3; ++xk)
#pragma unroll
            for (int yk = 0; yk < 3; ++yk) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < 6; ++b) acc += JE[xk][b] * J2[yk][b];
                t0m[xk][yk] = acc;
            }
    }
    double Cvc[4][4];
#pragma unroll
    for (int xk = 0; xk < 3; ++xk) {
        double acc = 0.0;
#pragma unroll
        for (int yk = 0; yk < 3; ++yk) {
            Cvc[xk][yk] = t0m[xk][yk];
            acc += (-t0m[xk][yk]) * r0[yk];
        }
        Cvc[xk][3] = acc;
        Cvc[3][xk] = acc;
    }
    {
        double cm1 = 0.0, cm3 = 0.0, hs = 0.0;
#pragma unroll
        for (int xk = 0; xk < 3; ++xk)
#pragma unroll
            for (int yk = 0; yk < 3; ++yk) {
                cm1 += v[xk] * C0[xk][yk] * v[yk];
                hs += C0[xk][yk] * t0m[xk][yk];
                cm3 += r0[xk] * t0m[xk][yk] * r0[yk];
            }
        Cvc[3][3] = cm1 + hs + cm3;
    }
    const double t = 1 / h;
    const double J3[3][4] = {{-v2x2_inv, 0.0, v[0] * sqr(v2x2_inv) * 2.0, 0.0},
                             {0.0, -v2x2_inv, v[1] * sqr(v2x2_inv) * 2.0, 0.0},
                             {v[0] * v2x2_inv * t, v[1] * v2x2_inv * t,
                              -h * sqr(v2x2_inv) * 2.0 - (2.0 * c + v[2]) * v2x2_inv * t, -t}};
    const double jqs = s * 1 / N;
    const double varq = (jqs * jqs) * cmc;  // Jq V Jq', Jq = mc' s / N
    double ccov[3][3];
    {
        double JC[3][4];
#pragma unroll
        for (int xk = 0; xk < 3; ++xk)
#pragma unroll
            for (int yk = 0; yk < 4; ++yk) {
                double acc = 0.0;
#pragma unroll
                for (int zk = 0; zk < 4; ++zk) acc += J3[xk][zk] * Cvc[zk][yk];
                JC[xk][yk] = acc;
            }
#pragma unroll
        for (int xk = 0; xk < 3; ++xk)
#pragma unroll
            for (int yk = 0; yk < 3; ++yk) {
                double acc = 0.0;
#pragma unroll
                for (int zk = 0; zk < 4; ++zk) acc += JC[xk][zk] * J3[yk][zk];
                ccov[xk][yk] = acc * sqr(s_inv) + (pu[xk] * pu[yk]) * varq;
            }
    }

    riemann_line_and_output<N>(x, y, z, ge, ff, X0, Y0, Rc, q, ccov, chi2_circle, B, out);
}

}  // namespace riemann
}  // namespace pt
