// pt_bl_fit.hpp - broken-line helix fit, device code (one track per lane).
//
// Method of the reference fit: BL_Fast_fit -> prepare_broken_line_data ->
// BL_Line_fit -> BL_Circle_fit -> copyFromCircle!
// (broken_line.jl:46-696, cholesky_inversion.jl:22-129,
//  broken_line_fit_on_gpu.jl:24-163, TrajectoryStateSOA.jl:20-44).
// Everything lives in registers: N <= 4 hits, so the band systems are at most
// 4x4 (line) and 5x5 (circle, with the curvature row).  The operation order is
// the reference's (build with -ffp-contract=off), so the all-double policy
// agrees with a serial double fit to the rounding of the transcendental
// functions.
//
// The arithmetic type of each part of the fit is chosen by a policy:
//   Pre  - the fast (three-point) circle, the hit radii about its centre and
//          the charge; near-collinear stiff tracks make the circle centre
//          large (|c| ~ R ~ 10^3-10^4 cm) and these differences lose all
//          digits in float
//   Geo  - arc lengths, the rotated (S, Z) frame, scattering variances and
//          the per-hit weights
//   Sys  - the band matrices, their inversion, the solution u, the line
//          parameters and both chi2 sums
//   Prop - circle parameters from the corrected points, the Jacobians, the
//          Karimaki translations and the covariance propagation
// PolicyFP64 is the reference's double-precision fit; PolicyMixed keeps
// double only for Pre and Sys (and the sagitta of the circle parameters).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PT_HD __host__ __device__ __forceinline__

namespace pt {

constexpr double kPiD = 3.141592653589793;

struct PolicyFP64 {
    typedef double Pre, Arc, Geo, Wt, Sys, Kar, Prop;
    static constexpr bool trig_free = false;
};
struct PolicyMixed {
    typedef double Pre, Arc, Geo, Sys, Kar;
    typedef float Wt, Prop;
    static constexpr bool trig_free = true;
};

template <class R> struct FitOut {
    R state[5];
    R cov[15];
    R pt, eta, chi2;
    R chi2_line, chi2_circle;
    uint32_t flags;  // 1: sqrt of a negative number, 2: q*d > 0, 4: NaN pre-fit
};

PT_HD double r_sqrt(double x) { return sqrt(x); }
PT_HD float r_sqrt(float x) { return sqrtf(x); }
PT_HD double r_abs(double x) { return fabs(x); }
PT_HD float r_abs(float x) { return fabsf(x); }

// broken_line.jl:256-268 (x < 0 gives atan(y/x) + pi, also for y < 0)
PT_HD double atan2_bl(double y, double x) {
    if (x > 0.0) return atan(y / x);
    if (x < 0.0) return atan(y / x) + kPiD;
    if (y > 0.0 && x == 0.0) return kPiD / 2.0;
    if (y < 0.0 && x == 0.0) return -kPiD / 2.0;
    return __builtin_nan("");
}
PT_HD float atan2_bl(float y, float x) {
    if (x > 0.0f) return atanf(y / x);
    if (x < 0.0f) return atanf(y / x) + (float)kPiD;
    if (y > 0.0f && x == 0.0f) return (float)(kPiD / 2.0);
    if (y < 0.0f && x == 0.0f) return (float)(-kPiD / 2.0);
    return __builtin_nanf("");
}
PT_HD double log_r(double x) { return log(x); }
PT_HD float log_r(float x) { return logf(x); }
PT_HD double cos_r(double x) { return cos(x); }
PT_HD float cos_r(float x) { return cosf(x); }
PT_HD double sin_r(double x) { return sin(x); }
PT_HD float sin_r(float x) { return sinf(x); }
PT_HD double asinh_r(double x) { return asinh(x); }
PT_HD float asinh_r(float x) { return asinhf(x); }

template <class R> PT_HD R jl_min(R a, R b) {
    if (a != a) return a;
    if (b != b) return b;
    return (b < a) ? b : a;
}

// broken_line.jl:46-79
template <class R> PT_HD R mult_scatt(R length, R B, R Rr, R slope) {
    R pt2 = jl_min(R(20.0), B * Rr);
    pt2 = pt2 * pt2;
    const R XXI_0 = R(0.06 / 16.0);
    const R fact = R(0.7 * ((13.6 / 1000) * (13.6 / 1000)));
    const R slope_term = R(1.0) + slope * slope;
    const R al = r_abs(length) * XXI_0;
    const R sq = R(1.0) + R(0.038) * log_r(al);
    return fact / (pt2 * slope_term) * al * (sq * sq);
}

// cholesky_inversion.jl:22-129; lower triangle computed, then symmetrised
template <class R> PT_HD void invert3(const R (&s)[3][3], R (&d)[3][3]) {
    R luc0 = R(1.0) / s[0][0];
    R luc1 = s[1][0];
    R luc2 = s[1][1] - luc0 * luc1 * luc1;
    luc2 = R(1.0) / luc2;
    R luc3 = s[2][0];
    R luc4 = s[2][1] - luc0 * luc1 * luc3;
    R luc5 = s[2][2] - (luc0 * luc3 * luc3 + luc2 * luc4 * luc4);
    luc5 = R(1.0) / luc5;
    R li21 = -luc0 * luc1;
    R li32 = -luc2 * luc4;
    R li31 = (luc1 * luc2 * luc4 - luc3) * luc0;
    d[0][0] = luc5 * li31 * li31 + luc2 * li21 * li21 + luc0;
    d[1][0] = luc5 * li31 * li32 + luc2 * li21;
    d[1][1] = luc5 * li32 * li32 + luc2;
    d[2][0] = luc5 * li31;
    d[2][1] = luc5 * li32;
    d[2][2] = luc5;
    d[0][1] = d[1][0];
    d[0][2] = d[2][0];
    d[1][2] = d[2][1];
}
template <class R> PT_HD void invert4(const R (&s)[4][4], R (&d)[4][4]) {
    R luc0 = R(1.0) / s[0][0];
    R luc1 = s[1][0];
    R luc2 = s[1][1] - luc0 * luc1 * luc1;
    luc2 = R(1.0) / luc2;
    R luc3 = s[2][0];
    R luc4 = s[2][1] - luc0 * luc1 * luc3;
    R luc5 = s[2][2] - (luc0 * luc3 * luc3 + luc2 * luc4 * luc4);
    luc5 = R(1.0) / luc5;
    R luc6 = s[3][0];
    R luc7 = s[3][1] - luc0 * luc1 * luc6;
    R luc8 = s[3][2] - luc0 * luc3 * luc6 - luc2 * luc4 * luc7;
    R luc9 = s[3][3] - (luc0 * luc6 * luc6 + luc2 * luc7 * luc7 + luc8 * luc8 * luc5);
    luc9 = R(1.0) / luc9;
    R li21 = -luc0 * luc1;
    R li32 = -luc2 * luc4;
    R li31 = (luc1 * luc2 * luc4 - luc3) * luc0;
    R li43 = -luc8 * luc5;
    R li42 = (luc4 * luc8 * luc5 - luc7) * luc2;
    R li41 = (-luc1 * luc2 * luc4 * luc8 * luc5 + luc1 * luc2 * luc7 + luc3 * luc8 * luc5 - luc6) * luc0;
    d[0][0] = luc9 * li41 * li41 + luc5 * li31 * li31 + luc2 * li21 * li21 + luc0;
    d[1][0] = luc9 * li41 * li42 + luc5 * li31 * li32 + luc2 * li21;
    d[1][1] = luc9 * li42 * li42 + luc5 * li32 * li32 + luc2;
    d[2][0] = luc9 * li41 * li43 + luc5 * li31;
    d[2][1] = luc9 * li42 * li43 + luc5 * li32;
    d[2][2] = luc9 * li43 * li43 + luc5;
    d[3][0] = luc9 * li41;
    d[3][1] = luc9 * li42;
    d[3][2] = luc9 * li43;
    d[3][3] = luc9;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i + 1; j < 4; ++j) d[i][j] = d[j][i];
}
template <class R> PT_HD void invert5(const R (&s)[5][5], R (&d)[5][5]) {
    R luc0 = R(1.0) / s[0][0];
    R luc1 = s[1][0];
    R luc2 = s[1][1] - luc0 * luc1 * luc1;
    luc2 = R(1.0) / luc2;
    R luc3 = s[2][0];
    R luc4 = s[2][1] - luc0 * luc1 * luc3;
    R luc5 = s[2][2] - (luc0 * luc3 * luc3 + luc2 * luc4 * luc4);
    luc5 = R(1.0) / luc5;
    R luc6 = s[3][0];
    R luc7 = s[3][1] - luc0 * luc1 * luc6;
    R luc8 = s[3][2] - luc0 * luc3 * luc6 - luc2 * luc4 * luc7;
    R luc9 = s[3][3] - (luc0 * luc6 * luc6 + luc2 * luc7 * luc7 + luc8 * luc8 * luc5);
    luc9 = R(1.0) / luc9;
    R luc10 = s[4][0];
    R luc11 = s[4][1] - luc0 * luc1 * luc10;
    R luc12 = s[4][2] - luc0 * luc3 * luc10 - luc2 * luc4 * luc11;
    R luc13 = s[4][3] - luc0 * luc6 * luc10 - luc2 * luc7 * luc11 - luc5 * luc8 * luc12;
    R luc14 = s[4][4] - (luc0 * luc10 * luc10 + luc2 * luc11 * luc11 + luc5 * luc12 * luc12 + luc9 * luc13 * luc13);
    luc14 = R(1.0) / luc14;
    R li21 = -luc0 * luc1;
    R li32 = -luc2 * luc4;
    R li31 = (luc1 * luc2 * luc4 - luc3) * luc0;
    R li43 = -luc8 * luc5;
    R li42 = (luc4 * luc8 * luc5 - luc7) * luc2;
    R li41 = (-luc1 * luc2 * luc4 * luc8 * luc5 + luc1 * luc2 * luc7 + luc3 * luc8 * luc5 - luc6) * luc0;
    R li54 = -luc13 * luc9;
    R li53 = (luc13 * luc8 * luc9 - luc12) * luc5;
    R li52 = (-luc4 * luc8 * luc13 * luc5 * luc9 + luc4 * luc12 * luc5 + luc7 * luc13 * luc9 - luc11) * luc2;
    R li51 = (luc1 * luc4 * luc8 * luc13 * luc2 * luc5 * luc9 - luc13 * luc8 * luc3 * luc9 * luc5 -
              luc12 * luc4 * luc1 * luc2 * luc5 - luc13 * luc7 * luc1 * luc9 * luc2 + luc11 * luc1 * luc2 +
              luc12 * luc3 * luc5 + luc13 * luc6 * luc9 - luc10) *
             luc0;
    d[0][0] = luc14 * li51 * li51 + luc9 * li41 * li41 + luc5 * li31 * li31 + luc2 * li21 * li21 + luc0;
    d[1][0] = luc14 * li51 * li52 + luc9 * li41 * li42 + luc5 * li31 * li32 + luc2 * li21;
    d[1][1] = luc14 * li52 * li52 + luc9 * li42 * li42 + luc5 * li32 * li32 + luc2;
    d[2][0] = luc14 * li51 * li53 + luc9 * li41 * li43 + luc5 * li31;
    d[2][1] = luc14 * li52 * li53 + luc9 * li42 * li43 + luc5 * li32;
    d[2][2] = luc14 * li53 * li53 + luc9 * li43 * li43 + luc5;
    d[3][0] = luc14 * li51 * li54 + luc9 * li41;
    d[3][1] = luc14 * li52 * li54 + luc9 * li42;
    d[3][2] = luc14 * li53 * li54 + luc9 * li43;
    d[3][3] = luc14 * li54 * li54 + luc9;
    d[4][0] = luc14 * li51;
    d[4][1] = luc14 * li52;
    d[4][2] = luc14 * li53;
    d[4][3] = luc14 * li54;
    d[4][4] = luc14;
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = i + 1; j < 5; ++j) d[i][j] = d[j][i];
}
template <class R> PT_HD void invert_n(const R (&s)[3][3], R (&d)[3][3]) { invert3(s, d); }
template <class R> PT_HD void invert_n(const R (&s)[4][4], R (&d)[4][4]) { invert4(s, d); }
template <class R> PT_HD void invert_n(const R (&s)[5][5], R (&d)[5][5]) { invert5(s, d); }

// broken_line.jl:217-250 ; band matrix of the broken-line cost (C + C')
template <class R, int N>
PT_HD void band_matrix(const R (&w)[N], const R (&S)[N], const R (&VB)[N], R (&out)[N][N]) {
    R C[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) C[i][j] = R(0.0);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const int I = i + 1;
        C[i][i] = w[i];
        if (I > 2) {
            const R d = S[i] - S[i - 1];
            C[i][i] = C[i][i] + R(1.0) / (VB[i - 1] * (d * d));
        }
        if (I > 1 && I < N) {
            const R q = (S[i + 1] - S[i - 1]) / ((S[i + 1] - S[i]) * (S[i] - S[i - 1]));
            C[i][i] = C[i][i] + (R(1.0) / VB[i]) * (q * q);
        }
        if (I < N - 1) {
            const R d = S[i + 1] - S[i];
            C[i][i] = C[i][i] + R(1.0) / (VB[i + 1] * (d * d));
        }
        if (I > 1 && I < N) {
            C[i][i + 1] = R(1.0) / (VB[i] * (S[i + 1] - S[i])) *
                          (-(S[i + 1] - S[i - 1]) / ((S[i + 1] - S[i]) * (S[i] - S[i - 1])));
        }
        if (I < N - 1) {
            C[i][i + 1] = C[i][i + 1] + R(1.0) / (VB[i + 1] * (S[i + 1] - S[i])) *
                                            (-(S[i + 2] - S[i]) / ((S[i + 2] - S[i + 1]) * (S[i + 1] - S[i])));
            C[i][i + 2] = R(1.0) / (VB[i + 1] * (S[i + 2] - S[i + 1]) * (S[i + 1] - S[i]));
        }
        C[i][i] = C[i][i] * R(0.5);
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) out[i][j] = C[i][j] + C[j][i];
}

// cov <- J * cov * J' for 2x2, element sums left to right
template <class R> PT_HD void sandwich2(const R (&J)[2][2], R (&cov)[2][2]) {
    R T[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) T[i][j] = J[i][0] * cov[0][j] + J[i][1] * cov[1][j];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) cov[i][j] = T[i][0] * J[j][0] + T[i][1] * J[j][1];
}
template <class R> PT_HD void sandwich3(const R (&J)[3][3], R (&cov)[3][3]) {
    R T[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) T[i][j] = J[i][0] * cov[0][j] + J[i][1] * cov[1][j] + J[i][2] * cov[2][j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) cov[i][j] = T[i][0] * J[j][0] + T[i][1] * J[j][1] + T[i][2] * J[j][2];
}

// translate_karimaki (broken_line.jl:103-132): parameters (phi, d, kappa)
// about a new origin (x0, y0).  The parameter part alone (the adjoint
// differentiates it); the Jacobian for the covariance below.
template <class R> struct KarimakiStep {
    R cp, sp, DP, DO, uu, Cc, BB, A, U, xi, v, lambda, mu, zeta;
};
template <class R>
PT_HD void karimaki_par(R (&par)[3], R x0, R y0, R cp, R sp, KarimakiStep<R> &k, uint32_t &flags) {
    k.cp = cp;  // cos(par[0]), sin(par[0])
    k.sp = sp;
    k.DP = x0 * k.cp + y0 * k.sp;
    k.DO = x0 * k.sp - y0 * k.cp + par[1];
    k.uu = R(1.0) + par[2] * par[1];
    k.Cc = -par[2] * y0 + k.uu * k.cp;
    k.BB = par[2] * x0 + k.uu * k.sp;
    k.A = R(2.0) * k.DO + par[2] * (k.DO * k.DO + k.DP * k.DP);
    const R arg = R(1.0) + par[2] * k.A;
    if (arg < R(0.0)) flags |= 1u;
    k.U = r_sqrt(arg);
    par[0] = atan2_bl(k.BB, k.Cc);
    par[1] = k.A / (R(1.0) + k.U);
}
template <class R, class KS> PT_HD void karimaki_cov(const KS &k, R kappa, R (&cov)[3][3]) {
    const R BB = (R)k.BB, Cc = (R)k.Cc, DO = (R)k.DO, DP = (R)k.DP, U = (R)k.U, A = (R)k.A, uu = (R)k.uu;
    const R xi = R(1.0) / (BB * BB + Cc * Cc);
    const R v = R(1.0) + kappa * DO;
    const R op = R(1.0) + U;
    const R lambda = (R(0.5) * A) / (U * (op * op));
    const R mu = R(1.0) / (U * (R(1.0) + U)) + kappa * lambda;
    const R zeta = DO * DO + DP * DP;
    R J[3][3];
    J[0][0] = xi * uu * v;
    J[0][1] = -xi * (kappa * kappa) * DP;
    J[0][2] = xi * DP;
    J[1][0] = R(2.0) * mu * uu * DP;
    J[1][1] = R(2.0) * mu * v;
    J[1][2] = mu * zeta - lambda * A;
    J[2][0] = R(0.0);
    J[2][1] = R(0.0);
    J[2][2] = R(1.0);
    sandwich3(J, cov);
}

// One track of N hits, every intermediate kept as a member so that the
// adjoint can run its reverse sweep over them (the members stay in registers
// once inlined; the plain fit never reads the ones it does not need).
// Inputs: hit positions (float, cm) and global covariances ge[6][N] (float,
// xx, xy, yy, xz, yz, zz); B: field in the fitter's units (Float32 bf).
template <class P, int N> struct BLFit {
    typedef typename P::Pre D;
    typedef typename P::Arc A_;
    typedef typename P::Geo G;
    typedef typename P::Wt W;
    typedef typename P::Sys S_;
    typedef typename P::Kar K;
    typedef typename P::Prop Q;
    static constexpr int mid = (N >> 1);

    uint32_t flags;
    int q;
    // fast fit
    D ff0, ff1, ff2, ff3, ang, dz;
    D fa1, fa2, fb1, fb2, fc1, fc2, fca, fba, fnc, fna, fnb, fd1, fd2, fe1, fe2;
    // prepared data
    D radx[N], rady[N], rad[N];
    D e1, e2, fac, rr;
    D cc[N], dd[N];
    G slope, ra, rb, s[N], S[N], Z[N];
    W VB[N];
    // line fit
    W lw[N], lj0[N], lj1[N];
    S_ lI[N][N], lu[N], lpar0, lpar1, lpar0_raw, lpar1_raw, lchi2;
    S_ lcv[2][2];
    // circle fit
    W vscale, VBs[N], cw[N], csl[N], ca_[N], cb_[N];
    D Zc[N];
    S_ cI[N + 1][N + 1], cu[N + 1], cchi2;
    K n0x, n0y, n1x, n1y, dpt1, dpt2, ept1, ept2, em0, em1, nrm;
    D arg, sarg;
    K cpar[3];
    D cpar1_d, cpar2_d;
    KarimakiStep<K> k1, k2;
    Q ms0;

    PT_HD void run(const float (&hx)[N], const float (&hy)[N], const float (&hz)[N], const float (&ge)[6][N],
                   double B, FitOut<Q> &o) {
        flags = 0;
        // ---- BL_Fast_fit (broken_line.jl:275-313)
        {
            const D h11 = hx[0], h21 = hy[0], h1m = hx[mid], h2m = hy[mid], h1n = hx[N - 1], h2n = hy[N - 1];
            fa1 = h1m - h11, fa2 = h2m - h21;
            fb1 = h1n - h1m, fb2 = h2n - h2m;
            fc1 = h11 - h1n, fc2 = h21 - h2n;
            fca = fc1 * fa2 - fc2 * fa1;
            fba = fb1 * fa2 - fb2 * fa1;
            fnc = fc1 * fc1 + fc2 * fc2;
            fna = fa1 * fa1 + fa2 * fa2;
            fnb = fb1 * fb1 + fb2 * fb2;
            const D tmp = D(0.5) / fca;
            ff0 = h11 - (fa2 * fnc + fc2 * fna) * tmp;
            ff1 = h21 + (fa1 * fnc + fc1 * fna) * tmp;
            ff2 = r_sqrt(fna * fnb * fnc) / (D(2.0) * r_abs(fba));
            fd1 = h11 - ff0, fd2 = h21 - ff1;
            fe1 = h1n - ff0, fe2 = h2n - ff1;
            const D de = fd1 * fe2 - fd2 * fe1;
            const D dotde = fd1 * fe1 + fd2 * fe2;
            dz = (D)hz[N - 1] - (D)hz[0];
            ang = atan2_bl(de, dotde);
            ff3 = ff2 * ang / dz;
            if (ff0 != ff0 || ff1 != ff1 || ff2 != ff2 || ff3 != ff3) flags |= 4u;
        }
        // ---- prepare_broken_line_data (broken_line.jl:150-207)
        {
            const D x1 = (D)hx[1] - (D)hx[0], y1 = (D)hy[1] - (D)hy[0];
            const D xn = (D)hx[N - 1] - (D)hx[N - 2], yn = (D)hy[N - 1] - (D)hy[N - 2];
            q = (x1 * yn - y1 * xn > 0.0) ? -1 : 1;
        }
        slope = (G)((D)(-q) / ff3);
        ra = G(1.0) / r_sqrt(G(1.0) + slope * slope);
        rb = slope * ra;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            radx[j] = (D)hx[j] - ff0;
            rady[j] = (D)hy[j] - ff1;
        }
        rr = r_sqrt(ff0 * ff0 + ff1 * ff1);
        {
            const D sc = -ff2 / rr;
            e1 = sc * ff0;
            e2 = sc * ff1;
        }
        fac = (D)q * ff2;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            cc[i] = radx[i] * e2 - rady[i] * e1;
            dd[i] = radx[i] * e1 + rady[i] * e2;
            s[i] = (G)(fac * (D)atan2_bl((A_)cc[i], (A_)dd[i]));
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
            S[i] = ra * s[i] + rb * (G)hz[i];
            Z[i] = (-rb) * s[i] + ra * (G)hz[i];
        }
        VB[0] = G(0.0);
        VB[N - 1] = G(0.0);
#pragma unroll
        for (int i = 1; i < N - 1; ++i)
            VB[i] = mult_scatt<W>((W)(S[i + 1] - S[i]), (W)B, (W)ff2, (W)slope) +
                    mult_scatt<W>((W)(S[i] - S[i - 1]), (W)B, (W)ff2, (W)slope);

        // ---- BL_Line_fit (broken_line.jl:546-696)
#pragma unroll
        for (int i = 0; i < N; ++i) {
            rad[i] = r_sqrt(radx[i] * radx[i] + rady[i] * rady[i]);
            const W v00 = ge[0][i];
            const W v01 = (fabsf(ge[1][i]) < 1e-12) ? W(0.0) : (W)ge[1][i];
            const W v02 = (fabsf(ge[3][i]) < 1e-12) ? W(0.0) : (W)ge[3][i];
            const W v11 = ge[2][i];
            const W v12 = (fabsf(ge[4][i]) < 1e-12) ? W(0.0) : (W)ge[4][i];
            const W v22 = ge[5][i];
            const D tmp = D(1.0) / rad[i];
            const W j00 = (W)(rady[i] * tmp), j01 = (W)(-radx[i] * tmp);
            const W rbw = (W)rb, raw = (W)ra;
            lj0[i] = j00;
            lj1[i] = j01;
            // Jacob row 1 = (j00, j01, 0), row 2 = (0, 0, 1); M1 = Rm * J with
            // Rm = [ra rb; -rb ra]; only element [2,2] of Rm J V J' Rm' is needed
            const W m1_0 = (-rbw) * j00 + raw * W(0.0);
            const W m1_1 = (-rbw) * j01 + raw * W(0.0);
            const W m1_2 = (-rbw) * W(0.0) + raw * W(1.0);
            const W m2_0 = m1_0 * v00 + m1_1 * v01 + m1_2 * v02;
            const W m2_1 = m1_0 * v01 + m1_1 * v11 + m1_2 * v12;
            const W m2_2 = m1_0 * v02 + m1_1 * v12 + m1_2 * v22;
            const W m3_0 = m2_0 * j00 + m2_1 * j01 + m2_2 * W(0.0);
            const W m3_1 = m2_0 * W(0.0) + m2_1 * W(0.0) + m2_2 * W(1.0);
            const W m4 = m3_0 * (-rbw) + m3_1 * raw;
            lw[i] = W(1.0) / m4;
        }
        {
            S_ w[N], Ss[N], Vb[N], C[N][N];
#pragma unroll
            for (int i = 0; i < N; ++i) {
                w[i] = lw[i];
                Ss[i] = S[i];
                Vb[i] = VB[i];
            }
            band_matrix<S_, N>(w, Ss, Vb, C);
            invert_n(C, lI);
#pragma unroll
            for (int i = 0; i < N; ++i) {
                S_ acc = lI[i][0] * (w[0] * (S_)Z[0]);
#pragma unroll
                for (int k = 1; k < N; ++k) acc = acc + lI[i][k] * (w[k] * (S_)Z[k]);
                lu[i] = acc;
            }
            const S_ dS = Ss[1] - Ss[0];
            lpar0_raw = (lu[1] - lu[0]) / dS;
            lpar1 = lu[0];
            const S_ idiff = S_(1.0) / dS;
            S_ cv[2][2];
            cv[0][0] = (lI[0][0] - S_(2.0) * lI[0][1] + lI[1][1]) * (idiff * idiff) +
                       (S_)mult_scatt<W>((W)(S[1] - S[0]), (W)B, (W)ff2, (W)slope);
            cv[0][1] = (lI[0][1] - lI[0][0]) * idiff;
            cv[1][0] = (lI[0][1] - lI[0][0]) * idiff;
            cv[1][1] = lI[0][0];
            lpar1 = lpar1 - lpar0_raw * Ss[0];
            lpar1_raw = lpar1;
            {
                const S_ J[2][2] = {{S_(1.0), S_(0.0)}, {-Ss[0], S_(1.0)}};
                sandwich2(J, cv);
            }
            const S_ tmp = (S_)ra - lpar0_raw * (S_)rb;
            const S_ j11 = S_(1.0) / tmp;
            const S_ j00 = j11 * j11;
            const S_ j10 = lpar1 * (S_)rb * j00;
            lpar1 = lpar1 * j11;
            lpar0 = ((S_)rb + lpar0_raw * (S_)ra) * j11;
            {
                const S_ J[2][2] = {{j00, S_(0.0)}, {j10, j11}};
                sandwich2(J, cv);
            }
            lcv[0][0] = cv[0][0];
            lcv[0][1] = cv[0][1];
            lcv[1][1] = cv[1][1];
            // the line's outputs are final here: storing them now frees their
            // registers for the circle fit
            o.state[3] = (Q)lpar0;
            o.state[4] = (Q)lpar1;
            o.cov[12] = (Q)lcv[0][0];
            o.cov[13] = (Q)lcv[0][1];
            o.cov[14] = (Q)lcv[1][1];
            o.eta = asinh_r((Q)lpar0);
            lchi2 = S_(0.0);
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const S_ diff = (S_)Z[i] - lu[i];
                lchi2 = lchi2 + w[i] * (diff * diff);
                if (i > 0 && i < N - 1) {
                    const S_ d1 = Ss[i] - Ss[i - 1], d2 = Ss[i + 1] - Ss[i], d4 = Ss[i + 1] - Ss[i - 1];
                    const S_ inner = (lu[i - 1] / d1) - (lu[i] * (d4 / (d2 * d1))) + (lu[i + 1] / (Ss[i + 1] - Ss[i]));
                    lchi2 = lchi2 + (inner * inner) / Vb[i];
                }
            }
        }

        // ---- BL_Circle_fit (broken_line.jl:330-528)
        vscale = W(1.0) + (W)slope * (W)slope;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            VBs[i] = VB[i] * vscale;
            Zc[i] = rad[i] - ff2;
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const W v00 = (fabsf(ge[0][i]) < 1e-12) ? W(0.0) : (W)ge[0][i];
            const W v01 = (fabsf(ge[1][i]) < 1e-12) ? W(0.0) : (W)ge[1][i];
            const W v11 = (fabsf(ge[2][i]) < 1e-12) ? W(0.0) : (W)ge[2][i];
            csl[i] = -(W)radx[i] / (W)rady[i];
            ca_[i] = W(1.0) / r_sqrt(W(1.0) + csl[i] * csl[i]);
            cb_[i] = csl[i] * ca_[i];
            const W a = ca_[i], b = cb_[i];
            // (RR*V*RR')[2,2] with RR = [a b; -b a]
            const W m1_0 = (-b) * v00 + a * v01;
            const W m1_1 = (-b) * v01 + a * v11;
            const W m2 = m1_0 * (-b) + m1_1 * a;
            cw[i] = W(1.0) / m2;
        }
        {
            S_ C[N + 1][N + 1];
            S_ ss[N], vb[N], w[N];
#pragma unroll
            for (int i = 0; i < N; ++i) {
                ss[i] = s[i];
                vb[i] = VBs[i];
                w[i] = cw[i];
            }
            {
                S_ Cn[N][N];
                band_matrix<S_, N>(w, ss, vb, Cn);
#pragma unroll
                for (int i = 0; i < N; ++i)
#pragma unroll
                    for (int j = 0; j < N; ++j) C[i][j] = Cn[i][j];
            }
            C[N][N] = S_(0.0);
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int I1 = i + 1;
                S_ cn = S_(0.0);
                if (I1 > 1 && I1 < N)
                    cn = cn + (-(ss[i + 1] - ss[i - 1])) * (ss[i + 1] - ss[i - 1]) /
                                  (S_(2.0) * vb[i] * (ss[i + 1] - ss[i]) * (ss[i] - ss[i - 1]));
                if (I1 > 2) cn = cn + (ss[i] - ss[i - 2]) / (S_(2.0) * vb[i - 1] * (ss[i] - ss[i - 1]));
                if (I1 < N - 1) cn = cn + (ss[i + 2] - ss[i]) / (S_(2.0) * vb[i + 1] * (ss[i + 1] - ss[i]));
                C[i][N] = cn;
                C[N][i] = cn;
                if (I1 > 1 && I1 < N) {
                    const S_ dd2 = ss[i + 1] - ss[i - 1];
                    C[N][N] = C[N][N] + (dd2 * dd2) / (S_(4.0) * vb[i]);
                }
            }
            invert_n(C, cI);
#pragma unroll
            for (int i = 0; i <= N; ++i) {
                S_ acc = cI[i][0] * (w[0] * (S_)Zc[0]);
#pragma unroll
                for (int k = 1; k < N; ++k) acc = acc + cI[i][k] * (w[k] * (S_)Zc[k]);
                acc = acc + cI[i][N] * S_(0.0);
                cu[i] = acc;
            }
            cchi2 = S_(0.0);
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const S_ diff = (S_)Zc[i] - cu[i];
                cchi2 = cchi2 + w[i] * (diff * diff);
                if (i > 0 && i < N - 1) {
                    const S_ dd1 = ss[i] - ss[i - 1], dd2 = ss[i + 1] - ss[i], dd3 = ss[i + 1] - ss[i - 1];
                    const S_ p1 = cu[i - 1] / dd1;
                    const S_ p2 = (cu[i] * dd3) / (dd1 * dd2);
                    const S_ p3 = cu[i + 1] / dd2;
                    const S_ p4 = (dd3 * cu[N]) / S_(2.0);
                    const S_ inner = p1 - p2 + p3 + p4;
                    cchi2 = cchi2 + (inner * inner) / vb[i];
                }
            }
        }
        n0x = (K)radx[0] / (K)rad[0];
        n0y = (K)rady[0] / (K)rad[0];
        n1x = (K)radx[1] / (K)rad[1];
        n1y = (K)rady[1] / (K)rad[1];
        {
            const K tmp1 = (K)(-(S_)Zc[0] + cu[0]);
            dpt1 = (K)hx[0] + tmp1 * n0x;
            dpt2 = (K)hy[0] + tmp1 * n0y;
            const K tmp2 = (K)(-(S_)Zc[1] + cu[1]);
            ept1 = (K)hx[1] + tmp2 * n1x;
            ept2 = (K)hy[1] + tmp2 * n1y;
        }
        em0 = ept1 - dpt1;
        em1 = ept2 - dpt2;
        cpar[0] = atan2_bl(em1, em0);
        nrm = r_sqrt(em0 * em0 + em1 * em1);
        // sagitta of the chord: R - sqrt(R^2 - |m|^2/4) cancels for stiff tracks
        arg = ff2 * ff2 - D(0.25) * ((D)nrm * (D)nrm);
        if (arg < D(0.0)) flags |= 1u;
        sarg = r_sqrt(arg);
        cpar1_d = (D)(-q) * (ff2 - sarg);
        cpar2_d = (D)q * (D(1.0) / ff2 + (D)cu[N]);
        cpar[1] = (K)cpar1_d;
        cpar[2] = (K)cpar2_d;
        if (!((D)q * cpar1_d <= D(0.0))) flags |= 2u;
        Q ccov[3][3];
        {
            const Q qem0 = (Q)em0, qem1 = (Q)em1, qn0x = (Q)n0x, qn0y = (Q)n0y, qn1x = (Q)n1x, qn1y = (Q)n1y;
            const Q t1 = (Q)nrm * (Q)nrm;
            const Q fq = (Q)((q >= 0) ? 0 : -1);  // floor(q / 2) for q = +-1
            Q J[3][3];
            J[0][0] = (qn0y * qem0 - qem1 * qn0x) / t1;
            J[0][1] = (qn1y * qem0 - qem1 * qn1x) / t1;
            J[0][2] = Q(0.0);
            const Q twoR = Q(2.0) * (Q)ff2;
            const Q arg2 = twoR * twoR - t1;
            if (arg2 < Q(0.0)) flags |= 1u;
            const Q den = r_sqrt(arg2);
            J[1][0] = fq * (qem0 * qn0x + qem1 * qn0y) / den;
            J[1][1] = fq * (qem0 * qn1x + qem1 * qn1y) / den;
            J[1][2] = Q(0.0);
            J[2][0] = Q(0.0);
            J[2][1] = Q(0.0);
            J[2][2] = (Q)q;
            Q cv[3][3] = {{(Q)cI[0][0], (Q)cI[0][1], (Q)cI[0][N]},
                          {(Q)cI[1][0], (Q)cI[1][1], (Q)cI[1][N]},
                          {(Q)cI[N][0], (Q)cI[N][1], (Q)cI[N][N]}};
            sandwich3(J, cv);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) ccov[i][j] = cv[i][j];
        }
        // the two translations (to the chord midpoint, then to the origin);
        // cos/sin of the angles from the vectors that define them when the
        // policy avoids the trigonometric calls (the same values in exact
        // arithmetic: atan2_bl returns the vector's angle modulo 2 pi)
        {
            K cp, sp;
            if (P::trig_free) {
                cp = em0 / nrm;
                sp = em1 / nrm;
            } else {
                cp = cos_r(cpar[0]);
                sp = sin_r(cpar[0]);
            }
            karimaki_par<K>(cpar, K(0.5) * em0, K(0.5) * em1, cp, sp, k1, flags);
        }
        karimaki_cov<Q>(k1, (Q)cpar[2], ccov);
        ms0 = mult_scatt<Q>((Q)(S[1] - S[0]), (Q)B, (Q)ff2, (Q)slope);
        ccov[0][0] = ccov[0][0] + (Q(1.0) + (Q)slope * (Q)slope) * ms0;
        {
            K cp, sp;
            if (P::trig_free) {
                const K hyp = r_sqrt(k1.BB * k1.BB + k1.Cc * k1.Cc);
                cp = k1.Cc / hyp;
                sp = k1.BB / hyp;
            } else {
                cp = cos_r(cpar[0]);
                sp = sin_r(cpar[0]);
            }
            karimaki_par<K>(cpar, dpt1, dpt2, cp, sp, k2, flags);
        }
        karimaki_cov<Q>(k2, (Q)cpar[2], ccov);

        // ---- copyFromCircle! (TrajectoryStateSOA.jl:20-44) and the kernelBLFit tail
        const Q b = (Q)(1.0 / B);
        o.state[0] = (Q)cpar[0];
        o.state[1] = (Q)cpar[1];
        o.state[2] = (Q)(cpar2_d * (D)(1.0 / B));
        o.cov[0] = ccov[0][0];
        o.cov[1] = ccov[0][1];
        o.cov[2] = b * ccov[0][2];
        o.cov[3] = Q(0.0);
        o.cov[4] = Q(0.0);
        o.cov[5] = ccov[1][1];
        o.cov[6] = b * ccov[1][2];
        o.cov[7] = Q(0.0);
        o.cov[8] = Q(0.0);
        o.cov[9] = (b * b) * ccov[2][2];
        o.cov[10] = Q(0.0);
        o.cov[11] = Q(0.0);
        o.pt = (Q)(B / r_abs(cpar2_d));
        o.chi2_line = (Q)lchi2;
        o.chi2_circle = (Q)cchi2;
        o.chi2 = (Q)((lchi2 + cchi2) / (S_)(2 * N - 5));
        o.flags = flags;
    }
};

// The reference's double-precision fit.
template <int N>
PT_HD void bl_fit_fp64(const float (&hx)[N], const float (&hy)[N], const float (&hz)[N], const float (&ge)[6][N],
                       double B, FitOut<double> &o) {
    BLFit<PolicyFP64, N> f;
    f.run(hx, hy, hz, ge, B, o);
}

// Float bulk, double fast fit, radii, band solve and sagitta.
template <int N>
PT_HD void bl_fit_mixed(const float (&hx)[N], const float (&hy)[N], const float (&hz)[N], const float (&ge)[6][N],
                        double B, FitOut<float> &o) {
    BLFit<PolicyMixed, N> f;
    f.run(hx, hy, hz, ge, B, o);
}

}  // namespace pt
