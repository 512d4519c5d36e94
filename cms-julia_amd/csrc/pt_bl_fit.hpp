// pt_bl_fit.hpp - broken-line helix fit, device code (one track per lane).
//
// Method of the reference fit: BL_Fast_fit -> prepare_broken_line_data ->
// BL_Line_fit -> BL_Circle_fit -> copyFromCircle!
// (broken_line.jl:46-696, cholesky_inversion.jl:22-129,
//  broken_line_fit_on_gpu.jl:24-163, TrajectoryStateSOA.jl:20-44).
// Everything lives in registers: N <= 4 hits, so the band systems are at most
// 4x4 (line) and 5x5 (circle, with the curvature row).  The operation order is
// the reference's (build with -ffp-contract=off), so the double version agrees
// with a serial double fit to rounding of the transcendental functions.
// Scalar type R = double gives the plain FP64 fit.  The mixed variant keeps
// the pre-fit curvature and the band solve in double and the rest in float.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

constexpr double kPiD = 3.141592653589793;

template <class R> struct FitOut {
    R state[5];
    R cov[15];
    R pt, eta, chi2;
    uint32_t flags;  // 1: sqrt of a negative number, 2: q*d > 0, 4: NaN pre-fit
};

template <class R> __device__ __forceinline__ R r_sqrt(R x) { return sqrt(x); }
template <> __device__ __forceinline__ float r_sqrt<float>(float x) { return sqrtf(x); }

// broken_line.jl:256-268 (x < 0 gives atan(y/x) + pi, also for y < 0)
__device__ __forceinline__ double atan2_bl(double y, double x) {
    if (x > 0.0) return atan(y / x);
    if (x < 0.0) return atan(y / x) + kPiD;
    if (y > 0.0 && x == 0.0) return kPiD / 2.0;
    if (y < 0.0 && x == 0.0) return -kPiD / 2.0;
    return __builtin_nan("");
}
__device__ __forceinline__ float atan2_bl(float y, float x) {
    if (x > 0.0f) return atanf(y / x);
    if (x < 0.0f) return atanf(y / x) + (float)kPiD;
    if (y > 0.0f && x == 0.0f) return (float)(kPiD / 2.0);
    if (y < 0.0f && x == 0.0f) return (float)(-kPiD / 2.0);
    return __builtin_nanf("");
}
__device__ __forceinline__ double log_r(double x) { return log(x); }
__device__ __forceinline__ float log_r(float x) { return logf(x); }
__device__ __forceinline__ double cos_r(double x) { return cos(x); }
__device__ __forceinline__ float cos_r(float x) { return cosf(x); }
__device__ __forceinline__ double sin_r(double x) { return sin(x); }
__device__ __forceinline__ float sin_r(float x) { return sinf(x); }

template <class R> __device__ __forceinline__ R jl_min(R a, R b) {
    if (a != a) return a;
    if (b != b) return b;
    return (b < a) ? b : a;
}

// broken_line.jl:46-79
template <class R> __device__ __forceinline__ R mult_scatt(R length, R B, R Rr, R slope) {
    R pt2 = jl_min(R(20.0), B * Rr);
    pt2 = pt2 * pt2;
    const R XXI_0 = R(0.06 / 16.0);
    const R fact = R(0.7 * ((13.6 / 1000) * (13.6 / 1000)));
    const R slope_term = R(1.0) + slope * slope;
    const R al = fabs(length) * XXI_0;
    const R sq = R(1.0) + R(0.038) * log_r(al);
    return fact / (pt2 * slope_term) * al * (sq * sq);
}

// cholesky_inversion.jl:22-129; lower triangle computed, then symmetrised
template <class R> __device__ __forceinline__ void invert3(const R (&s)[3][3], R (&d)[3][3]) {
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
template <class R> __device__ __forceinline__ void invert4(const R (&s)[4][4], R (&d)[4][4]) {
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
template <class R> __device__ __forceinline__ void invert5(const R (&s)[5][5], R (&d)[5][5]) {
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
template <class R> __device__ __forceinline__ void invert_n(const R (&s)[3][3], R (&d)[3][3]) { invert3(s, d); }
template <class R> __device__ __forceinline__ void invert_n(const R (&s)[4][4], R (&d)[4][4]) { invert4(s, d); }
template <class R> __device__ __forceinline__ void invert_n(const R (&s)[5][5], R (&d)[5][5]) { invert5(s, d); }

// broken_line.jl:217-250 ; band matrix of the broken-line cost (C + C')
template <class R, int N>
__device__ __forceinline__ void band_matrix(const R (&w)[N], const R (&S)[N], const R (&VB)[N], R (&out)[N][N]) {
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
template <class R> __device__ __forceinline__ void sandwich2(const R (&J)[2][2], R (&cov)[2][2]) {
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

// translate_karimaki (broken_line.jl:103-132) on (par, cov), Jacobian J*cov*J'
template <class R>
__device__ __forceinline__ void karimaki(R (&par)[3], R (&cov)[3][3], R x0, R y0, uint32_t &flags) {
    const R cp = cos_r(par[0]), sp = sin_r(par[0]);
    const R DP = x0 * cp + y0 * sp;
    const R DO = x0 * sp - y0 * cp + par[1];
    const R uu = R(1.0) + par[2] * par[1];
    const R Cc = -par[2] * y0 + uu * cp;
    const R BB = par[2] * x0 + uu * sp;
    const R A = R(2.0) * DO + par[2] * (DO * DO + DP * DP);
    const R arg = R(1.0) + par[2] * A;
    if (arg < R(0.0)) flags |= 1u;
    const R U = r_sqrt(arg);
    const R xi = R(1.0) / (BB * BB + Cc * Cc);
    const R v = R(1.0) + par[2] * DO;
    const R op = R(1.0) + U;
    const R lambda = (R(0.5) * A) / (U * (op * op));
    const R mu = R(1.0) / (U * (R(1.0) + U)) + par[2] * lambda;
    const R zeta = DO * DO + DP * DP;
    R J[3][3];
    J[0][0] = xi * uu * v;
    J[0][1] = -xi * (par[2] * par[2]) * DP;
    J[0][2] = xi * DP;
    J[1][0] = R(2.0) * mu * uu * DP;
    J[1][1] = R(2.0) * mu * v;
    J[1][2] = mu * zeta - lambda * A;
    J[2][0] = R(0.0);
    J[2][1] = R(0.0);
    J[2][2] = R(1.0);
    par[0] = atan2_bl(BB, Cc);
    par[1] = A / (R(1.0) + U);
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

// Full fit of one track of N hits.  hx/hy/hz: positions, ge[6][N]: global
// covariance (xx,xy,yy,xz,yz,zz) as float, B: field (1/GeV units).
template <int N>
__device__ __forceinline__ void bl_fit_fp64(const double (&hx)[N], const double (&hy)[N], const double (&hz)[N],
                                            const float (&ge)[6][N], double B, FitOut<double> &o) {
    typedef double R;
    uint32_t flags = 0;
    // ---- BL_Fast_fit (broken_line.jl:275-313)
    constexpr int mid = (N >> 1);
    R ff0, ff1, ff2, ff3;
    {
        const R h11 = hx[0], h21 = hy[0], h1m = hx[mid], h2m = hy[mid], h1n = hx[N - 1], h2n = hy[N - 1];
        const R a1 = h1m - h11, a2 = h2m - h21;
        const R b1 = h1n - h1m, b2 = h2n - h2m;
        const R c1 = h11 - h1n, c2 = h21 - h2n;
        const R ca = c1 * a2 - c2 * a1;
        const R ba = b1 * a2 - b2 * a1;
        const R nc = c1 * c1 + c2 * c2;
        const R na = a1 * a1 + a2 * a2;
        const R nb = b1 * b1 + b2 * b2;
        const R tmp = 0.5 / ca;
        ff0 = h11 - (a2 * nc + c2 * na) * tmp;
        ff1 = h21 + (a1 * nc + c1 * na) * tmp;
        ff2 = sqrt(na * nb * nc) / (2.0 * fabs(ba));
        const R d1 = h11 - ff0, d2 = h21 - ff1;
        const R e1 = h1n - ff0, e2 = h2n - ff1;
        const R de = d1 * e2 - d2 * e1;
        const R dotde = d1 * e1 + d2 * e2;
        const R dz = hz[N - 1] - hz[0];
        ff3 = ff2 * atan2_bl(de, dotde) / dz;
        if (ff0 != ff0 || ff1 != ff1 || ff2 != ff2 || ff3 != ff3) flags |= 4u;
    }
    // ---- prepare_broken_line_data (broken_line.jl:150-207)
    int q;
    R radx[N], rady[N], s[N], S[N], Z[N], VB[N];
    {
        const R x1 = hx[1] - hx[0], y1 = hy[1] - hy[0];
        const R xn = hx[N - 1] - hx[N - 2], yn = hy[N - 1] - hy[N - 2];
        q = (x1 * yn - y1 * xn > 0.0) ? -1 : 1;
    }
    const R slope = (R)(-q) / ff3;
    const R ra = 1.0 / sqrt(1.0 + slope * slope);
    const R rb = slope * ra;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        radx[j] = hx[j] - ff0;
        rady[j] = hy[j] - ff1;
    }
    {
        const R rr = sqrt(ff0 * ff0 + ff1 * ff1);
        const R sc = -ff2 / rr;
        const R e1 = sc * ff0, e2 = sc * ff1;
        const R fac = (R)q * ff2;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const R c = radx[i] * e2 - rady[i] * e1;
            const R d = radx[i] * e1 + rady[i] * e2;
            s[i] = fac * atan2_bl(c, d);
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        S[i] = ra * s[i] + rb * hz[i];
        Z[i] = (-rb) * s[i] + ra * hz[i];
    }
    VB[0] = 0.0;
    VB[N - 1] = 0.0;
#pragma unroll
    for (int i = 1; i < N - 1; ++i)
        VB[i] = mult_scatt<R>(S[i + 1] - S[i], B, ff2, slope) + mult_scatt<R>(S[i] - S[i - 1], B, ff2, slope);

    // ---- BL_Line_fit (broken_line.jl:546-696)
    R lpar0, lpar1, lc00, lc01, lc11, lchi2;
    {
        R w[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const R v00 = ge[0][i];
            const R v01 = (fabsf(ge[1][i]) < 1e-12) ? 0.0 : (R)ge[1][i];
            const R v02 = (fabsf(ge[3][i]) < 1e-12) ? 0.0 : (R)ge[3][i];
            const R v11 = ge[2][i];
            const R v12 = (fabsf(ge[4][i]) < 1e-12) ? 0.0 : (R)ge[4][i];
            const R v22 = ge[5][i];
            const R tmp = 1.0 / sqrt(radx[i] * radx[i] + rady[i] * rady[i]);
            const R j00 = rady[i] * tmp, j01 = -radx[i] * tmp;  // Jacob row 1 = (j00, j01, 0); row 2 = (0,0,1)
            // M1 = Rm * J  (Rm = [ra rb; -rb ra]); only row 2 of the chain is needed
            const R m1_0 = (-rb) * j00 + ra * 0.0;
            const R m1_1 = (-rb) * j01 + ra * 0.0;
            const R m1_2 = (-rb) * 0.0 + ra * 1.0;
            // M2 = M1 * V
            const R m2_0 = m1_0 * v00 + m1_1 * v01 + m1_2 * v02;
            const R m2_1 = m1_0 * v01 + m1_1 * v11 + m1_2 * v12;
            const R m2_2 = m1_0 * v02 + m1_1 * v12 + m1_2 * v22;
            // M3 = M2 * J'
            const R m3_0 = m2_0 * j00 + m2_1 * j01 + m2_2 * 0.0;
            const R m3_1 = m2_0 * 0.0 + m2_1 * 0.0 + m2_2 * 1.0;
            // M4[2][2] = M3 * Rm' : Rm'[0][1] = -rb, Rm'[1][1] = ra
            const R m4 = m3_0 * (-rb) + m3_1 * ra;
            w[i] = 1.0 / m4;
        }
        R C[N][N], I[N][N];
        band_matrix<R, N>(w, S, VB, C);
        invert_n(C, I);
        R u[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            R acc = I[i][0] * (w[0] * Z[0]);
#pragma unroll
            for (int k = 1; k < N; ++k) acc = acc + I[i][k] * (w[k] * Z[k]);
            u[i] = acc;
        }
        lpar0 = (u[1] - u[0]) / (S[1] - S[0]);
        lpar1 = u[0];
        const R idiff = 1.0 / (S[1] - S[0]);
        R cv[2][2];
        cv[0][0] = (I[0][0] - 2.0 * I[0][1] + I[1][1]) * (idiff * idiff) + mult_scatt<R>(S[1] - S[0], B, ff2, slope);
        cv[0][1] = (I[0][1] - I[0][0]) * idiff;
        cv[1][0] = (I[0][1] - I[0][0]) * idiff;
        cv[1][1] = I[0][0];
        lpar1 = lpar1 - lpar0 * S[0];
        {
            const R J[2][2] = {{1.0, 0.0}, {-S[0], 1.0}};
            sandwich2(J, cv);
        }
        const R tmp = ra - lpar0 * rb;
        const R j11 = 1.0 / tmp;
        const R j00 = j11 * j11;
        const R j10 = lpar1 * rb * j00;
        lpar1 = lpar1 * j11;
        lpar0 = (rb + lpar0 * ra) * j11;
        {
            const R J[2][2] = {{j00, 0.0}, {j10, j11}};
            sandwich2(J, cv);
        }
        lc00 = cv[0][0];
        lc01 = cv[0][1];
        lc11 = cv[1][1];
        lchi2 = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const R diff = Z[i] - u[i];
            lchi2 = lchi2 + w[i] * (diff * diff);
            if (i > 0 && i < N - 1) {
                const R d1 = S[i] - S[i - 1], d2 = S[i + 1] - S[i], d4 = S[i + 1] - S[i - 1];
                const R inner = (u[i - 1] / d1) - (u[i] * (d4 / (d2 * d1))) + (u[i + 1] / (S[i + 1] - S[i]));
                lchi2 = lchi2 + (inner * inner) / VB[i];
            }
        }
    }

    // ---- BL_Circle_fit (broken_line.jl:330-528)
    R cpar[3], ccov[3][3], cchi2;
    {
        const R vscale = 1.0 + slope * slope;
        R VBs[N], w[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            VBs[i] = VB[i] * vscale;
            Z[i] = sqrt(radx[i] * radx[i] + rady[i] * rady[i]) - ff2;
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const R v00 = (fabsf(ge[0][i]) < 1e-12) ? 0.0 : (R)ge[0][i];
            const R v01 = (fabsf(ge[1][i]) < 1e-12) ? 0.0 : (R)ge[1][i];
            const R v11 = (fabsf(ge[2][i]) < 1e-12) ? 0.0 : (R)ge[2][i];
            const R sl = -radx[i] / rady[i];
            const R a = 1.0 / sqrt(1.0 + sl * sl);
            const R b = sl * a;
            // (RR*V*RR')[2,2] with RR = [a b; -b a]
            const R m1_0 = (-b) * v00 + a * v01;
            const R m1_1 = (-b) * v01 + a * v11;
            const R m2 = m1_0 * (-b) + m1_1 * a;
            w[i] = 1.0 / m2;
        }
        R C[N + 1][N + 1];
        {
            R Cn[N][N];
            band_matrix<R, N>(w, s, VBs, Cn);
#pragma unroll
            for (int i = 0; i < N; ++i)
#pragma unroll
                for (int j = 0; j < N; ++j) C[i][j] = Cn[i][j];
        }
        C[N][N] = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const int I1 = i + 1;
            R cn = 0.0;
            if (I1 > 1 && I1 < N)
                cn = cn + (-(s[i + 1] - s[i - 1])) * (s[i + 1] - s[i - 1]) /
                              (2.0 * VBs[i] * (s[i + 1] - s[i]) * (s[i] - s[i - 1]));
            if (I1 > 2) cn = cn + (s[i] - s[i - 2]) / (2.0 * VBs[i - 1] * (s[i] - s[i - 1]));
            if (I1 < N - 1) cn = cn + (s[i + 2] - s[i]) / (2.0 * VBs[i + 1] * (s[i + 1] - s[i]));
            C[i][N] = cn;
            C[N][i] = cn;
            if (I1 > 1 && I1 < N) {
                const R dd = s[i + 1] - s[i - 1];
                C[N][N] = C[N][N] + (dd * dd) / (4.0 * VBs[i]);
            }
        }
        R I[N + 1][N + 1];
        invert_n(C, I);
        R u[N + 1];
#pragma unroll
        for (int i = 0; i <= N; ++i) {
            R acc = I[i][0] * (w[0] * Z[0]);
#pragma unroll
            for (int k = 1; k < N; ++k) acc = acc + I[i][k] * (w[k] * Z[k]);
            acc = acc + I[i][N] * 0.0;
            u[i] = acc;
        }
        R n0x, n0y, n1x, n1y;
        {
            const R t0 = sqrt(radx[0] * radx[0] + rady[0] * rady[0]);
            n0x = radx[0] / t0;
            n0y = rady[0] / t0;
            const R t1 = sqrt(radx[1] * radx[1] + rady[1] * rady[1]);
            n1x = radx[1] / t1;
            n1y = rady[1] / t1;
        }
        const R tmp1 = -Z[0] + u[0];
        const R d1 = hx[0] + tmp1 * n0x, d2 = hy[0] + tmp1 * n0y;
        const R tmp2 = -Z[1] + u[1];
        const R e1 = hx[1] + tmp2 * n1x, e2 = hy[1] + tmp2 * n1y;
        const R em0 = e1 - d1, em1 = e2 - d2;
        cpar[0] = atan2_bl(em1, em0);
        const R nrm = sqrt(em0 * em0 + em1 * em1);
        const R arg = ff2 * ff2 - 0.25 * (nrm * nrm);
        if (arg < 0.0) flags |= 1u;
        cpar[1] = (R)(-q) * (ff2 - sqrt(arg));
        cpar[2] = (R)q * (1.0 / ff2 + u[N]);
        if (!((R)q * cpar[1] <= 0.0)) flags |= 2u;
        const R t1 = nrm * nrm;
        const R fq = floor((R)q / 2.0);
        R J[3][3];
        J[0][0] = (n0y * em0 - em1 * n0x) / t1;
        J[0][1] = (n1y * em0 - em1 * n1x) / t1;
        J[0][2] = 0.0;
        const R twoR = 2.0 * ff2;
        const R arg2 = twoR * twoR - t1;
        if (arg2 < 0.0) flags |= 1u;
        const R den = sqrt(arg2);
        J[1][0] = fq * (em0 * n0x + em1 * n0y) / den;
        J[1][1] = fq * (em0 * n1x + em1 * n1y) / den;
        J[1][2] = 0.0;
        J[2][0] = 0.0;
        J[2][1] = 0.0;
        J[2][2] = (R)q;
        const R cv[3][3] = {{I[0][0], I[0][1], I[0][N]}, {I[1][0], I[1][1], I[1][N]}, {I[N][0], I[N][1], I[N][N]}};
        R T[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) T[i][j] = J[i][0] * cv[0][j] + J[i][1] * cv[1][j] + J[i][2] * cv[2][j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) ccov[i][j] = T[i][0] * J[j][0] + T[i][1] * J[j][1] + T[i][2] * J[j][2];
        karimaki<R>(cpar, ccov, 0.5 * em0, 0.5 * em1, flags);
        ccov[0][0] = ccov[0][0] + (1.0 + slope * slope) * mult_scatt<R>(S[1] - S[0], B, ff2, slope);
        karimaki<R>(cpar, ccov, d1, d2, flags);
        cchi2 = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const R diff = Z[i] - u[i];
            cchi2 = cchi2 + w[i] * (diff * diff);
            if (i > 0 && i < N - 1) {
                const R dd1 = s[i] - s[i - 1], dd2 = s[i + 1] - s[i], dd3 = s[i + 1] - s[i - 1];
                const R p1 = u[i - 1] / dd1;
                const R p2 = (u[i] * dd3) / (dd1 * dd2);
                const R p3 = u[i + 1] / dd2;
                const R p4 = (dd3 * u[N]) / 2.0;
                const R inner = p1 - p2 + p3 + p4;
                cchi2 = cchi2 + (inner * inner) / VBs[i];
            }
        }
    }
    // ---- copyFromCircle! (TrajectoryStateSOA.jl:20-44) and the kernelBLFit tail
    const R b = 1.0 / B;
    o.state[0] = cpar[0];
    o.state[1] = cpar[1];
    o.state[2] = cpar[2] * b;
    o.state[3] = lpar0;
    o.state[4] = lpar1;
    o.cov[0] = ccov[0][0];
    o.cov[1] = ccov[0][1];
    o.cov[2] = b * ccov[0][2];
    o.cov[3] = 0.0;
    o.cov[4] = 0.0;
    o.cov[5] = ccov[1][1];
    o.cov[6] = b * ccov[1][2];
    o.cov[7] = 0.0;
    o.cov[8] = 0.0;
    o.cov[9] = (b * b) * ccov[2][2];
    o.cov[10] = 0.0;
    o.cov[11] = 0.0;
    o.cov[12] = lc00;
    o.cov[13] = lc01;
    o.cov[14] = lc11;
    o.pt = B / fabs(cpar[2]);
    o.eta = asinh(lpar0);
    o.chi2 = (lchi2 + cchi2) / (R)(2 * N - 5);
    o.flags = flags;
}

}  // namespace pt
