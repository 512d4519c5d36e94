// oracle_fit.hpp - TEST INFRASTRUCTURE (the oracle), never linked into the product.
//
// Serial restatement of the reference broken-line fit, templated on the scalar
// type T so the same code runs in double (the FP64 oracle), with an
// op-counting scalar (flop counts) and with a forward-mode dual number (the
// ForwardDiff substitute for the adjoint checks).
//
// Follows, statement by statement and in the same association order (Julia
// never contracts a*b+c into an FMA; build with -ffp-contract=off):
//   broken_line.jl:46-79     mult_scatt
//   broken_line.jl:89-95     rotation_matrix
//   broken_line.jl:103-132   translate_karimaki
//   broken_line.jl:143-207   compute_e, prepare_broken_line_data
//   broken_line.jl:217-250   matrixc_u
//   broken_line.jl:256-268   atan2 (custom quadrant logic: x<0 -> atan(y/x)+pi)
//   broken_line.jl:275-313   BL_Fast_fit
//   broken_line.jl:330-528   BL_Circle_fit
//   broken_line.jl:546-696   BL_Line_fit
//   cholesky_inversion.jl:22-129,193-256  invert33/44/55 + symmetrize
//   broken_line_fit_on_gpu.jl:71-85,150-159  gather, toGlobal, output
//   TrajectoryStateSOA.jl:20-44  copyFromCircle!
//   SOARotation.jl:139-147   toGlobal (error rotation, float)
// Parity notes: Julia's libm (atan, log, sin, cos, asinh) and StaticArrays'
// small-matrix products may differ from this restatement in the last ulp;
// everything else is exact restatement.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

namespace oracle {

constexpr double kPi = 3.141592653589793;  // Float64(pi)

// Julia's Base.min for floats: NaN propagates.
template <class T> inline T jl_min(T a, T b) {
    using std::isnan;
    if (isnan(a)) return a;
    if (isnan(b)) return b;
    return (b < a) ? b : a;
}

// Flags raised where Julia would throw (DomainError on sqrt of a negative,
// @assert failures).  Collected per track.
struct FitFlags {
    uint32_t bits = 0;
};
constexpr uint32_t kFlagSqrtDomain = 1u;
constexpr uint32_t kFlagCharge = 2u;
constexpr uint32_t kFlagFastNaN = 4u;

template <class T> inline T checked_sqrt(T x, FitFlags &f) {
    using std::sqrt;
    if (x < 0.0) f.bits |= kFlagSqrtDomain;
    return sqrt(x);
}

// broken_line.jl:256-268
template <class T> inline T atan2_ref(T y, T x) {
    using std::atan;
    if (x > 0.0) return atan(y / x);
    if (x < 0.0) return atan(y / x) + kPi;
    if (y > 0.0 && x == 0.0) return T(kPi / 2.0);
    if (y < 0.0 && x == 0.0) return T(-kPi / 2.0);
    return T(std::numeric_limits<double>::quiet_NaN());
}

// Planted mistakes in mult_scatt, set only by the tests that show the
// scattering pins catch them (default: none).
struct MsPlant {
    double scale = 1.0;   // result multiplied by this
    bool log10 = false;   // log10 in place of the natural log
    bool no_cap = false;  // no 20 GeV cap on B*R
};
inline MsPlant &ms_plant() {
    static MsPlant p;
    return p;
}

// broken_line.jl:46-79
template <class T> inline T mult_scatt(T length, double B, T R, int /*layer*/, T slope) {
    using std::abs;
    using std::log;
    const MsPlant &plant = ms_plant();
    T pt2 = plant.no_cap ? T(B * R) : jl_min(T(20.0), B * R);
    pt2 = pt2 * pt2;
    const double XXI_0 = 0.06 / 16.0;
    const double geometry_factor = 0.7;
    const double c = 13.6 / 1000;
    const double fact = geometry_factor * (c * c);
    T slope_term = 1.0 + slope * slope;
    T abs_length_XXI_0 = abs(length) * XXI_0;
    T log_term = log(abs_length_XXI_0);
    if (plant.log10) log_term = log_term / std::log(10.0);
    T sq = 1.0 + 0.038 * log_term;
    T sqrt_term = sq * sq;
    T r = fact / (pt2 * slope_term) * abs_length_XXI_0 * sqrt_term;
    if (plant.scale != 1.0) r = r * plant.scale;
    return r;
}

// broken_line.jl:89-95 ; R = [a b; -b a]
template <class T> struct Rot2 {
    T a, b;  // R11 = a, R12 = b, R21 = -b, R22 = a
};
template <class T> inline Rot2<T> rotation_matrix(T slope) {
    using std::sqrt;
    T a = 1.0 / sqrt(1.0 + slope * slope);
    T b = slope * a;
    return {a, b};
}

template <class T> struct CircleFit {
    T par[3];
    T cov[3][3];
    int q;
    T chi2;
};
template <class T> struct LineFit {
    T par[2];
    T cov[2][2];
    T chi2;
};
template <class T, int N> struct Prepared {
    int q;
    T radii[2][N];
    T s[N], S[N], Z[N], VarBeta[N];
};

// Dense small-matrix product in the Julia (A*B)*C association, each element a
// left-to-right sum over k.
template <class T, int R, int K, int C>
inline void matmul(const T (&A)[R][K], const T (&B)[K][C], T (&out)[R][C]) {
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < C; ++j) {
            T acc = A[i][0] * B[0][j];
            for (int k = 1; k < K; ++k) acc = acc + A[i][k] * B[k][j];
            out[i][j] = acc;
        }
}
template <class T, int R, int C>
inline void transpose(const T (&A)[R][C], T (&out)[C][R]) {
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < C; ++j) out[j][i] = A[i][j];
}
// out = J * M * J'
template <class T, int R, int C>
inline void sandwich(const T (&J)[R][C], const T (&M)[C][C], T (&out)[R][R]) {
    T JM[R][C];
    matmul(J, M, JM);
    T Jt[C][R];
    transpose(J, Jt);
    matmul(JM, Jt, out);
}

// broken_line.jl:103-132
template <class T>
inline void translate_karimaki(CircleFit<T> &circle, T x0, T y0, T (&jacobian)[3][3], FitFlags &fl) {
    using std::cos;
    using std::sin;
    T DP = x0 * cos(circle.par[0]) + y0 * sin(circle.par[0]);
    T DO = x0 * sin(circle.par[0]) - y0 * cos(circle.par[0]) + circle.par[1];
    T uu = 1.0 + circle.par[2] * circle.par[1];
    T C = -circle.par[2] * y0 + uu * cos(circle.par[0]);
    T BB = circle.par[2] * x0 + uu * sin(circle.par[0]);
    T A = 2.0 * DO + circle.par[2] * (DO * DO + DP * DP);
    T U = checked_sqrt(1.0 + circle.par[2] * A, fl);
    T xi = 1.0 / (BB * BB + C * C);
    T v = 1.0 + circle.par[2] * DO;
    T onepu = 1.0 + U;
    T lambda = (0.5 * A) / (U * (onepu * onepu));
    T mu = 1.0 / (U * (1.0 + U)) + circle.par[2] * lambda;
    T zeta = DO * DO + DP * DP;

    jacobian[0][0] = xi * uu * v;
    jacobian[0][1] = -xi * (circle.par[2] * circle.par[2]) * DP;
    jacobian[0][2] = xi * DP;
    jacobian[1][0] = 2.0 * mu * uu * DP;
    jacobian[1][1] = 2.0 * mu * v;
    jacobian[1][2] = mu * zeta - lambda * A;
    jacobian[2][0] = 0.0;
    jacobian[2][1] = 0.0;
    jacobian[2][2] = 1.0;

    circle.par[0] = atan2_ref(BB, C);
    circle.par[1] = A / (1.0 + U);
    T out[3][3];
    sandwich(jacobian, circle.cov, out);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) circle.cov[i][j] = out[i][j];
}

// broken_line.jl:275-313 ; hits[c][k], c = x,y,z
template <class T, int N>
inline void BL_Fast_fit(const T (&hits)[3][N], T (&res)[4]) {
    using std::abs;
    using std::sqrt;
    constexpr int n = N;
    constexpr int mid = (n >> 1) + 1;  // 1-based
    T h11 = hits[0][0], h21 = hits[1][0];
    T h1m = hits[0][mid - 1], h2m = hits[1][mid - 1];
    T h1n = hits[0][n - 1], h2n = hits[1][n - 1];
    T a1 = h1m - h11, a2 = h2m - h21;
    T b1 = h1n - h1m, b2 = h2n - h2m;
    T c1 = h11 - h1n, c2 = h21 - h2n;
    T ca = c1 * a2 - c2 * a1;
    T ba = b1 * a2 - b2 * a1;
    T nc = c1 * c1 + c2 * c2;
    T na = a1 * a1 + a2 * a2;
    T nb = b1 * b1 + b2 * b2;
    T tmp = 0.5 / ca;
    res[0] = h11 - (a2 * nc + c2 * na) * tmp;
    res[1] = h21 + (a1 * nc + c1 * na) * tmp;
    res[2] = sqrt(na * nb * nc) / (2.0 * abs(ba));
    T cx = res[0], cy = res[1];
    T d1 = h11 - cx, d2 = h21 - cy;
    T e1 = h1n - cx, e2 = h2n - cy;
    T de = d1 * e2 - d2 * e1;
    T dotde = d1 * e1 + d2 * e2;
    T dz = hits[2][n - 1] - hits[2][0];
    res[3] = res[2] * atan2_ref(de, dotde) / dz;
}

// broken_line.jl:150-207
template <class T, int N>
inline void prepare_broken_line_data(const T (&hits)[3][N], const T (&ff)[4], double B,
                                     Prepared<T, N> &r) {
    using std::sqrt;
    constexpr int n = N;
    T f1 = ff[0], f2 = ff[1], f3 = ff[2], f4 = ff[3];
    T x1 = hits[0][1] - hits[0][0];
    T y1 = hits[1][1] - hits[1][0];
    T xn = hits[0][n - 1] - hits[0][n - 2];
    T yn = hits[1][n - 1] - hits[1][n - 2];
    r.q = (x1 * yn - y1 * xn > 0.0) ? -1 : 1;
    T slope = double(-r.q) / f4;
    Rot2<T> R = rotation_matrix(slope);
    for (int j = 0; j < n; ++j) {
        r.radii[0][j] = hits[0][j] - f1;
        r.radii[1][j] = hits[1][j] - f2;
    }
    // compute_e (broken_line.jl:143-148)
    T rr = sqrt(f1 * f1 + f2 * f2);
    T sc = -f3 / rr;
    T e1 = sc * f1, e2 = sc * f2;
    T fac = double(r.q) * f3;
    for (int i = 0; i < n; ++i) {
        T xi = r.radii[0][i], yi = r.radii[1][i];
        T c = xi * e2 - yi * e1;
        T d = xi * e1 + yi * e2;
        r.s[i] = fac * atan2_ref(c, d);
    }
    for (int i = 0; i < n; ++i) {
        T xi = r.s[i], yi = hits[2][i];
        r.S[i] = R.a * xi + R.b * yi;
        r.Z[i] = (-R.b) * xi + R.a * yi;
    }
    r.VarBeta[0] = 0.0;
    r.VarBeta[n - 1] = 0.0;
    for (int i = 1; i < n - 1; ++i) {
        T d1 = r.S[i + 1] - r.S[i];
        T d2 = r.S[i] - r.S[i - 1];
        r.VarBeta[i] = mult_scatt(d1, B, f3, i + 3, slope) + mult_scatt(d2, B, f3, i + 2, slope);
    }
}

// broken_line.jl:217-250 (0-based i; the reference's i is i+1)
template <class T, int N>
inline void matrixc_u(const T *w, const T *S, const T *VarBeta, T (&out)[N][N]) {
    constexpr int n = N;
    T C[N][N];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) C[i][j] = T(0.0);
    for (int i = 0; i < n; ++i) {
        const int I = i + 1;  // 1-based
        C[i][i] = w[i];
        if (I > 2) {
            T d = S[i] - S[i - 1];
            C[i][i] = C[i][i] + 1.0 / (VarBeta[i - 1] * (d * d));
        }
        if (I > 1 && I < n) {
            T q = (S[i + 1] - S[i - 1]) / ((S[i + 1] - S[i]) * (S[i] - S[i - 1]));
            C[i][i] = C[i][i] + (1.0 / VarBeta[i]) * (q * q);
        }
        if (I < n - 1) {
            T d = S[i + 1] - S[i];
            C[i][i] = C[i][i] + 1.0 / (VarBeta[i + 1] * (d * d));
        }
        if (I > 1 && I < n) {
            C[i][i + 1] = 1.0 / (VarBeta[i] * (S[i + 1] - S[i])) *
                          (-(S[i + 1] - S[i - 1]) / ((S[i + 1] - S[i]) * (S[i] - S[i - 1])));
        }
        if (I < n - 1) {
            C[i][i + 1] = C[i][i + 1] + 1.0 / (VarBeta[i + 1] * (S[i + 1] - S[i])) *
                                            (-(S[i + 2] - S[i]) / ((S[i + 2] - S[i + 1]) * (S[i + 1] - S[i])));
        }
        if (I < n - 1) {
            C[i][i + 2] = 1.0 / (VarBeta[i + 1] * (S[i + 2] - S[i + 1]) * (S[i + 1] - S[i]));
        }
        C[i][i] = C[i][i] * 0.5;
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) out[i][j] = C[i][j] + C[j][i];
}

// cholesky_inversion.jl:22-43 (+ symmetrize33!)
template <class T> inline void invert33(const T (&s)[3][3], T (&d)[3][3]) {
    T luc0 = 1.0 / s[0][0];
    T luc1 = s[1][0];
    T luc2 = s[1][1] - luc0 * luc1 * luc1;
    luc2 = 1.0 / luc2;
    T luc3 = s[2][0];
    T luc4 = s[2][1] - luc0 * luc1 * luc3;
    T luc5 = s[2][2] - (luc0 * luc3 * luc3 + luc2 * luc4 * luc4);
    luc5 = 1.0 / luc5;
    T li21 = -luc0 * luc1;
    T li32 = -luc2 * luc4;
    T li31 = (luc1 * luc2 * luc4 - luc3) * luc0;
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

// cholesky_inversion.jl:45-78 (+ symmetrize44!)
template <class T> inline void invert44(const T (&s)[4][4], T (&d)[4][4]) {
    T luc0 = 1.0 / s[0][0];
    T luc1 = s[1][0];
    T luc2 = s[1][1] - luc0 * luc1 * luc1;
    luc2 = 1.0 / luc2;
    T luc3 = s[2][0];
    T luc4 = s[2][1] - luc0 * luc1 * luc3;
    T luc5 = s[2][2] - (luc0 * luc3 * luc3 + luc2 * luc4 * luc4);
    luc5 = 1.0 / luc5;
    T luc6 = s[3][0];
    T luc7 = s[3][1] - luc0 * luc1 * luc6;
    T luc8 = s[3][2] - luc0 * luc3 * luc6 - luc2 * luc4 * luc7;
    T luc9 = s[3][3] - (luc0 * luc6 * luc6 + luc2 * luc7 * luc7 + luc8 * luc8 * luc5);
    luc9 = 1.0 / luc9;
    T li21 = -luc0 * luc1;
    T li32 = -luc2 * luc4;
    T li31 = (luc1 * luc2 * luc4 - luc3) * luc0;
    T li43 = -luc8 * luc5;
    T li42 = (luc4 * luc8 * luc5 - luc7) * luc2;
    T li41 = (-luc1 * luc2 * luc4 * luc8 * luc5 + luc1 * luc2 * luc7 + luc3 * luc8 * luc5 - luc6) * luc0;
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
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) d[i][j] = d[j][i];
}

// cholesky_inversion.jl:79-129 (+ symmetrize55!)
template <class T> inline void invert55(const T (&s)[5][5], T (&d)[5][5]) {
    T luc0 = 1.0 / s[0][0];
    T luc1 = s[1][0];
    T luc2 = s[1][1] - luc0 * luc1 * luc1;
    luc2 = 1.0 / luc2;
    T luc3 = s[2][0];
    T luc4 = s[2][1] - luc0 * luc1 * luc3;
    T luc5 = s[2][2] - (luc0 * luc3 * luc3 + luc2 * luc4 * luc4);
    luc5 = 1.0 / luc5;
    T luc6 = s[3][0];
    T luc7 = s[3][1] - luc0 * luc1 * luc6;
    T luc8 = s[3][2] - luc0 * luc3 * luc6 - luc2 * luc4 * luc7;
    T luc9 = s[3][3] - (luc0 * luc6 * luc6 + luc2 * luc7 * luc7 + luc8 * luc8 * luc5);
    luc9 = 1.0 / luc9;
    T luc10 = s[4][0];
    T luc11 = s[4][1] - luc0 * luc1 * luc10;
    T luc12 = s[4][2] - luc0 * luc3 * luc10 - luc2 * luc4 * luc11;
    T luc13 = s[4][3] - luc0 * luc6 * luc10 - luc2 * luc7 * luc11 - luc5 * luc8 * luc12;
    T luc14 = s[4][4] - (luc0 * luc10 * luc10 + luc2 * luc11 * luc11 + luc5 * luc12 * luc12 + luc9 * luc13 * luc13);
    luc14 = 1.0 / luc14;
    T li21 = -luc0 * luc1;
    T li32 = -luc2 * luc4;
    T li31 = (luc1 * luc2 * luc4 - luc3) * luc0;
    T li43 = -luc8 * luc5;
    T li42 = (luc4 * luc8 * luc5 - luc7) * luc2;
    T li41 = (-luc1 * luc2 * luc4 * luc8 * luc5 + luc1 * luc2 * luc7 + luc3 * luc8 * luc5 - luc6) * luc0;
    T li54 = -luc13 * luc9;
    T li53 = (luc13 * luc8 * luc9 - luc12) * luc5;
    T li52 = (-luc4 * luc8 * luc13 * luc5 * luc9 + luc4 * luc12 * luc5 + luc7 * luc13 * luc9 - luc11) * luc2;
    T li51 = (luc1 * luc4 * luc8 * luc13 * luc2 * luc5 * luc9 - luc13 * luc8 * luc3 * luc9 * luc5 -
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
    for (int i = 0; i < 5; ++i)
        for (int j = i + 1; j < 5; ++j) d[i][j] = d[j][i];
}

template <class T> inline void invert(const T (&s)[3][3], T (&d)[3][3]) { invert33(s, d); }
template <class T> inline void invert(const T (&s)[4][4], T (&d)[4][4]) { invert44(s, d); }
template <class T> inline void invert(const T (&s)[5][5], T (&d)[5][5]) { invert55(s, d); }

// Intermediate values exported for the independent checks in the tests.
template <class T, int N> struct LineDetail {
    T w[N];
    T C[N][N];
    T I[N][N];
    T u[N];
};
template <class T, int N> struct CircleDetail {
    T w[N];
    T C[N + 1][N + 1];
    T I[N + 1][N + 1];
    T u[N + 1];
    T VarBeta[N];  // scaled copy
    T Z[N];
};

// broken_line.jl:546-696 ; ge[c][k] float (xx,xy,yy,xz,yz,zz)
template <class T, int N>
inline void BL_Line_fit(const T (&ge)[6][N], const T (&ff)[4], double B, Prepared<T, N> &data, LineFit<T> &lr,
                        LineDetail<T, N> *det = nullptr) {
    using std::abs;
    using std::sqrt;
    constexpr int n = N;
    const T *S = data.S;
    const T *Z = data.Z;
    const T *VarBeta = data.VarBeta;
    T slope = double(-data.q) / ff[3];
    Rot2<T> R = rotation_matrix(slope);
    const double tol = 1e-12;
    T w[N];
    for (int i = 0; i < n; ++i) {
        T V[3][3];
        V[0][0] = ge[0][i];
        V[0][1] = V[1][0] = (abs(ge[1][i]) < tol) ? T(0.0) : ge[1][i];
        V[0][2] = V[2][0] = (abs(ge[3][i]) < tol) ? T(0.0) : ge[3][i];
        V[1][1] = ge[2][i];
        V[1][2] = V[2][1] = (abs(ge[4][i]) < tol) ? T(0.0) : ge[4][i];
        V[2][2] = ge[5][i];
        T tmp = 1.0 / sqrt(data.radii[0][i] * data.radii[0][i] + data.radii[1][i] * data.radii[1][i]);
        T J[2][3] = {{data.radii[1][i] * tmp, -data.radii[0][i] * tmp, T(0.0)}, {T(0.0), T(0.0), T(1.0)}};
        T Rm[2][2] = {{R.a, R.b}, {-R.b, R.a}};
        T M1[2][3], M2[2][3], M3[2][2], M4[2][2];
        matmul(Rm, J, M1);
        matmul(M1, V, M2);
        T Jt[3][2];
        transpose(J, Jt);
        matmul(M2, Jt, M3);
        T Rt[2][2];
        transpose(Rm, Rt);
        matmul(M3, Rt, M4);
        w[i] = 1.0 / M4[1][1];
    }
    T r_u[N];
    for (int i = 0; i < n; ++i) r_u[i] = w[i] * Z[i];
    T C[N][N];
    matrixc_u<T, N>(w, S, VarBeta, C);
    T I[N][N];
    invert(C, I);
    T u[N];
    for (int i = 0; i < n; ++i) {
        T acc = I[i][0] * r_u[0];
        for (int k = 1; k < n; ++k) acc = acc + I[i][k] * r_u[k];
        u[i] = acc;
    }
    if (det) {
        for (int i = 0; i < n; ++i) {
            det->w[i] = w[i];
            det->u[i] = u[i];
            for (int j = 0; j < n; ++j) {
                det->C[i][j] = C[i][j];
                det->I[i][j] = I[i][j];
            }
        }
    }
    lr.par[0] = (u[1] - u[0]) / (S[1] - S[0]);
    lr.par[1] = u[0];
    T idiff = 1.0 / (S[1] - S[0]);
    lr.cov[0][0] = (I[0][0] - 2.0 * I[0][1] + I[1][1]) * (idiff * idiff) +
                   mult_scatt(S[1] - S[0], B, ff[2], 2, slope);
    lr.cov[0][1] = (I[0][1] - I[0][0]) * idiff;
    lr.cov[1][0] = (I[0][1] - I[0][0]) * idiff;
    lr.cov[1][1] = I[0][0];

    T jac[2][2] = {{T(1.0), T(0.0)}, {-S[0], T(1.0)}};
    lr.par[1] = lr.par[1] - lr.par[0] * S[0];
    T out[2][2];
    sandwich(jac, lr.cov, out);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) lr.cov[i][j] = out[i][j];
    T tmp = R.a - lr.par[0] * R.b;
    jac[1][1] = 1.0 / tmp;
    jac[0][0] = jac[1][1] * jac[1][1];
    jac[0][1] = 0.0;
    jac[1][0] = lr.par[1] * R.b * jac[0][0];
    lr.par[1] = lr.par[1] * jac[1][1];
    lr.par[0] = (R.b + lr.par[0] * R.a) * jac[1][1];
    sandwich(jac, lr.cov, out);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) lr.cov[i][j] = out[i][j];

    lr.chi2 = 0.0;
    for (int i = 0; i < n; ++i) {
        const int I1 = i + 1;
        T diff = Z[i] - u[i];
        lr.chi2 = lr.chi2 + w[i] * (diff * diff);
        if (I1 > 1 && I1 < n) {
            T denom1 = S[i] - S[i - 1];
            T denom2 = S[i + 1] - S[i];
            T denom3 = S[i] - S[i - 1];
            T denom4 = S[i + 1] - S[i - 1];
            T inner = (u[i - 1] / denom1) - (u[i] * (denom4 / (denom2 * denom3))) + (u[i + 1] / (S[i + 1] - S[i]));
            lr.chi2 = lr.chi2 + (inner * inner) / VarBeta[i];
        }
    }
}

// broken_line.jl:330-528 ; hits[c][k] double, ge[c][k] (float values)
template <class T, int N>
inline void BL_Circle_fit(const T (&hits)[3][N], const T (&ge)[6][N], const T (&ff)[4], double B,
                          Prepared<T, N> &data, CircleFit<T> &cr, FitFlags &fl,
                          CircleDetail<T, N> *det = nullptr) {
    using std::abs;
    using std::floor;
    using std::sqrt;
    constexpr int n = N;
    cr.q = data.q;
    T *s = data.s;
    T *S = data.S;
    T *Z = data.Z;
    T slope = double(-cr.q) / ff[3];
    T VarBeta[N];
    T vscale = 1.0 + slope * slope;
    for (int i = 0; i < n; ++i) VarBeta[i] = data.VarBeta[i] * vscale;
    for (int i = 0; i < n; ++i)
        Z[i] = sqrt(data.radii[0][i] * data.radii[0][i] + data.radii[1][i] * data.radii[1][i]) - ff[2];
    const double tol = 1e-12;
    T w[N];
    for (int i = 0; i < n; ++i) {
        T V[2][2];
        V[0][0] = (abs(ge[0][i]) < tol) ? T(0.0) : ge[0][i];
        V[0][1] = V[1][0] = (abs(ge[1][i]) < tol) ? T(0.0) : ge[1][i];
        V[1][1] = (abs(ge[2][i]) < tol) ? T(0.0) : ge[2][i];
        Rot2<T> RR = rotation_matrix(-data.radii[0][i] / data.radii[1][i]);
        T Rm[2][2] = {{RR.a, RR.b}, {-RR.b, RR.a}};
        T Rt[2][2];
        transpose(Rm, Rt);
        T M1[2][2], M2[2][2];
        matmul(Rm, V, M1);
        matmul(M1, Rt, M2);
        w[i] = 1.0 / M2[1][1];
    }
    T r_u[N + 1];
    r_u[n] = 0.0;
    for (int i = 0; i < n; ++i) r_u[i] = w[i] * Z[i];
    T C[N + 1][N + 1];
    {
        T Cn[N][N];
        matrixc_u<T, N>(w, s, VarBeta, Cn);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) C[i][j] = Cn[i][j];
    }
    C[n][n] = 0.0;
    for (int i = 0; i < n; ++i) {
        const int I1 = i + 1;
        C[i][n] = 0.0;
        if (I1 > 1 && I1 < n) {
            C[i][n] = C[i][n] + (-(s[i + 1] - s[i - 1])) * (s[i + 1] - s[i - 1]) /
                                    (2.0 * VarBeta[i] * (s[i + 1] - s[i]) * (s[i] - s[i - 1]));
        }
        if (I1 > 2) {
            C[i][n] = C[i][n] + (s[i] - s[i - 2]) / (2.0 * VarBeta[i - 1] * (s[i] - s[i - 1]));
        }
        if (I1 < n - 1) {
            C[i][n] = C[i][n] + (s[i + 2] - s[i]) / (2.0 * VarBeta[i + 1] * (s[i + 1] - s[i]));
        }
        C[n][i] = C[i][n];
        if (I1 > 1 && I1 < n) {
            T dd = s[i + 1] - s[i - 1];
            C[n][n] = C[n][n] + (dd * dd) / (4.0 * VarBeta[i]);
        }
    }
    T I[N + 1][N + 1];
    invert(C, I);
    T u[N + 1];
    for (int i = 0; i <= n; ++i) {
        T acc = I[i][0] * r_u[0];
        for (int k = 1; k <= n; ++k) acc = acc + I[i][k] * r_u[k];
        u[i] = acc;
    }
    if (det) {
        for (int i = 0; i < n; ++i) {
            det->w[i] = w[i];
            det->VarBeta[i] = VarBeta[i];
            det->Z[i] = Z[i];
        }
        for (int i = 0; i <= n; ++i) {
            det->u[i] = u[i];
            for (int j = 0; j <= n; ++j) {
                det->C[i][j] = C[i][j];
                det->I[i][j] = I[i][j];
            }
        }
    }
    // normalise radii columns 1 and 2 in place (broken_line.jl:407-414)
    for (int c = 0; c < 2; ++c) {
        T t = sqrt(data.radii[0][c] * data.radii[0][c] + data.radii[1][c] * data.radii[1][c]);
        data.radii[0][c] = data.radii[0][c] / t;
        data.radii[1][c] = data.radii[1][c] / t;
    }
    T tmp1 = -Z[0] + u[0];
    T d1 = hits[0][0] + tmp1 * data.radii[0][0];
    T d2 = hits[1][0] + tmp1 * data.radii[1][0];
    T tmp2 = -Z[1] + u[1];
    T e1 = hits[0][1] + tmp2 * data.radii[0][1];
    T e2 = hits[1][1] + tmp2 * data.radii[1][1];
    T emd0 = e1 - d1, emd1 = e2 - d2;
    cr.par[0] = atan2_ref(emd1, emd0);
    T nrm = sqrt(emd0 * emd0 + emd1 * emd1);
    cr.par[1] = double(-cr.q) * (ff[2] - checked_sqrt(ff[2] * ff[2] - 0.25 * (nrm * nrm), fl));
    cr.par[2] = double(cr.q) * (1.0 / ff[2] + u[n]);
    if (!(double(cr.q) * cr.par[1] <= 0.0)) fl.bits |= kFlagCharge;
    T t1 = nrm * nrm;
    const double fq = std::floor(double(cr.q) / 2.0);
    T jac[3][3];
    jac[0][0] = (data.radii[1][0] * emd0 - emd1 * data.radii[0][0]) / t1;
    jac[0][1] = (data.radii[1][1] * emd0 - emd1 * data.radii[0][1]) / t1;
    jac[0][2] = 0.0;
    T twoR = 2.0 * ff[2];
    T den = checked_sqrt(twoR * twoR - t1, fl);
    jac[1][0] = fq * (emd0 * data.radii[0][0] + emd1 * data.radii[1][0]) / den;
    jac[1][1] = fq * (emd0 * data.radii[0][1] + emd1 * data.radii[1][1]) / den;
    jac[1][2] = 0.0;
    jac[2][0] = 0.0;
    jac[2][1] = 0.0;
    jac[2][2] = double(cr.q);
    T cv[3][3] = {{I[0][0], I[0][1], I[0][n]}, {I[1][0], I[1][1], I[1][n]}, {I[n][0], I[n][1], I[n][n]}};
    sandwich(jac, cv, cr.cov);
    translate_karimaki(cr, T(0.5 * emd0), T(0.5 * emd1), jac, fl);
    T sl2 = slope * slope;
    cr.cov[0][0] = cr.cov[0][0] + (1.0 + sl2) * mult_scatt(S[1] - S[0], B, ff[2], 2, slope);
    translate_karimaki(cr, d1, d2, jac, fl);

    cr.chi2 = 0.0;
    for (int i = 0; i < n; ++i) {
        const int I1 = i + 1;
        T diff = Z[i] - u[i];
        cr.chi2 = cr.chi2 + w[i] * (diff * diff);
        if (I1 > 1 && I1 < n) {
            T dd1 = s[i] - s[i - 1];
            T dd2 = s[i + 1] - s[i];
            T dd3 = s[i + 1] - s[i - 1];
            T p1 = u[i - 1] / dd1;
            T p2 = (u[i] * dd3) / (dd1 * dd2);
            T p3 = u[i + 1] / dd2;
            T p4 = (dd3 * u[n]) / 2.0;
            T inner = p1 - p2 + p3 + p4;
            cr.chi2 = cr.chi2 + (inner * inner) / VarBeta[i];
        }
    }
}

// SOARotation.jl:139-147, float arithmetic; r = (R11,R12,R13,R21,R22,R23,R31,R32,R33)
inline void to_global_err(const float *r, float cxx, float cxy, float cyy, float *gl) {
    const float xx = r[0], xy = r[1], xz = r[2], yx = r[3], yy = r[4], yz = r[5];
    gl[0] = xx * (xx * cxx + yx * cxy) + yx * (xx * cxy + yx * cyy);
    gl[1] = xx * (xy * cxx + yy * cxy) + yx * (xy * cxy + yy * cyy);
    gl[2] = xy * (xy * cxx + yy * cxy) + yy * (xy * cxy + yy * cyy);
    gl[3] = xx * (xz * cxx + yz * cxy) + yx * (xz * cxy + yz * cyy);
    gl[4] = xy * (xz * cxx + yz * cxy) + yy * (xz * cxy + yz * cyy);
    gl[5] = xz * (xz * cxx + yz * cxy) + yz * (xz * cxy + yz * cyy);
}

// One track: fast fit + prepare + line fit + circle fit (kernelBLFastFit +
// kernelBLFit body, broken_line_fit_on_gpu.jl:71-159).  Output in the
// TrajectoryStateSOA convention (TrajectoryStateSOA.jl:20-44).
template <class T> struct TrackOut {
    T state[5];
    T cov15[15];
    T pt, eta, chi2;
    T fast[4];
    T line_chi2, circle_chi2;
    int q;
    uint32_t flags;
};

template <class T, int N>
inline void fit_track(const T (&hits)[3][N], const T (&ge)[6][N], double B, TrackOut<T> &o,
                      LineDetail<T, N> *ldet = nullptr, CircleDetail<T, N> *cdet = nullptr,
                      Prepared<T, N> *prep_out = nullptr) {
    using std::abs;
    using std::asinh;
    using std::isnan;
    FitFlags fl;
    T ff[4];
    BL_Fast_fit<T, N>(hits, ff);
    for (int k = 0; k < 4; ++k) {
        o.fast[k] = ff[k];
        if (isnan(ff[k])) fl.bits |= kFlagFastNaN;
    }
    Prepared<T, N> data;
    prepare_broken_line_data<T, N>(hits, ff, B, data);
    if (prep_out) *prep_out = data;
    LineFit<T> line;
    BL_Line_fit<T, N>(ge, ff, B, data, line, ldet);
    CircleFit<T> circle;
    BL_Circle_fit<T, N>(hits, ge, ff, B, data, circle, fl, cdet);
    // copyFromCircle!(..., b = 1.0f0/B) -> Float64 division
    const double b = 1.0 / B;
    o.state[0] = circle.par[0];
    o.state[1] = circle.par[1];
    o.state[2] = circle.par[2] * b;
    o.state[3] = line.par[0];
    o.state[4] = line.par[1];
    o.cov15[0] = circle.cov[0][0];
    o.cov15[1] = circle.cov[0][1];
    o.cov15[2] = b * circle.cov[0][2];
    o.cov15[3] = 0.0;
    o.cov15[4] = 0.0;
    o.cov15[5] = circle.cov[1][1];
    o.cov15[6] = b * circle.cov[1][2];
    o.cov15[7] = 0.0;
    o.cov15[8] = 0.0;
    o.cov15[9] = (b * b) * circle.cov[2][2];
    o.cov15[10] = 0.0;
    o.cov15[11] = 0.0;
    o.cov15[12] = line.cov[0][0];
    o.cov15[13] = line.cov[0][1];
    o.cov15[14] = line.cov[1][1];
    o.pt = B / abs(circle.par[2]);
    o.eta = asinh(line.par[0]);
    o.chi2 = (line.chi2 + circle.chi2) / double(2 * N - 5);
    o.line_chi2 = line.chi2;
    o.circle_chi2 = circle.chi2;
    o.q = circle.q;
    o.flags = fl.bits;
}

}  // namespace oracle
