// pt_bl_adjoint.hpp - reverse-mode adjoint of the broken-line fit (device code,
// one track per lane).
//
// Given the forward fit of one track (BLFit, whose members hold every
// intermediate in registers) and cotangents of its outputs - the five track
// parameters (phi, tip, q/pT, cot(theta), zip) and the two chi2 sums - it
// returns the cotangents of the hit positions and of the hit covariances.
// The forward is recomputed, not stored: the band matrices are factorised
// again in the same registers (recompute mode).
//
// Differentiated along the branch the forward took: the charge q, the
// |ge| < 1e-12 zeroing, min(20, B R) in the scattering term, |length| and the
// quadrant of each atan2 are held fixed (their derivatives are zero or the
// local ones).  The covariance outputs are not differentiated.
//
// The band systems C u = r (line: C = W + sum_k a_k a_k' / VB_k; circle: the
// same with a curvature column) are handled with the implicit function
// theorem instead of differentiating the explicit inverse: with
// lambda = C^-1 u_bar, theta_bar = -d/dtheta [lambda' (C u - r)] at fixed u,
// and chi2 = (Z - u)' W (Z - u) + sum_k (a_k . u)^2 / VB_k at its minimum
// differentiates at fixed u (the u-gradient vanishes there).
#pragma once
#include "pt_bl_fit.hpp"

namespace pt {

// d(atan2(y, x)) = (x dy - y dx) / (x^2 + y^2), every branch of atan2_bl
template <class R> PT_HD void atan2_vjp(R y, R x, R gbar, R &ybar, R &xbar) {
    const R h = x * x + y * y;
    ybar += gbar * x / h;
    xbar -= gbar * y / h;
}

// unit vector n = v / |v|: v_bar += (n_bar - (n_bar . n) n) / |v|
template <class R> PT_HD void unit_vjp(R nx, R ny, R len, R nbx, R nby, R &vbx, R &vby) {
    const R dot = nbx * nx + nby * ny;
    vbx += (nbx - dot * nx) / len;
    vby += (nby - dot * ny) / len;
}

// one translate_karimaki step (parameters only): inputs (cp, sp, d, kappa,
// x0, y0), outputs the vector (Cc, BB) that defines the new angle and
// d' = A / (1 + U)
template <class R>
PT_HD void karimaki_vjp(const KarimakiStep<R> &k, R d, R kappa, R x0, R y0, R bCc, R bBB, R bdp, R &bcp, R &bsp,
                        R &bd, R &bkappa, R &bx0, R &by0) {
    const R opu = R(1.0) + k.U;
    R bA = bdp / opu;
    const R bU = -bdp * k.A / (opu * opu);
    const R barg = bU / (R(2.0) * k.U);  // U = sqrt(1 + kappa A)
    bkappa += barg * k.A;
    bA += barg * kappa;
    const R bDO = bA * (R(2.0) + R(2.0) * kappa * k.DO);  // A = 2 DO + kappa (DO^2 + DP^2)
    const R bDP = bA * (R(2.0) * kappa * k.DP);
    bkappa += bA * (k.DO * k.DO + k.DP * k.DP);
    bkappa += bBB * x0;  // BB = kappa x0 + uu sp
    bx0 += bBB * kappa;
    R buu = bBB * k.sp;
    bsp += bBB * k.uu;
    bkappa -= bCc * y0;  // Cc = -kappa y0 + uu cp
    by0 -= bCc * kappa;
    buu += bCc * k.cp;
    bcp += bCc * k.uu;
    bkappa += buu * d;  // uu = 1 + kappa d
    bd += buu * kappa;
    bx0 += bDO * k.sp;  // DO = x0 sp - y0 cp + d
    bsp += bDO * x0;
    by0 -= bDO * k.cp;
    bcp -= bDO * y0;
    bd += bDO;
    bx0 += bDP * k.cp;  // DP = x0 cp + y0 sp
    bcp += bDP * x0;
    by0 += bDP * k.sp;
    bsp += bDP * y0;
}

// broken_line.jl:46-79 mult_scatt: partial derivatives in length, radius and slope
template <class R> PT_HD void mult_scatt_vjp(R length, R B, R Rr, R slope, R bms, R &blen, R &bR, R &bslope) {
    const R BR = B * Rr;
    R pt2 = (BR < R(20.0)) ? BR : R(20.0);
    pt2 = pt2 * pt2;
    const R XXI_0 = R(0.06 / 16.0);
    const R fact = R(0.7 * ((13.6 / 1000) * (13.6 / 1000)));
    const R st = R(1.0) + slope * slope;
    const R al = r_abs(length) * XXI_0;
    const R sq = R(1.0) + R(0.038) * log_r(al);
    const R ms = fact / (pt2 * st) * al * (sq * sq);
    const R sgn = (length < R(0.0)) ? R(-1.0) : R(1.0);
    blen += bms * fact / (pt2 * st) * XXI_0 * sgn * (sq * sq + R(0.076) * sq);
    if (BR < R(20.0)) bR += bms * (R(-2.0) * ms / Rr);
    bslope += bms * (-ms * R(2.0) * slope / st);
}

// Reverse sweep through one band system.  pos: node positions (S or s),
// w, Zr: weights and measured residual coordinates, VB: scattering
// variances, u: solution, I: inverse of C, ubar: cotangent of u, cbar:
// cotangent of the chi2 sum.  NC = N (line) or N + 1 (circle, the last
// unknown is the curvature correction with column (pos[k+1] - pos[k-1]) / 2).
template <int N, int NC>
PT_HD void band_vjp(const double (&pos)[N], const double (&w)[N], const double (&Zr)[N], const double (&VB)[N],
                    const double (&u)[NC], const double (&I)[NC][NC], const double (&ubar)[NC], double cbar,
                    double (&bpos)[N], double (&bw)[N], double (&bZ)[N], double (&bVB)[N]) {
    double lam[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < NC; ++j) acc += I[i][j] * ubar[j];
        lam[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double res = Zr[i] - u[i];
        bw[i] += cbar * res * res + lam[i] * res;
        bZ[i] += cbar * 2.0 * w[i] * res + lam[i] * w[i];
    }
#pragma unroll
    for (int k = 1; k < N - 1; ++k) {
        const double d1 = pos[k] - pos[k - 1], d2 = pos[k + 1] - pos[k];
        double a[3] = {1.0 / d1, -(1.0 / d1 + 1.0 / d2), 1.0 / d2};
        double au = a[0] * u[k - 1] + a[1] * u[k] + a[2] * u[k + 1];
        double al = a[0] * lam[k - 1] + a[1] * lam[k] + a[2] * lam[k + 1];
        double aN = 0.0;
        if (NC > N) {
            aN = 0.5 * (pos[k + 1] - pos[k - 1]);
            au += aN * u[NC - 1];
            al += aN * lam[NC - 1];
        }
        const double vb = VB[k];
        bVB[k] += (al * au - cbar * au * au) / (vb * vb);
        // a_bar = 2 cbar (a.u) u / VB - (lambda (a.u) + u (a.lambda)) / VB
        double ab[3];
#pragma unroll
        for (int m = 0; m < 3; ++m)
            ab[m] = (2.0 * cbar * au * u[k - 1 + m] - (lam[k - 1 + m] * au + u[k - 1 + m] * al)) / vb;
        const double bd1 = (ab[1] - ab[0]) / (d1 * d1);
        const double bd2 = (ab[1] - ab[2]) / (d2 * d2);
        bpos[k] += bd1 - bd2;
        bpos[k - 1] -= bd1;
        bpos[k + 1] += bd2;
        if (NC > N) {
            const double abN = (2.0 * cbar * au * u[NC - 1] - (lam[NC - 1] * au + u[NC - 1] * al)) / vb;
            bpos[k + 1] += 0.5 * abN;
            bpos[k - 1] -= 0.5 * abN;
        }
    }
}

// Cotangents of one track's hits from cotangents of its outputs.
//   f      the forward fit of the track (mixed or double policy)
//   bstate cotangents of (phi, tip, q/pT, cot(theta), zip)
//   bcl, bcc cotangents of the line and circle chi2 sums
//   gx, gy, gz, gge: cotangents of the positions and of ge[6] per hit
template <class P, int N>
PT_HD void bl_fit_vjp(const BLFit<P, N> &f, const float (&hx)[N], const float (&hy)[N], const float (&hz)[N],
                      const float (&ge)[6][N], double B, const double (&bstate)[5], double bcl, double bcc,
                      double (&gx)[N], double (&gy)[N], double (&gz)[N], double (&gge)[6][N]) {
    typedef double R;
    constexpr int mid = (N >> 1);
    const R q = (R)f.q;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        gx[i] = gy[i] = gz[i] = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) gge[c][i] = 0.0;
    }
    R bff0 = 0, bff1 = 0, bff2 = 0, bff3 = 0, bslope = 0, bra = 0, brb = 0;
    R brx[N], bry[N], bs[N], bS[N], bZ[N], bVB[N], bVBs[N], blw[N], bcw[N], bZc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) brx[i] = bry[i] = bs[i] = bS[i] = bZ[i] = bVB[i] = bVBs[i] = blw[i] = bcw[i] = bZc[i] = 0;

    // ---- outputs: state = (phi2, d2, kappa / B, lpar0, lpar1)
    R bkappa = bstate[2] * (1.0 / B);
    // ---- second translation (to the origin), inputs from the first
    R bdpt1 = 0, bdpt2 = 0, bept1 = 0, bept2 = 0;
    R bd1in;  // cotangent of the first step's output d'
    R bCc1 = 0, bBB1 = 0;
    {
        const KarimakiStep<typename P::Kar> &k = f.k2;
        R bCc = 0, bBB = 0;
        atan2_vjp<R>(k.BB, k.Cc, bstate[0], bBB, bCc);
        R bcp = 0, bsp = 0, bd = 0, bx0 = 0, by0 = 0;
        const R d_in = (R)(f.k1.A / (1.0 + f.k1.U));
        karimaki_vjp<R>(KarimakiStep<R>{k.cp, k.sp, k.DP, k.DO, k.uu, k.Cc, k.BB, k.A, k.U, 0, 0, 0, 0, 0}, d_in,
                        (R)f.cpar2_d, (R)f.dpt1, (R)f.dpt2, bCc, bBB, bstate[1], bcp, bsp, bd, bkappa, bx0, by0);
        bdpt1 += bx0;
        bdpt2 += by0;
        bd1in = bd;
        // (cp, sp) = (Cc1, BB1) / |(Cc1, BB1)|
        const R hyp = sqrt((R)f.k1.BB * f.k1.BB + (R)f.k1.Cc * f.k1.Cc);
        unit_vjp<R>(k.cp, k.sp, hyp, bcp, bsp, bCc1, bBB1);
    }
    // ---- first translation (to the chord midpoint em / 2)
    R bem0 = 0, bem1 = 0, bcu0 = 0, bcu1 = 0, bcuN = 0;
    {
        const KarimakiStep<typename P::Kar> &k = f.k1;
        R bcp = 0, bsp = 0, bd = 0, bx0 = 0, by0 = 0;
        karimaki_vjp<R>(KarimakiStep<R>{k.cp, k.sp, k.DP, k.DO, k.uu, k.Cc, k.BB, k.A, k.U, 0, 0, 0, 0, 0},
                        (R)f.cpar1_d, (R)f.cpar2_d, 0.5 * (R)f.em0, 0.5 * (R)f.em1, bCc1, bBB1, bd1in, bcp, bsp, bd,
                        bkappa, bx0, by0);
        bem0 += 0.5 * bx0;
        bem1 += 0.5 * by0;
        unit_vjp<R>(k.cp, k.sp, (R)f.nrm, bcp, bsp, bem0, bem1);  // (cp, sp) = em / |em|
        // d = -q (R - sqrt(R^2 - |em|^2 / 4))
        bff2 += -q * bd;
        const R bsarg = q * bd;
        const R barg = bsarg / (2.0 * f.sarg);
        bff2 += 2.0 * f.ff2 * barg;
        const R bnrm = -0.5 * (R)f.nrm * barg;
        bem0 += bnrm * (R)f.em0 / (R)f.nrm;
        bem1 += bnrm * (R)f.em1 / (R)f.nrm;
    }
    // kappa = q (1 / R + u[N])
    bff2 += -q * bkappa / (f.ff2 * f.ff2);
    bcuN += q * bkappa;
    // em = e - d; d = P0 + (u0 - Zc0) n0; e = P1 + (u1 - Zc1) n1
    bept1 += bem0;
    bept2 += bem1;
    bdpt1 -= bem0;
    bdpt2 -= bem1;
    {
        gx[0] += bdpt1;
        gy[0] += bdpt2;
        const R n0x = f.radx[0] / f.rad[0], n0y = f.rady[0] / f.rad[0];
        const R tmp1 = f.cu[0] - f.Zc[0];
        const R bt1 = bdpt1 * n0x + bdpt2 * n0y;
        bcu0 += bt1;
        bZc[0] -= bt1;
        unit_vjp<R>(n0x, n0y, f.rad[0], tmp1 * bdpt1, tmp1 * bdpt2, brx[0], bry[0]);
        gx[1] += bept1;
        gy[1] += bept2;
        const R n1x = f.radx[1] / f.rad[1], n1y = f.rady[1] / f.rad[1];
        const R tmp2 = f.cu[1] - f.Zc[1];
        const R bt2 = bept1 * n1x + bept2 * n1y;
        bcu1 += bt2;
        bZc[1] -= bt2;
        unit_vjp<R>(n1x, n1y, f.rad[1], tmp2 * bept1, tmp2 * bept2, brx[1], bry[1]);
    }
    // ---- circle band system
    {
        double pos[N], w[N], Zr[N], vb[N], u[N + 1], I[N + 1][N + 1], ub[N + 1];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            pos[i] = f.s[i];
            w[i] = f.cw[i];
            Zr[i] = f.Zc[i];
            vb[i] = f.VBs[i];
        }
#pragma unroll
        for (int i = 0; i <= N; ++i) {
            u[i] = f.cu[i];
            ub[i] = 0.0;
#pragma unroll
            for (int j = 0; j <= N; ++j) I[i][j] = f.cI[i][j];
        }
        ub[0] = bcu0;
        ub[1] = bcu1;
        ub[N] += bcuN;
        band_vjp<N, N + 1>(pos, w, Zr, vb, u, I, ub, bcc, bs, bcw, bZc, bVBs);
    }
    // ---- line parameters from u (broken_line.jl:610-645)
    R bu0 = 0, bu1 = 0;
    {
        const R S0 = f.S[0], S1 = f.S[1], ra = f.ra, rb = f.rb;
        const R l0r = f.lpar0_raw, l1r = f.lpar1_raw;
        const R tmp = ra - l0r * rb, j11 = 1.0 / tmp;
        const R bl0 = bstate[3], bl1 = bstate[4];
        const R bj11 = bl1 * l1r + bl0 * (rb + l0r * ra);
        const R bl1r = bl1 * j11;
        brb += bl0 * j11;
        R bl0r = bl0 * ra * j11;
        bra += bl0 * l0r * j11;
        const R btmp = -bj11 / (tmp * tmp);
        bra += btmp;
        bl0r += -btmp * rb;
        brb += -btmp * l0r;
        bu0 += bl1r;  // l1r = u0 - l0r S0
        bl0r += -bl1r * S0;
        bS[0] += -bl1r * l0r;
        const R dS = S1 - S0;  // l0r = (u1 - u0) / dS
        bu1 += bl0r / dS;
        bu0 -= bl0r / dS;
        const R bdS = -bl0r * l0r / dS;
        bS[1] += bdS;
        bS[0] -= bdS;
    }
    // ---- line band system
    {
        double pos[N], w[N], Zr[N], vb[N], u[N], I[N][N], ub[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            pos[i] = f.S[i];
            w[i] = f.lw[i];
            Zr[i] = f.Z[i];
            vb[i] = f.VB[i];
            u[i] = f.lu[i];
            ub[i] = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) I[i][j] = f.lI[i][j];
        }
        ub[0] = bu0;
        ub[1] = bu1;
        band_vjp<N, N>(pos, w, Zr, vb, u, I, ub, bcl, bS, blw, bZ, bVB);
    }
    // ---- circle weights: w = 1 / (n' V n), n the radial unit vector
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool z0 = fabsf(ge[0][i]) < 1e-12, z1 = fabsf(ge[1][i]) < 1e-12, z2 = fabsf(ge[2][i]) < 1e-12;
        const R v00 = z0 ? 0.0 : (R)ge[0][i], v01 = z1 ? 0.0 : (R)ge[1][i], v11 = z2 ? 0.0 : (R)ge[2][i];
        const R nx = f.radx[i] / f.rad[i], ny = f.rady[i] / f.rad[i];
        const R w = f.cw[i];
        const R bm = -bcw[i] * w * w;
        if (!z0) gge[0][i] += bm * nx * nx;
        if (!z2) gge[2][i] += bm * ny * ny;
        if (!z1) gge[1][i] += 2.0 * bm * nx * ny;
        const R bnx = 2.0 * bm * (v00 * nx + v01 * ny), bny = 2.0 * bm * (v01 * nx + v11 * ny);
        unit_vjp<R>(nx, ny, f.rad[i], bnx, bny, brx[i], bry[i]);
    }
    // ---- line weights: w = 1 / (g' V g), g = (-rb t, ra), t = (rady, -radx) / |rad|
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool z1 = fabsf(ge[1][i]) < 1e-12, z3 = fabsf(ge[3][i]) < 1e-12, z4 = fabsf(ge[4][i]) < 1e-12;
        const R v00 = ge[0][i], v11 = ge[2][i], v22 = ge[5][i];
        const R v01 = z1 ? 0.0 : (R)ge[1][i], v02 = z3 ? 0.0 : (R)ge[3][i], v12 = z4 ? 0.0 : (R)ge[4][i];
        const R r = f.rad[i];
        const R t0 = f.rady[i] / r, t1 = -f.radx[i] / r;
        const R ra = f.ra, rb = f.rb;
        const R g0 = -rb * t0, g1 = -rb * t1, g2 = ra;
        const R w = f.lw[i];
        const R bm = -blw[i] * w * w;
        gge[0][i] += bm * g0 * g0;
        gge[2][i] += bm * g1 * g1;
        gge[5][i] += bm * g2 * g2;
        if (!z1) gge[1][i] += 2.0 * bm * g0 * g1;
        if (!z3) gge[3][i] += 2.0 * bm * g0 * g2;
        if (!z4) gge[4][i] += 2.0 * bm * g1 * g2;
        const R bg0 = 2.0 * bm * (v00 * g0 + v01 * g1 + v02 * g2);
        const R bg1 = 2.0 * bm * (v01 * g0 + v11 * g1 + v12 * g2);
        const R bg2 = 2.0 * bm * (v02 * g0 + v12 * g1 + v22 * g2);
        brb += -(bg0 * t0 + bg1 * t1);
        bra += bg2;
        const R bt0 = -rb * bg0, bt1 = -rb * bg1;
        const R dot = bt0 * t0 + bt1 * t1;
        // t = (y, -x) / r with (x, y) = (radx, rady)
        brx[i] += -bt1 / r - dot * f.radx[i] / (r * r);
        bry[i] += bt0 / r - dot * f.rady[i] / (r * r);
    }
    // ---- circle residuals Zc = |rad| - R
#pragma unroll
    for (int i = 0; i < N; ++i) {
        bff2 -= bZc[i];
        brx[i] += bZc[i] * f.radx[i] / f.rad[i];
        bry[i] += bZc[i] * f.rady[i] / f.rad[i];
    }
    // ---- scattering variances: VBs = VB (1 + slope^2), VB_i = ms(S_{i+1} - S_i) + ms(S_i - S_{i-1})
    {
        const R vscale = f.vscale;
        R bvscale = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            bVB[i] += bVBs[i] * vscale;
            bvscale += bVBs[i] * (R)f.VB[i];
        }
        bslope += 2.0 * (R)f.slope * bvscale;
#pragma unroll
        for (int i = 1; i < N - 1; ++i) {
            R bl = 0;
            mult_scatt_vjp<R>((R)f.S[i + 1] - (R)f.S[i], B, f.ff2, (R)f.slope, bVB[i], bl, bff2, bslope);
            bS[i + 1] += bl;
            bS[i] -= bl;
            bl = 0;
            mult_scatt_vjp<R>((R)f.S[i] - (R)f.S[i - 1], B, f.ff2, (R)f.slope, bVB[i], bl, bff2, bslope);
            bS[i] += bl;
            bS[i - 1] -= bl;
        }
    }
    // ---- (S, Z) = rotation of (s, z)
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const R ra = f.ra, rb = f.rb, s = f.s[i], z = hz[i];
        bs[i] += bS[i] * ra - bZ[i] * rb;
        gz[i] += bS[i] * rb + bZ[i] * ra;
        bra += bS[i] * s + bZ[i] * z;
        brb += bS[i] * z - bZ[i] * s;
    }
    // ra = 1 / sqrt(1 + slope^2), rb = slope ra; slope = -q / ff3
    {
        const R ra = f.ra, sl = f.slope, ra3 = ra * ra * ra;
        bslope += bra * (-sl * ra3) + brb * ra3;
        bff3 += bslope * q / (f.ff3 * f.ff3);
    }
    // ---- arc lengths s_i = q R atan2(cc_i, dd_i)
    {
        R bfac = 0, bE1 = 0, bE2 = 0;
        const R fac = f.fac;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            bfac += bs[i] * (R)f.s[i] / fac;
            R bcc = 0, bdd = 0;
            atan2_vjp<R>(f.cc[i], f.dd[i], bs[i] * fac, bcc, bdd);
            brx[i] += bcc * f.e2 + bdd * f.e1;
            bry[i] += -bcc * f.e1 + bdd * f.e2;
            bE2 += bcc * f.radx[i] + bdd * f.rady[i];
            bE1 += -bcc * f.rady[i] + bdd * f.radx[i];
        }
        bff2 += q * bfac;
        // E = sc (ff0, ff1), sc = -R / rr, rr = |(ff0, ff1)|
        const R rr = f.rr, sc = -f.ff2 / rr;
        const R bsc = bE1 * f.ff0 + bE2 * f.ff1;
        bff0 += bE1 * sc;
        bff1 += bE2 * sc;
        bff2 += -bsc / rr;
        const R brr = bsc * f.ff2 / (rr * rr);
        bff0 += brr * f.ff0 / rr;
        bff1 += brr * f.ff1 / rr;
    }
    // ---- radii about the fast-fit centre
#pragma unroll
    for (int i = 0; i < N; ++i) {
        gx[i] += brx[i];
        gy[i] += bry[i];
        bff0 -= brx[i];
        bff1 -= bry[i];
    }
    // ---- BL_Fast_fit (broken_line.jl:275-313)
    {
        const R ff2 = f.ff2, ff3 = f.ff3, ang = f.ang, dz = f.dz;
        bff2 += bff3 * ang / dz;
        const R bang = bff3 * ff2 / dz;
        const R bdz = -bff3 * ff3 / dz;
        gz[N - 1] += bdz;
        gz[0] -= bdz;
        const R d1 = f.fd1, d2 = f.fd2, e1 = f.fe1, e2 = f.fe2;
        const R de = d1 * e2 - d2 * e1, dotde = d1 * e1 + d2 * e2;
        R bde = 0, bdot = 0;
        atan2_vjp<R>(de, dotde, bang, bde, bdot);
        const R bd1 = bde * e2 + bdot * e1, bd2 = -bde * e1 + bdot * e2;
        const R be1 = -bde * d2 + bdot * d1, be2 = bde * d1 + bdot * d2;
        R bP0x = bd1, bP0y = bd2, bPnx = be1, bPny = be2, bPmx = 0, bPmy = 0;
        bff0 -= bd1 + be1;
        bff1 -= bd2 + be2;
        // ff0 = h11 - X tmp, ff1 = h21 + Y tmp, ff2 = sqrt(na nb nc) / (2 |ba|), tmp = 0.5 / ca
        const R a1 = f.fa1, a2 = f.fa2, b1 = f.fb1, b2 = f.fb2, c1 = f.fc1, c2 = f.fc2;
        const R na = f.fna, nb = f.fnb, nc = f.fnc, ca = f.fca, ba = f.fba;
        const R tmp = 0.5 / ca;
        const R X = a2 * nc + c2 * na, Y = a1 * nc + c1 * na;
        bP0x += bff0;
        bP0y += bff1;
        const R bX = -bff0 * tmp, bY = bff1 * tmp;
        const R btmp = -bff0 * X + bff1 * Y;
        R bca = -btmp * tmp / ca;
        const R Pi = na * nb * nc;
        const R bPi = bff2 * ff2 / (2.0 * Pi);
        const R bba = -bff2 * ff2 / ba;
        R ba1 = 0, ba2 = 0, bb1 = 0, bb2 = 0, bc1 = 0, bc2 = 0;
        R bna = bPi * nb * nc, bnb = bPi * na * nc, bnc = bPi * na * nb;
        ba2 += bX * nc;
        bnc += bX * a2;
        bc2 += bX * na;
        bna += bX * c2;
        ba1 += bY * nc;
        bnc += bY * a1;
        bc1 += bY * na;
        bna += bY * c1;
        bc1 += bca * a2;
        ba2 += bca * c1;
        bc2 -= bca * a1;
        ba1 -= bca * c2;
        bb1 += bba * a2;
        ba2 += bba * b1;
        bb2 -= bba * a1;
        ba1 -= bba * b2;
        bc1 += 2.0 * c1 * bnc;
        bc2 += 2.0 * c2 * bnc;
        ba1 += 2.0 * a1 * bna;
        ba2 += 2.0 * a2 * bna;
        bb1 += 2.0 * b1 * bnb;
        bb2 += 2.0 * b2 * bnb;
        // a = Pm - P0, b = Pn - Pm, c = P0 - Pn
        bPmx += ba1 - bb1;
        bPmy += ba2 - bb2;
        bP0x += -ba1 + bc1;
        bP0y += -ba2 + bc2;
        bPnx += bb1 - bc1;
        bPny += bb2 - bc2;
        gx[0] += bP0x;
        gy[0] += bP0y;
        gx[mid] += bPmx;
        gy[mid] += bPmy;
        gx[N - 1] += bPnx;
        gy[N - 1] += bPny;
    }
}

// Global covariance of a hit from its local variances (cxy = 0) and the
// module frame r = (R11..R33) (SOARotation.jl:139-147), and the cotangents
// of (cxx, cyy) from those of the six global entries.
PT_HD void local_err_vjp(const float *r, const double (&bge)[6], double &bcxx, double &bcyy) {
    const double xx = r[0], xy = r[1], xz = r[2], yx = r[3], yy = r[4], yz = r[5];
    bcxx = bge[0] * xx * xx + bge[1] * xx * xy + bge[2] * xy * xy + bge[3] * xx * xz + bge[4] * xy * xz +
           bge[5] * xz * xz;
    bcyy = bge[0] * yx * yx + bge[1] * yx * yy + bge[2] * yy * yy + bge[3] * yx * yz + bge[4] * yy * yz +
           bge[5] * yz * yz;
}

}  // namespace pt
