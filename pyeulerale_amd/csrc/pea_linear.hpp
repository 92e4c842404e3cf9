// Linearised JST/ALE face fluxes: the tangent of face_flux2_lds at one primal state, for many
// directions.
//
// The forward mode of face_flux2_lds (Dual numbers) evaluates the primal flux again for every
// direction.  For K directions at one state, the primal part of a face -- its normal and ALE speed
// and, per dissipative side, the chosen wave integral, sensor and switch branches (the biased
// maxima of f90:111-120 as taken at f90:207-213) -- is computed once into a FaceLin record.  Each
// direction then costs the linear map alone:
//   d(central)   = (d ph_L + d ph_R) / 2,  d ph_v = du_v q + u_v dq (+ n dp, or q dp for energy)
//   d(ALE)       = (du_L + du_R) vn / 2
//   d(D) per side = de2 d2 + e2 dd2 - (de4 d4 + e4 dd4)
// with dq = (n.dj - q drho)/rho, de2 = k2 (ds lam + s dlam), de4 = k4 dlam - c4 de2 (0 on the
// e4 = 0 branch), dlam = +-dq + d(c sn) of the chosen cell and ds the tangent of the chosen
// sensor.  The primal states, pressures and 1/rho come from LDS at each direction, as do the
// direction's states and pressures.  The decisions are those of face_flux2_lds evaluated with the
// same expressions, so the tangent is the one the Dual kernel computes, a rounding apart.
#pragma once
#include "pea_face.hpp"

namespace pea {

// One dissipative side's primal coefficients.
struct SideLin {
    double lam, s, ib;   // chosen wave integral |q| + c sn, chosen sensor, 1/(its denominator)
    double aC, bC;       // d(c sn) of the chosen cell = aC dp - bC drho
};

// Decision bits of one side (L in bits 0-7, R in bits 8-15).
enum : unsigned {
    kLinDiss = 1u,    // the side is dissipative
    kLinCellR = 2u,   // the chosen wave integral is the R cell's
    kLinNegQ = 4u,    // d|q| of the chosen cell is -dq
    kLinT2 = 8u,      // the chosen sensor is the triple (l, r, rr), else (ll, l, r)
    kLinNegA = 16u,   // the chosen sensor's numerator is negative (or -0)
    kLinZero4 = 32u,  // e4 = 0
};

struct FaceLin {
    double n0, n1, hvn;   // L's normal of the face, (face velocity . n) / 2
    SideLin L, R;
    unsigned bits;
};
constexpr int kLinWords = 14;   // doubles of a FaceLin in LDS (the bits in the last word)

// Numerator, denominator and value of the pressure sensor of the triple (a, b, c), as sensor()
// forms them (f90:209-210).
PEA_HD void sensor_parts(double pa, double pb, double pc, double& num, double& den, double& s) {
    num = pa - 2.0 * pb + pc;
    den = pa + 2.0 * pb + pc;
    s = qdiv(sabs(num), den);
}

// Coefficients of one side, in side_coefs' argument order: A is the side's own cell (its first
// wave integral), B the other; sensor 1 before sensor 2.  a_is_R: A is the face's R cell;
// sgnA/sgnB: the signs that d|q| takes for A and B; s1_is_T2: sensor 1 is the triple (l, r, rr).
PEA_HD unsigned side_lin_build(const pea_params& P, double lamA, double lamB, bool a_is_R,
                               double sgnA, double sgnB, double aCA, double bCA, double aCB,
                               double bCB, double s1, double num1, double den1, bool s1_is_T2,
                               double s2, double num2, double den2, SideLin& o) {
    const bool takeA = maxb_takes_first(lamA, lamB);
    o.lam = takeA ? lamA : lamB;
    const bool take1 = maxb_takes_first(s1, s2);
    o.s = take1 ? s1 : s2;
    const double e2 = P.k2 * o.s * o.lam;
    const bool zero4 = maxb_takes_first(0.0, P.k4 * o.lam - P.c4 * e2);
    const double num = take1 ? num1 : num2;
    o.ib = rcp_fast(take1 ? den1 : den2);
    o.aC = takeA ? aCA : aCB;
    o.bC = takeA ? bCA : bCB;
    unsigned b = kLinDiss;
    if (takeA == a_is_R) b |= kLinCellR;
    if ((takeA ? sgnA : sgnB) < 0.0) b |= kLinNegQ;
    if (take1 == s1_is_T2) b |= kLinT2;
    if (sign1(num) < 0.0) b |= kLinNegA;
    if (zero4) b |= kLinZero4;
    return b;
}

// The record of one face: arguments as face_flux2_lds's, plus iL, iR = 1/rho of L and R as
// rcp_fast gives them (the face kernels' quotient q = (n.j) / rho).  st(slot, v): primal state,
// slots 0..3 = ll, l, r, rr.
template <class St>
PEA_HD void face_lin_build(const pea_params& P, const St& st, double pLL, double pL, double pR,
                           double pRR, double cL, double cR, double iL, double iR,
                           const double* n, const double* nR, const double* gv, bool dissL,
                           bool dissR, FaceLin& f) {
    const double L1 = st(1, 1), L2 = st(1, 2), R1 = st(2, 1), R2 = st(2, 2);
    const double qL = (n[0] * L1 + n[1] * L2) * iL;
    const double qR = (n[0] * R1 + n[1] * R2) * iR;
    f.n0 = n[0];
    f.n1 = n[1];
    f.hvn = (gv[0] * n[0] + gv[1] * n[1]) / 2.0;
    f.bits = 0;
    f.L = SideLin{0.0, 0.0, 0.0, 0.0, 0.0};
    f.R = f.L;
    if (!(dissL || dissR)) return;
    const double sn = ssqrt(n[0] * n[0] + n[1] * n[1]);
    const double lamL = sabs(qL) + cL * sn, lamR = sabs(qR) + cR * sn;
    // d(c sn) = sn (gamma dp - x drho) rb h with x = gamma p rb, rb = 1/rho, h = 1/(2c): the
    // tangent of ssqrt(qdiv(gamma p, rho)) of cell_pc
    auto ccoef = [&](double c, double p, double rb, double& a, double& b) {
        const double h = 0.5 * rcp_fast(c);
        const double x = kGamma * p * rb;
        a = sn * (kGamma * rb * h);
        b = sn * (x * rb * h);
    };
    double aL, bL, aR, bR;
    ccoef(cL, pL, iL, aL, bL);
    ccoef(cR, pR, iR, aR, bR);
    if (dissL) {   // side_coefs(lamL, lamR, sensor(ll, l, r), sensor(l, r, rr))
        double n1, d1, s1, n2, d2, s2;
        sensor_parts(pLL, pL, pR, n1, d1, s1);
        sensor_parts(pL, pR, pRR, n2, d2, s2);
        f.bits |= side_lin_build(P, lamL, lamR, false, sign1(qL), sign1(qR), aL, bL, aR, bR, s1,
                                 n1, d1, false, s2, n2, d2, f.L);
    }
    if (dissR) {   // side_coefs(lamR, lamL, sensor(rr, r, l), sensor(r, l, ll)), R's own signs
        const double sL = -sign1(nR[0] * L1 + nR[1] * L2);
        const double sR = -sign1(nR[0] * R1 + nR[1] * R2);
        double n1, d1, s1, n2, d2, s2;
        sensor_parts(pRR, pR, pL, n1, d1, s1);
        sensor_parts(pR, pL, pLL, n2, d2, s2);
        f.bits |= side_lin_build(P, lamR, lamL, true, sR, sL, aR, bR, aL, bL, s1, n1, d1, true,
                                 s2, n2, d2, f.R)
                  << 8;
    }
}

// Tangents of a side's switch coefficients along one direction; e2, e4 are the primal ones.
PEA_HD void side_lin_apply(const pea_params& P, const SideLin& o, unsigned b, double dqL,
                           double dqR, double drL, double drR, const double* dp, double& e2,
                           double& e4, double& de2, double& de4) {
    const bool cr = b & kLinCellR;
    const double dq = cr ? dqR : dqL;
    const double dlam = ((b & kLinNegQ) ? -dq : dq) + (o.aC * (cr ? dp[2] : dp[1]) -
                                                      o.bC * (cr ? drR : drL));
    const bool t2 = b & kLinT2;
    const double x = t2 ? dp[1] : dp[0], y = t2 ? dp[2] : dp[1], z = t2 ? dp[3] : dp[2];
    const double da = x - 2.0 * y + z, db = x + 2.0 * y + z;
    const double ds = (((b & kLinNegA) ? -da : da) - o.s * db) * o.ib;
    e2 = P.k2 * o.s * o.lam;
    de2 = P.k2 * (ds * o.lam + o.s * dlam);
    if (b & kLinZero4) {
        e4 = 0.0;
        de4 = 0.0;
    } else {
        e4 = P.k4 * o.lam - P.c4 * e2;
        de4 = P.k4 * dlam - P.c4 * de2;
    }
}

// Tangent of the face's two fluxes (f_L, f_R of face_flux2_lds) along one real direction.
// st(slot, v), pL, pR, iL, iR: the primal states, the pressures and 1/rho of L and R;
// dst(slot, v): the direction's states, dp[4]: its pressures at ll, l, r, rr.
template <class St, class DSt, class OutL, class OutR>
PEA_HD void face_lin_apply(const pea_params& P, const FaceLin& f, const St& st, double pL,
                           double pR, double iL, double iR, const DSt& dst, const double* dp,
                           OutL outL, OutR outR) {
    const double L1 = st(1, 1), L2 = st(1, 2), R1 = st(2, 1), R2 = st(2, 2);
    const double qL = (f.n0 * L1 + f.n1 * L2) * iL;
    const double qR = (f.n0 * R1 + f.n1 * R2) * iR;
    const double dL0 = dst(1, 0), dR0 = dst(2, 0);
    const double dqL = (f.n0 * dst(1, 1) + f.n1 * dst(1, 2) - qL * dL0) * iL;
    const double dqR = (f.n0 * dst(2, 1) + f.n1 * dst(2, 2) - qR * dR0) * iR;
    const unsigned bL = f.bits & 0xffu, bR = f.bits >> 8;
    double e2L = 0.0, e4L = 0.0, de2L = 0.0, de4L = 0.0;
    double e2R = 0.0, e4R = 0.0, de2R = 0.0, de4R = 0.0;
    if (bL & kLinDiss) side_lin_apply(P, f.L, bL, dqL, dqR, dL0, dR0, dp, e2L, e4L, de2L, de4L);
    if (bR & kLinDiss) side_lin_apply(P, f.R, bR, dqL, dqR, dL0, dR0, dp, e2R, e4R, de2R, de4R);
    // Regrouped as face_flux2_cell: the tangent of central minus ALE,
    //   t_v = dL_v (qL/2 - vn/2) + dR_v (qR/2 - vn/2) + L_v dqL/2 + R_v dqR/2
    //         (+ n_v (dpL + dpR)/2 for v = 1, 2; + (dpL qL + dpR qR + pL dqL + pR dqR)/2 for v = 3),
    // and of each side's e2 d2 - e4 d4 with d4 = (rr - ll) - 3 d2:
    //   (de2 + 3 de4) d2 - de4 (rr - ll) + (e2 + 3 e4) dd2 - e4 (drr - dll).
    // f_L = t - D_L, f_R = D_R - t.
    const double aL = 0.5 * qL - f.hvn, aR = 0.5 * qR - f.hvn;
    const double hqL = 0.5 * dqL, hqR = 0.5 * dqR;
    const double hp = 0.5 * (dp[1] + dp[2]);
    const double ep = 0.5 * (dp[1] * qL + dp[2] * qR) + (pL * hqL + pR * hqR);
    const double cL2 = e2L + 3.0 * e4L, cR2 = e2R + 3.0 * e4R;
    const double dcL2 = de2L + 3.0 * de4L, dcR2 = de2R + 3.0 * de4R;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double Lv = st(1, v), Rv = st(2, v), dLv = dst(1, v), dRv = dst(2, v);
        double t = dLv * aL + dRv * aR + Lv * hqL + Rv * hqR;
        if (v == 1) t = t + f.n0 * hp;
        if (v == 2) t = t + f.n1 * hp;
        if (v == 3) t = t + ep;
        double fL = t, fR = -t;
        if ((bL | bR) & kLinDiss) {
            const double d2 = Rv - Lv, dd = st(3, v) - st(0, v);
            const double dd2 = dRv - dLv, ddd = dst(3, v) - dst(0, v);
            if (bL & kLinDiss) fL = t - (dcL2 * d2 - de4L * dd + (cL2 * dd2 - e4L * ddd));
            if (bR & kLinDiss) fR = (dcR2 * d2 - de4R * dd + (cR2 * dd2 - e4R * ddd)) - t;
        }
        outL(v, fL);
        outR(v, fR);
    }
}

}  // namespace pea
