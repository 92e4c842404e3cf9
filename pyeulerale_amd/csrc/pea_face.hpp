// Face-based evaluation of the JST/ALE flux: one evaluation per face serves both adjacent
// cells.
//
// The reference evaluates every face twice, once from each cell (spatial_discretization.f90:
// 395-433).  For the face between L and R with normal n (pointing L -> R) the two evaluations are
//   f_L = face_flux(LL, L, R, RR,  n, gv)      (L's f or o face)
//   f_R = face_flux(RR, R, L, LL, -n, gv)      (R's b or i face)
// Their central and ALE parts are exact negatives of each other in floating point (negating n
// negates q = n.j/rho and p n exactly; sums commute), so they are computed once.  The
// dissipation is not antisymmetric: the biased maximum (f90:111-120) picks by argument order and
// the sensors differ in summation order, so each side keeps its own sensors, Lambda choice and
// coefficients, evaluated with exactly the reference's expressions.  Pressures and sound speeds
// sqrt(gamma p / rho) are per-cell quantities computed once (same expressions as f90:128-147).
#pragma once
#include <type_traits>

#include "pea_physics.hpp"

namespace pea {

// |pa - 2 pb + pc| / (pa + 2 pb + pc): the pressure sensor of f90:209-210 for the triple
// (a, b, c) in this order.
template <class T>
PEA_HD T sensor(const T& pa, const T& pb, const T& pc) {
    return qdiv(sabs(pa - 2.0 * pb + pc), pa + 2.0 * pb + pc);
}

// |q| whose tangent is s * dq: the derivative convention of abs_c (f90:97-102) for a value
// q' = -q that R's side computes with its own normal; only the sign of an exact zero differs.
PEA_HD double abs_signed(double q, double) { return fabs(q); }
template <int K>
PEA_HD Dual<K> abs_signed(const Dual<K>& q, double s) {
    Dual<K> r;
    r.v = fabs(q.v);
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = s * q.d[k];
    return r;
}

// Switch coefficients of one side (f90:207-213): e2 and e4 with D = e2 d2 - e4 d4.
template <class T>
PEA_HD void side_coefs(const pea_params& P, const T& lam_l, const T& lam_r, const T& s1,
                       const T& s2, T& e2, T& e4) {
    const T lam = maxb(lam_l, lam_r);
    const T s = maxb(s1, s2);
    e2 = P.k2 * s * lam;
    e4 = maxb(lit<T>(0.0), P.k4 * lam - P.c4 * e2);
}

// Both fluxes of one face.  n is L's normal of the face (pointing to R), nR is R's own normal
// of it: the same edge traversed the other way, so nR == -n except for the signs of exact
// zeros, which decide the tangent of |n.j/rho| in R's evaluation (abs_c, f90:97-102).
// p* and c* are the cells' pressures and sound speeds; ll/rr (and their pressures) are only
// read by a dissipative side.  f_L and f_R are the contributions to the residual sums of L and
// R, those of face_flux() evaluated from each side.
// The stencil is read through an accessor and the results handed to sinks: st(slot, v) returns
// component v of slot 0..3 = ll, l, r, rr; outL(v, f) and outR(v, f) take the two sides'
// fluxes.  Only the scalar coefficients and one component of the stencil are live at a time
// (the states sit in LDS), which keeps the tangent types within registers.
template <class T, class G, class St, class OutL, class OutR>
PEA_HD void face_flux2_lds(const pea_params& P, const St& st, const T& pLL, const T& pL,
                           const T& pR, const T& pRR, const T& cL, const T& cR, const G* n,
                           const G* nR, const G* gv, bool dissL, bool dissR, OutL outL,
                           OutR outR) {
    T qL, qR;
    {
        const T L0 = st(1, 0), L1 = st(1, 1), L2 = st(1, 2);
        qL = qdiv(n[0] * L1 + n[1] * L2, L0);
        const T R0 = st(2, 0), R1 = st(2, 1), R2 = st(2, 2);
        qR = qdiv(n[0] * R1 + n[1] * R2, R0);
    }
    const G vn = gv[0] * n[0] + gv[1] * n[1];
    T e2L = lit<T>(0.0), e4L = lit<T>(0.0), e2R = lit<T>(0.0), e4R = lit<T>(0.0);
    if (dissL || dissR) {
        const G sn = ssqrt(n[0] * n[0] + n[1] * n[1]);
        if (dissL) {
            const T lamL = sabs(qL) + cL * sn;
            const T lamR = sabs(qR) + cR * sn;
            side_coefs(P, lamL, lamR, sensor(pLL, pL, pR), sensor(pL, pR, pRR), e2L, e4L);
        }
        if (dissR) {
            // R's view: q' = (nR . j) / rho = -q up to the sign of an exact zero
            const double sL = -sign1(val(nR[0]) * val(st(1, 1)) + val(nR[1]) * val(st(1, 2)));
            const double sR = -sign1(val(nR[0]) * val(st(2, 1)) + val(nR[1]) * val(st(2, 2)));
            const T lamL = abs_signed(qL, sL) + cL * sn;
            const T lamR = abs_signed(qR, sR) + cR * sn;
            side_coefs(P, lamR, lamL, sensor(pRR, pR, pL), sensor(pR, pL, pLL), e2R, e4R);
        }
    }
    // Central minus ALE and the dissipation regrouped per face as in face_flux2_cell:
    // t_v = L_v (qL/2 - vn/2) + R_v (qR/2 - vn/2) (+ n_v (pL + pR)/2, or (pL qL + pR qR)/2 for
    // v = 3), each side's e2 d2 - e4 d4 = (e2 + 3 e4) d2 - e4 (rr - ll), with the second
    // difference seen from L (R's side sees its negative; the reference sums R's in its own
    // order, a rounding apart).  f_L = t - D_L, f_R = D_R - t.
    const G hvn = vn / 2.0;
    const T aL = 0.5 * qL - hvn, aR = 0.5 * qR - hvn;
    const T hp = 0.5 * (pL + pR), ep = 0.5 * (pL * qL + pR * qR);
    const T cL2 = e2L + 3.0 * e4L, cR2 = e2R + 3.0 * e4R;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const T Lv = st(1, v), Rv = st(2, v);
        T t = Lv * aL + Rv * aR;
        if (v == 1 || v == 2) t = t + n[v - 1] * hp;
        if (v == 3) t = t + ep;
        T fL = t, fR = -t;
        if (dissL || dissR) {
            const T d2 = Rv - Lv, dd = st(3, v) - st(0, v);
            if (dissL) fL = t - (cL2 * d2 - e4L * dd);
            if (dissR) fR = (cR2 * d2 - e4R * dd) - t;
        }
        outL(v, fL);
        outR(v, fR);
    }
}

// face_flux2_lds in FP64 with per-cell quantities computed once per cell instead of in every
// face that reads them: iL, iR = 1/rho of L and R as rcp_fast gives them (so q = (n.j) iL is
// the same value qdiv forms), nuL, nuR = the pressure sensors of L and R along the face's grid
// line, |p_prev - 2 p + p_next| / (p_prev + 2 p + p_next) with prev the lower index.  L's side
// reads its sensors in the reference's order (f90:209-210) and gets them bitwise; R's side
// (f90 evaluates sensor(rr, r, l) and sensor(r, l, ll)) gets the same triples summed from the
// other end, which can differ by one rounding.  The flux terms are regrouped per face (below),
// a few roundings from the reference's association.
template <class St, class OutL, class OutR>
PEA_HD void face_flux2_cell(const pea_params& P, const St& st, double pL, double pR, double cL,
                            double cR, double iL, double iR, double nuL, double nuR,
                            const double* n, const double* gv, bool dissL, bool dissR,
                            OutL outL, OutR outR) {
    const double qL = (n[0] * st(1, 1) + n[1] * st(1, 2)) * iL;
    const double qR = (n[0] * st(2, 1) + n[1] * st(2, 2)) * iR;
    const double vn = gv[0] * n[0] + gv[1] * n[1];
    double e2L = 0.0, e4L = 0.0, e2R = 0.0, e4R = 0.0;
    if (dissL || dissR) {
        const double sn = sqrt_pos(n[0] * n[0] + n[1] * n[1]);
        const double lamL = fabs(qL) + cL * sn, lamR = fabs(qR) + cR * sn;
        if (dissL) side_coefs(P, lamL, lamR, nuL, nuR, e2L, e4L);
        if (dissR) side_coefs(P, lamR, lamL, nuR, nuL, e2R, e4R);
    }
    // Per component v the central flux minus the ALE term,
    //   t_v = (ph_L,v + ph_R,v)/2 - (L_v + R_v) vn/2
    //       = L_v (qL/2 - vn/2) + R_v (qR/2 - vn/2) + n_v (pL + pR)/2     (v = 1, 2)
    //                                               + (pL qL + pR qR)/2    (v = 3),
    // and each side's dissipation e2 d2 - e4 d4 with d4 = (rr - ll) - 3 d2 as
    // (e2 + 3 e4) d2 - e4 (rr - ll): the face's scalars are formed once, each component then
    // costs a few FMAs.  f_L = t - D_L and f_R = D_R - t.
    const double hvn = vn / 2.0;
    const double aL = 0.5 * qL - hvn, aR = 0.5 * qR - hvn;
    const double hp = 0.5 * (pL + pR), ep = 0.5 * (pL * qL + pR * qR);
    const double cL2 = e2L + 3.0 * e4L, cR2 = e2R + 3.0 * e4R;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double Lv = st(1, v), Rv = st(2, v);
        double t = Lv * aL + Rv * aR;
        if (v == 1 || v == 2) t = t + n[v - 1] * hp;
        if (v == 3) t = t + ep;
        double fL = t, fR = -t;
        if (dissL || dissR) {
            const double d2 = Rv - Lv, dd = st(3, v) - st(0, v);
            if (dissL) fL = t - (cL2 * d2 - e4L * dd);
            if (dissR) fR = (cR2 * d2 - e4R * dd) - t;
        }
        outL(v, fL);
        outR(v, fR);
    }
}

// ---------------------------------------------------------------------------- reverse
// Reverse of side_dissipation with seed db on D: accumulates into the adjoints of the side's
// ll, l, r, rr states (sb*), of its two wave integrals (lamb_l, lamb_r) and of the four
// pressures entering its sensors (pb_ll, pb_l, pb_r, pb_rr).  Switch decisions are recomputed.
// d2, dd: the face's differences r - l and rr - ll seen from its L side (the fourth difference
// is dd - 3 d2); flip = true for the R side, which sees their negatives.
template <class T, int L>
PEA_HD void side_dissipation_rev(const pea_params& P, const T& lam_l, const T& lam_r,
                                 const T& pll, const T& pl, const T& pr, const T& prr,
                                 const T* d2, const T* dd, bool flip,
                                 const T (*db)[L], T (*sbll)[L], T (*sbl)[L], T (*sbr)[L],
                                 T (*sbrr)[L], T* lamb_l, T* lamb_r, T* pbll, T* pbl, T* pbr,
                                 T* pbrr) {
    const bool take_l = maxb_takes_first(lam_l, lam_r);
    const T lam = take_l ? lam_l : lam_r;
    const T a1 = pll - 2.0 * pl + pr, b1 = pll + 2.0 * pl + pr;
    const T a2 = pl - 2.0 * pr + prr, b2 = pl + 2.0 * pr + prr;
    const T s1 = qdiv(sabs(a1), b1), s2 = qdiv(sabs(a2), b2);
    const bool take_s1 = maxb_takes_first(s1, s2);
    const T s = take_s1 ? s1 : s2;
    const T e2 = P.k2 * s * lam;
    const T t4 = P.k4 * lam - P.c4 * e2;
    const bool zero4 = maxb_takes_first(lit<T>(0.0), t4);
    const T e4 = zero4 ? lit<T>(0.0) : t4;
    const double sg1 = sign1(a1), sg2 = sign1(a2);
    const T ib1 = recip(b1), ib2 = recip(b2);
    // D = e2 d2 - e4 d4 with d4 = dd - 3 d2: the states see (e2 + 3 e4) d2 - e4 dd
    const T c2 = e2 + 3.0 * e4;
#pragma unroll
    for (int k = 0; k < L; ++k) {
        T e2b = adj_zero<T>(), sdd = adj_zero<T>();
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const T d = db[v][k], x = flip ? -d : d;
            e2b = e2b + x * d2[v];
            sdd = sdd + x * dd[v];
            const T t2 = d * c2, tq = d * e4;
            sbr[v][k] = sbr[v][k] + t2;
            sbl[v][k] = sbl[v][k] - t2;
            sbrr[v][k] = sbrr[v][k] - tq;
            sbll[v][k] = sbll[v][k] + tq;
        }
        T e4b = 3.0 * e2b - sdd;   // -(sum of x d4)
        const T t4b = zero4 ? lit<T>(0.0) : e4b;
        T lamb = P.k4 * t4b;
        e2b = e2b - P.c4 * t4b;
        const T sbar = P.k2 * lam * e2b;
        lamb = lamb + P.k2 * s * e2b;
        const T s1b = take_s1 ? sbar : lit<T>(0.0);
        const T s2b = take_s1 ? lit<T>(0.0) : sbar;
        const T a1b = sg1 * s1b * ib1, b1b = -(s1b * s1 * ib1);
        const T a2b = sg2 * s2b * ib2, b2b = -(s2b * s2 * ib2);
        pbll[k] = pbll[k] + a1b + b1b;
        pbl[k] = pbl[k] - 2.0 * a1b + 2.0 * b1b + a2b + b2b;
        pbr[k] = pbr[k] + a1b + b1b - 2.0 * a2b + 2.0 * b2b;
        pbrr[k] = pbrr[k] + a2b + b2b;
        lamb_l[k] = lamb_l[k] + (take_l ? lamb : lit<T>(0.0));
        lamb_r[k] = lamb_r[k] + (take_l ? lit<T>(0.0) : lamb);
    }
}

// Adjoint contributions of one face to its four cells (index 0..3 = LL, L, R, RR).
template <class T, int L>
struct FaceAdj {
    T u[4][4][L];   // direct state adjoints
    T p[4][L];      // pressure adjoints
    T c[2][L];      // sound-speed adjoints of L and R
};

template <class T, int L>
PEA_HD void face_adj_zero(FaceAdj<T, L>& a) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < L; ++k) {
#pragma unroll
            for (int v = 0; v < 4; ++v) a.u[j][v][k] = adj_zero<T>();
            a.p[j][k] = adj_zero<T>();
            if (j < 2) a.c[j][k] = adj_zero<T>();
        }
}

// Reverse of face_flux2_lds with seeds fbL on f_L and fbR on f_R (the faces' shares of the two
// cells' residual seeds, -w/area).  Accumulates into `a`; pressures and sound speeds are
// treated as inputs (their adjoints go to a.p / a.c).
template <class T, class G, int L>
PEA_HD void face_flux2_rev(const pea_params& P, const T* LL, const T* Lc, const T* R,
                           const T* RR, const T& pLL, const T& pL, const T& pR, const T& pRR,
                           const T& cL, const T& cR, const G* n, const G* nR, const G* gv,
                           bool dissL, bool dissR, const T (*fbL)[L], const T (*fbR)[L],
                           FaceAdj<T, L>& a) {
    const T iL = recip(Lc[0]), iR = recip(R[0]);
    const T qL = qdiv(n[0] * Lc[1] + n[1] * Lc[2], Lc[0]);
    const T qR = qdiv(n[0] * R[1] + n[1] * R[2], R[0]);
    const G vn = gv[0] * n[0] + gv[1] * n[1];
    T qLb[L], qRb[L];
#pragma unroll
    for (int k = 0; k < L; ++k) {
        qLb[k] = adj_zero<T>();
        qRb[k] = adj_zero<T>();
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            // f_L = cen - D_L - ale,  f_R = -cen - D_R + ale
            const T ceb = (fbL[v][k] - fbR[v][k]) / 2.0;    // seed on phi(L) and phi(R)
            const T aleb = (fbR[v][k] - fbL[v][k]) * vn / 2.0;
            a.u[1][v][k] = a.u[1][v][k] + aleb + ceb * qL;
            a.u[2][v][k] = a.u[2][v][k] + aleb + ceb * qR;
        }
        const T fL0 = (fbL[0][k] - fbR[0][k]) / 2.0, fL1 = (fbL[1][k] - fbR[1][k]) / 2.0,
                fL2 = (fbL[2][k] - fbR[2][k]) / 2.0, fL3 = (fbL[3][k] - fbR[3][k]) / 2.0;
        qLb[k] = fL0 * Lc[0] + fL1 * Lc[1] + fL2 * Lc[2] + fL3 * (Lc[3] + pL);
        qRb[k] = fL0 * R[0] + fL1 * R[1] + fL2 * R[2] + fL3 * (R[3] + pR);
        const T pn = fL1 * n[0] + fL2 * n[1];
        a.p[1][k] = a.p[1][k] + pn + fL3 * qL;
        a.p[2][k] = a.p[2][k] + pn + fL3 * qR;
    }
    if (dissL || dissR) {
        const G sn = ssqrt(n[0] * n[0] + n[1] * n[1]);
        T d2[4], dd[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            d2[v] = R[v] - Lc[v];
            dd[v] = RR[v] - LL[v];
        }
        if (dissL) {
            const double sL = sign1(qL), sR = sign1(qR);
            const T lamL = sabs(qL) + cL * sn, lamR = sabs(qR) + cR * sn;
            T db[4][L], lbl[L], lbr[L];
#pragma unroll
            for (int k = 0; k < L; ++k) {
                lbl[k] = lbr[k] = adj_zero<T>();
#pragma unroll
                for (int v = 0; v < 4; ++v) db[v][k] = -fbL[v][k];
            }
            side_dissipation_rev<T, L>(P, lamL, lamR, pLL, pL, pR, pRR, d2, dd, false, db,
                                       a.u[0], a.u[1], a.u[2], a.u[3], lbl, lbr, a.p[0],
                                       a.p[1], a.p[2], a.p[3]);
#pragma unroll
            for (int k = 0; k < L; ++k) {
                qLb[k] = qLb[k] + sL * lbl[k];
                qRb[k] = qRb[k] + sR * lbr[k];
                a.c[0][k] = a.c[0][k] + lbl[k] * sn;
                a.c[1][k] = a.c[1][k] + lbr[k] * sn;
            }
        }
        if (dissR) {
            // R's view: lam of q' = -q with R's own normal (signs of exact zeros from nR)
            const double sL = -sign1(val(nR[0]) * val(Lc[1]) + val(nR[1]) * val(Lc[2]));
            const double sR = -sign1(val(nR[0]) * val(R[1]) + val(nR[1]) * val(R[2]));
            const T lamL = sabs(qL) + cL * sn, lamR = sabs(qR) + cR * sn;
            T db[4][L], lbl[L], lbr[L];
#pragma unroll
            for (int k = 0; k < L; ++k) {
                lbl[k] = lbr[k] = adj_zero<T>();
#pragma unroll
                for (int v = 0; v < 4; ++v) db[v][k] = -fbR[v][k];
            }
            side_dissipation_rev<T, L>(P, lamR, lamL, pRR, pR, pL, pLL, d2, dd, true, db,
                                       a.u[3], a.u[2], a.u[1], a.u[0], lbr, lbl, a.p[3],
                                       a.p[2], a.p[1], a.p[0]);
#pragma unroll
            for (int k = 0; k < L; ++k) {
                qLb[k] = qLb[k] + sL * lbl[k];
                qRb[k] = qRb[k] + sR * lbr[k];
                a.c[0][k] = a.c[0][k] + lbl[k] * sn;
                a.c[1][k] = a.c[1][k] + lbr[k] * sn;
            }
        }
    }
    // q_X = (n0 X1 + n1 X2) / X0
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const T tl = qLb[k] * iL, tr = qRb[k] * iR;
        a.u[1][1][k] = a.u[1][1][k] + tl * n[0];
        a.u[1][2][k] = a.u[1][2][k] + tl * n[1];
        a.u[1][0][k] = a.u[1][0][k] - tl * qL;
        a.u[2][1][k] = a.u[2][1][k] + tr * n[0];
        a.u[2][2][k] = a.u[2][2][k] + tr * n[1];
        a.u[2][0][k] = a.u[2][0][k] - tr * qR;
    }
}

// Pressure (gamma-1)(E - j.j/rho/2) of f90:128-133 with the face kernels' quotient.
template <class T>
PEA_HD T pressure_q(const T* u) {
    return (kGamma - 1.0) * (u[3] - qdiv(u[1] * u[1] + u[2] * u[2], u[0]) / 2.0);
}

// Per-cell pressure and sound speed sqrt(gamma p / rho) (f90:128-133, 147).  A state that is
// not finite or not physical (rho <= 0 or p <= 0) gets a NaN pressure, so that the residual of
// every cell whose faces use it is NaN: the reference's complex arithmetic returns finite values
// there (an imaginary sound speed that the biased maximum of the dissipation may discard).
template <class T>
PEA_HD void cell_pc(const T* u, T& p, T& c) {
    p = pressure_q(u);
    if (!(val(u[0]) > 0.0 && val(p) > 0.0)) p = p * __builtin_nan("");
    c = ssqrt(qdiv(kGamma * p, u[0]));
}

}  // namespace pea
