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
#include "pea_physics.hpp"

namespace pea {

// |pa - 2 pb + pc| / (pa + 2 pb + pc): the pressure sensor of f90:209-210 for the triple
// (a, b, c) in this order.
template <class T>
PEA_HD T sensor(const T& pa, const T& pb, const T& pc) {
    return sabs(pa - 2.0 * pb + pc) / (pa + 2.0 * pb + pc);
}

// Dissipation of one side (f90:207-215): returns D with f -= D.
//   lam_self/lam_other: wave integrals of this side's l and r;  s1 = sensor(ll, l, r),
//   s2 = sensor(l, r, rr);  l, r, ll, rr: states as seen from this side.
template <class T>
PEA_HD void side_dissipation(const pea_params& P, const T& lam_l, const T& lam_r, const T& s1,
                             const T& s2, const T* ll, const T* l, const T* r, const T* rr,
                             T* D) {
    const T lam = maxb(lam_l, lam_r);
    const T s = maxb(s1, s2);
    const T e2 = P.k2 * s * lam;
    const T e4 = maxb(lit<T>(0.0), P.k4 * lam - P.c4 * e2);
#pragma unroll
    for (int v = 0; v < 4; ++v)
        D[v] = e2 * (r[v] - l[v]) - e4 * (rr[v] - 3.0 * r[v] + 3.0 * l[v] - ll[v]);
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

// Both fluxes of one face.  n is L's normal of the face (pointing to R), nR is R's own normal
// of it: the same edge traversed the other way, so nR == -n except for the signs of exact
// zeros, which decide the tangent of |n.j/rho| in R's evaluation (abs_c, f90:97-102).
// p* and c* are the cells' pressures and sound speeds; ll/rr (and their pressures) are only
// read by a dissipative side.  f_L and f_R are the contributions to the residual sums of L and
// R, those of face_flux() evaluated from each side.
template <class T, class G>
PEA_HD void face_flux2(const pea_params& P, const T* LL, const T* L, const T* R, const T* RR,
                       const T& pLL, const T& pL, const T& pR, const T& pRR, const T& cL,
                       const T& cR, const G* n, const G* nR, const G* gv, bool dissL,
                       bool dissR, T* fL, T* fR) {
    const T qL = (n[0] * L[1] + n[1] * L[2]) / L[0];
    const T qR = (n[0] * R[1] + n[1] * R[2]) / R[0];
    const G vn = gv[0] * n[0] + gv[1] * n[1];
    T cen[4], ale[4];
    {
        const T phL[4] = {L[0] * qL, L[1] * qL + n[0] * pL, L[2] * qL + n[1] * pL,
                          (L[3] + pL) * qL};
        const T phR[4] = {R[0] * qR, R[1] * qR + n[0] * pR, R[2] * qR + n[1] * pR,
                          (R[3] + pR) * qR};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            cen[v] = (phL[v] + phR[v]) / 2.0;
            ale[v] = (L[v] + R[v]) / 2.0 * vn;
        }
    }
    T DL[4], DR[4];
    if (dissL || dissR) {
        const G sn = ssqrt(n[0] * n[0] + n[1] * n[1]);
        if (dissL) {
            const T lamL = sabs(qL) + cL * sn;
            const T lamR = sabs(qR) + cR * sn;
            side_dissipation(P, lamL, lamR, sensor(pLL, pL, pR), sensor(pL, pR, pRR), LL, L, R,
                             RR, DL);
        }
        if (dissR) {
            // R's view: q' = (nR . j) / rho = -q up to the sign of an exact zero
            const double sL = -sign1(val(nR[0]) * val(L[1]) + val(nR[1]) * val(L[2]));
            const double sR = -sign1(val(nR[0]) * val(R[1]) + val(nR[1]) * val(R[2]));
            const T lamL = abs_signed(qL, sL) + cL * sn;
            const T lamR = abs_signed(qR, sR) + cR * sn;
            side_dissipation(P, lamR, lamL, sensor(pRR, pR, pL), sensor(pR, pL, pLL), RR, R, L,
                             LL, DR);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        T a = cen[v];
        if (dissL) a = a - DL[v];
        fL[v] = a - ale[v];
        T b = -cen[v];
        if (dissR) b = b - DR[v];
        fR[v] = b + ale[v];
    }
}

// Per-cell pressure and sound speed sqrt(gamma p / rho) (f90:128-133, 147).
template <class T>
PEA_HD void cell_pc(const T* u, T& p, T& c) {
    p = pressure(u);
    c = ssqrt(kGamma * p / u[0]);
}

}  // namespace pea
