// Hand-written reverse mode of the discretization (pea_physics.hpp), templated on the primal
// scalar type T (double for the VJP, Dual<K> for tangent-over-adjoint second-order products)
// and on L adjoint lanes (L = 2 carries the real and imaginary part of a complex seed).
//
// Every primal intermediate needed by the reverse sweep is recomputed from the stencil, the
// dissipation switches included (biased maximum decisions, f90:207-213); nothing is stored.
#pragma once
#include "pea_physics.hpp"

namespace pea {

// ub += reverse of p(u) with seed pb:  p = (g-1)(u3 - (u1^2+u2^2)/u0/2)
template <class T, int L>
PEA_HD void pressure_rev(const T* u, const T inv0, const T* pb, T (*ub)[L]) {
    constexpr double g1 = kGamma - 1.0;
    const T a1 = g1 * u[1] * inv0, a2 = g1 * u[2] * inv0;
    const T a0 = g1 * (u[1] * u[1] + u[2] * u[2]) * inv0 * inv0 / 2.0;
#pragma unroll
    for (int k = 0; k < L; ++k) {
        ub[3][k] = ub[3][k] + g1 * pb[k];
        ub[1][k] = ub[1][k] - a1 * pb[k];
        ub[2][k] = ub[2][k] - a2 * pb[k];
        ub[0][k] = ub[0][k] + a0 * pb[k];
    }
}

// Reverse of face_flux.  fb[v][k]: seed on the face flux.  Accumulates into
//   sb[0..3][v][k]  adjoints of (ll, l, r, rr)
//   nb[d][k]        adjoint of the face normal
//   gvb[d][k]       adjoint of the face velocity
template <class T, int L>
PEA_HD void face_flux_rev(const pea_params& P, const T* ll, const T* l, const T* r,
                          const T* rr, const T* n, const T* gv, bool diss, const T (*fb)[L],
                          T (*sb)[4][L], T (*nb)[L], T (*gvb)[L]) {
    const T il = 1.0 / l[0], ir = 1.0 / r[0];
    const T ql = (n[0] * l[1] + n[1] * l[2]) / l[0];
    const T qr = (n[0] * r[1] + n[1] * r[2]) / r[0];
    const T pl = pressure(l), pr = pressure(r);
    const T vn = gv[0] * n[0] + gv[1] * n[1];

    T plb[L], prb[L], qlb[L], qrb[L];
#pragma unroll
    for (int k = 0; k < L; ++k) {
        // ALE: f -= (l + r)/2 * vn
        T vnb = lit<T>(0.0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const T t = fb[v][k] * vn / 2.0;
            sb[1][v][k] = sb[1][v][k] - t;
            sb[2][v][k] = sb[2][v][k] - t;
            vnb = vnb - fb[v][k] * (l[v] + r[v]) / 2.0;
        }
        gvb[0][k] = gvb[0][k] + vnb * n[0];
        gvb[1][k] = gvb[1][k] + vnb * n[1];
        nb[0][k] = nb[0][k] + vnb * gv[0];
        nb[1][k] = nb[1][k] + vnb * gv[1];
        plb[k] = lit<T>(0.0);
        prb[k] = lit<T>(0.0);
        qlb[k] = lit<T>(0.0);
        qrb[k] = lit<T>(0.0);
    }

    if (diss) {
        const T nn = n[0] * n[0] + n[1] * n[1];
        const T sn = ssqrt(nn);
        const T al = kGamma * pl / l[0], ar = kGamma * pr / r[0];
        const T cl = ssqrt(al), cr = ssqrt(ar);
        const T lam_l = sabs(ql) + cl * sn, lam_r = sabs(qr) + cr * sn;
        const bool take_l = maxb_takes_first(lam_l, lam_r);
        const T lam = take_l ? lam_l : lam_r;
        const T pll = pressure(ll), prr = pressure(rr);
        const T a1 = pll - 2.0 * pl + pr, b1 = pll + 2.0 * pl + pr;
        const T a2 = pl - 2.0 * pr + prr, b2 = pl + 2.0 * pr + prr;
        const T s1 = sabs(a1) / b1, s2 = sabs(a2) / b2;
        const bool take_s1 = maxb_takes_first(s1, s2);
        const T s = take_s1 ? s1 : s2;
        const T e2 = P.k2 * s * lam;
        const T t4 = P.k4 * lam - P.c4 * e2;
        const bool zero4 = maxb_takes_first(lit<T>(0.0), t4);
        const T e4 = zero4 ? lit<T>(0.0) : t4;
        const double sg1 = sign1(a1), sg2 = sign1(a2), sgl = sign1(ql), sgr = sign1(qr);
        const T ib1 = 1.0 / b1, ib2 = 1.0 / b2;
        T pllb[L], prrb[L];
#pragma unroll
        for (int k = 0; k < L; ++k) {
            // f -= e2 (r - l) - e4 (rr - 3r + 3l - ll)
            T e2b = lit<T>(0.0), e4b = lit<T>(0.0);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const T db = -fb[v][k];
                e2b = e2b + db * (r[v] - l[v]);
                e4b = e4b - db * (rr[v] - 3.0 * r[v] + 3.0 * l[v] - ll[v]);
                const T t2 = db * e2, t4v = db * e4;
                sb[2][v][k] = sb[2][v][k] + t2 + 3.0 * t4v;
                sb[1][v][k] = sb[1][v][k] - t2 - 3.0 * t4v;
                sb[3][v][k] = sb[3][v][k] - t4v;
                sb[0][v][k] = sb[0][v][k] + t4v;
            }
            const T t4b = zero4 ? lit<T>(0.0) : e4b;
            T lamb = P.k4 * t4b;
            e2b = e2b - P.c4 * t4b;
            const T sbar = P.k2 * lam * e2b;
            lamb = lamb + P.k2 * s * e2b;
            const T s1b = take_s1 ? sbar : lit<T>(0.0);
            const T s2b = take_s1 ? lit<T>(0.0) : sbar;
            const T a1b = sg1 * s1b * ib1, b1b = -(s1b * s1 * ib1);
            const T a2b = sg2 * s2b * ib2, b2b = -(s2b * s2 * ib2);
            pllb[k] = a1b + b1b;
            plb[k] = plb[k] - 2.0 * a1b + 2.0 * b1b + a2b + b2b;
            prb[k] = prb[k] + a1b + b1b - 2.0 * a2b + 2.0 * b2b;
            prrb[k] = a2b + b2b;
            const T llb = take_l ? lamb : lit<T>(0.0);
            const T lrb = take_l ? lit<T>(0.0) : lamb;
            qlb[k] = qlb[k] + sgl * llb;
            qrb[k] = qrb[k] + sgr * lrb;
            // lam_x = |q_x| + sqrt(gamma p_x / rho_x) * sn
            const T snb = llb * cl + lrb * cr;
            const T alb = llb * sn / (2.0 * cl), arb = lrb * sn / (2.0 * cr);
            plb[k] = plb[k] + kGamma * alb * il;
            prb[k] = prb[k] + kGamma * arb * ir;
            sb[1][0][k] = sb[1][0][k] - alb * al * il;
            sb[2][0][k] = sb[2][0][k] - arb * ar * ir;
            const T nnb = snb / (2.0 * sn);
            nb[0][k] = nb[0][k] + 2.0 * n[0] * nnb;
            nb[1][k] = nb[1][k] + 2.0 * n[1] * nnb;
        }
        pressure_rev<T, L>(ll, 1.0 / ll[0], pllb, sb[0]);
        pressure_rev<T, L>(rr, 1.0 / rr[0], prrb, sb[3]);
    }

    // central part: f = (phi(l) + phi(r)) / 2
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const T* u = side ? r : l;
        const T q = side ? qr : ql;
        const T p = side ? pr : pl;
        const T iu = side ? ir : il;
        T pbs[L];
#pragma unroll
        for (int k = 0; k < L; ++k) {
            const T f0 = fb[0][k] / 2.0, f1 = fb[1][k] / 2.0, f2 = fb[2][k] / 2.0,
                    f3 = fb[3][k] / 2.0;
            T qb = (side ? qrb[k] : qlb[k]) + f0 * u[0] + f1 * u[1] + f2 * u[2] +
                   f3 * (u[3] + p);
            pbs[k] = (side ? prb[k] : plb[k]) + f1 * n[0] + f2 * n[1] + f3 * q;
            T(*ub)[L] = sb[1 + side];
            ub[0][k] = ub[0][k] + f0 * q;
            ub[1][k] = ub[1][k] + f1 * q;
            ub[2][k] = ub[2][k] + f2 * q;
            ub[3][k] = ub[3][k] + f3 * q;
            nb[0][k] = nb[0][k] + f1 * p;
            nb[1][k] = nb[1][k] + f2 * p;
            // q = (n0 u1 + n1 u2) / u0
            const T t = qb * iu;
            ub[1][k] = ub[1][k] + t * n[0];
            ub[2][k] = ub[2][k] + t * n[1];
            ub[0][k] = ub[0][k] - t * q;
            nb[0][k] = nb[0][k] + t * u[1];
            nb[1][k] = nb[1][k] + t * u[2];
        }
        pressure_rev<T, L>(u, iu, pbs, sb[1 + side]);
    }
}

// Reverse of wall_ghost: g = (u0, u - 2 n (n.(j - gv u0)) / (n.n), u3).
template <class T, int L>
PEA_HD void wall_ghost_rev(const T* u, const T* n, const T* gv, const T (*gb)[L], T (*ub)[L],
                           T (*nb)[L], T (*gvb)[L]) {
    const T s = n[0] * (u[1] - gv[0] * u[0]) + n[1] * (u[2] - gv[1] * u[0]);
    const T nn = n[0] * n[0] + n[1] * n[1];
    const T inn = 1.0 / nn;
#pragma unroll
    for (int k = 0; k < L; ++k) {
        ub[0][k] = ub[0][k] + gb[0][k];
        ub[1][k] = ub[1][k] + gb[1][k];
        ub[2][k] = ub[2][k] + gb[2][k];
        ub[3][k] = ub[3][k] + gb[3][k];
        // t_d = 2 n_d s / nn,  g_d = u_d - t_d
        const T tb0 = -gb[1][k], tb1 = -gb[2][k];
        const T sbar = 2.0 * (n[0] * tb0 + n[1] * tb1) * inn;
        const T nnb = -(2.0 * n[0] * s * tb0 + 2.0 * n[1] * s * tb1) * inn * inn;
        nb[0][k] = nb[0][k] + 2.0 * s * tb0 * inn + 2.0 * n[0] * nnb +
                   sbar * (u[1] - gv[0] * u[0]);
        nb[1][k] = nb[1][k] + 2.0 * s * tb1 * inn + 2.0 * n[1] * nnb +
                   sbar * (u[2] - gv[1] * u[0]);
        ub[1][k] = ub[1][k] + sbar * n[0];
        ub[2][k] = ub[2][k] + sbar * n[1];
        ub[0][k] = ub[0][k] - sbar * (n[0] * gv[0] + n[1] * gv[1]);
        gvb[0][k] = gvb[0][k] - sbar * n[0] * u[0];
        gvb[1][k] = gvb[1][k] - sbar * n[1] * u[0];
    }
}

// d(free stream)/d(mach): (0, sqrt(g) cos a, sqrt(g) sin a, mach g)
template <class T>
PEA_HD void free_stream_dmach(const pea_params& P, const T& mach, T* du) {
    const double a = kDegToRad * P.aoa_deg;
    du[0] = lit<T>(0.0);
    du[1] = lit<T>(sqrt(kGamma) * cos(a));
    du[2] = lit<T>(sqrt(kGamma) * sin(a));
    du[3] = mach * kGamma;
}

// Per-cell reverse state shared by the four face sweeps of one cell.
template <class T, int L>
struct CellAdj {
    CellGeom<T> g;
    T gh[4];
    T fb[4][L];         // seed on each face flux: -w / area
    T nb[4][2][L];      // adjoints of the four face normals
    T gvb[4][2][L];     // adjoints of the four face velocities
    T ghb[4][L];        // adjoint of the ghost state
    T areab[L];
};

template <class T, int L, class Src>
PEA_HD void cell_adj_init(const pea_params& P, const T& mach, const Src& src, const T* uc,
                          const T (*x)[2], const T (*xw)[2], int flag, const T (*w)[L],
                          bool want_area, CellAdj<T, L>& ca) {
    cell_geometry(x, xw, ca.g);
    if (flag == 1 || flag == 5) ghost_state<T, T, T>(P, mach, uc, ca.g, flag, ca.gh);
    const T ia = 1.0 / ca.g.area;
#pragma unroll
    for (int k = 0; k < L; ++k) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ca.fb[v][k] = -(w[v][k] * ia);
            ca.ghb[v][k] = lit<T>(0.0);
        }
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                ca.nb[f][d][k] = lit<T>(0.0);
                ca.gvb[f][d][k] = lit<T>(0.0);
            }
        ca.areab[k] = lit<T>(0.0);
    }
    if (want_area) {
        // R = -S/area  =>  dR/darea = -R/area ; areab = -sum_v w_v R_v / area
        T R[4];
        cell_ode_src(P, mach, src, uc, ca.g, flag, R);
#pragma unroll
        for (int k = 0; k < L; ++k) {
            T acc = lit<T>(0.0);
#pragma unroll
            for (int v = 0; v < 4; ++v) acc = acc + w[v][k] * R[v];
            ca.areab[k] = -(acc * ia);
        }
    }
}

// Reverse of face `fc` of a cell.  tb[slot][v][k] receives the adjoint of the face's four
// stencil arguments (ll, l, r, rr); ghost contributions are moved into ca.ghb and absent
// arguments get zero.
template <class T, int L, class Src>
PEA_HD void cell_face_rev(const pea_params& P, const Src& src, const T* uc, int flag, int fc,
                          CellAdj<T, L>& ca, T (*tb)[4][L]) {
    const FaceArgs fa = face_args(flag);
    // the four arguments by value, chosen with selects: a pointer chosen at run time among
    // private arrays would put them in scratch memory
    T arg[4][4];
    bool ghost_at[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int a = fa.a[fc][j];
        ghost_at[j] = a == 9;
        const bool own = a == C || a < 0;
        T nb_[4];
        if (!ghost_at[j] && !own) src.get(a, nb_);
#pragma unroll
        for (int v = 0; v < 4; ++v) arg[j][v] = ghost_at[j] ? ca.gh[v] : own ? uc[v] : nb_[v];
    }
    const bool diss = fa.a[fc][0] >= 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int k = 0; k < L; ++k) tb[j][v][k] = lit<T>(0.0);
    face_flux_rev<T, L>(P, arg[0], arg[1], arg[2], arg[3], ca.g.n[fc], ca.g.gv[fc], diss,
                        ca.fb, tb, ca.nb[fc], ca.gvb[fc]);
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (ghost_at[j]) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int k = 0; k < L; ++k) {
                    ca.ghb[v][k] = ca.ghb[v][k] + tb[j][v][k];
                    tb[j][v][k] = lit<T>(0.0);
                }
        }
}

// After the four faces: fold the ghost adjoint into the centre cell (tbc[v][k]) and the wall
// normal/velocity, and the far-field ghost into the Mach adjoint.
template <class T, int L>
PEA_HD void cell_ghost_rev(const pea_params& P, const T& mach, const T* uc, int flag,
                           CellAdj<T, L>& ca, T (*tbc)[L], T* machb) {
    if (flag == 1) {
        wall_ghost_rev<T, L>(uc, ca.g.n[FI_], ca.g.gv[FI_], ca.ghb, tbc, ca.nb[FI_],
                             ca.gvb[FI_]);
    } else if (flag == 5) {
        T du[4];
        free_stream_dmach(P, mach, du);
#pragma unroll
        for (int k = 0; k < L; ++k) {
            T acc = lit<T>(0.0);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                tbc[v][k] = tbc[v][k] - ca.ghb[v][k];
                acc = acc + 2.0 * ca.ghb[v][k] * du[v];
            }
            machb[k] = machb[k] + acc;
        }
    }
}

// Vertex and velocity adjoints of a cell from its normal, face-velocity and area adjoints.
// xb[vtx][d][k], wb[vtx][d][k] are overwritten.
template <class T, int L>
PEA_HD void cell_vertex_rev(const T (*x)[2], const CellAdj<T, L>& ca, T (*xb)[2][L],
                            T (*wb)[2][L]) {
    // normal(a->b) = (b_z - a_z, a_x - b_x) ; faces b: BO->BI, f: FI->FO, i: BI->FI, o: FO->BO
    const int ea[4] = {BO, FI, BI, FO}, eb[4] = {BI, FO, FI, BO};
#pragma unroll
    for (int vtx = 0; vtx < 4; ++vtx)
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
            for (int k = 0; k < L; ++k) {
                xb[vtx][d][k] = lit<T>(0.0);
                wb[vtx][d][k] = lit<T>(0.0);
            }
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int k = 0; k < L; ++k) {
            const int a = ea[f], b = eb[f];
            xb[b][1][k] = xb[b][1][k] + ca.nb[f][0][k];
            xb[a][1][k] = xb[a][1][k] - ca.nb[f][0][k];
            xb[a][0][k] = xb[a][0][k] + ca.nb[f][1][k];
            xb[b][0][k] = xb[b][0][k] - ca.nb[f][1][k];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const T h = ca.gvb[f][d][k] / 2.0;
                wb[a][d][k] = wb[a][d][k] + h;
                wb[b][d][k] = wb[b][d][k] + h;
            }
        }
    // area = |d1 x d2| / 2, d1 = fo - bi, d2 = bo - fi
    const T d1x = x[FO][0] - x[BI][0], d1z = x[FO][1] - x[BI][1];
    const T d2x = x[BO][0] - x[FI][0], d2z = x[BO][1] - x[FI][1];
    const double sg = sign1(d1x * d2z - d1z * d2x);
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const T cb = sg * ca.areab[k] / 2.0;
        const T d1xb = cb * d2z, d2zb = cb * d1x, d1zb = -(cb * d2x), d2xb = -(cb * d1z);
        xb[FO][0][k] = xb[FO][0][k] + d1xb;
        xb[BI][0][k] = xb[BI][0][k] - d1xb;
        xb[FO][1][k] = xb[FO][1][k] + d1zb;
        xb[BI][1][k] = xb[BI][1][k] - d1zb;
        xb[BO][0][k] = xb[BO][0][k] + d2xb;
        xb[FI][0][k] = xb[FI][0][k] - d2xb;
        xb[BO][1][k] = xb[BO][1][k] + d2zb;
        xb[FI][1][k] = xb[FI][1][k] - d2zb;
    }
}

}  // namespace pea
