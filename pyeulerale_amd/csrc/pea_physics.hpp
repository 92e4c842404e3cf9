// Primal spatial discretization, templated on the scalar type (pea_scalar.hpp).
//
// Finite-volume JST scheme with ALE face fluxes on a closed O-grid.  The arithmetic follows the
// reference expression by expression (spatial_discretization.f90:97-436), so that the FP64
// instantiation reproduces the reference's real part and the Cx instantiation its complex step.
#pragma once
#include "pea_scalar.hpp"
#include "../../include/pea.h"

namespace pea {

// stencil slots, in the reference's Jacobian block order (f90:699-707)
enum Slot { C = 0, B = 1, BB = 2, F = 3, FF = 4, I = 5, II = 6, O = 7, OO = 8 };
// cell vertex slots (f90:511-514)
enum Vtx { BI = 0, FI = 1, FO = 2, BO = 3 };

// radial boundary flag for 0-based radial index m (f90:450-464): 1 wall, 2 second ring,
// 3 interior, 4 second-to-last ring, 5 far field.  Priority as in the reference.
PEA_HD int boundary_flag(int M, int m) {
    if (m == 0) return 1;
    if (m == 1) return 2;
    if (m == M - 2) return 4;
    if (m == M - 1) return 5;
    return 3;
}

// p = (gamma-1)(E - j.j/rho/2)  (f90:128-133)
template <class T>
PEA_HD T pressure(const T* u) {
    return (kGamma - 1.0) * (u[3] - (u[1] * u[1] + u[2] * u[2]) / u[0] / 2.0);
}

// Normal of the edge a->b, |n| = edge length (f90:306-311).
template <class T>
PEA_HD void edge_normal(const T* a, const T* b, T* n) {
    n[0] = b[1] - a[1];
    n[1] = a[0] - b[0];
}

// Physical flux through an edge (f90:159-171); q = n.j/rho is returned for reuse.
// T: state scalar, G: geometry scalar (double when the geometry carries no tangent).
template <class T, class G>
PEA_HD void physical_flux(const T* u, const G* n, T* f, T& q, T& p) {
    q = (n[0] * u[1] + n[1] * u[2]) / u[0];
    p = pressure(u);
    f[0] = u[0] * q;
    f[1] = u[1] * q + n[0] * p;
    f[2] = u[2] * q + n[1] * p;
    f[3] = (u[3] + p) * q;
}

// JST numerical flux with the ALE term (f90:189-220).  With diss == false the face is central
// only (the reference's absent ll/rr optional arguments).
template <class T, class G>
PEA_HD void face_flux(const pea_params& P, const T* ll, const T* l, const T* r, const T* rr,
                      const G* n, const G* gv, bool diss, T* f) {
    T fl[4], fr[4], ql, pl, qr, pr;
    physical_flux(l, n, fl, ql, pl);
    physical_flux(r, n, fr, qr, pr);
#pragma unroll
    for (int v = 0; v < 4; ++v) f[v] = (fl[v] + fr[v]) / 2.0;
    if (diss) {
        const G sn = ssqrt(n[0] * n[0] + n[1] * n[1]);
        const T lam_l = sabs(ql) + ssqrt(kGamma * pl / l[0]) * sn;   // f90:146-147
        const T lam_r = sabs(qr) + ssqrt(kGamma * pr / r[0]) * sn;
        const T lam = maxb(lam_l, lam_r);
        const T pll = pressure(ll), prr = pressure(rr);
        const T s = maxb(sabs(pll - 2.0 * pl + pr) / (pll + 2.0 * pl + pr),
                         sabs(pl - 2.0 * pr + prr) / (pl + 2.0 * pr + prr));
        const T e2 = P.k2 * s * lam;
        const T e4 = maxb(lit<T>(0.0), P.k4 * lam - P.c4 * e2);
#pragma unroll
        for (int v = 0; v < 4; ++v)
            f[v] = f[v] - (e2 * (r[v] - l[v]) - e4 * (rr[v] - 3.0 * r[v] + 3.0 * l[v] - ll[v]));
    }
    const G vn = gv[0] * n[0] + gv[1] * n[1];
#pragma unroll
    for (int v = 0; v < 4; ++v) f[v] = f[v] - (l[v] + r[v]) / 2.0 * vn;
}

// Direction (cos, sin) of the free stream for the angle of attack in degrees (f90:250-259).
struct FsDir {
    double ca, sa;
};
inline FsDir fs_dir(const pea_params& P) {
    const double a = kDegToRad * P.aoa_deg;
    return FsDir{cos(a), sin(a)};
}

// Free-stream state for Mach number `mach` along direction (ca, sa) (f90:250-259).
template <class T>
PEA_HD void free_stream_dir(const T& mach, double ca, double sa, T* u) {
    const T spd = mach * sqrt(kGamma);
    u[0] = lit<T>(1.0);
    u[1] = spd * ca;
    u[2] = spd * sa;
    u[3] = spd * spd / 2.0 + 1.0 / (kGamma - 1.0);
}
template <class T>
PEA_HD void free_stream(const pea_params& P, const T& mach, T* u) {
    const double a = kDegToRad * P.aoa_deg;
    free_stream_dir(mach, cos(a), sin(a), u);
}

// Wall ghost state (f90:231-240).
template <class T, class G>
PEA_HD void wall_ghost(const T* u, const G* n, const G* gv, T* g) {
    const T s = n[0] * (u[1] - gv[0] * u[0]) + n[1] * (u[2] - gv[1] * u[0]);
    const G nn = n[0] * n[0] + n[1] * n[1];
    g[0] = u[0];
    g[1] = u[1] - 2.0 * n[0] * s / nn;
    g[2] = u[2] - 2.0 * n[1] * s / nn;
    g[3] = u[3];
}

// Cell geometry derived from the vertices: area, outward normals and face velocities.
template <class T>
struct CellGeom {
    T area;
    T n[4][2];    // faces b, f, i, o
    T gv[4][2];
};
enum Face { FB = 0, FF_ = 1, FI_ = 2, FO_ = 3 };

template <class T>
PEA_HD void cell_geometry(const T (*x)[2], const T (*w)[2], CellGeom<T>& g) {
    const T d1x = x[FO][0] - x[BI][0], d1z = x[FO][1] - x[BI][1];
    const T d2x = x[BO][0] - x[FI][0], d2z = x[BO][1] - x[FI][1];
    g.area = sabs(d1x * d2z - d1z * d2x) / 2.0;            // f90:287-293
    edge_normal(x[BO], x[BI], g.n[FB]);                    // f90:383-386
    edge_normal(x[FI], x[FO], g.n[FF_]);
    edge_normal(x[BI], x[FI], g.n[FI_]);
    edge_normal(x[FO], x[BO], g.n[FO_]);
#pragma unroll
    for (int d = 0; d < 2; ++d) {                          // f90:320-324, 387-390
        g.gv[FB][d] = (w[BO][d] + w[BI][d]) / 2.0;
        g.gv[FF_][d] = (w[FI][d] + w[FO][d]) / 2.0;
        g.gv[FI_][d] = (w[BI][d] + w[FI][d]) / 2.0;
        g.gv[FO_][d] = (w[FO][d] + w[BO][d]) / 2.0;
    }
}

// The four stencil arguments (ll, l, r, rr) of each face of a cell with boundary flag `flag`
// (f90:392-434).  Index 9 denotes the wall or far-field ghost state, -1 an absent argument
// (central-only face).
struct FaceArgs {
    int a[4][4];
};
PEA_HD FaceArgs face_args(int flag) {
    FaceArgs fa = {{{F, C, B, BB}, {B, C, F, FF}, {O, C, I, II}, {I, C, O, OO}}};
    switch (flag) {
    case 1:  // wall: inner face central against the ghost, ghost is ll of the outer face
        fa.a[FI_][0] = -1; fa.a[FI_][2] = 9; fa.a[FI_][3] = -1;
        fa.a[FO_][0] = 9;
        break;
    case 2:  // inner face central only
        fa.a[FI_][0] = -1; fa.a[FI_][3] = -1;
        break;
    case 4:  // outer face central only
        fa.a[FO_][0] = -1; fa.a[FO_][3] = -1;
        break;
    case 5:  // far field: ghost is ll of the inner face; outer face central against the ghost
        fa.a[FI_][0] = 9;
        fa.a[FO_][0] = -1; fa.a[FO_][2] = 9; fa.a[FO_][3] = -1;
        break;
    default:
        break;
    }
    return fa;
}

// Ghost state of a boundary cell (wall: flag 1, far field: flag 5).
template <class T, class G, class A>
PEA_HD void ghost_state(const pea_params& P, const A& mach, const T* uc, const CellGeom<G>& g,
                        int flag, T* gh) {
    if (flag == 1) {
        wall_ghost(uc, g.n[FI_], g.gv[FI_], gh);
    } else {
        A inf[4];
        free_stream(P, mach, inf);
#pragma unroll
        for (int v = 0; v < 4; ++v) gh[v] = 2.0 * inf[v] - uc[v];   // f90:269-274
    }
}

// Residual of one cell (f90:370-436): R = -(f_b + f_f + f_i + f_o) / area.
// `src.get(slot, u)` supplies the state of stencil slot `slot` (Slot order); states are fetched
// face by face so only one face's four arguments are live at a time.  `uc` is the centre state.
template <class T, class G, class A, class Src>
PEA_HD void cell_ode_src(const pea_params& P, const A& mach, const Src& src, const T* uc,
                         const CellGeom<G>& g, int flag, T* R) {
    T gh[4];
    if (flag == 1) {
        wall_ghost(uc, g.n[FI_], g.gv[FI_], gh);
    } else if (flag == 5) {
        A inf[4];
        free_stream(P, mach, inf);
#pragma unroll
        for (int v = 0; v < 4; ++v) gh[v] = 2.0 * inf[v] - uc[v];   // f90:269-274
    }
    const FaceArgs fa = face_args(flag);
    T sum[4];
#pragma unroll
    for (int fc = 0; fc < 4; ++fc) {
        // the four arguments by value, chosen with selects: a pointer chosen at run time among
        // private arrays would put them in scratch memory
        T arg[4][4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int a = fa.a[fc][k];
            const bool ghost = a == 9, own = a == C || a < 0;
            T nb[4];
            if (!ghost && !own) src.get(a, nb);
#pragma unroll
            for (int v = 0; v < 4; ++v) arg[k][v] = ghost ? gh[v] : own ? uc[v] : nb[v];
        }
        T f[4];
        face_flux(P, arg[0], arg[1], arg[2], arg[3], g.n[fc], g.gv[fc], fa.a[fc][0] >= 0, f);
#pragma unroll
        for (int v = 0; v < 4; ++v) sum[v] = fc == 0 ? f[v] : sum[v] + f[v];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) R[v] = -(sum[v] / g.area);
}

}  // namespace pea
