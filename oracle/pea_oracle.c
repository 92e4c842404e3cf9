/*
 * CPU oracle for the PyEulerALE spatial discretization -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C99 restatement of the reference's complex(8) arithmetic, used by the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark as the checker.  Nothing in the
 * product (the HIP library and the Python package) may link, load or call it.
 *
 * Array layout is the reference's Fortran order so that results compare one-to-one with the
 * compiled reference (oracle/_ref):
 *   vertices/velocities  X[d + 2*(m + (M+1)*n)]           d<2, m<=M, n<=N
 *   states/odes          U[v + 4*(m + M*n)]               v<4, m<M,  n<N
 *   forces               F[d + 2*n]
 *   dR/dU blocks         J[i + 4*j + 16*(m + M*n) + 16*C*s]  s<9  (row i = output var)
 *   dR/dX, dR/dV blocks  J[i + 4*j + 8*(m + M*n) + 8*C*s]    s<4, j<2
 *   dR/dMa               J[i + 4*(m + M*n)]
 *   dF/dU                J[d + 2*v + 8*n]
 *   dF/dX                J[d + 2*e + 4*n + 4*N*s]             s<2
 * Indices are 0-based here (the reference is 1-based).
 *
 * Each function is pinned by tests/test_oracle.py: fixtures written by the compiled reference
 * (oracle/_ref), fresh random cases against it, invariants and finite differences.
 */
#include <complex.h>
#include <math.h>
#include <string.h>
#include <stdlib.h>

typedef double complex cplx;

typedef struct {
    double aoa;   /* degrees */
    double k2, k4, c4;
} orc_params;

#define GAMMA 1.4
static const double DEG_TO_RAD = 0.0174532925199432957692369;  /* spatial_discretization.f90:54 */
static const double CS_STEP = 1e-12;                           /* spatial_discretization.f90:52 */

/* |x| with complex-step derivative sign(1,Re x)*Im x -- spatial_discretization.f90:97-102 */
static cplx cabs_cs(cplx x) {
    double r = creal(x), i = cimag(x);
    return CMPLX(fabs(r), i * copysign(1.0, r));
}

/* biased maximum: x only if Re x exceeds Re y by more than 1e-8 -- f90:111-120 */
static cplx cmax_b(cplx x, cplx y) { return (creal(x) > creal(y) + 1e-8) ? x : y; }

/* p = (gamma-1)(E - j.j/rho/2) -- f90:128-133 */
static cplx pressure(const cplx *u) {
    return (GAMMA - 1.0) * (u[3] - (u[1] * u[1] + u[2] * u[2]) / u[0] / 2.0);
}

/* Lambda = |n.j/rho| + sqrt(gamma p/rho) |n| -- f90:142-148 */
static cplx wave_integral(const cplx *u, const cplx *nrm) {
    return cabs_cs((nrm[0] * u[1] + nrm[1] * u[2]) / u[0]) +
           csqrt(GAMMA * pressure(u) / u[0]) * csqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1]);
}

/* physical flux through an edge with normal n -- f90:159-171 */
static void physical_flux(const cplx *u, const cplx *nrm, cplx *f) {
    cplx q = (nrm[0] * u[1] + nrm[1] * u[2]) / u[0];
    cplx p = pressure(u);
    f[0] = u[0] * q;
    f[1] = u[1] * q + nrm[0] * p;
    f[2] = u[2] * q + nrm[1] * p;
    f[3] = (u[3] + p) * q;
}

/* JST face flux with ALE term; ll/rr may be NULL (central only) -- f90:189-220 */
static void face_flux(const orc_params *prm, const cplx *ll, const cplx *l, const cplx *r,
                      const cplx *rr, const cplx *nrm, const cplx *gv, cplx *f) {
    cplx fl[4], fr[4];
    physical_flux(l, nrm, fl);
    physical_flux(r, nrm, fr);
    for (int v = 0; v < 4; ++v) f[v] = (fl[v] + fr[v]) / 2.0;
    if (ll && rr) {
        cplx lam = cmax_b(wave_integral(l, nrm), wave_integral(r, nrm));
        cplx pll = pressure(ll), pl = pressure(l), pr = pressure(r), prr = pressure(rr);
        cplx s = cmax_b(cabs_cs(pll - 2.0 * pl + pr) / (pll + 2.0 * pl + pr),
                        cabs_cs(pl - 2.0 * pr + prr) / (pl + 2.0 * pr + prr));
        cplx e2 = prm->k2 * s * lam;
        cplx e4 = cmax_b(0.0, prm->k4 * lam - prm->c4 * e2);
        for (int v = 0; v < 4; ++v)
            f[v] = f[v] - (e2 * (r[v] - l[v]) - e4 * (rr[v] - 3.0 * r[v] + 3.0 * l[v] - ll[v]));
    }
    cplx vn = gv[0] * nrm[0] + gv[1] * nrm[1];
    for (int v = 0; v < 4; ++v) f[v] = f[v] - (l[v] + r[v]) / 2.0 * vn;
}

/* free-stream conservative state -- f90:250-259 */
static void free_stream(const orc_params *prm, cplx mach, cplx *u) {
    cplx spd = mach * sqrt(GAMMA);
    double a = DEG_TO_RAD * prm->aoa;
    u[0] = 1.0;
    u[1] = spd * cos(a);
    u[2] = spd * sin(a);
    u[3] = 1.0 / (GAMMA - 1.0) + spd * spd / 2.0;
}

/* wall ghost: momentum reflected relative to the wall velocity -- f90:231-240 */
static void wall_ghost(const cplx *u, const cplx *nrm, const cplx *gv, cplx *g) {
    cplx s = nrm[0] * (u[1] - gv[0] * u[0]) + nrm[1] * (u[2] - gv[1] * u[0]);
    cplx nn = nrm[0] * nrm[0] + nrm[1] * nrm[1];
    g[0] = u[0];
    g[1] = u[1] - 2.0 * nrm[0] * s / nn;
    g[2] = u[2] - 2.0 * nrm[1] * s / nn;
    g[3] = u[3];
}

/* normal of the edge a->b -- f90:306-311 */
static void edge_normal(const cplx *a, const cplx *b, cplx *nrm) {
    nrm[0] = b[1] - a[1];
    nrm[1] = a[0] - b[0];
}

/* Cell stencil: 9 states, 4 vertices, 4 velocities (pointers into caller storage). */
typedef struct {
    const cplx *c, *b, *bb, *f, *ff, *i, *ii, *o, *oo;
    const cplx *xbi, *xfi, *xfo, *xbo;
    const cplx *vbi, *vfi, *vfo, *vbo;
} stencil;

/* radial boundary flag, 1-based m -- f90:450-464 */
static int boundary_flag(int M, int m1) {
    if (m1 == 1) return 1;
    if (m1 == 2) return 2;
    if (m1 == M - 1) return 4;
    if (m1 == M) return 5;
    return 3;
}

/* cell rate of change -- f90:370-436 */
static void cell_ode(const orc_params *prm, cplx mach, const stencil *s, int flag, cplx *ode) {
    cplx d1[2] = {s->xfo[0] - s->xbi[0], s->xfo[1] - s->xbi[1]};
    cplx d2[2] = {s->xbo[0] - s->xfi[0], s->xbo[1] - s->xfi[1]};
    cplx area = cabs_cs(d1[0] * d2[1] - d1[1] * d2[0]) / 2.0;
    cplx ni[2], nf[2], no[2], nb[2];
    edge_normal(s->xbi, s->xfi, ni);
    edge_normal(s->xfi, s->xfo, nf);
    edge_normal(s->xfo, s->xbo, no);
    edge_normal(s->xbo, s->xbi, nb);
    cplx gi[2] = {(s->vbi[0] + s->vfi[0]) / 2.0, (s->vbi[1] + s->vfi[1]) / 2.0};
    cplx gf[2] = {(s->vfi[0] + s->vfo[0]) / 2.0, (s->vfi[1] + s->vfo[1]) / 2.0};
    cplx go[2] = {(s->vfo[0] + s->vbo[0]) / 2.0, (s->vfo[1] + s->vbo[1]) / 2.0};
    cplx gb[2] = {(s->vbo[0] + s->vbi[0]) / 2.0, (s->vbo[1] + s->vbi[1]) / 2.0};
    cplx fb[4], ff[4], fi[4], fo[4], ghost[4];
    face_flux(prm, s->f, s->c, s->b, s->bb, nb, gb, fb);
    face_flux(prm, s->b, s->c, s->f, s->ff, nf, gf, ff);
    switch (flag) {
    case 1:
        wall_ghost(s->c, ni, gi, ghost);
        face_flux(prm, NULL, s->c, ghost, NULL, ni, gi, fi);
        face_flux(prm, ghost, s->c, s->o, s->oo, no, go, fo);
        break;
    case 2:
        face_flux(prm, NULL, s->c, s->i, NULL, ni, gi, fi);
        face_flux(prm, s->i, s->c, s->o, s->oo, no, go, fo);
        break;
    case 3:
        face_flux(prm, s->o, s->c, s->i, s->ii, ni, gi, fi);
        face_flux(prm, s->i, s->c, s->o, s->oo, no, go, fo);
        break;
    case 4:
        face_flux(prm, s->o, s->c, s->i, s->ii, ni, gi, fi);
        face_flux(prm, NULL, s->c, s->o, NULL, no, go, fo);
        break;
    default: {
        cplx inf[4];
        free_stream(prm, mach, inf);
        for (int v = 0; v < 4; ++v) ghost[v] = 2.0 * inf[v] - s->c[v];
        face_flux(prm, ghost, s->c, s->i, s->ii, ni, gi, fi);
        face_flux(prm, NULL, s->c, ghost, NULL, no, go, fo);
    }
    }
    for (int v = 0; v < 4; ++v) ode[v] = -(fb[v] + ff[v] + fi[v] + fo[v]) / area;
}

static int wrap(int n, int N) { return ((n % N) + N) % N; }

/* Gathers the stencil of cell (m,n) (0-based).  Out-of-range radial neighbours are left NULL;
 * the boundary flag guarantees they are never dereferenced. */
static void gather(int M, int N, const cplx *X, const cplx *V, const cplx *U, int m, int n,
                   stencil *s) {
#define UAT(mm, nn) (U + 4 * ((mm) + M * (nn)))
#define XAT(A, mm, nn) (A + 2 * ((mm) + (M + 1) * (nn)))
    s->c = UAT(m, n);
    s->b = UAT(m, wrap(n - 1, N));
    s->bb = UAT(m, wrap(n - 2, N));
    s->f = UAT(m, wrap(n + 1, N));
    s->ff = UAT(m, wrap(n + 2, N));
    s->i = m >= 1 ? UAT(m - 1, n) : NULL;
    s->ii = m >= 2 ? UAT(m - 2, n) : NULL;
    s->o = m + 1 < M ? UAT(m + 1, n) : NULL;
    s->oo = m + 2 < M ? UAT(m + 2, n) : NULL;
    s->xbi = XAT(X, m, n);
    s->xfi = XAT(X, m, n + 1);
    s->xfo = XAT(X, m + 1, n + 1);
    s->xbo = XAT(X, m + 1, n);
    s->vbi = XAT(V, m, n);
    s->vfi = XAT(V, m, n + 1);
    s->vfo = XAT(V, m + 1, n + 1);
    s->vbo = XAT(V, m + 1, n);
}

/* ---------------------------------------------------------------- exported API */

/* f90:471-482 */
void orc_set_free_stream_state(const orc_params *prm, double mach_re, double mach_im, int M, int N, cplx *U) {
    const cplx mach = CMPLX(mach_re, mach_im);
    cplx inf[4];
    free_stream(prm, mach, inf);
    for (long k = 0; k < (long)M * N; ++k) memcpy(U + 4 * k, inf, sizeof inf);
}

/* f90:496-537 */
void orc_compute_odes(const orc_params *prm, double mach_re, double mach_im, int M, int N, const cplx *X,
                      const cplx *V, const cplx *U, cplx *R) {
    const cplx mach = CMPLX(mach_re, mach_im);
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < M; ++m) {
            stencil s;
            gather(M, N, X, V, U, m, n, &s);
            cell_ode(prm, mach, &s, boundary_flag(M, m + 1), R + 4 * (m + M * n));
        }
}

/* f90:552-593: only the far-field ring depends on the Mach number */
void orc_compute_odes_wrt_mach(const orc_params *prm, double mach_re, double mach_im, int M, int N, const cplx *X,
                               const cplx *V, const cplx *U, double *J) {
    const cplx mach = CMPLX(mach_re, mach_im);
    memset(J, 0, sizeof(double) * 4 * (size_t)M * N);
    int m = M - 1;
    for (int n = 0; n < N; ++n) {
        stencil s;
        gather(M, N, X, V, U, m, n, &s);
        cplx ode[4];
        cell_ode(prm, mach + I * CS_STEP, &s, boundary_flag(M, M), ode);
        for (int v = 0; v < 4; ++v) J[v + 4 * (m + M * n)] = cimag(ode[v]) / CS_STEP;
    }
}

/* f90:666-750: 9 real 4x4 blocks per cell by complex step.  Blocks whose neighbour does not
 * exist (radial boundary) are left untouched, as in the reference. */
static void odes_wrt_states(const orc_params *prm, cplx mach, int M, int N, const cplx *X,
                            const cplx *V, const cplx *U, const unsigned char *mask, double *J) {
    const long C = (long)M * N;
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < M; ++m) {
            if (mask && !mask[m + (long)M * n]) continue;
            stencil s;
            gather(M, N, X, V, U, m, n, &s);
            int flag = boundary_flag(M, m + 1);
            const cplx *base[9] = {s.c, s.b, s.bb, s.f, s.ff, s.i, s.ii, s.o, s.oo};
            int present[9] = {1, 1, 1, 1, 1, m > 0, m > 1, m < M - 1, m < M - 2};
            for (int blk = 0; blk < 9; ++blk) {
                if (!present[blk]) continue;
                for (int j = 0; j < 4; ++j) {
                    cplx pert[4];
                    memcpy(pert, base[blk], sizeof pert);
                    pert[j] += I * CS_STEP;
                    stencil sp = s;
                    const cplx **slot[9] = {&sp.c, &sp.b, &sp.bb, &sp.f, &sp.ff,
                                            &sp.i, &sp.ii, &sp.o, &sp.oo};
                    *slot[blk] = pert;
                    cplx ode[4];
                    cell_ode(prm, mach, &sp, flag, ode);
                    for (int i = 0; i < 4; ++i)
                        J[i + 4 * j + 16 * (m + (long)M * n) + 16 * C * blk] =
                            cimag(ode[i]) / CS_STEP;
                }
            }
        }
}

void orc_compute_odes_wrt_states(const orc_params *prm, double mach_re, double mach_im, int M, int N,
                                 const cplx *X, const cplx *V, const cplx *U, double *J) {
    odes_wrt_states(prm, CMPLX(mach_re, mach_im), M, N, X, V, U, NULL, J);
}

/* The same blocks for the cells with mask[m + M*n] != 0 only (other cells untouched): lets the
 * checks evaluate the complex-step Jacobian on a sample of a large grid. */
void orc_compute_odes_wrt_states_masked(const orc_params *prm, double mach_re, double mach_im,
                                        int M, int N, const cplx *X, const cplx *V,
                                        const cplx *U, const unsigned char *mask, double *J) {
    odes_wrt_states(prm, CMPLX(mach_re, mach_im), M, N, X, V, U, mask, J);
}

/* f90:933-993 (which = 0) and f90:1087-1143 (which = 1): 4 real 4x2 blocks per cell */
static void odes_wrt_grid(const orc_params *prm, cplx mach, int M, int N, const cplx *X,
                          const cplx *V, const cplx *U, const unsigned char *mask, double *J,
                          int which) {
    const long C = (long)M * N;
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < M; ++m) {
            if (mask && !mask[m + (long)M * n]) continue;
            stencil s;
            gather(M, N, X, V, U, m, n, &s);
            int flag = boundary_flag(M, m + 1);
            for (int vtx = 0; vtx < 4; ++vtx)
                for (int d = 0; d < 2; ++d) {
                    stencil sp = s;
                    const cplx **slot_x[4] = {&sp.xbi, &sp.xfi, &sp.xfo, &sp.xbo};
                    const cplx **slot_v[4] = {&sp.vbi, &sp.vfi, &sp.vfo, &sp.vbo};
                    const cplx **slot = which ? slot_v[vtx] : slot_x[vtx];
                    cplx pert[2] = {(*slot)[0], (*slot)[1]};
                    pert[d] += I * CS_STEP;
                    *slot = pert;
                    cplx ode[4];
                    cell_ode(prm, mach, &sp, flag, ode);
                    for (int i = 0; i < 4; ++i)
                        J[i + 4 * d + 8 * (m + (long)M * n) + 8 * C * vtx] =
                            cimag(ode[i]) / CS_STEP;
                }
        }
}

void orc_compute_odes_wrt_vertices(const orc_params *prm, double mach_re, double mach_im, int M, int N,
                                   const cplx *X, const cplx *V, const cplx *U, double *J) {
    const cplx mach = CMPLX(mach_re, mach_im);
    odes_wrt_grid(prm, mach, M, N, X, V, U, NULL, J, 0);
}

/* Vertex (which = 0) or velocity (which = 1) blocks of the masked cells only. */
void orc_compute_odes_wrt_grid_masked(const orc_params *prm, double mach_re, double mach_im,
                                      int M, int N, const cplx *X, const cplx *V, const cplx *U,
                                      const unsigned char *mask, int which, double *J) {
    odes_wrt_grid(prm, CMPLX(mach_re, mach_im), M, N, X, V, U, mask, J, which);
}

void orc_compute_odes_wrt_velocities(const orc_params *prm, double mach_re, double mach_im, int M, int N,
                                     const cplx *X, const cplx *V, const cplx *U, double *J) {
    const cplx mach = CMPLX(mach_re, mach_im);
    odes_wrt_grid(prm, mach, M, N, X, V, U, NULL, J, 1);
}

/* f90:604-621 */
void orc_apply_odes_wrt_mach_fwd(int M, int N, const double *J, const cplx *dmach, cplx *dR) {
    for (long k = 0; k < 4L * M * N; ++k) dR[k] = J[k] * dmach[0];
}

/* f90:632-650 */
void orc_apply_odes_wrt_mach_rev(int M, int N, const double *J, const cplx *dR, cplx *dmach) {
    cplx acc = 0.0;
    for (long k = 0; k < (long)M * N; ++k) {
        cplx t = 0.0;
        for (int v = 0; v < 4; ++v) t += dR[4 * k + v] * J[4 * k + v];
        acc += t;
    }
    dmach[0] = acc;
}

/* neighbour cell index of block blk for cell (m,n); -1 if absent -- f90:836-861 */
static long nb_cell(int M, int N, int m, int n, int blk) {
    switch (blk) {
    case 0: return m + (long)M * n;
    case 1: return m + (long)M * wrap(n - 1, N);
    case 2: return m + (long)M * wrap(n - 2, N);
    case 3: return m + (long)M * wrap(n + 1, N);
    case 4: return m + (long)M * wrap(n + 2, N);
    case 5: return m > 0 ? (m - 1) + (long)M * n : -1;
    case 6: return m > 1 ? (m - 2) + (long)M * n : -1;
    case 7: return m < M - 1 ? (m + 1) + (long)M * n : -1;
    default: return m < M - 2 ? (m + 2) + (long)M * n : -1;
    }
}

/* f90:825-864 */
void orc_apply_odes_wrt_states_fwd(int M, int N, const double *J, const cplx *dU, cplx *dR) {
    const long C = (long)M * N;
    /* accumulation order of the reference: c, b, bb, f, ff, then ii, i, oo, o */
    static const int order[9] = {0, 1, 2, 3, 4, 6, 5, 8, 7};
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < M; ++m) {
            long cell = m + (long)M * n;
            cplx acc[4] = {0, 0, 0, 0};
            for (int t = 0; t < 9; ++t) {
                int blk = order[t];
                long nb = nb_cell(M, N, m, n, blk);
                if (nb < 0) continue;
                const double *B = J + 16 * cell + 16 * C * blk;
                for (int i = 0; i < 4; ++i) {
                    cplx s = 0.0;
                    for (int j = 0; j < 4; ++j) s += B[i + 4 * j] * dU[4 * nb + j];
                    acc[i] += s;
                }
            }
            for (int i = 0; i < 4; ++i) dR[4 * cell + i] = acc[i];
        }
}

/* f90:875-918 (scatter form, as the reference) */
void orc_apply_odes_wrt_states_rev(int M, int N, const double *J, const cplx *dR, cplx *dU) {
    const long C = (long)M * N;
    static const int order[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    memset(dU, 0, sizeof(cplx) * 4 * C);
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < M; ++m) {
            long cell = m + (long)M * n;
            for (int t = 0; t < 9; ++t) {
                int blk = order[t];
                long nb = nb_cell(M, N, m, n, blk);
                if (nb < 0) continue;
                const double *B = J + 16 * cell + 16 * C * blk;
                for (int j = 0; j < 4; ++j) {
                    cplx s = 0.0;
                    for (int i = 0; i < 4; ++i) s += dR[4 * cell + i] * B[i + 4 * j];
                    dU[4 * nb + j] += s;
                }
            }
        }
}

/* vertex index of slot vtx (bi, fi, fo, bo) of cell (m,n) */
static long vtx_index(int M, int m, int n, int vtx) {
    static const int dm[4] = {0, 0, 1, 1}, dn[4] = {0, 1, 1, 0};
    return (m + dm[vtx]) + (long)(M + 1) * (n + dn[vtx]);
}

/* f90:1005-1032 (also used for velocities) */
void orc_apply_odes_wrt_vertices_fwd(int M, int N, const double *J, const cplx *dX, cplx *dR) {
    const long C = (long)M * N;
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < M; ++m) {
            long cell = m + (long)M * n;
            cplx acc[4] = {0, 0, 0, 0};
            for (int vtx = 0; vtx < 4; ++vtx) {
                const double *B = J + 8 * cell + 8 * C * vtx;
                long vi = vtx_index(M, m, n, vtx);
                for (int i = 0; i < 4; ++i)
                    acc[i] += B[i] * dX[2 * vi] + B[i + 4] * dX[2 * vi + 1];
            }
            for (int i = 0; i < 4; ++i) dR[4 * cell + i] = acc[i];
        }
}

/* f90:1044-1072 */
void orc_apply_odes_wrt_vertices_rev(int M, int N, const double *J, const cplx *dR, cplx *dX) {
    const long C = (long)M * N;
    memset(dX, 0, sizeof(cplx) * 2 * (size_t)(M + 1) * (N + 1));
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < M; ++m) {
            long cell = m + (long)M * n;
            for (int vtx = 0; vtx < 4; ++vtx) {
                const double *B = J + 8 * cell + 8 * C * vtx;
                long vi = vtx_index(M, m, n, vtx);
                for (int d = 0; d < 2; ++d) {
                    cplx s = 0.0;
                    for (int i = 0; i < 4; ++i) s += dR[4 * cell + i] * B[i + 4 * d];
                    dX[2 * vi + d] += s;
                }
            }
        }
}

/* f90:1156-1174 */
void orc_compute_forces(int M, int N, const cplx *X, const cplx *U, cplx *F) {
    for (int n = 0; n < N; ++n) {
        cplx nrm[2];
        edge_normal(X + 2 * ((long)(M + 1) * n), X + 2 * ((long)(M + 1) * (n + 1)), nrm);
        cplx p = pressure(U + 4 * ((long)M * n));
        F[2 * n] = p * nrm[0];
        F[2 * n + 1] = p * nrm[1];
    }
}

/* f90:1188-1213 */
void orc_compute_forces_wrt_states(int M, int N, const cplx *X, const cplx *U, double *J) {
    for (int n = 0; n < N; ++n) {
        cplx nrm[2];
        edge_normal(X + 2 * ((long)(M + 1) * n), X + 2 * ((long)(M + 1) * (n + 1)), nrm);
        for (int j = 0; j < 4; ++j) {
            cplx u[4];
            memcpy(u, U + 4 * ((long)M * n), sizeof u);
            u[j] += I * CS_STEP;
            cplx p = pressure(u);
            J[0 + 2 * j + 8 * n] = cimag(p * nrm[0]) / CS_STEP;
            J[1 + 2 * j + 8 * n] = cimag(p * nrm[1]) / CS_STEP;
        }
    }
}

/* f90:1227-1256 */
void orc_compute_forces_wrt_vertices(int M, int N, const cplx *X, const cplx *U, double *J) {
    for (int n = 0; n < N; ++n) {
        cplx p = pressure(U + 4 * ((long)M * n));
        const cplx *a = X + 2 * ((long)(M + 1) * n), *b = X + 2 * ((long)(M + 1) * (n + 1));
        for (int d = 0; d < 2; ++d) {
            cplx ap[2] = {a[0], a[1]}, bp[2] = {b[0], b[1]}, nrm[2];
            ap[d] += I * CS_STEP;
            edge_normal(ap, b, nrm);
            J[0 + 2 * d + 4 * n] = cimag(p * nrm[0]) / CS_STEP;
            J[1 + 2 * d + 4 * n] = cimag(p * nrm[1]) / CS_STEP;
            bp[d] += I * CS_STEP;
            edge_normal(a, bp, nrm);
            J[0 + 2 * d + 4 * n + 4 * N] = cimag(p * nrm[0]) / CS_STEP;
            J[1 + 2 * d + 4 * n + 4 * N] = cimag(p * nrm[1]) / CS_STEP;
        }
    }
}

/* f90:1267-1288 */
void orc_apply_forces_wrt_vertices_fwd(int M, int N, const double *J, const cplx *dX, cplx *dF) {
    for (int n = 0; n < N; ++n) {
        const cplx *a = dX + 2 * ((long)(M + 1) * n), *b = dX + 2 * ((long)(M + 1) * (n + 1));
        for (int d = 0; d < 2; ++d) {
            const double *Ja = J + 4 * n, *Jb = J + 4 * n + 4 * N;
            dF[2 * n + d] = (Ja[d] * a[0] + Ja[d + 2] * a[1]) + (Jb[d] * b[0] + Jb[d + 2] * b[1]);
        }
    }
}

/* f90:1299-1322 */
void orc_apply_forces_wrt_vertices_rev(int M, int N, const double *J, const cplx *dF, cplx *dX) {
    memset(dX, 0, sizeof(cplx) * 2 * (size_t)(M + 1) * (N + 1));
    for (int n = 0; n < N; ++n) {
        cplx *a = dX + 2 * ((long)(M + 1) * n), *b = dX + 2 * ((long)(M + 1) * (n + 1));
        const double *Ja = J + 4 * n, *Jb = J + 4 * n + 4 * N;
        for (int e = 0; e < 2; ++e) {
            a[e] += dF[2 * n] * Ja[2 * e] + dF[2 * n + 1] * Ja[2 * e + 1];
            b[e] += dF[2 * n] * Jb[2 * e] + dF[2 * n + 1] * Jb[2 * e + 1];
        }
    }
}

/* f90:1333-1348 */
void orc_apply_forces_wrt_states_fwd(int M, int N, const double *J, const cplx *dU, cplx *dF) {
    for (int n = 0; n < N; ++n) {
        const cplx *u = dU + 4 * ((long)M * n);
        for (int d = 0; d < 2; ++d) {
            cplx s = 0.0;
            for (int j = 0; j < 4; ++j) s += J[d + 2 * j + 8 * n] * u[j];
            dF[2 * n + d] = s;
        }
    }
}

/* f90:1359-1375 */
void orc_apply_forces_wrt_states_rev(int M, int N, const double *J, const cplx *dF, cplx *dU) {
    memset(dU, 0, sizeof(cplx) * 4 * (size_t)M * N);
    for (int n = 0; n < N; ++n)
        for (int j = 0; j < 4; ++j)
            dU[4 * ((long)M * n) + j] = dF[2 * n] * J[2 * j + 8 * n] + dF[2 * n + 1] * J[1 + 2 * j + 8 * n];
}

/* Directional complex-step JVP of R along real directions: Im R(Ma+i h dMa, X+i h dX,
 * V+i h dV, U+i h dU)/h.  Any direction may be NULL (zero).  The method of the reference's
 * complex-step Jacobians (f90:45-53) applied matrix-free. */
void orc_cs_jvp(const orc_params *prm, double mach, int M, int N, const double *X,
                const double *V, const double *U, const double *dmach, const double *dX,
                const double *dV, const double *dU, double h, double *dR) {
    const long C = (long)M * N, nv = 2L * (M + 1) * (N + 1);
    cplx *Xc = malloc(sizeof(cplx) * nv), *Vc = malloc(sizeof(cplx) * nv);
    cplx *Uc = malloc(sizeof(cplx) * 4 * C), *Rc = malloc(sizeof(cplx) * 4 * C);
    for (long k = 0; k < nv; ++k) {
        Xc[k] = CMPLX(X[k], dX ? h * dX[k] : 0.0);
        Vc[k] = CMPLX(V[k], dV ? h * dV[k] : 0.0);
    }
    for (long k = 0; k < 4 * C; ++k) Uc[k] = CMPLX(U[k], dU ? h * dU[k] : 0.0);
    orc_compute_odes(prm, mach, dmach ? h * dmach[0] : 0.0, M, N, Xc, Vc, Uc, Rc);
    for (long k = 0; k < 4 * C; ++k) dR[k] = cimag(Rc[k]) / h;
    free(Xc); free(Vc); free(Uc); free(Rc);
}

/* O(nnz) export of the dR/dU blocks to 0-based BSR with ascending column indices, the
 * structure of convert_odes_wrt_states (f90:765-814) without its all-pairs scan.
 * data: (nnzb,4,4) row-major blocks [row i][col j]; returns nnzb. */
long orc_convert_odes_wrt_states(int M, int N, const double *J, double *data, int *indices,
                                 int *indptr) {
    const long C = (long)M * N;
    long ptr = 0;
    indptr[0] = 0;
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < M; ++m) {
            long row = m + (long)M * n;
            long cols[9];
            int blks[9], cnt = 0;
            for (int blk = 0; blk < 9; ++blk) {
                long nb = nb_cell(M, N, m, n, blk);
                if (nb < 0) continue;
                cols[cnt] = nb;
                blks[cnt++] = blk;
            }
            for (int a = 1; a < cnt; ++a)       /* insertion sort by column */
                for (int b = a; b > 0 && cols[b - 1] > cols[b]; --b) {
                    long tc = cols[b]; cols[b] = cols[b - 1]; cols[b - 1] = tc;
                    int tb = blks[b]; blks[b] = blks[b - 1]; blks[b - 1] = tb;
                }
            for (int a = 0; a < cnt; ++a) {
                const double *B = J + 16 * row + 16 * C * blks[a];
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) data[16 * ptr + 4 * i + j] = B[i + 4 * j];
                indices[ptr++] = (int)cols[a];
            }
            indptr[row + 1] = (int)ptr;
        }
    return ptr;
}
