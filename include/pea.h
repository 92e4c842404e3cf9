/*
 * pea.h -- C ABI of libpea_hip.so, the MI355X (gfx950) spatial discretization of PyEulerALE.
 *
 * Replaces the f2py module `py_euler_ale.euler_ale.spatial_discretization` (the reference's FFI,
 * src/spatial_discretization.f90:68-87, generated by meson.build:21-35).  Differences by design:
 *   - no module globals: angle of attack and the JST constants travel in `pea_params`
 *     (replaces the writable globals aoa, jst_k2, jst_k4, jst_c4 at f90:63-66);
 *   - all pointers are DEVICE pointers; every call is asynchronous on `stream` (a hipStream_t,
 *     NULL = default stream) and allocates nothing unless it says so;
 *   - layout is structure-of-arrays with the radial index fastest (coalesced on the GPU):
 *       states / odes       U[v*C + n*M + m]          (4 planes of M x N, C = M*N)
 *       vertices/velocities X[d*P + n*(M+1) + m]      (2 planes of (M+1) x (N+1), P=(M+1)(N+1))
 *       forces              F[d*N + n]
 *     Complex arrays (directions, adjoint seeds, products) are interleaved complex128 in the
 *     same element order (element e at doubles 2e, 2e+1).  Viewed as a tensor of shape
 *     (4, M, N) such a plane stack has strides (C, 1, M): the reference's index (v, m, n).
 *   - every entry point returns 0 on success or a negative pea_status; pea_last_error() gives
 *     the message (thread-local).  M >= 4 and N >= 5 are required (the reference reads out of
 *     range below that, f90:453-463, f90:786-803).
 *   - output semantics are "overwrite" for every output argument.
 */
#ifndef PEA_H
#define PEA_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pea_params {
    double aoa_deg;   /* free-stream angle of attack, degrees (f90:63) */
    double k2;        /* JST kappa_2 (f90:64), default 1     */
    double k4;        /* JST kappa_4 (f90:65), default 1/32  */
    double c4;        /* JST c_4     (f90:66), default 2     */
} pea_params;

/* One grid of a multi-grid launch: dimensions, flow parameters and device arrays in the
 * layouts above (R is written; dU/dR are the tangent direction and output of the multi-grid
 * JVP, complex interleaved, and may be NULL for the residual). */
typedef struct pea_grid {
    int M, N;
    double mach;
    pea_params prm;
    const double *X, *V, *U, *dU;
    double *R, *dR;
} pea_grid;

enum pea_status {
    PEA_OK = 0,
    PEA_ERR_DIMS = -1,    /* M < 4, N < 5 or sizes overflow */
    PEA_ERR_NULL = -2,    /* a required pointer is NULL */
    PEA_ERR_HIP = -3,     /* a HIP runtime call failed */
    PEA_ERR_ARG = -4      /* other invalid argument */
};

const char *pea_last_error(void);
int pea_version(void);

/* Tile shapes of the adjoint kernels (no counterpart in the reference: the tuner's control).
 * pea_num_tiles() shapes exist; pea_set_tile registers the shape used for grids of M x N cells
 * (M = N = 0: for every grid; tile = -1 clears); pea_get_tile reports the current choice. */
int pea_num_tiles(void);
int pea_set_tile(int M, int N, int tile);
int pea_get_tile(int M, int N);

/* Tile shapes of the face-based kernels (no counterpart in the reference: the autotuner's
 * control).  kind 0: residual (pea_compute_odes, pea_compute_odes_batch), 1: states tangent
 * (pea_jvp_odes along dU only, pea_jvp_odes_multi), 2: states adjoint (pea_vjp_odes to the
 * states only, pea_vjp_odes_states_multi); kinds 3-5: the same kernels in multi-grid launches
 * (pea_multi_prepare modes 0-2; process-wide, M = N = 0).  pea_num_face_tiles() shapes exist
 * per kind; pea_set_face_tile registers one for grids of M x N cells (M = N = 0: every grid;
 * tile = -1 clears); pea_get_face_tile returns the shape in use (and its cells per tile in *tm,
 * *tn when not NULL), -1 when no adjoint tile fits.  The shapes agree to round-off, not bitwise:
 * a face on a tile's edge is evaluated by other instructions than an interior face.  A
 * multi-grid launch is bitwise equal to single-grid calls with the same shape. */
int pea_num_face_tiles(void);
int pea_set_face_tile(int kind, int M, int N, int tile);
int pea_get_face_tile(int kind, int M, int N, int *tm, int *tn);

/* ------------------------------------------------------------------ primal */

/* Fill U with the free-stream state.  Replaces set_free_stream_state (f90:471-482). */
int pea_set_free_stream_state(const pea_params *prm, double mach, int M, int N, double *U,
                              void *stream);

/* Residual R(Ma, U, X, V), FP64.  Replaces compute_odes (f90:496-537). */
int pea_compute_odes(const pea_params *prm, double mach, int M, int N, const double *X,
                     const double *V, const double *U, double *R, void *stream);

/* Residuals of `nstates` states on one grid in one launch: U and R hold nstates consecutive
 * state arrays (stride 4*M*N doubles).  Same arithmetic as pea_compute_odes. */
int pea_compute_odes_batch(const pea_params *prm, double mach, int M, int N, int nstates,
                           const double *X, const double *V, const double *U, double *R,
                           void *stream);

/* Many grids of uneven size in one launch, one workgroup per tile (the reference evaluates one
 * grid per compute_odes call, f90:496-537).  Results are bitwise those of one single-grid call
 * per grid.
 * pea_multi_prepare: uploads the launch's descriptor table built from the host array `grids`
 *   (G entries) into `workspace` (device, pea_multi_workspace_bytes(G) bytes) and returns the
 *   tile count; it synchronises the stream.  tangent = 0: residuals (R written); tangent = 1:
 *   states-direction JVP (pea_jvp_odes with dU only; dR written, R too when non-NULL);
 *   tangent = 2: states adjoint (pea_vjp_odes to the states only), dU holding the seed w and
 *   dR receiving (dR/dU)^T w; every grid needs N >= the adjoint tile's width + 3 (18 for the
 *   default tile).
 * pea_multi_launch: one launch over the prepared table with the tile shape pea_multi_prepare
 *   chose for it (pea_set_face_tile kinds 3-5); it allocates and copies nothing, so it may be
 *   captured in a HIP graph and replayed while the table and arrays stay valid.
 * pea_compute_odes_multi / pea_jvp_odes_multigrid: prepare + launch. */
long long pea_multi_workspace_bytes(int G);
int pea_multi_prepare(int G, const pea_grid *grids, int tangent, void *workspace,
                      int *total_tiles, void *stream);
int pea_multi_launch(int G, int total_tiles, int tangent, const void *workspace, void *stream);
int pea_compute_odes_multi(int G, const pea_grid *grids, void *workspace, void *stream);
int pea_jvp_odes_multigrid(int G, const pea_grid *grids, void *workspace, void *stream);

/* Counts, on the device, the cells whose state is non-finite or non-physical (rho <= 0 or
 * p <= 0) into status[0] and, when R is not NULL, the cells with a non-finite residual into
 * status[1] (status: 2 ints on the device, overwritten).  The reference evaluates such states in
 * complex arithmetic and returns an imaginary wave speed (f90:142-148); here the residual of
 * every cell whose faces use such a state is NaN and this call reports it without a host
 * round trip. */
int pea_check_states(int M, int N, const double *U, const double *R, int *status, void *stream);

/* Surface forces F_n = p(U_{0,n}) normal(X_{0,n} -> X_{0,n+1}).  Replaces compute_forces
 * (f90:1156-1174).  If coef != NULL (device, 3 doubles) it also receives the reductions
 * (sum F_x, sum F_z, sum (z_p - zc) F_x - (x_p - xc) F_z) with x_p the face midpoints (the sums
 * behind core.py:1277-1363); xc, zc = moment reference point. */
int pea_compute_forces(int M, int N, const double *X, const double *U, double *F, double xc,
                       double zc, double *coef, void *stream);

/* ------------------------------------------------------------------ hand-written tangents */

/* dR = dR/dMa dMa + dR/dX dX + dR/dV dV + dR/dU dU at (Ma, X, V, U): hand-written forward
 * mode, complex directions (each may be NULL = zero; dmach is ONE complex value on the device).
 * R (optional, may be NULL) receives the primal residual of the same pass.
 * Replaces the products apply_odes_wrt_{mach,states,vertices}_fwd (f90:604-621, 825-864,
 * 1005-1032) after compute_odes_wrt_* (f90:552-593, 666-750, 933-993, 1087-1143), without
 * assembling any Jacobian. */
int pea_jvp_odes(const pea_params *prm, double mach, int M, int N, const double *X,
                 const double *V, const double *U, const double *dmach, const double *dX,
                 const double *dV, const double *dU, double *dR, double *R, void *stream);

/* k directions at once: dU, dX, dV are k consecutive complex arrays (stride one full array),
 * dmach k complex values; dR receives k complex arrays.  Any direction pointer may be NULL.
 * With dU only: one launch of the face kernel, one direction per grid z, any k in 1..65535,
 * each output bitwise equal to pea_jvp_odes of that direction.  Otherwise k in {1, 2, 4, 8}.
 * The reference applies one direction per call (apply_odes_wrt_states_fwd, f90:825-865). */
int pea_jvp_odes_multi(const pea_params *prm, double mach, int M, int N, int k,
                       const double *X, const double *V, const double *U, const double *dmach,
                       const double *dX, const double *dV, const double *dU, double *dR,
                       void *stream);

/* dR[j] = dR/dU dU[j] for k complex directions at one state (dU, dR: k consecutive complex
 * cell fields), the same products as pea_jvp_odes_multi along dU alone.  Each workgroup computes
 * its tile's primal part once (states, pressures, and per face the switch decisions and
 * coefficients of the dissipation) and applies it to `per` complex directions (per = 0: chosen
 * from the grid size), so the primal flux is not evaluated again per direction.  Agrees with pea_jvp_odes_multi to round-off.  The
 * reference's products apply assembled complex-step blocks per direction
 * (apply_odes_wrt_states_fwd, f90:825-864). */
int pea_jvp_odes_states_dirs(const pea_params *prm, double mach, int M, int N, int k,
                             const double *X, const double *V, const double *U, const double *dU,
                             double *dR, int per, void *stream);

/* dF = dF/dX dX + dF/dU dU (complex; dX or dU may be NULL).  Replaces
 * apply_forces_wrt_{vertices,states}_fwd (f90:1267-1288, 1333-1348). */
int pea_jvp_forces(int M, int N, const double *X, const double *U, const double *dX,
                   const double *dU, double *dF, void *stream);

/* ------------------------------------------------------------------ hand-written adjoints */

/* Transposed products with the complex seed w (4 planes, complex; NOT conjugated, as the
 * reference's reverse products, f90:907-915): gU = (dR/dU)^T w, gX = (dR/dX)^T w,
 * gV = (dR/dV)^T w, gmach = (dR/dMa)^T w (one complex value).  Any output may be NULL (skipped).
 * Gather form: deterministic, no atomics; dissipation switches are recomputed.
 * Replaces apply_odes_wrt_{states,vertices,mach}_rev (f90:875-918, 1044-1072, 632-650).
 * Uses a workspace of pea_vjp_workspace_bytes(M, N) bytes when gmach != NULL. */
int pea_vjp_odes(const pea_params *prm, double mach, int M, int N, const double *X,
                 const double *V, const double *U, const double *w, double *gU, double *gX,
                 double *gV, double *gmach, void *workspace, void *stream);

/* pea_vjp_odes without a workspace argument and without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The Mach partial sums go to a static 2 MiB device array per
 * GPU (131072 workgroup partials: grids whose pea_vjp_workspace_bytes exceeds 2 MiB are refused
 * with PEA_ERR_DIMS when gmach != NULL).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (on different streams) race; calls in stream order do not.  Results are
 * bitwise those of pea_vjp_odes. */
int pea_vjp_odes_noalloc(const pea_params *prm, double mach, int M, int N, const double *X,
                         const double *V, const double *U, const double *w, double *gU,
                         double *gX, double *gV, double *gmach, void *stream);

/* k seeds at once, states output only: w and gU are k consecutive complex arrays.  One launch of
 * the face adjoint (one seed per grid z) where N >= 18, else one adjoint per seed; each gU
 * bitwise equal to pea_vjp_odes of that seed.  The reference transposes one product per call
 * (apply_odes_wrt_states_rev, f90:875-918). */
int pea_vjp_odes_states_multi(const pea_params *prm, double mach, int M, int N, int k,
                              const double *X, const double *V, const double *U, const double *w,
                              double *gU, void *stream);
long long pea_vjp_workspace_bytes(int M, int N);

/* gU = (dF/dU)^T wF, gX = (dF/dX)^T wF (complex; either may be NULL).  Replaces
 * apply_forces_wrt_{states,vertices}_rev (f90:1359-1375, 1299-1322). */
int pea_vjp_forces(int M, int N, const double *X, const double *U, const double *wF,
                   double *gU, double *gX, void *stream);

/* ------------------------------------------------------------------ second order */

/* Tangent-over-adjoint: with seed w and tangent direction (dmach, dX, dV, dU) (complex, NULL =
 * zero) returns the directional derivative of the adjoint products,
 *   hU = d/de [ (dR/dU)^T w ](Ma + e dMa, X + e dX, V + e dV, U + e dU)   and likewise hX, hV,
 * hmach, bilinear in (w, direction) without conjugation.  The reference's second-order
 * gradients grad_C <dA, dA/dB dB> (core.py:718-1275) are these with w = conj(dA).
 * Workspace as pea_vjp_odes. */
int pea_hvp_odes(const pea_params *prm, double mach, int M, int N, const double *X,
                 const double *V, const double *U, const double *w, const double *dmach,
                 const double *dX, const double *dV, const double *dU, double *hU,
                 double *hX, double *hV, double *hmach, void *workspace, void *stream);

/* Tangent-over-adjoint of the force adjoints: directional derivative of
 * ((dF/dU)^T wF, (dF/dX)^T wF) along (dX, dU) (complex, NULL = zero), bilinear in (wF,
 * direction).  The reference's grad_C <dF, dF/dB dB> (core.py:1105-1275) with wF = conj(dF). */
int pea_hvp_forces(int M, int N, const double *X, const double *U, const double *wF,
                   const double *dX, const double *dU, double *hU, double *hX, void *stream);

/* ------------------------------------------------------------------ complex-step path */

/* Directional complex step along complex directions: Im R(. + i h d)/h, two complex residual
 * evaluations per complex direction (real and imaginary part).  h <= 0 selects 1e-12, the
 * reference's step (f90:52). */
int pea_cs_jvp_odes(const pea_params *prm, double mach, int M, int N, const double *X,
                    const double *V, const double *U, const double *dmach, const double *dX,
                    const double *dV, const double *dU, double h, double *dR, void *stream);

/* k states directions at once (dU: k consecutive complex arrays; dR likewise): the directional
 * complex step, one direction per grid y, as pea_cs_jvp_odes per direction. */
int pea_cs_jvp_odes_multi(const pea_params *prm, double mach, int M, int N, int k,
                          const double *X, const double *V, const double *U, const double *dU,
                          double h, double *dR, void *stream);

/* Complex-step Jacobian assembly at (Ma, X, V, U), the reference's linearize (core.py:278-292):
 *   JU  9 blocks/cell  JU[((s*4 + i)*4 + j)*C + n*M + m]  = dR_i/dU_j of neighbour s
 *       (s = c,b,bb,f,ff,i,ii,o,oo as f90:699-707; absent radial neighbours are written 0)
 *   JX, JV 4 blocks/cell  J[((s*4 + i)*2 + d)*C + n*M + m]  (s = bi,fi,fo,bo)
 *   JMa  J[i*C + n*M + m]  (zero except on the far-field ring)
 * Any output may be NULL.  Replaces compute_odes_wrt_{states,vertices,velocities,mach}. */
int pea_cs_linearize_odes(const pea_params *prm, double mach, int M, int N, const double *X,
                          const double *V, const double *U, double *JU, double *JX,
                          double *JV, double *JMa, void *stream);

/* Block products on the assembled Jacobians (complex vectors).  Replace
 * apply_odes_wrt_states_fwd/rev (f90:825-918) and apply_odes_wrt_vertices_fwd/rev
 * (f90:1005-1072, also used for the velocity blocks).  The reverse forms gather (no atomics). */
int pea_apply_odes_wrt_states_fwd(int M, int N, const double *JU, const double *dU, double *dR,
                                  void *stream);
int pea_apply_odes_wrt_states_rev(int M, int N, const double *JU, const double *w, double *gU,
                                  void *stream);
int pea_apply_odes_wrt_vertices_fwd(int M, int N, const double *JX, const double *dX,
                                    double *dR, void *stream);
int pea_apply_odes_wrt_vertices_rev(int M, int N, const double *JX, const double *w,
                                    double *gX, void *stream);

/* O(nnz) export of JU to 0-based BSR (data[nnzb][4][4] row-major, int32 indices/indptr,
 * ascending columns), nnzb = (9M-6)N, optionally subtracting the complex shift from the
 * diagonal blocks (data then complex: 2 doubles per entry, shift_im != 0 or complex_out).
 * Replaces convert_odes_wrt_states (f90:765-814, an all-pairs scan) + _assemble_csr
 * (core.py:512-531). */
int pea_export_bsr(int M, int N, const double *JU, double shift_re, double shift_im,
                   int complex_out, double *data, int *indices, int *indptr, void *stream);

/* ------------------------------------------------------------------ resolvent kernels */
/* Building blocks of the batched resolvent (dR/dU - s I)^-1 for S shifts at once (the
 * reference factorises one shift at a time with SuperLU, core.py:533-599).  `shifts` holds S
 * complex values on the device; vectors are S consecutive complex cell fields.  With
 * transposed != 0 every operator is that of the adjoint system (dR/dU)^T - conj(s) I
 * (core.py:594-598).  JU are the assembled blocks of pea_cs_linearize_odes. */
const char *pea_resolvent_last_error(void);

/* y_k = (J - s_k I) x_k for k < S, reading the blocks once for all vectors.  Grids of up to
 * 2^24 - 1 cells (PEA_ERR_DIMS beyond). */
int pea_block_multi(int M, int N, int S, const double *JU, const double *shifts,
                    int transposed, const double *x, double *y, void *stream);

/* Radial-line preconditioner: block LU of the block-pentadiagonal matrix of each radial line
 * (blocks ii, i, c - sI, o, oo) for every shift; F needs pea_line_factor_bytes(M, N, S).  F holds
 * all the solve uses: pea_line_solve ignores JU (may be null) and `transposed`, which the
 * factorisation already applied; r, x and F 16-byte aligned, x may alias r. */
long long pea_line_factor_bytes(int M, int N, int S);
int pea_line_factor(int M, int N, int S, const double *JU, const double *shifts,
                    int transposed, double *F, void *stream);
int pea_line_solve(int M, int N, int S, const double *JU, const double *F, int transposed,
                   const double *r, double *x, void *stream);
/* Lanes per radial line in pea_line_solve: 4 (one per block row) or 16 (one per block element);
 * 0 (the default) picks 4 when S*N >= 32768, else 16.  Process-wide; both give the same
 * solution to round-off.  No counterpart in the reference (it factorises with SuperLU). */
int pea_line_solve_lanes(int lanes);

/* Exact direct resolvent by block LU over radial rings (replaces the SuperLU factorisation of
 * core.py:562-564 and 596-598).  Ring k holds every cell at radial index k; its 4N unknowns
 * are ordered (n, v).  The system is block pentadiagonal over rings: a dense ring block and
 * couplings to rings k-2..k+2 that are block diagonal in n.  Dense S x 4N x q complex
 * matrices are row-major (colmajor = 0) or column-major (colmajor = 1) per batch entry.
 * pea_ring_band: D[s] (4N x 4N) = ring block k of the system, the shift on the diagonal,
 *   zero outside the periodic band.
 * pea_ring_couple: Z[s] = beta Z[s] + alpha C Y[s], C = coupling of ring k to ring k + dk
 *   (dk in -2, -1, 1, 2); Y = NULL stands for the identity (q = 4N), which writes C as a
 *   dense matrix.
 * pea_dense_lu_factor / pea_dense_lu_solve: batched partial-pivoting LU of S column-major
 *   n x n complex matrices (rocSOLVER getrf/getrs, strided batched); ipiv S*n and info S
 *   int32 on the device; B (n x nrhs per entry, column-major) is overwritten by the solution.
 * The eliminations between these calls are dense complex GEMMs. */
int pea_ring_band(int M, int N, int S, const double *JU, const double *shifts, int transposed,
                  int k, int colmajor, double *D, void *stream);
int pea_ring_couple(int M, int N, int S, const double *JU, int transposed, int k, int dk,
                    int colmajor, int q, double alpha, const double *Y, double beta, double *Z,
                    void *stream);
int pea_dense_lu_factor(int S, int n, double *A, int *ipiv, int *info, void *stream);
int pea_dense_lu_solve(int S, int n, int nrhs, const double *A, const int *ipiv, double *B,
                       void *stream);

/* Nested-dissection direct resolvent: a multifrontal block LU of dR/dU - s I (or of the
 * adjoint system (dR/dU)^T - conj(s) I), replacing the SuperLU factorisation and solve of
 * core.py:562-564 and 596-598 on grids where the ring LU does not fit (its factors need
 * 3 M (4N)^2 complex values per shift).  The elimination tree and every index map come from the
 * host (pyeulerale_amd/nested.py); these calls do one tree level, all its fronts in one batch.
 * A front is a dense L x L complex matrix, column-major, L = 4 (Pc + Bc): its separator cells
 * (padded to Pc) then its ring cells (padded to Bc), 4 unknowns per cell; cells[k (Pc+Bc) + q]
 * is the q-th cell of front k (-1 = padding).  All arrays are device arrays; int32 maps.
 * pea_nd_assemble: F = the children's update matrices gathered at their ring positions (child
 *   2k + c of front k indexes U_child, each (4 Bc_child)^2 column-major; child_pos
 *   ((2k + c)(Pc+Bc) + q) = position of front cell q in that child's ring or -1), plus identity
 *   on S's padding.  Every entry is written.
 * pea_nd_blocks: adds the original 4x4 blocks of JU (the assembled blocks of
 *   pea_cs_linearize_odes) listed in `entries` (ne x 5 int32: node + node0, front row cell,
 *   front column cell, row cell, block 0..8), minus the shift on block 0's diagonal; adjoint != 0
 *   adds the transposed block at the transposed position (the caller passes conj(s)).
 * pea_nd_factor: LU with partial pivoting of F11 (P x P, P = 4 Pc), F12 <- F11^-1 F12 and
 *   F22 <- F22 - F21 F12 (B = 4 Bc), in place for n fronts (rocSOLVER getrf/getrs, rocBLAS
 *   ZGEMM, strided batched); ipiv n x P, info n.  invert != 0: F11 is replaced by its inverse
 *   (getri), Z = F11^-1 F12 goes to Zws (n x P x B column-major, F12 untouched) and the update
 *   uses it (two GEMMs).
 * pea_nd_gather / pea_nd_scatter: out[k][r][4 q + v] <-> y[r ldy + v M N + cells[k(Pc+Bc) + off
 *   + q]] for q < cnt (gather writes 0 for padding; scatter skips it).
 * pea_nd_accumulate: y[r ldy + target[t]] -= sum_{e = ptr[t]}^{ptr[t+1]-1} W[(src_k[e] nrhs + r)
 *   B + src_j[e]] for every target t and right-hand side r, the sum in the listed order. */
const char *pea_nd_last_error(void);
int pea_nd_assemble(int n, int Pc, int Bc, const int *cells, const int *child,
                    const int *child_pos, int Bc_child, const double *U_child, double *F,
                    void *stream);
int pea_nd_blocks(int ne, const int *entries, int node0, int M, int N, const double *JU,
                  double shift_re, double shift_im, int adjoint, int Pc, int Bc, double *F,
                  void *stream);
int pea_nd_factor(int n, int P, int B, double *F, int *ipiv, int *info, int invert, double *Zws,
                  void *stream);
int pea_nd_gather(int n, int cnt, int off, int Lc, const int *cells, int nrhs, int M, int N,
                  const double *y, long long ldy, double *out, void *stream);
int pea_nd_scatter(int n, int cnt, int off, int Lc, const int *cells, int nrhs, int M, int N,
                   const double *T, double *y, long long ldy, void *stream);
int pea_nd_accumulate(int ntarget, int nrhs, const int *ptr, const int *src_k, const int *src_j,
                      const int *target, int B, const double *W, double *y, long long ldy,
                      void *stream);

/* GMRES: Givens update of column j of S Hessenberg matrices on the device (H (S, R+1, R), cs/sn
 * (S, R), g (S, R+1) complex row-major; est[s] = |g[s][j+1]|), so that Arnoldi steps need no
 * host round trip and can be captured in a HIP graph. */
int pea_gmres_rotate(int S, int R, int j, double *H, double *cs, double *sn, double *g,
                     double *est, void *stream);

/* Gram-Schmidt projections: h[s*bsh + i] = sum_k conj(V[s*bsv + i*ldv + k]) w[s*bsw + k] for
 * i < q, k < n, s < S (complex; strides in complex elements; V and w 16-byte aligned).  K is
 * split into nb chunks summed in a fixed order; P is a workspace of S*nb*q complex values. */
int pea_multidot(int S, int n, int q, const double *V, long long ldv, long long bsv,
                 const double *w, long long bsw, int nb, double *P, double *h, long long bsh,
                 void *stream);

/* Gram-Schmidt update: w[s*bsw + k] -= sum_{i<q} V[s*bsv + i*ldv + k] h[s*bsh + i] (complex,
 * strides in complex elements, V and w 16-byte aligned), each row of V read once.  mode 0: update
 * only; mode 1: also out[s*bso + i] = sum_k conj(V[s*bsv + i*ldv + k]) w_new[s*bsw + k] for i < q
 * (the next projections, same chunked order as pea_multidot); mode 2: also out[s*bso] =
 * |w_new[s]|^2 (imaginary part 0).  P: workspace of S*nb*q complex values (modes 1 and 2).
 * The reference solves each shift directly (spsolve, core.py:562 and 596) and has no
 * counterpart; this is part of the batched GMRES that stands in for those solves. */
int pea_gs_update(int S, int n, int q, const double *V, long long ldv, long long bsv,
                  const double *h, long long bsh, double *w, long long bsw, int mode, int nb,
                  double *P, double *out, long long bso, void *stream);

/* One Gram-Schmidt step of batched GMRES with selective reorthogonalisation (Daniel, Gragg,
 * Kaufman, Stewart 1976): h1 = V^H w and |w|^2 in one pass over the q basis rows; w -= V h1;
 * only for shifts where |h1|^2 > |w|^2/2 the projection is repeated (h2 = V^H w, w -= V h2).
 * Outputs h[s*bsh + i] = h1 + h2 (complex, i < q) and nrm2[s] = |w_new|^2 (double); flags[s] = 1
 * where the projection was repeated.  Workspace: P of S*nb*(q+1) complex values, hw of
 * 2*S*(q+1) complex values.  1 <= q <= 32; V and w 16-byte aligned.  Strides as pea_gs_update.
 * Part of the batched GMRES; the reference solves each shift directly (core.py:562, 596). */
int pea_gs_orthogonalize(int S, int n, int q, const double *V, long long ldv, long long bsv,
                         double *w, long long bsw, int nb, double *P, double *hw, int *flags,
                         double *h, long long bsh, double *nrm2, void *stream);

/* Block-Jacobi preconditioner: inverses of the 4x4 diagonal blocks (J_cc - s I);
 * F holds S*16*M*N complex values. */
int pea_jacobi_factor(int M, int N, int S, const double *JU, const double *shifts,
                      int transposed, double *F, void *stream);
int pea_jacobi_solve(int M, int N, int S, const double *F, const double *r, double *x,
                     void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PEA_H */
