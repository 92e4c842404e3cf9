// Kernels of the resolvent (dR/dU - s I)^-1 for many shifts s at once.
//
// * k_block_multi: products of the assembled dR/dU blocks (or their transposes) with S complex
//   vectors (operator of batched GMRES).  A workgroup stages the 9 blocks of 64 cells in LDS once
//   and applies them to all vectors, so the 1152 B/cell of blocks are read once per product
//   batch instead of once per vector.  The shift is applied in the same pass: y = J x - s x.
// * k_line_blocks / k_line_factor / k_line_solve_rows / k_line_solve_elems: radial-line
//   preconditioner.  For every angular index n and shift s the block-pentadiagonal matrix of the
//   radial line (blocks ii, i, c - sI, o, oo) is factorised by block LU (partial pivoting inside
//   each 4x4 diagonal block) and solved.  k_line_blocks transposes the line blocks lines-fastest;
//   the factorisation runs one thread per (line, shift), m sequential; the solve four lanes per
//   (line, shift) (one per block row) or sixteen (one per block element).  The transposed variant
//   factorises the line of (dR/dU)^T - conj(s) I for the adjoint solve.
// * k_jacobi_factor / k_jacobi_solve: block-Jacobi preconditioner (4x4 diagonal blocks).
// * k_ring_band / k_ring_couple: pieces of the exact block-banded LU over radial rings.  A ring
//   is every cell at one radial index m; ordered (n, v) its unknowns form a dense 4N x 4N
//   block.  The matrix is block pentadiagonal in m: the ring's own block (blocks c, b, bb, f,
//   ff, periodic in n, minus the shift) and four couplings to rings m-2..m+2 that are
//   block diagonal in n (blocks ii, i, o, oo).  k_ring_band writes the dense ring block for all
//   shifts; k_ring_couple forms Z = beta Z + alpha C Y with a coupling C, or C itself as a dense
//   matrix when Y is null.  The dense eliminations between them are library GEMM/LU calls.
// * k_multidot_partial / k_dgks_update / k_gs_update / k_multidot_final / k_dgks_finish:
//   Gram-Schmidt of batched GMRES.  A projection pass reads the basis once; an update pass
//   subtracts the projections and, reading the basis rows only once more, forms the next
//   projections or the norm of the result; a second update runs only for the shifts whose
//   first projection removed more than half of the vector's norm squared.
// * k_gmres_rotate: Givens updates of the Hessenberg matrices, one thread per shift.
#include <hip/hip_runtime.h>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdint>
#include <cstdio>
#include <mutex>

#include "../../include/pea.h"

namespace {

struct Z {
    double r, i;
};
__device__ __forceinline__ Z zadd(Z a, Z b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ Z zsub(Z a, Z b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ Z zmul(Z a, Z b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ Z zfma(Z a, Z b, Z c) {   // a*b + c
    return {fma(a.r, b.r, fma(-a.i, b.i, c.r)), fma(a.r, b.i, fma(a.i, b.r, c.i))};
}
// a*b + c for real a: two FMAs (zfma with a zero imaginary part would keep the 0*b terms)
__device__ __forceinline__ Z zfma_re(double a, Z b, Z c) {
    return {fma(a, b.r, c.r), fma(a, b.i, c.i)};
}
__device__ __forceinline__ Z zinv(Z a) {
    const double d = a.r * a.r + a.i * a.i;
    return {a.r / d, -a.i / d};
}
__device__ __forceinline__ double zabs2(Z a) { return a.r * a.r + a.i * a.i; }

// 4x4 complex block in registers
struct B4 {
    Z a[4][4];
};

__device__ __forceinline__ B4 bmul(const B4& x, const B4& y) {
    B4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            Z s = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 4; ++k) s = zfma(x.a[i][k], y.a[k][j], s);
            r.a[i][j] = s;
        }
    return r;
}
__device__ __forceinline__ B4 bsub(const B4& x, const B4& y) {
    B4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) r.a[i][j] = zsub(x.a[i][j], y.a[i][j]);
    return r;
}
// inverse by Gauss-Jordan with partial pivoting
__device__ B4 binv(const B4& x) {
    Z m[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            m[i][j] = x.a[i][j];
            m[i][j + 4] = Z{i == j ? 1.0 : 0.0, 0.0};
        }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int p = c;
        double best = zabs2(m[c][c]);
#pragma unroll
        for (int r = c + 1; r < 4; ++r)
            if (zabs2(m[r][c]) > best) {
                best = zabs2(m[r][c]);
                p = r;
            }
        // swap rows c and p by selects over the candidate rows: a row index known only at run
        // time would put m in scratch memory
#pragma unroll
        for (int r = c + 1; r < 4; ++r) {
            const bool sw = p == r;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const Z a = m[c][j], b = m[r][j];
                m[c][j] = sw ? b : a;
                m[r][j] = sw ? a : b;
            }
        }
        const Z piv = zinv(m[c][c]);
#pragma unroll
        for (int j = 0; j < 8; ++j) m[c][j] = zmul(m[c][j], piv);
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                const Z f = m[r][c];
#pragma unroll
                for (int j = 0; j < 8; ++j) m[r][j] = zsub(m[r][j], zmul(f, m[c][j]));
            }
    }
    B4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) r.a[i][j] = m[i][j + 4];
    return r;
}

// block s of cell (m, n) of the assembled Jacobian (real), optionally transposed, minus
// `shift` on the diagonal
__device__ __forceinline__ B4 load_jblock(const double* JU, int C, int cell, int s, bool tr,
                                          Z shift) {
    B4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ii = tr ? j : i, jj = tr ? i : j;
            r.a[i][j] = Z{JU[(long)((s * 4 + ii) * 4 + jj) * C + cell], 0.0};
            if (i == j) r.a[i][j] = zsub(r.a[i][j], shift);
        }
    return r;
}

// factor storage: lines in groups of kLineGroup (the lines one workgroup of the solve holds);
// per (shift, group) the steps m one after another, each with its 4 slots x kLineGroup lines x
// 16 elements:  F[((((sh*NG + g)*M + m)*4 + slot)*kLineGroup + n % kLineGroup)*16 + 4i + j] with
// g = n / kLineGroup, slots 0 = L2, 1 = L1, 2 = Dinv, 3 = U1.  One block of a workgroup's lines
// is a contiguous 4 KB (1 KB per wave), and the workgroup streams through one contiguous region
// as it walks its lines.
constexpr int kLineGroup = 16;
__host__ __device__ __forceinline__ long line_groups(int N) {
    return (N + kLineGroup - 1) / kLineGroup;
}
__device__ __forceinline__ long fidx(int M, int N, int sh, int m, int slot, int ij, int n) {
    const long g = n / kLineGroup, nl = n % kLineGroup;
    return (((((long)sh * line_groups(N) + g) * M + m) * 4 + slot) * kLineGroup + nl) * 16 + ij;
}
// after the S shifts' complex factors (in doubles): the real block U2 = (m, m+2) of the system,
// the same for every shift, element (i, j) of line n at U2[((g*M + m)*kLineGroup + n %
// kLineGroup)*16 + 4i + j]
__device__ __forceinline__ long u2idx(int M, int N, int S, int m, int ij, int n) {
    const long base = 2L * S * line_groups(N) * M * 4 * 16 * kLineGroup;
    const long g = n / kLineGroup, nl = n % kLineGroup;
    return base + (((long)g * M + m) * kLineGroup + nl) * 16 + ij;
}
// after U2 (in doubles): the system's line blocks (m, m + off), off = -2..2, without the shift,
// lines fastest: SB[((off + 2)*16 + 4i + j)*M*N + m*N + n], written by k_line_blocks for the
// factorisation
__host__ __device__ __forceinline__ long sb_offset(int M, int N, int S) {
    return 2L * S * line_groups(N) * M * 4 * 16 * kLineGroup + line_groups(N) * M * 16L * kLineGroup;
}
__device__ __forceinline__ B4 sb_block(const double* SB, long C, int N, int n, int m, int off,
                                       Z shift) {
    B4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            r.a[i][j] = Z{SB[(long)((off + 2) * 16 + 4 * i + j) * C + (long)m * N + n], 0.0};
            if (off == 0 && i == j) r.a[i][j] = zsub(r.a[i][j], shift);
        }
    return r;
}

// The system's radial-line blocks in the line-major layout of sb_block, zero where m + off is
// off the line: a tiled transpose of the blocks' (n, m) planes (m fastest) into (m, n) planes,
// both sides coalesced.  The transposed system's block (m, m + off) is block (m + off, m) of J
// transposed.  Grid (ceil(M/kTrM), ceil(N/kTrN), 80 planes).
constexpr int kTrM = 16, kTrN = 64;

__global__ void __launch_bounds__(256)
k_line_blocks(int M, int N, const double* __restrict__ JU, int tr, double* __restrict__ SB) {
    __shared__ double t[kTrN][kTrM + 1];
    const int plane = blockIdx.z;
    const int off = plane / 16 - 2, ij = plane % 16, i = ij / 4, j = ij % 4;
    const int m0 = blockIdx.x * kTrM, n0 = blockIdx.y * kTrN;
    const long C = (long)M * N;
    int slot, a = i, b = j, dm = 0;
    if (!tr) {
        slot = off == -2 ? 6 : off == -1 ? 5 : off == 0 ? 0 : off == 1 ? 7 : 8;
    } else {
        slot = off == -2 ? 8 : off == -1 ? 7 : off == 0 ? 0 : off == 1 ? 5 : 6;
        a = j;
        b = i;
        dm = off;
    }
    for (int k = threadIdx.x; k < kTrM * kTrN; k += 256) {
        const int ml = k % kTrM, nl = k / kTrM, m = m0 + ml, n = n0 + nl;
        double v = 0.0;
        if (m < M && n < N && m + off >= 0 && m + off < M)
            v = JU[((long)(slot * 4 + a) * 4 + b) * C + (long)n * M + m + dm];
        t[nl][ml] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kTrM * kTrN; k += 256) {
        const int nl = k % kTrN, ml = k / kTrN, m = m0 + ml, n = n0 + nl;
        if (m < M && n < N) SB[(long)plane * C + (long)m * N + n] = t[nl][ml];
    }
}
__device__ __forceinline__ void store_b(double* F, int M, int N, int sh, int m, int slot, int n,
                                        const B4& b) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long e = fidx(M, N, sh, m, slot, 4 * i + j, n);
            F[2 * e] = b.a[i][j].r;
            F[2 * e + 1] = b.a[i][j].i;
        }
}
__device__ __forceinline__ B4 load_b(const double* F, int M, int N, int sh, int m, int slot,
                                     int n) {
    B4 b;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long e = fidx(M, N, sh, m, slot, 4 * i + j, n);
            b.a[i][j] = Z{F[2 * e], F[2 * e + 1]};
        }
    return b;
}

// The blocks D^-1 and U1 of the two previous steps stay in LDS, in a ring of two per kind,
// [kind][element][lane] so that each element's 64 lanes are contiguous: reading them back from
// F would wait on this thread's own recent stores, 256 B apart across lanes.
constexpr int kFactorThreads = 64;

__global__ void __launch_bounds__(kFactorThreads)
k_line_factor(int M, int N, int S, const double* __restrict__ sh, int tr, double* __restrict__ F) {
    __shared__ Z ring[4][16][kFactorThreads];   // kinds: Dinv even/odd step, U1 even/odd step
    const int lane = threadIdx.x;
    const int n = blockIdx.x * kFactorThreads + lane;
    const int s = blockIdx.y;
    if (n >= N || s >= S) return;   // per thread: each lane touches only its own LDS column
    const long C = (long)M * N;
    const double* SB = F + sb_offset(M, N, S);
    auto rput = [&](int kind, const B4& b) {
#pragma unroll
        for (int ij = 0; ij < 16; ++ij) ring[kind][ij][lane] = b.a[ij / 4][ij % 4];
    };
    auto rget = [&](int kind) {
        B4 b;
#pragma unroll
        for (int ij = 0; ij < 16; ++ij) b.a[ij / 4][ij % 4] = ring[kind][ij][lane];
        return b;
    };
    // adjoint: (J - s I)^H = J^T - conj(s) I
    const Z shift = tr ? Z{sh[2 * s], -sh[2 * s + 1]} : Z{sh[2 * s], sh[2 * s + 1]};
    const Z zero = {0.0, 0.0};
    for (int m = 0; m < M; ++m) {
        const int cur = m & 1, prev = cur ^ 1;   // ring slot of steps m (and m - 2), m - 1
        B4 D = sb_block(SB, C, N, n, m, 0, shift);
        B4 L2, L1;
        if (m >= 2) {
            L2 = bmul(sb_block(SB, C, N, n, m, -2, zero), rget(cur));
            // D -= L2 * U2(m-2),  U2(m-2) = block (m-2, m)
            D = bsub(D, bmul(L2, sb_block(SB, C, N, n, m - 2, 2, zero)));
            store_b(F, M, N, s, m, 0, n, L2);
        }
        if (m >= 1) {
            B4 B1 = sb_block(SB, C, N, n, m, -1, zero);
            if (m >= 2) B1 = bsub(B1, bmul(L2, rget(2 + cur)));
            L1 = bmul(B1, rget(prev));
            D = bsub(D, bmul(L1, rget(2 + prev)));
            store_b(F, M, N, s, m, 1, n, L1);
        }
        if (m + 1 < M) {
            B4 U1 = sb_block(SB, C, N, n, m, 1, zero);
            if (m >= 1) U1 = bsub(U1, bmul(L1, sb_block(SB, C, N, n, m - 1, 2, zero)));
            store_b(F, M, N, s, m, 3, n, U1);
            rput(2 + cur, U1);
        }
        if (s == 0 && m + 2 < M) {
            const B4 U2 = sb_block(SB, C, N, n, m, 2, zero);
#pragma unroll
            for (int ij = 0; ij < 16; ++ij) F[u2idx(M, N, S, m, ij, n)] = U2.a[ij / 4][ij % 4].r;
        }
        const B4 Di = binv(D);
        store_b(F, M, N, s, m, 2, n, Di);
        rput(cur, Di);
    }
}

// Orders the LDS accesses of a wave's lanes: what one lane wrote, another lane reads after this.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// A double moved between lanes of a quad by DPP (quad_perm control CTRL).
template <int CTRL>
__device__ __forceinline__ double quad_dpp(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(
        (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}
// Sum over the quad, the same bits in all four lanes: (x0 + x1) + (x2 + x3) in every lane,
// addition being commutative.  quad_perm [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E.
__device__ __forceinline__ double quad_sum(double x) {
    x += quad_dpp<0xB1>(x);
    return x + quad_dpp<0x4E>(x);
}
__device__ __forceinline__ Z quad_sum(Z x) { return Z{quad_sum(x.r), quad_sum(x.i)}; }

// x = (line LU)^-1 r for every line and shift.  r and x: S consecutive complex cell fields
// (4 planes of M x N each); x may alias r.  Sixteen lanes per (line, shift), lane 4i + j holding
// element (i, j) of the factor blocks and component j of the previous steps' vectors; a wave
// holds 4 lines and a workgroup the kLineGroup = 16 lines of one factor group.  A row's sum over
// j is a quad butterfly (DPP), which leaves it in the four lanes of row i; component j then
// moves to the lanes of column j by one transpose within the line's 16 lanes (ds_bpermute).
// The workgroup's waves stay in step (a barrier per block of steps), so together they read each
// step's factors as one contiguous 4 KB per block.
// * The vectors run along m, the fastest index of the cell fields: each wave moves them through
//   LDS in blocks of kLineBlock steps, one whole 128 B line per (line, row).
// * The factor loads of a step do not depend on the recurrence: they are issued kLineAhead
//   steps early into a ring of register sets (the block loop is unrolled, so every set is
//   addressed at compile time).
constexpr int kLineThreads = 256;
constexpr int kLineAhead = 4;
int g_line_lanes = 0;   // lanes per line of the solve: 0 automatic, 4 or 16 (pea_line_solve_lanes)
constexpr int kLineBlock = 8;        // 8 steps x 16 B: one 128 B line per (line, row)
constexpr int kLinePad = kLineBlock + 1;

__global__ void __launch_bounds__(kLineThreads)
k_line_solve_elems(int M, int N, int S, const double* __restrict__ F, const double* r, double* x) {
    static_assert(kLineThreads / 16 == kLineGroup, "a workgroup holds one line group");
    static_assert(kLineBlock % kLineAhead == 0, "ring slots repeat within a block");
    __shared__ double2 sv[kLineThreads / 64][16][kLinePad];   // per wave: 4 lines x 4 rows
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n0 = blockIdx.x * kLineGroup + wv * 4;   // first line of the wave
    const int nl = lane >> 4, ij = lane & 15, i = ij >> 2, j = ij & 3;
    const int n = n0 + nl, s = blockIdx.y;
    // every wave runs to the end (barriers): lines past N, inside the last, padded group,
    // compute values that are never stored
    if (s >= S) return;
    const int tsrc = (lane & ~15) | (j << 2) | i;   // lane of row j, column i: the transpose
    const long C = (long)M * N;
    const double2* F2 = reinterpret_cast<const double2*>(F);
    const double2* r2 = reinterpret_cast<const double2*>(r);
    double2* x2 = reinterpret_cast<double2*>(x);
    double2(*buf)[kLinePad] = sv[wv];
    const Z zero = {0.0, 0.0};
    const int prow = nl * 4 + i;   // (line, row) pair of this lane in the staging block
    // block transfer: element e = lane + 64k of the wave's 16 x kLineBlock block is step e % 8
    // of pair e / 8, the pair being (line, row) = (e / 32, e / 8 % 4)
    auto vload = [&](const double2* a, int m0) {
#pragma unroll
        for (int k = 0; k < 16 * kLineBlock / 64; ++k) {
            const int e = lane + 64 * k, p = e / kLineBlock, q = e % kLineBlock;
            const int pn = n0 + p / 4, m = m0 + q;
            if (pn < N && m < M) buf[p][q] = a[(long)s * 4 * C + (long)(p & 3) * C + (long)pn * M + m];
        }
    };
    auto vstore = [&](double2* a, int m0) {
#pragma unroll
        for (int k = 0; k < 16 * kLineBlock / 64; ++k) {
            const int e = lane + 64 * k, p = e / kLineBlock, q = e % kLineBlock;
            const int pn = n0 + p / 4, m = m0 + q;
            if (pn < N && m < M) a[(long)s * 4 * C + (long)(p & 3) * C + (long)pn * M + m] = buf[p][q];
        }
    };
    auto fel = [&](int m, int slot) {   // element (i, j) of factor block (m, slot)
        const double2 v = F2[fidx(M, N, s, m, slot, ij, n)];
        return Z{v.x, v.y};
    };
    auto tr = [&](Z a) {   // component j, from the lanes of row j
        return Z{__shfl(a.r, tsrc), __shfl(a.i, tsrc)};
    };
    // forward: y_m = r_m - L2 y_{m-2} - L1 y_{m-1}  (y stored in x)
    {
        Z y1 = zero, y2 = zero;   // component j of y_{m-1}, y_{m-2}
        Z l2[kLineAhead], l1[kLineAhead];
        auto load = [&](int m, int k) {
            if (m < M) {
                if (m >= 2) l2[k] = fel(m, 0);
                if (m >= 1) l1[k] = fel(m, 1);
            }
        };
#pragma unroll
        for (int k = 0; k < kLineAhead; ++k) load(k, k);
        for (int m0 = 0; m0 < M; m0 += kLineBlock) {
            vload(r2, m0);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kLineBlock; ++k) {
                const int m = m0 + k, slot = k % kLineAhead;
                if (m < M) {
                    Z acc = zero;
                    if (j == 0) {
                        const double2 rv = buf[prow][k];
                        acc = Z{rv.x, rv.y};
                    }
                    if (m >= 2) acc = zsub(acc, zmul(l2[slot], y2));
                    if (m >= 1) acc = zsub(acc, zmul(l1[slot], y1));
                    const Z y = quad_sum(acc);            // row i of y_m, in all four lanes
                    if (j == 0) buf[prow][k] = make_double2(y.r, y.i);
                    y2 = y1;
                    y1 = tr(y);
                }
                load(m + kLineAhead, slot);
            }
            __syncthreads();
            vstore(x2, m0);
            __syncthreads();
        }
    }
    // backward: x_m = Dinv (y_m - U1 x_{m+1} - U2 x_{m+2}),  U2 = block (m, m+2) of the system
    {
        Z x1 = zero, xx2 = zero;   // component j of x_{m+1}, x_{m+2}
        Z u1[kLineAhead], di[kLineAhead];
        double u2[kLineAhead];
        auto load = [&](int m, int k) {
            if (m >= 0 && m < M) {
                if (m + 1 < M) u1[k] = fel(m, 3);
                if (m + 2 < M) u2[k] = F[u2idx(M, N, S, m, ij, n)];
                di[k] = fel(m, 2);
            }
        };
        const int mt = (M - 1) / kLineBlock * kLineBlock;   // first step of the top block
        // step m = m0 + kLineBlock - 1 - k of a block uses ring slot m % kLineAhead
#pragma unroll
        for (int k = 0; k < kLineAhead; ++k)
            load(mt + kLineBlock - 1 - k, (kLineBlock - 1 - k) % kLineAhead);
        for (int m0 = mt; m0 >= 0; m0 -= kLineBlock) {
            vload(x2, m0);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kLineBlock; ++k) {
                const int q = kLineBlock - 1 - k, m = m0 + q, slot = q % kLineAhead;
                if (m < M) {
                    Z acc = zero;
                    if (j == 0) {
                        const double2 yv = buf[prow][q];
                        acc = Z{yv.x, yv.y};
                    }
                    if (m + 1 < M) acc = zsub(acc, zmul(u1[slot], x1));
                    if (m + 2 < M) acc = zsub(acc, zmul(Z{u2[slot], 0.0}, xx2));
                    const Z a = tr(quad_sum(acc));          // component j of y_m - U1 x - U2 x
                    const Z xm = quad_sum(zmul(di[slot], a));   // row i of x_m
                    if (j == 0) buf[prow][q] = make_double2(xm.r, xm.i);
                    xx2 = x1;
                    x1 = tr(xm);
                }
                load(m - kLineAhead, slot);
            }
            __syncthreads();
            vstore(x2, m0);
            __syncthreads();
        }
    }
}

// Lane J of each quad of lanes, broadcast to the quad (DPP quad_perm [J, J, J, J]).
template <int J>
__device__ __forceinline__ double quad_bcast(double x) {
    constexpr int ctrl = J * 0x55;
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), ctrl, 0xF, 0xF, false);
    return __longlong_as_double(
        (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}
// the 4-vector whose component i lane i of the quad holds
__device__ __forceinline__ void quad_gather(Z mine, Z* y) {
    y[0] = Z{quad_bcast<0>(mine.r), quad_bcast<0>(mine.i)};
    y[1] = Z{quad_bcast<1>(mine.r), quad_bcast<1>(mine.i)};
    y[2] = Z{quad_bcast<2>(mine.r), quad_bcast<2>(mine.i)};
    y[3] = Z{quad_bcast<3>(mine.r), quad_bcast<3>(mine.i)};
}

// The same solve with four lanes per (line, shift), for many lines and shifts: lane i holds row
// i of the 4-vectors and of the factor blocks (elements 4i..4i+3, contiguous), and the whole
// vectors of the two previous steps are shared in the quad by DPP broadcasts.  A wave holds the
// kLineGroup lines of one factor group and reads a block of them as one contiguous 4 KB.
// Vectors move through LDS as in k_line_solve_elems; factor loads run kRowsAhead steps early
// (two: its row operands need 4x the registers of an element's).
constexpr int kRowsAhead = 2;

struct LineFwd {   // operands of one forward step
    Z l2[4], l1[4];
};
struct LineBwd {   // operands of one backward step
    Z u1[4], di[4];
    double u2[4];
};

__global__ void __launch_bounds__(kLineThreads)
k_line_solve_rows(int M, int N, int S, const double* __restrict__ F, const double* r, double* x) {
    static_assert(kLineThreads % 64 == 0 && 64 / 4 == kLineGroup, "a wave holds one line group");
    static_assert(kLineBlock % kRowsAhead == 0, "ring slots repeat within a block");
    __shared__ double2 sv[kLineThreads / 64][64][kLinePad];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n0 = (blockIdx.x * kLineThreads + wv * 64) / 4;   // first line of the wave
    const int n = n0 + lane / 4, i = lane & 3, s = blockIdx.y;
    if (n0 >= N || s >= S) return;   // whole waves; a wave's lines past N compute unused values
    const long C = (long)M * N;
    const double2* F2 = reinterpret_cast<const double2*>(F);
    const double2* r2 = reinterpret_cast<const double2*>(r);
    double2* x2 = reinterpret_cast<double2*>(x);
    double2(*buf)[kLinePad] = sv[wv];
    const Z zero = {0.0, 0.0};
    // block transfer: element e = lane + 64k of the wave's 64 x kLineBlock block is step e % 8
    // of pair e / 8, the pair being lane (line, row) = (e / 32, e / 8 % 4) of the compute phase
    auto vload = [&](const double2* a, int m0) {
#pragma unroll
        for (int k = 0; k < kLineBlock; ++k) {
            const int e = lane + 64 * k, p = e / kLineBlock, q = e % kLineBlock;
            const int pn = n0 + p / 4, m = m0 + q;
            if (pn < N && m < M) buf[p][q] = a[(long)s * 4 * C + (long)(p & 3) * C + (long)pn * M + m];
        }
    };
    auto vstore = [&](double2* a, int m0) {
#pragma unroll
        for (int k = 0; k < kLineBlock; ++k) {
            const int e = lane + 64 * k, p = e / kLineBlock, q = e % kLineBlock;
            const int pn = n0 + p / 4, m = m0 + q;
            if (pn < N && m < M) a[(long)s * 4 * C + (long)(p & 3) * C + (long)pn * M + m] = buf[p][q];
        }
    };
    auto frow = [&](int m, int slot, Z* a) {   // row i of factor block (m, slot)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double2 v = F2[fidx(M, N, s, m, slot, 4 * i + j, n)];
            a[j] = Z{v.x, v.y};
        }
    };
    // forward: y_m = r_m - L2 y_{m-2} - L1 y_{m-1}  (y stored in x)
    {
        Z y1[4] = {zero, zero, zero, zero}, y2[4] = {zero, zero, zero, zero};
        auto load = [&](int m, LineFwd& f) {
            if (m < M) {
                if (m >= 2) frow(m, 0, f.l2);
                if (m >= 1) frow(m, 1, f.l1);
            }
        };
        LineFwd ring[kRowsAhead];
#pragma unroll
        for (int k = 0; k < kRowsAhead; ++k) load(k, ring[k]);
        for (int m0 = 0; m0 < M; m0 += kLineBlock) {
            vload(r2, m0);
            wave_sync();
#pragma unroll
            for (int k = 0; k < kLineBlock; ++k) {
                const int m = m0 + k;
                LineFwd& f = ring[k % kRowsAhead];
                if (m < M) {
                    const double2 rv = buf[lane][k];
                    Z acc = {rv.x, rv.y};
                    if (m >= 2)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc = zsub(acc, zmul(f.l2[j], y2[j]));
                    if (m >= 1)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc = zsub(acc, zmul(f.l1[j], y1[j]));
                    buf[lane][k] = make_double2(acc.r, acc.i);
                    Z y[4];
                    quad_gather(acc, y);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        y2[j] = y1[j];
                        y1[j] = y[j];
                    }
                }
                load(m + kRowsAhead, f);
            }
            wave_sync();
            vstore(x2, m0);
            wave_sync();
        }
    }
    // backward: x_m = Dinv (y_m - U1 x_{m+1} - U2 x_{m+2}),  U2 = block (m, m+2) of the system
    {
        Z x1[4] = {zero, zero, zero, zero}, xx2[4] = {zero, zero, zero, zero};
        auto load = [&](int m, LineBwd& b) {
            if (m >= 0 && m < M) {
                if (m + 1 < M) frow(m, 3, b.u1);
                if (m + 2 < M)
#pragma unroll
                    for (int j = 0; j < 4; ++j) b.u2[j] = F[u2idx(M, N, S, m, 4 * i + j, n)];
                frow(m, 2, b.di);
            }
        };
        const int mt = (M - 1) / kLineBlock * kLineBlock;   // first step of the top block
        // step m = m0 + kLineBlock - 1 - k of a block uses ring slot m % kRowsAhead
        LineBwd ring[kRowsAhead];
#pragma unroll
        for (int k = 0; k < kRowsAhead; ++k)
            load(mt + kLineBlock - 1 - k, ring[(kLineBlock - 1 - k) % kRowsAhead]);
        for (int m0 = mt; m0 >= 0; m0 -= kLineBlock) {
            vload(x2, m0);
            wave_sync();
#pragma unroll
            for (int k = 0; k < kLineBlock; ++k) {
                const int q = kLineBlock - 1 - k, m = m0 + q;
                LineBwd& b = ring[q % kRowsAhead];
                if (m < M) {
                    const double2 yv = buf[lane][q];
                    Z acc = {yv.x, yv.y};
                    if (m + 1 < M)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc = zsub(acc, zmul(b.u1[j], x1[j]));
                    if (m + 2 < M)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc = zsub(acc, zmul(Z{b.u2[j], 0.0}, xx2[j]));
                    Z yq[4];
                    quad_gather(acc, yq);
                    Z xm = zero;
#pragma unroll
                    for (int j = 0; j < 4; ++j) xm = zfma(b.di[j], yq[j], xm);
                    buf[lane][q] = make_double2(xm.r, xm.i);
                    Z xq[4];
                    quad_gather(xm, xq);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        xx2[j] = x1[j];
                        x1[j] = xq[j];
                    }
                }
                load(m - kRowsAhead, b);
            }
            wave_sync();
            vstore(x2, m0);
            wave_sync();
        }
    }
}

// block-Jacobi: F[(s*16 + ij)*C + cell] complex inverse of (J_cc - s I) (or its transpose)
__global__ void __launch_bounds__(256)
k_jacobi_factor(int C, int S, const double* __restrict__ JU, const double* __restrict__ sh, int tr,
                double* __restrict__ F) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (c >= C || s >= S) return;
    const Z shift = tr ? Z{sh[2 * s], -sh[2 * s + 1]} : Z{sh[2 * s], sh[2 * s + 1]};
    const B4 Di = binv(load_jblock(JU, C, c, 0, tr, shift));
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long e = ((long)s * 16 + 4 * i + j) * C + c;
            F[2 * e] = Di.a[i][j].r;
            F[2 * e + 1] = Di.a[i][j].i;
        }
}

__global__ void __launch_bounds__(256)
k_jacobi_solve(int C, int S, const double* __restrict__ F, const double* r, double* x) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (c >= C || s >= S) return;
    Z y[4], o[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const long e = (long)s * 4 * C + (long)v * C + c;
        y[v] = Z{r[2 * e], r[2 * e + 1]};
        o[v] = Z{0.0, 0.0};
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long e = ((long)s * 16 + 4 * i + j) * C + c;
            o[i] = zfma(Z{F[2 * e], F[2 * e + 1]}, y[j], o[i]);
        }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const long e = (long)s * 4 * C + (long)v * C + c;
        x[2 * e] = o[v].r;
        x[2 * e + 1] = o[v].i;
    }
}

// y_k = J x_k - s_k x_k (or J^T x_k - conj(s_k) x_k) for k < S.  A workgroup holds 64 cells'
// nine blocks in LDS; wave w handles vectors w, w + 4, ..., one cell per lane.  Blocks of
// neighbours outside the grid (radially) are staged as zeros and read the cell's own vector
// instead, so both loops are straight-line code whose loads all issue before the first wait
// (fma(0, x, a) = a, so the sums are unchanged).
constexpr int kMvCells = 64, kMvSlots = 4, kMvPasses = 16 / kMvSlots;
// slot s of a cell's nine blocks: the cell itself, angular neighbours n-1, n-2, n+1, n+2
// (periodic), radial neighbours m-1, m-2, m+1, m+2
__device__ constexpr int kSlotDm[9] = {0, 0, 0, 0, 0, -1, -2, 1, 2};
__device__ constexpr int kSlotDn[9] = {0, -1, -2, 1, 2, 0, 0, 0, 0};

// two workgroups per CU (LDS 2 x 72 KB): the register allocation must fit two waves per SIMD
__global__ void __launch_bounds__(kMvCells * kMvSlots) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_block_multi(int M, int N, int S, const double* __restrict__ JU, const double* __restrict__ sh,
              int tr, const double* __restrict__ x, double* __restrict__ y) {
    static_assert(kMvCells == 64 && kMvSlots * kMvPasses == 16, "one wave per vector slot");
    __shared__ double sJ[9][16][kMvCells];
    const int C = M * N;
    const int lc = threadIdx.x % kMvCells, slot = threadIdx.x / kMvCells;
    const int c0 = blockIdx.x * kMvCells;
    auto wr = [&](int q) { return q < 0 ? q + N : (q >= N ? q - N : q); };
    {
        // forward: blocks of this cell; transposed: block s of the neighbour k = c - offset(s),
        // element (i, j) of its J^T block = (j, i) of the J block
        const int c = c0 + lc;
        const int cs = c < C ? c : 0;
        const int n = cs / M, m = cs - n * M;
        double v[9][kMvPasses];
#pragma unroll
        for (int s = 0; s < 9; ++s) {
            const int mk = tr ? m - kSlotDm[s] : m + kSlotDm[s];
            const bool ok = c < C && mk >= 0 && mk < M;
            const int k = tr ? (ok ? wr(n - kSlotDn[s]) * M + mk : cs) : cs;
#pragma unroll
            for (int h = 0; h < kMvPasses; ++h) {
                const int e = h * kMvSlots + slot;
                const int ee = tr ? (e % 4) * 4 + e / 4 : e;
                const double t = JU[(long)(s * 16 + ee) * C + k];
                v[s][h] = ok ? t : 0.0;
            }
        }
#pragma unroll
        for (int s = 0; s < 9; ++s)
#pragma unroll
            for (int h = 0; h < kMvPasses; ++h) sJ[s][h * kMvSlots + slot][lc] = v[s][h];
    }
    __syncthreads();
    const int c = c0 + lc;
    if (c >= C) return;
    const int n = c / M, m = c - n * M;
    int nb[9];
    // forward: neighbour s of c;  transposed: cell k = c - offset(s); the cell itself where
    // that lies outside the grid (its staged block is zero)
#pragma unroll
    for (int s = 0; s < 9; ++s) {
        const int mm = tr ? m - kSlotDm[s] : m + kSlotDm[s];
        const int nn = wr(tr ? n - kSlotDn[s] : n + kSlotDn[s]);
        nb[s] = (mm >= 0 && mm < M) ? nn * M + mm : c;
    }
    for (int k = slot; k < S; k += kMvSlots) {
        // keeps the block reads from LDS inside the loop: hoisted, they would hold 288 registers
        asm volatile("" ::: "memory");
        // one buffer descriptor per vector (k is the wave's), 32-bit lane offsets (cells < 2^24)
        // and the component as a scalar offset: no 64-bit address is formed per load
        const int kw = __builtin_amdgcn_readfirstlane(k);
        const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(x + (long)kw * 8 * C), (short)0, 64 * C, 0x00020000);
        Z v[9][4];
#pragma unroll
        for (int s = 0; s < 9; ++s)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double2 t = __builtin_bit_cast(
                    double2, __builtin_amdgcn_raw_buffer_load_b128(xr, nb[s] * 16, j * C * 16, 0));
                v[s][j] = Z{t.x, t.y};
            }
        // all 36 loads in flight before the first wait (the scheduler would otherwise sink
        // them to their uses, two at a time)
        __builtin_amdgcn_sched_barrier(0);
        Z acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
#pragma unroll
        for (int s = 0; s < 9; ++s)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i] = zfma_re(sJ[s][4 * i + j][lc], v[s][j], acc[i]);
        const Z shift = tr ? Z{sh[2 * k], -sh[2 * k + 1]} : Z{sh[2 * k], sh[2 * k + 1]};
        double* yk = y + (long)k * 8 * C;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const Z r = zsub(acc[i], zmul(shift, v[0][i]));
            const long e = (long)i * C + c;
            yk[2 * e] = r.r;
            yk[2 * e + 1] = r.i;
        }
    }
}


// Ring block of the system: D[s][4n+v][4n2+w] for ring k, row-major or column-major.  Untransposed it is the block of
// cell (k, n) on neighbour n2 = n + o (o in -2..2, periodic); transposed, the block of cell
// (k, n2) on n, with v and w swapped.  The shift (conj for the transposed system) is on the
// diagonal.  One thread per complex element; elements outside the band are written as zero.
__global__ void __launch_bounds__(256)
k_ring_band(int M, int N, int S, const double* __restrict__ JU, const double* __restrict__ shifts,
            int transposed, int k, int colmajor, double* __restrict__ D) {
    const long n4 = 4L * N;
    const long total = (long)S * n4 * n4;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    // consecutive threads write consecutive addresses in either layout
    const long fast = e % n4, slow = (e / n4) % n4;
    const long col = colmajor ? slow : fast;
    const long row = colmajor ? fast : slow;
    const int s = (int)(e / (n4 * n4));
    const int n = (int)(row >> 2), v = (int)(row & 3);
    const int n2 = (int)(col >> 2), w = (int)(col & 3);
    int d = n2 - n;
    if (d < 0) d += N;
    // offset of n2 relative to n -> block index (c, b, bb, f, ff = 0..4)
    int blk = -1;
    if (d == 0) blk = 0;
    else if (d == 1) blk = 3;
    else if (d == 2) blk = 4;
    else if (d == N - 1) blk = 1;
    else if (d == N - 2) blk = 2;
    double re = 0.0, im = 0.0;
    if (blk >= 0) {
        const long C = (long)M * N;
        if (!transposed) {
            re = JU[((long)(blk * 4 + v) * 4 + w) * C + (long)n * M + k];
        } else {
            // A^T[(n,v)][(n2,w)] = A[(n2,w)][(n,v)]: block of cell n2 on n = n2 - offset
            const int tb = blk == 0 ? 0 : blk == 3 ? 1 : blk == 4 ? 2 : blk == 1 ? 3 : 4;
            re = JU[((long)(tb * 4 + w) * 4 + v) * C + (long)n2 * M + k];
        }
        if (row == col) {
            re -= shifts[2 * s];
            im = transposed ? shifts[2 * s + 1] : -shifts[2 * s + 1];
        }
    }
    D[2 * e] = re;
    D[2 * e + 1] = im;
}

// Z[s] = beta Z[s] + alpha C Y[s] with C the coupling of ring k to ring k + dk (dk in
// {-2, -1, 1, 2}), block diagonal in n: C[(n,v)][(n,w)].  Untransposed, C is block ii, i, o
// or oo of ring k; transposed, the block of ring k + dk on ring k with v and w swapped.
// Y (S, 4N, q) complex, or null for the identity (q = 4N), Z (S, 4N, q) complex; both
// row-major or both column-major (each batch entry a 4N x q matrix).
__global__ void __launch_bounds__(256)
k_ring_couple(int M, int N, int S, const double* __restrict__ JU, int transposed, int k, int dk,
              int colmajor, int q, double alpha, const double* __restrict__ Y, double beta,
              double* __restrict__ Z) {
    const long n4 = 4L * N;
    const long total = (long)S * n4 * q;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const long s = e / (n4 * q);
    const long in = e - s * n4 * q;
    const long col = colmajor ? in / n4 : in % q;
    const long row = colmajor ? in % n4 : in / q;
    const int n = (int)(row >> 2), v = (int)(row & 3);
    const long C = (long)M * N;
    int blk, ring;
    if (!transposed) {
        blk = dk == -1 ? 5 : dk == -2 ? 6 : dk == 1 ? 7 : 8;
        ring = k;
    } else {
        blk = dk == -1 ? 7 : dk == -2 ? 8 : dk == 1 ? 5 : 6;
        ring = k + dk;
    }
    const double* cb = JU + (long)blk * 16 * C + (long)n * M + ring;
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const double c = transposed ? cb[(long)(w * 4 + v) * C] : cb[(long)(v * 4 + w) * C];
        if (Y) {
            const long yr = 4L * n + w;
            const long yi = s * n4 * q + (colmajor ? col * n4 + yr : yr * q + col);
            re = fma(c, Y[2 * yi], re);
            im = fma(c, Y[2 * yi + 1], im);
        } else if (col == 4L * n + w) {
            re = c;
        }
    }
    if (beta != 0.0) {
        re = fma(alpha, re, beta * Z[2 * e]);
        im = fma(alpha, im, beta * Z[2 * e + 1]);
    } else {
        re *= alpha;
        im *= alpha;
    }
    Z[2 * e] = re;
    Z[2 * e + 1] = im;
}


// GMRES: Givens update of column j of the S Hessenberg matrices, in place (one thread per
// shift).  H (S, R+1, R), cs and sn (S, R), g (S, R+1): complex, row-major; est (S): |g[j+1]|,
// the residual norm of the least-squares problem after step j.  Rotations are
// G = [[conj(c), conj(s)], [-s, c]] with c = a/|(a, b)|, s = b/|(a, b)|.
__global__ void __launch_bounds__(64)
k_gmres_rotate(int S, int R, int j, double* __restrict__ Hd, double* __restrict__ csd,
               double* __restrict__ snd, double* __restrict__ gd, double* __restrict__ est) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    Z* H = reinterpret_cast<Z*>(Hd) + (long)s * (R + 1) * R;
    Z* cs = reinterpret_cast<Z*>(csd) + (long)s * R;
    Z* sn = reinterpret_cast<Z*>(snd) + (long)s * R;
    Z* g = reinterpret_cast<Z*>(gd) + (long)s * (R + 1);
    auto conj = [](Z a) { return Z{a.r, -a.i}; };
    for (int i = 0; i < j; ++i) {
        const Z hi = H[i * R + j], hi1 = H[(i + 1) * R + j];
        H[i * R + j] = zadd(zmul(conj(cs[i]), hi), zmul(conj(sn[i]), hi1));
        H[(i + 1) * R + j] = zsub(zmul(cs[i], hi1), zmul(sn[i], hi));
    }
    const Z a = H[j * R + j], b = H[(j + 1) * R + j];
    const double den = sqrt(zabs2(a) + zabs2(b));
    if (den == 0.0) {
        cs[j] = Z{1.0, 0.0};
        sn[j] = Z{0.0, 0.0};
    } else {
        cs[j] = Z{a.r / den, a.i / den};
        sn[j] = Z{b.r / den, b.i / den};
    }
    H[j * R + j] = Z{den, 0.0};
    H[(j + 1) * R + j] = Z{0.0, 0.0};
    const Z gj = g[j];
    g[j] = zmul(conj(cs[j]), gj);
    const Z gn = zmul(sn[j], gj);
    g[j + 1] = Z{-gn.r, -gn.i};
    est[s] = sqrt(zabs2(g[j + 1]));
}

// Batched multi-dot for Gram-Schmidt: h[s][i] = sum_k conj(V[s][i][k]) w[s][k] for i < q, with
// V rows of length n (row stride ldv, batch stride bsv) and w of batch stride bsw.  Library
// GEMMs handle this shape (tiny output, K of millions) without splitting K; here stage 1 cuts K
// into nb chunks, one workgroup per (chunk, s), and stage 2 adds the chunks in a fixed order:
// deterministic, no atomics, no memset.
constexpr int kDotThreads = 256;

// Sums acc[i], i < nout, over the workgroup (a butterfly in each wave, then the waves in order)
// into out[2i], out[2i + 1].
template <int QMAX>
__device__ __forceinline__ void block_sums(const Z* acc, int nout, double (*s_red)[2 * QMAX],
                                           double* __restrict__ out) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int i = 0; i < QMAX; ++i) {
        if (i < nout) {
            double ar = acc[i].r, ai = acc[i].i;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                ar += __shfl_xor(ar, off);
                ai += __shfl_xor(ai, off);
            }
            if (lane == 0) {
                s_red[wv][2 * i] = ar;
                s_red[wv][2 * i + 1] = ai;
            }
        }
    }
    __syncthreads();
    if (tid < 2 * nout) {
        double t = 0.0;
#pragma unroll
        for (int u = 0; u < kDotThreads / 64; ++u) t += s_red[u][tid];
        out[tid] = t;
    }
}

// Stage 1: workgroup (chunk b, shift s) reads each element of w once and the q <= QMAX rows of V
// at it, and writes the chunk's q sums to P[((s*nb + b)*pq + i)] (complex; pq >= q is the row
// count of the whole projection when rows come in slices, P then offset to the slice).
// NORM: also |w|^2, written after the q sums (pq >= q + 1).
template <int QMAX, bool NORM = false>
__global__ void __launch_bounds__(kDotThreads)
k_multidot_partial(int n, int q, const double* __restrict__ V, long ldv, long bsv,
                   const double* __restrict__ w, long bsw, int nb, int pq, double* __restrict__ P) {
    __shared__ double s_red[kDotThreads / 64][2 * QMAX];
    __shared__ double s_redn[kDotThreads / 64][2];
    const int b = blockIdx.x, s = blockIdx.y, tid = threadIdx.x;
    const long chunk = ((long)n + nb - 1) / nb;
    const long k0 = (long)b * chunk, k1 = min((long)n, k0 + chunk);
    const double2* Vs = reinterpret_cast<const double2*>(V) + (long)s * bsv;
    const double2* ws = reinterpret_cast<const double2*>(w) + (long)s * bsw;
    Z acc[QMAX], accn = {0.0, 0.0};
#pragma unroll
    for (int i = 0; i < QMAX; ++i) acc[i] = Z{0.0, 0.0};
#pragma unroll 1
    for (long k = k0 + tid; k < k1; k += kDotThreads) {
        const double2 x = ws[k];
#pragma unroll
        for (int i = 0; i < QMAX; ++i)
            if (i < q) {   // conj(v) * x
                const double2 v = Vs[(long)i * ldv + k];
                acc[i].r = fma(v.x, x.x, fma(v.y, x.y, acc[i].r));
                acc[i].i = fma(v.x, x.y, fma(-v.y, x.x, acc[i].i));
            }
        if (NORM) accn.r = fma(x.x, x.x, fma(x.y, x.y, accn.r));
    }
    double* out = P + 2 * ((long)s * nb + b) * pq;
    block_sums<QMAX>(acc, q, s_red, out);
    if (NORM) block_sums<1>(&accn, 1, s_redn, out + 2 * q);
}

// Stage 2: one wave per output (s, i).  Lane l adds chunks l, l + 64, ...; a butterfly over the
// wave adds the lanes.  Both orders are fixed, so the sum is deterministic.
__global__ void __launch_bounds__(64)
k_multidot_final(int S, int q, int nb, const double* __restrict__ P, double* __restrict__ h,
                 long bsh) {
    const int t = blockIdx.x;
    if (t >= S * q) return;
    const int s = t / q, i = t - s * q, lane = threadIdx.x;
    double ar = 0.0, ai = 0.0;
    for (int b = lane; b < nb; b += 64) {
        const long e = ((long)s * nb + b) * q + i;
        ar += P[2 * e];
        ai += P[2 * e + 1];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ar += __shfl_xor(ar, off);
        ai += __shfl_xor(ai, off);
    }
    if (lane == 0) {
        h[2 * ((long)s * bsh + i)] = ar;
        h[2 * ((long)s * bsh + i) + 1] = ai;
    }
}

// Fused Gram-Schmidt pass: w[s] -= sum_{i<q} V[s][i] h[s][i], reading each row of V once, and in
// the same pass (mode 1) the partial projections conj(V[s][i]) . w_new of every row, or (mode 2)
// the partial squared norm |w_new|^2, one value per (chunk, s) for k_multidot_final.  Mode 0
// updates only.  The chunking of K is that of k_multidot_partial.  q <= QMAX: the rows of one
// element and their projections stay in registers.
template <int QMAX>
__global__ void __launch_bounds__(kDotThreads)
k_gs_update(int n, int q, const double* __restrict__ V, long ldv, long bsv,
            const double* __restrict__ h, long bsh, double* __restrict__ w, long bsw, int nb,
            int mode, double* __restrict__ P, const int* __restrict__ only = nullptr) {
    __shared__ double s_h[2 * QMAX];
    __shared__ double s_red[kDotThreads / 64][2 * QMAX];
    const int b = blockIdx.x, s = blockIdx.y, tid = threadIdx.x;
    if (only && !only[s]) return;   // whole workgroups: shifts the caller leaves out
    if (tid < 2 * q) s_h[tid] = h[2 * (long)s * bsh + tid];
    __syncthreads();
    const long chunk = ((long)n + nb - 1) / nb;
    const long k0 = (long)b * chunk, k1 = min((long)n, k0 + chunk);
    const double2* Vs = reinterpret_cast<const double2*>(V) + (long)s * bsv;
    double2* ws = reinterpret_cast<double2*>(w) + (long)s * bsw;
    Z acc[QMAX];
#pragma unroll
    for (int i = 0; i < QMAX; ++i) acc[i] = Z{0.0, 0.0};
#pragma unroll 1
    for (long k = k0 + tid; k < k1; k += kDotThreads) {
        const double2 x0 = ws[k];
        Z x = {x0.x, x0.y};
        Z v[QMAX];
#pragma unroll
        for (int i = 0; i < QMAX; ++i) {
            if (i < q) {
                const double2 t = Vs[(long)i * ldv + k];
                v[i] = Z{t.x, t.y};
            }
        }
#pragma unroll
        for (int i = 0; i < QMAX; ++i) {   // x -= v_i h_i, rows in order
            if (i < q) {
                const double hr = s_h[2 * i], hi = s_h[2 * i + 1];
                x = Z{fma(-v[i].r, hr, fma(v[i].i, hi, x.r)), fma(-v[i].r, hi, fma(-v[i].i, hr, x.i))};
            }
        }
        ws[k] = make_double2(x.r, x.i);
        if (mode == 1) {
#pragma unroll
            for (int i = 0; i < QMAX; ++i)
                if (i < q) {   // conj(v) * x
                    acc[i].r = fma(v[i].r, x.r, fma(v[i].i, x.i, acc[i].r));
                    acc[i].i = fma(v[i].r, x.i, fma(-v[i].i, x.r, acc[i].i));
                }
        } else if (mode == 2) {
            acc[0].r = fma(x.r, x.r, fma(x.i, x.i, acc[0].r));
        }
    }
    if (mode == 0) return;
    const int nout = mode == 1 ? q : 1;
    block_sums<QMAX>(acc, nout, s_red, P + 2 * ((long)s * nb + b) * nout);
}

template <int QMAX>
void launch_gs_update(int S, int n, int q, const double* V, long ldv, long bsv, const double* h,
                      long bsh, double* w, long bsw, int nb, int mode, double* P, hipStream_t st,
                      const int* only = nullptr) {
    k_gs_update<QMAX><<<dim3(nb, S), kDotThreads, 0, st>>>(n, q, V, ldv, bsv, h, bsh, w, bsw, nb,
                                                           mode, P, only);
}

// Selective reorthogonalisation (Daniel, Gragg, Kaufman and Stewart, Math. Comp. 30 (1976)):
// after w -= V h1, repeat the projection only if it removed more than half of |w|^2, i.e.
// |h1|^2 > |w|^2 / 2.  First update pass: hA holds h1 (q values) and |w|^2 (index q), stride
// q + 1.  Every workgroup evaluates the test the same way; the shift's flag goes to flags[s].
// The pass writes q + 1 partial sums per (chunk, s): the next projections V^H w_new if the
// shift is flagged (zeros otherwise), and |w_new|^2.
template <int QMAX>
__global__ void __launch_bounds__(kDotThreads)
k_dgks_update(int n, int q, const double* __restrict__ V, long ldv, long bsv,
              const double* __restrict__ hA, double* __restrict__ w, long bsw, int nb,
              double* __restrict__ P, int* __restrict__ flags) {
    __shared__ double s_h[2 * QMAX];
    __shared__ double s_red[kDotThreads / 64][2 * QMAX];
    __shared__ double s_redn[kDotThreads / 64][2];
    __shared__ int s_flag;
    const int b = blockIdx.x, s = blockIdx.y, tid = threadIdx.x;
    const double* hs = hA + 2 * (long)s * (q + 1);
    if (tid < 2 * q) s_h[tid] = hs[tid];
    if (tid == 0) {
        double h2 = 0.0;
        for (int i = 0; i < q; ++i) h2 += hs[2 * i] * hs[2 * i] + hs[2 * i + 1] * hs[2 * i + 1];
        s_flag = h2 > 0.5 * hs[2 * q];
        if (b == 0) flags[s] = s_flag;
    }
    __syncthreads();
    const bool reorth = s_flag;
    const long chunk = ((long)n + nb - 1) / nb;
    const long k0 = (long)b * chunk, k1 = min((long)n, k0 + chunk);
    const double2* Vs = reinterpret_cast<const double2*>(V) + (long)s * bsv;
    double2* ws = reinterpret_cast<double2*>(w) + (long)s * bsw;
    Z acc[QMAX], accn = {0.0, 0.0};
#pragma unroll
    for (int i = 0; i < QMAX; ++i) acc[i] = Z{0.0, 0.0};
#pragma unroll 1
    for (long k = k0 + tid; k < k1; k += kDotThreads) {
        const double2 x0 = ws[k];
        Z x = {x0.x, x0.y};
        Z v[QMAX];
#pragma unroll
        for (int i = 0; i < QMAX; ++i) {
            if (i < q) {
                const double2 t = Vs[(long)i * ldv + k];
                v[i] = Z{t.x, t.y};
            }
        }
#pragma unroll
        for (int i = 0; i < QMAX; ++i) {   // x -= v_i h_i, rows in order
            if (i < q) {
                const double hr = s_h[2 * i], hi = s_h[2 * i + 1];
                x = Z{fma(-v[i].r, hr, fma(v[i].i, hi, x.r)), fma(-v[i].r, hi, fma(-v[i].i, hr, x.i))};
            }
        }
        ws[k] = make_double2(x.r, x.i);
        if (reorth) {
#pragma unroll
            for (int i = 0; i < QMAX; ++i)
                if (i < q) {   // conj(v) * x
                    acc[i].r = fma(v[i].r, x.r, fma(v[i].i, x.i, acc[i].r));
                    acc[i].i = fma(v[i].r, x.i, fma(-v[i].i, x.r, acc[i].i));
                }
        }
        accn.r = fma(x.r, x.r, fma(x.i, x.i, accn.r));
    }
    double* out = P + 2 * ((long)s * nb + b) * (q + 1);
    block_sums<QMAX>(acc, q, s_red, out);
    block_sums<1>(&accn, 1, s_redn, out + 2 * q);
}

// Last stage, one wave per shift: h = h1 + h2 (h2 only where flagged) and |w|^2, from the
// second update's partial norms where flagged, else from the first update's (hB index q).
__global__ void __launch_bounds__(64)
k_dgks_finish(int S, int q, int nb, const double* __restrict__ P, const int* __restrict__ flags,
              const double* __restrict__ hA, const double* __restrict__ hB, double* __restrict__ h,
              long bsh, double* __restrict__ nrm2) {
    const int s = blockIdx.x, lane = threadIdx.x;
    if (s >= S) return;
    const bool reorth = flags[s];
    const double* a = hA + 2 * (long)s * (q + 1);
    const double* c = hB + 2 * (long)s * (q + 1);
    for (int i = lane; i < q; i += 64) {
        h[2 * ((long)s * bsh + i)] = a[2 * i] + (reorth ? c[2 * i] : 0.0);
        h[2 * ((long)s * bsh + i) + 1] = a[2 * i + 1] + (reorth ? c[2 * i + 1] : 0.0);
    }
    double t = 0.0;
    if (reorth)
        for (int b = lane; b < nb; b += 64) t += P[2 * ((long)s * nb + b)];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if (lane == 0) nrm2[s] = reorth ? t : c[2 * q];
}

template <int QMAX>
void dgks_passes(int S, int n, int q, const double* V, long ldv, long bsv, double* w, long bsw,
                 int nb, double* P, double* hA, double* hB, int* flags, hipStream_t st) {
    k_multidot_partial<QMAX, true><<<dim3(nb, S), kDotThreads, 0, st>>>(n, q, V, ldv, bsv, w, bsw,
                                                                        nb, q + 1, P);
    k_multidot_final<<<S * (q + 1), 64, 0, st>>>(S, q + 1, nb, P, hA, q + 1);
    k_dgks_update<QMAX><<<dim3(nb, S), kDotThreads, 0, st>>>(n, q, V, ldv, bsv, hA, w, bsw, nb, P,
                                                             flags);
    k_multidot_final<<<S * (q + 1), 64, 0, st>>>(S, q + 1, nb, P, hB, q + 1);
    launch_gs_update<QMAX>(S, n, q, V, ldv, bsv, hB, q + 1, w, bsw, nb, 2, P, st, flags);
}
constexpr int kGsMaxRows = 32;

template <int QMAX>
void launch_multidot(int S, int n, int q, const double* V, long ldv, long bsv, const double* w,
                     long bsw, int nb, int pq, double* P, hipStream_t st) {
    k_multidot_partial<QMAX><<<dim3(nb, S), kDotThreads, 0, st>>>(n, q, V, ldv, bsv, w, bsw, nb,
                                                                  pq, P);
}
// projections of q rows (any count): slices of at most kGsMaxRows rows, then the chunk sums
void multidot(int S, int n, int q, const double* V, long ldv, long bsv, const double* w, long bsw,
              int nb, double* P, double* h, long bsh, hipStream_t st) {
    for (int i0 = 0; i0 < q; i0 += kGsMaxRows) {
        const int qs = q - i0 < kGsMaxRows ? q - i0 : kGsMaxRows;
        const double* Vi = V + 2 * (long)i0 * ldv;
        double* Pi = P + 2 * (long)i0;
        if (qs <= 4) launch_multidot<4>(S, n, qs, Vi, ldv, bsv, w, bsw, nb, q, Pi, st);
        else if (qs <= 8) launch_multidot<8>(S, n, qs, Vi, ldv, bsv, w, bsw, nb, q, Pi, st);
        else if (qs <= 16) launch_multidot<16>(S, n, qs, Vi, ldv, bsv, w, bsw, nb, q, Pi, st);
        else launch_multidot<kGsMaxRows>(S, n, qs, Vi, ldv, bsv, w, bsw, nb, q, Pi, st);
    }
    k_multidot_final<<<S * q, 64, 0, st>>>(S, q, nb, P, h, bsh);
}

void gs_update_slice(int S, int n, int q, const double* V, long ldv, long bsv, const double* h,
                     long bsh, double* w, long bsw, int nb, int mode, double* P, hipStream_t st) {
    if (q <= 4) launch_gs_update<4>(S, n, q, V, ldv, bsv, h, bsh, w, bsw, nb, mode, P, st);
    else if (q <= 8) launch_gs_update<8>(S, n, q, V, ldv, bsv, h, bsh, w, bsw, nb, mode, P, st);
    else if (q <= 16) launch_gs_update<16>(S, n, q, V, ldv, bsv, h, bsh, w, bsw, nb, mode, P, st);
    else launch_gs_update<kGsMaxRows>(S, n, q, V, ldv, bsv, h, bsh, w, bsw, nb, mode, P, st);
}

// One rocBLAS handle per device for the dense factorisations, created on first use.
constexpr int kMaxDevices = 64;
rocblas_handle g_handles[kMaxDevices] = {};
std::mutex g_handle_mutex;

rocblas_handle dense_handle(hipStream_t stream, int* err) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *err = 1;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_handle_mutex);
    if (!g_handles[dev] && rocblas_create_handle(&g_handles[dev]) != rocblas_status_success) {
        g_handles[dev] = nullptr;
        *err = 1;
        return nullptr;
    }
    if (rocblas_set_stream(g_handles[dev], stream) != rocblas_status_success) {
        *err = 1;
        return nullptr;
    }
    *err = 0;
    return g_handles[dev];
}

thread_local char g_rerr[256] = "";
int rfail(int code, const char* msg) {
    snprintf(g_rerr, sizeof g_rerr, "%s", msg);
    return code;
}

}  // namespace

extern "C" {

const char* pea_resolvent_last_error(void) { return g_rerr; }

long long pea_line_factor_bytes(int M, int N, int S) {
    return (long long)line_groups(N) * kLineGroup * M * 16 * (S * 4 * 16 + 8) +
           80LL * M * N * 8;
}

int pea_line_factor(int M, int N, int S, const double* JU, const double* shifts, int transposed,
                    double* F, void* stream) {
    if (M < 4 || N < 5 || S < 1 || S > 65535) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!JU || !shifts || !F) return rfail(PEA_ERR_NULL, "null argument");
    k_line_blocks<<<dim3((M + kTrM - 1) / kTrM, (N + kTrN - 1) / kTrN, 80), 256, 0,
                    (hipStream_t)stream>>>(M, N, JU, transposed, F + sb_offset(M, N, S));
    k_line_factor<<<dim3((N + kFactorThreads - 1) / kFactorThreads, S), kFactorThreads, 0,
                    (hipStream_t)stream>>>(M, N, S, shifts, transposed, F);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_line_factor");
}

int pea_line_solve_lanes(int lanes) {
    if (lanes != 0 && lanes != 4 && lanes != 16) return rfail(PEA_ERR_DIMS, "lanes must be 0, 4 or 16");
    g_line_lanes = lanes;
    return PEA_OK;
}

int pea_line_solve(int M, int N, int S, const double* JU, const double* F, int transposed,
                   const double* r, double* x, void* stream) {
    if (M < 4 || N < 5 || S < 1 || S > 65535) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!F || !r || !x) return rfail(PEA_ERR_NULL, "null argument");
    if (((uintptr_t)F | (uintptr_t)r | (uintptr_t)x) & 15)
        return rfail(PEA_ERR_DIMS, "F, r and x must be 16-byte aligned");
    // the factorisation stored the system's blocks (transposed or not): the solve is the same
    (void)transposed;
    (void)JU;
    // four lanes per line while that still gives two waves per SIMD (S*N/16 >= 2048), else
    // sixteen: the rows form moves fewer values between lanes, the elements form has 4x the
    // parallelism for few shifts
    const int lanes = g_line_lanes ? g_line_lanes : ((long)S * N >= 32768 ? 4 : 16);
    if (lanes == 4)
        k_line_solve_rows<<<dim3((4 * N + kLineThreads - 1) / kLineThreads, S), kLineThreads, 0,
                             (hipStream_t)stream>>>(M, N, S, F, r, x);
    else
        k_line_solve_elems<<<dim3((16 * N + kLineThreads - 1) / kLineThreads, S), kLineThreads, 0,
                              (hipStream_t)stream>>>(M, N, S, F, r, x);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_line_solve");
}

int pea_jacobi_factor(int M, int N, int S, const double* JU, const double* shifts,
                      int transposed, double* F, void* stream) {
    if (M < 4 || N < 5 || S < 1 || S > 65535) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!JU || !shifts || !F) return rfail(PEA_ERR_NULL, "null argument");
    const int C = M * N;
    k_jacobi_factor<<<dim3((C + 255) / 256, S), 256, 0, (hipStream_t)stream>>>(C, S, JU, shifts,
                                                                             transposed, F);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_jacobi_factor");
}

int pea_jacobi_solve(int M, int N, int S, const double* F, const double* r, double* x,
                     void* stream) {
    if (M < 4 || N < 5 || S < 1 || S > 65535) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!F || !r || !x) return rfail(PEA_ERR_NULL, "null argument");
    const int C = M * N;
    k_jacobi_solve<<<dim3((C + 255) / 256, S), 256, 0, (hipStream_t)stream>>>(C, S, F, r, x);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_jacobi_solve");
}

int pea_block_multi(int M, int N, int S, const double* JU, const double* shifts, int transposed,
                    const double* x, double* y, void* stream) {
    if (M < 4 || N < 5 || S < 1) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!JU || !shifts || !x || !y) return rfail(PEA_ERR_NULL, "null argument");
    const int C = M * N;
    if (C >= (1 << 24)) return rfail(PEA_ERR_DIMS, "block product: at most 2^24 - 1 cells");
    k_block_multi<<<(C + kMvCells - 1) / kMvCells, kMvCells * kMvSlots, 0, (hipStream_t)stream>>>(
        M, N, S, JU, shifts, transposed, x, y);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_block_multi");
}

int pea_ring_band(int M, int N, int S, const double* JU, const double* shifts, int transposed,
                  int k, int colmajor, double* D, void* stream) {
    if (M < 4 || N < 5 || S < 1) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (k < 0 || k >= M) return rfail(PEA_ERR_DIMS, "ring index out of range");
    if (!JU || !shifts || !D) return rfail(PEA_ERR_NULL, "null argument");
    const long total = (long)S * 16L * N * N;
    const long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffL) return rfail(PEA_ERR_DIMS, "ring block too large");
    k_ring_band<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(M, N, S, JU, shifts,
                                                                   transposed, k, colmajor, D);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_ring_band");
}

int pea_ring_couple(int M, int N, int S, const double* JU, int transposed, int k, int dk,
                    int colmajor, int q, double alpha, const double* Y, double beta, double* Z,
                    void* stream) {
    if (M < 4 || N < 5 || S < 1 || q < 1) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!(dk == -2 || dk == -1 || dk == 1 || dk == 2))
        return rfail(PEA_ERR_DIMS, "ring offset must be -2, -1, 1 or 2");
    if (k < 0 || k >= M || k + dk < 0 || k + dk >= M)
        return rfail(PEA_ERR_DIMS, "ring index out of range");
    if (!Y && q != 4 * N) return rfail(PEA_ERR_DIMS, "identity operand needs q = 4N");
    if (!JU || !Z) return rfail(PEA_ERR_NULL, "null argument");
    const long total = (long)S * 4L * N * q;
    const long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffL) return rfail(PEA_ERR_DIMS, "operand too large");
    k_ring_couple<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(
        M, N, S, JU, transposed, k, dk, colmajor, q, alpha, Y, beta, Z);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_ring_couple");
}

int pea_dense_lu_factor(int S, int n, double* A, int* ipiv, int* info, void* stream) {
    if (S < 1 || n < 1) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!A || !ipiv || !info) return rfail(PEA_ERR_NULL, "null argument");
    int err = 0;
    rocblas_handle h = dense_handle((hipStream_t)stream, &err);
    if (err) return rfail(PEA_ERR_HIP, "rocBLAS handle");
    const rocblas_status st = rocsolver_zgetrf_strided_batched(
        h, n, n, reinterpret_cast<rocblas_double_complex*>(A), n, (rocblas_stride)n * n, ipiv,
        (rocblas_stride)n, info, S);
    return st == rocblas_status_success ? PEA_OK : rfail(PEA_ERR_HIP, "rocsolver getrf");
}

int pea_dense_lu_solve(int S, int n, int nrhs, const double* A, const int* ipiv, double* B,
                       void* stream) {
    if (S < 1 || n < 1 || nrhs < 1) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!A || !ipiv || !B) return rfail(PEA_ERR_NULL, "null argument");
    int err = 0;
    rocblas_handle h = dense_handle((hipStream_t)stream, &err);
    if (err) return rfail(PEA_ERR_HIP, "rocBLAS handle");
    const rocblas_status st = rocsolver_zgetrs_strided_batched(
        h, rocblas_operation_none, n, nrhs,
        reinterpret_cast<rocblas_double_complex*>(const_cast<double*>(A)), n,
        (rocblas_stride)n * n, ipiv, (rocblas_stride)n,
        reinterpret_cast<rocblas_double_complex*>(B), n, (rocblas_stride)n * nrhs, S);
    return st == rocblas_status_success ? PEA_OK : rfail(PEA_ERR_HIP, "rocsolver getrs");
}

int pea_gmres_rotate(int S, int R, int j, double* H, double* cs, double* sn, double* g,
                     double* est, void* stream) {
    if (S < 1 || R < 1 || j < 0 || j >= R) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!H || !cs || !sn || !g || !est) return rfail(PEA_ERR_NULL, "null argument");
    k_gmres_rotate<<<(S + 63) / 64, 64, 0, (hipStream_t)stream>>>(S, R, j, H, cs, sn, g, est);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_gmres_rotate");
}

int pea_multidot(int S, int n, int q, const double* V, long long ldv, long long bsv,
                 const double* w, long long bsw, int nb, double* P, double* h, long long bsh,
                 void* stream) {
    if (S < 1 || n < 1 || q < 1 || nb < 1 || S > 65535) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (!V || !w || !P || !h) return rfail(PEA_ERR_NULL, "null argument");
    if (((uintptr_t)V | (uintptr_t)w) & 15) return rfail(PEA_ERR_DIMS, "V and w must be 16-byte aligned");
    multidot(S, n, q, V, ldv, bsv, w, bsw, nb, P, h, bsh, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_multidot");
}

int pea_gs_update(int S, int n, int q, const double* V, long long ldv, long long bsv,
                  const double* h, long long bsh, double* w, long long bsw, int mode, int nb,
                  double* P, double* out, long long bso, void* stream) {
    if (S < 1 || n < 1 || q < 1 || nb < 1 || S > 65535) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (mode < 0 || mode > 2) return rfail(PEA_ERR_DIMS, "mode must be 0, 1 or 2");
    if (!V || !h || !w) return rfail(PEA_ERR_NULL, "null argument");
    if (mode != 0 && (!P || !out)) return rfail(PEA_ERR_NULL, "null workspace or output");
    if (((uintptr_t)V | (uintptr_t)w) & 15) return rfail(PEA_ERR_DIMS, "V and w must be 16-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    // rows in slices of at most kGsMaxRows; the norm comes with the last slice, the projections
    // of more rows than one slice holds from a separate multi-dot over the updated w
    for (int i0 = 0; i0 < q; i0 += kGsMaxRows) {
        const int qs = q - i0 < kGsMaxRows ? q - i0 : kGsMaxRows;
        const bool last = i0 + qs == q;
        const int md = !last ? 0 : (mode == 1 && q > kGsMaxRows) ? 0 : mode;
        gs_update_slice(S, n, qs, V + 2 * (long)i0 * ldv, ldv, bsv, h + 2 * (long)i0, bsh, w, bsw,
                        nb, md, P, st);
    }
    if (mode == 1 && q > kGsMaxRows) {
        multidot(S, n, q, V, ldv, bsv, w, bsw, nb, P, out, bso, st);
    } else if (mode != 0) {
        const int nout = mode == 1 ? q : 1;
        k_multidot_final<<<S * nout, 64, 0, st>>>(S, nout, nb, P, out, bso);
    }
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_gs_update");
}

int pea_gs_orthogonalize(int S, int n, int q, const double* V, long long ldv, long long bsv,
                         double* w, long long bsw, int nb, double* P, double* hw, int* flags,
                         double* h, long long bsh, double* nrm2, void* stream) {
    if (S < 1 || n < 1 || nb < 1 || S > 65535) return rfail(PEA_ERR_DIMS, "bad dimensions");
    if (q < 1 || q > kGsMaxRows) return rfail(PEA_ERR_DIMS, "q must be 1..32");
    if (!V || !w || !P || !hw || !flags || !h || !nrm2) return rfail(PEA_ERR_NULL, "null argument");
    if (((uintptr_t)V | (uintptr_t)w) & 15) return rfail(PEA_ERR_DIMS, "V and w must be 16-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    double* hA = hw;
    double* hB = hw + 2L * S * (q + 1);
    if (q <= 4) dgks_passes<4>(S, n, q, V, ldv, bsv, w, bsw, nb, P, hA, hB, flags, st);
    else if (q <= 8) dgks_passes<8>(S, n, q, V, ldv, bsv, w, bsw, nb, P, hA, hB, flags, st);
    else if (q <= 16) dgks_passes<16>(S, n, q, V, ldv, bsv, w, bsw, nb, P, hA, hB, flags, st);
    else dgks_passes<kGsMaxRows>(S, n, q, V, ldv, bsv, w, bsw, nb, P, hA, hB, flags, st);
    k_dgks_finish<<<S, 64, 0, st>>>(S, q, nb, P, flags, hA, hB, h, bsh, nrm2);
    return hipGetLastError() == hipSuccess ? PEA_OK : rfail(PEA_ERR_HIP, "k_dgks");
}

}  // extern "C"
