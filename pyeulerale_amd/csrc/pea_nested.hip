// Kernels of the nested-dissection direct resolvent (dR/dU - s I)^-1 (multifrontal block LU).
//
// The elimination tree, the front sizes and every index map come from the host
// (pyeulerale_amd/nested.py); these kernels do the numerical work of one tree level, all its
// nodes in one batch:
//
// * k_nd_assemble: writes the fronts of a level.  Front k is a dense L x L complex matrix,
//   column-major, over its cells (separator S first, padded to Pc cells, then its ring B, padded
//   to Bc cells), 4 unknowns per cell.  Every entry is the sum of the children's update matrices
//   at the entry's positions in the children's rings (a gather: no atomics, each entry written
//   once), plus 1 on the diagonal of S's padding.
// * k_nd_blocks: adds the original 4x4 blocks of dR/dU (or their transposes) whose first
//   eliminated cell lies in the node's S, the shift on the diagonal blocks.  Each front entry
//   receives at most one block, so the adds do not race.
// * pea_nd_factor: partial-pivoting LU of F11 (rocSOLVER getrf), Z = F11^-1 F12 (getrs) and the
//   update F22 - F21 Z (rocBLAS ZGEMM), strided batched over the level's fronts; or F11 inverted
//   from its LU (getri) and Z and the update as two GEMMs, so that solves are GEMMs as well.
// * k_nd_gather / k_nd_scatter / k_nd_accumulate: the solve's moves between the global vectors
//   (v * C + cell layout) and the fronts' S or B parts; the subtraction of the level's
//   contributions to shared ring cells sums them per target in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "../../include/pea.h"

namespace {

thread_local char g_nerr[256] = "";

int nfail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_nerr, sizeof g_nerr, fmt, ap);
    va_end(ap);
    return code;
}

constexpr int kMaxDevices = 64;
rocblas_handle g_nd_handles[kMaxDevices] = {};
std::mutex g_nd_mutex;

rocblas_handle nd_handle(hipStream_t stream) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_nd_mutex);
    if (!g_nd_handles[dev] && rocblas_create_handle(&g_nd_handles[dev]) != rocblas_status_success) {
        g_nd_handles[dev] = nullptr;
        return nullptr;
    }
    if (rocblas_set_stream(g_nd_handles[dev], stream) != rocblas_status_success) return nullptr;
    return g_nd_handles[dev];
}

unsigned grid_for(long long work, int threads) {
    const long long b = (work + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : (b > (1LL << 20) ? (1LL << 20) : b));
}

// Fronts of one level: F[k][i + L j], L = 4 Lc.  Child c of front k is child[2k + c] (index into
// the update store U of the level below, -1 = none); the front's cell q sits at position
// child_pos[(2k + c) Lc + q] of that child's ring (-1 = not in it).  U[kid] is Lu x Lu, Lu the
// child level's ring size in unknowns.
__global__ void __launch_bounds__(256)
k_nd_assemble(int n, int Pc, int Lc, const int* __restrict__ cells, const int* __restrict__ child,
              const int* __restrict__ child_pos, int Lu, const double2* __restrict__ U,
              double2* __restrict__ F) {
    const long long L = 4LL * Lc;
    const long long total = (long long)n * L * L;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % L);
        const long long r = t / L;
        const int j = (int)(r % L);
        const int k = (int)(r / L);
        const int ci = i >> 2, cj = j >> 2;
        double2 v = {0.0, 0.0};
        for (int c = 0; c < 2; ++c) {
            const int kid = child[2 * k + c];
            if (kid < 0) continue;
            const int* pos = child_pos + (2LL * k + c) * Lc;
            const int pi = pos[ci], pj = pos[cj];
            if (pi >= 0 && pj >= 0) {
                const double2 u = U[(long long)kid * Lu * Lu + (4LL * pi + (i & 3)) +
                                    (long long)Lu * (4LL * pj + (j & 3))];
                v.x += u.x;
                v.y += u.y;
            }
        }
        if (i == j && ci < Pc && cells[(long long)k * Lc + ci] < 0) v.x = 1.0;
        F[t] = v;
    }
}

// Original blocks: entry e = (node, front row cell, front column cell, row cell, block s); the
// block is dR_i(row cell)/dU_j(neighbour s), stored JU[((s*4 + i)*4 + j) C + cell].  The adjoint
// system places the transposed block at the transposed position.
__global__ void __launch_bounds__(256)
k_nd_blocks(int ne, const int* __restrict__ ent, int node0, long long C,
            const double* __restrict__ JU, double sre, double sim, int adjoint, long long L,
            double2* __restrict__ F) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int* q = ent + 5LL * e;
    const int k = q[0] - node0, fr = q[1], fc = q[2], cell = q[3], s = q[4];
    double2* Fk = F + (long long)k * L * L;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double v = JU[((long long)(s * 4 + i) * 4 + j) * C + cell];
            const long long row = adjoint ? 4LL * fc + j : 4LL * fr + i;
            const long long col = adjoint ? 4LL * fr + i : 4LL * fc + j;
            double2 d = Fk[row + L * col];
            d.x += v;
            if (s == 0 && i == j) {
                d.x -= sre;
                d.y -= sim;
            }
            Fk[row + L * col] = d;
        }
}

// out[k][r][4 j + v] = y[r ldy + v C + cells[k Lc + off + j]] for j < cnt (0 for padding)
__global__ void __launch_bounds__(256)
k_nd_gather(int n, int cnt, int off, int Lc, const int* __restrict__ cells, int nrhs, long long C,
            const double2* __restrict__ y, long long ldy, double2* __restrict__ out) {
    const long long w = 4LL * cnt;
    const long long total = (long long)n * nrhs * w;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int u = (int)(t % w);
        const long long kr = t / w;
        const int r = (int)(kr % nrhs);
        const int k = (int)(kr / nrhs);
        const int cell = cells[(long long)k * Lc + off + (u >> 2)];
        out[t] = cell >= 0 ? y[r * ldy + (u & 3) * C + cell] : double2{0.0, 0.0};
    }
}

// The inverse of k_nd_gather on the valid entries (each global entry written by one front).
__global__ void __launch_bounds__(256)
k_nd_scatter(int n, int cnt, int off, int Lc, const int* __restrict__ cells, int nrhs,
             long long C, const double2* __restrict__ T, double2* __restrict__ y, long long ldy) {
    const long long w = 4LL * cnt;
    const long long total = (long long)n * nrhs * w;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int u = (int)(t % w);
        const long long kr = t / w;
        const int r = (int)(kr % nrhs);
        const int k = (int)(kr / nrhs);
        const int cell = cells[(long long)k * Lc + off + (u >> 2)];
        if (cell >= 0) y[r * ldy + (u & 3) * C + cell] = T[t];
    }
}

// y[r ldy + target[t]] -= sum over e in [ptr[t], ptr[t+1]) of W[k_e][r][j_e], W of shape
// (n, nrhs, B): the level's ring contributions, summed per target in a fixed order.
__global__ void __launch_bounds__(256)
k_nd_accumulate(int ntarget, int nrhs, const int* __restrict__ ptr, const int* __restrict__ src_k,
                const int* __restrict__ src_j, const int* __restrict__ target, int B,
                const double2* __restrict__ W, double2* __restrict__ y, long long ldy) {
    const long long total = (long long)ntarget * nrhs;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int t = (int)(q % ntarget);
        const int r = (int)(q / ntarget);
        double2 acc = {0.0, 0.0};
        for (int e = ptr[t]; e < ptr[t + 1]; ++e) {
            const double2 w = W[((long long)src_k[e] * nrhs + r) * B + src_j[e]];
            acc.x += w.x;
            acc.y += w.y;
        }
        double2& d = y[r * ldy + target[t]];
        d.x -= acc.x;
        d.y -= acc.y;
    }
}

}  // namespace

extern "C" {

const char* pea_nd_last_error(void) { return g_nerr; }

int pea_nd_assemble(int n, int Pc, int Bc, const int* cells, const int* child,
                    const int* child_pos, int Bc_child, const double* U_child, double* F,
                    void* stream) {
    if (n < 1 || Pc < 1 || Bc < 0 || Bc_child < 0) return nfail(PEA_ERR_DIMS, "bad dimensions");
    if (!cells || !child || !child_pos || !F) return nfail(PEA_ERR_NULL, "null argument");
    const int Lc = Pc + Bc;
    const long long total = (long long)n * 16LL * Lc * Lc;
    hipLaunchKernelGGL(k_nd_assemble, dim3(grid_for(total, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, Pc, Lc, cells, child, child_pos, 4 * Bc_child,
                       reinterpret_cast<const double2*>(U_child), reinterpret_cast<double2*>(F));
    return hipGetLastError() == hipSuccess ? PEA_OK : nfail(PEA_ERR_HIP, "k_nd_assemble");
}

int pea_nd_blocks(int ne, const int* entries, int node0, int M, int N, const double* JU,
                  double shift_re, double shift_im, int adjoint, int Pc, int Bc, double* F,
                  void* stream) {
    if (ne < 0 || M < 4 || N < 5 || Pc < 1 || Bc < 0) return nfail(PEA_ERR_DIMS, "bad dimensions");
    if (ne == 0) return PEA_OK;
    if (!entries || !JU || !F) return nfail(PEA_ERR_NULL, "null argument");
    hipLaunchKernelGGL(k_nd_blocks, dim3((ne + 255) / 256), dim3(256), 0, (hipStream_t)stream, ne,
                       entries, node0, (long long)M * N, JU, shift_re, shift_im, adjoint,
                       4LL * (Pc + Bc), reinterpret_cast<double2*>(F));
    return hipGetLastError() == hipSuccess ? PEA_OK : nfail(PEA_ERR_HIP, "k_nd_blocks");
}

int pea_nd_factor(int n, int P, int B, double* F, int* ipiv, int* info, int invert, double* Zws,
                  void* stream) {
    if (n < 1 || P < 1 || B < 0) return nfail(PEA_ERR_DIMS, "bad dimensions");
    if (!F || !ipiv || !info || (invert && B > 0 && !Zws)) return nfail(PEA_ERR_NULL, "null argument");
    rocblas_handle h = nd_handle((hipStream_t)stream);
    if (!h) return nfail(PEA_ERR_HIP, "rocBLAS handle");
    const rocblas_int L = P + B;
    const rocblas_stride sF = (rocblas_stride)L * L;
    auto* A = reinterpret_cast<rocblas_double_complex*>(F);
    const rocblas_double_complex minus_one(-1.0, 0.0), one(1.0, 0.0), zero(0.0, 0.0);
    rocblas_status st = rocsolver_zgetrf_strided_batched(h, P, P, A, L, sF, ipiv, P, info, n);
    if (st != rocblas_status_success) return nfail(PEA_ERR_HIP, "rocsolver getrf (%d)", (int)st);
    if (invert) {
        // F11 <- F11^-1, then Z = F11^-1 F12 and the update are GEMMs
        st = rocsolver_zgetri_strided_batched(h, P, A, L, sF, ipiv, P, info, n);
        if (st != rocblas_status_success) return nfail(PEA_ERR_HIP, "rocsolver getri (%d)", (int)st);
        if (B == 0) return PEA_OK;
        auto* Z = reinterpret_cast<rocblas_double_complex*>(Zws);
        const rocblas_stride sZ = (rocblas_stride)P * B;
        st = rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, P, B,
                                           P, &one, A, L, sF, A + (long long)P * L, L, sF, &zero,
                                           Z, P, sZ, n);
        if (st != rocblas_status_success) return nfail(PEA_ERR_HIP, "rocblas zgemm (%d)", (int)st);
        st = rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, B, B,
                                           P, &minus_one, A + P, L, sF, Z, P, sZ, &one,
                                           A + P + (long long)P * L, L, sF, n);
        return st == rocblas_status_success ? PEA_OK
                                            : nfail(PEA_ERR_HIP, "rocblas zgemm (%d)", (int)st);
    }
    if (B == 0) return PEA_OK;
    st = rocsolver_zgetrs_strided_batched(h, rocblas_operation_none, P, B, A, L, sF, ipiv, P,
                                          A + (long long)P * L, L, sF, n);
    if (st != rocblas_status_success) return nfail(PEA_ERR_HIP, "rocsolver getrs (%d)", (int)st);
    st = rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, B, B, P,
                                       &minus_one, A + P, L, sF, A + (long long)P * L, L, sF, &one,
                                       A + P + (long long)P * L, L, sF, n);
    return st == rocblas_status_success ? PEA_OK : nfail(PEA_ERR_HIP, "rocblas zgemm (%d)", (int)st);
}

int pea_nd_gather(int n, int cnt, int off, int Lc, const int* cells, int nrhs, int M, int N,
                  const double* y, long long ldy, double* out, void* stream) {
    if (n < 1 || cnt < 0 || off < 0 || off + cnt > Lc || nrhs < 1 || M < 4 || N < 5)
        return nfail(PEA_ERR_DIMS, "bad dimensions");
    if (!cells || !y || !out) return nfail(PEA_ERR_NULL, "null argument");
    const long long total = (long long)n * nrhs * 4 * cnt;
    if (total == 0) return PEA_OK;
    hipLaunchKernelGGL(k_nd_gather, dim3(grid_for(total, 256)), dim3(256), 0, (hipStream_t)stream,
                       n, cnt, off, Lc, cells, nrhs, (long long)M * N,
                       reinterpret_cast<const double2*>(y), ldy, reinterpret_cast<double2*>(out));
    return hipGetLastError() == hipSuccess ? PEA_OK : nfail(PEA_ERR_HIP, "k_nd_gather");
}

int pea_nd_scatter(int n, int cnt, int off, int Lc, const int* cells, int nrhs, int M, int N,
                   const double* T, double* y, long long ldy, void* stream) {
    if (n < 1 || cnt < 0 || off < 0 || off + cnt > Lc || nrhs < 1 || M < 4 || N < 5)
        return nfail(PEA_ERR_DIMS, "bad dimensions");
    if (!cells || !T || !y) return nfail(PEA_ERR_NULL, "null argument");
    const long long total = (long long)n * nrhs * 4 * cnt;
    if (total == 0) return PEA_OK;
    hipLaunchKernelGGL(k_nd_scatter, dim3(grid_for(total, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, cnt, off, Lc, cells, nrhs, (long long)M * N,
                       reinterpret_cast<const double2*>(T), reinterpret_cast<double2*>(y), ldy);
    return hipGetLastError() == hipSuccess ? PEA_OK : nfail(PEA_ERR_HIP, "k_nd_scatter");
}

int pea_nd_accumulate(int ntarget, int nrhs, const int* ptr, const int* src_k, const int* src_j,
                      const int* target, int B, const double* W, double* y, long long ldy,
                      void* stream) {
    if (ntarget < 0 || nrhs < 1 || B < 0) return nfail(PEA_ERR_DIMS, "bad dimensions");
    if (ntarget == 0) return PEA_OK;
    if (!ptr || !src_k || !src_j || !target || !W || !y) return nfail(PEA_ERR_NULL, "null argument");
    const long long total = (long long)ntarget * nrhs;
    hipLaunchKernelGGL(k_nd_accumulate, dim3(grid_for(total, 256)), dim3(256), 0,
                       (hipStream_t)stream, ntarget, nrhs, ptr, src_k, src_j, target, B,
                       reinterpret_cast<const double2*>(W), reinterpret_cast<double2*>(y), ldy);
    return hipGetLastError() == hipSuccess ? PEA_OK : nfail(PEA_ERR_HIP, "k_nd_accumulate");
}

}  // extern "C"
