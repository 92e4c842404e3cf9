// libpea_hip.so: gfx950 kernels of the PyEulerALE spatial discretization and the C ABI declared
// in include/pea.h.
//
// Layout (all device memory, radial index fastest):
//   U[v*C + n*M + m]   X[d*P + n*(M+1) + m]   complex arrays interleaved (re, im) per element.
// Cell-centric kernels map one thread to one cell with m fastest, so neighbouring lanes touch
// neighbouring doubles in every plane; stencil reads of the angular neighbours hit L1/L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "pea_adjoint.hpp"
#include "pea_face.hpp"
#include "pea_linear.hpp"

using namespace pea;

// =============================================================================== errors
namespace {
thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(PEA_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define PEA_LAUNCH_CHECK(what)                                  \
    do {                                                        \
        hipError_t e_ = hipGetLastError();                      \
        if (e_ != hipSuccess) return hip_fail(e_, what);        \
    } while (0)

// The buffer loaders address one array with 32-bit byte offsets and a 32-bit record count
// (buf_rsrc below): a complex cell field spans 64 C bytes, so C < 2^25 cells keeps every offset
// and the count of the largest array below 2^31.
constexpr long long PEA_MAX_CELLS = (1LL << 25) - 1;

int check_dims(int M, int N) {
    if (M < 4 || N < 5) return fail(PEA_ERR_DIMS, "need M >= 4 and N >= 5 (got M=%d N=%d)", M, N);
    if ((long long)M * N > PEA_MAX_CELLS || (long long)(M + 1) * (N + 1) > 2 * PEA_MAX_CELLS)
        return fail(PEA_ERR_DIMS, "grid too large for 32-bit buffer offsets: M*N = %lld cells, "
                    "at most %lld (M=%d N=%d)", (long long)M * N, PEA_MAX_CELLS, M, N);
    return PEA_OK;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
}  // namespace

// =============================================================================== indexing
struct Grid {
    int M, N;
    int C;   // cells
    int P;   // vertices
};

static Grid make_grid(int M, int N) { return Grid{M, N, M * N, (M + 1) * (N + 1)}; }

PEA_HD int wrapn(int n, int N) { return n < 0 ? n + N : (n >= N ? n - N : n); }
// n mod N: one compare-and-add for n within a period of [0, N) (stencil offsets of a few
// cells), the integer remainder only for grids narrower than a tile's halo.
PEA_HD int wrap_any(int n, int N) {
    n = wrapn(n, N);
    if ((unsigned)n >= (unsigned)N) n = (n % N + N) % N;
    return n;
}

// MI355X dispatches workgroup b to XCD b mod 8, and each XCD has its own L2.  Tiles that share
// halo cells should run on one XCD: xcd_tile() maps the linear workgroup id to a tile id so that
// XCD x walks the contiguous tile range [start(x), start(x+1)) (a bijection for any total).
constexpr int kXcdCount = 8;
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned total) {
    const unsigned q = total / kXcdCount, r = total % kXcdCount, x = b % kXcdCount;
    return x * q + (x < r ? x : r) + b / kXcdCount;
}

// Tile (bx, by) and batch item bz (state, direction or seed of one grid) of this workgroup in a
// launch over nx x ny tiles and nz items.  Items come in chunks of up to kItemChunk; a chunk's
// tiles are numbered spatial tile after spatial tile, its items fastest, and xcd_tile() hands
// each XCD a contiguous range of that numbering.  The items of a tile share its geometry (and,
// for the products, its primal states), so one read of them into an XCD's L2 serves the whole
// chunk; the tiles next to it in m and n (which share its halo) follow within kItemChunk and
// kItemChunk * nx workgroups.  With the items slowest each XCD would take whole items, and
// every XCD would read all of the geometry: 1.8x the algorithmic reads for 8 states at
// 1025x513, against 1.2x this way.
constexpr unsigned kItemChunk = 8;
struct TileId {
    int bx, by, bz;
};
__device__ __forceinline__ TileId batch_tile(unsigned nx, unsigned ny, unsigned nz) {
    const unsigned S = nx * ny;
    const unsigned t = xcd_tile(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), S * nz);
    const unsigned c = t / (S * kItemChunk), r = t - c * S * kItemChunk;
    const unsigned cz = min(kItemChunk, nz - c * kItemChunk);   // items in this chunk
    const unsigned s = r / cz;
    return {(int)(s % nx), (int)(s / nx), (int)(c * kItemChunk + r - s * cz)};
}

// cell index (n*M + m) of stencil slot `slot` of cell (m, n); radial neighbours are only
// requested where they exist (boundary flags)
PEA_HD int slot_cell(const Grid& g, int m, int n, int slot) {
    switch (slot) {
    case B: return wrapn(n - 1, g.N) * g.M + m;
    case BB: return wrapn(n - 2, g.N) * g.M + m;
    case F: return wrapn(n + 1, g.N) * g.M + m;
    case FF: return wrapn(n + 2, g.N) * g.M + m;
    case I: return n * g.M + m - 1;
    case II: return n * g.M + m - 2;
    case O: return n * g.M + m + 1;
    case OO: return n * g.M + m + 2;
    default: return n * g.M + m;
    }
}

// vertex index of vertex slot vtx (bi, fi, fo, bo) of cell (m, n)
PEA_HD int vtx_index(const Grid& g, int m, int n, int vtx) {
    const int dm = (vtx == FO || vtx == BO) ? 1 : 0;
    const int dn = (vtx == FI || vtx == FO) ? 1 : 0;
    return (n + dn) * (g.M + 1) + m + dm;
}

// =============================================================================== loaders
// A loader turns device arrays into scalars of type T: states, vertices, velocities, Mach.

// Loads through a buffer descriptor: the element offset is one 32-bit lane value shared by
// every plane of an array, and the plane is a scalar offset, so no 64-bit address is formed per
// load.  The descriptor covers `bytes` (offsets beyond read 0); word 3 is the gfx9 raw-buffer
// format.  Arrays below 2 GiB: check_dims bounds the cell count accordingly.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const double* p, long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, int elem, int plane_elems) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, elem * 8,
                                                                          plane_elems * 8, 0));
}
// elements elem and elem + 1 of a plane (one 16-byte load at an 8-byte aligned offset)
__device__ __forceinline__ double2 buf_load_pair(__amdgpu_buffer_rsrc_t r, int elem,
                                                 int plane_elems) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, elem * 8,
                                                                            plane_elems * 8, 0));
}
// 2-vectors p and p + 1 of a vertex array (two planes of P doubles): one 16-byte load per plane
__device__ __forceinline__ void vec2_pair(const double* A, int P, int p, double* a, double* b) {
    const auto r = buf_rsrc(A, 16L * P);
    const double2 d0 = buf_load_pair(r, p, 0), d1 = buf_load_pair(r, p, P);
    a[0] = d0.x, b[0] = d0.y, a[1] = d1.x, b[1] = d1.y;
}
// complex element (re, im) of an interleaved complex array
__device__ __forceinline__ double2 buf_load_cx(__amdgpu_buffer_rsrc_t r, int elem, int plane_elems) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, elem * 16,
                                                                            plane_elems * 16, 0));
}

struct LdReal {
    const double *X, *V, *U;
    double mach;
    Grid g;
    __device__ void state(int c, double* u) const {
        const auto r = buf_rsrc(U, 32L * g.C);
#pragma unroll
        for (int v = 0; v < 4; ++v) u[v] = buf_load(r, c, v * g.C);
    }
    __device__ void vertex(int p, double* x) const {
        const auto r = buf_rsrc(X, 16L * g.P);
        x[0] = buf_load(r, p, 0);
        x[1] = buf_load(r, p, g.P);
    }
    __device__ void velocity(int p, double* w) const {
        const auto r = buf_rsrc(V, 16L * g.P);
        w[0] = buf_load(r, p, 0);
        w[1] = buf_load(r, p, g.P);
    }
    // vertices (velocities) p and p + 1, which are neighbours along m in one vertex column
    __device__ void vertex2(int p, double* xa, double* xb) const { vec2_pair(X, g.P, p, xa, xb); }
    __device__ void velocity2(int p, double* wa, double* wb) const {
        vec2_pair(V, g.P, p, wa, wb);
    }
    PEA_HD double machv() const { return mach; }
    // item b of a batch launch (grid z): the b-th of consecutive state fields
    PEA_HD LdReal at(int b) const {
        LdReal r = *this;
        r.U += (long)b * 4 * g.C;
        return r;
    }
};

// K tangent lanes: lane j is the real (j even) or imaginary (j odd) part of complex direction
// j/2; direction arrays hold K/2 consecutive complex arrays.  NULL = zero direction.
template <int K>
struct LdDual {
    const double *X, *V, *U;
    const double *dX, *dV, *dU, *dmach;
    double mach;
    Grid g;
    PEA_HD void state(int c, Dual<K>* u) const {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int e = v * g.C + c;
            u[v].v = U[e];
#pragma unroll
            for (int j = 0; j < K; ++j)
                u[v].d[j] = dU ? dU[2 * ((long)(j >> 1) * 4 * g.C + e) + (j & 1)] : 0.0;
        }
    }
    PEA_HD void vec2(const double* A, const double* dA, int p, Dual<K>* x) const {
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const int e = d * g.P + p;
            x[d].v = A[e];
#pragma unroll
            for (int j = 0; j < K; ++j)
                x[d].d[j] = dA ? dA[2 * ((long)(j >> 1) * 2 * g.P + e) + (j & 1)] : 0.0;
        }
    }
    PEA_HD void vertex(int p, Dual<K>* x) const { vec2(X, dX, p, x); }
    PEA_HD void velocity(int p, Dual<K>* w) const { vec2(V, dV, p, w); }
    PEA_HD LdDual at(int) const { return *this; }   // not launched as a batch
    PEA_HD Dual<K> machv() const {
        Dual<K> r;
        r.v = mach;
#pragma unroll
        for (int j = 0; j < K; ++j) r.d[j] = dmach ? dmach[2 * (j >> 1) + (j & 1)] : 0.0;
        return r;
    }
};

// Tangents on the states only: geometry and Mach stay plain FP64.
template <int K>
struct LdDualU {
    static_assert(K % 2 == 0, "tangent lanes come in (re, im) pairs of complex directions");
    const double *X, *V, *U;
    const double* dU;
    double mach;
    Grid g;
    __device__ void state(int c, Dual<K>* u) const {
        const auto r = buf_rsrc(U, 32L * g.C);
        const auto rd = buf_rsrc(dU, 64L * g.C * (K / 2));
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            u[v].v = buf_load(r, c, v * g.C);
#pragma unroll
            for (int d = 0; d < K / 2; ++d) {
                const double2 z = buf_load_cx(rd, c, d * 4 * g.C + v * g.C);
                u[v].d[2 * d] = z.x;
                u[v].d[2 * d + 1] = z.y;
            }
        }
    }
    __device__ void vertex(int p, double* x) const {
        const auto r = buf_rsrc(X, 16L * g.P);
        x[0] = buf_load(r, p, 0);
        x[1] = buf_load(r, p, g.P);
    }
    __device__ void velocity(int p, double* w) const {
        const auto r = buf_rsrc(V, 16L * g.P);
        w[0] = buf_load(r, p, 0);
        w[1] = buf_load(r, p, g.P);
    }
    // vertices (velocities) p and p + 1, which are neighbours along m in one vertex column
    __device__ void vertex2(int p, double* xa, double* xb) const { vec2_pair(X, g.P, p, xa, xb); }
    __device__ void velocity2(int p, double* wa, double* wb) const {
        vec2_pair(V, g.P, p, wa, wb);
    }
    PEA_HD double machv() const { return mach; }
    // item b of a batch launch (grid z): the b-th group of K/2 consecutive complex directions
    PEA_HD LdDualU at(int b) const {
        LdDualU r = *this;
        r.dU += (long)b * (K / 2) * 8 * g.C;
        return r;
    }
};

// Complex step on the states only (geometry and Mach real).
struct LdCxU {
    const double *X, *V, *U;
    const double* dU;
    double mach, h;
    int part;
    Grid g;
    __device__ void state(int c, Cx* u) const {
        const auto r = buf_rsrc(U, 32L * g.C);
        const auto rd = buf_rsrc(dU, 64L * g.C);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double2 z = buf_load_cx(rd, c, v * g.C);
            u[v] = Cx{buf_load(r, c, v * g.C), h * (part ? z.y : z.x)};
        }
    }
    __device__ void vertex(int p, double* x) const {
        const auto r = buf_rsrc(X, 16L * g.P);
        x[0] = buf_load(r, p, 0);
        x[1] = buf_load(r, p, g.P);
    }
    __device__ void velocity(int p, double* w) const {
        const auto r = buf_rsrc(V, 16L * g.P);
        w[0] = buf_load(r, p, 0);
        w[1] = buf_load(r, p, g.P);
    }
    // vertices (velocities) p and p + 1, which are neighbours along m in one vertex column
    __device__ void vertex2(int p, double* xa, double* xb) const { vec2_pair(X, g.P, p, xa, xb); }
    __device__ void velocity2(int p, double* wa, double* wb) const {
        vec2_pair(V, g.P, p, wa, wb);
    }
    PEA_HD double machv() const { return mach; }
    // direction b of a batch launch (grid y)
    PEA_HD LdCxU at(int b) const {
        LdCxU r = *this;
        r.dU += (long)b * 8 * g.C;
        return r;
    }
};

// Complex step along part `part` (0 real, 1 imaginary) of complex directions:
// value + i h * part(direction).
struct LdCx {
    const double *X, *V, *U;
    const double *dX, *dV, *dU, *dmach;
    double mach, h;
    int part;
    Grid g;
    PEA_HD void state(int c, Cx* u) const {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int e = v * g.C + c;
            u[v] = Cx{U[e], dU ? h * dU[2 * e + part] : 0.0};
        }
    }
    PEA_HD void vec2(const double* A, const double* dA, int p, Cx* x) const {
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const int e = d * g.P + p;
            x[d] = Cx{A[e], dA ? h * dA[2 * e + part] : 0.0};
        }
    }
    PEA_HD void vertex(int p, Cx* x) const { vec2(X, dX, p, x); }
    PEA_HD void velocity(int p, Cx* w) const { vec2(V, dV, p, w); }
    PEA_HD Cx machv() const { return Cx{mach, dmach ? h * dmach[part] : 0.0}; }
    PEA_HD LdCx at(int) const { return *this; }   // not launched as a batch
};

// Stencil source of one cell over a loader.
template <class Ld>
struct CellSrc {
    const Ld& ld;
    int m, n;
    template <class T>
    PEA_HD void get(int slot, T* u) const {
        ld.state(slot_cell(ld.g, m, n, slot), u);
    }
};

template <class T, class Ld>
PEA_HD void load_cell_geometry(const Ld& ld, int m, int n, T (*x)[2], T (*w)[2]) {
#pragma unroll
    for (int vtx = 0; vtx < 4; ++vtx) {
        const int p = vtx_index(ld.g, m, n, vtx);
        ld.vertex(p, x[vtx]);
        ld.velocity(p, w[vtx]);
    }
}

// =============================================================================== primal
__global__ void __launch_bounds__(256)
k_free_stream(pea_params P, double mach, int C, double* __restrict__ U) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    double u[4];
    free_stream(P, mach, u);
#pragma unroll
    for (int v = 0; v < 4; ++v) U[v * C + c] = u[v];
}

// One residual evaluation per thread; T state scalar, G geometry scalar.
template <class T, class G, class Ld, class Store>
__device__ __forceinline__ void residual_cell(const pea_params& P, const Ld& ld, int c,
                                              const Store& store) {
    const int n = c / ld.g.M, m = c - n * ld.g.M;
    G x[4][2], w[4][2];
    T uc[4];
    load_cell_geometry<G>(ld, m, n, x, w);
    ld.state(c, uc);
    CellGeom<G> geo;
    cell_geometry<G>(x, w, geo);
    T R[4];
    cell_ode_src(P, ld.machv(), CellSrc<Ld>{ld, m, n}, uc, geo, boundary_flag(ld.g.M, m), R);
    store(c, R);
}

struct StoreReal {
    double* R;
    int C;
    PEA_HD void operator()(int c, const double* r) const {
#pragma unroll
        for (int v = 0; v < 4; ++v) R[v * C + c] = r[v];
    }
};

// blockIdx.y selects one of several states on the same grid (consecutive 4*C arrays)
__global__ void __launch_bounds__(256)
k_odes(pea_params P, LdReal ld, double* __restrict__ R) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ld.g.C) return;
    const long off = (long)blockIdx.y * 4 * ld.g.C;
    LdReal l = ld;
    l.U = ld.U + off;
    residual_cell<double, double>(P, l, c, StoreReal{R + off, ld.g.C});
}

// ------------------------------------------------------------------ tangent (JVP)
template <int K>
struct StoreDual {
    double *dR, *R;
    int C;
    PEA_HD void operator()(int c, const Dual<K>* r) const {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long e = (long)v * C + c;
            if (R) R[e] = r[v].v;
#pragma unroll
            for (int j = 0; j < K; ++j) dR[2 * ((long)(j >> 1) * 4 * C + e) + (j & 1)] = r[v].d[j];
        }
    }
};

template <int K, class Ld>
__global__ void __launch_bounds__(256)
k_jvp(pea_params P, Ld ld, double* __restrict__ dR, double* __restrict__ R) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ld.g.C) return;
    using G = decltype(ld.machv());
    residual_cell<Dual<K>, G>(P, ld, c, StoreDual<K>{dR, R, ld.g.C});
}

// ------------------------------------------------------------------ face-based forward
// Residual (T = double) or residual + tangent (T = Dual<K>) of a TM x TN tile of cells, one
// thread per cell.  Every face touching the tile is evaluated once (face_flux2_lds): a thread
// owns its cell's f and o faces and hands the other side's flux to the neighbouring cell (a
// lane shuffle within the wave, LDS across waves); the faces the tile shares with its
// neighbouring tiles are evaluated by the first wave.  Per-cell pressures and sound speeds of
// the tile grown by two cells are computed once into LDS; the wall and far-field ghosts sit in
// the grown rows m = -1 and m = M.  Each cell sums its four face fluxes in a fixed order for its
// position in the tile (the reference's b, f, i, o of f90:395-433, a rounding apart).
// blockIdx.z selects one of several states on the same grid (consecutive 4*C arrays).
// A value of scalar type T (double, Dual<K>, Cx: plain aggregates of doubles) stored in LDS as
// structure of arrays: component k of element i at base[k * stride + i].
template <class T>
PEA_HD void lds_put(double* base, int stride, int i, const T& x) {
    constexpr int W = sizeof(T) / sizeof(double);
    const double* p = reinterpret_cast<const double*>(&x);
#pragma unroll
    for (int k = 0; k < W; ++k) base[k * stride + i] = p[k];
}
template <class T>
PEA_HD T lds_get(const double* base, int stride, int i) {
    constexpr int W = sizeof(T) / sizeof(double);
    T x;
    double* p = reinterpret_cast<double*>(&x);
#pragma unroll
    for (int k = 0; k < W; ++k) p[k] = base[k * stride + i];
    return x;
}

// x of the lane `delta` below this one in the wave (this lane's own x for the lowest lanes);
// every double of T moves separately.
template <class T>
__device__ __forceinline__ T shfl_up_t(const T& x, unsigned delta) {
    constexpr int W = sizeof(T) / sizeof(double);
    T r;
    const double* p = reinterpret_cast<const double*>(&x);
    double* q = reinterpret_cast<double*>(&r);
#pragma unroll
    for (int k = 0; k < W; ++k) q[k] = __shfl_up(p[k], delta, 64);
    return r;
}

// One tile of TM x TN cells at tile coordinates (bx, by) of state bz; the body of the
// single-grid and of the multi-grid kernels.
template <class T, class G, int TM, int TN, class Ld, class Store>
__device__ __forceinline__ void face_fwd_tile(const pea_params& P, const FsDir& fs, const Ld& ld,
                                              const Store& store, int bx, int by, int bz) {
    constexpr int NT = TM * TN, BM = TM + 4, BN = TN + 4, BOX = BM * BN;
    static_assert(TM + TN <= NT, "the tile-edge faces take one thread each");
    static_assert(64 % TM == 0 && NT % 64 == 0, "a wave holds whole tile columns");
    constexpr int CPW = 64 / TM, SB = NT / 64 * TM;   // columns per wave; b hand-off slots
    // LDS holds, as structure of arrays of the doubles of T: the states of the tile grown by two
    // cells (with the wall and far-field ghosts in the grown rows m = -1 and m = M), and the
    // fluxes handed to a cell in another wave: the b faces of each wave's first column and the
    // i faces of the tile's first row.  Every stencil read of the face evaluations is an LDS
    // read; the global loads happen once, coalesced, in the staging loop.  Fluxes for a cell in
    // the same wave move between lanes.
    constexpr int W = sizeof(T) / sizeof(double);
    __shared__ T sP[BOX], sC[BOX];
    __shared__ double sU[4 * W * BOX];
    // FP64: 1/rho and the pressure sensors along m and n of each box cell, computed once per
    // cell (face_flux2_cell); the tangent kernels evaluate them per face (face_flux2_lds)
    constexpr bool kCell = std::is_same<T, double>::value;
    constexpr int CB = kCell ? BOX : 1;
    __shared__ double sI[CB], sNm[CB], sNn[CB];
    __shared__ double sFb[4 * W * SB], sFi[4 * W * TN];
    const Grid g = ld.g;
    const Ld l = ld.at(bz);
    const int m0 = bx * TM, n0 = by * TN;
    const int tid = threadIdx.x;
    const int tm = tid % TM, tn = tid / TM;
    const int m = m0 + tm, n = n0 + tn;
    const bool live = m < g.M && n < g.N;
    const auto mach = l.machv();

    auto vtx = [&](int mm, int nn, G* x) { l.vertex(nn * (g.M + 1) + mm, x); };
    auto vel = [&](int mm, int nn, G* w) { l.velocity(nn * (g.M + 1) + mm, w); };
    // normal of the edge from vertex (ma, na) to vertex (mb, nb) and the mean velocity of its
    // two vertices
    auto edge = [&](int ma, int na, int mb, int nb, G* nr, G* gv) {
        G xa[2], xb[2], va[2], vb[2];
        vtx(ma, na, xa);
        vtx(mb, nb, xb);
        vel(ma, na, va);
        vel(mb, nb, vb);
        edge_normal(xa, xb, nr);
        gv[0] = (va[0] + vb[0]) / 2.0;
        gv[1] = (va[1] + vb[1]) / 2.0;
    };
    // the same from vertices and velocities already loaded
    auto edge_of = [](const G* xa, const G* xb, const G* va, const G* vb, G* nr, G* gv) {
        edge_normal(xa, xb, nr);
        gv[0] = (va[0] + vb[0]) / 2.0;
        gv[1] = (va[1] + vb[1]) / 2.0;
    };

    // 0. Plain FP64 geometry: the vertices and velocities of this cell's faces and area, and of
    // the tile-edge face a thread below TM + TN evaluates, are loaded before the staging loop,
    // so their latency overlaps the staging instead of following each face (clamped to the grid
    // for threads without a cell or face).  Geometry with tangents (G = Dual) would hold twice
    // the registers; it is loaded at each face, as vtx/vel calls there.
    constexpr bool kEarlyCell = std::is_same<G, double>::value, kEarlyEdge = kEarlyCell;
    G gx[4][2], gw[3][2];   // vertices BI, FI, FO, BO; velocities at FI, FO, BO
    G ex[2][2], ew[2][2];   // tile-edge face: the edge from vertex ex[0] to vertex ex[1]
    auto load_cell = [&](int mc, int nc) {
        if constexpr (kEarlyCell) {
            // bi, bo and fi, fo are neighbours along m: one 16-byte load per pair and plane
            const int p = nc * (g.M + 1) + mc, q = p + g.M + 1;
            G wbi[2];
            l.vertex2(p, gx[BI], gx[BO]);
            l.vertex2(q, gx[FI], gx[FO]);
            l.velocity2(q, gw[0], gw[1]);
            l.velocity2(p, wbi, gw[2]);
        } else {
            vtx(mc, nc, gx[BI]);
            vtx(mc, nc + 1, gx[FI]);
            vtx(mc + 1, nc + 1, gx[FO]);
            vtx(mc + 1, nc, gx[BO]);
            vel(mc, nc + 1, gw[0]);
            vel(mc + 1, nc + 1, gw[1]);
            vel(mc + 1, nc, gw[2]);
        }
    };
    // the tile-edge face's vertices: threads below TM, the angular face n0-1 | n0 of row
    // m0 + tid (edge bi -> bo); threads TM..TM+TN-1, the wall face of column n0 + tid - TM (edge
    // bi -> fi) or its radial face m0-1 | m0 (edge fo -> bo of the cell below)
    auto load_edge = [&](int mm, int nn) {
        int ma, na, mb, nb;
        if (tid < TM)
            ma = mm, na = n0, mb = mm + 1, nb = n0;
        else if (m0 == 0)
            ma = 0, na = nn, mb = 0, nb = nn + 1;
        else
            ma = m0, na = nn + 1, mb = m0, nb = nn;
        vtx(ma, na, ex[0]);
        vtx(mb, nb, ex[1]);
        vel(ma, na, ew[0]);
        vel(mb, nb, ew[1]);
    };
    if constexpr (kEarlyCell) load_cell(m < g.M ? m : g.M - 1, n < g.N ? n : g.N - 1);
    if constexpr (kEarlyEdge) {
        if (tid < TM + TN)
            load_edge(m0 + tid < g.M ? m0 + tid : g.M - 1,
                      n0 + tid - TM < g.N ? n0 + tid - TM : g.N - 1);
    }

    // 1. pressures, sound speeds and ghosts over the grown tile.  Box cell i holds the state of
    // cell (cm, cn), or a ghost made from the state of the first or last row; its state is
    // loaded from the row clamped to the grid.
    auto stage_load = [&](int i, T* u) {
        const int bm = i % BM, bn = i / BM;
        const int cm = m0 - 2 + bm;
        const int cn = wrap_any(n0 - 2 + bn, g.N);
        l.state(cn * g.M + (cm < 0 ? 0 : (cm >= g.M ? g.M - 1 : cm)), u);
    };
    auto stage_cell = [&](int i, const T* u) {
        const int bm = i % BM, bn = i / BM;
        const int cm = m0 - 2 + bm;
        const int cn = wrap_any(n0 - 2 + bn, g.N);
        if (cm >= 0 && cm < g.M) {
            cell_pc(u, sP[i], sC[i]);
            if constexpr (kCell) sI[i] = rcp_fast(val(u[0]));
#pragma unroll
            for (int v = 0; v < 4; ++v) lds_put(sU + v * W * BOX, BOX, i, u[v]);
        } else if (cm == -1 || cm == g.M) {
            T gh[4];
            if (cm < 0) {
                G nr[2], gv[2];
                edge(0, cn, 0, cn + 1, nr, gv);       // wall face bi -> fi
                wall_ghost(u, nr, gv, gh);
            } else {
                decltype(l.machv()) inf[4];
                free_stream_dir(mach, fs.ca, fs.sa, inf);
#pragma unroll
                for (int v = 0; v < 4; ++v) gh[v] = 2.0 * inf[v] - u[v];
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) lds_put(sU + v * W * BOX, BOX, i, gh[v]);
            sP[i] = pressure_q(gh);
        }
    };
    // Every pass's loads are issued before the first is used (one round trip instead of one per
    // pass; passes past the box load a valid cell and discard it).
    {
        constexpr int NPASS = (BOX + NT - 1) / NT;
        T u[NPASS][4];
#pragma unroll
        for (int k = 0; k < NPASS; ++k)
            stage_load(tid + k * NT < BOX ? tid + k * NT : BOX - 1, u[k]);
#pragma unroll
        for (int k = 0; k < NPASS; ++k)
            if (tid + k * NT < BOX) stage_cell(tid + k * NT, u[k]);
    }
    __syncthreads();
    if constexpr (kCell) {
        // sensors of the cells the tile's faces read as L or R (box interior; rows outside
        // [-1, M] hold no pressure and their sensors are never read)
        for (int i = tid; i < BOX; i += NT) {
            const int bm = i % BM, bn = i / BM;
            if (bm < 1 || bm > BM - 2 || bn < 1 || bn > BN - 2) continue;
            const double p = val(sP[i]);
            sNm[i] = sensor(val(sP[i - 1]), p, val(sP[i + 1]));
            sNn[i] = sensor(val(sP[i - BM]), p, val(sP[i + BM]));
        }
        __syncthreads();
    }

    auto bidx = [&](int mm, int nb) { return nb * BM + (mm - m0 + 2); };   // nb: box column
    // state of cell (mm, box column nb); mm = -1 and mm = M are the ghosts
    auto box = [&](int mm, int nb, T* u) {
        const int b = bidx(mm, nb);
#pragma unroll
        for (int v = 0; v < 4; ++v) u[v] = lds_get<T>(sU + v * W * BOX, BOX, b);
    };
    // stencil accessor over four box cells (slots ll, l, r, rr)
    auto stencil = [&](int bLL, int bL, int bR, int bRR) {
        const double* u = sU;
        return [u, bLL, bL, bR, bRR](int slot, int v) -> T {
            const int b = slot == 0 ? bLL : slot == 1 ? bL : slot == 2 ? bR : bRR;
            return lds_get<T>(u + v * W * BOX, BOX, b);
        };
    };
    // both fluxes of the face L | R (box cells bLL, bL, bR, bRR along one grid line; nu: the
    // sensors along that line)
    auto flux2 = [&](int bLL, int bL, int bR, int bRR, const double* nu, const G* nr,
                     const G* nr2, const G* gv, bool dL, bool dR, auto outL, auto outR) {
        if constexpr (kCell)
            face_flux2_cell(P, stencil(bLL, bL, bR, bRR), sP[bL], sP[bR], sC[bL], sC[bR], sI[bL],
                            sI[bR], nu[bL], nu[bR], nr, gv, dL, dR, outL, outR);
        else
            face_flux2_lds(P, stencil(bLL, bL, bR, bRR), sP[bLL], sP[bL], sP[bR], sP[bRR],
                           sC[bL], sC[bR], nr, nr2, gv, dL, dR, outL, outR);
    };
    T S[4];       // running sum of this cell's four face fluxes
    T fb[4], fo[4];   // fluxes of this cell's f and o faces for the cells at n + 1 and m + 1
    const int bn = tn + 2;
    const int cw = tn % CPW;   // column of this cell within its wave
    // ---- angular face n | n+1 (this cell's f face)
    if (live) {
        const int bc = bidx(m, bn);
        G nr[2], nr2[2], gv[2];
        if constexpr (!kEarlyCell) {
            vtx(m, n + 1, gx[FI]);
            vtx(m + 1, n + 1, gx[FO]);
            vel(m, n + 1, gw[0]);
            vel(m + 1, n + 1, gw[1]);
        }
        edge_of(gx[FI], gx[FO], gw[0], gw[1], nr, gv);   // fi -> fo
        edge_normal(gx[FO], gx[FI], nr2);                // R's bo -> bi
        const int bLL = bidx(m, bn - 1), bR = bidx(m, bn + 1), bRR = bidx(m, bn + 2);
        flux2(bLL, bc, bR, bRR, sNn, nr, nr2, gv, true, true,
              [&](int v, const T& f) { S[v] = f; }, [&](int v, const T& f) { fb[v] = f; });
        // the cell at n + 1 starts the next wave: hand over through LDS
        if (cw == CPW - 1 && tn + 1 < TN && n + 1 < g.N)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                lds_put(sFb + v * W * SB, SB, (tn + 1) / CPW * TM + tm, fb[v]);
    }
    // within the wave the cell at n + 1 is TM lanes up
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const T f = shfl_up_t(fb[v], TM);
        if (live && cw > 0) S[v] = S[v] + f;
    }
    // ---- radial face m | m+1 (this cell's o face), or the far-field face
    if (live) {
        const int bc = bidx(m, bn);
        if constexpr (!kEarlyCell) {
            vtx(m + 1, n + 1, gx[FO]);
            vtx(m + 1, n, gx[BO]);
            vel(m + 1, n + 1, gw[1]);
            vel(m + 1, n, gw[2]);
        }
        if (m < g.M - 1) {
            G nr[2], nr2[2], gv[2];
            edge_of(gx[FO], gx[BO], gw[1], gw[2], nr, gv);   // fo -> bo
            edge_normal(gx[BO], gx[FO], nr2);                // R's bi -> fi
            // m - 1 is the wall ghost row for m = 0, m + 2 the far-field row for m + 1 = M - 1
            const int bLL = bidx(m - 1, bn), bR = bidx(m + 1, bn), bRR = bidx(m + 2, bn);
            flux2(bLL, bc, bR, bRR, sNm, nr, nr2, gv, m != g.M - 2, m + 1 != 1,
                  [&](int v, const T& f) { S[v] = S[v] + f; },
                  [&](int v, const T& f) { fo[v] = f; });
        } else {
            T uc[4], gh[4], Fo[4];
            box(m, bn, uc);
            box(g.M, bn, gh);
            G nr[2], gv[2];
            edge_of(gx[FO], gx[BO], gw[1], gw[2], nr, gv);
            face_flux(P, uc, uc, gh, gh, nr, gv, false, Fo);
#pragma unroll
            for (int v = 0; v < 4; ++v) S[v] = S[v] + Fo[v];
        }
    }
    // the cell at m + 1 is the next lane
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const T f = shfl_up_t(fo[v], 1);
        if (live && tm > 0) S[v] = S[v] + f;
    }
    // ---- faces shared with the neighbouring tiles: the angular faces n0-1 | n0 of the tile's
    // rows (threads 0..TM-1) and the radial faces m0-1 | m0 of its columns (threads TM..TM+TN-1).
    // They take the first wave(s) only: a branch taken by one or two lanes of every wave would
    // cost every wave a full face evaluation.  Their results reach the cells through LDS.
    if (tid < TM) {   // angular face n0-1 | n0; normal L -> R is minus cell (mm, n0)'s b normal
        const int mm = m0 + tid;
        if (mm < g.M) {
            if constexpr (!kEarlyEdge) load_edge(mm, n0);
            G nr[2], nr2[2], gv[2];
            edge_of(ex[0], ex[1], ew[0], ew[1], nr, gv);   // bi -> bo
            edge_normal(ex[1], ex[0], nr2);                // the cell's bo -> bi
            const int bLL = bidx(mm, 0), bL = bidx(mm, 1), bR = bidx(mm, 2), bRR = bidx(mm, 3);
            flux2(bLL, bL, bR, bRR, sNn, nr, nr2, gv, false, true, [](int, const T&) {},
                  [&](int v, const T& f) { lds_put(sFb + v * W * SB, SB, tid, f); });
        }
    } else if (tid < TM + TN) {
        const int j = tid - TM, mm = m0, nn = n0 + j, bnn = j + 2;
        if (nn < g.N) {
            if constexpr (!kEarlyEdge) load_edge(mm, nn);
            if (mm == 0) {   // wall face: central against the wall ghost
                T uc[4], gh[4], Fi[4];
                box(mm, bnn, uc);
                box(-1, bnn, gh);
                G nr[2], gv[2];
                edge_of(ex[0], ex[1], ew[0], ew[1], nr, gv);   // bi -> fi
                face_flux(P, uc, uc, gh, gh, nr, gv, false, Fi);
#pragma unroll
                for (int v = 0; v < 4; ++v) lds_put(sFi + v * W * TN, TN, j, Fi[v]);
            } else {         // radial face m0-1 | m0; normal L -> R is minus the cell's i normal
                G nr[2], nr2[2], gv[2];
                edge_of(ex[0], ex[1], ew[0], ew[1], nr, gv);   // fo -> bo of L
                edge_normal(ex[1], ex[0], nr2);                // the cell's bi -> fi
                // mm - 2 is the wall ghost row for mm = 1, mm + 1 the far-field row for
                // mm = M - 1
                const int bLL = bidx(mm - 2, bnn), bL = bidx(mm - 1, bnn), bR = bidx(mm, bnn),
                          bRR = bidx(mm + 1, bnn);
                flux2(bLL, bL, bR, bRR, sNm, nr, nr2, gv, false, mm != 1, [](int, const T&) {},
                      [&](int v, const T& f) { lds_put(sFi + v * W * TN, TN, j, f); });
            }
        }
    }
    __syncthreads();
    if (live) {
        // faces handed over through LDS: the b face of a wave's first column, the i face of
        // the tile's first row
        if (cw == 0)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                S[v] = S[v] + lds_get<T>(sFb + v * W * SB, SB, tn / CPW * TM + tm);
        if (tm == 0)
#pragma unroll
            for (int v = 0; v < 4; ++v) S[v] = S[v] + lds_get<T>(sFi + v * W * TN, TN, tn);
        if constexpr (!kEarlyCell) load_cell(m, n);
        const G d1x = gx[FO][0] - gx[BI][0], d1z = gx[FO][1] - gx[BI][1];
        const G d2x = gx[BO][0] - gx[FI][0], d2z = gx[BO][1] - gx[FI][1];
        const G ia = recip(sabs(d1x * d2z - d1z * d2x) / 2.0);
        T R[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) R[v] = -(S[v] * ia);
        store.put(bz, n * g.M + m, R);
    }
}

template <class T, class G, int TM, int TN, class Ld, class Store>
__global__ void __launch_bounds__(TM * TN)
k_face_fwd(pea_params P, FsDir fs, Ld ld, Store store) {
    const TileId t = batch_tile(gridDim.x, gridDim.y, gridDim.z);
    face_fwd_tile<T, G, TM, TN>(P, fs, ld, store, t.bx, t.by, t.bz);
}

// ------------------------------------------------------------------ many directions at one state
// States tangent of many real directions at one state (pea_jvp_odes_states_dirs).  One workgroup
// per tile, as face_fwd_tile, but the primal part is done once per tile: the states, pressures and
// 1/rho of the grown tile go to LDS, and every face the tile evaluates gets its FaceLin record
// (pea_linear.hpp): in registers for the f and o faces a thread owns, in LDS for the tile-edge
// faces of the first wave.  Then the workgroup walks complex directions d0 .. d1-1: it stages a
// direction's states of the grown tile (real and imaginary parts) and their pressure tangents in
// LDS, applies each face's record to both parts, hands the other side's flux to the neighbouring
// cell as face_fwd_tile does, and stores -(sum of the four face fluxes)/area.  Every global
// access is a buffer access bounded by its array.
__device__ __forceinline__ void lin_put(double* base, int stride, int i, const FaceLin& f) {
    const double w[kLinWords] = {f.n0,   f.n1,   f.hvn,  f.L.lam, f.L.s, f.L.ib, f.L.aC,
                                 f.L.bC, f.R.lam, f.R.s, f.R.ib, f.R.aC, f.R.bC,
                                 __builtin_bit_cast(double, (unsigned long long)f.bits)};
#pragma unroll
    for (int k = 0; k < kLinWords; ++k) base[k * stride + i] = w[k];
}
__device__ __forceinline__ FaceLin lin_get(const double* base, int stride, int i) {
    double w[kLinWords];
#pragma unroll
    for (int k = 0; k < kLinWords; ++k) w[k] = base[k * stride + i];
    FaceLin f;
    f.n0 = w[0], f.n1 = w[1], f.hvn = w[2];
    f.L = SideLin{w[3], w[4], w[5], w[6], w[7]};
    f.R = SideLin{w[8], w[9], w[10], w[11], w[12]};
    f.bits = (unsigned)__builtin_bit_cast(unsigned long long, w[13]);
    return f;
}
typedef unsigned int pea_u32x4 __attribute__((ext_vector_type(4)));
// complex element `elem` of plane `plane_elems` (in complex elements) of an interleaved array
__device__ __forceinline__ void buf_store_cx(__amdgpu_buffer_rsrc_t r, double2 x, int elem,
                                             int plane_elems) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(pea_u32x4, x), r, elem * 16,
                                           plane_elems * 16, 0);
}

template <int TM, int TN>
__device__ __forceinline__ void face_dirs_tile(const pea_params& P, const FsDir& fs,
                                               const LdReal& l, const double* dU, double* dR,
                                               int ndir, int d0, int d1, int bx, int by) {
    constexpr int NT = TM * TN, BM = TM + 4, BN = TN + 4, BOX = BM * BN;
    static_assert(TM + TN <= NT, "the tile-edge faces take one thread each");
    static_assert(64 % TM == 0 && NT % 64 == 0, "a wave holds whole tile columns");
    constexpr int CPW = 64 / TM, SB = NT / 64 * TM, NPASS = (BOX + NT - 1) / NT;
    constexpr int NE = TM + TN;
    // LDS: primal states, pressures and 1/rho of the grown tile; the direction's states and
    // pressure tangents (sdP holds the primal sound speeds while the records are built); the
    // tile-edge records; the wall-ghost map of each box column; the flux hand-offs.
    __shared__ double sU[4 * BOX], sP[BOX], sI[BOX];
    __shared__ double sdU[2 * 4 * BOX], sdP[2 * BOX];
    __shared__ double sRec[kLinWords * NE];
    __shared__ double sW[5 * BN];
    __shared__ double sFb[2 * 4 * SB], sFi[2 * 4 * TN];
    double* const sC = sdP;
    const Grid g = l.g;
    const int m0 = bx * TM, n0 = by * TN;
    const int tid = threadIdx.x;
    const int tm = tid % TM, tn = tid / TM;
    const int m = m0 + tm, n = n0 + tn;
    const bool live = m < g.M && n < g.N;

    auto vtx = [&](int mm, int nn, double* x) { l.vertex(nn * (g.M + 1) + mm, x); };
    auto vel = [&](int mm, int nn, double* w) { l.velocity(nn * (g.M + 1) + mm, w); };
    auto edge_of = [](const double* xa, const double* xb, const double* va, const double* vb,
                      double* nr, double* gv) {
        edge_normal(xa, xb, nr);
        gv[0] = (va[0] + vb[0]) / 2.0;
        gv[1] = (va[1] + vb[1]) / 2.0;
    };
    // geometry of this cell's faces and area, and of the tile-edge face of threads below NE
    double gx[4][2], gw[3][2], ex[2][2], ew[2][2];
    {
        // bi, bo and fi, fo are neighbours along m: one 16-byte load per pair and plane
        const int mc = m < g.M ? m : g.M - 1, nc = n < g.N ? n : g.N - 1;
        const int p = nc * (g.M + 1) + mc, q = p + g.M + 1;
        double wbi[2];
        l.vertex2(p, gx[BI], gx[BO]);
        l.vertex2(q, gx[FI], gx[FO]);
        l.velocity2(q, gw[0], gw[1]);
        l.velocity2(p, wbi, gw[2]);
    }
    if (tid < NE) {
        const int mm = m0 + tid < g.M ? m0 + tid : g.M - 1;
        const int nn = n0 + tid - TM < g.N ? n0 + tid - TM : g.N - 1;
        int ma, na, mb, nb;
        if (tid < TM)
            ma = mm, na = n0, mb = mm + 1, nb = n0;
        else if (m0 == 0)
            ma = 0, na = nn, mb = 0, nb = nn + 1;
        else
            ma = m0, na = nn + 1, mb = m0, nb = nn;
        vtx(ma, na, ex[0]);
        vtx(mb, nb, ex[1]);
        vel(ma, na, ew[0]);
        vel(mb, nb, ew[1]);
    }

    // 1. primal staging.  Each thread keeps, for the box cells it stages, the coefficients of
    // the pressure tangent dp = (gamma-1)(dE - cu dj_x - cv dj_y + ck drho) (pressure_q's) and
    // what the cell is: 0 interior, 1 wall ghost, 2 far-field ghost, 3 outside the stencils.
    double cu[NPASS], cv[NPASS], ck[NPASS];
    int kind[NPASS], cell[NPASS];
    auto pcoef = [](const double* u, double& a, double& b, double& c) {
        const double rb = rcp_fast(u[0]);
        const double r = (u[1] * u[1] + u[2] * u[2]) * rb;
        a = u[1] * rb;
        b = u[2] * rb;
        c = 0.5 * r * rb;
    };
#pragma unroll
    for (int k = 0; k < NPASS; ++k) {
        const int i = tid + k * NT;
        kind[k] = 3;
        cell[k] = 0;
        cu[k] = cv[k] = ck[k] = 0.0;
        if (i >= BOX) continue;
        const int bm = i % BM, bn = i / BM;
        const int cm = m0 - 2 + bm;
        const int cn = wrap_any(n0 - 2 + bn, g.N);
        cell[k] = cn * g.M + (cm < 0 ? 0 : (cm >= g.M ? g.M - 1 : cm));
        double u[4];
        l.state(cell[k], u);
        double s[4] = {0.0, 0.0, 0.0, 0.0}, p = 0.0, c = 0.0;
        if (cm >= 0 && cm < g.M) {
            kind[k] = 0;
            cell_pc(u, p, c);
#pragma unroll
            for (int v = 0; v < 4; ++v) s[v] = u[v];
        } else if (cm == -1) {
            kind[k] = 1;
            double xa[2], xb[2], va[2], vb[2], nr[2], gv[2];
            vtx(0, cn, xa);
            vtx(0, cn + 1, xb);
            vel(0, cn, va);
            vel(0, cn + 1, vb);
            edge_of(xa, xb, va, vb, nr, gv);   // wall face bi -> fi
            wall_ghost(u, nr, gv, s);
            p = pressure_q(s);
            // the ghost is linear in the state: dg_{1,2} = dj - 2 n (n.(dj - gv drho)) / |n|^2
            sW[bn] = nr[0];
            sW[BN + bn] = nr[1];
            sW[2 * BN + bn] = gv[0];
            sW[3 * BN + bn] = gv[1];
            sW[4 * BN + bn] = rcp_fast(nr[0] * nr[0] + nr[1] * nr[1]);
        } else if (cm == g.M) {
            kind[k] = 2;
            double inf[4];
            free_stream_dir(l.machv(), fs.ca, fs.sa, inf);
#pragma unroll
            for (int v = 0; v < 4; ++v) s[v] = 2.0 * inf[v] - u[v];
            p = pressure_q(s);
        }
        if (kind[k] < 3) pcoef(s, cu[k], cv[k], ck[k]);
#pragma unroll
        for (int v = 0; v < 4; ++v) sU[v * BOX + i] = s[v];
        sP[i] = p;
        sC[i] = c;
        sI[i] = kind[k] < 3 ? rcp_fast(s[0]) : 0.0;
    }
    __syncthreads();

    auto bidx = [&](int mm, int nb) { return nb * BM + (mm - m0 + 2); };   // nb: box column
    auto prim = [&](int bLL, int bL, int bR, int bRR) {
        return [bLL, bL, bR, bRR](int slot, int v) -> double {
            const int b = slot == 0 ? bLL : slot == 1 ? bL : slot == 2 ? bR : bRR;
            return sU[v * BOX + b];
        };
    };
    auto build = [&](int bLL, int bL, int bR, int bRR, const double* nr, const double* nr2,
                     const double* gv, bool dL, bool dRs, FaceLin& f) {
        face_lin_build(P, prim(bLL, bL, bR, bRR), sP[bLL], sP[bL], sP[bR], sP[bRR], sC[bL],
                       sC[bR], sI[bL], sI[bR], nr, nr2, gv, dL, dRs, f);
    };

    // 2. records.  Stencil box indices of this thread's f face (angular n | n+1) and o face
    // (radial m | m+1, or the far-field face of m = M-1: central against the ghost row M).
    const int bn = tn + 2;
    const int bc = bidx(m, bn);
    const int fLL = bc - BM, fR = bc + BM, fRR = bc + 2 * BM;
    const bool far = m == g.M - 1;
    const int oLL = far ? bc : bc - 1, oR = bc + 1, oRR = far ? bc + 1 : bc + 2;
    FaceLin rf, ro;
    double ia = 0.0;
    if (live) {
        double nr[2], nr2[2], gv[2];
        edge_of(gx[FI], gx[FO], gw[0], gw[1], nr, gv);   // fi -> fo
        edge_normal(gx[FO], gx[FI], nr2);                // R's bo -> bi
        build(fLL, bc, fR, fRR, nr, nr2, gv, true, true, rf);
        edge_of(gx[FO], gx[BO], gw[1], gw[2], nr, gv);   // fo -> bo
        edge_normal(gx[BO], gx[FO], nr2);                // R's bi -> fi
        build(oLL, bc, oR, oRR, nr, nr2, gv, !far && m != g.M - 2, !far && m + 1 != 1, ro);
        const double d1x = gx[FO][0] - gx[BI][0], d1z = gx[FO][1] - gx[BI][1];
        const double d2x = gx[BO][0] - gx[FI][0], d2z = gx[BO][1] - gx[FI][1];
        ia = recip(sabs(d1x * d2z - d1z * d2x) / 2.0);
    } else {
        const double z[2] = {0.0, 0.0};
        build(bc, bc, bc, bc, z, z, z, false, false, rf);
        ro = rf;
    }
    // tile-edge faces: threads below TM take the angular face n0-1 | n0 of row m0 + tid,
    // threads TM .. NE-1 the radial face m0-1 | m0 of column n0 + tid - TM (the wall face for
    // m0 = 0: central against the wall ghost, its flux is the cell's own)
    int eLL = 0, eL = 0, eR = 0, eRR = 0;
    bool e_live = false, e_own = false;
    if (tid < TM) {
        const int mm = m0 + tid;
        e_live = mm < g.M;
        eLL = bidx(mm, 0), eL = bidx(mm, 1), eR = bidx(mm, 2), eRR = bidx(mm, 3);
        if (e_live) {
            FaceLin fe;
            double nr[2], nr2[2], gv[2];
            edge_of(ex[0], ex[1], ew[0], ew[1], nr, gv);   // bi -> bo
            edge_normal(ex[1], ex[0], nr2);                // the cell's bo -> bi
            build(eLL, eL, eR, eRR, nr, nr2, gv, false, true, fe);
            lin_put(sRec, NE, tid, fe);
        }
    } else if (tid < NE) {
        const int j = tid - TM, mm = m0, bnn = j + 2;
        e_live = n0 + j < g.N;
        if (m0 == 0) {
            e_own = true;
            eLL = eL = bidx(0, bnn);
            eR = eRR = bidx(-1, bnn);
        } else {
            eLL = bidx(mm - 2, bnn), eL = bidx(mm - 1, bnn), eR = bidx(mm, bnn),
            eRR = bidx(mm + 1, bnn);
        }
        if (e_live) {
            FaceLin fe;
            double nr[2], nr2[2], gv[2];
            edge_of(ex[0], ex[1], ew[0], ew[1], nr, gv);   // bi -> fi (wall) or fo -> bo of L
            edge_normal(ex[1], ex[0], nr2);
            build(eLL, eL, eR, eRR, nr, nr2, gv, false, !e_own && mm != 1, fe);
            lin_put(sRec, NE, tid, fe);
        }
    }
    __syncthreads();

    // 3. directions.  Complex direction d is staged whole, its real part in the first half of
    // sdU / sdP and its imaginary part in the second (one 16-byte load per element, at byte
    // 16 ((4 d + v) C + cell)); every face is applied to the real part, then to the imaginary
    // part, and the product is stored as whole complex elements.
    const auto rU = buf_rsrc(dU, 64L * g.C * ndir);
    const auto rR = buf_rsrc(dR, 64L * g.C * ndir);
    const int cme = n * g.M + m;
    const int cw = tn % CPW;
    for (int d = d0; d < d1; ++d) {
        // every pass's loads are issued before the first is used
        double2 zz[NPASS][4];
#pragma unroll
        for (int k = 0; k < NPASS; ++k)
#pragma unroll
            for (int v = 0; v < 4; ++v) zz[k][v] = buf_load_cx(rU, cell[k], (4 * d + v) * g.C);
#pragma unroll
        for (int k = 0; k < NPASS; ++k) {
            const int i = tid + k * NT;
            if (i >= BOX) continue;
            const double2* z = zz[k];
#pragma unroll
            for (int part = 0; part < 2; ++part) {
                double du[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) du[v] = part ? z[v].y : z[v].x;
                if (kind[k] == 1) {
                    const int bnn = i / BM;
                    const double w0 = sW[bnn], w1 = sW[BN + bnn];
                    const double sg = w0 * (du[1] - sW[2 * BN + bnn] * du[0]) +
                                      w1 * (du[2] - sW[3 * BN + bnn] * du[0]);
                    const double tt = 2.0 * sg * sW[4 * BN + bnn];
                    du[1] = du[1] - w0 * tt;
                    du[2] = du[2] - w1 * tt;
                } else if (kind[k] == 2) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) du[v] = -du[v];
                } else if (kind[k] == 3) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) du[v] = 0.0;
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) sdU[(part * 4 + v) * BOX + i] = du[v];
                sdP[part * BOX + i] =
                    (kGamma - 1.0) * (du[3] - cu[k] * du[1] - cv[k] * du[2] + ck[k] * du[0]);
            }
        }
        __syncthreads();

        double S2[2][4];
#pragma unroll
        for (int part = 0; part < 2; ++part) {
            const double* dUp = sdU + part * 4 * BOX;
            const double* dPp = sdP + part * BOX;
            double* Fb = sFb + part * 4 * SB;
            double* Fi = sFi + part * 4 * TN;
            auto dirs = [dUp](int bLL, int bL, int bR, int bRR) {
                return [dUp, bLL, bL, bR, bRR](int slot, int v) -> double {
                    const int b = slot == 0 ? bLL : slot == 1 ? bL : slot == 2 ? bR : bRR;
                    return dUp[v * BOX + b];
                };
            };
            auto apply = [&](const FaceLin& f, int bLL, int bL, int bR, int bRR, auto outL,
                             auto outR) {
                const double dp[4] = {dPp[bLL], dPp[bL], dPp[bR], dPp[bRR]};
                face_lin_apply(P, f, prim(bLL, bL, bR, bRR), sP[bL], sP[bR], sI[bL], sI[bR],
                               dirs(bLL, bL, bR, bRR), dp, outL, outR);
            };
            double* S = S2[part];
            double fb[4] = {0.0, 0.0, 0.0, 0.0}, fo[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int v = 0; v < 4; ++v) S[v] = 0.0;
            if (live) {
                apply(rf, fLL, bc, fR, fRR, [&](int v, double x) { S[v] = x; },
                      [&](int v, double x) { fb[v] = x; });
                if (cw == CPW - 1 && tn + 1 < TN && n + 1 < g.N)
#pragma unroll
                    for (int v = 0; v < 4; ++v) Fb[v * SB + (tn + 1) / CPW * TM + tm] = fb[v];
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double x = __shfl_up(fb[v], TM, 64);
                if (live && cw > 0) S[v] = S[v] + x;
            }
            if (live)
                apply(ro, oLL, bc, oR, oRR, [&](int v, double x) { S[v] = S[v] + x; },
                      [&](int v, double x) { fo[v] = x; });
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double x = __shfl_up(fo[v], 1, 64);
                if (live && tm > 0) S[v] = S[v] + x;
            }
            if (tid < NE && e_live) {
                const FaceLin fe = lin_get(sRec, NE, tid);
                double* dst = tid < TM ? Fb + tid : Fi + (tid - TM);
                const int stride = tid < TM ? SB : TN;
                if (e_own)
                    apply(fe, eLL, eL, eR, eRR, [&](int v, double x) { dst[v * stride] = x; },
                          [](int, double) {});
                else
                    apply(fe, eLL, eL, eR, eRR, [](int, double) {},
                          [&](int v, double x) { dst[v * stride] = x; });
            }
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double2 r;
                double* rp = reinterpret_cast<double*>(&r);
#pragma unroll
                for (int part = 0; part < 2; ++part) {
                    double x = S2[part][v];
                    if (cw == 0) x = x + sFb[(part * 4 + v) * SB + tn / CPW * TM + tm];
                    if (tm == 0) x = x + sFi[(part * 4 + v) * TN + tn];
                    rp[part] = -(x * ia);
                }
                buf_store_cx(rR, r, cme, (4 * d + v) * g.C);
            }
        }
    }
}

// Complex directions [d0, d1) of item z: chunk z of the ndir directions, `per` each.
template <int TM, int TN>
__global__ void __launch_bounds__(TM * TN)
k_face_dirs(pea_params P, FsDir fs, LdReal ld, const double* __restrict__ dU,
            double* __restrict__ dR, int ndir, int per) {
    const TileId t = batch_tile(gridDim.x, gridDim.y, gridDim.z);
    const int d0 = t.bz * per, d1 = min(ndir, d0 + per);
    face_dirs_tile<TM, TN>(P, fs, ld, dU, dR, ndir, d0, d1, t.bx, t.by);
}

struct StoreRealB {
    double* R;
    int C;
    PEA_HD void put(int b, int c, const double* r) const {
#pragma unroll
        for (int v = 0; v < 4; ++v) R[(long)b * 4 * C + (long)v * C + c] = r[v];
    }
};

template <int K>
struct StoreDualB {
    double *dR, *R;
    int C;
    // item b of a batch writes the b-th group of K/2 complex outputs; R (the primal, the same
    // for every item) only from item 0
    PEA_HD void put(int b, int c, const Dual<K>* r) const {
        double* out = dR + (long)b * (K / 2) * 8 * C;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long e = (long)v * C + c;
            if (R && b == 0) R[e] = r[v].v;
#pragma unroll
            for (int j = 0; j < K; ++j) out[2 * ((long)(j >> 1) * 4 * C + e) + (j & 1)] = r[v].d[j];
        }
    }
};

// ------------------------------------------------------------------ complex step
struct StoreCs {
    double* dR;
    double h;
    int C, part;
    PEA_HD void operator()(int c, const Cx* r) const {
#pragma unroll
        for (int v = 0; v < 4; ++v) dR[2 * ((long)v * C + c) + part] = r[v].i / h;
    }
};

template <class Ld>
__global__ void __launch_bounds__(256)
k_cs_jvp(pea_params P, Ld ld, double* __restrict__ dR) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ld.g.C) return;
    using G = decltype(ld.machv());
    double* out = dR + (long)blockIdx.y * 8 * ld.g.C;   // direction y of a batch
    Ld l0 = ld.at(blockIdx.y);
    l0.part = 0;
    residual_cell<Cx, G>(P, l0, c, StoreCs{out, ld.h, ld.g.C, 0});
    Ld l1 = ld.at(blockIdx.y);
    l1.part = 1;
    residual_cell<Cx, G>(P, l1, c, StoreCs{out, ld.h, ld.g.C, 1});
}

// Complex-step block assembly (the reference's linearize): one thread per cell, one complex
// residual evaluation per perturbed input component.
struct PertSrc {
    const LdReal& ld;
    int m, n, pslot, pvar;
    double h;
    PEA_HD void get(int slot, Cx* u) const {
        double r[4];
        ld.state(slot_cell(ld.g, m, n, slot), r);
#pragma unroll
        for (int v = 0; v < 4; ++v) u[v] = Cx{r[v], (slot == pslot && v == pvar) ? h : 0.0};
    }
};

__global__ void __launch_bounds__(128)
k_cs_linearize(pea_params P, LdReal ld, double* __restrict__ JU, double* __restrict__ JX,
               double* __restrict__ JV, double* __restrict__ JMa) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const Grid g = ld.g;
    if (c >= g.C) return;
    const int n = c / g.M, m = c - n * g.M;
    const int flag = boundary_flag(g.M, m);
    const double h = kCsStep;
    double xr[4][2], wr[4][2], ur[4];
    load_cell_geometry<double>(ld, m, n, xr, wr);
    ld.state(c, ur);
    Cx x[4][2], w[4][2], uc[4];
    for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 2; ++d) {
            x[a][d] = Cx{xr[a][d], 0.0};
            w[a][d] = Cx{wr[a][d], 0.0};
        }
    const Cx mach = Cx{ld.mach, 0.0};
    if (JU) {
        const bool present[9] = {true, true, true, true, true, m > 0, m > 1, m < g.M - 1,
                                 m < g.M - 2};
        for (int s = 0; s < 9; ++s)
            for (int j = 0; j < 4; ++j) {
                double col[4] = {0.0, 0.0, 0.0, 0.0};
                if (present[s]) {
                    for (int v = 0; v < 4; ++v) uc[v] = Cx{ur[v], (s == C && v == j) ? h : 0.0};
                    CellGeom<Cx> geo;
                    cell_geometry<Cx>(x, w, geo);
                    Cx R[4];
                    cell_ode_src(P, mach, PertSrc{ld, m, n, s, j, h}, uc, geo, flag, R);
                    for (int i = 0; i < 4; ++i) col[i] = R[i].i / h;
                }
                for (int i = 0; i < 4; ++i) JU[(long)((s * 4 + i) * 4 + j) * g.C + c] = col[i];
            }
    }
    for (int v = 0; v < 4; ++v) uc[v] = Cx{ur[v], 0.0};
    for (int which = 0; which < 2; ++which) {
        double* J = which ? JV : JX;
        if (!J) continue;
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 2; ++d) {
                Cx (*pt)[2] = which ? w : x;
                pt[a][d].i = h;
                CellGeom<Cx> geo;
                cell_geometry<Cx>(x, w, geo);
                Cx R[4];
                cell_ode_src(P, mach, PertSrc{ld, m, n, -1, -1, h}, uc, geo, flag, R);
                pt[a][d].i = 0.0;
                for (int i = 0; i < 4; ++i) J[(long)((a * 4 + i) * 2 + d) * g.C + c] = R[i].i / h;
            }
    }
    if (JMa) {
        double col[4] = {0.0, 0.0, 0.0, 0.0};
        if (flag == 5) {
            CellGeom<Cx> geo;
            cell_geometry<Cx>(x, w, geo);
            Cx R[4];
            cell_ode_src(P, Cx{ld.mach, h}, PertSrc{ld, m, n, -1, -1, h}, uc, geo, flag, R);
            for (int i = 0; i < 4; ++i) col[i] = R[i].i / h;
        }
        for (int i = 0; i < 4; ++i) JMa[(long)i * g.C + c] = col[i];
    }
}

// ------------------------------------------------------------------ block products
// neighbour cell of block s of cell (m, n), -1 if absent
PEA_HD int block_cell(const Grid& g, int m, int n, int s) {
    switch (s) {
    case I: return m > 0 ? slot_cell(g, m, n, s) : -1;
    case II: return m > 1 ? slot_cell(g, m, n, s) : -1;
    case O: return m < g.M - 1 ? slot_cell(g, m, n, s) : -1;
    case OO: return m < g.M - 2 ? slot_cell(g, m, n, s) : -1;
    default: return slot_cell(g, m, n, s);
    }
}

__global__ void __launch_bounds__(256)
k_apply_states_fwd(Grid g, const double* __restrict__ J, const double* __restrict__ d,
                   double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    const int n = c / g.M, m = c - n * g.M;
    const int order[9] = {C, B, BB, F, FF, II, I, OO, O};   // f90:856-861
    double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
    for (int t = 0; t < 9; ++t) {
        const int s = order[t];
        const int k = block_cell(g, m, n, s);
        if (k < 0) continue;
        double xr[4], xi[4];
        for (int j = 0; j < 4; ++j) {
            xr[j] = d[2 * ((long)j * g.C + k)];
            xi[j] = d[2 * ((long)j * g.C + k) + 1];
        }
        for (int i = 0; i < 4; ++i) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < 4; ++j) {
                const double b = J[(long)((s * 4 + i) * 4 + j) * g.C + c];
                sr += b * xr[j];
                si += b * xi[j];
            }
            ar[i] += sr;
            ai[i] += si;
        }
    }
    for (int i = 0; i < 4; ++i) {
        out[2 * ((long)i * g.C + c)] = ar[i];
        out[2 * ((long)i * g.C + c) + 1] = ai[i];
    }
}

// Gather form of the transposed block product: cell c collects block `s` of every cell k whose
// neighbour s is c.
__global__ void __launch_bounds__(256)
k_apply_states_rev(Grid g, const double* __restrict__ J, const double* __restrict__ w,
                   double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    const int n = c / g.M, m = c - n * g.M;
    // cell k = c - offset(s) has c as its neighbour s
    const int dm[9] = {0, 0, 0, 0, 0, -1, -2, 1, 2};
    const int dn[9] = {0, -1, -2, 1, 2, 0, 0, 0, 0};
    double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
    for (int s = 0; s < 9; ++s) {
        const int mk = m - dm[s];
        if (mk < 0 || mk >= g.M) continue;
        const int nk = wrapn(n - dn[s], g.N);
        const int k = nk * g.M + mk;
        double wr[4], wi[4];
        for (int i = 0; i < 4; ++i) {
            wr[i] = w[2 * ((long)i * g.C + k)];
            wi[i] = w[2 * ((long)i * g.C + k) + 1];
        }
        for (int j = 0; j < 4; ++j) {
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < 4; ++i) {
                const double b = J[(long)((s * 4 + i) * 4 + j) * g.C + k];
                sr += wr[i] * b;
                si += wi[i] * b;
            }
            ar[j] += sr;
            ai[j] += si;
        }
    }
    for (int j = 0; j < 4; ++j) {
        out[2 * ((long)j * g.C + c)] = ar[j];
        out[2 * ((long)j * g.C + c) + 1] = ai[j];
    }
}

__global__ void __launch_bounds__(256)
k_apply_vertices_fwd(Grid g, const double* __restrict__ J, const double* __restrict__ d,
                     double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    const int n = c / g.M, m = c - n * g.M;
    double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
    for (int a = 0; a < 4; ++a) {
        const int p = vtx_index(g, m, n, a);
        const double x0r = d[2 * p], x0i = d[2 * p + 1];
        const double x1r = d[2 * ((long)g.P + p)], x1i = d[2 * ((long)g.P + p) + 1];
        for (int i = 0; i < 4; ++i) {
            const double b0 = J[(long)((a * 4 + i) * 2 + 0) * g.C + c];
            const double b1 = J[(long)((a * 4 + i) * 2 + 1) * g.C + c];
            ar[i] += b0 * x0r + b1 * x1r;
            ai[i] += b0 * x0i + b1 * x1i;
        }
    }
    for (int i = 0; i < 4; ++i) {
        out[2 * ((long)i * g.C + c)] = ar[i];
        out[2 * ((long)i * g.C + c) + 1] = ai[i];
    }
}

// Gather: vertex p collects from the (up to) four cells sharing it.
__global__ void __launch_bounds__(256)
k_apply_vertices_rev(Grid g, const double* __restrict__ J, const double* __restrict__ w,
                     double* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const int nv = p / (g.M + 1), mv = p - nv * (g.M + 1);
    double ar[2] = {0, 0}, ai[2] = {0, 0};
    // cell (mv - dm, nv - dn) has this vertex as slot a
    const int dm[4] = {0, 0, 1, 1}, dn[4] = {0, 1, 1, 0};
    for (int a = 0; a < 4; ++a) {
        const int mk = mv - dm[a], nk = nv - dn[a];
        if (mk < 0 || mk >= g.M || nk < 0 || nk >= g.N) continue;
        const int k = nk * g.M + mk;
        for (int d = 0; d < 2; ++d) {
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < 4; ++i) {
                const double b = J[(long)((a * 4 + i) * 2 + d) * g.C + k];
                sr += w[2 * ((long)i * g.C + k)] * b;
                si += w[2 * ((long)i * g.C + k) + 1] * b;
            }
            ar[d] += sr;
            ai[d] += si;
        }
    }
    for (int d = 0; d < 2; ++d) {
        out[2 * ((long)d * g.P + p)] = ar[d];
        out[2 * ((long)d * g.P + p) + 1] = ai[d];
    }
}

// ------------------------------------------------------------------ BSR export
// Row `row` (cell n*M + m) lists its existing neighbours in ascending column order.
__global__ void __launch_bounds__(256)
k_export_bsr(Grid g, const double* __restrict__ J, double sre, double sim, int cplx,
             double* __restrict__ data, int* __restrict__ indices, int* __restrict__ indptr) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    const int n = c / g.M, m = c - n * g.M;
    // blocks before this row: 9 per cell minus the absent radial neighbours of earlier cells
    // (per column n: m=0 lacks i,ii; m=1 lacks ii; m=M-2 lacks oo; m=M-1 lacks o,oo)
    auto missing_upto = [&](int mm) {   // absent blocks of cells 0..mm-1 of one column
        int r = 0;
        if (mm > 0) r += 2;
        if (mm > 1) r += 1;
        if (mm > g.M - 2) r += 1;
        if (mm > g.M - 1) r += 2;
        return r;
    };
    const long base = 9L * c - (long)n * 6 - missing_upto(m);
    int cols[9], blks[9], cnt = 0;
    for (int s = 0; s < 9; ++s) {
        const int k = block_cell(g, m, n, s);
        if (k < 0) continue;
        cols[cnt] = k;
        blks[cnt++] = s;
    }
    for (int a = 1; a < cnt; ++a)
        for (int b = a; b > 0 && cols[b - 1] > cols[b]; --b) {
            int t = cols[b]; cols[b] = cols[b - 1]; cols[b - 1] = t;
            t = blks[b]; blks[b] = blks[b - 1]; blks[b - 1] = t;
        }
    for (int a = 0; a < cnt; ++a) {
        const long e = base + a;
        indices[e] = cols[a];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double v = J[(long)((blks[a] * 4 + i) * 4 + j) * g.C + c];
                double vi = 0.0;
                if (blks[a] == C && i == j) {
                    v -= sre;
                    vi = -sim;
                }
                if (cplx) {
                    data[2 * (16 * e + 4 * i + j)] = v;
                    data[2 * (16 * e + 4 * i + j) + 1] = vi;
                } else {
                    data[16 * e + 4 * i + j] = v;
                }
            }
    }
    indptr[c + 1] = (int)(base + cnt);
    if (c == 0) indptr[0] = 0;
}

// ------------------------------------------------------------------ forces
// One block; thread t loops over faces t, t+blockDim, ...; fixed-order LDS tree reductions.
__global__ void __launch_bounds__(1024)
k_forces(Grid g, const double* __restrict__ X, const double* __restrict__ U,
         double* __restrict__ F, double xc, double zc, double* __restrict__ coef) {
    __shared__ double red[3][1024];
    double sx = 0.0, sz = 0.0, sm = 0.0;
    for (int n = threadIdx.x; n < g.N; n += blockDim.x) {
        const int a = n * (g.M + 1), b = (n + 1) * (g.M + 1);
        double xa[2] = {X[a], X[g.P + a]}, xb[2] = {X[b], X[g.P + b]}, nr[2], u[4];
        edge_normal(xa, xb, nr);
        for (int v = 0; v < 4; ++v) u[v] = U[v * g.C + n * g.M];
        const double p = pressure(u);
        const double fx = p * nr[0], fz = p * nr[1];
        F[n] = fx;
        F[g.N + n] = fz;
        const double px = (xa[0] + xb[0]) / 2.0 - xc, pz = (xa[1] + xb[1]) / 2.0 - zc;
        sx += fx;
        sz += fz;
        sm += pz * fx - px * fz;
    }
    if (!coef) return;
    red[0][threadIdx.x] = sx;
    red[1][threadIdx.x] = sz;
    red[2][threadIdx.x] = sm;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) coef[q] = red[q][0];
}

__global__ void __launch_bounds__(256)
k_jvp_forces(Grid g, const double* __restrict__ X, const double* __restrict__ U,
             const double* __restrict__ dX, const double* __restrict__ dU,
             double* __restrict__ dF) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.N) return;
    const int a = n * (g.M + 1), b = (n + 1) * (g.M + 1), c = n * g.M;
    Dual<2> xa[2], xb[2], u[4], nr[2];
    for (int d = 0; d < 2; ++d) {
        xa[d].v = X[d * g.P + a];
        xb[d].v = X[d * g.P + b];
        for (int j = 0; j < 2; ++j) {
            xa[d].d[j] = dX ? dX[2 * ((long)d * g.P + a) + j] : 0.0;
            xb[d].d[j] = dX ? dX[2 * ((long)d * g.P + b) + j] : 0.0;
        }
    }
    for (int v = 0; v < 4; ++v) {
        u[v].v = U[(long)v * g.C + c];
        for (int j = 0; j < 2; ++j) u[v].d[j] = dU ? dU[2 * ((long)v * g.C + c) + j] : 0.0;
    }
    edge_normal(xa, xb, nr);
    const Dual<2> p = pressure(u);
    for (int d = 0; d < 2; ++d) {
        const Dual<2> f = p * nr[d];
        dF[2 * (d * g.N + n)] = f.d[0];
        dF[2 * (d * g.N + n) + 1] = f.d[1];
    }
}

// Adjoint seed and output conversion for each primal type.
template <class T>
struct AdjIO;
template <>
struct AdjIO<double> {
    static constexpr int L = 2;
    PEA_HD static void seed(const double* w, long e, double* s) {
        s[0] = w[2 * e];
        s[1] = w[2 * e + 1];
    }
    PEA_HD static void put(double* out, long e, const double* a) {
        out[2 * e] = a[0];
        out[2 * e + 1] = a[1];
    }
};
// tangent-over-adjoint: lanes k = real/imag part of the seed, tangent lanes j = real/imag part
// of the direction; the complex product is bilinear in (seed, direction).
template <>
struct AdjIO<Dual<2>> {
    static constexpr int L = 2;
    PEA_HD static void seed(const double* w, long e, Dual<2>* s) {
        s[0] = dconst<2>(w[2 * e]);
        s[1] = dconst<2>(w[2 * e + 1]);
    }
    PEA_HD static void put(double* out, long e, const Dual<2>* a) {
        out[2 * e] = a[0].d[0] - a[1].d[1];
        out[2 * e + 1] = a[0].d[1] + a[1].d[0];
    }
};
// The same product in two passes with one seed lane each (PART 0: real part of the seed,
// written; PART 1: imaginary part, added): half the adjoint state per thread.
template <int PART>
struct HvpIO {
    static constexpr int L = 1;
    PEA_HD static void seed(const double* w, long e, Dual<2>* s) {
        s[0] = dconst<2>(w[2 * e + PART]);
    }
    PEA_HD static void put(double* out, long e, const Dual<2>* a) {
        if (PART == 0) {
            out[2 * e] = a[0].d[0];
            out[2 * e + 1] = a[0].d[1];
        } else {
            out[2 * e] = out[2 * e] - a[0].d[1];
            out[2 * e + 1] = out[2 * e + 1] + a[0].d[0];
        }
    }
};

// Reverse of the surface forces F_n = p(U_{0,n}) normal(X_{0,n} -> X_{0,n+1}) (f90:1156-1174).
// gU on the wall ring (zero elsewhere), one thread per cell; T = double gives the adjoint,
// T = Dual<2> its tangent along the loader's directions.
template <class T, class Ld>
__global__ void __launch_bounds__(256)
k_vjp_forces_states(Ld ld, const double* __restrict__ wF, double* __restrict__ gU) {
    const Grid g = ld.g;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    const int n = c / g.M, m = c - n * g.M;
    T out[4][2];
    for (int v = 0; v < 4; ++v) out[v][0] = out[v][1] = lit<T>(0.0);
    if (m == 0) {
        T xa[2], xb[2], nr[2], u[4];
        ld.vertex(n * (g.M + 1), xa);
        ld.vertex((n + 1) * (g.M + 1), xb);
        edge_normal(xa, xb, nr);
        ld.state(c, u);
        const T iu = 1.0 / u[0];
        for (int k = 0; k < 2; ++k) {
            const T pb[1] = {wF[2 * n + k] * nr[0] + wF[2 * (g.N + n) + k] * nr[1]};
            T ub[4][1];
            for (int v = 0; v < 4; ++v) ub[v][0] = lit<T>(0.0);
            pressure_rev<T, 1>(u, iu, pb, ub);
            for (int v = 0; v < 4; ++v) out[v][k] = ub[v][0];
        }
    }
    for (int v = 0; v < 4; ++v) AdjIO<T>::put(gU, (long)v * g.C + c, out[v]);
}

// gX on the surface row (zero elsewhere), one thread per vertex.
template <class T, class Ld>
__global__ void __launch_bounds__(256)
k_vjp_forces_vertices(Ld ld, const double* __restrict__ wF, double* __restrict__ gX) {
    const Grid g = ld.g;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const int nv = p / (g.M + 1), mv = p - nv * (g.M + 1);
    T out[2][2];
    for (int d = 0; d < 2; ++d) out[d][0] = out[d][1] = lit<T>(0.0);
    if (mv == 0) {
        // normal(a->b) = (b_z - a_z, a_x - b_x);  F = p n
        for (int side = 0; side < 2; ++side) {   // 0: vertex is a of face nv, 1: b of nv-1
            const int f = side ? nv - 1 : nv;
            if (f < 0 || f >= g.N) continue;
            T u[4];
            ld.state(f * g.M, u);
            const T pr = pressure(u);
            const double sgn = side ? 1.0 : -1.0;
            for (int k = 0; k < 2; ++k) {
                const T w0 = wF[2 * f + k] * pr, w1 = wF[2 * (g.N + f) + k] * pr;
                out[1][k] = out[1][k] + sgn * w0;    // d n0 / d z
                out[0][k] = out[0][k] - sgn * w1;    // d n1 / d x
            }
        }
    }
    for (int d = 0; d < 2; ++d) AdjIO<T>::put(gX, (long)d * g.P + p, out[d]);
}

// =============================================================================== adjoint
// Tile of TM x TN cells per block.  The block sweeps the reverse of every cell whose stencil
// reaches the tile (the tile grown by two cells along both grid lines, plus the diagonal ring
// that shares tile vertices), face by face.  Each face reverse yields four 4-vectors for the
// cells at fixed offsets along one grid line; they are added into LDS accumulators in four
// barrier-separated phases, so within a phase every thread writes a distinct cell.  The
// accumulation order of every output is therefore fixed: results are bitwise deterministic
// and independent of the tile shape.
constexpr int kVjpThreads = 256;

template <class T, int TM, int TN, bool WU, bool WX, bool WV, bool WMA, class Ld,
          class IO = AdjIO<T>>
__global__ void __launch_bounds__(kVjpThreads)
k_vjp(pea_params P, Ld ld, const double* __restrict__ wseed, double* __restrict__ gU,
      double* __restrict__ gX, double* __restrict__ gV, double* __restrict__ partial) {
    constexpr int L = IO::L;
    constexpr int BM = TM + 4, BN = TN + 4, NBOX = BM * BN;
    constexpr int ROUNDS = (NBOX + kVjpThreads - 1) / kVjpThreads;
    constexpr int NVX = (WX || WV) ? (TM + 1) * (TN + 1) : 1;
    __shared__ T accU[WU ? TM * TN : 1][4][L];
    __shared__ T accX[WX ? NVX : 1][2][L];
    __shared__ T accV[WV ? NVX : 1][2][L];
    __shared__ T red[WMA ? kVjpThreads : 1][L];

    const Grid g = ld.g;
    const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
    const int tid = threadIdx.x;
    const bool last_m = m0 + TM >= g.M, last_n = n0 + TN >= g.N;

    if (WU)
        for (int i = tid; i < TM * TN * 4 * L; i += kVjpThreads)
            (&accU[0][0][0])[i] = lit<T>(0.0);
    if (WX)
        for (int i = tid; i < NVX * 2 * L; i += kVjpThreads) (&accX[0][0][0])[i] = lit<T>(0.0);
    if (WV)
        for (int i = tid; i < NVX * 2 * L; i += kVjpThreads) (&accV[0][0][0])[i] = lit<T>(0.0);
    __syncthreads();

    T machacc[L];
#pragma unroll
    for (int k = 0; k < L; ++k) machacc[k] = lit<T>(0.0);

    // offsets (dm, dn) of slot j of face fc
    constexpr int odm[4][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {1, 0, -1, -2}, {-1, 0, 1, 2}};
    constexpr int odn[4][4] = {{1, 0, -1, -2}, {-1, 0, 1, 2}, {0, 0, 0, 0}, {0, 0, 0, 0}};

    for (int r = 0; r < ROUNDS; ++r) {
        const int bidx = r * kVjpThreads + tid;
        const int mb = bidx % BM, nb = bidx / BM;
        const int m = m0 - 2 + mb;
        const int n = wrapn(n0 - 2 + nb, g.N);
        const bool cross = (mb >= 2 && mb < TM + 2) || (nb >= 2 && nb < TN + 2);
        const bool ring = (WX || WV) && mb >= 1 && mb <= TM + 1 && nb >= 1 && nb <= TN + 1;
        const bool active = bidx < NBOX && m >= 0 && m < g.M && (cross || ring);
        const bool in_tile = active && mb >= 2 && mb < TM + 2 && nb >= 2 && nb < TN + 2 &&
                             n0 - 2 + nb < g.N;
        const int flag = active ? boundary_flag(g.M, m) : 3;
        const int c = n * g.M + (active ? m : 0);
        const CellSrc<Ld> src{ld, active ? m : 0, n};

        CellAdj<T, L> ca;
        T uc[4];
        T x[4][2];
        if (active) {
            T xw[4][2];
            load_cell_geometry<T>(ld, m, n, x, xw);
            ld.state(c, uc);
            T ws[4][L];
#pragma unroll
            for (int v = 0; v < 4; ++v) IO::seed(wseed, (long)v * g.C + c, ws[v]);
            cell_adj_init<T, L>(P, ld.machv(), src, uc, x, xw, flag, ws, WX && gX, ca);
        }
#pragma unroll
        for (int fc = 0; fc < 4; ++fc) {
            T tb[4][4][L];
            if (active) {
                cell_face_rev<T, L>(P, src, uc, flag, fc, ca, tb);
                if (fc == 3) {
                    T mb_[L];
#pragma unroll
                    for (int k = 0; k < L; ++k) mb_[k] = lit<T>(0.0);
                    cell_ghost_rev<T, L>(P, ld.machv(), uc, flag, ca, tb[1], mb_);
                    if (in_tile)
#pragma unroll
                        for (int k = 0; k < L; ++k) machacc[k] = machacc[k] + mb_[k];
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (WU && active) {
                    const int mt = m + odm[fc][j];
                    const int nt = wrapn(n + odn[fc][j], g.N);
                    if (mt >= m0 && mt < m0 + TM && mt < g.M && nt >= n0 && nt < n0 + TN) {
                        const int t = (nt - n0) * TM + (mt - m0);
#pragma unroll
                        for (int v = 0; v < 4; ++v)
#pragma unroll
                            for (int k = 0; k < L; ++k)
                                accU[t][v][k] = accU[t][v][k] + tb[j][v][k];
                    }
                }
                __syncthreads();
            }
        }
        if (WX || WV) {
            T xb[4][2][L], wb[4][2][L];
            if (active) cell_vertex_rev<T, L>(x, ca, xb, wb);
#pragma unroll
            for (int vtx = 0; vtx < 4; ++vtx) {
                if (active) {
                    const int mv = m + ((vtx == FO || vtx == BO) ? 1 : 0);
                    const int nv = n + ((vtx == FI || vtx == FO) ? 1 : 0);
                    const bool own_m = (mv >= m0 && mv < m0 + TM) || (mv == g.M && last_m);
                    const bool own_n = (nv >= n0 && nv < n0 + TN) || (nv == g.N && last_n);
                    if (own_m && own_n) {
                        const int t = (nv - n0) * (TM + 1) + (mv - m0);
#pragma unroll
                        for (int d = 0; d < 2; ++d)
#pragma unroll
                            for (int k = 0; k < L; ++k) {
                                if (WX) accX[t][d][k] = accX[t][d][k] + xb[vtx][d][k];
                                if (WV) accV[t][d][k] = accV[t][d][k] + wb[vtx][d][k];
                            }
                    }
                }
                __syncthreads();
            }
        }
    }

    if (WU && gU)
        for (int t = tid; t < TM * TN; t += kVjpThreads) {
            const int mt = m0 + t % TM, nt = n0 + t / TM;
            if (mt < g.M && nt < g.N)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    IO::put(gU, (long)v * g.C + nt * g.M + mt, accU[t][v]);
        }
    if (WX || WV)
        for (int t = tid; t < NVX; t += kVjpThreads) {
            const int mv = m0 + t % (TM + 1), nv = n0 + t / (TM + 1);
            const bool own_m = (mv < m0 + TM && mv < g.M) || (mv == g.M && last_m);
            const bool own_n = (nv < n0 + TN && nv < g.N) || (nv == g.N && last_n);
            if (own_m && own_n)
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    const long e = (long)d * g.P + nv * (g.M + 1) + mv;
                    if (WX && gX) IO::put(gX, e, accX[t][d]);
                    if (WV && gV) IO::put(gV, e, accV[t][d]);
                }
        }
    if (WMA) {
#pragma unroll
        for (int k = 0; k < L; ++k) red[tid][k] = machacc[k];
        __syncthreads();
        for (int s = kVjpThreads / 2; s > 0; s >>= 1) {
            if (tid < s)
#pragma unroll
                for (int k = 0; k < L; ++k) red[tid][k] = red[tid][k] + red[tid + s][k];
            __syncthreads();
        }
        if (tid == 0 && partial)
            IO::put(partial, blockIdx.y * gridDim.x + blockIdx.x, red[0]);
    }
}

// ------------------------------------------------------------------ face-based adjoint
// (dR/dU)^T w for a TM x TN tile of cells (T = double, two lanes: real and imaginary part of
// the seed).  Every face whose stencil reaches the tile is reversed once for both of its cells
// (face_flux2_rev): first the angular faces of the tile's rows, then the radial faces of its
// columns.  A face's contributions to its four cells along one grid line are added to LDS
// accumulators in four barrier-separated phases (offset -1, 0, +1, +2 from L), so within a
// phase every face writes a distinct cell and the order of every sum is fixed.  Pressure and
// sound-speed adjoints are accumulated as scalars and turned into state adjoints once per cell
// at the end, together with the wall / far-field faces and their ghost states.
// Requires TN + 3 <= N (no face is reached twice through the periodic seam).
template <int TM, int TN, int NT>
__device__ __forceinline__ void face_vjp_tile(const pea_params& P, const FsDir& fs,
                                              const LdReal& ld, const double* __restrict__ wseed,
                                              double* __restrict__ gU, int m0, int n0) {
    using T = double;
    constexpr int L = 2;
    constexpr int BM = TM + 6, BN = TN + 6, BOX = BM * BN, NC = TM * TN;
    __shared__ T sP[BOX], sC[BOX];
    // structure-of-arrays: consecutive lanes touch consecutive doubles (no bank conflicts)
    __shared__ T sFB[4][L][BOX];          // -w / area of each cell of the grown tile
    // states of the grown tile, wall (m = -1) and far-field (m = M) ghosts in the grown rows,
    // staged once so that every stencil read below is an LDS read
    __shared__ T sU[4][BOX];
    __shared__ T accU[4][L][NC], accP[L][NC], accC[L][NC];
    __shared__ T gU_[2][TN][4][L], gP_[2][TN][L];   // ghost adjoints of the tile's columns
    const Grid g = ld.g;
    const int tid = threadIdx.x;

    auto vtx = [&](int mm, int nn, T* x) { ld.vertex(nn * (g.M + 1) + mm, x); };
    auto nrm = [&](int ma, int na, int mb, int nb, T* nr) {
        T xa[2], xb[2];
        vtx(ma, na, xa);
        vtx(mb, nb, xb);
        edge_normal(xa, xb, nr);
    };
    auto edge = [&](int ma, int na, int mb, int nb, T* nr, T* gv) {
        T va[2], vb[2];
        nrm(ma, na, mb, nb, nr);
        ld.velocity(na * (g.M + 1) + ma, va);
        ld.velocity(nb * (g.M + 1) + mb, vb);
        gv[0] = (va[0] + vb[0]) / 2.0;
        gv[1] = (va[1] + vb[1]) / 2.0;
    };
    auto bidx = [&](int mm, int nn_box) { return nn_box * BM + (mm - m0 + 3); };
    // box column of cell column nn (wrapped), for nn within three columns of the tile
    auto bcol = [&](int nn) {
        int d = nn - (n0 - 3);
        if (d < 0) d += g.N;
        if (d >= g.N) d -= g.N;
        return d;
    };
    auto in_tile_n = [&](int nn) { return nn >= n0 && nn < n0 + TN; };
    auto box = [&](int b, T* u) {
#pragma unroll
        for (int v = 0; v < 4; ++v) u[v] = sU[v][b];
    };
    struct Seed {
        T w[4][L];
    };
    auto seed = [&](int b) {
        Seed r;
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int kk = 0; kk < L; ++kk) r.w[v][kk] = sFB[v][kk][b];
        return r;
    };

    // 0. the vertices and velocities of this thread's angular face (step 2) and radial face
    // (step 3), loaded before the staging so that their latency overlaps it (indices clamped to
    // the grid for threads without a face).  An angular face across the seam also needs vertex
    // column 0, R's own.
    constexpr int NFA = TM * (TN + 3), NFR = (TM + 3) * TN;
    static_assert(NFA <= NT && NFR <= NT, "one angular and one radial face per thread");
    const int fa_tm = tid % TM, fa_j = tid / TM;
    const int fa_m = m0 + fa_tm, fa_Lc = wrap_any(n0 - 2 + fa_j, g.N);
    const bool fa_act = tid < NFA && fa_m < g.M;
    const bool fa_seam = fa_Lc + 1 == g.N;
    T axa[2], axb[2], ava[2], avb[2], sxa[2], sxb[2], sva[2], svb[2];
    {
        // the face's two vertices are neighbours along m: one 16-byte load per plane
        const int mm = fa_act ? fa_m : 0, cc = fa_act ? fa_Lc + 1 : 0;
        ld.vertex2(cc * (g.M + 1) + mm, axa, axb);
        ld.velocity2(cc * (g.M + 1) + mm, ava, avb);
        if (fa_act && fa_seam) {
            ld.vertex2(mm, sxa, sxb);
            ld.velocity2(mm, sva, svb);
        }
    }
    const int fr_i = tid % (TM + 3), fr_tn = tid / (TM + 3);
    const int fr_n = n0 + fr_tn, fr_Lm = m0 - 2 + fr_i;
    const bool fr_act = tid < NFR && fr_n < g.N && fr_Lm >= 0 && fr_Lm <= g.M - 2;
    T rxa[2], rxb[2], rva[2], rvb[2];
    {
        const int mm = fr_act ? fr_Lm + 1 : 0, nn = fr_act ? fr_n : 0;
        vtx(mm, nn + 1, rxa);
        vtx(mm, nn, rxb);
        ld.velocity((nn + 1) * (g.M + 1) + mm, rva);
        ld.velocity(nn * (g.M + 1) + mm, rvb);
    }

    // 1. per-cell pressures, sound speeds, seeds; ghosts; zero accumulators.  Every pass's
    // loads (the box cell's state, area vertices and seed, the row clamped to the grid) are
    // issued before the first is used: one round trip instead of one per pass (passes past the
    // box load a valid cell and discard it).
    constexpr int NPASS = (BOX + NT - 1) / NT;
    T su[NPASS][4], sx[NPASS][4][2], sw[NPASS][4][L];
#pragma unroll
    for (int k = 0; k < NPASS; ++k) {
        const int i = tid + k * NT < BOX ? tid + k * NT : BOX - 1;
        const int bm = i % BM, bn = i / BM;
        const int cm = m0 - 3 + bm, cmc = cm < 0 ? 0 : (cm >= g.M ? g.M - 1 : cm);
        const int cn = wrap_any(n0 - 3 + bn, g.N);
        const int c = cn * g.M + cmc;
        ld.state(c, su[k]);
        ld.vertex2(cn * (g.M + 1) + cmc, sx[k][BI], sx[k][BO]);
        ld.vertex2((cn + 1) * (g.M + 1) + cmc, sx[k][FI], sx[k][FO]);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            sw[k][v][0] = wseed[2 * ((long)v * g.C + c)];
            sw[k][v][1] = wseed[2 * ((long)v * g.C + c) + 1];
        }
    }
#pragma unroll
    for (int k = 0; k < NPASS; ++k) {
        const int i = tid + k * NT;
        if (i >= BOX) continue;
        const int bm = i % BM, bn = i / BM;
        const int cm = m0 - 3 + bm;
        const int cn = wrap_any(n0 - 3 + bn, g.N);
        if (cm >= 0 && cm < g.M) {
            const T* u = su[k];
            cell_pc(u, sP[i], sC[i]);
#pragma unroll
            for (int v = 0; v < 4; ++v) sU[v][i] = u[v];
            const T(&x)[4][2] = sx[k];
            const T d1x = x[FO][0] - x[BI][0], d1z = x[FO][1] - x[BI][1];
            const T d2x = x[BO][0] - x[FI][0], d2z = x[BO][1] - x[FI][1];
            const T ia = rcp_fast(sabs(d1x * d2z - d1z * d2x) / 2.0);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                sFB[v][0][i] = -(sw[k][v][0] * ia);
                sFB[v][1][i] = -(sw[k][v][1] * ia);
            }
        } else if (cm == -1 || cm == g.M) {
            const T* u = su[k];   // the state of row 0 or M - 1 (the row clamped to the grid)
            T gh[4];
            if (cm < 0) {
                T nr[2], gv[2];
                edge(0, cn, 0, cn + 1, nr, gv);
                wall_ghost(u, nr, gv, gh);
            } else {
                T inf[4];
                free_stream_dir(ld.mach, fs.ca, fs.sa, inf);
#pragma unroll
                for (int v = 0; v < 4; ++v) gh[v] = 2.0 * inf[v] - u[v];
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) sU[v][i] = gh[v];
            sP[i] = pressure_q(gh);
        }
    }
    for (int i = tid; i < NC * 4 * L; i += NT) (&accU[0][0][0])[i] = 0.0;
    for (int i = tid; i < NC * L; i += NT) {
        (&accP[0][0])[i] = 0.0;
        (&accC[0][0])[i] = 0.0;
    }
    for (int i = tid; i < 2 * TN * 4 * L; i += NT) (&gU_[0][0][0][0])[i] = 0.0;
    for (int i = tid; i < 2 * TN * L; i += NT) (&gP_[0][0][0])[i] = 0.0;
    __syncthreads();

    auto add = [&](int t, const FaceAdj<T, L>& a, int j) {
#pragma unroll
        for (int k = 0; k < L; ++k) {
#pragma unroll
            for (int v = 0; v < 4; ++v) accU[v][k][t] += a.u[j][v][k];
            accP[k][t] += a.p[j][k];
            if (j == 1 || j == 2) accC[k][t] += a.c[j - 1][k];
        }
    };

    // 2. angular faces (m, Lc) | (m, Lc+1) for the tile's rows, Lc from n0-2 to n0+TN
    {
        {
            const int tm = fa_tm, m = fa_m, Lc = fa_Lc;
            const bool act = fa_act;
            FaceAdj<T, L> a;
            if (act) {
                face_adj_zero(a);
                const int Rc = wrapn(Lc + 1, g.N), LLc = wrapn(Lc - 1, g.N),
                          RRc = wrapn(Lc + 2, g.N);
                const int bL = bidx(m, bcol(Lc)), bR = bidx(m, bcol(Rc));
                const int bLL = bidx(m, bcol(LLc)), bRR = bidx(m, bcol(RRc));
                T Ls[4], Rs[4], LLs[4], RRs[4];
                box(bL, Ls);
                box(bR, Rs);
                box(bLL, LLs);
                box(bRR, RRs);
                T nL[2], nRn[2], gv[2];
                edge_normal(axa, axb, nL);                  // L's fi -> fo
                gv[0] = (ava[0] + avb[0]) / 2.0;
                gv[1] = (ava[1] + avb[1]) / 2.0;
                if (fa_seam) edge_normal(sxb, sxa, nRn);    // R's bo -> bi (column 0)
                else edge_normal(axb, axa, nRn);
                const T zero[4][L] = {{-0.0, -0.0}, {-0.0, -0.0}, {-0.0, -0.0}, {-0.0, -0.0}};
                if (Lc + 1 == g.N) {
                    // seam: L's face uses vertex column N, R's face column 0 -- reverse each
                    // side with its own geometry
                    face_flux2_rev<T, T, L>(P, LLs, Ls, Rs, RRs, sP[bLL], sP[bL], sP[bR],
                                            sP[bRR], sC[bL], sC[bR], nL, nRn, gv, true, false,
                                            seed(bL).w, zero, a);
                    T nL0[2], gv0[2];
                    edge_normal(sxa, sxb, nL0);
                    gv0[0] = (sva[0] + svb[0]) / 2.0;
                    gv0[1] = (sva[1] + svb[1]) / 2.0;
                    face_flux2_rev<T, T, L>(P, LLs, Ls, Rs, RRs, sP[bLL], sP[bL], sP[bR],
                                            sP[bRR], sC[bL], sC[bR], nL0, nRn, gv0, false, true,
                                            zero, seed(bR).w, a);
                } else {
                    face_flux2_rev<T, T, L>(P, LLs, Ls, Rs, RRs, sP[bLL], sP[bL], sP[bR],
                                            sP[bRR], sC[bL], sC[bR], nL, nRn, gv, true, true,
                                            seed(bL).w, seed(bR).w, a);
                }
            }
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                if (act && tm < TM) {
                    const int nt = wrapn(Lc - 1 + jj, g.N);
                    if (in_tile_n(nt) && nt < g.N) add((nt - n0) * TM + tm, a, jj);
                }
                __syncthreads();
            }
        }
    }
    // 3. radial faces (Lm, n) | (Lm+1, n) for the tile's columns, Lm from m0-2 to m0+TM
    {
        {
            const int tn = fr_tn, Lm = fr_Lm;
            const bool act = fr_act;
            FaceAdj<T, L> a;
            if (act) {
                face_adj_zero(a);
                const int bn = tn + 3;
                T Ls[4], Rs[4], LLs[4], RRs[4];
                box(bidx(Lm, bn), Ls);
                box(bidx(Lm + 1, bn), Rs);
                box(bidx(Lm - 1, bn), LLs);    // the wall ghost for Lm = 0
                box(bidx(Lm + 2, bn), RRs);    // the far-field ghost for Lm + 1 = M - 1
                T nL[2], nRn[2], gv[2];
                edge_normal(rxa, rxb, nL);                   // L's fo -> bo
                gv[0] = (rva[0] + rvb[0]) / 2.0;
                gv[1] = (rva[1] + rvb[1]) / 2.0;
                edge_normal(rxb, rxa, nRn);                  // R's bi -> fi
                face_flux2_rev<T, T, L>(P, LLs, Ls, Rs, RRs, sP[bidx(Lm - 1, bn)],
                                        sP[bidx(Lm, bn)], sP[bidx(Lm + 1, bn)],
                                        sP[bidx(Lm + 2, bn)], sC[bidx(Lm, bn)],
                                        sC[bidx(Lm + 1, bn)], nL, nRn, gv, Lm != g.M - 2,
                                        Lm + 1 != 1, seed(bidx(Lm, bn)).w,
                                        seed(bidx(Lm + 1, bn)).w, a);
            }
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                if (act) {
                    const int mt = Lm - 1 + jj;
                    if (mt == -1) {
                        if (m0 == 0) {
#pragma unroll
                            for (int k = 0; k < L; ++k) {
#pragma unroll
                                for (int v = 0; v < 4; ++v) gU_[0][tn][v][k] += a.u[jj][v][k];
                                gP_[0][tn][k] += a.p[jj][k];
                            }
                        }
                    } else if (mt == g.M) {
                        if (m0 + TM >= g.M) {
#pragma unroll
                            for (int k = 0; k < L; ++k) {
#pragma unroll
                                for (int v = 0; v < 4; ++v) gU_[1][tn][v][k] += a.u[jj][v][k];
                                gP_[1][tn][k] += a.p[jj][k];
                            }
                        }
                    } else if (mt >= m0 && mt < m0 + TM) {
                        add(tn * TM + (mt - m0), a, jj);
                    }
                }
                __syncthreads();
            }
        }
    }
    // 4. per cell: wall / far-field faces and ghosts, then pressure and sound-speed adjoints
    for (int t = tid; t < NC; t += NT) {
        const int tm = t % TM, tn = t / TM;
        const int m = m0 + tm, n = n0 + tn;
        if (m >= g.M || n >= g.N) continue;
        const int c = n * g.M + m, bn = tn + 3, b = bidx(m, bn);
        T u[4];
        box(b, u);
        T ub[4][L], pb[L], cb[L];
#pragma unroll
        for (int k = 0; k < L; ++k) {
#pragma unroll
            for (int v = 0; v < 4; ++v) ub[v][k] = accU[v][k][t];
            pb[k] = accP[k][t];
            cb[k] = accC[k][t];
        }
        if (m == 0 || m == g.M - 1) {
            const int side = m == 0 ? 0 : 1;
            T gh[4];
            box(bidx(side ? g.M : -1, bn), gh);
            T nr[2], gv[2];
            if (side == 0) edge(0, n, 0, n + 1, nr, gv);           // wall face bi -> fi
            else edge(m + 1, n + 1, m + 1, n, nr, gv);             // far field fo -> bo
            // central face (cell, ghost) seen from the cell: f_L of face_flux2_lds with L = cell
            FaceAdj<T, L> a;
            face_adj_zero(a);
            const T zero[4][L] = {{-0.0, -0.0}, {-0.0, -0.0}, {-0.0, -0.0}, {-0.0, -0.0}};
            face_flux2_rev<T, T, L>(P, u, u, gh, gh, sP[b], sP[b], sP[bidx(side ? g.M : -1, bn)],
                                    sP[bidx(side ? g.M : -1, bn)], sC[b], sC[b], nr, nr, gv,
                                    false, false, seed(b).w, zero, a);
            T ghb[4][L], ghp[L];
#pragma unroll
            for (int k = 0; k < L; ++k) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    ub[v][k] += a.u[1][v][k];
                    ghb[v][k] = gU_[side][tn][v][k] + a.u[2][v][k];
                }
                pb[k] += a.p[1][k];
                ghp[k] = gP_[side][tn][k] + a.p[2][k];
            }
            pressure_rev<T, L>(gh, rcp_fast(gh[0]), ghp, ghb);
            if (side == 0) {
                T nb_[2][L], gvb_[2][L];
                wall_ghost_rev<T, L>(u, nr, gv, ghb, ub, nb_, gvb_);
            } else {
#pragma unroll
                for (int k = 0; k < L; ++k)
#pragma unroll
                    for (int v = 0; v < 4; ++v) ub[v][k] -= ghb[v][k];
            }
        }
        // c = sqrt(gamma p / rho)
        const T p = sP[b], cc = sC[b], iu = rcp_fast(u[0]);
        const T a_ = kGamma * p * iu, ic2 = rcp_fast(2.0 * cc);
#pragma unroll
        for (int k = 0; k < L; ++k) {
            const T ab = cb[k] * ic2;
            pb[k] += kGamma * ab * iu;
            ub[0][k] -= ab * a_ * iu;
        }
        pressure_rev<T, L>(u, iu, pb, ub);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            gU[2 * ((long)v * g.C + c)] = ub[v][0];
            gU[2 * ((long)v * g.C + c) + 1] = ub[v][1];
        }
    }
}

template <int TM, int TN, int NT>
__global__ void __launch_bounds__(NT)
k_face_vjp(pea_params P, FsDir fs, LdReal ld, const double* __restrict__ wseed,
           double* __restrict__ gU) {
    const TileId t = batch_tile(gridDim.x, gridDim.y, gridDim.z);
    const long z = (long)t.bz * 8 * ld.g.C;   // seed bz of a batch
    face_vjp_tile<TM, TN, NT>(P, fs, ld, wseed + z, gU + z, t.bx * TM, t.by * TN);
}

// Sums per-block complex partials in block order (deterministic).
__global__ void __launch_bounds__(256)
k_sum_partials(const double* __restrict__ partial, int count, double* __restrict__ out) {
    __shared__ double red[2][256];
    double sr = 0.0, si = 0.0;
    for (int i = threadIdx.x; i < count; i += 256) {
        sr += partial[2 * i];
        si += partial[2 * i + 1];
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        out[1] = red[1][0];
    }
}

// =============================================================================== host side
namespace {

unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// cell kernels: smaller blocks when the grid is small so that all CUs get work
int cell_block(long C) { return C >= 256L * 1024 ? 256 : 64; }

// Tile of the face-based forward kernels (cells per workgroup, one thread per cell, m fastest
// within a tile) for the tangent along several inputs; the residual and the states tangent
// choose among several shapes at run time (kFaceTiles below).
constexpr int kFwdTM = 16, kFwdTN = 16;
constexpr int kJvpTM = 16, kJvpTN = 16;

// ------------------------------------------------------------------ many grids, one launch
// Descriptor of one grid of a multi-grid launch as the kernel reads it: the public pea_grid
// plus the grid's first tile in the launch-wide tile numbering.
struct GridDesc {
    pea_grid d;
    FsDir fs;
    int tiles_m, tile0;
};


// One workgroup per tile of G grids.  Tiles are numbered grid after grid; xcd_tile() gives each
// XCD a contiguous range of them, so neighbouring tiles of a grid (which share their halos) are
// read through the same L2.  gridDim.x is the total tile count; a workgroup finds its grid by
// binary search over the first-tile table.  (A persistent loop over tiles would keep every
// tile-invariant address and constant live across the loop.)
template <bool TANGENT, int TM, int TN>
__global__ void __launch_bounds__(TM * TN)
k_face_fwd_multi(const GridDesc* __restrict__ desc, int G) {
    const int t = (int)xcd_tile(blockIdx.x, gridDim.x);
    int lo = 0, hi = G - 1;   // last grid whose first tile is <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (desc[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    const GridDesc e = desc[lo];
    const Grid g{e.d.M, e.d.N, e.d.M * e.d.N, (e.d.M + 1) * (e.d.N + 1)};
    const int tl = t - e.tile0;
    if (TANGENT)
        face_fwd_tile<Dual<2>, double, TM, TN>(
            e.d.prm, e.fs, LdDualU<2>{e.d.X, e.d.V, e.d.U, e.d.dU, e.d.mach, g},
            StoreDualB<2>{e.d.dR, e.d.R, g.C}, tl % e.tiles_m, tl / e.tiles_m, 0);
    else
        face_fwd_tile<double, double, TM, TN>(
            e.d.prm, e.fs, LdReal{e.d.X, e.d.V, e.d.U, e.d.mach, g}, StoreRealB{e.d.R, g.C},
            tl % e.tiles_m, tl / e.tiles_m, 0);
}
dim3 fwd_grid(const Grid& g, int nstates, int tm = kFwdTM, int tn = kFwdTN) {
    return dim3((g.M + tm - 1) / tm, (g.N + tn - 1) / tn, nstates);
}

struct TileShape {
    int tm, tn;
};
// Tile shapes the adjoint kernels are instantiated for (tn + 4 <= N is required so that a box
// never wraps onto itself).  The tangent-over-adjoint kernels (3x wider accumulators) use only
// the shapes of at most 144 cells.
constexpr TileShape kTiles[] = {{28, 28}, {60, 12}, {28, 12}, {12, 12}, {28, 4}, {12, 1}};
constexpr int kNumTiles = 6;

PEA_HD constexpr int tile_rounds(int i) {
    return ((kTiles[i].tm + 4) * (kTiles[i].tn + 4) + 255) / 256;
}

bool tile_fits(const Grid& g, int i, bool small) {
    return i >= 0 && i < kNumTiles && kTiles[i].tn + 4 <= g.N &&
           (!small || kTiles[i].tm * kTiles[i].tn <= 144);
}

long tile_blocks(const Grid& g, int i) {
    return (long)((g.M + kTiles[i].tm - 1) / kTiles[i].tm) *
           ((g.N + kTiles[i].tn - 1) / kTiles[i].tn);
}

// Per-grid choices registered by the tuner (M = N = 0: process-wide override).
struct TileChoice {
    int M, N, tile;
};
std::mutex g_tile_mutex;
TileChoice g_tiles[64];
int g_ntiles = 0;
int g_tile_global = -1;
bool g_face_vjp = true;   // face-based adjoint for the states-only product

// Default: minimise the estimated time max(rounds, blocks*rounds/512) -- 512 resident blocks
// over 256 CUs -- then the total work.
int pick_tile(const Grid& g, bool small) {
    {
        std::lock_guard<std::mutex> lk(g_tile_mutex);
        if (tile_fits(g, g_tile_global, small)) return g_tile_global;
        for (int i = 0; i < g_ntiles; ++i)
            if (g_tiles[i].M == g.M && g_tiles[i].N == g.N && tile_fits(g, g_tiles[i].tile, small))
                return g_tiles[i].tile;
    }
    int best = -1;
    double best_t = 0.0, best_w = 0.0;
    for (int i = 0; i < kNumTiles; ++i) {
        if (!tile_fits(g, i, small)) continue;
        const double work = (double)tile_blocks(g, i) * tile_rounds(i);
        const double t = std::max((double)tile_rounds(i), work / 512.0);
        if (best < 0 || t < best_t - 1e-9 || (t < best_t + 1e-9 && work < best_w)) {
            best = i;
            best_t = t;
            best_w = work;
        }
    }
    return best;
}

template <class T, int TM, int TN, class Ld, class IO>
int launch_vjp_pass(const pea_params& P, const Ld& ld, const double* w, double* gU, double* gX,
                    double* gV, double* partial, hipStream_t st) {
    const dim3 grid((ld.g.M + TM - 1) / TM, (ld.g.N + TN - 1) / TN);
    if (gU && !gX && !gV && !partial) {
        k_vjp<T, TM, TN, true, false, false, false, Ld, IO>
            <<<grid, kVjpThreads, 0, st>>>(P, ld, w, gU, gX, gV, partial);
    } else {
        k_vjp<T, TM, TN, true, true, true, true, Ld, IO>
            <<<grid, kVjpThreads, 0, st>>>(P, ld, w, gU, gX, gV, partial);
    }
    PEA_LAUNCH_CHECK("k_vjp");
    return PEA_OK;
}

// The adjoint (T = double) in one pass; the tangent-over-adjoint (T = Dual<2>) in two passes of
// one seed lane each, the second adding into the first's outputs (and block partials).
template <class T, int TM, int TN, class Ld>
int launch_vjp_shape(const pea_params& P, const Ld& ld, const double* w, double* gU, double* gX,
                     double* gV, double* partial, hipStream_t st) {
    if constexpr (std::is_same<T, double>::value) {
        return launch_vjp_pass<T, TM, TN, Ld, AdjIO<T>>(P, ld, w, gU, gX, gV, partial, st);
    } else {
        if (int rc = launch_vjp_pass<T, TM, TN, Ld, HvpIO<0>>(P, ld, w, gU, gX, gV, partial, st))
            return rc;
        return launch_vjp_pass<T, TM, TN, Ld, HvpIO<1>>(P, ld, w, gU, gX, gV, partial, st);
    }
}

// (dR/dU)^T w of G grids, one workgroup per tile as k_face_fwd_multi: a descriptor's dU holds
// the seed w and dR receives the product.  Every grid needs N >= the adjoint tile's width + 3.
template <int TM, int TN>
__global__ void __launch_bounds__(256)
k_face_vjp_multi(const GridDesc* __restrict__ desc, int G) {
    const int t = (int)xcd_tile(blockIdx.x, gridDim.x);
    int lo = 0, hi = G - 1;   // last grid whose first tile is <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (desc[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    const GridDesc e = desc[lo];
    const Grid g{e.d.M, e.d.N, e.d.M * e.d.N, (e.d.M + 1) * (e.d.N + 1)};
    const int tl = t - e.tile0;
    face_vjp_tile<TM, TN, 256>(
        e.d.prm, e.fs, LdReal{e.d.X, e.d.V, e.d.U, e.d.mach, g}, e.d.dU, e.d.dR,
        (tl % e.tiles_m) * TM, (tl / e.tiles_m) * TN);
}

// ------------------------------------------------------------------ face-kernel tiles
// Tile shapes of the face-based kernels among which the tuner chooses, per grid size (kind 0:
// residual, 1: states tangent, 2: states adjoint; kinds 3-5: the same kernels in multi-grid
// launches).  The first shape of each list is the default.  An adjoint tile TM x TN takes its
// TM (TN + 3) angular and (TM + 3) TN radial faces in one round of 256 threads each and needs
// N >= TN + 3.
constexpr TileShape kFaceTiles[3][4] = {
    {{16, 16}, {32, 8}, {32, 16}, {16, 32}},
    {{16, 16}, {16, 8}, {32, 8}, {8, 16}},
    {{14, 15}, {16, 13}, {18, 11}, {20, 9}},
};
constexpr int kNumFaceTiles = 4;
struct FaceChoice {
    int kind, M, N, tile;
};
std::mutex g_face_mutex;
FaceChoice g_face_tiles[256];
int g_nface_tiles = 0;
int g_face_global[6] = {-1, -1, -1, -1, -1, -1};

bool face_tile_fits(int kind, int tile, int N) {
    return tile >= 0 && tile < kNumFaceTiles && (kind % 3 != 2 || N >= kFaceTiles[2][tile].tn + 3);
}

// The registered tile of `kind` for an M x N grid (the process-wide choice first), else the
// first shape of the list that fits; -1 when no adjoint tile fits the grid.
int face_tile(int kind, int M, int N) {
    {
        std::lock_guard<std::mutex> lk(g_face_mutex);
        if (face_tile_fits(kind, g_face_global[kind], N)) return g_face_global[kind];
        for (int i = 0; i < g_nface_tiles; ++i)
            if (g_face_tiles[i].kind == kind && g_face_tiles[i].M == M && g_face_tiles[i].N == N &&
                face_tile_fits(kind, g_face_tiles[i].tile, N))
                return g_face_tiles[i].tile;
    }
    for (int t = 0; t < kNumFaceTiles; ++t)
        if (face_tile_fits(kind, t, N)) return t;
    return -1;
}

template <int TM, int TN>
void launch_fwd_real_t(const pea_params& P, const LdReal& ld, double* R, int nstates,
                       hipStream_t st) {
    k_face_fwd<double, double, TM, TN><<<fwd_grid(ld.g, nstates, TM, TN), TM * TN, 0, st>>>(
        P, fs_dir(P), ld, StoreRealB{R, ld.g.C});
}
// residual of nstates states, tile from the registry
void launch_fwd_real(const pea_params& P, const LdReal& ld, double* R, int nstates,
                     hipStream_t st) {
    switch (face_tile(0, ld.g.M, ld.g.N)) {
    case 1: launch_fwd_real_t<32, 8>(P, ld, R, nstates, st); break;
    case 2: launch_fwd_real_t<32, 16>(P, ld, R, nstates, st); break;
    case 3: launch_fwd_real_t<16, 32>(P, ld, R, nstates, st); break;
    default: launch_fwd_real_t<16, 16>(P, ld, R, nstates, st); break;
    }
}

template <int TM, int TN>
void launch_jvp_u_t(const pea_params& P, const LdDualU<2>& ld, const StoreDualB<2>& out, int k,
                    hipStream_t st) {
    k_face_fwd<Dual<2>, double, TM, TN><<<fwd_grid(ld.g, k, TM, TN), TM * TN, 0, st>>>(
        P, fs_dir(P), ld, out);
}
// states tangent of k complex directions (one per grid z), tile from the registry
void launch_jvp_u(const pea_params& P, const LdDualU<2>& ld, const StoreDualB<2>& out, int k,
                  hipStream_t st) {
    switch (face_tile(1, ld.g.M, ld.g.N)) {
    case 1: launch_jvp_u_t<16, 8>(P, ld, out, k, st); break;
    case 2: launch_jvp_u_t<32, 8>(P, ld, out, k, st); break;
    case 3: launch_jvp_u_t<8, 16>(P, ld, out, k, st); break;
    default: launch_jvp_u_t<16, 16>(P, ld, out, k, st); break;
    }
}

template <int TM, int TN>
void launch_vjp_u_t(const pea_params& P, const LdReal& ld, const double* w, double* gU, int k,
                    hipStream_t st) {
    const dim3 grid((ld.g.M + TM - 1) / TM, (ld.g.N + TN - 1) / TN, k);
    k_face_vjp<TM, TN, 256><<<grid, 256, 0, st>>>(P, fs_dir(P), ld, w, gU);
}
// states adjoint of k seeds (one per grid z); false when no adjoint tile fits the grid
bool launch_vjp_u(const pea_params& P, const LdReal& ld, const double* w, double* gU, int k,
                  hipStream_t st) {
    switch (face_tile(2, ld.g.M, ld.g.N)) {
    case 0: launch_vjp_u_t<14, 15>(P, ld, w, gU, k, st); return true;
    case 1: launch_vjp_u_t<16, 13>(P, ld, w, gU, k, st); return true;
    case 2: launch_vjp_u_t<18, 11>(P, ld, w, gU, k, st); return true;
    case 3: launch_vjp_u_t<20, 9>(P, ld, w, gU, k, st); return true;
    default: return false;
    }
}

// The tile a multi-grid workspace was prepared with (pea_multi_launch must use the same).
std::mutex g_multi_mutex;
struct MultiTile {
    const void* ws;
    int mode, tile;
};
std::vector<MultiTile> g_multi_tiles;

template <bool TANGENT, int TM, int TN>
void launch_fwd_multi_t(const GridDesc* d, int G, int total, hipStream_t st) {
    k_face_fwd_multi<TANGENT, TM, TN><<<(unsigned)total, TM * TN, 0, st>>>(d, G);
}
template <bool TANGENT>
void launch_fwd_multi(int tile, const GridDesc* d, int G, int total, hipStream_t st) {
    if (TANGENT) {
        switch (tile) {
        case 1: launch_fwd_multi_t<TANGENT, 16, 8>(d, G, total, st); break;
        case 2: launch_fwd_multi_t<TANGENT, 32, 8>(d, G, total, st); break;
        case 3: launch_fwd_multi_t<TANGENT, 8, 16>(d, G, total, st); break;
        default: launch_fwd_multi_t<TANGENT, 16, 16>(d, G, total, st); break;
        }
    } else {
        switch (tile) {
        case 1: launch_fwd_multi_t<TANGENT, 32, 8>(d, G, total, st); break;
        case 2: launch_fwd_multi_t<TANGENT, 32, 16>(d, G, total, st); break;
        case 3: launch_fwd_multi_t<TANGENT, 16, 32>(d, G, total, st); break;
        default: launch_fwd_multi_t<TANGENT, 16, 16>(d, G, total, st); break;
        }
    }
}
void launch_vjp_multi(int tile, const GridDesc* d, int G, int total, hipStream_t st) {
    switch (tile) {
    case 1: k_face_vjp_multi<16, 13><<<(unsigned)total, 256, 0, st>>>(d, G); break;
    case 2: k_face_vjp_multi<18, 11><<<(unsigned)total, 256, 0, st>>>(d, G); break;
    case 3: k_face_vjp_multi<20, 9><<<(unsigned)total, 256, 0, st>>>(d, G); break;
    default: k_face_vjp_multi<14, 15><<<(unsigned)total, 256, 0, st>>>(d, G); break;
    }
}

template <class T, class Ld>
int launch_vjp(const pea_params& P, const Ld& ld, const double* w, double* gU, double* gX,
               double* gV, double* gmach, void* ws, hipStream_t st) {
    if constexpr (std::is_same<T, double>::value && std::is_same<Ld, LdReal>::value) {
        if (gU && !gX && !gV && !gmach && g_face_vjp && launch_vjp_u(P, ld, w, gU, 1, st)) {
            PEA_LAUNCH_CHECK("k_face_vjp");
            return PEA_OK;
        }
    }
    constexpr bool small = sizeof(T) > sizeof(double);
    const int tile = pick_tile(ld.g, small);
    if (tile < 0) return fail(PEA_ERR_DIMS, "no adjoint tile fits N=%d", ld.g.N);
    double* partial = gmach ? static_cast<double*>(ws) : nullptr;
    if (gmach && !ws) return fail(PEA_ERR_NULL, "gmach requested without workspace");
    int rc;
    if constexpr (small) {
        switch (tile) {
        case 3: rc = launch_vjp_shape<T, 12, 12>(P, ld, w, gU, gX, gV, partial, st); break;
        case 4: rc = launch_vjp_shape<T, 28, 4>(P, ld, w, gU, gX, gV, partial, st); break;
        case 5: rc = launch_vjp_shape<T, 12, 1>(P, ld, w, gU, gX, gV, partial, st); break;
        default: return fail(PEA_ERR_ARG, "tile %d not available here", tile);
        }
    } else {
        switch (tile) {
        case 0: rc = launch_vjp_shape<T, 28, 28>(P, ld, w, gU, gX, gV, partial, st); break;
        case 1: rc = launch_vjp_shape<T, 60, 12>(P, ld, w, gU, gX, gV, partial, st); break;
        case 2: rc = launch_vjp_shape<T, 28, 12>(P, ld, w, gU, gX, gV, partial, st); break;
        case 3: rc = launch_vjp_shape<T, 12, 12>(P, ld, w, gU, gX, gV, partial, st); break;
        case 4: rc = launch_vjp_shape<T, 28, 4>(P, ld, w, gU, gX, gV, partial, st); break;
        default: rc = launch_vjp_shape<T, 12, 1>(P, ld, w, gU, gX, gV, partial, st); break;
        }
    }
    if (rc) return rc;
    if (gmach) {
        k_sum_partials<<<1, 256, 0, st>>>(partial, (int)tile_blocks(ld.g, tile), gmach);
        PEA_LAUNCH_CHECK("k_sum_partials");
    }
    return PEA_OK;
}

}  // namespace

// =============================================================================== state checks
// Non-physical states: the reference evaluates p < 0 or rho < 0 in complex arithmetic, where the
// wave speed sqrt(gamma p / rho) becomes imaginary (f90:142-148); in FP64 it is NaN, and NaN or
// Inf in a state reaches the residual of every cell whose faces use it.  This pass counts the
// cells with a non-finite or non-physical state (rho <= 0 or p <= 0) into status[0] and those
// with a non-finite residual into status[1] (one atomic per wavefront, only where a lane counts).
__global__ void k_zero_ints(int* __restrict__ p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(256)
k_check_states(Grid g, const double* __restrict__ U, const double* __restrict__ R,
               int* __restrict__ status) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    bool bad_u = false, bad_r = false;
    if (c < g.C) {
        const double rho = U[c], jx = U[g.C + c], jz = U[2L * g.C + c], e = U[3L * g.C + c];
        const double p = (kGamma - 1.0) * (e - (jx * jx + jz * jz) / rho / 2.0);
        bad_u = !(isfinite(rho) && isfinite(jx) && isfinite(jz) && isfinite(e)) ||
                !(rho > 0.0) || !(p > 0.0);
        if (R)
            for (int v = 0; v < 4; ++v) bad_r = bad_r || !isfinite(R[(long)v * g.C + c]);
    }
    const unsigned long long bu = __ballot(bad_u), br = __ballot(bad_r);
    if ((threadIdx.x & 63) == 0) {
        if (bu) atomicAdd(&status[0], (int)__popcll(bu));
        if (br) atomicAdd(&status[1], (int)__popcll(br));
    }
}

// =============================================================================== C ABI
extern "C" {

const char* pea_last_error(void) { return g_err; }
int pea_version(void) { return 1; }

int pea_num_tiles(void) { return kNumTiles; }

int pea_set_tile(int M, int N, int tile) {
    if (tile < -1 || tile >= kNumTiles) return fail(PEA_ERR_ARG, "tile %d out of range", tile);
    std::lock_guard<std::mutex> lk(g_tile_mutex);
    if (M == 0 && N == 0) {
        g_tile_global = tile;
        return PEA_OK;
    }
    if (int rc = check_dims(M, N)) return rc;
    if (tile >= 0 && !tile_fits(make_grid(M, N), tile, false))
        return fail(PEA_ERR_ARG, "tile %d does not fit N=%d", tile, N);
    for (int i = 0; i < g_ntiles; ++i)
        if (g_tiles[i].M == M && g_tiles[i].N == N) {
            g_tiles[i].tile = tile;
            return PEA_OK;
        }
    if (tile < 0) return PEA_OK;
    if (g_ntiles == 64) return fail(PEA_ERR_ARG, "tile registry full");
    g_tiles[g_ntiles++] = TileChoice{M, N, tile};
    return PEA_OK;
}

int pea_num_face_tiles(void) { return kNumFaceTiles; }

int pea_set_face_tile(int kind, int M, int N, int tile) {
    if (kind < 0 || kind > 5) return fail(PEA_ERR_ARG, "kind %d out of range", kind);
    if (tile < -1 || tile >= kNumFaceTiles) return fail(PEA_ERR_ARG, "tile %d out of range", tile);
    std::lock_guard<std::mutex> lk(g_face_mutex);
    if (M == 0 && N == 0) {
        g_face_global[kind] = tile;
        return PEA_OK;
    }
    if (kind >= 3) return fail(PEA_ERR_ARG, "multi-grid tiles are process-wide (M = N = 0)");
    if (int rc = check_dims(M, N)) return rc;
    if (tile >= 0 && !face_tile_fits(kind, tile, N))
        return fail(PEA_ERR_ARG, "tile %d does not fit N=%d", tile, N);
    for (int i = 0; i < g_nface_tiles; ++i)
        if (g_face_tiles[i].kind == kind && g_face_tiles[i].M == M && g_face_tiles[i].N == N) {
            g_face_tiles[i].tile = tile;
            return PEA_OK;
        }
    if (tile < 0) return PEA_OK;
    if (g_nface_tiles == 256) return fail(PEA_ERR_ARG, "tile registry full");
    g_face_tiles[g_nface_tiles++] = FaceChoice{kind, M, N, tile};
    return PEA_OK;
}

int pea_get_face_tile(int kind, int M, int N, int* tm, int* tn) {
    if (kind < 0 || kind > 5) return fail(PEA_ERR_ARG, "kind %d out of range", kind);
    const int t = kind >= 3 ? face_tile(kind, 0, 1 << 30) : face_tile(kind, M, N);
    if (t >= 0 && tm && tn) {
        *tm = kFaceTiles[kind % 3][t].tm;
        *tn = kFaceTiles[kind % 3][t].tn;
    }
    return t;
}

int pea_get_tile(int M, int N) {
    if (check_dims(M, N)) return -1;
    return pick_tile(make_grid(M, N), false);
}

int pea_set_free_stream_state(const pea_params* prm, double mach, int M, int N, double* U,
                              void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !U) return fail(PEA_ERR_NULL, "null argument");
    const int C = M * N, bs = cell_block(C);
    k_free_stream<<<blocks_for(C, bs), bs, 0, as_stream(stream)>>>(*prm, mach, C, U);
    PEA_LAUNCH_CHECK("k_free_stream");
    return PEA_OK;
}

int pea_compute_odes(const pea_params* prm, double mach, int M, int N, const double* X,
                     const double* V, const double* U, double* R, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !R) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    launch_fwd_real(*prm, LdReal{X, V, U, mach, g}, R, 1, as_stream(stream));
    PEA_LAUNCH_CHECK("k_face_fwd");
    return PEA_OK;
}

int pea_compute_odes_batch(const pea_params* prm, double mach, int M, int N, int nstates,
                           const double* X, const double* V, const double* U, double* R,
                           void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !R) return fail(PEA_ERR_NULL, "null argument");
    if (nstates < 1 || nstates > 65535) return fail(PEA_ERR_ARG, "nstates %d", nstates);
    const Grid g = make_grid(M, N);
    launch_fwd_real(*prm, LdReal{X, V, U, mach, g}, R, nstates, as_stream(stream));
    PEA_LAUNCH_CHECK("k_face_fwd");
    return PEA_OK;
}

long long pea_multi_workspace_bytes(int G) { return G < 1 ? 0 : (long long)G * sizeof(GridDesc); }

int pea_multi_prepare(int G, const pea_grid* grids, int tangent, void* workspace,
                      int* total_tiles, void* stream) {
    if (G < 1) return fail(PEA_ERR_DIMS, "no grids");
    if (!grids || !workspace || !total_tiles) return fail(PEA_ERR_NULL, "null argument");
    if (tangent < 0 || tangent > 2) return fail(PEA_ERR_ARG, "mode %d", tangent);
    const int tile = face_tile(3 + tangent, 0, 1 << 30);
    const TileShape ts = kFaceTiles[tangent][tile < 0 ? 0 : tile];
    const int tm = ts.tm, tn = ts.tn;
    std::vector<GridDesc> d(G);
    long long total = 0;
    for (int i = 0; i < G; ++i) {
        const pea_grid& q = grids[i];
        if (int rc = check_dims(q.M, q.N)) return rc;
        if (!q.X || !q.V || !q.U || (tangent ? (!q.dU || !q.dR) : !q.R))
            return fail(PEA_ERR_NULL, "null array in a grid descriptor");
        if (tangent == 2 && q.N < tn + 3)
            return fail(PEA_ERR_DIMS, "multi-grid adjoint needs N >= %d (grid %d: N=%d)",
                        tn + 3, i, q.N);
        d[i].d = q;
        d[i].fs = fs_dir(q.prm);
        d[i].tiles_m = (q.M + tm - 1) / tm;
        d[i].tile0 = (int)total;
        total += (long long)d[i].tiles_m * ((q.N + tn - 1) / tn);
        if (total > 0x7fffffff) return fail(PEA_ERR_DIMS, "too many tiles");
    }
    // synchronous upload: the host table may be released as soon as this returns
    hipStream_t st = as_stream(stream);
    if (hipMemcpyAsync(workspace, d.data(), G * sizeof(GridDesc), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(PEA_ERR_HIP, "descriptor upload");
    *total_tiles = (int)total;
    std::lock_guard<std::mutex> lk(g_multi_mutex);
    for (auto& e : g_multi_tiles)
        if (e.ws == workspace) {
            e.mode = tangent;
            e.tile = tile < 0 ? 0 : tile;
            return PEA_OK;
        }
    g_multi_tiles.push_back(MultiTile{workspace, tangent, tile < 0 ? 0 : tile});
    return PEA_OK;
}

int pea_multi_launch(int G, int total_tiles, int tangent, const void* workspace, void* stream) {
    if (G < 1 || total_tiles < 1) return fail(PEA_ERR_DIMS, "no tiles");
    if (!workspace) return fail(PEA_ERR_NULL, "null argument");
    if (tangent < 0 || tangent > 2) return fail(PEA_ERR_ARG, "mode %d", tangent);
    const GridDesc* d = static_cast<const GridDesc*>(workspace);
    int tile = -1;
    {
        std::lock_guard<std::mutex> lk(g_multi_mutex);
        for (const auto& e : g_multi_tiles)
            if (e.ws == workspace && e.mode == tangent) tile = e.tile;
    }
    if (tile < 0) return fail(PEA_ERR_ARG, "workspace not prepared for mode %d", tangent);
    if (tangent == 2) {
        launch_vjp_multi(tile, d, G, total_tiles, as_stream(stream));
        PEA_LAUNCH_CHECK("k_face_vjp_multi");
        return PEA_OK;
    }
    if (tangent)
        launch_fwd_multi<true>(tile, d, G, total_tiles, as_stream(stream));
    else
        launch_fwd_multi<false>(tile, d, G, total_tiles, as_stream(stream));
    PEA_LAUNCH_CHECK("k_face_fwd_multi");
    return PEA_OK;
}

int pea_compute_odes_multi(int G, const pea_grid* grids, void* workspace, void* stream) {
    int total = 0;
    if (int rc = pea_multi_prepare(G, grids, 0, workspace, &total, stream)) return rc;
    return pea_multi_launch(G, total, 0, workspace, stream);
}

int pea_jvp_odes_multigrid(int G, const pea_grid* grids, void* workspace, void* stream) {
    int total = 0;
    if (int rc = pea_multi_prepare(G, grids, 1, workspace, &total, stream)) return rc;
    return pea_multi_launch(G, total, 1, workspace, stream);
}

int pea_check_states(int M, int N, const double* U, const double* R, int* status,
                     void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!U || !status) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    k_zero_ints<<<1, 64, 0, as_stream(stream)>>>(status, 2);
    PEA_LAUNCH_CHECK("k_zero_ints");
    k_check_states<<<blocks_for(g.C, 256), 256, 0, as_stream(stream)>>>(g, U, R, status);
    PEA_LAUNCH_CHECK("k_check_states");
    return PEA_OK;
}

int pea_compute_forces(int M, int N, const double* X, const double* U, double* F, double xc,
                       double zc, double* coef, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!X || !U || !F) return fail(PEA_ERR_NULL, "null argument");
    k_forces<<<1, 1024, 0, as_stream(stream)>>>(make_grid(M, N), X, U, F, xc, zc, coef);
    PEA_LAUNCH_CHECK("k_forces");
    return PEA_OK;
}

int pea_jvp_odes(const pea_params* prm, double mach, int M, int N, const double* X,
                 const double* V, const double* U, const double* dmach, const double* dX,
                 const double* dV, const double* dU, double* dR, double* R, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !dR) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    const int bs = cell_block(g.C);
    (void)bs;
    if (dU && !dX && !dV && !dmach)
        launch_jvp_u(*prm, LdDualU<2>{X, V, U, dU, mach, g}, StoreDualB<2>{dR, R, g.C}, 1,
                     as_stream(stream));
    else
        k_face_fwd<Dual<2>, Dual<2>, kJvpTM, kJvpTN><<<fwd_grid(g, 1, kJvpTM, kJvpTN),
                                                       kJvpTM * kJvpTN, 0, as_stream(stream)>>>(
            *prm, fs_dir(*prm), LdDual<2>{X, V, U, dX, dV, dU, dmach, mach, g},
            StoreDualB<2>{dR, R, g.C});
    PEA_LAUNCH_CHECK("k_face_fwd");
    return PEA_OK;
}

int pea_jvp_odes_states_dirs(const pea_params* prm, double mach, int M, int N, int k,
                             const double* X, const double* V, const double* U, const double* dU,
                             double* dR, int per, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !dU || !dR) return fail(PEA_ERR_NULL, "null argument");
    if (k < 1) return fail(PEA_ERR_ARG, "k must be >= 1 (got %d)", k);
    if (per < 0) return fail(PEA_ERR_ARG, "per must be >= 0 (got %d)", per);
    const Grid g = make_grid(M, N);
    constexpr int TM = 16, TN = 16;
    const long tiles = (long)((M + TM - 1) / TM) * ((N + TN - 1) / TN);
    // The buffer descriptors hold byte counts below 2^31: at most kmax complex directions per
    // launch.
    const int kmax = (int)std::max(1LL, ((1LL << 31) - 1) / (64LL * g.C));
    hipStream_t st = as_stream(stream);
    for (int d0 = 0; d0 < k; d0 += kmax) {
        const int kd = std::min(kmax, k - d0);
        // default: all of a launch's directions in each workgroup when the tiles alone fill the
        // GPU two workgroups per CU twice; else split them over enough workgroups to do so
        int pr = per;
        if (pr == 0) {
            const long want = 2L * 2 * 256;
            const long chunks = std::max(1L, std::min((long)kd, (want + tiles - 1) / tiles));
            pr = (int)((kd + chunks - 1) / chunks);
        }
        pr = std::min(pr, kd);
        const int nz = (kd + pr - 1) / pr;
        if (nz > 65535) return fail(PEA_ERR_ARG, "too many direction chunks (%d)", nz);
        const dim3 grid((M + TM - 1) / TM, (N + TN - 1) / TN, nz);
        k_face_dirs<TM, TN><<<grid, TM * TN, 0, st>>>(*prm, fs_dir(*prm),
                                                        LdReal{X, V, U, mach, g},
                                                        dU + (long)d0 * 8 * g.C,
                                                        dR + (long)d0 * 8 * g.C, kd, pr);
        PEA_LAUNCH_CHECK("k_face_dirs");
    }
    return PEA_OK;
}

int pea_jvp_odes_multi(const pea_params* prm, double mach, int M, int N, int k,
                       const double* X, const double* V, const double* U, const double* dmach,
                       const double* dX, const double* dV, const double* dU, double* dR,
                       void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !dR) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    const int bs = cell_block(g.C);
    const unsigned nb = blocks_for(g.C, bs);
    hipStream_t st = as_stream(stream);
    const bool u_only = dU && !dX && !dV && !dmach;
    if (u_only) {   // one face-kernel launch, one complex direction per grid z
        if (k < 1 || k > 65535) return fail(PEA_ERR_ARG, "k must be 1..65535 (got %d)", k);
        launch_jvp_u(*prm, LdDualU<2>{X, V, U, dU, mach, g}, StoreDualB<2>{dR, nullptr, g.C}, k,
                     st);
        PEA_LAUNCH_CHECK("k_face_fwd");
        return PEA_OK;
    }
    if (k != 1 && k != 2 && k != 4 && k != 8)
        return fail(PEA_ERR_ARG, "k must be 1, 2, 4 or 8 (got %d)", k);
    // directions are processed two at a time (four tangent lanes per pass)
    const long su = 2L * 4 * 2 * g.C, sx = 2L * 2 * 2 * g.P;
    for (int h = 0; h < k; h += 2) {
        const long o = h / 2;
        double* out = dR + o * su;
        if (k == 1) {
            if (u_only)
                k_jvp<2><<<nb, bs, 0, st>>>(*prm, LdDualU<2>{X, V, U, dU, mach, g}, out, nullptr);
            else
                k_jvp<2><<<nb, bs, 0, st>>>(
                    *prm, LdDual<2>{X, V, U, dX, dV, dU, dmach, mach, g}, out, nullptr);
        } else if (u_only) {
            k_jvp<4><<<nb, bs, 0, st>>>(*prm, LdDualU<4>{X, V, U, dU + o * su, mach, g}, out,
                                        nullptr);
        } else {
            k_jvp<4><<<nb, bs, 0, st>>>(
                *prm,
                LdDual<4>{X, V, U, dX ? dX + o * sx : nullptr, dV ? dV + o * sx : nullptr,
                          dU ? dU + o * su : nullptr, dmach ? dmach + 4 * o : nullptr, mach, g},
                out, nullptr);
        }
    }
    PEA_LAUNCH_CHECK("k_jvp");
    return PEA_OK;
}

int pea_jvp_forces(int M, int N, const double* X, const double* U, const double* dX,
                   const double* dU, double* dF, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!X || !U || !dF) return fail(PEA_ERR_NULL, "null argument");
    k_jvp_forces<<<blocks_for(N, 64), 64, 0, as_stream(stream)>>>(make_grid(M, N), X, U, dX, dU,
                                                                   dF);
    PEA_LAUNCH_CHECK("k_jvp_forces");
    return PEA_OK;
}

long long pea_vjp_workspace_bytes(int M, int N) {
    if (check_dims(M, N)) return 0;
    const Grid g = make_grid(M, N);
    long most = 1;
    for (int i = 0; i < kNumTiles; ++i)
        if (tile_fits(g, i, false) && tile_blocks(g, i) > most) most = tile_blocks(g, i);
    return (long long)most * 2 * (long long)sizeof(double);
}

int pea_vjp_odes(const pea_params* prm, double mach, int M, int N, const double* X,
                 const double* V, const double* U, const double* w, double* gU, double* gX,
                 double* gV, double* gmach, void* workspace, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !w) return fail(PEA_ERR_NULL, "null argument");
    return launch_vjp<double>(*prm, LdReal{X, V, U, mach, make_grid(M, N)}, w, gU, gX, gV,
                              gmach, workspace, as_stream(stream));
}

}  // extern "C"

// Partial sums of the Mach adjoint for pea_vjp_odes_noalloc: one complex value per workgroup of
// the cell-based adjoint.  Each GPU holds its own copy for the life of the process; host code
// reaches it only through the address from hipGetSymbolAddress (g_scratch_ptr).
constexpr long long kScratchBytes = 2LL << 20;   // 2 MiB: 131072 partial sums
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device (with that device
// current, as every call is made).
int scratch_address(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return hip_fail(hipGetLastError(), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(PEA_ERR_ARG, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PEA_OK;
}
}  // namespace

extern "C" {

int pea_vjp_odes_noalloc(const pea_params* prm, double mach, int M, int N, const double* X,
                         const double* V, const double* U, const double* w, double* gU,
                         double* gX, double* gV, double* gmach, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !w) return fail(PEA_ERR_NULL, "null argument");
    void* ws = nullptr;
    if (gmach) {
        const long long need = pea_vjp_workspace_bytes(M, N);
        if (need > kScratchBytes)
            return fail(PEA_ERR_DIMS, "pea_vjp_odes_noalloc: M=%d N=%d needs %lld bytes of "
                        "partial sums, the static scratch holds %lld", M, N, need, kScratchBytes);
        if (int rc = scratch_address(&ws)) return rc;
    }
    return launch_vjp<double>(*prm, LdReal{X, V, U, mach, make_grid(M, N)}, w, gU, gX, gV,
                              gmach, ws, as_stream(stream));
}

int pea_vjp_odes_states_multi(const pea_params* prm, double mach, int M, int N, int k,
                              const double* X, const double* V, const double* U, const double* w,
                              double* gU, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !w || !gU) return fail(PEA_ERR_NULL, "null argument");
    if (k < 1 || k > 65535) return fail(PEA_ERR_ARG, "k must be 1..65535 (got %d)", k);
    const Grid g = make_grid(M, N);
    const LdReal ld{X, V, U, mach, g};
    hipStream_t st = as_stream(stream);
    if (g_face_vjp && launch_vjp_u(*prm, ld, w, gU, k, st)) {   // one launch, seed per grid z
        PEA_LAUNCH_CHECK("k_face_vjp");
        return PEA_OK;
    }
    for (int i = 0; i < k; ++i) {   // grids too narrow for the face adjoint's tile
        const long o = (long)i * 8 * g.C;
        if (int rc = launch_vjp<double>(*prm, ld, w + o, gU + o, nullptr, nullptr, nullptr,
                                        nullptr, st))
            return rc;
    }
    return PEA_OK;
}

int pea_vjp_forces(int M, int N, const double* X, const double* U, const double* wF,
                   double* gU, double* gX, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!X || !U || !wF) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    hipStream_t st = as_stream(stream);
    const LdReal ld{X, X, U, 0.0, g};
    if (gU) {
        k_vjp_forces_states<double><<<blocks_for(g.C, 256), 256, 0, st>>>(ld, wF, gU);
        PEA_LAUNCH_CHECK("k_vjp_forces_states");
    }
    if (gX) {
        k_vjp_forces_vertices<double><<<blocks_for(g.P, 256), 256, 0, st>>>(ld, wF, gX);
        PEA_LAUNCH_CHECK("k_vjp_forces_vertices");
    }
    return PEA_OK;
}

int pea_hvp_forces(int M, int N, const double* X, const double* U, const double* wF,
                   const double* dX, const double* dU, double* hU, double* hX, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!X || !U || !wF) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    hipStream_t st = as_stream(stream);
    const LdDual<2> ld{X, X, U, dX, nullptr, dU, nullptr, 0.0, g};
    if (hU) {
        k_vjp_forces_states<Dual<2>><<<blocks_for(g.C, 256), 256, 0, st>>>(ld, wF, hU);
        PEA_LAUNCH_CHECK("k_vjp_forces_states");
    }
    if (hX) {
        k_vjp_forces_vertices<Dual<2>><<<blocks_for(g.P, 256), 256, 0, st>>>(ld, wF, hX);
        PEA_LAUNCH_CHECK("k_vjp_forces_vertices");
    }
    return PEA_OK;
}

int pea_hvp_odes(const pea_params* prm, double mach, int M, int N, const double* X,
                 const double* V, const double* U, const double* w, const double* dmach,
                 const double* dX, const double* dV, const double* dU, double* hU,
                 double* hX, double* hV, double* hmach, void* workspace, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !w) return fail(PEA_ERR_NULL, "null argument");
    return launch_vjp<Dual<2>>(*prm, LdDual<2>{X, V, U, dX, dV, dU, dmach, mach,
                                               make_grid(M, N)},
                               w, hU, hX, hV, hmach, workspace, as_stream(stream));
}

int pea_cs_jvp_odes(const pea_params* prm, double mach, int M, int N, const double* X,
                    const double* V, const double* U, const double* dmach, const double* dX,
                    const double* dV, const double* dU, double h, double* dR, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !dR) return fail(PEA_ERR_NULL, "null argument");
    if (!(h > 0.0)) h = kCsStep;
    const Grid g = make_grid(M, N);
    const int bs = cell_block(g.C);
    if (dU && !dX && !dV && !dmach)
        k_cs_jvp<<<blocks_for(g.C, bs), bs, 0, as_stream(stream)>>>(
            *prm, LdCxU{X, V, U, dU, mach, h, 0, g}, dR);
    else
        k_cs_jvp<<<blocks_for(g.C, bs), bs, 0, as_stream(stream)>>>(
            *prm, LdCx{X, V, U, dX, dV, dU, dmach, mach, h, 0, g}, dR);
    PEA_LAUNCH_CHECK("k_cs_jvp");
    return PEA_OK;
}

int pea_cs_jvp_odes_multi(const pea_params* prm, double mach, int M, int N, int k,
                          const double* X, const double* V, const double* U, const double* dU,
                          double h, double* dR, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U || !dU || !dR) return fail(PEA_ERR_NULL, "null argument");
    if (k < 1 || k > 65535) return fail(PEA_ERR_ARG, "k must be 1..65535 (got %d)", k);
    if (!(h > 0.0)) h = kCsStep;
    const Grid g = make_grid(M, N);
    const int bs = cell_block(g.C);
    k_cs_jvp<<<dim3(blocks_for(g.C, bs), k), bs, 0, as_stream(stream)>>>(
        *prm, LdCxU{X, V, U, dU, mach, h, 0, g}, dR);
    PEA_LAUNCH_CHECK("k_cs_jvp");
    return PEA_OK;
}

int pea_cs_linearize_odes(const pea_params* prm, double mach, int M, int N, const double* X,
                          const double* V, const double* U, double* JU, double* JX,
                          double* JV, double* JMa, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!prm || !X || !V || !U) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    k_cs_linearize<<<blocks_for(g.C, 64), 64, 0, as_stream(stream)>>>(
        *prm, LdReal{X, V, U, mach, g}, JU, JX, JV, JMa);
    PEA_LAUNCH_CHECK("k_cs_linearize");
    return PEA_OK;
}

int pea_apply_odes_wrt_states_fwd(int M, int N, const double* JU, const double* dU, double* dR,
                                  void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!JU || !dU || !dR) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    k_apply_states_fwd<<<blocks_for(g.C, 256), 256, 0, as_stream(stream)>>>(g, JU, dU, dR);
    PEA_LAUNCH_CHECK("k_apply_states_fwd");
    return PEA_OK;
}

int pea_apply_odes_wrt_states_rev(int M, int N, const double* JU, const double* w, double* gU,
                                  void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!JU || !w || !gU) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    k_apply_states_rev<<<blocks_for(g.C, 256), 256, 0, as_stream(stream)>>>(g, JU, w, gU);
    PEA_LAUNCH_CHECK("k_apply_states_rev");
    return PEA_OK;
}

int pea_apply_odes_wrt_vertices_fwd(int M, int N, const double* JX, const double* dX,
                                    double* dR, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!JX || !dX || !dR) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    k_apply_vertices_fwd<<<blocks_for(g.C, 256), 256, 0, as_stream(stream)>>>(g, JX, dX, dR);
    PEA_LAUNCH_CHECK("k_apply_vertices_fwd");
    return PEA_OK;
}

int pea_apply_odes_wrt_vertices_rev(int M, int N, const double* JX, const double* w,
                                    double* gX, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!JX || !w || !gX) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    k_apply_vertices_rev<<<blocks_for(g.P, 256), 256, 0, as_stream(stream)>>>(g, JX, w, gX);
    PEA_LAUNCH_CHECK("k_apply_vertices_rev");
    return PEA_OK;
}

int pea_export_bsr(int M, int N, const double* JU, double shift_re, double shift_im,
                   int complex_out, double* data, int* indices, int* indptr, void* stream) {
    if (int rc = check_dims(M, N)) return rc;
    if (!JU || !data || !indices || !indptr) return fail(PEA_ERR_NULL, "null argument");
    const Grid g = make_grid(M, N);
    const int cplx = (complex_out || shift_im != 0.0) ? 1 : 0;
    k_export_bsr<<<blocks_for(g.C, 256), 256, 0, as_stream(stream)>>>(
        g, JU, shift_re, shift_im, cplx, data, indices, indptr);
    PEA_LAUNCH_CHECK("k_export_bsr");
    return PEA_OK;
}

}  // extern "C"
