// erfx_common.h -- shared device/host definitions of the HIP dycore library.
//
// All fields of a box share one layout (include/erfx.h): i fastest, ng ghost
// cells on every side, one extra face in each direction.  Kernels address every
// field with the same linear index, which lets a fused kernel stream several
// fields with one address computation and keeps every access x-coalesced.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/erfx.h"
#include "erfx_pow.h"

namespace erfx {

// ---- physical constants (Source/ERF_Constants.H:7-20) ---------------------
constexpr double PI = 3.14159265358979323846264338327950288;
constexpr double PIoTwo = PI / 2.0;
constexpr double R_d = 287.0;
constexpr double R_v = 461.505;
constexpr double Cp_d = 1004.5;
constexpr double p_0 = 1.0e5;
constexpr double Gamma = 1.4;
constexpr double CONST_GRAV = 9.81;
constexpr double ip_0 = 1. / p_0;
constexpr double L_v = 2.5e6;          // ERF_Constants.H:16
constexpr double lcond = 2.501e6;      // ERF_Constants.H:66
constexpr double qcw0 = 1.e-3;         // ERF_Constants.H:46
constexpr double alphaelq = 1.e-3;     // ERF_Constants.H:48
constexpr double Rd_on_Rv = R_d / R_v; // ERF_Constants.H:87

// ---- EOS (Source/Utils/ERF_EOS.H:89-93, 130-135), dry ---------------------
__device__ __forceinline__ double eos_p_of_rt(double rt) {
    return p_0 * erfx_pow(R_d * rt * (1.0 + (R_v / R_d) * 0.0) * ip_0, Gamma);
}
__device__ __forceinline__ double eos_exner_of_rt(double rt, double rdOcp) {
    return erfx_pow(R_d * rt * (1.0 + R_v / R_d * 0.0) * ip_0, Gamma * rdOcp);
}
// moist forms: the pressure of the moist potential temperature theta (1 + R_v/R_d qv)
// (ERF_EOS.H:76-79) and the temperature (ERF_EOS.H:42-51)
__device__ __forceinline__ double eos_p_of_rt_qv(double rt, double qv) {
    return p_0 * erfx_pow(R_d * rt * (1.0 + (R_v / R_d) * qv) * ip_0, Gamma);
}
__device__ __forceinline__ double eos_T(double rho, double rt, double qv) {
    double p_loc = p_0 * erfx_pow(R_d * rt * (1.0 + R_v / R_d * qv) * ip_0, Gamma);
    return p_loc / (R_d * rho * (1.0 + R_v / R_d * qv));
}

// ---- index space ------------------------------------------------------------
struct Lay {
    int nx, ny, nz, ng;
    long sy, sz, ne;
    __host__ __device__ __forceinline__ long ix(int i, int j, int k) const {
        return (long)(i + ng) + (long)(j + ng) * sy + (long)(k + ng) * sz;
    }
};

struct Box {
    int lo[3], hi[3];
    __host__ __device__ Box() : lo{0, 0, 0}, hi{-1, -1, -1} {}
    __host__ __device__ Box(int l0, int l1, int l2, int h0, int h1, int h2) : lo{l0, l1, l2}, hi{h0, h1, h2} {}
    __host__ __device__ bool ok() const { return lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]; }
    __host__ __device__ long npts() const {
        return ok() ? (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) : 0;
    }
};

// Work-groups are dispatched to the 8 XCDs of the chip round-robin by
// linear id, and every XCD has its own L2.  xcd_remap turns the hardware id w
// into a logical id so that each XCD receives one contiguous range of logical
// ids (a bijection on [0, nb)): tiles that share halos (the next x-tile, the
// next rows, the planes above and below) then run on the same XCD and share
// its L2 instead of each fetching the halo from memory.
constexpr unsigned NUM_XCD = 8;
__device__ __forceinline__ unsigned xcd_remap(unsigned w, unsigned nb) {
    const unsigned x = w % NUM_XCD, local = w / NUM_XCD;
    const unsigned q = nb / NUM_XCD, r = nb % NUM_XCD;
    return x * q + (x < r ? x : r) + local;
}

// Generic box launch: x-lanes along i (64-wide waves), 4 rows of j per block,
// tiles ordered x, y, z; every lane checks its point against the box, so any
// grid covering the box is safe.
// ---- batched launches (many boxes in one kernel) ---------------------------
// While a Recorder is active on the calling thread, launch_box, launch_cols and
// copy_fields append their work item (the tile ranges and the functor, by
// value) to it instead of launching.  A batch then launches the n-th item of
// every member box in ONE kernel: the work-group ids are split into
// consecutive ranges, one per member, and each work-group finds its member by
// a binary search over the ranges' first ids (a uniform value, so the search
// and the functor's captures are scalar loads).
struct BHdr {
    int lo[3], hi[3];
    unsigned gx, gy;      // tiles along x and y (column launches: along x only)
    unsigned t0, tiles;   // the member's first work-group id in the batch, and its count
};
// launches `total` work-groups of `width` lanes over n items stored at
// `items`, whose first work-group ids are starts[0..n-1]
using BatchLaunch = void (*)(const void* items, const unsigned* starts, int n, unsigned total, unsigned width,
                             hipStream_t s);
struct RecItem {
    BatchLaunch launch;   // one instance per item type, so it also identifies the kernel
    unsigned width;       // work-group size
    size_t size, align;   // of the item type
    std::vector<unsigned char> bytes;
};
struct Recorder {
    std::vector<RecItem> items;
    const char* refused = nullptr;   // set by a launch that cannot be recorded
};
// the calling thread's active recorder (host only), or null
extern thread_local Recorder* g_rec;
// for host functions whose launches cannot be batched: true (and the recorder
// marked) when recording
inline bool rec_refuse(const char* what) {
    if (!g_rec) return false;
    if (!g_rec->refused) g_rec->refused = what;
    return true;
}
template <class It>
inline void rec_push(BatchLaunch fn, unsigned width, const It& item) {
    static_assert(std::is_trivially_copyable<It>::value, "batched items are copied bytewise");
    RecItem r;
    r.launch = fn;
    r.width = width;
    r.size = sizeof(It);
    r.align = alignof(It);
    r.bytes.resize(sizeof(It));
    std::memcpy(r.bytes.data(), &item, sizeof(It));
    g_rec->items.push_back(std::move(r));
}
// the item of work-group t: the last one whose range starts at or before t
// (members with no work-groups share their start with the next and are
// skipped).  The starts are a compact array of their own, so the search reads
// one or two cache lines that every work-group of the launch shares.
// The items are read through a __restrict__ pointer to const: no store of the
// kernel can reach them, so the captures are loaded as scalar (uniform) loads
// where the functor uses them, as from a kernel argument.  (A local copy of
// the functor would move captured arrays indexed at run time into scratch.)
// Every functor call is forced inline: with two kernels per functor the
// inliner would otherwise leave large bodies as calls, with their spills.
template <class It>
__device__ __forceinline__ const It& batch_member(const It* it, const unsigned* __restrict__ starts, int n,
                                                  unsigned t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return it[lo];
}

template <class F>
__global__ void __launch_bounds__(256) box_kernel(Box b, unsigned gx, unsigned gy, F f) {
    const unsigned nb = gridDim.x;
    const unsigned t = xcd_remap(blockIdx.x, nb);
    const unsigned bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    int i = b.lo[0] + (int)(bx * 64 + threadIdx.x);
    int j = b.lo[1] + (int)(by * 4 + threadIdx.y);
    int k = b.lo[2] + (int)bz;
    if (i > b.hi[0] || j > b.hi[1] || k > b.hi[2]) return;
    [[clang::always_inline]] f(i, j, k);
}

template <class F>
struct BoxItem {
    BHdr h;
    F f;
};
template <class F>
__global__ void __launch_bounds__(256) box_kernel_batch(const BoxItem<F>* __restrict__ it, const unsigned* __restrict__ starts, int n) {
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const BoxItem<F>& m = batch_member(it, starts, n, t);
    const BHdr h = m.h;
    const F& f = m.f;
    const unsigned u = t - h.t0;
    const unsigned bx = u % h.gx, by = (u / h.gx) % h.gy, bz = u / (h.gx * h.gy);
    int i = h.lo[0] + (int)(bx * 64 + threadIdx.x);
    int j = h.lo[1] + (int)(by * 4 + threadIdx.y);
    int k = h.lo[2] + (int)bz;
    if (i > h.hi[0] || j > h.hi[1] || k > h.hi[2]) return;
    [[clang::always_inline]] f(i, j, k);
}
template <class F>
void launch_box_batch(const void* items, const unsigned* starts, int n, unsigned total, unsigned width,
                      hipStream_t s) {
    (void)width;
    hipLaunchKernelGGL(box_kernel_batch<F>, dim3(total), dim3(64, 4, 1), 0, s, static_cast<const BoxItem<F>*>(items),
                       starts, n);
}

template <class F>
inline void launch_box(const Box& b, hipStream_t s, F f) {
    const bool ok = b.ok();
    const unsigned gx = ok ? (unsigned)((b.hi[0] - b.lo[0] + 1 + 63) / 64) : 1u;
    const unsigned gy = ok ? (unsigned)((b.hi[1] - b.lo[1] + 1 + 3) / 4) : 1u;
    const unsigned gz = ok ? (unsigned)(b.hi[2] - b.lo[2] + 1) : 0u;
    if (g_rec) {   // recorded even when empty, so every member records the same sequence
        BoxItem<F> it{{{b.lo[0], b.lo[1], b.lo[2]}, {b.hi[0], b.hi[1], b.hi[2]}, gx, gy, 0u, gx * gy * gz}, f};
        rec_push(&launch_box_batch<F>, 256u, it);
        return;
    }
    if (!ok) return;
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(box_kernel<F>, dim3(gx * gy * gz), blk, 0, s, b, gx, gy, f);
}

// Column launch over a 2-D (i, j) range; the functor loops over k itself.
// Work-groups are `width` lanes along x (256 by default, so each field-level
// access of a group is one contiguous 2 KB row segment).
template <class F>
__global__ void __launch_bounds__(256) col_kernel(int ilo, int ihi, int jlo, int jhi, unsigned gx, F f) {
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    int i = ilo + (int)((t % gx) * blockDim.x + threadIdx.x);
    int j = jlo + (int)(t / gx);
    if (i > ihi || j > jhi) return;
    [[clang::always_inline]] f(i, j);
}

// The same box launch with the tiles ordered k fastest: consecutive
// work-groups take the same (x, y) tile at consecutive levels, so a kernel
// whose stencils reach several levels up and down finds them in L2 (with x
// fastest, the levels of a stencil are a whole plane of tiles apart).
template <class F>
__global__ void __launch_bounds__(256) box_kernel_z(Box b, unsigned gx, unsigned gz, F f) {
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const unsigned bz = t % gz, r = t / gz, bx = r % gx, by = r / gx;
    int i = b.lo[0] + (int)(bx * 64 + threadIdx.x);
    int j = b.lo[1] + (int)(by * 4 + threadIdx.y);
    int k = b.lo[2] + (int)bz;
    if (i > b.hi[0] || j > b.hi[1] || k > b.hi[2]) return;
    [[clang::always_inline]] f(i, j, k);
}
template <class F>
__global__ void __launch_bounds__(256) box_kernel_z_batch(const BoxItem<F>* __restrict__ it,
                                                          const unsigned* __restrict__ starts, int n) {
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const BoxItem<F>& m = batch_member(it, starts, n, t);
    const BHdr h = m.h;
    const F& f = m.f;
    const unsigned u = t - h.t0;
    const unsigned gz = (unsigned)(h.hi[2] - h.lo[2] + 1);
    const unsigned bz = u % gz, r = u / gz, bx = r % h.gx, by = r / h.gx;
    int i = h.lo[0] + (int)(bx * 64 + threadIdx.x);
    int j = h.lo[1] + (int)(by * 4 + threadIdx.y);
    int k = h.lo[2] + (int)bz;
    if (i > h.hi[0] || j > h.hi[1] || k > h.hi[2]) return;
    [[clang::always_inline]] f(i, j, k);
}
template <class F>
void launch_box_z_batch(const void* items, const unsigned* starts, int n, unsigned total, unsigned width,
                        hipStream_t s) {
    (void)width;
    hipLaunchKernelGGL(box_kernel_z_batch<F>, dim3(total), dim3(64, 4, 1), 0, s,
                       static_cast<const BoxItem<F>*>(items), starts, n);
}
template <class F>
inline void launch_box_z(const Box& b, hipStream_t s, F f) {
    const bool ok = b.ok();
    const unsigned gx = ok ? (unsigned)((b.hi[0] - b.lo[0] + 1 + 63) / 64) : 1u;
    const unsigned gy = ok ? (unsigned)((b.hi[1] - b.lo[1] + 1 + 3) / 4) : 1u;
    const unsigned gz = ok ? (unsigned)(b.hi[2] - b.lo[2] + 1) : 0u;
    if (g_rec) {
        BoxItem<F> it{{{b.lo[0], b.lo[1], b.lo[2]}, {b.hi[0], b.hi[1], b.hi[2]}, gx, gy, 0u, gx * gy * gz}, f};
        rec_push(&launch_box_z_batch<F>, 256u, it);
        return;
    }
    if (!ok) return;
    hipLaunchKernelGGL(box_kernel_z<F>, dim3(gx * gy * gz), dim3(64, 4, 1), 0, s, b, gx, gz, f);
}

// Work-group launch for block-level reductions: f(u) runs on all 256 lanes of
// work-group u of a 1-D grid.
template <class F>
__global__ void __launch_bounds__(256) grid_kernel(F f) {
    [[clang::always_inline]] f(blockIdx.x);
}
template <class F>
struct GridItem {
    BHdr h;
    F f;
};
template <class F>
__global__ void __launch_bounds__(256) grid_kernel_batch(const GridItem<F>* __restrict__ it, const unsigned* __restrict__ starts, int n) {
    const GridItem<F>& m = batch_member(it, starts, n, blockIdx.x);
    const unsigned t0 = m.h.t0;
    const F& f = m.f;
    [[clang::always_inline]] f(blockIdx.x - t0);
}
template <class F>
void launch_grid_batch(const void* items, const unsigned* starts, int n, unsigned total, unsigned width,
                       hipStream_t s) {
    (void)width;
    hipLaunchKernelGGL(grid_kernel_batch<F>, dim3(total), dim3(256), 0, s, static_cast<const GridItem<F>*>(items),
                       starts, n);
}
template <class F>
inline void launch_grid(unsigned blocks, hipStream_t s, F f) {
    if (g_rec) {
        GridItem<F> it{{{0, 0, 0}, {0, 0, 0}, 1u, 1u, 0u, blocks}, f};
        rec_push(&launch_grid_batch<F>, 256u, it);
        return;
    }
    if (blocks == 0) return;
    hipLaunchKernelGGL(grid_kernel<F>, dim3(blocks), dim3(256), 0, s, f);
}

template <class F>
struct ColItem {
    BHdr h;
    F f;
};
// Batched columns run over the member's (i, j) range flattened, i fastest, so
// a box narrower than the work-group fills its lanes with the next rows.
template <class F>
__global__ void __launch_bounds__(256) col_kernel_batch(const ColItem<F>* __restrict__ it, const unsigned* __restrict__ starts, int n) {
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const ColItem<F>& m = batch_member(it, starts, n, t);
    const BHdr h = m.h;
    const F& f = m.f;
    const unsigned nxr = (unsigned)(h.hi[0] - h.lo[0] + 1);
    const unsigned c = (t - h.t0) * blockDim.x + threadIdx.x;
    if (c >= nxr * (unsigned)(h.hi[1] - h.lo[1] + 1)) return;
    const int i = h.lo[0] + (int)(c % nxr);
    const int j = h.lo[1] + (int)(c / nxr);
    [[clang::always_inline]] f(i, j);
}
template <class F>
void launch_cols_batch(const void* items, const unsigned* starts, int n, unsigned total, unsigned width,
                       hipStream_t s) {
    hipLaunchKernelGGL(col_kernel_batch<F>, dim3(total), dim3(width, 1, 1), 0, s, static_cast<const ColItem<F>*>(items),
                       starts, n);
}

template <class F>
inline void launch_cols(int ilo, int ihi, int jlo, int jhi, hipStream_t s, F f, int width = 256) {
    const bool ok = ilo <= ihi && jlo <= jhi;
    const unsigned gx = ok ? (unsigned)((ihi - ilo + 1 + width - 1) / width) : 1u;
    const unsigned gy = ok ? (unsigned)(jhi - jlo + 1) : 0u;
    if (g_rec) {
        const unsigned long long cols = ok ? (unsigned long long)(ihi - ilo + 1) * (unsigned long long)(jhi - jlo + 1) : 0;
        ColItem<F> it{{{ilo, jlo, 0}, {ihi, jhi, 0}, gx, 1u, 0u, (unsigned)((cols + width - 1) / width)}, f};
        rec_push(&launch_cols_batch<F>, (unsigned)width, it);
        return;
    }
    if (!ok) return;
    hipLaunchKernelGGL(col_kernel<F>, dim3(gx * gy), dim3((unsigned)width, 1, 1), 0, s, ilo, ihi, jlo, jhi, gx, f);
}

// ---- column tiles: one work-group of NW waves per 64 columns -------------
// For column solves whose per-level work is independent except for a
// tridiagonal recurrence: the waves split the levels into chunks and form the
// per-level terms of their chunk in parallel, one wave runs the recurrence
// over the tile's 64 columns from LDS, and the waves finish their chunks in
// parallel again.  Lane l of every wave holds column tile * 64 + l of the
// (i, j) range flattened i fastest (a box narrower than 64 fills the lanes
// with the next rows), so every level's access of a wave is one contiguous
// row segment.  The functor is called on all lanes, `valid` false past the
// range's end, so that every lane reaches the work-group barriers inside it;
// `lds` holds three arrays of tile_rows<CM, NW>() rows of 64
// doubles, one value per level and column.
template <int CM, int NW>
constexpr int tile_rows() { return NW * CM + 1; }
template <int CM, int NW, int TC = 64>
constexpr int tile_lds_doubles() { return 3 * tile_rows<CM, NW>() * TC; }
// levels per wave chunk for nz levels, and the smallest instantiated CM >= it
// at 8 waves (0: nz above 96, no tile form: three arrays of 97 rows are
// 149 KB of the CU's 160 KB of LDS)
__host__ __device__ inline int tile_chunk(int nz, int nw) { return (nz + nw - 1) / nw; }
inline int tile_cm8(int nz) {
    const int c = tile_chunk(nz, 8);
    return c <= 4 ? 4 : c <= 8 ? 8 : c <= 12 ? 12 : 0;
}
// Ends one level of a wave's chunk: keeps the compiler from hoisting the next
// levels' loads above it (one level of loads per wave covers the latency with
// the tile's waves in flight, while hoisting several levels exceeds a lane's
// registers and spills).
__device__ __forceinline__ void tile_level_fence() { asm volatile("" ::: "memory"); }

// The tridiagonal sweeps of a tile over its LDS rows (row k of column `lane`
// at k * TC + lane, TC columns per tile).  Forward: X holds the right-hand sides of faces 0..nz, Y
// the sub-diagonal a_k and Z the inverted pivots b_k; X becomes s_0 = X_0 b_0,
// s_k = (X_k - a_k s_{k-1}) b_k.  Backward: Y holds c_k b_k; X becomes
// x_nz = s_nz, x_k = s_k - (c_k b_k) x_{k+1}.  Both keep the operation order
// of a per-column Thomas sweep, so the values are the same; the reads of the
// next U levels are issued before their dependent arithmetic.
template <int TC = 64>
__device__ __forceinline__ void tile_forward(double* X, const double* Y, const double* Z, int nz, int lane) {
    constexpr int U = 8;
    double s = X[lane] * Z[lane];
    X[lane] = s;
    int k = 1;
    for (; k + U <= nz + 1; k += U) {
        double a[U], b[U], r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int o = (k + u) * TC + lane;
            a[u] = Y[o]; b[u] = Z[o]; r[u] = X[o];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            s = (r[u] - a[u] * s) * b[u];
            X[(k + u) * TC + lane] = s;
        }
    }
    for (; k <= nz; ++k) {
        const int o = k * TC + lane;
        s = (X[o] - Y[o] * s) * Z[o];
        X[o] = s;
    }
}
template <int TC = 64>
__device__ __forceinline__ void tile_backward(double* X, const double* Y, int nz, int lane) {
    constexpr int U = 8;
    double s = X[nz * TC + lane];
    int k = nz - 1;
    for (; k - U + 1 >= 0; k -= U) {
        double cb[U], r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int o = (k - u) * TC + lane;
            cb[u] = Y[o]; r[u] = X[o];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            s = r[u] - cb[u] * s;
            X[(k - u) * TC + lane] = s;
        }
    }
    for (; k >= 0; --k) {
        const int o = k * TC + lane;
        s = X[o] - Y[o] * s;
        X[o] = s;
    }
}

// TW = 64: a tile is 64 consecutive columns of the flattened range; TW = 32,
// 16, 8: a tile is a TW x (64 / TW) block of columns (tiles ordered x
// fastest), whose lanes then find their j-neighbours in the same tile.  (The
// column kernels of the substep measure fastest with TW = 64 on 512 x 512 x 96:
// 2.18 ms against 2.76 ms at 32 x 2 and 2.86 ms at 16 x 4, which fetch 4% and
// 10% more -- a row segment of TW < 64 doubles spans a cache line more, in
// proportion, than a whole row does.)
template <int CM, int NW, int TW, class F>
__global__ void __launch_bounds__(64 * NW) tile_kernel(int ilo, int ihi, int jlo, int jhi, unsigned nxr,
                                                       unsigned ncols, unsigned gx, F f) {
    __shared__ double lds[tile_lds_doubles<CM, NW>()];
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int i, j;
    bool valid;
    if (TW == 64) {
        const unsigned c = t * 64u + (unsigned)lane;
        valid = c < ncols;
        i = ilo + (int)(c % nxr);
        j = jlo + (int)(c / nxr);
    } else {
        i = ilo + (int)((t % gx) * TW) + (lane % TW);
        j = jlo + (int)((t / gx) * (64 / TW)) + (lane / TW);
        valid = i <= ihi && j <= jhi;
    }
    [[clang::always_inline]] f(i, j, valid, lane, wave, lds);
}
template <class F>
struct TileItem {
    BHdr h;
    F f;
};
template <int CM, int NW, class F>
__global__ void __launch_bounds__(64 * NW) tile_kernel_batch(const TileItem<F>* __restrict__ it,
                                                             const unsigned* __restrict__ starts, int n) {
    __shared__ double lds[tile_lds_doubles<CM, NW>()];
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const TileItem<F>& m = batch_member(it, starts, n, t);
    const BHdr h = m.h;
    const F& f = m.f;
    const unsigned nxr = (unsigned)(h.hi[0] - h.lo[0] + 1);
    const unsigned ncols = nxr * (unsigned)(h.hi[1] - h.lo[1] + 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned c = (t - h.t0) * 64u + (unsigned)lane;
    const bool valid = c < ncols;
    const int i = h.lo[0] + (int)(c % nxr), j = h.lo[1] + (int)(c / nxr);
    [[clang::always_inline]] f(i, j, valid, lane, wave, lds);
}
template <int CM, int NW, class F>
void launch_tiles_batch(const void* items, const unsigned* starts, int n, unsigned total, unsigned width,
                        hipStream_t s) {
    (void)width;
    hipLaunchKernelGGL((tile_kernel_batch<CM, NW, F>), dim3(total), dim3(64 * NW), 0, s,
                       static_cast<const TileItem<F>*>(items), starts, n);
}
// Tiles over the columns (ilo..ihi) x (jlo..jhi); CM >= tile_chunk(nz, NW) of
// the functor's box.  A recorded (batched) launch always uses flattened tiles.
template <int CM, int NW, int TW = 64, class F>
inline void launch_tiles(int ilo, int ihi, int jlo, int jhi, hipStream_t s, F f) {
    static_assert(TW == 64 || TW == 32 || TW == 16 || TW == 8, "tile width");
    const bool ok = ilo <= ihi && jlo <= jhi;
    const unsigned nxr = ok ? (unsigned)(ihi - ilo + 1) : 1u;
    const unsigned nyr = ok ? (unsigned)(jhi - jlo + 1) : 1u;
    const unsigned ncols = ok ? nxr * nyr : 0u;
    if (g_rec) {
        const unsigned tiles = (ncols + 63u) / 64u;
        TileItem<F> it{{{ilo, jlo, 0}, {ihi, jhi, 0}, 1u, 1u, 0u, tiles}, f};
        rec_push(&launch_tiles_batch<CM, NW, F>, (unsigned)(64 * NW), it);
        return;
    }
    if (!ok) return;
    const unsigned gx = (nxr + TW - 1) / TW;
    const unsigned tiles = TW == 64 ? (ncols + 63u) / 64u : gx * ((nyr + 64 / TW - 1) / (64 / TW));
    hipLaunchKernelGGL((tile_kernel<CM, NW, TW, F>), dim3(tiles), dim3(64 * NW), 0, s, ilo, ihi, jlo, jhi, nxr,
                       ncols, gx, f);
}

// Narrow tiles for boxes of few columns: a work-group of 8 waves takes TC
// (16 or 32) consecutive columns of the flattened range and splits their
// levels into 512 / TC chunks; thread t works on column t % TC in chunk
// t / TC, so the functor sees `lane` = the column within the tile and `wave`
// = the chunk (chunk 0, the first TC threads, runs the serial sweeps).  The
// 4,096 columns of a 64 x 64 box make 256 tiles of 16 instead of 64 of 64.
template <int CM, int TC, class F>
__global__ void __launch_bounds__(512) tile_kernel_n(int ilo, int jlo, unsigned nxr, unsigned ncols, F f) {
    __shared__ double lds[tile_lds_doubles<CM, 512 / TC, TC>()];
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = (int)threadIdx.x % TC, chunk = (int)threadIdx.x / TC;
    const unsigned c = t * (unsigned)TC + (unsigned)lane;
    const bool valid = c < ncols;
    const int i = ilo + (int)(c % nxr), j = jlo + (int)(c / nxr);
    [[clang::always_inline]] f(i, j, valid, lane, chunk, lds);
}
template <int CM, int TC, class F>
inline void launch_tiles_n(int ilo, int ihi, int jlo, int jhi, hipStream_t s, F f) {
    static_assert(TC == 16 || TC == 32, "narrow tile width");
    if (ilo > ihi || jlo > jhi) return;
    const unsigned nxr = (unsigned)(ihi - ilo + 1);
    const unsigned ncols = nxr * (unsigned)(jhi - jlo + 1);
    hipLaunchKernelGGL((tile_kernel_n<CM, TC, F>), dim3((ncols + TC - 1) / TC), dim3(512), 0, s, ilo, jlo, nxr, ncols,
                       f);
}
// columns per tile for a box of ncols columns: 64 while that gives every CU a
// tile, else 32 or 16
inline int tile_columns(long ncols) {
    constexpr long CUS = 256;
    if ((ncols + 63) / 64 >= CUS) return 64;
    if ((ncols + 31) / 32 >= CUS) return 32;
    return 16;
}

// Variants of the acoustic column kernel (erf_fast_rhs_N): levels of loads
// kept in flight per lane and work-group width.  Every variant computes the
// same values; the sim's autotuner times them on the first launches of a run
// and keeps the fastest.
// Variant COL_TILE is the tile form (launch_tiles; launch_tiles_n for boxes
// of few columns, which may have up to 128 levels); COL_FALLBACK for boxes
// the tile form does not take.
constexpr int COL_NV = ERFX_COL_VARIANTS;
constexpr int COL_DEPTH[COL_NV] = {2, 3, 4, 2, 3, 4, 2, 3, 4, 0};
constexpr int COL_WIDTH[COL_NV] = {64, 64, 64, 128, 128, 128, 256, 256, 256, 0};
constexpr int COL_TILE = 9;
constexpr int COL_FALLBACK = 7;  // depth 3, 256 lanes
constexpr int COL_DEFAULT = COL_TILE;
struct ColTuner {
    static constexpr int SAMPLES = 3;
    hipEvent_t ev[2 * COL_NV * SAMPLES] = {};
    int next = 0;          // launches timed so far
    int chosen = -1;       // variant picked, -1 while timing
    float ms[COL_NV] = {}; // best time per variant
};

// ---- boundary-condition kinds (AMReX BCType + ERFBCType) -------------------
enum BCK { INT_DIR = 0, EXT_DIR, FOEXTRAP, REFLECT_EVEN, REFLECT_ODD, NEUMANN_INT };

// Per-launch timing of one chosen kernel with HIP events recorded on the
// kernel's own stream around each of its launches (host side only).
enum { KT_FAST_COLUMN = 0 };
struct KTimer {
    int kernel = -1;                 // KT_* id being timed, -1 = off
    hipEvent_t* ev = nullptr;        // start/stop pairs
    int cap = 0;                     // events available
    int used = 0;                    // events recorded
    double bytes = 0;                // algorithmic HBM bytes per launch
    const char* name = "";
};
inline void kt_start(KTimer* t, int id, hipStream_t s) {
    if (t && t->kernel == id && t->used + 2 <= t->cap) (void)hipEventRecord(t->ev[t->used], s);
}
inline void kt_stop(KTimer* t, int id, hipStream_t s, double bytes, const char* name) {
    if (t && t->kernel == id && t->used + 2 <= t->cap) {
        (void)hipEventRecord(t->ev[t->used + 1], s);
        t->used += 2;
        t->bytes = bytes;
        t->name = name;
    }
}

// Per-run constants derived from erfx_config.
struct Params {
    erfx_config cfg;
    Lay L;
    int ng_cons, ng_vel;
    double rdOcp, gravity;
    double dxInv[3];
    int bck[4][6];            // [x/y/z velocity, cons][ori]
    double vel_extdir[3][6];
    double cons_extdir[ERFX_NCONS_MOIST][6];   // ext_dir values of the cons components (inflow faces)
    bool periodic[3];
    bool use_diff, use_turb, consA;
    KTimer* timer = nullptr;  // optional launch timing (host pointer, never read on the device)
    const double* ray = nullptr;  // Rayleigh table (device), RAY_* blocks of nz+1 values
    bool moist = false;       // a Kessler model: cons comps 5..7 are RhoQ1..RhoQ3
    int n_qstate = 0;         // 3 with Kessler (ERF_Kessler.H:172)
    // terrain-following grid (use_terrain): node heights z_phys_nd and the
    // metrics detJ, face areas ax, ay, az and cell-centre heights (device
    // fields of the layout); null on a flat grid
    bool terrain = false;
    const double* z_nd = nullptr;
    const double* detJ = nullptr;
    const double* ax = nullptr;
    const double* ay = nullptr;
    const double* az = nullptr;
    const double* z_cc = nullptr;
    // h_xi / h_zeta on x-faces and h_eta / h_zeta on y-faces of the valid box,
    // the factors of the pressure-gradient metric term (static terrain)
    const double* mrx = nullptr;
    const double* mry = nullptr;
    double dz_min = 0.0;      // dz, times the smallest detJ of the valid cells on a terrain grid
    // terrain grid with Rayleigh damping: factors and flags per point (the
    // heights vary along a level), RPT_* blocks of ne values
    const double* ray_pt = nullptr;
    int col_variant = COL_DEFAULT;   // acoustic column kernel variant (COL_DEPTH / COL_WIDTH)
    ColTuner* tuner = nullptr;       // when timing variants (host pointer)
};

// prim components of the moisture species (prim(n) = cons(n+1) / rho)
constexpr int PRIM_Q1 = ERFX_RHOQ1 - 1, PRIM_Q2 = ERFX_RHOQ2 - 1, PRIM_Q3 = ERFX_RHOQ3 - 1;
// 1 + qv + qc on the face between cells a and b, the moist divisor of the
// pressure gradients (ERF_slow_rhs_pre.cpp:622-626; ERF_fast_rhs_N.cpp:215-219;
// ERF_make_fast_coeffs.cpp:183-188): q = 0.5 (qv_a + qv_b + qc_a + qc_b)
__device__ __forceinline__ double face_q(const double* pq1, const double* pq2, long a, long b) {
    return 0.5 * (pq1[a] + pq1[b] + pq2[a] + pq2[b]);
}

// blocks of the Rayleigh table
enum { RAY_UBAR = 0, RAY_VBAR, RAY_WBAR, RAY_THETABAR, RAY_FAC_CC, RAY_FAC_ND, RAY_ACT_CC, RAY_ACT_ND, RAY_NFIELDS };
enum { RPT_FAC_CC = 0, RPT_FAC_ND, RPT_ACT_CC, RPT_ACT_ND, RPT_NFIELDS };

// Rayleigh factor and activity at cell / z-node height of point q on level k:
// from the per-level table on a flat grid, per point on a terrain grid
struct RayAt {
    const double* ray;
    const double* pt;
    int n1;
    long ne;
    __device__ __forceinline__ bool act_cc(long q, int k) const {
        return pt ? pt[RPT_ACT_CC * ne + q] != 0.0 : ray[RAY_ACT_CC * n1 + k] != 0.0;
    }
    __device__ __forceinline__ double fac_cc(long q, int k) const {
        return pt ? pt[RPT_FAC_CC * ne + q] : ray[RAY_FAC_CC * n1 + k];
    }
    __device__ __forceinline__ bool act_nd(long q, int k) const {
        return pt ? pt[RPT_ACT_ND * ne + q] != 0.0 : ray[RAY_ACT_ND * n1 + k] != 0.0;
    }
    __device__ __forceinline__ double fac_nd(long q, int k) const {
        return pt ? pt[RPT_FAC_ND * ne + q] : ray[RAY_FAC_ND * n1 + k];
    }
};

void make_params(const erfx_config& cfg, const erfx_layout& lay, Params& P);

// error state
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

}  // namespace erfx

#define ERFX_HIP(call)                                                         \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) return erfx::hip_fail(e_, #call);                \
    } while (0)
