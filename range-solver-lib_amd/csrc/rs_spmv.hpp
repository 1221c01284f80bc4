// rs_spmv.hpp -- SpMV building blocks for gfx950, shared by every kernel that multiplies by
// the matrix (plain SpMV, fused CG kernels, residuals).  Three row mappings (the autotuner
// times all of them per matrix and keeps the fastest):
//
//  VecRows<G>    "vector CSR": G lanes per row (G | 64, so 64/G rows per wavefront); each
//                workgroup owns a contiguous, XCD-aware chunk of rows; lane-group sums by
//                shuffles.  Good when rows are long.
//  TileRows<CAP> "stream CSR": the rows are cut into tiles of <= 256 rows and <= CAP nonzeros
//                (computed once per matrix).  A workgroup streams a tile's column/value
//                entries with independent, fully coalesced loads, stages the products in LDS,
//                then one thread per row sums its segment from LDS.  No dependent
//                rowptr -> col -> x chain per row, so many more loads are in flight.
//
//  SellRows      SELL-64 copy of the matrix, one row per lane (below).
//  SellRowsOff   the same slices with int16 column offsets from the slice's first row (when
//                every column is within +-32767 of it, as on mesh matrices): 2 bytes per
//                entry instead of 4, and padding points at the lane's own row (value 0), so the
//                product loop has no branch.
//
// All call epi.row(r, s) exactly once per row with the row's dot product s (accumulated in
// AT: double for the FP64 path, float for the FP32 path; VecRows calls a group epilogue's
// row_group on each lane of the row's group instead); the epilogue fuses whatever the caller
// needs (store, residual, dot products).  Sums are in a fixed order (deterministic).
#pragma once

#include <type_traits>

#include "rs_device.hpp"

namespace rs {

constexpr uint32_t kTileRows = 256;

// An epilogue with kGroupRow = true is called by every lane of a VecRows group, as
// epi.row_group<G>(r, s, lane in group) with the group's sum on every lane, instead of
// epi.row(r, s) on lane 0 only; it stores each row once all the same.
template <class E, class = void>
struct GroupRow {
    static constexpr bool value = false;
};
template <class E>
struct GroupRow<E, std::void_t<decltype(E::kGroupRow)>> {
    static constexpr bool value = E::kGroupRow;
};

struct CsrView {
    uint32_t n;
    uint32_t chunk;            // VecRows: rows per workgroup
    const uint32_t *rowptr, *col;
    const double *val;
    const float *val32;
    const uint32_t *tiles;     // TileRows: tile start rows (ntiles + 1)
    uint32_t ntiles, tpb;      // tiles, tiles per workgroup
    const uint32_t *sell_ptr;  // SellRows: slice start offsets (nslices + 1)
    const uint32_t *sell_col;  // column-major within a slice, padding = kSellPad
    const int16_t *sell_off;   // the same as offsets from the slice's first row (or null)
    const double *sell_val;
    const float *sell_val32;
    uint32_t nslices, spb;     // slices, slices per workgroup
};

constexpr uint32_t kSellPad = 0xFFFFFFFFu;

// The compact SELL arrays (FP32 values and int16 column offsets) store the entries of a slice in
// pairs: entries 2p and 2p+1 of a lane side by side at base + 128 p + 2 lane, an odd last entry
// in a plain 64-wide row after the pairs.  A lane then loads two entries of its row with one
// 8-byte (FP32 values) or 4-byte (offsets) load instead of two 4- or 2-byte loads.  FP64 values
// and 32-bit columns keep the plain slots base + 64 j + lane, where a lane's loads are already 8
// or 4 bytes wide.  (The learned smoother's packed FP16 copy has its own quad layout.)
__host__ __device__ __forceinline__ uint32_t sell_pair_slot(uint32_t base, uint32_t width, uint32_t j, uint32_t lane)
{
    return j < (width & ~1u) ? base + (j >> 1) * 2 * kWave + 2 * lane + (j & 1u) : base + j * kWave + lane;
}

typedef float sell_f2 __attribute__((ext_vector_type(2)));
typedef uint32_t sell_u2 __attribute__((ext_vector_type(2)));

// Values of entries 2p and 2p+1 of a lane's row in a slice starting at b.
__device__ __forceinline__ void sell_two(const double *__restrict__ v, uint32_t b, uint32_t p, uint32_t lane, double &a0,
                                         double &a1)
{
    a0 = __builtin_nontemporal_load(&v[b + 2 * p * kWave + lane]);
    a1 = __builtin_nontemporal_load(&v[b + (2 * p + 1) * kWave + lane]);
}
__device__ __forceinline__ void sell_two(const float *__restrict__ v, uint32_t b, uint32_t p, uint32_t lane, float &a0,
                                         float &a1)
{
    const sell_f2 t = __builtin_nontemporal_load(reinterpret_cast<const sell_f2 *>(&v[b + p * 2 * kWave + 2 * lane]));
    a0 = t.x;
    a1 = t.y;
}

template <int G>
struct VecRows {
    static constexpr bool kTile = false;
    static constexpr int kParam = G;
    template <typename AT, typename VT, typename XT, class Epi>
    __device__ __forceinline__ static void run(const CsrView &A, const VT *__restrict__ val,
                                               const XT *__restrict__ x, Epi &epi)
    {
        const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
        const int gl = threadIdx.x % G;
        const uint32_t r0 = lb * A.chunk;
        const uint32_t r1 = min(A.n, r0 + A.chunk);
        const uint32_t *__restrict__ rowptr = A.rowptr;
        const uint32_t *__restrict__ col = A.col;
        for (uint32_t base = r0; base < r1; base += kBlock / G) {
            const uint32_t r = base + threadIdx.x / G;
            if (r < r1) {
                const uint32_t b = rowptr[r], e = rowptr[r + 1];
                AT s = 0;
                for (uint32_t j = b + gl; j < e; j += G) s += (AT)val[j] * (AT)x[col[j]];
                s = group_sum<G>(s);
                if constexpr (GroupRow<Epi>::value) epi.template row_group<G>(r, s, gl);
                else if (gl == 0) epi.row(r, s);
            }
        }
    }
};

template <int CAP>
struct TileRows {
    static constexpr bool kTile = true;
    static constexpr int kParam = CAP;
    template <typename AT, typename VT, typename XT, class Epi>
    __device__ __forceinline__ static void run(const CsrView &A, const VT *__restrict__ val,
                                               const XT *__restrict__ x, Epi &epi)
    {
        __shared__ AT prod[CAP];
        __shared__ uint32_t rp[kTileRows + 1];
        __shared__ AT red[kBlock / kWave];
        const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
        const uint32_t t0 = lb * A.tpb;
        const uint32_t t1 = min(A.ntiles, t0 + A.tpb);
        const uint32_t *__restrict__ col = A.col;
        for (uint32_t t = t0; t < t1; t++) {
            const uint32_t r0 = A.tiles[t], nr = A.tiles[t + 1] - r0;
            for (uint32_t i = threadIdx.x; i <= nr; i += kBlock) rp[i] = A.rowptr[r0 + i];
            __syncthreads();
            const uint32_t j0 = rp[0], j1 = rp[nr];
            if (j1 - j0 <= (uint32_t)CAP) {
                uint32_t j = j0 + threadIdx.x;
                // 4 independent entries per thread in flight
                for (; j + 3 * kBlock < j1; j += 4 * kBlock) {
                    uint32_t c0 = col[j], c1 = col[j + kBlock], c2 = col[j + 2 * kBlock], c3 = col[j + 3 * kBlock];
                    VT v0 = val[j], v1 = val[j + kBlock], v2 = val[j + 2 * kBlock], v3 = val[j + 3 * kBlock];
                    XT x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
                    prod[j - j0] = (AT)v0 * (AT)x0;
                    prod[j + kBlock - j0] = (AT)v1 * (AT)x1;
                    prod[j + 2 * kBlock - j0] = (AT)v2 * (AT)x2;
                    prod[j + 3 * kBlock - j0] = (AT)v3 * (AT)x3;
                }
                for (; j < j1; j += kBlock) prod[j - j0] = (AT)val[j] * (AT)x[col[j]];
                __syncthreads();
                if (threadIdx.x < nr) {
                    const uint32_t a = rp[threadIdx.x] - j0, e = rp[threadIdx.x + 1] - j0;
                    AT s = 0;
                    for (uint32_t k = a; k < e; k++) s += prod[k];
                    epi.row(r0 + threadIdx.x, s);
                }
            } else {
                // a single row longer than CAP: the whole workgroup reduces it
                AT s = 0;
                for (uint32_t j = j0 + threadIdx.x; j < j1; j += kBlock) s += (AT)val[j] * (AT)x[col[j]];
                s = group_sum<kWave>(s);
                if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = s;
                __syncthreads();
                if (threadIdx.x == 0) {
                    AT tot = 0;
                    for (int w = 0; w < kBlock / kWave; w++) tot += red[w];
                    epi.row(r0, tot);
                }
            }
            __syncthreads();
        }
    }
};

// SELL-64 (sliced ELLPACK, Kreutzer et al. 2014, sigma = 1): rows in slices of 64, one row per
// lane of a wavefront; the entries of a slice are stored column-major, so entry j of the 64
// rows is one contiguous 64-wide load for columns and values.  On mesh matrices the j-th
// neighbours of consecutive rows are consecutive too, so the x gather coalesces as well.
// Padding entries carry column kSellPad and are skipped.
__device__ __forceinline__ const double *sell_values(const CsrView &A, const double *) { return A.sell_val; }
__device__ __forceinline__ const float *sell_values(const CsrView &A, const float *) { return A.sell_val32; }

// U entries of a row per loop trip (U / 2 pairs), all loads of a trip independent (U = 4 or 8:
// the autotuner keeps the faster; deeper unrolling keeps more loads in flight per wave).  The
// entries are summed in ascending order, as the CSR row is.
template <int U>
struct SellRowsU {
    static constexpr bool kTile = false;
    static constexpr int kParam = 64;
    static_assert(U % 2 == 0, "entries come in pairs");
    template <typename AT, typename VT, typename XT, class Epi>
    __device__ __forceinline__ static void run(const CsrView &A, const VT *__restrict__ csr_val,
                                               const XT *__restrict__ x, Epi &epi)
    {
        const VT *__restrict__ val = sell_values(A, csr_val);
        const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
        const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
        const uint32_t s0 = lb * A.spb, s1 = min(A.nslices, s0 + A.spb);
        const uint32_t *__restrict__ col = A.sell_col;
        // the matrix streams through once per product: non-temporal loads keep the L2 for
        // the vectors (measured: FP64 SpMV 93.6 -> 78.7 us at 128^3, and the FP32 products
        // inside the solvers faster although a lone FP32 SpMV is ~5% slower)
        auto ld = [](const auto *p) { return __builtin_nontemporal_load(p); };
        for (uint32_t s = s0 + wid; s < s1; s += kBlock / kWave) {
            const uint32_t b = A.sell_ptr[s], e = A.sell_ptr[s + 1], width = (e - b) / kWave, np = width >> 1;
            AT acc = 0;
            if constexpr (sizeof(VT) == 8) {
                // FP64 values in the plain slots: U entries per trip by one running index (this
                // form measured 71.7 us at 128^3 against 80.0 us for the pair-indexed loop below)
                uint32_t j = b + lane;
                for (; j + (U - 1) * kWave < e; j += U * kWave) {
                    uint32_t c[U];
                    VT v[U];
                    XT y[U];
#pragma unroll
                    for (int u = 0; u < U; u++) c[u] = ld(&col[j + u * kWave]);
#pragma unroll
                    for (int u = 0; u < U; u++) v[u] = ld(&val[j + u * kWave]);
#pragma unroll
                    for (int u = 0; u < U; u++) y[u] = c[u] != kSellPad ? x[c[u]] : XT(0);
#pragma unroll
                    for (int u = 0; u < U; u++) acc += (AT)v[u] * (AT)y[u];
                }
                for (; j < e; j += kWave) {
                    const uint32_t cc = col[j];
                    if (cc != kSellPad) acc += (AT)val[j] * (AT)x[cc];
                }
                const uint32_t r = s * kWave + lane;
                if (r < A.n) epi.row(r, acc);
                continue;
            }
            uint32_t p = 0;
            for (; p + U / 2 <= np; p += U / 2) {
                uint32_t c[U];
                VT v[U];
                XT y[U];
#pragma unroll
                for (int u = 0; u < U; u++) c[u] = ld(&col[b + (2 * p + u) * kWave + lane]);
#pragma unroll
                for (int q = 0; q < U / 2; q++) sell_two(val, b, p + q, lane, v[2 * q], v[2 * q + 1]);
#pragma unroll
                for (int u = 0; u < U; u++) y[u] = c[u] != kSellPad ? x[c[u]] : XT(0);
#pragma unroll
                for (int u = 0; u < U; u++) acc += (AT)v[u] * (AT)y[u];
            }
            for (; p < np; p++) {
                const uint32_t c0 = col[b + 2 * p * kWave + lane], c1 = col[b + (2 * p + 1) * kWave + lane];
                VT v0, v1;
                sell_two(val, b, p, lane, v0, v1);
                if (c0 != kSellPad) acc += (AT)v0 * (AT)x[c0];
                if (c1 != kSellPad) acc += (AT)v1 * (AT)x[c1];
            }
            if (width & 1u) {
                const uint32_t k = b + (width - 1) * kWave + lane, cc = col[k];
                if (cc != kSellPad) acc += (AT)val[k] * (AT)x[cc];
            }
            const uint32_t r = s * kWave + lane;
            if (r < A.n) epi.row(r, acc);
        }
    }
};
using SellRows = SellRowsU<4>;

template <int U>
struct SellRowsOffU {
    static constexpr bool kTile = false;
    static constexpr int kParam = 64;
    static_assert(U % 2 == 0, "entries come in pairs");
    template <typename AT, typename VT, typename XT, class Epi>
    __device__ __forceinline__ static void run(const CsrView &A, const VT *__restrict__ csr_val,
                                               const XT *__restrict__ x, Epi &epi)
    {
        if (A.sell_off == nullptr) {
            // columns do not fit 16-bit offsets: the full-column slices
            SellRowsU<U>::template run<AT, VT, XT, Epi>(A, csr_val, x, epi);
            return;
        }
        const VT *__restrict__ val = sell_values(A, csr_val);
        const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
        const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
        const uint32_t s0 = lb * A.spb, s1 = min(A.nslices, s0 + A.spb);
        const int16_t *__restrict__ off = A.sell_off;
        // the offsets of entries 2p and 2p+1 as one 32-bit load (low half: entry 2p)
        auto ld2 = [&](uint32_t b, uint32_t p) {
            return __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(&off[b + p * 2 * kWave + 2 * lane]));
        };
        for (uint32_t s = s0 + wid; s < s1; s += kBlock / kWave) {
            const uint32_t b = A.sell_ptr[s], width = (A.sell_ptr[s + 1] - b) / kWave, np = width >> 1;
            const int32_t sb = (int32_t)(s * kWave);
            AT acc = 0;
            uint32_t p = 0;
            for (; p + U / 2 <= np; p += U / 2) {
                uint32_t o[U / 2];
                VT v[U];
                XT y[U];
#pragma unroll
                for (int q = 0; q < U / 2; q++) o[q] = ld2(b, p + q);
#pragma unroll
                for (int q = 0; q < U / 2; q++) sell_two(val, b, p + q, lane, v[2 * q], v[2 * q + 1]);
#pragma unroll
                for (int q = 0; q < U / 2; q++) {
                    y[2 * q] = x[sb + (int32_t)(int16_t)(o[q] & 0xffffu)];
                    y[2 * q + 1] = x[sb + (int32_t)(int16_t)(o[q] >> 16)];
                }
#pragma unroll
                for (int u = 0; u < U; u++) acc += (AT)v[u] * (AT)y[u];
            }
            for (; p < np; p++) {
                const uint32_t o = ld2(b, p);
                VT v0, v1;
                sell_two(val, b, p, lane, v0, v1);
                acc += (AT)v0 * (AT)x[sb + (int32_t)(int16_t)(o & 0xffffu)];
                acc += (AT)v1 * (AT)x[sb + (int32_t)(int16_t)(o >> 16)];
            }
            if (width & 1u) {
                const uint32_t k = b + (width - 1) * kWave + lane;
                acc += (AT)val[k] * (AT)x[sb + off[k]];
            }
            const uint32_t r = s * kWave + lane;
            if (r < A.n) epi.row(r, acc);
        }
    }
};
using SellRowsOff = SellRowsOffU<4>;

// Launch plan of one matrix: which shape, and its grid.
struct SpmvPlan {
    int id = -1;           // index into kShapes
    uint32_t nb = 0;       // workgroups
};

// Shape ids: 0..4 VecRows<4,8,16,32,64>, 5 TileRows<2048>, 6 TileRows<4096>, 7 SellRows,
// 8 SellRowsOff, 9 SellRowsU<8>, 10 SellRowsOffU<8>.  (16 entries per trip were measured too and
// lost on every matrix tried: more VGPRs, lower occupancy.)
constexpr int kNumShapes = 11;

#define RS_SHAPE_SWITCH(id, KERNEL, grid, stream, ...)                                         \
    do {                                                                                       \
        switch (id) {                                                                          \
            case 0: KERNEL<::rs::VecRows<4>><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break;   \
            case 1: KERNEL<::rs::VecRows<8>><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break;   \
            case 2: KERNEL<::rs::VecRows<16>><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break;  \
            case 3: KERNEL<::rs::VecRows<32>><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break;  \
            case 4: KERNEL<::rs::VecRows<64>><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break;  \
            case 5: KERNEL<::rs::TileRows<2048>><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break; \
            case 6: KERNEL<::rs::TileRows<4096>><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break; \
            case 7: KERNEL<::rs::SellRows><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break;       \
            case 8: KERNEL<::rs::SellRowsOff><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break;    \
            case 9: KERNEL<::rs::SellRowsU<8>><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break;   \
            default: KERNEL<::rs::SellRowsOffU<8>><<<grid, ::rs::kBlock, 0, stream>>>(__VA_ARGS__); break; \
        }                                                                                      \
    } while (0)

}  // namespace rs
