// rs_batch.hip -- many independent sparse systems solved together on gfx950: every system runs
// its own implicit-Euler time steps of Jacobi-CG, and all systems share every launch.
//
// Reference behaviour per system (Range-Software/range-solver-lib):
//   RSolverHeat::prepare clears x before every step (src/rsolverheat.cpp:83-88); assemblyMatrix
//   forms Ae = Me + alpha dt Ke and be = dt fe + (Me - (1 - alpha) dt Ke) T_prev
//   (src/rsolverheat.cpp:757-800); RSolverHeat::solve runs Jacobi-CG (:458-465) and writes x
//   back into nodeTemperature (:473-480), which is T_prev of the next step.
//   With alpha = 1 and coefficients fixed in time, A is the same in every step and
//   b_{k+1} = c + M_ff x_k, where c is b assembled with the free temperatures at zero and M_ff
//   the free-free block of the mass matrix.  The batch forms b on the device from c and M
//   instead of assembling again (the same sums up to rounding).
//   CG: RMatrixSolver::solveCG (src/rmatrixsolver.cpp:208-355) with the Jacobi preconditioner
//   of RMatrixPreconditioner (src/rmatrixpreconditioner.cpp:78-94); a step ends when the FP64
//   true relative residual ||b - A x|| / ||b|| is below the tolerance.
//
// Layout in HBM: the systems are concatenated into one block-diagonal matrix whose system row
// ranges start at multiples of 64 and are stored SELL-64: slice s (rows 64 s .. 64 s + 63)
// belongs to exactly one system, so every per-system reduction is a fixed-order sum over
// whole slices.  Padding entries point at their own row with value 0, so the product loop has
// no branch.  Per entry: column (uint32), A (FP64 and FP32 copy), M (FP64).  Per row: c, b,
// two solution banks (FP64; step k works in bank k & 1 and reads the previous step's solution
// from the other; step 0 reads the initial solution, kept apart), Jacobi 1/A_ii (FP64 and FP32), r/p/q in the working precision.
//
// One round of the batched loop = 7 launches over all systems' slices:
//   dir      p = z + beta p (z = D^-1 r on the fly)                   iterating systems
//   spmv     q = A p, slice partial p.q                                autotuned SELL shape
//   fin      alpha per system (one workgroup per system, its slices summed in order)
//   update   x += alpha p, r -= alpha q, slice partials |r|^2, r.z
//   fin      recursive error; replacement request or beta
//   event    FP64 b - A x for a replacement; b = c + M x_prev and a fresh start for a system
//            starting a step
//   fin      true relative residual: step done -> next step or finished; else continue
// A system waiting for its event, or finished, skips its slices.  Sixteen rounds are captured
// once in a hipGraph and replayed until every system has finished (device-side counter).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "rs_device.hpp"
#include "rs_internal.hpp"

namespace rs {

namespace {

constexpr uint32_t kSlice = kWave;           // rows per slice
constexpr uint32_t kFlagActive = 1u;         // not finished
constexpr uint32_t kFlagFirst = 2u;          // next direction is p = z
constexpr uint32_t kFlagReplace = 4u;        // FP64 residual replacement pending
constexpr uint32_t kFlagAdvance = 8u;        // step start pending
constexpr uint32_t kEvent = kFlagReplace | kFlagAdvance;
constexpr double kBatchDelta = 0.01;         // mixed: replace when |r| fell by this since the last replacement
constexpr uint32_t kBatchNoGain = 20;        // replacements without halving the true residual
constexpr uint32_t kBatchRounds = 16;        // rounds per captured graph
constexpr int kBatchShapes = 5;             // see RS_BT_SHAPE

struct BtSys {
    double alpha, beta, rho0, rho1;
    double bn, rmax, err, best;
    double clamp, pad0;                 // |p.q| floor: eps of the reference's scaled system
    uint32_t it, step, flags, reason;   // it: iterations in the current step
    uint32_t nrepl, no_gain, it_total, bank;
};

struct BtGlobal {
    uint32_t it, done, reason, pad;     // it: rounds (host reads these 16 bytes per replay)
    uint32_t finished, pad2[3];
};

struct BtArgs {
    uint32_t nslices, count, steps, n_outer;
    const uint32_t *slice_sys, *sys_s0, *sys_ns;
    const double *an;            // ||A||_F per system
    const uint32_t *sell_ptr, *sell_col;
    const int16_t *sell_off;     // columns as offsets from the slice's first row (when they fit)
    const double *sell_val;
    const float *sell_val32;
    const double *sell_mval;     // null: no mass matrix (b = c every step)
    const double *c;
    double *b, *x0, *x1;
    const double *xinit;         // initial solution (the "previous step" of step 0)
    const double *dinv;
    const float *dinv32;
    void *r, *p, *q;
    double *partials;
    BtSys *sys;
    BtGlobal *g;
    uint32_t *its;
    double *relres;
    double cvg, delta;
    int32_t warm;
};

template <typename T>
__device__ __forceinline__ const T *bt_vals(const BtArgs &a);
template <>
__device__ __forceinline__ const double *bt_vals<double>(const BtArgs &a) { return a.sell_val; }
template <>
__device__ __forceinline__ const float *bt_vals<float>(const BtArgs &a) { return a.sell_val32; }
template <typename T>
__device__ __forceinline__ const T *bt_dinv(const BtArgs &a);
template <>
__device__ __forceinline__ const double *bt_dinv<double>(const BtArgs &a) { return a.dinv; }
template <>
__device__ __forceinline__ const float *bt_dinv<float>(const BtArgs &a) { return a.dinv32; }

__device__ __forceinline__ bool bt_iterating(uint32_t fl) { return (fl & kFlagActive) && !(fl & kEvent); }

__device__ __forceinline__ double *bt_bank(const BtArgs &a, uint32_t step) { return (step & 1u) ? a.x1 : a.x0; }

template <typename P>
__device__ __forceinline__ P bt_nt(const P *p) { return __builtin_nontemporal_load(p); }

// Column of SELL entry k in slice s: the full uint32 column, or (when every column of the batch
// lies within +-32767 of its slice's first row, as on mesh matrices) an int16 offset from the
// slice's first row -- 2 bytes per entry instead of 4.
struct ColFull {
    const uint32_t *c;
    __device__ __forceinline__ uint32_t operator()(uint32_t k, uint32_t) const { return bt_nt(c + k); }
};
struct ColOff {
    const int16_t *c;
    __device__ __forceinline__ uint32_t operator()(uint32_t k, uint32_t sbase) const
    {
        return (uint32_t)((int32_t)sbase + (int32_t)bt_nt(c + k));
    }
};
template <bool OFF>
__device__ __forceinline__ auto bt_cols(const BtArgs &a)
{
    if constexpr (OFF) return ColOff{ a.sell_off };
    else return ColFull{ a.sell_col };
}

// Row `row` of a slice whose entries start at `base` (width entries per row), G lanes per
// row: lane sub takes entries sub, sub + G, ...; the G partial sums are combined by shuffles
// in a fixed order (result in every lane of the row).
template <int G, typename AT, class CR, typename VT, typename XT>
__device__ __forceinline__ AT bt_row(CR col, const VT *__restrict__ val, const XT *__restrict__ x, uint32_t base,
                                     uint32_t width, uint32_t row, uint32_t sub, uint32_t sbase)
{
    const uint32_t k0 = base + row;
    const VT *__restrict__ vp = val + k0;
    AT s = 0;
    uint32_t j = sub;
    for (; j + 3 * G < width; j += 4 * G) {
        uint32_t c0 = col(k0 + j * kSlice, sbase), c1 = col(k0 + (j + G) * kSlice, sbase);
        uint32_t c2 = col(k0 + (j + 2 * G) * kSlice, sbase), c3 = col(k0 + (j + 3 * G) * kSlice, sbase);
        VT v0 = bt_nt(vp + j * kSlice), v1 = bt_nt(vp + (j + G) * kSlice);
        VT v2 = bt_nt(vp + (j + 2 * G) * kSlice), v3 = bt_nt(vp + (j + 3 * G) * kSlice);
        XT y0 = x[c0], y1 = x[c1], y2 = x[c2], y3 = x[c3];
        s += (AT)v0 * (AT)y0;
        s += (AT)v1 * (AT)y1;
        s += (AT)v2 * (AT)y2;
        s += (AT)v3 * (AT)y3;
    }
    for (; j < width; j += G) s += (AT)bt_nt(vp + j * kSlice) * (AT)x[col(k0 + j * kSlice, sbase)];
#pragma unroll
    for (int off = kWave / 2; off >= kWave / G; off >>= 1) s += __shfl_xor(s, off, kWave);
    return s;
}

// Both products of one row with the same vector: (A y)_row and (M y)_row, FP64.
template <int G, class CR>
__device__ __forceinline__ void bt_row2(const BtArgs &a, CR col, const double *__restrict__ y, uint32_t base,
                                        uint32_t width, uint32_t row, uint32_t sub, uint32_t sbase, double &ay,
                                        double &my)
{
    const uint32_t k0 = base + row;
    const double *__restrict__ va = a.sell_val + k0;
    const double *__restrict__ vm = a.sell_mval + k0;
    double s = 0.0, m = 0.0;
    for (uint32_t j = sub; j < width; j += G) {
        const double yc = y[col(k0 + j * kSlice, sbase)];
        s += bt_nt(va + j * kSlice) * yc;
        m += bt_nt(vm + j * kSlice) * yc;
    }
#pragma unroll
    for (int off = kWave / 2; off >= kWave / G; off >>= 1) {
        s += __shfl_xor(s, off, kWave);
        m += __shfl_xor(m, off, kWave);
    }
    ay = s;
    my = m;
}

// ---------------------------------------------------------------------------- set-up kernels
__global__ void k_bt_copy_pattern(const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col, uint32_t n,
                                  uint32_t n_pad, uint64_t nnz, uint32_t roff, uint32_t joff,
                                  uint32_t *__restrict__ rpc, uint32_t *__restrict__ colc)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += stride)
        rpc[roff + i] = joff + (i < n ? rp[i] : (uint32_t)nnz);
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += stride)
        colc[joff + j] = roff + col[j];
}

// Entries of `rp`/`col` that differ from the batch's copy of system i's pattern.
__global__ void k_bt_check_pattern(const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col, uint32_t n,
                                   uint64_t nnz, uint32_t roff, uint32_t joff, const uint32_t *__restrict__ rpc,
                                   const uint32_t *__restrict__ colc, uint32_t *__restrict__ bad)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t miss = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        miss += rpc[roff + i] != joff + rp[i];
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += stride)
        miss += colc[joff + j] != roff + col[j];
    if (miss) atomicAdd(bad, miss);
}

// Width (longest row) of every slice.
__global__ void k_bt_widths(const uint32_t *__restrict__ rpc, uint32_t nslices, uint32_t *__restrict__ width)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    for (uint32_t s = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; s < nslices;
         s += gridDim.x * blockDim.x / kWave) {
        uint64_t r = (uint64_t)s * kSlice + lane;
        uint32_t len = rpc[r + 1] - rpc[r];
        for (int off = kWave / 2; off > 0; off >>= 1) len = max(len, (uint32_t)__shfl_xor((int)len, off, kWave));
        if (lane == 0) width[s] = len;
    }
}

// SELL-64 copy of the combined CSR, and the Jacobi inverse diagonal (constructJacobi: 1/A_ii,
// 0 when the diagonal is absent or zero).
__global__ void k_bt_sell_fill(const uint32_t *__restrict__ rpc, const uint32_t *__restrict__ colc,
                               const double *__restrict__ val, const double *__restrict__ mval,
                               const uint32_t *__restrict__ sell_ptr, uint32_t nslices, uint32_t *__restrict__ scol,
                               double *__restrict__ sval, float *__restrict__ sval32, double *__restrict__ smval,
                               double *__restrict__ dinv, float *__restrict__ dinv32, int16_t *__restrict__ soff,
                               uint32_t *__restrict__ far)
{
    const uint64_t rows = (uint64_t)nslices * kSlice;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = (uint32_t)(r / kSlice), lane = (uint32_t)(r % kSlice);
        const uint32_t base = sell_ptr[s], width = (sell_ptr[s + 1] - base) / kSlice;
        const uint32_t b = rpc[r], len = rpc[r + 1] - b;
        double d = 0.0;
        for (uint32_t j = 0; j < width; j++) {
            const uint32_t k = base + j * kSlice + lane;
            if (j < len) {
                const uint32_t c = colc[b + j];
                const double v = val[b + j];
                scol[k] = c;
                sval[k] = v;
                sval32[k] = (float)v;
                if (smval) smval[k] = mval[b + j];
                if (c == r && d == 0.0) d = v;
            } else {
                scol[k] = (uint32_t)r;
                sval[k] = 0.0;
                sval32[k] = 0.0f;
                if (smval) smval[k] = 0.0;
            }
        }
        for (uint32_t j = 0; j < width; j++) {
            const uint32_t k = base + j * kSlice + lane;
            const int64_t o = (int64_t)scol[k] - (int64_t)s * kSlice;
            if (o < -32767 || o > 32767) atomicAdd(far, 1u);
            else soff[k] = (int16_t)o;
        }
        const double inv = d != 0.0 ? 1.0 / d : 0.0;
        dinv[r] = inv;
        dinv32[r] = (float)inv;
    }
}

// ||A||_F of every system (one workgroup per system, fixed order).
__global__ __launch_bounds__(kBlock) void k_bt_frob(const double *__restrict__ val, const uint32_t *__restrict__ rpc,
                                                    const uint32_t *__restrict__ sys_s0,
                                                    const uint32_t *__restrict__ sys_ns, double *__restrict__ an)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t i = blockIdx.x;
    const uint64_t r0 = (uint64_t)sys_s0[i] * kSlice, r1 = r0 + (uint64_t)sys_ns[i] * kSlice;
    const uint32_t j0 = rpc[r0], j1 = rpc[r1];
    double v[1] = { 0.0 };
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += kBlock) v[0] += val[j] * val[j];
    block_sum<1, kBlock>(v, lds);
    if (threadIdx.x == 0) an[i] = sqrt(v[0]);
}

__global__ void k_bt_fill(double *__restrict__ a, uint64_t n, double v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

__global__ void k_bt_fill32(float *__restrict__ a, uint64_t n, float v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

__global__ void k_bt_fill_u32(uint32_t *__restrict__ a, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// Initial state: every system waits for the start of step 0.  tune = 1: every system iterates
// (used only to time the SpMV shapes).
__global__ void k_bt_init(BtArgs a, int tune)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.count; i += gridDim.x * blockDim.x) {
        BtSys s{};
        s.flags = tune ? kFlagActive : (kFlagActive | kFlagAdvance);
        s.best = INFINITY;
        s.bank = 0;
        a.sys[i] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.g = BtGlobal{};
}

// ---------------------------------------------------------------------------- iteration kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bt_dir(BtArgs a)
{
    if (a.g->done) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t nw = gridDim.x * (kBlock / kWave);
    const T *__restrict__ r = (const T *)a.r;
    T *__restrict__ p = (T *)a.p;
    const T *__restrict__ dinv = bt_dinv<T>(a);
    for (uint32_t s = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; s < a.nslices; s += nw) {
        const BtSys &S = a.sys[a.slice_sys[s]];
        const uint32_t fl = S.flags;
        if (!bt_iterating(fl)) continue;
        const uint64_t i = (uint64_t)s * kSlice + lane;
        const T z = dinv[i] * r[i];
        p[i] = (fl & kFlagFirst) ? z : z + (T)S.beta * p[i];
    }
}

template <int G, int W, typename T, bool OFF>
__global__ __launch_bounds__(W * kWave) void k_bt_spmv(BtArgs a)
{
    if (a.g->done) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t nw = gridDim.x * W;
    const uint32_t lw = xcd_remap(blockIdx.x, gridDim.x) * W + threadIdx.x / kWave;
    const T *__restrict__ p = (const T *)a.p;
    T *__restrict__ q = (T *)a.q;
    const T *__restrict__ val = bt_vals<T>(a);
    const auto cols = bt_cols<OFF>(a);
    constexpr uint32_t R = kWave / G;
    const uint32_t sub = lane / R;
    // the next slice's system and row range are loaded while this slice's rows run
    uint32_t nsys = lw < a.nslices ? a.slice_sys[lw] : 0;
    uint32_t nb0 = lw < a.nslices ? a.sell_ptr[lw] : 0, nb1 = lw < a.nslices ? a.sell_ptr[lw + 1] : 0;
    for (uint32_t s = lw; s < a.nslices; s += nw) {
        const uint32_t sys = nsys, base = nb0, width = (nb1 - nb0) / kSlice;
        const uint32_t sn = s + nw;
        if (sn < a.nslices) {
            nsys = a.slice_sys[sn];
            nb0 = a.sell_ptr[sn];
            nb1 = a.sell_ptr[sn + 1];
        }
        if (!bt_iterating(a.sys[sys].flags)) continue;
        const uint32_t sbase = s * kSlice;
        double acc = 0.0;
#pragma unroll
        for (int pass = 0; pass < G; pass++) {
            const uint32_t row = pass * R + lane % R;
            const T sum = bt_row<G, T>(cols, val, p, base, width, row, sub, sbase);
            if (sub == 0) {
                const uint64_t i = (uint64_t)s * kSlice + row;
                q[i] = sum;
                acc += (double)p[i] * (double)sum;
            }
        }
        acc = wave_sum(acc);
        if (lane == 0) a.partials[s] = acc;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bt_update(BtArgs a)
{
    if (a.g->done) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t nw = gridDim.x * (kBlock / kWave);
    T *__restrict__ r = (T *)a.r;
    const T *__restrict__ p = (const T *)a.p;
    const T *__restrict__ q = (const T *)a.q;
    const T *__restrict__ dinv = bt_dinv<T>(a);
    for (uint32_t s = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; s < a.nslices; s += nw) {
        const BtSys &S = a.sys[a.slice_sys[s]];
        if (!bt_iterating(S.flags)) continue;
        const double alpha = S.alpha;
        double *__restrict__ x = bt_bank(a, S.step);
        const uint64_t i = (uint64_t)s * kSlice + lane;
        const T pi = p[i];
        x[i] += alpha * (double)pi;
        const T ri = r[i] - (T)alpha * q[i];
        r[i] = ri;
        const T z = dinv[i] * ri;
        double v0 = wave_sum((double)ri * (double)ri);
        double v1 = wave_sum((double)ri * (double)z);
        if (lane == 0) {
            a.partials[s] = v0;
            a.partials[a.nslices + s] = v1;
        }
    }
}

// Events: a replacement (r = b - A x in FP64) or a step start (b = c + M x_prev, x = 0 as the
// reference's prepare does, or x = x_prev with warm starts; r = b - A x).
template <int G, int W, typename T, bool OFF>
__global__ __launch_bounds__(W * kWave) void k_bt_event(BtArgs a)
{
    if (a.g->done) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t nw = gridDim.x * W;
    const uint32_t lw = xcd_remap(blockIdx.x, gridDim.x) * W + threadIdx.x / kWave;
    T *__restrict__ r = (T *)a.r;
    const T *__restrict__ dinv = bt_dinv<T>(a);
    const auto cols = bt_cols<OFF>(a);
    constexpr uint32_t R = kWave / G;
    const uint32_t sub = lane / R;
    for (uint32_t s = lw; s < a.nslices; s += nw) {
        const uint32_t sbase = s * kSlice;
        const BtSys &S = a.sys[a.slice_sys[s]];
        const uint32_t fl = S.flags;
        if (!(fl & kFlagActive) || !(fl & kEvent)) continue;
        const bool adv = fl & kFlagAdvance;
        double *__restrict__ x = bt_bank(a, S.step);
        const double *__restrict__ xp = S.step == 0 ? a.xinit : bt_bank(a, S.step + 1);
        const uint32_t base = a.sell_ptr[s], width = (a.sell_ptr[s + 1] - base) / kSlice;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
        for (int pass = 0; pass < G; pass++) {
            const uint32_t row = pass * R + lane % R;
            const uint64_t i = (uint64_t)s * kSlice + row;
            double bi, ax = 0.0;
            if (adv) {
                double my = 0.0, ay = 0.0;
                if (a.sell_mval) bt_row2<G>(a, cols, xp, base, width, row, sub, sbase, ay, my);
                else if (a.warm) ay = bt_row<G, double>(cols, a.sell_val, xp, base, width, row, sub, sbase);
                bi = a.c[i] + my;
                if (a.warm) ax = ay;
                if (sub == 0) {
                    a.b[i] = bi;
                    x[i] = a.warm ? xp[i] : 0.0;
                }
            } else {
                ax = bt_row<G, double>(cols, a.sell_val, x, base, width, row, sub, sbase);
                bi = a.b[i];
            }
            if (sub == 0) {
                const double ri = bi - ax;
                const T rt = (T)ri;
                r[i] = rt;
                const T z = dinv[i] * rt;
                v0 += ri * ri;
                v1 += (double)rt * (double)z;
                v2 += bi * bi;
            }
        }
        v0 = wave_sum(v0);
        v1 = wave_sum(v1);
        v2 = wave_sum(v2);
        if (lane == 0) {
            a.partials[s] = v0;
            a.partials[a.nslices + s] = v1;
            a.partials[2ull * a.nslices + s] = v2;
        }
    }
}

enum BtFin { BT_ALPHA = 0, BT_UPDATE = 1, BT_EVENT = 2 };

__device__ __forceinline__ void bt_finish(const BtArgs &a, BtSys *S, uint32_t reason)
{
    S->flags = 0;
    S->reason = reason;
    S->bank = S->step & 1u;
    uint32_t old = atomicAdd(&a.g->finished, 1u);
    if (old + 1 == a.count) a.g->done = 1;
}

// Per-system finalize: one workgroup per system sums its slices' partials in slice order.
template <int OP>
__global__ __launch_bounds__(kBlock) void k_bt_fin(BtArgs a)
{
    if (a.g->done) return;
    if (OP == BT_EVENT && blockIdx.x == 0 && threadIdx.x == 0) a.g->it++;
    const uint32_t i = blockIdx.x;
    BtSys *S = &a.sys[i];
    const uint32_t fl = S->flags;
    if (!(fl & kFlagActive)) return;
    if ((OP == BT_EVENT) != ((fl & kEvent) != 0)) return;
    constexpr int NV = OP == BT_ALPHA ? 1 : (OP == BT_UPDATE ? 2 : 3);
    __shared__ double lds[NV * (kBlock / kWave)];
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = 0.0;
    const uint32_t s0 = a.sys_s0[i], ns = a.sys_ns[i];
    for (uint32_t t = threadIdx.x; t < ns; t += kBlock)
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] += a.partials[(uint64_t)k * a.nslices + s0 + t];
    block_sum<NV, kBlock>(v, lds);
    if (threadIdx.x != 0) return;
    const double eps = kEps;
    if (OP == BT_ALPHA) {
        // RMatrixSolver::solveCG: alpha = rho / p.q with p.q kept away from zero (:302-321)
        const double floor = S->clamp;
        double dot = v[0];
        if (dot > 0.0) dot = fmax(dot, floor);
        else if (dot < 0.0) dot = fmin(dot, -floor);
        else dot = floor;
        S->alpha = S->rho0 / dot;
        S->rho1 = S->rho0;
        S->flags = fl & ~kFlagFirst;
        return;
    }
    const double rn = sqrt(v[0]);
    if (OP == BT_UPDATE) {
        const double err = rn / fmax(S->bn, eps);
        S->err = err;
        S->it++;
        S->it_total++;
        if (!dev_isnormal(err) && err != 0.0) {
            bt_finish(a, S, RS_REASON_NONNORMAL);
        } else if (rn < a.delta * S->rmax || err < a.cvg) {
            S->flags = fl | kFlagReplace;
        } else if (S->it >= a.n_outer) {
            a.its[(uint64_t)i * a.steps + S->step] = S->it;
            a.relres[(uint64_t)i * a.steps + S->step] = err;
            bt_finish(a, S, RS_REASON_MAXIT);
        } else {
            S->beta = v[1] / S->rho1;
            S->rho0 = v[1];
        }
        return;
    }
    // BT_EVENT
    const bool adv = fl & kFlagAdvance;
    if (adv) {
        // RMatrixSolver::solve runs CG on the system scaled by s = 1e9 ||A||_F / ||b||
        // (src/rmatrixsolver.cpp:164-171); p.q scales by s^2, so the eps floor of the scaled
        // system is eps / s^2 here
        const double bn = sqrt(v[2]), an = a.an[i];
        const double sc = (bn != 0.0 && an != 0.0) ? 1.0e9 * an / bn : 1.0;
        S->bn = bn;
        S->clamp = fmax(eps / (sc * sc), 1.0e-300);
    }
    const double err = rn / fmax(S->bn, eps);
    S->err = err;
    S->rmax = rn;
    uint32_t nf = fl & ~kEvent;
    const uint64_t h = (uint64_t)i * a.steps + S->step;
    if (!adv) S->nrepl++;
    if (!dev_isnormal(err) && err != 0.0) {
        S->flags = nf;
        a.its[h] = S->it;
        a.relres[h] = err;
        bt_finish(a, S, RS_REASON_NONNORMAL);
        return;
    }
    if (err < a.cvg) {
        // the step is done: record it, then start the next step or finish
        a.its[h] = adv ? 0u : S->it;
        a.relres[h] = err;
        S->step++;
        if (S->step >= a.steps) {
            S->step--;
            S->flags = nf;
            bt_finish(a, S, RS_REASON_CONVERGED);
        } else {
            S->flags = nf | kFlagAdvance;
        }
        return;
    }
    if (adv) {
        S->it = 0;
        S->best = err;
        S->no_gain = 0;
        S->rho0 = v[1];
        S->flags = nf | kFlagFirst;
        return;
    }
    // a replacement that did not converge: keep the direction (reliable updates)
    if (err < 0.5 * S->best) {
        S->best = err;
        S->no_gain = 0;
    } else if (++S->no_gain >= kBatchNoGain) {
        S->flags = nf;
        a.its[h] = S->it;
        a.relres[h] = err;
        bt_finish(a, S, RS_REASON_STAGNATED);
        return;
    }
    if (S->it >= a.n_outer) {
        S->flags = nf;
        a.its[h] = S->it;
        a.relres[h] = err;
        bt_finish(a, S, RS_REASON_MAXIT);
        return;
    }
    S->beta = v[1] / S->rho1;
    S->rho0 = v[1];
    S->flags = nf;
}

#define RS_BT_SHAPE(id, KERNEL, T, OFF, nslices, stream, args)                                               \
    do {                                                                                                  \
        switch (id) {                                                                                     \
            case 0: KERNEL<1, 4, T, OFF><<<bt_grid(nslices, 4), 4 * kWave, 0, stream>>>(args); break;     \
            case 1: KERNEL<1, 8, T, OFF><<<bt_grid(nslices, 8), 8 * kWave, 0, stream>>>(args); break;     \
            case 2: KERNEL<1, 16, T, OFF><<<bt_grid(nslices, 16), 16 * kWave, 0, stream>>>(args); break;  \
            case 3: KERNEL<2, 4, T, OFF><<<bt_grid(nslices, 4), 4 * kWave, 0, stream>>>(args); break;     \
            default: KERNEL<4, 4, T, OFF><<<bt_grid(nslices, 4), 4 * kWave, 0, stream>>>(args); break;    \
        }                                                                                                 \
    } while (0)

uint32_t round8(uint32_t a) { return (a + kXcd - 1) / kXcd * kXcd; }

// Persistent grid: at most as many waves as the chip holds (32 per CU), one wave per slice.
uint32_t bt_grid(uint32_t nslices, uint32_t waves_per_block)
{
    uint32_t cap = (uint32_t)std::max(1, runtime().n_cu) * (32u / waves_per_block);
    uint32_t want = (nslices + waves_per_block - 1) / waves_per_block;
    return round8(std::max(1u, std::min(cap, want)));
}

uint32_t bt_vec_grid(uint32_t nslices) { return bt_grid(nslices, kBlock / kWave); }

}  // namespace

}  // namespace rs

using namespace rs;

struct rs_batch {
    uint32_t count = 0;
    std::vector<uint32_t> n, n_pad, roff, s0, ns;
    std::vector<uint64_t> nnz, joff;
    std::vector<uint8_t> have_pattern, have_matrix, have_mass, have_rhs;
    uint32_t N = 0, nslices = 0;
    uint64_t NNZ = 0, sell_entries = 0;
    // combined CSR
    DBuf<uint32_t> rowptr, col;
    DBuf<double> val, mval;
    // SELL-64 copy
    DBuf<uint32_t> sell_ptr, sell_col, slice_sys, sys_s0, sys_ns, widths, bad;
    DBuf<double> sell_val, sell_mval, dinv, an;
    DBuf<float> sell_val32, dinv32;
    DBuf<int16_t> sell_off;
    bool built = false, compact = false;
    double delta = kBatchDelta;
    // vectors and state
    DBuf<double> c, b, x0, x1, xinit, r, p, q, partials, relres;
    DBuf<uint32_t> its;
    DBuf<BtSys> sys;
    DBuf<BtGlobal> g;
    uint32_t last_steps = 0;
    int shape[2] = { -1, -1 };
    double shape_ms[2][kBatchShapes] = {};
    // captured rounds
    hipGraphExec_t graph = nullptr;
    std::string graph_key;
    void *hflags = nullptr;
    hipEvent_t ev[2] = { nullptr, nullptr };
    ~rs_batch()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (hipEvent_t &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (hflags) (void)hipHostFree(hflags);
    }
};

namespace {

int bt_build(rs_batch &B)
{
    if (B.built) return RS_OK;
    for (uint32_t i = 0; i < B.count; i++)
        if (!B.have_matrix[i]) return fail(RS_ERR_INVALID_ARG, "rs_batch: system " + std::to_string(i) + " has no matrix");
    bool any_mass = false, all_mass = true;
    for (uint32_t i = 0; i < B.count; i++) {
        any_mass |= B.have_mass[i] != 0;
        all_mass &= B.have_mass[i] != 0 || B.n[i] == 0;
    }
    if (any_mass && !all_mass) return fail(RS_ERR_INVALID_ARG, "rs_batch: a mass matrix is given for some systems only");
    hipStream_t st = runtime().stream;
    int rc;
    if ((rc = B.widths.alloc(std::max(1u, B.nslices))) || (rc = B.sell_ptr.alloc(B.nslices + 1ull))) return rc;
    if (B.nslices) k_bt_widths<<<bt_vec_grid(B.nslices), kBlock, 0, st>>>(B.rowptr.p, B.nslices, B.widths.p);
    RS_HIP(hipGetLastError());
    std::vector<uint32_t> w(B.nslices), sp(B.nslices + 1ull, 0);
    if (B.nslices) RS_HIP(hipMemcpyAsync(w.data(), B.widths.p, 4ull * B.nslices, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    uint64_t acc = 0;
    for (uint32_t s = 0; s < B.nslices; s++) {
        acc += (uint64_t)w[s] * kSlice;
        if (acc > 0xFFFFFFF0ull) return fail(RS_ERR_UNSUPPORTED, "rs_batch: SELL-64 copy exceeds 2^32 entries");
        sp[s + 1] = (uint32_t)acc;
    }
    B.sell_entries = acc;
    size_t E = std::max<uint64_t>(1, acc);
    if ((rc = B.sell_col.alloc(E)) || (rc = B.sell_val.alloc(E)) || (rc = B.sell_val32.alloc(E)) ||
        (rc = B.sell_off.alloc(E)) ||
        (rc = B.sell_mval.alloc(any_mass ? E : 1)) || (rc = B.dinv.alloc(std::max(1u, B.N))) ||
        (rc = B.dinv32.alloc(std::max(1u, B.N))))
        return rc;
    RS_HIP(hipMemcpyAsync(B.sell_ptr.p, sp.data(), 4ull * sp.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemsetAsync(B.bad.p, 0, sizeof(uint32_t), st));
    if (B.nslices)
        k_bt_sell_fill<<<bt_vec_grid(B.nslices), kBlock, 0, st>>>(B.rowptr.p, B.col.p, B.val.p,
                                                                  any_mass ? B.mval.p : nullptr, B.sell_ptr.p,
                                                                  B.nslices, B.sell_col.p, B.sell_val.p,
                                                                  B.sell_val32.p, any_mass ? B.sell_mval.p : nullptr,
                                                                  B.dinv.p, B.dinv32.p, B.sell_off.p, B.bad.p);
    RS_HIP(hipGetLastError());
    if ((rc = B.an.alloc(B.count))) return rc;
    k_bt_frob<<<B.count, kBlock, 0, st>>>(B.val.p, B.rowptr.p, B.sys_s0.p, B.sys_ns.p, B.an.p);
    RS_HIP(hipGetLastError());
    uint32_t far = 0;
    RS_HIP(hipMemcpyAsync(&far, B.bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    B.compact = far == 0 && std::getenv("RS_BATCH_FULL_COLUMNS") == nullptr;
    B.built = true;
    if (B.graph) {
        (void)hipGraphExecDestroy(B.graph);
        B.graph = nullptr;
    }
    return RS_OK;
}

BtArgs bt_args(rs_batch &B, bool mixed, uint32_t steps, uint32_t n_outer, double cvg, int32_t warm)
{
    BtArgs a{};
    a.nslices = B.nslices;
    a.count = B.count;
    a.steps = steps;
    a.n_outer = n_outer;
    a.slice_sys = B.slice_sys.p;
    a.sys_s0 = B.sys_s0.p;
    a.sys_ns = B.sys_ns.p;
    a.an = B.an.p;
    a.sell_ptr = B.sell_ptr.p;
    a.sell_col = B.sell_col.p;
    a.sell_off = B.compact ? B.sell_off.p : nullptr;
    a.sell_val = B.sell_val.p;
    a.sell_val32 = B.sell_val32.p;
    a.sell_mval = B.have_mass.size() && std::any_of(B.have_mass.begin(), B.have_mass.end(), [](uint8_t v) { return v; })
                      ? B.sell_mval.p
                      : nullptr;
    a.c = B.c.p;
    a.b = B.b.p;
    a.x0 = B.x0.p;
    a.x1 = B.x1.p;
    a.xinit = B.xinit.p;
    a.dinv = B.dinv.p;
    a.dinv32 = B.dinv32.p;
    a.r = B.r.p;
    a.p = B.p.p;
    a.q = B.q.p;
    a.partials = B.partials.p;
    a.sys = B.sys.p;
    a.g = B.g.p;
    a.its = B.its.p;
    a.relres = B.relres.p;
    a.cvg = cvg;
    a.delta = mixed ? B.delta : 0.0;
    a.warm = warm;
    return a;
}

void bt_spmv_launch(int id, bool mixed, uint32_t nslices, hipStream_t st, const BtArgs &a)
{
    const bool off = a.sell_off != nullptr;
    if (mixed && off) RS_BT_SHAPE(id, k_bt_spmv, float, true, nslices, st, a);
    else if (mixed) RS_BT_SHAPE(id, k_bt_spmv, float, false, nslices, st, a);
    else if (off) RS_BT_SHAPE(id, k_bt_spmv, double, true, nslices, st, a);
    else RS_BT_SHAPE(id, k_bt_spmv, double, false, nslices, st, a);
}

void bt_event_launch(int id, bool mixed, uint32_t nslices, hipStream_t st, const BtArgs &a)
{
    const bool off = a.sell_off != nullptr;
    if (mixed && off) RS_BT_SHAPE(id, k_bt_event, float, true, nslices, st, a);
    else if (mixed) RS_BT_SHAPE(id, k_bt_event, float, false, nslices, st, a);
    else if (off) RS_BT_SHAPE(id, k_bt_event, double, true, nslices, st, a);
    else RS_BT_SHAPE(id, k_bt_event, double, false, nslices, st, a);
}

// The batched SpMV autotuner: times every SELL shape (lanes per row x waves per workgroup)
// over all systems once per batch and precision and keeps the fastest.
int bt_tune(rs_batch &B, bool mixed)
{
    const int slot = mixed ? 1 : 0;
    if (B.shape[slot] >= 0) return RS_OK;
    if (B.nslices == 0) {
        B.shape[slot] = 0;
        return RS_OK;
    }
    hipStream_t st = runtime().stream;
    BtArgs a = bt_args(B, mixed, 1, 1, 0.0, 0);
    k_bt_init<<<(B.count + kBlock - 1) / kBlock, kBlock, 0, st>>>(a, 1);
    const uint64_t rows = (uint64_t)B.nslices * kSlice;
    if (mixed) k_bt_fill32<<<bt_vec_grid(B.nslices), kBlock, 0, st>>>(reinterpret_cast<float *>(B.p.p), rows, 1.0f);
    else k_bt_fill<<<bt_vec_grid(B.nslices), kBlock, 0, st>>>(B.p.p, rows, 1.0);
    RS_HIP(hipGetLastError());
    const uint32_t reps = B.sell_entries > 64000000ull ? 3 : 10;
    EventPair e0e1_;
    RS_HIP(e0e1_.create());
    hipEvent_t e0 = e0e1_.e[0], e1 = e0e1_.e[1];
    double best = 1e30;
    int best_id = 0;
    for (int id = 0; id < kBatchShapes; id++) {
        bt_spmv_launch(id, mixed, B.nslices, st, a);
        RS_HIP(hipEventRecord(e0, st));
        for (uint32_t k = 0; k < reps; k++) bt_spmv_launch(id, mixed, B.nslices, st, a);
        RS_HIP(hipEventRecord(e1, st));
        RS_HIP(hipEventSynchronize(e1));
        RS_HIP(hipGetLastError());
        float f = 0.0f;
        RS_HIP(hipEventElapsedTime(&f, e0, e1));
        B.shape_ms[slot][id] = f / reps;
        if (f / reps < best) {
            best = f / reps;
            best_id = id;
        }
    }
    B.shape[slot] = best_id;
    return RS_OK;
}

void bt_round(int id, bool mixed, uint32_t nslices, uint32_t count, hipStream_t st, const BtArgs &a)
{
    const uint32_t vg = bt_vec_grid(nslices);
    if (mixed) k_bt_dir<float><<<vg, kBlock, 0, st>>>(a);
    else k_bt_dir<double><<<vg, kBlock, 0, st>>>(a);
    bt_spmv_launch(id, mixed, nslices, st, a);
    k_bt_fin<BT_ALPHA><<<count, kBlock, 0, st>>>(a);
    if (mixed) k_bt_update<float><<<vg, kBlock, 0, st>>>(a);
    else k_bt_update<double><<<vg, kBlock, 0, st>>>(a);
    k_bt_fin<BT_UPDATE><<<count, kBlock, 0, st>>>(a);
    bt_event_launch(id, mixed, nslices, st, a);
    k_bt_fin<BT_EVENT><<<count, kBlock, 0, st>>>(a);
}

int bt_check_system(rs_batch *B, uint32_t i, const char *fn)
{
    if (!B) return fail(RS_ERR_INVALID_ARG, std::string(fn) + ": null batch");
    if (i >= B->count) return fail(RS_ERR_INVALID_ARG, std::string(fn) + ": system index out of range");
    return require_runtime();
}

// Copy (first call for system i) or check (later calls) system i's pattern.
int bt_pattern(rs_batch *B, uint32_t i, const rs_matrix *A, const char *fn)
{
    if (A->n != B->n[i] || A->nnz != B->nnz[i])
        return fail(RS_ERR_INVALID_ARG, std::string(fn) + ": matrix size differs from the size given at create");
    hipStream_t st = runtime().stream;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(4096, (A->nnz + kBlock - 1) / kBlock));
    if (!B->have_pattern[i]) {
        k_bt_copy_pattern<<<grid, kBlock, 0, st>>>(A->rowptr.p, A->col.p, A->n, B->n_pad[i], A->nnz, B->roff[i],
                                                   (uint32_t)B->joff[i], B->rowptr.p, B->col.p);
        RS_HIP(hipGetLastError());
        B->have_pattern[i] = 1;
        return RS_OK;
    }
    RS_HIP(hipMemsetAsync(B->bad.p, 0, sizeof(uint32_t), st));
    k_bt_check_pattern<<<grid, kBlock, 0, st>>>(A->rowptr.p, A->col.p, A->n, A->nnz, B->roff[i], (uint32_t)B->joff[i],
                                                B->rowptr.p, B->col.p, B->bad.p);
    RS_HIP(hipGetLastError());
    uint32_t bad = 0;
    RS_HIP(hipMemcpyAsync(&bad, B->bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    if (bad) return fail(RS_ERR_INVALID_ARG, std::string(fn) + ": pattern differs from the system's other matrix");
    return RS_OK;
}

}  // namespace

// ============================================================================ C ABI: batch
extern "C" int rs_batch_create(uint32_t count, const uint32_t *n_rows, const uint64_t *nnz, rs_batch **out)
{
    if (!out || !n_rows || !nnz || count == 0) return fail(RS_ERR_INVALID_ARG, "rs_batch_create: null or empty argument");
    int rc = require_runtime();
    if (rc) return rc;
    std::unique_ptr<rs_batch> B(new rs_batch());
    B->count = count;
    B->n.assign(n_rows, n_rows + count);
    B->nnz.assign(nnz, nnz + count);
    B->n_pad.resize(count);
    B->roff.resize(count);
    B->s0.resize(count);
    B->ns.resize(count);
    B->joff.resize(count);
    B->have_pattern.assign(count, 0);
    B->have_matrix.assign(count, 0);
    B->have_mass.assign(count, 0);
    B->have_rhs.assign(count, 0);
    uint64_t rows = 0, ent = 0;
    for (uint32_t i = 0; i < count; i++) {
        const uint64_t np = ((uint64_t)n_rows[i] + kSlice - 1) / kSlice * kSlice;
        B->roff[i] = (uint32_t)rows;
        B->n_pad[i] = (uint32_t)np;
        B->s0[i] = (uint32_t)(rows / kSlice);
        B->ns[i] = (uint32_t)(np / kSlice);
        B->joff[i] = ent;
        rows += np;
        ent += nnz[i];
        if (rows >= 0xFFFFFFC0ull || ent >= 0xFFFFFFF0ull)
            return fail(RS_ERR_UNSUPPORTED, "rs_batch_create: more than 2^32 rows or nonzeros in the batch");
    }
    B->N = (uint32_t)rows;
    B->nslices = (uint32_t)(rows / kSlice);
    B->NNZ = ent;
    const size_t N1 = std::max<uint64_t>(1, rows), E1 = std::max<uint64_t>(1, ent);
    if ((rc = B->rowptr.alloc(rows + 1)) || (rc = B->col.alloc(E1)) || (rc = B->val.alloc(E1)) ||
        (rc = B->c.alloc(N1)) || (rc = B->b.alloc(N1)) || (rc = B->x0.alloc(N1)) || (rc = B->x1.alloc(N1)) || (rc = B->xinit.alloc(N1)) ||
        (rc = B->r.alloc(N1)) || (rc = B->p.alloc(N1)) || (rc = B->q.alloc(N1)) ||
        (rc = B->partials.alloc(3ull * std::max(1u, B->nslices))) || (rc = B->sys.alloc(count)) ||
        (rc = B->g.alloc(1)) || (rc = B->slice_sys.alloc(std::max(1u, B->nslices))) ||
        (rc = B->sys_s0.alloc(count)) || (rc = B->sys_ns.alloc(count)) || (rc = B->bad.alloc(1)))
        return rc;
    hipStream_t st = runtime().stream;
    const uint32_t vg = bt_vec_grid(std::max(1u, B->nslices));
    k_bt_fill<<<vg, kBlock, 0, st>>>(B->c.p, N1, 0.0);
    k_bt_fill<<<vg, kBlock, 0, st>>>(B->b.p, N1, 0.0);
    k_bt_fill<<<vg, kBlock, 0, st>>>(B->x0.p, N1, 0.0);
    k_bt_fill<<<vg, kBlock, 0, st>>>(B->x1.p, N1, 0.0);
    k_bt_fill<<<vg, kBlock, 0, st>>>(B->xinit.p, N1, 0.0);
    k_bt_fill<<<vg, kBlock, 0, st>>>(B->r.p, N1, 0.0);
    k_bt_fill<<<vg, kBlock, 0, st>>>(B->p.p, N1, 0.0);
    k_bt_fill<<<vg, kBlock, 0, st>>>(B->q.p, N1, 0.0);
    RS_HIP(hipGetLastError());
    const uint32_t last = (uint32_t)ent;
    RS_HIP(hipMemcpyAsync(B->rowptr.p + rows, &last, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    std::vector<uint32_t> ss(std::max(1u, B->nslices), 0);
    for (uint32_t i = 0; i < count; i++)
        for (uint32_t s = 0; s < B->ns[i]; s++) ss[B->s0[i] + s] = i;
    RS_HIP(hipMemcpyAsync(B->slice_sys.p, ss.data(), 4ull * ss.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(B->sys_s0.p, B->s0.data(), 4ull * count, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(B->sys_ns.p, B->ns.data(), 4ull * count, hipMemcpyHostToDevice, st));
    RS_HIP(hipStreamSynchronize(st));
    // systems without rows need no matrix
    for (uint32_t i = 0; i < count; i++)
        if (n_rows[i] == 0 && nnz[i] == 0) B->have_pattern[i] = B->have_matrix[i] = 1;
    *out = B.release();
    return RS_OK;
}

extern "C" int rs_batch_set_matrix(rs_batch *B, uint32_t i, const rs_matrix *A)
{
    int rc = bt_check_system(B, i, "rs_batch_set_matrix");
    if (rc) return rc;
    if (!A) return fail(RS_ERR_INVALID_ARG, "rs_batch_set_matrix: null matrix");
    if ((rc = bt_pattern(B, i, A, "rs_batch_set_matrix"))) return rc;
    if (A->nnz)
        RS_HIP(hipMemcpyAsync(B->val.p + B->joff[i], A->val.p, 8ull * A->nnz, hipMemcpyDeviceToDevice, runtime().stream));
    RS_HIP(hipStreamSynchronize(runtime().stream));
    B->have_matrix[i] = 1;
    B->built = false;
    return RS_OK;
}

extern "C" int rs_batch_set_mass(rs_batch *B, uint32_t i, const rs_matrix *M)
{
    int rc = bt_check_system(B, i, "rs_batch_set_mass");
    if (rc) return rc;
    if (!M) return fail(RS_ERR_INVALID_ARG, "rs_batch_set_mass: null matrix");
    if (!B->mval.p && (rc = B->mval.alloc(std::max<uint64_t>(1, B->NNZ)))) return rc;
    if ((rc = bt_pattern(B, i, M, "rs_batch_set_mass"))) return rc;
    if (M->nnz)
        RS_HIP(hipMemcpyAsync(B->mval.p + B->joff[i], M->val.p, 8ull * M->nnz, hipMemcpyDeviceToDevice, runtime().stream));
    RS_HIP(hipStreamSynchronize(runtime().stream));
    B->have_mass[i] = 1;
    B->built = false;
    return RS_OK;
}

extern "C" int rs_batch_set_rhs(rs_batch *B, uint32_t i, const double *d_c, const double *d_x0)
{
    int rc = bt_check_system(B, i, "rs_batch_set_rhs");
    if (rc) return rc;
    if (!d_c && B->n[i]) return fail(RS_ERR_INVALID_ARG, "rs_batch_set_rhs: null right-hand side");
    hipStream_t st = runtime().stream;
    const size_t bytes = 8ull * B->n[i];
    if (bytes) {
        RS_HIP(hipMemcpyAsync(B->c.p + B->roff[i], d_c, bytes, hipMemcpyDeviceToDevice, st));
        // the initial solution is the "previous step" of step 0
        if (d_x0) RS_HIP(hipMemcpyAsync(B->xinit.p + B->roff[i], d_x0, bytes, hipMemcpyDeviceToDevice, st));
        else RS_HIP(hipMemsetAsync(B->xinit.p + B->roff[i], 0, bytes, st));
    }
    RS_HIP(hipStreamSynchronize(st));
    B->have_rhs[i] = 1;
    return RS_OK;
}

extern "C" int rs_batch_solve(rs_batch *B, const rs_conf *conf, uint32_t steps, int32_t warm_start,
                              rs_batch_info *info)
{
    if (!B || !conf || !info) return fail(RS_ERR_INVALID_ARG, "rs_batch_solve: null argument");
    if (steps == 0) return fail(RS_ERR_INVALID_ARG, "rs_batch_solve: steps must be >= 1");
    if (conf->solver_type != RS_SOLVER_CG) return fail(RS_ERR_UNSUPPORTED, "rs_batch_solve: the batch runs CG only");
    if (conf->stop_rule != RS_STOP_RELRES)
        return fail(RS_ERR_UNSUPPORTED, "rs_batch_solve: the batch stops on the relative residual (RS_STOP_RELRES)");
    if (conf->precision != RS_PRECISION_FP64 && conf->precision != RS_PRECISION_MIXED)
        return fail(RS_ERR_INVALID_ARG, "rs_batch_solve: unknown precision");
    if (conf->n_outer_iterations == 0) return fail(RS_ERR_INVALID_ARG, "rs_batch_solve: n_outer_iterations must be >= 1");
    int rc = require_runtime();
    if (rc) return rc;
    for (uint32_t i = 0; i < B->count; i++)
        if (B->n[i] && !B->have_rhs[i])
            return fail(RS_ERR_INVALID_ARG, "rs_batch_solve: system " + std::to_string(i) + " has no right-hand side");
    if ((rc = bt_build(*B))) return rc;
    const bool mixed = conf->precision == RS_PRECISION_MIXED;
    if ((rc = bt_tune(*B, mixed))) return rc;
    if (B->last_steps != steps) {
        if ((rc = B->its.alloc((uint64_t)B->count * steps)) || (rc = B->relres.alloc((uint64_t)B->count * steps))) return rc;
        B->last_steps = steps;
    }
    hipStream_t st = runtime().stream;
    k_bt_fill_u32<<<64, kBlock, 0, st>>>(B->its.p, (uint64_t)B->count * steps, 0u);
    k_bt_fill<<<64, kBlock, 0, st>>>(B->relres.p, (uint64_t)B->count * steps, 0.0);
    const int id = B->shape[mixed ? 1 : 0];
    BtArgs a = bt_args(*B, mixed, steps, conf->n_outer_iterations, conf->solver_cvg_value, warm_start ? 1 : 0);
    const char *ng = std::getenv("RS_NO_GRAPH");
    const bool eager = ng && ng[0] == '1';
    auto body = [&]() {
        for (uint32_t k = 0; k < kBatchRounds; k++) bt_round(id, mixed, B->nslices, B->count, st, a);
    };
    std::string key(reinterpret_cast<const char *>(&a), sizeof(a));
    key.append(reinterpret_cast<const char *>(&id), sizeof(id));
    if (!eager && !(B->graph && B->graph_key == key)) {
        if (B->graph) {
            (void)hipGraphExecDestroy(B->graph);
            B->graph = nullptr;
        }
        hipGraph_t gr;
        RS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        body();
        hipError_t e = hipStreamEndCapture(st, &gr);
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&B->graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) {
            B->graph = nullptr;
            return hip_fail(e, "hipGraphInstantiate");
        }
        B->graph_key = key;
    }
    if (!B->hflags) {
        RS_HIP(hipHostMalloc(&B->hflags, 2 * 16, hipHostMallocDefault));
        RS_HIP(hipEventCreateWithFlags(&B->ev[0], hipEventDisableTiming));
        RS_HIP(hipEventCreateWithFlags(&B->ev[1], hipEventDisableTiming));
    }
    EventPair t0t1_;
    RS_HIP(t0t1_.create());
    hipEvent_t t0 = t0t1_.e[0], t1 = t0t1_.e[1];
    RS_HIP(hipEventRecord(t0, st));
    k_bt_init<<<(B->count + kBlock - 1) / kBlock, kBlock, 0, st>>>(a, 0);
    RS_HIP(hipGetLastError());
    uint32_t *hf = static_cast<uint32_t *>(B->hflags);
    auto enqueue = [&](int slot) -> int {
        if (eager) {
            body();
            RS_HIP(hipGetLastError());
        } else {
            RS_HIP(hipGraphLaunch(B->graph, st));
        }
        RS_HIP(hipMemcpyAsync(hf + 4 * slot, B->g.p, 16, hipMemcpyDeviceToHost, st));
        RS_HIP(hipEventRecord(B->ev[slot], st));
        return RS_OK;
    };
    // two replays in flight: the host's flag check overlaps device work.  A system needs at
    // most n_outer + 2 rounds per step, so more rounds than that mean a broken state.
    const uint64_t max_rounds = (uint64_t)steps * (conf->n_outer_iterations + 2ull) + 2 * kBatchRounds;
    if ((rc = enqueue(0))) return rc;
    uint64_t launched = 1;
    for (;;) {
        const int slot = (int)(launched % 2);
        if ((rc = enqueue(slot))) break;
        launched++;
        const int prev = (int)((launched - 2) % 2);
        hipError_t e = hipEventSynchronize(B->ev[prev]);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipEventSynchronize");
            break;
        }
        if (hf[4 * prev + 1]) break;
        if (hf[4 * prev] > max_rounds) {
            rc = fail(RS_ERR_DEVICE, "rs_batch_solve: the batched loop did not finish within its round bound");
            break;
        }
    }
    RS_HIP(hipEventRecord(t1, st));
    RS_HIP(hipStreamSynchronize(st));
    if (rc) return rc;
    float ms = 0.0f;
    RS_HIP(hipEventElapsedTime(&ms, t0, t1));
    // summary
    std::vector<BtSys> S(B->count);
    std::vector<uint32_t> its((uint64_t)B->count * steps);
    std::vector<double> rr((uint64_t)B->count * steps);
    BtGlobal g{};
    RS_HIP(hipMemcpyAsync(S.data(), B->sys.p, sizeof(BtSys) * B->count, hipMemcpyDeviceToHost, st));
    RS_HIP(hipMemcpyAsync(its.data(), B->its.p, 4ull * its.size(), hipMemcpyDeviceToHost, st));
    RS_HIP(hipMemcpyAsync(rr.data(), B->relres.p, 8ull * rr.size(), hipMemcpyDeviceToHost, st));
    RS_HIP(hipMemcpyAsync(&g, B->g.p, sizeof(g), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    std::memset(info, 0, sizeof(*info));
    info->rounds = g.it;
    for (uint32_t i = 0; i < B->count; i++) {
        info->max_iterations = std::max(info->max_iterations, S[i].it_total);
        info->total_iterations += S[i].it_total;
        info->replacements += S[i].nrepl;
        info->systems_converged += S[i].reason == RS_REASON_CONVERGED;
    }
    for (double v : rr) info->max_relres = std::max(info->max_relres, v);
    info->solve_ms = ms;
    return RS_OK;
}

extern "C" int rs_batch_stats(const rs_batch *B, uint32_t *iterations, double *relres, int32_t *reasons)
{
    if (!B) return fail(RS_ERR_INVALID_ARG, "rs_batch_stats: null batch");
    if (!B->last_steps) return fail(RS_ERR_INVALID_ARG, "rs_batch_stats: no solve yet");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    const uint64_t m = (uint64_t)B->count * B->last_steps;
    if (iterations) RS_HIP(hipMemcpyAsync(iterations, B->its.p, 4ull * m, hipMemcpyDeviceToHost, st));
    if (relres) RS_HIP(hipMemcpyAsync(relres, B->relres.p, 8ull * m, hipMemcpyDeviceToHost, st));
    if (reasons) {
        std::vector<BtSys> S(B->count);
        RS_HIP(hipMemcpyAsync(S.data(), B->sys.p, sizeof(BtSys) * B->count, hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
        for (uint32_t i = 0; i < B->count; i++) reasons[i] = (int32_t)S[i].reason;
    }
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

extern "C" int rs_batch_solution(const rs_batch *B, uint32_t i, double *d_x)
{
    if (!B || i >= B->count || (!d_x && B->n[i])) return fail(RS_ERR_INVALID_ARG, "rs_batch_solution: bad argument");
    if (!B->last_steps) return fail(RS_ERR_INVALID_ARG, "rs_batch_solution: no solve yet");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    BtSys S{};
    RS_HIP(hipMemcpyAsync(&S, B->sys.p + i, sizeof(S), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    const double *src = (S.bank ? B->x1.p : B->x0.p) + B->roff[i];
    if (B->n[i]) RS_HIP(hipMemcpyAsync(d_x, src, 8ull * B->n[i], hipMemcpyDeviceToDevice, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

extern "C" int rs_batch_shape(rs_batch *B, int32_t precision, int32_t *shape, double *shape_ms)
{
    if (!B || !shape) return fail(RS_ERR_INVALID_ARG, "rs_batch_shape: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    for (uint32_t i = 0; i < B->count; i++)
        if (!B->have_matrix[i]) return fail(RS_ERR_INVALID_ARG, "rs_batch_shape: matrices missing");
    const bool mixed = precision == RS_PRECISION_MIXED;
    if ((rc = bt_build(*B)) || (rc = bt_tune(*B, mixed))) return rc;
    *shape = B->shape[mixed ? 1 : 0];
    if (shape_ms)
        for (int k = 0; k < kBatchShapes; k++) shape_ms[k] = B->shape_ms[mixed ? 1 : 0][k];
    return RS_OK;
}

extern "C" int rs_batch_force_shape(rs_batch *B, int32_t precision, int32_t shape)
{
    if (!B || shape < -1 || shape >= kBatchShapes) return fail(RS_ERR_INVALID_ARG, "rs_batch_force_shape: bad argument");
    B->shape[precision == RS_PRECISION_MIXED ? 1 : 0] = shape;
    return RS_OK;
}

extern "C" int rs_batch_info_layout(const rs_batch *B, uint32_t *padded_rows, uint64_t *nnz, uint64_t *sell_entries,
                                    int32_t *column_bytes)
{
    if (!B) return fail(RS_ERR_INVALID_ARG, "rs_batch_info_layout: null batch");
    if (column_bytes) *column_bytes = B->built ? (B->compact ? 2 : 4) : 0;
    if (padded_rows) *padded_rows = B->N;
    if (nnz) *nnz = B->NNZ;
    if (sell_entries) *sell_entries = B->built ? B->sell_entries : 0;
    return RS_OK;
}

extern "C" int rs_batch_set_replacement(rs_batch *B, double ratio)
{
    if (!B || !(ratio >= 0.0 && ratio < 1.0)) return fail(RS_ERR_INVALID_ARG, "rs_batch_set_replacement: ratio must be in [0, 1)");
    B->delta = ratio;
    return RS_OK;
}

extern "C" int rs_batch_destroy(rs_batch *B)
{
    delete B;
    return RS_OK;
}
