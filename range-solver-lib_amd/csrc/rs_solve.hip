// rs_solve.hip -- device CSR matrix, preconditioners and the Krylov solvers for gfx950.
//
// Reference behaviour restated (Range-Software/range-solver-lib):
//   SparseMatrixCache / ||A||_F       src/rmatrixsolver.cpp:14-109
//   RMatrixSolver::solve              src/rmatrixsolver.cpp:158-201 (equation scaling)
//   RMatrixSolver::solveCG            src/rmatrixsolver.cpp:208-355
//   RMatrixSolver::solveGMRES         src/rmatrixsolver.cpp:357-581
//   RMatrixPreconditioner             src/rmatrixpreconditioner.cpp:16-190
//   RIterationInfo                    src/riterationinfo.cpp:74-193
//
// Layout in HBM: CSR (uint32 row pointers, uint32 columns, FP64 values, FP32 copy on demand),
// SpMV tile boundaries, vectors as contiguous arrays, Jacobi as n inverse diagonal values,
// block Jacobi as n*w values (row m holds row m of its block's inverse, as the reference).
//
// Reductions are deterministic two-step sums: every workgroup of a producer kernel stores
// its partial sums in its own slots (plain stores), then a one-workgroup finalize kernel sums
// the slots in workgroup order and updates the solver scalars.  The kernel boundary makes the
// partials visible across the XCD L2s without per-workgroup release fences (an agent-scope
// release writes back the issuing XCD's L2, which serialised 2048 workgroups per reduction).
//
// One CG iteration = 5 launches:
//   dir        p = P r + beta p
//   spmv_dot   q = A p, partial p.q                 (SpMV shape picked by the autotuner)
//   fin DOT    alpha = rho / p.q
//   update     x += alpha p, r -= alpha q, partial |x|^2, |r|^2, r.P r
//   fin UPDATE error, trend, stopping rule, beta
// S iterations are captured once in a hipGraph and replayed; every kernel returns at once
// when the device-side stop flag is set, so the iteration count equals the reference's.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <memory>
#include <mutex>

#include "rs_device.hpp"
#include "rs_internal.hpp"
#include "rs_spmv.hpp"

namespace rs {

// ============================================================================ device state
struct SolveState {
    double rho0, rho1, beta, alpha;
    double bn, an, error, trend;
    double cvg, scale, rmax, t0;
    double t1, t2, t3, t4;
    uint32_t it, done, reason, n_outer;
    uint32_t stop_rule, hist_cap, mode, it0;   // it0: first iteration of the current CG segment
    uint32_t replace, no_gain, nrepl, pad;     // device-side residual replacement (flexible CG)
    double best;                               // lowest FP64 residual so far (stagnation test)
};
enum { kStopRunning = 0, kStopDone = 1, kStopPause = 2 };
constexpr double kReliableDelta = 0.1;   // FP64 residual replacement when |r| drops 10x
constexpr uint32_t kMaxReplacementsWithoutGain = 20;

// Block sum of NV values, thread 0 stores them at partials[k*gridDim.x + blockIdx.x].
template <int NV>
__device__ __forceinline__ void store_partials(double (&v)[NV], double *partials)
{
    __shared__ double lds[NV * (kBlock / kWave)];
    block_sum<NV, kBlock>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) partials[k * gridDim.x + blockIdx.x] = v[k];
    }
}

// ============================================================================ finalize
enum FinOp {
    FIN_NORM0 = 0,      // t0 = sqrt(s0)
    FIN_NORM01,         // t0 = sqrt(s0), t1 = sqrt(s1)
    FIN_SCALE,          // t0 = |b|, scale = 1e9 / |bn / an|   (src/rmatrixsolver.cpp:164-171)
    FIN_CG_INIT,        // bn = sqrt(s0)
    FIN_RHO,            // rho0 = s0
    FIN_DOT,            // alpha = rho0 / clamp(p.q), rho1 = rho0       (:302-321)
    FIN_UPDATE,         // stopping rule, beta                          (:333-352, :258-275)
    FIN_MP_UPDATE,      // mixed: recursive residual, pause for replacement
    FIN_MP_TRUE,        // mixed: FP64 true residual (flag: first)
    FIN_MP_RHO,         // mixed: rho after a replacement (flag: first)
    FIN_FCG_BETA,       // flexible CG: rho = r.z, beta = -(z.q_prev) / (p.q)_prev
    FIN_FCG_UPDATE,     // flexible CG: recursive residual, request a replacement
    FIN_MP_REPLACE,     // flexible CG: FP64 true residual of a requested replacement
};

constexpr int kFinBlock = 1024;

// The solver scalars from the finalized sums v (thread 0 of the finalizing workgroup).
__device__ __forceinline__ void fin_apply(int op, int flag, const double (&v)[4], SolveState *st, double *__restrict__ hist)
{
    const double eps = kEps;
    switch (op) {
        case FIN_NORM0: st->t0 = sqrt(v[0]); break;
        case FIN_NORM01: st->t0 = sqrt(v[0]); st->t1 = sqrt(v[1]); break;
        case FIN_SCALE: {
            double bn = sqrt(v[0]), an = st->an;
            st->t0 = bn;
            st->scale = (bn != 0.0 && an != 0.0) ? 1.0e9 / fabs(bn / an) : 1.0;
            break;
        }
        case FIN_CG_INIT: st->bn = sqrt(v[0]); break;
        case FIN_RHO: st->rho0 = v[0]; break;
        case FIN_DOT: {
            double dot = v[0];
            if (dot > 0.0) dot = fmax(dot, eps);
            else if (dot < 0.0) dot = fmin(dot, -eps);
            else dot = eps;
            st->alpha = st->rho0 / dot;
            st->rho1 = st->rho0;
            st->t3 = dot;
            break;
        }
        case FIN_UPDATE: {
            double xn = sqrt(v[0]), rn = sqrt(v[1]);
            double norm = st->an * xn + st->bn;
            if (fabs(norm) < eps) norm = eps;
            double err = st->stop_rule == RS_STOP_RELRES ? rn / fmax(st->bn, eps) : rn / norm;
            double trend = err - st->error;
            st->trend = trend;
            st->error = err;
            uint32_t it = st->it;
            if (it < st->hist_cap) {
                hist[2 * it] = err;
                hist[2 * it + 1] = trend;
            }
            st->it = it + 1;
            int reason = converged_reason(err, trend, st->cvg, eps);
            if (reason != RS_REASON_RUNNING) {
                st->reason = reason;
                st->done = kStopDone;
            } else if (it + 1 >= st->n_outer) {
                st->reason = RS_REASON_MAXIT;
                st->done = kStopDone;
            } else {
                st->beta = v[2] / st->rho1;
                st->rho0 = v[2];
            }
            break;
        }
        case FIN_MP_UPDATE: {
            double rn = sqrt(v[0]);
            double err = rn / fmax(st->bn, eps);
            double trend = err - st->error;
            st->trend = trend;
            st->error = err;
            uint32_t it = st->it;
            if (it < st->hist_cap) {
                hist[2 * it] = err;
                hist[2 * it + 1] = trend;
            }
            st->it = it + 1;
            if (!dev_isnormal(err) || !isfinite(v[1])) {
                st->reason = RS_REASON_NONNORMAL;
                st->done = kStopDone;
            } else if (it + 1 >= st->n_outer) {
                st->reason = RS_REASON_MAXIT;
                st->done = kStopDone;
            } else {
                st->beta = v[1] / st->rho1;
                st->rho0 = v[1];
                if (rn < kReliableDelta * st->rmax || err < st->cvg) st->done = kStopPause;
            }
            break;
        }
        case FIN_MP_TRUE: {
            double rn = sqrt(v[0]);
            if (flag) st->bn = sqrt(v[1]);
            double err = rn / fmax(st->bn, eps);
            st->rmax = rn;
            st->t2 = err;
            if (flag) {
                st->best = err;
                st->no_gain = 0;
                st->replace = 0;
            }
            if (!flag) {
                st->error = err;
                uint32_t it = st->it;
                if (it > 0 && it - 1 < st->hist_cap) hist[2 * (it - 1)] = err;
            }
            if (!dev_isnormal(err) && err != 0.0) {
                st->reason = RS_REASON_NONNORMAL;
                st->done = kStopDone;
            } else if (err < st->cvg) {
                st->reason = RS_REASON_CONVERGED;
                st->done = kStopDone;
            } else {
                st->done = kStopRunning;
            }
            break;
        }
        case FIN_MP_RHO:
            if (!flag) st->beta = v[0] / st->rho1;
            st->rho0 = v[0];
            break;
        case FIN_FCG_BETA:
            st->rho0 = v[0];
            st->beta = (st->it == st->it0 || st->t3 == 0.0) ? 0.0 : -v[1] / st->t3;
            break;
        case FIN_FCG_UPDATE: {
            double rn = sqrt(v[0]);
            double err = rn / fmax(st->bn, eps);
            double trend = err - st->error;
            st->trend = trend;
            st->error = err;
            uint32_t it = st->it;
            if (it < st->hist_cap) {
                hist[2 * it] = err;
                hist[2 * it + 1] = trend;
            }
            st->it = it + 1;
            if (!dev_isnormal(err)) {
                st->reason = RS_REASON_NONNORMAL;
                st->done = kStopDone;
            } else if (it + 1 >= st->n_outer) {
                st->reason = RS_REASON_MAXIT;
                st->done = kStopDone;
            } else if (rn < kReliableDelta * st->rmax || err < st->cvg) {
                st->replace = 1;
            }
            break;
        }
        case FIN_MP_REPLACE: {
            // the FP64 true residual replaces the recursive one; convergence is decided here;
            // kMaxReplacementsWithoutGain replacements without halving the FP64 residual stop
            // the solve as stagnated
            double rn = sqrt(v[0]);
            double err = rn / fmax(st->bn, eps);
            st->rmax = rn;
            st->t2 = err;
            st->error = err;
            uint32_t it = st->it;
            if (it > 0 && it - 1 < st->hist_cap) hist[2 * (it - 1)] = err;
            st->replace = 0;
            st->nrepl++;
            if (!dev_isnormal(err) && err != 0.0) {
                st->reason = RS_REASON_NONNORMAL;
                st->done = kStopDone;
            } else if (err < st->cvg) {
                st->reason = RS_REASON_CONVERGED;
                st->done = kStopDone;
            } else if (err < 0.5 * st->best) {
                st->best = err;
                st->no_gain = 0;
            } else if (++st->no_gain >= kMaxReplacementsWithoutGain) {
                st->reason = RS_REASON_STAGNATED;
                st->done = kStopDone;
            }
            break;
        }
    }
}


__global__ __launch_bounds__(kFinBlock) void k_fin(const double *__restrict__ partials, uint32_t nparts,
                                                   int nv, int op, int flag, SolveState *st,
                                                   double *__restrict__ hist)
{
    if ((op == FIN_DOT || op == FIN_UPDATE || op == FIN_MP_UPDATE || op == FIN_FCG_BETA || op == FIN_FCG_UPDATE) &&
        st->done)
        return;
    if (op == FIN_MP_REPLACE && (st->done || !st->replace)) return;
    __shared__ double lds[4 * (kFinBlock / kWave)];
    double v[4] = { 0.0, 0.0, 0.0, 0.0 };
    // all loads of a thread are independent and issued before the sums (up to 4 partials per
    // value and thread at the launch sizes used here)
    double ld[4][4];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int u = 0; u < 4; u++) {
            uint32_t i = threadIdx.x + u * kFinBlock;
            ld[k][u] = (k < nv && i < nparts) ? partials[(uint64_t)k * nparts + i] : 0.0;
        }
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = (ld[k][0] + ld[k][1]) + (ld[k][2] + ld[k][3]);
    for (uint32_t i = threadIdx.x + 4 * kFinBlock; i < nparts; i += kFinBlock)
        for (int k = 0; k < nv; k++) v[k] += partials[(uint64_t)k * nparts + i];
    block_sum<4, kFinBlock>(v, lds);
    if (threadIdx.x == 0) fin_apply(op, flag, v, st, hist);
}

// k_fin folded into the producing launch.  Thread 0 of every workgroup publishes the NV block
// sums at partials[k*gridDim.x + blockIdx.x] with device-coherent stores, waits until they are
// written and counts its workgroup in the arrival counter of its group (workgroup index mod
// kArriveGroups); the workgroup that completes a group counts the group in the last counter,
// and the workgroup that completes that one adds all partials in index order (the same order on
// every launch, so results are reproducible) and applies OP.  Each workgroup that completes a
// counter resets it, so the counters are zero between launches.  No workgroup waits for another,
// and every store that the last workgroup reads is device-coherent, so no cache write-back is
// needed.  The groups spread the arrivals over kArriveGroups cache lines: with one counter, the
// 2048 workgroups of a launch serialise on it.
constexpr uint32_t kArriveGroups = 8;
constexpr uint32_t kArriveStride = 64;                                  // words: 256 bytes apart
constexpr uint32_t kArriveWords = (kArriveGroups + 1) * kArriveStride;
constexpr uint64_t kPartialDoubles = 4 * 4096;                          // partial sums per solve
// the partial sums of one solve, then its arrival counters
constexpr uint64_t kPartialBytes = kPartialDoubles * sizeof(double) + kArriveWords * sizeof(uint32_t);

template <int NV, int OP>
__device__ __forceinline__ void finalize_in_launch(double (&v)[NV], double *partials, uint32_t *arrive,
                                                   SolveState *st, double *__restrict__ hist)
{
    __shared__ double lds[NV * (kBlock / kWave)];
    __shared__ uint32_t last;
    block_sum<NV, kBlock>(v, lds);
    const uint32_t nparts = gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++)
            __hip_atomic_store(partials + (uint64_t)k * nparts + blockIdx.x, v[k], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);
        const uint32_t g = blockIdx.x % kArriveGroups;
        const uint32_t members = (nparts - g + kArriveGroups - 1) / kArriveGroups;
        const uint32_t groups = nparts < kArriveGroups ? nparts : kArriveGroups;
        uint32_t *gc = arrive + g * kArriveStride, *tc = arrive + kArriveGroups * kArriveStride;
        bool fin = false;
        if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
            __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            fin = __hip_atomic_fetch_add(tc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == groups - 1;
            if (fin) __hip_atomic_store(tc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        last = fin;
    }
    __syncthreads();
    if (!last) return;
    // U partials per value and thread in flight before they are added (one pass at 2048 workgroups)
    constexpr int U = 8;
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) s[k] = 0.0;
    for (uint32_t base = 0; base < nparts; base += U * kBlock) {
        double t[NV][U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = base + u * kBlock + threadIdx.x;
#pragma unroll
            for (int k = 0; k < NV; k++)
                t[k][u] = i < nparts ? __hip_atomic_load(partials + (uint64_t)k * nparts + i, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)
                                     : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int k = 0; k < NV; k++) s[k] += t[k][u];
    }
    block_sum<NV, kBlock>(s, lds);
    if (threadIdx.x == 0) {
        double w[4] = { 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
        for (int k = 0; k < NV; k++) w[k] = s[k];
        fin_apply(OP, 0, w, st, hist);
    }
}

// ============================================================================ epilogues
struct EpiStore64 {
    double *y;
    __device__ void row(uint32_t r, double s) { y[r] = s; }
};
struct EpiStore32 {
    float *y;
    __device__ void row(uint32_t r, float s) { y[r] = s; }
};
// res = b - A x : accumulate |res|^2, |b|^2
struct EpiResidNorm {
    const double *b;
    double acc[2];
    __device__ void row(uint32_t r, double s)
    {
        double bi = b[r], ri = bi - s;
        acc[0] += ri * ri;
        acc[1] += bi * bi;
    }
};
// r = b - A x, |b|^2  (CG start)
struct EpiCgInit {
    const double *b;
    double *r;
    double acc[1];
    __device__ void row(uint32_t i, double s)
    {
        double bi = b[i];
        r[i] = bi - s;
        acc[0] += bi * bi;
    }
};
// q = A p, p.q
struct EpiDot64 {
    const double *p;
    double *q;
    double acc[1];
    __device__ void row(uint32_t r, double s)
    {
        q[r] = s;
        acc[0] += p[r] * s;
    }
};
struct EpiDot32 {
    const float *p;
    float *q;
    double acc[1];
    __device__ void row(uint32_t r, float s)
    {
        q[r] = s;
        acc[0] += (double)p[r] * (double)s;
    }
};
// r32 = (float)(b - A x) in FP64, |r|^2, |b|^2
struct EpiTrueResid32 {
    const double *b;
    float *r;
    double acc[2];
    __device__ void row(uint32_t i, double s)
    {
        double bi = b[i], ri = bi - s;
        r[i] = (float)ri;
        acc[0] += ri * ri;
        acc[1] += bi * bi;
    }
};
// ro = b - A x, |x|^2, |ro|^2  (GMRES outer residual)
struct EpiGmresResid {
    const double *b, *x;
    double *ro;
    double acc[2];
    __device__ void row(uint32_t i, double s)
    {
        double xi = x[i], ri = b[i] - s;
        ro[i] = ri;
        acc[0] += xi * xi;
        acc[1] += ri * ri;
    }
};

// ============================================================================ SpMV kernels
template <class Shape>
__global__ __launch_bounds__(kBlock) void k_spmv(CsrView A, const double *__restrict__ x, double *__restrict__ y)
{
    EpiStore64 e{ y };
    Shape::template run<double>(A, A.val, x, e);
}

// y = A x unless the GMRES inner cycle has stopped
template <class Shape>
__global__ __launch_bounds__(kBlock) void k_gm_spmv(CsrView A, const double *__restrict__ x, double *__restrict__ y,
                                                    const uint32_t *stop)
{
    if (*stop) return;
    EpiStore64 e{ y };
    Shape::template run<double>(A, A.val, x, e);
}

// mixed-precision GMRES: y = A x on the FP32 values and basis, unless the inner cycle has stopped
template <class Shape>
__global__ __launch_bounds__(kBlock) void k_gm_spmv32(CsrView A, const float *__restrict__ x, float *__restrict__ y,
                                                      const uint32_t *stop)
{
    if (*stop) return;
    EpiStore32 e{ y };
    Shape::template run<float>(A, A.val32, x, e);
}

template <class Shape>
__global__ __launch_bounds__(kBlock) void k_spmv32(CsrView A, const float *__restrict__ x, float *__restrict__ y)
{
    EpiStore32 e{ y };
    Shape::template run<float>(A, A.val32, x, e);
}

template <class Shape>
__global__ __launch_bounds__(kBlock) void k_resid_norm(CsrView A, const double *__restrict__ b,
                                                       const double *__restrict__ x, double *partials)
{
    EpiResidNorm e{ b, { 0.0, 0.0 } };
    Shape::template run<double>(A, A.val, x, e);
    store_partials<2>(e.acc, partials);
}

template <class Shape>
__global__ __launch_bounds__(kBlock) void k_gmres_resid(CsrView A, const double *__restrict__ b,
                                                        const double *__restrict__ x, double *__restrict__ ro,
                                                        double *partials)
{
    EpiGmresResid e{ b, x, ro, { 0.0, 0.0 } };
    Shape::template run<double>(A, A.val, x, e);
    store_partials<2>(e.acc, partials);
}

// ============================================================================ misc vector kernels
__global__ __launch_bounds__(kBlock) void k_norm2(const double *__restrict__ a, uint64_t n, double *partials)
{
    double v[1] = { 0.0 };
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        v[0] += a[i] * a[i];
    store_partials<1>(v, partials);
}

__global__ void k_to_f32(const double *__restrict__ a, float *__restrict__ b, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        b[i] = (float)a[i];
}

__global__ void k_fill(double *__restrict__ a, uint64_t n, double v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) a[i] = v;
}

__global__ void k_fill32(float *__restrict__ a, uint64_t n, float v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) a[i] = v;
}

// SELL-64 columns: row r = 64 s + lane writes its entries at sell_ptr[s] + 64 j + lane.
__global__ void k_sell_cols(CsrView A, uint32_t *__restrict__ sell_col)
{
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < (uint64_t)A.nslices * kWave;
         r += (uint64_t)gridDim.x * kBlock) {
        uint32_t s = (uint32_t)(r / kWave), lane = (uint32_t)(r % kWave);
        uint32_t base = A.sell_ptr[s], width = (A.sell_ptr[s + 1] - base) / kWave;
        uint32_t b = 0, len = 0;
        if (r < A.n) {
            b = A.rowptr[r];
            len = A.rowptr[r + 1] - b;
        }
        for (uint32_t j = 0; j < width; j++) sell_col[base + j * kWave + lane] = j < len ? A.col[b + j] : kSellPad;
    }
}

// int16 offsets of the SELL-64 columns from each slice's first row, in the pair slots;
// padding points at the lane's own row, or at row n-1 for the lanes of the last slice past the
// end (the offset shape reads x there unconditionally, and 0 * x must stay finite).  `far`
// counts columns that do not fit.
__global__ void k_sell_offs(CsrView A, const uint32_t *__restrict__ sell_col, int16_t *__restrict__ sell_off,
                            uint32_t *__restrict__ far)
{
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < (uint64_t)A.nslices * kWave;
         r += (uint64_t)gridDim.x * kBlock) {
        uint32_t s = (uint32_t)(r / kWave), lane = (uint32_t)(r % kWave);
        uint32_t base = A.sell_ptr[s], width = (A.sell_ptr[s + 1] - base) / kWave;
        uint32_t miss = 0;
        for (uint32_t j = 0; j < width; j++) {
            const uint32_t c = sell_col[base + j * kWave + lane];
            const int64_t o = c == kSellPad ? (int64_t)min(r, (uint64_t)A.n - 1) - (int64_t)s * kWave
                                            : (int64_t)c - (int64_t)s * kWave;
            if (o < -32767 || o > 32767) miss++;
            else sell_off[sell_pair_slot(base, width, j, lane)] = (int16_t)o;
        }
        if (miss) atomicAdd(far, miss);
    }
}

// FP32 SELL values in the pair slots from the FP64 SELL values in the plain slots.
__global__ void k_sell_to_f32(CsrView A, const double *__restrict__ sell_val, float *__restrict__ sell_val32)
{
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < (uint64_t)A.nslices * kWave;
         r += (uint64_t)gridDim.x * kBlock) {
        const uint32_t s = (uint32_t)(r / kWave), lane = (uint32_t)(r % kWave);
        const uint32_t base = A.sell_ptr[s], width = (A.sell_ptr[s + 1] - base) / kWave;
        for (uint32_t j = 0; j < width; j++)
            sell_val32[sell_pair_slot(base, width, j, lane)] = (float)sell_val[base + j * kWave + lane];
    }
}

// SELL-64 values from the CSR values (padding 0).
__global__ void k_sell_vals(CsrView A, double *__restrict__ sell_val)
{
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < (uint64_t)A.nslices * kWave;
         r += (uint64_t)gridDim.x * kBlock) {
        uint32_t s = (uint32_t)(r / kWave), lane = (uint32_t)(r % kWave);
        uint32_t base = A.sell_ptr[s], width = (A.sell_ptr[s + 1] - base) / kWave;
        uint32_t b = 0, len = 0;
        if (r < A.n) {
            b = A.rowptr[r];
            len = A.rowptr[r + 1] - b;
        }
        for (uint32_t j = 0; j < width; j++) sell_val[base + j * kWave + lane] = j < len ? A.val[b + j] : 0.0;
    }
}

// ============================================================================ preconditioner
// constructJacobi (src/rmatrixpreconditioner.cpp:78-94): 1/A_ii, 0 if absent or zero.
__global__ void k_jacobi_build(CsrView A, double *__restrict__ d)
{
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < A.n; i += gridDim.x * kBlock) {
        double v = 0.0;
        bool found = false;
        for (uint32_t j = A.rowptr[i]; j < A.rowptr[i + 1]; j++)
            if (A.col[j] == i) { v = A.val[j]; found = true; break; }
        d[i] = (found && v != 0.0) ? 1.0 / v : 0.0;
    }
}

// constructBlockJacobi (:96-148): gather the w x w diagonal block (missing -> 0) and invert
// it by Gauss-Jordan with partial pivoting (a singular block becomes zero).
template <int W>
__device__ __forceinline__ void invert_small(double (&m)[W][W], int w)
{
    double inv[W][W];
    for (int i = 0; i < w; i++)
        for (int j = 0; j < w; j++) inv[i][j] = (i == j) ? 1.0 : 0.0;
    for (int c = 0; c < w; c++) {
        int piv = c;
        for (int r = c + 1; r < w; r++)
            if (fabs(m[r][c]) > fabs(m[piv][c])) piv = r;
        if (m[piv][c] == 0.0) {
            for (int i = 0; i < w; i++)
                for (int j = 0; j < w; j++) m[i][j] = 0.0;
            return;
        }
        if (piv != c)
            for (int j = 0; j < w; j++) {
                double t = m[c][j]; m[c][j] = m[piv][j]; m[piv][j] = t;
                t = inv[c][j]; inv[c][j] = inv[piv][j]; inv[piv][j] = t;
            }
        double s = 1.0 / m[c][c];
        for (int j = 0; j < w; j++) { m[c][j] *= s; inv[c][j] *= s; }
        for (int r = 0; r < w; r++) {
            if (r == c) continue;
            double f = m[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < w; j++) { m[r][j] -= f * m[c][j]; inv[r][j] -= f * inv[c][j]; }
        }
    }
    for (int i = 0; i < w; i++)
        for (int j = 0; j < w; j++) m[i][j] = inv[i][j];
}

template <int WMAX>
__global__ void k_blockjacobi_build(CsrView A, uint32_t w, double *__restrict__ data)
{
    uint32_t nb = A.n / w;
    for (uint32_t ib = blockIdx.x * kBlock + threadIdx.x; ib < nb; ib += gridDim.x * kBlock) {
        double m[WMAX][WMAX];
        for (uint32_t k = 0; k < w; k++) {
            for (uint32_t l = 0; l < w; l++) m[k][l] = 0.0;
            uint32_t row = ib * w + k;
            for (uint32_t j = A.rowptr[row]; j < A.rowptr[row + 1]; j++) {
                uint32_t c = A.col[j];
                if (c >= ib * w && c < ib * w + w) m[k][c - ib * w] = A.val[j];
            }
        }
        invert_small<WMAX>(m, (int)w);
        for (uint32_t k = 0; k < w; k++)
            for (uint32_t l = 0; l < w; l++) data[(uint64_t)(ib * w + k) * w + l] = m[k][l];
    }
}

// z = P r on preconditioner block ib (W rows; W == 0: runtime width up to 16).
template <int W, typename T, typename PT>
__device__ __forceinline__ void pc_apply_block(int ptype, uint32_t w, const PT *__restrict__ pd,
                                               uint64_t ib, const T (&r)[W == 0 ? 16 : W],
                                               T (&z)[W == 0 ? 16 : W])
{
    const int ww = W == 0 ? (int)w : W;
    if (ptype == RS_PRECOND_BLOCK_JACOBI) {
        for (int k = 0; k < ww; k++) {
            T s = 0;
            for (int l = 0; l < ww; l++) s += (T)pd[(ib * ww + k) * ww + l] * r[l];
            z[k] = s;
        }
    } else if (ptype == RS_PRECOND_JACOBI) {
        for (int k = 0; k < ww; k++) z[k] = (T)pd[ib * ww + k] * r[k];
    } else {
        for (int k = 0; k < ww; k++) z[k] = r[k];
    }
}

// mixed-precision GMRES: z_k = P v_k on the FP32 preconditioner data and basis
template <int W>
__global__ void k_pc_apply32(int ptype, uint32_t w, const float *__restrict__ pd, uint32_t n,
                             const float *__restrict__ x, float *__restrict__ y, const uint32_t *stop)
{
    if (stop && *stop) return;
    constexpr int WM = W == 0 ? 16 : W;
    const uint32_t ww = W == 0 ? w : W;
    uint64_t nb = n / ww;
    for (uint64_t ib = (uint64_t)blockIdx.x * kBlock + threadIdx.x; ib < nb; ib += (uint64_t)gridDim.x * kBlock) {
        float r[WM], z[WM];
        for (uint32_t k = 0; k < ww; k++) r[k] = x[ib * ww + k];
        pc_apply_block<W, float, float>(ptype, ww, pd, ib, r, z);
        for (uint32_t k = 0; k < ww; k++) y[ib * ww + k] = z[k];
    }
}

// y = P x (RMatrixPreconditioner::compute, and GMRES's z_k = P v_k)
template <int W>
__global__ void k_pc_apply(int ptype, uint32_t w, const double *__restrict__ pd, uint32_t n,
                           const double *__restrict__ x, double *__restrict__ y, const uint32_t *stop)
{
    if (stop && *stop) return;
    constexpr int WM = W == 0 ? 16 : W;
    const uint32_t ww = W == 0 ? w : W;
    uint64_t nb = n / ww;
    for (uint64_t ib = (uint64_t)blockIdx.x * kBlock + threadIdx.x; ib < nb; ib += (uint64_t)gridDim.x * kBlock) {
        double r[WM], z[WM];
        for (uint32_t k = 0; k < ww; k++) r[k] = x[ib * ww + k];
        pc_apply_block<W, double, double>(ptype, ww, pd, ib, r, z);
        for (uint32_t k = 0; k < ww; k++) y[ib * ww + k] = z[k];
    }
}

// ============================================================================ CG (FP64)
struct CgArgs {
    CsrView A;
    const double *b;
    double *x, *r, *p, *q;
    int32_t ptype;
    uint32_t pw;
    const double *pd;
    double *partials;
    uint32_t *arrive;     // kArriveWords arrival counters, zero between launches
    SolveState *st;
    double *hist;
    uint32_t stream;      // the update and direction kernels stream their vectors non-temporally
};

// FP64 CG vectors of at least this size are streamed non-temporally by the update and direction
// kernels: their ~7 vector passes per iteration then leave the L2 and the 256 MB MALL to the
// matrix and to the direction the SpMV gathers (128^3 heat, 17 MB per vector: FP64 Jacobi-CG
// 116-117 -> 112-114 ms).  Smaller vectors stay cached: they fit the 8 XCDs' L2s, and streaming
// them costs (cantilever, 3.3 MB: 146 -> 152 ms).  The FP32 vectors of the mixed and flexible CG
// stay cached too (streaming them made the mixed solve 74 -> 78 ms).
constexpr uint64_t kStreamVectorBytes = 8ull << 20;

// (A runtime choice between the two loads of one address lets the compiler merge them into one
// plain load, so the choice is a template parameter and the kernels branch once, on entry.)
template <bool NT, typename T>
__device__ __forceinline__ T vload(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void vstore(T *p, T v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

__global__ void k_state_init(SolveState *st, double an, double cvg, double prev_error,
                             uint32_t n_outer, uint32_t stop_rule, uint32_t hist_cap, uint32_t mode)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->rho0 = st->rho1 = st->beta = st->alpha = 0.0;
    st->bn = 0.0;
    st->an = an;
    st->error = prev_error;
    st->trend = 0.0;
    st->cvg = cvg;
    st->scale = 1.0;
    st->rmax = 0.0;
    st->t0 = st->t1 = st->t2 = st->t3 = st->t4 = 0.0;
    st->it = 0;
    st->done = n_outer == 0 ? kStopDone : kStopRunning;
    st->reason = n_outer == 0 ? RS_REASON_MAXIT : RS_REASON_RUNNING;
    st->n_outer = n_outer;
    st->stop_rule = stop_rule;
    st->hist_cap = hist_cap;
    st->mode = mode;
    st->it0 = 0;
    st->replace = 0;
    st->no_gain = 0;
    st->nrepl = 0;
    st->best = 0.0;
}

// Start a new CG segment on the same solution (the learned solve's Jacobi-CG fallback).
__global__ void k_state_restart(SolveState *st)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->replace = 0;
    st->it0 = st->it;
    st->done = kStopRunning;
    st->reason = RS_REASON_RUNNING;
    st->beta = 0.0;
}

// y = b*s ; x *= s   (src/rmatrixsolver.cpp:181-184)
__global__ void k_apply_scale(const double *__restrict__ b, double *__restrict__ y,
                              double *__restrict__ x, uint64_t n, const SolveState *st)
{
    double s = st->scale;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        y[i] = b[i] * s;
        x[i] *= s;
    }
}

// x *= 1/s   (src/rmatrixsolver.cpp:198)
__global__ void k_unscale(double *__restrict__ x, uint64_t n, const SolveState *st)
{
    double s = 1.0 / st->scale;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        x[i] *= s;
}

// r = b - A x, partial |b|^2   (src/rmatrixsolver.cpp:236-246)
template <class Shape>
__global__ __launch_bounds__(kBlock) void k_cg_init(CgArgs a)
{
    EpiCgInit e{ a.b, a.r, { 0.0 } };
    Shape::template run<double>(a.A, a.A.val, a.x, e);
    store_partials<1>(e.acc, a.partials);
}

// partial rho0 = r . P r for the first iteration (src/rmatrixsolver.cpp:258-267)
template <int W>
__global__ __launch_bounds__(kBlock) void k_cg_rho(CgArgs a)
{
    constexpr int WM = W == 0 ? 16 : W;
    const uint32_t ww = W == 0 ? a.pw : W;
    const uint64_t nb = a.A.n / ww;
    double v[1] = { 0.0 };
    for (uint64_t ib = (uint64_t)blockIdx.x * kBlock + threadIdx.x; ib < nb; ib += (uint64_t)gridDim.x * kBlock) {
        double r[WM], z[WM];
        for (uint32_t k = 0; k < ww; k++) r[k] = a.r[ib * ww + k];
        pc_apply_block<W, double, double>(a.ptype, ww, a.pd, ib, r, z);
        for (uint32_t k = 0; k < ww; k++) v[0] += r[k] * z[k];
    }
    store_partials<1>(v, a.partials);
}

// p = z (first iteration) or p*beta + z   (src/rmatrixsolver.cpp:278-289)
template <int W, bool NT>
__device__ __forceinline__ void cg_dir(const CgArgs &a)
{
    constexpr int WM = W == 0 ? 16 : W;
    const uint32_t ww = W == 0 ? a.pw : W;
    const uint64_t nb = a.A.n / ww;
    const bool first = a.st->it == a.st->it0;
    const double beta = a.st->beta;
    for (uint64_t ib = (uint64_t)blockIdx.x * kBlock + threadIdx.x; ib < nb; ib += (uint64_t)gridDim.x * kBlock) {
        double r[WM], z[WM];
        for (uint32_t k = 0; k < ww; k++) r[k] = vload<NT>(a.r + ib * ww + k);
        pc_apply_block<W, double, double>(a.ptype, ww, a.pd, ib, r, z);
        for (uint32_t k = 0; k < ww; k++) {
            uint64_t i = ib * ww + k;
            a.p[i] = first ? z[k] : vload<NT>(a.p + i) * beta + z[k];
        }
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_cg_dir(CgArgs a)
{
    if (a.st->done) return;
    if (a.stream) cg_dir<W, true>(a);
    else cg_dir<W, false>(a);
}

// q = A p and p.q, then alpha   (src/rmatrixsolver.cpp:295-321)
template <class Shape>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_dot(CgArgs a)
{
    if (a.st->done) return;
    EpiDot64 e{ a.p, a.q, { 0.0 } };
    Shape::template run<double>(a.A, a.A.val, a.p, e);
    finalize_in_launch<1, FIN_DOT>(e.acc, a.partials, a.arrive, a.st, a.hist);
}

// x += alpha p, r -= alpha q, |x|^2, |r|^2 and r.P r, then the stopping rule and beta
// (src/rmatrixsolver.cpp:324-352)
template <int W, bool NT>
__device__ __forceinline__ void cg_update(const CgArgs &a)
{
    SolveState *st = a.st;
    constexpr int WM = W == 0 ? 16 : W;
    const uint32_t ww = W == 0 ? a.pw : W;
    const uint64_t nb = a.A.n / ww;
    const double alpha = st->alpha;
    double v[3] = { 0.0, 0.0, 0.0 };
    for (uint64_t ib = (uint64_t)blockIdx.x * kBlock + threadIdx.x; ib < nb; ib += (uint64_t)gridDim.x * kBlock) {
        double r[WM], z[WM];
        for (uint32_t k = 0; k < ww; k++) {
            uint64_t i = ib * ww + k;
            double xi = vload<NT>(a.x + i) + alpha * vload<NT>(a.p + i);
            double ri = vload<NT>(a.r + i) - alpha * vload<NT>(a.q + i);
            vstore<NT>(a.x + i, xi);
            vstore<NT>(a.r + i, ri);
            v[0] += xi * xi;
            v[1] += ri * ri;
            r[k] = ri;
        }
        pc_apply_block<W, double, double>(a.ptype, ww, a.pd, ib, r, z);
        for (uint32_t k = 0; k < ww; k++) v[2] += r[k] * z[k];
    }
    finalize_in_launch<3, FIN_UPDATE>(v, a.partials, a.arrive, st, a.hist);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_cg_update(CgArgs a)
{
    if (a.st->done) return;
    if (a.stream) cg_update<W, true>(a);
    else cg_update<W, false>(a);
}

// ============================================================================ CG (mixed)
// FP32 matrix values, FP32 direction/residual/correction vectors and FP32 preconditioner;
// FP64 accumulation of every dot product and norm; FP64 solution.  The recursive FP32
// residual is replaced by the FP64 true residual b - A x whenever it has dropped by
// kReliableDelta since the last replacement, or when it claims convergence (reliable
// updates: Sleijpen & van der Vorst 1996); convergence is decided on the FP64 residual.
struct MpArgs {
    CsrView A;
    const double *b;
    double *x;            // FP64 solution (scaled)
    float *r, *p, *q, *dx;
    int32_t ptype;
    uint32_t pw;
    const float *pd;
    double *partials;
    uint32_t *arrive;     // kArriveWords arrival counters, zero between launches
    SolveState *st;
    double *hist;
};

// Jacobi or no preconditioner (width 1): 4 rows per lane in 16-byte loads and stores, then the
// remaining n % 4 one per lane; the same expressions as pc_apply_block (z = d r).
__device__ __forceinline__ float4 mp_pc4(const MpArgs &a, uint64_t q, const float4 &r)
{
    if (a.ptype != RS_PRECOND_JACOBI) return r;
    const float4 d = reinterpret_cast<const float4 *>(a.pd)[q];
    return make_float4(d.x * r.x, d.y * r.y, d.z * r.z, d.w * r.w);
}

__device__ __forceinline__ float mp_pc1(const MpArgs &a, uint64_t i, float r)
{
    return a.ptype == RS_PRECOND_JACOBI ? a.pd[i] * r : r;
}

__device__ __forceinline__ void mp_dir_jacobi4(const MpArgs &a)
{
    const uint64_t n = a.A.n, nq = n / 4, stride = (uint64_t)gridDim.x * kBlock;
    const bool first = a.st->it == a.st->it0;
    const float beta = (float)a.st->beta;
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < nq; q += stride) {
        const float4 z = mp_pc4(a, q, reinterpret_cast<const float4 *>(a.r)[q]);
        float4 o = z;
        if (!first) {
            const float4 p = reinterpret_cast<const float4 *>(a.p)[q];
            o = make_float4(fmaf(p.x, beta, z.x), fmaf(p.y, beta, z.y), fmaf(p.z, beta, z.z), fmaf(p.w, beta, z.w));
        }
        reinterpret_cast<float4 *>(a.p)[q] = o;
    }
    for (uint64_t i = 4 * nq + (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const float z = mp_pc1(a, i, a.r[i]);
        a.p[i] = first ? z : fmaf(a.p[i], beta, z);
    }
}

__device__ __forceinline__ void mp_update_jacobi4(const MpArgs &a, double (&v)[2])
{
    const uint64_t n = a.A.n, nq = n / 4, stride = (uint64_t)gridDim.x * kBlock;
    const float alpha = (float)a.st->alpha;
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < nq; q += stride) {
        const float4 p = reinterpret_cast<const float4 *>(a.p)[q], dx = reinterpret_cast<const float4 *>(a.dx)[q];
        const float4 qq = reinterpret_cast<const float4 *>(a.q)[q], r = reinterpret_cast<const float4 *>(a.r)[q];
        reinterpret_cast<float4 *>(a.dx)[q] =
            make_float4(fmaf(alpha, p.x, dx.x), fmaf(alpha, p.y, dx.y), fmaf(alpha, p.z, dx.z), fmaf(alpha, p.w, dx.w));
        const float4 rn = make_float4(fmaf(-alpha, qq.x, r.x), fmaf(-alpha, qq.y, r.y), fmaf(-alpha, qq.z, r.z),
                                      fmaf(-alpha, qq.w, r.w));
        reinterpret_cast<float4 *>(a.r)[q] = rn;
        const float4 z = mp_pc4(a, q, rn);
        v[0] += (double)rn.x * (double)rn.x;
        v[0] += (double)rn.y * (double)rn.y;
        v[0] += (double)rn.z * (double)rn.z;
        v[0] += (double)rn.w * (double)rn.w;
        v[1] += (double)rn.x * (double)z.x;
        v[1] += (double)rn.y * (double)z.y;
        v[1] += (double)rn.z * (double)z.z;
        v[1] += (double)rn.w * (double)z.w;
    }
    for (uint64_t i = 4 * nq + (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        a.dx[i] = fmaf(alpha, a.p[i], a.dx[i]);
        const float ri = fmaf(-alpha, a.q[i], a.r[i]);
        a.r[i] = ri;
        v[0] += (double)ri * (double)ri;
        v[1] += (double)ri * (double)mp_pc1(a, i, ri);
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_mp_dir(MpArgs a)
{
    if (a.st->done) return;
    if constexpr (W == 1) {
        if (a.ptype != RS_PRECOND_BLOCK_JACOBI) {
            mp_dir_jacobi4(a);
            return;
        }
    }
    constexpr int WM = W == 0 ? 16 : W;
    const uint32_t ww = W == 0 ? a.pw : W;
    const uint64_t nb = a.A.n / ww;
    const bool first = a.st->it == a.st->it0;
    const float beta = (float)a.st->beta;
    for (uint64_t ib = (uint64_t)blockIdx.x * kBlock + threadIdx.x; ib < nb; ib += (uint64_t)gridDim.x * kBlock) {
        float r[WM], z[WM];
        for (uint32_t k = 0; k < ww; k++) r[k] = a.r[ib * ww + k];
        pc_apply_block<W, float, float>(a.ptype, ww, a.pd, ib, r, z);
        for (uint32_t k = 0; k < ww; k++) {
            uint64_t i = ib * ww + k;
            a.p[i] = first ? z[k] : fmaf(a.p[i], beta, z[k]);
        }
    }
}

template <class Shape>
__global__ __launch_bounds__(kBlock) void k_mp_spmv_dot(MpArgs a)
{
    if (a.st->done) return;
    EpiDot32 e{ a.p, a.q, { 0.0 } };
    Shape::template run<float>(a.A, a.A.val32, a.p, e);
    finalize_in_launch<1, FIN_DOT>(e.acc, a.partials, a.arrive, a.st, a.hist);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_mp_update(MpArgs a)
{
    SolveState *st = a.st;
    if (st->done) return;
    constexpr int WM = W == 0 ? 16 : W;
    const uint32_t ww = W == 0 ? a.pw : W;
    const uint64_t nb = a.A.n / ww;
    const float alpha = (float)st->alpha;
    double v[2] = { 0.0, 0.0 };
    if constexpr (W == 1) {
        if (a.ptype != RS_PRECOND_BLOCK_JACOBI) {
            mp_update_jacobi4(a, v);
            finalize_in_launch<2, FIN_MP_UPDATE>(v, a.partials, a.arrive, st, a.hist);
            return;
        }
    }
    for (uint64_t ib = (uint64_t)blockIdx.x * kBlock + threadIdx.x; ib < nb; ib += (uint64_t)gridDim.x * kBlock) {
        float r[WM], z[WM];
        for (uint32_t k = 0; k < ww; k++) {
            uint64_t i = ib * ww + k;
            a.dx[i] = fmaf(alpha, a.p[i], a.dx[i]);
            float ri = fmaf(-alpha, a.q[i], a.r[i]);
            a.r[i] = ri;
            v[0] += (double)ri * (double)ri;
            r[k] = ri;
        }
        pc_apply_block<W, float, float>(a.ptype, ww, a.pd, ib, r, z);
        for (uint32_t k = 0; k < ww; k++) v[1] += (double)r[k] * (double)z[k];
    }
    finalize_in_launch<2, FIN_MP_UPDATE>(v, a.partials, a.arrive, st, a.hist);
}

// x += dx, dx = 0
__global__ void k_mp_fold(double *__restrict__ x, float *__restrict__ dx, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        x[i] += (double)dx[i];
        dx[i] = 0.0f;
    }
}

// r32 = b - A x in FP64, partial |r|^2, |b|^2
template <class Shape>
__global__ __launch_bounds__(kBlock) void k_mp_true_resid(MpArgs a)
{
    EpiTrueResid32 e{ a.b, a.r, { 0.0, 0.0 } };
    Shape::template run<double>(a.A, a.A.val, a.x, e);
    store_partials<2>(e.acc, a.partials);
}

// The same two steps of a replacement, executed only when the stopping rule requested one
// (flexible CG keeps them inside its captured iterations instead of pausing for the host); the
// true-residual launch finalizes the replacement itself.
__global__ void k_mp_fold_cond(double *__restrict__ x, float *__restrict__ dx, uint64_t n, const SolveState *st)
{
    if (st->done || !st->replace) return;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        x[i] += (double)dx[i];
        dx[i] = 0.0f;
    }
}

template <class Shape>
__global__ __launch_bounds__(kBlock) void k_mp_true_resid_cond(MpArgs a)
{
    if (a.st->done || !a.st->replace) return;
    EpiTrueResid32 e{ a.b, a.r, { 0.0, 0.0 } };
    Shape::template run<double>(a.A, a.A.val, a.x, e);
    finalize_in_launch<2, FIN_MP_REPLACE>(e.acc, a.partials, a.arrive, a.st, a.hist);
}

// partial rho = r . P r after a residual replacement
template <int W>
__global__ __launch_bounds__(kBlock) void k_mp_rho(MpArgs a)
{
    constexpr int WM = W == 0 ? 16 : W;
    const uint32_t ww = W == 0 ? a.pw : W;
    const uint64_t nb = a.A.n / ww;
    double v[1] = { 0.0 };
    for (uint64_t ib = (uint64_t)blockIdx.x * kBlock + threadIdx.x; ib < nb; ib += (uint64_t)gridDim.x * kBlock) {
        float r[WM], z[WM];
        for (uint32_t k = 0; k < ww; k++) r[k] = a.r[ib * ww + k];
        pc_apply_block<W, float, float>(a.ptype, ww, a.pd, ib, r, z);
        for (uint32_t k = 0; k < ww; k++) v[0] += (double)r[k] * (double)z[k];
    }
    store_partials<1>(v, a.partials);
}

// ============================================================================ GMRES kernels
// The GMRES vector kernels take the basis type T: double (FP64), or float for the
// mixed-precision GMRES (FP32 basis, products and norms accumulated in FP64).  The Gram-Schmidt
// kernels read the basis in 16-byte loads (2 FP64 or 4 FP32 entries per lane; basis rows start at
// multiples of 16 bytes, ld % (16 / sizeof(T)) == 0): they are bound by the number of loads in
// flight, so 8- or 4-byte loads per entry would leave the FP32 basis, half the bytes, barely faster
// (measured on the duct flow: 134 against 161 us for the products at 4 bytes per entry).
template <typename T>
struct Load16 {
    static constexpr int kN = 16 / sizeof(T);
    __device__ __forceinline__ static void get(const T *p, T (&e)[kN])
    {
        if constexpr (sizeof(T) == 4) {
            const float4 v = *reinterpret_cast<const float4 *>(p);
            e[0] = v.x;
            e[1] = v.y;
            e[2] = v.z;
            e[3] = v.w;
        } else {
            const double2 v = *reinterpret_cast<const double2 *>(p);
            e[0] = v.x;
            e[1] = v.y;
        }
    }
    __device__ __forceinline__ static void put(T *p, const T (&e)[kN])
    {
        if constexpr (sizeof(T) == 4) *reinterpret_cast<float4 *>(p) = make_float4(e[0], e[1], e[2], e[3]);
        else *reinterpret_cast<double2 *>(p) = make_double2(e[0], e[1]);
    }
};

// h_i = w . v_i for i in [i0, i0+8) -> partials[i][block]; final sum in k_multidot_final
template <typename T>
__global__ __launch_bounds__(kBlock) void k_multidot(const T *__restrict__ w, const T *__restrict__ V, uint64_t ld,
                                                     uint32_t n, uint32_t cnt, double *partials, const uint32_t *stop)
{
    if (*stop) return;
    constexpr int N = Load16<T>::kN;
    __shared__ double lds[8 * (kBlock / kWave)];
    const uint32_t i0 = blockIdx.y * 8;
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = 0.0;
    const uint64_t nq = n / N, stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < nq; q += stride) {
        T wq[N];
        Load16<T>::get(w + N * q, wq);
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (i0 + k < cnt) {
                T vq[N];
                Load16<T>::get(V + (i0 + k) * ld + N * q, vq);
#pragma unroll
                for (int e = 0; e < N; e++) v[k] += (double)wq[e] * (double)vq[e];
            }
    }
    for (uint64_t j = N * nq + (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
        const double wj = (double)w[j];
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (i0 + k < cnt) v[k] += wj * (double)V[(i0 + k) * ld + j];
    }
    block_sum<8, kBlock>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (i0 + k < cnt) partials[(uint64_t)(i0 + k) * gridDim.x + blockIdx.x] = v[k];
    }
}

__global__ __launch_bounds__(kBlock) void k_multidot_final(const double *partials, uint32_t nblk,
                                                           double *out, const uint32_t *stop)
{
    if (*stop) return;
    __shared__ double lds[kBlock / kWave];
    double v[1] = { 0.0 };
    for (uint32_t j = threadIdx.x; j < nblk; j += kBlock) v[0] += partials[(uint64_t)blockIdx.x * nblk + j];
    block_sum<1, kBlock>(v, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

// w -= sum_i h_i v_i ; partial |w|^2  (classical Gram-Schmidt, src/rmatrixsolver.cpp:486-508);
// the sums in FP64, the stored w in T and its norm from the stored values
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gs_subtract(T *__restrict__ w, const T *__restrict__ V, uint64_t ld,
                                                        uint32_t n, uint32_t cnt, const double *__restrict__ h,
                                                        double *partials, const uint32_t *stop)
{
    if (*stop) return;
    constexpr int N = Load16<T>::kN;
    double v[1] = { 0.0 };
    const uint64_t nq = n / N, stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < nq; q += stride) {
        T wq[N];
        Load16<T>::get(w + N * q, wq);
        double a[N];
#pragma unroll
        for (int e = 0; e < N; e++) a[e] = (double)wq[e];
        // FP32: eight basis rows' loads in flight, subtracted in ascending order (FP64 rows are
        // read at ~6.3 TB/s either way, and their unrolled form measured slower: 137.5 vs 129 us)
        constexpr uint32_t U = sizeof(T) == 4 ? 8 : 1;
        uint32_t i = 0;
        for (; i + U <= cnt; i += U) {
            T vq[U][N];
#pragma unroll
            for (uint32_t k = 0; k < U; k++) Load16<T>::get(V + (i + k) * ld + N * q, vq[k]);
#pragma unroll
            for (uint32_t k = 0; k < U; k++) {
                const double hi = h[i + k];
#pragma unroll
                for (int e = 0; e < N; e++) a[e] -= hi * (double)vq[k][e];
            }
        }
        for (; i < cnt; i++) {
            T vq[N];
            Load16<T>::get(V + i * ld + N * q, vq);
            const double hi = h[i];
#pragma unroll
            for (int e = 0; e < N; e++) a[e] -= hi * (double)vq[e];
        }
#pragma unroll
        for (int e = 0; e < N; e++) {
            wq[e] = (T)a[e];
            v[0] += (double)wq[e] * (double)wq[e];
        }
        Load16<T>::put(w + N * q, wq);
    }
    for (uint64_t j = N * nq + (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
        double value = (double)w[j];
        for (uint32_t i = 0; i < cnt; i++) value -= h[i] * (double)V[i * ld + j];
        const T stored = (T)value;
        w[j] = stored;
        v[0] += (double)stored * (double)stored;
    }
    store_partials<1>(v, partials);
}

// v = (h == 0) ? 0 : w / h with h = st->t0
template <typename T>
__global__ void k_normalize(const T *__restrict__ w, T *__restrict__ v, uint64_t n, const SolveState *st,
                            const uint32_t *stop)
{
    if (*stop) return;
    double h = st->t0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        v[i] = (h == 0.0) ? (T)0 : (T)((double)w[i] / h);
}

// v0 = ro / beta (0 if beta == 0)
template <typename T>
__global__ void k_scale_by(const double *__restrict__ a, T *__restrict__ out, uint64_t n, double beta)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        out[i] = (beta == 0.0) ? (T)0 : (T)(a[i] / beta);
}

// x += sum_{j<cnt} Z_j y_j   (src/rmatrixsolver.cpp:571-578), cnt = the inner cycle's length
template <typename T>
__global__ void k_update_x(double *__restrict__ x, const T *__restrict__ Z, uint64_t ld,
                           uint32_t n, const uint32_t *__restrict__ cnt_dev, const double *__restrict__ y)
{
    const uint32_t cnt = *cnt_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        double xi = x[i];
        for (uint32_t j = 0; j < cnt; j++) xi += (double)Z[j * ld + i] * y[j];
        x[i] = xi;
    }
}

__global__ void k_zero_u32(uint32_t *__restrict__ a, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0u;
}

// GMRES inner state (device, one outer cycle): H ((k+1) x k, row-major), the rotations c, s,
// the rotated right-hand side pp and y (k+1 each), and beta = |r_0|; flag[0] = stop,
// flag[1] = the cycle's length.
struct GmView {
    double *H, *c, *s, *pp, *y;
    uint32_t k;
    __host__ __device__ GmView(double *base, uint32_t k_) : k(k_)
    {
        H = base;
        c = H + (uint64_t)(k + 1) * k;
        s = c + k + 1;
        pp = s + k + 1;
        y = pp + k + 1;
    }
    __host__ __device__ double &h(uint32_t i, uint32_t j) const { return H[(uint64_t)i * k + j]; }
    __host__ __device__ double &beta() const { return y[k + 1]; }
    static uint64_t doubles(uint32_t k) { return (uint64_t)(k + 1) * k + 4ull * (k + 1) + 1; }
};

constexpr uint32_t kGmLds = 2048;     // inner cycles up to this length keep their vectors in LDS
constexpr uint32_t kGmBlock = 256;

// Column it of H from the Gram-Schmidt coefficients and |w|, the previous rotations applied,
// the new rotation, and the stopping test of src/rmatrixsolver.cpp:521-550.  The column and the
// rotations are staged in LDS by the workgroup; one thread runs the rotation chain with the
// running entry in a register.  The host loop's expressions without contraction, so the
// rotations agree with it bit for bit.
__global__ __launch_bounds__(kGmBlock) void k_gm_givens(double *__restrict__ base, uint32_t k, uint32_t *__restrict__ flag,
                                                        const double *__restrict__ hcol,
                                                        const SolveState *__restrict__ st, uint32_t it)
{
#pragma clang fp contract(off)
    if (flag[0]) return;
    __shared__ double col[kGmLds], cs[kGmLds], sn[kGmLds];
    GmView g(base, k);
    const bool lds = it < kGmLds;
    const double *cc = lds ? cs : g.c, *ss = lds ? sn : g.s;
    const uint64_t cstride = lds ? 1 : k;   // column it of H: stride k in place
    double *cvi = lds ? col : g.H + it;
    if (lds) {
        for (uint32_t i = threadIdx.x; i <= it; i += kGmBlock) {
            col[i] = hcol[i];
            if (i < it) {
                cs[i] = g.c[i];
                sn[i] = g.s[i];
            }
        }
    } else {
        for (uint32_t i = threadIdx.x; i <= it; i += kGmBlock) g.h(i, it) = hcol[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double prev = cvi[0];
        uint32_t i = 1;
        if (lds)   // eight steps per trip, their operands loaded ahead of the chain
            for (; i + 7 <= it; i += 8) {
                double cur[8], c[8], s[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    cur[u] = col[i + u];
                    c[u] = cs[i - 1 + u];
                    s[u] = sn[i - 1 + u];
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    col[i - 1 + u] = c[u] * prev + s[u] * cur[u];
                    prev = -s[u] * prev + c[u] * cur[u];
                }
            }
        for (; i <= it; i++) {
            const double cur = cvi[i * cstride], c = cc[i - 1], s = ss[i - 1];
            cvi[(i - 1) * cstride] = c * prev + s * cur;
            prev = -s * prev + c * cur;
        }
        const double sub = st->t0;   // h(it + 1, it)
        const double gg = __dsqrt_rn(prev * prev + sub * sub);
        double cn = 0.0, sg = 0.0;
        if (gg != 0.0) {
            cn = prev / gg;
            sg = sub / gg;
        }
        g.c[it] = cn;
        g.s[it] = sg;
        cvi[it * cstride] = gg;
        g.h(it + 1, it) = 0.0;
        g.pp[it + 1] = -sg * g.pp[it];
        g.pp[it] = cn * g.pp[it];
        if (fabs(g.pp[it + 1]) < g.beta() * 1.0e-4 || it + 1 == k) {
            flag[0] = 1u;
            flag[1] = it + 1;
        }
    }
    __syncthreads();
    if (lds)
        for (uint32_t i = threadIdx.x; i <= it; i += kGmBlock) g.h(i, it) = col[i];
}

// Back substitution of the cycle (src/rmatrixsolver.cpp:557-566) by columns: y_i from pp_i,
// then every pp_j (j < i) less y_i h(j, i) in parallel -- for each pp_j the same subtractions in
// the same order (i = m-1 down to j+1) as the row loop, so the same bits.
__global__ __launch_bounds__(kGmBlock) void k_gm_backsub(double *__restrict__ base, uint32_t k,
                                                         const uint32_t *__restrict__ flag)
{
#pragma clang fp contract(off)
    __shared__ double lpp[kGmLds], ly[kGmLds];
    GmView g(base, k);
    const int m = (int)flag[1];
    if (m == 0) return;
    const bool lds = m <= (int)kGmLds;
    double *pp = lds ? lpp : g.pp, *y = lds ? ly : g.y;
    if (lds)
        for (int j = threadIdx.x; j < m; j += kGmBlock) lpp[j] = g.pp[j];
    __syncthreads();
    for (int i = m - 1; i >= 0; i--) {
        if (threadIdx.x == 0) {
            const double d = g.h(i, i);
            y[i] = d == 0.0 ? 0.0 : pp[i] / d;
        }
        __syncthreads();
        const double yi = y[i];
        for (int j = threadIdx.x; j < i; j += kGmBlock) pp[j] -= yi * g.h(j, i);
        __syncthreads();
    }
    if (lds)
        for (int j = threadIdx.x; j < m; j += kGmBlock) g.y[j] = ly[j];
}

// ============================================================================ launch helpers
static uint32_t round_up(uint32_t a, uint32_t m) { return (a + m - 1) / m * m; }

static uint32_t target_blocks() { return (uint32_t)std::max(8, runtime().n_cu * 8); }

static uint32_t vec_blocks(uint64_t items)
{
    uint64_t want = (items + kBlock - 1) / kBlock;
    return round_up((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, target_blocks())), kXcd);
}

static const int kVecLanes[5] = { 4, 8, 16, 32, 64 };
static const uint32_t kTileCap[2] = { 2048, 4096 };
constexpr int kSellPlan = 7;      // SellRows; the shapes after it read the same slices
static_assert(kNumShapes == 11, "plan_ms holds 11 shapes");
static_assert(RS_SPMV_SHAPES == kNumShapes, "the header's shape count matches the library's");

static bool is_sell_plan(int id) { return id >= kSellPlan; }

// Workgroups of plan `id` for matrix A, and the view its kernels read.
static CsrView plan_view(const Matrix &A, int id, uint32_t &nb)
{
    CsrView v{};
    v.n = A.n;
    v.rowptr = A.rowptr.p;
    v.col = A.col.p;
    v.val = A.val.p;
    v.val32 = A.val32.p;
    if (id < 5) {
        uint32_t rpp = kBlock / kVecLanes[id];
        uint32_t passes = (A.n + rpp - 1) / rpp;
        nb = round_up(std::max(1u, std::min(target_blocks(), passes)), kXcd);
        v.chunk = round_up((A.n + nb - 1) / nb, rpp);
        if (v.chunk == 0) v.chunk = rpp;
    } else if (id < 7) {
        int t = id - 5;
        v.tiles = A.tiles[t].p;
        v.ntiles = A.ntiles[t];
        nb = round_up(std::max(1u, std::min(target_blocks(), v.ntiles)), kXcd);
        v.tpb = (v.ntiles + nb - 1) / nb;
    } else {
        v.sell_ptr = A.sell_ptr.p;
        v.sell_col = A.sell_col.p;
        v.sell_off = A.sell_off_ok ? A.sell_off.p : nullptr;
        v.sell_val = A.sell_val.p;
        v.sell_val32 = A.sell_val32.p;
        v.nslices = A.nslices;
        uint32_t waves = kBlock / kWave;
        nb = round_up(std::max(1u, std::min(target_blocks(), (A.nslices + waves - 1) / waves)), kXcd);
        v.spb = (A.nslices + nb - 1) / nb;
    }
    return v;
}

// Plan before tuning: vector rows with lanes ~ row length.
static int default_plan(const Matrix &A)
{
    double avg = A.n ? (double)A.nnz / A.n : 1.0;
    int g = 0;
    while (g < 4 && kVecLanes[g] * 1.5 < avg) g++;
    return g;
}

#define RS_DISPATCH_W(w, KERNEL, grid, ...)                                                   \
    do {                                                                                      \
        hipStream_t _s = runtime().stream;                                                    \
        if ((w) == 1) KERNEL<1><<<grid, kBlock, 0, _s>>>(__VA_ARGS__);                        \
        else if ((w) == 3) KERNEL<3><<<grid, kBlock, 0, _s>>>(__VA_ARGS__);                   \
        else KERNEL<0><<<grid, kBlock, 0, _s>>>(__VA_ARGS__);                                 \
    } while (0)

static void fin(const double *partials, uint32_t nparts, int nv, int op, int flag, SolveState *st, double *hist)
{
    k_fin<<<1, kFinBlock, 0, runtime().stream>>>(partials, nparts, nv, op, flag, st, hist);
}

// Workspace of standalone reductions (matrix norm).
struct NormWork {
    DBuf<double> partials;
    DBuf<SolveState> st;
};
static NormWork &norm_work()
{
    static NormWork s;
    return s;
}
static int ensure_norm_work()
{
    NormWork &s = norm_work();
    int rc;
    if (!s.partials.p && (rc = s.partials.alloc(8 * 4096))) return rc;
    if (!s.st.p && (rc = s.st.alloc(1))) return rc;
    return RS_OK;
}

int matrix_frobenius(Matrix &A)
{
    int rc = ensure_norm_work();
    if (rc) return rc;
    NormWork &s = norm_work();
    uint32_t nb = vec_blocks(A.nnz);
    hipStream_t st = runtime().stream;
    k_norm2<<<nb, kBlock, 0, st>>>(A.val.p, A.nnz, s.partials.p);
    fin(s.partials.p, nb, 1, FIN_NORM0, 0, s.st.p, nullptr);
    RS_HIP(hipGetLastError());
    SolveState h;
    RS_HIP(hipMemcpyAsync(&h, s.st.p, sizeof(h), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    A.frob = h.t0;
    return RS_OK;
}

int matrix_refresh_fp32(Matrix &A)
{
    if (A.val32_valid) return RS_OK;
    int rc = A.val32.alloc(A.nnz);
    if (rc) return rc;
    if (A.nnz) k_to_f32<<<vec_blocks(A.nnz), kBlock, 0, runtime().stream>>>(A.val.p, A.val32.p, A.nnz);
    RS_HIP(hipGetLastError());
    A.val32_valid = true;
    return RS_OK;
}

int matrix_set_layout(Matrix &A, const uint32_t *host_rowptr)
{
    // SELL-64 slice offsets (host), columns (device)
    A.nslices = (A.n + kWave - 1) / kWave;
    std::vector<uint32_t> sp(A.nslices + 1ull, 0);
    for (uint32_t s = 0; s < A.nslices; s++) {
        uint32_t w = 0;
        for (uint32_t r = s * kWave; r < std::min<uint64_t>(A.n, (uint64_t)(s + 1) * kWave); r++)
            w = std::max(w, host_rowptr[r + 1] - host_rowptr[r]);
        uint64_t next = (uint64_t)sp[s] + (uint64_t)w * kWave;
        if (next > 0xFFFFFFF0ull) return fail(RS_ERR_UNSUPPORTED, "SELL-64 copy exceeds 2^32 entries");
        sp[s + 1] = (uint32_t)next;
    }
    int rc;
    if ((rc = A.sell_ptr.alloc(sp.size())) || (rc = A.sell_col.alloc(std::max<uint32_t>(sp.back(), 1)))) return rc;
    RS_HIP(hipMemcpyAsync(A.sell_ptr.p, sp.data(), sizeof(uint32_t) * sp.size(), hipMemcpyHostToDevice, runtime().stream));
    {
        uint32_t nb = 0;
        CsrView v = plan_view(A, kSellPlan, nb);
        if (A.nslices) k_sell_cols<<<vec_blocks((uint64_t)A.nslices * kWave), kBlock, 0, runtime().stream>>>(v, A.sell_col.p);
        RS_HIP(hipGetLastError());
        // int16 offsets for SellRowsOff when every column fits
        A.sell_off_ok = false;
        DBuf<uint32_t> far;
        if ((rc = A.sell_off.alloc(std::max<uint32_t>(sp.back(), 1))) || (rc = far.alloc(1))) return rc;
        RS_HIP(hipMemsetAsync(far.p, 0, sizeof(uint32_t), runtime().stream));
        if (A.nslices)
            k_sell_offs<<<vec_blocks((uint64_t)A.nslices * kWave), kBlock, 0, runtime().stream>>>(v, A.sell_col.p,
                                                                                              A.sell_off.p, far.p);
        RS_HIP(hipGetLastError());
        uint32_t nfar = 0;
        RS_HIP(hipMemcpyAsync(&nfar, far.p, sizeof(uint32_t), hipMemcpyDeviceToHost, runtime().stream));
        RS_HIP(hipStreamSynchronize(runtime().stream));
        A.sell_off_ok = nfar == 0;
    }
    A.sell_valid = A.sell32_valid = false;
    for (int t = 0; t < 2; t++) {
        std::vector<uint32_t> tiles = build_tiles(host_rowptr, A.n, kTileCap[t], kTileRows);
        int rc = A.tiles[t].alloc(tiles.size());
        if (rc) return rc;
        A.ntiles[t] = (uint32_t)tiles.size() - 1;
        RS_HIP(hipMemcpyAsync(A.tiles[t].p, tiles.data(), sizeof(uint32_t) * tiles.size(), hipMemcpyHostToDevice,
                              runtime().stream));
    }
    RS_HIP(hipStreamSynchronize(runtime().stream));
    A.plan[0] = A.plan[1] = -1;
    return RS_OK;
}

// Bring the copies a plan reads up to date with the FP64 CSR values.
int matrix_sync_plan(Matrix &A, int plan, bool fp32)
{
    int rc;
    if (fp32 && !is_sell_plan(plan) && (rc = matrix_refresh_fp32(A))) return rc;
    if (!is_sell_plan(plan)) return RS_OK;
    hipStream_t st = runtime().stream;
    uint32_t nb = 0;
    CsrView v = plan_view(A, kSellPlan, nb);
    uint64_t total = A.sell_col.n;
    if (!A.sell_valid) {
        if ((rc = A.sell_val.alloc(total))) return rc;
        v.sell_val = A.sell_val.p;
        if (A.nslices) k_sell_vals<<<vec_blocks((uint64_t)A.nslices * kWave), kBlock, 0, st>>>(v, A.sell_val.p);
        RS_HIP(hipGetLastError());
        A.sell_valid = true;
        A.sell32_valid = false;
    }
    if (fp32 && !A.sell32_valid) {
        if ((rc = A.sell_val32.alloc(total))) return rc;
        if (A.nslices)
            k_sell_to_f32<<<vec_blocks((uint64_t)A.nslices * kWave), kBlock, 0, st>>>(v, A.sell_val.p, A.sell_val32.p);
        RS_HIP(hipGetLastError());
        A.sell32_valid = true;
    }
    return RS_OK;
}

// Average device time of `reps` launches of SpMV plan `id`.
static int time_plan(Matrix &A, int id, bool fp32, uint32_t reps, const void *x, void *y, double &ms)
{
    int rc = matrix_sync_plan(A, id, fp32);
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    uint32_t nb = 0;
    CsrView v = plan_view(A, id, nb);
    auto launch = [&]() {
        if (fp32) RS_SHAPE_SWITCH(id, k_spmv32, nb, st, v, (const float *)x, (float *)y);
        else RS_SHAPE_SWITCH(id, k_spmv, nb, st, v, (const double *)x, (double *)y);
    };
    launch();
    EventPair e0e1_;
    RS_HIP(e0e1_.create());
    hipEvent_t e0 = e0e1_.e[0], e1 = e0e1_.e[1];
    RS_HIP(hipEventRecord(e0, st));
    for (uint32_t i = 0; i < reps; i++) launch();
    RS_HIP(hipEventRecord(e1, st));
    RS_HIP(hipEventSynchronize(e1));
    RS_HIP(hipGetLastError());
    float f = 0.0f;
    RS_HIP(hipEventElapsedTime(&f, e0, e1));
    ms = f / reps;
    return RS_OK;
}

// The SpMV autotuner: times every shape once per matrix and precision and keeps the fastest
// (the pattern of a matrix handle never changes and values do not affect the choice).
int matrix_tune(Matrix &A, bool fp32)
{
    int slot = fp32 ? 1 : 0;
    if (A.plan[slot] >= 0) return RS_OK;
    if (A.n == 0 || A.tiles[0].p == nullptr) {
        A.plan[slot] = default_plan(A);
        return RS_OK;
    }
    int rc;
    if (fp32 && (rc = matrix_refresh_fp32(A))) return rc;
    DBuf<double> x, y;
    if ((rc = x.alloc(A.n)) || (rc = y.alloc(A.n))) return rc;
    k_fill<<<vec_blocks(A.n), kBlock, 0, runtime().stream>>>(x.p, A.n, 1.0);
    if (fp32) k_fill32<<<vec_blocks(A.n), kBlock, 0, runtime().stream>>>((float *)x.p, A.n, 1.0f);
    uint32_t reps = A.nnz > 4000000 ? 4 : 10;
    double best = 1e30;
    int best_id = default_plan(A);
    for (int id = 0; id < kNumShapes; id++) {
        double ms = 0.0;
        if ((rc = time_plan(A, id, fp32, reps, x.p, y.p, ms))) return rc;
        A.plan_ms[slot][id] = ms;
        if (ms < best) {
            best = ms;
            best_id = id;
        }
    }
    A.plan[slot] = best_id;
    return RS_OK;
}

}  // namespace rs

using namespace rs;

// ============================================================================ C ABI: matrix
extern "C" int rs_matrix_create(uint32_t n_rows, const uint32_t *row_ptr, const uint32_t *col,
                                const double *val, rs_matrix **out)
{
    if (!out || !row_ptr || (n_rows && (!col || !val) && row_ptr[n_rows] > 0))
        return fail(RS_ERR_INVALID_ARG, "rs_matrix_create: null argument");
    if (row_ptr[0] != 0) return fail(RS_ERR_INVALID_ARG, "rs_matrix_create: row_ptr[0] != 0");
    for (uint32_t i = 0; i < n_rows; i++)
        if (row_ptr[i + 1] < row_ptr[i]) return fail(RS_ERR_INVALID_ARG, "rs_matrix_create: row_ptr not monotone");
    uint64_t nnz = row_ptr[n_rows];
    for (uint64_t j = 0; j < nnz; j++)
        if (col[j] >= n_rows) return fail(RS_ERR_INVALID_ARG, "rs_matrix_create: column index out of range");
    int rc = require_runtime();
    if (rc) return rc;
    std::unique_ptr<rs_matrix> A(new rs_matrix());
    A->n = n_rows;
    A->nnz = nnz;
    if ((rc = A->rowptr.alloc(n_rows + 1ull))) return rc;
    if ((rc = A->col.alloc(nnz))) return rc;
    if ((rc = A->val.alloc(nnz))) return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(A->rowptr.p, row_ptr, sizeof(uint32_t) * (n_rows + 1ull), hipMemcpyHostToDevice, st));
    if (nnz) {
        RS_HIP(hipMemcpyAsync(A->col.p, col, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice, st));
        RS_HIP(hipMemcpyAsync(A->val.p, val, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
        if ((rc = matrix_frobenius(*A))) return rc;
    }
    if ((rc = matrix_set_layout(*A, row_ptr))) return rc;
    RS_HIP(hipStreamSynchronize(st));
    *out = A.release();
    return RS_OK;
}

namespace rs {

int matrix_create_device(uint32_t n, DBuf<uint32_t> &rowptr, DBuf<uint32_t> &col, uint64_t nnz, rs_matrix **out)
{
    if (!out || rowptr.n != n + 1ull || (nnz && col.n < nnz)) return fail(RS_ERR_INVALID_ARG, "matrix_create_device: bad pattern");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    std::unique_ptr<rs_matrix> A(new rs_matrix());
    A->n = n;
    A->nnz = nnz;
    std::swap(A->rowptr.p, rowptr.p);
    std::swap(A->rowptr.n, rowptr.n);
    std::swap(A->col.p, col.p);
    std::swap(A->col.n, col.n);
    if ((rc = A->val.alloc(nnz))) return rc;
    if (nnz) RS_HIP(hipMemsetAsync(A->val.p, 0, sizeof(double) * nnz, st));
    A->frob = 0.0;
    std::vector<uint32_t> hrp(n + 1ull);
    RS_HIP(hipMemcpyAsync(hrp.data(), A->rowptr.p, sizeof(uint32_t) * hrp.size(), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    if (hrp[0] != 0 || hrp[n] != nnz) return fail(RS_ERR_INVALID_ARG, "matrix_create_device: row pointers do not span the entries");
    if ((rc = matrix_set_layout(*A, hrp.data()))) return rc;
    RS_HIP(hipStreamSynchronize(st));
    *out = A.release();
    return RS_OK;
}

}  // namespace rs

extern "C" int rs_matrix_update_values(rs_matrix *A, const double *val)
{
    if (!A || (!val && A->nnz)) return fail(RS_ERR_INVALID_ARG, "rs_matrix_update_values: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    if (A->nnz) {
        RS_HIP(hipMemcpyAsync(A->val.p, val, sizeof(double) * A->nnz, hipMemcpyHostToDevice, runtime().stream));
        if ((rc = matrix_frobenius(*A))) return rc;
    }
    matrix_values_changed(*A);
    return RS_OK;
}

extern "C" int rs_matrix_update_values_device(rs_matrix *A, const double *d_val)
{
    if (!A || (!d_val && A->nnz)) return fail(RS_ERR_INVALID_ARG, "rs_matrix_update_values_device: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    if (A->nnz) {
        if (d_val != A->val.p)
            RS_HIP(hipMemcpyAsync(A->val.p, d_val, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice, runtime().stream));
        if ((rc = matrix_frobenius(*A))) return rc;
    }
    matrix_values_changed(*A);
    return RS_OK;
}

extern "C" int rs_matrix_info(const rs_matrix *A, uint32_t *n_rows, uint64_t *nnz, double *frobenius)
{
    if (!A) return fail(RS_ERR_INVALID_ARG, "rs_matrix_info: null matrix");
    if (n_rows) *n_rows = A->n;
    if (nnz) *nnz = A->nnz;
    if (frobenius) *frobenius = A->frob;
    return RS_OK;
}

extern "C" int rs_matrix_download(const rs_matrix *A, uint32_t *row_ptr, uint32_t *col, double *val)
{
    if (!A) return fail(RS_ERR_INVALID_ARG, "rs_matrix_download: null matrix");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    if (row_ptr) RS_HIP(hipMemcpyAsync(row_ptr, A->rowptr.p, sizeof(uint32_t) * (A->n + 1ull), hipMemcpyDeviceToHost, st));
    if (col && A->nnz) RS_HIP(hipMemcpyAsync(col, A->col.p, sizeof(uint32_t) * A->nnz, hipMemcpyDeviceToHost, st));
    if (val && A->nnz) RS_HIP(hipMemcpyAsync(val, A->val.p, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

extern "C" int rs_matrix_spmv_device(rs_matrix *A, const double *d_x, double *d_y)
{
    if (!A || !d_x || !d_y) return fail(RS_ERR_INVALID_ARG, "rs_matrix_spmv_device: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    if (A->n == 0) return RS_OK;
    if ((rc = matrix_tune(*A, false)) || (rc = matrix_sync_plan(*A, A->plan[0], false))) return rc;
    uint32_t nb = 0;
    CsrView v = plan_view(*A, A->plan[0], nb);
    RS_SHAPE_SWITCH(A->plan[0], k_spmv, nb, runtime().stream, v, d_x, d_y);
    RS_HIP(hipGetLastError());
    return RS_OK;
}

extern "C" int rs_matrix_spmv(rs_matrix *A, const double *x, double *y)
{
    if (!A || !x || !y) return fail(RS_ERR_INVALID_ARG, "rs_matrix_spmv: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    DBuf<double> dx, dy;
    if ((rc = dx.alloc(A->n)) || (rc = dy.alloc(A->n))) return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * A->n, hipMemcpyHostToDevice, st));
    if ((rc = rs_matrix_spmv_device(A, dx.p, dy.p))) return rc;
    RS_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * A->n, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

extern "C" int rs_matrix_spmv_timed(rs_matrix *A, int32_t precision, uint32_t reps, int32_t plan, double *avg_ms)
{
    if (!A || !avg_ms || reps == 0 || plan >= kNumShapes) return fail(RS_ERR_INVALID_ARG, "rs_matrix_spmv_timed: bad argument");
    int rc = require_runtime();
    if (rc) return rc;
    if (A->n == 0) {
        *avg_ms = 0.0;
        return RS_OK;
    }
    const bool fused = precision == RS_TIMED_CG_PRODUCT;
    bool fp32 = precision == RS_PRECISION_MIXED;
    if ((rc = matrix_tune(*A, fp32))) return rc;
    int id = plan < 0 ? A->plan[fp32 ? 1 : 0] : plan;
    DBuf<double> x, y;
    if ((rc = x.alloc(A->n)) || (rc = y.alloc(A->n))) return rc;
    hipStream_t st = runtime().stream;
    k_fill<<<vec_blocks(A->n), kBlock, 0, st>>>(x.p, A->n, 1.0);
    if (fp32) k_fill32<<<vec_blocks(A->n), kBlock, 0, st>>>((float *)x.p, A->n, 1.0f);
    if (!fused) return time_plan(*A, id, fp32, reps, x.p, y.p, *avg_ms);
    // the FP64 CG iteration's product as the solver launches it: q = A p with the per-workgroup
    // partial sums of p.q, on a running solver state
    DBuf<double> partials;
    DBuf<SolveState> state;
    if ((rc = partials.alloc(kPartialBytes / sizeof(double))) || (rc = state.alloc(1)) ||
        (rc = matrix_sync_plan(*A, id, false)))
        return rc;
    k_state_init<<<1, 64, 0, st>>>(state.p, A->frob, 0.0, 0.0, 1, RS_STOP_RELRES, 0, 0);
    uint32_t nb = 0;
    CgArgs a{};
    a.A = plan_view(*A, id, nb);
    a.p = x.p;
    a.q = y.p;
    a.partials = partials.p;
    a.arrive = reinterpret_cast<uint32_t *>(partials.p + kPartialDoubles);
    a.st = state.p;
    k_zero_u32<<<1, 64, 0, st>>>(a.arrive, kArriveWords);
    RS_SHAPE_SWITCH(id, k_cg_spmv_dot, nb, st, a);
    EventPair ev;
    RS_HIP(ev.create());
    RS_HIP(hipEventRecord(ev.e[0], st));
    for (uint32_t i = 0; i < reps; i++) RS_SHAPE_SWITCH(id, k_cg_spmv_dot, nb, st, a);
    RS_HIP(hipEventRecord(ev.e[1], st));
    RS_HIP(hipEventSynchronize(ev.e[1]));
    RS_HIP(hipGetLastError());
    float f = 0.0f;
    RS_HIP(hipEventElapsedTime(&f, ev.e[0], ev.e[1]));
    *avg_ms = f / reps;
    return RS_OK;
}

extern "C" int rs_matrix_spmv_plan(rs_matrix *A, int32_t precision, int32_t *plan, double *plan_ms)
{
    if (!A || !plan) return fail(RS_ERR_INVALID_ARG, "rs_matrix_spmv_plan: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    bool fp32 = precision == RS_PRECISION_MIXED;
    if ((rc = matrix_tune(*A, fp32))) return rc;
    *plan = A->plan[fp32 ? 1 : 0];
    if (plan_ms)
        for (int i = 0; i < kNumShapes; i++) plan_ms[i] = A->plan_ms[fp32 ? 1 : 0][i];
    return RS_OK;
}

extern "C" int rs_matrix_force_plan(rs_matrix *A, int32_t precision, int32_t plan)
{
    if (!A || plan < -1 || plan >= kNumShapes) return fail(RS_ERR_INVALID_ARG, "rs_matrix_force_plan: bad argument");
    A->plan[precision == RS_PRECISION_MIXED ? 1 : 0] = plan;
    return RS_OK;
}

extern "C" int rs_matrix_destroy(rs_matrix *A)
{
    if (!A) return RS_OK;
    if (A->borrowed) return fail(RS_ERR_INVALID_ARG, "rs_matrix_destroy: matrix is owned by an assembly");
    delete A;
    return RS_OK;
}

// ============================================================================ C ABI: precond
struct rs_precond {
    int32_t type = RS_PRECOND_NONE;
    uint32_t n = 0, width = 0;
    DBuf<double> data;
    DBuf<float> data32;
    bool data32_valid = false;
};

namespace rs {
// Block width the reference uses for block size w: decremented until it divides n
// (src/rmatrixpreconditioner.cpp:98-105).
static uint32_t block_width(uint32_t n, uint32_t w)
{
    while (w > 1 && n % w != 0) w--;
    return w;
}

// Launch the Jacobi / block-Jacobi build of width w into `data` (n or n*w doubles).
static int precond_build_at(const Matrix &A, int32_t type, uint32_t w, double *data)
{
    if (A.n == 0 || type == RS_PRECOND_NONE) return RS_OK;
    uint32_t nb_unused = 0;
    CsrView v = plan_view(A, 0, nb_unused);
    hipStream_t st = runtime().stream;
    if (type == RS_PRECOND_JACOBI) k_jacobi_build<<<vec_blocks(A.n), kBlock, 0, st>>>(v, data);
    else if (w == 3) k_blockjacobi_build<3><<<vec_blocks(A.n / w), kBlock, 0, st>>>(v, w, data);
    else if (w <= 4) k_blockjacobi_build<4><<<vec_blocks(A.n / w), kBlock, 0, st>>>(v, w, data);
    else k_blockjacobi_build<16><<<vec_blocks(A.n / w), kBlock, 0, st>>>(v, w, data);
    RS_HIP(hipGetLastError());
    return RS_OK;
}

static int precond_build(const Matrix &A, int32_t type, uint32_t block_size, rs_precond &P)
{
    P.type = type;
    P.n = A.n;
    P.data32_valid = false;
    int rc;
    if (type == RS_PRECOND_NONE) {
        P.width = 1;
        return RS_OK;
    }
    if (type == RS_PRECOND_JACOBI) {
        P.width = 1;
        if ((rc = P.data.alloc(std::max(A.n, 1u)))) return rc;
        return precond_build_at(A, type, 1, P.data.p);
    }
    if (type == RS_PRECOND_BLOCK_JACOBI) {
        if (block_size == 0) return fail(RS_ERR_INVALID_ARG, "block-Jacobi: block size 0");
        if (block_size > 16) return fail(RS_ERR_UNSUPPORTED, "block-Jacobi: block size above 16");
        const uint32_t w = block_width(A.n, block_size);
        P.width = w;
        if ((rc = P.data.alloc(std::max<uint64_t>((uint64_t)A.n * w, 1)))) return rc;
        return precond_build_at(A, type, w, P.data.p);
    }
    return fail(RS_ERR_INVALID_PRECOND, "Invalid matrix preconditioner type '" + std::to_string(type) + "'");
}

static int precond_fp32(rs_precond &P)
{
    if (P.type == RS_PRECOND_NONE || P.data32_valid) return RS_OK;
    int rc = P.data32.alloc(P.data.n);
    if (rc) return rc;
    k_to_f32<<<vec_blocks(P.data.n), kBlock, 0, runtime().stream>>>(P.data.p, P.data32.p, P.data.n);
    RS_HIP(hipGetLastError());
    P.data32_valid = true;
    return RS_OK;
}
}  // namespace rs

extern "C" int rs_precond_create(const rs_matrix *A, int32_t type, uint32_t block_size, rs_precond **out)
{
    if (!A || !out) return fail(RS_ERR_INVALID_ARG, "rs_precond_create: null argument");
    if (type < RS_PRECOND_NONE || type > RS_PRECOND_BLOCK_JACOBI)
        return fail(RS_ERR_INVALID_PRECOND, "Invalid matrix preconditioner type '" + std::to_string(type) + "'");
    int rc = require_runtime();
    if (rc) return rc;
    std::unique_ptr<rs_precond> P(new rs_precond());
    if ((rc = precond_build(*A, type, block_size, *P))) return rc;
    RS_HIP(hipStreamSynchronize(runtime().stream));
    *out = P.release();
    return RS_OK;
}

extern "C" int rs_precond_width(const rs_precond *P, uint32_t *width)
{
    if (!P || !width) return fail(RS_ERR_INVALID_ARG, "rs_precond_width: null argument");
    *width = P->width;
    return RS_OK;
}

extern "C" int rs_precond_download(const rs_precond *P, double *data)
{
    if (!P || !data) return fail(RS_ERR_INVALID_ARG, "rs_precond_download: null argument");
    if (P->type == RS_PRECOND_NONE) return RS_OK;
    int rc = require_runtime();
    if (rc) return rc;
    RS_HIP(hipMemcpyAsync(data, P->data.p, sizeof(double) * (uint64_t)P->n * P->width, hipMemcpyDeviceToHost, runtime().stream));
    RS_HIP(hipStreamSynchronize(runtime().stream));
    return RS_OK;
}

extern "C" int rs_precond_compute(const rs_precond *P, const double *x, double *y)
{
    if (!P || !x || !y) return fail(RS_ERR_INVALID_ARG, "rs_precond_compute: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    if (P->n == 0) return RS_OK;
    DBuf<double> dx, dy;
    if ((rc = dx.alloc(P->n)) || (rc = dy.alloc(P->n))) return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * P->n, hipMemcpyHostToDevice, st));
    uint32_t w = P->width ? P->width : 1;
    RS_DISPATCH_W(w, k_pc_apply, vec_blocks(P->n / w), P->type, w, P->data.p, P->n, dx.p, dy.p, nullptr);
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * P->n, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

extern "C" int rs_precond_copy(const rs_precond *src, rs_precond **out)
{
    if (!src || !out) return fail(RS_ERR_INVALID_ARG, "rs_precond_copy: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    std::unique_ptr<rs_precond> P(new rs_precond());
    P->type = src->type;
    P->n = src->n;
    P->width = src->width;
    hipStream_t st = runtime().stream;
    if (src->data.n) {
        if ((rc = P->data.alloc(src->data.n))) return rc;
        RS_HIP(hipMemcpyAsync(P->data.p, src->data.p, sizeof(double) * src->data.n, hipMemcpyDeviceToDevice, st));
    }
    if (src->data32_valid && src->data32.n) {
        if ((rc = P->data32.alloc(src->data32.n))) return rc;
        RS_HIP(hipMemcpyAsync(P->data32.p, src->data32.p, sizeof(float) * src->data32.n, hipMemcpyDeviceToDevice, st));
        P->data32_valid = true;
    }
    RS_HIP(hipStreamSynchronize(st));
    *out = P.release();
    return RS_OK;
}

extern "C" int rs_precond_destroy(rs_precond *P)
{
    delete P;
    return RS_OK;
}

// ============================================================================ C ABI: solver
struct rs_solver {
    rs_conf conf{};
    std::string out_file;
    double error = 0.0;     // RIterationInfo::error, carried across solves of this object
    double trend = 0.0;
    uint32_t last_iterations = 0;
    std::vector<double> history;
    // device workspace
    DBuf<double> y, r, p, q, ro, w, V, Z, hbuf, gm;
    DBuf<uint32_t> gmflag;      // GMRES inner cycle: stop, length
    DBuf<float> r32, p32, q32, dx32, z32;
    DBuf<float> V32, Z32, w32;  // mixed-precision GMRES: FP32 basis, preconditioned basis, w
    DBuf<double> partials, hist;
    DBuf<SolveState> st;
    // captured iteration graph (or, with RS_NO_GRAPH=1 in the environment, the same S
    // iterations launched one kernel at a time)
    hipGraphExec_t graph = nullptr;
    std::string graph_key;   // raw bytes of every kernel argument and launch shape captured
    bool eager = false;
    // replay loop: pinned copies of the stop flags and their events, kept for the solver's life
    void *hflags = nullptr;
    hipEvent_t ev[2] = { nullptr, nullptr };
    ~rs_solver()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (hipEvent_t &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (hflags) (void)hipHostFree(hflags);
    }
};

namespace rs {

static constexpr uint32_t kGraphIters = 8;
static constexpr uint32_t kHistCap = 1u << 20;

static int ensure_ws(rs_solver &S, uint32_t n, bool mixed, bool gmres)
{
    int rc;
    size_t N = std::max(n, 1u);
    if ((rc = S.y.alloc(N)) || (rc = S.partials.alloc(kPartialBytes / sizeof(double))) || (rc = S.st.alloc(1)))
        return rc;
    uint32_t hc = std::min(kHistCap, std::max(1u, S.conf.n_outer_iterations));
    if ((rc = S.hist.alloc(2ull * hc))) return rc;
    if (!mixed && !gmres) {
        if ((rc = S.r.alloc(N)) || (rc = S.p.alloc(N)) || (rc = S.q.alloc(N))) return rc;
    }
    if (mixed && !gmres) {
        if ((rc = S.r32.alloc(N)) || (rc = S.p32.alloc(N)) || (rc = S.q32.alloc(N)) || (rc = S.dx32.alloc(N))) return rc;
    }
    if (gmres) {
        // the mixed-precision GMRES keeps its basis in FP32 buffers of its own (solve_gmres)
        uint64_t k = S.conf.n_inner_iterations;
        if ((rc = S.ro.alloc(N)) || (rc = S.hbuf.alloc(2 * (k + 2)))) return rc;
        const uint64_t ld = (N + 1) / 2 * 2;   // basis rows at multiples of 16 bytes
        if (!mixed && ((rc = S.w.alloc(N)) || (rc = S.V.alloc((k + 1) * ld)) || (rc = S.Z.alloc((k + 1) * ld)))) return rc;
    }
    return RS_OK;
}

// RIterationInfo::printIteration / writeToFile (src/riterationinfo.cpp:133-193), replayed
// on the host from the device history after the solve.
static void emit_log(const rs_solver &S, const char *title, uint32_t iters, uint32_t hist_n)
{
    const rs_conf &c = S.conf;
    if (c.output_frequency == 0) return;
    std::printf("-------------------------------------------------------------------------------\n");
    std::printf("%s\n", title);
    std::printf("-------------------------------------------------------------------------------\n");
    std::printf(">    Iteration     |   Residual   |    Trend     |                             \n");
    std::printf("-------------------------------------------------------------------------------\n");
    FILE *f = nullptr;
    bool first_write = false;
    if (!S.out_file.empty()) {
        FILE *probe = std::fopen(S.out_file.c_str(), "r");
        first_write = probe == nullptr;
        if (probe) std::fclose(probe);
        f = std::fopen(S.out_file.c_str(), "a");
        if (f && first_write) std::fprintf(f, "Iteration number|Solver residual|Solver convergence\n");
    }
    for (uint32_t it = 0; it < iters && it < hist_n; it++) {
        double e = S.history[2 * it], t = S.history[2 * it + 1];
        bool conv = !std::isnormal(e) || !std::isnormal(t) || e < c.solver_cvg_value || std::fabs(t) < kEps * kEps;
        if (!conv && it % c.output_frequency != 0) continue;
        std::printf("> %6u of %-6u |% 12e |% 12e |\n", it + 1, c.n_outer_iterations, e, t);
        if (f) std::fprintf(f, "%u|%.7e|%.7e\n", it + 1, e, t);
    }
    if (f) std::fclose(f);
    std::printf("-------------------------------------------------------------------------------\n");
}

// Replay the captured iteration graph until the device-side stop flag is raised.  `body` is the
// function the graph was captured from; it is called directly only in the eager mode, and the
// caller keeps it (and the locals it refers to) alive for the whole loop.  Two
// replays are kept in flight so that the host's flag check overlaps device work, except when
// the error history projects convergence within the replay already in flight: then the next
// replay waits until that one's state is read.  A replay enqueued after convergence runs all
// of its iterations idle (each kernel still dispatches and returns), so this only changes how
// many idle iterations follow the stop, never the iterations or the results.
static int run_graph_loop(rs_solver &S, uint32_t &done_out, const std::function<void()> &body)
{
    hipStream_t st = runtime().stream;
    if (!S.hflags) {
        RS_HIP(hipHostMalloc(&S.hflags, 2 * sizeof(SolveState), hipHostMallocDefault));
        RS_HIP(hipEventCreateWithFlags(&S.ev[0], hipEventDisableTiming));
        RS_HIP(hipEventCreateWithFlags(&S.ev[1], hipEventDisableTiming));
    }
    SolveState *hs = static_cast<SolveState *>(S.hflags);
    hipEvent_t *ev = S.ev;
    auto enqueue = [&](int slot) -> int {
        if (S.eager) {
            body();
            RS_HIP(hipGetLastError());
        } else {
            RS_HIP(hipGraphLaunch(S.graph, st));
        }
        RS_HIP(hipMemcpyAsync(&hs[slot], S.st.p, sizeof(SolveState), hipMemcpyDeviceToHost, st));
        RS_HIP(hipEventRecord(ev[slot], st));
        return RS_OK;
    };
    uint64_t launched = 0, seen = 0;   // replays enqueued; replays whose state was read
    bool ahead = true;                 // keep a second replay in flight
    bool have_prev = false;
    double e_prev = 0.0;
    uint32_t it_prev = 0;
    int rc = enqueue(0);
    launched = 1;
    while (rc == RS_OK) {
        if (launched == seen || (ahead && launched == seen + 1)) {
            if ((rc = enqueue((int)(launched % 2)))) break;
            launched++;
        }
        const int slot = (int)(seen % 2);
        hipError_t e = hipEventSynchronize(ev[slot]);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipEventSynchronize");
            break;
        }
        seen++;
        const SolveState h = hs[slot];
        if (h.done != kStopRunning) {
            done_out = h.done;
            break;
        }
        // iterations still needed at the error's mean log rate since the previous read
        ahead = true;
        if (have_prev && h.it > it_prev && e_prev > 0.0 && h.error > 0.0 && h.cvg > 0.0 && std::isfinite(h.error)) {
            const double rate = std::log(h.error / e_prev) / (double)(h.it - it_prev);
            if (rate < 0.0 && std::log(h.cvg / h.error) / rate <= (double)kGraphIters) ahead = false;
        }
        have_prev = true;
        e_prev = h.error;
        it_prev = h.it;
    }
    (void)hipStreamSynchronize(st);
    return rc;
}

template <typename Args>
static std::string graph_key_of(int kind, const Args &a, int plan, uint32_t nb, uint32_t w, uint32_t vb)
{
    std::string k(reinterpret_cast<const char *>(&a), sizeof(Args));
    uint32_t extra[5] = { (uint32_t)kind, (uint32_t)plan, nb, w, vb };
    k.append(reinterpret_cast<const char *>(extra), sizeof(extra));
    return k;
}

static int capture(rs_solver &S, const std::function<void()> &body, const std::string &key)
{
    const char *ng = std::getenv("RS_NO_GRAPH");
    S.eager = ng && ng[0] == '1';
    if (S.eager) return RS_OK;
    if (S.graph && S.graph_key == key) return RS_OK;
    if (S.graph) {
        (void)hipGraphExecDestroy(S.graph);
        S.graph = nullptr;
    }
    hipStream_t st = runtime().stream;
    hipGraph_t g;
    RS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    body();
    hipError_t e = hipStreamEndCapture(st, &g);
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&S.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        S.graph = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    S.graph_key = key;
    return RS_OK;
}

// kGraphIters FP64 CG iterations (src/rmatrixsolver.cpp:248-353), 3 launches each:
//   dir, spmv_dot (finalizes alpha), update (finalizes the stopping rule and beta).
static void cg64_iterations(const CgArgs &a, int plan, uint32_t nb, uint32_t w, uint32_t vb, hipStream_t stream)
{
    for (uint32_t i = 0; i < kGraphIters; i++) {
        RS_DISPATCH_W(w, k_cg_dir, vb, a);
        RS_SHAPE_SWITCH(plan, k_cg_spmv_dot, nb, stream, a);
        RS_DISPATCH_W(w, k_cg_update, vb, a);
    }
}

static int solve_cg64(rs_solver &S, rs_matrix &A, rs_precond &P, double *d_x, uint32_t &done)
{
    int rc = matrix_tune(A, false);
    if (rc) return rc;
    const int plan = A.plan[0];
    if ((rc = matrix_sync_plan(A, plan, false))) return rc;
    uint32_t nb = 0;
    CgArgs a{};
    a.A = plan_view(A, plan, nb);
    a.b = S.y.p;
    a.x = d_x;
    a.r = S.r.p;
    a.p = S.p.p;
    a.q = S.q.p;
    a.ptype = P.type;
    a.pw = P.width ? P.width : 1;
    a.pd = P.data.p;
    a.partials = S.partials.p;
    a.arrive = reinterpret_cast<uint32_t *>(S.partials.p + kPartialDoubles);
    a.st = S.st.p;
    a.hist = S.hist.p;
    a.stream = (uint64_t)A.n * sizeof(double) >= kStreamVectorBytes;
    uint32_t w = a.pw;
    uint32_t vb = vec_blocks(A.n / w);
    hipStream_t stream = runtime().stream;
    k_zero_u32<<<1, 64, 0, stream>>>(a.arrive, kArriveWords);
    RS_SHAPE_SWITCH(plan, k_cg_init, nb, stream, a);
    fin(S.partials.p, nb, 1, FIN_CG_INIT, 0, S.st.p, S.hist.p);
    RS_DISPATCH_W(w, k_cg_rho, vb, a);
    fin(S.partials.p, vb, 1, FIN_RHO, 0, S.st.p, S.hist.p);
    RS_HIP(hipGetLastError());
    const std::function<void()> body = [&] { cg64_iterations(a, plan, nb, w, vb, stream); };
    rc = capture(S, body, graph_key_of(1, a, plan, nb, w, vb));
    if (rc) return rc;
    return run_graph_loop(S, done, body);
}

static int solve_cg_mixed(rs_solver &S, rs_matrix &A, rs_precond &P, double *d_x, uint32_t &done,
                          uint32_t &refinements)
{
    int rc;
    if ((rc = matrix_refresh_fp32(A)) || (rc = precond_fp32(P)) || (rc = matrix_tune(A, true)) ||
        (rc = matrix_tune(A, false)))
        return rc;
    const int plan32 = A.plan[1], plan64 = A.plan[0];
    if ((rc = matrix_sync_plan(A, plan32, true)) || (rc = matrix_sync_plan(A, plan64, false))) return rc;
    uint32_t nb32 = 0, nb64 = 0;
    MpArgs a{};
    a.A = plan_view(A, plan32, nb32);
    a.b = S.y.p;
    a.x = d_x;
    a.r = S.r32.p;
    a.p = S.p32.p;
    a.q = S.q32.p;
    a.dx = S.dx32.p;
    a.ptype = P.type;
    a.pw = P.width ? P.width : 1;
    a.pd = P.data32.p;
    a.partials = S.partials.p;
    a.arrive = reinterpret_cast<uint32_t *>(S.partials.p + kPartialDoubles);
    a.st = S.st.p;
    a.hist = S.hist.p;
    MpArgs a64 = a;
    a64.A = plan_view(A, plan64, nb64);
    uint32_t w = a.pw;
    uint32_t vb = vec_blocks(A.n / w);
    uint32_t nvb = vec_blocks(A.n);
    hipStream_t stream = runtime().stream;
    k_fill32<<<nvb, kBlock, 0, stream>>>(S.dx32.p, A.n, 0.0f);
    k_zero_u32<<<1, 64, 0, stream>>>(a.arrive, kArriveWords);
    RS_SHAPE_SWITCH(plan64, k_mp_true_resid, nb64, stream, a64);
    fin(S.partials.p, nb64, 2, FIN_MP_TRUE, 1, S.st.p, S.hist.p);
    RS_DISPATCH_W(w, k_mp_rho, vb, a);
    fin(S.partials.p, vb, 1, FIN_MP_RHO, 1, S.st.p, S.hist.p);
    RS_HIP(hipGetLastError());
    const std::function<void()> body = [&] {
        for (uint32_t i = 0; i < kGraphIters; i++) {
            RS_DISPATCH_W(w, k_mp_dir, vb, a);
            RS_SHAPE_SWITCH(plan32, k_mp_spmv_dot, nb32, stream, a);   // finalizes alpha
            RS_DISPATCH_W(w, k_mp_update, vb, a);                      // finalizes the stopping rule, beta

        }
    };
    rc = capture(S, body, graph_key_of(2, a, plan32, nb32, w, vb));
    if (rc) return rc;
    refinements = 0;
    SolveState h;
    RS_HIP(hipMemcpyAsync(&h, S.st.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    RS_HIP(hipStreamSynchronize(stream));
    done = h.done;
    double best = h.t2;
    uint32_t no_gain = 0;
    while (done == kStopRunning) {
        if ((rc = run_graph_loop(S, done, body))) return rc;
        if (done != kStopPause) break;
        refinements++;
        // replacements that stop lowering the FP64 residual mean the FP32 recursion has lost
        // contact with the true residual: stop instead of spinning
        if (no_gain >= kMaxReplacementsWithoutGain) {
            k_mp_fold<<<nvb, kBlock, 0, stream>>>(d_x, S.dx32.p, A.n);
            RS_HIP(hipGetLastError());
            h.reason = RS_REASON_STAGNATED;
            RS_HIP(hipMemcpyAsync(&S.st.p->reason, &h.reason, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
            RS_HIP(hipStreamSynchronize(stream));
            return RS_OK;
        }
        k_mp_fold<<<nvb, kBlock, 0, stream>>>(d_x, S.dx32.p, A.n);
        RS_SHAPE_SWITCH(plan64, k_mp_true_resid, nb64, stream, a64);
        fin(S.partials.p, nb64, 2, FIN_MP_TRUE, 0, S.st.p, S.hist.p);
        RS_DISPATCH_W(w, k_mp_rho, vb, a);
        fin(S.partials.p, vb, 1, FIN_MP_RHO, 0, S.st.p, S.hist.p);
        RS_HIP(hipGetLastError());
        RS_HIP(hipMemcpyAsync(&h, S.st.p, sizeof(h), hipMemcpyDeviceToHost, stream));
        RS_HIP(hipStreamSynchronize(stream));
        done = h.done;
        if (h.t2 < 0.5 * best) {
            best = h.t2;
            no_gain = 0;
        } else {
            no_gain++;
        }
    }
    // fold the last FP32 corrections into the FP64 solution
    k_mp_fold<<<nvb, kBlock, 0, stream>>>(d_x, S.dx32.p, A.n);
    RS_HIP(hipGetLastError());
    return RS_OK;
}

// Restarted GMRES (src/rmatrixsolver.cpp:357-581).  mixed: the inner cycle keeps its basis V,
// the preconditioned basis Z and w in FP32 and runs the SpMV and the preconditioner on FP32
// values (the Gram-Schmidt products, norms, Hessenberg column, rotations and back substitution
// stay FP64); each restart's residual b - A x and the solution stay FP64, so the outer stopping
// test is the FP64 one.  The basis reads of the Gram-Schmidt steps, which bound an inner
// iteration, are halved.
static int solve_gmres(rs_solver &S, rs_matrix &A, rs_precond &P, double *d_x, uint32_t &outer,
                       uint32_t &inner_total, bool mixed)
{
    hipStream_t stream = runtime().stream;
    const uint32_t n = A.n;
    const uint32_t nouter = S.conf.n_outer_iterations, ninner = S.conf.n_inner_iterations;
    if (ninner == 0) return fail(RS_ERR_INVALID_ARG, "GMRES: nInnerIterations is 0");
    int rc = matrix_tune(A, false);
    if (rc) return rc;
    const int plan = A.plan[0];
    if ((rc = matrix_sync_plan(A, plan, false))) return rc;
    uint32_t nb = 0;
    CsrView v = plan_view(A, plan, nb);
    int plan32 = -1;
    uint32_t nb32 = 0;
    CsrView v32{};
    const uint64_t ld32 = (n + 3ull) / 4 * 4;   // FP32 basis rows start at multiples of 4 entries (16 bytes)
    if (mixed) {
        const size_t N = std::max(n, 1u);
        if ((rc = matrix_refresh_fp32(A)) || (rc = matrix_tune(A, true)) || (rc = precond_fp32(P))) return rc;
        plan32 = A.plan[1];
        if ((rc = matrix_sync_plan(A, plan32, true))) return rc;
        v32 = plan_view(A, plan32, nb32);
        if ((rc = S.V32.alloc((ninner + 1ull) * ld32)) || (rc = S.Z32.alloc((ninner + 1ull) * ld32)) ||
            (rc = S.w32.alloc(N)))
            return rc;
    }
    uint32_t nvb = vec_blocks(n);
    uint32_t w = P.width ? P.width : 1;
    const uint64_t ld = (n + 1ull) / 2 * 2;     // FP64 basis rows start at multiples of 2 entries (16 bytes)
    SolveState *dst = S.st.p;
    SolveState h;
    // the inner cycle runs on the device (Gram-Schmidt, Hessenberg column, Givens rotation,
    // stopping test, back substitution); the host launches it in batches of kGmBatch
    // iterations and reads only the stop flag between batches
    constexpr uint32_t kGmBatch = 8;
    if ((rc = S.gm.alloc(GmView::doubles(ninner))) || (rc = S.gmflag.alloc(2))) return rc;
    const GmView gv(S.gm.p, ninner);
    std::vector<double> init(ninner + 2, 0.0);   // pp (k + 1 values) ... and beta, uploaded per cycle
    // bn = |b| (src/rmatrixsolver.cpp:375)
    k_norm2<<<nvb, kBlock, 0, stream>>>(S.y.p, n, S.partials.p);
    fin(S.partials.p, nvb, 1, FIN_NORM0, 0, dst, nullptr);
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(&h, dst, sizeof(h), hipMemcpyDeviceToHost, stream));
    RS_HIP(hipStreamSynchronize(stream));
    const double bn = h.t0, An = A.frob;
    uint32_t nbm = std::min<uint32_t>(nvb, 512);
    DBuf<double> mpart;
    if ((rc = mpart.alloc((size_t)(ninner + 1) * nbm))) return rc;
    outer = 0;
    inner_total = 0;
    const uint32_t *stop = S.gmflag.p;
    for (uint32_t ito = 0; ito < nouter; ito++) {
        RS_SHAPE_SWITCH(plan, k_gmres_resid, nb, stream, v, S.y.p, d_x, S.ro.p, S.partials.p);
        fin(S.partials.p, nb, 2, FIN_NORM01, 0, dst, nullptr);
        RS_HIP(hipGetLastError());
        RS_HIP(hipMemcpyAsync(&h, dst, sizeof(h), hipMemcpyDeviceToHost, stream));
        RS_HIP(hipStreamSynchronize(stream));
        double xn = h.t0, beta = h.t1;
        double norm = An * xn + bn;                         // src/rmatrixsolver.cpp:415-421
        if (std::fabs(norm) < kEps) norm = kEps;
        double err = S.conf.stop_rule == RS_STOP_RELRES ? beta / std::max(bn, kEps) : beta / norm;
        S.trend = err - S.error;
        S.error = err;
        if (ito < S.history.size() / 2) {
            S.history[2 * ito] = err;
            S.history[2 * ito + 1] = S.trend;
        }
        outer = ito + 1;
        bool conv = !std::isnormal(S.error) || !std::isnormal(S.trend) || S.error < S.conf.solver_cvg_value ||
                    std::fabs(S.trend) < kEps * kEps;
        if (conv) break;
        if (mixed) k_scale_by<float><<<nvb, kBlock, 0, stream>>>(S.ro.p, S.V32.p, n, beta);   // row 0
        else k_scale_by<double><<<nvb, kBlock, 0, stream>>>(S.ro.p, S.V.p, n, beta);
        // pp = (beta, 0, ...), beta; stop = 0
        init[0] = beta;
        init[ninner + 1] = beta;
        RS_HIP(hipMemcpyAsync(gv.pp, init.data(), sizeof(double) * (ninner + 1), hipMemcpyHostToDevice, stream));
        RS_HIP(hipMemcpyAsync(&gv.beta(), &init[ninner + 1], sizeof(double), hipMemcpyHostToDevice, stream));
        k_zero_u32<<<1, 64, 0, stream>>>(S.gmflag.p, 2);
        uint32_t hflag[2] = { 0, 0 };
        for (uint32_t iti = 0; iti < ninner;) {
            for (const uint32_t end = std::min(ninner, iti + kGmBatch); iti < end; iti++) {
                const uint32_t cnt = iti + 1;
                const dim3 g(nbm, (cnt + 7) / 8);
                if (mixed) {
                    float *vi = S.V32.p + (uint64_t)iti * ld32, *zi = S.Z32.p + (uint64_t)iti * ld32;
                    RS_DISPATCH_W(w, k_pc_apply32, vec_blocks(n / w), P.type, w, P.data32.p, n, vi, zi, stop);
                    RS_SHAPE_SWITCH(plan32, k_gm_spmv32, nb32, stream, v32, zi, S.w32.p, stop);
                    k_multidot<float><<<g, kBlock, 0, stream>>>(S.w32.p, S.V32.p, ld32, n, cnt, mpart.p, stop);
                    k_multidot_final<<<cnt, kBlock, 0, stream>>>(mpart.p, nbm, S.hbuf.p, stop);
                    k_gs_subtract<float><<<nvb, kBlock, 0, stream>>>(S.w32.p, S.V32.p, ld32, n, cnt, S.hbuf.p,
                                                                      S.partials.p, stop);
                    fin(S.partials.p, nvb, 1, FIN_NORM0, 0, dst, nullptr);
                    k_normalize<float><<<nvb, kBlock, 0, stream>>>(S.w32.p, S.V32.p + (uint64_t)(iti + 1) * ld32, n,
                                                                    dst, stop);
                } else {
                    double *vi = S.V.p + (uint64_t)iti * ld, *zi = S.Z.p + (uint64_t)iti * ld;
                    RS_DISPATCH_W(w, k_pc_apply, vec_blocks(n / w), P.type, w, P.data.p, n, vi, zi, stop);
                    RS_SHAPE_SWITCH(plan, k_gm_spmv, nb, stream, v, zi, S.w.p, stop);
                    k_multidot<double><<<g, kBlock, 0, stream>>>(S.w.p, S.V.p, ld, n, cnt, mpart.p, stop);
                    k_multidot_final<<<cnt, kBlock, 0, stream>>>(mpart.p, nbm, S.hbuf.p, stop);
                    k_gs_subtract<double><<<nvb, kBlock, 0, stream>>>(S.w.p, S.V.p, ld, n, cnt, S.hbuf.p,
                                                                       S.partials.p, stop);
                    fin(S.partials.p, nvb, 1, FIN_NORM0, 0, dst, nullptr);
                    k_normalize<double><<<nvb, kBlock, 0, stream>>>(S.w.p, S.V.p + (uint64_t)(iti + 1) * ld, n, dst,
                                                                     stop);
                }
                k_gm_givens<<<1, kGmBlock, 0, stream>>>(S.gm.p, ninner, S.gmflag.p, S.hbuf.p, dst, iti);
            }
            RS_HIP(hipGetLastError());
            RS_HIP(hipMemcpyAsync(hflag, S.gmflag.p, sizeof(hflag), hipMemcpyDeviceToHost, stream));
            RS_HIP(hipStreamSynchronize(stream));
            if (hflag[0]) break;
        }
        inner_total += hflag[1];
        k_gm_backsub<<<1, kGmBlock, 0, stream>>>(S.gm.p, ninner, S.gmflag.p);
        if (mixed) k_update_x<float><<<nvb, kBlock, 0, stream>>>(d_x, S.Z32.p, ld32, n, S.gmflag.p + 1, gv.y);
        else k_update_x<double><<<nvb, kBlock, 0, stream>>>(d_x, S.Z.p, ld, n, S.gmflag.p + 1, gv.y);
        RS_HIP(hipGetLastError());
        RS_HIP(hipStreamSynchronize(stream));
    }
    return RS_OK;
}

}  // namespace rs
struct rs_learned;
namespace rs {
static int solve_fcg_learned(rs_solver &S, rs_matrix &A, rs_precond &P, rs_learned &L, double *d_x, uint32_t &done,
                             uint32_t &refinements);
static int solve_fcg64_learned(rs_solver &S, rs_matrix &A, rs_learned &L, double *d_x, uint32_t &done);

static int solve_device_ext(rs_solver &S, rs_matrix &A, const double *d_b, double *d_x, int32_t ptype,
                            uint32_t block_size, rs_learned *L, rs_info *info)
{
    const rs_conf &c = S.conf;
    if (c.solver_type != RS_SOLVER_CG && c.solver_type != RS_SOLVER_GMRES)
        return fail(RS_ERR_UNKNOWN_SOLVER, "Unknown matrix solver type '" + std::to_string(c.solver_type) + "'");
    const bool learned = L != nullptr;
    if (!learned && (ptype < RS_PRECOND_NONE || ptype > RS_PRECOND_BLOCK_JACOBI))
        return fail(RS_ERR_INVALID_PRECOND, "Invalid matrix preconditioner type '" + std::to_string(ptype) + "'");
    if (learned) ptype = RS_PRECOND_JACOBI;   // the smoother part and the fallback
    // the learned cycle always runs in FP32; PRECISION_MIXED also keeps the flexible CG's
    // vectors and product in FP32 (FP64 residual replacement), PRECISION_FP64 keeps them in FP64
    bool mixed = c.precision == RS_PRECISION_MIXED;
    bool gmres = c.solver_type == RS_SOLVER_GMRES;
    int rc;
    // tune outside the timed region (once per matrix)
    if (A.n && (rc = matrix_tune(A, false))) return rc;
    if (A.n && mixed && (rc = matrix_tune(A, true))) return rc;
    // RMatrixPreconditioner P(A, type, blockSize)  (src/rmatrixsolver.cpp:160)
    rs_precond P;
    if ((rc = ensure_ws(S, A.n, mixed, gmres))) return rc;
    hipStream_t stream = runtime().stream;
    EventPair e0e1_;
    RS_HIP(e0e1_.create());
    hipEvent_t e0 = e0e1_.e[0], e1 = e0e1_.e[1];
    RS_HIP(hipEventRecord(e0, stream));
    // the learned path's iterations do not read P: it is built only if Jacobi-CG takes over
    if (!learned && (rc = precond_build(A, ptype, block_size, P))) return rc;
    const uint32_t n = A.n;
    uint32_t nvb = vec_blocks(std::max(n, 1u));
    uint32_t hist_cap = (uint32_t)(S.hist.n / 2);
    k_state_init<<<1, 64, 0, stream>>>(S.st.p, A.frob, c.solver_cvg_value, S.error, c.n_outer_iterations,
                                       mixed ? RS_STOP_RELRES : (uint32_t)c.stop_rule, hist_cap, mixed ? 1 : 0);
    // equation scaling (src/rmatrixsolver.cpp:164-184)
    k_norm2<<<nvb, kBlock, 0, stream>>>(d_b, n, S.partials.p);
    fin(S.partials.p, nvb, 1, FIN_SCALE, 0, S.st.p, S.hist.p);
    k_apply_scale<<<nvb, kBlock, 0, stream>>>(d_b, S.y.p, d_x, n, S.st.p);
    RS_HIP(hipGetLastError());
    uint32_t done = 0, iters = 0, inner = 0, refinements = 0, applications = 0, fallback = 0;
    S.history.assign(2ull * std::min(hist_cap, std::max(1u, c.n_outer_iterations)), 0.0);
    if (n > 0) {
        if (gmres) {
            rc = solve_gmres(S, A, P, d_x, iters, inner, mixed);
        } else if (learned) {
            rc = mixed ? solve_fcg_learned(S, A, P, *L, d_x, done, refinements)
                       : solve_fcg64_learned(S, A, *L, d_x, done);
            if (rc) return rc;
            SolveState hs;
            RS_HIP(hipMemcpyAsync(&hs, S.st.p, sizeof(hs), hipMemcpyDeviceToHost, stream));
            RS_HIP(hipStreamSynchronize(stream));
            applications = hs.it;
            if (hs.reason != RS_REASON_CONVERGED) {
                // Jacobi-CG takes over from the current solution
                fallback = 1;
                if ((rc = precond_build(A, ptype, block_size, P))) return rc;
                k_state_restart<<<1, 64, 0, stream>>>(S.st.p);
                uint32_t more = 0;
                rc = mixed ? solve_cg_mixed(S, A, P, d_x, done, more) : solve_cg64(S, A, P, d_x, done);
                refinements += more;
            }
        } else if (mixed) {
            rc = solve_cg_mixed(S, A, P, d_x, done, refinements);
        } else {
            rc = solve_cg64(S, A, P, d_x, done);
        }
        if (rc) return rc;
    }
    k_unscale<<<nvb, kBlock, 0, stream>>>(d_x, n, S.st.p);
    RS_HIP(hipGetLastError());
    RS_HIP(hipEventRecord(e1, stream));
    RS_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    RS_HIP(hipEventElapsedTime(&ms, e0, e1));
    SolveState h;
    RS_HIP(hipMemcpyAsync(&h, S.st.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    RS_HIP(hipStreamSynchronize(stream));
    int reason;
    double equation_scale = h.scale;
    if (!gmres) {
        iters = h.it;
        S.error = h.error;
        S.trend = h.trend;
        reason = (int)h.reason;
        if (!S.history.empty() && iters)
            RS_HIP(hipMemcpy(S.history.data(), S.hist.p,
                             sizeof(double) * 2 * std::min<uint64_t>(iters, S.history.size() / 2), hipMemcpyDeviceToHost));
    } else {
        reason = RS_REASON_MAXIT;
        bool conv = !std::isnormal(S.error) || !std::isnormal(S.trend) || S.error < c.solver_cvg_value ||
                    std::fabs(S.trend) < kEps * kEps;
        if (conv) {
            if (!std::isnormal(S.error) || !std::isnormal(S.trend)) reason = RS_REASON_NONNORMAL;
            else if (S.error < c.solver_cvg_value) reason = RS_REASON_CONVERGED;
            else reason = RS_REASON_STAGNATED;
        }
    }
    if (n == 0) iters = 0;
    S.last_iterations = iters;
    // true residual on the caller's (unscaled) system, FP64
    double true_relres = 0.0, b_norm = 0.0;
    if (n > 0) {
        uint32_t nb = 0;
        const int plan = A.plan[0] >= 0 ? A.plan[0] : default_plan(A);
        if ((rc = matrix_sync_plan(A, plan, false))) return rc;
        CsrView v = plan_view(A, plan, nb);
        RS_SHAPE_SWITCH(plan, k_resid_norm, nb, stream, v, d_b, d_x, S.partials.p);
        fin(S.partials.p, nb, 2, FIN_NORM01, 0, S.st.p, S.hist.p);
        RS_HIP(hipGetLastError());
        RS_HIP(hipMemcpyAsync(&h, S.st.p, sizeof(h), hipMemcpyDeviceToHost, stream));
        RS_HIP(hipStreamSynchronize(stream));
        true_relres = h.t1 > 0.0 ? h.t0 / h.t1 : h.t0;
        b_norm = h.t1;
    }
    emit_log(S, gmres ? "GMRES" : "CG", iters, (uint32_t)(S.history.size() / 2));
    if (info) {
        info->iterations = iters;
        info->inner_iterations = gmres ? inner : refinements;
        info->reason = reason;
        info->converged = reason == RS_REASON_CONVERGED;
        info->error = S.error;
        info->trend = S.trend;
        info->equation_scale = equation_scale;
        info->a_norm = A.frob;
        info->b_norm = b_norm;
        info->true_relres = true_relres;
        info->solve_ms = ms;
        info->precond_applications = applications;
        info->fallback = fallback;
    }
    return RS_OK;
}

static int solve_device(rs_solver &S, rs_matrix &A, const double *d_b, double *d_x, int32_t ptype,
                        uint32_t block_size, rs_info *info)
{
    return solve_device_ext(S, A, d_b, d_x, ptype, block_size, nullptr, info);
}

}  // namespace rs

// ============================================================================ allocation-free solve
// Workspace of rs_solver_solve_device_noalloc: one file-scope device array, allocated with the
// code object on every GPU the library runs on and held until the process ends.  Host code
// never passes `scratch` itself to a kernel (on the host it is only a stand-in for the
// symbol): rs_init looks up its device address once per device (scratch_setup) and every
// launch uses that address.  Calls on one GPU share the array: overlapping calls race, and
// each call initialises every part it reads (state, partial sums, vectors, preconditioner).
static constexpr uint64_t kNoAllocScratchBytes = 128ull << 20;   // 134,217,728 bytes
static constexpr uint32_t kNoAllocHistCap = 65536;                // error history entries kept
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace rs {

static constexpr int kMaxScratchDevices = 64;
static void *g_scratch_ptr[kMaxScratchDevices] = {};
static std::mutex g_scratch_mu;

// Device address of `scratch` on `device`, which must be the current device.
int scratch_setup(int device)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return fail(RS_ERR_UNSUPPORTED, "allocation-free workspace: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return RS_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(RS_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[device] = p;
    return RS_OK;
}

static int scratch_address(void *&out)
{
    int dev = -1;
    RS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices || !g_scratch_ptr[dev]) {
        int rc = scratch_setup(dev);   // a device the library was not initialised on
        if (rc) return rc;
    }
    out = g_scratch_ptr[dev];
    return RS_OK;
}

// Bytes of `scratch` a solve of n unknowns needs (state, partial sums, history, y r p q, the
// preconditioner's n * w values), laid out at 256-byte boundaries.
struct NoAllocLayout {
    uint64_t st = 0, partials = 0, hist = 0, y = 0, r = 0, p = 0, q = 0, pd = 0, total = 0;
    uint32_t hist_cap = 0;
};

static NoAllocLayout noalloc_layout(uint64_t n, uint32_t w, bool precond, uint32_t n_outer)
{
    NoAllocLayout L;
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = off;
        off += (bytes + 255) & ~255ull;
        return o;
    };
    const uint64_t N = std::max<uint64_t>(n, 1);
    L.hist_cap = std::min(kNoAllocHistCap, std::max(1u, n_outer));
    L.st = take(sizeof(SolveState));
    L.partials = take(kPartialBytes);
    L.hist = take(sizeof(double) * 2 * L.hist_cap);
    L.y = take(sizeof(double) * N);
    L.r = take(sizeof(double) * N);
    L.p = take(sizeof(double) * N);
    L.q = take(sizeof(double) * N);
    L.pd = precond ? take(sizeof(double) * N * w) : 0;
    L.total = off;
    return L;
}

// RMatrixSolver::solve (src/rmatrixsolver.cpp:158-201) with solveCG (:208-355), FP64, on the
// caller's device vectors, using only `scratch`: the same kernels in the same order as
// solve_device_ext's FP64 CG, launched eagerly (a graph instantiation allocates), with the
// stop flag read back every kGraphIters iterations.  The SpMV shape is the matrix's tuned one
// when its layout needs nothing more, else vector CSR (no tuning: the tuner allocates).
static int solve_device_noalloc(rs_solver &S, rs_matrix &A, const double *d_b, double *d_x, int32_t ptype,
                                uint32_t block_size, rs_info *info)
{
    const rs_conf &c = S.conf;
    if (c.solver_type != RS_SOLVER_CG && c.solver_type != RS_SOLVER_GMRES)
        return fail(RS_ERR_UNKNOWN_SOLVER, "Unknown matrix solver type '" + std::to_string(c.solver_type) + "'");
    if (ptype < RS_PRECOND_NONE || ptype > RS_PRECOND_BLOCK_JACOBI)
        return fail(RS_ERR_INVALID_PRECOND, "Invalid matrix preconditioner type '" + std::to_string(ptype) + "'");
    if (c.solver_type != RS_SOLVER_CG || c.precision != RS_PRECISION_FP64)
        return fail(RS_ERR_UNSUPPORTED, "rs_solver_solve_device_noalloc: FP64 CG only");
    uint32_t w = 1;
    if (ptype == RS_PRECOND_BLOCK_JACOBI) {
        if (block_size == 0) return fail(RS_ERR_INVALID_ARG, "block-Jacobi: block size 0");
        if (block_size > 16) return fail(RS_ERR_UNSUPPORTED, "block-Jacobi: block size above 16");
        w = block_width(A.n, block_size);
    }
    const NoAllocLayout lay = noalloc_layout(A.n, w, ptype != RS_PRECOND_NONE, c.n_outer_iterations);
    if (lay.total > kNoAllocScratchBytes)
        return fail(RS_ERR_INVALID_ARG, "rs_solver_solve_device_noalloc: " + std::to_string(A.n) + " unknowns need " +
                                            std::to_string(lay.total) + " bytes of workspace, above its " +
                                            std::to_string(kNoAllocScratchBytes) + "-byte maximum");
    void *base_v = nullptr;
    int rc = scratch_address(base_v);
    if (rc) return rc;
    unsigned char *base = static_cast<unsigned char *>(base_v);
    SolveState *st = reinterpret_cast<SolveState *>(base + lay.st);
    double *partials = reinterpret_cast<double *>(base + lay.partials);
    double *hist = reinterpret_cast<double *>(base + lay.hist);
    double *y = reinterpret_cast<double *>(base + lay.y);
    double *pd = ptype != RS_PRECOND_NONE ? reinterpret_cast<double *>(base + lay.pd) : nullptr;
    // an SpMV shape whose layout is already there (a SELL shape needs its value copy allocated)
    int plan = A.plan[0];
    if (plan < 0 || (is_sell_plan(plan) && (!A.sell_val.p || A.sell_val.n != A.sell_col.n))) plan = default_plan(A);
    if ((rc = matrix_sync_plan(A, plan, false))) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t stream = runtime().stream;
    if ((rc = precond_build_at(A, ptype, w, pd))) return rc;
    const uint32_t n = A.n;
    const uint32_t nvb = vec_blocks(std::max(n, 1u));
    k_state_init<<<1, 64, 0, stream>>>(st, A.frob, c.solver_cvg_value, S.error, c.n_outer_iterations,
                                       (uint32_t)c.stop_rule, lay.hist_cap, 0);
    k_norm2<<<nvb, kBlock, 0, stream>>>(d_b, n, partials);
    fin(partials, nvb, 1, FIN_SCALE, 0, st, hist);
    k_apply_scale<<<nvb, kBlock, 0, stream>>>(d_b, y, d_x, n, st);
    RS_HIP(hipGetLastError());
    uint32_t nb = 0;
    CgArgs a{};
    a.A = plan_view(A, plan, nb);
    a.b = y;
    a.x = d_x;
    a.r = reinterpret_cast<double *>(base + lay.r);
    a.p = reinterpret_cast<double *>(base + lay.p);
    a.q = reinterpret_cast<double *>(base + lay.q);
    a.ptype = ptype;
    a.pw = w;
    a.pd = pd;
    a.partials = partials;
    a.arrive = reinterpret_cast<uint32_t *>(partials + kPartialDoubles);
    a.st = st;
    a.hist = hist;
    a.stream = (uint64_t)n * sizeof(double) >= kStreamVectorBytes;
    const uint32_t vb = vec_blocks(n / w);
    SolveState h{};
    if (n > 0) {
        k_zero_u32<<<1, 64, 0, stream>>>(a.arrive, kArriveWords);
        RS_SHAPE_SWITCH(plan, k_cg_init, nb, stream, a);
        fin(partials, nb, 1, FIN_CG_INIT, 0, st, hist);
        RS_DISPATCH_W(w, k_cg_rho, vb, a);
        fin(partials, vb, 1, FIN_RHO, 0, st, hist);
        RS_HIP(hipGetLastError());
        for (;;) {
            cg64_iterations(a, plan, nb, w, vb, stream);
            RS_HIP(hipGetLastError());
            RS_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, stream));
            RS_HIP(hipStreamSynchronize(stream));
            if (h.done != kStopRunning) break;
        }
    }
    k_unscale<<<nvb, kBlock, 0, stream>>>(d_x, n, st);
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, stream));
    RS_HIP(hipStreamSynchronize(stream));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const uint32_t iters = n > 0 ? h.it : 0;
    const double equation_scale = h.scale;
    S.error = h.error;
    S.trend = h.trend;
    S.last_iterations = 0;   // no host history is kept
    double true_relres = 0.0, b_norm = 0.0;
    if (n > 0) {
        RS_SHAPE_SWITCH(plan, k_resid_norm, nb, stream, a.A, d_b, d_x, partials);
        fin(partials, nb, 2, FIN_NORM01, 0, st, hist);
        RS_HIP(hipGetLastError());
        SolveState r{};
        RS_HIP(hipMemcpyAsync(&r, st, sizeof(r), hipMemcpyDeviceToHost, stream));
        RS_HIP(hipStreamSynchronize(stream));
        true_relres = r.t1 > 0.0 ? r.t0 / r.t1 : r.t0;
        b_norm = r.t1;
    }
    if (info) {
        info->iterations = iters;
        info->inner_iterations = 0;
        info->reason = (int)h.reason;
        info->converged = h.reason == RS_REASON_CONVERGED;
        info->error = h.error;
        info->trend = h.trend;
        info->equation_scale = equation_scale;
        info->a_norm = A.frob;
        info->b_norm = b_norm;
        info->true_relres = true_relres;
        info->solve_ms = ms;
        info->precond_applications = 0;
        info->fallback = 0;
    }
    return RS_OK;
}

}  // namespace rs

extern "C" int rs_solver_create(const rs_conf *conf, rs_solver **out)
{
    if (!conf || !out) return fail(RS_ERR_INVALID_ARG, "rs_solver_create: null argument");
    rs_solver *S = new rs_solver();
    S->conf = *conf;
    if (conf->output_file_name) S->out_file = conf->output_file_name;
    S->conf.output_file_name = nullptr;
    *out = S;
    return RS_OK;
}

extern "C" int rs_solver_copy(const rs_solver *src, rs_solver **out)
{
    if (!src || !out) return fail(RS_ERR_INVALID_ARG, "rs_solver_copy: null argument");
    rs_solver *S = new rs_solver();
    S->conf = src->conf;
    S->out_file = src->out_file;
    S->error = src->error;
    S->trend = src->trend;
    *out = S;
    return RS_OK;
}

extern "C" int rs_solver_disable_convergence_log(rs_solver *S)
{
    if (!S) return fail(RS_ERR_INVALID_ARG, "rs_solver_disable_convergence_log: null solver");
    S->out_file.clear();
    return RS_OK;
}

extern "C" int rs_solver_solve_device(rs_solver *S, rs_matrix *A, const double *d_b, double *d_x,
                                      int32_t precond_type, uint32_t block_size, rs_info *info)
{
    if (!S || !A || (A->n && (!d_b || !d_x))) return fail(RS_ERR_INVALID_ARG, "rs_solver_solve_device: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    return solve_device(*S, *A, d_b, d_x, precond_type, block_size, info);
}

extern "C" int rs_solver_solve_device_noalloc(rs_solver *S, rs_matrix *A, const double *d_b, double *d_x,
                                              int32_t precond_type, uint32_t block_size, rs_info *info)
{
    if (!S || !A || (A->n && (!d_b || !d_x)))
        return fail(RS_ERR_INVALID_ARG, "rs_solver_solve_device_noalloc: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    return solve_device_noalloc(*S, *A, d_b, d_x, precond_type, block_size, info);
}

extern "C" int rs_solver_solve(rs_solver *S, rs_matrix *A, const double *b, double *x,
                               int32_t precond_type, uint32_t block_size, rs_info *info)
{
    if (!S || !A || (A->n && (!b || !x))) return fail(RS_ERR_INVALID_ARG, "rs_solver_solve: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    DBuf<double> db, dx;
    if ((rc = db.alloc(A->n)) || (rc = dx.alloc(A->n))) return rc;
    hipStream_t st = runtime().stream;
    if (A->n) {
        RS_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * A->n, hipMemcpyHostToDevice, st));
        RS_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * A->n, hipMemcpyHostToDevice, st));
    }
    rc = solve_device(*S, *A, db.p, dx.p, precond_type, block_size, info);
    if (rc) return rc;
    if (A->n) {
        RS_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * A->n, hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
    }
    return RS_OK;
}

extern "C" int rs_solver_history(const rs_solver *S, double *out, uint32_t cap, uint32_t *count)
{
    if (!S || !count) return fail(RS_ERR_INVALID_ARG, "rs_solver_history: null argument");
    uint32_t n = std::min<uint32_t>(S->last_iterations, (uint32_t)(S->history.size() / 2));
    *count = n;
    if (out) std::memcpy(out, S->history.data(), sizeof(double) * 2 * std::min(n, cap));
    return RS_OK;
}

extern "C" int rs_solver_destroy(rs_solver *S)
{
    delete S;
    return RS_OK;
}

#include "rs_learned.inc.hip"
