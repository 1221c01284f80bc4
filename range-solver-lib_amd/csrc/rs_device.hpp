// rs_device.hpp -- device-side building blocks shared by the gfx950 kernels:
// wave/block reductions, the deterministic grid reduction (fixed-order partials + last
// workgroup finalises), XCD-aware workgroup remapping and the convergence test.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "range_solver_amd.h"

namespace rs {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kXcd = 8;   // MI355X: 8 XCDs; workgroup i is dispatched to XCD i % 8

// Workgroups are dealt round-robin to the 8 XCDs.  Remap so that XCD x owns the
// contiguous logical range [x*nb/8, (x+1)*nb/8): neighbouring rows (and the x entries their
// SpMV gathers) then stay in one XCD's L2.  nb must be a multiple of 8.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    return (b % kXcd) * (nb / kXcd) + b / kXcd;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over aligned groups of G lanes (result in every lane of the group).
template <int G>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

template <int G>
__device__ __forceinline__ float group_sum(float v)
{
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block sum of NV values; result valid in thread 0.  lds must hold NV*BLOCK/64 doubles.
template <int NV, int BLOCK>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds)
{
    constexpr int NW = BLOCK / kWave;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) lds[k * NW + wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; w++) s += lds[k * NW + w];
            v[k] = s;
        }
    }
    __syncthreads();
}

// std::isnormal for doubles: finite, nonzero and not subnormal.
__device__ __forceinline__ bool dev_isnormal(double x)
{
    double a = fabs(x);
    return a >= 2.2250738585072014e-308 && a <= 1.7976931348623157e+308;
}

// RIterationInfo::hasConverged (src/riterationinfo.cpp:90-109) with the reason it stopped.
__device__ __forceinline__ int converged_reason(double error, double trend, double cvg, double eps)
{
    if (!dev_isnormal(error)) return RS_REASON_NONNORMAL;
    if (!dev_isnormal(trend)) return RS_REASON_NONNORMAL;
    if (error < cvg) return RS_REASON_CONVERGED;
    if (fabs(trend) < eps * eps) return RS_REASON_STAGNATED;
    return RS_REASON_RUNNING;
}

}  // namespace rs
