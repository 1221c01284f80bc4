// rs_radiosity.hip -- RSolverRadiativeHeat (src/rsolverradiativeheat.cpp) with the view-factor
// matrix loaded (read from its file or computed by RHemiCube): the patch means of the element
// emissivity and temperatures, the dense net-radiation system and the split of the patch heats
// onto the patch elements by area.
//
//   prepare  :145-241  A_ij = d_ij / E_j - F_ij (1 - E_j) / E_j   (0 when E_j == 0)
//                      b_i  = -sum_j (d_ij - F_ij) sigma T_j^4 + (1 - sum_j F_ij) sigma Tamb_i^4
//   solve    :243-266  CG + Jacobi on (A, b) -- the ordinary rs_solver path
//   process  :268-303  element heat = patch heat * element area / patch area
//
// The reference adds every (i, j) with RSparseMatrix::addValue, so the matrix is dense; here it
// is a CSR whose rows hold every column, written in place by one workgroup per row (coalesced
// reads of the view-factor row, one fixed-order block reduction for b_i).
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "rs_device.hpp"
#include "rs_internal.hpp"

namespace rs {

constexpr double kStefanBoltzmann = 5.670367e-8;   // RSolverGeneric::sigma (src/rsolvergeneric.cpp:12)

// Patch means over the patch's elements (src/rsolverradiativeheat.cpp:175-205).
__global__ void k_rad_means(const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ elems, uint32_t np,
                            const double *__restrict__ em, const double *__restrict__ te,
                            const double *__restrict__ am, double *__restrict__ E, double *__restrict__ T,
                            double *__restrict__ Ta)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= np) return;
    double se = 0.0, st = 0.0, sa = 0.0;
    const uint32_t k0 = ptr[p], k1 = ptr[p + 1];
    for (uint32_t k = k0; k < k1; k++) {
        const uint32_t e = elems[k];
        se += em[e];
        st += te[e];
        sa += am[e];
    }
    if (k1 > k0) {
        const double c = (double)(k1 - k0);
        se /= c;
        st /= c;
        sa /= c;
    }
    E[p] = se;
    T[p] = st;
    Ta[p] = sa;
}

__device__ __forceinline__ double pow4(double t)
{
    const double t2 = t * t;
    return t2 * t2;
}

// One workgroup per patch row i: A_i* and b_i.
__global__ __launch_bounds__(kBlock) void k_rad_rows(const double *__restrict__ F, uint32_t np,
                                                     const double *__restrict__ E, const double *__restrict__ T,
                                                     const double *__restrict__ Ta, double *__restrict__ val,
                                                     double *__restrict__ b)
{
    __shared__ double lds[2 * kBlock / kWave];
    const uint32_t i = blockIdx.x;
    const uint64_t row = (uint64_t)i * np;
    double s[2] = { 0.0, 0.0 };   // sum_j F_ij, -sum_j (d_ij - F_ij) sigma T_j^4
    for (uint32_t j = threadIdx.x; j < np; j += kBlock) {
        const double dij = i == j ? 1.0 : 0.0;
        const double Fij = F[row + j], Ej = E[j];
        s[0] += Fij;
        s[1] -= (dij - Fij) * kStefanBoltzmann * pow4(T[j]);
        val[row + j] = Ej != 0.0 ? dij / Ej - Fij * (1.0 - Ej) / Ej : 0.0;
    }
    block_sum<2, kBlock>(s, lds);
    if (threadIdx.x == 0) b[i] = s[1] + (1.0 - s[0]) * kStefanBoltzmann * pow4(Ta[i]);
}

__global__ void k_rad_cols(uint32_t np, uint64_t nnz, uint32_t *__restrict__ col)
{
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < nnz; k += (uint64_t)gridDim.x * kBlock)
        col[k] = (uint32_t)(k % np);
}

// Patch heat onto the patch elements by area (src/rsolverradiativeheat.cpp:276-303).
__global__ void k_rad_process(const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ elems, uint32_t np,
                              const double *__restrict__ area, const double *__restrict__ heat,
                              double *__restrict__ out)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= np) return;
    const uint32_t k0 = ptr[p], k1 = ptr[p + 1];
    double a = 0.0;
    for (uint32_t k = k0; k < k1; k++) a += area[elems[k]];
    for (uint32_t k = k0; k < k1; k++) {
        const uint32_t e = elems[k];
        out[e] = heat[p] * (area[e] / a);
    }
}

__global__ void k_rad_zero(double *__restrict__ a, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) a[i] = 0.0;
}

static uint32_t rad_blocks(uint64_t n)
{
    const uint64_t want = (n + kBlock - 1) / kBlock;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)std::max(8, runtime().n_cu * 8)));
}

}  // namespace rs

using namespace rs;

struct rs_radiosity {
    uint32_t np = 0, ne = 0;
    DBuf<uint32_t> ptr, elems;
    DBuf<double> area, E, T, Ta;
    rs_matrix A;
    double last_ms = 0.0;
};

extern "C" int rs_radiosity_create(uint32_t n_patches, uint32_t n_nodes, const double *xyz, uint32_t n_elems,
                                   const uint32_t *tris, const uint32_t *patch_ptr, const uint32_t *patch_elems,
                                   rs_radiosity **out)
{
    if (!out || !xyz || !tris || !patch_ptr || n_patches == 0)
        return fail(RS_ERR_INVALID_ARG, "rs_radiosity_create: bad arguments");
    if ((uint64_t)n_patches * n_patches > 0xFFFFFFF0ull)
        return fail(RS_ERR_UNSUPPORTED, "rs_radiosity_create: n_patches^2 exceeds 2^32 (uint32 row pointers)");
    if (patch_ptr[0] != 0) return fail(RS_ERR_INVALID_ARG, "rs_radiosity_create: patch_ptr[0] must be 0");
    for (uint32_t p = 0; p < n_patches; p++)
        if (patch_ptr[p + 1] < patch_ptr[p]) return fail(RS_ERR_INVALID_ARG, "rs_radiosity_create: patch_ptr decreases");
    const uint32_t nref = patch_ptr[n_patches];
    if (nref && !patch_elems) return fail(RS_ERR_INVALID_ARG, "rs_radiosity_create: null patch_elems");
    std::vector<uint8_t> seen(n_elems, 0);
    for (uint32_t k = 0; k < nref; k++) {
        if (patch_elems[k] >= n_elems) return fail(RS_ERR_INVALID_ARG, "rs_radiosity_create: element id out of range");
        if (seen[patch_elems[k]]++) return fail(RS_ERR_INVALID_ARG, "rs_radiosity_create: element in two patches");
    }
    for (uint64_t k = 0; k < 3ull * n_elems; k++)
        if (tris[k] >= n_nodes) return fail(RS_ERR_INVALID_ARG, "rs_radiosity_create: node id out of range");
    int rc = require_runtime();
    if (rc) return rc;
    std::unique_ptr<rs_radiosity> R(new rs_radiosity());
    R->np = n_patches;
    R->ne = n_elems;
    // element areas (RElement::findArea of the surface triangles)
    std::vector<double> area(n_elems);
    for (uint32_t e = 0; e < n_elems; e++) {
        const double *p0 = xyz + 3ull * tris[3ull * e], *p1 = xyz + 3ull * tris[3ull * e + 1],
                     *p2 = xyz + 3ull * tris[3ull * e + 2];
        const double u[3] = { p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2] };
        const double v[3] = { p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2] };
        const double c[3] = { u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0] };
        area[e] = 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    }
    hipStream_t st = runtime().stream;
    if ((rc = R->ptr.alloc(n_patches + 1ull)) || (rc = R->elems.alloc(std::max(nref, 1u))) ||
        (rc = R->area.alloc(std::max(n_elems, 1u))) || (rc = R->E.alloc(n_patches)) || (rc = R->T.alloc(n_patches)) ||
        (rc = R->Ta.alloc(n_patches)))
        return rc;
    RS_HIP(hipMemcpyAsync(R->ptr.p, patch_ptr, sizeof(uint32_t) * (n_patches + 1ull), hipMemcpyHostToDevice, st));
    if (nref) RS_HIP(hipMemcpyAsync(R->elems.p, patch_elems, sizeof(uint32_t) * nref, hipMemcpyHostToDevice, st));
    if (n_elems) RS_HIP(hipMemcpyAsync(R->area.p, area.data(), sizeof(double) * n_elems, hipMemcpyHostToDevice, st));
    rs_matrix &A = R->A;
    A.n = n_patches;
    A.nnz = (uint64_t)n_patches * n_patches;
    A.borrowed = true;
    std::vector<uint32_t> rowptr(n_patches + 1ull);
    for (uint32_t i = 0; i <= n_patches; i++) rowptr[i] = i * n_patches;
    if ((rc = A.rowptr.alloc(rowptr.size())) || (rc = A.col.alloc(A.nnz)) || (rc = A.val.alloc(A.nnz))) return rc;
    RS_HIP(hipMemcpyAsync(A.rowptr.p, rowptr.data(), sizeof(uint32_t) * rowptr.size(), hipMemcpyHostToDevice, st));
    k_rad_cols<<<rad_blocks(A.nnz), kBlock, 0, st>>>(n_patches, A.nnz, A.col.p);
    k_rad_zero<<<rad_blocks(A.nnz), kBlock, 0, st>>>(A.val.p, A.nnz);
    RS_HIP(hipGetLastError());
    if ((rc = matrix_set_layout(A, rowptr.data()))) return rc;
    RS_HIP(hipStreamSynchronize(st));
    *out = R.release();
    return RS_OK;
}

extern "C" int rs_radiosity_matrix(rs_radiosity *R, rs_matrix **A)
{
    if (!R || !A) return fail(RS_ERR_INVALID_ARG, "rs_radiosity_matrix: null");
    *A = &R->A;
    return RS_OK;
}

extern "C" int rs_radiosity_last_ms(const rs_radiosity *R, double *ms)
{
    if (!R || !ms) return fail(RS_ERR_INVALID_ARG, "rs_radiosity_last_ms: null");
    *ms = R->last_ms;
    return RS_OK;
}

extern "C" int rs_assemble_radiosity_device(rs_radiosity *R, const double *d_view_factors,
                                            const double *d_emissivity, const double *d_temperature,
                                            const double *d_ambient, double *d_b)
{
    if (!R || !d_view_factors || !d_emissivity || !d_temperature || !d_ambient || !d_b)
        return fail(RS_ERR_INVALID_ARG, "rs_assemble_radiosity_device: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    EventPair evp;
    RS_HIP(evp.create());
    hipEvent_t *ev = evp.e;
    RS_HIP(hipEventRecord(ev[0], st));
    k_rad_means<<<(R->np + kBlock - 1) / kBlock, kBlock, 0, st>>>(R->ptr.p, R->elems.p, R->np, d_emissivity,
                                                                  d_temperature, d_ambient, R->E.p, R->T.p, R->Ta.p);
    k_rad_rows<<<R->np, kBlock, 0, st>>>(d_view_factors, R->np, R->E.p, R->T.p, R->Ta.p, R->A.val.p, d_b);
    RS_HIP(hipGetLastError());
    RS_HIP(hipEventRecord(ev[1], st));
    RS_HIP(hipEventSynchronize(ev[1]));
    float f = 0.0f;
    RS_HIP(hipEventElapsedTime(&f, ev[0], ev[1]));
    R->last_ms = f;
    if ((rc = matrix_frobenius(R->A))) return rc;
    matrix_values_changed(R->A);
    return RS_OK;
}

extern "C" int rs_radiosity_process_device(rs_radiosity *R, const double *d_patch_heat, double *d_elem_heat)
{
    if (!R || !d_patch_heat || !d_elem_heat) return fail(RS_ERR_INVALID_ARG, "rs_radiosity_process_device: null");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    if (R->ne) k_rad_zero<<<rad_blocks(R->ne), kBlock, 0, st>>>(d_elem_heat, R->ne);
    k_rad_process<<<(R->np + kBlock - 1) / kBlock, kBlock, 0, st>>>(R->ptr.p, R->elems.p, R->np, R->area.p,
                                                                    d_patch_heat, d_elem_heat);
    RS_HIP(hipGetLastError());
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

extern "C" int rs_radiosity_destroy(rs_radiosity *R)
{
    delete R;
    return RS_OK;
}
