// rs_assembly.hip -- deterministic graph-coloured tetrahedral assembly on gfx950.
//
// Reference behaviour restated (Range-Software/range-solver-lib):
//   RSolverHeat::prepare volume loop    src/rsolverheat.cpp:361-455
//   RSolverHeat::assemblyMatrix         src/rsolverheat.cpp:755-820
//   RSolverStress::prepare volume loop  src/rsolverstress.cpp:573-715
//   RSolverStress surface traction      src/rsolverstress.cpp:384-571 (force per unit area)
//   RSolverStress::assemblyMatrix       src/rsolverstress.cpp:1654-1772 (static, no rotation)
//   node write-back                     src/rsolverheat.cpp:473-480, rsolverstress.cpp:800-836
//
// The reference serialises its scatter with `omp critical` around RSparseMatrix::addValue.
// Here the CSR pattern and an element -> CSR-position scatter map are built once; elements
// are sorted by a node-disjoint colour, and each colour is one launch whose threads own
// disjoint rows, so they add into the CSR values and b without atomics and in a fixed
// order (bit-reproducible from run to run).
//
// Element matrices use the closed forms of the linear tetrahedron (constant gradients):
//   heat    Ke = k V grad(N_m).grad(N_n), fe = Q V / 4, Me = rho c V (1 + delta_mn) / 20
//   stress  K_mn[i][j] = V (lambda g_mi g_nj + mu g_mj g_ni + mu delta_ij g_m.g_n)
// which equal the reference's integration-point sums (B De B^T detJ w) exactly for linear
// tets.  Stress uses 16 lanes per element (one 3x3 node block each) with the element's
// vertex coordinates staged in LDS.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "rs_device.hpp"
#include "rs_internal.hpp"

namespace rs {

struct HeatArgs {
    const double *xyz;
    const uint4 *conn;          // colour-sorted elements
    const uint32_t *perm;       // colour-sorted index -> original element id
    const uint32_t *map;        // 16 CSR positions per element (RS_EOD when row/col fixed)
    const uint32_t *book;       // node -> dof
    const double *k, *q, *rhoc; // per original element
    const double *t_node;       // nodeTemperature
    double *val, *b;
    int transient;
    double alpha, dt;
};

__global__ void k_zero(double *__restrict__ a, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) a[i] = 0.0;
}

// Gradients of the 4 linear shape functions and the volume of tet (X0..X3).  No contraction:
// every kernel that inlines it gets the same bits (the geometry records of the row-gather
// assemblies are computed by other kernels than the colour launches they must match).
__device__ __forceinline__ double tet_gradients(const double (&X)[4][3], double (&g)[4][3])
{
#pragma clang fp contract(off)
    double a[3], b[3], c[3];
    for (int i = 0; i < 3; i++) {
        a[i] = X[1][i] - X[0][i];
        b[i] = X[2][i] - X[0][i];
        c[i] = X[3][i] - X[0][i];
    }
    // J = [a b c] (columns); rows of J^-1 are grad N1..N3: (b x c, c x a, a x b) / det
    double bc[3] = { b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0] };
    double ca[3] = { c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0] };
    double ab[3] = { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] };
    double det = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    double inv = 1.0 / det;
    for (int i = 0; i < 3; i++) {
        g[1][i] = bc[i] * inv;
        g[2][i] = ca[i] * inv;
        g[3][i] = ab[i] * inv;
        g[0][i] = -(g[1][i] + g[2][i] + g[3][i]);
    }
    return det / 6.0;
}

// a.b of two 3-vectors with its roundings spelled out (a0 b0, then two fused multiply-adds) and
// no further contraction: the same bits in every kernel that inlines it, whatever the compiler
// would fuse around an open-coded sum in one kernel and not in another.  The row-gather and
// colour-launch assemblies rely on it to add identical element terms.
__device__ __forceinline__ double dot3(const double *a, const double *b)
{
#pragma clang fp contract(off)
    return fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0]));
}

__global__ __launch_bounds__(kBlock) void k_heat_colour(HeatArgs a, uint32_t e0, uint32_t e1)
{
    uint32_t e = e0 + blockIdx.x * kBlock + threadIdx.x;
    if (e >= e1) return;
    uint4 nd4 = a.conn[e];
    uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
    uint32_t orig = a.perm[e];
    double X[4][3], g[4][3];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int i = 0; i < 3; i++) X[m][i] = a.xyz[3 * (uint64_t)nd[m] + i];
    double V = tet_gradients(X, g);
    double kV = a.k[orig] * V;
    double fq = a.q[orig] * V * 0.25;
    double Ae[4][4], be[4];
    if (a.transient) {
        double mc = a.rhoc[orig] * V / 20.0;
        double T[4];
#pragma unroll
        for (int n = 0; n < 4; n++) T[n] = a.t_node[nd[n]];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            be[m] = a.dt * fq;
#pragma unroll
            for (int n = 0; n < 4; n++) {
                double K = kV * dot3(g[m], g[n]);
                double M = mc * (m == n ? 2.0 : 1.0);
                Ae[m][n] = M + a.alpha * a.dt * K;
                be[m] += (M - (1.0 - a.alpha) * a.dt * K) * T[n];
            }
        }
    } else {
#pragma unroll
        for (int m = 0; m < 4; m++) {
            be[m] = fq;
#pragma unroll
            for (int n = 0; n < 4; n++) Ae[m][n] = kV * dot3(g[m], g[n]);
        }
    }
    uint32_t dof[4];
#pragma unroll
    for (int m = 0; m < 4; m++) dof[m] = a.book[nd[m]];
    // Dirichlet lifting (src/rsolverheat.cpp:787-798)
#pragma unroll
    for (int m = 0; m < 4; m++) {
        if (dof[m] != RS_EOD) continue;
        double tm = a.t_node[nd[m]];
#pragma unroll
        for (int n = 0; n < 4; n++) be[n] -= Ae[n][m] * tm;
    }
    const uint4 *mp = reinterpret_cast<const uint4 *>(a.map + 16ull * e);
#pragma unroll
    for (int m = 0; m < 4; m++) {
        if (dof[m] == RS_EOD) continue;
        a.b[dof[m]] += be[m];
        uint4 pos = mp[m];
        uint32_t pp[4] = { pos.x, pos.y, pos.z, pos.w };
#pragma unroll
        for (int n = 0; n < 4; n++)
            if (pp[n] != RS_EOD) a.val[pp[n]] += Ae[m][n];
    }
}

// Row-gather variant: one thread per free DOF row sums the contributions of the elements
// incident to its node, in colour-sorted element order -- the order in which the colour
// launches add into that row, so the values are bit-identical to k_heat_colour's -- and writes
// the row once.
//
// What does not change between assemblies of one mesh is computed once, at create, into one
// record per (row, incident element) ("row records"): the four gradient products
// grad N_m . grad N_n of the row's local node m, the volume, the element's original index (and
// m in its top two bits), the row slots of the element's four columns (one byte each, 0xFF for
// a fixed column) and the four node ids.  Records are laid out like SELL-64: the records of the
// 64 rows of a slice are interleaved (record j of lane l at base + 64 j + l), so every field is
// one coalesced load per wave.  An assembly then streams the records and gathers only the
// coefficients k, q, rho*c and the prescribed temperatures.
constexpr int kRowCap = 16;
constexpr uint32_t kSlotFixed = 0xFFu;

struct RowRecords {
    const double *grad;         // 4 per record: grad N_m . grad N_n, n = 0..3
    const double *vol;
    const uint32_t *elem;       // original element index | m << 30
    const uint32_t *slots;      // byte n: slot of column n in the row, kSlotFixed if fixed
    const uint4 *nodes;
    const uint64_t *base;       // per slice: first record
    const uint32_t *count;      // per row: records
};

typedef double rec_d2 __attribute__((ext_vector_type(2)));

struct HeatRowArgs {
    HeatArgs h;
    const uint32_t *rowptr;
    RowRecords rec;
    uint32_t n_rows;
};

// Records of every free row (once per mesh): the same tet_gradients and products as the
// assembly kernels, so the assembled values do not change.
__global__ __launch_bounds__(kBlock) void k_heat_records(HeatArgs h, const uint32_t *__restrict__ rowptr,
                                                         const uint32_t *__restrict__ dof_node,
                                                         const uint32_t *__restrict__ inc_ptr,
                                                         const uint32_t *__restrict__ inc, uint32_t n_rows,
                                                         double *__restrict__ grad, double *__restrict__ vol,
                                                         uint32_t *__restrict__ elem, uint32_t *__restrict__ slots,
                                                         uint4 *__restrict__ nodes, const uint64_t *__restrict__ base,
                                                         uint32_t *__restrict__ count)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_rows) return;
    const uint32_t r0 = rowptr[i];
    const uint32_t v = dof_node[i];
    const uint32_t k0 = inc_ptr[v], k1 = inc_ptr[v + 1];
    const uint64_t b = base[i / kWave] + (i % kWave);
    count[i] = k1 - k0;
    for (uint32_t k = k0; k < k1; k++) {
        const uint32_t w = inc[k], e = w >> 2, m = w & 3u;
        const uint4 nd4 = h.conn[e];
        const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
        double X[4][3], g[4][3];
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) X[q][c] = h.xyz[3 * (uint64_t)nd[q] + c];
        const double V = tet_gradients(X, g);
        const uint64_t r = b + (uint64_t)(k - k0) * kWave;
        const uint4 pos = reinterpret_cast<const uint4 *>(h.map + 16ull * e)[m];
        const uint32_t pp[4] = { pos.x, pos.y, pos.z, pos.w };
        uint32_t sl = 0;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            grad[4 * r + n] = dot3(g[m], g[n]);
            sl |= (pp[n] == RS_EOD ? kSlotFixed : (pp[n] - r0)) << (8 * n);
        }
        vol[r] = V;
        elem[r] = h.perm[e] | (m << 30);
        slots[r] = sl;
        nodes[r] = nd4;
    }
}

__global__ __launch_bounds__(kBlock) void k_heat_rows(HeatRowArgs a)
{
    __shared__ double acc[kRowCap][kBlock];
    const uint32_t t = threadIdx.x;
    // neighbouring row ranges on one XCD: the k, q entries that rows of adjacent node lines
    // gather stay in that XCD's L2
    const uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * kBlock + t;
    if (i >= a.n_rows) return;
    const HeatArgs &h = a.h;
    const RowRecords &R = a.rec;
    const uint32_t r0 = a.rowptr[i], len = a.rowptr[i + 1] - r0;
#pragma unroll
    for (int j = 0; j < kRowCap; j++) acc[j][t] = 0.0;
    double bi = 0.0;
    const uint64_t b = R.base[i / kWave] + (i % kWave);
    const uint32_t cnt = R.count[i];
    // one record: its fields, loaded before any of them is used
    struct Rec {
        double gg[4], V, kk, qq;
        uint32_t orig, m, sl;
    };
    auto load = [&](uint32_t k, Rec &c) {
        const uint64_t r = b + (uint64_t)k * kWave;
        // the records stream once per assembly (non-temporal), so the gathered k and q lines,
        // each shared by the rows of an element's four nodes, keep their place in L2
        const rec_d2 g01 = __builtin_nontemporal_load(reinterpret_cast<const rec_d2 *>(R.grad) + 2 * r);
        const rec_d2 g23 = __builtin_nontemporal_load(reinterpret_cast<const rec_d2 *>(R.grad) + 2 * r + 1);
        c.gg[0] = g01.x;
        c.gg[1] = g01.y;
        c.gg[2] = g23.x;
        c.gg[3] = g23.y;
        c.V = __builtin_nontemporal_load(R.vol + r);
        const uint32_t em = __builtin_nontemporal_load(R.elem + r);
        c.orig = em & 0x3FFFFFFFu;
        c.m = em >> 30;
        c.sl = __builtin_nontemporal_load(R.slots + r);
    };
    auto process = [&](uint32_t k, const Rec &c) {
        const uint64_t r = b + (uint64_t)k * kWave;
        const double V = c.V, kV = c.kk * V, fq = c.qq * V * 0.25;
        const uint32_t sl = c.sl, m = c.m;
        double Ae[4], be;
        const bool lift = ((sl & 0xFFu) == kSlotFixed) | (((sl >> 8) & 0xFFu) == kSlotFixed) |
                          (((sl >> 16) & 0xFFu) == kSlotFixed) | ((sl >> 24) == kSlotFixed);
        uint4 nd4 = make_uint4(0, 0, 0, 0);
        if (h.transient || lift) nd4 = R.nodes[r];
        const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
        if (h.transient) {
            const double mc = h.rhoc[c.orig] * V / 20.0;
            be = h.dt * fq;
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const double K = kV * c.gg[n];
                const double M = mc * ((uint32_t)n == m ? 2.0 : 1.0);
                Ae[n] = M + h.alpha * h.dt * K;
                be += (M - (1.0 - h.alpha) * h.dt * K) * h.t_node[nd[n]];
            }
        } else {
            be = fq;
#pragma unroll
            for (int n = 0; n < 4; n++) Ae[n] = kV * c.gg[n];
        }
        // Dirichlet lifting (src/rsolverheat.cpp:787-798), in the colour kernel's order
#pragma unroll
        for (int n = 0; n < 4; n++)
            if (((sl >> (8 * n)) & 0xFFu) == kSlotFixed) be -= Ae[n] * h.t_node[nd[n]];
        bi += be;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const uint32_t slot = (sl >> (8 * n)) & 0xFFu;
            if (slot != kSlotFixed) acc[slot][t] += Ae[n];
        }
    };
    // two records per trip: both records and both coefficient gathers are in flight before the
    // first is processed; the sums still run in record order
    for (uint32_t k = 0; k < cnt; k += 2) {
        const bool two = k + 1 < cnt;
        Rec c0, c1;
        load(k, c0);
        load(two ? k + 1 : k, c1);
        c0.kk = h.k[c0.orig];
        c0.qq = h.q[c0.orig];
        c1.kk = h.k[c1.orig];
        c1.qq = h.q[c1.orig];
        process(k, c0);
        if (two) process(k + 1, c1);
    }
    for (uint32_t j = 0; j < len; j++) h.val[r0 + j] = acc[j][t];
    h.b[i] = bi;
}

// ------------------------------------------------------------------ post-processing
// The caller's element order of the connectivity (the assembly keeps it colour-sorted), for
// the element-local post-processing: conn0[perm[s]] = conn[s].
__global__ void k_unsort_conn(const uint4 *__restrict__ conn, const uint32_t *__restrict__ perm, uint32_t ne,
                              uint4 *__restrict__ conn0)
{
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < ne; s += gridDim.x * kBlock) conn0[perm[s]] = conn[s];
}

// RSolverHeat::process, volume elements (src/rsolverheat.cpp:646-692): per tet the reference
// integrates grad N_m * detJ / nInp over its points, so q = -detJ k grad T (detJ = 6 V, the
// signed determinant of the reference-tet map).  One thread per element in the caller's order:
// connectivity, conductivity and flux stream coalesced.
__global__ __launch_bounds__(kBlock) void k_heat_flux(const double *__restrict__ xyz, const uint4 *__restrict__ conn0,
                                                      uint32_t ne, const double *__restrict__ k,
                                                      const double *__restrict__ t_node, double *__restrict__ flux)
{
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    const uint4 nd4 = conn0[e];
    const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
    double X[4][3], g[4][3];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int c = 0; c < 3; c++) X[m][c] = xyz[3 * (uint64_t)nd[m] + c];
    const double detJ = 6.0 * tet_gradients(X, g);
    const uint32_t orig = e;
    const double ke = k[orig];
    double Q[3] = { 0.0, 0.0, 0.0 };
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const double T = t_node[nd[m]];
#pragma unroll
        for (int c = 0; c < 3; c++) Q[c] -= g[m][c] * detJ * ke * T;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) flux[3 * (uint64_t)orig + c] = Q[c];
}

// RSolverMagnetostatics::prepare + assemblyMatrix (src/rsolvermagnetostatics.cpp:44-145,
// 234-287): three copies of the negative Laplacian, Ke[3m+c][3n+c] = -(g_m . g_n) V, source
// fe[3m+c] = e0 V / 4 (g_m x J_m)_c with the node current density J, lifting of the prescribed
// field on fixed nodes.  One launch per colour, added into the 3-DOF CSR through the scatter map.
struct MagnetoArgs {
    const double *xyz;
    const uint4 *conn;
    const uint32_t *map;        // 144 positions per element, [(4m+n)*9 + 3i+j]
    const uint32_t *book;       // 3*node+c -> dof
    const double *current, *field;
    double *val, *b;
};

// The magnetostatic element's terms, without contraction (shared by the colour launches and
// the row gather): K_mn = -(grad N_m . grad N_n) V on the three component diagonals, and the
// source of node m, component c: V/4 eps0 (curl-type product of grad N_m with J_m).
__device__ __forceinline__ double magneto_k(const double *gm, const double *gn, double V)
{
#pragma clang fp contract(off)
    return -(dot3(gm, gn) * V);
}
__device__ __forceinline__ void magneto_source(const double *gm, const double *Jm, double V, double (&fe)[3])
{
#pragma clang fp contract(off)
    const double fv = V * 0.25 * 8.854187817e-12;
    fe[0] = fv * (-gm[2] * Jm[1] + gm[1] * Jm[2]);
    fe[1] = fv * (gm[2] * Jm[0] - gm[0] * Jm[2]);
    fe[2] = fv * (-gm[1] * Jm[0] + gm[0] * Jm[1]);
}

__global__ __launch_bounds__(kBlock) void k_magneto_colour(MagnetoArgs a, uint32_t e0, uint32_t e1)
{
#pragma clang fp contract(off)
    const uint32_t e = e0 + blockIdx.x * kBlock + threadIdx.x;
    if (e >= e1) return;
    const uint4 nd4 = a.conn[e];
    const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
    double X[4][3], g[4][3];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int c = 0; c < 3; c++) X[m][c] = a.xyz[3 * (uint64_t)nd[m] + c];
    const double V = tet_gradients(X, g);
    double K[4][4], fe[12];
#pragma unroll
    for (int m = 0; m < 4; m++) {
#pragma unroll
        for (int n = 0; n < 4; n++) K[m][n] = magneto_k(g[m], g[n], V);
        double f3[3];
        magneto_source(g[m], a.current + 3 * (uint64_t)nd[m], V, f3);
#pragma unroll
        for (int c = 0; c < 3; c++) fe[3 * m + c] = f3[c];
    }
    bool free_[4];
#pragma unroll
    for (int m = 0; m < 4; m++) free_[m] = a.book[3 * (uint64_t)nd[m]] != RS_EOD;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        if (free_[m]) continue;
        const double *H = a.field + 3 * (uint64_t)nd[m];
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int c = 0; c < 3; c++) fe[3 * n + c] -= K[n][m] * H[c];
    }
    const uint32_t *mp = a.map + 144ull * e;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        if (!free_[m]) continue;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            a.b[a.book[3 * (uint64_t)nd[m] + c]] += fe[3 * m + c];
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const uint32_t pos = mp[(4 * m + n) * 9 + 4 * c];
                if (pos != RS_EOD) a.val[pos] += K[m][n];
            }
        }
    }
}

// RSolverFluid (src/rsolverfluid.cpp): the stabilised tetrahedron of
// computeElementConstantDerivative (:1845-2201) -- SUPG / PSPG parameters switching at element
// Reynolds number 3 (:1947-1983), LSIC |v| h / 2 (:1986) -- with the element averages of
// prepare (:306-331) and h_n = cbrt(6 V / pi) (computeElementScales :2235-2268), scattered
// into the (vx, vy, vz, p) CSR as assemblyMatrix does (:2399-2419: enabled positions only, no
// lifting, the system is for increments).
//
// Sixteen lanes per element, one per node pair (m, n).  Each lane rebuilds the element
// quantities (the 16 lanes of an element read the same addresses, served as broadcasts) and
// owns the 4x4 block (m, n) of Ae; the lanes with n = 0 own node m's four right-hand-side
// rows.  The linear tetrahedron's integration constants in closed form: sum w N_m = 1/24,
// sum w N_m N_n = (1 + d_mn) / 120, sum w = 1/6, detJ = 6 V.
struct FluidArgs {
    const double *xyz;
    const uint4 *conn;          // colour-sorted elements
    const uint32_t *perm;       // colour-sorted index -> original element id
    const uint32_t *map;        // 64 per element: [(4m+n)*4 + i] = CSR position of node n's first
                                // enabled column in row (m, i), RS_EOD if none or the row is disabled
    const uint4 *book;          // (vx, vy, vz, p) positions of a node -> dof
    const double *rho, *mu, *grav;   // per original element (grav: 3 per element)
    const double *state;        // 4 per node: vx vy vz p
    const double *v_old, *acc;  // 3 per node
    double *val, *b;
    double stream_velocity, alpha, dt;
    int unsteady;
};

constexpr uint32_t kFluidLanes = 16;
constexpr double kPi = 3.14159265358979323846;

__global__ __launch_bounds__(kBlock) void k_fluid_colour(FluidArgs a, uint32_t e0, uint32_t e1)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t e = e0 + t / kFluidLanes;
    if (e >= e1) return;
    const uint32_t lane = t % kFluidLanes, m = lane >> 2, n = lane & 3u;
    const uint4 nd4 = a.conn[e];
    const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
    const uint32_t orig = a.perm[e];
    double X[4][3], B[4][3];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int c = 0; c < 3; c++) X[q][c] = a.xyz[3 * (uint64_t)nd[q] + c];
    const double V = tet_gradients(X, B);
    const double detJ = 6.0 * V;
    const double wt = 1.0 / 6.0, iN = 1.0 / 24.0, inv6 = 1.0 / 6.0;
    const double ro = a.rho[orig], invro = 1.0 / ro, u = a.mu[orig];
    const double hn = cbrt(6.0 * fabs(V) / kPi);
    double st[4][4], ve[3] = { 0, 0, 0 }, veo[3] = { 0, 0, 0 }, p = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int c = 0; c < 4; c++) st[q][c] = a.state[4 * (uint64_t)nd[q] + c];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            ve[c] += st[q][c];
            veo[c] += a.v_old[3 * (uint64_t)nd[q] + c];
        }
        p += st[q][3];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) { ve[c] *= 0.25; veo[c] *= 0.25; }
    p *= 0.25;
    const double mvh = sqrt(veo[0] * veo[0] + veo[1] * veo[1] + veo[2] * veo[2]);
    double sdir[3] = { 0, 0, 0 };
    if (mvh > 0.0) {
#pragma unroll
        for (int c = 0; c < 3; c++) sdir[c] = veo[c] / mvh;
    }
    // G[a][c] = d v_c / d x_a, gp = grad p, vdiv[q] = ve . grad N_q
    double vdiv[4], G[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } }, gp[3] = { 0, 0, 0 }, h = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        vdiv[q] = ve[0] * B[q][0] + ve[1] * B[q][1] + ve[2] * B[q][2];
#pragma unroll
        for (int d = 0; d < 3; d++) {
#pragma unroll
            for (int c = 0; c < 3; c++) G[d][c] += B[q][d] * st[q][c];
            gp[d] += B[q][d] * st[q][3];
        }
        h += fabs(sdir[0] * B[q][0] + sdir[1] * B[q][1] + sdir[2] * B[q][2]);
    }
    if (h != 0.0) h = 2.0 / h;
    const double roD2u = 0.5 * ro / u;
    const double Re = roD2u * mvh * h, Ren = roD2u * a.stream_velocity * hn;
    double Tsupg = 0.0, Tpspg = 0.0;
    if (mvh > 0.0) Tsupg = (Re > 0.0 && Re <= 3.0) ? h * Re * (1.0 / mvh) * inv6 : h * (1.0 / mvh) * 0.5;
    if (a.stream_velocity > 0.0) {
        const double inv = 1.0 / a.stream_velocity;
        Tpspg = (Ren > 0.0 && Ren <= 3.0) ? hn * Ren * inv * inv6 : hn * inv * 0.5;
    }
    const double Tlsic = mvh * h * 0.5;
    double vv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) vv[c] = ve[0] * G[0][c] + ve[1] * G[1][c] + ve[2] * G[2][c];
    const double divV = G[0][0] + G[1][1] + G[2][2];
    const bool unsteady = a.unsteady != 0;
    const double dt = a.dt, alphaDt = a.alpha * a.dt;

    // ---- block (m, n) of Ae
    const double iNiN = (m == n ? 2.0 : 1.0) / 120.0;
    const double BB = B[m][0] * B[n][0] + B[m][1] * B[n][1] + B[m][2] * B[n][2];
    const double kScale = ro * u * wt, cpeScale = ro * iNiN;
    const double kte = Tsupg * ro * vdiv[m] * vdiv[n] * wt;
    const double ktpe = Tsupg * ro * vdiv[m] * iN, ktppe = Tsupg * ro * iN;
    const double ytpe = Tsupg * iN, epe = Tlsic * ro * wt;
    const double ce = ro * iN * vdiv[n];
    double an[3] = { 0, 0, 0 };
    if (unsteady) {
#pragma unroll
        for (int c = 0; c < 3; c++) an[c] = a.acc[3 * (uint64_t)nd[n] + c];
    }
    double blk[4][4];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double Bmi_nj = B[m][i] * B[n][j];
            const double Bmj_ni = B[m][j] * B[n][i];
            const double k = kScale * ((i == j ? BB : 0.0) + Bmj_ni);
            const double cpe = cpeScale * G[j][i];
            const double steady = cpe + k + (i == j ? kte : 0.0) + ktpe * G[j][i] + ktppe * B[m][j] * vv[i]
                                + ytpe * B[m][j] * gp[i] + Bmi_nj * epe;
            double v = 0.0;
            if (unsteady) {
                if (i == j) v += ro * iNiN + Tsupg * ro * vdiv[m] * iN;
                v += alphaDt * (Tsupg * ro * iN * B[m][j] * an[i]);
                if (i == j) v += alphaDt * ce;
                v += alphaDt * (steady - Bmi_nj * epe) + dt * Bmi_nj * epe;
            } else {
                if (i == j) v += ce;
                v += steady;
            }
            blk[i][j] = v * detJ;
        }
    const double yte = Tsupg * vdiv[m] * wt, s12 = unsteady ? -dt : -1.0;
#pragma unroll
    for (int i = 0; i < 3; i++) blk[i][3] = s12 * (B[m][i] * iN + yte * B[n][i]) * detJ;
    const double bte = Tpspg * iN, ye = Tpspg * vdiv[n] * wt, ype = Tpspg * iN;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const double ypej = ype * (B[m][0] * G[j][0] + B[m][1] * G[j][1] + B[m][2] * G[j][2]);
        const double v = unsteady ? bte * B[m][j] + dt * (B[n][j] * iN) + alphaDt * (ye * B[m][j] + ypej)
                                  : B[n][j] * iN + ye * B[m][j] + ypej;
        blk[3][j] = v * detJ;
    }
    const double theta = Tpspg * (BB * invro) * wt;
    blk[3][3] = (unsteady ? dt * theta : theta) * detJ;

    const uint4 bk4 = a.book[nd[n]];
    const uint32_t bn[4] = { bk4.x, bk4.y, bk4.z, bk4.w };
    const uint32_t *mp = a.map + 64ull * e + (4 * m + n) * 4;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t pos = mp[i];
        if (pos == RS_EOD) continue;
        uint32_t k = 0;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (bn[j] != RS_EOD) a.val[pos + k++] += blk[i][j];
    }

    // ---- right-hand side of node m (:2150-2196)
    if (n != 0) return;
    double am[3] = { 0, 0, 0 };
    if (unsteady) {
#pragma unroll
        for (int c = 0; c < 3; c++) am[c] = a.acc[3 * (uint64_t)nd[m] + c];
    }
    const double *g = a.grav + 3 * (uint64_t)orig;
    const double gr[3] = { g[0], g[1], g[2] };
    const double vdivm = vdiv[m];
    double be[4];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double cv = iN * ro * vv[i];
        const double kv = u * (B[m][0] * G[0][i] + B[m][1] * G[1][i] + B[m][2] * G[2][i]
                              + B[m][0] * G[i][0] + B[m][1] * G[i][1] + B[m][2] * G[i][2]) * wt;
        const double gv = B[m][i] * p * wt;
        const double ktv = Tsupg * ro * vdivm * wt * vv[i];
        const double ytv = Tsupg * vdivm * wt * gp[i];
        const double ev = Tlsic * ro * divV * wt * B[m][i];
        const double fv = ro * iN * gr[i] + Tsupg * ro * vdivm * wt * gr[i];
        const double res = cv + kv - gv + ktv - ytv + ev;
        be[i] = unsteady ? dt * (fv - (iN * ro * am[i] + Tsupg * ro * vdivm * wt * am[i] + res)) : fv - res;
    }
    const double Bg = B[m][0] * gr[0] + B[m][1] * gr[1] + B[m][2] * gr[2];
    const double Ba = B[m][0] * am[0] + B[m][1] * am[1] + B[m][2] * am[2];
    const double Bv = B[m][0] * vv[0] + B[m][1] * vv[1] + B[m][2] * vv[2];
    const double Bp = B[m][0] * gp[0] + B[m][1] * gp[1] + B[m][2] * gp[2];
    be[3] = unsteady ? dt * (Tpspg * Bg * wt - (Tpspg * Ba * wt + iN * divV + Tpspg * Bv * wt + Tpspg * (Bp * invro) * wt))
                     : Tpspg * Bg * wt - (iN * divV + Tpspg * Bv * wt + Tpspg * (Bp * invro) * wt);
    const uint4 bm4 = a.book[nd[m]];
    const uint32_t bm[4] = { bm4.x, bm4.y, bm4.z, bm4.w };
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (bm[i] != RS_EOD) a.b[bm[i]] += be[i] * detJ;
}

// RSolverFluid::solve write-back (src/rsolverfluid.cpp:519-624): state += x through the node
// book; a steady run keeps the pre-update velocities as v_old, an unsteady one sets the
// acceleration (v - v_old) / dt.  Per block: sums of |v|^2 and p^2 before and after, in a fixed
// order (the host adds the block partials in block order).
__global__ __launch_bounds__(kBlock) void k_fluid_update(const uint4 *__restrict__ book, uint32_t nn,
                                                         const double *__restrict__ x, int unsteady, double inv_dt,
                                                         double *__restrict__ state, double *__restrict__ v_old,
                                                         double *__restrict__ acc, double *__restrict__ part)
{
    __shared__ double lds[4 * kBlock / kWave];
    double s[4] = { 0, 0, 0, 0 };
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nn; i += gridDim.x * kBlock) {
        const uint4 bk4 = book[i];
        const uint32_t bk[4] = { bk4.x, bk4.y, bk4.z, bk4.w };
        double v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) v[c] = state[4 * (uint64_t)i + c];
        s[0] += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        s[2] += v[3] * v[3];
        if (!unsteady) {
#pragma unroll
            for (int c = 0; c < 3; c++) v_old[3 * (uint64_t)i + c] = v[c];
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (bk[c] != RS_EOD) v[c] += x[bk[c]];
#pragma unroll
        for (int c = 0; c < 4; c++) state[4 * (uint64_t)i + c] = v[c];
        if (unsteady) {
#pragma unroll
            for (int c = 0; c < 3; c++) acc[3 * (uint64_t)i + c] = (v[c] - v_old[3 * (uint64_t)i + c]) * inv_dt;
        }
        s[1] += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        s[3] += v[3] * v[3];
    }
    block_sum<4, kBlock>(s, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) part[4 * blockIdx.x + k] = s[k];
    }
}

// RSolverFluidHeat / RSolverFluidParticle (src/rsolverfluidheat.cpp:669-880,
// src/rsolverfluidparticle.cpp:552-763): SUPG advection-diffusion of a node field by the element
// velocity -- Galerkin advection ca iN_m (v . grad N_n), the reference's negated diffusion
// -k (grad N_m . grad N_n), streamline diffusion tau ca (v . grad N_m)(v . grad N_n) and the
// SUPG term tau (v . grad N_m) that does not depend on n; tau = h / (2 |v|) min(Pe / 3, 1) --
// with the heat-style lifting of fixed nodes.  Same layout as the heat colour kernel.
struct TransportArgs {
    const double *xyz;
    const uint4 *conn;
    const uint32_t *perm;
    const uint32_t *map;
    const uint32_t *book;
    const double *ca, *k, *vel;     // per original element (vel: 3 per element)
    const double *q, *field;        // per node
    double *val, *b;
    double alpha, dt;
    int unsteady;
};

// The transport element's row-independent quantities (RSolverFluidHeat, src/rsolverfluidheat.cpp:
// 669-880): v.B_n, the stream-line length h = 2 / sum_m |s.B_m| and the SUPG parameter, with
// the roundings of dot3 and no contraction, shared by the colour and row-gather kernels.
struct TransportElem {
    double vdiv[4], Tsupg;
};
__device__ __forceinline__ TransportElem transport_elem(const double (&B)[4][3], const double (&ve)[3], double k)
{
#pragma clang fp contract(off)
    TransportElem t;
    const double mvh = sqrt(dot3(ve, ve));
    double s[3] = { 0, 0, 0 };
    if (mvh > 0.0) {
#pragma unroll
        for (int c = 0; c < 3; c++) s[c] = ve[c] / mvh;
    }
    double h = 0.0;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        t.vdiv[m] = dot3(ve, B[m]);
        h += fabs(dot3(s, B[m]));
    }
    if (h != 0.0) h = 2.0 / h;
    const double Re = k == 0.0 ? 0.0 : mvh * h / (2.0 * k);
    t.Tsupg = 0.0;
    if (mvh > 0.0) {
        t.Tsupg = h / (2.0 * mvh);
        if (Re > 0.0 && Re <= 3.0) t.Tsupg *= Re / 3.0;
    }
    return t;
}

// Entry (m, n) of the element matrix and its right-hand-side term, without contraction:
// ss = ca iN vd_n - k wt (B_m.B_n) + Tsupg ca vd_m vd_n wt + Tsupg vd_m wt; A = (me + a dt ss) detJ
// (steady: ss detJ); rhs term (me - (1 - a) dt ss) T_n.
__device__ __forceinline__ double transport_entry(double ca, double k, double vdm, double vdn, double gmn, double Tsupg,
                                                  bool unsteady, bool diag, double alpha, double dt, double detJ,
                                                  double Tn, double &rhs)
{
#pragma clang fp contract(off)
    const double wt = 1.0 / 6.0, iN = 1.0 / 24.0;
    const double me = unsteady ? ca * ((diag ? 2.0 : 1.0) / 120.0) : 0.0;
    const double ss = ca * iN * vdn - k * wt * gmn + Tsupg * ca * vdm * vdn * wt + Tsupg * vdm * wt;
    rhs = unsteady ? (me - (1.0 - alpha) * dt * ss) * Tn : 0.0;
    return (unsteady ? me + alpha * dt * ss : ss) * detJ;
}

__global__ __launch_bounds__(kBlock) void k_transport_colour(TransportArgs a, uint32_t e0, uint32_t e1)
{
#pragma clang fp contract(off)
    const uint32_t e = e0 + blockIdx.x * kBlock + threadIdx.x;
    if (e >= e1) return;
    const uint4 nd4 = a.conn[e];
    const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
    const uint32_t orig = a.perm[e];
    double X[4][3], B[4][3];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int c = 0; c < 3; c++) X[q][c] = a.xyz[3 * (uint64_t)nd[q] + c];
    const double V = tet_gradients(X, B);
    const double detJ = 6.0 * V, iN = 1.0 / 24.0;
    const double ca = a.ca[orig], k = a.k[orig];
    const double ve[3] = { a.vel[3 * (uint64_t)orig], a.vel[3 * (uint64_t)orig + 1], a.vel[3 * (uint64_t)orig + 2] };
    const TransportElem te = transport_elem(B, ve, k);
    const bool unsteady = a.unsteady != 0;
    double T[4];
#pragma unroll
    for (int n = 0; n < 4; n++) T[n] = a.field[nd[n]];
    double Ae[4][4], be[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const double fv = a.q[nd[m]] * iN;
        double sum = 0.0;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            double rhs;
            Ae[m][n] = transport_entry(ca, k, te.vdiv[m], te.vdiv[n], dot3(B[m], B[n]), te.Tsupg, unsteady, m == n,
                                       a.alpha, a.dt, detJ, T[n], rhs);
            if (unsteady) sum += rhs;
        }
        be[m] = (unsteady ? a.dt * fv + sum : fv) * detJ;
    }
    uint32_t dof[4];
#pragma unroll
    for (int m = 0; m < 4; m++) dof[m] = a.book[nd[m]];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        if (dof[m] != RS_EOD) continue;
#pragma unroll
        for (int n = 0; n < 4; n++) be[n] -= Ae[n][m] * T[m];
    }
    const uint4 *mp = reinterpret_cast<const uint4 *>(a.map + 16ull * e);
#pragma unroll
    for (int m = 0; m < 4; m++) {
        if (dof[m] == RS_EOD) continue;
        a.b[dof[m]] += be[m];
        const uint4 pos = mp[m];
        const uint32_t pp[4] = { pos.x, pos.y, pos.z, pos.w };
#pragma unroll
        for (int n = 0; n < 4; n++)
            if (pp[n] != RS_EOD) a.val[pp[n]] += Ae[m][n];
    }
}

// Row gather of the transport element, on the heat assembly's row records (the gradient
// products, volume, element and column slots of every (row, incident element) pair, once per
// mesh).  Per assembly, k_transport_elem first computes what does not depend on the row, in the
// caller's element order: v.B_n (n = 0..3), the SUPG parameter, ca and k -- 64 bytes per element;
// k_transport_rows then sums, per free row, its entries over the incident elements in colour
// order with k_transport_colour's expressions, and writes the row and b_i once.  Same terms in
// the same order as the colour launches, so the same bits.
constexpr int kTElemWords = 8;

__global__ __launch_bounds__(kBlock) void k_transport_elem(TransportArgs a, uint32_t ne, double *__restrict__ telem)
{
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < ne; e += gridDim.x * kBlock) {
        const uint4 nd4 = a.conn[e];
        const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
        const uint32_t orig = a.perm[e];
        double X[4][3], B[4][3];
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) X[q][c] = a.xyz[3 * (uint64_t)nd[q] + c];
        (void)tet_gradients(X, B);
        const double ca = a.ca[orig], k = a.k[orig];
        const double ve[3] = { a.vel[3 * (uint64_t)orig], a.vel[3 * (uint64_t)orig + 1], a.vel[3 * (uint64_t)orig + 2] };
        const TransportElem te = transport_elem(B, ve, k);
        double4 *o = reinterpret_cast<double4 *>(telem + (uint64_t)kTElemWords * orig);
        o[0] = make_double4(te.vdiv[0], te.vdiv[1], te.vdiv[2], te.vdiv[3]);
        o[1] = make_double4(te.Tsupg, ca, k, 0.0);
    }
}

struct TransportRowArgs {
    const uint32_t *rowptr, *dof_node;
    RowRecords rec;
    uint32_t n_rows;
    const double *telem;
    const double *q, *field;        // per node
    double *val, *b;
    double alpha, dt;
    int unsteady;
};

__global__ __launch_bounds__(kBlock) void k_transport_rows(TransportRowArgs a)
{
#pragma clang fp contract(off)
    __shared__ double acc[kRowCap][kBlock];
    const uint32_t t = threadIdx.x;
    const uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * kBlock + t;
    if (i >= a.n_rows) return;
    const RowRecords &R = a.rec;
    const uint32_t r0 = a.rowptr[i], len = a.rowptr[i + 1] - r0;
#pragma unroll
    for (int j = 0; j < kRowCap; j++) acc[j][t] = 0.0;
    const bool unsteady = a.unsteady != 0;
    const double iN = 1.0 / 24.0;
    const double fv = a.q[a.dof_node[i]] * iN;
    double bi = 0.0;
    const uint64_t b = R.base[i / kWave] + (i % kWave);
    const uint32_t cnt = R.count[i];
    for (uint32_t kk = 0; kk < cnt; kk++) {
        const uint64_t r = b + (uint64_t)kk * kWave;
        const rec_d2 g01 = __builtin_nontemporal_load(reinterpret_cast<const rec_d2 *>(R.grad) + 2 * r);
        const rec_d2 g23 = __builtin_nontemporal_load(reinterpret_cast<const rec_d2 *>(R.grad) + 2 * r + 1);
        const double gg[4] = { g01.x, g01.y, g23.x, g23.y };
        const double detJ = 6.0 * __builtin_nontemporal_load(R.vol + r);
        const uint32_t em = __builtin_nontemporal_load(R.elem + r), orig = em & 0x3FFFFFFFu, m = em >> 30;
        const uint32_t sl = __builtin_nontemporal_load(R.slots + r);
        const double4 *te = reinterpret_cast<const double4 *>(a.telem + (uint64_t)kTElemWords * orig);
        const double4 t0 = te[0], t1 = te[1];
        const double vdiv[4] = { t0.x, t0.y, t0.z, t0.w };
        const double Tsupg = t1.x, ca = t1.y, k = t1.z;
        const double vm = m == 0 ? vdiv[0] : m == 1 ? vdiv[1] : m == 2 ? vdiv[2] : vdiv[3];
        const bool lift = ((sl & 0xFFu) == kSlotFixed) | (((sl >> 8) & 0xFFu) == kSlotFixed) |
                          (((sl >> 16) & 0xFFu) == kSlotFixed) | ((sl >> 24) == kSlotFixed);
        double T[4] = { 0.0, 0.0, 0.0, 0.0 };
        if (unsteady || lift) {
            const uint4 nd4 = R.nodes[r];
            T[0] = a.field[nd4.x];
            T[1] = a.field[nd4.y];
            T[2] = a.field[nd4.z];
            T[3] = a.field[nd4.w];
        }
        double Ae[4], sum = 0.0;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            double rhs;
            Ae[n] = transport_entry(ca, k, vm, vdiv[n], gg[n], Tsupg, unsteady, (uint32_t)n == m, a.alpha, a.dt, detJ,
                                    T[n], rhs);
            if (unsteady) sum += rhs;
        }
        double be = (unsteady ? a.dt * fv + sum : fv) * detJ;
        // lifting of the prescribed columns, in the colour kernel's order
#pragma unroll
        for (int n = 0; n < 4; n++)
            if (((sl >> (8 * n)) & 0xFFu) == kSlotFixed) be -= Ae[n] * T[n];
        bi += be;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const uint32_t slot = (sl >> (8 * n)) & 0xFFu;
            if (slot != kSlotFixed) acc[slot][t] += Ae[n];
        }
    }
    for (uint32_t j = 0; j < len; j++) a.val[r0 + j] = acc[j][t];
    a.b[i] = bi;
}

// RSolverFluidHeat::solve / RSolverFluidParticle::solve write-back: field = max(x, 0) on the
// enabled nodes; block partials of |field|^2 before and after.
__global__ __launch_bounds__(kBlock) void k_transport_update(const uint32_t *__restrict__ book, uint32_t nn,
                                                             const double *__restrict__ x,
                                                             double *__restrict__ field, double *__restrict__ part)
{
    __shared__ double lds[2 * kBlock / kWave];
    double s[2] = { 0, 0 };
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nn; i += gridDim.x * kBlock) {
        double f = field[i];
        s[0] += f * f;
        const uint32_t d = book[i];
        if (d != RS_EOD) {
            f = fmax(x[d], 0.0);
            field[i] = f;
        }
        s[1] += f * f;
    }
    block_sum<2, kBlock>(s, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s[0];
        part[2 * blockIdx.x + 1] = s[1];
    }
}

// RSolverElectrostatics (src/rsolverelectrostatics.cpp): the volume assembly (:300-345) is
// the heat form with coefficient eps_r * e0 and source -rho (fe -= rho N detJ w); e0 as
// RSolverGeneric::e0 (src/rsolvergeneric.cpp:13).
constexpr double kVacuumPermittivity = 8.854187817e-12;

__global__ void k_electro_coeffs(const double *__restrict__ eps_r, const double *__restrict__ rho, uint32_t ne,
                                 double *__restrict__ k, double *__restrict__ q)
{
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < ne; e += gridDim.x * kBlock) {
        k[e] = eps_r[e] * kVacuumPermittivity;
        q[e] = -rho[e];
    }
}

// RSolverElectrostatics::process, volume elements (:578-657): B_m = grad N_m * detJ (integrated
// grad N detJ / nInp), E = -sum B_m phi_m, J = sigma E, energy eps_r e0 |E|^2 / 2, resistivity
// |E| / |J|, Joule heat sigma |E|^2 L with L = 2 / sum_m |s . B_m|, s = E / |E|.
struct ElectroPostArgs {
    const double *xyz;
    const uint4 *conn0;
    uint32_t ne;
    const double *eps_r, *sigma, *phi;
    double *field, *current, *energy, *resistivity, *joule;
};

__global__ __launch_bounds__(kBlock) void k_electro_post(ElectroPostArgs a)
{
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= a.ne) return;
    const uint4 nd4 = a.conn0[e];
    const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
    double X[4][3], g[4][3];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int c = 0; c < 3; c++) X[m][c] = a.xyz[3 * (uint64_t)nd[m] + c];
    const double detJ = 6.0 * tet_gradients(X, g);
    double B[4][3];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int c = 0; c < 3; c++) B[m][c] = g[m][c] * detJ;
    double E[3] = { 0.0, 0.0, 0.0 };
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const double ph = a.phi[nd[m]];
#pragma unroll
        for (int c = 0; c < 3; c++) E[c] -= B[m][c] * ph;
    }
    const double sig = a.sigma[e];
    const double dE = E[0] * E[0] + E[1] * E[1] + E[2] * E[2];
    const double ere = sqrt(dE);
    const double J[3] = { E[0] * sig, E[1] * sig, E[2] * sig };
    const double jre = sqrt(J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
    double s[3] = { 0.0, 0.0, 0.0 };
    if (ere > kEps)
        for (int c = 0; c < 3; c++) s[c] = E[c] / ere;
    double L = 0.0;
#pragma unroll
    for (int m = 0; m < 4; m++) L += fabs(s[0] * B[m][0] + s[1] * B[m][1] + s[2] * B[m][2]);
    if (L > kEps) L = 2.0 / L;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        a.field[3 * (uint64_t)e + c] = E[c];
        a.current[3 * (uint64_t)e + c] = J[c];
    }
    a.energy[e] = a.eps_r[e] * kVacuumPermittivity * dE / 2.0;
    a.resistivity[e] = jre > kEps ? ere / jre : 0.0;
    a.joule[e] = sig * dE * L;
}

// RSolverStress::process, volume elements, static case (src/rsolverstress.cpp:1238-1408):
// element stresses sigma = De eps (Voigt xx, yy, zz, yz, xz, xy, engineering shear), the normal
// and shear measures and their sum (the reference's "von Mises"), and node forces Ke u_e =
// V Be sigma added to the element's nodes (one launch per colour: no shared nodes, a fixed
// order).
struct StressPostArgs {
    const double *xyz;
    const uint4 *conn;
    const uint32_t *perm;
    const double *E, *nu, *u;
    double *normal, *shear, *vm, *force;
};

__global__ __launch_bounds__(kBlock) void k_stress_post(StressPostArgs a, uint32_t e0, uint32_t e1)
{
    const uint32_t e = e0 + blockIdx.x * kBlock + threadIdx.x;
    if (e >= e1) return;
    const uint4 nd4 = a.conn[e];
    const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
    double X[4][3], g[4][3], u[4][3];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            X[m][c] = a.xyz[3 * (uint64_t)nd[m] + c];
            u[m][c] = a.u[3 * (uint64_t)nd[m] + c];
        }
    const double V = tet_gradients(X, g);
    const uint32_t orig = a.perm[e];
    const double E = a.E[orig], v = a.nu[orig];
    const double f = E / ((1.0 + v) * (1.0 - 2.0 * v));
    double ex = 0.0, ey = 0.0, ez = 0.0, gyz = 0.0, gxz = 0.0, gxy = 0.0;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        ex += g[m][0] * u[m][0];
        ey += g[m][1] * u[m][1];
        ez += g[m][2] * u[m][2];
        gyz += g[m][2] * u[m][1] + g[m][1] * u[m][2];
        gxz += g[m][2] * u[m][0] + g[m][0] * u[m][2];
        gxy += g[m][1] * u[m][0] + g[m][0] * u[m][1];
    }
    const double sx = f * ((1.0 - v) * ex + v * ey + v * ez);
    const double sy = f * (v * ex + (1.0 - v) * ey + v * ez);
    const double sz = f * (v * ex + v * ey + (1.0 - v) * ez);
    const double sh = f * (1.0 - 2.0 * v) / 2.0;
    const double tyz = sh * gyz, txz = sh * gxz, txy = sh * gxy;
    const double QeN = sqrt(sx * sx + sy * sy + sz * sz - (sx * sy + sy * sz + sz * sx));
    const double QeS = sqrt(3.0 * (tyz * tyz + txz * txz + txy * txy));
    a.normal[orig] = QeN;
    a.shear[orig] = QeS;
    a.vm[orig] = QeN + QeS;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        double *F = a.force + 3 * (uint64_t)nd[m];
        F[0] += V * (g[m][0] * sx + g[m][2] * txz + g[m][1] * txy);
        F[1] += V * (g[m][1] * sy + g[m][2] * tyz + g[m][0] * txy);
        F[2] += V * (g[m][2] * sz + g[m][1] * tyz + g[m][0] * txz);
    }
}

struct StressArgs {
    const double *xyz;
    const uint4 *conn;
    const uint32_t *perm;
    const uint32_t *map;        // 144 positions per element, [(4m+n)*9 + 3i+j]
    const uint32_t *book;       // 3*node+c -> dof
    const double *E, *nu, *rho;
    const double *alpha_t, *dT; // thermal expansion, element minus environment temperature (or null)
    double gx, gy, gz;
    const double *u_node;       // nodeDisplacement: prescribed values, and the previous step (transient)
    double *val, *b;
    double *mval;               // modal: values of M (same pattern as A)
    int problem;                // RS_STRESS_STATIC / _TRANSIENT / _MODAL
    double alpha, dt;
    const double *rot;          // local rotations: 3x3 row-major per node (null: none)
    const uint8_t *rot_active;  // per node
};

__device__ __forceinline__ void node_rotation(const StressArgs &a, uint32_t node, double (&R)[3][3])
{
    const bool on = a.rot_active[node] != 0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) R[i][j] = on ? a.rot[9 * (uint64_t)node + 3 * i + j] : (i == j ? 1.0 : 0.0);
}

// 16 lanes per element: lane (m, n) assembles the 3x3 block coupling node m to node n.
__global__ __launch_bounds__(kBlock) void k_stress_colour(StressArgs a, uint32_t e0, uint32_t e1)
{
    constexpr int EPB = kBlock / 16;  // elements per workgroup
    __shared__ double sX[EPB][12];
    const int le = threadIdx.x / 16, lane = threadIdx.x % 16;
    const int m = lane / 4, n = lane % 4;
    uint32_t e = e0 + blockIdx.x * EPB + le;
    bool live = e < e1;
    uint4 nd4 = live ? a.conn[e] : make_uint4(0, 0, 0, 0);
    uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
    if (live && lane < 12) sX[le][lane] = a.xyz[3 * (uint64_t)nd[lane / 3] + lane % 3];
    __syncthreads();
    if (!live) return;
    double X[4][3], g[4][3];
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int i = 0; i < 3; i++) X[p][i] = sX[le][3 * p + i];
    double V = tet_gradients(X, g);
    uint32_t orig = a.perm[e];
    double E = a.E[orig], v = a.nu[orig];
    double lam = E * v / ((1.0 + v) * (1.0 - 2.0 * v));
    double mu = E / (2.0 * (1.0 + v));
    double gmn = g[m][0] * g[n][0] + g[m][1] * g[n][1] + g[m][2] * g[n][2];
    double K[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            K[i][j] = V * (lam * g[m][i] * g[n][j] + mu * g[m][j] * g[n][i] + (i == j ? mu * gmn : 0.0));
    // consistent mass rho V (1 + d_mn) / 20 on the component diagonal (src/rsolverstress.cpp:665-680)
    const bool transient = a.problem == RS_STRESS_TRANSIENT, modal = a.problem == RS_STRESS_MODAL;
    const double Mmn = (transient || modal) ? a.rho[orig] * V * (m == n ? 2.0 : 1.0) / 20.0 : 0.0;
    // assemblyMatrix (:1654-1772): transient Ae = Me + alpha dt Ke, modal Ae = Ke and Be = Me
    double Ae[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Ae[i][j] = transient ? (i == j ? Mmn : 0.0) + a.alpha * a.dt * K[i][j] : K[i][j];
    // node n's displacements into the rows of node m: the transient history term (all components,
    // unrotated blocks)
    double hist[3] = { 0.0, 0.0, 0.0 };
    const double *un = a.u_node + 3 * (uint64_t)nd[n];
    if (transient) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) hist[i] += ((i == j ? Mmn : 0.0) - (1.0 - a.alpha) * a.dt * K[i][j]) * un[j];
    }
    // applyLocalRotations (:1774-1833): Ae -> Rm^T Ae Rn, Be -> Rm^T Be Rn; be -> Rm be (below)
    double Be[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Be[i][j] = (modal && i == j) ? Mmn : 0.0;
    double Rm[3][3], Rn[3][3];
    if (a.rot) {
        node_rotation(a, nd[m], Rm);
        node_rotation(a, nd[n], Rn);
        double t[3][3], t2[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) t[i][j] = Rm[0][i] * Ae[0][j] + Rm[1][i] * Ae[1][j] + Rm[2][i] * Ae[2][j];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) Ae[i][j] = t[i][0] * Rn[0][j] + t[i][1] * Rn[1][j] + t[i][2] * Rn[2][j];
        if (modal) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) t2[i][j] = Rm[0][i] * Be[0][j] + Rm[1][i] * Be[1][j] + Rm[2][i] * Be[2][j];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) Be[i][j] = t2[i][0] * Rn[0][j] + t2[i][1] * Rn[1][j] + t2[i][2] * Rn[2][j];
        }
    }
    // lifting of the prescribed components of node n with the (rotated) Ae + Be (:1719-1739)
    double lift[3] = { 0.0, 0.0, 0.0 };
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (a.book[3 * (uint64_t)nd[n] + j] != RS_EOD) continue;
        const double u = un[j];
#pragma unroll
        for (int i = 0; i < 3; i++) lift[i] += (Ae[i][j] + Be[i][j]) * u;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        hist[i] = group_sum<4>(hist[i]);
        lift[i] = group_sum<4>(lift[i]);
    }
    const uint32_t *mp = a.map + 144ull * e + lane * 9;
    // node m's own loads: gravity (:683-685) and thermal expansion (:688-693), whose sum over the
    // three normal components of Be De is alpha dT E / (1 - 2 nu) grad N_m for the isotropic De
    double w = a.rho[orig] * V * 0.25;
    double fe[3] = { a.gx * w, a.gy * w, a.gz * w };
    if (a.alpha_t) {
        const double th = a.alpha_t[orig] * a.dT[orig] * E / (1.0 - 2.0 * v) * V;
#pragma unroll
        for (int i = 0; i < 3; i++) fe[i] += th * g[m][i];
    }
    double be[3];
#pragma unroll
    for (int i = 0; i < 3; i++) be[i] = (transient ? a.dt * fe[i] : fe[i]) + hist[i];
    if (a.rot) {
        const double r0 = Rm[0][0] * be[0] + Rm[0][1] * be[1] + Rm[0][2] * be[2];
        const double r1 = Rm[1][0] * be[0] + Rm[1][1] * be[1] + Rm[1][2] * be[2];
        const double r2 = Rm[2][0] * be[0] + Rm[2][1] * be[1] + Rm[2][2] * be[2];
        be[0] = r0;
        be[1] = r1;
        be[2] = r2;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        uint32_t row = a.book[3 * (uint64_t)nd[m] + i];
        if (row == RS_EOD) continue;
        if (n == 0) a.b[row] += be[i] - lift[i];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            uint32_t pos = mp[3 * i + j];
            if (pos == RS_EOD) continue;
            a.val[pos] += Ae[i][j];
            // M receives the unrotated Me: assemblyMatrix scatters its Me argument (:1762-1765)
            if (modal && i == j) a.mval[pos] += Mmn;
        }
    }
}

// ------------------------------------------------------------------ stress: row gather
// The static, unrotated stress assembly as a gather over rows instead of colour launches that
// read-modify-write the CSR values.  One wavefront per node with a free component, one lane per
// incident element: every lane loads its element (geometry record, Lame constants, loads) and
// computes the 3 x 3 blocks K_mn coupling the node (local m) to the element's four nodes, with
// the colour kernel's expressions, into LDS; then lane (b, i) sums row i's entries of column
// block b over the elements in incidence order, which is colour order, so every entry receives
// its terms in the colour kernel's order; the right-hand side is added the same way (load minus
// lifting, the lifting summed over the four nodes as the colour kernel's lane sum does).  The
// rows are written whole: no zero fill of the matrix, no scatter.  Per mesh: the element
// geometry records and, per node, its column blocks (free neighbours and the offsets of their
// first free columns).  b is added to (the traction kernels run first).
constexpr int kStressRowCap = 48;     // entries of a row: up to 16 node blocks of 3 columns
constexpr int kStressMaxInc = 32;     // incident elements of a node (more: the colour launches)
constexpr int kStressMaxNbr = kStressRowCap / 3;

struct StressRowArgs {
    const uint32_t *nodes;      // nodes with a free component, ascending
    uint32_t n_nodes;
    const uint32_t *inc_ptr, *inc;   // per node: incident elements (colour-sorted index << 2 | local node)
    const uint32_t *nbr_ptr;    // per node with a free component: its column blocks
    const uint2 *nbr;           // (node, offset of its first free column | free mask << 8)
    const double *geo;          // per colour-sorted element: kGeoWords values (geometry record)
    const double *coef;         // per original element: lambda, mu, rho, thermal factor / V
    const uint32_t *book;
    const uint32_t *rowptr;
    double gx, gy, gz;
    int thermal;
    const double *u_node;
    double *val, *b;
};

// Geometry record of an element (built once per mesh): the 12 gradient values g[q][d] at 3q + d,
// the volume at 12, the free-component masks of its four nodes in the low 4 bytes of word 13 and
// its original element id in the high 4, and the four node ids in the last 16 bytes: one 128-byte
// line per element.
constexpr uint32_t kGeoWords = 16;

__global__ void k_stress_geometry(const double *__restrict__ xyz, const uint4 *__restrict__ conn,
                                  const uint32_t *__restrict__ perm, const uint8_t *__restrict__ fmask, uint32_t ne,
                                  double *__restrict__ geo)
{
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < ne; e += gridDim.x * kBlock) {
        const uint4 nd4 = conn[e];
        const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
        double X[4][3], g[4][3];
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int i = 0; i < 3; i++) X[p][i] = xyz[3 * (uint64_t)nd[p] + i];
        double *r = geo + (uint64_t)kGeoWords * e;
        r[12] = tet_gradients(X, g);
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int i = 0; i < 3; i++) r[3 * p + i] = g[p][i];
        const uint32_t fm = fmask[nd[0]] | (fmask[nd[1]] << 8) | (fmask[nd[2]] << 16) | ((uint32_t)fmask[nd[3]] << 24);
        reinterpret_cast<uint2 *>(r)[13] = make_uint2(fm, perm[e]);
        reinterpret_cast<uint4 *>(r)[7] = nd4;
    }
}

// Per assembly: each element's Lame constants, density and thermal factor (before the volume),
// as the colour kernel computes them, in the caller's element order (coalesced reads and
// writes; the row kernel finds an element's entry through the original id in its record).
__global__ void k_stress_coef(const double *__restrict__ E, const double *__restrict__ nu,
                              const double *__restrict__ rho, const double *__restrict__ alpha_t,
                              const double *__restrict__ dT, uint32_t ne, double *__restrict__ coef)
{
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < ne; e += gridDim.x * kBlock) {
        const double Ee = E[e], v = nu[e];
        double4 c;
        c.x = Ee * v / ((1.0 + v) * (1.0 - 2.0 * v));
        c.y = Ee / (2.0 * (1.0 + v));
        c.z = rho[e];
        c.w = alpha_t ? alpha_t[e] * dT[e] * Ee / (1.0 - 2.0 * v) : 0.0;
        reinterpret_cast<double4 *>(coef)[e] = c;
    }
}

// One wavefront per node, four nodes per workgroup.  Phase 1: lane q < cnt loads incident
// element q into LDS (gradients, volume, Lame constants, nodes, the node's local index) and
// computes its right-hand-side terms.  Phase 2: lane (b, i) owns row i's entries of column block
// b and sums, over the elements in incidence order, K_mn[i][j] of the elements that contain block
// b's node (local n), computed from LDS with the colour kernel's expressions, in registers; lanes
// 0-2 add the right-hand-side terms to b.
constexpr int kStressWaves = kBlock / kWave;
constexpr int kElemWords = 16;        // per element in LDS: g[12], V, lambda, mu, (unused)

__device__ __forceinline__ double stress_k(const double *x, uint32_t m, uint32_t n, uint32_t i, uint32_t j)
{
    // K_mn[i][j] = V (lambda g_mi g_nj + mu g_mj g_ni + delta_ij mu g_m.g_n), as k_stress_colour
    const double gmn = x[3 * m] * x[3 * n] + x[3 * m + 1] * x[3 * n + 1] + x[3 * m + 2] * x[3 * n + 2];
    const double V = x[12], lam = x[13], mu = x[14];
    return V * (lam * x[3 * m + i] * x[3 * n + j] + mu * x[3 * m + j] * x[3 * n + i] + (i == j ? mu * gmn : 0.0));
}

__global__ __launch_bounds__(kBlock) void k_stress_rows(StressRowArgs a)
{
    __shared__ double el[kStressWaves][kStressMaxInc][kElemWords];
    __shared__ double bcon[kStressWaves][3][kStressMaxInc];   // load - lifting of rows i, element q
    __shared__ uint4 enode[kStressWaves][kStressMaxInc];
    __shared__ uint32_t elocal[kStressWaves][kStressMaxInc];
    const uint32_t w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const uint32_t k = xcd_remap(blockIdx.x, gridDim.x) * kStressWaves + w;
    const bool live = k < a.n_nodes;                  // whole wavefronts
    const uint32_t node = live ? a.nodes[k] : 0;
    const uint32_t p0 = live ? a.inc_ptr[node] : 0, cnt = live ? a.inc_ptr[node + 1] - p0 : 0;
    if (lane < cnt) {
        const uint32_t wd = a.inc[p0 + lane], e = wd >> 2, m = wd & 3u;
        const double2 *r = reinterpret_cast<const double2 *>(a.geo + (uint64_t)kGeoWords * e);
        double2 t[7];
#pragma unroll
        for (int q = 0; q < 7; q++) t[q] = r[q];
        const uint4 nd4 = reinterpret_cast<const uint4 *>(r)[7];
        // word 13: the nodes' free-component masks and the original element id
        const uint64_t w13 = (uint64_t)__double_as_longlong(t[6].y);
        const uint32_t fm = (uint32_t)w13;
        const double4 cf = reinterpret_cast<const double4 *>(a.coef)[w13 >> 32];
        double *x = el[w][lane];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            x[2 * q] = t[q].x;
            x[2 * q + 1] = t[q].y;
        }
        x[12] = t[6].x;
        x[13] = cf.x;
        x[14] = cf.y;
        enode[w][lane] = nd4;
        elocal[w][lane] = m;
        double g[4][3];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            g[(2 * q) / 3][(2 * q) % 3] = t[q].x;
            g[(2 * q + 1) / 3][(2 * q + 1) % 3] = t[q].y;
        }
        const double V = t[6].x, lam = cf.x, mu = cf.y;
        double gm[3];
#pragma unroll
        for (int d = 0; d < 3; d++) gm[d] = m == 0 ? g[0][d] : m == 1 ? g[1][d] : m == 2 ? g[2][d] : g[3][d];
        const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
        // the colour kernel's right-hand side of node m: load minus the lifting of the prescribed
        // components, the lifting summed over j per node and then as (n0 + n2) + (n1 + n3)
        double lift[4][3];
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const double gmn = gm[0] * g[n][0] + gm[1] * g[n][1] + gm[2] * g[n][2];
            const uint32_t fmn = (fm >> (8 * n)) & 0xFFu;
#pragma unroll
            for (int i = 0; i < 3; i++) lift[n][i] = 0.0;
            if (fmn == 7u) continue;                  // every component free: nothing to lift
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (!(fmn & (1u << j))) {
                        const double K = V * (lam * gm[i] * g[n][j] + mu * gm[j] * g[n][i] + (i == j ? mu * gmn : 0.0));
                        lift[n][i] += K * a.u_node[3 * (uint64_t)nd[n] + j];
                    }
        }
        const double wq = cf.z * V * 0.25;
        const double gv[3] = { a.gx, a.gy, a.gz };
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double fe = gv[i] * wq;
            if (a.thermal) fe += cf.w * V * gm[i];
            bcon[w][i][lane] = fe - ((lift[0][i] + lift[2][i]) + (lift[1][i] + lift[3][i]));
        }
    }
    __syncthreads();
    const uint32_t b0 = live ? a.nbr_ptr[k] : 0, nb = live ? a.nbr_ptr[k + 1] - b0 : 0;
    if (lane < 3 * nb) {
        const uint32_t bk = lane / 3, i = lane % 3;
        const uint32_t row = a.book[3 * (uint64_t)node + i];
        if (row != RS_EOD) {
            const uint2 nb2 = a.nbr[b0 + bk];
            const uint32_t nn = nb2.x, off = nb2.y & 0xFFu, fm = nb2.y >> 8;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (uint32_t q = 0; q < cnt; q++) {
                const uint4 e4 = enode[w][q];
                const uint32_t n = e4.x == nn ? 0u : e4.y == nn ? 1u : e4.z == nn ? 2u : e4.w == nn ? 3u : 4u;
                if (n < 4u) {
                    const double *x = el[w][q];
                    const uint32_t m = elocal[w][q];
                    s0 += stress_k(x, m, n, i, 0);
                    s1 += stress_k(x, m, n, i, 1);
                    s2 += stress_k(x, m, n, i, 2);
                }
            }
            double *v = a.val + a.rowptr[row] + off;
            uint32_t c = 0;
            if (fm & 1u) v[c++] = s0;
            if (fm & 2u) v[c++] = s1;
            if (fm & 4u) v[c] = s2;
        }
    }
    if (lane < 3 && live) {
        const uint32_t row = a.book[3 * (uint64_t)node + lane];
        if (row != RS_EOD) {
            double bi = a.b[row];
            for (uint32_t q = 0; q < cnt; q++) bi += bcon[w][lane][q];
            a.b[row] = bi;
        }
    }
}

// Row gather of the magnetostatic element (3 unknowns per node, the component diagonals of each
// 3 x 3 block), on the stress assembly's node lists and geometry records: one thread per node
// with a free component sums K_mn of its incident elements (colour order) per neighbour block in
// LDS, and the sources and the lifting of prescribed neighbours; then writes its three rows
// whole (zeros off the diagonals) and b.  A node counts as prescribed when its first component
// is, as in k_magneto_colour; same terms in the same order, so the same bits.
struct MagnetoRowArgs {
    const uint32_t *nodes;
    uint32_t n_nodes;
    const uint32_t *inc_ptr, *inc;
    const uint32_t *nbr_ptr;
    const uint2 *nbr;           // (node, offset of its first free column | free mask << 8)
    const uint8_t *fmask;       // per node: free components
    const double *geo;
    const uint32_t *book, *rowptr;
    const double *current, *field;
    double *val, *b;
};

__global__ __launch_bounds__(kBlock) void k_magneto_rows(MagnetoRowArgs a)
{
#pragma clang fp contract(off)
    __shared__ double acc[kStressMaxNbr][kBlock];
    __shared__ uint32_t nbn[kStressMaxNbr][kBlock];
    const uint32_t t = threadIdx.x;
    const uint32_t k = xcd_remap(blockIdx.x, gridDim.x) * kBlock + t;
    if (k >= a.n_nodes) return;
    const uint32_t v = a.nodes[k];
    const uint32_t b0 = a.nbr_ptr[k], nb = a.nbr_ptr[k + 1] - b0;
    for (uint32_t j = 0; j < nb; j++) {
        nbn[j][t] = a.nbr[b0 + j].x;
        acc[j][t] = 0.0;
    }
    const bool vfree = (a.fmask[v] & 1u) != 0;
    double bsum[3] = { 0.0, 0.0, 0.0 };
    if (vfree) {
        const double *Jv = a.current + 3 * (uint64_t)v;
        for (uint32_t q = a.inc_ptr[v]; q < a.inc_ptr[v + 1]; q++) {
            const uint32_t wd = a.inc[q], e = wd >> 2, m = wd & 3u;
            const double *r = a.geo + (uint64_t)kGeoWords * e;
            double g[4][3];
#pragma unroll
            for (int i = 0; i < 12; i++) g[i / 3][i % 3] = r[i];
            const double V = r[12];
            const uint32_t fm = reinterpret_cast<const uint2 *>(r)[13].x;
            const uint4 nd4 = reinterpret_cast<const uint4 *>(r)[7];
            const uint32_t nd[4] = { nd4.x, nd4.y, nd4.z, nd4.w };
            double gm[3];
#pragma unroll
            for (int d = 0; d < 3; d++) gm[d] = m == 0 ? g[0][d] : m == 1 ? g[1][d] : m == 2 ? g[2][d] : g[3][d];
            double fe[3];
            magneto_source(gm, Jv, V, fe);
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const double K = magneto_k(gm, g[n], V);
                if (!((fm >> (8 * n)) & 1u)) {          // prescribed neighbour: lifting
                    const double *H = a.field + 3 * (uint64_t)nd[n];
#pragma unroll
                    for (int c = 0; c < 3; c++) fe[c] -= K * H[c];
                } else {
                    uint32_t j = 0;
                    while (j + 1 < nb && nbn[j][t] != nd[n]) j++;
                    acc[j][t] += K;
                }
            }
#pragma unroll
            for (int c = 0; c < 3; c++) bsum[c] += fe[c];
        }
    }
    for (int c = 0; c < 3; c++) {
        const uint32_t row = a.book[3 * (uint64_t)v + c];
        if (row == RS_EOD) continue;
        double *dst = a.val + a.rowptr[row];
        for (uint32_t j = 0; j < nb; j++) {
            const uint32_t w = a.nbr[b0 + j].y, off = w & 0xFFu, fmu = w >> 8;
            const double kv = (fmu & 1u) ? acc[j][t] : 0.0;
            uint32_t p = off;
            for (int c2 = 0; c2 < 3; c2++)
                if ((fmu >> c2) & 1u) dst[p++] = c2 == c ? kv : 0.0;
        }
        a.b[row] = bsum[c];
    }
}

// RSolverStress::setDisplacement (src/rsolverstress.cpp:800-836): du = x on the enabled
// components, 0 on the disabled ones; an active local rotation writes iR^T du
// (RLocalRotation::rotateResultsVector, src/rlocalrotation.cpp:60-77).
__global__ void k_stress_set_displacement(const uint32_t *__restrict__ book, uint32_t nn, const double *__restrict__ x,
                                          const double *__restrict__ irot, const uint8_t *__restrict__ active,
                                          double *__restrict__ u)
{
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nn; i += gridDim.x * kBlock) {
        double du[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const uint32_t p = book[3 * (uint64_t)i + c];
            du[c] = p != RS_EOD ? x[p] : 0.0;
        }
        if (irot && active[i]) {
            const double *iR = irot + 9 * (uint64_t)i;
#pragma unroll
            for (int c = 0; c < 3; c++) u[3 * (uint64_t)i + c] = iR[c] * du[0] + iR[3 + c] * du[1] + iR[6 + c] * du[2];
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) u[3 * (uint64_t)i + c] = du[c];
        }
    }
}

// Surface force per unit area on linear triangles: each node gets t * area / 3
// (src/rsolverstress.cpp:498-507 with N integrated exactly).
__global__ void k_traction_colour(const double *__restrict__ xyz, const uint32_t *__restrict__ tris,
                                  const double *__restrict__ t, const uint32_t *__restrict__ book,
                                  double *__restrict__ b, uint32_t f0, uint32_t f1, double scale)
{
    uint32_t f = f0 + blockIdx.x * kBlock + threadIdx.x;
    if (f >= f1) return;
    uint32_t nd[3] = { tris[3 * f], tris[3 * f + 1], tris[3 * f + 2] };
    double e1[3], e2[3];
    for (int i = 0; i < 3; i++) {
        e1[i] = xyz[3 * (uint64_t)nd[1] + i] - xyz[3 * (uint64_t)nd[0] + i];
        e2[i] = xyz[3 * (uint64_t)nd[2] + i] - xyz[3 * (uint64_t)nd[0] + i];
    }
    double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
    double area3 = 0.5 * sqrt(cx * cx + cy * cy + cz * cz) / 3.0 * scale;
    for (int m = 0; m < 3; m++)
        for (int c = 0; c < 3; c++) {
            uint32_t row = book[3 * (uint64_t)nd[m] + c];
            if (row != RS_EOD) b[row] += t[3 * f + c] * area3;
        }
}

// node[i*dpn+c] = x[book[i*dpn+c]] for enabled positions (src/rsolverheat.cpp:473-480)
__global__ void k_dofs_to_nodes(const uint32_t *__restrict__ book, uint64_t npos, const double *__restrict__ x,
                                double *__restrict__ node)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < npos; i += (uint64_t)gridDim.x * kBlock) {
        uint32_t d = book[i];
        if (d != RS_EOD) node[i] = x[d];
    }
}

__global__ void k_gather(const double *__restrict__ src, const uint32_t *__restrict__ perm, double *__restrict__ dst,
                         uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        dst[i] = src[perm[i]];
}

static uint32_t grid_for(uint64_t items, uint32_t per_block)
{
    return (uint32_t)std::max<uint64_t>(1, (items + per_block - 1) / per_block);
}

// A grid whose size is a multiple of the XCD count, for kernels that map blockIdx through
// xcd_remap (extra workgroups find no rows and exit).
static uint32_t grid_xcd(uint64_t items, uint32_t per_block)
{
    const uint32_t g = grid_for(items, per_block);
    return (g + kXcd - 1) / kXcd * kXcd;
}

static uint32_t vblocks(uint64_t n)
{
    uint64_t want = (n + kBlock - 1) / kBlock;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)std::max(8, runtime().n_cu * 8)));
}

}  // namespace rs

using namespace rs;

struct rs_assembly {
    uint32_t n_nodes = 0, n_elems = 0, dpn = 1, n_dof = 0, n_colours = 0;
    std::vector<uint32_t> book;           // host copy
    std::vector<uint32_t> colour_off;     // n_colours+1
    DBuf<double> xyz;
    DBuf<uint32_t> conn, perm, map, dbook;
    DBuf<uint32_t> conn0;                   // connectivity in the caller's order (post-processing)
    DBuf<double> coeff;                     // mapped element coefficients (electrostatics)
    DBuf<double> telem;                     // per-assembly transport element quantities (row gather)
    int whole_nodes = -1;                   // 3-DOF book prescribes whole nodes only (-1: not checked)
    DBuf<double> part;                      // block partials of the fluid write-back norms
    DBuf<uint32_t> dof_node, inc_ptr, inc;  // row-gather heat assembly
    DBuf<double> rec_grad, rec_vol;         // row records (built once per mesh)
    DBuf<uint32_t> rec_elem, rec_slots, rec_count;
    DBuf<uint4> rec_nodes;
    DBuf<uint64_t> rec_base;
    bool rows = false;                      // every row fits kRowCap entries
    bool use_rows = true;                   // row-gather kernels (else the colour launches)
    // static stress row gather: nodes with a free component, their column blocks, per node its
    // free-component mask, element geometry records and per-assembly coefficients
    DBuf<uint32_t> s_nodes, s_nbr_ptr, s_nbr;
    DBuf<uint8_t> s_fmask;
    DBuf<double> s_geo, s_coef;
    uint32_t s_count = 0;
    bool srows = false;
    rs_matrix A;                          // borrowed by callers
    rs_matrix M;                          // modal stress: the mass matrix (A's pattern)
    bool has_mass = false;
    DBuf<double> rot, irot;               // stress local rotations: R and R^-1 per node (row-major)
    DBuf<uint8_t> rot_active;
    bool has_rot = false;
    DBuf<double> fields, t_node, bvec;    // staging for the host-array entry points
    DBuf<uint32_t> tris;
    DBuf<double> trac;
    double last_ms = 0.0;
};

static int timed_begin(EventPair &ev)
{
    RS_HIP(ev.create());
    RS_HIP(hipEventRecord(ev.e[0], runtime().stream));
    return RS_OK;
}

static int timed_end(EventPair &ev, double &ms)
{
    RS_HIP(hipEventRecord(ev.e[1], runtime().stream));
    RS_HIP(hipEventSynchronize(ev.e[1]));
    float f = 0.0f;
    RS_HIP(hipEventElapsedTime(&f, ev.e[0], ev.e[1]));
    ms = f;
    return RS_OK;
}

extern "C" int rs_assembly_create(uint32_t n_nodes, const double *xyz, uint32_t n_tets, const uint32_t *tets,
                                  uint32_t dofs_per_node, const uint8_t *fixed, rs_assembly **out)
{
    if (!xyz || !tets || !out || (dofs_per_node != 1 && dofs_per_node != 3 && dofs_per_node != 4))
        return fail(RS_ERR_INVALID_ARG, "rs_assembly_create: bad arguments (dofs_per_node must be 1, 3 or 4)");
    for (uint64_t i = 0; i < 4ull * n_tets; i++)
        if (tets[i] >= n_nodes) return fail(RS_ERR_INVALID_ARG, "rs_assembly_create: node id out of range");
    int rc = require_runtime();
    if (rc) return rc;
    std::unique_ptr<rs_assembly> S(new rs_assembly());
    S->n_nodes = n_nodes;
    S->n_elems = n_tets;
    S->dpn = dofs_per_node;
    const uint32_t dpn = dofs_per_node;
    S->book.resize((uint64_t)n_nodes * dpn);
    if ((rc = rs_node_book(n_nodes, n_tets, 4, tets, dpn, fixed, S->book.data(), &S->n_dof))) return rc;
    Pattern pat;
    if ((rc = build_pattern(n_nodes, n_tets, 4, tets, dpn, S->book.data(), S->n_dof, pat))) return rc;
    std::vector<uint16_t> colour;
    if ((rc = colour_elements(n_nodes, n_tets, 4, tets, colour, S->n_colours))) return rc;
    // counting sort of the elements by colour (stable: spatial order kept inside a colour)
    S->colour_off.assign(S->n_colours + 1, 0);
    for (uint32_t e = 0; e < n_tets; e++) S->colour_off[colour[e] + 1]++;
    for (uint32_t c = 0; c < S->n_colours; c++) S->colour_off[c + 1] += S->colour_off[c];
    std::vector<uint32_t> perm(n_tets), conn_s(4ull * n_tets);
    {
        std::vector<uint32_t> fill(S->colour_off.begin(), S->colour_off.end() - 1);
        for (uint32_t e = 0; e < n_tets; e++) perm[fill[colour[e]]++] = e;
    }
    const uint32_t per = dpn == 1 ? 16 : dpn == 3 ? 144 : 64;
    std::vector<uint32_t> map((uint64_t)per * n_tets);
    const uint32_t *rp = pat.rowptr.data(), *cl = pat.col.data();
    const std::vector<uint32_t> &book = S->book;
    parallel_for(n_tets, [&](uint64_t b, uint64_t e_end) {
        for (uint64_t s = b; s < e_end; s++) {
            const uint32_t *nd = tets + 4ull * perm[s];
            std::memcpy(&conn_s[4 * s], nd, 16);
            auto find = [&](uint32_t row, uint32_t colv) -> uint32_t {
                const uint32_t *lo = cl + rp[row], *hi = cl + rp[row + 1];
                const uint32_t *it = std::lower_bound(lo, hi, colv);
                return (it != hi && *it == colv) ? (uint32_t)(it - cl) : RS_EOD;
            };
            if (dpn == 1) {
                for (int m = 0; m < 4; m++)
                    for (int n = 0; n < 4; n++) {
                        uint32_t r = book[nd[m]], c = book[nd[n]];
                        map[16 * s + 4 * m + n] = (r == RS_EOD || c == RS_EOD) ? RS_EOD : find(r, c);
                    }
            } else if (dpn == 4) {
                // the enabled columns of one node are adjacent in a row (the node book is monotone
                // in (node, component)), so one position per (m, n, row component) locates them
                for (int m = 0; m < 4; m++)
                    for (int n = 0; n < 4; n++) {
                        uint32_t c0 = RS_EOD;
                        for (int j = 0; j < 4 && c0 == RS_EOD; j++) c0 = book[4ull * nd[n] + j];
                        for (int i = 0; i < 4; i++) {
                            uint32_t r = book[4ull * nd[m] + i];
                            map[64 * s + (4 * m + n) * 4 + i] = (r == RS_EOD || c0 == RS_EOD) ? RS_EOD : find(r, c0);
                        }
                    }
            } else {
                for (int m = 0; m < 4; m++)
                    for (int n = 0; n < 4; n++)
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++) {
                                uint32_t r = book[3ull * nd[m] + i], c = book[3ull * nd[n] + j];
                                map[144 * s + (4 * m + n) * 9 + 3 * i + j] =
                                    (r == RS_EOD || c == RS_EOD) ? RS_EOD : find(r, c);
                            }
            }
        }
    });
    // node -> incident elements (colour-sorted index, local node) for the row-gather heat kernel
    std::vector<uint32_t> inc_ptr, inc, dof_node;
    uint32_t max_row = 0;
    for (uint32_t r = 0; r < S->n_dof; r++) max_row = std::max(max_row, pat.rowptr[r + 1] - pat.rowptr[r]);
    S->rows = dpn == 1 && max_row <= (uint32_t)kRowCap && n_tets < (1u << 30);
    if (S->rows) {
        inc_ptr.assign(n_nodes + 1ull, 0);
        for (uint64_t s = 0; s < n_tets; s++)
            for (int m = 0; m < 4; m++) inc_ptr[conn_s[4 * s + m] + 1]++;
        for (uint32_t v = 0; v < n_nodes; v++) inc_ptr[v + 1] += inc_ptr[v];
        inc.resize(4ull * n_tets);
        std::vector<uint32_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
        for (uint64_t s = 0; s < n_tets; s++)
            for (uint32_t m = 0; m < 4; m++) inc[fill[conn_s[4 * s + m]]++] = (uint32_t)(s << 2) | m;
        dof_node.assign(std::max(S->n_dof, 1u), 0);
        for (uint32_t v = 0; v < n_nodes; v++)
            if (book[v] != RS_EOD) dof_node[book[v]] = v;
    }
    // static stress row gather: node incidence (colour order), free-component masks, and each
    // node's column blocks (the free neighbours and the offsets of their first free columns)
    std::vector<uint32_t> s_nodes, s_nbr_ptr, s_nbr;
    std::vector<uint8_t> s_fmask;
    S->srows = dpn == 3 && max_row <= (uint32_t)kStressRowCap && n_tets < (1u << 30);
    if (S->srows) {
        inc_ptr.assign(n_nodes + 1ull, 0);
        for (uint64_t s = 0; s < n_tets; s++)
            for (int m = 0; m < 4; m++) inc_ptr[conn_s[4 * s + m] + 1]++;
        for (uint32_t v = 0; v < n_nodes; v++) inc_ptr[v + 1] += inc_ptr[v];
        inc.resize(4ull * n_tets);
        std::vector<uint32_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
        for (uint64_t s = 0; s < n_tets; s++)
            for (uint32_t m = 0; m < 4; m++) inc[fill[conn_s[4 * s + m]]++] = (uint32_t)(s << 2) | m;
        s_fmask.assign(n_nodes, 0);
        std::vector<uint32_t> first_dof(n_nodes, RS_EOD);
        for (uint32_t v = 0; v < n_nodes; v++)
            for (int c = 2; c >= 0; c--)
                if (book[3ull * v + c] != RS_EOD) {
                    s_fmask[v] |= (uint8_t)(1u << c);
                    first_dof[v] = book[3ull * v + c];
                }
        for (uint32_t v = 0; v < n_nodes; v++)
            if (s_fmask[v]) s_nodes.push_back(v);
        bool ok = true;
        // column blocks of each node's rows: the nodes with a free component that share an
        // element with it, ascending (the order of their columns in the rows)
        s_nbr_ptr.assign(s_nodes.size() + 1, 0);
        for (size_t k = 0; k < s_nodes.size() && ok; k++) {
            const uint32_t v = s_nodes[k];
            if (inc_ptr[v + 1] - inc_ptr[v] > (uint32_t)kStressMaxInc) ok = false;
            std::vector<uint32_t> nb;
            for (uint32_t q = inc_ptr[v]; q < inc_ptr[v + 1]; q++)
                for (int n = 0; n < 4; n++) {
                    const uint32_t u = conn_s[4ull * (inc[q] >> 2) + n];
                    if (s_fmask[u]) nb.push_back(u);
                }
            std::sort(nb.begin(), nb.end());
            nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
            if (nb.size() > (size_t)kStressMaxNbr) ok = false;
            const uint32_t r = first_dof[v];
            const uint32_t *lo = cl + rp[r], *hi = cl + rp[r + 1];
            for (uint32_t u : nb) {
                const uint32_t *it = std::lower_bound(lo, hi, first_dof[u]);
                if (it == hi || *it != first_dof[u]) ok = false;
                s_nbr.push_back(u);
                s_nbr.push_back((uint32_t)(it - lo) | ((uint32_t)s_fmask[u] << 8));
            }
            s_nbr_ptr[k + 1] = (uint32_t)(s_nbr.size() / 2);
        }
        S->srows = ok;
        S->s_count = (uint32_t)s_nodes.size();
    }
    hipStream_t st = runtime().stream;
    uint64_t nnz = pat.col.size();
    if ((rc = S->xyz.alloc(3ull * n_nodes)) || (rc = S->conn.alloc(4ull * n_tets)) || (rc = S->perm.alloc(n_tets)) ||
        (rc = S->map.alloc(map.size())) || (rc = S->dbook.alloc(S->book.size())))
        return rc;
    rs_matrix &A = S->A;
    A.n = S->n_dof;
    A.nnz = nnz;
    A.borrowed = true;
    if ((rc = A.rowptr.alloc(S->n_dof + 1ull)) || (rc = A.col.alloc(nnz)) || (rc = A.val.alloc(nnz))) return rc;
    RS_HIP(hipMemcpyAsync(S->xyz.p, xyz, sizeof(double) * 3 * n_nodes, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->conn.p, conn_s.data(), sizeof(uint32_t) * conn_s.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->perm.p, perm.data(), sizeof(uint32_t) * perm.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->map.p, map.data(), sizeof(uint32_t) * map.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->dbook.p, S->book.data(), sizeof(uint32_t) * S->book.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(A.rowptr.p, pat.rowptr.data(), sizeof(uint32_t) * pat.rowptr.size(), hipMemcpyHostToDevice, st));
    if (nnz) {
        RS_HIP(hipMemcpyAsync(A.col.p, pat.col.data(), sizeof(uint32_t) * nnz, hipMemcpyHostToDevice, st));
        k_zero<<<vblocks(nnz), kBlock, 0, st>>>(A.val.p, nnz);
    }
    if (S->rows) {
        if ((rc = S->inc_ptr.alloc(inc_ptr.size())) || (rc = S->inc.alloc(std::max<size_t>(inc.size(), 1))) ||
            (rc = S->dof_node.alloc(dof_node.size())))
            return rc;
        RS_HIP(hipMemcpyAsync(S->inc_ptr.p, inc_ptr.data(), sizeof(uint32_t) * inc_ptr.size(), hipMemcpyHostToDevice, st));
        if (!inc.empty()) RS_HIP(hipMemcpyAsync(S->inc.p, inc.data(), sizeof(uint32_t) * inc.size(), hipMemcpyHostToDevice, st));
        RS_HIP(hipMemcpyAsync(S->dof_node.p, dof_node.data(), sizeof(uint32_t) * dof_node.size(), hipMemcpyHostToDevice, st));
        // row records: slices of 64 rows, padded to the slice's longest incidence list
        const uint32_t nsl = (S->n_dof + kWave - 1) / kWave;
        std::vector<uint64_t> rb(nsl + 1ull, 0);
        for (uint32_t sidx = 0; sidx < nsl; sidx++) {
            uint32_t w = 0;
            for (uint32_t r = sidx * kWave; r < std::min(S->n_dof, (sidx + 1) * kWave); r++) {
                const uint32_t v = dof_node[r];
                w = std::max(w, inc_ptr[v + 1] - inc_ptr[v]);
            }
            rb[sidx + 1] = rb[sidx] + (uint64_t)w * kWave;
        }
        const uint64_t nrec = std::max<uint64_t>(rb[nsl], 1);
        if ((rc = S->rec_grad.alloc(4 * nrec)) || (rc = S->rec_vol.alloc(nrec)) || (rc = S->rec_elem.alloc(nrec)) ||
            (rc = S->rec_slots.alloc(nrec)) || (rc = S->rec_nodes.alloc(nrec)) || (rc = S->rec_base.alloc(rb.size())) ||
            (rc = S->rec_count.alloc(std::max(S->n_dof, 1u))))
            return rc;
        RS_HIP(hipMemcpyAsync(S->rec_base.p, rb.data(), sizeof(uint64_t) * rb.size(), hipMemcpyHostToDevice, st));
        HeatArgs ha{};
        ha.xyz = S->xyz.p;
        ha.conn = reinterpret_cast<const uint4 *>(S->conn.p);
        ha.perm = S->perm.p;
        ha.map = S->map.p;
        if (S->n_dof)
            k_heat_records<<<grid_for(S->n_dof, kBlock), kBlock, 0, st>>>(
                ha, A.rowptr.p, S->dof_node.p, S->inc_ptr.p, S->inc.p, S->n_dof, S->rec_grad.p, S->rec_vol.p,
                S->rec_elem.p, S->rec_slots.p, S->rec_nodes.p, S->rec_base.p, S->rec_count.p);
        RS_HIP(hipGetLastError());
    }
    if (S->srows) {
        if ((rc = S->inc_ptr.alloc(inc_ptr.size())) || (rc = S->inc.alloc(std::max<size_t>(inc.size(), 1))) ||
            (rc = S->s_nodes.alloc(std::max<size_t>(s_nodes.size(), 1))) || (rc = S->s_nbr_ptr.alloc(s_nbr_ptr.size())) ||
            (rc = S->s_nbr.alloc(std::max<size_t>(s_nbr.size(), 2))) ||
            (rc = S->s_fmask.alloc(std::max(n_nodes, 1u))) || (rc = S->s_geo.alloc((uint64_t)kGeoWords * std::max(n_tets, 1u))) ||
            (rc = S->s_coef.alloc(4ull * std::max(n_tets, 1u))))
            return rc;
        RS_HIP(hipMemcpyAsync(S->inc_ptr.p, inc_ptr.data(), sizeof(uint32_t) * inc_ptr.size(), hipMemcpyHostToDevice, st));
        if (!inc.empty()) RS_HIP(hipMemcpyAsync(S->inc.p, inc.data(), sizeof(uint32_t) * inc.size(), hipMemcpyHostToDevice, st));
        if (!s_nodes.empty())
            RS_HIP(hipMemcpyAsync(S->s_nodes.p, s_nodes.data(), sizeof(uint32_t) * s_nodes.size(), hipMemcpyHostToDevice, st));
        RS_HIP(hipMemcpyAsync(S->s_nbr_ptr.p, s_nbr_ptr.data(), sizeof(uint32_t) * s_nbr_ptr.size(), hipMemcpyHostToDevice, st));
        if (!s_nbr.empty())
            RS_HIP(hipMemcpyAsync(S->s_nbr.p, s_nbr.data(), sizeof(uint32_t) * s_nbr.size(), hipMemcpyHostToDevice, st));
        if (n_nodes) RS_HIP(hipMemcpyAsync(S->s_fmask.p, s_fmask.data(), n_nodes, hipMemcpyHostToDevice, st));
        if (n_tets)
            k_stress_geometry<<<vblocks(n_tets), kBlock, 0, st>>>(S->xyz.p, reinterpret_cast<const uint4 *>(S->conn.p),
                                                                 S->perm.p, S->s_fmask.p, n_tets, S->s_geo.p);
        RS_HIP(hipGetLastError());
    }
    if ((rc = matrix_set_layout(A, pat.rowptr.data()))) return rc;
    RS_HIP(hipStreamSynchronize(st));
    *out = S.release();
    return RS_OK;
}

// Assembly method: 1 = row gather (heat rows of <= 16 entries, static stress rows of <= 48), 0 = colour launches.
extern "C" int rs_assembly_set_method(rs_assembly *S, int32_t row_gather)
{
    if (!S) return fail(RS_ERR_INVALID_ARG, "rs_assembly_set_method: null");
    S->use_rows = row_gather != 0;
    return RS_OK;
}

extern "C" int rs_assembly_info(const rs_assembly *S, uint32_t *n_dof, uint64_t *nnz, uint32_t *n_colours)
{
    if (!S) return fail(RS_ERR_INVALID_ARG, "rs_assembly_info: null");
    if (n_dof) *n_dof = S->n_dof;
    if (nnz) *nnz = S->A.nnz;
    if (n_colours) *n_colours = S->n_colours;
    return RS_OK;
}

extern "C" int rs_assembly_last_ms(const rs_assembly *S, double *ms)
{
    if (!S || !ms) return fail(RS_ERR_INVALID_ARG, "rs_assembly_last_ms: null");
    *ms = S->last_ms;
    return RS_OK;
}

extern "C" int rs_assembly_book(const rs_assembly *S, uint32_t *book)
{
    if (!S || !book) return fail(RS_ERR_INVALID_ARG, "rs_assembly_book: null");
    std::memcpy(book, S->book.data(), sizeof(uint32_t) * S->book.size());
    return RS_OK;
}

extern "C" int rs_assembly_device_mesh(const rs_assembly *S, const double **d_xyz, const uint32_t **d_book,
                                      uint32_t *n_nodes, uint32_t *dofs_per_node)
{
    if (!S || !d_xyz || !d_book || !n_nodes || !dofs_per_node) return fail(RS_ERR_INVALID_ARG, "rs_assembly_device_mesh: null");
    *d_xyz = S->xyz.p;
    *d_book = S->dbook.p;
    *n_nodes = S->n_nodes;
    *dofs_per_node = S->dpn;
    return RS_OK;
}

extern "C" int rs_assembly_matrix(rs_assembly *S, rs_matrix **A)
{
    if (!S || !A) return fail(RS_ERR_INVALID_ARG, "rs_assembly_matrix: null");
    *A = &S->A;
    return RS_OK;
}

extern "C" int rs_assemble_heat_device(rs_assembly *S, const double *d_k, const double *d_q, const double *d_rhoc,
                                       const double *d_t_node, int32_t transient, double alpha, double dt, double *d_b)
{
    if (!S || !d_k || !d_q || !d_t_node || !d_b || (transient && !d_rhoc))
        return fail(RS_ERR_INVALID_ARG, "rs_assemble_heat_device: null argument");
    if (S->dpn != 1) return fail(RS_ERR_INVALID_ARG, "rs_assemble_heat_device: assembly has 3 dofs per node");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    EventPair ev;
    if ((rc = timed_begin(ev))) return rc;
    HeatArgs a;
    a.xyz = S->xyz.p;
    a.conn = reinterpret_cast<const uint4 *>(S->conn.p);
    a.perm = S->perm.p;
    a.map = S->map.p;
    a.book = S->dbook.p;
    a.k = d_k;
    a.q = d_q;
    a.rhoc = d_rhoc;
    a.t_node = d_t_node;
    a.val = S->A.val.p;
    a.b = d_b;
    a.transient = transient;
    a.alpha = alpha;
    a.dt = dt;
    if (S->rows && S->use_rows) {
        RowRecords rec{ S->rec_grad.p, S->rec_vol.p, S->rec_elem.p, S->rec_slots.p, S->rec_nodes.p, S->rec_base.p,
                        S->rec_count.p };
        HeatRowArgs ra{ a, S->A.rowptr.p, rec, S->n_dof };
        if (S->n_dof) k_heat_rows<<<grid_xcd(S->n_dof, kBlock), kBlock, 0, st>>>(ra);
    } else {
        if (S->A.nnz) k_zero<<<vblocks(S->A.nnz), kBlock, 0, st>>>(S->A.val.p, S->A.nnz);
        if (S->n_dof) k_zero<<<vblocks(S->n_dof), kBlock, 0, st>>>(d_b, S->n_dof);
        for (uint32_t c = 0; c < S->n_colours; c++) {
            uint32_t e0 = S->colour_off[c], e1 = S->colour_off[c + 1];
            if (e1 > e0) k_heat_colour<<<grid_for(e1 - e0, kBlock), kBlock, 0, st>>>(a, e0, e1);
        }
    }
    RS_HIP(hipGetLastError());
    if ((rc = timed_end(ev, S->last_ms))) return rc;
    if ((rc = matrix_frobenius(S->A))) return rc;
    matrix_values_changed(S->A);
    return RS_OK;
}

extern "C" int rs_assemble_heat(rs_assembly *S, const double *k, const double *q, const double *rhoc,
                                const double *t_node, int32_t transient, double alpha, double dt, double *b)
{
    if (!S || !k || !q || !t_node || !b || (transient && !rhoc))
        return fail(RS_ERR_INVALID_ARG, "rs_assemble_heat: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    uint64_t ne = S->n_elems;
    if ((rc = S->fields.alloc(3 * ne + 1)) || (rc = S->t_node.alloc(S->n_nodes + 1ull)) || (rc = S->bvec.alloc(S->n_dof + 1ull)))
        return rc;
    RS_HIP(hipMemcpyAsync(S->fields.p, k, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->fields.p + ne, q, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    if (transient) RS_HIP(hipMemcpyAsync(S->fields.p + 2 * ne, rhoc, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->t_node.p, t_node, sizeof(double) * S->n_nodes, hipMemcpyHostToDevice, st));
    rc = rs_assemble_heat_device(S, S->fields.p, S->fields.p + ne, S->fields.p + 2 * ne, S->t_node.p, transient, alpha,
                                 dt, S->bvec.p);
    if (rc) return rc;
    RS_HIP(hipMemcpyAsync(b, S->bvec.p, sizeof(double) * S->n_dof, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

// Modal analysis: the mass matrix M shares A's pattern (its own copy of rowptr/col and layouts).
static int ensure_mass(rs_assembly *S)
{
    if (S->has_mass) return RS_OK;
    rs_matrix &M = S->M;
    const rs_matrix &A = S->A;
    hipStream_t st = runtime().stream;
    M.n = A.n;
    M.nnz = A.nnz;
    M.borrowed = true;
    int rc;
    if ((rc = M.rowptr.alloc(A.n + 1ull)) || (rc = M.col.alloc(A.nnz)) || (rc = M.val.alloc(A.nnz))) return rc;
    std::vector<uint32_t> rowptr(A.n + 1ull);
    RS_HIP(hipMemcpyAsync(M.rowptr.p, A.rowptr.p, sizeof(uint32_t) * rowptr.size(), hipMemcpyDeviceToDevice, st));
    if (A.nnz) RS_HIP(hipMemcpyAsync(M.col.p, A.col.p, sizeof(uint32_t) * A.nnz, hipMemcpyDeviceToDevice, st));
    RS_HIP(hipMemcpyAsync(rowptr.data(), A.rowptr.p, sizeof(uint32_t) * rowptr.size(), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    if ((rc = matrix_set_layout(M, rowptr.data()))) return rc;
    S->has_mass = true;
    return RS_OK;
}

// RSolverStress::prepare (src/rsolverstress.cpp:54-724: traction surfaces :384-571, volume
// :573-715) + assemblyMatrix (:1654-1772) for a static, transient or modal problem.
static int stress_assemble(rs_assembly *S, const rs_stress_conf &conf, const double *E, const double *nu,
                           const double *rho, const double *alpha_t, const double *dT, const double *u_node,
                           uint32_t n_tri, const uint32_t *tris, const double *traction, double *b, const char *who)
{
    if (!S || !E || !nu || !rho || !u_node || !b || (n_tri && (!tris || !traction)) || (!alpha_t != !dT))
        return fail(RS_ERR_INVALID_ARG, std::string(who) + ": null argument");
    if (S->dpn != 3) return fail(RS_ERR_INVALID_ARG, std::string(who) + ": assembly needs 3 dofs per node");
    if (conf.problem < RS_STRESS_STATIC || conf.problem > RS_STRESS_MODAL)
        return fail(RS_ERR_INVALID_ARG, std::string(who) + ": unknown problem type");
    for (uint64_t i = 0; i < 3ull * n_tri; i++)
        if (tris[i] >= S->n_nodes) return fail(RS_ERR_INVALID_ARG, std::string(who) + ": triangle node out of range");
    int rc = require_runtime();
    if (rc) return rc;
    const bool transient = conf.problem == RS_STRESS_TRANSIENT, modal = conf.problem == RS_STRESS_MODAL;
    if (modal && (rc = ensure_mass(S))) return rc;
    hipStream_t st = runtime().stream;
    uint64_t ne = S->n_elems;
    const bool thermal = alpha_t != nullptr;
    if ((rc = S->fields.alloc((thermal ? 5 : 3) * ne + 1)) || (rc = S->t_node.alloc(3ull * S->n_nodes + 1)) ||
        (rc = S->bvec.alloc(S->n_dof + 1ull)))
        return rc;
    RS_HIP(hipMemcpyAsync(S->fields.p, E, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->fields.p + ne, nu, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->fields.p + 2 * ne, rho, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    if (thermal) {
        RS_HIP(hipMemcpyAsync(S->fields.p + 3 * ne, alpha_t, sizeof(double) * ne, hipMemcpyHostToDevice, st));
        RS_HIP(hipMemcpyAsync(S->fields.p + 4 * ne, dT, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    }
    RS_HIP(hipMemcpyAsync(S->t_node.p, u_node, sizeof(double) * 3 * S->n_nodes, hipMemcpyHostToDevice, st));
    // colour the traction triangles (node-disjoint) on the host
    std::vector<uint16_t> fcol;
    uint32_t nfc = 0;
    std::vector<uint32_t> fsorted(3ull * n_tri), foff;
    std::vector<double> tsorted(3ull * n_tri);
    if (n_tri) {
        if ((rc = colour_elements(S->n_nodes, n_tri, 3, tris, fcol, nfc))) return rc;
        foff.assign(nfc + 1, 0);
        for (uint32_t f = 0; f < n_tri; f++) foff[fcol[f] + 1]++;
        for (uint32_t c = 0; c < nfc; c++) foff[c + 1] += foff[c];
        std::vector<uint32_t> fill(foff.begin(), foff.end() - 1);
        for (uint32_t f = 0; f < n_tri; f++) {
            uint32_t s = fill[fcol[f]]++;
            for (int i = 0; i < 3; i++) {
                fsorted[3 * s + i] = tris[3 * f + i];
                tsorted[3 * s + i] = traction[3 * f + i];
            }
        }
        if ((rc = S->tris.alloc(3ull * n_tri)) || (rc = S->trac.alloc(3ull * n_tri))) return rc;
        RS_HIP(hipMemcpyAsync(S->tris.p, fsorted.data(), sizeof(uint32_t) * 3 * n_tri, hipMemcpyHostToDevice, st));
        RS_HIP(hipMemcpyAsync(S->trac.p, tsorted.data(), sizeof(double) * 3 * n_tri, hipMemcpyHostToDevice, st));
    }
    EventPair ev;
    if ((rc = timed_begin(ev))) return rc;
    StressArgs a;
    a.xyz = S->xyz.p;
    a.conn = reinterpret_cast<const uint4 *>(S->conn.p);
    a.perm = S->perm.p;
    a.map = S->map.p;
    a.book = S->dbook.p;
    a.E = S->fields.p;
    a.nu = S->fields.p + ne;
    a.rho = S->fields.p + 2 * ne;
    a.alpha_t = thermal ? S->fields.p + 3 * ne : nullptr;
    a.dT = thermal ? S->fields.p + 4 * ne : nullptr;
    a.gx = conf.gravity[0];
    a.gy = conf.gravity[1];
    a.gz = conf.gravity[2];
    a.u_node = S->t_node.p;
    a.val = S->A.val.p;
    a.b = S->bvec.p;
    a.mval = modal ? S->M.val.p : nullptr;
    a.problem = conf.problem;
    a.alpha = conf.alpha;
    a.dt = conf.dt;
    a.rot = S->has_rot ? S->rot.p : nullptr;
    a.rot_active = S->has_rot ? S->rot_active.p : nullptr;
    // the static, unrotated problem gathers rows (every value written once); the others scatter
    // per colour into zeroed values
    const bool gather = S->srows && S->use_rows && conf.problem == RS_STRESS_STATIC && !S->has_rot;
    if (S->A.nnz && !gather) k_zero<<<vblocks(S->A.nnz), kBlock, 0, st>>>(S->A.val.p, S->A.nnz);
    if (modal && S->M.nnz) k_zero<<<vblocks(S->M.nnz), kBlock, 0, st>>>(S->M.val.p, S->M.nnz);
    if (S->n_dof) k_zero<<<vblocks(S->n_dof), kBlock, 0, st>>>(S->bvec.p, S->n_dof);
    // surfaces first, then volumes, as the reference's prepare() (src/rsolverstress.cpp:384, 573);
    // a transient step scales the surface loads by dt like every fe (:1662-1668)
    for (uint32_t c = 0; c < nfc; c++) {
        uint32_t f0 = foff[c], f1 = foff[c + 1];
        if (f1 > f0)
            k_traction_colour<<<grid_for(f1 - f0, kBlock), kBlock, 0, st>>>(S->xyz.p, S->tris.p, S->trac.p, S->dbook.p,
                                                                            S->bvec.p, f0, f1,
                                                                            transient ? conf.dt : 1.0);
    }
    if (gather) {
        StressRowArgs r{};
        r.nodes = S->s_nodes.p;
        r.n_nodes = S->s_count;
        r.inc_ptr = S->inc_ptr.p;
        r.inc = S->inc.p;
        r.nbr_ptr = S->s_nbr_ptr.p;
        r.nbr = reinterpret_cast<const uint2 *>(S->s_nbr.p);
        r.geo = S->s_geo.p;
        r.coef = S->s_coef.p;
        r.book = a.book;
        r.rowptr = S->A.rowptr.p;
        r.thermal = thermal ? 1 : 0;
        r.gx = a.gx;
        r.gy = a.gy;
        r.gz = a.gz;
        r.u_node = a.u_node;
        r.val = a.val;
        r.b = a.b;
        if (ne) k_stress_coef<<<vblocks(ne), kBlock, 0, st>>>(a.E, a.nu, a.rho, a.alpha_t, a.dT, (uint32_t)ne,
                                                            S->s_coef.p);
        if (S->s_count) k_stress_rows<<<grid_xcd(S->s_count, kStressWaves), kBlock, 0, st>>>(r);
    } else {
        constexpr uint32_t EPB = kBlock / 16;
        for (uint32_t c = 0; c < S->n_colours; c++) {
            uint32_t e0 = S->colour_off[c], e1 = S->colour_off[c + 1];
            if (e1 > e0) k_stress_colour<<<grid_for(e1 - e0, EPB), kBlock, 0, st>>>(a, e0, e1);
        }
    }
    RS_HIP(hipGetLastError());
    if ((rc = timed_end(ev, S->last_ms))) return rc;
    if ((rc = matrix_frobenius(S->A))) return rc;
    matrix_values_changed(S->A);
    if (modal) {
        if ((rc = matrix_frobenius(S->M))) return rc;
        matrix_values_changed(S->M);
    }
    RS_HIP(hipMemcpyAsync(b, S->bvec.p, sizeof(double) * S->n_dof, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

extern "C" int rs_assemble_stress(rs_assembly *S, const double *E, const double *nu, const double *rho,
                                  const double *gravity, const double *u_node, uint32_t n_tri, const uint32_t *tris,
                                  const double *traction, double *b)
{
    if (!gravity) return fail(RS_ERR_INVALID_ARG, "rs_assemble_stress: null argument");
    rs_stress_conf conf{ RS_STRESS_STATIC, 1.0, 1.0, { gravity[0], gravity[1], gravity[2] } };
    return stress_assemble(S, conf, E, nu, rho, nullptr, nullptr, u_node, n_tri, tris, traction, b,
                           "rs_assemble_stress");
}

extern "C" int rs_assemble_stress_ex(rs_assembly *S, const rs_stress_conf *conf, const double *E, const double *nu,
                                     const double *rho, const double *thermal_expansion, const double *dT,
                                     const double *u_node, uint32_t n_tri, const uint32_t *tris,
                                     const double *traction, double *b)
{
    if (!conf) return fail(RS_ERR_INVALID_ARG, "rs_assemble_stress_ex: null argument");
    return stress_assemble(S, *conf, E, nu, rho, thermal_expansion, dT, u_node, n_tri, tris, traction, b,
                           "rs_assemble_stress_ex");
}

// RSolverGeneric local rotations (src/rsolvergeneric.cpp:353-365, RLocalRotation::activate): R per
// node (row-major), its inverse for the write-back; NULL R clears them.
extern "C" int rs_assembly_set_rotations(rs_assembly *S, const double *R, const uint8_t *active)
{
    if (!S) return fail(RS_ERR_INVALID_ARG, "rs_assembly_set_rotations: null");
    if (S->dpn != 3) return fail(RS_ERR_INVALID_ARG, "rs_assembly_set_rotations: assembly needs 3 dofs per node");
    if (!R) {
        S->has_rot = false;
        return RS_OK;
    }
    if (!active) return fail(RS_ERR_INVALID_ARG, "rs_assembly_set_rotations: null active flags");
    const uint64_t nn = S->n_nodes;
    std::vector<double> inv(9 * nn, 0.0);
    for (uint64_t i = 0; i < nn; i++) {
        if (!active[i]) continue;
        const double *r = R + 9 * i;
        const double det = r[0] * (r[4] * r[8] - r[5] * r[7]) - r[1] * (r[3] * r[8] - r[5] * r[6]) +
                           r[2] * (r[3] * r[7] - r[4] * r[6]);
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
            return fail(RS_ERR_INVALID_ARG, "rs_assembly_set_rotations: singular rotation matrix");
        double *o = inv.data() + 9 * i;
        o[0] = (r[4] * r[8] - r[5] * r[7]) / det;
        o[1] = (r[2] * r[7] - r[1] * r[8]) / det;
        o[2] = (r[1] * r[5] - r[2] * r[4]) / det;
        o[3] = (r[5] * r[6] - r[3] * r[8]) / det;
        o[4] = (r[0] * r[8] - r[2] * r[6]) / det;
        o[5] = (r[2] * r[3] - r[0] * r[5]) / det;
        o[6] = (r[3] * r[7] - r[4] * r[6]) / det;
        o[7] = (r[1] * r[6] - r[0] * r[7]) / det;
        o[8] = (r[0] * r[4] - r[1] * r[3]) / det;
    }
    int rc = require_runtime();
    if (rc) return rc;
    if ((rc = S->rot.alloc(9 * nn)) || (rc = S->irot.alloc(9 * nn)) || (rc = S->rot_active.alloc(nn))) return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(S->rot.p, R, sizeof(double) * 9 * nn, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->irot.p, inv.data(), sizeof(double) * 9 * nn, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(S->rot_active.p, active, nn, hipMemcpyHostToDevice, st));
    RS_HIP(hipStreamSynchronize(st));
    S->has_rot = true;
    return RS_OK;
}

extern "C" int rs_stress_set_displacement_device(rs_assembly *S, const double *d_x, double *d_u_node)
{
    if (!S || !d_x || !d_u_node) return fail(RS_ERR_INVALID_ARG, "rs_stress_set_displacement_device: null");
    if (S->dpn != 3)
        return fail(RS_ERR_INVALID_ARG, "rs_stress_set_displacement_device: assembly needs 3 dofs per node");
    int rc = require_runtime();
    if (rc) return rc;
    k_stress_set_displacement<<<vblocks(S->n_nodes), kBlock, 0, runtime().stream>>>(
        S->dbook.p, S->n_nodes, d_x, S->has_rot ? S->irot.p : nullptr, S->has_rot ? S->rot_active.p : nullptr,
        d_u_node);
    RS_HIP(hipGetLastError());
    return RS_OK;
}

extern "C" int rs_assembly_mass_matrix(rs_assembly *S, rs_matrix **M)
{
    if (!S || !M) return fail(RS_ERR_INVALID_ARG, "rs_assembly_mass_matrix: null");
    if (!S->has_mass) return fail(RS_ERR_INVALID_ARG, "rs_assembly_mass_matrix: no modal assembly yet");
    *M = &S->M;
    return RS_OK;
}

extern "C" int rs_assembly_dofs_to_nodes(rs_assembly *S, const double *d_x, double *d_node)
{
    if (!S || !d_x || !d_node) return fail(RS_ERR_INVALID_ARG, "rs_assembly_dofs_to_nodes: null");
    int rc = require_runtime();
    if (rc) return rc;
    uint64_t npos = (uint64_t)S->n_nodes * S->dpn;
    k_dofs_to_nodes<<<vblocks(npos), kBlock, 0, runtime().stream>>>(S->dbook.p, npos, d_x, d_node);
    RS_HIP(hipGetLastError());
    return RS_OK;
}

extern "C" int rs_assembly_destroy(rs_assembly *S)
{
    delete S;
    return RS_OK;
}

static int ensure_conn0(rs_assembly *S);

extern "C" int rs_heat_flux_device(rs_assembly *S, const double *d_k, const double *d_t_node, double *d_flux)
{
    if (!S || !d_k || !d_t_node || !d_flux) return fail(RS_ERR_INVALID_ARG, "rs_heat_flux_device: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    EventPair ev;
    if ((rc = timed_begin(ev))) return rc;
    if ((rc = ensure_conn0(S))) return rc;
    if (S->n_elems)
        k_heat_flux<<<grid_for(S->n_elems, kBlock), kBlock, 0, runtime().stream>>>(
            S->xyz.p, reinterpret_cast<const uint4 *>(S->conn0.p), S->n_elems, d_k, d_t_node, d_flux);
    RS_HIP(hipGetLastError());
    return timed_end(ev, S->last_ms);
}

extern "C" int rs_stress_process_device(rs_assembly *S, const double *d_E, const double *d_nu, const double *d_u_node,
                                        double *d_normal, double *d_shear, double *d_von_mises, double *d_node_force)
{
    if (!S || !d_E || !d_nu || !d_u_node || !d_normal || !d_shear || !d_von_mises || !d_node_force)
        return fail(RS_ERR_INVALID_ARG, "rs_stress_process_device: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    EventPair ev;
    if ((rc = timed_begin(ev))) return rc;
    k_zero<<<vblocks(3ull * S->n_nodes), kBlock, 0, st>>>(d_node_force, 3ull * S->n_nodes);
    StressPostArgs a{ S->xyz.p, reinterpret_cast<const uint4 *>(S->conn.p), S->perm.p, d_E, d_nu, d_u_node,
                      d_normal, d_shear, d_von_mises, d_node_force };
    for (uint32_t c = 0; c < S->n_colours; c++) {
        const uint32_t e0 = S->colour_off[c], e1 = S->colour_off[c + 1];
        if (e1 > e0) k_stress_post<<<grid_for(e1 - e0, kBlock), kBlock, 0, st>>>(a, e0, e1);
    }
    RS_HIP(hipGetLastError());
    return timed_end(ev, S->last_ms);
}

// The caller's element order of the connectivity, built on first use.
static int ensure_conn0(rs_assembly *S)
{
    if (!S->n_elems || S->conn0.p) return RS_OK;
    int rc = S->conn0.alloc(4ull * S->n_elems);
    if (rc) return rc;
    k_unsort_conn<<<vblocks(S->n_elems), kBlock, 0, runtime().stream>>>(
        reinterpret_cast<const uint4 *>(S->conn.p), S->perm.p, S->n_elems, reinterpret_cast<uint4 *>(S->conn0.p));
    RS_HIP(hipGetLastError());
    return RS_OK;
}

extern "C" int rs_assemble_electrostatics_device(rs_assembly *S, const double *d_eps_r, const double *d_rho,
                                                 const double *d_potential, double *d_b)
{
    if (!S || !d_eps_r || !d_rho || !d_potential || !d_b)
        return fail(RS_ERR_INVALID_ARG, "rs_assemble_electrostatics_device: null argument");
    if (S->dpn != 1) return fail(RS_ERR_INVALID_ARG, "rs_assemble_electrostatics_device: assembly has 3 dofs per node");
    int rc = require_runtime();
    if (rc) return rc;
    const uint64_t ne = S->n_elems;
    if ((rc = S->coeff.alloc(2 * ne + 1))) return rc;
    if (ne) k_electro_coeffs<<<vblocks(ne), kBlock, 0, runtime().stream>>>(d_eps_r, d_rho, (uint32_t)ne, S->coeff.p,
                                                                            S->coeff.p + ne);
    RS_HIP(hipGetLastError());
    return rs_assemble_heat_device(S, S->coeff.p, S->coeff.p + ne, nullptr, d_potential, 0, 1.0, 1.0, d_b);
}

extern "C" int rs_electrostatics_process_device(rs_assembly *S, const double *d_eps_r, const double *d_sigma,
                                                const double *d_potential, double *d_field, double *d_current,
                                                double *d_energy, double *d_resistivity, double *d_joule)
{
    if (!S || !d_eps_r || !d_sigma || !d_potential || !d_field || !d_current || !d_energy || !d_resistivity || !d_joule)
        return fail(RS_ERR_INVALID_ARG, "rs_electrostatics_process_device: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    EventPair ev;
    if ((rc = timed_begin(ev))) return rc;
    if ((rc = ensure_conn0(S))) return rc;
    ElectroPostArgs a{ S->xyz.p, reinterpret_cast<const uint4 *>(S->conn0.p), S->n_elems, d_eps_r, d_sigma,
                       d_potential, d_field, d_current, d_energy, d_resistivity, d_joule };
    if (S->n_elems) k_electro_post<<<grid_for(S->n_elems, kBlock), kBlock, 0, runtime().stream>>>(a);
    RS_HIP(hipGetLastError());
    return timed_end(ev, S->last_ms);
}

extern "C" int rs_assemble_magnetostatics_device(rs_assembly *S, const double *d_current_node,
                                                 const double *d_field_node, double *d_b)
{
    if (!S || !d_current_node || !d_field_node || !d_b)
        return fail(RS_ERR_INVALID_ARG, "rs_assemble_magnetostatics_device: null argument");
    if (S->dpn != 3) return fail(RS_ERR_INVALID_ARG, "rs_assemble_magnetostatics_device: assembly needs 3 dofs per node");
    if (S->whole_nodes < 0) {
        // the reference's magnetostatic node book is per node (src/rsolvermagnetostatics.cpp:
        // 241-261): a node's field is prescribed or free as a whole
        S->whole_nodes = 1;
        for (uint32_t v = 0; v < S->n_nodes && S->whole_nodes; v++) {
            const uint32_t nfix = (S->book[3ull * v] == RS_EOD) + (S->book[3ull * v + 1] == RS_EOD) +
                                  (S->book[3ull * v + 2] == RS_EOD);
            if (nfix != 0 && nfix != 3) S->whole_nodes = 0;
        }
    }
    if (!S->whole_nodes)
        return fail(RS_ERR_INVALID_ARG,
                    "rs_assemble_magnetostatics_device: a node has some field components prescribed and others free; "
                    "magnetostatics prescribes whole nodes");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    EventPair ev;
    if ((rc = timed_begin(ev))) return rc;
    if (S->srows && S->use_rows) {
        MagnetoRowArgs r{ S->s_nodes.p, S->s_count, S->inc_ptr.p, S->inc.p, S->s_nbr_ptr.p,
                          reinterpret_cast<const uint2 *>(S->s_nbr.p), S->s_fmask.p, S->s_geo.p, S->dbook.p,
                          S->A.rowptr.p, d_current_node, d_field_node, S->A.val.p, d_b };
        if (S->n_dof) k_zero<<<vblocks(S->n_dof), kBlock, 0, st>>>(d_b, S->n_dof);   // rows of no free node
        if (S->s_count) k_magneto_rows<<<grid_xcd(S->s_count, kBlock), kBlock, 0, st>>>(r);
    } else {
        if (S->A.nnz) k_zero<<<vblocks(S->A.nnz), kBlock, 0, st>>>(S->A.val.p, S->A.nnz);
        if (S->n_dof) k_zero<<<vblocks(S->n_dof), kBlock, 0, st>>>(d_b, S->n_dof);
        MagnetoArgs a{ S->xyz.p, reinterpret_cast<const uint4 *>(S->conn.p), S->map.p, S->dbook.p, d_current_node,
                       d_field_node, S->A.val.p, d_b };
        for (uint32_t c = 0; c < S->n_colours; c++) {
            const uint32_t e0 = S->colour_off[c], e1 = S->colour_off[c + 1];
            if (e1 > e0) k_magneto_colour<<<grid_for(e1 - e0, kBlock), kBlock, 0, st>>>(a, e0, e1);
        }
    }
    RS_HIP(hipGetLastError());
    if ((rc = timed_end(ev, S->last_ms))) return rc;
    if ((rc = matrix_frobenius(S->A))) return rc;
    matrix_values_changed(S->A);
    return RS_OK;
}

// RSolverFluid::prepare, volume elements (src/rsolverfluid.cpp:289-497 with
// computeElementConstantDerivative :1845-2201 and assemblyMatrix :2399-2419).
extern "C" int rs_assemble_fluid_device(rs_assembly *S, const rs_fluid_conf *conf, const double *d_rho,
                                        const double *d_mu, const double *d_gravity, const double *d_state,
                                        const double *d_v_old, const double *d_acc, double *d_b)
{
    if (!S || !conf || !d_rho || !d_mu || !d_gravity || !d_state || !d_v_old || !d_b || (conf->unsteady && !d_acc))
        return fail(RS_ERR_INVALID_ARG, "rs_assemble_fluid_device: null argument");
    if (S->dpn != 4) return fail(RS_ERR_INVALID_ARG, "rs_assemble_fluid_device: assembly needs 4 dofs per node");
    if (!(conf->stream_velocity > 0.0))
        return fail(RS_ERR_INVALID_ARG, "rs_assemble_fluid_device: stream velocity must be positive");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    EventPair ev;
    if ((rc = timed_begin(ev))) return rc;
    if (S->A.nnz) k_zero<<<vblocks(S->A.nnz), kBlock, 0, st>>>(S->A.val.p, S->A.nnz);
    if (S->n_dof) k_zero<<<vblocks(S->n_dof), kBlock, 0, st>>>(d_b, S->n_dof);
    FluidArgs a{ S->xyz.p, reinterpret_cast<const uint4 *>(S->conn.p), S->perm.p, S->map.p,
                 reinterpret_cast<const uint4 *>(S->dbook.p), d_rho, d_mu, d_gravity, d_state, d_v_old,
                 conf->unsteady ? d_acc : d_v_old, S->A.val.p, d_b, conf->stream_velocity, conf->alpha, conf->dt,
                 conf->unsteady ? 1 : 0 };
    for (uint32_t c = 0; c < S->n_colours; c++) {
        const uint32_t e0 = S->colour_off[c], e1 = S->colour_off[c + 1];
        if (e1 > e0) k_fluid_colour<<<grid_for((uint64_t)(e1 - e0) * kFluidLanes, kBlock), kBlock, 0, st>>>(a, e0, e1);
    }
    RS_HIP(hipGetLastError());
    if ((rc = timed_end(ev, S->last_ms))) return rc;
    if ((rc = matrix_frobenius(S->A))) return rc;
    matrix_values_changed(S->A);
    return RS_OK;
}

// RSolverFluid::solve after the linear solve (src/rsolverfluid.cpp:519-624).
extern "C" int rs_fluid_update_device(rs_assembly *S, const double *d_x, int32_t unsteady, double dt,
                                      double *d_state, double *d_v_old, double *d_acc, double *norms)
{
    if (!S || !d_x || !d_state || !d_v_old || (unsteady && !d_acc))
        return fail(RS_ERR_INVALID_ARG, "rs_fluid_update_device: null argument");
    if (S->dpn != 4) return fail(RS_ERR_INVALID_ARG, "rs_fluid_update_device: assembly needs 4 dofs per node");
    if (unsteady && !(dt > 0.0)) return fail(RS_ERR_INVALID_ARG, "rs_fluid_update_device: dt must be positive");
    int rc = require_runtime();
    if (rc) return rc;
    const uint32_t nb = vblocks(S->n_nodes);
    if (S->part.n < 4ull * nb && (rc = S->part.alloc(4ull * nb))) return rc;
    hipStream_t st = runtime().stream;
    k_fluid_update<<<nb, kBlock, 0, st>>>(reinterpret_cast<const uint4 *>(S->dbook.p), S->n_nodes, d_x,
                                          unsteady ? 1 : 0, unsteady ? 1.0 / dt : 0.0, d_state, d_v_old,
                                          unsteady ? d_acc : d_v_old, S->part.p);
    RS_HIP(hipGetLastError());
    std::vector<double> part(4ull * nb);
    RS_HIP(hipMemcpyAsync(part.data(), S->part.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    double sum[4] = { 0, 0, 0, 0 };
    for (uint32_t blk = 0; blk < nb; blk++)
        for (int k = 0; k < 4; k++) sum[k] += part[4ull * blk + k];
    if (norms) {
        norms[0] = std::sqrt(sum[0]);
        norms[1] = std::sqrt(sum[1]);
        norms[2] = std::sqrt(sum[2]);
        norms[3] = std::sqrt(sum[3]);
    }
    return RS_OK;
}

// RSolverFluidHeat / RSolverFluidParticle prepare (volume elements) on a 1-DOF assembly.
extern "C" int rs_assemble_transport_device(rs_assembly *S, const rs_transport_conf *conf, const double *d_ca,
                                            const double *d_k, const double *d_velocity, const double *d_node_source,
                                            const double *d_node_field, double *d_b)
{
    if (!S || !conf || !d_ca || !d_k || !d_velocity || !d_node_source || !d_node_field || !d_b)
        return fail(RS_ERR_INVALID_ARG, "rs_assemble_transport_device: null argument");
    if (S->dpn != 1) return fail(RS_ERR_INVALID_ARG, "rs_assemble_transport_device: assembly needs 1 dof per node");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    EventPair ev;
    if ((rc = timed_begin(ev))) return rc;
    TransportArgs a{ S->xyz.p, reinterpret_cast<const uint4 *>(S->conn.p), S->perm.p, S->map.p, S->dbook.p, d_ca, d_k,
                     d_velocity, d_node_source, d_node_field, S->A.val.p, d_b, conf->alpha, conf->dt,
                     conf->unsteady ? 1 : 0 };
    if (S->rows && S->use_rows) {
        if (S->telem.n < (uint64_t)kTElemWords * S->n_elems &&
            (rc = S->telem.alloc((uint64_t)kTElemWords * std::max(S->n_elems, 1u))))
            return rc;
        if (S->n_elems) k_transport_elem<<<vblocks(S->n_elems), kBlock, 0, st>>>(a, S->n_elems, S->telem.p);
        RowRecords rec{ S->rec_grad.p, S->rec_vol.p, S->rec_elem.p, S->rec_slots.p, S->rec_nodes.p, S->rec_base.p,
                        S->rec_count.p };
        TransportRowArgs ra{ S->A.rowptr.p, S->dof_node.p, rec, S->n_dof, S->telem.p, d_node_source, d_node_field,
                             S->A.val.p, d_b, conf->alpha, conf->dt, conf->unsteady ? 1 : 0 };
        if (S->n_dof) k_transport_rows<<<grid_xcd(S->n_dof, kBlock), kBlock, 0, st>>>(ra);
    } else {
        if (S->A.nnz) k_zero<<<vblocks(S->A.nnz), kBlock, 0, st>>>(S->A.val.p, S->A.nnz);
        if (S->n_dof) k_zero<<<vblocks(S->n_dof), kBlock, 0, st>>>(d_b, S->n_dof);
        for (uint32_t c = 0; c < S->n_colours; c++) {
            const uint32_t e0 = S->colour_off[c], e1 = S->colour_off[c + 1];
            if (e1 > e0) k_transport_colour<<<grid_for(e1 - e0, kBlock), kBlock, 0, st>>>(a, e0, e1);
        }
    }
    RS_HIP(hipGetLastError());
    if ((rc = timed_end(ev, S->last_ms))) return rc;
    if ((rc = matrix_frobenius(S->A))) return rc;
    matrix_values_changed(S->A);
    return RS_OK;
}

extern "C" int rs_transport_update_device(rs_assembly *S, const double *d_x, double *d_node_field, double *norms)
{
    if (!S || !d_x || !d_node_field) return fail(RS_ERR_INVALID_ARG, "rs_transport_update_device: null argument");
    if (S->dpn != 1) return fail(RS_ERR_INVALID_ARG, "rs_transport_update_device: assembly needs 1 dof per node");
    int rc = require_runtime();
    if (rc) return rc;
    const uint32_t nb = vblocks(S->n_nodes);
    if (S->part.n < 2ull * nb && (rc = S->part.alloc(2ull * nb))) return rc;
    hipStream_t st = runtime().stream;
    k_transport_update<<<nb, kBlock, 0, st>>>(S->dbook.p, S->n_nodes, d_x, d_node_field, S->part.p);
    RS_HIP(hipGetLastError());
    std::vector<double> part(2ull * nb);
    RS_HIP(hipMemcpyAsync(part.data(), S->part.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    double sum[2] = { 0, 0 };
    for (uint32_t blk = 0; blk < nb; blk++) {
        sum[0] += part[2ull * blk];
        sum[1] += part[2ull * blk + 1];
    }
    if (norms) {
        norms[0] = std::sqrt(sum[0]);
        norms[1] = std::sqrt(sum[1]);
    }
    return RS_OK;
}
