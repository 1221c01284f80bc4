// rs_learned.inc.hip -- learned multilevel preconditioner (no counterpart in the reference).
// Included at the end of rs_solve.hip: it shares that translation unit's solver state, graph
// replay, finalize kernel and SpMV kernels.
//
// Preconditioner: one V-cycle over an aggregation hierarchy, applied inside flexible CG.
//   levels    caller-given aggregations (for example nested coordinate bins of 2x2x2 nodes);
//             piecewise-constant prolongation P_l; Galerkin operators A_{l+1} = P_l^T A_l P_l
//             summed in FP64 in a fixed order (one gather list per coarse entry), so the
//             set-up is deterministic
//   smoother  Chebyshev polynomial in D^-1 A_l over [lambda/30, 1.05 lambda], lambda the
//             Gershgorin bound of D^-1 A_l (Adams, Brezina, Hu, Tuminaro 2003); FP32 values
//             and vectors, each step one fused SpMV kernel
//   coarsest  the learned part: a 6-layer, 64-wide message-passing network over the coarsest
//             graph (<= 512 nodes) maps the coarsest residual r_c to a correction e.  Other
//             modes: coarse Jacobi (what the untrained network computes) and the exact dense
//             inverse (the ceiling a coarse solve can reach).
// Network: features [u/sigma, log d - mean, 1, 0] with u = r_c / diag(A_c), sigma = rms(u);
// messages are neighbour features weighted by -A_c[I,J]/A_c[I,I]; Y = H Ws + M Wn + b,
// H += relu(Y); BF16 weights and activations on gfx950's K = 32 MFMA (v_mfma_f32_16x16x32_bf16;
// the K = 16 form for an odd last 16 of K), FP32 accumulation; output e = u + sigma * (H wo + bo).
// Inference: the embedding and one launch per layer, one workgroup per 16-node strip of the
// coarsest graph.
// Training (one workgroup per sample, activations in global memory): forward, the loss
// (e.A_c.e - 2 e.r_c)/s + 1 with s = r_c.A_c^-1.r_c -- the A_c-norm error fraction the
// correction leaves -- backward, per-sample gradients, a fixed-order batch mean and Adam.
// The solve is flexible CG (Notay 2000, FCG(1)) in mixed precision with FP64 residual
// replacement; Jacobi-CG finishes the solve if the flexible iteration stops unconverged.

namespace rs {

constexpr int kHid = 64;
constexpr int kNin = 4;
constexpr int kLayers = 6;
constexpr uint32_t kMaxCoarse = 512;
constexpr uint32_t kMaxDeg = 4;          // Chebyshev steps per smoothing
constexpr uint32_t kLayerStride = 2 * kHid * kHid + kHid;
constexpr uint32_t kOffWin = 0, kOffBin = kNin * kHid, kOffL0 = kNin * kHid + kHid;
constexpr uint32_t kOffWo = kOffL0 + kLayers * kLayerStride, kOffBo = kOffWo + kHid;
constexpr uint32_t kNumParams = kOffBo + 1;   // 49,921
constexpr uint32_t kBfPerLayer = 2 * kHid * kHid;
__host__ __device__ constexpr uint32_t off_ws(int l) { return kOffL0 + l * kLayerStride; }
__host__ __device__ constexpr uint32_t off_wn(int l) { return off_ws(l) + kHid * kHid; }
__host__ __device__ constexpr uint32_t off_b(int l) { return off_ws(l) + 2 * kHid * kHid; }

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

// One 16x16 output tile: acc += A(16 x K) B(K x 16), K a multiple of 16, operands fetched as
// BF16 bit patterns through a(row, k) and b(k, col): 32 of K per gfx950 K = 32 MFMA (lane l
// supplies A[l%16][8(l/16)+i] and B[8(l/16)+i][l%16], i < 8), a last 16 on the K = 16 form
// (A[l%16][4(l/16)+i], B[4(l/16)+i][l%16], i < 4); acc[i] is C[4(l/16)+i][l%16].
template <class FA, class FB>
__device__ __forceinline__ f32x4 mfma_tile(FA a, FB b, int K, f32x4 acc)
{
    const int lane = threadIdx.x & 63, rc = lane & 15, q = lane >> 4;
    int k0 = 0;
    for (; k0 + 32 <= K; k0 += 32) {
        bf16x8 av, bv;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            av[i] = __builtin_bit_cast(__bf16, (unsigned short)a(rc, k0 + 8 * q + i));
            bv[i] = __builtin_bit_cast(__bf16, (unsigned short)b(k0 + 8 * q + i, rc));
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    if (k0 < K) {
        bf16x4 av, bv;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            av[i] = (short)a(rc, k0 + 4 * q + i);
            bv[i] = (short)b(k0 + 4 * q + i, rc);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// Sum over the workgroup, result in every thread.
__device__ __forceinline__ float block_sum_f(float v, float *red, int nthreads)
{
    v = group_sum<kWave>(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    float s = 0.0f;
    for (int w = 0; w < nthreads / kWave; w++) s += red[w];
    __syncthreads();
    return s;
}

// Workgroup maximum of m, stored by thread 0 at partials[blockIdx.x].
__device__ __forceinline__ void block_max_store(double m, double *partials)
{
    __shared__ double red[kBlock / kWave];
    for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; w++) t = fmax(t, red[w]);
        partials[blockIdx.x] = t;
    }
}

constexpr int kMaxFinBlock = 1024;
__global__ __launch_bounds__(kMaxFinBlock) void k_max_fin(const double *partials, uint32_t nparts, double *out)
{
    // the maximum does not depend on the order: 1024 threads with independent loads, then the waves
    __shared__ double w[kMaxFinBlock / kWave];
    double m = 0.0;
#pragma unroll 4
    for (uint32_t i = threadIdx.x; i < nparts; i += kMaxFinBlock) m = fmax(m, partials[i]);
    for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) w[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kMaxFinBlock / kWave; k++) t = fmax(t, w[k]);
        *out = t;
    }
}

// ---------------------------------------------------------------------------- hierarchy set-up
// Coarse values: cval[k] = sum of the finer entries listed for coarse entry k, in list order.
__global__ void k_galerkin_gather(const double *__restrict__ fval, const uint32_t *__restrict__ gptr,
                                  const uint32_t *__restrict__ gsrc, uint64_t nnzc, double *__restrict__ cval)
{
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < nnzc; k += (uint64_t)gridDim.x * kBlock) {
        uint32_t j = gptr[k];
        const uint32_t e = gptr[k + 1];
        double s = 0.0;
        // four list entries' loads in flight at once, summed in list order
        for (; j + 4 <= e; j += 4) {
            const uint32_t i0 = gsrc[j], i1 = gsrc[j + 1], i2 = gsrc[j + 2], i3 = gsrc[j + 3];
            const double v0 = fval[i0], v1 = fval[i1], v2 = fval[i2], v3 = fval[i3];
            s += v0;
            s += v1;
            s += v2;
            s += v3;
        }
        for (; j < e; j++) s += fval[gsrc[j]];
        cval[k] = s;
    }
}

// FP32 diagonal and inverse diagonal (0 where the diagonal is absent or zero) and the block
// maxima of the Gershgorin bound sum_j |a_ij| / |a_ii| of D^-1 A; 16 lanes per row.
__global__ __launch_bounds__(kBlock) void k_level_diag(const uint32_t *__restrict__ rowptr,
                                                       const uint32_t *__restrict__ col,
                                                       const double *__restrict__ val, uint32_t n,
                                                       float *__restrict__ diag, float *__restrict__ dinv,
                                                       double *__restrict__ partials)
{
    constexpr int G = 16;
    const uint32_t gl = threadIdx.x % G;
    double m = 0.0;
    for (uint64_t i = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / G; i < n; i += (uint64_t)gridDim.x * (kBlock / G)) {
        double d = 0.0, s = 0.0;
        for (uint32_t j = rowptr[i] + gl; j < rowptr[i + 1]; j += G) {
            const double a = val[j];
            s += fabs(a);
            if (col[j] == i) d += a;
        }
        d = group_sum<G>(d);
        s = group_sum<G>(s);
        if (gl == 0) {
            diag[i] = (float)d;
            dinv[i] = d != 0.0 ? (float)(1.0 / d) : 0.0f;
        }
        if (d != 0.0) m = fmax(m, s / fabs(d));
    }
    block_max_store(m, partials);
}

// Coarsest graph for the network: diagonal and message weights -A[I,J]/A[I,I] (0 on the diagonal).
// 16 lanes per row; a row's one diagonal entry plus zeros sums to itself exactly.
__global__ __launch_bounds__(kBlock) void k_coarse_weights(const uint32_t *__restrict__ cptr,
                                                           const uint32_t *__restrict__ ccol,
                                                           const double *__restrict__ cval, uint32_t nc,
                                                           float *__restrict__ cw, float *__restrict__ cdiag)
{
    constexpr int G = 16;
    const uint32_t gl = threadIdx.x % G;
    for (uint64_t I = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / G; I < nc; I += (uint64_t)gridDim.x * (kBlock / G)) {
        const uint32_t p0 = cptr[I], p1 = cptr[I + 1];
        double d = 0.0;
        for (uint32_t p = p0 + gl; p < p1; p += G)
            if (ccol[p] == I) d += cval[p];
        d = group_sum<G>(d);
        if (gl == 0) cdiag[I] = (float)d;
        for (uint32_t p = p0 + gl; p < p1; p += G)
            cw[p] = (ccol[p] == I || d == 0.0) ? 0.0f : (float)(-cval[p] / d);
    }
}

// Fine vector restricted straight to the coarsest level (block per coarsest node, FP64 sum).
__global__ __launch_bounds__(kBlock) void k_restrict_total(const double *__restrict__ r, const uint32_t *__restrict__ mptr,
                                                           const uint32_t *__restrict__ mrows, float *__restrict__ rc)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t I = blockIdx.x;
    double v[1] = { 0.0 };
    for (uint32_t k = mptr[I] + threadIdx.x; k < mptr[I + 1]; k += kBlock) v[0] += r[mrows[k]];
    block_sum<1, kBlock>(v, lds);
    if (threadIdx.x == 0) rc[I] = (float)v[0];
}

// ---------------------------------------------------------------------------- cycle kernels
// Prolongation fused into the kernel that finishes a coarse correction: every finer row i of
// node I gets zout[i] = zpre[i] + e_I (zpre nullptr: zero); mptr nullptr: nothing to do.
struct Prolong {
    const uint32_t *mptr, *mrows;    // finer rows of each node of this level
    const float *zpre;
    float *zout;
    // One lane: the members eight at a time, every load of a batch issued before its stores (the
    // member rows of a node are distinct, so a store never feeds a later load of the batch).
    __device__ __forceinline__ void apply(uint32_t I, float e) const
    {
        if (!mptr) return;
        const uint32_t *__restrict__ mr = mrows;
        const float *__restrict__ zp = zpre;
        float *__restrict__ zo = zout;
        uint32_t k = mptr[I];
        const uint32_t k1 = mptr[I + 1];
        for (; k + 8 <= k1; k += 8) {
            uint32_t i[8];
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) i[u] = mr[k + u];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = (zp ? zp[i[u]] : 0.0f) + e;
#pragma unroll
            for (int u = 0; u < 8; u++) zo[i[u]] = v[u];
        }
        for (; k < k1; k++) {
            const uint32_t i = mr[k];
            zo[i] = (zp ? zp[i] : 0.0f) + e;
        }
    }
    // The G lanes of a row group (each holding e): member k on lane k mod G.
    template <int G>
    __device__ __forceinline__ void apply_group(uint32_t I, float e, uint32_t gl) const
    {
        if (!mptr) return;
        const uint32_t k1 = mptr[I + 1];
        for (uint32_t k = mptr[I] + gl; k < k1; k += G) {
            const uint32_t i = mrows[k];
            zout[i] = (zpre ? zpre[i] : 0.0f) + e;
        }
    }
};

// The first Chebyshev step of the next level fused into the restriction: z = d = c2 D^-1 r.
struct FirstStep {
    const float *coef, *dinv;        // coef nullptr: nothing to do
    float *z, *d;
};

struct CycArgs {
    CsrView A;                 // level matrix (FP32 values)
    const float *r;            // right-hand side of the level
    const float *zin;          // iterate the SpMV reads
    float *zout, *d, *res;
    const float *dinv;
    const float *coef;         // (c1, c2) of this Chebyshev step
    const SolveState *st;      // nullptr outside a solve
    const float *q;            // finest level, last step: the flexible CG's previous A p
    double *partials;          // finest level, last step: partial r.z and z.q
    Prolong up;                // last step of a coarse level: its correction into the finer level
    uint32_t n;
};

__device__ __forceinline__ bool cyc_idle(const SolveState *st) { return st && st->done; }

// First Chebyshev step from a zero iterate: d = c2 D^-1 r, z = d.
// (4 entries per lane in 16-byte loads and stores, as the flexible CG's vector kernels: with
// 4-byte accesses these FP32 passes are bound by the number of loads, not by their bytes)
__global__ void k_cyc_first(CycArgs a)
{
    if (cyc_idle(a.st)) return;
    const float c2 = a.coef[1];
    const uint64_t nq = a.n / 4, stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < nq; q += stride) {
        const float4 dv = reinterpret_cast<const float4 *>(a.dinv)[q], rv = reinterpret_cast<const float4 *>(a.r)[q];
        const float4 z = make_float4(c2 * dv.x * rv.x, c2 * dv.y * rv.y, c2 * dv.z * rv.z, c2 * dv.w * rv.w);
        reinterpret_cast<float4 *>(a.zout)[q] = z;
        reinterpret_cast<float4 *>(a.d)[q] = z;
    }
    for (uint64_t i = 4 * nq + (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
        float z = c2 * a.dinv[i] * a.r[i];
        a.zout[i] = z;
        a.d[i] = z;
    }
}

// Chebyshev step: res = r - A z_in, d = c1 d + c2 D^-1 res, z_out = z_in + d (c1 = 0: first
// step from a non-zero iterate, d not read).
struct EpiCheb {
    const float *r, *zin, *dinv, *q;
    float *zout, *d;
    float c1, c2;
    bool dots;
    Prolong up;
    double acc[2];
    __device__ void row(uint32_t i, float s)
    {
        const float ri = r[i];
        float dd = c2 * dinv[i] * (ri - s);
        if (c1 != 0.0f) dd = fmaf(c1, d[i], dd);
        d[i] = dd;
        const float z = zin[i] + dd;
        zout[i] = z;
        up.apply(i, z);
        if (dots) {
            acc[0] += (double)ri * (double)z;
            acc[1] += (double)z * (double)q[i];
        }
    }
    // Shapes with a lane group per row: every lane holds the row's sum (a butterfly sum, the same
    // bits on every lane) and computes the same z; lane 0 stores, the group spreads the
    // prolongation over its lanes.
    static constexpr bool kGroupRow = true;
    template <int G>
    __device__ void row_group(uint32_t i, float s, uint32_t gl)
    {
        const float ri = r[i];
        float dd = c2 * dinv[i] * (ri - s);
        if (c1 != 0.0f) dd = fmaf(c1, d[i], dd);
        const float z = zin[i] + dd;
        if (gl == 0) {
            d[i] = dd;
            zout[i] = z;
            if (dots) {
                acc[0] += (double)ri * (double)z;
                acc[1] += (double)z * (double)q[i];
            }
        }
        up.template apply_group<G>(i, z, gl);
    }
};

template <class Shape>
__global__ __launch_bounds__(kBlock) void k_cyc_step(CycArgs a)
{
    if (cyc_idle(a.st)) return;
    EpiCheb e{ a.r, a.zin, a.dinv, a.q, a.zout, a.d, a.coef[0], a.coef[1], a.partials != nullptr, a.up, { 0.0, 0.0 } };
    Shape::template run<float>(a.A, a.A.val32, a.zin, e);
    if (a.partials) store_partials<2>(e.acc, a.partials);
}

struct EpiResid32 {
    const float *r;
    float *res;
    __device__ void row(uint32_t i, float s) { res[i] = r[i] - s; }
};

template <class Shape>
__global__ __launch_bounds__(kBlock) void k_cyc_resid(CycArgs a)
{
    if (cyc_idle(a.st)) return;
    EpiResid32 e{ a.r, a.res };
    Shape::template run<float>(a.A, a.A.val32, a.zin, e);
}

// ---------------------------------------------------------------------------- compressed smoother copy
// Row-scaled copy S = D^-1 A of a level matrix in the SELL-64 slices, FP16 values and int16
// column offsets (col - row) packed into one uint32 per entry: 4 bytes per entry instead of 6 for
// the FP32 SELL copy with offsets.  The entries are stored in quads: entries 4q .. 4q+3 of a lane
// side by side at base + 256 q + 4 lane, the last width % 4 in plain 64-wide rows after the
// quads, so that a lane loads four entries of its row with one 16-byte load.  The smoother only
// needs D^-1 A (and D for the residual); the Krylov operator stays FP32.
__host__ __device__ __forceinline__ uint32_t s16_slot(uint32_t base, uint32_t width, uint32_t j, uint32_t lane)
{
    return j < (width & ~3u) ? base + (j >> 2) * 4 * kWave + 4 * lane + (j & 3u) : base + j * kWave + lane;
}

// *bad is set when an offset does not fit in 16 bits or a diagonal is zero.
__global__ void k_s16_build(CsrView A, uint32_t *__restrict__ ent, uint32_t *__restrict__ bad)
{
    // built from the FP64 SELL slots, so that the 64 lanes of a wavefront (64 consecutive rows)
    // read each entry index coalesced; the diagonal sums the row's entries with c == r in entry
    // order, as the CSR row does
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < (uint64_t)A.nslices * kWave;
         r += (uint64_t)gridDim.x * kBlock) {
        const uint32_t s = (uint32_t)(r / kWave), lane = (uint32_t)(r % kWave);
        const uint32_t base = A.sell_ptr[s], width = (A.sell_ptr[s + 1] - base) / kWave;
        double d = 1.0;
        if (r < A.n) {
            d = 0.0;
            for (uint32_t j = 0; j < width; j++) {
                const uint32_t k = base + j * kWave + lane;
                if (A.sell_col[k] == r) d += A.sell_val[k];
            }
            if (d == 0.0) atomicOr(bad, 1u);
        }
        for (uint32_t j = 0; j < width; j++) {
            const uint32_t k = base + j * kWave + lane, c = A.sell_col[k];
            uint32_t e = 0;
            if (c != kSellPad) {
                const int64_t off = (int64_t)c - (int64_t)r;
                if (off < -32768 || off > 32767) atomicOr(bad, 1u);
                const _Float16 h = (_Float16)(d != 0.0 ? A.sell_val[k] / d : 0.0);
                e = (uint32_t)(uint16_t)(int16_t)off | ((uint32_t)__builtin_bit_cast(uint16_t, h) << 16);
            }
            ent[s16_slot(base, width, j, lane)] = e;
        }
    }
}

typedef uint32_t s16_u4 __attribute__((ext_vector_type(4)));

struct S16View {
    uint32_t n, nslices, spb;
    const uint32_t *sell_ptr, *ent;
};

// s = S x for the lanes' rows (one row per lane, slices of 64), then epi.row(r, s).
template <class Epi>
__device__ __forceinline__ void s16_run(const S16View &S, const float *__restrict__ x, Epi &epi)
{
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const uint32_t s0 = lb * S.spb, s1 = min(S.nslices, s0 + S.spb);
    const uint32_t *__restrict__ ent = S.ent;
    for (uint32_t s = s0 + wid; s < s1; s += kBlock / kWave) {
        const uint32_t b = S.sell_ptr[s], width = (S.sell_ptr[s + 1] - b) / kWave, nq = width >> 2;
        const uint32_t r = s * kWave + lane, rr = min(r, S.n - 1);   // padding lanes read row n-1
        float acc = 0.0f;
        auto term = [&](uint32_t w) {
            const int32_t off = (int16_t)(w & 0xffffu);
            const float v = (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
            return v * x[(int32_t)rr + off];
        };
        // entries 4q .. 4q+3 as one 16-byte load, two quads per trip, summed in entry order
        auto quad = [&](uint32_t q) {
            return __builtin_nontemporal_load(reinterpret_cast<const s16_u4 *>(&ent[b + q * 4 * kWave + 4 * lane]));
        };
        if (width <= 16) {
            // the whole row in one round of loads (mesh rows: ~15 entries), summed in entry order
            s16_u4 w[4];
            uint32_t t[3];
            const uint32_t tail = width - 4 * nq;
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) w[k] = k < nq ? quad(k) : s16_u4{ 0u, 0u, 0u, 0u };
#pragma unroll
            for (uint32_t k = 0; k < 3; k++) t[k] = k < tail ? ent[b + (4 * nq + k) * kWave + lane] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < 4; k++)
                if (k < nq) {
                    acc += term(w[k].x);
                    acc += term(w[k].y);
                    acc += term(w[k].z);
                    acc += term(w[k].w);
                }
#pragma unroll
            for (uint32_t k = 0; k < 3; k++)
                if (k < tail) acc += term(t[k]);
            if (r < S.n) epi.row(r, acc);
            continue;
        }
        uint32_t q = 0;
        for (; q + 2 <= nq; q += 2) {
            const s16_u4 w0 = quad(q), w1 = quad(q + 1);
            acc += term(w0.x);
            acc += term(w0.y);
            acc += term(w0.z);
            acc += term(w0.w);
            acc += term(w1.x);
            acc += term(w1.y);
            acc += term(w1.z);
            acc += term(w1.w);
        }
        for (; q < nq; q++) {
            const s16_u4 w0 = quad(q);
            acc += term(w0.x);
            acc += term(w0.y);
            acc += term(w0.z);
            acc += term(w0.w);
        }
        for (uint32_t j = nq * 4; j < width; j++) acc += term(ent[b + j * kWave + lane]);
        if (r < S.n) epi.row(r, acc);
    }
}

// res = r - D (S z)
struct EpiResidS16 {
    const float *r, *diag;
    float *res;
    __device__ void row(uint32_t i, float s) { res[i] = r[i] - diag[i] * s; }
};

// d = c1 d + c2 (D^-1 r - S z_in), z_out = z_in + d; optional partial r.z_out, z_out.q
struct EpiChebS16 {
    const float *r, *zin, *dinv, *q;
    float *zout, *d;
    float c1, c2;
    bool dots;
    Prolong up;
    double acc[2];
    __device__ void row(uint32_t i, float s)
    {
        const float ri = r[i];
        float dd = c2 * (dinv[i] * ri - s);
        if (c1 != 0.0f) dd = fmaf(c1, d[i], dd);
        d[i] = dd;
        const float z = zin[i] + dd;
        zout[i] = z;
        up.apply(i, z);
        if (dots) {
            acc[0] += (double)ri * (double)z;
            acc[1] += (double)z * (double)q[i];
        }
    }
};

__global__ __launch_bounds__(kBlock) void k_cyc_resid_s16(S16View S, CycArgs a, const float *__restrict__ diag)
{
    if (cyc_idle(a.st)) return;
    EpiResidS16 e{ a.r, diag, a.res };
    s16_run(S, a.zin, e);
}

__global__ __launch_bounds__(kBlock) void k_cyc_step_s16(S16View S, CycArgs a)
{
    if (cyc_idle(a.st)) return;
    EpiChebS16 e{ a.r, a.zin, a.dinv, a.q, a.zout, a.d, a.coef[0], a.coef[1], a.partials != nullptr, a.up, { 0.0, 0.0 } };
    s16_run(S, a.zin, e);
    if (a.partials) store_partials<2>(e.acc, a.partials);
}

// r_next[I] = sum of res over the members of I (G lanes per aggregate, FP64 sum, fixed order).
template <int G>
__global__ __launch_bounds__(kBlock) void k_cyc_restrict(const float *__restrict__ res, const uint32_t *__restrict__ mptr,
                                                         const uint32_t *__restrict__ mrows, uint32_t nnext,
                                                         float *__restrict__ rnext, const SolveState *st, FirstStep f)
{
    if (cyc_idle(st)) return;
    const uint32_t gl = threadIdx.x % G;
    const uint64_t stride = (uint64_t)gridDim.x * (kBlock / G);
    const float c2 = f.coef ? f.coef[1] : 0.0f;
    for (uint64_t I = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / G; I < nnext; I += stride) {
        double s = 0.0;
        for (uint32_t k = mptr[I] + gl; k < mptr[I + 1]; k += G) s += (double)res[mrows[k]];
        s = group_sum<G>(s);
        if (gl == 0) {
            const float rv = (float)s;
            rnext[I] = rv;
            if (f.coef) {
                const float z = c2 * f.dinv[I] * rv;
                f.z[I] = z;
                f.d[I] = z;
            }
        }
    }
}

__global__ void k_coarse_jacobi(const float *__restrict__ r, const float *__restrict__ dinv, float *__restrict__ z,
                                uint32_t n, const SolveState *st, Prolong up)
{
    if (cyc_idle(st)) return;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float v = dinv[i] * r[i];
        z[i] = v;
        up.apply(i, v);
    }
}

// z = C r with the dense FP32 inverse C (row-major n x n), one wavefront per row.
__global__ __launch_bounds__(kBlock) void k_coarse_exact(const float *__restrict__ cinv, const float *__restrict__ r,
                                                         float *__restrict__ z, uint32_t n, const SolveState *st, Prolong up)
{
    if (cyc_idle(st)) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    for (uint32_t row = (blockIdx.x * kBlock + threadIdx.x) / kWave; row < n; row += gridDim.x * (kBlock / kWave)) {
        float s = 0.0f;
        for (uint32_t j = lane; j < n; j += kWave) s += cinv[(uint64_t)row * n + j] * r[j];
        s = group_sum<kWave>(s);
        if (lane == 0) z[row] = s;
        up.apply_group<kWave>(row, s, lane);
    }
}

// ---------------------------------------------------------------- exact coarsest inverse
// The coarsest operator (SPD, n <= 512) is inverted on the device by the blocked symmetric
// sweep operator over 64 x 64 pivot blocks (Goodnight 1979, "A tutorial on the SWEEP
// operator"): sweeping pivot block k maps A_kk -> -D, A_ik -> A_ik D, A_kj -> D A_kj and
// A_ij -> A_ij - A_ik D A_kj with D = A_kk^-1; after every block has been swept the matrix holds
// -A^-1.  FP64 throughout, the result rounded to FP32 for the cycle.  The matrix is padded to a
// multiple of 64 with an identity block (decoupled, so its inverse is the identity there).
constexpr uint32_t kSweepB = 64;
constexpr uint32_t kSweepLd = kSweepB + 1;   // LDS row stride (one pad double)

__global__ void k_dense_pad(double *__restrict__ m, uint32_t N, uint32_t n)
{
    for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < (uint64_t)N * N; e += (uint64_t)gridDim.x * kBlock) {
        const uint32_t i = (uint32_t)(e / N), j = (uint32_t)(e % N);
        m[e] = (i == j && i >= n) ? 1.0 : 0.0;
    }
}

// 16 lanes per row.
__global__ __launch_bounds__(kBlock) void k_dense_scatter(const uint32_t *__restrict__ rowptr,
                                                          const uint32_t *__restrict__ col,
                                                          const double *__restrict__ val, uint32_t n, uint32_t N,
                                                          double *__restrict__ m)
{
    constexpr uint32_t G = 16;
    const uint32_t gl = threadIdx.x % G;
    for (uint64_t i = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / G; i < n; i += (uint64_t)gridDim.x * (kBlock / G))
        for (uint32_t p = rowptr[i] + gl; p < rowptr[i + 1]; p += G) m[i * N + col[p]] = val[p];
}

// D = A_kk^-1 by sweeping the 64 scalar pivots of the block (one workgroup; each thread keeps a
// 4 x 4 block of the pivot block in registers, and before every pivot step the owners of row c
// and column c publish them in LDS).  A pivot that is not positive (not SPD) raises *bad.
__global__ __launch_bounds__(kBlock) void k_sweep_pivot(const double *__restrict__ m, uint32_t N, uint32_t k,
                                                        double *__restrict__ d, uint32_t *bad)
{
    __shared__ double rowc[kSweepB], colc[kSweepB];
    const uint32_t rb = threadIdx.x / 16, cb = threadIdx.x % 16;   // rows 4 rb + i, columns 4 cb + j
    const uint64_t o = (uint64_t)k * kSweepB;
    double v[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) v[i][j] = m[(o + 4 * rb + i) * N + o + 4 * cb + j];
    // Every element is computed by selects, not branches (a branch per element made the step
    // ~600 instructions): the generic update v - c_i r_j / p, row c: r_j / p, column c: c_i / p,
    // pivot: -1 / p, each with the same operations as written.
    for (uint32_t c = 0; c < kSweepB; c++) {
        const uint32_t cq = c / 4, cr = c % 4;
        auto pick = [cr](double a0, double a1, double a2, double a3) {
            return cr == 0 ? a0 : cr == 1 ? a1 : cr == 2 ? a2 : a3;
        };
        if (cq == rb)
#pragma unroll
            for (int j = 0; j < 4; j++) rowc[4 * cb + j] = pick(v[0][j], v[1][j], v[2][j], v[3][j]);
        if (cq == cb)
#pragma unroll
            for (int i = 0; i < 4; i++) colc[4 * rb + i] = pick(v[i][0], v[i][1], v[i][2], v[i][3]);
        __syncthreads();
        const double p = rowc[c];
        if (threadIdx.x == 0 && !(p > 0.0)) *bad = 1u;
        const double ip = p > 0.0 ? 1.0 / p : 0.0;
        double ci[4], rj[4], civ[4], rjv[4];
        bool rowi[4], colj[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ci[i] = colc[4 * rb + i];
            civ[i] = ci[i] * ip;
            rowi[i] = 4 * rb + i == c;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            rj[j] = rowc[4 * cb + j];
            rjv[j] = rj[j] * ip;
            colj[j] = 4 * cb + j == c;
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const double g = v[i][j] - ci[i] * rj[j] * ip;
                const double r = rowi[i] ? rjv[j] : g;
                v[i][j] = colj[j] ? (rowi[i] ? -ip : civ[i]) : r;
            }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) d[(4 * rb + i) * kSweepB + 4 * cb + j] = -v[i][j];
}

// C = X Y for two 64 x 64 LDS tiles; each thread owns a 4 x 4 block of C.
__device__ __forceinline__ void tile_mul(const double *x, const double *y, double (&c)[4][4])
{
    const uint32_t rb = threadIdx.x / 16, cb = threadIdx.x % 16;
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) c[r][q] = 0.0;
    for (uint32_t kk = 0; kk < kSweepB; kk++) {
        double xa[4], yb[4];
#pragma unroll
        for (int r = 0; r < 4; r++) xa[r] = x[(4 * rb + r) * kSweepLd + kk];
#pragma unroll
        for (int q = 0; q < 4; q++) yb[q] = y[kk * kSweepLd + 4 * cb + q];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) c[r][q] = fma(xa[r], yb[q], c[r][q]);
    }
}

__device__ __forceinline__ void tile_load(const double *__restrict__ m, uint32_t N, uint32_t ti, uint32_t tj, double *t)
{
    for (uint32_t e = threadIdx.x; e < kSweepB * kSweepB; e += kBlock)
        t[(e / kSweepB) * kSweepLd + e % kSweepB] = m[((uint64_t)ti * kSweepB + e / kSweepB) * N + (uint64_t)tj * kSweepB + e % kSweepB];
}

// One sweep of pivot block k over every 64 x 64 tile (one workgroup per tile), reading `in` and
// writing `out` (every tile reads the pre-sweep row and column k).
__global__ __launch_bounds__(kBlock) void k_sweep_tiles(const double *__restrict__ in, double *__restrict__ out, uint32_t N,
                                                        uint32_t k, const double *__restrict__ d)
{
    __shared__ double X[kSweepB * kSweepLd], Y[kSweepB * kSweepLd], T[kSweepB * kSweepLd];
    const uint32_t nb = N / kSweepB;
    const uint32_t ti = blockIdx.x / nb, tj = blockIdx.x % nb;
    const uint32_t rb = threadIdx.x / 16, cb = threadIdx.x % 16;
    double c[4][4];
    auto store = [&](double sign, const double *add) {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t i = 4 * rb + r, j = 4 * cb + q;
                const double base = add ? add[i * kSweepLd + j] : 0.0;
                out[((uint64_t)ti * kSweepB + i) * N + (uint64_t)tj * kSweepB + j] = base + sign * c[r][q];
            }
    };
    for (uint32_t e = threadIdx.x; e < kSweepB * kSweepB; e += kBlock)
        Y[(e / kSweepB) * kSweepLd + e % kSweepB] = d[e];              // D
    if (ti == k && tj == k) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) c[r][q] = Y[(4 * rb + r) * kSweepLd + 4 * cb + q];
        store(-1.0, nullptr);
    } else if (ti == k) {                                               // D A_kj
        tile_load(in, N, k, tj, X);
        __syncthreads();
        tile_mul(Y, X, c);
        store(1.0, nullptr);
    } else if (tj == k) {                                               // A_ik D
        tile_load(in, N, ti, k, X);
        __syncthreads();
        tile_mul(X, Y, c);
        store(1.0, nullptr);
    } else {                                                            // A_ij - (A_ik D) A_kj
        tile_load(in, N, ti, k, X);
        __syncthreads();
        tile_mul(X, Y, c);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) T[(4 * rb + r) * kSweepLd + 4 * cb + q] = c[r][q];
        tile_load(in, N, k, tj, X);
        tile_load(in, N, ti, tj, Y);
        __syncthreads();
        tile_mul(T, X, c);
        store(-1.0, Y);
    }
}

// cinv = -m (the swept matrix) in FP32, n x n row-major.
__global__ void k_sweep_result(const double *__restrict__ m, uint32_t N, uint32_t n, float *__restrict__ cinv)
{
    for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < (uint64_t)n * n; e += (uint64_t)gridDim.x * kBlock)
        cinv[e] = (float)(-m[(e / n) * N + e % n]);
}

// Keep the coarsest right-hand side of iteration it (training data).
__global__ void k_record(const float *__restrict__ r, float *__restrict__ rec, uint32_t n, uint32_t cap,
                         const SolveState *st)
{
    if (!st || st->done) return;
    const uint32_t it = st->it;
    if (it >= cap) return;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) rec[(uint64_t)it * n + i] = r[i];
}

// ---------------------------------------------------------------------------- inference
// The coarsest graph for inference as 16 x 16 BF16 blocks of the message weights over strips
// of 16 nodes: blocks bptr[s]..bptr[s+1] of row strip s, column strip bcol[b], entries
// bval[b][c][k] = w(16 s + c, 16 bcol[b] + k) (row-major per block, zero where no edge).
// Activations live in two global ping-pong buffers (L2-resident, 64 KB each), node-major h and
// feature-major ht (row stride npad), so that every MFMA operand is one 8-byte load.
struct GnnArgs {
    uint32_t nc, npad;
    const float *rc, *diag;
    const uint32_t *bptr, *bcol;
    const unsigned short *bval;
    const float *params;
    const unsigned short *wbt;       // transposed BF16 layer matrices [layer][Ws^T | Wn^T][64][64]
    unsigned short *h[2], *ht[2];
    float *u;                        // u = r_c / diag (npad)
    float *norm;                     // [sigma, mean log diag]
    float *e;
    const SolveState *st;
    Prolong up;                      // the correction into the finer level (last layer)
};

__device__ __forceinline__ bf16x4 bf16x4_of(uint2 w) { return __builtin_bit_cast(bf16x4, w); }

// gfx950's K = 32 BF16 MFMA (v_mfma_f32_16x16x32_bf16): lane l supplies A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16] (i = 0..7, one 16-byte load each); the accumulator layout is that of the
// K = 16 form, acc[i] = C[4(l/16)+i][l%16].
__device__ __forceinline__ bf16x8 bf16x8_of(uint4 w) { return __builtin_bit_cast(bf16x8, w); }
__device__ __forceinline__ f32x4 mfma32(uint4 a, uint4 b, f32x4 acc)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf16x8_of(a), bf16x8_of(b), acc, 0, 0, 0);
}

constexpr int kInferMaxPairs = 5;   // pairs of graph blocks per strip prefetched into registers

// Input features and the first activations of one strip of 16 nodes (64 threads, lane = feature).
__global__ __launch_bounds__(kWave) void k_gnn_embed(GnnArgs g)
{
    if (g.st && g.st->done) return;
    const int lane = threadIdx.x;
    const uint32_t nc = g.nc, npad = g.npad, r0 = blockIdx.x * 16;
    // sigma = rms(u) and the mean of log diag over all nodes (every workgroup, same order)
    float su = 0.0f, sl = 0.0f;
    for (uint32_t I = lane; I < nc; I += kWave) {
        const float d = g.diag[I], uu = g.rc[I] / d;
        su += uu * uu;
        sl += logf(d);
    }
    su = group_sum<kWave>(su);
    sl = group_sum<kWave>(sl);
    const float sigma = sqrtf(su / (float)nc) + 1e-30f, meanlog = sl / (float)nc;
    if (blockIdx.x == 0 && lane == 0) {
        g.norm[0] = sigma;
        g.norm[1] = meanlog;
    }
    const float *P = g.params;
    const float w0 = P[kOffWin + lane], w1 = P[kOffWin + kHid + lane], c = P[kOffWin + 2 * kHid + lane] + P[kOffBin + lane];
    // the strip's 16 nodes on lanes 0..15 (one load round), then every lane (a feature) reads
    // them by shuffles; lane f's 16 values of ht are contiguous and go out as two 16-byte stores
    const uint32_t Il = r0 + (lane & 15);
    float ul = 0.0f, ll = 0.0f;
    if (lane < 16) {
        if (Il < nc) {
            const float d = g.diag[Il];
            ul = g.rc[Il] / d;
            ll = logf(d);
        }
        g.u[Il] = ul;
    }
    unsigned short hs[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const float uu = __shfl(ul, t, kWave), ld = __shfl(ll, t, kWave);
        const float z = r0 + t < nc ? (uu / sigma) * w0 + (ld - meanlog) * w1 + c : 0.0f;
        hs[t] = f2bf(fmaxf(z, 0.0f));
        g.h[0][(r0 + t) * kHid + lane] = hs[t];
    }
    uint4 v[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        v[k].x = hs[8 * k + 0] | ((uint32_t)hs[8 * k + 1] << 16);
        v[k].y = hs[8 * k + 2] | ((uint32_t)hs[8 * k + 3] << 16);
        v[k].z = hs[8 * k + 4] | ((uint32_t)hs[8 * k + 5] << 16);
        v[k].w = hs[8 * k + 6] | ((uint32_t)hs[8 * k + 7] << 16);
    }
    uint4 *dst = reinterpret_cast<uint4 *>(g.ht[0] + (size_t)lane * npad + r0);
    dst[0] = v[0];
    dst[1] = v[1];
}

// Layer l for one strip of 16 nodes, 4 wavefronts.  Wavefront w computes the messages of feature
// block w (M^T = H^T G^T over the strip's graph blocks, MFMA; the accumulator layout of M^T is
// the A-operand layout of M) and then the output column tile w of Y = H Ws + M Wn.
__global__ __launch_bounds__(4 * kWave) void k_gnn_layer(GnnArgs g, int l)
{
    if (g.st && g.st->done) return;
    __shared__ __align__(16) unsigned short sM[16 * kHid];
    __shared__ float sO[4][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x / 64;
    const int rc = lane & 15, kq = (lane >> 4) * 4;
    const uint32_t nc = g.nc, npad = g.npad, s = blockIdx.x, r0 = s * 16;
    const unsigned short *H = g.h[l & 1], *Ht = g.ht[l & 1];
    unsigned short *Hn = g.h[(l + 1) & 1], *Htn = g.ht[(l + 1) & 1];
    const float *P = g.params;
    // operands of the output tile first: they do not depend on the messages (K = 32 MFMA: two
    // 32-feature halves, 8 consecutive features per lane)
    const unsigned short *Wl = g.wbt + (size_t)l * kBfPerLayer;
    const int ct = w, q = lane >> 4, k8 = 8 * q;
    uint4 ws[2], wn[2], ha[2];
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
        ws[kb] = *reinterpret_cast<const uint4 *>(Wl + (ct * 16 + rc) * kHid + 32 * kb + k8);
        wn[kb] = *reinterpret_cast<const uint4 *>(Wl + 4096 + (ct * 16 + rc) * kHid + 32 * kb + k8);
        ha[kb] = *reinterpret_cast<const uint4 *>(H + (r0 + rc) * kHid + 32 * kb + k8);
    }
    float hold[4];
#pragma unroll
    for (int i = 0; i < 4; i++) hold[i] = bf2f(H[(r0 + kq + i) * kHid + ct * 16 + rc]);
    const float bias = P[off_b(l) + ct * 16 + rc];
    // messages of feature block kt = w, graph blocks two at a time (lanes q < 2 take the pair's
    // first block, q >= 2 its second, 8 of its 16 nodes each); an odd last block on the K = 16 form
    const uint32_t b0 = g.bptr[s], nb = g.bptr[s + 1] - b0, npair = nb / 2;
    const int kt = w, h8 = 8 * (q & 1), second = q >> 1;
    f32x4 macc = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
    if (npair) {
        uint4 bv[kInferMaxPairs], av[kInferMaxPairs];
#pragma unroll
        for (int j = 0; j < kInferMaxPairs; j++) {
            const uint32_t b = b0 + 2 * min((uint32_t)j, npair - 1) + second;
            bv[j] = *reinterpret_cast<const uint4 *>(g.bval + (uint64_t)b * 256 + rc * 16 + h8);
            av[j] = *reinterpret_cast<const uint4 *>(Ht + (kt * 16 + rc) * npad + g.bcol[b] * 16 + h8);
        }
#pragma unroll
        for (int j = 0; j < kInferMaxPairs; j++)
            if ((uint32_t)j < npair) macc = mfma32(av[j], bv[j], macc);
        for (uint32_t j = kInferMaxPairs; j < npair; j++) {
            const uint32_t b = b0 + 2 * j + second;
            const uint4 bb = *reinterpret_cast<const uint4 *>(g.bval + (uint64_t)b * 256 + rc * 16 + h8);
            const uint4 aa = *reinterpret_cast<const uint4 *>(Ht + (kt * 16 + rc) * npad + g.bcol[b] * 16 + h8);
            macc = mfma32(aa, bb, macc);
        }
    }
    if (nb & 1u) {
        const uint32_t b = b0 + nb - 1;
        const uint2 bb = *reinterpret_cast<const uint2 *>(g.bval + (uint64_t)b * 256 + rc * 16 + kq);
        const uint2 aa = *reinterpret_cast<const uint2 *>(Ht + (kt * 16 + rc) * npad + g.bcol[b] * 16 + kq);
        macc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16x4_of(aa), bf16x4_of(bb), macc, 0, 0, 0);
    }
    // lane holds M[r0 + rc][16 kt + kq + i]
    {
        bf16x4 mv;
#pragma unroll
        for (int i = 0; i < 4; i++) mv[i] = (short)f2bf(macc[i]);
        *reinterpret_cast<bf16x4 *>(&sM[rc * kHid + kt * 16 + kq]) = mv;
    }
    __syncthreads();
    f32x4 acc = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
        const uint4 ma = *reinterpret_cast<const uint4 *>(&sM[rc * kHid + 32 * kb + k8]);
        acc = mfma32(ha[kb], ws[kb], acc);
        acc = mfma32(ma, wn[kb], acc);
    }
    const uint32_t f = ct * 16 + rc;
    float hv[4];
    unsigned short b16[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t row = r0 + kq + i;
        b16[i] = row < nc ? f2bf(hold[i] + fmaxf(acc[i] + bias, 0.0f)) : (unsigned short)0;
        hv[i] = bf2f(b16[i]);
        Hn[row * kHid + f] = b16[i];
    }
    // rows r0 + kq .. + 3 of feature f are contiguous in ht: one 8-byte store
    *reinterpret_cast<uint2 *>(Htn + (size_t)f * npad + r0 + kq) =
        make_uint2(b16[0] | ((uint32_t)b16[1] << 16), b16[2] | ((uint32_t)b16[3] << 16));
    if (l == kLayers - 1) {
        // e = u + sigma (H wo + bo) for the strip's nodes
        const float wo = P[kOffWo + f];
        float part[4];
#pragma unroll
        for (int i = 0; i < 4; i++) part[i] = group_sum<16>(hv[i] * wo);
        if (rc == 0)
#pragma unroll
            for (int i = 0; i < 4; i++) sO[w][kq + i] = part[i];
        __syncthreads();
        if (threadIdx.x < 64) {
            // four lanes per node: the same ev on each, the prolongation spread over them
            const uint32_t t = threadIdx.x >> 2, sub = threadIdx.x & 3u, row = r0 + t;
            if (row < nc) {
                const float o = ((sO[0][t] + sO[1][t]) + (sO[2][t] + sO[3][t])) + P[kOffBo];
                const float ev = g.u[row] + g.norm[0] * o;
                if (sub == 0) g.e[row] = ev;
                g.up.apply_group<4>(row, ev, sub);
            }
        }
    }
}

static void launch_gnn(const GnnArgs &g, hipStream_t s)
{
    const uint32_t strips = g.npad / 16;
    k_gnn_embed<<<strips, kWave, 0, s>>>(g);
    for (int l = 0; l < kLayers; l++) k_gnn_layer<<<strips, 4 * kWave, 0, s>>>(g, l);
}

__global__ void k_fill_u32(uint32_t *__restrict__ a, uint64_t n, uint32_t v)
{
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kBlock) a[k] = v;
}

// Graph blocks from the coarsest CSR weights: bval[bidx[p]] = bf16(cw[p]) (bval zeroed first).
__global__ void k_fill_u16(unsigned short *__restrict__ a, uint64_t n)
{
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kBlock) a[k] = 0;
}
__global__ void k_gnn_blocks(const float *__restrict__ cw, const uint32_t *__restrict__ bidx, uint64_t nnz,
                             unsigned short *__restrict__ bval)
{
    for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < nnz; p += (uint64_t)gridDim.x * kBlock)
        bval[bidx[p]] = f2bf(cw[p]);
}

}  // namespace rs

namespace rs {

// ---------------------------------------------------------------------------- training
// One workgroup per sample; every sample of a launch shares the coarse pattern (structurally
// symmetric; tpos[p] is the position of entry (J, I) for entry p = (I, J)) and has its own
// operator values, coarse residual and normaliser s.
struct TrainArgs {
    uint32_t n, npad, nnz;
    const uint32_t *cptr, *ccol, *tpos;
    const float *cval;                  // [B][nnz]
    const float *rc;                    // [B][n]
    const float *s;                     // [B]
    const float *params;
    const unsigned short *wbf;
    float *grads;                       // [B][kNumParams]
    float *loss;                        // [B]
    unsigned short *hs;                 // [B][kLayers + 1][npad][kHid] activations
    unsigned short *mb, *dy;            // [B][npad][kHid]
    float *dh, *dm;                     // [B][npad][kHid]
};

constexpr int kTrainThreads = 512;

__global__ __launch_bounds__(kTrainThreads) void k_gnn_train(TrainArgs t)
{
    constexpr int kWaves = kTrainThreads / kWave;
    __shared__ unsigned short W[kBfPerLayer];
    __shared__ float u[kMaxCoarse], ld[kMaxCoarse], ev[kMaxCoarse], gov[kMaxCoarse], dg[kMaxCoarse];
    __shared__ float red[kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid / 64;
    const uint32_t n = t.n, npad = t.npad, b = blockIdx.x;
    const uint64_t act = (uint64_t)npad * kHid;
    const float *P = t.params;
    const uint32_t *cptr = t.cptr, *ccol = t.ccol;
    const float *cv = t.cval + (uint64_t)b * t.nnz;
    const float *rcg = t.rc + (uint64_t)b * n;
    float *G = t.grads + (uint64_t)b * kNumParams;
    unsigned short *Hs = t.hs + (uint64_t)b * (kLayers + 1) * act;
    unsigned short *Mb = t.mb + (uint64_t)b * act;
    unsigned short *dY = t.dy + (uint64_t)b * act;
    float *dH = t.dh + (uint64_t)b * act;
    float *dM = t.dm + (uint64_t)b * act;
    for (uint32_t I = tid; I < npad; I += kTrainThreads) {
        float d = 0.0f;
        if (I < n)
            for (uint32_t p = cptr[I]; p < cptr[I + 1]; p++)
                if (ccol[p] == I) d += cv[p];
        dg[I] = d;
    }
    __syncthreads();
    float su = 0.0f, sl = 0.0f;
    for (uint32_t I = tid; I < npad; I += kTrainThreads) {
        float uu = 0.0f, l = 0.0f;
        if (I < n) {
            uu = rcg[I] / dg[I];
            l = logf(dg[I]);
        }
        u[I] = uu;
        ld[I] = l;
        su += uu * uu;
        sl += l;
    }
    const float sigma = sqrtf(block_sum_f(su, red, kTrainThreads) / (float)n) + 1e-30f;
    const float meanlog = block_sum_f(sl, red, kTrainThreads) / (float)n;
    auto feat = [&](uint32_t I, int k) -> float {
        if (I >= n) return 0.0f;
        return k == 0 ? u[I] / sigma : k == 1 ? ld[I] - meanlog : k == 2 ? 1.0f : 0.0f;
    };
    auto z0 = [&](uint32_t I, uint32_t f) -> float {
        if (I >= n) return 0.0f;
        return feat(I, 0) * P[kOffWin + f] + feat(I, 1) * P[kOffWin + kHid + f] + P[kOffWin + 2 * kHid + f] +
               P[kOffBin + f];
    };
    for (uint32_t idx = tid; idx < act; idx += kTrainThreads) Hs[idx] = f2bf(fmaxf(z0(idx / kHid, idx % kHid), 0.0f));
    __syncthreads();
    // M = wn H, wn[I][J] = -A[I][J]/A[I][I] off the diagonal
    // one wavefront per node, lane = feature
    auto aggregate = [&](const unsigned short *Hl) {
        for (uint32_t I = wave; I < npad; I += kWaves) {
            const uint32_t Iu = __builtin_amdgcn_readfirstlane(I);
            float m = 0.0f;
            if (Iu < n) {
                const float d = dg[Iu];
                for (uint32_t p = cptr[Iu]; p < cptr[Iu + 1]; p++) {
                    const uint32_t J = ccol[p];
                    if (J != Iu) m += bf2f(f2bf(-cv[p] / d)) * bf2f(Hl[J * kHid + lane]);
                }
            }
            Mb[Iu * kHid + lane] = f2bf(m);
        }
    };
    // pre-activation Y - b of a 16-row strip, four 16-column tiles
    auto layer_y = [&](const unsigned short *Hl, uint32_t r0, f32x4 (&acc)[4]) {
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            acc[ct] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
            acc[ct] = mfma_tile([&](int r, int k) { return Hl[(r0 + r) * kHid + k]; },
                                [&](int k, int c) { return W[k * kHid + ct * 16 + c]; }, kHid, acc[ct]);
            acc[ct] = mfma_tile([&](int r, int k) { return Mb[(r0 + r) * kHid + k]; },
                                [&](int k, int c) { return W[kHid * kHid + k * kHid + ct * 16 + c]; }, kHid, acc[ct]);
        }
    };
    for (int l = 0; l < kLayers; l++) {
        const unsigned short *Hl = Hs + l * act;
        unsigned short *Hn = Hs + (l + 1) * act;
        for (uint32_t k = tid; k < kBfPerLayer; k += kTrainThreads) W[k] = t.wbf[l * kBfPerLayer + k];
        aggregate(Hl);
        __syncthreads();
        for (uint32_t strip = wave; strip < npad / 16; strip += kWaves) {
            f32x4 acc[4];
            layer_y(Hl, strip * 16, acc);
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    uint32_t row = strip * 16 + (lane >> 4) * 4 + i, f = ct * 16 + (lane & 15);
                    float y = acc[ct][i] + P[off_b(l) + f];
                    float h = bf2f(Hl[row * kHid + f]) + fmaxf(y, 0.0f);
                    Hn[row * kHid + f] = row < n ? f2bf(h) : (unsigned short)0;
                }
        }
        __syncthreads();
    }
    const unsigned short *H6 = Hs + kLayers * act;
    for (uint32_t I = tid; I < npad; I += kTrainThreads) {
        float o = P[kOffBo];
        for (int f = 0; f < kHid; f++) o += bf2f(H6[I * kHid + f]) * P[kOffWo + f];
        ev[I] = I < n ? u[I] + sigma * o : 0.0f;
    }
    __syncthreads();
    // loss and the gradient at the output
    float part = 0.0f;
    const float s = t.s[b];
    for (uint32_t I = tid; I < npad; I += kTrainThreads) {
        float g = 0.0f;
        if (I < n) {
            float ae = 0.0f;
            for (uint32_t p = cptr[I]; p < cptr[I + 1]; p++) ae += cv[p] * ev[ccol[p]];
            g = 2.0f * (ae - rcg[I]) / s * sigma;
            part += ev[I] * ae - 2.0f * ev[I] * rcg[I];
        }
        gov[I] = g;
    }
    const float L = block_sum_f(part, red, kTrainThreads) / s + 1.0f;
    if (tid == 0) t.loss[b] = L;
    __syncthreads();
    if (tid < kHid) {
        float gw = 0.0f;
        for (uint32_t I = 0; I < n; I++) gw += gov[I] * bf2f(H6[I * kHid + tid]);
        G[kOffWo + tid] = gw;
    } else if (tid == kHid) {
        float gb = 0.0f;
        for (uint32_t I = 0; I < n; I++) gb += gov[I];
        G[kOffBo] = gb;
    }
    for (uint32_t idx = tid; idx < act; idx += kTrainThreads) dH[idx] = gov[idx / kHid] * P[kOffWo + idx % kHid];
    __syncthreads();
    for (int l = kLayers - 1; l >= 0; l--) {
        const unsigned short *Hl = Hs + l * act;
        for (uint32_t k = tid; k < kBfPerLayer; k += kTrainThreads) W[k] = t.wbf[l * kBfPerLayer + k];
        aggregate(Hl);
        __syncthreads();
        // dY = dH where Y > 0 (rows < n), rounded to BF16
        for (uint32_t strip = wave; strip < npad / 16; strip += kWaves) {
            f32x4 acc[4];
            layer_y(Hl, strip * 16, acc);
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    uint32_t row = strip * 16 + (lane >> 4) * 4 + i, f = ct * 16 + (lane & 15);
                    float y = acc[ct][i] + P[off_b(l) + f];
                    float g = (y > 0.0f && row < n) ? dH[row * kHid + f] : 0.0f;
                    dY[row * kHid + f] = f2bf(g);
                }
        }
        __syncthreads();
        // dWs = H^T dY and dWn = M^T dY: 16 tiles (input-feature strip x output tile)
        for (int tile = wave; tile < 16; tile += kWaves) {
            const uint32_t fi0 = (tile >> 2) * 16, fo0 = (tile & 3) * 16;
            f32x4 gs = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f }, gn = gs;
            gs = mfma_tile([&](int r, int k) { return Hl[k * kHid + fi0 + r]; },
                           [&](int k, int c) { return dY[k * kHid + fo0 + c]; }, (int)npad, gs);
            gn = mfma_tile([&](int r, int k) { return Mb[k * kHid + fi0 + r]; },
                           [&](int k, int c) { return dY[k * kHid + fo0 + c]; }, (int)npad, gn);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint32_t fi = fi0 + (lane >> 4) * 4 + i, fo = fo0 + (lane & 15);
                G[off_ws(l) + fi * kHid + fo] = gs[i];
                G[off_wn(l) + fi * kHid + fo] = gn[i];
            }
        }
        if (tid < kHid) {
            float gb = 0.0f;
            for (uint32_t I = 0; I < n; I++) gb += bf2f(dY[I * kHid + tid]);
            G[off_b(l) + tid] = gb;
        }
        // dM = dY Wn^T ; dH += dY Ws^T
        for (uint32_t strip = wave; strip < npad / 16; strip += kWaves) {
            const uint32_t r0 = strip * 16;
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                f32x4 am = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f }, as = am;
                am = mfma_tile([&](int r, int k) { return dY[(r0 + r) * kHid + k]; },
                               [&](int k, int c) { return W[kHid * kHid + (ct * 16 + c) * kHid + k]; }, kHid, am);
                as = mfma_tile([&](int r, int k) { return dY[(r0 + r) * kHid + k]; },
                               [&](int k, int c) { return W[(ct * 16 + c) * kHid + k]; }, kHid, as);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    uint32_t row = r0 + (lane >> 4) * 4 + i, f = ct * 16 + (lane & 15);
                    dM[row * kHid + f] = am[i];
                    dH[row * kHid + f] += as[i];
                }
            }
        }
        __syncthreads();
        // dH[I] += sum_J wn[J][I] dM[J]
        for (uint32_t I = wave; I < n; I += kWaves) {
            const uint32_t Iu = __builtin_amdgcn_readfirstlane(I);
            float a = 0.0f;
            for (uint32_t p = cptr[Iu]; p < cptr[Iu + 1]; p++) {
                const uint32_t J = ccol[p];
                if (J != Iu) a += bf2f(f2bf(-cv[t.tpos[p]] / dg[J])) * dM[J * kHid + lane];
            }
            dH[Iu * kHid + lane] += a;
        }
        __syncthreads();
    }
    // embedding: rows 0..3 of Win, then bin
    for (uint32_t task = tid; task < (kNin + 1) * kHid; task += kTrainThreads) {
        const uint32_t k = task / kHid, f = task % kHid;
        float g = 0.0f;
        for (uint32_t I = 0; I < n; I++) {
            float dz = z0(I, f) > 0.0f ? dH[I * kHid + f] : 0.0f;
            g += (k < (uint32_t)kNin ? feat(I, k) : 1.0f) * dz;
        }
        if (k < (uint32_t)kNin) G[kOffWin + k * kHid + f] = g;
        else G[kOffBin + f] = g;
    }
}

// mean gradient over the batch, fixed order
__global__ void k_grad_mean(const float *__restrict__ grads, uint32_t B, float *__restrict__ g)
{
    for (uint32_t p = blockIdx.x * kBlock + threadIdx.x; p < kNumParams; p += gridDim.x * kBlock) {
        float s = 0.0f;
        for (uint32_t b = 0; b < B; b++) s += grads[(uint64_t)b * kNumParams + p];
        g[p] = s / (float)B;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights.
__global__ void k_adam(float *__restrict__ theta, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (uint32_t p = blockIdx.x * kBlock + threadIdx.x; p < kNumParams; p += gridDim.x * kBlock) {
        float gp = g[p];
        float mp = b1 * m[p] + (1.0f - b1) * gp;
        float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
        m[p] = mp;
        v[p] = vp;
        theta[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
    }
}

// BF16 transposed layer matrices for inference: wbt[l][h][c][k] = W_h[l][k][c] (h = 0: Ws, 1: Wn)
__global__ void k_weights_bf16_t(const float *__restrict__ theta, unsigned short *__restrict__ wbt)
{
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < kLayers * kBfPerLayer; k += gridDim.x * kBlock) {
        const uint32_t l = k / kBfPerLayer, o = k % kBfPerLayer, h = o >> 12, c = (o >> 6) & 63, r = o & 63;
        wbt[k] = f2bf(theta[off_ws(l) + (h << 12) + r * kHid + c]);
    }
}

// BF16 copies of the layer matrices: wbf[l][0:4096] = Ws[l], wbf[l][4096:8192] = Wn[l]
__global__ void k_weights_bf16(const float *__restrict__ theta, unsigned short *__restrict__ wbf)
{
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < kLayers * kBfPerLayer; k += gridDim.x * kBlock) {
        uint32_t l = k / kBfPerLayer, o = k % kBfPerLayer;
        wbf[k] = f2bf(theta[off_ws(l) + o]);
    }
}

// ---------------------------------------------------------------------------- rigid-body aggregation
// For 3-DOF elasticity the transfer from level l to level l + 1 can take the six rigid-body modes
// of every aggregate as its coarse space: the tentative prolongation of smoothed aggregation
// (Vanek, Mandel, Brezina 1996), unsmoothed here.  Row i of aggregate J carries the row
// q[6 i .. 6 i + 5] of the orthonormal basis Q_J of the modes of J's members; level l + 1 has
// the rows 6 J + a.  The modes are taken about a reference point p_J (the node of the first fine
// DOF of J) and orthonormalised by modified Gram-Schmidt, every column twice, B_J = Q_J R_J; the
// rows of R_J shifted back to the origin's frame are the next level's near-null-space rows.

constexpr uint32_t kRigidPerLane = 4;
constexpr uint32_t kRigidMaxMembers = kRigidPerLane * kWave;   // rows per aggregate
constexpr double kRigidDeadTol = 1e-10;                        // relative norm of a dependent mode

// Reference point (node coordinates) and component of every DOF: level 0 of a rigid hierarchy.
__global__ void k_rigid_dofs(const uint32_t *__restrict__ book, uint64_t npos, const double *__restrict__ xyz,
                             double *__restrict__ ref, uint32_t *__restrict__ comp)
{
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < npos; q += (uint64_t)gridDim.x * kBlock) {
        const uint32_t d = book[q];
        if (d == RS_EOD) continue;
        const uint64_t node = q / 3;
        ref[3ull * d] = xyz[3 * node];
        ref[3ull * d + 1] = xyz[3 * node + 1];
        ref[3ull * d + 2] = xyz[3 * node + 2];
        comp[d] = (uint32_t)(q % 3);
    }
}

struct RigidQrArgs {
    uint32_t nagg;
    const uint32_t *mptr, *mrows;   // members of each aggregate, rows ascending
    const double *ref;              // reference point of every row of the level (3 per row)
    const uint32_t *comp;           // level 0: component of every row; nullptr on coarser levels
    const double *bsrc;             // coarser levels: near-null-space rows (6 per row, origin's frame)
    double *q64;                    // the basis rows of the level (6 per row)
    float *q32;
    double *bnext, *refnext;        // next level: near-null-space rows (6 per row), reference points
    uint32_t *dead;                 // next level: 1 on the rows of dependent modes
    uint32_t *bad;                  // an aggregate with more than kRigidMaxMembers rows
};

// One wavefront per aggregate J (kRigidPerLane members per lane, in row order lane + 64 u):
// B_J = Q_J R_J by modified Gram-Schmidt, each column orthogonalised twice against the previous
// live columns; wavefront sums by xor butterflies (every lane holds the same bits).  A column whose
// norm falls to kRigidDeadTol of its norm before is dead: zero basis column, zero row of R.
__global__ __launch_bounds__(kBlock) void k_rigid_qr(RigidQrArgs a)
{
    const uint32_t lane = threadIdx.x % kWave;
    const uint64_t stride = (uint64_t)gridDim.x * (kBlock / kWave);
    // J is the same on every lane of a wavefront
    for (uint64_t J = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / kWave; J < a.nagg; J += stride) {
        const uint32_t m0 = a.mptr[J], m = a.mptr[J + 1] - m0;
        if (m > kRigidMaxMembers) {
            if (lane == 0) atomicOr(a.bad, 1u);
            continue;
        }
        const uint32_t first = a.mrows[m0];
        const double px = a.ref[3ull * first], py = a.ref[3ull * first + 1], pz = a.ref[3ull * first + 2];
        double B[kRigidPerLane][6];
        uint32_t row[kRigidPerLane];
#pragma unroll
        for (uint32_t u = 0; u < kRigidPerLane; u++) {
            const uint32_t k = lane + u * kWave;
            row[u] = k < m ? a.mrows[m0 + k] : 0u;
#pragma unroll
            for (int c = 0; c < 6; c++) B[u][c] = 0.0;
            if (k >= m) continue;
            const uint64_t i = row[u];
            if (a.comp) {
                // the modes at a DOF of component cm, relative to p: translations, rotations about z, x, y
                const double rx = a.ref[3 * i] - px, ry = a.ref[3 * i + 1] - py, rz = a.ref[3 * i + 2] - pz;
                const uint32_t cm = a.comp[i];
                B[u][0] = cm == 0 ? 1.0 : 0.0;
                B[u][1] = cm == 1 ? 1.0 : 0.0;
                B[u][2] = cm == 2 ? 1.0 : 0.0;
                B[u][3] = cm == 0 ? -ry : cm == 1 ? rx : 0.0;
                B[u][4] = cm == 1 ? -rz : cm == 2 ? ry : 0.0;
                B[u][5] = cm == 2 ? -rx : cm == 0 ? rz : 0.0;
            } else {
                // origin's frame -> p's frame: each rotation loses the translation p x w
                const double *b = a.bsrc + 6 * i;
                B[u][0] = b[0];
                B[u][1] = b[1];
                B[u][2] = b[2];
                B[u][3] = b[3] + py * b[0] - px * b[1];
                B[u][4] = b[4] + pz * b[1] - py * b[2];
                B[u][5] = b[5] - pz * b[0] + px * b[2];
            }
        }
        double R[6][6];
        bool dead[6];
#pragma unroll
        for (int c = 0; c < 6; c++) {
            dead[c] = false;
#pragma unroll
            for (int k = 0; k < 6; k++) R[c][k] = 0.0;
        }
#pragma unroll
        for (int c = 0; c < 6; c++) {
            double t = 0.0;
#pragma unroll
            for (uint32_t u = 0; u < kRigidPerLane; u++) t += B[u][c] * B[u][c];
            const double n0 = sqrt(group_sum<kWave>(t));
#pragma unroll
            for (int pass = 0; pass < 2; pass++) {
#pragma unroll
                for (int cp = 0; cp < c; cp++) {
                    if (dead[cp]) continue;
                    double d = 0.0;
#pragma unroll
                    for (uint32_t u = 0; u < kRigidPerLane; u++) d += B[u][cp] * B[u][c];
                    d = group_sum<kWave>(d);
#pragma unroll
                    for (uint32_t u = 0; u < kRigidPerLane; u++) B[u][c] -= d * B[u][cp];
                    R[cp][c] += d;
                }
            }
            t = 0.0;
#pragma unroll
            for (uint32_t u = 0; u < kRigidPerLane; u++) t += B[u][c] * B[u][c];
            const double nv = sqrt(group_sum<kWave>(t));
            if (n0 == 0.0 || nv <= kRigidDeadTol * n0) {
                dead[c] = true;
#pragma unroll
                for (uint32_t u = 0; u < kRigidPerLane; u++) B[u][c] = 0.0;
            } else {
#pragma unroll
                for (uint32_t u = 0; u < kRigidPerLane; u++) B[u][c] /= nv;
                R[c][c] = nv;
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kRigidPerLane; u++) {
            if (lane + u * kWave >= m) continue;
            const uint64_t i = row[u];
#pragma unroll
            for (int c = 0; c < 6; c++) {
                a.q64[6 * i + c] = B[u][c];
                a.q32[6 * i + c] = (float)B[u][c];
            }
        }
        if (lane < 6) {
            // row `lane` of R_J in the origin's frame: R T(p), each rotation gains the translation p x w
            const int r = (int)lane;
            double Rr[6];
#pragma unroll
            for (int k = 0; k < 6; k++) Rr[k] = 0.0;
#pragma unroll
            for (int rr = 0; rr < 6; rr++)
                if (rr == r)
#pragma unroll
                    for (int k = 0; k < 6; k++) Rr[k] = R[rr][k];
            double *o = a.bnext + 6 * (6 * J + r);
            o[0] = Rr[0];
            o[1] = Rr[1];
            o[2] = Rr[2];
            o[3] = Rr[3] - py * Rr[0] + px * Rr[1];
            o[4] = Rr[4] - pz * Rr[1] + py * Rr[2];
            o[5] = Rr[5] + pz * Rr[0] - px * Rr[2];
            double *pr = a.refnext + 3 * (6 * J + r);
            pr[0] = px;
            pr[1] = py;
            pr[2] = pz;
            bool dr = false;
#pragma unroll
            for (int rr = 0; rr < 6; rr++)
                if (rr == r) dr = dead[rr];
            a.dead[6 * J + r] = dr ? 1u : 0u;
        }
    }
}

// Row of every entry of a CSR pattern.
__global__ void k_entry_rows(const uint32_t *__restrict__ rowptr, uint32_t n, uint32_t *__restrict__ erow)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        for (uint32_t p = rowptr[i]; p < rowptr[i + 1]; p++) erow[p] = (uint32_t)i;
}

// out[6 I + a] = m[I]: the map of a level's rows when its nodes carry 6 rows each.
__global__ void k_expand6(const uint32_t *__restrict__ m, uint64_t n6, uint32_t *__restrict__ out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n6; i += (uint64_t)gridDim.x * kBlock)
        out[i] = m[i / 6];
}

// The next level's pattern from the aggregate pattern (cptr, ccol): every aggregate entry
// (J, J') becomes the 6 x 6 block of rows 6 J + a and columns 6 J' + b, columns ascending
// within a row; crow[k] is the aggregate row of entry k.
__global__ void k_rigid_expand(const uint32_t *__restrict__ cptr, const uint32_t *__restrict__ ccol, uint32_t nn,
                               uint32_t *__restrict__ erowptr, uint32_t *__restrict__ ecol, uint32_t *__restrict__ crow)
{
    for (uint64_t J = (uint64_t)blockIdx.x * kBlock + threadIdx.x; J < nn; J += (uint64_t)gridDim.x * kBlock) {
        const uint32_t k0 = cptr[J], len = cptr[J + 1] - k0;
        for (uint32_t r = 0; r < 6; r++) {
            const uint32_t base = 36 * k0 + 6 * r * len;
            erowptr[6 * J + r] = base;
            for (uint32_t k = 0; k < len; k++)
                for (uint32_t b = 0; b < 6; b++) ecol[base + 6 * k + b] = 6 * ccol[k0 + k] + b;
        }
        for (uint32_t k = 0; k < len; k++) crow[k0 + k] = (uint32_t)J;
        if (J + 1 == nn) erowptr[6ull * nn] = 36 * cptr[nn];
    }
}

struct RigidGalArgs {
    const double *fval;                     // level l values
    const uint32_t *fcol, *frow;            // level l columns and rows of its entries
    const double *q64;                      // level l basis rows
    const uint32_t *gptr, *gsrc;            // aggregate entry k -> level l entries, in list order
    const uint32_t *cptr, *ccol, *crow;     // aggregate pattern
    const uint32_t *erowptr;                // level l + 1 row pointers (the expanded pattern)
    const uint32_t *dead;                   // level l + 1 rows of dependent modes
    uint64_t cnnz;
    double *cval;                           // level l + 1 values
};

// Level l + 1 values Q^T A Q block by block: a group of G lanes (32 or 64) per aggregate entry
// (J, J'); G = 32 where the entries list few level-l entries (~80 on a mesh's finest level), 64
// where they list hundreds (the coarser levels).  Lane t of the group takes the listed entries
// (i, j) t, t + G, ... in list order and keeps the 36 products q_i[a] a_ij q_j[b]; a
// reduce-scatter over the group then sums them in a fixed order: the xor-G/2 and xor-G/4
// exchanges leave each lane 9 of the 36 sums (index base (t & G/2 ? 18 : 0) + (t & G/4 ? 9 : 0)),
// partial over the G/4 lanes of its subgroup, and a subgroup sum completes them (63 or 54 FP64
// shuffles, against 36 x 6 for a butterfly per sum).  Lane c < 9 of a subgroup writes sum c (with
// 8 lanes, lane 0 also sum 8).  The diagonal of a dead row is 1.
template <int G>
__global__ __launch_bounds__(kBlock) void k_rigid_galerkin(RigidGalArgs a)
{
    static_assert(G == 32 || G == 64, "32 or 64 lanes per aggregate entry");
    constexpr uint32_t kSub = G / 4;
    const uint32_t hl = threadIdx.x % G;
    const uint64_t stride = (uint64_t)gridDim.x * (kBlock / G);
    // k is the same on every lane of a group, so every exchange below (xor offsets below G)
    // stays among lanes of one k
    for (uint64_t k = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / G; k < a.cnnz; k += stride) {
        double acc[36];
#pragma unroll
        for (int c = 0; c < 36; c++) acc[c] = 0.0;
        const uint32_t e = a.gptr[k + 1];
        for (uint32_t t = a.gptr[k] + hl; t < e; t += G) {
            const uint32_t p = a.gsrc[t];
            const double v = a.fval[p];
            const double *qi = a.q64 + 6ull * a.frow[p];
            const double *qj = a.q64 + 6ull * a.fcol[p];
            double vi[6], vj[6];
#pragma unroll
            for (int c = 0; c < 6; c++) {
                vi[c] = qi[c];
                vj[c] = qj[c];
            }
#pragma unroll
            for (int ra = 0; ra < 6; ra++) {
                const double w = vi[ra] * v;
#pragma unroll
                for (int cb = 0; cb < 6; cb++) acc[6 * ra + cb] += w * vj[cb];
            }
        }
        const bool hi = (hl & (G / 2)) != 0, lo = (hl & kSub) != 0;
        double v18[18];
#pragma unroll
        for (int c = 0; c < 18; c++) {
            const double send = hi ? acc[c] : acc[18 + c];
            const double keep = hi ? acc[18 + c] : acc[c];
            v18[c] = keep + __shfl_xor(send, G / 2, kWave);
        }
        double v9[9];
#pragma unroll
        for (int c = 0; c < 9; c++) {
            const double send = lo ? v18[c] : v18[9 + c];
            const double keep = lo ? v18[9 + c] : v18[c];
            v9[c] = keep + __shfl_xor(send, kSub, kWave);
        }
        const uint32_t sub = hl % kSub;
        double out = 0.0, out8 = 0.0;
#pragma unroll
        for (int c = 0; c < 9; c++) {
            const double t = group_sum<kSub>(v9[c]);
            if (sub == (uint32_t)c) out = t;
            if (c == 8) out8 = t;
        }
        const uint32_t base = (hi ? 18u : 0u) + (lo ? 9u : 0u);
        const uint32_t J = a.crow[k];
        const bool diag_block = a.ccol[k] == J;
        const uint32_t col0 = 6 * (uint32_t)(k - a.cptr[J]);
        auto put = [&](uint32_t idx, double val) {
            const uint32_t ra = idx / 6, cb = idx % 6;
            if (diag_block && ra == cb && a.dead[6 * J + ra]) val = 1.0;
            a.cval[a.erowptr[6 * J + ra] + col0 + cb] = val;
        };
        if (sub < 9) put(base + sub, out);
        if (kSub < 9 && sub == 0) put(base + 8, out8);
    }
}

// r_next[6 J + a] = sum over the rows i of J of q_i[a] res_i (G lanes per aggregate, FP64 sums
// in a fixed order), with the next level's first Chebyshev step fused as in k_cyc_restrict.
template <int G>
__global__ __launch_bounds__(kBlock) void k_rigid_restrict(const float *__restrict__ res, const uint32_t *__restrict__ mptr,
                                                           const uint32_t *__restrict__ mrows, const float *__restrict__ q32,
                                                           uint32_t nagg, float *__restrict__ rnext, const SolveState *st,
                                                           FirstStep f)
{
    if (cyc_idle(st)) return;
    const uint32_t gl = threadIdx.x % G;
    const uint64_t stride = (uint64_t)gridDim.x * (kBlock / G);
    const float c2 = f.coef ? f.coef[1] : 0.0f;
    for (uint64_t J = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / G; J < nagg; J += stride) {
        double s[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
        for (uint32_t k = mptr[J] + gl; k < mptr[J + 1]; k += G) {
            const uint64_t i = mrows[k];
            const double r = (double)res[i];
#pragma unroll
            for (int c = 0; c < 6; c++) s[c] += (double)q32[6 * i + c] * r;
        }
#pragma unroll
        for (int c = 0; c < 6; c++) s[c] = group_sum<G>(s[c]);
        if (gl == 0) {
#pragma unroll
            for (int c = 0; c < 6; c++) {
                const uint64_t I = 6 * J + c;
                const float rv = (float)s[c];
                rnext[I] = rv;
                if (f.coef) {
                    const float z = c2 * f.dinv[I] * rv;
                    f.z[I] = z;
                    f.d[I] = z;
                }
            }
        }
    }
}

// z_out[i] = z_pre[i] + sum_a q_i[a] e[6 agg(i) + a] (z_pre nullptr: zero).
__global__ void k_rigid_prolong(uint32_t n, const uint32_t *__restrict__ agg, const float *__restrict__ q32,
                                const float *__restrict__ e, const float *__restrict__ zpre, float *__restrict__ zout,
                                const SolveState *st)
{
    if (cyc_idle(st)) return;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const float *q = q32 + 6 * i;
        const float *ej = e + 6ull * agg[i];
        float s = q[0] * ej[0];
#pragma unroll
        for (int c = 1; c < 6; c++) s = fmaf(q[c], ej[c], s);
        zout[i] = (zpre ? zpre[i] : 0.0f) + s;
    }
}

// ---------------------------------------------------------------------------- flexible CG
struct FcgArgs {
    uint32_t n;
    float *r, *z, *p, *q, *dx;
    double *partials;
    uint32_t *arrive;     // kArriveWords arrival counters, zero between launches
    SolveState *st;
    double *hist;
};

// p = z (first iteration of a segment) or z + beta p
// (both in 16-byte loads and stores, 4 entries per lane, then the remaining n % 4 one per lane)
__global__ void k_fcg_p(FcgArgs a)
{
    if (a.st->done) return;
    const bool first = a.st->it == a.st->it0;
    const float beta = (float)a.st->beta;
    const uint64_t nq = a.n / 4, stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < nq; q += stride) {
        const float4 z = reinterpret_cast<const float4 *>(a.z)[q];
        float4 o = z;
        if (!first) {
            const float4 p = reinterpret_cast<const float4 *>(a.p)[q];
            o = make_float4(fmaf(beta, p.x, z.x), fmaf(beta, p.y, z.y), fmaf(beta, p.z, z.z), fmaf(beta, p.w, z.w));
        }
        reinterpret_cast<float4 *>(a.p)[q] = o;
    }
    for (uint64_t i = 4 * nq + (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride)
        a.p[i] = first ? a.z[i] : fmaf(beta, a.p[i], a.z[i]);
}

// dx += alpha p, r -= alpha q, |r|^2, then the stopping rule and a replacement request
__global__ __launch_bounds__(kBlock) void k_fcg_update(FcgArgs a)
{
    if (a.st->done) return;
    const float alpha = (float)a.st->alpha;
    double v[1] = { 0.0 };
    const uint64_t nq = a.n / 4, stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < nq; q += stride) {
        const float4 p = reinterpret_cast<const float4 *>(a.p)[q], dx = reinterpret_cast<const float4 *>(a.dx)[q];
        const float4 qq = reinterpret_cast<const float4 *>(a.q)[q], r = reinterpret_cast<const float4 *>(a.r)[q];
        reinterpret_cast<float4 *>(a.dx)[q] =
            make_float4(fmaf(alpha, p.x, dx.x), fmaf(alpha, p.y, dx.y), fmaf(alpha, p.z, dx.z), fmaf(alpha, p.w, dx.w));
        const float4 rn = make_float4(fmaf(-alpha, qq.x, r.x), fmaf(-alpha, qq.y, r.y), fmaf(-alpha, qq.z, r.z),
                                      fmaf(-alpha, qq.w, r.w));
        reinterpret_cast<float4 *>(a.r)[q] = rn;
        v[0] += (double)rn.x * (double)rn.x;
        v[0] += (double)rn.y * (double)rn.y;
        v[0] += (double)rn.z * (double)rn.z;
        v[0] += (double)rn.w * (double)rn.w;
    }
    for (uint64_t i = 4 * nq + (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
        a.dx[i] = fmaf(alpha, a.p[i], a.dx[i]);
        float ri = fmaf(-alpha, a.q[i], a.r[i]);
        a.r[i] = ri;
        v[0] += (double)ri * (double)ri;
    }
    finalize_in_launch<1, FIN_FCG_UPDATE>(v, a.partials, a.arrive, a.st, a.hist);
}

// ---------------------------------------------------------------------------- flexible CG, FP64
// The Krylov vectors in FP64, the cycle in FP32: p = z (first iteration of a segment) or
// p beta + z, with z the cycle's FP32 output.
__global__ void k_fcg64_p(const float *__restrict__ z, double *__restrict__ p, uint32_t n, const SolveState *st)
{
    if (st->done) return;
    const bool first = st->it == st->it0;
    const double beta = st->beta;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        p[i] = first ? (double)z[i] : p[i] * beta + (double)z[i];
}

// x += alpha p, r -= alpha q in FP64; the FP32 copies of r (the cycle's right-hand side) and of
// q (for z.q of the next beta); |x|^2 and |r|^2, then the stopping rule of the configuration
// (the FP64 CG's, src/riterationinfo.cpp:90-109).
struct Fcg64Args {
    uint32_t n;
    double *x, *r;
    const double *p, *q;
    float *r32, *q32;
    double *partials;
    uint32_t *arrive;
    SolveState *st;
    double *hist;
};

__global__ __launch_bounds__(kBlock) void k_fcg64_update(Fcg64Args a)
{
    if (a.st->done) return;
    const double alpha = a.st->alpha;
    double v[3] = { 0.0, 0.0, 0.0 };
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * kBlock) {
        const double qi = a.q[i];
        const double xi = a.x[i] + alpha * a.p[i];
        const double ri = a.r[i] - alpha * qi;
        a.x[i] = xi;
        a.r[i] = ri;
        a.r32[i] = (float)ri;
        a.q32[i] = (float)qi;
        v[0] += xi * xi;
        v[1] += ri * ri;
    }
    // v[2] = 0: FIN_UPDATE's beta is the plain CG's; the next FIN_FCG_BETA replaces it
    finalize_in_launch<3, FIN_UPDATE>(v, a.partials, a.arrive, a.st, a.hist);
}

// r32 = r, q32 = 0 before the first iteration
__global__ void k_fcg64_start(const double *__restrict__ r, float *__restrict__ r32, float *__restrict__ q32, uint32_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        r32[i] = (float)r[i];
        q32[i] = 0.0f;
    }
}

}  // namespace rs

// ============================================================================ hierarchy (host)
namespace rs {

struct CycLevel {
    Matrix *M = nullptr;                 // level 0: the solver's matrix; coarser levels: own
    rs_matrix *own = nullptr;
    uint32_t n = 0;
    DBuf<float> diag, dinv, r, za, zb, d, res;
    DBuf<uint32_t> s16;                  // compressed smoother copy (SELL-64 slots), when used
    bool use_s16 = false;
    int s16_pref = 0;                    // timed against the FP32 plan: 0 not yet, 1 faster, 2 slower
    // to the next level
    uint32_t nnext = 0, rgroup = 1;
    DBuf<uint32_t> agg, mptr, mrows, gptr, gsrc;
    // rigid-body transfer to the next level (6 rows per aggregate): this level's reference points
    // (3 per row), components (level 0) or near-null-space rows (6 per row, coarser levels),
    // dependent-mode flags (1 per row, coarser levels), basis rows, entry rows and the aggregate
    // pattern of the next level's blocks
    bool rigid = false;
    DBuf<double> ref, bsrc, q64;
    DBuf<float> q32;
    DBuf<uint32_t> comp, dead, erow, ncptr, nccol, ncrow;
    ~CycLevel()
    {
        delete own;
    }
};

}  // namespace rs

struct rs_learned {
    uint32_t n = 0, nc = 0, nlev = 0;                // fine rows, coarsest nodes, levels incl. fine
    const rs_matrix *owner = nullptr;                // the matrix whose pattern the hierarchy follows
    std::vector<std::unique_ptr<rs::CycLevel>> lv;   // lv[0] fine ... lv[nlev - 1] coarsest
    rs::DBuf<uint32_t> tot_mptr, tot_mrows;          // fine rows of each coarsest node
    rs::DBuf<uint32_t> gbptr, gbcol, gbidx;          // coarsest graph as 16 x 16 blocks (inference)
    rs::DBuf<unsigned short> gbval;
    rs::DBuf<unsigned short> gh, ght;                // inference activations, two buffers each
    rs::DBuf<float> gu, gnorm;
    rs::DBuf<float> cw, cdiag, coef, cinv, rec, tmp_rc, tmp_r;
    rs::DBuf<double> part, gersh;
    rs::DBuf<double> sweep[2], sweep_d;              // exact coarsest inverse: swept matrix, pivot inverse
    rs::DBuf<uint32_t> sweep_bad;
    rs::DBuf<float> theta;
    rs::DBuf<unsigned short> wbf;
    bool have_params = false;
    uint32_t pre = 1, post = 1;                      // smoothing steps on the finest level
    uint32_t cpre = 1, cpost = 1;                    // and on the coarser levels
    int coarse = RS_COARSE_NETWORK;
    uint32_t rec_cap = 0;
    uint32_t s16_min_rows = 65536;                   // levels at least this large smooth with the FP16 copy
    bool s16_auto = true;                            // ... when it times faster than the FP32 plan
    rs::DBuf<uint32_t> s16bad;
    uint64_t setup_version = ~0ull;
    bool inverse_valid = false;
    std::vector<double> lam;                         // Gershgorin bound per level
    const float *r0 = nullptr;                       // finest right-hand side of the cycle
    bool rigid = false;                              // rigid-body transfers (6 rows per aggregate)
};

namespace rs {

// Chebyshev coefficients (c1, c2) of steps 0..kMaxDeg-1 over [lambda/30, 1.05 lambda].
static void cheb_coefs(double lambda, float *out)
{
    const double hi = 1.05 * lambda, lo = lambda / 30.0;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sig = theta / delta;
    double rho = 1.0 / sig;
    out[0] = 0.0f;
    out[1] = (float)(1.0 / theta);
    for (uint32_t k = 1; k < kMaxDeg; k++) {
        double rn = 1.0 / (2.0 * sig - rho);
        out[2 * k] = (float)(rn * rho);
        out[2 * k + 1] = (float)(2.0 * rn / delta);
        rho = rn;
    }
}

static int coarsest_dense(const rs_learned &L, std::vector<double> &dense)
{
    const Matrix &C = *L.lv.back()->M;
    std::vector<uint32_t> cp(C.n + 1ull), cc(C.nnz);
    std::vector<double> cv(C.nnz);
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(cp.data(), C.rowptr.p, sizeof(uint32_t) * cp.size(), hipMemcpyDeviceToHost, st));
    RS_HIP(hipMemcpyAsync(cc.data(), C.col.p, sizeof(uint32_t) * C.nnz, hipMemcpyDeviceToHost, st));
    RS_HIP(hipMemcpyAsync(cv.data(), C.val.p, sizeof(double) * C.nnz, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    dense.assign((size_t)C.n * C.n, 0.0);
    for (uint32_t I = 0; I < C.n; I++)
        for (uint32_t p = cp[I]; p < cp[I + 1]; p++) dense[(size_t)I * C.n + cc[p]] += cv[p];
    return RS_OK;
}

// Values of every level for the current values of A: Galerkin gathers, inverse diagonals,
// Chebyshev coefficients, the coarsest graph weights (and its inverse in exact mode).
static int learned_setup(rs_learned &L, rs_matrix &A)
{
    if (&A != L.owner) return fail(RS_ERR_INVALID_ARG, "learned preconditioner: built for another matrix");
    if (L.setup_version == A.version && (L.coarse != RS_COARSE_EXACT || L.inverse_valid)) return RS_OK;
    hipStream_t st = runtime().stream;
    int rc;
    L.lv[0]->M = &A;
    for (uint32_t l = 0; l + 1 < L.nlev; l++) {
        CycLevel &F = *L.lv[l], &C = *L.lv[l + 1];
        if (F.rigid) {
            const uint64_t cnnz = F.ncrow.n;
            RigidGalArgs g{ F.M->val.p, F.M->col.p, F.erow.p, F.q64.p, F.gptr.p, F.gsrc.p, F.ncptr.p, F.nccol.p,
                            F.ncrow.p, C.M->rowptr.p, C.dead.p, cnnz, C.M->val.p };
            // 32 lanes per aggregate entry while the entries list at most ~3 trips of 32 level-l
            // entries on average, 64 above
            if (F.gsrc.n <= 96 * std::max<uint64_t>(cnnz, 1))
                k_rigid_galerkin<32><<<vec_blocks(std::max<uint64_t>(cnnz, 1) * 32), kBlock, 0, st>>>(g);
            else
                k_rigid_galerkin<64><<<vec_blocks(std::max<uint64_t>(cnnz, 1) * 64), kBlock, 0, st>>>(g);
        } else {
            k_galerkin_gather<<<vec_blocks(std::max<uint64_t>(C.M->nnz, 1)), kBlock, 0, st>>>(F.M->val.p, F.gptr.p,
                                                                                             F.gsrc.p, C.M->nnz, C.M->val.p);
        }
        matrix_values_changed(*C.M);
    }
    const uint32_t kParts = 4096;
    for (uint32_t l = 0; l < L.nlev; l++) {
        CycLevel &V = *L.lv[l];
        // 16 lanes per row: a workgroup per 16 rows, up to the partials' capacity
        uint32_t nb = std::min(vec_blocks(std::max<uint64_t>(V.n, 1) * 16), kParts);
        k_level_diag<<<nb, kBlock, 0, st>>>(V.M->rowptr.p, V.M->col.p, V.M->val.p, V.n, V.diag.p, V.dinv.p,
                                            L.part.p + (uint64_t)l * kParts);
        k_max_fin<<<1, kMaxFinBlock, 0, st>>>(L.part.p + (uint64_t)l * kParts, nb, L.gersh.p + l);
    }
    const Matrix &C = *L.lv.back()->M;
    k_coarse_weights<<<vec_blocks((uint64_t)L.nc * 16), kBlock, 0, st>>>(C.rowptr.p, C.col.p, C.val.p, L.nc, L.cw.p, L.cdiag.p);
    k_fill_u16<<<vec_blocks(L.gbval.n), kBlock, 0, st>>>(L.gbval.p, L.gbval.n);
    k_gnn_blocks<<<vec_blocks(std::max<uint64_t>(C.nnz, 1)), kBlock, 0, st>>>(L.cw.p, L.gbidx.p, C.nnz, L.gbval.p);
    RS_HIP(hipGetLastError());
    // compressed smoother copies of the large levels
    std::vector<uint32_t> bad(L.nlev, 1u);
    k_fill_u32<<<1, 64, 0, st>>>(L.s16bad.p, L.nlev, 0u);
    for (uint32_t l = 0; l + 1 < L.nlev; l++) {
        CycLevel &V = *L.lv[l];
        if (L.s16_min_rows == 0 || V.n < L.s16_min_rows || V.M->sell_col.n == 0) continue;
        if (L.s16_auto && V.s16_pref == 2) continue;
        if ((rc = V.s16.alloc(V.M->sell_col.n)) || (rc = matrix_sync_plan(*V.M, kSellPlan, false))) return rc;
        uint32_t nb = 0;
        const CsrView sv = plan_view(*V.M, kSellPlan, nb);
        k_s16_build<<<vec_blocks((uint64_t)V.M->nslices * kWave), kBlock, 0, st>>>(sv, V.s16.p, L.s16bad.p + l);
    }
    RS_HIP(hipGetLastError());
    L.lam.assign(L.nlev, 0.0);
    RS_HIP(hipMemcpyAsync(L.lam.data(), L.gersh.p, sizeof(double) * L.nlev, hipMemcpyDeviceToHost, st));
    RS_HIP(hipMemcpyAsync(bad.data(), L.s16bad.p, sizeof(uint32_t) * L.nlev, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    for (uint32_t l = 0; l < L.nlev; l++) {
        CycLevel &V = *L.lv[l];
        V.use_s16 = l + 1 < L.nlev && L.s16_min_rows != 0 && V.n >= L.s16_min_rows && V.s16.p && bad[l] == 0 &&
                    !(L.s16_auto && V.s16_pref == 2);
    }
    std::vector<float> coef((size_t)L.nlev * kMaxDeg * 2);
    for (uint32_t l = 0; l < L.nlev; l++) cheb_coefs(L.lam[l] > 0.0 ? L.lam[l] : 1.0, &coef[(size_t)l * kMaxDeg * 2]);
    RS_HIP(hipMemcpyAsync(L.coef.p, coef.data(), sizeof(float) * coef.size(), hipMemcpyHostToDevice, st));
    L.inverse_valid = false;
    if (L.coarse == RS_COARSE_EXACT) {
        // the coarsest operator's inverse, on the device (blocked sweep, 2 launches per 64-block)
        const uint32_t n = L.nc, N = (n + kSweepB - 1) / kSweepB * kSweepB, nbk = N / kSweepB;
        if ((rc = L.sweep[0].alloc((size_t)N * N)) || (rc = L.sweep[1].alloc((size_t)N * N)) ||
            (rc = L.sweep_d.alloc(kSweepB * kSweepB)) || (rc = L.sweep_bad.alloc(1)) ||
            (rc = L.cinv.alloc((size_t)n * n)))
            return rc;
        k_fill_u32<<<1, 64, 0, st>>>(L.sweep_bad.p, 1, 0u);
        k_dense_pad<<<vec_blocks((uint64_t)N * N), kBlock, 0, st>>>(L.sweep[0].p, N, n);
        k_dense_scatter<<<vec_blocks((uint64_t)n * 16), kBlock, 0, st>>>(C.rowptr.p, C.col.p, C.val.p, n, N, L.sweep[0].p);
        for (uint32_t k = 0; k < nbk; k++) {
            k_sweep_pivot<<<1, kBlock, 0, st>>>(L.sweep[k & 1].p, N, k, L.sweep_d.p, L.sweep_bad.p);
            k_sweep_tiles<<<nbk * nbk, kBlock, 0, st>>>(L.sweep[k & 1].p, L.sweep[(k + 1) & 1].p, N, k, L.sweep_d.p);
        }
        k_sweep_result<<<vec_blocks((uint64_t)n * n), kBlock, 0, st>>>(L.sweep[nbk & 1].p, N, n, L.cinv.p);
        RS_HIP(hipGetLastError());
        uint32_t bad_pivot = 0;
        RS_HIP(hipMemcpyAsync(&bad_pivot, L.sweep_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
        if (bad_pivot) return fail(RS_ERR_UNSUPPORTED, "learned preconditioner: the coarsest operator is not positive definite");
        L.inverse_valid = true;
    }
    RS_HIP(hipStreamSynchronize(st));
    L.setup_version = A.version;
    return RS_OK;
}

// FP32 SpMV plans of every level (autotuned once per level).
// and, under the default threshold, whether a level's compressed smoother copy is faster than
// its FP32 plan: the two residual kernels timed back to back on the level's own vectors, once per
// hierarchy (long rows favour the FP32 shapes with several lanes per row: the rigid-body
// level 1 of the cantilever, 83 entries per row).
static int learned_plans(rs_learned &L)
{
    int rc;
    for (uint32_t l = 0; l < L.nlev; l++) {
        Matrix &M = *L.lv[l]->M;
        if ((rc = matrix_refresh_fp32(M)) || (rc = matrix_tune(M, true)) || (rc = matrix_sync_plan(M, M.plan[1], true)))
            return rc;
    }
    if (!L.s16_auto) return RS_OK;
    hipStream_t st = runtime().stream;
    for (uint32_t l = 0; l + 1 < L.nlev; l++) {
        CycLevel &V = *L.lv[l];
        if (!V.use_s16 || V.s16_pref != 0) continue;
        const int plan = V.M->plan[1];
        uint32_t nb = 0, nbs = 0;
        CycArgs a{};
        a.A = plan_view(*V.M, plan, nb);
        a.r = V.r.p;
        a.zin = V.za.p;
        a.res = V.res.p;
        a.dinv = V.dinv.p;
        a.n = V.n;
        const CsrView v16 = plan_view(*V.M, kSellPlan, nbs);
        const S16View sv{ V.n, v16.nslices, v16.spb, v16.sell_ptr, V.s16.p };
        k_fill32<<<vec_blocks(V.n), kBlock, 0, st>>>(V.r.p, V.n, 1.0f);
        k_fill32<<<vec_blocks(V.n), kBlock, 0, st>>>(V.za.p, V.n, 1.0f);
        RS_HIP(hipGetLastError());
        double ms[2] = { 0.0, 0.0 };
        for (int k = 0; k < 2; k++) {
            auto launch = [&]() {
                if (k == 0) k_cyc_resid_s16<<<nbs, kBlock, 0, st>>>(sv, a, V.diag.p);
                else RS_SHAPE_SWITCH(plan, k_cyc_resid, nb, st, a);
            };
            launch();
            EventPair ev;
            RS_HIP(ev.create());
            RS_HIP(hipEventRecord(ev.e[0], st));
            for (int i = 0; i < 10; i++) launch();
            RS_HIP(hipEventRecord(ev.e[1], st));
            RS_HIP(hipEventSynchronize(ev.e[1]));
            RS_HIP(hipGetLastError());
            float f = 0.0f;
            RS_HIP(hipEventElapsedTime(&f, ev.e[0], ev.e[1]));
            ms[k] = f;
        }
        // the copy rounds the smoother's values to FP16, so it is taken only for a clear gain:
        // with a near tie the timing noise alone would pick it on some runs and not on others
        V.s16_pref = ms[0] <= 0.9 * ms[1] ? 1 : 2;
        if (V.s16_pref == 2) V.use_s16 = false;
    }
    return RS_OK;
}

static GnnArgs gnn_args(rs_learned &L, const SolveState *st)
{
    CycLevel &V = *L.lv.back();
    GnnArgs g{};
    g.nc = L.nc;
    g.npad = (L.nc + 15) / 16 * 16;
    g.rc = V.r.p;
    g.diag = L.cdiag.p;
    g.bptr = L.gbptr.p;
    g.bcol = L.gbcol.p;
    g.bval = L.gbval.p;
    g.params = L.theta.p;
    g.wbt = L.wbf.p;
    const size_t act = (size_t)g.npad * kHid;
    g.h[0] = L.gh.p;
    g.h[1] = L.gh.p + act;
    g.ht[0] = L.ght.p;
    g.ht[1] = L.ght.p + act;
    g.u = L.gu.p;
    g.norm = L.gnorm.p;
    g.e = V.za.p;
    g.st = st;
    return g;
}

static void launch_restrict(uint32_t G, uint32_t nb, const float *res, const CycLevel &V, float *rnext, const SolveState *st,
                            FirstStep f)
{
    hipStream_t s = runtime().stream;
    switch (G) {
        case 1: k_cyc_restrict<1><<<nb, kBlock, 0, s>>>(res, V.mptr.p, V.mrows.p, V.nnext, rnext, st, f); break;
        case 4: k_cyc_restrict<4><<<nb, kBlock, 0, s>>>(res, V.mptr.p, V.mrows.p, V.nnext, rnext, st, f); break;
        case 16: k_cyc_restrict<16><<<nb, kBlock, 0, s>>>(res, V.mptr.p, V.mrows.p, V.nnext, rnext, st, f); break;
        default: k_cyc_restrict<64><<<nb, kBlock, 0, s>>>(res, V.mptr.p, V.mrows.p, V.nnext, rnext, st, f); break;
    }
}

// Workgroups of a Chebyshev step on level V: its partial-sum slot count when it is the finest
// level's last step (the SELL grid for the compressed copy, else the level's FP32 plan grid).
static uint32_t step_blocks(const CycLevel &V)
{
    uint32_t nb = 0;
    plan_view(*V.M, V.use_s16 ? kSellPlan : V.M->plan[1], nb);
    return nb;
}

// One Chebyshev step on level V (SpMV in the level's compressed copy or its FP32 plan).
static void launch_step(const CycLevel &V, const CycArgs &a, hipStream_t s)
{
    uint32_t nb = 0;
    if (V.use_s16) {
        const CsrView v16 = plan_view(*V.M, kSellPlan, nb);
        k_cyc_step_s16<<<nb, kBlock, 0, s>>>(S16View{ V.n, v16.nslices, v16.spb, v16.sell_ptr, V.s16.p }, a);
    } else {
        const int plan = V.M->plan[1];
        plan_view(*V.M, plan, nb);
        RS_SHAPE_SWITCH(plan, k_cyc_step, nb, s, a);
    }
}

// Launches the V-cycle from level l.  `up` says where the level's finished correction goes
// (the finer level's post-smoothing input, prolongation fused); the finest level returns its
// result buffer.  On the finest level with q and partials given, the last step also leaves the
// partial sums r.z and z.q for the flexible CG.  Levels below the finest receive their first
// pre-smoothing step from the restriction kernel of the level above.
static const float *cycle_level(rs_learned &L, uint32_t l, const SolveState *st, const float *q, double *partials,
                                Prolong up)
{
    CycLevel &V = *L.lv[l];
    hipStream_t s = runtime().stream;
    const uint32_t nvb = vec_blocks(std::max(V.n, 1u));
    const float *r = l == 0 ? L.r0 : V.r.p;
    if (l + 1 == L.nlev) {
        if (L.coarse == RS_COARSE_NETWORK) {
            GnnArgs g = gnn_args(L, st);
            g.up = up;
            launch_gnn(g, s);
        } else if (L.coarse == RS_COARSE_EXACT) {
            k_coarse_exact<<<vec_blocks((uint64_t)V.n * kWave), kBlock, 0, s>>>(L.cinv.p, r, V.za.p, V.n, st, up);
        } else {
            k_coarse_jacobi<<<nvb, kBlock, 0, s>>>(r, V.dinv.p, V.za.p, V.n, st, up);
        }
        return V.za.p;
    }
    CycLevel &C = *L.lv[l + 1];
    const int plan = V.M->plan[1];
    uint32_t nb = 0;
    CycArgs a{};
    a.A = plan_view(*V.M, plan, nb);
    a.r = r;
    a.d = V.d.p;
    a.res = V.res.p;
    a.dinv = V.dinv.p;
    a.st = st;
    a.n = V.n;
    const float *coef = L.coef.p + (size_t)l * kMaxDeg * 2;
    float *buf[2] = { V.za.p, V.zb.p };
    int cur = -1;
    const uint32_t npre = l == 0 ? L.pre : L.cpre, npost = l == 0 ? L.post : L.cpost;
    for (uint32_t k = 0; k < npre; k++) {
        a.coef = coef + 2 * k;
        if (k == 0) {
            if (l == 0) {
                a.zout = buf[0];
                k_cyc_first<<<nvb, kBlock, 0, s>>>(a);
            }
            cur = 0;
        } else {
            a.zin = buf[cur];
            a.zout = buf[cur ^ 1];
            launch_step(V, a, s);
            cur ^= 1;
        }
    }
    const float *rres = r;
    if (cur >= 0) {
        a.zin = buf[cur];
        if (V.use_s16) {
            uint32_t nbs = 0;
            const CsrView v16 = plan_view(*V.M, kSellPlan, nbs);
            k_cyc_resid_s16<<<nbs, kBlock, 0, s>>>(S16View{ V.n, v16.nslices, v16.spb, v16.sell_ptr, V.s16.p }, a, V.diag.p);
        } else {
            RS_SHAPE_SWITCH(plan, k_cyc_resid, nb, s, a);
        }
        rres = V.res.p;
    }
    FirstStep fs{};
    if (L.cpre >= 1 && l + 2 < L.nlev) fs = FirstStep{ L.coef.p + (size_t)(l + 1) * kMaxDeg * 2, C.dinv.p, C.za.p, C.d.p };
    const uint32_t G = V.rgroup;
    const int nxt = cur >= 0 ? cur ^ 1 : 0;
    if (V.rigid) {
        // restriction Q^T res, the coarser cycle, then z_pre + Q e in buf[nxt] (one thread per row)
        const uint32_t nagg = V.nnext / 6;
        const uint32_t nbr = vec_blocks((uint64_t)nagg * G);
        switch (G) {
            case 4: k_rigid_restrict<4><<<nbr, kBlock, 0, s>>>(rres, V.mptr.p, V.mrows.p, V.q32.p, nagg, C.r.p, st, fs); break;
            case 16: k_rigid_restrict<16><<<nbr, kBlock, 0, s>>>(rres, V.mptr.p, V.mrows.p, V.q32.p, nagg, C.r.p, st, fs); break;
            default: k_rigid_restrict<64><<<nbr, kBlock, 0, s>>>(rres, V.mptr.p, V.mrows.p, V.q32.p, nagg, C.r.p, st, fs); break;
        }
        const float *e = cycle_level(L, l + 1, st, nullptr, nullptr, Prolong{});
        k_rigid_prolong<<<nvb, kBlock, 0, s>>>(V.n, V.agg.p, V.q32.p, e, cur >= 0 ? buf[cur] : nullptr, buf[nxt], st);
    } else {
        launch_restrict(G, vec_blocks((uint64_t)V.nnext * G), rres, V, C.r.p, st, fs);
        // the coarser level leaves z_pre + P e in buf[nxt]
        cycle_level(L, l + 1, st, nullptr, nullptr, Prolong{ V.mptr.p, V.mrows.p, cur >= 0 ? buf[cur] : nullptr, buf[nxt] });
    }
    cur = nxt;
    for (uint32_t k = 0; k < npost; k++) {
        const bool last = k + 1 == npost;
        a.coef = coef + 2 * k;
        a.zin = buf[cur];
        a.zout = buf[cur ^ 1];
        a.q = last ? q : nullptr;
        a.partials = last ? partials : nullptr;
        a.up = last ? up : Prolong{};
        launch_step(V, a, s);
        cur ^= 1;
    }
    return buf[cur];
}

static std::string learned_key(const rs_learned &L)
{
    std::vector<uint64_t> k;
    for (uint32_t l = 0; l < L.nlev; l++) {
        const Matrix &M = *L.lv[l]->M;
        k.push_back((uint64_t)(uintptr_t)&M);
        k.push_back((uint64_t)M.plan[1]);
        k.push_back((uint64_t)(uintptr_t)M.val32.p);
        k.push_back((uint64_t)(uintptr_t)M.sell_val32.p);
        k.push_back(L.lv[l]->use_s16 ? (uint64_t)(uintptr_t)L.lv[l]->s16.p : 0ull);
    }
    k.push_back(L.pre);
    k.push_back(L.post);
    k.push_back(L.cpre);
    k.push_back(L.cpost);
    k.push_back((uint64_t)L.coarse);
    k.push_back(L.rec_cap);
    k.push_back((uint64_t)(uintptr_t)L.rec.p);
    k.push_back((uint64_t)(uintptr_t)L.cinv.p);
    k.push_back((uint64_t)(uintptr_t)L.r0);
    return std::string(reinterpret_cast<const char *>(k.data()), k.size() * sizeof(uint64_t));
}

// Flexible CG with the learned multilevel preconditioner (mixed precision, FP64 residual
// replacement).
static int solve_fcg_learned(rs_solver &S, rs_matrix &A, rs_precond &P, rs_learned &L, double *d_x, uint32_t &done,
                             uint32_t &refinements)
{
    int rc;
    if ((rc = matrix_refresh_fp32(A)) || (rc = precond_fp32(P)) || (rc = matrix_tune(A, true)) ||
        (rc = matrix_tune(A, false)))
        return rc;
    const int plan32 = A.plan[1], plan64 = A.plan[0];
    if ((rc = matrix_sync_plan(A, plan32, true)) || (rc = matrix_sync_plan(A, plan64, false))) return rc;
    if ((rc = learned_setup(L, A)) || (rc = learned_plans(L))) return rc;
    if ((rc = S.z32.alloc(std::max(A.n, 1u)))) return rc;
    uint32_t nb32 = 0, nb64 = 0;
    MpArgs a{};
    a.A = plan_view(A, plan32, nb32);
    a.b = S.y.p;
    a.x = d_x;
    a.r = S.r32.p;
    a.p = S.p32.p;
    a.q = S.q32.p;
    a.dx = S.dx32.p;
    a.ptype = RS_PRECOND_JACOBI;
    a.pw = 1;
    a.pd = P.data32.p;
    a.partials = S.partials.p;
    a.arrive = reinterpret_cast<uint32_t *>(S.partials.p + kPartialDoubles);
    a.st = S.st.p;
    a.hist = S.hist.p;
    MpArgs a64 = a;
    a64.A = plan_view(A, plan64, nb64);
    FcgArgs f{};
    f.n = A.n;
    f.r = S.r32.p;
    f.p = S.p32.p;
    f.q = S.q32.p;
    f.dx = S.dx32.p;
    f.partials = S.partials.p;
    f.arrive = a.arrive;
    f.st = S.st.p;
    f.hist = S.hist.p;
    const uint32_t nvb = vec_blocks(A.n);
    hipStream_t stream = runtime().stream;
    L.r0 = S.r32.p;
    // r.z and z.q come from the finest level's last post-smoothing step, whose grid is not the
    // FP32 SpMV plan's when that level smooths with its compressed copy
    const uint32_t nbz = step_blocks(*L.lv[0]);
    k_fill32<<<nvb, kBlock, 0, stream>>>(S.dx32.p, A.n, 0.0f);
    k_fill32<<<nvb, kBlock, 0, stream>>>(S.q32.p, A.n, 0.0f);
    k_zero_u32<<<1, 64, 0, stream>>>(a.arrive, kArriveWords);
    RS_SHAPE_SWITCH(plan64, k_mp_true_resid, nb64, stream, a64);
    fin(S.partials.p, nb64, 2, FIN_MP_TRUE, 1, S.st.p, S.hist.p);
    RS_HIP(hipGetLastError());
    const std::function<void()> body = [&] {
        for (uint32_t i = 0; i < kGraphIters; i++) {
            const float *z = cycle_level(L, 0, S.st.p, S.q32.p, S.partials.p, Prolong{});
            if (L.rec_cap) {
                CycLevel &Cl = *L.lv.back();
                k_record<<<vec_blocks(L.nc), kBlock, 0, stream>>>(Cl.r.p, L.rec.p, L.nc, L.rec_cap, S.st.p);
            }
            fin(S.partials.p, nbz, 2, FIN_FCG_BETA, 0, S.st.p, S.hist.p);
            FcgArgs fz = f;
            fz.z = const_cast<float *>(z);
            k_fcg_p<<<nvb, kBlock, 0, stream>>>(fz);
            RS_SHAPE_SWITCH(plan32, k_mp_spmv_dot, nb32, stream, a);   // finalizes alpha
            k_fcg_update<<<nvb, kBlock, 0, stream>>>(f);                // finalizes the stopping rule
            // FP64 residual replacement when requested, inside the graph (no host round trip);
            // the true-residual launch finalizes it
            k_mp_fold_cond<<<nvb, kBlock, 0, stream>>>(d_x, S.dx32.p, A.n, S.st.p);
            RS_SHAPE_SWITCH(plan64, k_mp_true_resid_cond, nb64, stream, a64);
        }
    };
    rc = capture(S, body, graph_key_of(3, a, plan32, nb32, nb64, nvb) + learned_key(L) + std::to_string(nbz));
    if (rc) return rc;
    SolveState h;
    RS_HIP(hipMemcpyAsync(&h, S.st.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    RS_HIP(hipStreamSynchronize(stream));
    done = h.done;
    if (done == kStopRunning && (rc = run_graph_loop(S, done, body))) return rc;
    RS_HIP(hipMemcpyAsync(&h, S.st.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    RS_HIP(hipStreamSynchronize(stream));
    refinements = h.nrepl;
    // a segment that ended on a non-finite residual leaves x as of its last replacement
    if (h.reason != RS_REASON_NONNORMAL) k_mp_fold<<<nvb, kBlock, 0, stream>>>(d_x, S.dx32.p, A.n);
    RS_HIP(hipGetLastError());
    return RS_OK;
}

// Flexible CG with FP64 Krylov vectors and FP64 SpMV around the FP32 cycle: the cycle's
// rounding does not enter the recursion, so operators whose FP32 rounding is amplified
// (elasticity: near rigid-body cancellation) keep the iteration count of an FP64 cycle.  Per
// iteration: the cycle (its last step leaves r.z and z.q), FIN_FCG_BETA, p, the FP64 product
// (k_cg_spmv_dot finalizes alpha = r.z / p.q) and the update (finalizes the stopping rule).
static int solve_fcg64_learned(rs_solver &S, rs_matrix &A, rs_learned &L, double *d_x, uint32_t &done)
{
    int rc;
    if ((rc = matrix_refresh_fp32(A)) || (rc = matrix_tune(A, true)) || (rc = matrix_tune(A, false))) return rc;
    const int plan64 = A.plan[0];
    if ((rc = matrix_sync_plan(A, A.plan[1], true)) || (rc = matrix_sync_plan(A, plan64, false))) return rc;
    if ((rc = learned_setup(L, A)) || (rc = learned_plans(L))) return rc;
    const size_t N = std::max(A.n, 1u);
    if ((rc = S.z32.alloc(N)) || (rc = S.r32.alloc(N)) || (rc = S.q32.alloc(N))) return rc;
    uint32_t nb64 = 0;
    CgArgs a{};
    a.A = plan_view(A, plan64, nb64);
    a.b = S.y.p;
    a.x = d_x;
    a.r = S.r.p;
    a.p = S.p.p;
    a.q = S.q.p;
    a.ptype = RS_PRECOND_NONE;
    a.pw = 1;
    a.partials = S.partials.p;
    a.arrive = reinterpret_cast<uint32_t *>(S.partials.p + kPartialDoubles);
    a.st = S.st.p;
    a.hist = S.hist.p;
    Fcg64Args u{ A.n, d_x, S.r.p, S.p.p, S.q.p, S.r32.p, S.q32.p, a.partials, a.arrive, S.st.p, S.hist.p };
    const uint32_t nvb = vec_blocks(A.n);
    hipStream_t stream = runtime().stream;
    L.r0 = S.r32.p;
    const uint32_t nbz = step_blocks(*L.lv[0]);
    k_zero_u32<<<1, 64, 0, stream>>>(a.arrive, kArriveWords);
    RS_SHAPE_SWITCH(plan64, k_cg_init, nb64, stream, a);
    fin(S.partials.p, nb64, 1, FIN_CG_INIT, 0, S.st.p, S.hist.p);
    k_fcg64_start<<<nvb, kBlock, 0, stream>>>(S.r.p, S.r32.p, S.q32.p, A.n);
    RS_HIP(hipGetLastError());
    const std::function<void()> body = [&] {
        for (uint32_t i = 0; i < kGraphIters; i++) {
            const float *z = cycle_level(L, 0, S.st.p, S.q32.p, S.partials.p, Prolong{});
            if (L.rec_cap) {
                CycLevel &Cl = *L.lv.back();
                k_record<<<vec_blocks(L.nc), kBlock, 0, stream>>>(Cl.r.p, L.rec.p, L.nc, L.rec_cap, S.st.p);
            }
            fin(S.partials.p, nbz, 2, FIN_FCG_BETA, 0, S.st.p, S.hist.p);
            k_fcg64_p<<<nvb, kBlock, 0, stream>>>(z, S.p.p, A.n, S.st.p);
            RS_SHAPE_SWITCH(plan64, k_cg_spmv_dot, nb64, stream, a);   // finalizes alpha
            k_fcg64_update<<<nvb, kBlock, 0, stream>>>(u);               // finalizes the stopping rule
        }
    };
    rc = capture(S, body, graph_key_of(4, a, plan64, nb64, nbz, nvb) + learned_key(L));
    if (rc) return rc;
    return run_graph_loop(S, done, body);
}

}  // namespace rs

extern "C" int rs_aggregates_by_bins(uint32_t n_nodes, const double *xyz, uint32_t dofs_per_node, const uint32_t *book,
                                     uint32_t nbx, uint32_t nby, uint32_t nbz, uint32_t *agg_of_dof, uint32_t *n_agg)
{
    if (!xyz || !book || !agg_of_dof || !n_agg || !nbx || !nby || !nbz || !dofs_per_node)
        return fail(RS_ERR_INVALID_ARG, "rs_aggregates_by_bins: bad arguments");
    double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
    for (uint64_t i = 0; i < n_nodes; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], xyz[3 * i + a]);
            hi[a] = std::max(hi[a], xyz[3 * i + a]);
        }
    const uint32_t nbins[3] = { nbx, nby, nbz };
    std::vector<uint32_t> bin_id((uint64_t)nbx * nby * nbz, RS_EOD);
    uint32_t next = 0;
    // ids in first-appearance order of the free DOFs, so coarse ids follow the fine numbering
    for (uint64_t i = 0; i < n_nodes; i++) {
        uint32_t c[3];
        for (int a = 0; a < 3; a++) {
            double t = hi[a] > lo[a] ? (xyz[3 * i + a] - lo[a]) / (hi[a] - lo[a]) : 0.0;
            c[a] = std::min<uint32_t>(nbins[a] - 1, (uint32_t)(t * nbins[a]));
        }
        uint64_t bidx = c[0] + (uint64_t)nbx * (c[1] + (uint64_t)nby * c[2]);
        for (uint32_t k = 0; k < dofs_per_node; k++) {
            uint32_t d = book[i * dofs_per_node + k];
            if (d == RS_EOD) continue;
            if (bin_id[bidx] == RS_EOD) bin_id[bidx] = next++;
            agg_of_dof[d] = bin_id[bidx];
        }
    }
    *n_agg = next;
    return RS_OK;
}

extern "C" int rs_hierarchy_by_bins_ex(uint32_t n_nodes, const double *xyz, uint32_t dofs_per_node, const uint32_t *book,
                                       uint32_t nbx, uint32_t nby, uint32_t nbz, uint32_t max_coarse, uint32_t max_levels,
                                       int32_t aggregation, uint32_t *agg, uint64_t agg_capacity, uint32_t *n_agg,
                                       uint32_t *n_levels)
{
    if (!xyz || !book || !agg || !n_agg || !n_levels || !nbx || !nby || !nbz || !dofs_per_node || !max_coarse ||
        !max_levels || (aggregation != RS_AGG_NODES && aggregation != RS_AGG_COMPONENTS))
        return fail(RS_ERR_INVALID_ARG, "rs_hierarchy_by_bins: bad arguments");
    const bool by_component = aggregation == RS_AGG_COMPONENTS;
    // enabled positions, their nodes and components
    std::vector<uint32_t> node_of, comp_of;
    for (uint64_t i = 0; i < (uint64_t)n_nodes * dofs_per_node; i++)
        if (book[i] != RS_EOD) {
            if (book[i] >= node_of.size()) {
                node_of.resize(book[i] + 1ull, RS_EOD);
                comp_of.resize(book[i] + 1ull, 0);
            }
            node_of[book[i]] = (uint32_t)(i / dofs_per_node);
            comp_of[book[i]] = (uint32_t)(i % dofs_per_node);
        }
    const uint32_t n_dof = (uint32_t)node_of.size();
    for (uint32_t d = 0; d < n_dof; d++)
        if (node_of[d] == RS_EOD) return fail(RS_ERR_INVALID_ARG, "rs_hierarchy_by_bins: book is not a numbering 0..n-1");
    double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
    for (uint64_t i = 0; i < n_nodes; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], xyz[3 * i + a]);
            hi[a] = std::max(hi[a], xyz[3 * i + a]);
        }
    // rep[J]: a fine position of node J of the current level (its coordinates place J in a bin)
    std::vector<uint32_t> rep(n_dof);
    for (uint32_t d = 0; d < n_dof; d++) rep[d] = d;
    uint32_t cur = n_dof, levels = 0;
    uint64_t used = 0;
    for (uint32_t k = 0; cur > max_coarse || levels == 0; k++) {
        if (levels == max_levels) return fail(RS_ERR_UNSUPPORTED, "rs_hierarchy_by_bins: coarsest level still too large");
        const uint32_t sh = std::min(k, 31u);
        const uint32_t nb[3] = { std::max(1u, nbx >> sh), std::max(1u, nby >> sh), std::max(1u, nbz >> sh) };
        if (used + cur > agg_capacity) return fail(RS_ERR_INVALID_ARG, "rs_hierarchy_by_bins: agg capacity too small");
        std::vector<uint32_t> id((uint64_t)nb[0] * nb[1] * nb[2] * (by_component ? dofs_per_node : 1u), RS_EOD),
            next_rep;
        uint32_t count = 0;
        for (uint32_t J = 0; J < cur; J++) {
            const double *p = xyz + 3ull * node_of[rep[J]];
            uint64_t c[3];
            for (int a = 0; a < 3; a++) {
                double t = hi[a] > lo[a] ? (p[a] - lo[a]) / (hi[a] - lo[a]) : 0.0;
                c[a] = std::min<uint64_t>(nb[a] - 1, (uint64_t)(t * nb[a]));
            }
            uint64_t b = c[0] + nb[0] * (c[1] + (uint64_t)nb[1] * c[2]);
            if (by_component) b = b * dofs_per_node + comp_of[rep[J]];
            if (id[b] == RS_EOD) {
                id[b] = count++;
                next_rep.push_back(rep[J]);
            }
            agg[used + J] = id[b];
        }
        if (count == cur) {
            if (nb[0] == 1 && nb[1] == 1 && nb[2] == 1)
                return fail(RS_ERR_UNSUPPORTED, "rs_hierarchy_by_bins: cannot coarsen further");
            continue;   // the bins are still as fine as the nodes: halve again
        }
        n_agg[levels++] = count;
        used += cur;
        cur = count;
        rep.swap(next_rep);
    }
    *n_levels = levels;
    return RS_OK;
}

extern "C" int rs_hierarchy_by_bins(uint32_t n_nodes, const double *xyz, uint32_t dofs_per_node, const uint32_t *book,
                                    uint32_t nbx, uint32_t nby, uint32_t nbz, uint32_t max_coarse, uint32_t max_levels,
                                    uint32_t *agg, uint64_t agg_capacity, uint32_t *n_agg, uint32_t *n_levels)
{
    return rs_hierarchy_by_bins_ex(n_nodes, xyz, dofs_per_node, book, nbx, nby, nbz, max_coarse, max_levels,
                                   RS_AGG_NODES, agg, agg_capacity, n_agg, n_levels);
}

namespace rs {

// The hierarchy from device level maps (maps[l] sends the nodes of level l to level l + 1):
// members, next-level patterns and gather lists on the GPU; only the coarsest pattern (at most
// 512 rows) comes to the host, for the network's 16 x 16 blocks.
static int learned_finish(std::unique_ptr<rs_learned> &L, DBuf<uint32_t> *total, rs_learned **out);

static int learned_build(const rs_matrix *A, std::vector<std::unique_ptr<DBuf<uint32_t>>> &maps,
                         const std::vector<uint32_t> &n_agg, rs_learned **out)
{
    const uint32_t n_levels = (uint32_t)maps.size();
    if (n_levels == 0 || n_levels != n_agg.size()) return fail(RS_ERR_INVALID_ARG, "learned hierarchy: no levels");
    const uint32_t nc = n_agg.back();
    if (nc == 0 || nc > kMaxCoarse) return fail(RS_ERR_UNSUPPORTED, "rs_learned_create: coarsest level must have 1..512 nodes");
    int rc;
    std::unique_ptr<rs_learned> L(new rs_learned());
    L->n = A->n;
    L->nc = nc;
    L->nlev = n_levels + 1;
    L->owner = A;
    L->lv.emplace_back(new CycLevel());
    L->lv[0]->M = const_cast<rs_matrix *>(A);
    L->lv[0]->n = A->n;
    DBuf<uint32_t> total;
    if ((rc = total.alloc(std::max(A->n, 1u))) || (rc = hier_identity(total.p, A->n))) return rc;
    HierWork work;   // the finest level sizes the lists' scratch, the coarser levels reuse it
    for (uint32_t l = 0; l < n_levels; l++) {
        CycLevel &V = *L->lv[l];
        const uint32_t nn = n_agg[l];
        V.nnext = nn;
        std::swap(V.agg.p, maps[l]->p);
        std::swap(V.agg.n, maps[l]->n);
        DBuf<uint32_t> cptr, ccol;
        uint32_t cnnz = 0;
        if ((rc = hier_members(V.agg.p, V.n, nn, V.mptr, V.mrows, &work))) return rc;
        if ((rc = hier_level_lists(V.M->rowptr.p, V.M->col.p, V.n, V.M->nnz, V.agg.p, nn, cptr, ccol, V.gptr, V.gsrc,
                                   cnnz, &work)))
            return rc;
        std::unique_ptr<CycLevel> C(new CycLevel());
        if ((rc = matrix_create_device(nn, cptr, ccol, cnnz, &C->own))) return rc;
        C->M = C->own;
        C->n = nn;
        const double avg = (double)V.n / nn;
        V.rgroup = avg <= 6 ? 1 : avg <= 24 ? 4 : avg <= 96 ? 16 : 64;
        if ((rc = hier_compose(total.p, A->n, V.agg.p))) return rc;
        L->lv.push_back(std::move(C));
    }
    return learned_finish(L, &total, out);
}

// Level buffers, the fine rows of each coarsest node (scalar hierarchies: total maps every fine
// row to its coarsest node) and the coarsest graph's blocks for the network.
static int learned_finish(std::unique_ptr<rs_learned> &L, DBuf<uint32_t> *total, rs_learned **out)
{
    int rc;
    hipStream_t st = runtime().stream;
    const uint32_t nc = L->nc;
    for (auto &lp : L->lv) {
        CycLevel &V = *lp;
        const size_t n = std::max(V.n, 1u);
        if ((rc = V.diag.alloc(n)) || (rc = V.dinv.alloc(n)) || (rc = V.r.alloc(n)) || (rc = V.za.alloc(n)) || (rc = V.zb.alloc(n)) ||
            (rc = V.d.alloc(n)) || (rc = V.res.alloc(n)))
            return rc;
    }
    // fine rows of each coarsest node
    if (total && (rc = hier_members(total->p, L->n, nc, L->tot_mptr, L->tot_mrows))) return rc;
    // coarsest graph blocks, from the coarsest pattern
    const Matrix &Cc = *L->lv.back()->M;
    std::vector<uint32_t> rp(nc + 1ull), ci(Cc.nnz);
    RS_HIP(hipMemcpyAsync(rp.data(), Cc.rowptr.p, sizeof(uint32_t) * rp.size(), hipMemcpyDeviceToHost, st));
    if (Cc.nnz) RS_HIP(hipMemcpyAsync(ci.data(), Cc.col.p, sizeof(uint32_t) * Cc.nnz, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    const uint32_t nstrips = (nc + 15) / 16;
    std::vector<uint32_t> gbptr(nstrips + 1ull, 0), gbcol, gbidx(ci.size());
    for (uint32_t s = 0; s < nstrips; s++) {
        std::vector<uint32_t> ts;
        for (uint32_t I = s * 16; I < std::min(nc, s * 16 + 16); I++)
            for (uint32_t p = rp[I]; p < rp[I + 1]; p++) ts.push_back(ci[p] / 16);
        std::sort(ts.begin(), ts.end());
        ts.erase(std::unique(ts.begin(), ts.end()), ts.end());
        const uint32_t b0 = (uint32_t)gbcol.size();
        gbcol.insert(gbcol.end(), ts.begin(), ts.end());
        gbptr[s + 1] = (uint32_t)gbcol.size();
        for (uint32_t I = s * 16; I < std::min(nc, s * 16 + 16); I++)
            for (uint32_t p = rp[I]; p < rp[I + 1]; p++) {
                const uint32_t t = ci[p] / 16;
                const uint32_t b = b0 + (uint32_t)(std::lower_bound(ts.begin(), ts.end(), t) - ts.begin());
                gbidx[p] = b * 256 + (I % 16) * 16 + ci[p] % 16;
            }
    }
    rs_learned &R = *L;
    if ((rc = R.gbptr.alloc(gbptr.size())) || (rc = R.gbcol.alloc(std::max<size_t>(gbcol.size(), 1))) ||
        (rc = R.gbidx.alloc(std::max<size_t>(gbidx.size(), 1))) || (rc = R.gbval.alloc(std::max<size_t>(gbcol.size(), 1) * 256)) ||
        (rc = R.gh.alloc(2ull * nstrips * 16 * kHid)) || (rc = R.ght.alloc(2ull * nstrips * 16 * kHid)) ||
        (rc = R.gu.alloc(nstrips * 16)) || (rc = R.gnorm.alloc(2)) || (rc = R.s16bad.alloc(R.nlev)))
        return rc;
    RS_HIP(hipMemcpyAsync(R.gbptr.p, gbptr.data(), sizeof(uint32_t) * gbptr.size(), hipMemcpyHostToDevice, st));
    if (!gbcol.empty())
        RS_HIP(hipMemcpyAsync(R.gbcol.p, gbcol.data(), sizeof(uint32_t) * gbcol.size(), hipMemcpyHostToDevice, st));
    if (!gbidx.empty())
        RS_HIP(hipMemcpyAsync(R.gbidx.p, gbidx.data(), sizeof(uint32_t) * gbidx.size(), hipMemcpyHostToDevice, st));
    if ((rc = R.cw.alloc(std::max<uint64_t>(Cc.nnz, 1))) || (rc = R.cdiag.alloc(nc)) ||
        (rc = R.coef.alloc((size_t)R.nlev * kMaxDeg * 2)) || (rc = R.tmp_rc.alloc(nc)) ||
        (rc = R.part.alloc((size_t)R.nlev * 4096)) || (rc = R.gersh.alloc(R.nlev)) || (rc = R.theta.alloc(kNumParams)) ||
        (rc = R.wbf.alloc(kLayers * kBfPerLayer)))
        return rc;
    RS_HIP(hipStreamSynchronize(st));
    *out = L.release();
    return RS_OK;
}

// The rigid-body hierarchy of a 3-DOF matrix from node-aggregate maps (maps[0]: DOFs -> aggregates,
// maps[l]: aggregates of level l -> level l + 1) and the mesh in device memory.  Level l + 1 has
// 6 rows per aggregate of maps[l]; everything is built on the GPU and depends on the mesh only.
static int learned_build_rigid(const rs_matrix *A, std::vector<std::unique_ptr<DBuf<uint32_t>>> &maps,
                               const std::vector<uint32_t> &n_agg, uint32_t n_nodes, const double *d_xyz,
                               const uint32_t *d_book, rs_learned **out)
{
    const uint32_t n_levels = (uint32_t)maps.size();
    if (n_levels == 0 || n_levels != n_agg.size()) return fail(RS_ERR_INVALID_ARG, "rigid hierarchy: no levels");
    const uint64_t nc64 = 6ull * n_agg.back();
    if (nc64 == 0 || nc64 > kMaxCoarse) return fail(RS_ERR_UNSUPPORTED, "rigid hierarchy: coarsest level must have 1..512 rows");
    int rc;
    hipStream_t st = runtime().stream;
    std::unique_ptr<rs_learned> L(new rs_learned());
    L->n = A->n;
    L->nc = (uint32_t)nc64;
    L->nlev = n_levels + 1;
    L->owner = A;
    L->rigid = true;
    L->lv.emplace_back(new CycLevel());
    CycLevel &V0 = *L->lv[0];
    V0.M = const_cast<rs_matrix *>(A);
    V0.n = A->n;
    if ((rc = V0.ref.alloc(3ull * std::max(A->n, 1u))) || (rc = V0.comp.alloc(std::max(A->n, 1u)))) return rc;
    k_rigid_dofs<<<vec_blocks(3ull * n_nodes), kBlock, 0, st>>>(d_book, 3ull * n_nodes, d_xyz, V0.ref.p, V0.comp.p);
    DBuf<uint32_t> bad;
    if ((rc = bad.alloc(1))) return rc;
    k_fill_u32<<<1, 64, 0, st>>>(bad.p, 1, 0u);
    RS_HIP(hipGetLastError());
    HierWork work;   // the finest level sizes the lists' scratch, the coarser levels reuse it
    for (uint32_t l = 0; l < n_levels; l++) {
        CycLevel &V = *L->lv[l];
        const uint32_t nn = n_agg[l];
        V.rigid = true;
        V.nnext = 6 * nn;
        if (l == 0) {
            std::swap(V.agg.p, maps[0]->p);
            std::swap(V.agg.n, maps[0]->n);
        } else {
            if ((rc = V.agg.alloc(std::max(V.n, 1u)))) return rc;
            k_expand6<<<vec_blocks(V.n), kBlock, 0, st>>>(maps[l]->p, V.n, V.agg.p);
        }
        uint32_t cnnz = 0;
        DBuf<uint32_t> cptr, ccol;
        if ((rc = hier_members(V.agg.p, V.n, nn, V.mptr, V.mrows, &work))) return rc;
        if ((rc = hier_level_lists(V.M->rowptr.p, V.M->col.p, V.n, V.M->nnz, V.agg.p, nn, cptr, ccol, V.gptr, V.gsrc,
                                   cnnz, &work)))
            return rc;
        if (36ull * cnnz >= 0xFFFFFFFFull) return fail(RS_ERR_UNSUPPORTED, "rigid hierarchy: coarse level too large");
        // members per aggregate within one wavefront's reach
        std::vector<uint32_t> hm(nn + 1ull);
        RS_HIP(hipMemcpyAsync(hm.data(), V.mptr.p, sizeof(uint32_t) * hm.size(), hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
        uint32_t mmax = 0;
        for (uint32_t J = 0; J < nn; J++) mmax = std::max(mmax, hm[J + 1] - hm[J]);
        if (mmax > kRigidMaxMembers)
            return fail(RS_ERR_UNSUPPORTED, "rigid hierarchy: an aggregate has more than 256 rows (use finer bins)");
        std::unique_ptr<CycLevel> C(new CycLevel());
        C->n = 6 * nn;
        if ((rc = C->bsrc.alloc(36ull * nn)) || (rc = C->ref.alloc(18ull * nn)) || (rc = C->dead.alloc(6ull * nn)) ||
            (rc = V.q64.alloc(6ull * std::max(V.n, 1u))) || (rc = V.q32.alloc(6ull * std::max(V.n, 1u))))
            return rc;
        RigidQrArgs qa{ nn, V.mptr.p, V.mrows.p, V.ref.p, l == 0 ? V.comp.p : nullptr, l == 0 ? nullptr : V.bsrc.p,
                        V.q64.p, V.q32.p, C->bsrc.p, C->ref.p, C->dead.p, bad.p };
        k_rigid_qr<<<vec_blocks((uint64_t)nn * kWave), kBlock, 0, st>>>(qa);
        DBuf<uint32_t> erowptr, ecol;
        if ((rc = erowptr.alloc(6ull * nn + 1)) || (rc = ecol.alloc(std::max<uint64_t>(36ull * cnnz, 1))) ||
            (rc = V.ncrow.alloc(std::max(cnnz, 1u))) || (rc = V.erow.alloc(std::max<uint64_t>(V.M->nnz, 1))))
            return rc;
        k_rigid_expand<<<vec_blocks(nn), kBlock, 0, st>>>(cptr.p, ccol.p, nn, erowptr.p, ecol.p, V.ncrow.p);
        k_entry_rows<<<vec_blocks(V.n), kBlock, 0, st>>>(V.M->rowptr.p, V.n, V.erow.p);
        RS_HIP(hipGetLastError());
        V.ncrow.n = cnnz;
        std::swap(V.ncptr.p, cptr.p);
        std::swap(V.ncptr.n, cptr.n);
        std::swap(V.nccol.p, ccol.p);
        std::swap(V.nccol.n, ccol.n);
        if ((rc = matrix_create_device(6 * nn, erowptr, ecol, 36ull * cnnz, &C->own))) return rc;
        C->M = C->own;
        const double avg = (double)V.n / nn;
        V.rgroup = avg <= 24 ? 4 : avg <= 96 ? 16 : 64;
        L->lv.push_back(std::move(C));
    }
    uint32_t hbad = 0;
    RS_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    if (hbad) return fail(RS_ERR_UNSUPPORTED, "rigid hierarchy: an aggregate has more than 256 rows");
    return learned_finish(L, nullptr, out);
}

}  // namespace rs

extern "C" int rs_learned_create(const rs_matrix *A, uint32_t n_levels, const uint32_t *agg, const uint32_t *n_agg,
                                 rs_learned **out)
{
    if (!A || !agg || !n_agg || !out || n_levels == 0 || n_levels > 16)
        return fail(RS_ERR_INVALID_ARG, "rs_learned_create: bad arguments");
    const uint32_t nc = n_agg[n_levels - 1];
    if (nc == 0 || nc > kMaxCoarse) return fail(RS_ERR_UNSUPPORTED, "rs_learned_create: coarsest level must have 1..512 nodes");
    {
        uint64_t off = 0;
        uint32_t nprev = A->n;
        for (uint32_t l = 0; l < n_levels; l++) {
            if (n_agg[l] == 0 || n_agg[l] > nprev) return fail(RS_ERR_INVALID_ARG, "rs_learned_create: bad level size");
            std::vector<uint8_t> seen(n_agg[l], 0);
            for (uint32_t i = 0; i < nprev; i++) {
                if (agg[off + i] >= n_agg[l]) return fail(RS_ERR_INVALID_ARG, "rs_learned_create: aggregate id out of range");
                seen[agg[off + i]] = 1;
            }
            for (uint32_t I = 0; I < n_agg[l]; I++)
                if (!seen[I]) return fail(RS_ERR_INVALID_ARG, "rs_learned_create: empty aggregate");
            off += nprev;
            nprev = n_agg[l];
        }
    }
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    std::vector<std::unique_ptr<DBuf<uint32_t>>> maps;
    uint64_t off = 0;
    uint32_t nprev = A->n;
    for (uint32_t l = 0; l < n_levels; l++) {
        maps.emplace_back(new DBuf<uint32_t>());
        if ((rc = maps.back()->alloc(nprev))) return rc;
        if (nprev) RS_HIP(hipMemcpyAsync(maps.back()->p, agg + off, sizeof(uint32_t) * nprev, hipMemcpyHostToDevice, st));
        off += nprev;
        nprev = n_agg[l];
    }
    RS_HIP(hipStreamSynchronize(st));
    return learned_build(A, maps, std::vector<uint32_t>(n_agg, n_agg + n_levels), out);
}

extern "C" int rs_learned_create_by_bins_ex(const rs_matrix *A, uint32_t n_nodes, const double *d_xyz,
                                            uint32_t dofs_per_node, const uint32_t *d_book, uint32_t nbx, uint32_t nby,
                                            uint32_t nbz, uint32_t max_coarse, uint32_t max_levels, int32_t aggregation,
                                            rs_learned **out)
{
    if (!A || !d_xyz || !d_book || !out || !n_nodes || !dofs_per_node || !nbx || !nby || !nbz || !max_coarse ||
        !max_levels || max_levels > 16 ||
        (aggregation != RS_AGG_NODES && aggregation != RS_AGG_COMPONENTS && aggregation != RS_AGG_RIGID))
        return fail(RS_ERR_INVALID_ARG, "rs_learned_create_by_bins: bad arguments");
    if (aggregation == RS_AGG_RIGID && (dofs_per_node != 3 || max_coarse < 6))
        return fail(RS_ERR_INVALID_ARG, "rs_learned_create_by_bins: rigid-body aggregation needs 3 DOFs per node "
                                        "and max_coarse >= 6");
    if (max_coarse > kMaxCoarse) return fail(RS_ERR_UNSUPPORTED, "rs_learned_create_by_bins: max_coarse above 512");
    int rc = require_runtime();
    if (rc) return rc;
    std::vector<std::unique_ptr<DBuf<uint32_t>>> maps;
    std::vector<uint32_t> n_agg;
    const uint32_t bins[3] = { nbx, nby, nbz };
    if (aggregation == RS_AGG_RIGID) {
        // node aggregates (the DOFs of a node together), at most max_coarse / 6 on the coarsest level
        if ((rc = hier_bins(n_nodes, d_xyz, dofs_per_node, d_book, A->n, bins, max_coarse / 6, max_levels, false, maps,
                            n_agg)))
            return rc;
        return learned_build_rigid(A, maps, n_agg, n_nodes, d_xyz, d_book, out);
    }
    if ((rc = hier_bins(n_nodes, d_xyz, dofs_per_node, d_book, A->n, bins, max_coarse, max_levels,
                        aggregation == RS_AGG_COMPONENTS, maps, n_agg)))
        return rc;
    return learned_build(A, maps, n_agg, out);
}

extern "C" int rs_learned_create_by_bins(const rs_matrix *A, uint32_t n_nodes, const double *d_xyz,
                                         uint32_t dofs_per_node, const uint32_t *d_book, uint32_t nbx, uint32_t nby,
                                         uint32_t nbz, uint32_t max_coarse, uint32_t max_levels, rs_learned **out)
{
    return rs_learned_create_by_bins_ex(A, n_nodes, d_xyz, dofs_per_node, d_book, nbx, nby, nbz, max_coarse,
                                        max_levels, RS_AGG_NODES, out);
}

extern "C" int rs_learned_level_map(const rs_learned *L, uint32_t level, uint32_t *agg, uint64_t capacity)
{
    if (!L || !agg || level + 1 >= L->nlev) return fail(RS_ERR_INVALID_ARG, "rs_learned_level_map: bad arguments");
    const CycLevel &V = *L->lv[level];
    if (capacity < V.n) return fail(RS_ERR_INVALID_ARG, "rs_learned_level_map: capacity below the level's rows");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    if (V.n) RS_HIP(hipMemcpyAsync(agg, V.agg.p, sizeof(uint32_t) * V.n, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}
extern "C" int rs_learned_set_params(rs_learned *L, const float *params, uint64_t count)
{
    if (!L || !params) return fail(RS_ERR_INVALID_ARG, "rs_learned_set_params: null argument");
    if (count != kNumParams) return fail(RS_ERR_INVALID_ARG, "rs_learned_set_params: expected 49921 parameters");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(L->theta.p, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, st));
    k_weights_bf16_t<<<vec_blocks(kLayers * kBfPerLayer), kBlock, 0, st>>>(L->theta.p, L->wbf.p);
    RS_HIP(hipGetLastError());
    RS_HIP(hipStreamSynchronize(st));
    L->have_params = true;
    return RS_OK;
}

extern "C" int rs_learned_set_cycle(rs_learned *L, uint32_t pre, uint32_t post, int coarse_mode)
{
    if (!L) return fail(RS_ERR_INVALID_ARG, "rs_learned_set_cycle: null");
    if (pre > kMaxDeg || post == 0 || post > kMaxDeg)
        return fail(RS_ERR_INVALID_ARG, "rs_learned_set_cycle: pre in 0..4 and post in 1..4 smoothing steps");
    if (coarse_mode < RS_COARSE_NETWORK || coarse_mode > RS_COARSE_EXACT)
        return fail(RS_ERR_INVALID_ARG, "rs_learned_set_cycle: unknown coarse mode");
    L->pre = L->cpre = pre;
    L->post = L->cpost = post;
    L->coarse = coarse_mode;
    return RS_OK;
}

extern "C" int rs_learned_set_coarse_cycle(rs_learned *L, uint32_t pre, uint32_t post)
{
    if (!L) return fail(RS_ERR_INVALID_ARG, "rs_learned_set_coarse_cycle: null");
    if (pre > kMaxDeg || post == 0 || post > kMaxDeg)
        return fail(RS_ERR_INVALID_ARG, "rs_learned_set_coarse_cycle: pre in 0..4 and post in 1..4 smoothing steps");
    L->cpre = pre;
    L->cpost = post;
    return RS_OK;
}

extern "C" int rs_learned_set_compressed(rs_learned *L, uint32_t min_rows)
{
    if (!L) return fail(RS_ERR_INVALID_ARG, "rs_learned_set_compressed: null");
    L->s16_min_rows = min_rows;
    L->s16_auto = false;   // an explicit threshold is followed as given
    L->setup_version = ~0ull;   // rebuilt at the next set-up
    return RS_OK;
}

extern "C" int rs_learned_info(const rs_learned *L, uint32_t *n_levels, uint32_t *level_rows, uint64_t *level_nnz,
                               uint64_t *n_params)
{
    if (!L) return fail(RS_ERR_INVALID_ARG, "rs_learned_info: null");
    if (n_levels) *n_levels = L->nlev;
    for (uint32_t l = 0; l < L->nlev; l++) {
        if (level_rows) level_rows[l] = L->lv[l]->n;
        if (level_nnz) level_nnz[l] = L->lv[l]->M->nnz;
    }
    if (n_params) *n_params = kNumParams;
    return RS_OK;
}

extern "C" int rs_learned_setup(rs_learned *L, rs_matrix *A)
{
    if (!L || !A) return fail(RS_ERR_INVALID_ARG, "rs_learned_setup: null");
    int rc = require_runtime();
    if (rc) return rc;
    if ((rc = learned_setup(*L, *A))) return rc;
    RS_HIP(hipStreamSynchronize(runtime().stream));
    return RS_OK;
}

extern "C" int rs_learned_level_lambda(const rs_learned *L, double *lambda)
{
    if (!L || !lambda) return fail(RS_ERR_INVALID_ARG, "rs_learned_level_lambda: null");
    if (L->lam.size() != L->nlev) return fail(RS_ERR_INVALID_ARG, "rs_learned_level_lambda: call rs_learned_setup first");
    for (uint32_t l = 0; l < L->nlev; l++) lambda[l] = L->lam[l];
    return RS_OK;
}

// Download the matrix of level l (CSR, FP64 values of the last set-up).
extern "C" int rs_learned_level_matrix(const rs_learned *L, uint32_t level, uint32_t *row_ptr, uint32_t *col, double *val)
{
    if (!L || level >= L->nlev) return fail(RS_ERR_INVALID_ARG, "rs_learned_level_matrix: bad level");
    return rs_matrix_download(static_cast<const rs_matrix *>(L->lv[level]->M), row_ptr, col, val);
}

extern "C" int rs_learned_coarse_matrix(const rs_learned *L, double *dense)
{
    if (!L || !dense) return fail(RS_ERR_INVALID_ARG, "rs_learned_coarse_matrix: null");
    int rc = require_runtime();
    if (rc) return rc;
    std::vector<double> d;
    if ((rc = coarsest_dense(*L, d))) return rc;
    std::copy(d.begin(), d.end(), dense);
    return RS_OK;
}

// The rigid-body basis rows of level `level` (6 per row of the level, FP64).
extern "C" int rs_learned_level_basis(const rs_learned *L, uint32_t level, double *q, uint64_t capacity)
{
    if (!L || !q || level + 1 >= L->nlev) return fail(RS_ERR_INVALID_ARG, "rs_learned_level_basis: bad arguments");
    const CycLevel &V = *L->lv[level];
    if (!V.rigid) return fail(RS_ERR_INVALID_ARG, "rs_learned_level_basis: not a rigid-body level");
    if (capacity < 6ull * V.n) return fail(RS_ERR_INVALID_ARG, "rs_learned_level_basis: capacity below 6 x the level's rows");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    if (V.n) RS_HIP(hipMemcpyAsync(q, V.q64.p, sizeof(double) * 6 * V.n, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

// A fine vector (device, FP64) restricted straight to the coarsest level.
extern "C" int rs_learned_restrict_device(rs_learned *L, const double *d_r, float *rc_host)
{
    if (!L || !d_r || !rc_host) return fail(RS_ERR_INVALID_ARG, "rs_learned_restrict_device: null");
    if (L->rigid) return fail(RS_ERR_UNSUPPORTED, "rs_learned_restrict_device: not for rigid-body hierarchies");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    k_restrict_total<<<L->nc, kBlock, 0, st>>>(d_r, L->tot_mptr.p, L->tot_mrows.p, L->tmp_rc.p);
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(rc_host, L->tmp_rc.p, sizeof(float) * L->nc, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

// Network output for a host coarsest residual (the coarsest operator of the last set-up).
extern "C" int rs_learned_apply_coarse(rs_learned *L, const float *rc_host, float *e_host)
{
    if (!L || !rc_host || !e_host) return fail(RS_ERR_INVALID_ARG, "rs_learned_apply_coarse: null");
    if (L->coarse == RS_COARSE_NETWORK && !L->have_params)
        return fail(RS_ERR_INVALID_ARG, "rs_learned_apply_coarse: parameters not set");
    if (L->lam.empty() || (L->coarse == RS_COARSE_EXACT && !L->inverse_valid))
        return fail(RS_ERR_INVALID_ARG, "rs_learned_apply_coarse: call rs_learned_setup first");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    CycLevel &V = *L->lv.back();
    RS_HIP(hipMemcpyAsync(V.r.p, rc_host, sizeof(float) * L->nc, hipMemcpyHostToDevice, st));
    // the configured coarsest solver: the network, coarse Jacobi or the exact inverse
    if (L->coarse == RS_COARSE_NETWORK) launch_gnn(gnn_args(*L, nullptr), st);
    else if (L->coarse == RS_COARSE_EXACT)
        k_coarse_exact<<<vec_blocks((uint64_t)L->nc * kWave), kBlock, 0, st>>>(L->cinv.p, V.r.p, V.za.p, L->nc, nullptr,
                                                                              Prolong{});
    else k_coarse_jacobi<<<vec_blocks(L->nc), kBlock, 0, st>>>(V.r.p, V.dinv.p, V.za.p, L->nc, nullptr, Prolong{});
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(e_host, V.za.p, sizeof(float) * L->nc, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

// Average duration of one network inference (embedding + 6 layer launches) over `reps`
// back-to-back inferences on the coarsest residual of the last cycle (stream events on the
// library stream).
extern "C" int rs_learned_infer_timed(rs_learned *L, uint32_t reps, double *avg_ms)
{
    if (!L || !avg_ms || reps == 0) return fail(RS_ERR_INVALID_ARG, "rs_learned_infer_timed: bad arguments");
    if (!L->have_params || L->lam.empty()) return fail(RS_ERR_INVALID_ARG, "rs_learned_infer_timed: set parameters and set up first");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    const GnnArgs g = gnn_args(*L, nullptr);
    launch_gnn(g, st);
    EventPair e0e1_;
    RS_HIP(e0e1_.create());
    hipEvent_t e0 = e0e1_.e[0], e1 = e0e1_.e[1];
    RS_HIP(hipEventRecord(e0, st));
    for (uint32_t i = 0; i < reps; i++) launch_gnn(g, st);
    RS_HIP(hipEventRecord(e1, st));
    RS_HIP(hipEventSynchronize(e1));
    RS_HIP(hipGetLastError());
    float ms = 0.0f;
    RS_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    return RS_OK;
}

// One V-cycle z = B r on host vectors of the fine size (set-up and plans as for a solve).
extern "C" int rs_learned_apply(rs_learned *L, rs_matrix *A, const float *r_host, float *z_host)
{
    if (!L || !A || !r_host || !z_host) return fail(RS_ERR_INVALID_ARG, "rs_learned_apply: null");
    if (L->coarse == RS_COARSE_NETWORK && !L->have_params)
        return fail(RS_ERR_INVALID_ARG, "rs_learned_apply: parameters not set");
    int rc = require_runtime();
    if (rc) return rc;
    if ((rc = learned_setup(*L, *A)) || (rc = learned_plans(*L))) return rc;
    hipStream_t st = runtime().stream;
    CycLevel &V = *L->lv[0];
    RS_HIP(hipMemcpyAsync(V.r.p, r_host, sizeof(float) * L->n, hipMemcpyHostToDevice, st));
    L->r0 = V.r.p;
    const float *z = cycle_level(*L, 0, nullptr, nullptr, nullptr, Prolong{});
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(z_host, z, sizeof(float) * L->n, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

// Keep the coarsest right-hand side of the first `capacity` iterations of the following
// learned solves (0 switches recording off).
extern "C" int rs_learned_record(rs_learned *L, uint32_t capacity)
{
    if (!L) return fail(RS_ERR_INVALID_ARG, "rs_learned_record: null");
    int rc = require_runtime();
    if (rc) return rc;
    if (capacity && (rc = L->rec.alloc((size_t)capacity * L->nc))) return rc;
    L->rec_cap = capacity;
    return RS_OK;
}

extern "C" int rs_learned_records(const rs_learned *L, uint32_t count, float *out)
{
    if (!L || !out || count > L->rec_cap) return fail(RS_ERR_INVALID_ARG, "rs_learned_records: bad arguments");
    int rc = require_runtime();
    if (rc) return rc;
    if (count) {
        RS_HIP(hipMemcpyAsync(out, L->rec.p, sizeof(float) * count * L->nc, hipMemcpyDeviceToHost, runtime().stream));
        RS_HIP(hipStreamSynchronize(runtime().stream));
    }
    return RS_OK;
}

extern "C" int rs_learned_destroy(rs_learned *L)
{
    delete L;
    return RS_OK;
}

extern "C" int rs_solver_solve_learned_device(rs_solver *S, rs_matrix *A, rs_learned *L, const double *d_b, double *d_x,
                                              rs_info *info)
{
    if (!S || !A || !L || (A->n && (!d_b || !d_x))) return fail(RS_ERR_INVALID_ARG, "rs_solver_solve_learned_device: null");
    if (L->coarse == RS_COARSE_NETWORK && !L->have_params)
        return fail(RS_ERR_INVALID_ARG, "learned preconditioner: parameters not set");
    if (S->conf.solver_type != RS_SOLVER_CG) return fail(RS_ERR_UNSUPPORTED, "learned preconditioner: flexible CG only");
    int rc = require_runtime();
    if (rc) return rc;
    return solve_device_ext(*S, *A, d_b, d_x, RS_PRECOND_LEARNED, 1, L, info);
}

// ---------------------------------------------------------------------------- trainer
struct rs_trainer {
    uint32_t max_batch = 0, max_nodes = 0;
    uint64_t max_nnz = 0, step = 0;
    rs::DBuf<float> theta, m, v, g, grads, loss, cval, rc, s, dh, dm;
    rs::DBuf<unsigned short> wbf, hs, mb, dy;
    rs::DBuf<uint32_t> cptr, ccol, tpos;
};

extern "C" int rs_trainer_create(const float *params, uint64_t count, uint32_t max_batch, uint32_t max_nodes,
                                 uint64_t max_nnz, rs_trainer **out)
{
    if (!params || !out || count != kNumParams || max_batch == 0 || max_nodes == 0 || max_nodes > kMaxCoarse ||
        max_nnz == 0)
        return fail(RS_ERR_INVALID_ARG, "rs_trainer_create: bad arguments (49921 parameters, <= 512 nodes)");
    int rc = require_runtime();
    if (rc) return rc;
    std::unique_ptr<rs_trainer> T(new rs_trainer());
    T->max_batch = max_batch;
    T->max_nodes = max_nodes;
    T->max_nnz = max_nnz;
    const uint64_t npad = (max_nodes + 15) / 16 * 16, act = npad * kHid;
    if ((rc = T->theta.alloc(kNumParams)) || (rc = T->m.alloc(kNumParams)) || (rc = T->v.alloc(kNumParams)) ||
        (rc = T->g.alloc(kNumParams)) || (rc = T->grads.alloc((uint64_t)max_batch * kNumParams)) ||
        (rc = T->loss.alloc(max_batch)) || (rc = T->cval.alloc(max_batch * max_nnz)) ||
        (rc = T->rc.alloc((uint64_t)max_batch * max_nodes)) || (rc = T->s.alloc(max_batch)) ||
        (rc = T->wbf.alloc(kLayers * kBfPerLayer)) || (rc = T->hs.alloc(max_batch * (kLayers + 1) * act)) ||
        (rc = T->mb.alloc(max_batch * act)) || (rc = T->dy.alloc(max_batch * act)) ||
        (rc = T->dh.alloc(max_batch * act)) || (rc = T->dm.alloc(max_batch * act)) ||
        (rc = T->cptr.alloc(max_nodes + 1ull)) || (rc = T->ccol.alloc(max_nnz)) || (rc = T->tpos.alloc(max_nnz)))
        return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(T->theta.p, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, st));
    k_fill32<<<vec_blocks(kNumParams), kBlock, 0, st>>>(T->m.p, kNumParams, 0.0f);
    k_fill32<<<vec_blocks(kNumParams), kBlock, 0, st>>>(T->v.p, kNumParams, 0.0f);
    k_weights_bf16<<<vec_blocks(kLayers * kBfPerLayer), kBlock, 0, st>>>(T->theta.p, T->wbf.p);
    RS_HIP(hipGetLastError());
    RS_HIP(hipStreamSynchronize(st));
    *out = T.release();
    return RS_OK;
}

// Per-sample losses and the batch-mean gradient (no update).  The B samples share the coarse
// pattern (cptr, ccol: n rows, nnz entries, structurally symmetric, diagonal present); cval:
// [B][nnz] operator values, rcv: [B][n] coarse residuals, s: [B] normalisers r.A^-1.r.
extern "C" int rs_trainer_gradient(rs_trainer *T, uint32_t n, uint64_t nnz, const uint32_t *cptr, const uint32_t *ccol,
                                   uint32_t B, const float *cval, const float *rcv, const float *s, float *loss_out,
                                   float *grad_out)
{
    if (!T || !cptr || !ccol || !cval || !rcv || !s || n == 0 || n > T->max_nodes || nnz == 0 || nnz > T->max_nnz ||
        B == 0 || B > T->max_batch || cptr[0] != 0 || cptr[n] != nnz)
        return fail(RS_ERR_INVALID_ARG, "rs_trainer_gradient: bad arguments");
    // transpose positions (the pattern must be structurally symmetric with its diagonal)
    std::vector<uint32_t> tpos(nnz);
    for (uint32_t I = 0; I < n; I++) {
        if (cptr[I + 1] < cptr[I]) return fail(RS_ERR_INVALID_ARG, "rs_trainer_gradient: cptr not monotone");
        bool diag = false;
        for (uint32_t p = cptr[I]; p < cptr[I + 1]; p++) {
            const uint32_t J = ccol[p];
            if (J >= n) return fail(RS_ERR_INVALID_ARG, "rs_trainer_gradient: column out of range");
            diag |= J == I;
            uint32_t q = cptr[J];
            while (q < cptr[J + 1] && ccol[q] != I) q++;
            if (q == cptr[J + 1]) return fail(RS_ERR_INVALID_ARG, "rs_trainer_gradient: pattern not symmetric");
            tpos[p] = q;
        }
        if (!diag) return fail(RS_ERR_INVALID_ARG, "rs_trainer_gradient: missing diagonal entry");
    }
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(T->cptr.p, cptr, sizeof(uint32_t) * (n + 1ull), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(T->ccol.p, ccol, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(T->tpos.p, tpos.data(), sizeof(uint32_t) * nnz, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(T->cval.p, cval, sizeof(float) * B * nnz, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(T->rc.p, rcv, sizeof(float) * B * n, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(T->s.p, s, sizeof(float) * B, hipMemcpyHostToDevice, st));
    TrainArgs t{};
    t.n = n;
    t.npad = (n + 15) / 16 * 16;
    t.nnz = (uint32_t)nnz;
    t.cptr = T->cptr.p;
    t.ccol = T->ccol.p;
    t.tpos = T->tpos.p;
    t.cval = T->cval.p;
    t.rc = T->rc.p;
    t.s = T->s.p;
    t.params = T->theta.p;
    t.wbf = T->wbf.p;
    t.grads = T->grads.p;
    t.loss = T->loss.p;
    t.hs = T->hs.p;
    t.mb = T->mb.p;
    t.dy = T->dy.p;
    t.dh = T->dh.p;
    t.dm = T->dm.p;
    k_gnn_train<<<B, kTrainThreads, 0, st>>>(t);
    k_grad_mean<<<vec_blocks(kNumParams), kBlock, 0, st>>>(T->grads.p, B, T->g.p);
    RS_HIP(hipGetLastError());
    if (loss_out) RS_HIP(hipMemcpyAsync(loss_out, T->loss.p, sizeof(float) * B, hipMemcpyDeviceToHost, st));
    if (grad_out) RS_HIP(hipMemcpyAsync(grad_out, T->g.p, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

// One Adam step on a batch; returns the mean loss.
extern "C" int rs_trainer_step(rs_trainer *T, uint32_t n, uint64_t nnz, const uint32_t *cptr, const uint32_t *ccol,
                               uint32_t B, const float *cval, const float *rcv, const float *s, float lr, float *mean_loss)
{
    std::vector<float> losses(std::max(B, 1u));
    int rc = rs_trainer_gradient(T, n, nnz, cptr, ccol, B, cval, rcv, s, losses.data(), nullptr);
    if (rc) return rc;
    T->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)T->step), c2 = 1.0f - std::pow(b2, (float)T->step);
    hipStream_t st = runtime().stream;
    k_adam<<<vec_blocks(kNumParams), kBlock, 0, st>>>(T->theta.p, T->m.p, T->v.p, T->g.p, lr, b1, b2, 1e-8f, c1, c2);
    k_weights_bf16<<<vec_blocks(kLayers * kBfPerLayer), kBlock, 0, st>>>(T->theta.p, T->wbf.p);
    RS_HIP(hipGetLastError());
    RS_HIP(hipStreamSynchronize(st));
    if (mean_loss) {
        double m = 0.0;
        for (uint32_t b = 0; b < B; b++) m += losses[b];
        *mean_loss = (float)(m / B);
    }
    return RS_OK;
}

extern "C" int rs_trainer_get_params(const rs_trainer *T, float *params, uint64_t count)
{
    if (!T || !params || count != kNumParams) return fail(RS_ERR_INVALID_ARG, "rs_trainer_get_params: bad arguments");
    int rc = require_runtime();
    if (rc) return rc;
    RS_HIP(hipMemcpyAsync(params, T->theta.p, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, runtime().stream));
    RS_HIP(hipStreamSynchronize(runtime().stream));
    return RS_OK;
}

extern "C" int rs_trainer_destroy(rs_trainer *T)
{
    delete T;
    return RS_OK;
}
