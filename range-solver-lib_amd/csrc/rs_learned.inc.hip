// rs_learned.inc.hip -- learned two-level preconditioner (no counterpart in the reference).
// Included at the end of rs_solve.hip: it shares that translation unit's solver state, graph
// replay, finalize kernel and SpMV kernels.
//
// Preconditioner:  z = D^-1 r + P e,   e = N_theta(P^T r; A_c)
//   P      piecewise-constant prolongation over caller-given aggregates (coordinate bins)
//   A_c    Galerkin coarse operator P^T A P, accumulated in 64-bit fixed point (integer
//          atomics are associative, so the result does not depend on scheduling)
//   N      6-layer, 64-wide message-passing network over the coarse graph: messages are the
//          neighbour features weighted by -A_c[I,J]/A_c[I,I]; BF16 weights and activations on
//          MFMA (v_mfma_f32_16x16x16_bf16), FP32 accumulation; output e = u + sigma * o, a
//          learned correction of the coarse Jacobi estimate u = r_c / diag(A_c).
// Inference runs in one workgroup with the whole coarse graph resident in LDS (<= 512 coarse
// nodes: 2 x 64 KB of BF16 features).  The solve is flexible CG (Notay's FCG(1): beta from the
// A-orthogonalisation against the previous direction) in mixed precision with FP64 residual
// replacement; Jacobi-CG finishes the solve if the flexible iteration stalls.
//
// Training kernels: one workgroup per (residual, error) sample on graphs of <= 64 coarse
// nodes: forward with the activations kept in LDS, the loss (e.A_c.e - 2 e.r_c)/s + 1 (the
// A-norm error fraction the correction leaves, s = e_k.r_k), backward, per-sample gradients;
// then a fixed-order batch mean and Adam.

namespace rs {

constexpr int kHid = 64;
constexpr int kNin = 4;
constexpr int kLayers = 6;
constexpr uint32_t kMaxCoarse = 512;
constexpr uint32_t kTrainN = 64;
constexpr uint32_t kLayerStride = 2 * kHid * kHid + kHid;
constexpr uint32_t kOffWin = 0, kOffBin = kNin * kHid, kOffL0 = kNin * kHid + kHid;
constexpr uint32_t kOffWo = kOffL0 + kLayers * kLayerStride, kOffBo = kOffWo + kHid;
constexpr uint32_t kNumParams = kOffBo + 1;   // 49,921
constexpr uint32_t kBfPerLayer = 2 * kHid * kHid;
__host__ __device__ constexpr uint32_t off_ws(int l) { return kOffL0 + l * kLayerStride; }
__host__ __device__ constexpr uint32_t off_wn(int l) { return off_ws(l) + kHid * kHid; }
__host__ __device__ constexpr uint32_t off_b(int l) { return off_ws(l) + 2 * kHid * kHid; }

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ unsigned short f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

// One 16x16 output tile: acc += A(16 x K) B(K x 16), K a multiple of 16, operands fetched as
// BF16 bit patterns through a(row, k) and b(k, col).  Lane l supplies A[l%16][4(l/16)+i] and
// B[4(l/16)+i][l%16]; acc[i] is C[4(l/16)+i][l%16].
template <class FA, class FB>
__device__ __forceinline__ f32x4 mfma_tile(FA a, FB b, int K, f32x4 acc)
{
    const int lane = threadIdx.x & 63, rc = lane & 15, kq = (lane >> 4) * 4;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 av, bv;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            av[i] = (short)a(rc, k0 + kq + i);
            bv[i] = (short)b(k0 + kq + i, rc);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

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

// ---------------------------------------------------------------------------- setup kernels
__global__ __launch_bounds__(kBlock) void k_maxabs(const double *__restrict__ v, uint64_t n, double *partials)
{
    __shared__ double red[kBlock / kWave];
    double m = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        m = fmax(m, fabs(v[i]));
    for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; w++) t = fmax(t, red[w]);
        partials[blockIdx.x] = t;
    }
}

__global__ void k_maxabs_fin(const double *partials, uint32_t nparts, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t = 0.0;
    for (uint32_t i = 0; i < nparts; i++) t = fmax(t, partials[i]);
    *out = t;
}

__global__ void k_zero_u64(unsigned long long *__restrict__ a, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) a[i] = 0ull;
}

// Fixed-point scale: |entries| <= 2^44, so 2^19 of headroom for sums of up to half a million
// terms; resolution 2^-44 * max|A|.
__device__ __forceinline__ double galerkin_scale(double maxabs) { return ldexp(1.0, 44) / fmax(maxabs, 1e-300); }

// A_c[I][J] += A_ij for I = agg(i), J = agg(j); entries of a row that fall in the same coarse
// column are combined before the atomics.
__global__ void k_galerkin(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ col,
                           const double *__restrict__ val, uint32_t n, const uint32_t *__restrict__ agg,
                           uint32_t nc, const double *__restrict__ maxabs, unsigned long long *__restrict__ acc)
{
    const double scale = galerkin_scale(*maxabs);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t I = agg[i];
        uint32_t cj[8];
        long long cv[8];
        int nk = 0;
        for (uint32_t j = rowptr[i]; j < rowptr[i + 1]; j++) {
            uint32_t J = agg[col[j]];
            long long qv = __double2ll_rn(val[j] * scale);
            int k = 0;
            while (k < nk && cj[k] != J) k++;
            if (k < nk) {
                cv[k] += qv;
            } else if (nk < 8) {
                cj[nk] = J;
                cv[nk] = qv;
                nk++;
            } else {
                atomicAdd(&acc[(uint64_t)I * nc + J], (unsigned long long)qv);
            }
        }
        for (int k = 0; k < nk; k++) atomicAdd(&acc[(uint64_t)I * nc + cj[k]], (unsigned long long)cv[k]);
    }
}

// Coarse CSR values, diagonal, message weights -A_c[I,J]/A_c[I,I], and the dense copy.
__global__ void k_galerkin_finish(const unsigned long long *__restrict__ acc, uint32_t nc,
                                  const double *__restrict__ maxabs, const uint32_t *__restrict__ cptr,
                                  const uint32_t *__restrict__ ccol, double *__restrict__ cval,
                                  float *__restrict__ cw, float *__restrict__ cdiag, double *__restrict__ dense)
{
    const double inv = 1.0 / galerkin_scale(*maxabs);
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < (uint64_t)nc * nc; k += (uint64_t)gridDim.x * kBlock)
        dense[k] = (double)(long long)acc[k] * inv;
    for (uint32_t I = blockIdx.x * kBlock + threadIdx.x; I < nc; I += gridDim.x * kBlock) {
        double d = (double)(long long)acc[(uint64_t)I * nc + I] * inv;
        cdiag[I] = (float)d;
        for (uint32_t p = cptr[I]; p < cptr[I + 1]; p++) {
            uint32_t J = ccol[p];
            double a = (double)(long long)acc[(uint64_t)I * nc + J] * inv;
            cval[p] = a;
            cw[p] = (J == I || d == 0.0) ? 0.0f : (float)(-a / d);
        }
    }
}

// r_c[I] = sum of r over the rows of aggregate I (fixed order: member list, then block sum).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(const T *__restrict__ r, const uint32_t *__restrict__ mptr,
                                                     const uint32_t *__restrict__ mrows, float *__restrict__ rc,
                                                     const SolveState *st)
{
    if (st && st->done) return;
    __shared__ double lds[kBlock / kWave];
    const uint32_t I = blockIdx.x;
    double v[1] = { 0.0 };
    for (uint32_t k = mptr[I] + threadIdx.x; k < mptr[I + 1]; k += kBlock) v[0] += (double)r[mrows[k]];
    block_sum<1, kBlock>(v, lds);
    if (threadIdx.x == 0) rc[I] = (float)v[0];
}

// ---------------------------------------------------------------------------- inference
struct GnnArgs {
    uint32_t nc, npad;
    const float *rc, *diag;
    const uint32_t *cptr, *ccol;
    const float *cw;
    const float *params;
    const unsigned short *wbf;
    float *e;
    const SolveState *st;
};

constexpr int kInferThreads = 1024;

__global__ __launch_bounds__(kInferThreads) void k_gnn_infer(GnnArgs g)
{
    if (g.st && g.st->done) return;
    __shared__ unsigned short H[kMaxCoarse * kHid];
    __shared__ unsigned short M[kMaxCoarse * kHid];
    __shared__ unsigned short W[kBfPerLayer];
    __shared__ float u[kMaxCoarse], ld[kMaxCoarse];
    __shared__ float red[kInferThreads / kWave];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid / 64;
    const uint32_t nc = g.nc, npad = g.npad;
    const float *P = g.params;
    float su = 0.0f, sl = 0.0f;
    for (uint32_t I = tid; I < npad; I += kInferThreads) {
        float uu = 0.0f, l = 0.0f;
        if (I < nc) {
            float d = g.diag[I];
            uu = g.rc[I] / d;
            l = logf(d);
        }
        u[I] = uu;
        ld[I] = l;
        su += uu * uu;
        sl += l;
    }
    const float sigma = sqrtf(block_sum_f(su, red, kInferThreads) / (float)nc) + 1e-30f;
    const float meanlog = block_sum_f(sl, red, kInferThreads) / (float)nc;
    for (uint32_t idx = tid; idx < npad * kHid; idx += kInferThreads) {
        uint32_t I = idx / kHid, f = idx % kHid;
        float z = 0.0f;
        if (I < nc)
            z = (u[I] / sigma) * P[kOffWin + f] + (ld[I] - meanlog) * P[kOffWin + kHid + f] +
                P[kOffWin + 2 * kHid + f] + P[kOffBin + f];
        H[idx] = f2bf(fmaxf(z, 0.0f));
    }
    __syncthreads();
    for (int l = 0; l < kLayers; l++) {
        for (uint32_t k = tid; k < kBfPerLayer; k += kInferThreads) W[k] = g.wbf[l * kBfPerLayer + k];
        for (uint32_t idx = tid; idx < npad * kHid; idx += kInferThreads) {
            uint32_t I = idx / kHid, f = idx % kHid;
            float m = 0.0f;
            if (I < nc)
                for (uint32_t p = g.cptr[I]; p < g.cptr[I + 1]; p++) m += g.cw[p] * bf2f(H[g.ccol[p] * kHid + f]);
            M[idx] = f2bf(m);
        }
        __syncthreads();
        const float *b = P + off_b(l);
        for (uint32_t strip = wave; strip < npad / 16; strip += kInferThreads / kWave) {
            const uint32_t r0 = strip * 16;
            f32x4 acc[4];
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                acc[ct] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
                acc[ct] = mfma_tile([&](int r, int k) { return H[(r0 + r) * kHid + k]; },
                                    [&](int k, int c) { return W[k * kHid + ct * 16 + c]; }, kHid, acc[ct]);
                acc[ct] = mfma_tile([&](int r, int k) { return M[(r0 + r) * kHid + k]; },
                                    [&](int k, int c) { return W[kHid * kHid + k * kHid + ct * 16 + c]; }, kHid, acc[ct]);
            }
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    uint32_t row = r0 + (lane >> 4) * 4 + i, f = ct * 16 + (lane & 15);
                    float y = acc[ct][i] + b[f];
                    float h = bf2f(H[row * kHid + f]) + fmaxf(y, 0.0f);
                    H[row * kHid + f] = row < nc ? f2bf(h) : (unsigned short)0;
                }
        }
        __syncthreads();
    }
    for (uint32_t I = tid; I < nc; I += kInferThreads) {
        float o = P[kOffBo];
        for (int f = 0; f < kHid; f++) o += bf2f(H[I * kHid + f]) * P[kOffWo + f];
        g.e[I] = u[I] + sigma * o;
    }
}

// ---------------------------------------------------------------------------- training
struct TrainArgs {
    uint32_t n;               // coarse nodes of every sample (<= 64); rows n..63 are padding
    const float *Ac;          // [B][64][64] dense coarse operators (padding zero)
    const float *rc;          // [B][64]
    const float *s;           // [B] e_k . r_k
    const float *params;
    const unsigned short *wbf;
    float *grads;             // [B][kNumParams]
    float *loss;              // [B]
};

__global__ __launch_bounds__(kBlock) void k_gnn_train(TrainArgs t)
{
    constexpr uint32_t N = kTrainN;
    __shared__ unsigned short Hs[(kLayers + 1) * N * kHid];
    __shared__ float Ac[N * N];
    __shared__ float dH[N * kHid];
    __shared__ float dM[N * kHid];
    __shared__ unsigned short Mb[N * kHid];
    __shared__ unsigned short dY[N * kHid];
    __shared__ unsigned short W[kBfPerLayer];
    __shared__ float u[N], ld[N], ev[N], gov[N];
    __shared__ float red[kBlock / kWave];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid / 64;
    const uint32_t n = t.n, b = blockIdx.x;
    const float *P = t.params;
    const float *Ag = t.Ac + (uint64_t)b * N * N;
    const float *rcg = t.rc + (uint64_t)b * N;
    float *G = t.grads + (uint64_t)b * kNumParams;
    for (uint32_t k = tid; k < N * N; k += kBlock) Ac[k] = Ag[k];
    __syncthreads();
    // inputs
    float su = 0.0f, sl = 0.0f;
    for (uint32_t I = tid; I < N; I += kBlock) {
        float uu = 0.0f, l = 0.0f;
        if (I < n) {
            float d = Ac[I * N + I];
            uu = rcg[I] / d;
            l = logf(d);
        }
        u[I] = uu;
        ld[I] = l;
        su += uu * uu;
        sl += l;
    }
    const float sigma = sqrtf(block_sum_f(su, red, kBlock) / (float)n) + 1e-30f;
    const float meanlog = block_sum_f(sl, red, kBlock) / (float)n;
    auto feat = [&](uint32_t I, int k) -> float {
        if (I >= n) return 0.0f;
        return k == 0 ? u[I] / sigma : k == 1 ? ld[I] - meanlog : k == 2 ? 1.0f : 0.0f;
    };
    auto z0 = [&](uint32_t I, uint32_t f) -> float {
        if (I >= n) return 0.0f;
        return feat(I, 0) * P[kOffWin + f] + feat(I, 1) * P[kOffWin + kHid + f] + P[kOffWin + 2 * kHid + f] +
               P[kOffBin + f];
    };
    auto wn = [&](uint32_t I, uint32_t J) -> float {   // -A[I][J]/A[I][I], 0 on the diagonal
        return (I == J || I >= n || J >= n) ? 0.0f : -Ac[I * N + J] / Ac[I * N + I];
    };
    for (uint32_t idx = tid; idx < N * kHid; idx += kBlock)
        Hs[idx] = f2bf(fmaxf(z0(idx / kHid, idx % kHid), 0.0f));
    __syncthreads();
    // forward; Y is recomputed in the backward pass from the stored H
    auto layer_y = [&](int l, const unsigned short *Hl, f32x4 (&acc)[4]) {
        const uint32_t r0 = wave * 16;
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            acc[ct] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
            acc[ct] = mfma_tile([&](int r, int k) { return Hl[(r0 + r) * kHid + k]; },
                                [&](int k, int c) { return W[k * kHid + ct * 16 + c]; }, kHid, acc[ct]);
            acc[ct] = mfma_tile([&](int r, int k) { return Mb[(r0 + r) * kHid + k]; },
                                [&](int k, int c) { return W[kHid * kHid + k * kHid + ct * 16 + c]; }, kHid, acc[ct]);
        }
    };
    auto aggregate = [&](const unsigned short *Hl) {
        for (uint32_t idx = tid; idx < N * kHid; idx += kBlock) {
            uint32_t I = idx / kHid, f = idx % kHid;
            float m = 0.0f;
            for (uint32_t J = 0; J < n; J++) m += wn(I, J) * bf2f(Hl[J * kHid + f]);
            Mb[idx] = f2bf(m);
        }
    };
    for (int l = 0; l < kLayers; l++) {
        const unsigned short *Hl = Hs + l * N * kHid;
        unsigned short *Hn = Hs + (l + 1) * N * kHid;
        for (uint32_t k = tid; k < kBfPerLayer; k += kBlock) W[k] = t.wbf[l * kBfPerLayer + k];
        aggregate(Hl);
        __syncthreads();
        f32x4 acc[4];
        layer_y(l, Hl, acc);
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint32_t row = wave * 16 + (lane >> 4) * 4 + i, f = ct * 16 + (lane & 15);
                float y = acc[ct][i] + P[off_b(l) + f];
                float h = bf2f(Hl[row * kHid + f]) + fmaxf(y, 0.0f);
                Hn[row * kHid + f] = row < n ? f2bf(h) : (unsigned short)0;
            }
        __syncthreads();
    }
    const unsigned short *H6 = Hs + kLayers * N * kHid;
    for (uint32_t I = tid; I < N; I += kBlock) {
        float o = P[kOffBo];
        for (int f = 0; f < kHid; f++) o += bf2f(H6[I * kHid + f]) * P[kOffWo + f];
        ev[I] = I < n ? u[I] + sigma * o : 0.0f;
    }
    __syncthreads();
    // loss and output-head gradient
    float part = 0.0f;
    const float s = t.s[b];
    for (uint32_t I = tid; I < N; I += kBlock) {
        float ae = 0.0f;
        for (uint32_t J = 0; J < n; J++) ae += Ac[I * N + J] * ev[J];
        float ge = I < n ? 2.0f * (ae - rcg[I]) / s : 0.0f;
        gov[I] = ge * sigma;
        part += I < n ? ev[I] * ae - 2.0f * ev[I] * rcg[I] : 0.0f;
    }
    float L = block_sum_f(part, red, kBlock) / s + 1.0f;
    if (tid == 0) t.loss[b] = L;
    __syncthreads();
    for (uint32_t f = tid; f < (uint32_t)kHid; f += kBlock) {
        float gw = 0.0f;
        for (uint32_t I = 0; I < n; I++) gw += gov[I] * bf2f(H6[I * kHid + f]);
        G[kOffWo + f] = gw;
    }
    if (tid == 0) {
        float gb = 0.0f;
        for (uint32_t I = 0; I < n; I++) gb += gov[I];
        G[kOffBo] = gb;
    }
    for (uint32_t idx = tid; idx < N * kHid; idx += kBlock) dH[idx] = gov[idx / kHid] * P[kOffWo + idx % kHid];
    __syncthreads();
    // backward through the layers
    for (int l = kLayers - 1; l >= 0; l--) {
        const unsigned short *Hl = Hs + l * N * kHid;
        for (uint32_t k = tid; k < kBfPerLayer; k += kBlock) W[k] = t.wbf[l * kBfPerLayer + k];
        aggregate(Hl);
        __syncthreads();
        f32x4 acc[4];
        layer_y(l, Hl, acc);
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint32_t row = wave * 16 + (lane >> 4) * 4 + i, f = ct * 16 + (lane & 15);
                float y = acc[ct][i] + P[off_b(l) + f];
                float g = (y > 0.0f && row < n) ? dH[row * kHid + f] : 0.0f;
                dY[row * kHid + f] = f2bf(g);
            }
        __syncthreads();
        // dWs = H^T dY, dWn = M^T dY (rows of the output: input features, strip per wave)
        {
            const uint32_t r0 = wave * 16;
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                f32x4 gs = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f }, gn = gs;
                gs = mfma_tile([&](int r, int k) { return Hl[k * kHid + r0 + r]; },
                               [&](int k, int c) { return dY[k * kHid + ct * 16 + c]; }, N, gs);
                gn = mfma_tile([&](int r, int k) { return Mb[k * kHid + r0 + r]; },
                               [&](int k, int c) { return dY[k * kHid + ct * 16 + c]; }, N, gn);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    uint32_t fi = r0 + (lane >> 4) * 4 + i, fo = ct * 16 + (lane & 15);
                    G[off_ws(l) + fi * kHid + fo] = gs[i];
                    G[off_wn(l) + fi * kHid + fo] = gn[i];
                }
            }
        }
        for (uint32_t f = tid; f < (uint32_t)kHid; f += kBlock) {
            float gb = 0.0f;
            for (uint32_t I = 0; I < N; I++) gb += bf2f(dY[I * kHid + f]);
            G[off_b(l) + f] = gb;
        }
        // dM = dY Wn^T ; dH += dY Ws^T   (node strip per wave)
        {
            const uint32_t r0 = wave * 16;
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
        // dH[I] += sum_J wn(J, I) dM[J]
        for (uint32_t idx = tid; idx < N * kHid; idx += kBlock) {
            uint32_t I = idx / kHid, f = idx % kHid;
            float a = 0.0f;
            for (uint32_t J = 0; J < n; J++) a += wn(J, I) * dM[J * kHid + f];
            dH[idx] += a;
        }
        __syncthreads();
    }
    // embedding
    for (uint32_t f = tid; f < (uint32_t)kHid; f += kBlock) {
        float gw[kNin] = { 0.0f, 0.0f, 0.0f, 0.0f }, gb = 0.0f;
        for (uint32_t I = 0; I < n; I++) {
            float dz = z0(I, f) > 0.0f ? dH[I * kHid + f] : 0.0f;
            for (int k = 0; k < kNin; k++) gw[k] += feat(I, k) * dz;
            gb += dz;
        }
        for (int k = 0; k < kNin; k++) G[kOffWin + k * kHid + f] = gw[k];
        G[kOffBin + f] = gb;
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

// Adam (Kingma & Ba 2015) on the FP32 master weights; refreshes the BF16 layer weights.
__global__ void k_adam(float *__restrict__ theta, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2,
                       unsigned short *__restrict__ wbf)
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

// BF16 copies of the layer matrices: wbf[l][0:4096] = Ws[l], wbf[l][4096:8192] = Wn[l]
__global__ void k_weights_bf16(const float *__restrict__ theta, unsigned short *__restrict__ wbf)
{
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < kLayers * kBfPerLayer; k += gridDim.x * kBlock) {
        uint32_t l = k / kBfPerLayer, o = k % kBfPerLayer;
        wbf[k] = f2bf(theta[off_ws(l) + o]);
    }
}

// r = b - y (y = A x), e = x* - x ; partial e.r  (training pairs)
__global__ __launch_bounds__(kBlock) void k_pair(const double *__restrict__ b, const double *__restrict__ y,
                                                 const double *__restrict__ x, const double *__restrict__ xs,
                                                 double *__restrict__ r, uint32_t n, double *partials)
{
    double v[1] = { 0.0 };
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        double ri = b[i] - y[i];
        r[i] = ri;
        v[0] += (xs[i] - x[i]) * ri;
    }
    store_partials<1>(v, partials);
}

// ---------------------------------------------------------------------------- flexible CG
struct FcgArgs {
    uint32_t n;
    float *r, *z, *p, *q, *dx;
    const float *d32;            // Jacobi inverse diagonal (FP32)
    const uint32_t *agg;
    const float *ec;             // network output per aggregate
    double *partials;
    SolveState *st;
};

// z = D^-1 r + P e ; partial r.z and z.q_prev
__global__ __launch_bounds__(kBlock) void k_fcg_z(FcgArgs a)
{
    if (a.st->done) return;
    double v[2] = { 0.0, 0.0 };
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * kBlock) {
        float ri = a.r[i];
        float zi = a.d32[i] * ri + a.ec[a.agg[i]];
        a.z[i] = zi;
        v[0] += (double)ri * (double)zi;
        v[1] += (double)zi * (double)a.q[i];
    }
    store_partials<2>(v, a.partials);
}

// p = z (first iteration) or z + beta p
__global__ void k_fcg_p(FcgArgs a)
{
    if (a.st->done) return;
    const bool first = a.st->it == 0;
    const float beta = (float)a.st->beta;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * kBlock)
        a.p[i] = first ? a.z[i] : fmaf(beta, a.p[i], a.z[i]);
}

// dx += alpha p, r -= alpha q, partial |r|^2
__global__ __launch_bounds__(kBlock) void k_fcg_update(FcgArgs a)
{
    if (a.st->done) return;
    const float alpha = (float)a.st->alpha;
    double v[1] = { 0.0 };
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * kBlock) {
        a.dx[i] = fmaf(alpha, a.p[i], a.dx[i]);
        float ri = fmaf(-alpha, a.q[i], a.r[i]);
        a.r[i] = ri;
        v[0] += (double)ri * (double)ri;
    }
    store_partials<1>(v, a.partials);
}

}  // namespace rs

// ============================================================================ C ABI: learned
struct rs_learned {
    uint32_t n = 0, nc = 0;
    std::vector<uint32_t> agg_h;
    DBuf<uint32_t> agg, mptr, mrows, cptr, ccol;
    DBuf<double> cval, dense, maxabs, mpart;
    DBuf<float> cw, cdiag, rc, ec;
    DBuf<unsigned long long> acc;
    DBuf<float> theta;
    DBuf<unsigned short> wbf;
    DBuf<double> tmp_y, tmp_r, tmp_part;
    bool have_params = false;
    uint64_t setup_version = ~0ull;
    const rs_matrix *setup_matrix = nullptr;
    double setup_ms = 0.0;
};

namespace rs {

static int learned_setup(rs_learned &L, rs_matrix &A)
{
    if (L.setup_matrix == &A && L.setup_version == A.version) return RS_OK;
    if (A.n != L.n) return fail(RS_ERR_INVALID_ARG, "learned preconditioner: matrix size differs from its aggregation");
    hipStream_t st = runtime().stream;
    int rc;
    uint32_t nb = vec_blocks(std::max<uint64_t>(A.nnz, 1));
    if ((rc = L.mpart.alloc(nb))) return rc;
    k_maxabs<<<nb, kBlock, 0, st>>>(A.val.p, A.nnz, L.mpart.p);
    k_maxabs_fin<<<1, 64, 0, st>>>(L.mpart.p, nb, L.maxabs.p);
    uint64_t nn = (uint64_t)L.nc * L.nc;
    k_zero_u64<<<vec_blocks(nn), kBlock, 0, st>>>(L.acc.p, nn);
    k_galerkin<<<vec_blocks(A.n), kBlock, 0, st>>>(A.rowptr.p, A.col.p, A.val.p, A.n, L.agg.p, L.nc, L.maxabs.p, L.acc.p);
    k_galerkin_finish<<<vec_blocks(nn), kBlock, 0, st>>>(L.acc.p, L.nc, L.maxabs.p, L.cptr.p, L.ccol.p, L.cval.p, L.cw.p,
                                                          L.cdiag.p, L.dense.p);
    RS_HIP(hipGetLastError());
    L.setup_matrix = &A;
    L.setup_version = A.version;
    return RS_OK;
}

static GnnArgs gnn_args(rs_learned &L, const SolveState *st)
{
    GnnArgs g{};
    g.nc = L.nc;
    g.npad = (L.nc + 15) / 16 * 16;
    g.rc = L.rc.p;
    g.diag = L.cdiag.p;
    g.cptr = L.cptr.p;
    g.ccol = L.ccol.p;
    g.cw = L.cw.p;
    g.params = L.theta.p;
    g.wbf = L.wbf.p;
    g.e = L.ec.p;
    g.st = st;
    return g;
}

// Flexible CG with the learned preconditioner (mixed precision, FP64 residual replacement).
static int solve_fcg_learned(rs_solver &S, rs_matrix &A, rs_precond &P, rs_learned &L, double *d_x, uint32_t &done,
                             uint32_t &refinements)
{
    int rc;
    if ((rc = matrix_refresh_fp32(A)) || (rc = precond_fp32(P)) || (rc = matrix_tune(A, true)) ||
        (rc = matrix_tune(A, false)))
        return rc;
    const int plan32 = A.plan[1], plan64 = A.plan[0];
    if ((rc = matrix_sync_plan(A, plan32, true)) || (rc = matrix_sync_plan(A, plan64, false))) return rc;
    if ((rc = learned_setup(L, A))) return rc;
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
    a.st = S.st.p;
    a.hist = S.hist.p;
    MpArgs a64 = a;
    a64.A = plan_view(A, plan64, nb64);
    FcgArgs f{};
    f.n = A.n;
    f.r = S.r32.p;
    f.z = S.z32.p;
    f.p = S.p32.p;
    f.q = S.q32.p;
    f.dx = S.dx32.p;
    f.d32 = P.data32.p;
    f.agg = L.agg.p;
    f.ec = L.ec.p;
    f.partials = S.partials.p;
    f.st = S.st.p;
    GnnArgs g = gnn_args(L, S.st.p);
    const uint32_t nvb = vec_blocks(A.n);
    hipStream_t stream = runtime().stream;
    k_fill32<<<nvb, kBlock, 0, stream>>>(S.dx32.p, A.n, 0.0f);
    k_fill32<<<nvb, kBlock, 0, stream>>>(S.q32.p, A.n, 0.0f);
    RS_SHAPE_SWITCH(plan64, k_mp_true_resid, nb64, stream, a64);
    fin(S.partials.p, nb64, 2, FIN_MP_TRUE, 1, S.st.p, S.hist.p);
    RS_HIP(hipGetLastError());
    rc = capture(S, [&] {
        for (uint32_t i = 0; i < kGraphIters; i++) {
            k_restrict<float><<<L.nc, kBlock, 0, stream>>>(S.r32.p, L.mptr.p, L.mrows.p, L.rc.p, S.st.p);
            k_gnn_infer<<<1, kInferThreads, 0, stream>>>(g);
            k_fcg_z<<<nvb, kBlock, 0, stream>>>(f);
            fin(S.partials.p, nvb, 2, FIN_FCG_BETA, 0, S.st.p, S.hist.p);
            k_fcg_p<<<nvb, kBlock, 0, stream>>>(f);
            RS_SHAPE_SWITCH(plan32, k_mp_spmv_dot, nb32, stream, a);
            fin(S.partials.p, nb32, 1, FIN_DOT, 0, S.st.p, S.hist.p);
            k_fcg_update<<<nvb, kBlock, 0, stream>>>(f);
            fin(S.partials.p, nvb, 1, FIN_FCG_UPDATE, 0, S.st.p, S.hist.p);
        }
    }, graph_key_of(3, a, plan32, nb32, L.nc, nvb) + graph_key_of(3, f, plan64, nb64, 0, 0));
    if (rc) return rc;
    refinements = 0;
    SolveState h;
    RS_HIP(hipMemcpyAsync(&h, S.st.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    RS_HIP(hipStreamSynchronize(stream));
    done = h.done;
    double best = h.t2;
    uint32_t no_gain = 0;
    while (done == kStopRunning) {
        if ((rc = run_graph_loop(S, done))) return rc;
        if (done != kStopPause) break;
        refinements++;
        k_mp_fold<<<nvb, kBlock, 0, stream>>>(d_x, S.dx32.p, A.n);
        if (no_gain >= kMaxReplacementsWithoutGain) {
            RS_HIP(hipGetLastError());
            h.reason = RS_REASON_STAGNATED;
            h.done = kStopDone;
            RS_HIP(hipMemcpyAsync(&S.st.p->done, &h.done, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
            RS_HIP(hipStreamSynchronize(stream));
            done = kStopDone;
            return RS_OK;
        }
        RS_SHAPE_SWITCH(plan64, k_mp_true_resid, nb64, stream, a64);
        fin(S.partials.p, nb64, 2, FIN_MP_TRUE, 0, S.st.p, S.hist.p);
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
    // a segment that ended on a non-finite residual leaves x as of its last replacement
    RS_HIP(hipMemcpyAsync(&h, S.st.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    RS_HIP(hipStreamSynchronize(stream));
    if (h.reason != RS_REASON_NONNORMAL) k_mp_fold<<<nvb, kBlock, 0, stream>>>(d_x, S.dx32.p, A.n);
    RS_HIP(hipGetLastError());
    return RS_OK;
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

extern "C" int rs_learned_create(const rs_matrix *A, const uint32_t *agg_of_row, uint32_t n_agg, rs_learned **out)
{
    if (!A || !agg_of_row || !out || n_agg == 0) return fail(RS_ERR_INVALID_ARG, "rs_learned_create: bad arguments");
    if (n_agg > kMaxCoarse) return fail(RS_ERR_UNSUPPORTED, "rs_learned_create: more than 512 aggregates");
    for (uint32_t i = 0; i < A->n; i++)
        if (agg_of_row[i] >= n_agg) return fail(RS_ERR_INVALID_ARG, "rs_learned_create: aggregate id out of range");
    int rc = require_runtime();
    if (rc) return rc;
    std::unique_ptr<rs_learned> L(new rs_learned());
    L->n = A->n;
    L->nc = n_agg;
    L->agg_h.assign(agg_of_row, agg_of_row + A->n);
    // member lists (rows of each aggregate in increasing order)
    std::vector<uint32_t> mptr(n_agg + 1, 0), mrows(A->n);
    for (uint32_t i = 0; i < A->n; i++) mptr[agg_of_row[i] + 1]++;
    for (uint32_t I = 0; I < n_agg; I++) mptr[I + 1] += mptr[I];
    {
        std::vector<uint32_t> fill(mptr.begin(), mptr.end() - 1);
        for (uint32_t i = 0; i < A->n; i++) mrows[fill[agg_of_row[i]]++] = i;
    }
    for (uint32_t I = 0; I < n_agg; I++)
        if (mptr[I + 1] == mptr[I]) return fail(RS_ERR_INVALID_ARG, "rs_learned_create: empty aggregate");
    // coarse pattern from the fine pattern
    std::vector<uint32_t> rowptr(A->n + 1ull), col(A->nnz);
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(rowptr.data(), A->rowptr.p, sizeof(uint32_t) * rowptr.size(), hipMemcpyDeviceToHost, st));
    if (A->nnz) RS_HIP(hipMemcpyAsync(col.data(), A->col.p, sizeof(uint32_t) * A->nnz, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    std::vector<std::vector<uint8_t>> adj(n_agg, std::vector<uint8_t>(n_agg, 0));
    for (uint32_t i = 0; i < A->n; i++)
        for (uint32_t j = rowptr[i]; j < rowptr[i + 1]; j++) adj[agg_of_row[i]][agg_of_row[col[j]]] = 1;
    std::vector<uint32_t> cptr(n_agg + 1, 0), ccol;
    for (uint32_t I = 0; I < n_agg; I++) {
        adj[I][I] = 1;
        for (uint32_t J = 0; J < n_agg; J++)
            if (adj[I][J]) ccol.push_back(J);
        cptr[I + 1] = (uint32_t)ccol.size();
    }
    rs_learned &R = *L;
    uint64_t nn = (uint64_t)n_agg * n_agg;
    if ((rc = R.agg.alloc(std::max(A->n, 1u))) || (rc = R.mptr.alloc(mptr.size())) || (rc = R.mrows.alloc(std::max(A->n, 1u))) ||
        (rc = R.cptr.alloc(cptr.size())) || (rc = R.ccol.alloc(ccol.size())) || (rc = R.cval.alloc(ccol.size())) ||
        (rc = R.cw.alloc(ccol.size())) || (rc = R.cdiag.alloc(n_agg)) || (rc = R.dense.alloc(nn)) ||
        (rc = R.maxabs.alloc(1)) || (rc = R.acc.alloc(nn)) || (rc = R.rc.alloc(n_agg)) || (rc = R.ec.alloc(n_agg)) ||
        (rc = R.theta.alloc(kNumParams)) || (rc = R.wbf.alloc(kLayers * kBfPerLayer)))
        return rc;
    RS_HIP(hipMemcpyAsync(R.agg.p, agg_of_row, sizeof(uint32_t) * A->n, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(R.mptr.p, mptr.data(), sizeof(uint32_t) * mptr.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(R.mrows.p, mrows.data(), sizeof(uint32_t) * mrows.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(R.cptr.p, cptr.data(), sizeof(uint32_t) * cptr.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(R.ccol.p, ccol.data(), sizeof(uint32_t) * ccol.size(), hipMemcpyHostToDevice, st));
    RS_HIP(hipStreamSynchronize(st));
    *out = L.release();
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
    k_weights_bf16<<<vec_blocks(kLayers * kBfPerLayer), kBlock, 0, st>>>(L->theta.p, L->wbf.p);
    RS_HIP(hipGetLastError());
    RS_HIP(hipStreamSynchronize(st));
    L->have_params = true;
    return RS_OK;
}

extern "C" int rs_learned_info(const rs_learned *L, uint32_t *n_coarse, uint64_t *n_params, uint64_t *coarse_nnz)
{
    if (!L) return fail(RS_ERR_INVALID_ARG, "rs_learned_info: null");
    if (n_coarse) *n_coarse = L->nc;
    if (n_params) *n_params = kNumParams;
    if (coarse_nnz) *coarse_nnz = L->ccol.n;
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

extern "C" int rs_learned_coarse_matrix(const rs_learned *L, double *dense)
{
    if (!L || !dense) return fail(RS_ERR_INVALID_ARG, "rs_learned_coarse_matrix: null");
    int rc = require_runtime();
    if (rc) return rc;
    RS_HIP(hipMemcpyAsync(dense, L->dense.p, sizeof(double) * L->nc * L->nc, hipMemcpyDeviceToHost, runtime().stream));
    RS_HIP(hipStreamSynchronize(runtime().stream));
    return RS_OK;
}

extern "C" int rs_learned_restrict_device(rs_learned *L, const double *d_r, float *rc_host)
{
    if (!L || !d_r || !rc_host) return fail(RS_ERR_INVALID_ARG, "rs_learned_restrict_device: null");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    k_restrict<double><<<L->nc, kBlock, 0, st>>>(d_r, L->mptr.p, L->mrows.p, L->rc.p, nullptr);
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(rc_host, L->rc.p, sizeof(float) * L->nc, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

// Network output for a host coarse residual (uses the coarse operator of the last setup).
extern "C" int rs_learned_apply_coarse(rs_learned *L, const float *rc_host, float *e_host)
{
    if (!L || !rc_host || !e_host) return fail(RS_ERR_INVALID_ARG, "rs_learned_apply_coarse: null");
    if (!L->have_params) return fail(RS_ERR_INVALID_ARG, "rs_learned_apply_coarse: parameters not set");
    if (!L->setup_matrix) return fail(RS_ERR_INVALID_ARG, "rs_learned_apply_coarse: call rs_learned_setup first");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(L->rc.p, rc_host, sizeof(float) * L->nc, hipMemcpyHostToDevice, st));
    GnnArgs g = gnn_args(*L, nullptr);
    k_gnn_infer<<<1, kInferThreads, 0, st>>>(g);
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(e_host, L->ec.p, sizeof(float) * L->nc, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

// A (residual, error) pair of a CG iterate x for the system A x* = b: r = b - A x (FP64),
// r_c = P^T r and s = (x* - x) . r = ||x* - x||_A^2.  All vectors are device arrays.
extern "C" int rs_learned_pair(rs_learned *L, rs_matrix *A, const double *d_b, const double *d_x,
                               const double *d_xstar, float *rc_host, double *s_out)
{
    if (!L || !A || !d_b || !d_x || !d_xstar || !rc_host || !s_out) return fail(RS_ERR_INVALID_ARG, "rs_learned_pair: null");
    if (A->n != L->n) return fail(RS_ERR_INVALID_ARG, "rs_learned_pair: size mismatch");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    uint32_t nb = vec_blocks(std::max(L->n, 1u));
    if ((rc = L->tmp_y.alloc(std::max(L->n, 1u))) || (rc = L->tmp_r.alloc(std::max(L->n, 1u))) ||
        (rc = L->tmp_part.alloc(nb)))
        return rc;
    if ((rc = rs_matrix_spmv_device(A, d_x, L->tmp_y.p))) return rc;
    k_pair<<<nb, kBlock, 0, st>>>(d_b, L->tmp_y.p, d_x, d_xstar, L->tmp_r.p, L->n, L->tmp_part.p);
    k_restrict<double><<<L->nc, kBlock, 0, st>>>(L->tmp_r.p, L->mptr.p, L->mrows.p, L->rc.p, nullptr);
    RS_HIP(hipGetLastError());
    std::vector<double> parts(nb);
    RS_HIP(hipMemcpyAsync(parts.data(), L->tmp_part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    RS_HIP(hipMemcpyAsync(rc_host, L->rc.p, sizeof(float) * L->nc, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    double sum = 0.0;
    for (double v : parts) sum += v;
    *s_out = sum;
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
    if (!L->have_params) return fail(RS_ERR_INVALID_ARG, "learned preconditioner: parameters not set");
    if (S->conf.solver_type != RS_SOLVER_CG) return fail(RS_ERR_UNSUPPORTED, "learned preconditioner: flexible CG only");
    int rc = require_runtime();
    if (rc) return rc;
    return solve_device_ext(*S, *A, d_b, d_x, RS_PRECOND_LEARNED, 1, L, info);
}

// ---------------------------------------------------------------------------- trainer
struct rs_trainer {
    uint32_t max_batch = 0;
    uint64_t step = 0;
    DBuf<float> theta, m, v, g, grads, loss, Ac, rc, s;
    DBuf<unsigned short> wbf;
};

extern "C" int rs_trainer_create(const float *params, uint64_t count, uint32_t max_batch, rs_trainer **out)
{
    if (!params || !out || count != kNumParams || max_batch == 0)
        return fail(RS_ERR_INVALID_ARG, "rs_trainer_create: bad arguments (expects 49921 parameters)");
    int rc = require_runtime();
    if (rc) return rc;
    std::unique_ptr<rs_trainer> T(new rs_trainer());
    T->max_batch = max_batch;
    if ((rc = T->theta.alloc(kNumParams)) || (rc = T->m.alloc(kNumParams)) || (rc = T->v.alloc(kNumParams)) ||
        (rc = T->g.alloc(kNumParams)) || (rc = T->grads.alloc((uint64_t)max_batch * kNumParams)) ||
        (rc = T->loss.alloc(max_batch)) || (rc = T->Ac.alloc((uint64_t)max_batch * kTrainN * kTrainN)) ||
        (rc = T->rc.alloc((uint64_t)max_batch * kTrainN)) || (rc = T->s.alloc(max_batch)) ||
        (rc = T->wbf.alloc(kLayers * kBfPerLayer)))
        return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(T->theta.p, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemsetAsync(T->m.p, 0, sizeof(float) * kNumParams, st));
    RS_HIP(hipMemsetAsync(T->v.p, 0, sizeof(float) * kNumParams, st));
    k_weights_bf16<<<vec_blocks(kLayers * kBfPerLayer), kBlock, 0, st>>>(T->theta.p, T->wbf.p);
    RS_HIP(hipGetLastError());
    RS_HIP(hipStreamSynchronize(st));
    *out = T.release();
    return RS_OK;
}

// Per-sample losses and the batch-mean gradient (no update).  Ac: [B][64][64] (padding 0),
// rc: [B][64], s: [B]; every sample has n <= 64 coarse nodes.
extern "C" int rs_trainer_gradient(rs_trainer *T, uint32_t n, uint32_t B, const float *Ac, const float *rcv,
                                   const float *s, float *loss_out, float *grad_out)
{
    if (!T || !Ac || !rcv || !s || n == 0 || n > kTrainN || B == 0 || B > T->max_batch)
        return fail(RS_ERR_INVALID_ARG, "rs_trainer_gradient: bad arguments");
    int rc = require_runtime();
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    RS_HIP(hipMemcpyAsync(T->Ac.p, Ac, sizeof(float) * B * kTrainN * kTrainN, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(T->rc.p, rcv, sizeof(float) * B * kTrainN, hipMemcpyHostToDevice, st));
    RS_HIP(hipMemcpyAsync(T->s.p, s, sizeof(float) * B, hipMemcpyHostToDevice, st));
    TrainArgs t{ n, T->Ac.p, T->rc.p, T->s.p, T->theta.p, T->wbf.p, T->grads.p, T->loss.p };
    k_gnn_train<<<B, kBlock, 0, st>>>(t);
    k_grad_mean<<<vec_blocks(kNumParams), kBlock, 0, st>>>(T->grads.p, B, T->g.p);
    RS_HIP(hipGetLastError());
    if (loss_out) RS_HIP(hipMemcpyAsync(loss_out, T->loss.p, sizeof(float) * B, hipMemcpyDeviceToHost, st));
    if (grad_out) RS_HIP(hipMemcpyAsync(grad_out, T->g.p, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

// One Adam step on a batch; returns the mean loss.
extern "C" int rs_trainer_step(rs_trainer *T, uint32_t n, uint32_t B, const float *Ac, const float *rcv, const float *s,
                               float lr, float *mean_loss)
{
    std::vector<float> losses(B);
    int rc = rs_trainer_gradient(T, n, B, Ac, rcv, s, losses.data(), nullptr);
    if (rc) return rc;
    T->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f - std::pow(b1, (float)T->step), c2 = 1.0f - std::pow(b2, (float)T->step);
    hipStream_t st = runtime().stream;
    k_adam<<<vec_blocks(kNumParams), kBlock, 0, st>>>(T->theta.p, T->m.p, T->v.p, T->g.p, lr, b1, b2, 1e-8f, c1, c2, T->wbf.p);
    k_weights_bf16<<<vec_blocks(kLayers * kBfPerLayer), kBlock, 0, st>>>(T->theta.p, T->wbf.p);
    RS_HIP(hipGetLastError());
    RS_HIP(hipStreamSynchronize(st));
    if (mean_loss) {
        double m = 0.0;
        for (float l : losses) m += l;
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
