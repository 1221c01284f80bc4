// rs_eigen.hip -- REigenValueSolver on the GPU: the generalized problem K x = lambda M x by
// inverse iteration on K^-1 M (Lanczos, Arnoldi) or Rayleigh-quotient iteration, every inner
// solve with the library's CG on the device.
//
// Reference behaviour restated (Range-Software/range-solver-lib):
//   REigenValueSolver::solve (post-processing d = |1/d|, swap, sort, row reorder)
//                                                    src/reigenvaluesolver.cpp:41-117
//   solveLanczos                                     :119-205
//   solveArnoldi                                     :207-287
//   solveRayleigh                                    :289-418
//   qlDecomposition                                  :420-496
//   qrDecomposition (Gram-Schmidt QR iteration)      :498-592
// Kept as the reference does them: the start vector from std::rand() / RAND_MAX (the process's
// C library stream, so a caller that seeds it with srand gets the reference's vector), every
// inner solve warm-started from the vector's previous contents, the quotient by the norm as
// a division, and qlDecomposition's early exit: its split search starts at m = l, so the
// "m >= l" test after it always holds and the tridiagonal's diagonal is returned unrotated
// (:440-452).  The eigenvector rows are reordered with the reference's swap loop (:98-110).
//
// Layout: vectors are FP64 device arrays; the Arnoldi basis Qa is n x ne column-major (column j
// contiguous); the small dense matrices (H, the QR factors) live on the host.  Dot products are
// deterministic two-step sums.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "rs_device.hpp"
#include "rs_internal.hpp"

namespace rs {
namespace {

__global__ __launch_bounds__(kBlock) void k_ev_dot(const double *__restrict__ a, const double *__restrict__ b,
                                                   uint64_t n, double *__restrict__ partials)
{
    __shared__ double lds[kBlock / kWave];
    double v[1] = { 0.0 };
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        v[0] += a[i] * b[i];
    block_sum<1, kBlock>(v, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(kBlock) void k_ev_sum(const double *__restrict__ partials, uint32_t nb,
                                                   double *__restrict__ out)
{
    __shared__ double lds[kBlock / kWave];
    double v[1] = { 0.0 };
    for (uint32_t i = threadIdx.x; i < nb; i += kBlock) v[0] += partials[i];
    block_sum<1, kBlock>(v, lds);
    if (threadIdx.x == 0) *out = v[0];
}

// y -= a * x
__global__ void k_ev_axmy(double *__restrict__ y, const double *__restrict__ x, double a, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        y[i] -= a * x[i];
}

// y -= x * a (the product rounded first, as "bTmp *= e; b -= bTmp")
__global__ void k_ev_sub_scaled(double *__restrict__ y, const double *__restrict__ x, double a, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const double t = x[i] * a;
        y[i] -= t;
    }
}

// y = x / a
__global__ void k_ev_div(double *__restrict__ y, const double *__restrict__ x, double a, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        y[i] = x[i] / a;
}

// y *= a
__global__ void k_ev_scale(double *__restrict__ y, double a, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) y[i] *= a;
}

// Rayleigh shift, one pass over K's entries (same pattern as M2): v = K_ij * mu; M2_ij += v when
// v != 0 (addValue), and with add: f_i = sum_j v * b_j.
__global__ void k_ev_shift(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ col,
                           const double *__restrict__ kval, double mu, double *__restrict__ m2val,
                           const double *__restrict__ b, double *__restrict__ f, uint32_t n, int add)
{
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double fi = 0.0;
        for (uint32_t j = rowptr[i]; j < rowptr[i + 1]; j++) {
            const double v = add ? kval[j] * mu : -1.0 * kval[j] * mu;
            if (v != 0.0) m2val[j] += v;
            if (add) fi += v * b[col[j]];
        }
        if (add) f[i] = fi;
    }
}

// f_i = sum_j M2_ij b_j in row order (the reference's explicit loop)
__global__ void k_ev_rowdot(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ col,
                            const double *__restrict__ val, const double *__restrict__ b, double *__restrict__ f,
                            uint32_t n)
{
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double fi = 0.0;
        for (uint32_t j = rowptr[i]; j < rowptr[i + 1]; j++) fi += val[j] * b[col[j]];
        f[i] = fi;
    }
}

// ev (ne rows of n) = (Qa Qq)^T: ev[i][k] = sum_j Qa[k][j] Qq[j][i]
__global__ void k_ev_vectors(const double *__restrict__ Qa, const double *__restrict__ Qq, uint32_t ne, uint64_t n,
                             double *__restrict__ ev)
{
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kBlock)
        for (uint32_t i = 0; i < ne; i++) {
            double s = 0.0;
            for (uint32_t j = 0; j < ne; j++) s += Qa[(uint64_t)j * n + k] * Qq[(uint64_t)j * ne + i];
            ev[(uint64_t)i * n + k] = s;
        }
}

uint32_t ev_blocks(uint64_t n)
{
    const uint64_t want = (n + kBlock - 1) / kBlock;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)std::max(8, runtime().n_cu * 8)));
}

struct EvWork {
    uint64_t n = 0;
    DBuf<double> partials, scalar;
    int init(uint64_t len)
    {
        n = len;
        int rc;
        if ((rc = partials.alloc(ev_blocks(len))) || (rc = scalar.alloc(1))) return rc;
        return RS_OK;
    }
    int dot(const double *a, const double *b, double &out)
    {
        hipStream_t st = runtime().stream;
        const uint32_t nb = ev_blocks(n);
        k_ev_dot<<<nb, kBlock, 0, st>>>(a, b, n, partials.p);
        k_ev_sum<<<1, kBlock, 0, st>>>(partials.p, nb, scalar.p);
        RS_HIP(hipGetLastError());
        RS_HIP(hipMemcpyAsync(&out, scalar.p, sizeof(double), hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
        return RS_OK;
    }
    int norm(const double *a, double &out)
    {
        int rc = dot(a, a, out);
        out = std::sqrt(out);
        return rc;
    }
};

// The start vector: std::rand() / RAND_MAX per entry, normalised (RRVector::normalize).
int random_unit(EvWork &w, double *d_v, std::vector<double> &h)
{
    for (auto &x : h) x = double(std::rand()) / double(RAND_MAX);
    RS_HIP(hipMemcpyAsync(d_v, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, runtime().stream));
    double nrm = 0.0;
    int rc = w.norm(d_v, nrm);
    if (rc) return rc;
    if (nrm == 0.0) return fail(RS_ERR_INVALID_ARG, "rs_eigen_solve: zero start vector");
    k_ev_div<<<ev_blocks(w.n), kBlock, 0, runtime().stream>>>(d_v, d_v, nrm, w.n);
    RS_HIP(hipGetLastError());
    return RS_OK;
}

// Every inner solve is Jacobi-CG as in the reference; with a multilevel preconditioner built
// for K (rs_eigen_solve_learned) the solves with K run flexible CG with its V-cycle instead.
struct InnerSolver {
    rs_solver *S = nullptr;
    rs_learned *L = nullptr;
    const rs_matrix *K = nullptr;
    uint64_t iterations = 0;
    ~InnerSolver()
    {
        if (S) rs_solver_destroy(S);
    }
    int create(const rs_conf *conf)
    {
        int rc = rs_solver_create(conf, &S);
        if (rc) return rc;
        return rs_solver_disable_convergence_log(S);
    }
    int solve(rs_matrix *A, const double *b, double *x)
    {
        rs_info info;
        int rc = (L && A == K) ? rs_solver_solve_learned_device(S, A, L, b, x, &info)
                               : rs_solver_solve_device(S, A, b, x, RS_PRECOND_JACOBI, 1, &info);
        iterations += info.iterations;
        return rc;
    }
};

// qlDecomposition as the reference runs it (src/reigenvaluesolver.cpp:420-496): e is shifted
// down by one; the split search starts at m = l, so its exit test "m >= l" holds at once for
// every l and no rotation is applied.
void ql_reference(std::vector<double> &d, std::vector<double> &e)
{
    const size_t n = d.size();
    for (size_t i = 1; i < n; i++) e[i - 1] = e[i];
    if (n) e[n - 1] = 0.0;
}

// qrDecomposition (src/reigenvaluesolver.cpp:498-592): Gram-Schmidt QR of A, A = R Q, V = V Q,
// until every subdiagonal entry is below the tolerance or n_iterations; d = diag(A).
int qr_reference(const std::vector<double> &H, uint32_t n, uint32_t n_iterations, double cvg, std::vector<double> &d,
                 std::vector<double> &V)
{
    auto at = [n](std::vector<double> &M, uint32_t i, uint32_t j) -> double & { return M[(size_t)i * n + j]; };
    std::vector<double> A(H), Q((size_t)n * n, 0.0), R((size_t)n * n, 0.0), a(n), e(n), u(n);
    d.assign(n, 0.0);
    V.assign((size_t)n * n, 0.0);
    for (uint32_t it = 0; it < n_iterations; it++) {
        for (uint32_t i = 0; i < n; i++) {
            for (uint32_t j = 0; j < n; j++) a[j] = at(A, j, i);
            u = a;
            for (uint32_t j = 0; j < i; j++) {
                for (uint32_t k = 0; k < n; k++) e[k] = at(Q, k, j);
                double dot = 0.0;
                for (uint32_t k = 0; k < n; k++) dot += a[k] * e[k];
                for (uint32_t k = 0; k < n; k++) u[k] -= dot * e[k];
            }
            double norm = 0.0;
            for (uint32_t k = 0; k < n; k++) norm += u[k] * u[k];
            norm = std::sqrt(norm);
            if (norm == 0.0) return fail(RS_ERR_INVALID_ARG, "rs_eigen_solve: QR decomposition hit a zero column");
            e = u;
            for (uint32_t k = 0; k < n; k++) e[k] *= 1.0 / norm;
            for (uint32_t j = 0; j < n; j++) at(Q, j, i) = e[j];
        }
        for (uint32_t i = 0; i < n; i++) {
            for (uint32_t j = 0; j < n; j++) e[j] = at(Q, j, i);
            for (uint32_t j = i; j < n; j++) {
                for (uint32_t k = 0; k < n; k++) a[k] = at(A, k, j);
                double dot = 0.0;
                for (uint32_t k = 0; k < n; k++) dot += a[k] * e[k];
                at(R, i, j) = dot;
            }
        }
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t j = 0; j < n; j++) {
                double s = 0.0;
                for (uint32_t k = 0; k < n; k++) s += at(R, i, k) * at(Q, k, j);
                at(A, i, j) = s;
            }
        if (it == 0) {
            V = Q;
        } else {
            std::vector<double> Vo(V);
            for (uint32_t i = 0; i < n; i++)
                for (uint32_t j = 0; j < n; j++) {
                    double s = 0.0;
                    for (uint32_t k = 0; k < n; k++) s += at(Vo, i, k) * at(Q, k, j);
                    at(V, i, j) = s;
                }
        }
        bool converged = true;
        for (uint32_t i = 1; i < n; i++)
            if (std::fabs(at(A, i, i - 1)) > cvg) converged = false;
        if (converged) break;
    }
    for (uint32_t i = 0; i < n; i++) d[i] = at(A, i, i);
    return RS_OK;
}

int check_pair(const rs_matrix *M, const rs_matrix *K)
{
    if (M->n != K->n) return fail(RS_ERR_INVALID_ARG, "rs_eigen_solve: M and K differ in size");
    if (M->n == 0) return fail(RS_ERR_INVALID_ARG, "rs_eigen_solve: empty system");
    return RS_OK;
}

}  // namespace
}  // namespace rs

using namespace rs;

static int eigen_solve(const rs_eigen_conf *ec, const rs_conf *mc, rs_matrix *M, rs_matrix *K, rs_learned *L,
                       double *d, uint32_t d_capacity, uint32_t *n_d, double *ev, uint64_t ev_capacity,
                       uint32_t *n_ev_rows, rs_eigen_info *info)
{
    if (!ec || !mc || !M || !K || !d || !n_d || !n_ev_rows)
        return fail(RS_ERR_INVALID_ARG, "rs_eigen_solve: null argument");
    int rc = require_runtime();
    if (rc) return rc;
    if ((rc = check_pair(M, K))) return rc;
    if (ec->method < RS_EIGEN_LANCZOS || ec->method > RS_EIGEN_RAYLEIGH)
        return fail(RS_ERR_INVALID_ARG, "Unknown eigen-value solver method.");
    const uint64_t n = M->n;
    hipStream_t st = runtime().stream;
    EvWork w;
    if ((rc = w.init(n))) return rc;
    InnerSolver solver;
    if ((rc = solver.create(mc))) return rc;
    solver.L = L;
    solver.K = K;
    std::vector<double> h(n);
    const uint32_t grid = ev_blocks(n);
    EventPair t0t1_;
    RS_HIP(t0t1_.create());
    hipEvent_t t0 = t0t1_.e[0], t1 = t0t1_.e[1];
    RS_HIP(hipEventRecord(t0, st));
    std::vector<double> dv;             // eigenvalues before the post-processing
    std::vector<double> evh;            // rows of n
    uint32_t nrows = 0, outer = 0;
    if (ec->method == RS_EIGEN_LANCZOS) {
        const uint32_t ne = (uint32_t)std::min<uint64_t>(ec->n_eigen_values, K->n);
        DBuf<double> v, vo, wv, b, t;
        if ((rc = v.alloc(n)) || (rc = vo.alloc(n)) || (rc = wv.alloc(n)) || (rc = b.alloc(n)) || (rc = t.alloc(n)))
            return rc;
        RS_HIP(hipMemsetAsync(vo.p, 0, 8 * n, st));
        RS_HIP(hipMemsetAsync(wv.p, 0, 8 * n, st));
        if ((rc = random_unit(w, v.p, h))) return rc;
        std::vector<double> e(ne, 0.0);
        dv.assign(ne, 0.0);
        for (uint32_t i = 0; i < ne; i++) {
            // K w = M v - e(i) K vo
            if ((rc = rs_matrix_spmv_device(M, v.p, b.p))) return rc;
            if (i > 0) {
                if ((rc = rs_matrix_spmv_device(K, vo.p, t.p))) return rc;
                k_ev_sub_scaled<<<grid, kBlock, 0, st>>>(b.p, t.p, e[i], n);
            }
            if ((rc = solver.solve(K, b.p, wv.p))) return rc;
            if ((rc = w.dot(wv.p, v.p, dv[i]))) return rc;
            if (i + 1 < ne) {
                k_ev_axmy<<<grid, kBlock, 0, st>>>(wv.p, v.p, dv[i], n);
                if ((rc = w.norm(wv.p, e[i + 1]))) return rc;
                if (e[i + 1] == 0.0) return fail(RS_ERR_INVALID_ARG, "Lanczos method failed: invariant subspace (e == 0)");
                RS_HIP(hipMemcpyAsync(vo.p, v.p, 8 * n, hipMemcpyDeviceToDevice, st));
                k_ev_div<<<grid, kBlock, 0, st>>>(v.p, wv.p, e[i + 1], n);
            }
            RS_HIP(hipGetLastError());
            outer++;
        }
        ql_reference(dv, e);
    } else if (ec->method == RS_EIGEN_ARNOLDI) {
        const uint32_t ne = (uint32_t)std::min<uint64_t>(ec->n_eigen_values, K->n);
        if (ne == 0) return fail(RS_ERR_INVALID_ARG, "rs_eigen_solve: no eigenvalues requested");
        DBuf<double> Qa, q, f, Qq, evd;
        if ((rc = Qa.alloc((uint64_t)n * ne)) || (rc = q.alloc(n)) || (rc = f.alloc(n))) return rc;
        RS_HIP(hipMemsetAsync(Qa.p, 0, 8ull * n * ne, st));
        RS_HIP(hipMemsetAsync(q.p, 0, 8 * n, st));
        // q: random normalised, first column of Qa
        if ((rc = random_unit(w, q.p, h))) return rc;
        RS_HIP(hipMemcpyAsync(Qa.p, q.p, 8 * n, hipMemcpyDeviceToDevice, st));
        std::vector<double> H((size_t)ne * ne, 0.0);
        for (uint32_t i = 1; i < ne + 1; i++) {
            const double *qo = Qa.p + (uint64_t)(i - 1) * n;
            if ((rc = rs_matrix_spmv_device(M, qo, f.p))) return rc;
            if ((rc = solver.solve(K, f.p, q.p))) return rc;
            for (uint32_t j = 0; j < i; j++) {
                const double *qj = Qa.p + (uint64_t)j * n;
                double hj = 0.0;
                if ((rc = w.dot(qj, q.p, hj))) return rc;
                H[(size_t)j * ne + (i - 1)] = hj;
                k_ev_axmy<<<grid, kBlock, 0, st>>>(q.p, qj, hj, n);
            }
            if (i < ne) {
                double hn = 0.0;
                if ((rc = w.norm(q.p, hn))) return rc;
                if (hn == 0.0) return fail(RS_ERR_INVALID_ARG, "Arnoldi method failed: invariant subspace (H == 0)");
                H[(size_t)i * ne + (i - 1)] = hn;
                k_ev_div<<<grid, kBlock, 0, st>>>(Qa.p + (uint64_t)i * n, q.p, hn, n);
            }
            RS_HIP(hipGetLastError());
            outer++;
        }
        std::vector<double> Vq;
        if ((rc = qr_reference(H, ne, ec->n_iterations, ec->solver_cvg_value, dv, Vq))) return rc;
        if ((rc = Qq.alloc((size_t)ne * ne)) || (rc = evd.alloc((uint64_t)ne * n))) return rc;
        RS_HIP(hipMemcpyAsync(Qq.p, Vq.data(), 8ull * ne * ne, hipMemcpyHostToDevice, st));
        k_ev_vectors<<<grid, kBlock, 0, st>>>(Qa.p, Qq.p, ne, n, evd.p);
        RS_HIP(hipGetLastError());
        evh.resize((size_t)ne * n);
        RS_HIP(hipMemcpyAsync(evh.data(), evd.p, 8ull * ne * n, hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
        nrows = ne;
    } else {
        // Rayleigh quotient iteration with M2 = M (+ mu K in each iteration); K and M share a pattern
        if (M->nnz != K->nnz) return fail(RS_ERR_UNSUPPORTED, "rs_eigen_solve: Rayleigh needs K and M with one pattern");
        std::vector<uint32_t> rpm(n + 1), rpk(n + 1), cm(M->nnz), ck(K->nnz);
        RS_HIP(hipMemcpyAsync(rpm.data(), M->rowptr.p, 4 * (n + 1), hipMemcpyDeviceToHost, st));
        RS_HIP(hipMemcpyAsync(rpk.data(), K->rowptr.p, 4 * (n + 1), hipMemcpyDeviceToHost, st));
        if (M->nnz) {
            RS_HIP(hipMemcpyAsync(cm.data(), M->col.p, 4 * M->nnz, hipMemcpyDeviceToHost, st));
            RS_HIP(hipMemcpyAsync(ck.data(), K->col.p, 4 * K->nnz, hipMemcpyDeviceToHost, st));
        }
        RS_HIP(hipStreamSynchronize(st));
        if (rpm != rpk || cm != ck) return fail(RS_ERR_UNSUPPORTED, "rs_eigen_solve: Rayleigh needs K and M with one pattern");
        std::vector<double> mv(M->nnz);
        if (M->nnz) RS_HIP(hipMemcpyAsync(mv.data(), M->val.p, 8 * M->nnz, hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
        rs_matrix *M2 = nullptr;
        if ((rc = rs_matrix_create((uint32_t)n, rpm.data(), cm.data(), mv.data(), &M2))) return rc;
        std::unique_ptr<rs_matrix, int (*)(rs_matrix *)> M2h(M2, rs_matrix_destroy);
        DBuf<double> m2val, b, f, c, dd;
        if ((rc = m2val.alloc(std::max<uint64_t>(M->nnz, 1))) || (rc = b.alloc(n)) || (rc = f.alloc(n)) ||
            (rc = c.alloc(n)) || (rc = dd.alloc(n)))
            return rc;
        if (M->nnz) RS_HIP(hipMemcpyAsync(m2val.p, M->val.p, 8 * M->nnz, hipMemcpyDeviceToDevice, st));
        RS_HIP(hipMemsetAsync(c.p, 0, 8 * n, st));
        RS_HIP(hipMemsetAsync(dd.p, 0, 8 * n, st));
        if ((rc = random_unit(w, b.p, h))) return rc;
        double mu = 1.0e9 * double(std::rand()) / double(RAND_MAX);
        const uint32_t rg = ev_blocks(n);
        for (uint32_t it = 0; it < ec->n_iterations; it++) {
            const double muo = mu;
            // M2 += K mu, f = (K mu) b
            k_ev_shift<<<rg, kBlock, 0, st>>>(K->rowptr.p, K->col.p, K->val.p, mu, m2val.p, b.p, f.p, (uint32_t)n, 1);
            RS_HIP(hipGetLastError());
            if ((rc = rs_matrix_update_values_device(M2, m2val.p))) return rc;
            // (M + K mu) c = K b mu
            if ((rc = solver.solve(M2, f.p, c.p))) return rc;
            double norm = 0.0;
            if ((rc = w.norm(c.p, norm))) return rc;
            if (norm == 0.0) return fail(RS_ERR_INVALID_ARG, "Rayleigh method failed: zero iterate");
            k_ev_scale<<<rg, kBlock, 0, st>>>(f.p, 1.0 / norm, n);
            RS_HIP(hipGetLastError());
            // M b = f (the reference solves with M here)
            if ((rc = solver.solve(M, f.p, b.p))) return rc;
            // M2 -= K mu
            k_ev_shift<<<rg, kBlock, 0, st>>>(K->rowptr.p, K->col.p, K->val.p, mu, m2val.p, b.p, f.p, (uint32_t)n, 0);
            // f = M2 b
            k_ev_rowdot<<<rg, kBlock, 0, st>>>(K->rowptr.p, K->col.p, m2val.p, b.p, f.p, (uint32_t)n);
            RS_HIP(hipGetLastError());
            if ((rc = rs_matrix_update_values_device(M2, m2val.p))) return rc;
            // K d = f
            if ((rc = solver.solve(K, f.p, dd.p))) return rc;
            double bb = 0.0, db = 0.0;
            if ((rc = w.dot(b.p, b.p, bb)) || (rc = w.dot(dd.p, b.p, db))) return rc;
            if (bb == 0.0) return fail(RS_ERR_INVALID_ARG, "Rayleigh method failed: zero iterate");
            mu = db / bb;
            outer++;
            if (it > 0 && std::fabs(muo - mu) < ec->solver_cvg_value) break;
        }
        dv.assign(1, mu);
        evh.resize(n);
        RS_HIP(hipMemcpyAsync(evh.data(), b.p, 8 * n, hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
        nrows = 1;
    }
    RS_HIP(hipEventRecord(t1, st));
    RS_HIP(hipEventSynchronize(t1));
    float ms = 0.0f;
    RS_HIP(hipEventElapsedTime(&ms, t0, t1));
    // REigenValueSolver::solve post-processing (src/reigenvaluesolver.cpp:89-116)
    if (dv.size() > 1) {
        for (double &x : dv) x = std::fabs(1.0 / x);
        std::swap(dv[0], dv[1]);
        std::vector<uint32_t> idx(dv.size());
        for (uint32_t i = 0; i < idx.size(); i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return dv[a] < dv[b]; });
        std::vector<double> sorted(dv.size());
        for (size_t i = 0; i < idx.size(); i++) sorted[i] = dv[idx[i]];
        dv.swap(sorted);
        for (uint32_t i = 0; i < nrows; i++) {
            const uint32_t n1 = i, n2 = idx[i];
            for (uint64_t j = 0; j < n; j++) std::swap(evh[(uint64_t)n1 * n + j], evh[(uint64_t)n2 * n + j]);
            idx[n2] = n2;
        }
    }
    if (dv.size() > d_capacity) return fail(RS_ERR_INVALID_ARG, "rs_eigen_solve: d_capacity too small");
    if (nrows && ev && (uint64_t)nrows * n > ev_capacity)
        return fail(RS_ERR_INVALID_ARG, "rs_eigen_solve: ev_capacity too small");
    std::copy(dv.begin(), dv.end(), d);
    *n_d = (uint32_t)dv.size();
    *n_ev_rows = nrows;
    if (ev && nrows) std::copy(evh.begin(), evh.end(), ev);
    if (info) {
        info->outer_iterations = outer;
        info->cg_iterations = solver.iterations;
        info->solve_ms = ms;
    }
    return RS_OK;
}

extern "C" int rs_eigen_solve(const rs_eigen_conf *ec, const rs_conf *mc, rs_matrix *M, rs_matrix *K, double *d,
                              uint32_t d_capacity, uint32_t *n_d, double *ev, uint64_t ev_capacity,
                              uint32_t *n_ev_rows, rs_eigen_info *info)
{
    return eigen_solve(ec, mc, M, K, nullptr, d, d_capacity, n_d, ev, ev_capacity, n_ev_rows, info);
}

extern "C" int rs_eigen_solve_learned(const rs_eigen_conf *ec, const rs_conf *mc, rs_matrix *M, rs_matrix *K,
                                      rs_learned *L, double *d, uint32_t d_capacity, uint32_t *n_d, double *ev,
                                      uint64_t ev_capacity, uint32_t *n_ev_rows, rs_eigen_info *info)
{
    if (!L) return fail(RS_ERR_INVALID_ARG, "rs_eigen_solve_learned: null preconditioner");
    if (mc && mc->solver_type != RS_SOLVER_CG)
        return fail(RS_ERR_UNSUPPORTED, "rs_eigen_solve_learned: the inner solver is flexible CG");
    return eigen_solve(ec, mc, M, K, L, d, d_capacity, n_d, ev, ev_capacity, n_ev_rows, info);
}
