/*
 * rs_oracle.c -- CPU restatement of Range's assembly + sparse-solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product.  Each function cites the reference
 * file:line it restates; the operation order inside the solver loops is kept so that
 * iteration counts and convergence histories can be compared with the GPU path.
 *
 * Parity unpinned: the value of RConstants::eps (OR_EPS), the behaviour of RRMatrix::invert
 * on a singular block (here: the block becomes zero), the tet quadrature rule of
 * range-model-lib (here: the exact 4-point degree-2 rule, which gives the closed-form
 * consistent mass; stiffness and load do not depend on the rule).
 */
#include "rs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------------------------ */
/* RIterationInfo::hasConverged -- src/riterationinfo.cpp:90-109                          */
/* ------------------------------------------------------------------------------------ */
int or_converged_reason(double error, double trend, double cvg)
{
    if (!isnormal(error)) return OR_REASON_NONNORMAL;
    if (!isnormal(trend)) return OR_REASON_NONNORMAL;
    if (error < cvg) return OR_REASON_CONVERGED;
    if (fabs(trend) < pow(OR_EPS, 2)) return OR_REASON_STAGNATED;
    return OR_REASON_RUNNING;
}

int or_has_converged(double error, double trend, double cvg)
{
    return or_converged_reason(error, trend, cvg) != OR_REASON_RUNNING;
}

/* ------------------------------------------------------------------------------------ */
/* CSR helpers -- src/rmatrixsolver.cpp:23-54 (cache build, Frobenius norm), :111-119    */
/* ------------------------------------------------------------------------------------ */
double or_frobenius(unsigned n, const uint32_t *rowptr, const double *val)
{
    double norm2 = 0.0;
    int64_t nnz = rowptr[n];
#pragma omp parallel for reduction(+:norm2)
    for (int64_t j = 0; j < nnz; j++) norm2 += val[j] * val[j];
    return sqrt(norm2);
}

static inline double multiply_row(const uint32_t *rowptr, const uint32_t *col, const double *val,
                                  uint32_t row, const double *x)
{
    double v = 0.0;
    for (uint32_t j = rowptr[row]; j < rowptr[row + 1]; j++) v += val[j] * x[col[j]];
    return v;
}

void or_spmv(unsigned n, const uint32_t *rowptr, const uint32_t *col, const double *val,
             const double *x, double *y)
{
#pragma omp parallel for
    for (int64_t i = 0; i < (int64_t)n; i++) y[i] = multiply_row(rowptr, col, val, (uint32_t)i, x);
}

static int find_column(const uint32_t *rowptr, const uint32_t *col, uint32_t row, uint32_t c,
                       uint32_t *pos)
{
    for (uint32_t j = rowptr[row]; j < rowptr[row + 1]; j++) {
        if (col[j] == c) { *pos = j; return 1; }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Preconditioner -- src/rmatrixpreconditioner.cpp:16-190                                */
/* ------------------------------------------------------------------------------------ */

/* Gauss-Jordan inverse with partial pivoting, in place, w <= 16.  Stand-in for
 * RRMatrix::invert (range-base-lib); singular -> zero block (parity unpinned). */
static void invert_block(double *a, unsigned w)
{
    double m[16][32];
    for (unsigned i = 0; i < w; i++) {
        for (unsigned j = 0; j < w; j++) { m[i][j] = a[i * w + j]; m[i][w + j] = (i == j) ? 1.0 : 0.0; }
    }
    for (unsigned c = 0; c < w; c++) {
        unsigned piv = c;
        for (unsigned r = c + 1; r < w; r++) if (fabs(m[r][c]) > fabs(m[piv][c])) piv = r;
        if (m[piv][c] == 0.0) { for (unsigned i = 0; i < w * w; i++) a[i] = 0.0; return; }
        if (piv != c) {
            for (unsigned j = 0; j < 2 * w; j++) { double t = m[c][j]; m[c][j] = m[piv][j]; m[piv][j] = t; }
        }
        double inv = 1.0 / m[c][c];
        for (unsigned j = 0; j < 2 * w; j++) m[c][j] *= inv;
        for (unsigned r = 0; r < w; r++) {
            if (r == c) continue;
            double f = m[r][c];
            if (f == 0.0) continue;
            for (unsigned j = 0; j < 2 * w; j++) m[r][j] -= f * m[c][j];
        }
    }
    for (unsigned i = 0; i < w; i++) for (unsigned j = 0; j < w; j++) a[i * w + j] = m[i][w + j];
}

int or_precond_build(unsigned n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                     int type, unsigned block_size, double *data, unsigned *width_out)
{
    if (type == OR_PRECOND_NONE) { *width_out = 0; return 0; }
    if (type == OR_PRECOND_JACOBI) {
        /* constructJacobi :78-94 (serial in the reference) */
        for (unsigned i = 0; i < n; i++) {
            uint32_t pos;
            data[i] = 0.0;
            if (find_column(rowptr, col, i, i, &pos)) {
                double v = val[pos];
                data[i] = (v == 0.0) ? 0.0 : 1.0 / v;
            }
        }
        *width_out = 1;
        return 0;
    }
    if (type == OR_PRECOND_BLOCK_JACOBI) {
        /* constructBlockJacobi :96-148 -- width reduced until it divides n */
        unsigned w = block_size;
        if (w == 0 || w > 16) return -1;
        while (n % w != 0) w--;
        unsigned nb = n / w;
#pragma omp parallel for
        for (int64_t ib = 0; ib < (int64_t)nb; ib++) {
            double blk[256];
            for (unsigned k = 0; k < w; k++) {
                uint32_t m = (uint32_t)ib * w + k;
                for (unsigned l = 0; l < w; l++) {
                    uint32_t c = (uint32_t)ib * w + l, pos;
                    blk[k * w + l] = find_column(rowptr, col, m, c, &pos) ? val[pos] : 0.0;
                }
            }
            invert_block(blk, w);
            for (unsigned k = 0; k < w; k++)
                for (unsigned l = 0; l < w; l++) data[((uint32_t)ib * w + k) * w + l] = blk[k * w + l];
        }
        *width_out = w;
        return 0;
    }
    return -2;
}

void or_precond_apply(int type, unsigned n, unsigned width, const double *data,
                      const double *x, double *y)
{
    if (type == OR_PRECOND_JACOBI) {
        for (unsigned i = 0; i < n; i++) y[i] = data[i] * x[i];            /* :150-163 */
    } else if (type == OR_PRECOND_BLOCK_JACOBI) {
        unsigned nb = n / width;                                            /* :165-190 */
        for (unsigned ib = 0; ib < nb; ib++) {
            for (unsigned k = 0; k < width; k++) {
                unsigned m = ib * width + k;
                double v = 0.0;
                for (unsigned l = 0; l < width; l++) v += data[m * width + l] * x[ib * width + l];
                y[m] = v;
            }
        }
    } else {
        memcpy(y, x, sizeof(double) * n);                                  /* :55-66 */
    }
}

/* Apply inside an enclosing OpenMP parallel region (orphaned worksharing, as the reference). */
static void precond_apply_omp(int type, unsigned n, unsigned width, const double *data,
                              const double *x, double *y)
{
    if (type == OR_PRECOND_JACOBI) {
#pragma omp for
        for (int64_t i = 0; i < (int64_t)n; i++) y[i] = data[i] * x[i];
    } else if (type == OR_PRECOND_BLOCK_JACOBI) {
        int64_t nb = n / width;
#pragma omp for
        for (int64_t ib = 0; ib < nb; ib++) {
            for (unsigned k = 0; k < width; k++) {
                uint64_t m = (uint64_t)ib * width + k;
                double v = 0.0;
                for (unsigned l = 0; l < width; l++) v += data[m * width + l] * x[(uint64_t)ib * width + l];
                y[m] = v;
            }
        }
    } else {
#pragma omp for
        for (int64_t i = 0; i < (int64_t)n; i++) y[i] = x[i];
    }
}

/* ------------------------------------------------------------------------------------ */
/* Solvers -- src/rmatrixsolver.cpp:158-581                                               */
/* ------------------------------------------------------------------------------------ */
typedef struct iterinfo {
    double cvg, error, trend;
    unsigned iteration;
    double *history; unsigned history_cap;
} iterinfo;

/* RIterationInfo::setError -- src/riterationinfo.cpp:74-78 */
static void set_error(iterinfo *it, double error)
{
    it->trend = error - it->error;
    it->error = error;
    if (it->history && it->iteration < it->history_cap) {
        it->history[2 * it->iteration] = it->error;
        it->history[2 * it->iteration + 1] = it->trend;
    }
}

/* solveCG -- src/rmatrixsolver.cpp:208-355 */
static unsigned solve_cg(unsigned m, const uint32_t *rowptr, const uint32_t *col, const double *val,
                         double An, const double *b, double *x, int ptype, unsigned width,
                         const double *pdata, unsigned n_outer, int stop_rule, iterinfo *info)
{
    double *r = calloc(m ? m : 1, sizeof(double));
    double *z = calloc(m ? m : 1, sizeof(double));
    double *p = calloc(m ? m : 1, sizeof(double));
    double *q = calloc(m ? m : 1, sizeof(double));
    double ro0 = 0.0, ro1 = 0.0, beta = 0.0, dot = 0.0, alpha = 0.0, xn = 0.0, rn = 0.0, bn = 0.0;
    unsigned iterations = 0;
    int converged = 0;

#pragma omp parallel default(shared)
    {
#pragma omp for reduction(+:bn)
        for (int64_t i = 0; i < (int64_t)m; i++) {
            r[i] = b[i] - multiply_row(rowptr, col, val, (uint32_t)i, x);       /* :236-241 */
            bn += b[i] * b[i];
        }
#pragma omp barrier
#pragma omp master
        bn = sqrt(bn);

        for (unsigned it = 0; it < n_outer; it++) {
#pragma omp barrier
#pragma omp master
            info->iteration = it;
#pragma omp barrier
            precond_apply_omp(ptype, m, width, pdata, r, z);                   /* :258 */
#pragma omp master
            ro0 = 0.0;
#pragma omp barrier
#pragma omp for reduction(+:ro0)
            for (int64_t i = 0; i < (int64_t)m; i++) ro0 += r[i] * z[i];        /* :263-267 */
#pragma omp master
            {
                if (it > 0) {                                                   /* :271-275 */
                    beta = ro0 / ro1;
                    ro1 = fmax(ro1, OR_EPS);
                }
            }
#pragma omp barrier
#pragma omp for
            for (int64_t i = 0; i < (int64_t)m; i++) p[i] = (it == 0) ? z[i] : p[i] * beta + z[i];
#pragma omp master
            dot = 0.0;
#pragma omp barrier
#pragma omp for reduction(+:dot)
            for (int64_t i = 0; i < (int64_t)m; i++) {                          /* :295-300 */
                q[i] = multiply_row(rowptr, col, val, (uint32_t)i, p);
                dot += p[i] * q[i];
            }
#pragma omp master
            {
                if (dot > 0.0) dot = fmax(dot, OR_EPS);                         /* :304-315 */
                else if (dot < 0.0) dot = fmin(dot, -OR_EPS);
                else dot = OR_EPS;
                alpha = ro0 / dot;
                ro1 = ro0;
                xn = 0.0;
                rn = 0.0;
            }
#pragma omp barrier
#pragma omp for reduction(+:xn,rn)
            for (int64_t i = 0; i < (int64_t)m; i++) {                          /* :324-331 */
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
                xn += x[i] * x[i];
                rn += r[i] * r[i];
            }
#pragma omp master
            {
                xn = sqrt(xn);
                rn = sqrt(rn);
                double norm = An * xn + bn;                                     /* :338-344 */
                if (fabs(norm) < OR_EPS) norm = OR_EPS;
                double err = (stop_rule == OR_STOP_RELRES) ? rn / fmax(bn, OR_EPS) : rn / norm;
                set_error(info, err);
                iterations = it + 1;
                converged = or_has_converged(info->error, info->trend, info->cvg);
            }
#pragma omp barrier
            if (converged) break;                                               /* :349-352 */
        }
    }
    free(r); free(z); free(p); free(q);
    return iterations;
}

/* solveGMRES -- src/rmatrixsolver.cpp:357-581.  The classical Gram-Schmidt dot products of
 * :461-476 are summed per thread and then in thread order (the reference uses a critical
 * section, whose order is not fixed). */
static unsigned solve_gmres(unsigned mA, const uint32_t *rowptr, const uint32_t *col,
                            const double *val, double An, const double *b, double *x, int ptype,
                            unsigned width, const double *pdata, unsigned nouter, unsigned ninner,
                            int stop_rule, iterinfo *info, unsigned *inner_total)
{
    size_t N = mA ? mA : 1;
    double *ro = calloc(N, sizeof(double));
    double *w = calloc(N, sizeof(double));
    double *p = calloc(ninner + 1, sizeof(double));
    double *y = calloc(ninner + 1, sizeof(double));
    double *c = calloc(ninner + 1, sizeof(double));
    double *s = calloc(ninner + 1, sizeof(double));
    double *v = calloc((size_t)(ninner + 1) * N, sizeof(double));
    double *z = calloc((size_t)(ninner + 1) * N, sizeof(double));
    double *h = calloc((size_t)(ninner + 1) * ninner, sizeof(double));
    double *hv = calloc(ninner + 1, sizeof(double));
    int nthr = omp_get_max_threads();
    double *hloc = calloc((size_t)nthr * (ninner + 1), sizeof(double));
#define H(i, j) h[(size_t)(i) * ninner + (j)]
#define V(i) (v + (size_t)(i) * N)
#define Z(i) (z + (size_t)(i) * N)
    double bn = 0.0;
    for (unsigned i = 0; i < mA; i++) bn += b[i] * b[i];
    bn = sqrt(bn);
    unsigned iterations = 0, total = 0;

    for (unsigned ito = 0; ito < nouter; ito++) {
        info->iteration = ito;
        memset(h, 0, sizeof(double) * (size_t)(ninner + 1) * ninner);
        double beta = 0.0, xn = 0.0;
#pragma omp parallel for reduction(+:xn,beta)
        for (int64_t i = 0; i < (int64_t)mA; i++) {                             /* :401-407 */
            ro[i] = b[i] - multiply_row(rowptr, col, val, (uint32_t)i, x);
            xn += x[i] * x[i];
            beta += ro[i] * ro[i];
        }
        xn = sqrt(xn);
        beta = sqrt(beta);
        double norm = An * xn + bn;                                             /* :415-421 */
        if (fabs(norm) < OR_EPS) norm = OR_EPS;
        double err = (stop_rule == OR_STOP_RELRES) ? beta / fmax(bn, OR_EPS) : beta / norm;
        set_error(info, err);
        iterations = ito + 1;
        if (or_has_converged(info->error, info->trend, info->cvg)) break;

#pragma omp parallel for
        for (int64_t i = 0; i < (int64_t)mA; i++) V(0)[i] = (beta == 0.0) ? 0.0 : ro[i] / beta;
        memset(p, 0, sizeof(double) * (ninner + 1));
        p[0] = beta;

        unsigned iti;
        for (iti = 0; iti < ninner; iti++) {
            total++;
#pragma omp parallel
            precond_apply_omp(ptype, mA, width, pdata, V(iti), Z(iti));         /* :446 */
#pragma omp parallel for
            for (int64_t i = 0; i < (int64_t)mA; i++) w[i] = multiply_row(rowptr, col, val, (uint32_t)i, Z(iti));
            memset(hloc, 0, sizeof(double) * (size_t)nthr * (ninner + 1));
#pragma omp parallel
            {
                double *hl = hloc + (size_t)omp_get_thread_num() * (ninner + 1);
#pragma omp for
                for (int64_t j = 0; j < (int64_t)mA; j++) {                    /* :461-469 */
                    double wj = w[j];
                    for (unsigned i = 0; i <= iti; i++) hl[i] += wj * V(i)[j];
                }
            }
            for (unsigned i = 0; i <= iti; i++) {
                hv[i] = 0.0;
                for (int t = 0; t < nthr; t++) hv[i] += hloc[(size_t)t * (ninner + 1) + i];
                H(i, iti) = hv[i];                                              /* :480-483 */
            }
            double wnorm = 0.0;
#pragma omp parallel for reduction(+:wnorm)
            for (int64_t j = 0; j < (int64_t)mA; j++) {                         /* :487-504 */
                double value = w[j];
                for (unsigned i = 0; i <= iti; i++) value -= hv[i] * V(i)[j];
                w[j] = value;
                wnorm += value * value;
            }
            H(iti + 1, iti) = sqrt(wnorm);
            double hn = H(iti + 1, iti);
#pragma omp parallel for
            for (int64_t i = 0; i < (int64_t)mA; i++) V(iti + 1)[i] = (hn == 0.0) ? 0.0 : w[i] / hn;
            for (unsigned i = 1; i <= iti; i++) {                               /* :521-526 */
                double hsave = H(i - 1, iti);
                H(i - 1, iti) = c[i - 1] * hsave + s[i - 1] * H(i, iti);
                H(i, iti) = -s[i - 1] * hsave + c[i - 1] * H(i, iti);
            }
            double g = sqrt(H(iti, iti) * H(iti, iti) + H(iti + 1, iti) * H(iti + 1, iti));
            if (g == 0.0) { c[iti] = s[iti] = 0.0; }
            else { c[iti] = H(iti, iti) / g; s[iti] = H(iti + 1, iti) / g; }
            H(iti, iti) = g;
            H(iti + 1, iti) = 0.0;
            p[iti + 1] = -s[iti] * p[iti];
            p[iti] = c[iti] * p[iti];
            if (fabs(p[iti + 1]) < beta * 1.0e-4) { iti++; break; }             /* :546-550 */
        }
        /* back substitution :557-566 */
        y[iti - 1] = (H(iti - 1, iti - 1) == 0.0) ? 0.0 : p[iti - 1] / H(iti - 1, iti - 1);
        for (int i = (int)iti - 2; i >= 0; i--) {
            for (int j = (int)iti - 1; j > i; j--) p[i] -= y[j] * H(i, j);
            y[i] = (H(i, i) == 0.0) ? 0.0 : p[i] / H(i, i);
        }
#pragma omp parallel for
        for (int64_t i = 0; i < (int64_t)mA; i++) {                             /* :571-578 */
            for (unsigned j = 0; j <= iti - 1; j++) x[i] += Z(j)[i] * y[j];
        }
    }
#undef H
#undef V
#undef Z
    free(ro); free(w); free(p); free(y); free(c); free(s); free(v); free(z); free(h); free(hv); free(hloc);
    *inner_total = total;
    return iterations;
}

/* RMatrixSolver::solve -- src/rmatrixsolver.cpp:158-201 */
int or_solve(unsigned n, const uint32_t *rowptr, const uint32_t *col, const double *val,
             const double *b, double *x, int precond_type, unsigned block_size,
             const or_conf *conf, double prev_error, or_info *out, double *history,
             unsigned history_cap)
{
    unsigned width = 0;
    double *pdata = NULL;
    if (precond_type < 0 || precond_type > OR_PRECOND_BLOCK_JACOBI) return -2;
    if (conf->solver_type != OR_SOLVER_CG && conf->solver_type != OR_SOLVER_GMRES) return -3;
    if (precond_type != OR_PRECOND_NONE) {
        unsigned w = precond_type == OR_PRECOND_BLOCK_JACOBI ? block_size : 1;
        pdata = malloc(sizeof(double) * (size_t)(n ? n : 1) * (w ? w : 1));
        if (or_precond_build(n, rowptr, col, val, precond_type, block_size, pdata, &width) != 0) {
            free(pdata);
            return -2;
        }
    }
    double An = or_frobenius(n, rowptr, val);
    double bn = 0.0;
    for (unsigned i = 0; i < n; i++) bn += b[i] * b[i];
    bn = sqrt(bn);
    double scale = 1.0;
    if (bn != 0.0 && An != 0.0) scale = 1.0e9 / fabs(bn / An);                  /* :166-171 */

    double *y = malloc(sizeof(double) * (n ? n : 1));
    for (unsigned i = 0; i < n; i++) { y[i] = b[i] * scale; x[i] *= scale; }   /* :181-184 */

    iterinfo info = { conf->cvg, prev_error, 0.0, 0, history, history_cap };
    unsigned iters, inner = 0;
    if (conf->solver_type == OR_SOLVER_CG)
        iters = solve_cg(n, rowptr, col, val, An, y, x, precond_type, width, pdata, conf->n_outer,
                         conf->stop_rule, &info);
    else
        iters = solve_gmres(n, rowptr, col, val, An, y, x, precond_type, width, pdata, conf->n_outer,
                            conf->n_inner, conf->stop_rule, &info, &inner);
    double inv = 1.0 / scale;
    for (unsigned i = 0; i < n; i++) x[i] *= inv;                              /* :198 */

    if (out) {
        out->iterations = iters;
        out->inner_total = inner;
        out->error = info.error;
        out->trend = info.trend;
        out->equation_scale = scale;
        out->a_norm = An;
        out->b_norm = bn;
        int reason = or_converged_reason(info.error, info.trend, conf->cvg);
        out->reason = reason != OR_REASON_RUNNING ? reason : OR_REASON_MAXIT;
    }
    free(y);
    free(pdata);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Node book -- src/rsolvergeneric.cpp:498-582, src/rsolverstress.cpp:1564-1652           */
/* ------------------------------------------------------------------------------------ */
uint32_t or_node_book(uint32_t nn, uint32_t ne, uint32_t npe, const uint32_t *conn,
                      unsigned dpn, const uint8_t *fixed, uint32_t *book)
{
    uint8_t *computable = calloc(nn ? nn : 1, 1);
    for (uint64_t e = 0; e < ne; e++)
        for (uint32_t k = 0; k < npe; k++) computable[conn[e * npe + k]] = 1;
    uint32_t next = 0;
    for (uint64_t i = 0; i < (uint64_t)nn * dpn; i++) {                         /* :565-582 */
        int disabled = fixed[i] || !computable[i / dpn];
        book[i] = disabled ? OR_EOD : next++;
    }
    free(computable);
    return next;
}

/* ------------------------------------------------------------------------------------ */
/* Row-wise sparse matrix with RSparseMatrix::addValue (find-or-insert, accumulate)      */
/* ------------------------------------------------------------------------------------ */
typedef struct or_row { uint32_t n, cap; uint32_t *idx; double *val; } or_row;
struct or_sparse { uint32_t n_rows; or_row *rows; };

or_sparse *or_sparse_create(uint32_t n_rows)
{
    or_sparse *A = calloc(1, sizeof(or_sparse));
    A->n_rows = n_rows;
    A->rows = calloc(n_rows ? n_rows : 1, sizeof(or_row));
    return A;
}

void or_sparse_destroy(or_sparse *A)
{
    if (!A) return;
    for (uint32_t i = 0; i < A->n_rows; i++) { free(A->rows[i].idx); free(A->rows[i].val); }
    free(A->rows);
    free(A);
}

static void add_value(or_sparse *A, uint32_t r, uint32_t c, double v)
{
    or_row *row = &A->rows[r];
    uint32_t lo = 0, hi = row->n;
    while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (row->idx[mid] < c) lo = mid + 1; else hi = mid; }
    if (lo < row->n && row->idx[lo] == c) { row->val[lo] += v; return; }
    if (row->n == row->cap) {
        row->cap = row->cap ? 2 * row->cap : 16;
        row->idx = realloc(row->idx, sizeof(uint32_t) * row->cap);
        row->val = realloc(row->val, sizeof(double) * row->cap);
    }
    memmove(row->idx + lo + 1, row->idx + lo, sizeof(uint32_t) * (row->n - lo));
    memmove(row->val + lo + 1, row->val + lo, sizeof(double) * (row->n - lo));
    row->idx[lo] = c;
    row->val[lo] = v;
    row->n++;
}

uint64_t or_sparse_nnz(const or_sparse *A)
{
    uint64_t s = 0;
    for (uint32_t i = 0; i < A->n_rows; i++) s += A->rows[i].n;
    return s;
}

void or_sparse_to_csr(const or_sparse *A, uint32_t *rowptr, uint32_t *col, double *val)
{
    rowptr[0] = 0;
    for (uint32_t i = 0; i < A->n_rows; i++) rowptr[i + 1] = rowptr[i] + A->rows[i].n;
#pragma omp parallel for
    for (int64_t i = 0; i < (int64_t)A->n_rows; i++) {
        memcpy(col + rowptr[i], A->rows[i].idx, sizeof(uint32_t) * A->rows[i].n);
        memcpy(val + rowptr[i], A->rows[i].val, sizeof(double) * A->rows[i].n);
    }
}

/* ------------------------------------------------------------------------------------ */
/* Linear tetrahedron: shape functions, exact 4-point rule, RElement::findJacobian        */
/* ------------------------------------------------------------------------------------ */
static const double TET_DN[4][3] = { { -1, -1, -1 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
#define TQ_A 0.5854101966249685
#define TQ_B 0.1381966011250105
static const double TET_IP[4][3] = { { TQ_B, TQ_B, TQ_B }, { TQ_A, TQ_B, TQ_B },
                                     { TQ_B, TQ_A, TQ_B }, { TQ_B, TQ_B, TQ_A } };
static const double TET_W = 1.0 / 24.0;

static void tet_shape(const double ip[3], double N[4])
{
    N[0] = 1.0 - ip[0] - ip[1] - ip[2];
    N[1] = ip[0];
    N[2] = ip[1];
    N[3] = ip[2];
}

/* Returns detJ of dx/dxi and J[a][b] = (dx/dxi)^-1[b][a], so that
 * B[m][a] = sum_b dN[m][b] * J[a][b] is grad N_m (src/rsolverheat.cpp:405-411). */
static double tet_jacobian(const double X[4][3], double J[3][3])
{
    double G[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            G[a][b] = 0.0;
            for (int m = 0; m < 4; m++) G[a][b] += X[m][a] * TET_DN[m][b];
        }
    double det = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1])
               - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0])
               + G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
    double inv[3][3];
    inv[0][0] = (G[1][1] * G[2][2] - G[1][2] * G[2][1]) / det;
    inv[0][1] = (G[0][2] * G[2][1] - G[0][1] * G[2][2]) / det;
    inv[0][2] = (G[0][1] * G[1][2] - G[0][2] * G[1][1]) / det;
    inv[1][0] = (G[1][2] * G[2][0] - G[1][0] * G[2][2]) / det;
    inv[1][1] = (G[0][0] * G[2][2] - G[0][2] * G[2][0]) / det;
    inv[1][2] = (G[0][2] * G[1][0] - G[0][0] * G[1][2]) / det;
    inv[2][0] = (G[1][0] * G[2][1] - G[1][1] * G[2][0]) / det;
    inv[2][1] = (G[0][1] * G[2][0] - G[0][0] * G[2][1]) / det;
    inv[2][2] = (G[0][0] * G[1][1] - G[0][1] * G[1][0]) / det;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) J[a][b] = inv[b][a];
    return det;
}

/* ------------------------------------------------------------------------------------ */
/* Heat -- src/rsolverheat.cpp:361-455 (element loop), :755-820 (assemblyMatrix)          */
/* ------------------------------------------------------------------------------------ */
int or_heat_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                     const double *k, const double *q, const double *rhoc,
                     const double *t_node, const uint32_t *book, uint32_t n_dof,
                     int transient, double alpha, double dt, or_sparse *A, double *b)
{
    (void)nn;
    memset(b, 0, sizeof(double) * n_dof);
#pragma omp parallel for
    for (int64_t e = 0; e < (int64_t)ne; e++) {
        const uint32_t *nd = tets + 4 * e;
        double X[4][3], J[3][3], B[4][3], N[4];
        double Ke[4][4] = { { 0 } }, Me[4][4] = { { 0 } }, fe[4] = { 0 };
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) X[m][a] = xyz[3 * (size_t)nd[m] + a];
        for (int ip = 0; ip < 4; ip++) {                                        /* :397-436 */
            tet_shape(TET_IP[ip], N);
            double detJ = tet_jacobian(X, J);
            for (int m = 0; m < 4; m++)
                for (int a = 0; a < 3; a++)
                    B[m][a] = TET_DN[m][0] * J[a][0] + TET_DN[m][1] * J[a][1] + TET_DN[m][2] * J[a][2];
            for (int m = 0; m < 4; m++) {
                for (int n = 0; n < 4; n++) {
                    Ke[m][n] += (B[m][0] * B[n][0] + B[m][1] * B[n][1] + B[m][2] * B[n][2])
                              * k[e] * detJ * TET_W;
                    if (transient) Me[m][n] += N[m] * N[n] * rhoc[e] * detJ * TET_W;
                }
                fe[m] += q[e] * N[m] * detJ * TET_W;
            }
        }
        /* assemblyMatrix :755-820 */
        double Ae[4][4], be[4];
        if (transient) {
            for (int m = 0; m < 4; m++) {
                be[m] = dt * fe[m];
                for (int n = 0; n < 4; n++) {
                    Ae[m][n] = Me[m][n] + alpha * dt * Ke[m][n];
                    be[m] += (Me[m][n] - (1.0 - alpha) * dt * Ke[m][n]) * t_node[nd[n]];
                }
            }
        } else {
            memcpy(Ae, Ke, sizeof(Ae));
            memcpy(be, fe, sizeof(be));
        }
        for (int m = 0; m < 4; m++) {                                           /* :787-798 */
            if (book[nd[m]] == OR_EOD)
                for (int n = 0; n < 4; n++) be[n] -= Ae[n][m] * t_node[nd[m]];
        }
#pragma omp critical
        {
            for (int m = 0; m < 4; m++) {                                       /* :802-819 */
                uint32_t mp = book[nd[m]];
                if (mp == OR_EOD) continue;
                b[mp] += be[m];
                for (int n = 0; n < 4; n++) {
                    uint32_t np = book[nd[n]];
                    if (np != OR_EOD) add_value(A, mp, np, Ae[m][n]);
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Stress -- src/rsolverstress.cpp:573-715 (volume), :384-571 (traction), :1654-1772      */
/* ------------------------------------------------------------------------------------ */
static void stress_scatter(or_sparse *A, double *b, const uint32_t *book, const uint32_t *nd,
                           int npe, const double *Ae, const double *be, int with_matrix)
{
    int nd3 = 3 * npe;
    for (int m = 0; m < npe; m++) {
        for (int i = 0; i < 3; i++) {
            uint32_t mp = book[3 * (size_t)nd[m] + i];
            if (mp == OR_EOD) continue;
            b[mp] += be[3 * m + i];
            if (!with_matrix) continue;
            for (int n = 0; n < npe; n++)
                for (int j = 0; j < 3; j++) {
                    uint32_t np = book[3 * (size_t)nd[n] + j];
                    if (np != OR_EOD) add_value(A, mp, np, Ae[(3 * m + i) * nd3 + 3 * n + j]);
                }
        }
    }
}

int or_stress_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                       const double *E, const double *nu, const double *rho,
                       const double *gravity, const double *u_node,
                       uint32_t ntri, const uint32_t *tris, const double *traction,
                       const uint32_t *book, uint32_t n_dof, or_sparse *A, double *b)
{
    return or_stress_assemble_ex(nn, xyz, ne, tets, E, nu, rho, NULL, NULL, gravity, u_node, ntri, tris, traction,
                                 OR_STRESS_STATIC, 1.0, 1.0, NULL, NULL, book, n_dof, A, NULL, b);
}

/* The whole volume loop of RSolverStress::prepare (src/rsolverstress.cpp:573-715: Ke, the
 * consistent mass Me when the time solver is on or the analysis is modal (:665-680, needsMass
 * :56), gravity and thermal-expansion loads :683-693) and assemblyMatrix (:1654-1772): transient
 * Ae = Me + alpha dt Ke, be = dt fe + (Me - (1 - alpha) dt Ke) u (:1669-1703); modal Ae = Ke,
 * Be = Me (:1704-1709), M receives Me; static Ae = Ke; lifting subtracts (Ae + Be) u of every
 * prescribed position (:1719-1739).  alpha_t and dT (thermal expansion, element temperature
 * minus environment temperature) may be NULL (no thermal load).  Surface tractions enter as fe,
 * times dt in the transient case.  Local rotations (rot: 3x3 row-major per node, rot_active per
 * node; NULL for none) act as applyLocalRotations (:1774-1833): Ae -> T^T Ae T and Be -> T^T Be T,
 * but be -> T be, with T = diag(R of the active nodes, I otherwise), before the lifting; M receives
 * the unrotated Me. */
int or_stress_assemble_ex(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                          const double *E, const double *nu, const double *rho, const double *alpha_t,
                          const double *dT, const double *gravity, const double *u_node,
                          uint32_t ntri, const uint32_t *tris, const double *traction,
                          int problem, double alpha, double dt, const double *rot, const uint8_t *rot_active,
                          const uint32_t *book, uint32_t n_dof, or_sparse *A, or_sparse *M, double *b)
{
    (void)nn;
    const int transient = problem == OR_STRESS_TRANSIENT, modal = problem == OR_STRESS_MODAL;
    const int needs_mass = transient || modal;
    memset(b, 0, sizeof(double) * n_dof);
    /* Surface elements: force per unit area, :445-515 with surfaceThickness == 0, so the
     * element matrices are zero and only fe is scattered. */
    for (uint32_t t = 0; t < ntri; t++) {
        const uint32_t *nd = tris + 3 * (size_t)t;
        double P[3][3];
        for (int m = 0; m < 3; m++) for (int a = 0; a < 3; a++) P[m][a] = xyz[3 * (size_t)nd[m] + a];
        double e1[3], e2[3], cr[3];
        for (int a = 0; a < 3; a++) { e1[a] = P[1][a] - P[0][a]; e2[a] = P[2][a] - P[0][a]; }
        cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
        cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
        cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double detJ = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);    /* = 2*area */
        static const double tip[3][2] = { { 1.0 / 6, 1.0 / 6 }, { 2.0 / 3, 1.0 / 6 }, { 1.0 / 6, 2.0 / 3 } };
        double fe[9] = { 0 };
        for (int ip = 0; ip < 3; ip++) {
            double N[3] = { 1.0 - tip[ip][0] - tip[ip][1], tip[ip][0], tip[ip][1] };
            for (int m = 0; m < 3; m++) {
                double integ = N[m] * detJ * (1.0 / 6.0);
                for (int c = 0; c < 3; c++) fe[3 * m + c] += traction[3 * (size_t)t + c] * integ;
            }
        }
        if (transient) for (int k = 0; k < 9; k++) fe[k] *= dt;
        stress_scatter(A, b, book, nd, 3, NULL, fe, 0);
    }
#pragma omp parallel for
    for (int64_t e = 0; e < (int64_t)ne; e++) {
        const uint32_t *nd = tets + 4 * e;
        double X[4][3], J[3][3], B[4][3], N[4];
        double De[6][6] = { { 0 } }, Be[12][6], BeD[12][6];
        double Ke[144] = { 0 }, Me[144] = { 0 }, fe[12] = { 0 };
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) X[m][a] = xyz[3 * (size_t)nd[m] + a];
        double Ee = E[e], v = nu[e];
        De[0][0] = 1.0 - v; De[0][1] = v;       De[0][2] = v;                  /* :619-623 */
        De[1][0] = v;       De[1][1] = 1.0 - v; De[1][2] = v;
        De[2][0] = v;       De[2][1] = v;       De[2][2] = 1.0 - v;
        De[3][3] = De[4][4] = De[5][5] = (1.0 - 2.0 * v) / 2.0;
        double f = Ee / ((1.0 + v) * (1.0 - 2.0 * v));
        for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) De[i][j] *= f;
        const double th = (alpha_t && dT) ? alpha_t[e] * dT[e] : 0.0;
        for (int ip = 0; ip < 4; ip++) {
            tet_shape(TET_IP[ip], N);
            double detJ = tet_jacobian(X, J);
            for (int m = 0; m < 4; m++)
                for (int a = 0; a < 3; a++)
                    B[m][a] = TET_DN[m][0] * J[a][0] + TET_DN[m][1] * J[a][1] + TET_DN[m][2] * J[a][2];
            memset(Be, 0, sizeof(Be));
            for (int m = 0; m < 4; m++) {                                       /* :643-651 */
                Be[3 * m + 0][0] = B[m][0];
                Be[3 * m + 1][1] = B[m][1];
                Be[3 * m + 2][2] = B[m][2];
                Be[3 * m + 1][3] = B[m][2]; Be[3 * m + 2][3] = B[m][1];
                Be[3 * m + 0][4] = B[m][2]; Be[3 * m + 2][4] = B[m][0];
                Be[3 * m + 0][5] = B[m][1]; Be[3 * m + 1][5] = B[m][0];
            }
            for (int i = 0; i < 12; i++)
                for (int j = 0; j < 6; j++) {
                    double s = 0.0;
                    for (int l = 0; l < 6; l++) s += Be[i][l] * De[l][j];
                    BeD[i][j] = s;
                }
            for (int i = 0; i < 12; i++)                                        /* :654-663 */
                for (int j = 0; j < 12; j++) {
                    double s = 0.0;
                    for (int l = 0; l < 6; l++) s += BeD[i][l] * Be[j][l];
                    Ke[i * 12 + j] += s * detJ * TET_W;
                }
            for (int m = 0; m < 4; m++) {
                if (needs_mass)                                                 /* :665-680 */
                    for (int n = 0; n < 4; n++) {
                        const double val = N[m] * N[n] * rho[e] * detJ * TET_W;
                        for (int c = 0; c < 3; c++) Me[(3 * m + c) * 12 + 3 * n + c] += val;
                    }
                for (int c = 0; c < 3; c++) fe[3 * m + c] += gravity[c] * rho[e] * N[m] * detJ * TET_W;  /* :683-685 */
                for (int n = 0; n < 3; n++)                                     /* :688-693 */
                    for (int c = 0; c < 3; c++) fe[3 * m + c] += th * BeD[3 * m + c][n] * detJ * TET_W;
            }
        }
        /* assemblyMatrix */
        double Ae[144], Bm[144], be[12];
        for (int i = 0; i < 144; i++) {
            Ae[i] = transient ? Me[i] + alpha * dt * Ke[i] : Ke[i];
            Bm[i] = modal ? Me[i] : 0.0;
        }
        for (int i = 0; i < 12; i++) {
            be[i] = transient ? dt * fe[i] : fe[i];
            if (transient)
                for (int n = 0; n < 4; n++)
                    for (int c = 0; c < 3; c++)
                        be[i] += (Me[i * 12 + 3 * n + c] - (1.0 - alpha) * dt * Ke[i * 12 + 3 * n + c])
                               * u_node[3 * (size_t)nd[n] + c];
        }
        if (rot) {                                                              /* :1713-1716 */
            double T[144] = { 0 };
            int any = 0;
            for (int i = 0; i < 12; i++) T[i * 12 + i] = 1.0;
            for (int m = 0; m < 4; m++) {
                if (!rot_active[nd[m]]) continue;
                any = 1;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) T[(3 * m + i) * 12 + 3 * m + j] = rot[9 * (size_t)nd[m] + 3 * i + j];
            }
            if (any) {
                double tmp[144], out[144], bt[12];
                double *mats[2] = { Ae, Bm };
                for (int k = 0; k < 2; k++) {
                    double *X = mats[k];
                    for (int i = 0; i < 12; i++)                                /* tmp = T^T X */
                        for (int j = 0; j < 12; j++) {
                            double sum = 0.0;
                            for (int l = 0; l < 12; l++) sum += T[l * 12 + i] * X[l * 12 + j];
                            tmp[i * 12 + j] = sum;
                        }
                    for (int i = 0; i < 12; i++)                                /* out = tmp T */
                        for (int j = 0; j < 12; j++) {
                            double sum = 0.0;
                            for (int l = 0; l < 12; l++) sum += tmp[i * 12 + l] * T[l * 12 + j];
                            out[i * 12 + j] = sum;
                        }
                    memcpy(X, out, sizeof(out));
                }
                for (int i = 0; i < 12; i++) {                                  /* be = T be */
                    double sum = 0.0;
                    for (int l = 0; l < 12; l++) sum += T[i * 12 + l] * be[l];
                    bt[i] = sum;
                }
                memcpy(be, bt, sizeof(bt));
                /* M keeps the unrotated Me: assemblyMatrix scatters its Me argument, not the
                 * rotated copy Be (:1762-1765) */
            }
        }
        for (int m = 0; m < 4; m++)                                             /* :1719-1739 */
            for (int c = 0; c < 3; c++) {
                uint32_t dof = 3 * nd[m] + c;
                if (book[dof] != OR_EOD) continue;
                double u = u_node[dof];
                for (int n = 0; n < 12; n++) be[n] -= (Ae[n * 12 + 3 * m + c] + Bm[n * 12 + 3 * m + c]) * u;
            }
#pragma omp critical
        {
            stress_scatter(A, b, book, nd, 4, Ae, be, 1);
            if (modal && M) {
                double zero[12] = { 0 };
                stress_scatter(M, zero, book, nd, 4, Me, zero, 1);
            }
        }
    }
    return 0;
}

/* RSolverStress::setDisplacement (src/rsolverstress.cpp:800-836): du = x on the enabled
 * components and 0 on the disabled ones; an active local rotation writes iR^T du (iR = R^-1,
 * RLocalRotation::rotateResultsVector, src/rlocalrotation.cpp:60-77). */
void or_stress_set_displacement(uint32_t nn, const uint32_t *book, const double *x, const double *rot,
                                const uint8_t *rot_active, double *u)
{
    for (uint32_t i = 0; i < nn; i++) {
        double du[3];
        for (int c = 0; c < 3; c++) {
            const uint32_t p = book[3 * (size_t)i + c];
            du[c] = p != OR_EOD ? x[p] : 0.0;
        }
        if (rot && rot_active[i]) {
            const double *R = rot + 9 * (size_t)i;
            double iR[9];
            const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6])
                             + R[2] * (R[3] * R[7] - R[4] * R[6]);
            iR[0] = (R[4] * R[8] - R[5] * R[7]) / det;
            iR[1] = (R[2] * R[7] - R[1] * R[8]) / det;
            iR[2] = (R[1] * R[5] - R[2] * R[4]) / det;
            iR[3] = (R[5] * R[6] - R[3] * R[8]) / det;
            iR[4] = (R[0] * R[8] - R[2] * R[6]) / det;
            iR[5] = (R[2] * R[3] - R[0] * R[5]) / det;
            iR[6] = (R[3] * R[7] - R[4] * R[6]) / det;
            iR[7] = (R[1] * R[6] - R[0] * R[7]) / det;
            iR[8] = (R[0] * R[4] - R[1] * R[3]) / det;
            for (int c = 0; c < 3; c++)
                u[3 * (size_t)i + c] = iR[0 * 3 + c] * du[0] + iR[1 * 3 + c] * du[1] + iR[2 * 3 + c] * du[2];
        } else {
            for (int c = 0; c < 3; c++) u[3 * (size_t)i + c] = du[c];
        }
    }
}

/* RSolverMagnetostatics::prepare + assemblyMatrix (src/rsolvermagnetostatics.cpp:44-145,
 * 234-287): per tet and integration point Ke[3m+c][3n+c] -= (B_m . B_n) detJ w for each
 * component c, fe[3m+c] += N_m detJ w e0 (B_m x J_m)_c with J_m the node current density;
 * lifting of the prescribed field H on fixed nodes; unknowns 3 * position + c. */
int or_magnetostatics_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                               const double *current, const double *field, const uint32_t *book, uint32_t n_dof,
                               or_sparse *A, double *b)
{
    (void)nn; (void)n_dof;
    const double e0 = 8.854187817e-12;
    for (int64_t e = 0; e < (int64_t)ne; e++) {
        const uint32_t *nd = tets + 4 * e;
        double X[4][3], J[3][3], B[4][3], N[4];
        double Ke[144] = { 0 }, fe[12] = { 0 };
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) X[m][a] = xyz[3 * (size_t)nd[m] + a];
        for (int ip = 0; ip < 4; ip++) {
            tet_shape(TET_IP[ip], N);
            double detJ = tet_jacobian(X, J);
            for (int m = 0; m < 4; m++)
                for (int a = 0; a < 3; a++)
                    B[m][a] = TET_DN[m][0] * J[a][0] + TET_DN[m][1] * J[a][1] + TET_DN[m][2] * J[a][2];
            for (int m = 0; m < 4; m++) {
                const double *Jm = current + 3 * (size_t)nd[m];
                for (int n = 0; n < 4; n++) {
                    double kv = (B[m][0] * B[n][0] + B[m][1] * B[n][1] + B[m][2] * B[n][2]) * detJ * TET_W;
                    for (int c = 0; c < 3; c++) Ke[(3 * m + c) * 12 + 3 * n + c] -= kv;
                }
                double fv = N[m] * detJ * TET_W * e0;
                double jsx = -B[m][2] * Jm[1] + B[m][1] * Jm[2];
                double jsy = B[m][2] * Jm[0] - B[m][0] * Jm[2];
                double jsz = -B[m][1] * Jm[0] + B[m][0] * Jm[1];
                fe[3 * m + 0] += fv * jsx;
                fe[3 * m + 1] += fv * jsy;
                fe[3 * m + 2] += fv * jsz;
            }
        }
        for (int m = 0; m < 4; m++) {
            if (book[3 * (size_t)nd[m]] != OR_EOD) continue;
            const double *H = field + 3 * (size_t)nd[m];
            for (int n = 0; n < 12; n++)
                for (int c = 0; c < 3; c++) fe[n] -= Ke[n * 12 + 3 * m + c] * H[c];
        }
        stress_scatter(A, b, book, nd, 4, Ke, fe, 1);
    }
    return 0;
}

/* ------------------------------------------------------------------ fluid
 * RSolverFluid::prepare for the volume elements (src/rsolverfluid.cpp:306-331 element
 * averages of velocity and pressure, :425-437 element loop), computeElementConstantDerivative
 * (:1845-2201, the stabilised tetrahedron: SUPG, PSPG and LSIC parameters :1947-1987),
 * computeElementScales (:2235-2268, tetrahedron h_n = cbrt(6 V / pi)) and assemblyMatrix
 * (:2399-2419).  Four unknowns per node (vx, vy, vz, p), book[4 node + c]; the system is for
 * the increments, so disabled positions are simply left out (no lifting).  The element
 * constants iN, iNiN and wt (RElement::getMassVector, getMassMatrix, getTotalWeightFactor)
 * are the sums over the integration points of w N_m, w N_m N_n and w. */
int or_fluid_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                      const double *rho, const double *mu, const double *grav, const double *state,
                      const double *v_old, const double *acc, double stream_velocity, double alpha,
                      double dt, int unsteady, const uint32_t *book, uint32_t n_dof, or_sparse *A, double *b)
{
    (void)nn;
    memset(b, 0, sizeof(double) * n_dof);
    double iN[4] = { 0 }, iNiN[4][4] = { { 0 } }, wt = 0.0;
    for (int ip = 0; ip < 4; ip++) {
        double N[4];
        tet_shape(TET_IP[ip], N);
        wt += TET_W;
        for (int m = 0; m < 4; m++) {
            iN[m] += TET_W * N[m];
            for (int n = 0; n < 4; n++) iNiN[m][n] += TET_W * N[m] * N[n];
        }
    }
    const double inv6 = 1.0 / 6.0;
    const double mVh = stream_velocity, invmVh = 1.0 / stream_velocity;
    const double alphaDt = alpha * dt;
    for (int64_t e = 0; e < (int64_t)ne; e++) {
        const uint32_t *nd = tets + 4 * e;
        double X[4][3], J[3][3], B[4][3];
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) X[m][a] = xyz[3 * (size_t)nd[m] + a];
        const double detJ = tet_jacobian(X, J);
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++)
                B[m][a] = TET_DN[m][0] * J[a][0] + TET_DN[m][1] * J[a][1] + TET_DN[m][2] * J[a][2];
        const double ro = rho[e], invro = 1.0 / ro, u = mu[e];
        const double hn = cbrt(6.0 * (fabs(detJ) / 6.0) / 3.14159265358979323846); /* RConstants::pi */
        double ve[3] = { 0, 0, 0 }, veo[3] = { 0, 0, 0 }, p = 0.0, ax[4], ay[4], az[4];
        for (int m = 0; m < 4; m++) {
            const double *st = state + 4 * (size_t)nd[m];
            for (int c = 0; c < 3; c++) { ve[c] += st[c]; veo[c] += v_old[3 * (size_t)nd[m] + c]; }
            p += st[3];
            ax[m] = acc[3 * (size_t)nd[m] + 0];
            ay[m] = acc[3 * (size_t)nd[m] + 1];
            az[m] = acc[3 * (size_t)nd[m] + 2];
        }
        for (int c = 0; c < 3; c++) { ve[c] *= 0.25; veo[c] *= 0.25; }
        p *= 0.25;
        const double *g = grav + 3 * e;
        const double mvh = sqrt(veo[0] * veo[0] + veo[1] * veo[1] + veo[2] * veo[2]);
        const double invmvh = 1.0 / mvh;
        double s[3] = { 0, 0, 0 };
        if (mvh > 0.0) for (int c = 0; c < 3; c++) s[c] = veo[c] / mvh;
        double vdiv[4], vex[3] = { 0, 0, 0 }, vey[3] = { 0, 0, 0 }, vez[3] = { 0, 0, 0 };
        double px = 0, py = 0, pz = 0, h = 0;
        for (int m = 0; m < 4; m++) {                                          /* :1913-1940 */
            const double *st = state + 4 * (size_t)nd[m];
            vdiv[m] = ve[0] * B[m][0] + ve[1] * B[m][1] + ve[2] * B[m][2];
            for (int c = 0; c < 3; c++) {
                vex[c] += B[m][0] * st[c];
                vey[c] += B[m][1] * st[c];
                vez[c] += B[m][2] * st[c];
            }
            px += B[m][0] * st[3];
            py += B[m][1] * st[3];
            pz += B[m][2] * st[3];
            h += fabs(s[0] * B[m][0] + s[1] * B[m][1] + s[2] * B[m][2]);
        }
        if (h != 0.0) h = 2.0 / h;
        const double roD2u = 0.5 * ro / u;
        const double Re = roD2u * mvh * h, Ren = roD2u * mVh * hn;
        double Tsupg = 0.0, Tpspg = 0.0;
        if (mvh > 0.0) Tsupg = (Re > 0.0 && Re <= 3.0) ? h * Re * invmvh * inv6 : h * invmvh * 0.5;
        if (mVh > 0.0) Tpspg = (Ren > 0.0 && Ren <= 3.0) ? hn * Ren * invmVh * inv6 : hn * invmVh * 0.5;
        const double Tlsic = mvh * h * 0.5;
        const double vv[3] = { ve[0] * vex[0] + ve[1] * vey[0] + ve[2] * vez[0],
                               ve[0] * vex[1] + ve[1] * vey[1] + ve[2] * vez[1],
                               ve[0] * vex[2] + ve[1] * vey[2] + ve[2] * vez[2] };
        const double divV = vex[0] + vey[1] + vez[2];
        /* gradient of velocity component c along axis a: G[a][c] */
        const double G[3][3] = { { vex[0], vex[1], vex[2] }, { vey[0], vey[1], vey[2] }, { vez[0], vez[1], vez[2] } };
        const double gp[3] = { px, py, pz };
        const double acc_n[4][3] = { { ax[0], ay[0], az[0] }, { ax[1], ay[1], az[1] },
                                     { ax[2], ay[2], az[2] }, { ax[3], ay[3], az[3] } };
        double Ae[16][16], be[16];
        memset(Ae, 0, sizeof(Ae));
        memset(be, 0, sizeof(be));
        for (int m = 0; m < 4; m++) {
            for (int n = 0; n < 4; n++) {
                const double BB = B[m][0] * B[n][0] + B[m][1] * B[n][1] + B[m][2] * B[n][2];
                const double kScale = ro * u * wt, cpeScale = ro * iNiN[m][n];
                const double kte = Tsupg * ro * vdiv[m] * vdiv[n] * wt;
                const double ktpe = Tsupg * ro * vdiv[m] * iN[n], ktppe = Tsupg * ro * iN[n];
                const double ytpe = Tsupg * iN[n], epe = Tlsic * ro * wt;
                const double ce = ro * iN[m] * vdiv[n];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        /* row (m, i), column (n, j); Bij = B[m][i] B[n][j] as in :2010-2018 */
                        const double Bmi_nj = B[m][i] * B[n][j];
                        const double Bmj_ni = B[m][j] * B[n][i];
                        const double k = kScale * ((i == j ? BB : 0.0) + Bmj_ni);
                        const double cpe = cpeScale * G[j][i];
                        double steady = cpe + k + (i == j ? kte : 0.0) + ktpe * G[j][i]
                                      + ktppe * B[m][j] * vv[i] + ytpe * B[m][j] * gp[i] + Bmi_nj * epe;
                        double a = 0.0;
                        if (unsteady) {
                            const double mass = ro * iNiN[m][n];
                            const double ctilde = Tsupg * ro * vdiv[m] * iN[n];
                            const double ctpe = Tsupg * ro * iN[n];
                            if (i == j) a += mass + ctilde;
                            a += alphaDt * (ctpe * B[m][j] * acc_n[n][i]);
                        }
                        if (i == j) a += unsteady ? alphaDt * ce : ce;
                        if (unsteady) a += alphaDt * (steady - Bmi_nj * epe) + dt * Bmi_nj * epe;
                        else a += steady;
                        Ae[4 * m + i][4 * n + j] += a;
                    }
                const double yte = Tsupg * vdiv[m] * wt, s12 = unsteady ? -dt : -1.0;
                for (int i = 0; i < 3; i++) Ae[4 * m + i][4 * n + 3] += s12 * (B[m][i] * iN[n] + yte * B[n][i]);
                const double bte = Tpspg * iN[n], ye = Tpspg * vdiv[n] * wt, ype = Tpspg * iN[n];
                for (int j = 0; j < 3; j++) {
                    const double ypej = ype * (B[m][0] * G[j][0] + B[m][1] * G[j][1] + B[m][2] * G[j][2]);
                    if (unsteady)
                        Ae[4 * m + 3][4 * n + j] += bte * B[m][j] + dt * (B[n][j] * iN[m]) + alphaDt * (ye * B[m][j] + ypej);
                    else
                        Ae[4 * m + 3][4 * n + j] += B[n][j] * iN[m] + ye * B[m][j] + ypej;
                }
                const double theta = Tpspg * (BB * invro) * wt;
                Ae[4 * m + 3][4 * n + 3] += unsteady ? dt * theta : theta;
            }
            /* right-hand side :2150-2196 */
            const double iNm = iN[m], vdivm = vdiv[m];
            const double am[3] = { ax[m], ay[m], az[m] };
            for (int i = 0; i < 3; i++) {
                const double cv = iNm * ro * vv[i];
                const double kv = u * (B[m][0] * G[0][i] + B[m][1] * G[1][i] + B[m][2] * G[2][i]
                                      + B[m][0] * G[i][0] + B[m][1] * G[i][1] + B[m][2] * G[i][2]) * wt;
                const double gv = B[m][i] * p * wt;
                const double ktv = Tsupg * ro * vdivm * wt * vv[i];
                const double ytv = Tsupg * vdivm * wt * gp[i];
                const double ev = Tlsic * ro * divV * wt * B[m][i];
                const double fv = ro * iNm * g[i] + Tsupg * ro * vdivm * wt * g[i];
                const double res = cv + kv - gv + ktv - ytv + ev;
                if (unsteady) be[4 * m + i] += dt * (fv - (iNm * ro * am[i] + Tsupg * ro * vdivm * wt * am[i] + res));
                else be[4 * m + i] += fv - res;
            }
            const double Bg = B[m][0] * g[0] + B[m][1] * g[1] + B[m][2] * g[2];
            const double Ba = B[m][0] * am[0] + B[m][1] * am[1] + B[m][2] * am[2];
            const double Bv = B[m][0] * vv[0] + B[m][1] * vv[1] + B[m][2] * vv[2];
            const double Bp = B[m][0] * px + B[m][1] * py + B[m][2] * pz;
            if (unsteady)
                be[4 * m + 3] += dt * (Tpspg * Bg * wt - (Tpspg * Ba * wt + iNm * divV + Tpspg * Bv * wt
                                                          + Tpspg * (Bp * invro) * wt));
            else
                be[4 * m + 3] += Tpspg * Bg * wt - (iNm * divV + Tpspg * Bv * wt + Tpspg * (Bp * invro) * wt);
        }
        for (int r = 0; r < 16; r++) {
            be[r] *= detJ;
            for (int c = 0; c < 16; c++) Ae[r][c] *= detJ;
        }
        for (int r = 0; r < 16; r++) {                                          /* :2399-2419 */
            const uint32_t mp = book[4 * (size_t)nd[r / 4] + r % 4];
            if (mp == OR_EOD) continue;
            b[mp] += be[r];
            for (int c = 0; c < 16; c++) {
                const uint32_t np = book[4 * (size_t)nd[c / 4] + c % 4];
                if (np != OR_EOD) add_value(A, mp, np, Ae[r][c]);
            }
        }
    }
    return 0;
}

/* RSolverFluid::solve after the linear solve (src/rsolverfluid.cpp:519-624): the increments
 * are added to the node velocities and pressures; in a steady run the old velocities are the
 * ones before the update, in an unsteady run the acceleration is (v - v_old) / dt.  norms
 * receives |v| before and after, |p| before and after (Euclidean over all nodes). */
void or_fluid_update(uint32_t nn, const uint32_t *book, const double *x, int unsteady, double dt,
                     double *state, double *v_old, double *acc, double norms[4])
{
    double u0 = 0, p0 = 0, u1 = 0, p1 = 0;
    for (uint32_t i = 0; i < nn; i++) {
        for (int c = 0; c < 3; c++) u0 += state[4 * (size_t)i + c] * state[4 * (size_t)i + c];
        p0 += state[4 * (size_t)i + 3] * state[4 * (size_t)i + 3];
    }
    if (!unsteady)
        for (uint32_t i = 0; i < nn; i++)
            for (int c = 0; c < 3; c++) v_old[3 * (size_t)i + c] = state[4 * (size_t)i + c];
    for (uint32_t i = 0; i < nn; i++)
        for (int c = 0; c < 4; c++) {
            const uint32_t d = book[4 * (size_t)i + c];
            if (d != OR_EOD) state[4 * (size_t)i + c] += x[d];
        }
    if (unsteady)
        for (uint32_t i = 0; i < nn; i++)
            for (int c = 0; c < 3; c++)
                acc[3 * (size_t)i + c] = (state[4 * (size_t)i + c] - v_old[3 * (size_t)i + c]) * (1.0 / dt);
    for (uint32_t i = 0; i < nn; i++) {
        for (int c = 0; c < 3; c++) u1 += state[4 * (size_t)i + c] * state[4 * (size_t)i + c];
        p1 += state[4 * (size_t)i + 3] * state[4 * (size_t)i + 3];
    }
    norms[0] = sqrt(u0);
    norms[1] = sqrt(u1);
    norms[2] = sqrt(p0);
    norms[3] = sqrt(p1);
}

/* ------------------------------------------------------------------ scalar transport
 * RSolverFluidHeat and RSolverFluidParticle share one element (src/rsolverfluidheat.cpp:669-803,
 * src/rsolverfluidparticle.cpp:552-686): SUPG advection-diffusion of a node field (temperature /
 * concentration) by the element velocity, with capacity ca (rho c for heat, 1 for particles),
 * diffusivity k (conduction / diffusion), node source q (nodeHeat / nodeRate), and the heat-style
 * lifting of fixed nodes (assemblyMatrix :845-880 / :728-763).  cte stays zero (the container is
 * cleared and never set).  book has one entry per node. */
int or_transport_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets, const double *ca_e,
                          const double *k_e, const double *vel_e, const double *q_node, const double *field,
                          double alpha, double dt, int unsteady, const uint32_t *book, uint32_t n_dof,
                          or_sparse *A, double *b)
{
    (void)nn;
    memset(b, 0, sizeof(double) * n_dof);
    double iN[4] = { 0 }, iNiN[4][4] = { { 0 } }, wt = 0.0;
    for (int ip = 0; ip < 4; ip++) {
        double N[4];
        tet_shape(TET_IP[ip], N);
        wt += TET_W;
        for (int m = 0; m < 4; m++) {
            iN[m] += TET_W * N[m];
            for (int n = 0; n < 4; n++) iNiN[m][n] += TET_W * N[m] * N[n];
        }
    }
    for (int64_t e = 0; e < (int64_t)ne; e++) {
        const uint32_t *nd = tets + 4 * e;
        double X[4][3], J[3][3], B[4][3];
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) X[m][a] = xyz[3 * (size_t)nd[m] + a];
        const double detJ = tet_jacobian(X, J);
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++)
                B[m][a] = TET_DN[m][0] * J[a][0] + TET_DN[m][1] * J[a][1] + TET_DN[m][2] * J[a][2];
        const double ca = ca_e[e], k = k_e[e];
        const double *ve = vel_e + 3 * e;
        const double mvh = sqrt(ve[0] * ve[0] + ve[1] * ve[1] + ve[2] * ve[2]);
        double s[3] = { 0, 0, 0 };
        if (mvh > 0.0) for (int c = 0; c < 3; c++) s[c] = ve[c] / mvh;
        double vdiv[4], h = 0.0;
        for (int m = 0; m < 4; m++) {
            vdiv[m] = ve[0] * B[m][0] + ve[1] * B[m][1] + ve[2] * B[m][2];
            h += fabs(s[0] * B[m][0] + s[1] * B[m][1] + s[2] * B[m][2]);
        }
        if (h != 0.0) h = 2.0 / h;
        const double Re = k == 0.0 ? 0.0 : mvh * h / (2.0 * k);
        double Tsupg = 0.0;
        if (mvh > 0.0) {
            Tsupg = h / (2.0 * mvh);
            if (Re > 0.0 && Re <= 3.0) Tsupg *= Re / 3.0;
        }
        double Ae[4][4], be[4];
        for (int m = 0; m < 4; m++) {
            const double fv = q_node[nd[m]] * iN[m];
            double sum = 0.0;
            for (int n = 0; n < 4; n++) {
                const double me = unsteady ? ca * iNiN[m][n] : 0.0;
                const double ce = ca * iN[m] * vdiv[n];
                const double ke = -k * wt * (B[m][0] * B[n][0] + B[m][1] * B[n][1] + B[m][2] * B[n][2]);
                const double kte = Tsupg * ca * vdiv[m] * vdiv[n] * wt;
                const double yte = Tsupg * vdiv[m] * wt;
                const double cte = 0.0;
                const double ss = ce + ke + kte + yte;
                Ae[m][n] = unsteady ? me + cte + alpha * dt * ss : ss;
                if (unsteady) sum += (me + cte - (1.0 - alpha) * dt * ss) * field[nd[n]];
            }
            be[m] = unsteady ? dt * fv + sum : fv;
        }
        for (int m = 0; m < 4; m++) {
            be[m] *= detJ;
            for (int n = 0; n < 4; n++) Ae[m][n] *= detJ;
        }
        for (int m = 0; m < 4; m++)                                              /* lifting */
            if (book[nd[m]] == OR_EOD)
                for (int n = 0; n < 4; n++) be[n] -= Ae[n][m] * field[nd[m]];
        for (int m = 0; m < 4; m++) {
            const uint32_t mp = book[nd[m]];
            if (mp == OR_EOD) continue;
            b[mp] += be[m];
            for (int n = 0; n < 4; n++) {
                const uint32_t np = book[nd[n]];
                if (np != OR_EOD) add_value(A, mp, np, Ae[m][n]);
            }
        }
    }
    return 0;
}

/* RSolverFluidHeat::solve / RSolverFluidParticle::solve write-back (src/rsolverfluidheat.cpp:
 * 320-344): field = max(x, 0) on the enabled nodes; norms: |field| before and after. */
void or_transport_update(uint32_t nn, const uint32_t *book, const double *x, double *field, double norms[2])
{
    double a = 0.0, c = 0.0;
    for (uint32_t i = 0; i < nn; i++) a += field[i] * field[i];
    for (uint32_t i = 0; i < nn; i++)
        if (book[i] != OR_EOD) field[i] = fmax(x[book[i]], 0.0);
    for (uint32_t i = 0; i < nn; i++) c += field[i] * field[i];
    norms[0] = sqrt(a);
    norms[1] = sqrt(c);
}

/* ------------------------------------------------------------------ radiative heat
 * RSolverRadiativeHeat::prepare with the view-factor matrix loaded (src/rsolverradiativeheat.cpp:
 * 145-241): patch means of the element emissivity, temperature and ambient temperature over the
 * patch's elements (:175-205), then every (i, j) of the dense system
 *   A_ij = d_ij / E_j - F_ij (1 - E_j) / E_j   (0 when E_j == 0)
 *   b_i  = -sum_j (d_ij - F_ij) sigma T_j^4 + (1 - sum_j F_ij) sigma Tamb_i^4     (:211-238)
 * F is dense row-major (RViewFactorRow::getViewFactors().getValues(nPatches)); A is dense
 * row-major n_patches^2.  sigma = RSolverGeneric::sigma (src/rsolvergeneric.cpp:12). */
void or_radiosity_assemble(uint32_t np, const double *F, const uint32_t *patch_ptr, const uint32_t *patch_elems,
                           const double *emissivity, const double *temperature, const double *ambient,
                           double *A, double *b)
{
    const double sigma = 5.670367e-8;
    double *E = calloc(np ? np : 1, sizeof(double));
    double *T = calloc(np ? np : 1, sizeof(double));
    double *Ta = calloc(np ? np : 1, sizeof(double));
    for (uint32_t p = 0; p < np; p++) {
        for (uint32_t k = patch_ptr[p]; k < patch_ptr[p + 1]; k++) {
            uint32_t e = patch_elems[k];
            E[p] += emissivity[e];
            T[p] += temperature[e];
            Ta[p] += ambient[e];
        }
        uint32_t cnt = patch_ptr[p + 1] - patch_ptr[p];
        if (cnt > 0) { E[p] /= (double)cnt; T[p] /= (double)cnt; Ta[p] /= (double)cnt; }
    }
#pragma omp parallel for
    for (int64_t i = 0; i < (int64_t)np; i++) {
        double Fsum = 0.0, bi = 0.0;
        for (uint32_t j = 0; j < np; j++) {
            const double dij = (i == (int64_t)j) ? 1.0 : 0.0;
            const double Fij = F[(size_t)i * np + j], Ej = E[j];
            Fsum += Fij;
            double Aij = 0.0;
            if (Ej != 0.0) Aij = dij / Ej - Fij * (1.0 - Ej) / Ej;
            bi -= (dij - Fij) * sigma * pow(T[j], 4);
            A[(size_t)i * np + j] = Aij;
        }
        bi += (1.0 - Fsum) * sigma * pow(Ta[i], 4);
        b[i] = bi;
    }
    free(E); free(T); free(Ta);
}

/* RSolverRadiativeHeat::process (src/rsolverradiativeheat.cpp:276-303): each element of a patch
 * receives the patch heat times its share of the patch area; elements in no patch get 0. */
void or_radiosity_process(uint32_t np, uint32_t ne, const uint32_t *patch_ptr, const uint32_t *patch_elems,
                          const double *elem_area, const double *patch_heat, double *elem_heat)
{
    memset(elem_heat, 0, sizeof(double) * ne);
    for (uint32_t p = 0; p < np; p++) {
        double area = 0.0;
        for (uint32_t k = patch_ptr[p]; k < patch_ptr[p + 1]; k++) area += elem_area[patch_elems[k]];
        for (uint32_t k = patch_ptr[p]; k < patch_ptr[p + 1]; k++) {
            uint32_t e = patch_elems[k];
            elem_heat[e] = patch_heat[p] * (elem_area[e] / area);
        }
    }
}

/* ------------------------------------------------------------------ post-processing
 * RSolverHeat::process, volume elements (src/rsolverheat.cpp:646-692): per tet
 * B[m] = sum over the integration points of grad N_m * detJ / nInp; q = -sum_m B[m] k T_m. */
int or_heat_flux(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets, const double *k,
                 const double *t_node, double *flux)
{
    (void)nn;
#pragma omp parallel for
    for (int64_t e = 0; e < (int64_t)ne; e++) {
        const uint32_t *nd = tets + 4 * e;
        double X[4][3], J[3][3], B[4][3] = { { 0 } };
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) X[m][a] = xyz[3 * (size_t)nd[m] + a];
        for (int ip = 0; ip < 4; ip++) {
            double detJ = tet_jacobian(X, J);
            for (int m = 0; m < 4; m++)
                for (int a = 0; a < 3; a++)
                    B[m][a] += (TET_DN[m][0] * J[a][0] + TET_DN[m][1] * J[a][1] + TET_DN[m][2] * J[a][2]) * detJ / 4.0;
        }
        double Q[3] = { 0.0, 0.0, 0.0 };
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) Q[a] -= B[m][a] * k[e] * t_node[nd[m]];
        for (int a = 0; a < 3; a++) flux[3 * (size_t)e + a] = Q[a];
    }
    return 0;
}

/* RSolverElectrostatics::process, volume elements (src/rsolverelectrostatics.cpp:578-657). */
int or_electrostatics_process(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets, const double *eps_r,
                              const double *sigma, const double *phi, double *field, double *current, double *energy,
                              double *resistivity, double *joule)
{
    (void)nn;
    const double e0 = 8.854187817e-12;                    /* RSolverGeneric::e0, src/rsolvergeneric.cpp:13 */
#pragma omp parallel for
    for (int64_t e = 0; e < (int64_t)ne; e++) {
        const uint32_t *nd = tets + 4 * e;
        double X[4][3], J[3][3], B[4][3] = { { 0 } };
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) X[m][a] = xyz[3 * (size_t)nd[m] + a];
        for (int ip = 0; ip < 4; ip++) {
            double detJ = tet_jacobian(X, J);
            for (int m = 0; m < 4; m++)
                for (int a = 0; a < 3; a++)
                    B[m][a] += (TET_DN[m][0] * J[a][0] + TET_DN[m][1] * J[a][1] + TET_DN[m][2] * J[a][2]) * detJ / 4.0;
        }
        double E[3] = { 0.0, 0.0, 0.0 }, Jc[3];
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) E[a] -= B[m][a] * phi[nd[m]];
        double dE = E[0] * E[0] + E[1] * E[1] + E[2] * E[2];
        double ere = sqrt(dE);
        for (int a = 0; a < 3; a++) Jc[a] = E[a] * sigma[e];
        double jre = sqrt(Jc[0] * Jc[0] + Jc[1] * Jc[1] + Jc[2] * Jc[2]);
        double sv[3] = { 0.0, 0.0, 0.0 }, L = 0.0;
        if (ere > OR_EPS)
            for (int a = 0; a < 3; a++) sv[a] = E[a] / ere;
        for (int m = 0; m < 4; m++) L += fabs(sv[0] * B[m][0] + sv[1] * B[m][1] + sv[2] * B[m][2]);
        if (L > OR_EPS) L = 2.0 / L;
        for (int a = 0; a < 3; a++) {
            field[3 * (size_t)e + a] = E[a];
            current[3 * (size_t)e + a] = Jc[a];
        }
        energy[e] = eps_r[e] * e0 * dE / 2.0;
        resistivity[e] = jre > OR_EPS ? ere / jre : 0.0;
        joule[e] = sigma[e] * dE * L;
    }
    return 0;
}

/* RSolverStress::process, volume elements, static case (src/rsolverstress.cpp:1238-1408):
 * Ke and the element stresses Qe = sum_ip sum_m BeD^T u_m / nInp (no thermal load), normal
 * sqrt(sx^2 + sy^2 + sz^2 - (sx sy + sy sz + sz sx)), shear sqrt(3 (tyz^2 + txz^2 + txy^2)),
 * "von Mises" = normal + shear as the reference sums them, node forces += Ke u_e. */
int or_stress_process(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets, const double *E,
                      const double *nu, const double *u_node, double *normal, double *shear, double *vm,
                      double *node_force)
{
    memset(node_force, 0, sizeof(double) * 3 * (size_t)nn);
    for (int64_t e = 0; e < (int64_t)ne; e++) {
        const uint32_t *nd = tets + 4 * e;
        double X[4][3], J[3][3], B[4][3];
        double De[6][6] = { { 0 } }, Be[12][6], BeD[12][6];
        double Ke[144] = { 0 }, Qe[6] = { 0 }, xe[12];
        for (int m = 0; m < 4; m++)
            for (int a = 0; a < 3; a++) {
                X[m][a] = xyz[3 * (size_t)nd[m] + a];
                xe[3 * m + a] = u_node[3 * (size_t)nd[m] + a];
            }
        double Ee = E[e], v = nu[e];
        De[0][0] = 1.0 - v; De[0][1] = v;       De[0][2] = v;
        De[1][0] = v;       De[1][1] = 1.0 - v; De[1][2] = v;
        De[2][0] = v;       De[2][1] = v;       De[2][2] = 1.0 - v;
        De[3][3] = De[4][4] = De[5][5] = (1.0 - 2.0 * v) / 2.0;
        double f = Ee / ((1.0 + v) * (1.0 - 2.0 * v));
        for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) De[i][j] *= f;
        for (int ip = 0; ip < 4; ip++) {
            double detJ = tet_jacobian(X, J);
            for (int m = 0; m < 4; m++)
                for (int a = 0; a < 3; a++)
                    B[m][a] = TET_DN[m][0] * J[a][0] + TET_DN[m][1] * J[a][1] + TET_DN[m][2] * J[a][2];
            memset(Be, 0, sizeof(Be));
            for (int m = 0; m < 4; m++) {
                Be[3 * m + 0][0] = B[m][0];
                Be[3 * m + 1][1] = B[m][1];
                Be[3 * m + 2][2] = B[m][2];
                Be[3 * m + 1][3] = B[m][2]; Be[3 * m + 2][3] = B[m][1];
                Be[3 * m + 0][4] = B[m][2]; Be[3 * m + 2][4] = B[m][0];
                Be[3 * m + 0][5] = B[m][1]; Be[3 * m + 1][5] = B[m][0];
            }
            for (int i = 0; i < 12; i++)
                for (int j = 0; j < 6; j++) {
                    double sum = 0.0;
                    for (int l = 0; l < 6; l++) sum += Be[i][l] * De[l][j];
                    BeD[i][j] = sum;
                }
            for (int i = 0; i < 12; i++)
                for (int j = 0; j < 12; j++) {
                    double sum = 0.0;
                    for (int l = 0; l < 6; l++) sum += BeD[i][l] * Be[j][l];
                    Ke[i * 12 + j] += sum * detJ * TET_W;
                }
            double iv = 1.0 / 4.0;
            for (int m = 0; m < 4; m++)
                for (int n = 0; n < 6; n++)
                    Qe[n] += BeD[3 * m + 0][n] * xe[3 * m + 0] * iv + BeD[3 * m + 1][n] * xe[3 * m + 1] * iv +
                             BeD[3 * m + 2][n] * xe[3 * m + 2] * iv;
        }
        double fe[12];
        for (int i = 0; i < 12; i++) {
            double sum = 0.0;
            for (int j = 0; j < 12; j++) sum += Ke[i * 12 + j] * xe[j];
            fe[i] = sum;
        }
        double QeN = sqrt(Qe[0] * Qe[0] + Qe[1] * Qe[1] + Qe[2] * Qe[2] - (Qe[0] * Qe[1] + Qe[1] * Qe[2] + Qe[2] * Qe[0]));
        double QeS = sqrt(3.0 * (Qe[3] * Qe[3] + Qe[4] * Qe[4] + Qe[5] * Qe[5]));
        normal[e] = QeN;
        shear[e] = QeS;
        vm[e] = QeN + QeS;
        for (int m = 0; m < 4; m++)
            for (int c = 0; c < 3; c++) node_force[3 * (size_t)nd[m] + c] += fe[3 * m + c];
    }
    return 0;
}

/* ------------------------------------------------------------------ REigenValueSolver
 * src/reigenvaluesolver.cpp:41-592: K x = lambda M x through K^-1 M.  Start vectors from
 * rand() / RAND_MAX as the reference; inner solves by or_solve (CG, Jacobi) with the conf,
 * warm-started from the vector's previous contents, RIterationInfo's error carried from one
 * solve to the next (one RMatrixSolver object). */
static double vdot(unsigned n, const double *a, const double *b)
{
    double s = 0.0;
    for (unsigned i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

static int eig_inner(unsigned n, const uint32_t *rp, const uint32_t *col, const double *val, const double *b,
                     double *x, const or_conf *conf, double *prev_error, unsigned long long *its)
{
    or_info info;
    int rc = or_solve(n, rp, col, val, b, x, OR_PRECOND_JACOBI, 1, conf, *prev_error, &info, NULL, 0);
    *prev_error = info.error;
    *its += info.iterations;
    return rc;
}

static void random_unit(unsigned n, double *v)
{
    for (unsigned i = 0; i < n; i++) v[i] = (double)rand() / (double)RAND_MAX;
    double nrm = sqrt(vdot(n, v, v));
    for (unsigned i = 0; i < n; i++) v[i] /= nrm;
}

/* qlDecomposition, :420-496: the split search starts at m = l and the loop exits on
 * "m >= l" at once, so only the renumbering of e takes effect. */
static void ql_reference(unsigned n, double *d, double *e)
{
    (void)d;
    for (unsigned i = 1; i < n; i++) e[i - 1] = e[i];
    if (n) e[n - 1] = 0.0;
}

/* qrDecomposition, :498-592 */
static int qr_reference(unsigned n, const double *H, unsigned n_iter, double cvg, double *d, double *V)
{
    size_t nn = (size_t)n * n;
    double *A = malloc(sizeof(double) * nn), *Q = calloc(nn, sizeof(double)), *R = calloc(nn, sizeof(double));
    double *a = malloc(sizeof(double) * n), *e = malloc(sizeof(double) * n), *u = malloc(sizeof(double) * n);
    double *T = malloc(sizeof(double) * nn);
    memcpy(A, H, sizeof(double) * nn);
    memset(V, 0, sizeof(double) * nn);
    int rc = 0;
    for (unsigned it = 0; it < n_iter; it++) {
        for (unsigned i = 0; i < n; i++) {
            for (unsigned j = 0; j < n; j++) a[j] = A[(size_t)j * n + i];
            memcpy(u, a, sizeof(double) * n);
            for (unsigned j = 0; j < i; j++) {
                for (unsigned k = 0; k < n; k++) e[k] = Q[(size_t)k * n + j];
                double dot = vdot(n, a, e);
                for (unsigned k = 0; k < n; k++) u[k] -= dot * e[k];
            }
            double norm = sqrt(vdot(n, u, u));
            if (norm == 0.0) { rc = -1; goto done; }
            for (unsigned k = 0; k < n; k++) e[k] = u[k] * (1.0 / norm);
            for (unsigned j = 0; j < n; j++) Q[(size_t)j * n + i] = e[j];
        }
        for (unsigned i = 0; i < n; i++) {
            for (unsigned j = 0; j < n; j++) e[j] = Q[(size_t)j * n + i];
            for (unsigned j = i; j < n; j++) {
                for (unsigned k = 0; k < n; k++) a[k] = A[(size_t)k * n + j];
                R[(size_t)i * n + j] = vdot(n, a, e);
            }
        }
        for (unsigned i = 0; i < n; i++)
            for (unsigned j = 0; j < n; j++) {
                double sum = 0.0;
                for (unsigned k = 0; k < n; k++) sum += R[(size_t)i * n + k] * Q[(size_t)k * n + j];
                A[(size_t)i * n + j] = sum;
            }
        if (it == 0) {
            memcpy(V, Q, sizeof(double) * nn);
        } else {
            memcpy(T, V, sizeof(double) * nn);
            for (unsigned i = 0; i < n; i++)
                for (unsigned j = 0; j < n; j++) {
                    double sum = 0.0;
                    for (unsigned k = 0; k < n; k++) sum += T[(size_t)i * n + k] * Q[(size_t)k * n + j];
                    V[(size_t)i * n + j] = sum;
                }
        }
        int converged = 1;
        for (unsigned i = 1; i < n; i++)
            if (fabs(A[(size_t)i * n + i - 1]) > cvg) converged = 0;
        if (converged) break;
    }
    for (unsigned i = 0; i < n; i++) d[i] = A[(size_t)i * n + i];
done:
    free(A); free(Q); free(R); free(a); free(e); free(u); free(T);
    return rc;
}

int or_eigen_solve(int method, unsigned n_eigen, unsigned n_iter, double eig_cvg, const or_conf *conf, unsigned n,
                   const uint32_t *rp, const uint32_t *col, const double *mval, const double *kval, double *d,
                   unsigned *n_d, double *ev, unsigned *n_ev_rows, unsigned long long *cg_its)
{
    double prev = 0.0;
    unsigned long long its = 0;
    unsigned nrows = 0, nd = 0;
    int rc = 0;
    if (n == 0) return -1;
    if (method == 0) {                                   /* solveLanczos :119-205 */
        unsigned ne = n_eigen < n ? n_eigen : n;
        double *v = calloc(n, sizeof(double)), *vo = calloc(n, sizeof(double)), *w = calloc(n, sizeof(double));
        double *b = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n), *e = calloc(ne ? ne : 1, sizeof(double));
        random_unit(n, v);
        for (unsigned i = 0; i < ne; i++) {
            or_spmv(n, rp, col, mval, v, b);
            if (i > 0) {
                or_spmv(n, rp, col, kval, vo, t);
                for (unsigned j = 0; j < n; j++) t[j] *= e[i];
                for (unsigned j = 0; j < n; j++) b[j] -= t[j];
            }
            if ((rc = eig_inner(n, rp, col, kval, b, w, conf, &prev, &its))) break;
            d[i] = vdot(n, w, v);
            if (i + 1 < ne) {
                for (unsigned j = 0; j < n; j++) w[j] -= d[i] * v[j];
                e[i + 1] = sqrt(vdot(n, w, w));
                if (e[i + 1] == 0.0) { rc = -4; break; }
                memcpy(vo, v, sizeof(double) * n);
                for (unsigned j = 0; j < n; j++) v[j] = w[j] / e[i + 1];
            }
        }
        if (!rc) ql_reference(ne, d, e);
        nd = ne;
        free(v); free(vo); free(w); free(b); free(t); free(e);
    } else if (method == 1) {                            /* solveArnoldi :207-287 */
        unsigned ne = n_eigen < n ? n_eigen : n;
        double *Qa = calloc((size_t)n * ne, sizeof(double)), *q = calloc(n, sizeof(double)), *f = malloc(sizeof(double) * n);
        double *H = calloc((size_t)ne * ne, sizeof(double)), *Qq = malloc(sizeof(double) * ne * ne);
        random_unit(n, q);
        memcpy(Qa, q, sizeof(double) * n);               /* column 0 (column-major here) */
        for (unsigned i = 1; i < ne + 1 && !rc; i++) {
            or_spmv(n, rp, col, mval, Qa + (size_t)(i - 1) * n, f);
            if ((rc = eig_inner(n, rp, col, kval, f, q, conf, &prev, &its))) break;
            for (unsigned j = 0; j < i; j++) {
                const double *qj = Qa + (size_t)j * n;
                double h = vdot(n, qj, q);
                H[(size_t)j * ne + i - 1] = h;
                for (unsigned k = 0; k < n; k++) q[k] -= h * qj[k];
            }
            if (i < ne) {
                double h = sqrt(vdot(n, q, q));
                if (h == 0.0) { rc = -4; break; }
                H[(size_t)i * ne + i - 1] = h;
                for (unsigned k = 0; k < n; k++) Qa[(size_t)i * n + k] = q[k] / h;
            }
        }
        if (!rc) rc = qr_reference(ne, H, n_iter, eig_cvg, d, Qq);
        if (!rc && ev)
            for (unsigned i = 0; i < ne; i++)
                for (unsigned k = 0; k < n; k++) {
                    double sum = 0.0;
                    for (unsigned j = 0; j < ne; j++) sum += Qa[(size_t)j * n + k] * Qq[(size_t)j * ne + i];
                    ev[(size_t)i * n + k] = sum;
                }
        nd = ne;
        nrows = ne;
        free(Qa); free(q); free(f); free(H); free(Qq);
    } else if (method == 2) {                            /* solveRayleigh :289-418 (K, M share rp/col) */
        size_t nnz = rp[n];
        double *m2 = malloc(sizeof(double) * (nnz ? nnz : 1));
        double *b = malloc(sizeof(double) * n), *f = malloc(sizeof(double) * n), *c = calloc(n, sizeof(double));
        double *dd = calloc(n, sizeof(double));
        memcpy(m2, mval, sizeof(double) * nnz);
        random_unit(n, b);
        double mu = 1.0e9 * (double)rand() / (double)RAND_MAX;
        for (unsigned it = 0; it < n_iter; it++) {
            double muo = mu;
            for (unsigned i = 0; i < n; i++) {
                f[i] = 0.0;
                for (uint32_t j = rp[i]; j < rp[i + 1]; j++) {
                    double value = kval[j] * mu;
                    if (value != 0.0) m2[j] += value;
                    f[i] += value * b[col[j]];
                }
            }
            if ((rc = eig_inner(n, rp, col, m2, f, c, conf, &prev, &its))) break;
            double norm = sqrt(vdot(n, c, c));
            if (norm == 0.0) { rc = -4; break; }
            for (unsigned i = 0; i < n; i++) f[i] *= 1.0 / norm;
            if ((rc = eig_inner(n, rp, col, mval, f, b, conf, &prev, &its))) break;
            for (unsigned i = 0; i < n; i++)
                for (uint32_t j = rp[i]; j < rp[i + 1]; j++) {
                    double value = -1.0 * kval[j] * mu;
                    if (value != 0.0) m2[j] += value;
                }
            for (unsigned i = 0; i < n; i++) {
                f[i] = 0.0;
                for (uint32_t j = rp[i]; j < rp[i + 1]; j++) f[i] += m2[j] * b[col[j]];
            }
            if ((rc = eig_inner(n, rp, col, kval, f, dd, conf, &prev, &its))) break;
            double bb = vdot(n, b, b);
            if (bb == 0.0) { rc = -4; break; }
            mu = vdot(n, dd, b) / bb;
            if (it > 0 && fabs(muo - mu) < eig_cvg) break;
        }
        d[0] = mu;
        if (ev) memcpy(ev, b, sizeof(double) * n);
        nd = 1;
        nrows = 1;
        free(m2); free(b); free(f); free(c); free(dd);
    } else {
        return -2;
    }
    if (rc) return rc;
    /* REigenValueSolver::solve post-processing :89-116 (RUtil::qSort: ascending, stable) */
    if (nd > 1) {
        for (unsigned i = 0; i < nd; i++) d[i] = fabs(1.0 / d[i]);
        double tmp = d[0]; d[0] = d[1]; d[1] = tmp;
        unsigned *idx = malloc(sizeof(unsigned) * nd);
        double *sorted = malloc(sizeof(double) * nd);
        for (unsigned i = 0; i < nd; i++) idx[i] = i;
        for (unsigned i = 1; i < nd; i++) {              /* insertion sort: stable */
            unsigned v = idx[i], j = i;
            while (j > 0 && d[idx[j - 1]] > d[v]) { idx[j] = idx[j - 1]; j--; }
            idx[j] = v;
        }
        for (unsigned i = 0; i < nd; i++) sorted[i] = d[idx[i]];
        memcpy(d, sorted, sizeof(double) * nd);
        if (ev)
            for (unsigned i = 0; i < nrows; i++) {
                unsigned n1 = i, n2 = idx[i];
                for (unsigned j = 0; j < n; j++) {
                    double t = ev[(size_t)n1 * n + j];
                    ev[(size_t)n1 * n + j] = ev[(size_t)n2 * n + j];
                    ev[(size_t)n2 * n + j] = t;
                }
                idx[n2] = n2;
            }
        free(idx); free(sorted);
    }
    *n_d = nd;
    *n_ev_rows = nrows;
    if (cg_its) *cg_its = its;
    return 0;
}

