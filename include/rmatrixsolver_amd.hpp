// rmatrixsolver_amd.hpp -- header-only C++ mirror of the reference's solver call surface,
// implemented over the C ABI of range_solver_amd.h (the GPU library).
//
// Mirrors (Range-Software/range-solver-lib):
//   RMatrixSolver          include/rmatrixsolver.h:9-52
//   RMatrixPreconditioner  include/rmatrixpreconditioner.h:7-64
// and minimal stand-ins for the range-model-lib / range-base-lib types those signatures use
// (RSparseMatrix with addValue semantics, RRVector, RMatrixSolverConf, RError), all in
// namespace rsamd.  Semantics kept: solve() resizes x to n with zeros and uses its content as
// the initial guess; unknown solver / preconditioner types throw RError; non-convergence does
// not throw; the flat CSR of a matrix is cached per matrix object and refreshed when the
// pattern is unchanged (src/rmatrixsolver.cpp:56-109), up to 16 objects (least recently used
// freed first; releaseDeviceMatrix frees one on request).
#pragma once

#include <algorithm>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "range_solver_amd.h"

namespace rsamd {

class RError : public std::runtime_error {
public:
    enum class Type { Application, Device };
    RError(Type type, const std::string &msg) : std::runtime_error(msg), type_(type) {}
    Type getType() const { return type_; }
    std::string getMessage() const { return what(); }

private:
    Type type_;
};

inline void check(int rc)
{
    if (rc == RS_OK) return;
    std::string msg = rs_last_error();
    if (rc == RS_ERR_UNKNOWN_SOLVER || rc == RS_ERR_INVALID_PRECOND || rc == RS_ERR_INVALID_ARG)
        throw RError(RError::Type::Application, msg);
    throw RError(RError::Type::Device, msg);
}

typedef enum _RMatrixPreconditionerType {
    R_MATRIX_PRECONDITIONER_NONE = RS_PRECOND_NONE,
    R_MATRIX_PRECONDITIONER_JACOBI = RS_PRECOND_JACOBI,
    R_MATRIX_PRECONDITIONER_BLOCK_JACOBI = RS_PRECOND_BLOCK_JACOBI,
    R_MATRIX_PRECONDITIONER_N_TYPES
} RMatrixPreconditionerType;

// RRVector stand-in: a resizable column of doubles.
class RRVector : public std::vector<double> {
public:
    RRVector() = default;
    explicit RRVector(unsigned n, double v = 0.0) : std::vector<double>(n, v) {}
    unsigned getNRows() const { return (unsigned)size(); }
    void resize(unsigned n, double v = 0.0) { std::vector<double>::resize(n, v); }
    void fill(double v) { std::fill(begin(), end(), v); }
};

// RRMatrix stand-in: a dense row-major matrix (rows x columns).
class RRMatrix {
public:
    RRMatrix() = default;
    RRMatrix(unsigned r, unsigned c, double v = 0.0) : r_(r), c_(c), a_((size_t)r * c, v) {}
    unsigned getNRows() const { return r_; }
    unsigned getNColumns() const { return c_; }
    void resize(unsigned r, unsigned c, double v = 0.0)
    {
        r_ = r;
        c_ = c;
        a_.assign((size_t)r * c, v);
    }
    double *operator[](unsigned i) { return a_.data() + (size_t)i * c_; }
    const double *operator[](unsigned i) const { return a_.data() + (size_t)i * c_; }
    double *data() { return a_.data(); }

private:
    unsigned r_ = 0, c_ = 0;
    std::vector<double> a_;
};

// RSparseMatrix stand-in: rows of (column, value) kept sorted; addValue accumulates.
class RSparseMatrix {
public:
    void setNRows(unsigned n) { rows_.resize(n); }
    unsigned getNRows() const { return (unsigned)rows_.size(); }
    void clear() { rows_.clear(); }
    void addValue(unsigned r, unsigned c, double v)
    {
        if (r >= rows_.size()) rows_.resize(r + 1);
        auto &row = rows_[r];
        auto it = std::lower_bound(row.begin(), row.end(), c,
                                   [](const std::pair<unsigned, double> &e, unsigned k) { return e.first < k; });
        if (it != row.end() && it->first == c) it->second += v;
        else row.insert(it, { c, v });
    }
    bool findColumnPosition(unsigned r, unsigned c, unsigned &pos) const
    {
        const auto &row = rows_[r];
        for (unsigned j = 0; j < row.size(); j++)
            if (row[j].first == c) { pos = j; return true; }
        return false;
    }
    double getValue(unsigned r, unsigned pos) const { return rows_[r][pos].second; }
    const std::vector<std::pair<unsigned, double>> &getRow(unsigned r) const { return rows_[r]; }

private:
    std::vector<std::vector<std::pair<unsigned, double>>> rows_;
};

// RMatrixSolverConf stand-in (range-model-lib).
class RMatrixSolverConf {
public:
    enum Type { CG = RS_SOLVER_CG, GMRES = RS_SOLVER_GMRES };
    explicit RMatrixSolverConf(Type t = CG) : type_(t) {}
    Type getType() const { return type_; }
    void setNInnerIterations(unsigned n) { nInner_ = n; }
    void setNOuterIterations(unsigned n) { nOuter_ = n; }
    void setSolverCvgValue(double v) { cvg_ = v; }
    void setOutputFrequency(unsigned f) { freq_ = f; }
    void setOutputFileName(const std::string &f) { file_ = f; }
    unsigned getNInnerIterations() const { return nInner_; }
    unsigned getNOuterIterations() const { return nOuter_; }
    double getSolverCvgValue() const { return cvg_; }
    unsigned getOutputFrequency() const { return freq_; }
    const std::string &getOutputFileName() const { return file_; }

private:
    Type type_;
    unsigned nInner_ = 1000, nOuter_ = 1000, freq_ = 1;
    double cvg_ = 1e-14;
    std::string file_;
};

namespace detail {
// Device CSR per matrix object, refreshed when the pattern is unchanged.
struct CsrCache {
    rs_matrix *h = nullptr;
    std::vector<uint32_t> rowptr, col;
    uint64_t last_use = 0;
    ~CsrCache()
    {
        if (h) rs_matrix_destroy(h);
    }
};

// The reference keeps one flat CSR per matrix object for the life of the process
// (src/rmatrixsolver.cpp:99-108).  Here the device copies are bounded: past kMaxCachedMatrices
// objects the least recently used copy is freed (a later solve with that object uploads it
// again), and releaseDeviceMatrix / releaseDeviceMatrices free them on request.
constexpr size_t kMaxCachedMatrices = 16;

struct CsrCaches {
    std::map<const RSparseMatrix *, std::unique_ptr<CsrCache>> map;
    uint64_t clock = 0;
};

inline CsrCaches &csr_caches()
{
    static CsrCaches c;
    return c;
}

inline rs_matrix *device_matrix(const RSparseMatrix &A)
{
    CsrCaches &all = csr_caches();
    auto &slot = all.map[&A];
    if (!slot) slot.reset(new CsrCache());
    CsrCache &c = *slot;
    c.last_use = ++all.clock;
    if (all.map.size() > kMaxCachedMatrices) {
        auto victim = all.map.end();
        for (auto it = all.map.begin(); it != all.map.end(); ++it)
            if (it->first != &A && (victim == all.map.end() || it->second->last_use < victim->second->last_use))
                victim = it;
        if (victim != all.map.end()) all.map.erase(victim);
    }
    unsigned n = A.getNRows();
    std::vector<uint32_t> rowptr(n + 1, 0), col;
    std::vector<double> val;
    for (unsigned i = 0; i < n; i++) {
        for (const auto &e : A.getRow(i)) {
            col.push_back(e.first);
            val.push_back(e.second);
        }
        rowptr[i + 1] = (uint32_t)col.size();
    }
    if (c.h && rowptr == c.rowptr && col == c.col) {
        check(rs_matrix_update_values(c.h, val.data()));
        return c.h;
    }
    if (c.h) rs_matrix_destroy(c.h);
    c.h = nullptr;
    check(rs_matrix_create(n, rowptr.data(), col.data(), val.data(), &c.h));
    c.rowptr.swap(rowptr);
    c.col.swap(col);
    return c.h;
}

inline size_t cached_matrices() { return csr_caches().map.size(); }
}  // namespace detail

// Free the device copy kept for matrix object A (for example before A is destroyed).
inline void releaseDeviceMatrix(const RSparseMatrix &A) { detail::csr_caches().map.erase(&A); }
// Free every device copy kept by the solvers of this header.
inline void releaseDeviceMatrices() { detail::csr_caches().map.clear(); }

// RMatrixPreconditioner (include/rmatrixpreconditioner.h:36-48)
class RMatrixPreconditioner {
public:
    RMatrixPreconditioner(const RSparseMatrix &matrix, RMatrixPreconditionerType type = R_MATRIX_PRECONDITIONER_NONE,
                          unsigned blockSize = 1)
    {
        check(rs_precond_create(detail::device_matrix(matrix), (int32_t)type, blockSize, &h_));
        n_ = matrix.getNRows();
    }
    // copy and assignment copy the type and the preconditioner values, as the reference's
    // _init does (src/rmatrixpreconditioner.cpp:6-14)
    RMatrixPreconditioner(const RMatrixPreconditioner &o) : n_(o.n_) { check(rs_precond_copy(o.h_, &h_)); }
    RMatrixPreconditioner &operator=(const RMatrixPreconditioner &o)
    {
        if (this != &o) {
            rs_precond *h = nullptr;
            check(rs_precond_copy(o.h_, &h));
            rs_precond_destroy(h_);
            h_ = h;
            n_ = o.n_;
        }
        return *this;
    }
    ~RMatrixPreconditioner() { rs_precond_destroy(h_); }
    void compute(const RRVector &x, RRVector &y) const
    {
        y.resize(n_);
        check(rs_precond_compute(h_, x.data(), y.data()));
    }

private:
    rs_precond *h_ = nullptr;
    unsigned n_ = 0;
};

// RMatrixSolver (include/rmatrixsolver.h:27-42)
class RMatrixSolver {
public:
    explicit RMatrixSolver(const RMatrixSolverConf &conf) : conf_(conf) { create(); }
    RMatrixSolver(const RMatrixSolver &o) : conf_(o.conf_) { check(rs_solver_copy(o.h_, &h_)); }
    RMatrixSolver &operator=(const RMatrixSolver &o)
    {
        if (this != &o) {
            rs_solver_destroy(h_);
            conf_ = o.conf_;
            check(rs_solver_copy(o.h_, &h_));
        }
        return *this;
    }
    ~RMatrixSolver() { rs_solver_destroy(h_); }

    void solve(const RSparseMatrix &A, const RRVector &b, RRVector &x,
               RMatrixPreconditionerType type = R_MATRIX_PRECONDITIONER_NONE, unsigned blockSize = 1)
    {
        rs_matrix *M = detail::device_matrix(A);
        x.resize(b.getNRows(), 0.0);
        check(rs_solver_solve(h_, M, b.data(), x.data(), (int32_t)type, blockSize, &info_));
    }
    void disableConvergenceLogFile() { check(rs_solver_disable_convergence_log(h_)); }
    const rs_info &lastInfo() const { return info_; }

private:
    void create()
    {
        rs_conf c{};
        c.solver_type = (int32_t)conf_.getType();
        c.n_outer_iterations = conf_.getNOuterIterations();
        c.n_inner_iterations = conf_.getNInnerIterations();
        c.output_frequency = conf_.getOutputFrequency();
        c.solver_cvg_value = conf_.getSolverCvgValue();
        c.stop_rule = RS_STOP_REFERENCE;
        c.precision = RS_PRECISION_FP64;
        c.output_file_name = conf_.getOutputFileName().empty() ? nullptr : conf_.getOutputFileName().c_str();
        check(rs_solver_create(&c, &h_));
    }
    RMatrixSolverConf conf_;
    rs_solver *h_ = nullptr;
    rs_info info_{};
};

// REigenValueSolverConf stand-in (range-model-lib): method, eigenvalues, iterations, tolerance.
class REigenValueSolverConf {
public:
    enum Method { Lanczos = RS_EIGEN_LANCZOS, Arnoldi = RS_EIGEN_ARNOLDI, Rayleigh = RS_EIGEN_RAYLEIGH };
    void setMethod(Method m) { method_ = m; }
    void setNEigenValues(unsigned n) { nEigen_ = n; }
    void setNIterations(unsigned n) { nIter_ = n; }
    void setSolverCvgValue(double v) { cvg_ = v; }
    Method getMethod() const { return method_; }
    unsigned getNEigenValues() const { return nEigen_; }
    unsigned getNIterations() const { return nIter_; }
    double getSolverCvgValue() const { return cvg_; }

private:
    Method method_ = Arnoldi;
    unsigned nEigen_ = 1, nIter_ = 100;
    double cvg_ = 1e-10;
};

// REigenValueSolver (include/reigenvaluesolver.h:9-63): solve(M, K, d, ev) on the GPU.
class REigenValueSolver {
public:
    REigenValueSolver(const REigenValueSolverConf &conf, const RMatrixSolverConf &matrixSolverConf)
        : conf_(conf), mconf_(matrixSolverConf)
    {
    }
    void solve(const RSparseMatrix &M, const RSparseMatrix &K, RRVector &d, RRMatrix &ev)
    {
        rs_matrix *Mh = detail::device_matrix(M);
        rs_matrix *Kh = detail::device_matrix(K);
        const unsigned n = K.getNRows();
        const unsigned cap = std::max(1u, std::min(conf_.getNEigenValues(), n));
        rs_eigen_conf ec{ (int32_t)conf_.getMethod(), conf_.getNEigenValues(), conf_.getNIterations(),
                          conf_.getSolverCvgValue() };
        rs_conf c{};
        c.solver_type = RS_SOLVER_CG;
        c.n_outer_iterations = mconf_.getNOuterIterations();
        c.n_inner_iterations = mconf_.getNInnerIterations();
        c.solver_cvg_value = mconf_.getSolverCvgValue();
        c.stop_rule = RS_STOP_REFERENCE;
        c.precision = RS_PRECISION_FP64;
        std::vector<double> dv(cap), evv((size_t)cap * n);
        uint32_t nd = 0, nr = 0;
        check(rs_eigen_solve(&ec, &c, Mh, Kh, dv.data(), cap, &nd, evv.data(), evv.size(), &nr, &info_));
        d.resize(nd);
        for (unsigned i = 0; i < nd; i++) d[i] = dv[i];
        ev.resize(nr, nr ? n : 0);
        for (unsigned i = 0; i < nr; i++)
            for (unsigned j = 0; j < n; j++) ev[i][j] = evv[(size_t)i * n + j];
    }
    const rs_eigen_info &lastInfo() const { return info_; }

private:
    REigenValueSolverConf conf_;
    RMatrixSolverConf mconf_;
    rs_eigen_info info_{};
};

}  // namespace rsamd
