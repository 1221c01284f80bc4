// Drop-in check of the C++ call surface: the known-answer cases of the reference's
// tests/tst_matrix_solver.cpp:33-77 (CG, preconditioner NONE, cvg 1e-13, 1000 iterations,
// tolerance 1e-6) through rsamd::RMatrixSolver (and the first through GMRES), plus the error behaviour of an unknown
// preconditioner type, the copy semantics of RMatrixPreconditioner, the bounded device-matrix
// cache, and an eigenvalue solve through rsamd::REigenValueSolver.  Exit status 0
// when every check passes.
#include <cmath>
#include <cstdio>
#include <vector>

#include "rmatrixsolver_amd.hpp"

using namespace rsamd;

static int failures = 0;
#define EXPECT(cond)                                                       \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);  \
            failures++;                                                    \
        }                                                                  \
    } while (0)

static RMatrixSolverConf cg_conf()
{
    RMatrixSolverConf conf(RMatrixSolverConf::CG);
    conf.setNInnerIterations(1000);
    conf.setNOuterIterations(1000);
    conf.setSolverCvgValue(1.0e-13);
    conf.setOutputFrequency(0);
    return conf;
}

int main()
{
    const double tol = 1.0e-6;
    {
        RSparseMatrix a;
        a.setNRows(2);
        a.addValue(0, 0, 4.0);
        a.addValue(0, 1, 1.0);
        a.addValue(1, 0, 1.0);
        a.addValue(1, 1, 3.0);
        RRVector b(2, 0.0);
        b[0] = 1.0;
        b[1] = 2.0;
        RRVector x;
        RMatrixSolver solver(cg_conf());
        solver.solve(a, b, x);
        EXPECT(x.getNRows() == 2u);
        EXPECT(std::fabs(x[0] - 1.0 / 11.0) < tol);
        EXPECT(std::fabs(x[1] - 7.0 / 11.0) < tol);
        std::printf("spd 2x2: x = [%.12f, %.12f] in %u iterations\n", x[0], x[1], solver.lastInfo().iterations);
        // the same known answer through GMRES (RMatrixSolverConf::GMRES, src/rmatrixsolver.cpp:357-581;
        // the reference's tests solve with CG only)
        RMatrixSolverConf gconf(RMatrixSolverConf::GMRES);
        gconf.setNInnerIterations(10);
        gconf.setNOuterIterations(100);
        gconf.setSolverCvgValue(1.0e-13);
        gconf.setOutputFrequency(0);
        RMatrixSolver gsolver(gconf);
        RRVector xg;
        gsolver.solve(a, b, xg, R_MATRIX_PRECONDITIONER_JACOBI, 1);
        EXPECT(std::fabs(xg[0] - 1.0 / 11.0) < tol);
        EXPECT(std::fabs(xg[1] - 7.0 / 11.0) < tol);
    }
    {
        RSparseMatrix a;
        a.setNRows(3);
        a.addValue(0, 0, 2.0);
        a.addValue(1, 1, 4.0);
        a.addValue(2, 2, 5.0);
        RRVector b(3, 0.0);
        b[0] = 6.0;
        b[1] = 8.0;
        b[2] = 10.0;
        RRVector x;
        RMatrixSolver solver(cg_conf());
        solver.solve(a, b, x);
        EXPECT(std::fabs(x[0] - 3.0) < tol);
        EXPECT(std::fabs(x[1] - 2.0) < tol);
        EXPECT(std::fabs(x[2] - 2.0) < tol);
        // same object, same pattern, new values: the cached CSR takes the refresh path
        a.addValue(0, 0, 2.0);  // A_00 = 4
        RRVector y;
        solver.solve(a, b, y, R_MATRIX_PRECONDITIONER_JACOBI, 1);
        EXPECT(std::fabs(y[0] - 1.5) < tol);
        // copy constructor keeps the configuration
        RMatrixSolver copy(solver);
        RRVector z;
        copy.solve(a, b, z, R_MATRIX_PRECONDITIONER_BLOCK_JACOBI, 3);
        EXPECT(std::fabs(z[2] - 2.0) < tol);
        RMatrixPreconditioner P(a, R_MATRIX_PRECONDITIONER_JACOBI, 1);
        RRVector pz;
        P.compute(b, pz);
        EXPECT(std::fabs(pz[1] - 2.0) < 1e-15);
        // copy constructor and assignment carry the type and the values
        // (include/rmatrixpreconditioner.h:39,45); the copies stay independent
        RMatrixPreconditioner Pc(P);
        RRVector pc;
        Pc.compute(b, pc);
        EXPECT(pc.getNRows() == 3u && pc[0] == pz[0] && pc[1] == pz[1] && pc[2] == pz[2]);
        RMatrixPreconditioner Pa(a, R_MATRIX_PRECONDITIONER_NONE, 1);
        RRVector pa;
        Pa.compute(b, pa);
        EXPECT(pa[0] == b[0]);
        Pa = P;
        Pa.compute(b, pa);
        EXPECT(pa[0] == pz[0] && pa[2] == pz[2]);
        {
            RMatrixPreconditioner Pb(a, R_MATRIX_PRECONDITIONER_BLOCK_JACOBI, 3);
            Pa = Pb;
        }  // Pb destroyed: Pa keeps its own copy
        Pa.compute(b, pa);
        EXPECT(std::fabs(pa[0] - 1.5) < 1e-15 && std::fabs(pa[1] - 2.0) < 1e-15 && std::fabs(pa[2] - 2.0) < 1e-15);
        P.compute(b, pz);
        EXPECT(pz[0] == pc[0]);
        bool threw = false;
        try {
            solver.solve(a, b, y, (RMatrixPreconditionerType)7, 1);
        } catch (const RError &e) {
            threw = e.getType() == RError::Type::Application;
        }
        EXPECT(threw);
    }
    {
        // REigenValueSolver (Arnoldi) on K = diag of three distinct values (10 each), M = I: the
        // 3-vector Krylov space is exact, so d = {2, 5, 11} after the reference's post-processing
        RSparseMatrix K, M;
        const unsigned n = 30;
        K.setNRows(n);
        M.setNRows(n);
        for (unsigned i = 0; i < n; i++) {
            K.addValue(i, i, i < 10 ? 2.0 : (i < 20 ? 5.0 : 11.0));
            M.addValue(i, i, 1.0);
        }
        REigenValueSolverConf econf;
        econf.setMethod(REigenValueSolverConf::Arnoldi);
        econf.setNEigenValues(3);
        econf.setNIterations(2000);
        econf.setSolverCvgValue(1e-14);
        RMatrixSolverConf mconf = cg_conf();
        REigenValueSolver es(econf, mconf);
        RRVector d;
        RRMatrix ev;
        es.solve(M, K, d, ev);
        EXPECT(d.getNRows() == 3u && ev.getNRows() == 3u && ev.getNColumns() == n);
        EXPECT(std::fabs(d[0] - 2.0) < 1e-8 && std::fabs(d[1] - 5.0) < 1e-8 && std::fabs(d[2] - 11.0) < 1e-8);
        std::printf("arnoldi: d = [%.10f, %.10f, %.10f]\n", d[0], d[1], d[2]);
    }
    {
        // the device copies of solved matrices are bounded: 40 matrix objects solved in turn
        // leave at most 16 copies, and each solve is still right after its copy was freed
        std::vector<RSparseMatrix> mats(40);
        RMatrixSolver solver(cg_conf());
        bool ok = true;
        for (int round = 0; round < 2; round++)
            for (unsigned k = 0; k < mats.size(); k++) {
                RSparseMatrix &m = mats[k];
                if (round == 0) {
                    m.setNRows(2);
                    m.addValue(0, 0, 1.0 + k);
                    m.addValue(1, 1, 2.0);
                }
                RRVector b(2, 2.0), x;
                solver.solve(m, b, x);
                ok = ok && std::fabs(x[0] - 2.0 / (1.0 + k)) < 1e-12 && std::fabs(x[1] - 1.0) < 1e-12;
            }
        EXPECT(ok);
        EXPECT(detail::cached_matrices() <= detail::kMaxCachedMatrices);
        releaseDeviceMatrix(mats[39]);
        EXPECT(detail::cached_matrices() <= detail::kMaxCachedMatrices - 1);
        releaseDeviceMatrices();
        EXPECT(detail::cached_matrices() == 0u);
    }
    std::printf("%s (%d failures)\n", failures ? "FAIL" : "PASS", failures);
    return failures ? 1 : 0;
}
