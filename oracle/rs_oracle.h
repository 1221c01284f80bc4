/*
 * rs_oracle.h -- CPU restatement of Range's element-assembly + sparse-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the timed CPU
 * baseline.  The product (range-solver-lib_amd/) never links, loads or calls it.
 *
 * Every function restates a reference function; the citation names the file:line it
 * follows (paths relative to the Range-Software/range-solver-lib tree).  Plain C11 +
 * OpenMP, mirroring the reference's own OpenMP loops so the timing is representative.
 *
 * Parity pins (tests/test_oracle.py): tst_matrix_solver.cpp vectors, direct sparse solves,
 * CG finite-termination on k distinct eigenvalues, closed-form tet matrices computed by a
 * different formulation, K*1=0, rigid-body kernels, patch tests.  Functions with no pin:
 * RConstants::eps (value chosen here, OR_EPS), RRMatrix::invert on singular blocks.
 */
#ifndef RS_ORACLE_H
#define RS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_EOD 0xFFFFFFFFu
/* RConstants::eps lives in range-base-lib (absent): value chosen here. */
#define OR_EPS 1.0e-14

enum { OR_PRECOND_NONE = 0, OR_PRECOND_JACOBI = 1, OR_PRECOND_BLOCK_JACOBI = 2 };
enum { OR_SOLVER_CG = 0, OR_SOLVER_GMRES = 1 };
enum { OR_STOP_REFERENCE = 0, OR_STOP_RELRES = 1 };
enum { OR_REASON_RUNNING = 0, OR_REASON_CONVERGED = 1, OR_REASON_STAGNATED = 2,
       OR_REASON_NONNORMAL = 3, OR_REASON_MAXIT = 4 };

typedef struct or_conf {
    int solver_type;          /* OR_SOLVER_* */
    unsigned n_outer;         /* RMatrixSolverConf::getNOuterIterations */
    unsigned n_inner;         /* RMatrixSolverConf::getNInnerIterations (GMRES restart) */
    double cvg;               /* RMatrixSolverConf::getSolverCvgValue */
    int stop_rule;            /* OR_STOP_REFERENCE: rn/(An*xn+bn); OR_STOP_RELRES: rn/bn */
} or_conf;

typedef struct or_info {
    unsigned iterations;      /* iterations performed (CG) / outer iterations (GMRES) */
    unsigned inner_total;     /* GMRES: total inner steps */
    double error, trend;      /* last RIterationInfo error / trend */
    double equation_scale, a_norm, b_norm;
    int reason;
} or_info;

/* RIterationInfo::hasConverged, src/riterationinfo.cpp:90-109 */
int or_has_converged(double error, double trend, double cvg);
int or_converged_reason(double error, double trend, double cvg);

/* RMatrixPreconditioner ctor + construct*, src/rmatrixpreconditioner.cpp:16-148.
 * data must hold n*block_size doubles; *width_out receives the effective block width. */
int or_precond_build(unsigned n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                     int type, unsigned block_size, double *data, unsigned *width_out);
/* RMatrixPreconditioner::compute, src/rmatrixpreconditioner.cpp:51-76,150-190 */
void or_precond_apply(int type, unsigned n, unsigned width, const double *data,
                      const double *x, double *y);

/* ||A||_F over stored entries, src/rmatrixsolver.cpp:37-51 */
double or_frobenius(unsigned n, const uint32_t *rowptr, const double *val);

/* RMatrixSolver::solve + solveCG / solveGMRES, src/rmatrixsolver.cpp:158-581.
 * x is in/out (initial guess).  prev_error seeds RIterationInfo's stored error (0 for a
 * fresh solver).  history (nullable) receives (error, trend) per iteration. */
int or_solve(unsigned n, const uint32_t *rowptr, const uint32_t *col, const double *val,
             const double *b, double *x, int precond_type, unsigned block_size,
             const or_conf *conf, double prev_error, or_info *info, double *history,
             unsigned history_cap);

/* y = A x, src/rmatrixsolver.cpp:111-119 */
void or_spmv(unsigned n, const uint32_t *rowptr, const uint32_t *col, const double *val,
             const double *x, double *y);

/* ---- DOF numbering (node book) ---------------------------------------------------- */
/* RSolverGeneric::generateNodeBook + rebuildNodeBook, src/rsolvergeneric.cpp:498-582,
 * and RSolverStress::generateNodeBook, src/rsolverstress.cpp:1564-1652.
 * fixed[nn*dpn] marks disabled positions; positions of nodes touched by no element are
 * disabled too.  book[nn*dpn] receives positions or OR_EOD.  Returns the enabled count. */
uint32_t or_node_book(uint32_t nn, uint32_t ne, uint32_t npe, const uint32_t *conn,
                      unsigned dpn, const uint8_t *fixed, uint32_t *book);

/* ---- Assembly into a row-wise sparse matrix (RSparseMatrix::addValue semantics) ----- */
typedef struct or_sparse or_sparse;
or_sparse *or_sparse_create(uint32_t n_rows);
void or_sparse_destroy(or_sparse *A);
uint64_t or_sparse_nnz(const or_sparse *A);
/* Export as CSR with sorted columns; arrays sized n+1 / nnz / nnz. */
void or_sparse_to_csr(const or_sparse *A, uint32_t *rowptr, uint32_t *col, double *val);

/* RSolverHeat::prepare volume loop + assemblyMatrix, src/rsolverheat.cpp:361-455, 755-820.
 * rhoc (density*capacity per element) is used only when transient != 0.
 * t_node: nodeTemperature (Dirichlet values at fixed nodes; T_old when transient).
 * b must hold n_dof doubles (zeroed here). */
int or_heat_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                     const double *k, const double *q, const double *rhoc,
                     const double *t_node, const uint32_t *book, uint32_t n_dof,
                     int transient, double alpha, double dt, or_sparse *A, double *b);

/* RSolverStress::prepare volume loop + surface traction + assemblyMatrix,
 * src/rsolverstress.cpp:573-715, 384-571, 1654-1772 (static, no local rotations).
 * book: 3*nn positions; u_node: 3*nn prescribed displacements (x,y,z interleaved);
 * gravity: 3 doubles (elementGravity, with rho); tris/traction: surface triangles and the
 * force per unit area on each (3 doubles per triangle). */
int or_stress_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                       const double *E, const double *nu, const double *rho,
                       const double *gravity, const double *u_node,
                       uint32_t ntri, const uint32_t *tris, const double *traction,
                       const uint32_t *book, uint32_t n_dof, or_sparse *A, double *b);
enum { OR_STRESS_STATIC = 0, OR_STRESS_TRANSIENT = 1, OR_STRESS_MODAL = 2 };
/* The full RSolverStress volume loop + assemblyMatrix (src/rsolverstress.cpp:573-715, 1654-1772):
 * thermal load (alpha_t, dT per element, may be NULL), consistent mass, transient theta-scheme,
 * modal M (may be NULL unless modal). */
int or_stress_assemble_ex(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                          const double *E, const double *nu, const double *rho, const double *alpha_t,
                          const double *dT, const double *gravity, const double *u_node,
                          uint32_t ntri, const uint32_t *tris, const double *traction,
                          int problem, double alpha, double dt, const double *rot, const uint8_t *rot_active,
                          const uint32_t *book, uint32_t n_dof, or_sparse *A, or_sparse *M, double *b);
/* RSolverStress::setDisplacement with local rotations (src/rsolverstress.cpp:800-836). */
void or_stress_set_displacement(uint32_t nn, const uint32_t *book, const double *x, const double *rot,
                                const uint8_t *rot_active, double *u);

/* RSolverMagnetostatics prepare + assemblyMatrix (volume), src/rsolvermagnetostatics.cpp:44-145,
 * 234-287; book has 3 entries per node (all three fixed together). */
int or_magnetostatics_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                               const double *current, const double *field, const uint32_t *book, uint32_t n_dof,
                               or_sparse *A, double *b);
/* RSolverFluid prepare (volume) + computeElementConstantDerivative + assemblyMatrix,
 * src/rsolverfluid.cpp:289-497, 1845-2201, 2399-2419; state = (vx, vy, vz, p) per node, v_old and
 * acc 3 per node, gravity 3 per element, book 4 per node. */
int or_fluid_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets,
                      const double *rho, const double *mu, const double *grav, const double *state,
                      const double *v_old, const double *acc, double stream_velocity, double alpha,
                      double dt, int unsteady, const uint32_t *book, uint32_t n_dof, or_sparse *A, double *b);
/* RSolverFluid::solve write-back (src/rsolverfluid.cpp:519-624). */
void or_fluid_update(uint32_t nn, const uint32_t *book, const double *x, int unsteady, double dt,
                     double *state, double *v_old, double *acc, double norms[4]);
/* RSolverFluidHeat / RSolverFluidParticle element + assemblyMatrix (src/rsolverfluidheat.cpp:167-299,
 * 669-880; src/rsolverfluidparticle.cpp:552-763); vel_e 3 per element; book 1 per node. */
int or_transport_assemble(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets, const double *ca_e,
                          const double *k_e, const double *vel_e, const double *q_node, const double *field,
                          double alpha, double dt, int unsteady, const uint32_t *book, uint32_t n_dof,
                          or_sparse *A, double *b);
void or_transport_update(uint32_t nn, const uint32_t *book, const double *x, double *field, double norms[2]);
/* RSolverRadiativeHeat::prepare / process with a loaded view-factor matrix,
 * src/rsolverradiativeheat.cpp:145-241, 276-303 (dense row-major F and A). */
void or_radiosity_assemble(uint32_t np, const double *F, const uint32_t *patch_ptr, const uint32_t *patch_elems,
                           const double *emissivity, const double *temperature, const double *ambient,
                           double *A, double *b);
void or_radiosity_process(uint32_t np, uint32_t ne, const uint32_t *patch_ptr, const uint32_t *patch_elems,
                          const double *elem_area, const double *patch_heat, double *elem_heat);
/* RSolverHeat::process (volume), src/rsolverheat.cpp:646-692: 3 flux values per tet. */
int or_heat_flux(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets, const double *k,
                 const double *t_node, double *flux);
/* RSolverElectrostatics::process (volume), src/rsolverelectrostatics.cpp:578-657. */
int or_electrostatics_process(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets, const double *eps_r,
                              const double *sigma, const double *phi, double *field, double *current, double *energy,
                              double *resistivity, double *joule);
/* RSolverStress::process (volume, static), src/rsolverstress.cpp:1238-1408. */
int or_stress_process(uint32_t nn, const double *xyz, uint32_t ne, const uint32_t *tets, const double *E,
                      const double *nu, const double *u_node, double *normal, double *shear, double *vm,
                      double *node_force);

/* REigenValueSolver::solve, src/reigenvaluesolver.cpp:41-592.  method 0 Lanczos, 1 Arnoldi,
 * 2 Rayleigh (K and M share rp/col).  d holds min(n_eigen, n) values (1 for Rayleigh); ev
 * (optional) n_ev_rows rows of n.  Start vectors from rand() / RAND_MAX. */
int or_eigen_solve(int method, unsigned n_eigen, unsigned n_iter, double eig_cvg, const or_conf *conf, unsigned n,
                   const uint32_t *rp, const uint32_t *col, const double *mval, const double *kval, double *d,
                   unsigned *n_d, double *ev, unsigned *n_ev_rows, unsigned long long *cg_its);

#ifdef __cplusplus
}
#endif
#endif
