/*
 * range_solver_amd.h -- C ABI of the MI355X (gfx950) element-assembly + sparse-solve path.
 *
 * Plain C: integers, doubles, raw pointers and opaque handles; no C++ or torch types cross
 * this boundary.  Every entry point returns an rs_status; rs_last_error() gives the text of
 * the last failure on the calling thread.  Non-convergence is not an error (it is reported
 * in rs_info, as the reference reports nothing and simply stops at nOuterIterations).
 *
 * Reference interfaces replaced (paths in Range-Software/range-solver-lib):
 *   RMatrixSolver            include/rmatrixsolver.h:9-52, src/rmatrixsolver.cpp:123-581
 *   RMatrixPreconditioner    include/rmatrixpreconditioner.h:7-64, src/rmatrixpreconditioner.cpp
 *   SparseMatrixCache        src/rmatrixsolver.cpp:14-109 (the flat CSR + ||A||_F)
 *   RIterationInfo           include/riterationinfo.h:9-91, src/riterationinfo.cpp:74-193
 *   RSolverGeneric node book src/rsolvergeneric.cpp:498-582; RSolverStress src/rsolverstress.cpp:1564-1652
 *   RSolverHeat assembly     src/rsolverheat.cpp:361-455 (element loop), :755-820 (assemblyMatrix)
 *   RSolverStress assembly   src/rsolverstress.cpp:384-715 (surface + volume), :1654-1772
 *   write-back               src/rsolverheat.cpp:473-480, src/rsolverstress.cpp:800-836
 */
#ifndef RANGE_SOLVER_AMD_H
#define RANGE_SOLVER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RS_EOD 0xFFFFFFFFu /* disabled position in a node book (RConstants::eod) */

typedef enum rs_status {
    RS_OK = 0,
    RS_ERR_INVALID_ARG = 1,
    RS_ERR_UNKNOWN_SOLVER = 2,      /* src/rmatrixsolver.cpp:194-195 */
    RS_ERR_INVALID_PRECOND = 3,     /* src/rmatrixpreconditioner.cpp:31-32, 73-74 */
    RS_ERR_DEVICE = 4,
    RS_ERR_NO_DEVICE = 5,
    RS_ERR_OUT_OF_MEMORY = 6,
    RS_ERR_NOT_INITIALIZED = 7,
    RS_ERR_UNSUPPORTED = 8
} rs_status;

/* Same values as RMatrixPreconditionerType, include/rmatrixpreconditioner.h:7-16;
 * LEARNED is new and follows BLOCK_JACOBI. */
typedef enum rs_precond_type {
    RS_PRECOND_NONE = 0,
    RS_PRECOND_JACOBI = 1,
    RS_PRECOND_BLOCK_JACOBI = 2,
    RS_PRECOND_LEARNED = 3
} rs_precond_type;

/* RMatrixSolverConf::Type (range-model-lib): the two types RMatrixSolver::solve dispatches,
 * src/rmatrixsolver.cpp:186-196. */
typedef enum rs_solver_type { RS_SOLVER_CG = 0, RS_SOLVER_GMRES = 1 } rs_solver_type;

/* Stopping rule.  REFERENCE is the reference's rule: error = ||r|| / (||A||_F ||x|| + ||b||)
 * (src/rmatrixsolver.cpp:333-344) fed to RIterationInfo::hasConverged.  RELRES uses
 * error = ||r|| / ||b|| with the same hasConverged tests. */
typedef enum rs_stop_rule { RS_STOP_REFERENCE = 0, RS_STOP_RELRES = 1 } rs_stop_rule;

/* FP64 is the drop-in.  MIXED: FP32 matrix values, FP32 SpMV and preconditioner, FP64 dot
 * products, norms and solution.  CG: FP32 vectors with FP64 residual replacement (RELRES
 * rule).  GMRES: the inner cycle's basis, preconditioned basis and w in FP32, its Gram-Schmidt
 * products, Hessenberg column and rotations in FP64; each restart's residual and the stopping
 * test in FP64 with the configured rule. */
typedef enum rs_precision { RS_PRECISION_FP64 = 0, RS_PRECISION_MIXED = 1 } rs_precision;

typedef enum rs_reason {
    RS_REASON_RUNNING = 0,
    RS_REASON_CONVERGED = 1,  /* error < cvg */
    RS_REASON_STAGNATED = 2,  /* |trend| < eps^2 */
    RS_REASON_NONNORMAL = 3,  /* error or trend not a normal number (0, NaN, Inf, subnormal) */
    RS_REASON_MAXIT = 4       /* nOuterIterations reached */
} rs_reason;

/* Mirrors RMatrixSolverConf as used by the reference (type, nOuterIterations,
 * nInnerIterations, solverCvgValue, outputFrequency, outputFileName). */
typedef struct rs_conf {
    int32_t solver_type;          /* rs_solver_type */
    uint32_t n_outer_iterations;
    uint32_t n_inner_iterations;  /* GMRES restart length */
    uint32_t output_frequency;    /* 0 = silent (RIterationInfo::printIteration) */
    double solver_cvg_value;
    int32_t stop_rule;            /* rs_stop_rule */
    int32_t precision;            /* rs_precision */
    const char *output_file_name; /* convergence file appended as the reference does; NULL = none */
} rs_conf;

typedef struct rs_info {
    uint32_t iterations;          /* CG iterations / GMRES outer iterations performed */
    uint32_t inner_iterations;    /* GMRES inner steps in total; MIXED CG: FP64 residual replacements */
    int32_t reason;               /* rs_reason */
    int32_t converged;            /* reason is CONVERGED */
    double error;                 /* last RIterationInfo error */
    double trend;                 /* last RIterationInfo trend */
    double equation_scale;        /* 1e9*||A||_F/||b|| (src/rmatrixsolver.cpp:166-171) */
    double a_norm;                /* ||A||_F */
    double b_norm;                /* ||b|| */
    double true_relres;           /* ||b - A x|| / ||b||, recomputed in FP64 after the solve */
    double solve_ms;              /* device time of the solve, host-visible stream events */
    uint32_t precond_applications;/* learned preconditioner: network evaluations */
    uint32_t fallback;            /* learned solve: 1 when it finished with Jacobi-CG */
} rs_info;

/* ---------------------------------------------------------------- runtime ------------ */
int rs_init(int device);          /* select the device, create the library stream */
int rs_finalize(void);
const char *rs_last_error(void);
int rs_device_count(int *count);
int rs_version(int *major, int *minor);

/* Device memory helpers for device-resident callers. */
int rs_dmalloc(void **ptr, size_t bytes);
int rs_dfree(void *ptr);
int rs_memcpy_htod(void *dst, const void *src, size_t bytes);
int rs_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int rs_memcpy_dtod(void *dst, const void *src, size_t bytes);
int rs_synchronize(void);

/* ---------------------------------------------------------------- matrix ------------- */
/* Device CSR: the SparseMatrixCache of src/rmatrixsolver.cpp:14-54 (uint32 rowPtr and
 * column indexes, double values, ||A||_F over stored entries), kept in HBM. */
typedef struct rs_matrix rs_matrix;
int rs_matrix_create(uint32_t n_rows, const uint32_t *row_ptr, const uint32_t *col,
                     const double *val, rs_matrix **out);
/* refreshSparseMatrixCache, src/rmatrixsolver.cpp:56-97: same pattern, new values. */
int rs_matrix_update_values(rs_matrix *A, const double *val);
int rs_matrix_update_values_device(rs_matrix *A, const double *d_val);
int rs_matrix_info(const rs_matrix *A, uint32_t *n_rows, uint64_t *nnz, double *frobenius);
int rs_matrix_download(const rs_matrix *A, uint32_t *row_ptr, uint32_t *col, double *val);
int rs_matrix_spmv(rs_matrix *A, const double *x, double *y);               /* host vectors */
int rs_matrix_spmv_device(rs_matrix *A, const double *d_x, double *d_y);
/* SpMV shapes: 0..4 = vector CSR with 4/8/16/32/64 lanes per row; 5, 6 = LDS-staged tile CSR
 * with <= 2048 / <= 4096 nonzeros per tile; 7 = SELL-64 copy (one row per lane); 8 = SELL-64 with
 * int16 column offsets from the slice's first row; 9, 10 = 7, 8 with 8 entries of a row per loop
 * trip.  The library times every shape once per matrix and precision (autotuner) and uses the
 * fastest. */
#define RS_SPMV_SHAPES 11
/* Measurement: average device time (ms, stream events) of `reps` back-to-back SpMV launches
 * (precision RS_PRECISION_FP64: FP64 values/vectors; RS_PRECISION_MIXED: FP32;
 * RS_TIMED_CG_PRODUCT: the FP64 CG iteration's fused product q = A p with its partial p.q sums,
 * as the solver launches it) of shape `plan` (-1 = the autotuned FP64 / FP32 shape). */
#define RS_TIMED_CG_PRODUCT 2
int rs_matrix_spmv_timed(rs_matrix *A, int32_t precision, uint32_t reps, int32_t plan,
                         double *avg_ms);
/* The autotuned shape for `precision` and (optional) the tuning time of every shape
 * (RS_SPMV_SHAPES doubles, ms). */
int rs_matrix_spmv_plan(rs_matrix *A, int32_t precision, int32_t *plan, double *plan_ms);
/* Override the autotuner for `precision` (plan -1: tune again on next use). */
int rs_matrix_force_plan(rs_matrix *A, int32_t precision, int32_t plan);
int rs_matrix_destroy(rs_matrix *A);

/* ---------------------------------------------------------------- preconditioner ----- */
/* RMatrixPreconditioner(const RSparseMatrix&, type, blockSize), include/rmatrixpreconditioner.h:36 */
typedef struct rs_precond rs_precond;
int rs_precond_create(const rs_matrix *A, int32_t type, uint32_t block_size, rs_precond **out);
/* RMatrixPreconditioner::compute(x, y), include/rmatrixpreconditioner.h:48 (host vectors) */
int rs_precond_compute(const rs_precond *P, const double *x, double *y);
int rs_precond_width(const rs_precond *P, uint32_t *width);
int rs_precond_download(const rs_precond *P, double *data);                 /* n*width values */
/* RMatrixPreconditioner(const RMatrixPreconditioner&) and operator=, include/rmatrixpreconditioner.h:39,45
 * (src/rmatrixpreconditioner.cpp:6-14,36-50): a new handle holding its own copy of the type, the
 * block width and the device data; later changes to either handle do not reach the other. */
int rs_precond_copy(const rs_precond *src, rs_precond **out);
int rs_precond_destroy(rs_precond *P);

/* ---------------------------------------------------------------- solver ------------- */
/* RMatrixSolver(const RMatrixSolverConf&), include/rmatrixsolver.h:27 */
typedef struct rs_solver rs_solver;
int rs_solver_create(const rs_conf *conf, rs_solver **out);
/* RMatrixSolver copy constructor / assignment, include/rmatrixsolver.h:30-36 */
int rs_solver_copy(const rs_solver *src, rs_solver **out);
/* RMatrixSolver::disableConvergenceLogFile, include/rmatrixsolver.h:42 */
int rs_solver_disable_convergence_log(rs_solver *S);
/* RMatrixSolver::solve(A, b, x, type, blockSize), include/rmatrixsolver.h:39.
 * b, x are host arrays of A's row count; x holds the initial guess and receives the
 * solution (the reference resizes x to n with zeros when it is shorter). */
int rs_solver_solve(rs_solver *S, rs_matrix *A, const double *b, double *x,
                    int32_t precond_type, uint32_t block_size, rs_info *info);
/* Same, with b and x already resident in device memory. */
int rs_solver_solve_device(rs_solver *S, rs_matrix *A, const double *d_b, double *d_x,
                           int32_t precond_type, uint32_t block_size, rs_info *info);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): rs_solver_solve_device for
 * FP64 CG (preconditioner NONE, JACOBI or BLOCK_JACOBI) that calls no hipMalloc/hipFree.  Its
 * workspace is one library-held device array of 134,217,728 bytes per GPU (held until the process
 * ends): 256-byte-aligned state, partial sums and history, plus 4 n + n*w doubles (w = the block
 * width, 1 for Jacobi, 0 for NONE), so up to ~3.3 M unknowns with Jacobi; larger systems are
 * refused with RS_ERR_INVALID_ARG.  Calls on one GPU share that array: calls that overlap on a
 * GPU race.  Same results as rs_solver_solve_device on the same SpMV shape (the matrix's tuned
 * shape if its layout is already built, else vector CSR; no tuning here).  Launches are eager (no
 * graph), no convergence log or history is written (rs_solver_history reports none), and GMRES
 * or mixed precision return RS_ERR_UNSUPPORTED.  Call rs_init first: it looks the array up. */
int rs_solver_solve_device_noalloc(rs_solver *S, rs_matrix *A, const double *d_b, double *d_x,
                                   int32_t precond_type, uint32_t block_size, rs_info *info);
/* (error, trend) per iteration of the last solve, 2*count doubles. */
int rs_solver_history(const rs_solver *S, double *out, uint32_t cap, uint32_t *count);
int rs_solver_destroy(rs_solver *S);

/* ---------------------------------------------------------------- model preparation -- */
/* Host-only (no device needed).  Structured box of nx*ny*nz hexahedral cells, each split
 * into 6 positively oriented tetrahedra around the cell diagonal (Kuhn split).
 * xyz: 3*(nx+1)*(ny+1)*(nz+1) doubles, node id = i + (nx+1)*(j + (ny+1)*k).
 * tets: 4*6*nx*ny*nz node ids, tet id = 6*cell + t, cell = ci + nx*(cj + ny*ck). */
int rs_mesh_kuhn_box(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz,
                     double *xyz, uint32_t *tets);
/* Triangular tet faces lying on the plane xyz[axis] == value (|.| <= tol).  Pass
 * tris == NULL to count only. */
int rs_mesh_plane_faces(uint32_t n_nodes, const double *xyz, uint32_t n_tets,
                        const uint32_t *tets, int32_t axis, double value, double tol,
                        uint32_t *tris, uint64_t cap, uint64_t *count);
/* Counter-based seeded fields (SplitMix64; Box-Muller for normals); bit-identical on any
 * host.  uniform: U(0,1); lognormal: exp(mu + sigma*N(0,1)). */
int rs_field_uniform(uint64_t seed, uint64_t n, double *out);
int rs_field_lognormal(uint64_t seed, uint64_t n, double mu, double sigma, double *out);
/* RSolverGeneric::generateNodeBook/rebuildNodeBook (dofs_per_node 1) and
 * RSolverStress::generateNodeBook (dofs_per_node 3) and RSolverFluid::generateNodeBook
 * (dofs_per_node 4, src/rsolverfluid.cpp:1086-1207): fixed[n_nodes*dofs_per_node] marks
 * prescribed positions; positions of nodes in no element are disabled too.  book receives
 * 0..n-1 in increasing position order or RS_EOD.  *n_enabled receives n. */
int rs_node_book(uint32_t n_nodes, uint32_t n_elems, uint32_t nodes_per_elem,
                 const uint32_t *conn, uint32_t dofs_per_node, const uint8_t *fixed,
                 uint32_t *book, uint32_t *n_enabled);

/* ---------------------------------------------------------------- assembly ----------- */
/* An assembly owns the device mesh, the node book, the CSR pattern of the free DOFs (the
 * matrix the reference builds with RSparseMatrix::addValue), the element -> CSR scatter map
 * and a node-disjoint colouring of the elements (each colour is one conflict-free launch,
 * so the result is deterministic without atomics). */
typedef struct rs_assembly rs_assembly;
int rs_assembly_create(uint32_t n_nodes, const double *xyz, uint32_t n_tets,
                       const uint32_t *tets, uint32_t dofs_per_node, const uint8_t *fixed,
                       rs_assembly **out);
int rs_assembly_info(const rs_assembly *S, uint32_t *n_dof, uint64_t *nnz,
                     uint32_t *n_colours);
int rs_assembly_book(const rs_assembly *S, uint32_t *book);
/* The assembly's device copies of the mesh coordinates (3 * n_nodes) and of the DOF book
 * (n_nodes * dofs_per_node), valid while S lives (input of rs_learned_create_by_bins). */
int rs_assembly_device_mesh(const rs_assembly *S, const double **d_xyz, const uint32_t **d_book,
                            uint32_t *n_nodes, uint32_t *dofs_per_node);
/* Device time (ms, stream events) of the last rs_assemble_* call, colour launches only. */
int rs_assembly_last_ms(const rs_assembly *S, double *ms);
/* The assembled matrix (owned by the assembly; do not destroy). */
int rs_assembly_matrix(rs_assembly *S, rs_matrix **A);
/* Heat (RSolverHeat::prepare + assemblyMatrix).  Per-element conduction k, volumetric heat
 * q, and (transient only) density*capacity rhoc; t_node = nodeTemperature (prescribed values
 * at fixed nodes, previous step elsewhere).  Host arrays; writes the matrix values and b
 * (n_dof doubles, host or device per the _device variant). */
int rs_assemble_heat(rs_assembly *S, const double *k, const double *q, const double *rhoc,
                     const double *t_node, int32_t transient, double alpha, double dt,
                     double *b);
/* Device-resident variant: all arrays are device pointers. */
int rs_assemble_heat_device(rs_assembly *S, const double *d_k, const double *d_q,
                            const double *d_rhoc, const double *d_t_node, int32_t transient,
                            double alpha, double dt, double *d_b);
/* Stress (RSolverStress::prepare + assemblyMatrix, static): per-element E, nu, rho;
 * gravity (3); u_node (3 per node: prescribed displacements); n_tri surface triangles
 * with a force per unit area each (3 doubles).  Host arrays. */
int rs_assemble_stress(rs_assembly *S, const double *E, const double *nu, const double *rho,
                       const double *gravity, const double *u_node, uint32_t n_tri,
                       const uint32_t *tris, const double *traction, double *b);
/* Assembly kernels: row_gather 1 (default) sums each free row's incident elements in one place
 * and writes the row whole -- heat and scalar transport: one thread per row (rows of at most 16
 * entries); static, unrotated stress: one wavefront per node (rows of at most 48 entries, at
 * most 32 incident elements); otherwise, and for the transient, modal or rotated stress
 * problems, the colour launches are used.  0 = one launch per node-disjoint colour with a
 * scatter.  Both add the matrix values in the same element order. */
int rs_assembly_set_method(rs_assembly *S, int32_t row_gather);
/* Write-back (src/rsolverheat.cpp:473-480): node[i*dpn+c] = x[book] where enabled. */
int rs_assembly_dofs_to_nodes(rs_assembly *S, const double *d_x, double *d_node);
/* Post-processing of the volume elements (device arrays, element outputs in the caller's element
 * order):
 * RSolverHeat::process, src/rsolverheat.cpp:646-692 -- 3 flux values per tet, q = -sum_m B_m k T_m
 * with B_m = grad N_m * detJ (the reference's integration of grad N detJ / nInp). */
int rs_heat_flux_device(rs_assembly *S, const double *d_k, const double *d_t_node, double *d_flux);
/* RSolverElectrostatics (src/rsolverelectrostatics.cpp): prepare's volume assembly (:300-345) --
 * Ke = B.B eps_r e0 V, fe = -rho V / 4, potential lifting -- and process (:578-657) -- per tet
 * the electric field (3), current density (3), energy, resistivity and Joule heat. */
int rs_assemble_electrostatics_device(rs_assembly *S, const double *d_eps_r, const double *d_rho,
                                      const double *d_potential, double *d_b);
int rs_electrostatics_process_device(rs_assembly *S, const double *d_eps_r, const double *d_sigma,
                                     const double *d_potential, double *d_field, double *d_current,
                                     double *d_energy, double *d_resistivity, double *d_joule);
/* RSolverStress problem kinds: static (Ae = Ke), transient (time solver on: Ae = Me + alpha dt Ke,
 * be = dt fe + (Me - (1 - alpha) dt Ke) u) and modal (R_PROBLEM_STRESS_MODAL: Ae = Ke, M = Me). */
enum { RS_STRESS_STATIC = 0, RS_STRESS_TRANSIENT = 1, RS_STRESS_MODAL = 2 };
typedef struct rs_stress_conf {
    int32_t problem;
    double alpha;               /* time-march approximation coefficient (transient) */
    double dt;                  /* current time-step size (transient) */
    double gravity[3];
} rs_stress_conf;
/* The full RSolverStress::prepare for tetrahedra (src/rsolverstress.cpp:54-724, 1654-1772): host
 * arrays as rs_assemble_stress, plus the thermal-expansion load (thermal_expansion and dT = element
 * temperature - environment temperature, per element; both NULL for none, :688-693), the consistent
 * mass (:665-680) for the transient and modal kinds, and the lifting with Ae + Be (:1719-1739).
 * u_node holds nodeDisplacement: the prescribed values and, when transient, the previous step. */
int rs_assemble_stress_ex(rs_assembly *S, const rs_stress_conf *conf, const double *E, const double *nu,
                          const double *rho, const double *thermal_expansion, const double *dT,
                          const double *u_node, uint32_t n_tri, const uint32_t *tris, const double *traction,
                          double *b);
/* The mass matrix of the last modal rs_assemble_stress_ex (owned by the assembly; same pattern as
 * rs_assembly_matrix); RSolverStress::solveEigenValue passes (M, A) to REigenValueSolver. */
int rs_assembly_mass_matrix(rs_assembly *S, rs_matrix **M);
/* Local rotations of a 3-DOF assembly (RSolverGeneric, src/rsolvergeneric.cpp:353-365;
 * RLocalRotation): R (9 doubles per node, row-major) for the nodes with active[node] != 0; NULL R
 * removes them.  rs_assemble_stress_ex then applies applyLocalRotations
 * (src/rsolverstress.cpp:1774-1833): Ae -> T^T Ae T, Be -> T^T Be T and be -> T be (as the
 * reference does) before the lifting; the modal M keeps the unrotated Me. */
int rs_assembly_set_rotations(rs_assembly *S, const double *R, const uint8_t *active);
/* RSolverStress::setDisplacement (src/rsolverstress.cpp:800-836): d_u_node (3 per node) = x on the
 * enabled components, 0 on the disabled ones, rotated back with iR^T (iR = R^-1) where a rotation is
 * active. */
int rs_stress_set_displacement_device(rs_assembly *S, const double *d_x, double *d_u_node);
/* RSolverMagnetostatics::prepare + assemblyMatrix (src/rsolvermagnetostatics.cpp:44-145, 234-287)
 * on a 3-DOF assembly whose nodes are fixed with all three components: per component the
 * negative Laplacian, source e0 V/4 (grad N_m x J_m) from the node current density (3 per node),
 * lifting of the prescribed field (3 per node). */
int rs_assemble_magnetostatics_device(rs_assembly *S, const double *d_current_node, const double *d_field_node,
                                      double *d_b);
/* RSolverFluid time-step parameters: streamVelocity (computeStreamVelocity, src/rsolverfluid.cpp:
 * 2453-2536, > 0), the time-march coefficient alpha and step dt, and whether the time solver is
 * enabled. */
typedef struct rs_fluid_conf {
    double stream_velocity;
    double alpha;
    double dt;
    int32_t unsteady;
} rs_fluid_conf;
/* RSolverFluid::prepare, volume elements (src/rsolverfluid.cpp:289-497, 1845-2201, 2399-2419) on a
 * 4-DOF assembly (positions 4 node + (vx, vy, vz, p)): the stabilised (SUPG/PSPG/LSIC) linear
 * tetrahedron for the increments, enabled positions only.  rho, mu: per element; gravity: 3 per
 * element; state: 4 per node (vx, vy, vz, p); v_old, acc: 3 per node (acc read only when
 * unsteady).  Surface free-pressure elements (:392-428) and local rotations are not applied. */
int rs_assemble_fluid_device(rs_assembly *S, const rs_fluid_conf *conf, const double *d_rho, const double *d_mu,
                             const double *d_gravity, const double *d_state, const double *d_v_old,
                             const double *d_acc, double *d_b);
/* RSolverFluid::solve after the linear solve (src/rsolverfluid.cpp:519-624): state += x on the
 * enabled positions; steady: v_old = the velocities before the update; unsteady: acc = (v - v_old)
 * / dt.  norms (host, 4, may be NULL): |v| before, |v| after, |p| before, |p| after. */
int rs_fluid_update_device(rs_assembly *S, const double *d_x, int32_t unsteady, double dt, double *d_state,
                           double *d_v_old, double *d_acc, double *norms);
/* RSolverFluidHeat / RSolverFluidParticle time-step parameters (time-march alpha, dt, time solver on). */
typedef struct rs_transport_conf {
    double alpha;
    double dt;
    int32_t unsteady;
} rs_transport_conf;
/* RSolverFluidHeat::prepare / RSolverFluidParticle::prepare, volume elements
 * (src/rsolverfluidheat.cpp:167-299, 669-880; src/rsolverfluidparticle.cpp:121-253, 552-763) on a
 * 1-DOF assembly: SUPG advection-diffusion of the node field (temperature / concentration) with
 * per-element capacity ca (rho c; 1 for particles), diffusivity k (conduction / diffusion) and
 * velocity (3 per element, the element vector of the node velocities), node source (nodeHeat /
 * nodeRate), lifting of the fixed nodes with their field values.  The solve is GMRES + Jacobi. */
int rs_assemble_transport_device(rs_assembly *S, const rs_transport_conf *conf, const double *d_ca,
                                 const double *d_k, const double *d_velocity, const double *d_node_source,
                                 const double *d_node_field, double *d_b);
/* Write-back (src/rsolverfluidheat.cpp:320-344, src/rsolverfluidparticle.cpp:274-298): field =
 * max(x, 0) on the enabled nodes; norms (host, 2, may be NULL): |field| before and after. */
int rs_transport_update_device(rs_assembly *S, const double *d_x, double *d_node_field, double *norms);

/* ---------------------------------------------------------------- radiative heat
 * RSolverRadiativeHeat (src/rsolverradiativeheat.cpp) with the view-factor matrix loaded (from its
 * file or computed by RHemiCube).  A radiosity handle holds the patch book (patch p owns the
 * surface elements patch_elems[patch_ptr[p] .. patch_ptr[p+1]), RPatchBook), the element areas
 * (tris: 3 nodes per surface element) and the dense n_patches x n_patches system matrix. */
typedef struct rs_radiosity rs_radiosity;
int rs_radiosity_create(uint32_t n_patches, uint32_t n_nodes, const double *xyz, uint32_t n_elems,
                        const uint32_t *tris, const uint32_t *patch_ptr, const uint32_t *patch_elems,
                        rs_radiosity **out);
/* The system matrix (owned by the handle; do not destroy). */
int rs_radiosity_matrix(rs_radiosity *R, rs_matrix **A);
/* RSolverRadiativeHeat::prepare (:145-241): patch means of the element emissivity, temperature and
 * ambient temperature (n_elems each), then A_ij = d_ij/E_j - F_ij (1 - E_j)/E_j (0 if E_j = 0) and
 * b_i = -sum_j (d_ij - F_ij) sigma T_j^4 + (1 - sum_j F_ij) sigma Tamb_i^4 from the dense row-major
 * view factors (n_patches^2).  The solve is rs_solver_solve_device with CG + Jacobi (:243-266). */
int rs_assemble_radiosity_device(rs_radiosity *R, const double *d_view_factors, const double *d_emissivity,
                                 const double *d_temperature, const double *d_ambient, double *d_b);
/* RSolverRadiativeHeat::process (:268-303): element heat = patch heat * element area / patch area
 * (n_elems outputs; elements in no patch get 0). */
int rs_radiosity_process_device(rs_radiosity *R, const double *d_patch_heat, double *d_elem_heat);
/* Device time (ms) of the last rs_assemble_radiosity_device. */
int rs_radiosity_last_ms(const rs_radiosity *R, double *ms);
int rs_radiosity_destroy(rs_radiosity *R);

/* RSolverStress::process, src/rsolverstress.cpp:1238-1408, static case without thermal load --
 * per tet the normal and shear stress measures and their sum (the reference's von Mises), and the
 * node forces sum_e Ke u_e (3 per node; reactions on the fixed nodes). */
int rs_stress_process_device(rs_assembly *S, const double *d_E, const double *d_nu, const double *d_u_node,
                             double *d_normal, double *d_shear, double *d_von_mises, double *d_node_force);
int rs_assembly_destroy(rs_assembly *S);

/* ---------------------------------------------------------------- eigenvalues
 * REigenValueSolver (include/reigenvaluesolver.h:9-63, src/reigenvaluesolver.cpp:41-592): the
 * generalized problem K x = lambda M x through K^-1 M (every inner solve is CG with Jacobi on the
 * device, configured by the rs_conf as RMatrixSolverConf), methods as
 * REigenValueSolverConf::Method.  LANCZOS returns ne values and no vectors, ARNOLDI ne values
 * and ne vector rows of n, RAYLEIGH one value (1/lambda of the dominant mode of K^-1 M, as the
 * reference returns it) and one vector.  With more than one value they are post-processed as
 * the reference does: d = |1/d|, d[0] and d[1] swapped, sorted ascending, vector rows
 * reordered by its swap loop.  The start vector (and Rayleigh's start shift) come from
 * std::rand() / RAND_MAX like the reference's, so srand() by the caller fixes them.  The
 * reference's QL step returns the Lanczos diagonal unrotated (its split search exits at once,
 * src/reigenvaluesolver.cpp:440-452); that is reproduced. */
typedef enum rs_eigen_method { RS_EIGEN_LANCZOS = 0, RS_EIGEN_ARNOLDI = 1, RS_EIGEN_RAYLEIGH = 2 } rs_eigen_method;
typedef struct rs_eigen_conf {
    int32_t method;               /* rs_eigen_method */
    uint32_t n_eigen_values;      /* REigenValueSolverConf::nEigenValues */
    uint32_t n_iterations;        /* QR iterations (Arnoldi) / Rayleigh iterations */
    double solver_cvg_value;      /* QR subdiagonal tolerance / Rayleigh |mu - mu_old| */
} rs_eigen_conf;
typedef struct rs_eigen_info {
    uint32_t outer_iterations;    /* Lanczos/Arnoldi steps or Rayleigh iterations */
    uint64_t cg_iterations;       /* inner CG iterations in total */
    double solve_ms;              /* device time of the iterations */
} rs_eigen_info;
int rs_eigen_solve(const rs_eigen_conf *econf, const rs_conf *mconf, rs_matrix *M, rs_matrix *K, double *d,
                   uint32_t d_capacity, uint32_t *n_d, double *ev, uint64_t ev_capacity, uint32_t *n_ev_rows,
                   rs_eigen_info *info);
/* rs_eigen_solve with every inner solve with K (all of Lanczos's and Arnoldi's, Rayleigh's
 * K d = f) run as rs_solver_solve_learned_device with the multilevel preconditioner L, which
 * must have been created for K (mconf: CG; its precision and stopping rule as for that call).
 * Rayleigh's solves with M + mu K and with M stay Jacobi-CG.  No counterpart in the reference,
 * whose inner solves are all Jacobi-CG (src/reigenvaluesolver.cpp:166,250,357,365,392): the
 * eigenvalues agree with rs_eigen_solve's to the inner solves' tolerance. */
typedef struct rs_learned rs_learned;
int rs_eigen_solve_learned(const rs_eigen_conf *econf, const rs_conf *mconf, rs_matrix *M, rs_matrix *K,
                           rs_learned *L, double *d, uint32_t d_capacity, uint32_t *n_d, double *ev,
                           uint64_t ev_capacity, uint32_t *n_ev_rows, rs_eigen_info *info);

/* ---------------------------------------------------------------- batched systems
 * Many independent systems solved together (no counterpart in the reference, which solves
 * one system per RMatrixSolver::solve call, src/rmatrixsolver.cpp:158-201).  Each system runs
 * `steps` implicit-Euler steps of Jacobi-CG as RSolverHeat does per time step
 * (src/rsolverheat.cpp:83-88 clears x, :757-800 forms the step's system, :458-480 solves and
 * writes back): step k solves A x_k = c + M x_{k-1} (x_{-1} = the initial solution), where A
 * is the step matrix (Me + dt Ke with alpha = 1), M the free-free mass matrix (omit it for a
 * steady system: b = c every step) and c the right-hand side assembled with the free
 * temperatures at zero.  Every step starts from x = 0 (warm_start = 0, as the reference) or
 * from the previous step's solution.  A step ends when the FP64 ||b - A x|| / ||b|| is below
 * solver_cvg_value (conf->stop_rule must be RS_STOP_RELRES, solver_type RS_SOLVER_CG).
 * MIXED: FP32 matrix, vectors and preconditioner inside, FP64 solution, dot products and
 * residual replacement.  All systems share each launch; the SpMV shape is autotuned per batch
 * (shape ids: 0..2 one lane per row with 4/8/16 waves per workgroup, 3 two lanes per row,
 * 4 four lanes per row). */
typedef struct rs_batch rs_batch;
typedef struct rs_batch_info {
    uint32_t rounds;              /* rounds of the batched loop (one CG iteration of every iterating system) */
    uint32_t max_iterations;      /* largest sum over one system's steps */
    uint64_t total_iterations;    /* sum over systems and steps */
    uint32_t systems_converged;   /* systems whose every step converged */
    uint32_t replacements;        /* FP64 residual replacements, all systems */
    double max_relres;            /* largest true relative residual at the end of a step */
    double solve_ms;              /* device time of the batched loop */
} rs_batch_info;
int rs_batch_create(uint32_t count, const uint32_t *n_rows, const uint64_t *nnz, rs_batch **out);
int rs_batch_set_matrix(rs_batch *B, uint32_t i, const rs_matrix *A);      /* copies A (device) */
int rs_batch_set_mass(rs_batch *B, uint32_t i, const rs_matrix *M);        /* same pattern as A */
int rs_batch_set_rhs(rs_batch *B, uint32_t i, const double *d_c, const double *d_x0); /* d_x0 NULL: 0 */
int rs_batch_solve(rs_batch *B, const rs_conf *conf, uint32_t steps, int32_t warm_start, rs_batch_info *info);
/* per system and step (count*steps, row-major): CG iterations and the final true relative
 * residual; reasons: rs_reason per system (count) */
int rs_batch_stats(const rs_batch *B, uint32_t *iterations, double *relres, int32_t *reasons);
int rs_batch_solution(const rs_batch *B, uint32_t i, double *d_x);       /* last step's x */
int rs_batch_shape(rs_batch *B, int32_t precision, int32_t *shape, double *shape_ms);
int rs_batch_force_shape(rs_batch *B, int32_t precision, int32_t shape);  /* -1: autotune */
/* column_bytes: 2 when every column fits an int16 offset from its 64-row slice (mesh
 * matrices), else 4 */
int rs_batch_info_layout(const rs_batch *B, uint32_t *padded_rows, uint64_t *nnz, uint64_t *sell_entries,
                         int32_t *column_bytes);
/* MIXED: an FP64 residual replacement whenever the recursive residual fell below ratio times
 * the last replaced one (default 0.01; 0 = only when it claims convergence) */
int rs_batch_set_replacement(rs_batch *B, double ratio);
int rs_batch_destroy(rs_batch *B);

/* ---------------------------------------------------------------- learned preconditioner
 * No counterpart in the reference (a new preconditioner type, RS_PRECOND_LEARNED): one V-cycle
 * over an aggregation hierarchy (piecewise-constant or rigid-body-mode prolongations, Galerkin
 * operators P^T A P, Chebyshev smoothing in FP32) whose coarsest level (<= 512 rows) is solved
 * by a 6-layer, 64-wide message-passing network, coarse Jacobi or its exact inverse, applied
 * inside flexible CG (Krylov vectors in FP32 or FP64) with Jacobi-CG as the fallback. */
enum {
    RS_COARSE_NETWORK = 0,   /* the learned network */
    RS_COARSE_JACOBI = 1,    /* coarse Jacobi: what the untrained network computes */
    RS_COARSE_EXACT = 2      /* dense inverse of the coarsest operator */
};
/* How the bins group the DOFs of a node (dofs_per_node > 1). */
enum {
    RS_AGG_NODES = 0,       /* the DOFs of a node share their bin's aggregate */
    RS_AGG_COMPONENTS = 1,  /* one aggregate per bin and component: the unknown-based
                               aggregation for vector problems (elasticity), whose coarse
                               spaces keep the translations of each component */
    RS_AGG_RIGID = 2        /* rs_learned_create_by_bins_ex only, 3 DOFs per node: node
                               aggregates whose coarse spaces are their six rigid-body modes
                               (translations and rotations), orthonormalised per aggregate; 6
                               coarse rows per aggregate, coarsest at most max_coarse rows */
};
/* Aggregates from node coordinates: nbx*nby*nbz bins over the bounding box; agg_of_dof gets
 * the bin of each enabled position of the node book (ids compacted to 0..n_agg-1). */
int rs_aggregates_by_bins(uint32_t n_nodes, const double *xyz, uint32_t dofs_per_node,
                          const uint32_t *book, uint32_t nbx, uint32_t nby, uint32_t nbz,
                          uint32_t *agg_of_dof, uint32_t *n_agg);
/* Nested coordinate bins: level 1 uses nbx*nby*nbz bins, every further level halves them,
 * until a level has at most max_coarse nodes.  agg receives the level maps one after the
 * other (level k maps the nodes of level k-1, level 0 being the enabled positions, to those
 * of level k; capacity agg_capacity entries, 2 * n_enabled always suffices), n_agg[k-1] the
 * node count of level k, *n_levels the number of coarse levels (<= max_levels). */
int rs_hierarchy_by_bins(uint32_t n_nodes, const double *xyz, uint32_t dofs_per_node,
                         const uint32_t *book, uint32_t nbx, uint32_t nby, uint32_t nbz,
                         uint32_t max_coarse, uint32_t max_levels, uint32_t *agg,
                         uint64_t agg_capacity, uint32_t *n_agg, uint32_t *n_levels);
/* The same with a choice of aggregation (RS_AGG_NODES as above, or RS_AGG_COMPONENTS: bin b
 * and component c form aggregate key dofs_per_node * b + c; every coarse node keeps the
 * component of its members). */
int rs_hierarchy_by_bins_ex(uint32_t n_nodes, const double *xyz, uint32_t dofs_per_node,
                            const uint32_t *book, uint32_t nbx, uint32_t nby, uint32_t nbz,
                            uint32_t max_coarse, uint32_t max_levels, int32_t aggregation,
                            uint32_t *agg, uint64_t agg_capacity, uint32_t *n_agg,
                            uint32_t *n_levels);
/* Hierarchy for the pattern of A: n_levels coarse levels given by their maps (as above). */
int rs_learned_create(const rs_matrix *A, uint32_t n_levels, const uint32_t *agg,
                      const uint32_t *n_agg, rs_learned **out);
/* The same hierarchy built on the GPU from the mesh in device memory: the nested coordinate
 * bins of rs_hierarchy_by_bins (identical level maps) and every list of rs_learned_create.
 * d_xyz: 3 * n_nodes coordinates; d_book: n_nodes * dofs_per_node positions, numbering the
 * rows of A (0 .. n-1) or RS_EOD.  max_coarse <= 512, max_levels <= 16. */
int rs_learned_create_by_bins(const rs_matrix *A, uint32_t n_nodes, const double *d_xyz,
                              uint32_t dofs_per_node, const uint32_t *d_book, uint32_t nbx,
                              uint32_t nby, uint32_t nbz, uint32_t max_coarse,
                              uint32_t max_levels, rs_learned **out);
/* The same with a choice of aggregation: RS_AGG_NODES / RS_AGG_COMPONENTS give the level maps of
 * rs_hierarchy_by_bins_ex; RS_AGG_RIGID (dofs_per_node 3, max_coarse >= 6) builds the rigid-body
 * hierarchy on the node aggregates (at most max_coarse / 6 of them on the coarsest level; at most
 * 256 rows per aggregate), whose level maps send each row to its aggregate (row 6 J + a of a
 * coarser level belongs to the aggregate of node J). */
int rs_learned_create_by_bins_ex(const rs_matrix *A, uint32_t n_nodes, const double *d_xyz,
                                 uint32_t dofs_per_node, const uint32_t *d_book, uint32_t nbx,
                                 uint32_t nby, uint32_t nbz, uint32_t max_coarse,
                                 uint32_t max_levels, int32_t aggregation, rs_learned **out);
/* Rigid-body levels (RS_AGG_RIGID): the orthonormal basis rows of level `level` (6 doubles per
 * row: row i of aggregate J holds row i of Q_J, the coarse rows 6 J .. 6 J + 5). */
int rs_learned_level_basis(const rs_learned *L, uint32_t level, double *q, uint64_t capacity);
/* Level map of level `level` (0 .. n_levels-2) to the next: one entry per row of the level. */
int rs_learned_level_map(const rs_learned *L, uint32_t level, uint32_t *agg, uint64_t capacity);
/* Network parameters, FP32, 49,921 values (layout: W_in[4][64], b_in[64], 6 x {W_s[64][64],
 * W_n[64][64], b[64]}, w_out[64], b_out). */
int rs_learned_set_params(rs_learned *L, const float *params, uint64_t count);
/* Smoothing steps before (0..4) and after (1..4) the coarse correction, coarsest solver. */
int rs_learned_set_cycle(rs_learned *L, uint32_t pre, uint32_t post, int coarse_mode);
/* Smoothing steps on the levels below the finest (rs_learned_set_cycle sets both to its own
 * pre/post; this call then changes the coarser levels' only). */
int rs_learned_set_coarse_cycle(rs_learned *L, uint32_t pre, uint32_t post);
/* Levels with at least min_rows rows smooth with a compressed copy of D^-1 A (FP16 values,
 * int16 column offsets in the SELL-64 slots; 4 bytes per entry); 0 keeps every level on its
 * FP32 values.  Levels whose offsets do not fit in 16 bits stay FP32.  Default: levels of at
 * least 65536 rows use the copy where a residual pass with it timed at least 10% faster than with
 * the level's FP32 plan (decided once per hierarchy); after this call min_rows is followed as given. */
int rs_learned_set_compressed(rs_learned *L, uint32_t min_rows);
/* Levels including the finest; level_rows / level_nnz hold n_levels entries when given. */
int rs_learned_info(const rs_learned *L, uint32_t *n_levels, uint32_t *level_rows,
                    uint64_t *level_nnz, uint64_t *n_params);
/* Level operators, diagonals and smoother bounds for the current values of A (done by the
 * solve as well). */
int rs_learned_setup(rs_learned *L, rs_matrix *A);
int rs_learned_level_lambda(const rs_learned *L, double *lambda);        /* n_levels bounds */
int rs_learned_level_matrix(const rs_learned *L, uint32_t level, uint32_t *row_ptr,
                            uint32_t *col, double *val);
int rs_learned_coarse_matrix(const rs_learned *L, double *dense);         /* n_coarse^2 */
int rs_learned_restrict_device(rs_learned *L, const double *d_r, float *rc_host);
/* The configured coarsest solver (network, coarse Jacobi or exact inverse) on rc (n_coarse values). */
int rs_learned_apply_coarse(rs_learned *L, const float *rc_host, float *e_host);
/* Average duration (ms) of the network inference kernel over reps back-to-back launches. */
int rs_learned_infer_timed(rs_learned *L, uint32_t reps, double *avg_ms);
/* One V-cycle z = B r (host vectors of length n). */
int rs_learned_apply(rs_learned *L, rs_matrix *A, const float *r_host, float *z_host);
/* Training data: the coarsest right-hand side of each of the first `capacity` iterations of
 * the following solves is kept on the device; rs_learned_records copies `count` of them. */
int rs_learned_record(rs_learned *L, uint32_t capacity);
int rs_learned_records(const rs_learned *L, uint32_t count, float *out);
int rs_learned_destroy(rs_learned *L);
/* Flexible CG with the learned preconditioner (solver conf: CG; b, x device arrays).  The
 * cycle runs in FP32 either way; conf precision RS_PRECISION_MIXED keeps the Krylov vectors
 * and product in FP32 with FP64 residual replacement (stopping rule: relres), RS_PRECISION_FP64
 * keeps them in FP64 (the configured stopping rule). */
int rs_solver_solve_learned_device(rs_solver *S, rs_matrix *A, rs_learned *L, const double *d_b,
                                   double *d_x, rs_info *info);

/* Training of the network on the GPU (forward, backward and Adam kernels).  A batch holds B
 * samples on one coarse pattern (cptr, ccol: n <= 512 rows, nnz entries, structurally
 * symmetric with its diagonal): operator values cval[B][nnz], coarse residuals rc[B][n] and
 * s[B] = rc . A^-1 . rc. */
typedef struct rs_trainer rs_trainer;
int rs_trainer_create(const float *params, uint64_t count, uint32_t max_batch, uint32_t max_nodes,
                      uint64_t max_nnz, rs_trainer **out);
int rs_trainer_gradient(rs_trainer *T, uint32_t n, uint64_t nnz, const uint32_t *cptr,
                        const uint32_t *ccol, uint32_t B, const float *cval, const float *rc,
                        const float *s, float *loss_out, float *grad_out);
int rs_trainer_step(rs_trainer *T, uint32_t n, uint64_t nnz, const uint32_t *cptr,
                    const uint32_t *ccol, uint32_t B, const float *cval, const float *rc,
                    const float *s, float lr, float *mean_loss);
int rs_trainer_get_params(const rs_trainer *T, float *params, uint64_t count);
int rs_trainer_destroy(rs_trainer *T);

#ifdef __cplusplus
}
#endif
#endif
