"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

EOD = 0xFFFFFFFF
PRECOND_NONE, PRECOND_JACOBI, PRECOND_BLOCK_JACOBI = 0, 1, 2
SOLVER_CG, SOLVER_GMRES = 0, 1
STOP_REFERENCE, STOP_RELRES = 0, 1
EPS = 1.0e-14


class OrConf(C.Structure):
    _fields_ = [("solver_type", C.c_int), ("n_outer", C.c_uint), ("n_inner", C.c_uint),
                ("cvg", C.c_double), ("stop_rule", C.c_int)]


class OrInfo(C.Structure):
    _fields_ = [("iterations", C.c_uint), ("inner_total", C.c_uint), ("error", C.c_double),
                ("trend", C.c_double), ("equation_scale", C.c_double), ("a_norm", C.c_double),
                ("b_norm", C.c_double), ("reason", C.c_int)]


_lib = None
_P = C.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    lib.or_has_converged.argtypes = [C.c_double, C.c_double, C.c_double]
    lib.or_has_converged.restype = C.c_int
    lib.or_converged_reason.argtypes = [C.c_double, C.c_double, C.c_double]
    lib.or_converged_reason.restype = C.c_int
    lib.or_precond_build.argtypes = [C.c_uint, _P, _P, _P, C.c_int, C.c_uint, _P, C.POINTER(C.c_uint)]
    lib.or_precond_apply.argtypes = [C.c_int, C.c_uint, C.c_uint, _P, _P, _P]
    lib.or_precond_apply.restype = None
    lib.or_frobenius.argtypes = [C.c_uint, _P, _P]
    lib.or_frobenius.restype = C.c_double
    lib.or_solve.argtypes = [C.c_uint, _P, _P, _P, _P, _P, C.c_int, C.c_uint, C.POINTER(OrConf),
                             C.c_double, C.POINTER(OrInfo), _P, C.c_uint]
    lib.or_spmv.argtypes = [C.c_uint, _P, _P, _P, _P, _P]
    lib.or_electrostatics_process.argtypes = [C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    lib.or_magnetostatics_assemble.argtypes = [C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P, C.c_uint32, _P, _P]
    lib.or_fluid_assemble.argtypes = [C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P, _P, _P, _P, C.c_double, C.c_double,
                                      C.c_double, C.c_int, _P, C.c_uint32, _P, _P]
    lib.or_fluid_update.argtypes = [C.c_uint32, _P, _P, C.c_int, C.c_double, _P, _P, _P, _P]
    lib.or_fluid_update.restype = None
    lib.or_radiosity_assemble.argtypes = [C.c_uint32, _P, _P, _P, _P, _P, _P, _P, _P]
    lib.or_radiosity_assemble.restype = None
    lib.or_radiosity_process.argtypes = [C.c_uint32, C.c_uint32, _P, _P, _P, _P, _P]
    lib.or_radiosity_process.restype = None
    lib.or_transport_assemble.argtypes = [C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P, _P, _P, C.c_double, C.c_double,
                                          C.c_int, _P, C.c_uint32, _P, _P]
    lib.or_transport_update.argtypes = [C.c_uint32, _P, _P, _P, _P]
    lib.or_transport_update.restype = None
    lib.or_heat_flux.argtypes = [C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P]
    lib.or_stress_process.argtypes = [C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P, _P, _P, _P, _P]
    lib.or_eigen_solve.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_double, C.POINTER(OrConf), C.c_uint, _P, _P,
                                   _P, _P, _P, C.POINTER(C.c_uint), _P, C.POINTER(C.c_uint),
                                   C.POINTER(C.c_ulonglong)]
    lib.or_spmv.restype = None
    lib.or_node_book.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, _P, C.c_uint, _P, _P]
    lib.or_node_book.restype = C.c_uint32
    lib.or_sparse_create.argtypes = [C.c_uint32]
    lib.or_sparse_create.restype = _P
    lib.or_sparse_destroy.argtypes = [_P]
    lib.or_sparse_destroy.restype = None
    lib.or_sparse_nnz.argtypes = [_P]
    lib.or_sparse_nnz.restype = C.c_uint64
    lib.or_sparse_to_csr.argtypes = [_P, _P, _P, _P]
    lib.or_sparse_to_csr.restype = None
    lib.or_heat_assemble.argtypes = [C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P, _P, _P, C.c_uint32,
                                     C.c_int, C.c_double, C.c_double, _P, _P]
    lib.or_stress_assemble.argtypes = [C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P, _P, _P,
                                       C.c_uint32, _P, _P, _P, C.c_uint32, _P, _P]
    lib.or_stress_assemble_ex.argtypes = [C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P, _P, _P, _P, _P, C.c_uint32, _P,
                                          _P, C.c_int, C.c_double, C.c_double, _P, _P, _P, C.c_uint32, _P, _P, _P]
    lib.or_stress_set_displacement.argtypes = [C.c_uint32, _P, _P, _P, _P, _P]
    lib.or_stress_set_displacement.restype = None
    _lib = lib
    return lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def _csr(m):
    rp = np.ascontiguousarray(m.indptr, np.uint32)
    col = np.ascontiguousarray(m.indices, np.uint32)
    val = np.ascontiguousarray(m.data, np.float64)
    return rp, col, val


def has_converged(error: float, trend: float, cvg: float) -> bool:
    return bool(load().or_has_converged(error, trend, cvg))


def converged_reason(error: float, trend: float, cvg: float) -> int:
    return int(load().or_converged_reason(error, trend, cvg))


def frobenius(m) -> float:
    rp, _, val = _csr(m.tocsr())
    return load().or_frobenius(m.shape[0], _p(rp), _p(val))


def spmv(m, x) -> np.ndarray:
    rp, col, val = _csr(m.tocsr())
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(m.shape[0])
    load().or_spmv(m.shape[0], _p(rp), _p(col), _p(val), _p(x), _p(y))
    return y


def precond_build(m, ptype: int, block_size: int = 1):
    m = m.tocsr()
    n = m.shape[0]
    rp, col, val = _csr(m)
    data = np.zeros(max(1, n * max(1, block_size)))
    w = C.c_uint()
    rc = load().or_precond_build(n, _p(rp), _p(col), _p(val), ptype, block_size, _p(data), C.byref(w))
    if rc != 0:
        raise ValueError("invalid preconditioner")
    return data[: n * max(1, w.value)], w.value


def precond_apply(ptype: int, n: int, width: int, data, x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(n)
    data = np.ascontiguousarray(data, np.float64)
    load().or_precond_apply(ptype, n, width, _p(data), _p(x), _p(y))
    return y


def solve(m, b, x0=None, precond=PRECOND_NONE, block_size=1, solver=SOLVER_CG, n_outer=1000,
          n_inner=1000, cvg=1e-13, stop_rule=STOP_REFERENCE, prev_error=0.0, history=False):
    """RMatrixSolver::solve restated (src/rmatrixsolver.cpp:158-581).  Returns (x, info[, hist])."""
    m = m.tocsr()
    n = m.shape[0]
    rp, col, val = _csr(m)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64, copy=True)
    conf = OrConf(solver, n_outer, n_inner, cvg, stop_rule)
    info = OrInfo()
    hist = np.zeros(2 * max(1, n_outer)) if history else None
    rc = load().or_solve(n, _p(rp), _p(col), _p(val), _p(b), _p(x), precond, block_size,
                         C.byref(conf), prev_error, C.byref(info), _p(hist),
                         n_outer if history else 0)
    if rc != 0:
        raise ValueError(f"oracle solve failed ({rc})")
    d = {k: getattr(info, k) for k, _ in info._fields_}
    if history:
        return x, d, hist[: 2 * info.iterations].reshape(-1, 2)
    return x, d


def node_book(n_nodes: int, conn, dofs_per_node: int, fixed):
    conn = np.ascontiguousarray(conn, np.uint32)
    npe = conn.shape[1]
    fixed = np.ascontiguousarray(fixed, np.uint8).reshape(-1)
    book = np.empty(n_nodes * dofs_per_node, np.uint32)
    n = load().or_node_book(n_nodes, conn.shape[0], npe, _p(conn), dofs_per_node, _p(fixed), _p(book))
    return book, int(n)


def _export(A, n_dof):
    import scipy.sparse as sp
    lib = load()
    nnz = lib.or_sparse_nnz(A)
    rp = np.empty(n_dof + 1, np.uint32)
    col = np.empty(nnz, np.uint32)
    val = np.empty(nnz, np.float64)
    lib.or_sparse_to_csr(A, _p(rp), _p(col), _p(val))
    lib.or_sparse_destroy(A)
    return sp.csr_matrix((val, col.astype(np.int64), rp.astype(np.int64)), shape=(n_dof, n_dof))


def heat_assemble(xyz, tets, k, q, t_node, book, n_dof, rhoc=None, transient=False, alpha=1.0,
                  dt=1.0):
    """RSolverHeat prepare + assemblyMatrix restated; returns (A as scipy CSR, b)."""
    lib = load()
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
    k, q, t_node = (np.ascontiguousarray(a, np.float64) for a in (k, q, t_node))
    rc = np.ascontiguousarray(rhoc, np.float64) if rhoc is not None else None
    book = np.ascontiguousarray(book, np.uint32)
    A = lib.or_sparse_create(n_dof)
    b = np.zeros(max(1, n_dof))
    lib.or_heat_assemble(xyz.size // 3, _p(xyz), tets.size // 4, _p(tets), _p(k), _p(q), _p(rc),
                         _p(t_node), _p(book), n_dof, int(bool(transient)), alpha, dt, A, _p(b))
    return _export(A, n_dof), b[:n_dof]


def stress_assemble(xyz, tets, E, nu, rho, gravity, u_node, tris, traction, book, n_dof):
    """RSolverStress prepare + assemblyMatrix (static) restated; returns (A, b)."""
    lib = load()
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
    E, nu, rho, u_node = (np.ascontiguousarray(a, np.float64) for a in (E, nu, rho, u_node))
    g = np.ascontiguousarray(gravity, np.float64)
    ntri = 0 if tris is None else int(np.asarray(tris).size // 3)
    tr = np.ascontiguousarray(tris, np.uint32).reshape(-1) if ntri else None
    tt = np.ascontiguousarray(traction, np.float64).reshape(-1) if ntri else None
    book = np.ascontiguousarray(book, np.uint32)
    A = lib.or_sparse_create(n_dof)
    b = np.zeros(max(1, n_dof))
    lib.or_stress_assemble(xyz.size // 3, _p(xyz), tets.size // 4, _p(tets), _p(E), _p(nu), _p(rho),
                           _p(g), _p(u_node), ntri, _p(tr), _p(tt), _p(book), n_dof, A, _p(b))
    return _export(A, n_dof), b[:n_dof]


STRESS_STATIC, STRESS_TRANSIENT, STRESS_MODAL = 0, 1, 2


def stress_assemble_ex(xyz, tets, E, nu, rho, gravity, u_node, book, n_dof, alpha_t=None, dT=None, tris=None,
                       traction=None, problem=STRESS_STATIC, alpha=1.0, dt=1.0, rot=None, rot_active=None):
    """The full RSolverStress volume loop + assemblyMatrix restated (thermal load, consistent
    mass, transient and modal branches); returns (A, M or None, b)."""
    lib = load()
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
    E, nu, rho, u_node = (np.ascontiguousarray(a, np.float64) for a in (E, nu, rho, u_node))
    at = np.ascontiguousarray(alpha_t, np.float64) if alpha_t is not None else None
    dt_ = np.ascontiguousarray(dT, np.float64) if dT is not None else None
    g = np.ascontiguousarray(gravity, np.float64)
    ntri = 0 if tris is None else int(np.asarray(tris).size // 3)
    tr = np.ascontiguousarray(tris, np.uint32).reshape(-1) if ntri else None
    tt = np.ascontiguousarray(traction, np.float64).reshape(-1) if ntri else None
    book = np.ascontiguousarray(book, np.uint32)
    A = lib.or_sparse_create(n_dof)
    M = lib.or_sparse_create(n_dof) if problem == STRESS_MODAL else None
    b = np.zeros(max(1, n_dof))
    R = np.ascontiguousarray(rot, np.float64).reshape(-1) if rot is not None else None
    ra = np.ascontiguousarray(rot_active, np.uint8) if rot is not None else None
    lib.or_stress_assemble_ex(xyz.size // 3, _p(xyz), tets.size // 4, _p(tets), _p(E), _p(nu), _p(rho), _p(at),
                              _p(dt_), _p(g), _p(u_node), ntri, _p(tr), _p(tt), int(problem), float(alpha),
                              float(dt), _p(R), _p(ra), _p(book), n_dof, A, M, _p(b))
    return _export(A, n_dof), (_export(M, n_dof) if M is not None else None), b[:n_dof]


def stress_set_displacement(book, x, rot=None, rot_active=None):
    """RSolverStress::setDisplacement restated: node displacements (3 per node)."""
    lib = load()
    book = np.ascontiguousarray(book, np.uint32)
    x = np.ascontiguousarray(x, np.float64)
    nn = book.size // 3
    R = np.ascontiguousarray(rot, np.float64).reshape(-1) if rot is not None else None
    ra = np.ascontiguousarray(rot_active, np.uint8) if rot is not None else None
    u = np.empty(3 * nn)
    lib.or_stress_set_displacement(nn, _p(book), _p(x), _p(R), _p(ra), _p(u))
    return u


def transient_heat(xyz, tets, k, q, rhoc, t_node0, book, n_dof, dt, steps, cvg=1e-13,
                   n_outer=100000, stop_rule=STOP_RELRES):
    """The RSolverHeat time loop restated, one system: every step assembles A and b with the
    previous step's nodeTemperature (src/rsolverheat.cpp:757-800, alpha = 1), clears x
    (prepare, :83-88), runs Jacobi-CG (RSolverHeat::solve, :458-465) and writes x back into
    the free nodes (:473-480).  Returns (per-step solutions, per-step CG iterations)."""
    t = np.array(t_node0, np.float64, copy=True)
    book = np.asarray(book, np.uint32)
    free = book != 0xFFFFFFFF
    pos = book[free].astype(np.int64)
    xs, its = [], []
    for _ in range(steps):
        A, b = heat_assemble(xyz, tets, k, q, t, book, n_dof, rhoc=rhoc, transient=True, alpha=1.0, dt=dt)
        x, info = solve(A, b, None, PRECOND_JACOBI, 1, SOLVER_CG, n_outer, cvg=cvg, stop_rule=stop_rule)
        t[free] = x[pos]
        xs.append(x)
        its.append(int(info["iterations"]))
    return xs, its


def eigen_solve(M, K, method, n_eigen=1, n_iter=100, eig_cvg=1e-10, cvg=1e-13, n_outer=100000,
                stop_rule=STOP_REFERENCE):
    """REigenValueSolver::solve restated (src/reigenvaluesolver.cpp:41-592).  M and K are put
    on one (union) pattern.  Returns (d, ev rows or None, inner CG iterations)."""
    import scipy.sparse as sp
    M, K = M.tocsr(), K.tocsr()
    M.sort_indices()
    K.sort_indices()
    if np.array_equal(M.indptr, K.indptr) and np.array_equal(M.indices, K.indices):
        P, mv, kv = M, M.data.astype(np.float64), K.data.astype(np.float64)
    else:
        # the union pattern, both matrices' values on it (explicit zeros elsewhere)
        P = (abs(M) + abs(K)).tocsr()
        P.sort_indices()
        coo = P.tocoo()

        def on(A):
            A = A.tocoo()
            pos = {(int(r), int(c)): v for r, c, v in zip(A.row, A.col, A.data)}
            return np.array([pos.get((int(r), int(c)), 0.0) for r, c in zip(coo.row, coo.col)], np.float64)
        mv, kv = on(M), on(K)
    rp, col = P.indptr.astype(np.uint32), P.indices.astype(np.uint32)
    n = P.shape[0]
    ne = max(1, min(n_eigen, n))
    d = np.zeros(ne)
    ev = np.zeros(ne * n)
    nd, nr = C.c_uint(), C.c_uint()
    its = C.c_ulonglong()
    conf = OrConf(SOLVER_CG, n_outer, 1000, cvg, stop_rule)
    rc = load().or_eigen_solve(method, n_eigen, n_iter, eig_cvg, C.byref(conf), n, _p(rp), _p(col), _p(mv),
                               _p(kv), _p(d), C.byref(nd), _p(ev), C.byref(nr), C.byref(its))
    if rc != 0:
        raise ValueError(f"oracle eigen solve failed ({rc})")
    rows = ev[: nr.value * n].reshape(nr.value, n) if nr.value else None
    return d[: nd.value], rows, its.value


def heat_flux(xyz, tets, k, t_node):
    """RSolverHeat::process, volume elements (src/rsolverheat.cpp:646-692): (ne, 3) fluxes."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
    k, t_node = np.ascontiguousarray(k, np.float64), np.ascontiguousarray(t_node, np.float64)
    ne = tets.size // 4
    out = np.empty(3 * ne)
    load().or_heat_flux(xyz.size // 3, _p(xyz), ne, _p(tets), _p(k), _p(t_node), _p(out))
    return out.reshape(ne, 3)


def stress_process(xyz, tets, E, nu, u_node):
    """RSolverStress::process, volume elements, static (src/rsolverstress.cpp:1238-1408):
    (normal, shear, von Mises as the reference sums them, node forces (nn, 3))."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
    E, nu, u = (np.ascontiguousarray(a, np.float64).reshape(-1) for a in (E, nu, u_node))
    nn, ne = xyz.size // 3, tets.size // 4
    nrm, sh, vm, f = np.empty(ne), np.empty(ne), np.empty(ne), np.empty(3 * nn)
    load().or_stress_process(nn, _p(xyz), ne, _p(tets), _p(E), _p(nu), _p(u), _p(nrm), _p(sh), _p(vm), _p(f))
    return nrm, sh, vm, f.reshape(nn, 3)


E0 = 8.854187817e-12   # RSolverGeneric::e0, src/rsolvergeneric.cpp:13


def electrostatics_assemble(xyz, tets, eps_r, rho, phi_node, book, n_dof):
    """RSolverElectrostatics::prepare, volume elements (src/rsolverelectrostatics.cpp:300-345):
    Ke = B.B eps_r e0 detJ w, fe = -rho N detJ w, the potential lifting of assemblyMatrix
    (:793-833) -- the heat form with k = eps_r e0 and Q = -rho."""
    eps_r, rho = np.asarray(eps_r, np.float64), np.asarray(rho, np.float64)
    return heat_assemble(xyz, tets, eps_r * E0, -rho, phi_node, book, n_dof)


def electrostatics_process(xyz, tets, eps_r, sigma, phi_node):
    """RSolverElectrostatics::process, volume elements: (field (ne,3), current (ne,3), energy,
    resistivity, Joule heat)."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
    eps_r, sigma, phi = (np.ascontiguousarray(a, np.float64) for a in (eps_r, sigma, phi_node))
    nn, ne = xyz.size // 3, tets.size // 4
    f, j = np.empty(3 * ne), np.empty(3 * ne)
    en, rs_, jh = np.empty(ne), np.empty(ne), np.empty(ne)
    load().or_electrostatics_process(nn, _p(xyz), ne, _p(tets), _p(eps_r), _p(sigma), _p(phi), _p(f), _p(j),
                                     _p(en), _p(rs_), _p(jh))
    return f.reshape(ne, 3), j.reshape(ne, 3), en, rs_, jh


def fluid_assemble(xyz, tets, rho, mu, gravity, state, v_old, acc, stream_velocity, alpha, dt, unsteady, book,
                   n_dof):
    """RSolverFluid::prepare (volume) + computeElementConstantDerivative + assemblyMatrix restated;
    state is (vx, vy, vz, p) per node, v_old and acc 3 per node, gravity 3 per element.
    Returns (A as scipy CSR, b)."""
    lib = load()
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
    ne = tets.size // 4
    args = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), shp)).reshape(-1)
            for v, shp in ((rho, (ne,)), (mu, (ne,)), (gravity, (ne, 3)))]
    st = np.ascontiguousarray(state, np.float64).reshape(-1)
    vo = np.ascontiguousarray(v_old, np.float64).reshape(-1)
    ac = np.ascontiguousarray(acc, np.float64).reshape(-1)
    book = np.ascontiguousarray(book, np.uint32)
    A = lib.or_sparse_create(n_dof)
    b = np.zeros(max(1, n_dof))
    lib.or_fluid_assemble(xyz.size // 3, _p(xyz), ne, _p(tets), _p(args[0]), _p(args[1]), _p(args[2]), _p(st),
                          _p(vo), _p(ac), float(stream_velocity), float(alpha), float(dt), int(bool(unsteady)),
                          _p(book), n_dof, A, _p(b))
    return _export(A, n_dof), b[:n_dof]


def fluid_update(book, x, unsteady, dt, state, v_old, acc):
    """RSolverFluid::solve write-back restated; updates state, v_old, acc in place (float64,
    contiguous) and returns the norms (|v| before, |v| after, |p| before, |p| after)."""
    lib = load()
    book = np.ascontiguousarray(book, np.uint32)
    x = np.ascontiguousarray(x, np.float64)
    for a in (state, v_old, acc):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    norms = np.zeros(4)
    lib.or_fluid_update(state.size // 4, _p(book), _p(x), int(bool(unsteady)), float(dt), _p(state), _p(v_old),
                        _p(acc), _p(norms))
    return norms


def transport_assemble(xyz, tets, ca, k, velocity, q_node, field, alpha, dt, unsteady, book, n_dof):
    """RSolverFluidHeat / RSolverFluidParticle element + assemblyMatrix restated; ca, k per element,
    velocity 3 per element, q_node and field per node.  Returns (A as scipy CSR, b)."""
    lib = load()
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
    ne = tets.size // 4
    ca, k = (np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (ne,))) for a in (ca, k))
    vel = np.ascontiguousarray(np.broadcast_to(np.asarray(velocity, np.float64), (ne, 3))).reshape(-1)
    q = np.ascontiguousarray(q_node, np.float64)
    f = np.ascontiguousarray(field, np.float64)
    book = np.ascontiguousarray(book, np.uint32)
    A = lib.or_sparse_create(n_dof)
    b = np.zeros(max(1, n_dof))
    lib.or_transport_assemble(xyz.size // 3, _p(xyz), ne, _p(tets), _p(ca), _p(k), _p(vel), _p(q), _p(f),
                              float(alpha), float(dt), int(bool(unsteady)), _p(book), n_dof, A, _p(b))
    return _export(A, n_dof), b[:n_dof]


def transport_update(book, x, field):
    """Write-back restated: field (float64, contiguous, updated in place) = max(x, 0) on the
    enabled nodes; returns (|field| before, |field| after)."""
    lib = load()
    book = np.ascontiguousarray(book, np.uint32)
    x = np.ascontiguousarray(x, np.float64)
    assert field.dtype == np.float64 and field.flags.c_contiguous
    norms = np.zeros(2)
    lib.or_transport_update(field.size, _p(book), _p(x), _p(field), _p(norms))
    return norms


SIGMA = 5.670367e-8


def radiosity_assemble(F, patch_ptr, patch_elems, emissivity, temperature, ambient):
    """RSolverRadiativeHeat::prepare restated (dense F, n_patches x n_patches); returns the dense
    matrix A and b."""
    lib = load()
    F = np.ascontiguousarray(F, np.float64)
    npch = F.shape[0]
    pp = np.ascontiguousarray(patch_ptr, np.uint32)
    pe = np.ascontiguousarray(patch_elems, np.uint32)
    em, te, am = (np.ascontiguousarray(a, np.float64) for a in (emissivity, temperature, ambient))
    A = np.empty((npch, npch))
    b = np.empty(npch)
    lib.or_radiosity_assemble(npch, _p(F), _p(pp), _p(pe), _p(em), _p(te), _p(am), _p(A), _p(b))
    return A, b


def radiosity_process(patch_ptr, patch_elems, elem_area, patch_heat):
    """RSolverRadiativeHeat::process restated; returns the element heats."""
    lib = load()
    pp = np.ascontiguousarray(patch_ptr, np.uint32)
    pe = np.ascontiguousarray(patch_elems, np.uint32)
    ar = np.ascontiguousarray(elem_area, np.float64)
    ph = np.ascontiguousarray(patch_heat, np.float64)
    out = np.empty(ar.size)
    lib.or_radiosity_process(ph.size, ar.size, _p(pp), _p(pe), _p(ar), _p(ph), _p(out))
    return out


def magnetostatics_assemble(xyz, tets, current_node, field_node, book, n_dof):
    """RSolverMagnetostatics::prepare + assemblyMatrix restated; returns (A as scipy CSR, b)."""
    lib = load()
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
    cur = np.ascontiguousarray(current_node, np.float64).reshape(-1)
    fld = np.ascontiguousarray(field_node, np.float64).reshape(-1)
    book = np.ascontiguousarray(book, np.uint32)
    A = lib.or_sparse_create(n_dof)
    b = np.zeros(max(1, n_dof))
    lib.or_magnetostatics_assemble(xyz.size // 3, _p(xyz), tets.size // 4, _p(tets), _p(cur), _p(fld), _p(book),
                                   n_dof, A, _p(b))
    return _export(A, n_dof), b[:n_dof]

