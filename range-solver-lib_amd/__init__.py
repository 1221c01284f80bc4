"""range_solver_lib_amd -- Python binding of the MI355X element-assembly + sparse-solve library.

A thin ctypes layer over the C ABI declared in include/range_solver_amd.h (the product is the
HIP/C++ library librangesolver_amd.so beside this file).  The classes mirror the reference's
call surface (Range-Software/range-solver-lib):

    RMatrixSolver(conf).solve(A, b, x, precond, block)   include/rmatrixsolver.h:27-42
    RMatrixPreconditioner(A, type, block).compute(x)     include/rmatrixpreconditioner.h:36-48
    SparseMatrix (the CSR cache of src/rmatrixsolver.cpp:14-109, resident in HBM)
    Assembly (RSolverHeat/RSolverStress prepare + assemblyMatrix on the GPU)

Loading fails loudly when the shared library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librangesolver_amd.so")

RS_EOD = 0xFFFFFFFF

PRECOND_NONE, PRECOND_JACOBI, PRECOND_BLOCK_JACOBI, PRECOND_LEARNED = 0, 1, 2, 3
SOLVER_CG, SOLVER_GMRES = 0, 1
STOP_REFERENCE, STOP_RELRES = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1
TIMED_CG_PRODUCT = 2     # spmv_timed: the FP64 CG iteration's fused product (q = A p, partial p.q)
SPMV_SHAPES = 11
SPMV_SHAPE_NAMES = ["vector-csr g4", "vector-csr g8", "vector-csr g16", "vector-csr g32",
                    "vector-csr g64", "tile-csr 2048", "tile-csr 4096", "sell-64", "sell-64 int16",
                    "sell-64 u8", "sell-64 int16 u8"]
SPMV_COLUMN_BYTES = [4, 4, 4, 4, 4, 4, 4, 4, 2, 4, 2]
REASON_NAMES = {0: "running", 1: "converged", 2: "stagnated", 3: "nonnormal", 4: "maxit"}

STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "UNKNOWN_SOLVER", 3: "INVALID_PRECOND",
                4: "DEVICE", 5: "NO_DEVICE", 6: "OUT_OF_MEMORY", 7: "NOT_INITIALIZED",
                8: "UNSUPPORTED"}
(RS_OK, RS_ERR_INVALID_ARG, RS_ERR_UNKNOWN_SOLVER, RS_ERR_INVALID_PRECOND, RS_ERR_DEVICE, RS_ERR_NO_DEVICE,
 RS_ERR_OUT_OF_MEMORY, RS_ERR_NOT_INITIALIZED, RS_ERR_UNSUPPORTED) = range(9)


class RSError(RuntimeError):
    """Raised for a non-zero rs_status (the reference throws RError in the same cases)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


class Conf(C.Structure):
    _fields_ = [("solver_type", C.c_int32), ("n_outer_iterations", C.c_uint32),
                ("n_inner_iterations", C.c_uint32), ("output_frequency", C.c_uint32),
                ("solver_cvg_value", C.c_double), ("stop_rule", C.c_int32),
                ("precision", C.c_int32), ("output_file_name", C.c_char_p)]


class Info(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("inner_iterations", C.c_uint32),
                ("reason", C.c_int32), ("converged", C.c_int32), ("error", C.c_double),
                ("trend", C.c_double), ("equation_scale", C.c_double), ("a_norm", C.c_double),
                ("b_norm", C.c_double), ("true_relres", C.c_double), ("solve_ms", C.c_double),
                ("precond_applications", C.c_uint32), ("fallback", C.c_uint32)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["reason_name"] = REASON_NAMES.get(self.reason, str(self.reason))
        return d


class EigenConf(C.Structure):
    _fields_ = [("method", C.c_int32), ("n_eigen_values", C.c_uint32), ("n_iterations", C.c_uint32),
                ("solver_cvg_value", C.c_double)]


class EigenInfo(C.Structure):
    _fields_ = [("outer_iterations", C.c_uint32), ("cg_iterations", C.c_uint64), ("solve_ms", C.c_double)]


class FluidConf(C.Structure):
    """RSolverFluid time-step parameters: stream velocity (> 0), time-march alpha, step dt, unsteady."""
    _fields_ = [("stream_velocity", C.c_double), ("alpha", C.c_double), ("dt", C.c_double),
                ("unsteady", C.c_int32)]


STRESS_STATIC, STRESS_TRANSIENT, STRESS_MODAL = 0, 1, 2


class StressConf(C.Structure):
    """RSolverStress problem kind (static / transient / modal), time-march alpha, dt, gravity."""
    _fields_ = [("problem", C.c_int32), ("alpha", C.c_double), ("dt", C.c_double), ("gravity", C.c_double * 3)]


class TransportConf(C.Structure):
    """RSolverFluidHeat / RSolverFluidParticle time-step parameters."""
    _fields_ = [("alpha", C.c_double), ("dt", C.c_double), ("unsteady", C.c_int32)]


class BatchInfo(C.Structure):
    _fields_ = [("rounds", C.c_uint32), ("max_iterations", C.c_uint32),
                ("total_iterations", C.c_uint64), ("systems_converged", C.c_uint32),
                ("replacements", C.c_uint32), ("max_relres", C.c_double), ("solve_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)

# name -> (argtypes); every entry point returns int (rs_status) except rs_last_error.
_SIGS = {
    "rs_init": [C.c_int], "rs_finalize": [], "rs_device_count": [C.POINTER(C.c_int)],
    "rs_version": [C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "rs_dmalloc": [C.POINTER(_P), C.c_size_t], "rs_dfree": [_P],
    "rs_memcpy_htod": [_P, _P, C.c_size_t], "rs_memcpy_dtoh": [_P, _P, C.c_size_t],
    "rs_memcpy_dtod": [_P, _P, C.c_size_t], "rs_synchronize": [],
    "rs_matrix_create": [C.c_uint32, _P, _P, _P, C.POINTER(_P)],
    "rs_matrix_update_values": [_P, _P], "rs_matrix_update_values_device": [_P, _P],
    "rs_matrix_info": [_P, _u32p, C.POINTER(C.c_uint64), _f64p],
    "rs_matrix_download": [_P, _P, _P, _P], "rs_matrix_spmv": [_P, _P, _P],
    "rs_matrix_spmv_device": [_P, _P, _P], "rs_matrix_destroy": [_P],
    "rs_matrix_spmv_timed": [_P, C.c_int32, C.c_uint32, C.c_int32, _f64p],
    "rs_matrix_spmv_plan": [_P, C.c_int32, C.POINTER(C.c_int32), _P],
    "rs_matrix_force_plan": [_P, C.c_int32, C.c_int32],
    "rs_precond_create": [_P, C.c_int32, C.c_uint32, C.POINTER(_P)],
    "rs_precond_compute": [_P, _P, _P], "rs_precond_width": [_P, _u32p],
    "rs_precond_download": [_P, _P], "rs_precond_copy": [_P, C.POINTER(_P)], "rs_precond_destroy": [_P],
    "rs_solver_create": [C.POINTER(Conf), C.POINTER(_P)], "rs_solver_copy": [_P, C.POINTER(_P)],
    "rs_solver_disable_convergence_log": [_P],
    "rs_solver_solve": [_P, _P, _P, _P, C.c_int32, C.c_uint32, C.POINTER(Info)],
    "rs_solver_solve_device": [_P, _P, _P, _P, C.c_int32, C.c_uint32, C.POINTER(Info)],
    "rs_solver_solve_device_noalloc": [_P, _P, _P, _P, C.c_int32, C.c_uint32, C.POINTER(Info)],
    "rs_solver_history": [_P, _P, C.c_uint32, _u32p], "rs_solver_destroy": [_P],
    "rs_mesh_kuhn_box": [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, _P, _P],
    "rs_mesh_plane_faces": [C.c_uint32, _P, C.c_uint32, _P, C.c_int32, C.c_double, C.c_double, _P,
                            C.c_uint64, C.POINTER(C.c_uint64)],
    "rs_field_uniform": [C.c_uint64, C.c_uint64, _P],
    "rs_field_lognormal": [C.c_uint64, C.c_uint64, C.c_double, C.c_double, _P],
    "rs_node_book": [C.c_uint32, C.c_uint32, C.c_uint32, _P, C.c_uint32, _P, _P, _u32p],
    "rs_assembly_create": [C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _P, C.POINTER(_P)],
    "rs_assembly_info": [_P, _u32p, C.POINTER(C.c_uint64), _u32p],
    "rs_assembly_book": [_P, _P], "rs_assembly_last_ms": [_P, _f64p],
    "rs_assembly_matrix": [_P, C.POINTER(_P)],
    "rs_assemble_heat": [_P, _P, _P, _P, _P, C.c_int32, C.c_double, C.c_double, _P],
    "rs_assemble_heat_device": [_P, _P, _P, _P, _P, C.c_int32, C.c_double, C.c_double, _P],
    "rs_assemble_stress": [_P, _P, _P, _P, _P, _P, C.c_uint32, _P, _P, _P],
    "rs_assemble_stress_ex": [_P, C.POINTER(StressConf), _P, _P, _P, _P, _P, _P, C.c_uint32, _P, _P, _P],
    "rs_assembly_mass_matrix": [_P, C.POINTER(_P)],
    "rs_assembly_set_rotations": [_P, _P, _P], "rs_stress_set_displacement_device": [_P, _P, _P],
    "rs_assembly_dofs_to_nodes": [_P, _P, _P], "rs_assembly_destroy": [_P],
    "rs_assembly_set_method": [_P, C.c_int32],
    "rs_heat_flux_device": [_P, _P, _P, _P],
    "rs_assemble_electrostatics_device": [_P, _P, _P, _P, _P],
    "rs_assemble_magnetostatics_device": [_P, _P, _P, _P],
    "rs_assemble_fluid_device": [_P, C.POINTER(FluidConf), _P, _P, _P, _P, _P, _P, _P],
    "rs_fluid_update_device": [_P, _P, C.c_int32, C.c_double, _P, _P, _P, _f64p],
    "rs_assemble_transport_device": [_P, C.POINTER(TransportConf), _P, _P, _P, _P, _P, _P],
    "rs_transport_update_device": [_P, _P, _P, _f64p],
    "rs_radiosity_create": [C.c_uint32, C.c_uint32, _P, C.c_uint32, _P, _P, _P, C.POINTER(_P)],
    "rs_radiosity_matrix": [_P, C.POINTER(_P)], "rs_radiosity_last_ms": [_P, _f64p],
    "rs_assemble_radiosity_device": [_P, _P, _P, _P, _P, _P],
    "rs_radiosity_process_device": [_P, _P, _P], "rs_radiosity_destroy": [_P],
    "rs_electrostatics_process_device": [_P, _P, _P, _P, _P, _P, _P, _P, _P],
    "rs_stress_process_device": [_P, _P, _P, _P, _P, _P, _P, _P],
    "rs_aggregates_by_bins": [C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                              _P, _u32p],
    "rs_hierarchy_by_bins": [C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                             C.c_uint32, C.c_uint32, _P, C.c_uint64, _P, _u32p],
    "rs_learned_create": [_P, C.c_uint32, _P, _P, C.POINTER(_P)],
    "rs_learned_create_by_bins": [_P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32, C.c_uint32, C.POINTER(_P)],
    "rs_hierarchy_by_bins_ex": [C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_uint32, C.c_uint32, C.c_int32, _P, C.c_uint64, _P, _u32p],
    "rs_learned_create_by_bins_ex": [_P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(_P)],
    "rs_learned_level_map": [_P, C.c_uint32, _P, C.c_uint64],
    "rs_learned_level_basis": [_P, C.c_uint32, _P, C.c_uint64],
    "rs_learned_set_coarse_cycle": [_P, C.c_uint32, C.c_uint32],
    "rs_assembly_device_mesh": [_P, C.POINTER(_P), C.POINTER(_P), _u32p, _u32p],
    "rs_learned_set_params": [_P, _P, C.c_uint64],
    "rs_learned_set_cycle": [_P, C.c_uint32, C.c_uint32, C.c_int],
    "rs_learned_set_compressed": [_P, C.c_uint32],
    "rs_learned_info": [_P, _u32p, _P, _P, C.POINTER(C.c_uint64)],
    "rs_learned_setup": [_P, _P], "rs_learned_level_lambda": [_P, _P],
    "rs_learned_level_matrix": [_P, C.c_uint32, _P, _P, _P],
    "rs_learned_coarse_matrix": [_P, _P],
    "rs_learned_restrict_device": [_P, _P, _P], "rs_learned_apply_coarse": [_P, _P, _P],
    "rs_learned_apply": [_P, _P, _P, _P], "rs_learned_infer_timed": [_P, C.c_uint32, _f64p],
    "rs_learned_record": [_P, C.c_uint32], "rs_learned_records": [_P, C.c_uint32, _P],
    "rs_learned_destroy": [_P],
    "rs_solver_solve_learned_device": [_P, _P, _P, _P, _P, C.POINTER(Info)],
    "rs_trainer_create": [_P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(_P)],
    "rs_trainer_gradient": [_P, C.c_uint32, C.c_uint64, _P, _P, C.c_uint32, _P, _P, _P, _P, _P],
    "rs_trainer_step": [_P, C.c_uint32, C.c_uint64, _P, _P, C.c_uint32, _P, _P, _P, C.c_float,
                        C.POINTER(C.c_float)],
    "rs_trainer_get_params": [_P, _P, C.c_uint64], "rs_trainer_destroy": [_P],
    "rs_batch_create": [C.c_uint32, _P, _P, C.POINTER(_P)],
    "rs_batch_set_matrix": [_P, C.c_uint32, _P], "rs_batch_set_mass": [_P, C.c_uint32, _P],
    "rs_batch_set_rhs": [_P, C.c_uint32, _P, _P],
    "rs_batch_solve": [_P, C.POINTER(Conf), C.c_uint32, C.c_int32, C.POINTER(BatchInfo)],
    "rs_batch_stats": [_P, _P, _P, _P], "rs_batch_solution": [_P, C.c_uint32, _P],
    "rs_batch_shape": [_P, C.c_int32, C.POINTER(C.c_int32), _P],
    "rs_batch_force_shape": [_P, C.c_int32, C.c_int32],
    "rs_batch_info_layout": [_P, _u32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_int32)],
    "rs_batch_set_replacement": [_P, C.c_double],
    "rs_batch_destroy": [_P],
    "rs_eigen_solve": [C.POINTER(EigenConf), C.POINTER(Conf), _P, _P, _P, C.c_uint32, _u32p, _P, C.c_uint64,
                       _u32p, C.POINTER(EigenInfo)],
    "rs_eigen_solve_learned": [C.POINTER(EigenConf), C.POINTER(Conf), _P, _P, _P, _P, C.c_uint32, _u32p, _P,
                               C.c_uint64, _u32p, C.POINTER(EigenInfo)],
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load librangesolver_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `make -C range-solver-lib_amd` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(path)
    for name, argtypes in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = C.c_int
    lib.rs_last_error.argtypes = []
    lib.rs_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return sorted(list(_SIGS) + ["rs_last_error"])


def _check(rc: int) -> None:
    if rc != 0:
        msg = _lib.rs_last_error().decode(errors="replace")
        raise RSError(rc, msg)


def _ptr(a: np.ndarray | None) -> C.c_void_p | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the C ABI must be contiguous"
    return C.c_void_p(a.ctypes.data)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _need(name: str, size: int, want: int) -> None:
    """Arrays cross the C ABI as bare pointers: a wrong length is refused here, before the
    library reads or writes past the end of the caller's buffer."""
    if size != want:
        raise ValueError(f"{name} has {size} entries, expected {want}")


def _need_dev(name: str, d: "DeviceArray", want: int, dtype=np.float64) -> None:
    if d.dtype != np.dtype(dtype) or d.n < want:
        raise ValueError(f"{name} holds {d.n} {d.dtype} values, expected at least {want} {np.dtype(dtype)}")


# ------------------------------------------------------------------------------ runtime
def init(device: int = 0) -> None:
    load()
    _check(_lib.rs_init(device))


def device_count() -> int:
    load()
    n = C.c_int(0)
    _check(_lib.rs_device_count(C.byref(n)))
    return n.value


class DeviceArray:
    """A plain device allocation (rs_dmalloc) holding a 1-D numpy-typed array."""

    def __init__(self, n: int, dtype=np.float64):
        load()
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.ptr = C.c_void_p()
        _check(_lib.rs_dmalloc(C.byref(self.ptr), max(1, self.n * self.dtype.itemsize)))

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _check(_lib.rs_memcpy_htod(self.ptr, _ptr(a), a.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        _check(_lib.rs_memcpy_dtoh(_ptr(out), self.ptr, out.nbytes))
        return out

    def __del__(self):
        if _lib is not None and getattr(self, "ptr", None) is not None and self.ptr.value:
            _lib.rs_dfree(self.ptr)
            self.ptr = C.c_void_p()


def synchronize() -> None:
    _check(_lib.rs_synchronize())


# ------------------------------------------------------------------------------ matrix
class SparseMatrix:
    """Device CSR with ||A||_F (SparseMatrixCache, src/rmatrixsolver.cpp:14-54)."""

    def __init__(self, n: int, rowptr, col, val, _handle=None, _owner=None):
        load()
        self._owner = _owner
        if _handle is not None:
            self.h = _handle
            self._borrowed = True
        else:
            self._borrowed = False
            self.rowptr, self.col, self.val = _u32(rowptr), _u32(col), _f64(val)
            _need("rowptr", self.rowptr.size, int(n) + 1)
            nnz = int(self.rowptr[-1])
            _need("col", self.col.size, nnz)
            _need("val", self.val.size, nnz)
            self.h = C.c_void_p()
            _check(_lib.rs_matrix_create(n, _ptr(self.rowptr), _ptr(self.col), _ptr(self.val),
                                         C.byref(self.h)))

    @classmethod
    def from_scipy(cls, m) -> "SparseMatrix":
        m = m.tocsr()
        return cls(m.shape[0], m.indptr, m.indices, m.data)

    def info(self):
        n, nnz, fro = C.c_uint32(), C.c_uint64(), C.c_double()
        _check(_lib.rs_matrix_info(self.h, C.byref(n), C.byref(nnz), C.byref(fro)))
        return n.value, nnz.value, fro.value

    @property
    def n(self) -> int:
        return self.info()[0]

    @property
    def frobenius(self) -> float:
        return self.info()[2]

    def update_values(self, val) -> None:
        v = _f64(val)
        _need("val", v.size, self.info()[1])
        _check(_lib.rs_matrix_update_values(self.h, _ptr(v)))

    def download(self):
        n, nnz, _ = self.info()
        rp = np.empty(n + 1, np.uint32)
        col = np.empty(nnz, np.uint32)
        val = np.empty(nnz, np.float64)
        _check(_lib.rs_matrix_download(self.h, _ptr(rp), _ptr(col), _ptr(val)))
        return rp, col, val

    def to_scipy(self):
        import scipy.sparse as sp
        rp, col, val = self.download()
        n = rp.size - 1
        return sp.csr_matrix((val, col, rp), shape=(n, n))

    def spmv(self, x) -> np.ndarray:
        x = _f64(x)
        _need("x", x.size, self.n)
        y = np.empty_like(x)
        _check(_lib.rs_matrix_spmv(self.h, _ptr(x), _ptr(y)))
        return y

    def spmv_device(self, dx: DeviceArray, dy: DeviceArray) -> None:
        _need_dev("dx", dx, self.n)
        _need_dev("dy", dy, self.n)
        _check(_lib.rs_matrix_spmv_device(self.h, dx.ptr, dy.ptr))

    def spmv_timed(self, precision: int = PRECISION_FP64, reps: int = 50, plan: int = -1) -> float:
        """Average device milliseconds of one SpMV launch (stream events); plan -1 = tuned."""
        ms = C.c_double()
        _check(_lib.rs_matrix_spmv_timed(self.h, precision, reps, plan, C.byref(ms)))
        return ms.value

    def force_plan(self, precision: int, plan: int) -> None:
        """Override the autotuner's SpMV shape for `precision` (plan -1: tune again)."""
        _check(_lib.rs_matrix_force_plan(self.h, precision, plan))

    def spmv_plan(self, precision: int = PRECISION_FP64):
        """(autotuned shape id, tuning ms per shape)."""
        plan = C.c_int32()
        ms = np.zeros(SPMV_SHAPES)
        _check(_lib.rs_matrix_spmv_plan(self.h, precision, C.byref(plan), _ptr(ms)))
        return plan.value, ms

    def __del__(self):
        if _lib is not None and not getattr(self, "_borrowed", True) and getattr(self, "h", None):
            _lib.rs_matrix_destroy(self.h)
            self.h = None


class RMatrixPreconditioner:
    """RMatrixPreconditioner(A, type, blockSize).compute(x) -- include/rmatrixpreconditioner.h:36-48."""

    def __init__(self, A: SparseMatrix, ptype: int = PRECOND_NONE, block_size: int = 1):
        self.A = A
        self.h = C.c_void_p()
        _check(_lib.rs_precond_create(A.h, ptype, block_size, C.byref(self.h)))

    @property
    def width(self) -> int:
        w = C.c_uint32()
        _check(_lib.rs_precond_width(self.h, C.byref(w)))
        return w.value

    def data(self) -> np.ndarray:
        out = np.empty(self.A.n * max(1, self.width), np.float64)
        _check(_lib.rs_precond_download(self.h, _ptr(out)))
        return out

    def copy(self) -> "RMatrixPreconditioner":
        """RMatrixPreconditioner(const RMatrixPreconditioner&): its own copy of type and values."""
        other = RMatrixPreconditioner.__new__(RMatrixPreconditioner)
        other.A = self.A
        other.h = C.c_void_p()
        _check(_lib.rs_precond_copy(self.h, C.byref(other.h)))
        return other

    def compute(self, x) -> np.ndarray:
        x = _f64(x)
        _need("x", x.size, self.A.n)
        y = np.empty_like(x)
        _check(_lib.rs_precond_compute(self.h, _ptr(x), _ptr(y)))
        return y

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.rs_precond_destroy(self.h)
            self.h = None


@dataclass
class MatrixSolverConf:
    """RMatrixSolverConf as used by the reference (range-model-lib)."""
    solver_type: int = SOLVER_CG
    n_outer_iterations: int = 1000
    n_inner_iterations: int = 1000
    solver_cvg_value: float = 1e-13
    output_frequency: int = 0
    output_file_name: str | None = None
    stop_rule: int = STOP_REFERENCE
    precision: int = PRECISION_FP64

    def to_c(self) -> Conf:
        return Conf(self.solver_type, self.n_outer_iterations, self.n_inner_iterations,
                    self.output_frequency, self.solver_cvg_value, self.stop_rule, self.precision,
                    self.output_file_name.encode() if self.output_file_name else None)


class RMatrixSolver:
    """RMatrixSolver(conf).solve(A, b, x, precond, blockSize) -- include/rmatrixsolver.h:27-42."""

    def __init__(self, conf: MatrixSolverConf):
        load()
        self.conf = conf
        c = conf.to_c()
        self.h = C.c_void_p()
        _check(_lib.rs_solver_create(C.byref(c), C.byref(self.h)))
        self.info: Info | None = None

    def copy(self) -> "RMatrixSolver":
        other = RMatrixSolver.__new__(RMatrixSolver)
        other.conf = self.conf
        other.info = None
        other.h = C.c_void_p()
        _check(_lib.rs_solver_copy(self.h, C.byref(other.h)))
        return other

    def disable_convergence_log_file(self) -> None:
        _check(_lib.rs_solver_disable_convergence_log(self.h))

    def solve(self, A: SparseMatrix, b, x=None, precond: int = PRECOND_NONE,
              block_size: int = 1) -> np.ndarray:
        b = _f64(b)
        n = A.n
        _need("b", b.size, n)
        # the reference resizes x to n, keeping the leading entries as the initial guess
        x0 = np.zeros(n) if x is None else _f64(x)
        if x0.size != n:
            x0 = np.concatenate([x0[:n], np.zeros(max(0, n - x0.size))])
        x0 = np.array(x0, dtype=np.float64, copy=True)
        info = Info()
        _check(_lib.rs_solver_solve(self.h, A.h, _ptr(b), _ptr(x0), precond, block_size,
                                    C.byref(info)))
        self.info = info
        return x0

    def solve_device(self, A: SparseMatrix, db: DeviceArray, dx: DeviceArray,
                     precond: int = PRECOND_NONE, block_size: int = 1) -> Info:
        n = A.n
        _need_dev("db", db, n)
        _need_dev("dx", dx, n)
        info = Info()
        _check(_lib.rs_solver_solve_device(self.h, A.h, db.ptr, dx.ptr, precond, block_size,
                                           C.byref(info)))
        self.info = info
        return info

    def solve_device_noalloc(self, A: SparseMatrix, db: DeviceArray, dx: DeviceArray,
                             precond: int = PRECOND_NONE, block_size: int = 1) -> Info:
        """rs_solver_solve_device_noalloc: FP64 CG in the library's fixed device workspace."""
        n = A.n
        _need_dev("db", db, n)
        _need_dev("dx", dx, n)
        info = Info()
        _check(_lib.rs_solver_solve_device_noalloc(self.h, A.h, db.ptr, dx.ptr, precond, block_size,
                                                   C.byref(info)))
        self.info = info
        return info

    def solve_learned_device(self, A: SparseMatrix, L: "LearnedPreconditioner", db: DeviceArray,
                             dx: DeviceArray) -> Info:
        """Flexible CG with the learned two-level preconditioner (mixed precision)."""
        n = A.n
        _need_dev("db", db, n)
        _need_dev("dx", dx, n)
        info = Info()
        _check(_lib.rs_solver_solve_learned_device(self.h, A.h, L.h, db.ptr, dx.ptr, C.byref(info)))
        self.info = info
        return info

    def history(self) -> np.ndarray:
        cnt = C.c_uint32()
        _check(_lib.rs_solver_history(self.h, None, 0, C.byref(cnt)))
        out = np.empty(2 * max(1, cnt.value), np.float64)
        _check(_lib.rs_solver_history(self.h, _ptr(out), cnt.value, C.byref(cnt)))
        return out[: 2 * cnt.value].reshape(-1, 2)

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.rs_solver_destroy(self.h)
            self.h = None


# ------------------------------------------------------------------------------ model prep (host)
def kuhn_box(nx: int, ny: int, nz: int, lx: float = 1.0, ly: float = 1.0, lz: float = 1.0):
    load()
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    xyz = np.empty(3 * nn, np.float64)
    tets = np.empty(24 * nx * ny * nz, np.uint32)
    _check(_lib.rs_mesh_kuhn_box(nx, ny, nz, lx, ly, lz, _ptr(xyz), _ptr(tets)))
    return xyz.reshape(-1, 3), tets.reshape(-1, 4)


def plane_faces(xyz, tets, axis: int, value: float, tol: float = 1e-12) -> np.ndarray:
    load()
    xyz, tets = _f64(xyz), _u32(tets)
    cnt = C.c_uint64()
    _check(_lib.rs_mesh_plane_faces(xyz.size // 3, _ptr(xyz), tets.size // 4, _ptr(tets), axis,
                                    value, tol, None, 0, C.byref(cnt)))
    out = np.empty(3 * cnt.value, np.uint32)
    _check(_lib.rs_mesh_plane_faces(xyz.size // 3, _ptr(xyz), tets.size // 4, _ptr(tets), axis,
                                    value, tol, _ptr(out), cnt.value, C.byref(cnt)))
    return out.reshape(-1, 3)


def field_uniform(seed: int, n: int) -> np.ndarray:
    load()
    out = np.empty(n, np.float64)
    _check(_lib.rs_field_uniform(seed, n, _ptr(out)))
    return out


def field_lognormal(seed: int, n: int, mu: float, sigma: float) -> np.ndarray:
    load()
    out = np.empty(n, np.float64)
    _check(_lib.rs_field_lognormal(seed, n, mu, sigma, _ptr(out)))
    return out


def node_book(n_nodes: int, conn, dofs_per_node: int, fixed) -> tuple[np.ndarray, int]:
    load()
    conn = _u32(conn)
    npe = conn.shape[1] if conn.ndim == 2 else 4
    fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
    book = np.empty(n_nodes * dofs_per_node, np.uint32)
    n = C.c_uint32()
    _check(_lib.rs_node_book(n_nodes, conn.size // npe, npe, _ptr(conn), dofs_per_node,
                             _ptr(fixed), _ptr(book), C.byref(n)))
    return book, n.value


# ------------------------------------------------------------------------------ radiative heat
class Radiosity:
    """RSolverRadiativeHeat with a loaded view-factor matrix: patch book, element areas and the
    dense patch system (rs_radiosity_*)."""

    def __init__(self, xyz, tris, patch_ptr, patch_elems):
        load()
        self.xyz = _f64(xyz).reshape(-1)
        self.tris = _u32(tris).reshape(-1)
        pp, pe = _u32(patch_ptr).reshape(-1), _u32(patch_elems).reshape(-1)
        self.n_patches = pp.size - 1
        self.n_elems = self.tris.size // 3
        self.h = C.c_void_p()
        _check(_lib.rs_radiosity_create(self.n_patches, self.xyz.size // 3, _ptr(self.xyz), self.n_elems,
                                        _ptr(self.tris), _ptr(pp), _ptr(pe), C.byref(self.h)))
        mh = C.c_void_p()
        _check(_lib.rs_radiosity_matrix(self.h, C.byref(mh)))
        self.matrix = SparseMatrix(0, None, None, None, _handle=mh, _owner=self)

    def assemble_device(self, dview_factors, demissivity, dtemperature, dambient, db) -> None:
        _check(_lib.rs_assemble_radiosity_device(self.h, dview_factors.ptr, demissivity.ptr, dtemperature.ptr,
                                                 dambient.ptr, db.ptr))

    def process_device(self, dpatch_heat, delem_heat) -> None:
        _check(_lib.rs_radiosity_process_device(self.h, dpatch_heat.ptr, delem_heat.ptr))

    @property
    def last_ms(self) -> float:
        v = C.c_double()
        _check(_lib.rs_radiosity_last_ms(self.h, C.byref(v)))
        return v.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.rs_radiosity_destroy(self.h)
            self.h = None


# ------------------------------------------------------------------------------ assembly
class Assembly:
    """Device mesh + node book + CSR pattern + scatter map + element colouring."""

    def __init__(self, xyz, tets, dofs_per_node: int, fixed):
        load()
        self.xyz = _f64(xyz).reshape(-1)
        self.tets = _u32(tets).reshape(-1)
        self.nn = self.xyz.size // 3
        self.ne = self.tets.size // 4
        self.dpn = dofs_per_node
        fixed = np.ascontiguousarray(fixed, dtype=np.uint8).reshape(-1)
        assert fixed.size == self.nn * dofs_per_node
        self.h = C.c_void_p()
        _check(_lib.rs_assembly_create(self.nn, _ptr(self.xyz), self.ne, _ptr(self.tets),
                                       dofs_per_node, _ptr(fixed), C.byref(self.h)))
        nd, nnz, nc = C.c_uint32(), C.c_uint64(), C.c_uint32()
        _check(_lib.rs_assembly_info(self.h, C.byref(nd), C.byref(nnz), C.byref(nc)))
        self.n_dof, self.nnz, self.n_colours = nd.value, nnz.value, nc.value
        mh = C.c_void_p()
        _check(_lib.rs_assembly_matrix(self.h, C.byref(mh)))
        self.matrix = SparseMatrix(0, None, None, None, _handle=mh, _owner=self)

    def set_method(self, row_gather: bool) -> None:
        """Heat, transport and static stress assembly kernels: row gather (default) or one launch
        per colour."""
        _check(_lib.rs_assembly_set_method(self.h, 1 if row_gather else 0))

    def book(self) -> np.ndarray:
        out = np.empty(self.nn * self.dpn, np.uint32)
        _check(_lib.rs_assembly_book(self.h, _ptr(out)))
        return out

    @property
    def last_ms(self) -> float:
        v = C.c_double()
        _check(_lib.rs_assembly_last_ms(self.h, C.byref(v)))
        return v.value

    def assemble_heat(self, k, q, t_node, rhoc=None, transient=False, alpha=1.0, dt=1.0):
        k, q, t_node = _f64(k), _f64(q), _f64(t_node)
        rc = _f64(rhoc) if rhoc is not None else None
        b = np.empty(self.n_dof, np.float64)
        _check(_lib.rs_assemble_heat(self.h, _ptr(k), _ptr(q), _ptr(rc), _ptr(t_node),
                                     int(bool(transient)), alpha, dt, _ptr(b)))
        return b

    def assemble_heat_device(self, dk, dq, dt_node, db, drhoc=None, transient=False, alpha=1.0,
                             dt=1.0):
        _check(_lib.rs_assemble_heat_device(self.h, dk.ptr, dq.ptr,
                                            drhoc.ptr if drhoc is not None else None,
                                            dt_node.ptr, int(bool(transient)), alpha, dt, db.ptr))

    def assemble_stress(self, E, nu, rho, gravity, u_node, tris=None, traction=None):
        E, nu, rho, u_node = _f64(E), _f64(nu), _f64(rho), _f64(u_node)
        g = _f64(gravity)
        ntri = 0 if tris is None else int(np.asarray(tris).size // 3)
        tr = _u32(tris) if ntri else None
        tt = _f64(traction) if ntri else None
        b = np.empty(self.n_dof, np.float64)
        _check(_lib.rs_assemble_stress(self.h, _ptr(E), _ptr(nu), _ptr(rho), _ptr(g),
                                       _ptr(u_node), ntri, _ptr(tr), _ptr(tt), _ptr(b)))
        return b

    def assemble_stress_ex(self, conf: "StressConf", E, nu, rho, u_node, thermal_expansion=None, dT=None, tris=None,
                           traction=None):
        """The full RSolverStress prepare (thermal load, consistent mass, transient / modal);
        host arrays; returns b.  After a modal call mass_matrix() is the assembled M."""
        E, nu, rho, u_node = _f64(E), _f64(nu), _f64(rho), _f64(u_node)
        at = _f64(thermal_expansion) if thermal_expansion is not None else None
        dt_ = _f64(dT) if dT is not None else None
        ntri = 0 if tris is None else int(np.asarray(tris).size // 3)
        tr = _u32(tris) if ntri else None
        tt = _f64(traction) if ntri else None
        b = np.empty(self.n_dof, np.float64)
        _check(_lib.rs_assemble_stress_ex(self.h, C.byref(conf), _ptr(E), _ptr(nu), _ptr(rho), _ptr(at), _ptr(dt_),
                                          _ptr(u_node), ntri, _ptr(tr), _ptr(tt), _ptr(b)))
        return b

    def set_rotations(self, R, active) -> None:
        """Local rotations (3x3 per node, row-major) for the nodes with active != 0; R None clears."""
        if R is None:
            _check(_lib.rs_assembly_set_rotations(self.h, None, None))
            return
        self._rot = _f64(R).reshape(-1)
        self._rot_active = np.ascontiguousarray(active, dtype=np.uint8).reshape(-1)
        _check(_lib.rs_assembly_set_rotations(self.h, _ptr(self._rot), _ptr(self._rot_active)))

    def stress_set_displacement_device(self, dx, du_node) -> None:
        """RSolverStress::setDisplacement: x -> node displacements (3 per node), rotated back."""
        _check(_lib.rs_stress_set_displacement_device(self.h, dx.ptr, du_node.ptr))

    def mass_matrix(self) -> "SparseMatrix":
        mh = C.c_void_p()
        _check(_lib.rs_assembly_mass_matrix(self.h, C.byref(mh)))
        return SparseMatrix(0, None, None, None, _handle=mh, _owner=self)

    def heat_flux_device(self, dk: DeviceArray, dt_node: DeviceArray, dflux: DeviceArray) -> None:
        """RSolverHeat::process (volume elements): 3 flux values per element."""
        _check(_lib.rs_heat_flux_device(self.h, dk.ptr, dt_node.ptr, dflux.ptr))

    def assemble_electrostatics_device(self, deps_r, drho, dphi_node, db) -> None:
        """RSolverElectrostatics prepare (volume): the heat form with eps_r e0 and source -rho."""
        _check(_lib.rs_assemble_electrostatics_device(self.h, deps_r.ptr, drho.ptr, dphi_node.ptr, db.ptr))

    def assemble_magnetostatics_device(self, dcurrent_node, dfield_node, db) -> None:
        """RSolverMagnetostatics prepare (volume): three negative Laplacians, curl source."""
        _check(_lib.rs_assemble_magnetostatics_device(self.h, dcurrent_node.ptr, dfield_node.ptr, db.ptr))

    def assemble_fluid_device(self, conf: "FluidConf", drho, dmu, dgravity, dstate, dv_old, dacc, db) -> None:
        """RSolverFluid prepare (volume): the stabilised tetrahedron for the increments of
        (vx, vy, vz, p); state 4 per node, v_old and acc 3 per node, gravity 3 per element."""
        _check(_lib.rs_assemble_fluid_device(self.h, C.byref(conf), drho.ptr, dmu.ptr, dgravity.ptr, dstate.ptr,
                                             dv_old.ptr, dacc.ptr if dacc is not None else None, db.ptr))

    def fluid_update_device(self, dx, unsteady: bool, dt: float, dstate, dv_old, dacc) -> np.ndarray:
        """RSolverFluid::solve write-back; returns (|v| before, |v| after, |p| before, |p| after)."""
        norms = np.zeros(4)
        _check(_lib.rs_fluid_update_device(self.h, dx.ptr, int(bool(unsteady)), dt, dstate.ptr, dv_old.ptr,
                                           dacc.ptr if dacc is not None else None,
                                           norms.ctypes.data_as(_f64p)))
        return norms

    def assemble_transport_device(self, conf: "TransportConf", dca, dk, dvelocity, dsource, dfield, db) -> None:
        """RSolverFluidHeat / RSolverFluidParticle prepare (volume): SUPG advection-diffusion."""
        _check(_lib.rs_assemble_transport_device(self.h, C.byref(conf), dca.ptr, dk.ptr, dvelocity.ptr, dsource.ptr,
                                                 dfield.ptr, db.ptr))

    def transport_update_device(self, dx, dfield) -> np.ndarray:
        """Write-back field = max(x, 0); returns (|field| before, |field| after)."""
        norms = np.zeros(2)
        _check(_lib.rs_transport_update_device(self.h, dx.ptr, dfield.ptr, norms.ctypes.data_as(_f64p)))
        return norms

    def electrostatics_process_device(self, deps_r, dsigma, dphi_node, dfield, dcurrent, denergy, dres,
                                      djoule) -> None:
        """RSolverElectrostatics::process (volume): field, current density, energy, resistivity, Joule."""
        _check(_lib.rs_electrostatics_process_device(self.h, deps_r.ptr, dsigma.ptr, dphi_node.ptr, dfield.ptr,
                                                     dcurrent.ptr, denergy.ptr, dres.ptr, djoule.ptr))

    def stress_process_device(self, dE, dnu, du_node, dnormal, dshear, dvm, dforce) -> None:
        """RSolverStress::process (volume elements, static): normal, shear, their sum, node forces."""
        _check(_lib.rs_stress_process_device(self.h, dE.ptr, dnu.ptr, du_node.ptr, dnormal.ptr, dshear.ptr,
                                             dvm.ptr, dforce.ptr))

    def dofs_to_nodes(self, dx: DeviceArray, dnode: DeviceArray) -> None:
        _check(_lib.rs_assembly_dofs_to_nodes(self.h, dx.ptr, dnode.ptr))

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.rs_assembly_destroy(self.h)
            self.h = None


# ------------------------------------------------------------------------------ eigenvalues
EIGEN_LANCZOS, EIGEN_ARNOLDI, EIGEN_RAYLEIGH = 0, 1, 2


@dataclass
class EigenValueSolverConf:
    """REigenValueSolverConf as the reference uses it (method, nEigenValues, nIterations,
    solverCvgValue)."""
    method: int = EIGEN_ARNOLDI
    n_eigen_values: int = 1
    n_iterations: int = 100
    solver_cvg_value: float = 1e-10


class REigenValueSolver:
    """REigenValueSolver(conf, matrixSolverConf).solve(M, K, d, ev) -- include/reigenvaluesolver.h:9-63,
    on the GPU (rs_eigen_solve).  Returns (d, ev rows or None); info in self.info."""

    def __init__(self, conf: EigenValueSolverConf, matrix_conf: MatrixSolverConf):
        load()
        self.conf = conf
        self.matrix_conf = matrix_conf
        self.info: EigenInfo | None = None

    def solve(self, M: SparseMatrix, K: SparseMatrix, preconditioner: "LearnedPreconditioner | None" = None):
        """preconditioner: a multilevel preconditioner built for K, used for every inner solve
        with K (rs_eigen_solve_learned; no counterpart in the reference)."""
        n = M.n
        ne = max(1, min(self.conf.n_eigen_values, n))
        d = np.zeros(ne)
        ev = np.zeros(ne * n)
        nd, nr = C.c_uint32(), C.c_uint32()
        ec = EigenConf(self.conf.method, self.conf.n_eigen_values, self.conf.n_iterations,
                       self.conf.solver_cvg_value)
        mc = self.matrix_conf.to_c()
        info = EigenInfo()
        if preconditioner is None:
            _check(_lib.rs_eigen_solve(C.byref(ec), C.byref(mc), M.h, K.h, _ptr(d), ne, C.byref(nd), _ptr(ev),
                                       ev.size, C.byref(nr), C.byref(info)))
        else:
            if preconditioner.A is not K:
                raise RSError(RS_ERR_INVALID_ARG, "REigenValueSolver: the preconditioner was built for another matrix")
            _check(_lib.rs_eigen_solve_learned(C.byref(ec), C.byref(mc), M.h, K.h, preconditioner.h, _ptr(d), ne,
                                               C.byref(nd), _ptr(ev), ev.size, C.byref(nr), C.byref(info)))
        self.info = info
        rows = ev[: nr.value * n].reshape(nr.value, n) if nr.value else None
        return d[: nd.value], rows


def seed_c_rand(seed: int) -> None:
    """srand() of the process's C library: fixes the start vectors of REigenValueSolver, which
    come from rand() / RAND_MAX as in the reference."""
    C.CDLL(None).srand(C.c_uint(seed))


# ------------------------------------------------------------------------------ batched systems
BATCH_SHAPES = 5
BATCH_SHAPE_NAMES = ["sell-64 1 lane/row 4 waves", "sell-64 1 lane/row 8 waves",
                     "sell-64 1 lane/row 16 waves", "sell-64 2 lanes/row", "sell-64 4 lanes/row"]


class Batch:
    """Many independent systems solved together (rs_batch_*): each runs `steps` implicit-Euler
    steps of Jacobi-CG, A x_k = c + M x_{k-1}, as RSolverHeat does once per time step
    (src/rsolverheat.cpp:83-88, :757-800, :458-480); every launch covers all systems."""

    def __init__(self, n_rows, nnz):
        load()
        n_rows = np.ascontiguousarray(n_rows, np.uint32)
        nnz = np.ascontiguousarray(nnz, np.uint64)
        assert n_rows.size == nnz.size
        self.count = int(n_rows.size)
        self.n_rows = n_rows
        self.h = C.c_void_p()
        _check(_lib.rs_batch_create(self.count, _ptr(n_rows), _ptr(nnz), C.byref(self.h)))
        self.steps = 0

    def set_matrix(self, i: int, A: SparseMatrix) -> None:
        _check(_lib.rs_batch_set_matrix(self.h, i, A.h))

    def set_mass(self, i: int, M: SparseMatrix) -> None:
        _check(_lib.rs_batch_set_mass(self.h, i, M.h))

    def set_rhs(self, i: int, dc: DeviceArray, dx0: DeviceArray | None = None) -> None:
        _check(_lib.rs_batch_set_rhs(self.h, i, dc.ptr, dx0.ptr if dx0 is not None else None))

    def solve(self, conf: MatrixSolverConf, steps: int, warm_start: bool = False) -> BatchInfo:
        c = conf.to_c()
        info = BatchInfo()
        _check(_lib.rs_batch_solve(self.h, C.byref(c), steps, int(bool(warm_start)), C.byref(info)))
        self.steps = steps
        return info

    def stats(self):
        """(iterations[count, steps], relres[count, steps], reasons[count])."""
        its = np.empty(self.count * self.steps, np.uint32)
        rr = np.empty(self.count * self.steps, np.float64)
        reasons = np.empty(self.count, np.int32)
        _check(_lib.rs_batch_stats(self.h, _ptr(its), _ptr(rr), _ptr(reasons)))
        return its.reshape(self.count, self.steps), rr.reshape(self.count, self.steps), reasons

    def solution(self, i: int, dx: DeviceArray | None = None):
        out = dx if dx is not None else DeviceArray(int(self.n_rows[i]))
        _check(_lib.rs_batch_solution(self.h, i, out.ptr))
        return out if dx is not None else out.download()

    def shape(self, precision: int = PRECISION_MIXED):
        sid = C.c_int32()
        ms = np.zeros(BATCH_SHAPES, np.float64)
        _check(_lib.rs_batch_shape(self.h, precision, C.byref(sid), _ptr(ms)))
        return sid.value, ms

    def force_shape(self, precision: int, shape: int) -> None:
        _check(_lib.rs_batch_force_shape(self.h, precision, shape))

    def layout(self):
        rows, nnz, ent, cb = C.c_uint32(), C.c_uint64(), C.c_uint64(), C.c_int32()
        _check(_lib.rs_batch_info_layout(self.h, C.byref(rows), C.byref(nnz), C.byref(ent), C.byref(cb)))
        return {"padded_rows": rows.value, "nnz": nnz.value, "sell_entries": ent.value, "column_bytes": cb.value}

    def set_replacement(self, ratio: float) -> None:
        """MIXED: FP64 residual replacement when the residual fell by `ratio` (default 0.01)."""
        _check(_lib.rs_batch_set_replacement(self.h, ratio))

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.rs_batch_destroy(self.h)
            self.h = None


# ------------------------------------------------------------------------------ learned preconditioner
N_PARAMS = 49921
MAX_COARSE = 512
COARSE_NETWORK, COARSE_JACOBI, COARSE_EXACT = 0, 1, 2
# rs_*_by_bins_ex: a node's DOFs share an aggregate / one aggregate per component / rigid-body modes
AGG_NODES, AGG_COMPONENTS, AGG_RIGID = 0, 1, 2


NET_HID, NET_LAYERS, NET_IN = 64, 6, 4
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "heat_coarse512.npz")


def init_params(seed: int) -> np.ndarray:
    """Initial network parameters: He-scaled input and layer maps, zero biases and a zero
    output head, so that the untrained network computes coarse Jacobi (e = r_c / diag)."""
    rng = np.random.default_rng(seed)
    h = NET_HID
    parts = [rng.standard_normal(NET_IN * h) * np.sqrt(2.0 / NET_IN), np.zeros(h)]
    for _ in range(NET_LAYERS):
        parts += [rng.standard_normal(h * h) * 0.5 / np.sqrt(h), rng.standard_normal(h * h) * 0.5 / np.sqrt(h),
                  np.zeros(h)]
    parts += [np.zeros(h), np.zeros(1)]
    theta = np.concatenate(parts).astype(np.float32)
    assert theta.size == N_PARAMS
    return theta


def load_weights(path: str = WEIGHTS_PATH):
    """Trained parameters and their training summary (None, None when absent)."""
    if not os.path.exists(path):
        return None, None
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        summary = str(z["summary"]) if "summary" in z.files else None
    return params, summary


def aggregates_by_bins(xyz, book, dofs_per_node: int, bins) -> tuple[np.ndarray, int]:
    """Coordinate-bin aggregates of the enabled DOFs (bins = (nbx, nby, nbz))."""
    load()
    xyz = _f64(xyz).reshape(-1)
    book = _u32(book)
    n_dof = int((book != RS_EOD).sum())
    agg = np.empty(max(1, n_dof), np.uint32)
    n = C.c_uint32()
    _check(_lib.rs_aggregates_by_bins(xyz.size // 3, _ptr(xyz), dofs_per_node, _ptr(book), bins[0],
                                      bins[1], bins[2], _ptr(agg), C.byref(n)))
    return agg[:n_dof], n.value


def hierarchy_by_bins(xyz, book, dofs_per_node: int, bins, max_coarse: int = MAX_COARSE,
                      max_levels: int = 16, by_component: bool = False) -> list[np.ndarray]:
    """Nested coordinate bins, halved per level until a level has <= max_coarse nodes.
    Returns the level maps: maps[k] sends the nodes of level k (level 0: the DOFs) to level k+1.
    by_component: one aggregate per bin and DOF component (RS_AGG_COMPONENTS)."""
    load()
    xyz = _f64(xyz).reshape(-1)
    book = _u32(book)
    n_dof = int((book != RS_EOD).sum())
    cap = 2 * max(n_dof, 1) + 16
    agg = np.empty(cap, np.uint32)
    n_agg = np.empty(max_levels, np.uint32)
    nl = C.c_uint32()
    _check(_lib.rs_hierarchy_by_bins_ex(xyz.size // 3, _ptr(xyz), dofs_per_node, _ptr(book), bins[0],
                                        bins[1], bins[2], max_coarse, max_levels,
                                        AGG_COMPONENTS if by_component else AGG_NODES, _ptr(agg), cap,
                                        _ptr(n_agg), C.byref(nl)))
    maps, off, n = [], 0, n_dof
    for k in range(nl.value):
        maps.append(agg[off:off + n].copy())
        off += n
        n = int(n_agg[k])
    return maps


class LearnedPreconditioner:
    """V-cycle over an aggregation hierarchy with a learned coarsest solver (flexible CG).

    levels: a list of level maps (levels[k] sends the nodes of level k to those of level k+1),
    or a single map (one coarse level)."""

    def __init__(self, A: SparseMatrix, levels, n_agg=None, params=None):
        load()
        if isinstance(levels, np.ndarray) and levels.ndim == 1 and n_agg is not None and np.isscalar(n_agg):
            levels = [levels]
        maps = [_u32(m) for m in levels]
        sizes = [int(m.max()) + 1 for m in maps] if n_agg is None else \
            ([int(n_agg)] if np.isscalar(n_agg) else [int(v) for v in n_agg])
        self.A = A
        self._agg = np.ascontiguousarray(np.concatenate(maps)) if maps else np.zeros(1, np.uint32)
        self._nagg = np.asarray(sizes, np.uint32)
        self.h = C.c_void_p()
        _check(_lib.rs_learned_create(A.h, len(maps), _ptr(self._agg), _ptr(self._nagg), C.byref(self.h)))
        self._init_info(params)

    @classmethod
    def from_bins(cls, A: SparseMatrix, asm, bins, max_coarse: int = 512, max_levels: int = 16,
                  params=None, by_component: bool = False, rigid: bool = False) -> "LearnedPreconditioner":
        """The hierarchy of hierarchy_by_bins(xyz, book, ...) over the assembly's mesh, built on the
        GPU from the assembly's device coordinates and DOF book (no host work over the mesh).
        by_component: one aggregate per bin and DOF component (vector problems); rigid: node
        aggregates with their six rigid-body modes as coarse space (3-DOF elasticity)."""
        load()
        self = cls.__new__(cls)
        self.A = A
        xp, bp = _P(), _P()
        nn, dpn = C.c_uint32(), C.c_uint32()
        _check(_lib.rs_assembly_device_mesh(asm.h, C.byref(xp), C.byref(bp), C.byref(nn), C.byref(dpn)))
        self.h = C.c_void_p()
        _check(_lib.rs_learned_create_by_bins_ex(A.h, nn.value, xp, dpn.value, bp, int(bins[0]), int(bins[1]),
                                                 int(bins[2]), max_coarse, max_levels,
                                                 AGG_RIGID if rigid else AGG_COMPONENTS if by_component else AGG_NODES,
                                                 C.byref(self.h)))
        self._init_info(params)
        return self

    def level_basis(self, level: int) -> np.ndarray:
        """Rigid-body levels: the basis rows of level `level` (rows x 6)."""
        n = int(self.level_rows[level])
        out = np.empty(max(6 * n, 1))
        _check(_lib.rs_learned_level_basis(self.h, level, _ptr(out), out.size))
        return out[:6 * n].reshape(n, 6)

    def level_map(self, level: int) -> np.ndarray:
        """Map of level `level` (0 = the matrix rows) to the next level."""
        out = np.empty(max(int(self.level_rows[level]), 1), np.uint32)
        _check(_lib.rs_learned_level_map(self.h, level, _ptr(out), out.size))
        return out[:int(self.level_rows[level])]

    def _init_info(self, params) -> None:
        nl = C.c_uint32()
        rows = np.zeros(32, np.uint32)
        nnz = np.zeros(32, np.uint64)
        npar = C.c_uint64()
        _check(_lib.rs_learned_info(self.h, C.byref(nl), _ptr(rows), _ptr(nnz), C.byref(npar)))
        self.n_levels = nl.value
        self.level_rows = rows[:nl.value].copy()
        self.level_nnz = nnz[:nl.value].copy()
        self.n_coarse = int(self.level_rows[-1])
        self.n_params = npar.value
        if params is not None:
            self.set_params(params)

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(_lib.rs_learned_set_params(self.h, _ptr(p), p.size))

    def set_cycle(self, pre: int = 1, post: int = 1, coarse: int = COARSE_NETWORK) -> None:
        _check(_lib.rs_learned_set_cycle(self.h, pre, post, coarse))

    def set_coarse_cycle(self, pre: int = 1, post: int = 1) -> None:
        """Smoothing steps on the levels below the finest (set_cycle resets them to its own)."""
        _check(_lib.rs_learned_set_coarse_cycle(self.h, pre, post))

    def set_compressed(self, min_rows: int) -> None:
        """Levels with >= min_rows rows smooth with the FP16 copy of D^-1 A (0: none)."""
        _check(_lib.rs_learned_set_compressed(self.h, min_rows))

    def setup(self, A: SparseMatrix | None = None) -> None:
        _check(_lib.rs_learned_setup(self.h, (A or self.A).h))

    def level_lambda(self) -> np.ndarray:
        out = np.empty(self.n_levels)
        _check(_lib.rs_learned_level_lambda(self.h, _ptr(out)))
        return out

    def level_matrix(self, level: int):
        import scipy.sparse as sp
        n, nnz = int(self.level_rows[level]), int(self.level_nnz[level])
        rp = np.empty(n + 1, np.uint32)
        col = np.empty(max(nnz, 1), np.uint32)
        val = np.empty(max(nnz, 1))
        _check(_lib.rs_learned_level_matrix(self.h, level, _ptr(rp), _ptr(col), _ptr(val)))
        return sp.csr_matrix((val[:nnz], col[:nnz], rp), shape=(n, n))

    def coarse_matrix(self) -> np.ndarray:
        out = np.empty(self.n_coarse * self.n_coarse)
        _check(_lib.rs_learned_coarse_matrix(self.h, _ptr(out)))
        return out.reshape(self.n_coarse, self.n_coarse)

    def restrict(self, dr: DeviceArray) -> np.ndarray:
        out = np.empty(self.n_coarse, np.float32)
        _check(_lib.rs_learned_restrict_device(self.h, dr.ptr, _ptr(out)))
        return out

    def apply_coarse(self, rc) -> np.ndarray:
        r = np.ascontiguousarray(rc, dtype=np.float32)
        out = np.empty(self.n_coarse, np.float32)
        _check(_lib.rs_learned_apply_coarse(self.h, _ptr(r), _ptr(out)))
        return out

    def infer_timed(self, reps: int = 100) -> float:
        """Average inference kernel time (ms), back-to-back launches."""
        ms = C.c_double()
        _check(_lib.rs_learned_infer_timed(self.h, reps, C.byref(ms)))
        return ms.value

    def apply(self, r, A: SparseMatrix | None = None) -> np.ndarray:
        """One V-cycle z = B r."""
        rr = np.ascontiguousarray(r, dtype=np.float32)
        out = np.empty_like(rr)
        _check(_lib.rs_learned_apply(self.h, (A or self.A).h, _ptr(rr), _ptr(out)))
        return out

    def record(self, capacity: int) -> None:
        _check(_lib.rs_learned_record(self.h, capacity))

    def records(self, count: int) -> np.ndarray:
        out = np.empty((max(count, 0), self.n_coarse), np.float32)
        _check(_lib.rs_learned_records(self.h, count, _ptr(out)))
        return out

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.rs_learned_destroy(self.h)
            self.h = None


class Trainer:
    """GPU training of the network: forward/backward/Adam kernels on batches of coarse graphs."""

    def __init__(self, params, max_batch: int = 64, max_nodes: int = MAX_COARSE, max_nnz: int = 65536):
        load()
        p = np.ascontiguousarray(params, dtype=np.float32)
        self.max_batch = max_batch
        self.h = C.c_void_p()
        _check(_lib.rs_trainer_create(_ptr(p), p.size, max_batch, max_nodes, max_nnz, C.byref(self.h)))

    @staticmethod
    def dense_pattern(n: int):
        cptr = (np.arange(n + 1) * n).astype(np.uint32)
        ccol = np.tile(np.arange(n, dtype=np.uint32), n)
        return cptr, ccol

    def gradient_csr(self, cptr, ccol, cvals, rc, s):
        cptr, ccol = _u32(cptr), _u32(ccol)
        cv = np.ascontiguousarray(cvals, dtype=np.float32)
        R = np.ascontiguousarray(rc, dtype=np.float32)
        S = np.ascontiguousarray(s, dtype=np.float32)
        B, n = R.shape
        loss = np.empty(B, np.float32)
        grad = np.empty(N_PARAMS, np.float32)
        _check(_lib.rs_trainer_gradient(self.h, n, ccol.size, _ptr(cptr), _ptr(ccol), B, _ptr(cv), _ptr(R),
                                        _ptr(S), _ptr(loss), _ptr(grad)))
        return loss, grad

    def step_csr(self, cptr, ccol, cvals, rc, s, lr: float) -> float:
        cptr, ccol = _u32(cptr), _u32(ccol)
        cv = np.ascontiguousarray(cvals, dtype=np.float32)
        R = np.ascontiguousarray(rc, dtype=np.float32)
        S = np.ascontiguousarray(s, dtype=np.float32)
        B, n = R.shape
        m = C.c_float()
        _check(_lib.rs_trainer_step(self.h, n, ccol.size, _ptr(cptr), _ptr(ccol), B, _ptr(cv), _ptr(R), _ptr(S),
                                    lr, C.byref(m)))
        return m.value

    def gradient(self, Ac, rc, s):
        """Dense coarse operators Ac[B][n][n] (every entry of the n x n pattern)."""
        Ac = np.asarray(Ac, dtype=np.float32)
        cptr, ccol = self.dense_pattern(Ac.shape[1])
        return self.gradient_csr(cptr, ccol, Ac.reshape(Ac.shape[0], -1), rc, s)

    def step(self, Ac, rc, s, lr: float) -> float:
        Ac = np.asarray(Ac, dtype=np.float32)
        cptr, ccol = self.dense_pattern(Ac.shape[1])
        return self.step_csr(cptr, ccol, Ac.reshape(Ac.shape[0], -1), rc, s, lr)

    def params(self) -> np.ndarray:
        out = np.empty(N_PARAMS, np.float32)
        _check(_lib.rs_trainer_get_params(self.h, _ptr(out), out.size))
        return out

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.rs_trainer_destroy(self.h)
            self.h = None
