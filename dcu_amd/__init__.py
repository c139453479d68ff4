"""Python host mirror of libspai_mi355x (static SPAI + SPAI-preconditioned
BiCGSTAB on one MI355X).

Thin ctypes bindings over the C ABI declared in include/spai_mi355x.h.  The
computation always runs in the HIP library; importing this package does not
need a GPU, but every compute call fails loudly when the library is missing or
no HIP device is present.  There is no CPU fallback.

Interfaces mirrored (reference file:line):
  System.build     StaticSPAIv20 / StaticSPAIv20_MultiDCU
                   (src/spai_single_dcu.cpp:1224, src/spai_multi_dcu.cpp:1952)
  System.solve     cublas2_pbicgstabv2 (include/bicgstab/bicgstab_solver.h:33)
  Problem.read_mtx readMatrixToCSC (include/common/read.h:1296)
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DCU_AMD_LIBRARY may name another build of the same library
LIB_PATH = os.environ.get("DCU_AMD_LIBRARY") or os.path.join(_HERE, "_lib", "libspai_mi355x.so")
CLI_PATH = os.path.join(_HERE, "_bin", "spai_mi355x")

PRECISION_FP64 = 0
PRECISION_MIXED = 1
SCHEDULE_BALANCED_COLUMNS = 0
SCHEDULE_BALANCED_NNZ = 1

# M's sparsity pattern (the reference's GetSparsityPattern flags, plus A's own)
PATTERN_A = -1
PATTERN_AT = 0
PATTERN_I_PLUS_A = 1
PATTERN_I_PLUS_A_PLUS_AT = 2

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_NOMEM = -3
ERR_UNSUPPORTED = -4
ERR_BREAKDOWN = -5
ERR_NOT_BUILT = -6

SOLVE_RUNNING = 0
SOLVE_CONVERGED = 1
SOLVE_MAX_ITER = 2
SOLVE_BREAKDOWN = 3
SOLVE_CONVERGED_AT_START = 4

# every symbol include/spai_mi355x.h and include/spai_multi_dcu_api.h declare
C_SYMBOLS = [
    "spai_status_string", "spai_last_error", "spai_version",
    "spai_context_create", "spai_context_destroy", "spai_context_device_info",
    "spai_system_create", "spai_system_create_batched", "spai_system_create_device",
    "spai_system_destroy", "spai_system_get_info", "spai_system_row_offsets",
    "spai_system_set_values", "spai_system_build", "spai_system_get_m",
    "spai_system_get_row_set_sizes", "spai_solve_options_default", "spai_system_solve",
    "spai_system_outcomes", "spai_system_solve_noalloc", "spai_system_residual_history", "spai_system_adjoint", "spai_system_time_spmv",
    "spai_system_time_iteration", "spai_system_layout", "spai_system_set_pattern", "spai_system_m_pattern",
    "spai_system_quality", "spai_system_product",
    "spai_tile_shape_count", "spai_tile_shape_get", "spai_system_autotune", "spai_system_set_tile_shapes",
    "spai_system_get_tile_shapes",
    "spai_calibration_copy", "spai_csc_to_csr_device",
    "spai_problem_generate", "spai_problem_generate_power_law", "spai_problem_generate_batched",
    "spai_problem_generate_grf", "spai_problem_convection_diffusion", "spai_problem_read_mtx", "spai_problem_free",
    "spai_problem_num_systems", "spai_problem_system", "spai_problem_rhs", "spai_problem_target",
    "spai_problem_name", "spai_problem_notes",
]
# C++ drop-ins with the reference's signatures (Itanium-mangled names)
CXX_SYMBOLS = {
    "initMultiDCU": "_Z12initMultiDCUi",
    "cleanupMultiDCU": "_Z15cleanupMultiDCUP16MultiDCU_Context",
    "RunStaticSPAI_MultiDCU_MPI": "_Z26RunStaticSPAI_MultiDCU_MPIP16MultiDCU_ContextPK10CSC_MatrixPS1_ii",
    "RunDynamicSPAI_MultiDCU_MPI": "_Z27RunDynamicSPAI_MultiDCU_MPIP16MultiDCU_ContextPK10CSC_MatrixPS1_ii",
    "cublas2_pbicgstabv2": "_Z19cublas2_pbicgstabv2P10CSR_MatrixS0_PdS1_di",
    "cuCSC2CSR": "_Z9cuCSC2CSRiiiPdPiS0_S_S0_S0_",
}


class SpaiError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status


class CSC_Matrix(ctypes.Structure):
    """Layout of CSC_Matrix / CSR_Matrix (include/common/dataType.h:24-42)."""
    _fields_ = [
        ("mIndex", ctypes.POINTER(ctypes.c_int)),
        ("mData", ctypes.POINTER(ctypes.c_double)),
        ("mPtr", ctypes.POINTER(ctypes.c_int)),
        ("nRow", ctypes.c_int),
        ("nCol", ctypes.c_int),
        ("n", ctypes.c_int),
        ("nonzeroes", ctypes.c_int),
    ]


CSR_Matrix = CSC_Matrix


class SystemInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_int), ("nsys", ctypes.c_int),
                ("m_precision", ctypes.c_int), ("setup_ms", ctypes.c_float)]


class BuildReport(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_float), ("tier_columns", ctypes.c_int * 5), ("max_gathered", ctypes.c_int),
                ("max_n2", ctypes.c_int), ("empty_columns", ctypes.c_int), ("sum_gathered", ctypes.c_longlong),
                ("static_tier", ctypes.c_int), ("tile_shape", ctypes.c_int * 3), ("phase_ms", ctypes.c_float * 4)]


TUNE_TIERS = 3
TUNE_MAX_CANDIDATES = 8


class TileShape(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_int), ("entries_per_lane", ctypes.c_int), ("max_cols", ctypes.c_int),
                ("rows_per_lane", ctypes.c_int)]


class Tuning(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int * TUNE_TIERS), ("columns", ctypes.c_int * TUNE_TIERS),
                ("ms", (ctypes.c_float * TUNE_MAX_CANDIDATES) * TUNE_TIERS), ("tune_ms", ctypes.c_float)]


class SolveOptions(ctypes.Structure):
    _fields_ = [("rel_tol", ctypes.c_double), ("abs_tol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("graph_iters", ctypes.c_int), ("check_every", ctypes.c_int),
                ("precondition", ctypes.c_int), ("record_history", ctypes.c_int),
                ("true_residual_restart", ctypes.c_int)]


class SolveReport(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_float), ("iterations_max", ctypes.c_int), ("iterations_sum", ctypes.c_longlong),
                ("host_checks", ctypes.c_int), ("all_converged", ctypes.c_int)]


class Outcome(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int), ("residual", ctypes.c_double),
                ("residual0", ctypes.c_double), ("true_residual", ctypes.c_double), ("tol", ctypes.c_double),
                ("restarts", ctypes.c_int)]


_lib: Optional[ctypes.CDLL] = None


def library_path() -> str:
    return LIB_PATH


def load_library() -> ctypes.CDLL:
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SpaiError(-6, f"libspai_mi355x.so not built at {LIB_PATH}: run __graft_entry__.build()")
    # One HIP runtime per process: torch ships its own libamdhip64; loaded
    # after this library, it would start a second runtime that cannot open the
    # GPU.  Importing torch first makes this library bind to torch's runtime.
    if os.environ.get("DCU_AMD_NO_TORCH") != "1":
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    lib = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    sig = {
        "spai_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "spai_last_error": (ctypes.c_char_p, []),
        "spai_version": (ctypes.c_char_p, []),
        "spai_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "spai_context_destroy": (ctypes.c_int, [vp]),
        "spai_context_device_info": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int, ip]),
        "spai_system_create": (ctypes.c_int, [vp, ctypes.POINTER(CSC_Matrix), ctypes.POINTER(vp)]),
        "spai_system_create_batched": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(CSC_Matrix),
                                                      ctypes.POINTER(vp)]),
        "spai_system_create_device": (ctypes.c_int, [vp, ctypes.POINTER(CSC_Matrix), ctypes.POINTER(vp)]),
        "spai_system_destroy": (ctypes.c_int, [vp]),
        "spai_system_get_info": (ctypes.c_int, [vp, ctypes.POINTER(SystemInfo)]),
        "spai_system_row_offsets": (ctypes.c_int, [vp, ip]),
        "spai_system_set_values": (ctypes.c_int, [vp, vp, ctypes.c_int]),
        "spai_system_build": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(BuildReport)]),
        "spai_system_get_m": (ctypes.c_int, [vp, vp, ctypes.c_int]),
        "spai_system_get_row_set_sizes": (ctypes.c_int, [vp, ip]),
        "spai_solve_options_default": (None, [ctypes.POINTER(SolveOptions)]),
        "spai_system_solve": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.POINTER(SolveOptions),
                                             ctypes.POINTER(SolveReport)]),
        "spai_system_solve_noalloc": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(SolveOptions),
                                                     ctypes.POINTER(SolveReport)]),
        "spai_system_outcomes": (ctypes.c_int, [vp, ctypes.POINTER(Outcome)]),
        "spai_system_residual_history": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_int]),
        "spai_system_adjoint": (ctypes.c_int, [vp, vp, vp, vp, vp, ctypes.c_int, ctypes.POINTER(SolveOptions),
                                               ctypes.POINTER(SolveReport)]),
        "spai_system_time_spmv": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
        "spai_system_time_iteration": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int,
                                                      ctypes.POINTER(ctypes.c_float)]),
        "spai_system_layout": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong)]),
        "spai_system_set_pattern": (ctypes.c_int, [vp, ctypes.c_int]),
        "spai_tile_shape_count": (ctypes.c_int, [ctypes.c_int]),
        "spai_tile_shape_get": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(TileShape)]),
        "spai_system_autotune": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.POINTER(Tuning)]),
        "spai_system_set_tile_shapes": (ctypes.c_int, [vp, ip]),
        "spai_system_get_tile_shapes": (ctypes.c_int, [vp, ip]),
        "spai_system_quality": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), vp]),
        "spai_system_product": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_longlong), vp, vp, vp]),
        "spai_system_m_pattern": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int), vp, vp]),
        "spai_calibration_copy": (ctypes.c_int, [vp, ctypes.c_longlong, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_float)]),
        "spai_csc_to_csr_device": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp,
                                                  vp]),
        "spai_problem_generate": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]),
        "spai_problem_generate_power_law": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]),
        "spai_problem_generate_batched": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]),
        "spai_problem_generate_grf": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]),
        "spai_problem_convection_diffusion": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                             ctypes.c_double, ctypes.POINTER(vp)]),
        "spai_problem_read_mtx": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(vp)]),
        "spai_problem_free": (ctypes.c_int, [vp]),
        "spai_problem_num_systems": (ctypes.c_int, [vp]),
        "spai_problem_system": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(CSC_Matrix)]),
        "spai_problem_rhs": (dp, [vp, ctypes.c_int]),
        "spai_problem_target": (dp, [vp, ctypes.c_int]),
        "spai_problem_name": (ctypes.c_char_p, [vp]),
        "spai_problem_notes": (ctypes.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def tile_shape_count(tier: int) -> int:
    """Candidate group shapes of register tier `tier` (0..2)."""
    return int(load_library().spai_tile_shape_count(int(tier)))


def tile_shape(tier: int, candidate: int) -> dict:
    """Lanes per column, gathered entries per lane, largest |J_k| and tile
    rows per lane of one candidate."""
    t = TileShape()
    _check(load_library().spai_tile_shape_get(int(tier), int(candidate), ctypes.byref(t)))
    return {"lanes": t.lanes, "entries_per_lane": t.entries_per_lane, "max_cols": t.max_cols,
            "rows_per_lane": t.rows_per_lane}


def _check(status: int) -> None:
    if status != 0:
        lib = load_library()
        raise SpaiError(status, f"{lib.spai_status_string(status).decode()}: {lib.spai_last_error().decode()}")


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None else a.ctypes.data


def csc_struct(n: int, colptr: np.ndarray, rowidx: np.ndarray, val: np.ndarray) -> CSC_Matrix:
    m = CSC_Matrix()
    m.n = m.nRow = m.nCol = n
    m.nonzeroes = int(colptr[n])
    m.mPtr = colptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    m.mIndex = rowidx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    m.mData = val.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return m


@dataclass
class CSC:
    """Host CSC matrix (0-based, square) as numpy arrays."""
    n: int
    colptr: np.ndarray
    rowidx: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.colptr[self.n])

    def struct(self) -> CSC_Matrix:
        self.colptr = np.ascontiguousarray(self.colptr, dtype=np.int32)
        self.rowidx = np.ascontiguousarray(self.rowidx, dtype=np.int32)
        self.val = np.ascontiguousarray(self.val, dtype=np.float64)
        return csc_struct(self.n, self.colptr, self.rowidx, self.val)


class Problem:
    """Seeded synthetic workloads and MatrixMarket input (host side)."""

    def __init__(self, handle: int):
        self._h = ctypes.c_void_p(handle)
        self._lib = load_library()

    @classmethod
    def generate(cls, config: int, seed: int) -> "Problem":
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.spai_problem_generate(int(config), int(seed), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def power_law(cls, n: int, seed: int) -> "Problem":
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.spai_problem_generate_power_law(int(n), int(seed), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def batched(cls, nsys: int, seed: int) -> "Problem":
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.spai_problem_generate_batched(int(nsys), int(seed), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def grf(cls, m: int, seed: int) -> "Problem":
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.spai_problem_generate_grf(int(m), int(seed), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def convection_diffusion(cls, m: int, peclet: float, vx: float = 1.0, vy: float = 0.5) -> "Problem":
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.spai_problem_convection_diffusion(int(m), float(peclet), float(vx), float(vy), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def read_mtx(cls, path: str) -> "Problem":
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.spai_problem_read_mtx(path.encode(), ctypes.byref(h)))
        return cls(h.value)

    def __del__(self):
        try:
            if getattr(self, "_h", None) and self._h.value:
                self._lib.spai_problem_free(self._h)
                self._h = None
        except Exception:  # interpreter shutdown
            pass

    @property
    def name(self) -> str:
        return self._lib.spai_problem_name(self._h).decode()

    @property
    def notes(self) -> str:
        return self._lib.spai_problem_notes(self._h).decode()

    @property
    def num_systems(self) -> int:
        return self._lib.spai_problem_num_systems(self._h)

    def system(self, s: int = 0) -> CSC:
        m = CSC_Matrix()
        _check(self._lib.spai_problem_system(self._h, int(s), ctypes.byref(m)))
        n, nnz = m.n, m.nonzeroes
        colptr = np.ctypeslib.as_array(m.mPtr, shape=(n + 1,)).copy()
        rowidx = np.ctypeslib.as_array(m.mIndex, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(m.mData, shape=(max(nnz, 1),))[:nnz].copy()
        return CSC(n, colptr, rowidx, val)

    def systems(self) -> List[CSC]:
        return [self.system(s) for s in range(self.num_systems)]

    def _vec(self, f, s: int, n: int) -> Optional[np.ndarray]:
        p = f(self._h, int(s))
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def rhs(self, s: int = 0) -> np.ndarray:
        return self._vec(self._lib.spai_problem_rhs, s, self.system(s).n)

    def target(self, s: int = 0) -> Optional[np.ndarray]:
        return self._vec(self._lib.spai_problem_target, s, self.system(s).n)


class Context:
    """One HIP device and stream."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = ctypes.c_void_p()
        _check(self._lib.spai_context_create(int(device), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h and self._h.value:
            self._lib.spai_context_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def calibration_copy(self, nbytes: int, reps: int = 10) -> float:
        """Average ms of a plain device copy reading and writing nbytes."""
        ms = ctypes.c_float()
        _check(self._lib.spai_calibration_copy(self._h, int(nbytes), int(reps), ctypes.byref(ms)))
        return ms.value

    def device_info(self) -> Tuple[str, int]:
        buf = ctypes.create_string_buffer(128)
        cus = ctypes.c_int()
        _check(self._lib.spai_context_device_info(self._h, buf, 128, ctypes.byref(cus)))
        return buf.value.decode(), cus.value


def _options(rel_tol=1e-8, abs_tol=0.0, max_iter=1000, use_graph=False, graph_iters=8, check_every=1,
             precondition=True, record_history=False, true_residual_restart=False) -> SolveOptions:
    o = SolveOptions()
    o.rel_tol = float(rel_tol)
    o.abs_tol = float(abs_tol)
    o.max_iter = int(max_iter)
    o.use_graph = int(bool(use_graph))
    o.graph_iters = int(graph_iters)
    o.check_every = int(check_every)
    o.precondition = int(bool(precondition))
    o.record_history = int(bool(record_history))
    o.true_residual_restart = int(bool(true_residual_restart))
    return o


class System:
    """A matrix (or a block-diagonal batch of matrices) resident on the device."""

    def __init__(self, ctx: Context, matrices: Sequence[CSC] | CSC):
        self._ctx = ctx
        self._lib = ctx._lib
        h = ctypes.c_void_p()
        if isinstance(matrices, CSC):
            self._keep = [matrices]
            m = matrices.struct()
            _check(self._lib.spai_system_create(ctx._h, ctypes.byref(m), ctypes.byref(h)))
        else:
            self._keep = list(matrices)
            arr = (CSC_Matrix * len(self._keep))(*[c.struct() for c in self._keep])
            _check(self._lib.spai_system_create_batched(ctx._h, len(self._keep), arr, ctypes.byref(h)))
        self._h = h
        info = SystemInfo()
        _check(self._lib.spai_system_get_info(self._h, ctypes.byref(info)))
        self.n, self.nnz, self.nsys, self.setup_ms = info.n, info.nnz, info.nsys, info.setup_ms
        off = np.zeros(self.nsys + 1, dtype=np.int32)
        _check(self._lib.spai_system_row_offsets(self._h, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        self.row_offsets = off

    @classmethod
    def from_device(cls, ctx: Context, n: int, colptr_ptr: int, rowidx_ptr: int, val_ptr: int, nnz: int) -> "System":
        self = cls.__new__(cls)
        self._ctx = ctx
        self._lib = ctx._lib
        self._keep = []
        m = CSC_Matrix()
        m.n = m.nRow = m.nCol = n
        m.nonzeroes = nnz
        m.mPtr = ctypes.cast(ctypes.c_void_p(colptr_ptr), ctypes.POINTER(ctypes.c_int))
        m.mIndex = ctypes.cast(ctypes.c_void_p(rowidx_ptr), ctypes.POINTER(ctypes.c_int))
        m.mData = ctypes.cast(ctypes.c_void_p(val_ptr), ctypes.POINTER(ctypes.c_double))
        h = ctypes.c_void_p()
        _check(self._lib.spai_system_create_device(ctx._h, ctypes.byref(m), ctypes.byref(h)))
        self._h = h
        self.n, self.nnz, self.nsys, self.setup_ms = n, nnz, 1, 0.0
        self.row_offsets = np.array([0, n], dtype=np.int32)
        return self

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.spai_system_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_values(self, values, on_device: bool = False) -> None:
        if on_device:
            _check(self._lib.spai_system_set_values(self._h, ctypes.c_void_p(int(values)), 1))
        else:
            v = np.ascontiguousarray(values, dtype=np.float64)
            assert v.size == self.nnz
            _check(self._lib.spai_system_set_values(self._h, v.ctypes.data, 0))

    def build(self, precision: int = PRECISION_FP64, schedule: int = SCHEDULE_BALANCED_NNZ,
              col_begin: int = 0, col_end: int = -1) -> dict:
        rep = BuildReport()
        _check(self._lib.spai_system_build(self._h, int(precision), int(schedule), int(col_begin), int(col_end),
                                           ctypes.byref(rep)))
        return {"ms": rep.ms, "tier_columns": list(rep.tier_columns), "max_gathered": rep.max_gathered,
                "max_n2": rep.max_n2, "empty_columns": rep.empty_columns, "sum_gathered": rep.sum_gathered,
                "static_tier": rep.static_tier, "tile_shape": list(rep.tile_shape),
                "phase_ms": dict(zip(("plan", "register_tiers", "lds_tier", "global_tier"),
                                     [float(v) for v in rep.phase_ms]))}

    def autotune(self, precision: int = PRECISION_MIXED, schedule: int = SCHEDULE_BALANCED_NNZ,
                 reps: int = 5) -> dict:
        """Times every candidate group shape of each populated register tier
        on this system's columns and keeps the fastest; M is left built with
        them.  Returns the chosen shape per tier, the columns per tier and the
        best tier-kernel ms per candidate (None where not timed)."""
        t = Tuning()
        _check(self._lib.spai_system_autotune(self._h, int(precision), int(schedule), int(reps), ctypes.byref(t)))
        return {"shape": list(t.shape), "columns": list(t.columns), "tune_ms": t.tune_ms,
                "ms": [[(v if v >= 0 else None) for v in t.ms[i][:tile_shape_count(i)]] for i in range(TUNE_TIERS)]}

    def set_tile_shapes(self, shape) -> None:
        a = np.ascontiguousarray(shape, dtype=np.int32)
        assert a.size == TUNE_TIERS
        _check(self._lib.spai_system_set_tile_shapes(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def tile_shapes(self) -> List[int]:
        a = np.zeros(TUNE_TIERS, dtype=np.int32)
        _check(self._lib.spai_system_get_tile_shapes(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return [int(v) for v in a]

    def set_pattern(self, pattern: int) -> None:
        """M's pattern for the next builds: PATTERN_A (default), PATTERN_AT,
        PATTERN_I_PLUS_A or PATTERN_I_PLUS_A_PLUS_AT."""
        _check(self._lib.spai_system_set_pattern(self._h, int(pattern)))

    def m_pattern(self) -> Tuple[np.ndarray, np.ndarray]:
        """M's pattern as CSC (colptr, rowidx)."""
        nnz = ctypes.c_int()
        _check(self._lib.spai_system_m_pattern(self._h, ctypes.byref(nnz), None, None))
        cp = np.empty(self.n + 1, dtype=np.int32)
        ri = np.empty(max(nnz.value, 1), dtype=np.int32)
        _check(self._lib.spai_system_m_pattern(self._h, None, cp.ctypes.data, ri.ctypes.data))
        return cp, ri[: nnz.value]

    def quality(self) -> Tuple[float, np.ndarray]:
        """(||A M - I||_F, ||A m_k - e_k||_2 per column) of the built M."""
        f = ctypes.c_double()
        cols = np.empty(self.n)
        _check(self._lib.spai_system_quality(self._h, ctypes.byref(f), cols.ctypes.data))
        return float(f.value), cols

    def product(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """C = A M as CSC (colptr, rowidx, values)."""
        nnz = ctypes.c_longlong()
        _check(self._lib.spai_system_product(self._h, ctypes.byref(nnz), None, None, None))
        cp = np.empty(self.n + 1, dtype=np.int32)
        ri = np.empty(max(nnz.value, 1), dtype=np.int32)
        va = np.empty(max(nnz.value, 1))
        _check(self._lib.spai_system_product(self._h, ctypes.byref(nnz), cp.ctypes.data, ri.ctypes.data,
                                             va.ctypes.data))
        return cp, ri[: nnz.value], va[: nnz.value]

    def m_values(self) -> np.ndarray:
        """M's values in its pattern's CSC order (A's CSC order by default)."""
        nnz = ctypes.c_int()
        _check(self._lib.spai_system_m_pattern(self._h, ctypes.byref(nnz), None, None))
        out = np.empty(max(nnz.value, 1), dtype=np.float64)
        _check(self._lib.spai_system_get_m(self._h, out.ctypes.data, 0))
        return out[: nnz.value]

    def row_set_sizes(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        _check(self._lib.spai_system_get_row_set_sizes(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def outcomes(self) -> List[dict]:
        arr = (Outcome * self.nsys)()
        _check(self._lib.spai_system_outcomes(self._h, arr))
        return [{"status": o.status, "iterations": o.iterations, "residual": o.residual,
                 "residual0": o.residual0, "true_residual": o.true_residual, "tol": o.tol,
                 "restarts": o.restarts} for o in arr]

    def residual_history(self, s: int = 0, maxlen: int = 100000) -> np.ndarray:
        out = np.empty(maxlen)
        k = self._lib.spai_system_residual_history(self._h, int(s), out.ctypes.data, int(maxlen))
        if k < 0:
            _check(k)
        return out[:k]

    def solve(self, b: np.ndarray, x0: Optional[np.ndarray] = None, **kw) -> Tuple[np.ndarray, dict, List[dict]]:
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        opt = _options(**kw)
        rep = SolveReport()
        _check(self._lib.spai_system_solve(self._h, b.ctypes.data, x.ctypes.data, 0, ctypes.byref(opt),
                                           ctypes.byref(rep)))
        return x, _report(rep), self.outcomes()

    def solve_device(self, b_ptr: int, x_ptr: int, **kw) -> dict:
        opt = _options(**kw)
        rep = SolveReport()
        _check(self._lib.spai_system_solve(self._h, ctypes.c_void_p(int(b_ptr)), ctypes.c_void_p(int(x_ptr)), 1,
                                           ctypes.byref(opt), ctypes.byref(rep)))
        return _report(rep)

    def solve_noalloc(self, b_ptr: int, x_ptr: int, **kw) -> dict:
        """spai_system_solve_noalloc on device pointers: the solve with its
        work memory in the library's static device scratch (no device
        allocation); no graphs or histories."""
        opt = _options(**kw)
        rep = SolveReport()
        _check(self._lib.spai_system_solve_noalloc(self._h, ctypes.c_void_p(int(b_ptr)), ctypes.c_void_p(int(x_ptr)),
                                                   ctypes.byref(opt), ctypes.byref(rep)))
        return _report(rep)

    def adjoint(self, x: np.ndarray, g: np.ndarray, **kw) -> Tuple[np.ndarray, np.ndarray, dict]:
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        dA = np.empty(max(self.nnz, 1))
        db = np.empty(self.n)
        opt = _options(**kw)
        rep = SolveReport()
        _check(self._lib.spai_system_adjoint(self._h, x.ctypes.data, g.ctypes.data, dA.ctypes.data, db.ctypes.data,
                                             0, ctypes.byref(opt), ctypes.byref(rep)))
        return dA[: self.nnz], db, _report(rep)

    def adjoint_device(self, x_ptr: int, g_ptr: int, dA_ptr: int, db_ptr: int, **kw) -> dict:
        opt = _options(**kw)
        rep = SolveReport()
        vp = ctypes.c_void_p
        _check(self._lib.spai_system_adjoint(self._h, vp(int(x_ptr)), vp(int(g_ptr)), vp(int(dA_ptr)),
                                             vp(int(db_ptr)), 1, ctypes.byref(opt), ctypes.byref(rep)))
        return _report(rep)

    def time_spmv(self, which: int = 0, reps: int = 20) -> float:
        ms = ctypes.c_float()
        _check(self._lib.spai_system_time_spmv(self._h, int(which), int(reps), ctypes.byref(ms)))
        return ms.value

    def time_iteration(self, reps: int = 20, variant: int = -1) -> List[float]:
        """Average ms per kernel over `reps` consecutive all-active iterations:
        [K1 or its streaming half, its M-apply half (0 if fused), K2, K3, K4];
        variant 0 fused K1, 1 split K1, -1 the solver's choice."""
        ms = (ctypes.c_float * 5)()
        _check(self._lib.spai_system_time_iteration(self._h, int(reps), int(variant), ms))
        return [float(v) for v in ms]

    def layout(self) -> Tuple[str, int]:
        """("sell64" | "sell64_off16" | "sell64_diag" | "row_tiles" |
        "row_tiles_blocked", stored entries per operator).  sell64_off16
        stores 16-bit column - row offsets, sell64_diag one offset per column
        block of a slice; row_tiles_blocked splits the products by column
        blocks."""
        kind, entries = ctypes.c_int(), ctypes.c_longlong()
        _check(self._lib.spai_system_layout(self._h, ctypes.byref(kind), ctypes.byref(entries)))
        names = {0: "row_tiles", 1: "sell64", 2: "sell64_off16", 3: "sell64_diag", 4: "row_tiles_blocked"}
        return names[kind.value], int(entries.value)


def _report(rep: SolveReport) -> dict:
    return {"ms": rep.ms, "iterations_max": rep.iterations_max, "iterations_sum": rep.iterations_sum,
            "host_checks": rep.host_checks, "all_converged": bool(rep.all_converged)}


def csc_to_csr_host(A: CSC) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Host CSC -> CSR (rows ordered by column) for building test operands."""
    n = A.n
    counts = np.bincount(A.rowidx, minlength=n)
    rowptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(counts, out=rowptr[1:])
    cols = np.repeat(np.arange(n, dtype=np.int32), np.diff(A.colptr))
    order = np.lexsort((cols, A.rowidx))
    return rowptr, cols[order].astype(np.int32), A.val[order]
