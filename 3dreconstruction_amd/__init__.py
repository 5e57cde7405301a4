"""MI355X bundle-adjustment engine: Python host API over the C ABI (``include/mba.h``).

The package name starts with a digit, so import it with
``importlib.import_module("3dreconstruction_amd")``.

Two layers:

* ``BAProblem`` / ``solve`` / ``evaluate`` / ``schur_structure`` / ``Context``
  -- array-level calls straight onto the C ABI;
* ``Problem`` / ``SolverOptions`` / ``Solve`` / ``Summary`` / ``HuberLoss`` --
  the shape of the Ceres calls libs/sfm makes for bundle adjustment
  (libs/sfm/src/sfm_incremental_engine.cpp:1413-1483): add one residual block
  per observation with a shared loss, mark camera blocks constant, set the
  solver options, solve, read ``initial_cost`` / ``final_cost`` /
  ``IsSolutionUsable()``; parameter blocks are numpy arrays updated in place
  when the solution is usable.

The engine runs only through ``libmba.so`` (HIP, gfx950).  If the library is
missing every call raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmba.so")

# ------------------------------------------------------------- constants --
OK = 0
ERR_INVALID_ARGUMENT, ERR_UNSUPPORTED, ERR_DEVICE, ERR_OUT_OF_MEMORY, ERR_NOT_FINITE = -1, -2, -3, -4, -5
CONVERGENCE, NO_CONVERGENCE, FAILURE = 0, 1, 2
LAYOUT_PACKED9 = 0
LAYOUT_GROUPED = 1
ORDER_AUTOMATIC, ORDER_POINTS_FIRST = 0, 1
PRECOND_JACOBI, PRECOND_SCHUR_JACOBI, PRECOND_CLUSTER_JACOBI, PRECOND_CLUSTER_TRIDIAGONAL = 0, 1, 2, 3
EXPLICIT_SCHUR, IMPLICIT_SCHUR_PCG = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1

# Ceres names for the linear solvers libs/sfm selects.
SPARSE_SCHUR = EXPLICIT_SCHUR
DENSE_SCHUR = EXPLICIT_SCHUR
ITERATIVE_SCHUR = IMPLICIT_SCHUR_PCG
JACOBI, SCHUR_JACOBI, CLUSTER_JACOBI, CLUSTER_TRIDIAGONAL = (PRECOND_JACOBI, PRECOND_SCHUR_JACOBI,
                                                           PRECOND_CLUSTER_JACOBI, PRECOND_CLUSTER_TRIDIAGONAL)

_TERMINATION_NAMES = {CONVERGENCE: "CONVERGENCE", NO_CONVERGENCE: "NO_CONVERGENCE", FAILURE: "FAILURE"}

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class MbaError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"mba status {status}: {message}")
        self.status = status


class _Problem(C.Structure):
    _fields_ = [("layout", C.c_int32), ("num_cameras", C.c_int32), ("num_points", C.c_int32),
                ("num_intrinsics", C.c_int32), ("num_observations", C.c_int64), ("cameras", _f64p),
                ("points", _f64p), ("observations", _f64p), ("camera_index", _i32p),
                ("point_index", _i32p), ("camera_constant", _u8p), ("intrinsics", _f64p),
                ("intrinsic_index", _i32p), ("intrinsic_constant", _u8p), ("intrinsic_subset", _u8p),
                ("camera_params", C.c_int32), ("ordering", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("max_num_iterations", C.c_int32), ("max_num_consecutive_invalid_steps", C.c_int32),
                ("max_linear_solver_iterations", C.c_int32), ("min_linear_solver_iterations", C.c_int32),
                ("jacobi_scaling", C.c_int32), ("linear_solver", C.c_int32), ("precision", C.c_int32),
                ("device", C.c_int32), ("function_tolerance", C.c_double),
                ("gradient_tolerance", C.c_double), ("parameter_tolerance", C.c_double),
                ("initial_trust_region_radius", C.c_double), ("max_trust_region_radius", C.c_double),
                ("min_trust_region_radius", C.c_double), ("min_relative_decrease", C.c_double),
                ("min_lm_diagonal", C.c_double), ("max_lm_diagonal", C.c_double), ("eta", C.c_double),
                ("pcg_r_tolerance", C.c_double), ("huber_a", C.c_double),
                ("refinement_iterations", C.c_int32), ("preconditioner", C.c_int32),
                ("dump_directory", C.c_char_p), ("dump_iterations", C.POINTER(C.c_int32)),
                ("num_dump_iterations", C.c_int32), ("point_rows_per_lane", C.c_int32),
                ("camera_chunk_rows", C.c_int32), ("pcg_graph_launch", C.c_int32)]

    def set_dump(self, directory, iterations):
        """trust_region_minimizer_iterations_to_dump / trust_region_problem_dump_directory
        (TEXTFILE).  The arrays are kept alive by this object."""
        its = np.ascontiguousarray(list(iterations), np.int32)
        self._dump_keep = (os.fsencode(directory) if directory is not None else None, its)
        self.dump_directory = self._dump_keep[0]
        self.dump_iterations = its.ctypes.data_as(C.POINTER(C.c_int32)) if its.size else None
        self.num_dump_iterations = int(its.size)


class Iteration(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("step_is_valid", C.c_int32), ("step_is_successful", C.c_int32),
                ("linear_solver_iterations", C.c_int32), ("cost", C.c_double), ("cost_change", C.c_double),
                ("gradient_max_norm", C.c_double), ("gradient_norm", C.c_double),
                ("step_norm", C.c_double), ("relative_decrease", C.c_double),
                ("trust_region_radius", C.c_double), ("cumulative_time_s", C.c_double)]


class _Summary(C.Structure):
    _fields_ = [("initial_cost", C.c_double), ("final_cost", C.c_double), ("fixed_cost", C.c_double),
                ("preprocessor_time_s", C.c_double), ("jacobian_evaluation_time_s", C.c_double),
                ("residual_evaluation_time_s", C.c_double), ("linear_solver_time_s", C.c_double),
                ("total_time_s", C.c_double), ("num_successful_steps", C.c_int32),
                ("num_unsuccessful_steps", C.c_int32), ("num_iterations", C.c_int32),
                ("termination_type", C.c_int32), ("num_e_blocks", C.c_int32), ("num_f_blocks", C.c_int32),
                ("num_linear_solver_iterations", C.c_int32), ("is_solution_usable", C.c_int32)]


# ------------------------------------------------------------- library --
_lib = None

#: every exported symbol of include/mba.h
EXPORTS = [
    "mba_default_options", "mba_solve", "mba_solve_batched", "mba_evaluate", "mba_evaluate_blocks",
    "mba_schur_structure",
    "mba_context_create", "mba_context_destroy", "mba_context_jacobian_schur",
    "mba_context_read_schur", "mba_context_read_scale", "mba_context_schur_matvec",
    "mba_context_solve", "mba_context_download", "mba_context_synchronize", "mba_context_stream",
    "mba_context_set_timing", "mba_context_last_timing", "mba_status_string", "mba_last_error",
    "mba_abi_version", "mba_scene_generate", "mba_scene_generate_grouped",
    "mba_rotation_matrix_to_angle_axis", "mba_angle_axis_to_rotation_matrix", "mba_read_libmv",
    "mba_read_bal", "mba_write_bal", "mba_evaluate_noalloc", "mba_context_lm_pass",
]


def library_path() -> str:
    return LIB_PATH


def load_library():
    """Load libmba.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(python -m 3dreconstruction_amd.build is not importable; run build.py)")
    lib = C.CDLL(LIB_PATH)
    lib.mba_status_string.restype = C.c_char_p
    lib.mba_last_error.restype = C.c_char_p
    lib.mba_context_stream.restype = C.c_void_p
    lib.mba_context_create.argtypes = [C.POINTER(_Problem), C.POINTER(Options), C.POINTER(C.c_void_p)]
    lib.mba_solve_batched.argtypes = [C.POINTER(Options), C.POINTER(_Problem), C.c_int32, C.POINTER(_Summary)]
    for name in ["mba_context_destroy", "mba_context_synchronize"]:
        getattr(lib, name).argtypes = [C.c_void_p]
    lib.mba_context_jacobian_schur.argtypes = [C.c_void_p, C.c_double]
    lib.mba_context_lm_pass.argtypes = [C.c_void_p, C.c_double, _f64p]
    lib.mba_context_set_timing.argtypes = [C.c_void_p, C.c_int32]
    lib.mba_context_last_timing.argtypes = [C.c_void_p, _f64p, C.c_int32]
    lib.mba_context_read_schur.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p, _f64p]
    lib.mba_context_read_scale.argtypes = [C.c_void_p, _f64p, _f64p]
    lib.mba_context_schur_matvec.argtypes = [C.c_void_p, _f64p, _f64p]
    lib.mba_context_solve.argtypes = [C.c_void_p, C.POINTER(_Summary), C.POINTER(Iteration), C.c_int32]
    lib.mba_context_download.argtypes = [C.c_void_p, C.POINTER(_Problem)]
    lib.mba_context_stream.argtypes = [C.c_void_p]
    lib.mba_evaluate.argtypes = [C.POINTER(_Problem), C.c_int32, C.c_double, _f64p, _f64p, _f64p, _f64p, C.c_int32]
    lib.mba_evaluate_blocks.argtypes = [C.POINTER(_Problem), C.c_int32, C.c_double, _f64p, _f64p, _f64p, _f64p,
                                        _f64p, C.c_int32]
    lib.mba_evaluate_noalloc.argtypes = [C.c_int32, C.c_int32, C.c_int64] + [C.c_void_p] * 5 + [
        C.c_int32, C.c_double] + [C.c_void_p] * 4 + [C.c_int32, C.c_void_p]
    lib.mba_scene_generate.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _i64p, _f64p, _f64p, _f64p,
                                       _f64p, _f64p, _i32p, _i32p]
    lib.mba_scene_generate_grouped.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _i64p, _f64p, _f64p, _f64p, _f64p,
                                               _i32p, _i32p, _i32p]
    lib.mba_rotation_matrix_to_angle_axis.argtypes = [_f64p, _f64p, C.c_int64]
    lib.mba_angle_axis_to_rotation_matrix.argtypes = [_f64p, _f64p, C.c_int64]
    lib.mba_read_libmv.argtypes = [C.c_char_p, _i32p, _i32p, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p, _i32p,
                                   _i32p, _i32p, _i32p]
    lib.mba_read_bal.argtypes = [C.c_char_p, _i32p, _i32p, _i64p, _f64p, _f64p, _f64p, _i32p, _i32p]
    lib.mba_write_bal.argtypes = [C.c_char_p, C.POINTER(_Problem)]
    _lib = lib
    return lib


def _check(status):
    if status != OK:
        lib = load_library()
        raise MbaError(status, lib.mba_last_error().decode(errors="replace"))


def _p(a, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)


def default_options(**kw) -> Options:
    """mba_default_options with fields overridden by keyword; `dump_directory`
    and `dump_iterations` go through Options.set_dump."""
    o = Options()
    _check(load_library().mba_default_options(C.byref(o)))
    if "dump_directory" in kw or "dump_iterations" in kw:
        o.set_dump(kw.pop("dump_directory", None), kw.pop("dump_iterations", ()))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


# ------------------------------------------------------------ problems --
class BAProblem:
    """Arrays of a bundle-adjustment problem (caller-owned).  Packed layout:
    cameras (Nc, 9).  Grouped layout (give `intrinsics`): cameras (Nc, 6)
    [aa, t], intrinsics (Ni, 6) [f, ppx, ppy, k1, k2, k3] indexed per
    observation by `intrinsic_index`."""

    def __init__(self, cameras, points, observations, camera_index, point_index, camera_constant=None,
                 intrinsics=None, intrinsic_index=None, intrinsic_constant=None, intrinsic_subset=None,
                 camera_params=0, ordering=0):
        self.cameras = np.ascontiguousarray(cameras, np.float64)
        self.points = np.ascontiguousarray(points, np.float64)
        self.observations = np.ascontiguousarray(observations, np.float64)
        self.camera_index = np.ascontiguousarray(camera_index, np.int32)
        self.point_index = np.ascontiguousarray(point_index, np.int32)
        self.camera_constant = (None if camera_constant is None
                                else np.ascontiguousarray(camera_constant, np.uint8))
        self.layout = LAYOUT_PACKED9 if intrinsics is None else LAYOUT_GROUPED
        self.intrinsics = None if intrinsics is None else np.ascontiguousarray(intrinsics, np.float64)
        self.intrinsic_index = (None if intrinsic_index is None
                                else np.ascontiguousarray(intrinsic_index, np.int32))
        self.intrinsic_constant = (None if intrinsic_constant is None
                                   else np.ascontiguousarray(intrinsic_constant, np.uint8))
        self.intrinsic_subset = (None if intrinsic_subset is None
                                 else np.ascontiguousarray(intrinsic_subset, np.uint8))
        self.camera_params = int(camera_params)
        self.ordering = int(ordering)
        width = 9 if self.layout == LAYOUT_PACKED9 else 6
        if self.cameras.ndim != 2 or self.cameras.shape[1] != width:
            raise ValueError(f"cameras must be (Nc, {width})")
        if self.layout == LAYOUT_GROUPED:
            if self.intrinsics.ndim != 2 or self.intrinsics.shape[1] != 6:
                raise ValueError("intrinsics must be (Ni, 6)")
            if self.intrinsic_index is None or self.intrinsic_index.shape != self.camera_index.shape:
                raise ValueError("intrinsic_index must give a group per observation")
        if self.points.ndim != 2 or self.points.shape[1] != 3:
            raise ValueError("points must be (Np, 3)")
        if self.observations.shape != (self.camera_index.shape[0], 2):
            raise ValueError("observations must be (Nobs, 2) matching the index arrays")

    @property
    def num_cameras(self):
        return self.cameras.shape[0]

    @property
    def num_points(self):
        return self.points.shape[0]

    @property
    def num_observations(self):
        return self.observations.shape[0]

    @property
    def num_intrinsics(self):
        return 0 if self.intrinsics is None else self.intrinsics.shape[0]

    def _struct(self) -> _Problem:
        s = _Problem()
        s.layout = self.layout
        s.num_cameras, s.num_points, s.num_observations = self.num_cameras, self.num_points, self.num_observations
        s.num_intrinsics = self.num_intrinsics
        s.cameras, s.points, s.observations = _p(self.cameras), _p(self.points), _p(self.observations)
        s.camera_index, s.point_index = _p(self.camera_index, _i32p), _p(self.point_index, _i32p)
        s.camera_constant = _p(self.camera_constant, _u8p)
        s.intrinsics = _p(self.intrinsics)
        s.intrinsic_index = _p(self.intrinsic_index, _i32p)
        s.intrinsic_constant = _p(self.intrinsic_constant, _u8p)
        s.intrinsic_subset = _p(self.intrinsic_subset, _u8p)
        s.camera_params = self.camera_params
        s.ordering = self.ordering
        return s


class Summary:
    """Solver::Summary subset (include/ceres/solver.h:753-831)."""

    def __init__(self, s: _Summary, iterations):
        for name, _ in _Summary._fields_:
            setattr(self, name, getattr(s, name))
        self.iterations = iterations

    def IsSolutionUsable(self) -> bool:
        return self.termination_type in (CONVERGENCE, NO_CONVERGENCE)

    def FullReport(self) -> str:
        return (f"MI355X bundle adjustment: blocks E={self.num_e_blocks} F={self.num_f_blocks}\n"
                f"  initial cost {self.initial_cost:.6e}  final cost {self.final_cost:.6e}\n"
                f"  steps: {self.num_successful_steps} successful, {self.num_unsuccessful_steps} unsuccessful\n"
                f"  linear solver iterations {self.num_linear_solver_iterations}\n"
                f"  time: total {self.total_time_s:.4f}s, jacobian {self.jacobian_evaluation_time_s:.4f}s, "
                f"residual {self.residual_evaluation_time_s:.4f}s, linear {self.linear_solver_time_s:.4f}s\n"
                f"  termination: {_TERMINATION_NAMES.get(self.termination_type, self.termination_type)}")


def solve(problem: BAProblem, options: Options | None = None, max_records: int = 1000) -> Summary:
    """mba_solve: parameters updated in place when the solution is usable."""
    lib = load_library()
    o = options if options is not None else default_options()
    s = problem._struct()
    summ = _Summary()
    its = (Iteration * max_records)()
    _check(lib.mba_solve(C.byref(o), C.byref(s), C.byref(summ), its, C.c_int32(max_records)))
    n = min(summ.num_iterations, max_records)
    return Summary(summ, [its[i] for i in range(n)])


def solve_batched(problems, options: Options | None = None) -> list:
    """mba_solve_batched: independent windows solved together; each usable
    window's parameters are updated in place.  Returns one Summary each."""
    lib = load_library()
    o = options if options is not None else default_options()
    n = len(problems)
    arr = (_Problem * max(n, 1))()
    for i, p in enumerate(problems):
        arr[i] = p._struct()
    sums = (_Summary * max(n, 1))()
    _check(lib.mba_solve_batched(C.byref(o), arr, C.c_int32(n), sums))
    return [Summary(sums[i], []) for i in range(n)]


def evaluate(problem: BAProblem, precision=PRECISION_FP64, huber_a=0.0, jacobians=True, device=0):
    """CostFunction::Evaluate for every observation (input order) on the GPU."""
    lib = load_library()
    n = problem.num_observations
    r = np.zeros((n, 2))
    jc = np.zeros((n, 2, 9)) if jacobians else None
    jp = np.zeros((n, 2, 3)) if jacobians else None
    cost = C.c_double()
    s = problem._struct()
    _check(lib.mba_evaluate(C.byref(s), C.c_int32(precision), C.c_double(huber_a), _p(r), _p(jc), _p(jp),
                            C.byref(cost), C.c_int32(device)))
    return r, jc, jp, cost.value


def evaluate_noalloc(num_cameras, num_points, num_observations, cameras, points, observations, camera_index,
                     point_index, residuals=0, jac_camera=0, jac_point=0, cost=0, precision=PRECISION_FP64,
                     huber_a=0.0, device=0, stream=0):
    """mba_evaluate_noalloc: arguments are device addresses (int, 0 = NULL)
    of caller-owned device buffers; asynchronous on `stream` (a hipStream_t
    address, 0 = the null stream).  Raises MbaError on a rejected call."""
    lib = load_library()
    _check(lib.mba_evaluate_noalloc(int(num_cameras), int(num_points), int(num_observations), cameras or None,
                                    points or None, observations or None, camera_index or None,
                                    point_index or None, int(precision), float(huber_a), residuals or None,
                                    jac_camera or None, jac_point or None, cost or None, int(device),
                                    stream or None))


def evaluate_blocks(problem: BAProblem, precision=PRECISION_FP64, huber_a=0.0, device=0):
    """Residuals and per-block Jacobians on the GPU: (r, J_intrinsics or None,
    J_camera, J_point, cost)."""
    lib = load_library()
    n = problem.num_observations
    r = np.zeros((n, 2))
    grouped = problem.layout == LAYOUT_GROUPED
    ji = np.zeros((n, 2, 6)) if grouped else None
    jc = np.zeros((n, 2, 6 if grouped else 9))
    jp = np.zeros((n, 2, 3))
    cost = C.c_double()
    s = problem._struct()
    _check(lib.mba_evaluate_blocks(C.byref(s), C.c_int32(precision), C.c_double(huber_a), _p(r), _p(ji), _p(jc),
                                   _p(jp), C.byref(cost), C.c_int32(device)))
    return r, ji, jc, jp, cost.value


def schur_structure(problem: BAProblem, with_pairs=True):
    """(num_e, column-block order, residual order, S block pairs)."""
    lib = load_library()
    s = problem._struct()
    nb = problem.num_intrinsics + problem.num_cameras + problem.num_points
    order = np.zeros(nb, np.int32)
    res = np.zeros(problem.num_observations, np.int64)
    nblk, ne, npairs = C.c_int32(), C.c_int32(), C.c_int64()
    _check(lib.mba_schur_structure(C.byref(s), _p(order, _i32p), C.byref(nblk), C.byref(ne), _p(res, _i64p),
                                   None, C.c_int64(0), C.byref(npairs) if with_pairs else None))
    pairs = None
    if with_pairs:
        pairs = np.zeros((npairs.value, 2), np.int32)
        _check(lib.mba_schur_structure(C.byref(s), None, None, None, None, _p(pairs, _i32p),
                                       C.c_int64(npairs.value), None))
    return ne.value, order[:nblk.value], res, pairs


class Context:
    """A preprocessed, HBM-resident problem (mba_context_*)."""

    def __init__(self, problem: BAProblem, options: Options | None = None):
        self._lib = load_library()
        self.problem = problem
        self.options = options if options is not None else default_options()
        self._s = problem._struct()
        h = C.c_void_p()
        _check(self._lib.mba_context_create(C.byref(self._s), C.byref(self.options), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.mba_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return self._lib.mba_context_stream(self.handle) or 0

    def jacobian_schur(self, radius: float):
        _check(self._lib.mba_context_jacobian_schur(self.handle, C.c_double(radius)))

    def lm_pass(self, radius: float) -> float:
        """The pass as the LM loop runs it (FP64 cost first; a mixed context's
        FP32 kernels take those residuals); returns the FP64 cost."""
        cost = C.c_double()
        _check(self._lib.mba_context_lm_pass(self.handle, C.c_double(radius), C.byref(cost)))
        return cost.value

    def synchronize(self):
        _check(self._lib.mba_context_synchronize(self.handle))

    def set_timing(self, on: bool):
        _check(self._lib.mba_context_set_timing(self.handle, 1 if on else 0))

    def last_timing(self):
        """[point pass, camera pass, pass total, S assembly, Cholesky,
        triangular solves] in ms (the last three from the last explicit
        solve, 0 if none ran)."""
        ms = np.zeros(6)
        _check(self._lib.mba_context_last_timing(self.handle, _p(ms), 6))
        return ms

    def read_schur(self):
        nc, np_ = self.problem.num_cameras, self.problem.num_points
        out = dict(U=np.zeros((nc, 45)), gc=np.zeros((nc, 9)), rhs=np.zeros((nc, 9)),
                   V=np.zeros((np_, 6)), gp=np.zeros((np_, 3)))
        _check(self._lib.mba_context_read_schur(self.handle, _p(out["U"]), _p(out["gc"]), _p(out["rhs"]),
                                                _p(out["V"]), _p(out["gp"])))
        return out

    def read_scale(self):
        cs = np.zeros((self.problem.num_cameras, 9))
        ps = np.zeros((self.problem.num_points, 3))
        _check(self._lib.mba_context_read_scale(self.handle, _p(cs), _p(ps)))
        return cs, ps

    def schur_matvec(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        _check(self._lib.mba_context_schur_matvec(self.handle, _p(x), _p(y)))
        return y

    def solve(self, max_records=1000) -> Summary:
        summ = _Summary()
        its = (Iteration * max_records)()
        _check(self._lib.mba_context_solve(self.handle, C.byref(summ), its, C.c_int32(max_records)))
        n = min(summ.num_iterations, max_records)
        return Summary(summ, [its[i] for i in range(n)])

    def download(self):
        _check(self._lib.mba_context_download(self.handle, C.byref(self._s)))


def generate_scene(preset: int, window: int = 0) -> dict:
    """Seeded benchmark scene (preset 1..5) generated by the library."""
    lib = load_library()
    nc, np_, no = C.c_int32(), C.c_int32(), C.c_int64()
    _check(lib.mba_scene_generate(preset, window, C.byref(nc), C.byref(np_), C.byref(no),
                                  None, None, None, None, None, None, None))
    d = dict(cam_true=np.zeros((nc.value, 9)), cam_init=np.zeros((nc.value, 9)),
             pt_true=np.zeros((np_.value, 3)), pt_init=np.zeros((np_.value, 3)),
             obs=np.zeros((no.value, 2)), cam_idx=np.zeros(no.value, np.int32),
             pt_idx=np.zeros(no.value, np.int32))
    _check(lib.mba_scene_generate(preset, window, C.byref(nc), C.byref(np_), C.byref(no),
                                  _p(d["cam_true"]), _p(d["cam_init"]), _p(d["pt_true"]), _p(d["pt_init"]),
                                  _p(d["obs"]), _p(d["cam_idx"], _i32p), _p(d["pt_idx"], _i32p)))
    return d


def generate_grouped_scene(preset: int = 6) -> dict:
    """Seeded grouped-intrinsics scene (incremental-engine scale)."""
    lib = load_library()
    nc, np_, ni, no = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64()
    _check(lib.mba_scene_generate_grouped(preset, C.byref(nc), C.byref(np_), C.byref(ni), C.byref(no),
                                          None, None, None, None, None, None, None))
    d = dict(cam_init=np.zeros((nc.value, 6)), intr_init=np.zeros((ni.value, 6)), pt_init=np.zeros((np_.value, 3)),
             obs=np.zeros((no.value, 2)), cam_idx=np.zeros(no.value, np.int32),
             intr_idx=np.zeros(no.value, np.int32), pt_idx=np.zeros(no.value, np.int32))
    _check(lib.mba_scene_generate_grouped(preset, C.byref(nc), C.byref(np_), C.byref(ni), C.byref(no),
                                          _p(d["cam_init"]), _p(d["intr_init"]), _p(d["pt_init"]), _p(d["obs"]),
                                          _p(d["cam_idx"], _i32p), _p(d["intr_idx"], _i32p),
                                          _p(d["pt_idx"], _i32p)))
    return d


# ------------------------------------------------ conversions and files --
def rotation_matrix_to_angle_axis(R):
    """RotationMatrixToAngleAxis of (n, 3, 3) matrices (row-major numpy
    arrays; the C ABI takes column-major)."""
    R = np.asarray(R, np.float64).reshape(-1, 3, 3)
    Rc = np.ascontiguousarray(np.transpose(R, (0, 2, 1)))
    aa = np.zeros((R.shape[0], 3))
    _check(load_library().mba_rotation_matrix_to_angle_axis(_p(Rc), _p(aa), R.shape[0]))
    return aa


def angle_axis_to_rotation_matrix(aa):
    aa = np.ascontiguousarray(np.asarray(aa, np.float64).reshape(-1, 3))
    Rc = np.zeros((aa.shape[0], 3, 3))
    _check(load_library().mba_angle_axis_to_rotation_matrix(_p(aa), _p(Rc), aa.shape[0]))
    return np.transpose(Rc, (0, 2, 1)).copy()


def read_libmv(path) -> dict:
    """A libmv problem file as grouped-layout arrays (one intrinsic group)."""
    lib = load_library()
    b = os.fsencode(path)
    nc, np_, no, sp = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int32()
    _check(lib.mba_read_libmv(b, C.byref(nc), C.byref(np_), C.byref(no), C.byref(sp), None, None, None, None,
                              None, None, None, None))
    d = dict(intrinsics=np.zeros((1, 6)), cameras=np.zeros((nc.value, 6)), points=np.zeros((np_.value, 3)),
             obs=np.zeros((no.value, 2)), cam_idx=np.zeros(no.value, np.int32), pt_idx=np.zeros(no.value, np.int32),
             image_of_camera=np.zeros(nc.value, np.int32), track_of_point=np.zeros(np_.value, np.int32))
    _check(lib.mba_read_libmv(b, C.byref(nc), C.byref(np_), C.byref(no), C.byref(sp), _p(d["intrinsics"]),
                              _p(d["cameras"]), _p(d["points"]), _p(d["obs"]), _p(d["cam_idx"], _i32p),
                              _p(d["pt_idx"], _i32p), _p(d["image_of_camera"], _i32p),
                              _p(d["track_of_point"], _i32p)))
    d["image_space"] = bool(sp.value)
    d["intr_idx"] = np.zeros(no.value, np.int32)
    return d


def read_bal(path) -> dict:
    """A BAL text problem as packed-layout arrays (observations negated: see
    mba_read_bal)."""
    lib = load_library()
    b = os.fsencode(path)
    nc, np_, no = C.c_int32(), C.c_int32(), C.c_int64()
    _check(lib.mba_read_bal(b, C.byref(nc), C.byref(np_), C.byref(no), None, None, None, None, None))
    d = dict(cameras=np.zeros((nc.value, 9)), points=np.zeros((np_.value, 3)), obs=np.zeros((no.value, 2)),
             cam_idx=np.zeros(no.value, np.int32), pt_idx=np.zeros(no.value, np.int32))
    _check(lib.mba_read_bal(b, C.byref(nc), C.byref(np_), C.byref(no), _p(d["cameras"]), _p(d["points"]),
                            _p(d["obs"]), _p(d["cam_idx"], _i32p), _p(d["pt_idx"], _i32p)))
    return d


def write_bal(path, problem: BAProblem):
    s = problem._struct()
    _check(load_library().mba_write_bal(os.fsencode(path), C.byref(s)))


def read_trust_region_problem(base) -> dict:
    """A TEXTFILE trust-region problem dump (<base>_A.txt, _D.txt, _b.txt,
    _x.txt and <base>.m, the layout of Ceres's
    DumpLinearLeastSquaresProblemToTextFile): num_rows, num_cols, A as
    (rows, cols, values) in file order, D, b, x (whichever exist)."""
    out = {}
    with open(str(base) + ".m") as f:
        for line in f:
            for key in ("num_rows", "num_cols"):
                if line.startswith(f"lsqp.{key} = "):
                    out[key] = int(line.split("=")[1].strip().rstrip(";"))
    a = np.loadtxt(str(base) + "_A.txt", ndmin=2)
    out["rows"] = a[:, 0].astype(np.int64)
    out["cols"] = a[:, 1].astype(np.int64)
    out["values"] = a[:, 2].copy()
    for key in ("D", "b", "x"):
        path = f"{base}_{key}.txt"
        if os.path.exists(path):
            out[key] = np.loadtxt(path, ndmin=1)
    return out


# ------------------------------------------------ Ceres-shaped host API --
class HuberLoss:
    """ceres::HuberLoss(a): threshold a pixels (include/ceres/loss_function.h:175-184)."""

    def __init__(self, a: float):
        self.a = float(a)


class ReprojectionError:
    """The libs/sfm reprojection functor for one observation
    (pinhole_brown_rt_ceres_functor.h:42-110), packed 9-parameter camera:
    AddResidualBlock(cost, loss, camera[9], point[3])."""

    def __init__(self, observed_xy):
        self.observed = np.asarray(observed_xy, np.float64).reshape(2)


class ErrorFunc_Refine_Camera_3DPoints(ReprojectionError):  # noqa: N801 (reference name)
    """The incremental engine's functor (pinhole_brown_rt_ceres_functor.h:42-110):
    AddResidualBlock(cost, loss, intrinsics[6], camera[6], point[3]) with
    intrinsics [f, ppx, ppy, k1, k2, k3] and camera [aa, t]."""


class SubsetParameterization:
    """ceres::SubsetParameterization(size, constant_parameters)
    (include/ceres/local_parameterization.h:173; checks of
    internal/ceres/local_parameterization.cc:88-113)."""

    def __init__(self, size, constant_parameters):
        held = sorted(int(k) for k in constant_parameters)
        if size != 6:
            raise ValueError("SubsetParameterization applies to 6-entry intrinsic blocks")
        if not held:
            raise ValueError("the set of constant parameters should contain at least one element")
        if len(set(held)) != len(held):
            raise ValueError("the set of constant parameters cannot contain duplicates")
        if len(held) >= size or held[0] < 0 or held[-1] >= size:
            raise ValueError("hold between 1 and size - 1 entries in [0, size)")
        self.mask = sum(1 << k for k in held)


class SolverOptions:
    """Solver::Options fields libs/sfm sets (sfm_incremental_engine.cpp:1449-1468)."""

    def __init__(self):
        o = default_options()
        self.linear_solver_type = SPARSE_SCHUR
        self.precision = PRECISION_FP64
        self.max_num_iterations = o.max_num_iterations
        self.function_tolerance = o.function_tolerance
        self.gradient_tolerance = o.gradient_tolerance
        self.parameter_tolerance = o.parameter_tolerance
        self.initial_trust_region_radius = o.initial_trust_region_radius
        self.max_linear_solver_iterations = o.max_linear_solver_iterations
        self.eta = o.eta
        self.pcg_r_tolerance = o.pcg_r_tolerance
        self.device = 0
        self.minimizer_progress_to_stdout = False
        self.preconditioner_type = JACOBI
        self.trust_region_minimizer_iterations_to_dump = []
        self.trust_region_problem_dump_directory = "/tmp"

    def _to_c(self, huber_a) -> Options:
        return default_options(
            linear_solver=self.linear_solver_type, precision=self.precision,
            max_num_iterations=self.max_num_iterations, function_tolerance=self.function_tolerance,
            gradient_tolerance=self.gradient_tolerance, parameter_tolerance=self.parameter_tolerance,
            initial_trust_region_radius=self.initial_trust_region_radius,
            max_linear_solver_iterations=self.max_linear_solver_iterations, eta=self.eta,
            pcg_r_tolerance=self.pcg_r_tolerance, device=self.device, huber_a=huber_a,
            preconditioner=self.preconditioner_type,
            dump_directory=self.trust_region_problem_dump_directory,
            dump_iterations=self.trust_region_minimizer_iterations_to_dump)


class Problem:
    """ceres::Problem for bundle adjustment, in either layout (numpy float64
    arrays as parameter blocks, updated in place by Solve):
      AddResidualBlock(ReprojectionError, loss, camera[9], point[3]), or
      AddResidualBlock(ErrorFunc_Refine_Camera_3DPoints, loss, intrinsics[6],
                       camera[6], point[3]) (the incremental engine's call)."""

    def __init__(self):
        self._cams, self._pts, self._intr = [], [], []
        self._cam_id, self._pt_id, self._intr_id = {}, {}, {}
        self._obs, self._ci, self._pi, self._ii = [], [], [], []
        self._loss = None
        self._unrobust = False  # some block was added with loss_function None
        self._constant, self._intr_constant = set(), set()
        self._subset = {}
        self._layout = None

    def _block(self, arr, size, table, ids):
        if not isinstance(arr, np.ndarray) or arr.dtype != np.float64 or arr.size != size:
            raise ValueError(f"parameter block must be a float64 array of {size}")
        key = arr.__array_interface__["data"][0]
        if key not in ids:
            ids[key] = len(table)
            table.append(arr)
        return ids[key]

    def AddResidualBlock(self, cost_function: ReprojectionError, loss_function, *blocks):
        grouped = len(blocks) == 3
        if len(blocks) not in (2, 3) or grouped != isinstance(cost_function, ErrorFunc_Refine_Camera_3DPoints):
            raise ValueError("(camera[9], point[3]) with ReprojectionError, or (intrinsics[6], camera[6], "
                             "point[3]) with ErrorFunc_Refine_Camera_3DPoints")
        layout = LAYOUT_GROUPED if grouped else LAYOUT_PACKED9
        if self._layout is not None and layout != self._layout:
            raise ValueError("all residual blocks of a problem use one camera layout")
        self._layout = layout
        # The engine applies one loss to every observation (libs/sfm shares one
        # HuberLoss instance across its blocks); a problem mixing robustified
        # and plain blocks, or two thresholds, is refused rather than solved
        # with a loss some blocks did not ask for.
        if loss_function is None:
            if self._loss is not None:
                raise ValueError("residual blocks with and without a loss function cannot be mixed")
            self._unrobust = True
        else:
            if self._unrobust:
                raise ValueError("residual blocks with and without a loss function cannot be mixed")
            if self._loss is not None and loss_function is not self._loss and loss_function.a != self._loss.a:
                raise ValueError("all residual blocks must share one loss function")
            self._loss = loss_function
        if grouped:
            self._ii.append(self._block(blocks[0], 6, self._intr, self._intr_id))
            blocks = blocks[1:]
        c = self._block(blocks[0], 6 if grouped else 9, self._cams, self._cam_id)
        p = self._block(blocks[1], 3, self._pts, self._pt_id)
        self._obs.append(cost_function.observed)
        self._ci.append(c)
        self._pi.append(p)

    def _key(self, block):
        return block.__array_interface__["data"][0]

    def SetParameterBlockConstant(self, block):
        key = self._key(block)
        if key in self._cam_id:
            self._constant.add(self._cam_id[key])
        elif key in self._intr_id:
            self._intr_constant.add(self._intr_id[key])
        else:
            raise ValueError("unknown camera or intrinsics block")

    def SetParameterization(self, block, parameterization: SubsetParameterization):
        key = self._key(block)
        if key not in self._intr_id:
            raise ValueError("parameterizations apply to intrinsics blocks")
        self._subset[self._intr_id[key]] = parameterization.mask

    def NumResidualBlocks(self):
        return len(self._obs)

    def _arrays(self) -> BAProblem:
        w = 6 if self._layout == LAYOUT_GROUPED else 9
        cams = np.array([c.reshape(w) for c in self._cams], np.float64).reshape(-1, w)
        pts = np.array([p.reshape(3) for p in self._pts], np.float64).reshape(-1, 3)
        const = None
        if self._constant:
            const = np.zeros(len(self._cams), np.uint8)
            const[list(self._constant)] = 1
        kw = {}
        if self._layout == LAYOUT_GROUPED:
            ni = len(self._intr)
            ic = np.zeros(ni, np.uint8)
            ic[list(self._intr_constant)] = 1
            sub = np.zeros(ni, np.uint8)
            for g, m in self._subset.items():
                sub[g] = m
            kw = dict(intrinsics=np.array([b.reshape(6) for b in self._intr], np.float64).reshape(-1, 6),
                      intrinsic_index=np.array(self._ii, np.int32), intrinsic_constant=ic, intrinsic_subset=sub)
        return BAProblem(cams, pts, np.array(self._obs, np.float64).reshape(-1, 2),
                         np.array(self._ci, np.int32), np.array(self._pi, np.int32), const, **kw)


def Solve(options: SolverOptions, problem: Problem, summary_out: list | None = None) -> Summary:
    """ceres::Solve: writes the parameter blocks back when usable."""
    arrays = problem._arrays()
    huber = problem._loss.a if problem._loss is not None else 0.0
    summary = solve(arrays, options._to_c(huber))
    if summary.IsSolutionUsable():
        w = arrays.cameras.shape[1]
        for i, c in enumerate(problem._cams):
            c.reshape(w)[:] = arrays.cameras[i]
        for i, p in enumerate(problem._pts):
            p.reshape(3)[:] = arrays.points[i]
        for i, b in enumerate(problem._intr):
            b.reshape(6)[:] = arrays.intrinsics[i]
    if summary_out is not None:
        summary_out.append(summary)
    return summary
