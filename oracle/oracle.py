"""ctypes bridge to the FP64 CPU oracle (``liborc.so``) and, where it has been
built in this container, to the reference's own Ceres (``_ref/libceres_ref.so``).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` -- as the checker, never as the
product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_SO = os.path.join(HERE, "liborc.so")
REF_SO = os.path.join(HERE, "_ref", "libceres_ref.so")

LAYOUT_PACKED9 = 0
LAYOUT_INTR6_EXT6 = 1
SPARSE_SCHUR = 0
ITERATIVE_SCHUR = 1
CONVERGENCE, NO_CONVERGENCE, FAILURE = 0, 1, 2

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class SceneSpec(C.Structure):
    _fields_ = [("num_cameras", C.c_int32), ("num_points", C.c_int32),
                ("track_kind", C.c_int32), ("track_min", C.c_int32),
                ("track_max", C.c_int32), ("pad_", C.c_int32),
                ("alpha", C.c_double), ("noise_px", C.c_double),
                ("outlier_frac", C.c_double), ("outlier_px", C.c_double),
                ("perturb", C.c_double), ("seed", C.c_uint64)]


class Problem(C.Structure):
    _fields_ = [("layout", C.c_int32), ("num_cameras", C.c_int32),
                ("num_intrinsics", C.c_int32), ("num_points", C.c_int32),
                ("num_obs", C.c_int64), ("cameras", _f64p), ("intrinsics", _f64p),
                ("points", _f64p), ("obs", _f64p), ("cam_idx", _i32p),
                ("intr_idx", _i32p), ("pt_idx", _i32p), ("cam_constant", _u8p)]


class Options(C.Structure):
    _fields_ = [("max_num_iterations", C.c_int32), ("linear_solver", C.c_int32),
                ("max_num_consecutive_invalid_steps", C.c_int32),
                ("max_linear_solver_iterations", C.c_int32),
                ("min_linear_solver_iterations", C.c_int32),
                ("jacobi_scaling", C.c_int32), ("num_threads", C.c_int32),
                ("pad_", C.c_int32), ("huber_a", C.c_double),
                ("function_tolerance", C.c_double), ("gradient_tolerance", C.c_double),
                ("parameter_tolerance", C.c_double),
                ("initial_trust_region_radius", C.c_double),
                ("max_trust_region_radius", C.c_double),
                ("min_trust_region_radius", C.c_double),
                ("min_relative_decrease", C.c_double), ("min_lm_diagonal", C.c_double),
                ("max_lm_diagonal", C.c_double), ("eta", C.c_double),
                ("r_tolerance", C.c_double)]


class Iteration(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("step_is_valid", C.c_int32),
                ("step_is_successful", C.c_int32), ("linear_solver_iterations", C.c_int32),
                ("cost", C.c_double), ("cost_change", C.c_double),
                ("gradient_max_norm", C.c_double), ("gradient_norm", C.c_double),
                ("step_norm", C.c_double), ("relative_decrease", C.c_double),
                ("trust_region_radius", C.c_double)]


class Summary(C.Structure):
    _fields_ = [("initial_cost", C.c_double), ("final_cost", C.c_double),
                ("fixed_cost", C.c_double), ("total_time_s", C.c_double),
                ("num_successful_steps", C.c_int32), ("num_unsuccessful_steps", C.c_int32),
                ("num_iterations", C.c_int32), ("termination_type", C.c_int32),
                ("num_e_blocks", C.c_int32), ("num_f_blocks", C.c_int32)]


_orc = None
_ref = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_SO):
            build()
        _orc = C.CDLL(ORC_SO)
        _orc.orc_scene_preset.restype = SceneSpec
        _orc.orc_scene_count.restype = C.c_int64
        _orc.orc_scene_fill.restype = C.c_int64
        _orc.orc_scene_fill_groups.restype = C.c_int64
        _orc.orc_schur_ordering.restype = C.c_int32
        _orc.orc_schur_block_pairs.restype = C.c_int64
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        _ref.ref_schur_ordering.restype = C.c_int
        _ref.ref_solve.restype = C.c_int
        _ref.ref_evaluate_obs.restype = C.c_int
    return _ref


def ptr(a, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)


# ------------------------------------------------------------------ scenes --
def preset(config: int, window: int = 0) -> SceneSpec:
    return lib().orc_scene_preset(C.c_int(config), C.c_int(window))


def make_spec(num_cameras, num_points, track_min=2, track_max=6, power_alpha=None,
              noise_px=0.5, outlier_frac=0.0, outlier_px=20.0, perturb=0.01, seed=1) -> SceneSpec:
    s = SceneSpec()
    s.num_cameras, s.num_points = num_cameras, num_points
    s.track_kind = 0 if power_alpha is None else 1
    s.track_min, s.track_max = track_min, track_max
    s.alpha = power_alpha or 0.0
    s.noise_px, s.outlier_frac, s.outlier_px, s.perturb = noise_px, outlier_frac, outlier_px, perturb
    s.seed = seed
    return s


def generate(spec: SceneSpec) -> dict:
    """Scene arrays (observations in AddResidualBlock order)."""
    L = lib()
    tl = np.zeros(spec.num_points, np.int32)
    nobs = L.orc_scene_count(C.byref(spec), ptr(tl, _i32p))
    d = dict(
        cam_true=np.zeros((spec.num_cameras, 9)), cam_init=np.zeros((spec.num_cameras, 9)),
        pt_true=np.zeros((spec.num_points, 3)), pt_init=np.zeros((spec.num_points, 3)),
        obs=np.zeros((nobs, 2)), cam_idx=np.zeros(nobs, np.int32),
        pt_idx=np.zeros(nobs, np.int32), offsets=np.zeros(spec.num_points + 1, np.int64))
    L.orc_scene_fill(C.byref(spec), ptr(tl, _i32p), ptr(d["cam_true"]), ptr(d["cam_init"]),
                     ptr(d["pt_true"]), ptr(d["pt_init"]), ptr(d["obs"]),
                     ptr(d["cam_idx"], _i32p), ptr(d["pt_idx"], _i32p),
                     ptr(d["offsets"], _i64p))
    d["num_cameras"], d["num_points"], d["num_obs"] = spec.num_cameras, spec.num_points, int(nobs)
    return d


def generate_groups(spec: SceneSpec, groups: int) -> dict:
    """Grouped-intrinsics scene (the incremental engine's layout): cameras
    [aa, t], intrinsic groups [f, ppx, ppy, k1, k2, k3], camera c in group
    c mod groups."""
    L = lib()
    tl = np.zeros(spec.num_points, np.int32)
    nobs = L.orc_scene_count(C.byref(spec), ptr(tl, _i32p))
    d = dict(
        cam_true=np.zeros((spec.num_cameras, 6)), cam_init=np.zeros((spec.num_cameras, 6)),
        intr_true=np.zeros((groups, 6)), intr_init=np.zeros((groups, 6)),
        pt_true=np.zeros((spec.num_points, 3)), pt_init=np.zeros((spec.num_points, 3)),
        obs=np.zeros((nobs, 2)), cam_idx=np.zeros(nobs, np.int32), intr_idx=np.zeros(nobs, np.int32),
        pt_idx=np.zeros(nobs, np.int32), offsets=np.zeros(spec.num_points + 1, np.int64))
    L.orc_scene_fill_groups(C.byref(spec), C.c_int32(groups), ptr(tl, _i32p), ptr(d["cam_true"]),
                            ptr(d["cam_init"]), ptr(d["intr_true"]), ptr(d["intr_init"]),
                            ptr(d["pt_true"]), ptr(d["pt_init"]), ptr(d["obs"]), ptr(d["cam_idx"], _i32p),
                            ptr(d["intr_idx"], _i32p), ptr(d["pt_idx"], _i32p), ptr(d["offsets"], _i64p))
    d["num_cameras"], d["num_points"], d["num_obs"] = spec.num_cameras, spec.num_points, int(nobs)
    d["num_intrinsics"] = groups
    return d


def group_scene(d, cam_constant=None, intr=None, intr_constant=None, intr_subset=None) -> "Scene":
    """Oracle Scene (layout 1) of a grouped-intrinsics scene.  intr_constant /
    intr_subset (per-group constant flags, SubsetParameterization masks) are
    honoured by the reference harness only."""
    sc = Scene(d, cam_constant=cam_constant, layout=1,
               intrinsics=d["intr_init"] if intr is None else intr, intr_idx=d["intr_idx"])
    sc.intr_constant = None if intr_constant is None else np.ascontiguousarray(intr_constant, np.uint8)
    sc.intr_subset = None if intr_subset is None else np.ascontiguousarray(intr_subset, np.uint8)
    return sc


class Scene:
    """Owns the arrays a Problem struct points into."""

    def __init__(self, d, cameras=None, points=None, cam_constant=None, layout=0,
                 intrinsics=None, intr_idx=None):
        self.layout = layout
        self.cameras = np.ascontiguousarray(d["cam_init"] if cameras is None else cameras, np.float64).copy()
        self.points = np.ascontiguousarray(d["pt_init"] if points is None else points, np.float64).copy()
        self.obs = np.ascontiguousarray(d["obs"], np.float64)
        self.cam_idx = np.ascontiguousarray(d["cam_idx"], np.int32)
        self.pt_idx = np.ascontiguousarray(d["pt_idx"], np.int32)
        self.cam_constant = None if cam_constant is None else np.ascontiguousarray(cam_constant, np.uint8)
        self.intrinsics = None if intrinsics is None else np.ascontiguousarray(intrinsics, np.float64).copy()
        self.intr_idx = None if intr_idx is None else np.ascontiguousarray(intr_idx, np.int32)
        self.intr_constant = None
        self.intr_subset = None
        self.num_cameras = self.cameras.shape[0]
        self.num_points = self.points.shape[0]
        self.num_obs = self.obs.shape[0]

    def problem(self) -> Problem:
        p = Problem()
        p.layout = self.layout
        p.num_cameras, p.num_points, p.num_obs = self.num_cameras, self.num_points, self.num_obs
        p.num_intrinsics = 0 if self.intrinsics is None else self.intrinsics.shape[0]
        p.cameras, p.points, p.obs = ptr(self.cameras), ptr(self.points), ptr(self.obs)
        p.intrinsics = ptr(self.intrinsics)
        p.cam_idx, p.pt_idx = ptr(self.cam_idx, _i32p), ptr(self.pt_idx, _i32p)
        p.intr_idx = ptr(self.intr_idx, _i32p)
        p.cam_constant = ptr(self.cam_constant, _u8p)
        return p

    def num_blocks(self):
        ni = 0 if self.intrinsics is None else self.intrinsics.shape[0]
        return ni + self.num_cameras + self.num_points


# ------------------------------------------------------------------ oracle --
def default_options(**kw) -> Options:
    o = Options()
    lib().orc_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def solve(scene: Scene, opts: Options, max_records=1000):
    its = (Iteration * max_records)()
    s = Summary()
    p = scene.problem()
    lib().orc_solve(C.byref(p), C.byref(opts), C.byref(s), its, C.c_int32(max_records))
    n = min(s.num_iterations, max_records)
    return s, [its[i] for i in range(n)]


def evaluate(scene: Scene, huber_a=0.0, jacobians=True):
    n = scene.num_obs
    r = np.zeros((n, 2))
    cs = 9 if scene.layout == 0 else 6  # layout 1: the extrinsic block
    jc = np.zeros((n, 2, cs)) if jacobians else None
    jp = np.zeros((n, 2, 3)) if jacobians else None
    cost = C.c_double()
    p = scene.problem()
    lib().orc_evaluate(C.byref(p), C.c_double(huber_a), ptr(r), ptr(jc), ptr(jp), C.byref(cost))
    return r, jc, jp, cost.value


def evaluate_obs(layout, cam, pt, obs, intr=None):
    r = np.zeros(2)
    if layout == 0:
        j0, j1 = np.zeros((2, 9)), np.zeros((2, 3))
        lib().orc_evaluate_obs(0, None, ptr(cam), ptr(pt), ptr(obs), ptr(r), ptr(j0), ptr(j1), None)
        return r, j0, j1
    j0, j1, j2 = np.zeros((2, 6)), np.zeros((2, 6)), np.zeros((2, 3))
    lib().orc_evaluate_obs(1, ptr(intr), ptr(cam), ptr(pt), ptr(obs), ptr(r), ptr(j0), ptr(j1), ptr(j2))
    return r, j0, j1, j2


def schur_ordering(scene: Scene):
    nb = scene.num_blocks()
    order = np.zeros(nb, np.int32)
    res = np.zeros(scene.num_obs, np.int64)
    nblk, nres = C.c_int32(), C.c_int64()
    p = scene.problem()
    ne = lib().orc_schur_ordering(C.byref(p), ptr(order, _i32p), C.byref(nblk), ptr(res, _i64p), C.byref(nres))
    return int(ne), order[:nblk.value].copy(), res[:nres.value].copy()


def schur_block_pairs(scene: Scene):
    p = scene.problem()
    n = lib().orc_schur_block_pairs(C.byref(p), None, C.c_int64(0))
    pairs = np.zeros((n, 2), np.int32)
    lib().orc_schur_block_pairs(C.byref(p), ptr(pairs, _i32p), C.c_int64(n))
    return pairs


def jacobian_schur(scene: Scene, huber_a=0.0, scale=None, D=None, num_threads=1):
    nc, np_ = scene.num_cameras, scene.num_points
    out = dict(U=np.zeros((nc, 45)), gc=np.zeros((nc, 9)), V=np.zeros((np_, 6)),
               gp=np.zeros((np_, 3)), rhs=np.zeros((nc, 9)))
    cost = C.c_double()
    p = scene.problem()
    ok = lib().orc_jacobian_schur(C.byref(p), C.c_double(huber_a), ptr(scale), ptr(D),
                                  ptr(out["U"]), ptr(out["gc"]), ptr(out["V"]), ptr(out["gp"]),
                                  ptr(out["rhs"]), C.byref(cost), C.c_int32(num_threads))
    out["cost"] = cost.value
    out["ok"] = bool(ok)
    return out


# ------------------------------------------------------- reference (Ceres) --
def _ref_args(scene: Scene):
    ni = 0 if scene.intrinsics is None else scene.intrinsics.shape[0]
    return [C.c_int(scene.layout), C.c_int(scene.num_cameras), C.c_int(ni),
            C.c_int(scene.num_points), C.c_long(scene.num_obs), ptr(scene.cameras),
            ptr(scene.intrinsics), ptr(scene.points), ptr(scene.obs),
            ptr(scene.cam_idx, _i32p), ptr(scene.intr_idx, _i32p), ptr(scene.pt_idx, _i32p),
            ptr(scene.cam_constant, _u8p)]


def ref_schur_ordering(scene: Scene, with_pairs=True):
    nb = scene.num_blocks()
    order = np.zeros(nb, np.int32)
    res = np.zeros(scene.num_obs, np.int64)
    nblk, nres, npairs = C.c_int(), C.c_long(), C.c_long()
    args = _ref_args(scene)
    ne = ref_lib().ref_schur_ordering(*args, ptr(order, _i32p), C.byref(nblk), ptr(res, _i64p),
                                      C.byref(nres), None, C.c_long(0),
                                      C.byref(npairs) if with_pairs else None)
    pairs = None
    if with_pairs:
        pairs = np.zeros((npairs.value, 2), np.int32)
        ref_lib().ref_schur_ordering(*args, ptr(order, _i32p), C.byref(nblk), ptr(res, _i64p),
                                     C.byref(nres), ptr(pairs, _i32p), C.c_long(npairs.value), None)
    return int(ne), order[:nblk.value].copy(), res[:nres.value].copy(), pairs


def ref_solve(scene: Scene, huber_a=0.0, linear_solver=0, max_iters=50, num_threads=1, max_rec=1000):
    summ = np.zeros(8)
    its = np.zeros((max_rec, 9))
    usable = ref_lib().ref_solve(*_ref_args(scene), ptr(scene.intr_constant, _u8p), ptr(scene.intr_subset, _u8p),
                                 C.c_double(huber_a), C.c_int(linear_solver),
                                 C.c_int(max_iters), C.c_int(num_threads), ptr(summ), ptr(its),
                                 C.c_int(max_rec))
    n = int(summ[5])
    return dict(initial_cost=summ[0], final_cost=summ[1], successful=int(summ[2]),
                unsuccessful=int(summ[3]), termination=int(summ[4]), total_time=summ[6],
                linear_solver_time=summ[7], usable=bool(usable), iterations=its[:min(n, max_rec)].copy())


def ref_jacobian_schur_pass(scene: Scene, huber_a=0.0, num_threads=1, reps=1):
    """One timed Jacobian + Schur pass of the reference's SPARSE_SCHUR path
    (ProgramEvaluator::Evaluate, ScaleColumns, SchurEliminator::Eliminate) on
    num_threads threads: (seconds per pass, set-up seconds, residual blocks)."""
    sec = np.zeros(2)
    f = ref_lib().ref_jacobian_schur_pass
    f.restype = C.c_long
    n = f(*_ref_args(scene), C.c_double(huber_a), C.c_int(num_threads), C.c_int(reps), ptr(sec))
    if n < 0:
        raise RuntimeError(f"ref_jacobian_schur_pass failed ({n})")
    return float(sec[0]), float(sec[1]), int(n)


def ref_evaluate_obs(layout, cam, pt, obs, intr=None):
    r = np.zeros(2)
    if layout == 0:
        j0, j1 = np.zeros((2, 9)), np.zeros((2, 3))
        ref_lib().ref_evaluate_obs(0, None, ptr(cam), ptr(pt), ptr(obs), ptr(r), ptr(j0), ptr(j1), None)
        return r, j0, j1
    j0, j1, j2 = np.zeros((2, 6)), np.zeros((2, 6)), np.zeros((2, 3))
    ref_lib().ref_evaluate_obs(1, ptr(intr), ptr(cam), ptr(pt), ptr(obs), ptr(r), ptr(j0), ptr(j1), ptr(j2))
    return r, j0, j1, j2


# ------------------------------------------ reference formats and rotations --
def ref_set_points_first(on: bool):
    """Solver::Options::linear_solver_ordering with the points in group 0."""
    ref_lib().ref_set_points_first(C.c_int(1 if on else 0))


def ref_set_schur_jacobi(on: bool):
    """Solver::Options::preconditioner_type = SCHUR_JACOBI in ref_solve."""
    ref_lib().ref_set_schur_jacobi(C.c_int(1 if on else 0))


def ref_set_preconditioner(kind: int):
    """Solver::Options::preconditioner_type of ref_solve: 0 JACOBI, 1 SCHUR_JACOBI,
    2 CLUSTER_JACOBI, 3 CLUSTER_TRIDIAGONAL."""
    ref_lib().ref_set_preconditioner(C.c_int(kind))


def ref_set_lm_tolerances(eta=0.0, function_tolerance=0.0):
    """Solver::Options::eta and ::function_tolerance of ref_solve; 0 keeps
    Ceres's defaults (1e-1 and 1e-6)."""
    ref_lib().ref_set_lm_tolerances(C.c_double(eta), C.c_double(function_tolerance))


def ref_set_dump(directory, iterations):
    """trust_region_minimizer_iterations_to_dump / trust_region_problem_dump_directory
    (TEXTFILE) of ref_solve; an empty list turns dumps off."""
    its = (C.c_int * max(1, len(iterations)))(*iterations)
    ref_lib().ref_set_dump(None if directory is None else os.fsencode(directory), its, C.c_int(len(iterations)))


def ref_last_message() -> str:
    """Solver::Summary::message of the last ref_solve."""
    n = ref_lib().ref_last_message(None, C.c_int(0))
    buf = C.create_string_buffer(n + 1)
    ref_lib().ref_last_message(buf, C.c_int(n + 1))
    return buf.value.decode()


def ref_rotation_matrix_to_angle_axis(R):
    """ceres::RotationMatrixToAngleAxis of (n, 3, 3) row-major matrices."""
    R = np.asarray(R, np.float64).reshape(-1, 3, 3)
    Rc = np.ascontiguousarray(np.transpose(R, (0, 2, 1)))
    aa = np.zeros((R.shape[0], 3))
    ref_lib().ref_rotation_matrix_to_angle_axis(ptr(Rc), ptr(aa), C.c_long(R.shape[0]))
    return aa


def ref_bal_roundtrip(path_in, path_out):
    ref_lib().ref_bal_roundtrip(os.fsencode(path_in), os.fsencode(path_out))


def ref_bal_arrays(path):
    sizes = np.zeros(3, np.int32)
    ref_lib().ref_bal_arrays(os.fsencode(path), ptr(sizes, _i32p), None, None, None, None)
    nc, np_, no = (int(v) for v in sizes)
    d = dict(params=np.zeros(9 * nc + 3 * np_), obs=np.zeros((no, 2)), cam_idx=np.zeros(no, np.int32),
             pt_idx=np.zeros(no, np.int32))
    ref_lib().ref_bal_arrays(os.fsencode(path), ptr(sizes, _i32p), ptr(d["params"]), ptr(d["obs"]),
                             ptr(d["cam_idx"], _i32p), ptr(d["pt_idx"], _i32p))
    d["cameras"] = d["params"][:9 * nc].reshape(nc, 9)
    d["points"] = d["params"][9 * nc:].reshape(np_, 3)
    return d


def ref_bal_solve(path, huber_a=0.0, linear_solver=0, max_iters=20, max_rec=1000):
    summ = np.zeros(8)
    its = np.zeros((max_rec, 9))
    d = ref_bal_arrays(path)
    params = np.zeros_like(d["params"])
    ref_lib().ref_bal_solve(os.fsencode(path), C.c_double(huber_a), C.c_int(linear_solver), C.c_int(max_iters),
                            ptr(summ), ptr(its), C.c_int(max_rec), ptr(params))
    n = int(summ[5])
    nc = d["cameras"].shape[0]
    return dict(initial_cost=summ[0], final_cost=summ[1], successful=int(summ[2]), unsuccessful=int(summ[3]),
                termination=int(summ[4]), iterations=its[:min(n, max_rec)].copy(),
                cameras=params[:9 * nc].reshape(nc, 9), points=params[9 * nc:].reshape(-1, 3))


def parse_libmv(path):
    """The libmv problem format read independently of the library
    (libmv_bundle_adjuster.cc:255-437): returns the file's raw contents."""
    import struct
    b = open(path, "rb").read()
    e = {ord("v"): "<", ord("V"): ">"}[b[0]]
    out = dict(image_space=b[1] == ord("P"))
    o = 2
    out["intr8"] = np.array(struct.unpack_from(e + "8f", b, o), np.float64)
    o += 32
    nc, = struct.unpack_from(e + "i", b, o)
    o += 4
    cams = np.frombuffer(b, np.dtype([("image", e + "i4"), ("R", e + "f4", 9), ("t", e + "f4", 3)]), nc, o)
    o += cams.nbytes
    npt, = struct.unpack_from(e + "i", b, o)
    o += 4
    pts = np.frombuffer(b, np.dtype([("track", e + "i4"), ("X", e + "f4", 3)]), npt, o)
    o += pts.nbytes
    nm, = struct.unpack_from(e + "i", b, o)
    o += 4
    mk = np.frombuffer(b, np.dtype([("image", e + "i4"), ("track", e + "i4"), ("x", e + "f4"), ("y", e + "f4")]),
                       nm, o)
    out.update(cams=cams, pts=pts, markers=mk)
    return out
