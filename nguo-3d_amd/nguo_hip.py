"""ctypes binding of libnguo_hip.so (the C ABI in include/nguo_hip.h).

This is the Python-side stand-in for the reference's Rust FFI crate: the same entry points a
Rust `extern "C"` block would bind, used by the tests, the benchmark
and the smoke check. It loads the in-tree library and fails loudly when it is missing; there
is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnguo_hip.so")

NGUO_OK = 0
NGUO_E_INVALID = -1
NGUO_E_DEVICE = -2
NGUO_E_NOMEM = -3
NGUO_E_CAPACITY = -4
NGUO_E_UNSUPPORTED = -5

EXPORTED = [
    "nguo_last_error", "nguo_ctx_create", "nguo_ctx_destroy", "nguo_ctx_info",
    "nguo_ls_create", "nguo_ls_destroy", "nguo_ls_solve", "nguo_ls_last_stats", "nguo_ls_spmv",
    "nguo_step_create", "nguo_step_destroy", "nguo_step_advance", "nguo_step_get_positions",
    "nguo_step_get_velocities", "nguo_step_set_state", "nguo_step_get_metrics",
    "nguo_scene_generate", "nguo_scene_view", "nguo_scene_free",
    "nguo_step_get_newton_trace", "nguo_step_stencils", "nguo_step_collider_sides",
    "nguo_lbvh_morton", "nguo_lbvh_sort_pairs", "nguo_lbvh_karras", "nguo_lbvh_aabb",
    "nguo_step_bench_spmv", "nguo_step_bench_precond", "nguo_bench_spmv_grid", "nguo_ls_precond_apply",
    "nguo_batch_create", "nguo_batch_destroy", "nguo_batch_advance", "nguo_batch_instance",
    "nguo_step_set_rest_scale", "nguo_step_adjoint", "nguo_batch_set_solve", "nguo_bench_sort",
    "nguo_step_set_trace", "nguo_lbvh_sort_pairs_noalloc", "nguo_step_energy", "nguo_step_set_precision",
    "nguo_step_set_kinematic_collider_positions",
]


class Collider(C.Structure):
    _fields_ = [("kind", C.c_int32), ("a", C.c_float * 3), ("b", C.c_float * 3), ("radius", C.c_float)]


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("gravity", C.c_float * 3), ("newton_iters", C.c_uint32),
        ("bend_stiffness", C.c_float), ("shell_mu", C.c_float), ("shell_lambda", C.c_float),
        ("shell_model", C.c_int32), ("eiganalysis_eps", C.c_float), ("contact_dhat", C.c_float),
        ("contact_kappa", C.c_float), ("friction", C.c_float), ("friction_eps", C.c_float),
        ("barrier_kind", C.c_int32), ("self_collision", C.c_int32), ("max_dx", C.c_float),
        ("target_toi", C.c_float), ("min_newton_steps", C.c_uint32), ("ccd_max_iter", C.c_uint32),
        ("cg_tol", C.c_float), ("cg_max_iter", C.c_uint32), ("floor_y", C.c_float),
        ("precond", C.c_int32), ("schwarz_domain", C.c_uint32), ("schwarz_weight", C.c_float),
        ("stretch_k", C.c_float),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_verts", C.c_uint32), ("positions", C.POINTER(C.c_float)), ("velocities", C.POINTER(C.c_float)),
        ("inv_mass", C.POINTER(C.c_float)), ("n_tris", C.c_uint32), ("tris", C.POINTER(C.c_uint32)),
        ("static_collider", C.POINTER(C.c_uint8)), ("n_shell_faces", C.c_uint32),
        ("shell_faces", C.POINTER(C.c_uint32)), ("n_rods", C.c_uint32), ("rod_verts", C.POINTER(C.c_uint32)),
        ("rod_rest", C.POINTER(C.c_float)), ("rod_k", C.POINTER(C.c_float)), ("n_rod_hinges", C.c_uint32),
        ("rod_hinge_verts", C.POINTER(C.c_uint32)), ("rod_hinge_rest", C.POINTER(C.c_float)),
        ("rod_hinge_stiff", C.POINTER(C.c_float)), ("n_colliders", C.c_uint32),
        ("colliders", C.POINTER(Collider)), ("params", Params),
    ]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("sweeps", C.c_uint32), ("residual", C.c_double)]


class StepMetrics(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("max_reresid", C.c_double), ("max_cg_iters", C.c_uint32),
        ("newton_iters", C.c_uint32), ("num_stencils", C.c_uint32), ("max_velocity", C.c_float),
        ("penetration_free", C.c_int32), ("all_finite", C.c_int32), ("last_toi", C.c_float),
        ("num_contacts", C.c_uint32), ("max_sigma", C.c_float), ("device_alloc", C.c_uint32), ("device_free", C.c_uint32),
        ("sum_cg_iters", C.c_uint32),
    ]


class NguoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nguo error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Loads the in-tree libnguo_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    lib.nguo_last_error.restype = C.c_char_p
    lib.nguo_ctx_create.argtypes = [C.c_int, P(vp)]
    lib.nguo_ctx_destroy.argtypes = [vp]
    lib.nguo_ctx_destroy.restype = None
    lib.nguo_ctx_info.argtypes = [vp, C.c_char_p, C.c_size_t, P(C.c_int)]
    lib.nguo_ls_create.argtypes = [vp, C.c_int, C.c_uint32, C.c_float, P(vp)]
    lib.nguo_ls_destroy.argtypes = [vp]
    lib.nguo_ls_destroy.restype = None
    lib.nguo_ls_solve.argtypes = [vp, C.c_uint32, P(C.c_uint32), P(C.c_uint32), P(C.c_float), P(C.c_float),
                                  C.c_float, C.c_uint32, P(C.c_float), P(Stats)]
    lib.nguo_ls_last_stats.argtypes = [vp, P(Stats)]
    lib.nguo_ls_spmv.argtypes = [vp, C.c_uint32, P(C.c_uint32), P(C.c_uint32), P(C.c_float), P(C.c_float), P(C.c_float)]
    lib.nguo_ls_precond_apply.argtypes = [vp, C.c_uint32, P(C.c_uint32), P(C.c_uint32), P(C.c_float), P(C.c_float), P(C.c_float)]
    lib.nguo_scene_generate.argtypes = [C.c_int, C.c_uint64, C.c_double, P(vp)]
    lib.nguo_scene_view.argtypes = [vp, P(SceneDesc)]
    lib.nguo_scene_free.argtypes = [vp]
    lib.nguo_scene_free.restype = None
    if hasattr(lib, "nguo_step_create"):
        lib.nguo_step_create.argtypes = [vp, P(SceneDesc), P(vp)]
        lib.nguo_step_destroy.argtypes = [vp]
        lib.nguo_step_destroy.restype = None
        lib.nguo_step_advance.argtypes = [vp, C.c_uint32]
        lib.nguo_step_get_positions.argtypes = [vp, P(C.c_float)]
        lib.nguo_step_get_velocities.argtypes = [vp, P(C.c_float)]
        lib.nguo_step_set_state.argtypes = [vp, P(C.c_float), P(C.c_float)]
        lib.nguo_step_get_metrics.argtypes = [vp, P(StepMetrics)]
        lib.nguo_step_set_kinematic_collider_positions.argtypes = [vp, P(C.c_float)]
        lib.nguo_step_get_newton_trace.argtypes = [vp, C.c_uint32, P(C.c_float), P(C.c_uint32), P(C.c_float)]
        lib.nguo_step_set_trace.argtypes = [vp, C.c_int]
        lib.nguo_step_set_precision.argtypes = [vp, C.c_int]
        lib.nguo_step_energy.argtypes = [vp, P(C.c_float), P(C.c_double), P(C.c_double)]
        lib.nguo_step_stencils.argtypes = [vp, P(C.c_float), C.c_uint32, P(C.c_uint32), P(C.c_float), P(C.c_float),
                                           P(C.c_float), P(C.c_uint32)]
        lib.nguo_step_collider_sides.argtypes = [vp, P(C.c_float), P(C.c_float)]
        lib.nguo_lbvh_morton.argtypes = [vp, C.c_uint32, P(C.c_float), C.c_uint32, P(C.c_uint32), P(C.c_uint32)]
        lib.nguo_lbvh_sort_pairs.argtypes = [vp, C.c_uint32, P(C.c_uint32), P(C.c_uint32)]
        lib.nguo_lbvh_sort_pairs_noalloc.argtypes = [vp, C.c_uint32, vp, vp]
        lib.nguo_lbvh_karras.argtypes = [vp, C.c_uint32, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        lib.nguo_lbvh_aabb.argtypes = [vp, C.c_uint32, P(C.c_uint32), P(C.c_uint32), P(C.c_float), P(C.c_float)]
        lib.nguo_batch_create.argtypes = [vp, C.c_uint32, P(SceneDesc), P(vp)]
        lib.nguo_batch_destroy.argtypes = [vp]
        lib.nguo_batch_destroy.restype = None
        lib.nguo_batch_advance.argtypes = [vp, C.c_uint32]
        lib.nguo_batch_instance.argtypes = [vp, C.c_uint32, P(vp)]
        lib.nguo_step_bench_spmv.argtypes = [vp, C.c_uint32, C.c_int, P(C.c_double), P(C.c_double)]
        lib.nguo_step_bench_precond.argtypes = [vp, C.c_uint32, P(C.c_double), P(C.c_double)]
        lib.nguo_step_set_rest_scale.argtypes = [vp, P(C.c_float), C.c_uint32]
        lib.nguo_batch_set_solve.argtypes = [vp, C.c_int]
        lib.nguo_bench_sort.argtypes = [vp, C.c_uint32, C.c_uint32, P(C.c_double), P(C.c_double)]
        lib.nguo_step_adjoint.argtypes = [vp, P(C.c_float), C.c_float, C.c_int, P(C.c_float), P(Stats)]
        lib.nguo_bench_spmv_grid.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_double), P(C.c_double), P(C.c_double)]
    _lib = lib
    return lib


def check(code):
    if code != NGUO_OK:
        raise NguoError(code, load().nguo_last_error().decode())


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Scene:
    """A generated workload: owns the native buffer and exposes numpy views of it."""

    def __init__(self, config, seed=0, param=1.0):
        lib = load()
        self._buf = C.c_void_p()
        check(lib.nguo_scene_generate(int(config), int(seed), float(param), C.byref(self._buf)))
        self.desc = SceneDesc()
        check(lib.nguo_scene_view(self._buf, C.byref(self.desc)))
        d = self.desc
        n = d.n_verts
        self.n_verts = n
        # copies: a numpy view must not outlive the native buffer it points into
        self.positions = np.ctypeslib.as_array(d.positions, shape=(n, 3)).copy()
        self.inv_mass = np.ctypeslib.as_array(d.inv_mass, shape=(n,)).copy()
        self.tris = np.ctypeslib.as_array(d.tris, shape=(d.n_tris, 3)).copy()
        self.static = np.ctypeslib.as_array(d.static_collider, shape=(n,)).copy()
        self.params = d.params

    def __del__(self):
        if getattr(self, "_buf", None) and _lib is not None:
            _lib.nguo_scene_free(self._buf)
            self._buf = None


class ArrayScene:
    """A scene built from numpy arrays (the caller's own lowering), with the attributes of
    Scene: positions [n, 3], inv_mass [n], tris [m, 3], static [n] (or None), colliders (a
    list of Collider) and params (a Params). The arrays are kept alive with the descriptor."""

    def __init__(self, positions, inv_mass, tris, params, static=None, colliders=(), velocities=None, shell_faces=None,
                 rods=None, rod_hinges=None):
        """rods: (verts [r, 2], rest [r], k [r]); rod_hinges: (verts [h, 3] as (j, i, k), rest [h], stiff [h])."""
        self.positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self.n_verts = len(self.positions)
        self.inv_mass = np.ascontiguousarray(inv_mass, dtype=np.float32).reshape(-1)
        self.tris = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
        self.static = None if static is None else np.ascontiguousarray(static, dtype=np.uint8).reshape(-1)
        self._vel = np.zeros_like(self.positions) if velocities is None else np.ascontiguousarray(velocities, dtype=np.float32).reshape(-1, 3)
        self._coll = (Collider * max(1, len(colliders)))(*colliders)
        u32p, f32p = C.POINTER(C.c_uint32), C.POINTER(C.c_float)
        d = SceneDesc()
        d.n_verts = self.n_verts
        d.positions = self.positions.ctypes.data_as(f32p)
        d.velocities = self._vel.ctypes.data_as(f32p)
        d.inv_mass = self.inv_mass.ctypes.data_as(f32p)
        d.n_tris = len(self.tris)
        d.tris = self.tris.ctypes.data_as(u32p)
        if self.static is not None:
            d.static_collider = self.static.ctypes.data_as(C.POINTER(C.c_uint8))
        if shell_faces is not None and len(shell_faces):
            self._shell = np.ascontiguousarray(shell_faces, dtype=np.uint32).reshape(-1, 3)
            d.n_shell_faces = len(self._shell)
            d.shell_faces = self._shell.ctypes.data_as(u32p)
        if rods is not None and len(rods[0]):
            self._rods = [np.ascontiguousarray(rods[0], dtype=np.uint32).reshape(-1, 2),
                          np.ascontiguousarray(rods[1], dtype=np.float32).reshape(-1), np.ascontiguousarray(rods[2], dtype=np.float32).reshape(-1)]
            d.n_rods = len(self._rods[0])
            d.rod_verts, d.rod_rest, d.rod_k = (self._rods[0].ctypes.data_as(u32p), self._rods[1].ctypes.data_as(f32p),
                                                self._rods[2].ctypes.data_as(f32p))
        if rod_hinges is not None and len(rod_hinges[0]):
            self._rh = [np.ascontiguousarray(rod_hinges[0], dtype=np.uint32).reshape(-1, 3),
                        np.ascontiguousarray(rod_hinges[1], dtype=np.float32).reshape(-1), np.ascontiguousarray(rod_hinges[2], dtype=np.float32).reshape(-1)]
            d.n_rod_hinges = len(self._rh[0])
            d.rod_hinge_verts, d.rod_hinge_rest, d.rod_hinge_stiff = (self._rh[0].ctypes.data_as(u32p), self._rh[1].ctypes.data_as(f32p),
                                                                      self._rh[2].ctypes.data_as(f32p))
        d.n_colliders = len(colliders)
        d.colliders = C.cast(self._coll, C.POINTER(Collider))
        d.params = params
        self.desc = d
        self.params = d.params


class Context:
    def __init__(self, device=0):
        lib = load()
        self.h = C.c_void_p()
        check(lib.nguo_ctx_create(int(device), C.byref(self.h)))

    def info(self):
        name = C.create_string_buffer(128)
        cu = C.c_int()
        check(load().nguo_ctx_info(self.h, name, 128, C.byref(cu)))
        return name.value.decode(), cu.value

    def close(self):
        if self.h:
            load().nguo_ctx_destroy(self.h)
            self.h = None


class LinearSolver:
    """The LinearSolver seam (engine-core/src/linsolve.rs:21-28) on the device."""

    def __init__(self, ctx, precond=0, domain=32, weight=0.0):
        self.h = C.c_void_p()
        check(load().nguo_ls_create(ctx.h, int(precond), int(domain), float(weight), C.byref(self.h)))

    def precond_apply(self, row_ptr, col_idx, values, r):
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        col_idx = np.ascontiguousarray(col_idx, dtype=np.uint32)
        values = np.ascontiguousarray(values, dtype=np.float32)
        r = np.ascontiguousarray(r, dtype=np.float32)
        z = np.zeros_like(r)
        check(load().nguo_ls_precond_apply(self.h, len(r), uptr(row_ptr), uptr(col_idx), fptr(values), fptr(r), fptr(z)))
        return z

    def solve(self, row_ptr, col_idx, values, b, tol, max_iter):
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        col_idx = np.ascontiguousarray(col_idx, dtype=np.uint32)
        values = np.ascontiguousarray(values, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        x = np.zeros_like(b)
        st = Stats()
        check(load().nguo_ls_solve(self.h, len(b), uptr(row_ptr), uptr(col_idx), fptr(values), fptr(b),
                                   float(tol), int(max_iter), fptr(x), C.byref(st)))
        return x, st

    def spmv(self, row_ptr, col_idx, values, x):
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        col_idx = np.ascontiguousarray(col_idx, dtype=np.uint32)
        values = np.ascontiguousarray(values, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros_like(x)
        check(load().nguo_ls_spmv(self.h, len(x), uptr(row_ptr), uptr(col_idx), fptr(values), fptr(x), fptr(y)))
        return y

    def close(self):
        if self.h:
            load().nguo_ls_destroy(self.h)
            self.h = None


def lbvh_morton(ctx, pos, tris):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    codes = np.zeros(len(tris), np.uint32)
    check(load().nguo_lbvh_morton(ctx.h, len(pos), fptr(pos), len(tris), uptr(tris), uptr(codes)))
    return codes


def lbvh_sort_pairs(ctx, keys, vals):
    keys = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    vals = np.ascontiguousarray(vals, dtype=np.uint32).copy()
    check(load().nguo_lbvh_sort_pairs(ctx.h, len(keys), uptr(keys), uptr(vals)))
    return keys, vals


NOALLOC_SORT_MAX = 1 << 20  # NGUO_NOALLOC_SORT_MAX


def lbvh_sort_pairs_noalloc(ctx, n, keys_dev, vals_dev):
    """Sorts n (u32 key, u32 value) pairs in place in device memory (addresses as ints),
    allocating nothing (nguo_lbvh_sort_pairs_noalloc)."""
    check(load().nguo_lbvh_sort_pairs_noalloc(ctx.h, int(n), C.c_void_p(keys_dev), C.c_void_p(vals_dev)))


def lbvh_karras(ctx, codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint32)
    n = len(codes)
    left = np.zeros(max(n - 1, 1), np.uint32)
    right = np.zeros(max(n - 1, 1), np.uint32)
    check(load().nguo_lbvh_karras(ctx.h, n, uptr(codes), uptr(left), uptr(right)))
    return left[: max(n - 1, 0)], right[: max(n - 1, 0)]


def lbvh_aabb(ctx, left, right, leaf_box):
    left = np.ascontiguousarray(left, dtype=np.uint32)
    right = np.ascontiguousarray(right, dtype=np.uint32)
    leaf_box = np.ascontiguousarray(leaf_box, dtype=np.float32)
    n = len(leaf_box)
    node = np.zeros((max(n - 1, 1), 6), np.float32)
    check(load().nguo_lbvh_aabb(ctx.h, n, uptr(left), uptr(right), fptr(leaf_box), fptr(node)))
    return node[: max(n - 1, 0)]


def bench_spmv_grid(ctx, nx, ny, reps=20):
    ms, by, nnzb = C.c_double(), C.c_double(), C.c_double()
    check(load().nguo_bench_spmv_grid(ctx.h, nx, ny, reps, C.byref(ms), C.byref(by), C.byref(nnzb)))
    return ms.value, by.value, nnzb.value


def bench_sort(ctx, n, reps=10):
    """(average ms, algorithmic bytes) of one 4-pass radix sort of n random pairs."""
    ms, by = C.c_double(), C.c_double()
    check(load().nguo_bench_sort(ctx.h, int(n), int(reps), C.byref(ms), C.byref(by)))
    return ms.value, by.value


def check_invariants(steps, cg_tol, cg_max_iter, strain_limit=0.0):
    """The solver invariants over a run's per-step metrics, as engine-core/src/metrics.rs:84-183
    check_invariants evaluates them: returns (all_ok, [(name, ok, detail), ...]) in the
    reference's order. `steps` holds StepMetrics (or objects with the same fields)."""
    checks = []
    bad = sum(1 for m in steps if not m.penetration_free)
    checks.append(("penetration_free", bad == 0, f"{bad} step(s) with penetration/CCD failure"))
    nonfin = sum(1 for m in steps if not m.all_finite)
    checks.append(("all_frames_finite", nonfin == 0, f"{nonfin} non-finite step(s)"))
    if strain_limit <= 0.0:
        checks.append(("strain_limit_respected", True, "strain limiting disabled (auto-pass)"))
    else:
        peak = max([0.0] + [float(m.max_sigma) for m in steps])
        ceiling = (1.0 + strain_limit) * 1.02
        checks.append(("strain_limit_respected", peak <= ceiling, f"peak sigma {peak:.4f} vs ceiling {ceiling:.4f}"))
    worst = max([0.0] + [float(m.max_reresid) for m in steps])
    tol = float(cg_tol) * 1.5
    checks.append(("cg_converged", worst <= tol, f"worst reresid {worst:.3e} vs {tol:.3e}"))
    max_it = max([0] + [int(m.max_cg_iters) for m in steps])
    checks.append(("cg_iter_under_cap", max_it < cg_max_iter, f"max cg iters {max_it} vs cap {cg_max_iter}"))
    half = steps[len(steps) // 2:]
    ta = sum(int(m.device_alloc) for m in half)
    tf = sum(int(m.device_free) for m in half)
    checks.append(("zero_steady_state_alloc", ta == 0 and tf == 0, f"steady-state device alloc {ta} free {tf}"))
    return all(c[1] for c in checks), checks


def drape_panels(scene, n_around=8, n_height=8, axes=None):
    """Panel index [0, n_around * n_height) of every scene shell face, from the face centroids
    of the rest shape: n_around sectors of the angle about the vertical (y) axis through the
    cloth centroid times n_height bands of the height. axes = (i, j) bins coordinates i and j
    instead (flat sheets). These panels carry the drape adjoint's rest-scale parameters."""
    nf = scene.desc.n_shell_faces
    faces = np.ctypeslib.as_array(scene.desc.shell_faces, (3 * nf,)).reshape(nf, 3)
    pos = np.ctypeslib.as_array(scene.desc.positions, (3 * scene.n_verts,)).reshape(-1, 3).astype(np.float64)
    cen = pos[faces].mean(axis=1)

    def bins(c, n):
        lo, hi = c.min(), c.max()
        return np.clip(((c - lo) / max(hi - lo, 1e-12) * n).astype(np.int64), 0, n - 1)

    if axes is not None:
        return bins(cen[:, axes[0]], n_around) * n_height + bins(cen[:, axes[1]], n_height)
    mid = cen.mean(axis=0)
    ang = np.arctan2(cen[:, 2] - mid[2], cen[:, 0] - mid[0])
    sector = np.clip(((ang + np.pi) / (2 * np.pi) * n_around).astype(np.int64), 0, n_around - 1)
    return sector * n_height + bins(cen[:, 1], n_height)


class Stepper:
    """The whole implicit step (GpuImplicitSolver / ImplicitSolver::step) on the device."""

    def __init__(self, ctx, scene, trace=False):
        self.h = C.c_void_p()
        self.n = scene.n_verts
        self.nfaces = scene.desc.n_shell_faces
        check(load().nguo_step_create(ctx.h, C.byref(scene.desc), C.byref(self.h)))
        if trace:
            self.set_trace(True)

    PCG_FP32_REFINED, PCG_FP64 = 0, 1

    def set_precision(self, precision):
        """Newton solves in FP32 PCG + FP64 refinement (0) or FP64 PCG (1)."""
        check(load().nguo_step_set_precision(self.h, int(precision)))

    def set_trace(self, on):
        """Keep the positions after every Newton iteration of the following steps."""
        check(load().nguo_step_set_trace(self.h, 1 if on else 0))

    @classmethod
    def from_handle(cls, handle, scene):
        """A view of a step object owned elsewhere (a batch instance)."""
        st = cls.__new__(cls)
        st.h = handle
        st.n = scene.n_verts
        st.nfaces = scene.desc.n_shell_faces
        st._borrowed = True
        return st

    def advance(self, nsteps=1):
        check(load().nguo_step_advance(self.h, int(nsteps)))

    def positions(self):
        out = np.zeros((self.n, 3), dtype=np.float32)
        check(load().nguo_step_get_positions(self.h, fptr(out)))
        return out

    def velocities(self):
        out = np.zeros((self.n, 3), dtype=np.float32)
        check(load().nguo_step_get_velocities(self.h, fptr(out)))
        return out

    def set_state(self, x, v):
        x = np.ascontiguousarray(x, dtype=np.float32)
        v = np.ascontiguousarray(v, dtype=np.float32)
        check(load().nguo_step_set_state(self.h, fptr(x), fptr(v)))

    def metrics(self):
        m = StepMetrics()
        check(load().nguo_step_get_metrics(self.h, C.byref(m)))
        return m

    def newton_trace(self, positions=True):
        """Positions after every Newton iteration of the last step (traced steps only), and
        their CCD fractions; positions=False returns ([], fractions) for any step."""
        cnt = C.c_uint32()
        check(load().nguo_step_get_newton_trace(self.h, 0, None, C.byref(cnt), None))
        xs, tois = [], []
        for k in range(cnt.value):
            t = C.c_float()
            if positions:
                out = np.zeros((self.n, 3), dtype=np.float32)
                check(load().nguo_step_get_newton_trace(self.h, k, fptr(out), None, C.byref(t)))
                xs.append(out)
            else:
                check(load().nguo_step_get_newton_trace(self.h, k, None, None, C.byref(t)))
            tois.append(t.value)
        return xs, tois

    def stencils(self, x):
        """Active contact stencils at x: (idx [n,4], d [n], g [n,12], kappa [n])."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        cnt = C.c_uint32()
        check(load().nguo_step_stencils(self.h, fptr(x), 0, None, None, None, None, C.byref(cnt)))
        n = cnt.value
        idx = np.zeros((max(n, 1), 4), np.uint32)
        d = np.zeros(max(n, 1), np.float32)
        g = np.zeros((max(n, 1), 12), np.float32)
        k = np.zeros(max(n, 1), np.float32)
        check(load().nguo_step_stencils(self.h, fptr(x), n, uptr(idx), fptr(d), fptr(g), fptr(k), C.byref(cnt)))
        return idx[:n], d[:n], g[:n], k[:n]

    def bench_spmv(self, reps=200, iteration_form=True):
        """(average ms per launch, algorithmic bytes per launch) of the PCG operator; the
        iteration form is the kernel instance the PCG iteration launches."""
        ms, by = C.c_double(), C.c_double()
        check(load().nguo_step_bench_spmv(self.h, int(reps), int(bool(iteration_form)), C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def bench_precond(self, reps=200):
        """(average ms per launch, algorithmic bytes per launch) of the preconditioner's
        launch in the PCG iteration (Schwarz apply with the PCG update)."""
        ms, by = C.c_double(), C.c_double()
        check(load().nguo_step_bench_precond(self.h, int(reps), C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def set_rest_scale(self, face_scale):
        """Rest-shape scale per scene shell face (the drape adjoint's parameters)."""
        sc = np.ascontiguousarray(face_scale, dtype=np.float32)
        check(load().nguo_step_set_rest_scale(self.h, fptr(sc), len(sc)))

    ADJ_NEWTON, ADJ_EXACT_ELASTIC, ADJ_EXACT = 0, 1, 2

    def adjoint(self, dLdx, tol=1e-6, exact_elastic=False, hessian=None):
        """dL/ds per shell face through the last step's equilibrium, given dL/dx [n, 3], and
        the adjoint solve's stats. exact_elastic: unprojected shell / rod-segment blocks;
        hessian (overrides it): ADJ_NEWTON, ADJ_EXACT_ELASTIC or ADJ_EXACT (the exact Jacobian
        of the step's gradient with the contact parameters held at x*)."""
        g = np.ascontiguousarray(dLdx, dtype=np.float32).reshape(-1)
        out = np.zeros(self.nfaces, np.float32)
        st = Stats()
        mode = hessian if hessian is not None else (self.ADJ_EXACT_ELASTIC if exact_elastic else self.ADJ_NEWTON)
        check(load().nguo_step_adjoint(self.h, fptr(g), float(tol), int(mode), fptr(out), C.byref(st)))
        return out, st

    ENERGY_TERMS = ("inertia", "shells", "hinges", "rods", "rod_bends", "barriers", "friction", "colliders", "springs")

    def energy(self, x):
        """(total, terms[9]) of the step energy of the last advanced step at positions x."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        terms = np.zeros(len(self.ENERGY_TERMS), np.float64)
        tot = C.c_double()
        check(load().nguo_step_energy(self.h, fptr(x), terms.ctypes.data_as(C.POINTER(C.c_double)), C.byref(tot)))
        return tot.value, terms

    def set_kinematic(self, x):
        """Move the static mesh collider to x's collider rows before the next step
        (ImplicitSolver::set_kinematic_collider_positions)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        check(load().nguo_step_set_kinematic_collider_positions(self.h, fptr(x)))

    def collider_sides(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        side = np.zeros(self.n, np.float32)
        check(load().nguo_step_collider_sides(self.h, fptr(x), fptr(side)))
        return side

    def close(self):
        if self.h and not getattr(self, "_borrowed", False):
            load().nguo_step_destroy(self.h)
            self.h = None


class Batch:
    """Independent scenes advanced concurrently, one stream each (nguo_batch_*)."""

    def __init__(self, ctx, scenes):
        self.scenes = list(scenes)  # keeps the scene buffers alive
        arr = (SceneDesc * len(self.scenes))(*[sc.desc for sc in self.scenes])
        self.h = C.c_void_p()
        check(load().nguo_batch_create(ctx.h, len(self.scenes), arr, C.byref(self.h)))
        self.steppers = []
        for i, sc in enumerate(self.scenes):
            hs = C.c_void_p()
            check(load().nguo_batch_instance(self.h, i, C.byref(hs)))
            self.steppers.append(Stepper.from_handle(hs, sc))

    def advance(self, steps=1):
        check(load().nguo_batch_advance(self.h, int(steps)))

    def set_solve(self, batched):
        """True: the instances' Newton solves run as one batched PCG; False: each alone."""
        check(load().nguo_batch_set_solve(self.h, 1 if batched else 0))

    def close(self):
        if self.h:
            for st in self.steppers:
                st.h = None  # owned by the batch
            load().nguo_batch_destroy(self.h)
            self.h = None
