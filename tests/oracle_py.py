"""ctypes loader of oracle/liboracle.so, the CPU restatement of the reference (test-only).

Only the tests, the smoke check and the benchmark's CPU-baseline leg use this module: it is
the checker for the HIP path, never part of it.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None

D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
U = C.POINTER(C.c_uint32)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise FileNotFoundError(f"{LIB} missing: run `make -C oracle`")
        lib = C.CDLL(LIB)
        vp = C.c_void_p
        lib.oracle_solver_create.restype = vp
        lib.oracle_solver_create.argtypes = [vp, C.c_int]
        lib.oracle_solver_destroy.argtypes = [vp]
        lib.oracle_solver_nverts.restype = C.c_uint32
        lib.oracle_solver_nverts.argtypes = [vp]
        lib.oracle_solver_step.argtypes = [vp, C.c_uint32]
        lib.oracle_solver_positions.argtypes = [vp, F]
        lib.oracle_solver_positions_d.argtypes = [vp, D]
        lib.oracle_solver_velocities.argtypes = [vp, F]
        lib.oracle_solver_set_state.argtypes = [vp, F, F]
        lib.oracle_solver_set_rest_scale.argtypes = [vp, D, C.c_uint32]
        lib.oracle_solver_set_rest_scale.restype = None
        lib.oracle_solver_adjoint.argtypes = [vp, D, C.c_uint32, C.c_double, C.c_int, D, D, D]
        lib.oracle_solver_adjoint.restype = C.c_uint32
        lib.oracle_solver_metrics.argtypes = [vp, D]
        lib.oracle_solver_newton_trace.restype = C.c_int
        lib.oracle_solver_newton_trace.argtypes = [vp, C.c_uint32, F, D]
        lib.oracle_solver_stencils.restype = C.c_int64
        lib.oracle_solver_stencils.argtypes = [vp, F, C.c_uint32, U, F, F]
        lib.oracle_solver_set_kinematic.restype = C.c_int
        lib.oracle_solver_set_kinematic.argtypes = [vp, F]
        lib.oracle_solver_collider_sides.restype = C.c_int
        lib.oracle_solver_collider_sides.argtypes = [vp, F, F]
        lib.oracle_solver_assemble.restype = C.c_int64
        lib.oracle_solver_assemble.argtypes = [vp, F, F, C.c_double, F, C.c_uint32]
        lib.oracle_solver_csr.argtypes = [vp, U, U, F, F]
        lib.oracle_solver_gradient.argtypes = [vp, F, F, C.c_double, D]
        lib.oracle_solver_gradient_d.argtypes = [vp, D, D, C.c_double, D]
        lib.oracle_solver_gradient_d.restype = None
        lib.oracle_solver_ccd.restype = C.c_double
        lib.oracle_solver_ccd.argtypes = [vp, F, F]
        lib.oracle_solver_self_intersects.restype = C.c_int
        lib.oracle_solver_self_intersects.argtypes = [vp, F]
        lib.oracle_solver_newton_iteration.restype = C.c_double
        lib.oracle_solver_newton_iteration.argtypes = [vp, F, F, D, U]
        lib.oracle_solver_x_tilde.argtypes = [vp, F]
        lib.oracle_cg.argtypes = [C.c_uint32, U, U, F, F, C.c_float, C.c_uint32, F, U, D]
        lib.oracle_spmv.argtypes = [C.c_uint32, U, U, F, F, F]
        lib.oracle_cg_schwarz.argtypes = [C.c_uint32, U, U, F, F, C.c_float, C.c_uint32, C.c_uint32, C.c_double, F, U, D]
        lib.oracle_schwarz.restype = C.c_uint32
        lib.oracle_schwarz.argtypes = [C.c_uint32, U, U, F, C.c_uint32, C.c_double, U, U, D, D]
        lib.oracle_point_tri_distance_grad.restype = C.c_double
        lib.oracle_point_tri_distance_grad.argtypes = [C.c_int, D, D]
        lib.oracle_segment_distance_grad.restype = C.c_double
        lib.oracle_segment_distance_grad.argtypes = [C.c_int, D, D]
        lib.oracle_segment_distance.restype = C.c_double
        lib.oracle_segment_distance.argtypes = [C.c_int, D]
        lib.oracle_point_tri_distance.restype = C.c_double
        lib.oracle_point_tri_distance.argtypes = [C.c_int, D]
        lib.oracle_edge_triangle_intersect.restype = C.c_int
        lib.oracle_edge_triangle_intersect.argtypes = [C.c_int, D]
        lib.oracle_accd_toi.restype = C.c_double
        lib.oracle_accd_toi.argtypes = [C.c_int, C.c_int, D, D, C.c_double, C.c_double, C.c_double, C.c_uint32]
        lib.oracle_barrier.restype = C.c_int
        lib.oracle_barrier.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, D]
        lib.oracle_rayleigh_stiffness.restype = C.c_double
        lib.oracle_rayleigh_stiffness.argtypes = [C.c_int, D, D, D, C.c_int, D, C.c_double]
        lib.oracle_friction.restype = C.c_int
        lib.oracle_friction.argtypes = [C.c_int, D, D, D, C.c_double, C.c_double, D, D]
        lib.oracle_lame.argtypes = [C.c_double, C.c_double, D, D]
        lib.oracle_shell_element.restype = C.c_double
        lib.oracle_shell_element.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, D, D, D, D]
        lib.oracle_dihedral.restype = C.c_double
        lib.oracle_dihedral.argtypes = [C.c_int, D, D]
        lib.oracle_strand.restype = C.c_double
        lib.oracle_strand.argtypes = [C.c_int, D, C.c_double, D, D]
        lib.oracle_rod_stretch.argtypes = [C.c_int, D, C.c_double, C.c_double, D, D]
        lib.oracle_spring.argtypes = [C.c_int, D, C.c_double, C.c_double, C.c_int, D, D]
        lib.oracle_solver_freeze.argtypes = [vp, D, D, C.c_double]
        lib.oracle_solver_energy.restype = C.c_double
        lib.oracle_solver_energy.argtypes = [vp, D, D, C.c_double, C.c_int, D]
        lib.oracle_solver_model_gradient.restype = C.c_int
        lib.oracle_solver_model_gradient.argtypes = [vp, D, D, C.c_double, C.c_int, D]
        lib.oracle_solver_hess_vec.restype = C.c_int
        lib.oracle_solver_hess_vec.argtypes = [vp, D, D, C.c_double, C.c_int, D, D]
        lib.oracle_solver_x_tilde_d.argtypes = [vp, D]
        lib.oracle_solver_freeze_end.argtypes = [vp, D]
        lib.oracle_solver_equilibrium.restype = C.c_int
        lib.oracle_solver_equilibrium.argtypes = [vp, D, C.c_double, C.c_uint32, D, D]
        lib.oracle_solver_adjoint_frozen.restype = C.c_int
        lib.oracle_solver_adjoint_frozen.argtypes = [vp, D, C.c_uint32, C.c_double, D, D, D]
        lib.oracle_solver_frozen_stencils.restype = C.c_uint32
        lib.oracle_solver_frozen_stencils.argtypes = [vp, C.c_uint32, U, D, D, D, C.POINTER(C.c_int32), D, C.POINTER(C.c_int32), D, D, D]
        lib.oracle_morton_codes.argtypes = [C.c_uint32, F, C.c_uint32, U, U]
        lib.oracle_sort_pairs.argtypes = [C.c_uint32, U, U]
        lib.oracle_karras.argtypes = [C.c_uint32, U, U, U]
        lib.oracle_propagate_aabb.argtypes = [C.c_uint32, U, U, F, F]
        _lib = lib
    return _lib


def dp(a):
    return a.ctypes.data_as(D)


def fp(a):
    return a.ctypes.data_as(F)


def up(a):
    return a.ctypes.data_as(U)


def d64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class OracleSolver:
    """The oracle ImplicitSolver on a generated scene (fp64=True gives the FP64 twin)."""

    def __init__(self, scene, fp64=False):
        self.lib = load()
        self.scene = scene  # keeps the arrays alive
        self.h = self.lib.oracle_solver_create(C.byref(scene.desc), 1 if fp64 else 0)
        self.n = self.lib.oracle_solver_nverts(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_solver_destroy(self.h)
            self.h = None

    def step(self, k=1):
        self.lib.oracle_solver_step(self.h, k)

    def positions(self):
        out = np.zeros((self.n, 3), np.float32)
        self.lib.oracle_solver_positions(self.h, fp(out))
        return out

    def positions_d(self):
        out = np.zeros((self.n, 3), np.float64)
        self.lib.oracle_solver_positions_d(self.h, dp(out))
        return out

    def velocities(self):
        out = np.zeros((self.n, 3), np.float32)
        self.lib.oracle_solver_velocities(self.h, fp(out))
        return out

    def set_state(self, x, v):
        x, v = f32(x), f32(v)
        self.lib.oracle_solver_set_state(self.h, fp(x), fp(v))

    def set_rest_scale(self, face_scale):
        """Per shell face rest-shape scale (one value per scene shell face)."""
        sc = d64(face_scale)
        self.lib.oracle_solver_set_rest_scale(self.h, dp(sc), len(sc))

    def adjoint(self, dLdx, nfaces, cg_tol=1e-10, exact=False, xstar=None):
        """dL/ds per shell face for the last step's equilibrium, given dL/dx [n, 3]; exact =
        unprojected elastic blocks; xstar = evaluation point (default: the step's end state)."""
        g = d64(dLdx).reshape(-1)
        out = np.zeros(nfaces)
        res = np.zeros(1)
        xs = None if xstar is None else d64(xstar).reshape(-1)
        it = self.lib.oracle_solver_adjoint(self.h, dp(g), nfaces, cg_tol, 1 if exact else 0, None if xs is None else dp(xs), dp(out), dp(res))
        return out, int(it), float(res[0])

    def metrics(self):
        out = np.zeros(8)
        self.lib.oracle_solver_metrics(self.h, dp(out))
        keys = ["time", "max_reresid", "max_cg_iters", "newton_iters", "last_toi", "penetration_free", "all_finite", "n_edges"]
        return dict(zip(keys, out))

    def newton_trace(self):
        xs, tois = [], []
        k = 0
        while True:
            out = np.zeros((self.n, 3), np.float32)
            t = C.c_double()
            r = self.lib.oracle_solver_newton_trace(self.h, k, fp(out), C.byref(t))
            if r < 0:
                break
            xs.append(out)
            tois.append(t.value)
            k += 1
        return xs, tois

    def stencils(self, x):
        x = f32(x)
        cnt = self.lib.oracle_solver_stencils(self.h, fp(x), 0, None, None, None)
        idx = np.zeros((max(cnt, 1), 4), np.uint32)
        d = np.zeros(max(cnt, 1), np.float32)
        g = np.zeros((max(cnt, 1), 12), np.float32)
        self.lib.oracle_solver_stencils(self.h, fp(x), cnt, up(idx), fp(d), fp(g))
        return idx[:cnt], d[:cnt], g[:cnt]

    def set_kinematic(self, x):
        """Teleport the static collider to x's collider rows (True = applied)."""
        return bool(self.lib.oracle_solver_set_kinematic(self.h, fp(f32(x))))

    def collider_sides(self, x):
        x = f32(x)
        side = np.ones(self.n, np.float32)
        ok = self.lib.oracle_solver_collider_sides(self.h, fp(x), fp(side))
        return side if ok else None

    def assemble(self, x, xt, inv_h2):
        x, xt = f32(x), f32(xt)
        nnz = self.lib.oracle_solver_assemble(self.h, fp(x), fp(xt), inv_h2, None, 0)
        n3 = 3 * self.n
        rp = np.zeros(n3 + 1, np.uint32)
        ci = np.zeros(nnz, np.uint32)
        va = np.zeros(nnz, np.float32)
        rhs = np.zeros(n3, np.float32)
        self.lib.oracle_solver_csr(self.h, up(rp), up(ci), fp(va), fp(rhs))
        return rp, ci, va, rhs

    def kappas(self, x, inv_h2):
        """Rayleigh contact stiffness of every active stencil at x (the stencils() order)."""
        x = f32(x)
        cap = 1 << 20
        out = np.zeros(cap, np.float32)
        self.lib.oracle_solver_assemble(self.h, fp(x), fp(x), inv_h2, fp(out), cap)
        return out

    def x_tilde(self):
        out = np.zeros((self.n, 3), np.float32)
        self.lib.oracle_solver_x_tilde(self.h, fp(out))
        return out

    def x_tilde_d(self):
        out = np.zeros((self.n, 3), np.float64)
        self.lib.oracle_solver_x_tilde_d(self.h, dp(out))
        return out

    # Assembly switches of the derivative checks (oracle_solver_model_gradient / hess_vec);
    # FROZEN_PARAMS = GN_MODEL | UNPROJECTED | 8: the frozen-parameter model with its exact Jacobian.
    GN_MODEL, UNPROJECTED, NO_ROD_BENDS = 1, 2, 4
    FROZEN_PARAMS = 1 | 2 | 8
    TERMS = ("inertia", "shells", "hinges", "rods", "rod_bends", "barriers", "friction", "colliders", "springs")

    def freeze(self, x, xt, inv_h2):
        """Hold the contact quantities of an assembly at x for energy / model_gradient."""
        self.lib.oracle_solver_freeze(self.h, dp(d64(x)), dp(d64(xt)), inv_h2)

    def freeze_end(self, x=None):
        """Hold the quantities of an assembly at the last step's end state (or at x), with the
        step's target and its start as the friction reference, for equilibrium /
        adjoint_frozen / FROZEN_PARAMS."""
        self.lib.oracle_solver_freeze_end(self.h, None if x is None else dp(d64(x).reshape(-1)))

    def equilibrium(self, x=None, tol=1e-13, max_iter=50):
        """(x, Newton updates, final largest free gradient entry) of the frozen-parameter
        equilibrium under the current rest scales, by Newton from x (default: the state)."""
        out = np.zeros(3 * self.n)
        res = np.zeros(1)
        xi = None if x is None else d64(x).reshape(-1)
        it = self.lib.oracle_solver_equilibrium(self.h, None if xi is None else dp(xi), tol, max_iter, dp(out), dp(res))
        assert it >= 0, "freeze_end first"
        return out.reshape(-1, 3), it, float(res[0])

    def adjoint_frozen(self, dLdx, nfaces, xeval, cg_tol=1e-13):
        """(dL/ds per shell face, CG iterations, residual) through the frozen-parameter
        equilibrium at xeval, with its exact Jacobian."""
        g = d64(dLdx).reshape(-1)
        out = np.zeros(nfaces)
        res = np.zeros(1)
        it = self.lib.oracle_solver_adjoint_frozen(self.h, dp(g), nfaces, cg_tol, dp(d64(xeval).reshape(-1)), dp(out), dp(res))
        assert it >= 0, "freeze_end first"
        return out, it, float(res[0])

    def energy(self, x, xt, inv_h2, frozen=False):
        """(total, terms[9]) of the step energy at x; frozen: with the held contact quantities."""
        terms = np.zeros(len(self.TERMS))
        e = self.lib.oracle_solver_energy(self.h, dp(d64(x)), dp(d64(xt)), inv_h2, 1 if frozen else 0, dp(terms))
        return e, terms

    def frozen_stencils(self):
        """The held contact quantities per stencil (after freeze): dict of arrays."""
        ns = self.lib.oracle_solver_frozen_stencils(self.h, 0, None, None, None, None, None, None, None, None, None, None)
        m = max(ns, 1)
        out = dict(idx=np.zeros((m, 4), np.uint32), d=np.zeros(m), g=np.zeros((m, 12)), kappa=np.zeros(m), kind=np.zeros(m, np.int32),
                   ghat=np.zeros(m), fric=np.zeros(m, np.int32), lam=np.zeros(m), n=np.zeros((m, 3)), w=np.zeros((m, 4)))
        I32 = C.POINTER(C.c_int32)
        self.lib.oracle_solver_frozen_stencils(self.h, ns, up(out["idx"]), dp(out["d"]), dp(out["g"]), dp(out["kappa"]),
                                               out["kind"].ctypes.data_as(I32), dp(out["ghat"]), out["fric"].ctypes.data_as(I32),
                                               dp(out["lam"]), dp(out["n"]), dp(out["w"]))
        return {k: v[:ns] for k, v in out.items()}

    def model_gradient(self, x, xt, inv_h2, flags=0):
        g = np.zeros(3 * self.n)
        assert self.lib.oracle_solver_model_gradient(self.h, dp(d64(x)), dp(d64(xt)), inv_h2, flags, dp(g)) == 0
        return g.reshape(-1, 3)

    def hess_vec(self, x, xt, inv_h2, u, flags=0):
        hu = np.zeros(3 * self.n)
        assert self.lib.oracle_solver_hess_vec(self.h, dp(d64(x)), dp(d64(xt)), inv_h2, flags, dp(d64(u).reshape(-1)), dp(hu)) == 0
        return hu.reshape(-1, 3)

    def ccd(self, x, dx):
        x, dx = f32(x), f32(dx)
        return self.lib.oracle_solver_ccd(self.h, fp(x), fp(dx))

    def self_intersects(self, x):
        x = f32(x)
        return bool(self.lib.oracle_solver_self_intersects(self.h, fp(x)))

    def newton_iteration(self, x, xt):
        x = f32(x).copy()
        xt = f32(xt)
        res = C.c_double()
        it = C.c_uint32()
        t = self.lib.oracle_solver_newton_iteration(self.h, fp(x), fp(xt), C.byref(res), C.byref(it))
        return x, t, res.value, it.value


def cg(row_ptr, col_idx, values, b, tol, max_iter):
    lib = load()
    row_ptr, col_idx, values, b = u32(row_ptr), u32(col_idx), f32(values), f32(b)
    x = np.zeros_like(b)
    it = C.c_uint32()
    res = C.c_double()
    lib.oracle_cg(len(b), up(row_ptr), up(col_idx), fp(values), fp(b), tol, max_iter, fp(x), C.byref(it), C.byref(res))
    return x, it.value, res.value


def spmv(row_ptr, col_idx, values, x):
    lib = load()
    row_ptr, col_idx, values, x = u32(row_ptr), u32(col_idx), f32(values), f32(x)
    y = np.zeros_like(x)
    lib.oracle_spmv(len(x), up(row_ptr), up(col_idx), fp(values), fp(x), fp(y))
    return y


def cg_schwarz(row_ptr, col_idx, values, b, tol, max_iter, domain, weight=0.0):
    lib = load()
    row_ptr, col_idx, values, b = u32(row_ptr), u32(col_idx), f32(values), f32(b)
    x = np.zeros_like(b)
    it = C.c_uint32()
    res = C.c_double()
    lib.oracle_cg_schwarz(len(b), up(row_ptr), up(col_idx), fp(values), fp(b), tol, max_iter, domain, weight, fp(x), C.byref(it), C.byref(res))
    return x, it.value, res.value


def schwarz(row_ptr, col_idx, values, domain, r=None, weight=0.0):
    """(domains as a list of vertex arrays, z = M^-1 r or None)."""
    lib = load()
    row_ptr, col_idx, values = u32(row_ptr), u32(col_idx), f32(values)
    n = len(row_ptr) - 1
    nv = n // 3
    dp = np.zeros(nv + 1, np.uint32)
    dv = np.zeros(nv, np.uint32)
    z = np.zeros(n) if r is not None else None
    nd = lib.oracle_schwarz(n, up(row_ptr), up(col_idx), fp(values), domain, weight, up(dp), up(dv),
                            None if r is None else d64(r).ctypes.data_as(D), None if r is None else z.ctypes.data_as(D))
    doms = [dv[dp[i]:dp[i + 1]] for i in range(nd)]
    return doms, z
