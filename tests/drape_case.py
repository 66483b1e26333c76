"""Drape-adjoint workloads shared by the CPU and GPU tests: a fixed start state, 64 panel
rest-scale parameters, a target from a second parameter draw, the loss
L = mean over free vertices of |x*(theta) - target|^2 and its FP64 central differences
through the oracle's FP64 step."""
import numpy as np

from conftest import nguo_hip
import oracle_py as O

FD_H = 1e-4


def sheet_scene():
    """10 x 10 sheet (config 100) without contact or hinge bending: every term of its step
    gradient has an exact second derivative among the elastic blocks."""
    s = nguo_hip.Scene(100, seed=1)
    p = s.desc.params
    p.self_collision = 0
    p.bend_stiffness = 0.0
    p.newton_iters = 300
    p.cg_tol = 1e-10
    p.cg_max_iter = 100000
    return s


def tube_scene(newton_iters=300, cg_tol=1e-10):
    """Small tube over a hip capsule with a seam rod (config 101): contact and friction."""
    s = nguo_hip.Scene(101, seed=1)
    p = s.desc.params
    p.newton_iters = newton_iters
    p.cg_tol = cg_tol
    p.cg_max_iter = 100000
    return s


class OracleDrape:
    """The loss and its central differences on the oracle's FP64 step from a fixed start."""

    def __init__(self, scene, warm, panel, theta_target):
        self.scene = scene
        self.o = O.OracleSolver(scene, fp64=True)
        if warm:
            self.o.step(warm)
        self.xs, self.vs = self.o.positions(), self.o.velocities()
        self.panel = panel
        self.free = scene.inv_mass > 0
        self.nf = scene.desc.n_shell_faces
        self.target = self.run(theta_target)

    def run(self, theta):
        self.o.set_state(self.xs, self.vs)
        self.o.set_rest_scale(theta[self.panel])
        self.o.step(1)
        return self.o.positions_d()

    def loss(self, theta):
        x = self.run(theta)
        d = (x - self.target)[self.free]
        return float((d * d).sum(axis=1).mean()), x

    def dloss_dx(self, x):
        g = np.zeros_like(x)
        g[self.free] = 2.0 * (x - self.target)[self.free] / self.free.sum()
        return g

    def central_fd(self, theta, k, h=FD_H):
        e = np.zeros_like(theta)
        e[k] = h
        return (self.loss(theta + e)[0] - self.loss(theta - e)[0]) / (2.0 * h)

    def adjoint(self, theta, exact, xstar=None):
        """Per-panel dL/dtheta from the oracle's FP64 adjoint at the step's end (or xstar)."""
        _, x = self.loss(theta)
        g_face, _, res = self.o.adjoint(self.dloss_dx(x), self.nf, 1e-13, exact=exact, xstar=xstar)
        return np.bincount(self.panel, weights=g_face, minlength=64), res


def draws():
    """(theta, theta_target, the 8 checked parameters): N(1, 0.02) with seeds 4 and 5, then 8
    distinct panels with seed 6."""
    theta = np.random.default_rng(4).normal(1.0, 0.02, 64)
    theta_t = np.random.default_rng(5).normal(1.0, 0.02, 64)
    pick = np.random.default_rng(6).choice(64, 8, replace=False)
    return theta, theta_t, pick


class FrozenDrape:
    """The drape equilibrium with the step's contact parameters held (the frozen-parameter
    model): one FP64 oracle step from the fixed start under theta gives x*; the stencil set,
    Rayleigh stiffnesses, barrier kinds and friction quantities (normal, weights, lambda) of an
    assembly at x* are held, and x(theta') is the root of the step gradient with them (distances,
    angles and normals of x itself), by Newton from x* to a largest free gradient entry of
    ~1e-13. The loss and its central differences run through that root; the adjoint is the
    oracle's exact Jacobian of the same gradient."""

    def __init__(self, scene, warm, panel, theta, theta_target):
        self.scene = scene
        self.o = o = O.OracleSolver(scene, fp64=True)
        o.step(warm)
        xs, vs = o.positions(), o.velocities()
        o.set_state(xs, vs)
        o.set_rest_scale(theta_target[panel])
        o.step(1)
        self.target = o.positions_d()
        o.set_state(xs, vs)
        o.set_rest_scale(theta[panel])
        o.step(1)
        self.step_metrics = o.metrics()
        self.xstar = o.positions_d()
        o.freeze_end()
        self.panel = panel
        self.free = scene.inv_mass > 0
        self.nf = scene.desc.n_shell_faces

    def solve(self, theta):
        self.o.set_rest_scale(theta[self.panel])
        return self.o.equilibrium(self.xstar)

    def loss(self, theta):
        x = self.solve(theta)[0]
        d = (x - self.target)[self.free]
        return float((d * d).sum(axis=1).mean()), x

    def dloss_dx(self, x):
        g = np.zeros_like(x)
        g[self.free] = 2.0 * (x - self.target)[self.free] / self.free.sum()
        return g

    def central_fd(self, theta, k, h):
        e = np.zeros_like(theta)
        e[k] = h
        return (self.loss(theta + e)[0] - self.loss(theta - e)[0]) / (2.0 * h)

    def adjoint(self, theta):
        """(per-panel dL/dtheta, CG residual, the equilibrium) through the frozen-parameter root."""
        _, x = self.loss(theta)
        g_face, _, res = self.o.adjoint_frozen(self.dloss_dx(x), self.nf, x)
        return np.bincount(self.panel, weights=g_face, minlength=64), res, x
