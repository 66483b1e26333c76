"""The oracle's step energy and what it pins (CPU). The step energy is the incremental
potential a Newton iteration of engine-core/src/implicit.rs:1951-2191 decreases: inertia,
shells, hinges, rod stretch, strand bending, edge springs, barriers, friction and analytic
colliders, each the
term whose gradient the assembly adds (implicit.rs:2280-3155). With the quantities an
assembly holds fixed taken at a point x_ref (stencil set, Rayleigh stiffnesses, friction normal
force, normals and weights), two checks pin the whole assembled system at x_ref:
  1. the assembled gradient is the gradient of the energy (central differences in FP64);
  2. the assembled Hessian, with shells and rod segments unprojected, is the Jacobian of the
     gradient with the Gauss-Newton directions also held (barrier and hinge gradients,
     friction lambda and projector), again by central differences. Strand bending is left out
     of this one: its Wu-Kim Hessian is eigen-projected by design (rod_bend.rs:1-20) and is
     checked element by element elsewhere.
Each check is shown to fail on a planted mistake: the gradient of the same state with one
term dropped (friction, hinge bending) in the oracle."""
import numpy as np
import pytest

import oracle_py as O
from cases import scene

H = 1e-8       # central-difference step (m)
GRAD_TOL = 1e-6  # relative to |g| |u| (the bound of g.u): measured <= 6e-10 on these states
HESS_TOL = 1e-6  # relative 2-norm: measured <= 3e-8
FLAGS_H = O.OracleSolver.GN_MODEL | O.OracleSolver.UNPROJECTED | O.OracleSolver.NO_ROD_BENDS


def newton_point(cfg, steps, **params):
    """(FP64 twin at the start of a step, x_ref = that step's first Newton iterate, x_tilde,
    1/h^2): contacts active and friction sliding at x_ref."""
    o32 = O.OracleSolver(scene(cfg))
    o32.step(steps)
    x, v = o32.positions(), o32.velocities()
    s = scene(cfg, **params)
    o = O.OracleSolver(s, fp64=True)
    o.set_state(x, v)
    xt = o.x_tilde_d()
    o.step(1)
    xr = o.newton_trace()[0][0].astype(np.float64)
    o.set_state(x, v)  # the friction reference is the step's start
    return o, s, xr, xt, 1.0 / float(s.params.dt) ** 2, (x, v)


def directions(o, s, xr, rng):
    free = s.inv_mass > 0
    idx, _, _ = o.stencils(xr)
    part = np.zeros(s.n_verts, bool)
    part[idx.reshape(-1)] = True
    out = []
    for mask in (free, free & part):
        if mask.sum() == 0:
            continue
        u = np.zeros((s.n_verts, 3))
        u[mask] = rng.normal(size=(mask.sum(), 3))
        out.append(u)
    return out


def fd_energy(o, xr, xt, inv_h2, u):
    ep, _ = o.energy(xr + H * u, xt, inv_h2, frozen=True)
    em, _ = o.energy(xr - H * u, xt, inv_h2, frozen=True)
    return (ep - em) / (2 * H)


CASES = [(100, 12), (101, 10), (101, 25), ("collider", 14), ("springs", 14)]


@pytest.mark.parametrize("cfg,steps", CASES)
def test_assembled_gradient_is_the_energy_gradient(cfg, steps):
    o, s, xr, xt, inv_h2, _ = newton_point(cfg, steps)
    o.freeze(xr, xt, inv_h2)
    e_frozen, terms = o.energy(xr, xt, inv_h2, frozen=True)
    e_own, _ = o.energy(xr, xt, inv_h2, frozen=False)
    assert e_frozen == e_own  # at x_ref the held quantities are x_ref's own
    names = dict(zip(O.OracleSolver.TERMS, terms))
    print({k: f"{v:.3e}" for k, v in names.items()})
    if cfg == "collider":
        assert names["colliders"] > 0 and names["barriers"] == 0
    elif cfg == "springs":
        assert names["springs"] > 0 and names["shells"] == 0 and names["barriers"] > 0 and names["hinges"] > 0
    else:
        assert names["barriers"] > 0 and names["friction"] > 0 and names["shells"] > 0 and names["hinges"] > 0
    g = o.model_gradient(xr, xt, inv_h2)
    # the frozen model's gradient at x_ref is the plain assembled gradient
    gm = o.model_gradient(xr, xt, inv_h2, O.OracleSolver.GN_MODEL)
    assert np.abs(gm - g).max() <= 1e-12 * np.abs(g).max()
    for u in directions(o, s, xr, np.random.default_rng(1)):
        fd, gu = fd_energy(o, xr, xt, inv_h2, u), float((g * u).sum())
        scale = np.linalg.norm(g) * np.linalg.norm(u)
        print(f"fd {fd:.10e} g.u {gu:.10e} rel {abs(fd - gu) / scale:.2e}")
        assert abs(fd - gu) <= GRAD_TOL * scale


@pytest.mark.parametrize("cfg,steps", CASES)
def test_assembled_hessian_is_the_jacobian_of_the_frozen_gradient(cfg, steps):
    o, s, xr, xt, inv_h2, _ = newton_point(cfg, steps)
    o.freeze(xr, xt, inv_h2)
    free = s.inv_mass > 0
    for u in directions(o, s, xr, np.random.default_rng(2)):
        hu = o.hess_vec(xr, xt, inv_h2, u, FLAGS_H)
        fd = (o.model_gradient(xr + H * u, xt, inv_h2, FLAGS_H) - o.model_gradient(xr - H * u, xt, inv_h2, FLAGS_H)) / (2 * H)
        rel = np.linalg.norm(fd[free] - hu[free]) / np.linalg.norm(hu[free])
        print(f"|H u - fd| / |H u| {rel:.2e}")
        assert rel <= HESS_TOL


@pytest.mark.parametrize("cfg,steps,dropped", [(101, 10, {"friction": 0.0}), (101, 10, {"bend_stiffness": 0.0}),
                                               ("springs", 14, {"stretch_k": 0.0})])
def test_checks_fail_on_a_dropped_term(cfg, steps, dropped):
    # the same state assembled with one term missing: the energy check (and the Hessian check
    # against the full gradient) must reject it
    o, s, xr, xt, inv_h2, (x, v) = newton_point(cfg, steps)
    o.freeze(xr, xt, inv_h2)
    bad = O.OracleSolver(scene(cfg, **dropped), fp64=True)
    bad.set_state(x, v)
    g_bad = bad.model_gradient(xr, xt, inv_h2)
    bad.freeze(xr, xt, inv_h2)
    worst_g = worst_h = 0.0
    free = s.inv_mass > 0
    for u in directions(o, s, xr, np.random.default_rng(3)):
        fd, gu = fd_energy(o, xr, xt, inv_h2, u), float((g_bad * u).sum())
        worst_g = max(worst_g, abs(fd - gu) / (np.linalg.norm(g_bad) * np.linalg.norm(u)))
        hu = bad.hess_vec(xr, xt, inv_h2, u, FLAGS_H)
        fdh = (o.model_gradient(xr + H * u, xt, inv_h2, FLAGS_H) - o.model_gradient(xr - H * u, xt, inv_h2, FLAGS_H)) / (2 * H)
        worst_h = max(worst_h, np.linalg.norm(fdh[free] - hu[free]) / np.linalg.norm(hu[free]))
    print(f"dropped {dropped}: gradient rel {worst_g:.2e}, Hessian rel {worst_h:.2e}")
    assert worst_g > 100 * GRAD_TOL and worst_h > 100 * HESS_TOL
