"""The oracle's drape adjoint (the FP64 reference gradient) pinned by FP64 central finite
differences of the loss: through the oracle's own FP64 step (h = 1e-4) on the contact-free
sheet, and through the frozen-parameter equilibrium (h = 1e-5) on the tube in contact."""
import numpy as np
import pytest

from conftest import nguo_hip
import cases
import drape_case as D

FD_REL = 1e-4  # |adjoint - central FD| / |central FD| per checked parameter


@pytest.fixture(scope="module")
def sheet():
    s = D.sheet_scene()
    theta, theta_t, pick = D.draws()
    dr = D.OracleDrape(s, 5, nguo_hip.drape_panels(s, axes=(0, 2)), theta_t)
    fd = np.array([dr.central_fd(theta, k) for k in pick])
    return dr, theta, pick, fd


def test_exact_elastic_adjoint_matches_central_fd(sheet):
    dr, theta, pick, fd = sheet
    g, res = dr.adjoint(theta, exact=True)
    assert res < 1e-10
    assert np.all(np.abs(fd) > 0)
    rel = np.abs(g[pick] - fd) / np.abs(fd)
    assert rel.max() <= FD_REL, rel


def test_projected_hessian_adjoint_is_not_the_derivative(sheet):
    # The Newton iteration's PSD-projected shell blocks are not the second derivative: the
    # same adjoint with them misses the finite differences, so the check above can fail.
    dr, theta, pick, fd = sheet
    g, _ = dr.adjoint(theta, exact=False)
    rel = np.abs(g[pick] - fd) / np.abs(fd)
    assert rel.max() > 1e-2


def test_adjoint_is_linear_in_the_loss_gradient(sheet):
    dr, theta, _, _ = sheet
    _, x = dr.loss(theta)
    rng = np.random.default_rng(9)
    g1, g2 = rng.normal(size=x.shape), rng.normal(size=x.shape)
    a1 = dr.o.adjoint(g1, dr.nf, 1e-13, exact=True)[0]
    a2 = dr.o.adjoint(g2, dr.nf, 1e-13, exact=True)[0]
    a12 = dr.o.adjoint(2.0 * g1 - 3.0 * g2, dr.nf, 1e-13, exact=True)[0]
    assert np.abs(a12 - (2.0 * a1 - 3.0 * a2)).max() <= 1e-8 * np.abs(a12).max()


def test_unit_rest_scale_leaves_the_step_unchanged():
    s = D.tube_scene(newton_iters=4, cg_tol=1e-6)
    a, b = D.O.OracleSolver(s), D.O.OracleSolver(s)
    b.set_rest_scale(np.ones(s.desc.n_shell_faces))
    a.step(2)
    b.step(2)
    assert np.array_equal(a.positions(), b.positions())


def test_drape_panels_cover_64_panels():
    s = nguo_hip.Scene(101, seed=1)
    p = nguo_hip.drape_panels(s)
    assert p.shape == (s.desc.n_shell_faces,)
    assert p.min() == 0 and p.max() == 63
    assert np.all(np.bincount(p, minlength=64) > 0)


# ---------------------------------------------------------------- in contact (config 101)
FD_H_CONTACT = 1e-5  # at h = 1e-4 the O(h^2) term of the central difference reaches 1.1e-4


@pytest.fixture(scope="module")
def tube():
    """Config 101 after 7 steps: the next step converges (70 Newton iterations) penetration
    free with 73 contact stencils, 68 of them with friction."""
    s = D.tube_scene()
    theta, theta_t, pick = D.draws()
    dr = D.FrozenDrape(s, 7, nguo_hip.drape_panels(s), theta, theta_t)
    fd = np.array([dr.central_fd(theta, k, FD_H_CONTACT) for k in pick])
    return dr, theta, pick, fd


def test_tube_state_is_a_converged_contact_drape(tube):
    dr, theta, _, _ = tube
    m = dr.step_metrics
    assert m["newton_iters"] < 300 and m["penetration_free"] == 1.0 and m["last_toi"] == 1.0, m
    st = dr.o.frozen_stencils()
    assert len(st["d"]) >= 50 and st["fric"].sum() >= 50
    x, it, res = dr.solve(theta)
    # the held-parameter root sits within a micron of the step's end state
    assert res < 1e-10 and it <= 10 and np.abs(x - dr.xstar).max() < 1e-5, (res, it)


def test_frozen_jacobian_matches_gradient_differences(tube):
    # the assembled Jacobian (barrier b'(d) d(grad d)/dx, hinge (theta - rest) d(grad theta)/dx,
    # exact strand blocks, unprojected elastic blocks) against central differences of the
    # assembled gradient of the same model, along random free directions
    dr, _, _, _ = tube
    o, x = dr.o, dr.xstar
    xt = o.x_tilde_d()
    h = float(np.float32(dr.scene.desc.params.dt))
    F = o.FROZEN_PARAMS
    rng = np.random.default_rng(3)
    for _ in range(3):
        u = rng.normal(size=x.shape)
        u[~dr.free] = 0
        eps = 1e-7
        fd = (o.model_gradient(x + eps * u, xt, 1.0 / (h * h), F) - o.model_gradient(x - eps * u, xt, 1.0 / (h * h), F)) / (2 * eps)
        fd[~dr.free] = 0
        hu = o.hess_vec(x, xt, 1.0 / (h * h), u, F)
        assert np.abs(hu - fd).max() <= 2e-5 * np.abs(fd).max()


def test_frozen_parameter_adjoint_matches_central_fd_in_contact(tube):
    dr, theta, pick, fd = tube
    g, res, _ = dr.adjoint(theta)
    assert res < 1e-10
    assert np.all(np.abs(fd) > 0)
    rel = np.abs(g[pick] - fd) / np.abs(fd)
    assert rel.max() <= FD_REL, rel


def test_newton_hessian_adjoint_misses_in_contact(tube):
    # the Newton Hessian (rank-one barrier and hinge terms, Wu-Kim strand blocks), even with
    # the elastic blocks unprojected, is not the Jacobian in contact
    dr, theta, pick, fd = tube
    _, _, x = dr.adjoint(theta)
    g_face, _, _ = dr.o.adjoint(dr.dloss_dx(x), dr.nf, 1e-13, exact=True, xstar=x)
    g = np.bincount(dr.panel, weights=g_face, minlength=64)
    rel = np.abs(g[pick] - fd) / np.abs(fd)
    assert rel.max() > 1e-2, rel


@pytest.fixture(scope="module")
def sphere():
    """The 10 x 10 sheet on the analytic sphere after 20 steps (65 Newton iterations to converge
    the next one; the sphere's barrier and friction hold the sheet, no self contact): the
    curvature term b'(d) (I - n n^T) / r of the sphere's barrier and the hinges' second
    derivatives. Checked on the first 8 panels that hold faces of the flat sheet."""
    s = cases.collider_scene()
    p = s.desc.params
    p.newton_iters = 300
    p.cg_tol = 1e-10
    p.cg_max_iter = 100000
    panel = nguo_hip.drape_panels(s)
    pick = np.flatnonzero(np.bincount(panel, minlength=64) > 0)[:8]
    theta, theta_t, _ = D.draws()
    dr = D.FrozenDrape(s, 20, panel, theta, theta_t)
    fd = np.array([dr.central_fd(theta, k, FD_H_CONTACT) for k in pick])
    return dr, theta, pick, fd


def test_frozen_parameter_adjoint_on_the_sphere_collider(sphere):
    dr, theta, pick, fd = sphere
    assert dr.step_metrics["newton_iters"] < 300
    g, res, x = dr.adjoint(theta)
    assert res < 1e-10 and np.all(np.abs(fd) > 0)
    rel = np.abs(g[pick] - fd) / np.abs(fd)
    assert rel.max() <= FD_REL, rel
    g_face, _, _ = dr.o.adjoint(dr.dloss_dx(x), dr.nf, 1e-13, exact=True, xstar=x)
    gn = np.bincount(dr.panel, weights=g_face, minlength=64)
    assert (np.abs(gn[pick] - fd) / np.abs(fd)).max() > 1e-2
