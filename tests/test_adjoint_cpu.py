"""The oracle's drape adjoint (the FP64 reference gradient) pinned by FP64 central finite
differences of the loss through the oracle's own FP64 step (h = 1e-4)."""
import numpy as np
import pytest

from conftest import nguo_hip
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
