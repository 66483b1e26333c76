"""The device drape adjoint (nguo_step_set_rest_scale / nguo_step_adjoint) against the
oracle: scaled rest shapes step like the oracle's, the adjoint matches the oracle's FP64
reference adjoint at the same end state, and FP64 central finite differences of the loss."""
import numpy as np
import pytest

from conftest import nguo_hip
import cases
import drape_case as D
import oracle_py as O

REF_REL = 1e-4  # |g_device - g_fp64| / |g_fp64| (2-norm over the 64 panels)
FD_REL = 1e-4   # per checked parameter, against FP64 central differences (h = 1e-4)


@pytest.mark.gpu
def test_unit_rest_scale_is_bitwise_identity(ctx):
    s = nguo_hip.Scene(101, seed=1)
    a, b = nguo_hip.Stepper(ctx, s), nguo_hip.Stepper(ctx, s)
    b.set_rest_scale(np.ones(s.desc.n_shell_faces, np.float32))
    a.advance(2)
    b.advance(2)
    assert np.array_equal(a.positions(), b.positions())
    a.close()
    b.close()


@pytest.mark.gpu
def test_scaled_rest_shape_step_matches_oracle(ctx):
    s = nguo_hip.Scene(101, seed=1)
    theta, _, _ = D.draws()
    scale = theta[nguo_hip.drape_panels(s)]
    o32, o64 = O.OracleSolver(s), O.OracleSolver(s, fp64=True)
    st = nguo_hip.Stepper(ctx, s)
    for o in (o32, o64):
        o.set_rest_scale(scale)
    st.set_rest_scale(scale.astype(np.float32))
    for _ in range(3):
        o32.step(1)
        o64.step(1)
        st.advance(1)
        x, x32, x64 = st.positions(), o32.positions(), o64.positions()
        spread = np.abs(x32 - x64).max()
        bar = max(1e-5, 1.5 * spread)
        assert np.abs(x - x32).max() <= bar
    st.close()


@pytest.mark.gpu
@pytest.mark.parametrize("exact", [False, True])
def test_adjoint_matches_fp64_reference_in_contact(ctx, exact):
    # Small tube in contact with the hip capsule: one step from a drape state on both sides,
    # then the device adjoint and the oracle's FP64 adjoint at the device's end state with the
    # same dL/dx.
    s = D.tube_scene(newton_iters=8, cg_tol=1e-6)
    panel = nguo_hip.drape_panels(s)
    theta, _, _ = D.draws()
    o = O.OracleSolver(s, fp64=True)
    o.step(6)
    xs, vs = o.positions(), o.velocities()
    st = nguo_hip.Stepper(ctx, s)
    for obj in (o, st):
        obj.set_rest_scale(theta[panel].astype(np.float32) if obj is st else theta[panel])
        obj.set_state(xs, vs)
    o.step(1)
    st.advance(1)
    assert st.metrics().num_stencils > 0, "the state must be in contact"
    x = st.positions()
    rng = np.random.default_rng(11)
    free = s.inv_mass > 0
    dLdx = np.zeros((s.n_verts, 3))
    dLdx[free] = rng.normal(size=(free.sum(), 3)) / free.sum()
    g_dev, stats = st.adjoint(dLdx, tol=1e-7, exact_elastic=exact)
    assert stats.residual <= 1e-7
    g_ref, _, res = o.adjoint(dLdx, s.desc.n_shell_faces, 1e-13, exact=exact, xstar=x.astype(np.float64))
    assert res < 1e-10
    pd = np.bincount(panel, weights=g_dev.astype(np.float64), minlength=64)
    pr = np.bincount(panel, weights=g_ref, minlength=64)
    rel = np.linalg.norm(pd - pr) / np.linalg.norm(pr)
    print(f"panels rel {rel:.3e}, faces rel {np.linalg.norm(g_dev - g_ref) / np.linalg.norm(g_ref):.3e}, iters {stats.iters}")
    assert rel <= REF_REL
    st.close()


@pytest.mark.gpu
def test_adjoint_matches_fp64_reference_on_tube_garment(ctx):
    # The 150 x 133 tube garment with its 64 panels, one step from a device drape state (6
    # steps in, in contact with the body): the device adjoint against the oracle's FP64
    # adjoint at the device's end state, with the same dL/dx.
    s = nguo_hip.Scene(1, seed=1)
    s.desc.params.cg_tol = 1e-10
    panel = nguo_hip.drape_panels(s)
    assert panel.max() == 63
    theta, _, _ = D.draws()
    st = nguo_hip.Stepper(ctx, s)
    st.advance(6)
    xs, vs = st.positions(), st.velocities()
    o = O.OracleSolver(s, fp64=True)
    for obj in (o, st):
        obj.set_rest_scale(theta[panel].astype(np.float32) if obj is st else theta[panel])
        obj.set_state(xs, vs)
    o.step(1)
    st.advance(1)
    assert st.metrics().num_stencils > 0, "the state must be in contact"
    # both adjoints are taken at the device's end state, so the step's own FP32-vs-FP64 spread
    # (CCD-fraction sensitive on this garment) does not enter the comparison
    x = st.positions()
    print(f"end state |gpu - oracle64| {np.abs(x - o.positions()).max():.2e}")
    rng = np.random.default_rng(12)
    free = s.inv_mass > 0
    dLdx = np.zeros((s.n_verts, 3))
    dLdx[free] = rng.normal(size=(free.sum(), 3)) / free.sum()
    g_dev, stats = st.adjoint(dLdx, tol=1e-7)
    assert stats.residual <= 1e-7
    g_ref, _, res = o.adjoint(dLdx, s.desc.n_shell_faces, 1e-12, xstar=x.astype(np.float64))
    assert res < 1e-11
    pd = np.bincount(panel, weights=g_dev.astype(np.float64), minlength=64)
    pr = np.bincount(panel, weights=g_ref, minlength=64)
    rel = np.linalg.norm(pd - pr) / np.linalg.norm(pr)
    print(f"panels rel {rel:.3e}, faces rel {np.linalg.norm(g_dev - g_ref) / np.linalg.norm(g_ref):.3e}, iters {stats.iters}")
    assert rel <= REF_REL
    st.close()


@pytest.mark.gpu
def test_adjoint_matches_central_fd(ctx):
    s = D.sheet_scene()
    theta, theta_t, pick = D.draws()
    panel = nguo_hip.drape_panels(s, axes=(0, 2))
    dr = D.OracleDrape(s, 5, panel, theta_t)
    _, x64 = dr.loss(theta)
    st = nguo_hip.Stepper(ctx, s)
    st.set_rest_scale(theta[panel].astype(np.float32))
    st.set_state(dr.xs, dr.vs)
    st.advance(1)
    assert np.abs(st.positions() - x64).max() <= 1e-5
    g_face, stats = st.adjoint(dr.dloss_dx(x64), tol=1e-7, exact_elastic=True)
    g = np.bincount(panel, weights=g_face.astype(np.float64), minlength=64)
    fd = np.array([dr.central_fd(theta, k) for k in pick])
    rel = np.abs(g[pick] - fd) / np.abs(fd)
    print("fd rel", rel)
    assert rel.max() <= FD_REL
    st.close()


@pytest.mark.gpu
def test_exact_adjoint_reduces_to_exact_elastic_without_contact(ctx):
    # the contact-free sheet with hinge bending off: the exact Jacobian's added blocks are all
    # zero, so the exact form takes one solve and returns the exact-elastic gradient bitwise
    s = D.sheet_scene()
    theta, _, _ = D.draws()
    panel = nguo_hip.drape_panels(s, axes=(0, 2))
    st = nguo_hip.Stepper(ctx, s)
    st.advance(5)
    st.set_rest_scale(theta[panel].astype(np.float32))
    st.advance(1)
    free = s.inv_mass > 0
    dLdx = np.zeros((s.n_verts, 3))
    dLdx[free] = np.random.default_rng(16).normal(size=(free.sum(), 3)) / free.sum()
    g_el, st_el = st.adjoint(dLdx, tol=1e-7, hessian=st.ADJ_EXACT_ELASTIC)
    g_ex, st_ex = st.adjoint(dLdx, tol=1e-7, hessian=st.ADJ_EXACT)
    assert st_ex.sweeps == 1 and st_ex.iters == st_el.iters
    assert np.array_equal(g_ex, g_el)
    st.close()


@pytest.mark.gpu
def test_adjoint_needs_a_step(ctx):
    s = nguo_hip.Scene(100, seed=1)
    st = nguo_hip.Stepper(ctx, s)
    g = np.zeros((s.n_verts, 3), np.float32)
    with pytest.raises(nguo_hip.NguoError):
        st.adjoint(g)
    st.advance(1)
    st.adjoint(g)  # zero loss gradient: fine after a step
    st.set_rest_scale(np.ones(s.desc.n_shell_faces, np.float32))
    with pytest.raises(nguo_hip.NguoError):
        st.adjoint(g)
    with pytest.raises(nguo_hip.NguoError):
        st.set_rest_scale(np.ones(3, np.float32))
    st.close()


def _exact_vs_frozen_reference(ctx, s, warm, seed):
    """`warm` oracle steps, then one step from that state on both sides, then the device's
    exact adjoint (NGUO_ADJ_EXACT) against the oracle's frozen-parameter adjoint (parameters
    held at the device's end state, exact Jacobian, FP64 CG) at the device's end state, with
    the same dL/dx."""
    panel = nguo_hip.drape_panels(s)
    theta, _, _ = D.draws()
    st = nguo_hip.Stepper(ctx, s)
    o = O.OracleSolver(s, fp64=True)
    o.step(warm)
    xs, vs = o.positions(), o.velocities()
    for obj in (o, st):
        obj.set_rest_scale(theta[panel].astype(np.float32) if obj is st else theta[panel])
        obj.set_state(xs, vs)
    o.step(1)
    st.advance(1)
    m = st.metrics()
    ns = m.num_stencils
    print(f"differentiated step: newton {m.newton_iters}, last toi {m.last_toi}, penetration free {m.penetration_free}")
    assert ns > 0 or m.num_contacts > 0, "the state must be in contact"
    x = st.positions().astype(np.float64)
    rng = np.random.default_rng(seed)
    free = s.inv_mass > 0
    dLdx = np.zeros((s.n_verts, 3))
    dLdx[free] = rng.normal(size=(free.sum(), 3)) / free.sum()
    g_dev, stats = st.adjoint(dLdx, tol=1e-7, hessian=st.ADJ_EXACT)
    o.freeze_end(x)
    g_ref, _, res = o.adjoint_frozen(dLdx, s.desc.n_shell_faces, x, cg_tol=1e-12)
    assert res < 1e-10
    g_newton, _ = st.adjoint(dLdx, tol=1e-7)
    pd = np.bincount(panel, weights=g_dev.astype(np.float64), minlength=64)
    pr = np.bincount(panel, weights=g_ref, minlength=64)
    pn = np.bincount(panel, weights=g_newton.astype(np.float64), minlength=64)
    rel = np.linalg.norm(pd - pr) / np.linalg.norm(pr)
    reln = np.linalg.norm(pn - pr) / np.linalg.norm(pr)
    print(f"stencils {ns}: exact panels rel {rel:.3e} (Newton Hessian {reln:.3e}), solves {stats.sweeps}, "
          f"PCG iterations {stats.iters}, final update {stats.residual:.2e}")
    st.close()
    return rel, reln, stats


@pytest.mark.gpu
def test_exact_adjoint_matches_frozen_reference_in_contact(ctx):
    # config 101 seven steps in (the state whose frozen-parameter adjoint the CPU tests pin
    # against central differences at 1.2e-6)
    s = D.tube_scene()
    rel, reln, stats = _exact_vs_frozen_reference(ctx, s, 7, 13)
    assert stats.residual <= 1e-6
    assert rel <= REF_REL
    assert reln > 10 * REF_REL  # the terms the Newton Hessian leaves out are resolved


@pytest.mark.gpu
def test_exact_adjoint_matches_frozen_reference_on_the_sphere_collider(ctx):
    # The 10 x 10 sheet on the analytic sphere 20 steps in: the sphere's curvature term and the
    # hinges' second derivatives (k_adj_spheres, k_adj_hinges), no contact stencils. On this
    # state the FP32 evaluation of the terms every form shares (the sphere's gap, the friction's
    # slip) moves the gradient by ~2e-4 (the exact-elastic form shows the same against the
    # oracle; the oracle's own FP32 mirror is 3.7% off its FP64 twin here), so the bar is that
    # shared part plus REF_REL, and the exact terms themselves (exact minus exact-elastic, on
    # both sides) are compared separately.
    s = cases.collider_scene()
    p = s.desc.params
    p.newton_iters = 300
    p.cg_tol = 1e-10
    p.cg_max_iter = 100000
    panel = nguo_hip.drape_panels(s)
    theta, _, _ = D.draws()
    o = O.OracleSolver(s, fp64=True)
    o.step(20)
    xs, vs = o.positions(), o.velocities()
    st = nguo_hip.Stepper(ctx, s)
    for obj in (o, st):
        obj.set_rest_scale(theta[panel].astype(np.float32) if obj is st else theta[panel])
        obj.set_state(xs, vs)
    o.step(1)
    st.advance(1)
    assert st.metrics().num_contacts > 0
    x = st.positions().astype(np.float64)
    free = s.inv_mass > 0
    dLdx = np.zeros((s.n_verts, 3))
    dLdx[free] = np.random.default_rng(15).normal(size=(free.sum(), 3)) / free.sum()
    g_ex, stats = st.adjoint(dLdx, tol=1e-7, hessian=st.ADJ_EXACT)
    g_el, _ = st.adjoint(dLdx, tol=1e-7, hessian=st.ADJ_EXACT_ELASTIC)
    o.freeze_end(x)
    r_ex, _, res = o.adjoint_frozen(dLdx, s.desc.n_shell_faces, x, cg_tol=1e-13)
    r_el, _, res2 = o.adjoint(dLdx, s.desc.n_shell_faces, 1e-13, exact=True, xstar=x)
    assert res < 1e-10 and res2 < 1e-10

    def pan(g):
        return np.bincount(panel, weights=np.asarray(g, np.float64), minlength=64)

    rel_ex = np.linalg.norm(pan(g_ex) - pan(r_ex)) / np.linalg.norm(pan(r_ex))
    rel_el = np.linalg.norm(pan(g_el) - pan(r_el)) / np.linalg.norm(pan(r_el))
    d_dev, d_ref = pan(g_ex) - pan(g_el), pan(r_ex) - pan(r_el)
    rel_terms = np.linalg.norm(d_dev - d_ref) / np.linalg.norm(d_ref)
    size = np.linalg.norm(d_ref) / np.linalg.norm(pan(r_ex))
    print(f"exact {rel_ex:.3e}, exact-elastic {rel_el:.3e}, the exact terms {rel_terms:.3e} (they are {size:.3e} of the gradient), "
          f"solves {stats.sweeps}")
    assert stats.residual <= 1e-6
    assert rel_ex <= REF_REL + rel_el
    assert size > 10 * REF_REL and rel_terms <= 1e-2
    st.close()


@pytest.mark.gpu
def test_exact_adjoint_is_refused_where_the_jacobian_is_indefinite(ctx):
    # The tube garment one step (the scene's 4-iteration Newton cap) from a device drape state
    # 6 steps in: the step ends short of its root and the exact Jacobian there is indefinite.
    # The device refuses the exact form rather than return a gradient, and the oracle's FP64
    # CG on the same Jacobian (parameters held at the device's end state) breaks down too.
    s = nguo_hip.Scene(1, seed=1)
    panel = nguo_hip.drape_panels(s)
    theta, _, _ = D.draws()
    st = nguo_hip.Stepper(ctx, s)
    st.advance(6)
    xs, vs = st.positions(), st.velocities()
    o = O.OracleSolver(s, fp64=True)
    for obj in (o, st):
        obj.set_rest_scale(theta[panel].astype(np.float32) if obj is st else theta[panel])
        obj.set_state(xs, vs)
    o.step(1)
    st.advance(1)
    x = st.positions().astype(np.float64)
    free = s.inv_mass > 0
    dLdx = np.zeros((s.n_verts, 3))
    dLdx[free] = np.random.default_rng(14).normal(size=(free.sum(), 3)) / free.sum()
    with pytest.raises(nguo_hip.NguoError) as e:
        st.adjoint(dLdx, tol=1e-7, hessian=st.ADJ_EXACT)
    assert "did not converge" in str(e.value) or "not positive definite" in str(e.value), str(e.value)
    o.freeze_end(x)
    _, it, res = o.adjoint_frozen(dLdx, s.desc.n_shell_faces, x, cg_tol=1e-12)
    print(f"oracle CG on the exact Jacobian: {it} iterations, residual {res:.3e}")
    assert res > 1e-3
    # the Newton-Hessian adjoint still solves there
    g, stats = st.adjoint(dLdx, tol=1e-7)
    assert np.all(np.isfinite(g)) and stats.residual <= 1e-7
    st.close()


FD_REL_FP32_STATE = 2e-4  # |g - fd| / max |fd| over the 8 panels, for the device's FP32 end state


@pytest.mark.gpu
def test_exact_adjoint_matches_central_fd_in_contact(ctx):
    # The device's exact adjoint at its own end state on config 101 seven steps in, against
    # FP64 central differences (h = 1e-5) through the frozen-parameter root with the
    # parameters held at the device's end state (the oracle's Newton re-solves the root from
    # there at theta +- h on the 8 checked panels). The loss is linear, L = c . x with c =
    # 2 (x* - target) / n fixed (the distance loss's gradient at x*, target the drape of a
    # second draw), so the device and the differences differentiate the same function. The
    # device's FP32 end state sits ~1.6e-6 m from the FP64 root, which alone moves the
    # derivative by ~1e-4 of its largest component (the oracle's adjoint at the root itself
    # meets the differences to 4e-8 there), hence the bar relative to the largest derivative.
    s = D.tube_scene()
    panel = nguo_hip.drape_panels(s)
    theta, theta_t, pick = D.draws()
    o = O.OracleSolver(s, fp64=True)
    o.step(7)
    xs, vs = o.positions(), o.velocities()
    o.set_rest_scale(theta_t[panel])
    o.step(1)
    target = o.positions_d()
    st = nguo_hip.Stepper(ctx, s)
    for obj in (o, st):
        obj.set_rest_scale(theta[panel].astype(np.float32) if obj is st else theta[panel])
        obj.set_state(xs, vs)
    o.step(1)
    st.advance(1)
    assert st.metrics().num_stencils > 0
    x = st.positions().astype(np.float64)
    free = s.inv_mass > 0
    c = np.zeros((s.n_verts, 3))
    c[free] = 2.0 * (x - target)[free] / free.sum()
    g_face, stats = st.adjoint(c, tol=1e-7, hessian=st.ADJ_EXACT)
    g = np.bincount(panel, weights=g_face.astype(np.float64), minlength=64)
    o.freeze_end(x)
    h = 1e-5
    fd = []
    for k in pick:
        e = np.zeros(64)
        e[k] = h
        lv = []
        for th in (theta + e, theta - e):
            o.set_rest_scale(th[panel])
            xr, it, res = o.equilibrium(x)
            assert res < 1e-10
            lv.append(float((c * xr).sum()))
        fd.append((lv[0] - lv[1]) / (2 * h))
    fd = np.array(fd)
    err = np.abs(g[pick] - fd).max() / np.abs(fd).max()
    print(f"max |g - fd| / max |fd| = {err:.3e}; per panel", np.abs(g[pick] - fd) / np.abs(fd))
    assert err <= FD_REL_FP32_STATE
    st.close()
