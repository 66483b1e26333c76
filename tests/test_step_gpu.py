"""The resident implicit step on the device against the oracle (engine-core/src/implicit.rs
ImplicitSolver::step restated in oracle/): per-step parity from identical inputs, bit-exact
sorted contact stencil lists, collider sides, closed-form free fall, invariants."""
import numpy as np
import pytest

from conftest import nguo_hip
import oracle_py as O
from cases import check_stencils, inside_state, max_stretch_ratio, pinned_state, scene

POS_TOL = 1e-5  # |x_gpu - x_oracle|_inf / max(1, |x|_inf) after each Newton iteration / step


def rel_err(a, b):
    return np.abs(a - b).max() / max(1.0, np.abs(b).max())


def make(ctx, cfg):
    s = nguo_hip.Scene(cfg, seed=cfg)
    return s, nguo_hip.Stepper(ctx, s)


@pytest.mark.gpu
def test_free_fall_closed_form(ctx):
    s = nguo_hip.Scene(100, seed=0)
    s.desc.params.self_collision = 0
    st = nguo_hip.Stepper(ctx, s)
    x0 = st.positions().astype(np.float64)
    st.advance(1)
    x1 = st.positions().astype(np.float64)
    h = s.params.dt
    free = s.inv_mass > 0
    want = x0.copy()
    want[free, 1] += h * h * -9.8
    assert np.abs(x1 - want).max() < 2e-6
    m = st.metrics()
    assert m.all_finite == 1 and m.max_reresid <= 1e-6
    st.close()


def oracle_state(cfg, steps, cg_tol=None):
    s = scene(cfg) if cg_tol is None else scene(cfg, cg_tol=cg_tol)
    o = O.OracleSolver(s)
    if steps:
        o.step(steps)
    return s, o


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps", [(100, 30), (0, 16), (0, 30), (101, 25)])
def test_stencils_bitexact(ctx, cfg, steps):
    s, o = oracle_state(cfg, steps)
    check_stencils(ctx, s, o, o.positions(), o.velocities())


@pytest.mark.gpu
def test_stencils_bitexact_inside_collider(ctx):
    # a vertex 0.4 dhat behind a collider facet (side -1) meets every collider facet of its
    # grid cells at a negative signed gap; one 2.5 dhat behind meets none (side +1)
    s, x, v, vs, vd, _, _ = inside_state()
    o = O.OracleSolver(s)
    side = o.collider_sides(x)
    assert side[vs] == -1 and side[vd] == 1
    idx, d = check_stencils(ctx, s, o, x, v)
    static = s.static.astype(bool)
    assert sum(1 for r, dd in zip(idx, d) if r[0] == vs and static[r[1:]].all() and dd < 0) >= 2


@pytest.mark.gpu
def test_stencils_bitexact_pinned_cloth(ctx):
    # stencils with pinned cloth participants: their Rayleigh stiffness reads the elastic
    # blocks between pinned and free vertices (the mixed-element table)
    s, x, v, pins = pinned_state()
    o = O.OracleSolver(s)
    o.set_state(x, v)
    idx, d = check_stencils(ctx, s, o, x, v)
    assert sum(1 for r in idx if np.isin(r, pins).any()) >= 3


# Both sides solve each Newton system to 1e-10 relative residual here, so the comparison
# measures the implementation (assembly, contact, CCD) and not the solver tolerance: at the
# production 1e-6 two correct solvers already differ by up to cond(H) * 1e-6 in dx.
# Every Newton iterate is compared with the oracle's FP64 twin at a fixed 1e-5. Against the
# FP32 mirror the bar is 1e-5 widened to 1.5x the mirror-to-twin spread where the state itself
# is that sensitive to FP32 rounding (a stiff rod chain pressed into contact).
# One cause is allowed to break the FP64 bar, and only where it is shown present: the
# reference's Rayleigh stiffness (implicit.rs:755-795) takes its contact normal from the
# distance gradient at the stencil's first vertex and falls back to contact_kappa when that
# gradient vanishes, which for an edge-edge stencil happens when the first vertex's witness
# weight is exactly 0. FP32 arithmetic (the reference's, the mirror's and the device's,
# bitwise) rounds such a weight to 0 where FP64 keeps ~1e-9, so the FP32 and FP64 schemes give
# that stencil stiffnesses ~650x apart. Such a state records its FP64 gap as an expected
# failure with the measured figure, instead of widening the bar.
@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps", [(100, 0), (100, 12), (100, 30), (0, 12), (0, 20), (0, 33), (101, 10), (101, 25), ("collider", 14),
                                       ("springs", 14), ("pinned", 0)])
def test_one_step_parity(ctx, cfg, steps):
    if cfg == "pinned":  # cloth vertices pinned in contact: elements join pinned and free vertices
        s, x, v, _ = pinned_state()
        s.desc.params.cg_tol = 1e-10
        o, o64 = O.OracleSolver(s), O.OracleSolver(s, fp64=True)
        o.set_state(x, v)
    else:
        s, o = oracle_state(cfg, steps, cg_tol=1e-10)
        x, v = o.positions(), o.velocities()
        o64 = O.OracleSolver(scene(cfg, cg_tol=1e-10), fp64=True)
    o64.set_state(x, v)
    st = nguo_hip.Stepper(ctx, s, trace=True)
    st.set_state(x, v)
    o.step(1)
    o64.step(1)
    st.advance(1)
    ox, otoi = o.newton_trace()
    dx, _ = o64.newton_trace()
    gx, gtoi = st.newton_trace()
    assert len(gx) == len(ox) == len(dx), (len(gx), len(ox), len(dx))
    worst64 = 0.0
    for k in range(len(ox)):
        e32, e64, spread = rel_err(gx[k], ox[k]), rel_err(gx[k], dx[k]), rel_err(ox[k], dx[k])
        print(f"newton {k}: |gpu-oracle32| {e32:.2e} |gpu-oracle64| {e64:.2e} |oracle32-oracle64| {spread:.2e}")
        assert e32 <= max(POS_TOL, 1.5 * spread), (k, e32, e64, spread)
        worst64 = max(worst64, e64)
    m = st.metrics()
    # the pierce metric agrees with the oracle's own test on the same positions (the
    # 101 tube's seam rod reaches a piercing state under the reference scheme too); without
    # self contact the metric is the analytic colliders' side test, compared with the mirror's
    if s.params.self_collision:
        want = int(not o.self_intersects(st.positions()))
    else:
        want = int(o.metrics()["penetration_free"])
    assert m.all_finite == 1 and m.penetration_free == want
    assert m.max_reresid <= 1e-9
    st.close()
    if worst64 > POS_TOL:
        inv_h2 = 1.0 / (s.params.dt * s.params.dt)
        n32, n64 = len(o.stencils(x)[1]), len(o64.stencils(x)[1])
        k32 = o.kappas(x, inv_h2)[:n32]
        k64 = o64.kappas(x, inv_h2)[:n64]
        fb = np.float32(s.params.contact_kappa)
        split = int(np.sum((k32 == fb) != (k64 == fb))) if n32 == n64 else -1
        assert split > 0, f"FP64 gap {worst64:.2e} without a stiffness branch split"
        pytest.xfail(f"FP64 twin {worst64:.2e} > {POS_TOL:.0e}: {split} stencil(s) take the other Rayleigh branch in FP64")


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps", [(0, 33), (101, 25)])
def test_one_step_parity_production_tolerance(ctx, cfg, steps):
    # at the workload's 1e-6: L2-relative position agreement after the step
    s, o = oracle_state(cfg, steps)
    x, v = o.positions(), o.velocities()
    o64 = O.OracleSolver(nguo_hip.Scene(cfg, seed=cfg), fp64=True)
    o64.set_state(x, v)
    st = nguo_hip.Stepper(ctx, s)
    st.set_state(x, v)
    o.step(1)
    o64.step(1)
    st.advance(1)
    a, b, c = (p.astype(np.float64) for p in (st.positions(), o.positions(), o64.positions()))
    l2 = lambda u, w: np.linalg.norm(u - w) / np.linalg.norm(w)
    print(f"L2 |gpu-oracle32| {l2(a, b):.2e} |gpu-oracle64| {l2(a, c):.2e} |oracle32-oracle64| {l2(b, c):.2e}")
    assert l2(a, b) <= max(POS_TOL, 1.5 * l2(b, c))
    st.close()


@pytest.mark.gpu
def test_sheet_drape_invariants_36_steps(ctx):
    s, st = make(ctx, 0)
    steps = []
    for k in range(36):
        st.advance(1)
        m = st.metrics()
        steps.append(m)
        assert m.all_finite == 1 and m.penetration_free == 1, k
        assert m.max_reresid <= 1e-6
        # max_sigma over the mesh edges, the sphere's included (their ratio stays 1)
        assert m.max_sigma == max_stretch_ratio(st.positions(), s.positions, s.tris), (k, m.max_sigma)
    # the reference's acceptance gate (metrics.rs:84-183) over the device metrics, including
    # no device allocation once the capacities have settled
    ok, checks = nguo_hip.check_invariants(steps, s.params.cg_tol, s.params.cg_max_iter)
    assert ok, checks
    assert sum(m.device_alloc for m in steps) > 0  # the first steps do allocate
    x = st.positions()
    free = s.inv_mass > 0
    assert x[free, 1].min() < 0.0  # draped over the sphere
    assert np.all(st.collider_sides(x) == 1)
    st.close()


@pytest.mark.gpu
def test_tube_garment_step_parity_from_device_state(ctx):
    s = nguo_hip.Scene(1, seed=1)
    s.desc.params.cg_tol = 1e-10
    st = nguo_hip.Stepper(ctx, s)
    st.advance(6)
    x, v = st.positions(), st.velocities()
    o = O.OracleSolver(s)
    o.set_state(x, v)
    o.step(1)
    st.advance(1)
    assert rel_err(st.positions(), o.positions()) <= POS_TOL
    m = st.metrics()
    assert m.all_finite == 1 and m.penetration_free == 1 and m.max_reresid <= 1e-9
    st.close()


@pytest.mark.gpu
@pytest.mark.parametrize("batched,nscenes", [(False, 3), (True, 3), (True, 10)])
def test_batch_matches_independent_steppers(ctx, batched, nscenes):
    # graded sizes advanced concurrently (nguo_batch_*), their Newton solves each on its own
    # stream or as one batched PCG (10 instances: two batched launches of at most 8): every
    # instance takes exactly the steps a lone nguo_step takes, bit for bit
    scenes = [nguo_hip.Scene(100, seed=0), nguo_hip.Scene(101, seed=101)] + [nguo_hip.Scene(100, seed=5 + k) for k in range(nscenes - 2)]
    b = nguo_hip.Batch(ctx, scenes)
    b.set_solve(batched)
    b.advance(4)
    for sc, bst in zip(scenes, b.steppers):
        st = nguo_hip.Stepper(ctx, sc)
        st.advance(4)
        assert np.array_equal(bst.positions(), st.positions())
        assert np.array_equal(bst.velocities(), st.velocities())
        mb, ms = bst.metrics(), st.metrics()
        assert mb.newton_iters == ms.newton_iters and mb.max_cg_iters == ms.max_cg_iters
        st.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps", [(0, 20), (100, 30)])
def test_fp64_pcg_step_parity(ctx, cfg, steps):
    # the FP64 PCG (nguo_step_set_precision) on sheet states in contact: every Newton iterate
    # against the FP32 mirror and the FP64 twin at 1e-10, as the default solver is held
    s, o = oracle_state(cfg, steps, cg_tol=1e-10)
    x, v = o.positions(), o.velocities()
    o64 = O.OracleSolver(scene(cfg, cg_tol=1e-10), fp64=True)
    o64.set_state(x, v)
    st = nguo_hip.Stepper(ctx, s, trace=True)
    st.set_precision(nguo_hip.Stepper.PCG_FP64)
    st.set_state(x, v)
    o.step(1)
    o64.step(1)
    st.advance(1)
    ox, _ = o.newton_trace()
    dx, _ = o64.newton_trace()
    gx, _ = st.newton_trace()
    assert len(gx) == len(ox) == len(dx)
    for k in range(len(ox)):
        e32, e64, spread = rel_err(gx[k], ox[k]), rel_err(gx[k], dx[k]), rel_err(ox[k], dx[k])
        print(f"newton {k}: |gpu-oracle32| {e32:.2e} |gpu-oracle64| {e64:.2e} |oracle32-oracle64| {spread:.2e}")
        assert e64 <= POS_TOL and e32 <= max(POS_TOL, 1.5 * spread), (k, e32, e64, spread)
    m = st.metrics()
    assert m.all_finite == 1 and m.max_reresid <= 1e-9 and m.num_stencils > 0
    st.close()
