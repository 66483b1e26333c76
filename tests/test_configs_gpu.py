"""The device on the larger generated workloads against the oracle: the tangled concentric
tubes at their start state, each of the eight graded tube garments one step from a drape
state, and the graded batch against lone steppers."""
import numpy as np
import pytest

from conftest import nguo_hip
import oracle_py as O
from cases import check_stencils

POS_TOL = 1e-5  # |x_gpu - x_oracle|_inf / max(1, |x|_inf)
GRADED = [0.85, 0.90, 0.95, 1.00, 1.05, 1.10, 1.15, 1.20]  # circumference factors of the graded tubes


def rel_err(a, b):
    return np.abs(a - b).max() / max(1.0, np.abs(b).max())


@pytest.mark.gpu
def test_tangled_tubes_start_state(ctx):
    # Two concentric tubes, the inner one interpenetrating the outer over ~5 % of it. The
    # reference's implicit step assumes an intersection-free start (implicit.rs:612-617): edge
    # pairs that cross at t = 0 sit at gap ~0, so every search direction that moves them meets
    # an ACCD time of impact ~1e-15 (contact.rs:378-431) and the step cannot advance. Both
    # sides must agree on the contact set at the start and both must stall on it.
    s = nguo_hip.Scene(2, seed=2)
    o = O.OracleSolver(s)
    x, v = o.positions(), o.velocities()
    assert o.self_intersects(x)
    idx, d = check_stencils(ctx, s, o, x, v)
    assert len(idx) > 10000 and (d < 1e-6).sum() > 1000, "crossing pairs sit at gap ~0"
    x1, t, _, _ = o.newton_iteration(x, o.x_tilde())
    st = nguo_hip.Stepper(ctx, s, trace=True)
    st.advance(1)
    xs, tois = st.newton_trace()
    print(f"first Newton iterate: oracle toi {t:.3e}, device toi {tois[0]:.3e}, "
          f"|x1 - x0| oracle {np.abs(x1 - x).max():.3e}, device {np.abs(xs[0] - x).max():.3e}")
    assert t < 1e-6 and tois[0] < 1e-6
    assert rel_err(xs[0], x1) <= POS_TOL
    m = st.metrics()
    assert m.all_finite == 1 and m.penetration_free == 0
    st.close()


# Each graded size one step from its own device drape state (6 steps in, in contact with the
# body), every Newton system solved to 1e-10 on all three sides. Every Newton iterate is
# compared with the FP32 mirror (the reference's own arithmetic) at 1e-5, widened to 1.5x the
# mirror-to-twin spread where the state is that sensitive to rounding, and with the FP64 twin
# at a fixed 1e-5. One cause may break a bar, and only where it is shown present: the state
# amplifies rounding-level differences beyond 1e-5 at or before the first iterate over it,
# measured as the mirror-to-twin spread or as the spread between the mirror and a second FP32
# mirror started from the same state moved by one unit in the last place (random sign per
# coordinate). On these
# seam-rod states that happens through the CCD fraction: it is discontinuous in the search
# direction (ACCD stops a pair at a fixed fraction of its gap, contact.rs:378-431, and the
# smallest fraction over the pairs switches pairs), so two schemes that differ by rounding take
# fractions 1e-4 to 1e-1 apart and the iterates part by that much times the step. Such a state
# records its FP64 gap as an expected failure with the measured figures.
@pytest.mark.gpu
@pytest.mark.parametrize("f", GRADED)
def test_graded_tube_step_parity(ctx, f):
    s = nguo_hip.Scene(3, seed=3, param=f)
    s.desc.params.cg_tol = 1e-10
    st = nguo_hip.Stepper(ctx, s, trace=True)
    st.advance(6)
    x, v = st.positions(), st.velocities()
    o, o64, op = O.OracleSolver(s), O.OracleSolver(s, fp64=True), O.OracleSolver(s)
    for oo in (o, o64):
        oo.set_state(x, v)
        oo.step(1)
    # the same state moved by one unit in the last place, sign per coordinate (the pinned body
    # keeps its positions: its trees are built from them)
    sign = np.random.default_rng(int(100 * f)).integers(0, 2, x.shape).astype(bool)
    xp = np.where(sign, np.nextafter(x, np.float32(np.inf)), np.nextafter(x, np.float32(-np.inf))).astype(np.float32)
    xp[s.inv_mass == 0] = x[s.inv_mass == 0]
    op.set_state(xp, v)
    op.step(1)
    st.advance(1)
    ox, otoi = o.newton_trace()
    dx, dtoi = o64.newton_trace()
    px, _ = op.newton_trace()
    gx, gtoi = st.newton_trace()
    m = st.metrics()
    print(f"x{f}: {s.n_verts} vertices, {m.num_stencils} stencils, CCD fractions device {np.round(gtoi, 4).tolist()} "
          f"mirror {np.round(otoi, 4).tolist()} twin {np.round(dtoi, 4).tolist()}")
    assert m.num_stencils > 0, "the state must be in contact"
    assert len(gx) == len(ox) == len(dx), (len(gx), len(ox), len(dx))
    first_over, sens = None, []
    for k in range(len(ox)):
        e32, e64, spread = rel_err(gx[k], ox[k]), rel_err(gx[k], dx[k]), rel_err(ox[k], dx[k])
        ulp = rel_err(px[k], ox[k]) if k < len(px) else np.inf
        sens.append(max(spread, ulp))
        print(f"  newton {k}: |gpu-oracle32| {e32:.2e} |gpu-oracle64| {e64:.2e} |oracle32-oracle64| {spread:.2e} "
              f"|oracle32(x+ulp)-oracle32| {ulp:.2e}")
        if (e32 > max(POS_TOL, 1.5 * spread) or e64 > POS_TOL) and first_over is None:
            first_over = (k, e32, e64)
    assert m.all_finite == 1 and m.penetration_free == 1 and m.max_reresid <= 1e-9
    st.close()
    if first_over is not None:
        k, e32, e64 = first_over
        over = [j for j in range(k + 1) if sens[j] > POS_TOL]
        assert over, (f"Newton iteration {k}: {e32:.2e} from the FP32 mirror, {e64:.2e} from the FP64 twin, while the state's "
                      f"rounding sensitivity stays within {POS_TOL:.0e}")
        j = over[0]
        pytest.xfail(f"Newton iteration {k}: {e32:.2e} from the FP32 mirror, {e64:.2e} from the FP64 twin; the state's rounding "
                     f"sensitivity is {sens[j]:.2e} at iteration {j} (CCD fractions mirror {otoi[j]:.4f}, twin {dtoi[j]:.4f})")


@pytest.mark.gpu
@pytest.mark.parametrize("batched", [False, True])
def test_graded_batch_matches_lone_steppers(ctx, batched):
    # the eight graded tubes advanced together (each Newton solve on its own stream, or one
    # lockstep batched PCG over all eight): every instance bit for bit a lone stepper's steps
    scenes = [nguo_hip.Scene(3, seed=3, param=f) for f in GRADED]
    b = nguo_hip.Batch(ctx, scenes)
    b.set_solve(batched)
    b.advance(6)
    for sc, bst in zip(scenes, b.steppers):
        st = nguo_hip.Stepper(ctx, sc)
        st.advance(6)
        assert np.array_equal(bst.positions(), st.positions())
        assert np.array_equal(bst.velocities(), st.velocities())
        mb, ms = bst.metrics(), st.metrics()
        assert mb.newton_iters == ms.newton_iters and mb.sum_cg_iters == ms.sum_cg_iters
        assert ms.num_stencils > 0
        st.close()
    b.close()
