"""The kinematic static collider on the device against the oracle: the sheet resting on the
icosphere while the body is moved between steps (nguo_step_set_kinematic_collider_positions,
engine-core/src/implicit.rs:1324-1398), each step taken from the oracle's state and compared
after the same move."""
import numpy as np
import pytest

from conftest import nguo_hip
import oracle_py as O
from cases import scene

POS_TOL = 1e-5


@pytest.mark.gpu
def test_moving_body_steps_match_the_oracle(ctx):
    s = scene(100, cg_tol=1e-10)
    o = O.OracleSolver(s)
    o.step(30)  # the sheet at rest on the sphere, contacts active
    x, v = o.positions(), o.velocities()
    st = nguo_hip.Stepper(ctx, s)
    st.set_state(x, v)
    body = s.static.astype(bool)
    # 0.1 mm per step, well inside the barrier's reach at the sheet's 0.3-0.5 mm gaps. A
    # teleport of 0.6 mm closes some gap to ~0 (the reference's documented ceiling): the
    # step's CCD fraction is then 1e-11 in the FP32 oracle and 0.06 in its FP64 twin, and no
    # two implementations of the scheme agree there
    move = np.array([1e-4, -5e-5, 0.0], np.float32)
    worst = 0.0
    for k in range(6):
        # each step from the oracle's state (trajectories are compared a step at a time)
        st.set_state(o.positions(), o.velocities())
        xb = o.positions()
        xb[body] += move
        assert o.set_kinematic(xb)
        st.set_kinematic(xb)
        o.step(1)
        st.advance(1)
        gx, ox = st.positions(), o.positions()
        assert np.array_equal(gx[body], ox[body])  # the body sits where it was put
        err = np.abs(gx - ox).max() / max(1.0, np.abs(ox).max())
        worst = max(worst, err)
        print(f"step {k}: |gpu - oracle| {err:.2e}, stencils {st.metrics().num_stencils}")
        assert err <= POS_TOL
    assert st.metrics().num_stencils > 0
    # the sheet was dragged along by friction with the moving body
    free = s.inv_mass > 0
    assert (gx[free, 0] - x[free, 0]).mean() > 0
    # bad input changes nothing on the device either
    bad = gx.copy()
    bad[np.nonzero(body)[0][0], 0] = np.nan
    with pytest.raises(nguo_hip.NguoError):
        st.set_kinematic(bad)
    assert np.array_equal(st.positions(), gx)
    st.close()
