"""The kinematic static collider in the oracle (engine-core/src/implicit.rs:1324-1398
set_kinematic_collider_positions, :430-444 kinematic_dx, :2065-2070 its clearing): the body
rows teleport with their velocity, the move enters the friction slip of the next step only,
and bad input is rejected with nothing changed. CPU only."""
import os

import numpy as np

import oracle_py as O
import scene_json as S
from cases import scene


def rested(mu=0.3, steps=30):
    s = scene(100, friction=mu)
    o = O.OracleSolver(s)
    o.step(steps)
    return s, o


def test_teleport_sets_rows_and_velocities():
    s, o = rested()
    st = s.static.astype(bool)
    x0, v0 = o.positions(), o.velocities()
    x = x0.copy()
    x[st] += np.array([1e-3, -5e-4, 2e-4], np.float32)
    x[~st] += 1.0  # other rows are not read
    assert o.set_kinematic(x)
    x1, v1 = o.positions(), o.velocities()
    assert np.array_equal(x1[st], x[st]) and np.array_equal(x1[~st], x0[~st])
    h = np.float32(s.params.dt)
    assert np.allclose(v1[st], (x[st] - x0[st]) / h, rtol=1e-6) and np.array_equal(v1[~st], v0[~st])
    o.step(1)
    assert np.array_equal(o.velocities()[st], np.zeros_like(v1[st]))  # the step re-zeroes pinned rows


def test_move_is_the_friction_slip_of_the_next_step_only():
    # at the step's start (x = the state) a contact's slip is only the body's move, so the
    # friction potential there is positive after a tangential move and zero without one
    s, o = rested()
    st = s.static.astype(bool)
    inv_h2 = 1.0 / float(s.params.dt) ** 2
    x = o.positions_d()
    e0, t0 = o.energy(x, o.x_tilde_d(), inv_h2)
    assert t0[O.OracleSolver.TERMS.index("friction")] == 0.0 and t0[O.OracleSolver.TERMS.index("barriers")] > 0.0
    xm = o.positions()
    xm[st, 0] += np.float32(1e-3)
    assert o.set_kinematic(xm)
    x = o.positions_d()
    e1, t1 = o.energy(x, o.x_tilde_d(), inv_h2)
    assert t1[O.OracleSolver.TERMS.index("friction")] > 0.0
    o.step(1)  # consumed: the next step's start sees a stationary body again
    x = o.positions_d()
    _, t2 = o.energy(x, o.x_tilde_d(), inv_h2)
    assert t2[O.OracleSolver.TERMS.index("friction")] == 0.0


def test_friction_drags_the_sheet_with_the_body():
    drift = {}
    for mu in (0.3, 0.0):
        s, o = rested(mu)
        st, free = s.static.astype(bool), s.inv_mass > 0
        xs = o.positions()
        for _ in range(10):
            x = o.positions()
            x[st, 0] += np.float32(1e-3)
            assert o.set_kinematic(x)
            o.step(1)
        drift[mu] = float((o.positions()[free, 0] - xs[free, 0]).mean())
    assert drift[0.3] > 2.0 * drift[0.0] > 0.0, drift


def test_bad_input_is_rejected_and_changes_nothing():
    s, o = rested(steps=2)
    st = np.nonzero(s.static)[0]
    x0 = o.positions()
    bad = x0.copy()
    bad[st[5], 1] = np.nan
    assert not o.set_kinematic(bad)
    assert np.array_equal(o.positions(), x0)
    # a scene without a static collider
    hs, _ = S.lower(os.path.join(os.path.dirname(__file__), "golden", "scenes", "hanging_cloth.json"))
    oh = O.OracleSolver(hs)
    assert not oh.set_kinematic(oh.positions())
