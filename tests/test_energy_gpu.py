"""Step energy on the device (nguo_step_energy) against the oracle after every Newton iterate
of one step taken from identical inputs (engine-core/src/implicit.rs:1951-2191), at 1e-10 PCG
residual so the comparison measures the implementation, not the solver tolerance.

Two figures per iterate, both printed:
  * the evaluator: the device energy at the device's iterate against the oracle's FP64 energy
    at the same positions and the same (FP32) inertial target (EVAL_TOL; both evaluate the
    FP64 scheme at FP32 positions);
  * the trajectory: the device's iterate against the FP32 mirror's iterate (the reference's
    arithmetic) and the FP64 twin's, under one energy: the FP64 step energy towards the FP64
    inertial target with the contact quantities (stencils, Rayleigh stiffnesses, friction
    normal force, normals, weights) held at the oracle's iterate, whose gradient there is the
    assembled gradient g. The energies agree within 1e-5 relative beyond twice the
    first-order change |g . dx| that the measured position difference dx makes (the positions
    themselves are held to 1e-5 by the step parity test). An iterate that is not yet the
    minimum has energy that moves with g . dx: the reference's own FP32 iterates sit up to
    5e-2 from the FP64 twin's in energy on a stiff rod chain in contact. The step energy with
    each iterate's own contact quantities is printed too ("own"): its stiffnesses carry m/d^2,
    so it moves by 2 dd/d of the barrier energy on top."""
import numpy as np
import pytest

from conftest import nguo_hip
import oracle_py as O
from cases import scene

EVAL_TOL = 1e-7  # same positions and target: measured <= 1e-8 per term
E_TOL = 1e-5
CASES = [(100, 12), (100, 30), (0, 12), (0, 20), (101, 10), (101, 25), ("collider", 14), ("springs", 14)]


def rel(a, b):
    return abs(a - b) / abs(b)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps", CASES)
def test_step_energy_per_newton_iterate(ctx, cfg, steps):
    o32 = O.OracleSolver(scene(cfg))
    o32.step(steps)
    x, v = o32.positions(), o32.velocities()
    mir = O.OracleSolver(scene(cfg, cg_tol=1e-10))
    mir.set_state(x, v)
    xt32 = mir.x_tilde().astype(np.float64)  # the inertial target the FP32 scheme (and the device) steps to
    mir.step(1)
    mx, _ = mir.newton_trace()
    twin = O.OracleSolver(scene(cfg, cg_tol=1e-10), fp64=True)
    twin.set_state(x, v)
    xt = twin.x_tilde_d()
    twin.step(1)
    tx, _ = twin.newton_trace()
    twin.set_state(x, v)  # the friction reference of the step is its start
    s = scene(cfg, cg_tol=1e-10)
    st = nguo_hip.Stepper(ctx, s, trace=True)
    st.set_state(x, v)
    st.advance(1)
    gx, _ = st.newton_trace()
    inv_h2 = 1.0 / float(s.params.dt) ** 2
    assert len(gx) == len(mx) == len(tx)
    for k in range(len(gx)):
        ed, td = st.energy(gx[k])
        es, ts = twin.energy(gx[k], xt32, inv_h2)
        eg, _ = twin.energy(gx[k], xt, inv_h2)
        terms = " ".join(f"{n}={abs(a - b) / max(abs(b), 1e-300):.1e}" for n, a, b in zip(st.ENERGY_TERMS, td, ts) if b != 0.0)
        line = f"newton {k}: E {ed:.9e} | evaluator {rel(ed, es):.2e} ({terms})"
        assert rel(ed, es) <= EVAL_TOL, (k, rel(ed, es))
        for name, xr in (("mirror", mx[k]), ("FP64 twin", tx[k])):
            own, _ = twin.energy(xr, xt, inv_h2)
            twin.freeze(xr, xt, inv_h2)
            er, _ = twin.energy(xr, xt, inv_h2, frozen=True)
            eh, _ = twin.energy(gx[k], xt, inv_h2, frozen=True)
            dx = gx[k].astype(np.float64) - xr.astype(np.float64)
            g1 = abs(float((twin.model_gradient(xr, xt, inv_h2) * dx).sum())) / abs(er)
            line += f" | vs {name} iterate {rel(eh, er):.2e} (first order {g1:.2e}, own {rel(eg, own):.2e})"
            assert rel(eh, er) <= E_TOL + 2.0 * g1, (k, name, rel(eh, er), g1)
        print(line)
    st.close()
