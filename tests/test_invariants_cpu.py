"""The invariant checker of the host mirror (engine-core/src/metrics.rs:84-183
check_invariants) on hand-made step metrics, with the reference's own unit cases
(metrics.rs:186-256): a good run passes, a diverged run fails exactly the four invariants it
breaks, the strain-limit gate with its 2% slack; plus steady-state device allocations."""
from conftest import nguo_hip


def good_step(t):
    m = nguo_hip.StepMetrics()
    m.time = t
    m.max_reresid = 1e-4
    m.max_cg_iters = 50
    m.newton_iters = 4
    m.max_sigma = 1.001
    m.max_velocity = 0.5
    m.num_contacts = 3
    m.penetration_free = 1
    m.all_finite = 1
    m.device_alloc = 0
    m.device_free = 0
    return m


def by(checks, name):
    return next(c for c in checks if c[0] == name)[1]


def test_good_run_passes_all_invariants():
    ok, checks = nguo_hip.check_invariants([good_step(k) for k in range(10)], 1e-3, 10000, 0.0)
    assert ok, checks
    assert [c[0] for c in checks] == ["penetration_free", "all_frames_finite", "strain_limit_respected", "cg_converged",
                                      "cg_iter_under_cap", "zero_steady_state_alloc"]


def test_diverged_run_fails():
    steps = []
    for k in range(10):
        s = good_step(k)
        if k >= 5:
            s.all_finite = 0
            s.penetration_free = 0
            s.max_reresid = 0.9
            s.max_cg_iters = 100
        steps.append(s)
    ok, checks = nguo_hip.check_invariants(steps, 1e-3, 100, 0.0)
    assert not ok
    for name in ("all_frames_finite", "penetration_free", "cg_converged", "cg_iter_under_cap"):
        assert not by(checks, name), name
    assert by(checks, "strain_limit_respected") and by(checks, "zero_steady_state_alloc")


def test_strain_limit_gate():
    s = good_step(0.0)
    s.max_sigma = 1.10  # ceiling 1.05 * 1.02 = 1.071
    assert not nguo_hip.check_invariants([s], 1e-3, 10000, 0.05)[0]
    s.max_sigma = 1.06
    assert nguo_hip.check_invariants([s], 1e-3, 10000, 0.05)[0]


def test_steady_state_allocation_fails_only_in_the_second_half():
    steps = [good_step(k) for k in range(10)]
    steps[0].device_alloc = 40  # warm-up allocations are allowed
    assert nguo_hip.check_invariants(steps, 1e-3, 10000)[0]
    steps[7].device_alloc = 1
    ok, checks = nguo_hip.check_invariants(steps, 1e-3, 10000)
    assert not ok and not by(checks, "zero_steady_state_alloc")
