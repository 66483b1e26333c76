"""LinearSolver seam on the device (nguo_ls_*) against the oracle's f64 block-Jacobi CG
(engine-core/src/linsolve.rs:350-411) and f32 CSR SpMV (csr.rs:98-110)."""
import numpy as np
import pytest

from conftest import nguo_hip
import oracle_py as O


def laplacian_1d(n):
    # reference manufactured system: tridiagonal 2/-1 (linsolve.rs:855-867), in 3 dof blocks
    rp, ci, va = [0], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                ci.append(j)
                va.append(v)
        rp.append(len(ci))
    return np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, np.float32)


def newton_system(cfg, steps_before=0):
    s = nguo_hip.Scene(cfg, seed=cfg)
    o = O.OracleSolver(s)
    if steps_before:
        o.step(steps_before)
    x = o.positions()
    xt = o.x_tilde()
    dt = s.params.dt
    return o.assemble(x, xt, 1.0 / (dt * dt))


def host_residual(rp, ci, va, rhs, x):
    r = rhs.astype(np.float64).copy()
    for row in range(len(rhs)):
        sl = slice(rp[row], rp[row + 1])
        r[row] -= np.dot(va[sl].astype(np.float64), x[ci[sl]].astype(np.float64))
    return np.linalg.norm(r) / np.linalg.norm(rhs.astype(np.float64))


@pytest.mark.gpu
def test_spmv_matches_oracle(ctx):
    rp, ci, va, rhs = newton_system(0)
    ls = nguo_hip.LinearSolver(ctx)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(len(rhs)).astype(np.float32)
    y = ls.spmv(rp, ci, va, x)
    y_ref = O.spmv(rp, ci, va, x)
    y64 = np.zeros(len(rhs))
    for r in range(len(rhs)):
        y64[r] = np.dot(va[rp[r]:rp[r + 1]].astype(np.float64), x[ci[rp[r]:rp[r + 1]]].astype(np.float64))
    scale = np.abs(va).max() * np.abs(x).max() * 16
    assert np.abs(y - y64).max() <= 1e-6 * scale
    assert np.abs(y - y_ref).max() <= 1e-6 * scale
    ls.close()


def hub_laplacian(n, stride=5, w=0.05):
    # a 1-D Laplacian plus dof 0 coupled to every stride-th dof (weight -w, diagonally
    # dominant): the first block row is far longer than the others, so the slices of the
    # block matrix keep widths of their own instead of one common width
    rows = [dict() for _ in range(n)]
    for i in range(n):
        rows[i][i] = 2.0
        if i > 0:
            rows[i][i - 1] = -1.0
        if i + 1 < n:
            rows[i][i + 1] = -1.0
    for j in range(stride, n, stride):
        rows[0][j] = rows[0].get(j, 0.0) - w
        rows[j][0] = rows[j].get(0, 0.0) - w
        rows[j][j] += w
        rows[0][0] += w
    rp, ci, va = [0], [], []
    for r in rows:
        for j in sorted(r):
            ci.append(j)
            va.append(r[j])
        rp.append(len(ci))
    return np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("stride", [5, 1000])
def test_spmv_and_pcg_irregular_rows(ctx, stride):
    # stride 5: one block row of ~180 blocks among rows of 2-3 (per-slice widths); stride
    # 1000: no hub coupling, every slice the same width (one common width)
    n = 3 * 300
    rp, ci, va = hub_laplacian(n, stride)
    ls = nguo_hip.LinearSolver(ctx)
    rng = np.random.default_rng(23)
    x = rng.standard_normal(n).astype(np.float32)
    y = ls.spmv(rp, ci, va, x)
    y64 = np.array([np.dot(va[rp[r]:rp[r + 1]].astype(np.float64), x[ci[rp[r]:rp[r + 1]]].astype(np.float64)) for r in range(n)])
    assert np.abs(y - y64).max() <= 1e-6 * np.abs(va).max() * np.abs(x).max() * 16
    xs = np.linspace(-1, 1, n).astype(np.float32)
    b = O.spmv(rp, ci, va, xs)
    xsol, st = ls.solve(rp, ci, va, b, 1e-6, 100000)
    assert st.residual <= 1e-6
    assert host_residual(rp, ci, va, b, xsol) <= 2e-6
    ls.close()


@pytest.mark.gpu
def test_pcg_manufactured_laplacian(ctx):
    n = 3 * 200
    rp, ci, va = laplacian_1d(n)
    xs = np.linspace(-1, 1, n).astype(np.float32)
    b = O.spmv(rp, ci, va, xs)
    ls = nguo_hip.LinearSolver(ctx)
    x, st = ls.solve(rp, ci, va, b, 1e-6, 100000)
    assert st.residual <= 1e-6
    assert np.abs(x - xs).max() < 1e-2
    x0, st0 = ls.solve(rp, ci, va, np.zeros(n, np.float32), 1e-6, 1000)
    assert np.all(x0 == 0)  # zero rhs -> zero solution (engine-gpu/src/pcg.rs:838-845)
    ls.close()


@pytest.mark.gpu
def test_pcg_stops_at_the_iteration_cap(ctx):
    # the 1-D Laplacian of 600 unknowns needs far more than these iterations: each solve ends
    # on its cap, and the iteration count the solver reports is exactly the cap. Caps inside
    # a queued chunk, at a chunk's start and at a chunk boundary (host convergence checks
    # every 8 iterations at first). PCG minimises the A-norm of the error over a growing
    # Krylov space, so that norm falls with every cap.
    n = 3 * 200
    rp, ci, va = laplacian_1d(n)
    xs = np.linspace(-1, 1, n).astype(np.float32)
    b = O.spmv(rp, ci, va, xs)
    ls = nguo_hip.LinearSolver(ctx)

    def a_norm_err(x):
        e = x.astype(np.float64) - xs.astype(np.float64)
        ae = np.array([np.dot(va[rp[r]:rp[r + 1]].astype(np.float64), e[ci[rp[r]:rp[r + 1]]]) for r in range(n)])
        return float(np.sqrt(max(np.dot(e, ae), 0.0)))

    errs = [a_norm_err(np.zeros(n, np.float32))]
    for cap in (5, 8, 13, 16):
        x, st = ls.solve(rp, ci, va, b, 1e-9, cap)
        assert st.iters == cap, (cap, st.iters)
        assert np.all(np.isfinite(x))
        assert host_residual(rp, ci, va, b, x) > 1e-9  # stopped by the cap, not converged
        errs.append(a_norm_err(x))
    assert all(b2 < a for a, b2 in zip(errs, errs[1:])), errs
    ls.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps", [(0, 0), (0, 14)])
def test_pcg_matches_oracle_cg(ctx, cfg, steps):
    rp, ci, va, rhs = newton_system(cfg, steps)
    x_ref, it_ref, res_ref = O.cg(rp, ci, va, rhs, 1e-6, 10000)
    ls = nguo_hip.LinearSolver(ctx)
    x, st = ls.solve(rp, ci, va, rhs, 1e-6, 10000)
    # FP64 residual of the device solution, evaluated on the host
    n = len(rhs)
    r = rhs.astype(np.float64).copy()
    for row in range(n):
        sl = slice(rp[row], rp[row + 1])
        r[row] -= np.dot(va[sl].astype(np.float64), x[ci[sl]].astype(np.float64))
    rel = np.linalg.norm(r) / np.linalg.norm(rhs.astype(np.float64))
    assert rel <= 2e-6, (rel, st.residual)
    assert st.residual <= 1e-6
    scale = np.abs(x_ref).max()
    assert np.abs(x - x_ref).max() <= 1e-3 * scale, (np.abs(x - x_ref).max(), scale)
    ls.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps,domain", [(0, 0, 32), (0, 14, 32), (0, 14, 7), (100, 3, 1)])
def test_schwarz_apply_matches_oracle(ctx, cfg, steps, domain):
    # additive Schwarz apply z = sum_d R_d^T A_d^-1 R_d r (linsolve.rs:134-240): same BFS
    # domains, same floor and fallback; the device keeps an FP32 explicit inverse, so the
    # bar is a relative FP32 tolerance on z
    rp, ci, va, rhs = newton_system(cfg, steps)
    r = np.random.default_rng(21).standard_normal(len(rhs)).astype(np.float32)
    _, z_ref = O.schwarz(rp, ci, va, domain, r)
    ls = nguo_hip.LinearSolver(ctx, precond=1, domain=domain)
    z = ls.precond_apply(rp, ci, va, r)
    err = np.linalg.norm(z - z_ref) / np.linalg.norm(z_ref)
    assert err < 2e-4, err
    ls.close()


@pytest.mark.gpu
@pytest.mark.parametrize("domain", [32, 5])
def test_schwarz_apply_fallback_on_indefinite_domain(ctx, domain):
    # a domain whose block is not positive definite fails the Cholesky pivot test
    # (pivot <= 1e-14, linsolve.rs:80-99) and falls back to the diagonal with the floor
    # 1e-6 tr/m + 1e-8 added: 1 / a_pp, or 1 where a_pp <= 1e-14 (linsolve.rs:232-250). Two
    # vertices are decoupled from their neighbours and given a negative diagonal, so their
    # pivots fail whatever the elimination order; the other domains are unaffected.
    rp, ci, va, rhs = newton_system(0, 14)
    va = va.copy()
    bad = {100: -1.0e6, 700: -5.0}
    for row in range(len(rhs)):
        for k in range(rp[row], rp[row + 1]):
            vr, vc = row // 3, ci[k] // 3
            if vr in bad or vc in bad:
                va[k] = bad[vr] if (row == ci[k]) else 0.0
    r = np.random.default_rng(23).standard_normal(len(rhs)).astype(np.float32)
    doms, z_ref = O.schwarz(rp, ci, va, domain, r)
    for v0 in bad:
        dom = next(d for d in doms if v0 in d)
        rows = np.concatenate([np.arange(3 * u, 3 * u + 3) for u in dom])
        diag = np.array([float(va[k]) for row in rows for k in range(rp[row], rp[row + 1]) if ci[k] == row])
        floor = 1e-6 * (diag.sum() / len(rows)) + 1e-8
        dinv = np.where(diag + floor > 1e-14, 1.0 / (diag + floor), 1.0)
        # the oracle took the fallback on that domain
        assert np.abs(z_ref[rows] - r[rows] * dinv).max() <= 1e-9 * np.abs(z_ref[rows]).max()
    ls = nguo_hip.LinearSolver(ctx, precond=1, domain=domain)
    z = ls.precond_apply(rp, ci, va, r)
    err = np.linalg.norm(z - z_ref) / np.linalg.norm(z_ref)
    assert err < 2e-4, err
    for v0 in bad:
        dom = next(d for d in doms if v0 in d)
        rows = np.concatenate([np.arange(3 * u, 3 * u + 3) for u in dom])
        assert np.abs(z[rows] - z_ref[rows]).max() <= 1e-6 * np.abs(z_ref[rows]).max()
    ls.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps", [(0, 0), (0, 14), (101, 0)])
def test_schwarz_pcg_matches_oracle(ctx, cfg, steps):
    rp, ci, va, rhs = newton_system(cfg, steps)
    x_ref, it_ref, _ = O.cg_schwarz(rp, ci, va, rhs, 1e-6, 10000, 32)
    _, it_bj, _ = O.cg(rp, ci, va, rhs, 1e-6, 10000)
    ls = nguo_hip.LinearSolver(ctx, precond=1, domain=32)
    x, st = ls.solve(rp, ci, va, rhs, 1e-6, 10000)
    # the device iterates in FP64 and hands back FP32: the host FP64 residual of the returned
    # vector is bounded by the rounding floor of FP32 x, which the oracle's rounded
    # solution shows for the same system
    rel, rel_ref = host_residual(rp, ci, va, rhs, x), host_residual(rp, ci, va, rhs, x_ref)
    assert rel <= max(2e-6, 2.0 * rel_ref), (rel, rel_ref, st.residual)
    assert st.residual <= 1e-6
    assert np.abs(x - x_ref).max() <= 1e-3 * np.abs(x_ref).max()
    assert st.iters < it_bj, (st.iters, it_bj)
    assert st.iters <= 1.5 * it_ref + 10, (st.iters, it_ref)  # FP32 sweeps + FP64 refinement
    ls.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps,domain", [(0, 14, 32), (101, 0, 32), (101, 20, 16)])
def test_two_level_schwarz_apply_matches_oracle(ctx, cfg, steps, domain):
    # z = S0 r + w C^T S1 (C r) (linsolve.rs:522-657); coarse operator, coarse domains and
    # restriction as the oracle's Schwarz2, FP32 inverses and sums on the device
    rp, ci, va, rhs = newton_system(cfg, steps)
    r = np.random.default_rng(22).standard_normal(len(rhs)).astype(np.float32)
    _, z_ref = O.schwarz(rp, ci, va, domain, r, weight=0.7)
    _, z1 = O.schwarz(rp, ci, va, domain, r, weight=0.0)
    assert np.linalg.norm(z_ref - z1) > 1e-3 * np.linalg.norm(z_ref)  # the coarse term is not negligible here
    ls = nguo_hip.LinearSolver(ctx, precond=1, domain=domain, weight=0.7)
    z = ls.precond_apply(rp, ci, va, r)
    err = np.linalg.norm(z - z_ref) / np.linalg.norm(z_ref)
    assert err < 5e-4, err
    ls.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,steps", [(101, 0), (101, 20)])
def test_two_level_schwarz_pcg(ctx, cfg, steps):
    rp, ci, va, rhs = newton_system(cfg, steps)
    x_ref, it_ref, _ = O.cg_schwarz(rp, ci, va, rhs, 1e-6, 10000, 32, 0.7)
    _, it_one, _ = O.cg_schwarz(rp, ci, va, rhs, 1e-6, 10000, 32, 0.0)
    ls = nguo_hip.LinearSolver(ctx, precond=1, domain=32, weight=0.7)
    x, st = ls.solve(rp, ci, va, rhs, 1e-6, 10000)
    rel, rel_ref = host_residual(rp, ci, va, rhs, x), host_residual(rp, ci, va, rhs, x_ref)
    assert rel <= max(2e-6, 2.0 * rel_ref), (rel, rel_ref, st.residual)
    assert st.residual <= 1e-6
    assert np.abs(x - x_ref).max() <= 1e-3 * np.abs(x_ref).max()
    assert it_ref < it_one  # the oracle's two-level count beats one level on the tube
    assert st.iters <= 1.5 * it_ref + 10, (st.iters, it_ref)
    ls.close()
