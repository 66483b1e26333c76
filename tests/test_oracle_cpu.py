"""Pins the oracle (oracle/, the CPU restatement of the reference) before it is trusted:
the reference's own known answers (tests/golden/reference_known_answers.json), central
finite differences in every branch, closed forms, invariants and brute force on tiny inputs.
CPU only."""
import json
import os

import numpy as np
import pytest

from conftest import nguo_hip
import oracle_py as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))
lib = O.load()


def ptd(p12, fp64=1):
    return lib.oracle_point_tri_distance(fp64, O.dp(O.d64(p12)))


def ptd_grad(p12, fp64=1):
    g = np.zeros(12)
    d = lib.oracle_point_tri_distance_grad(fp64, O.dp(O.d64(p12)), O.dp(g))
    return d, g


def ssd(p12, fp64=1):
    return lib.oracle_segment_distance(fp64, O.dp(O.d64(p12)))


def ssd_grad(p12, fp64=1):
    g = np.zeros(12)
    d = lib.oracle_segment_distance_grad(fp64, O.dp(O.d64(p12)), O.dp(g))
    return d, g


def fd_grad(f, x, h=1e-6):
    x = np.asarray(x, dtype=np.float64)
    g = np.zeros_like(x)
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (f(xp) - f(xm)) / (2 * h)
    return g


# ------------------------------------------------------------------ known answers
@pytest.mark.parametrize("fp64", [0, 1])
def test_point_triangle_known_answers(fp64):
    kt = GOLD["point_triangle_distance"]
    tri = np.array(kt["triangle"], float).ravel()
    for c in kt["cases"]:
        p12 = np.concatenate([c["p"], tri])
        assert abs(ptd(p12, fp64) - c["d"]) < c["tol"]
        d, _ = ptd_grad(p12, fp64)
        assert abs(d - c["d"]) < c["tol"]


@pytest.mark.parametrize("fp64", [0, 1])
def test_segment_known_answers(fp64):
    for c in GOLD["segment_segment_distance"]["cases"]:
        p12 = np.array(c["seg"], float).ravel()
        assert abs(ssd(p12, fp64) - c["d"]) < c["tol"]
        d, _ = ssd_grad(p12, fp64)  # the witness-based distance agrees on these cases
        assert abs(d - c["d"]) < c["tol"]


@pytest.mark.parametrize("fp64", [0, 1])
def test_accd_known_answers(fp64):
    ka = GOLD["accd"]
    c = ka["edge_through_edge"]
    x, dx = np.array(c["x"], float).ravel(), np.array(c["dx"], float).ravel()
    toi = lib.oracle_accd_toi(fp64, 1, O.dp(x), O.dp(dx), 0.0, c["max_t"], c["reduction"], c["iters"])
    assert toi < c["toi_below"]
    m = x.copy()
    m[:6] += dx[:6] * toi
    assert ssd(m, fp64) > 0
    c = ka["point_into_triangle"]
    x, dx = np.array(c["x"], float).ravel(), np.array(c["dx"], float).ravel()
    toi = lib.oracle_accd_toi(fp64, 0, O.dp(x), O.dp(dx), 0.0, c["max_t"], c["reduction"], c["iters"])
    assert c["toi_range"][0] <= toi < c["toi_range"][1]
    m = x.copy()
    m[:3] += dx[:3] * toi
    assert ptd(m, fp64) > 0
    c = ka["no_relative_motion"]
    x, dx = np.array(c["x"], float).ravel(), np.array(c["dx"], float).ravel()
    assert lib.oracle_accd_toi(fp64, 0, O.dp(x), O.dp(dx), 0.0, c["max_t"], c["reduction"], c["iters"]) == c["toi_exact"]


@pytest.mark.parametrize("kind", [0, 1])
def test_accd_horizon_beyond_the_query_cap(kind):
    # The device caps each CCD query's search padding at 1.3 W, W = |dx| of the query plus
    # the largest |dx|: every point of one primitive moves relative to the other's by at most
    # W, so the distance stays above 0.0385 d0 up to t = 1.25 and the FP64 advancement reaches
    # its horizon within 68 iterations. Random pairs whose boxes are 1.3 W .. 2 W apart, with
    # the candidate deforming by up to twice the largest |dx|, all return the horizon with a
    # limit of 68; head-on pairs at 1.1 W do not, which shows the check can fail. Beyond
    # 1.3 x twice the largest |dx| the horizon comes at the first iterate (the cap used with an
    # iteration limit below 80).
    rng = np.random.default_rng(17 + kind)
    nq = 1 if kind == 0 else 2

    def pair(gap_factor, head_on, iters=68, wide=False):
        maxd = rng.uniform(0.01, 0.2)
        dq = rng.uniform(0.0, 1.0) * maxd
        cand = rng.uniform(0.0, 1.0, size=(4 - nq, 3)) * rng.uniform(0.05, 1.0)
        w = 2.0 * maxd if wide else dq + maxd
        q = rng.uniform(0.0, 1.0, size=(nq, 3)) * 0.2
        q[:, 0] += cand[:, 0].max() - q[:, 0].min() + gap_factor * w
        q[:, 1:] += cand[:, 1:].mean(axis=0) - q[:, 1:].mean(axis=0)
        unit = lambda n: (lambda v: v / np.linalg.norm(v, axis=1, keepdims=True))(rng.normal(size=(n, 3)))
        if head_on:
            dxq = np.tile([[-dq, 0.0, 0.0]], (nq, 1))
            dxc = np.tile([[maxd, 0.0, 0.0]], (4 - nq, 1))
        else:
            dxq = unit(nq) * dq * rng.uniform(0.0, 1.0, size=(nq, 1))
            dxc = unit(4 - nq) * maxd * rng.uniform(0.5, 1.0, size=(4 - nq, 1))
        x = np.concatenate([q, cand]).ravel()
        dx = np.concatenate([dxq, dxc]).ravel()
        return lib.oracle_accd_toi(1, kind, O.dp(x), O.dp(dx), 0.0, 1.25, 0.01, iters)

    for _ in range(400):
        assert pair(rng.uniform(1.3001, 2.0), False) == 1.25
    assert min(pair(1.1, True) for _ in range(20)) < 1.25
    # below the limit the cap is 1.3 x twice the largest |dx|: the horizon at the first iterate
    for _ in range(200):
        assert pair(rng.uniform(1.3001, 2.0), False, iters=1, wide=True) == 1.25


def test_lame_known_answers():
    import ctypes as C
    for c in GOLD["lame"]["cases"]:
        mu, lam = C.c_double(), C.c_double()
        lib.oracle_lame(c["young"], c["poisson"], C.byref(mu), C.byref(lam))
        assert abs(mu.value - c["mu"]) < c["tol_mu"]
        assert abs(lam.value - c["lambda"]) < c["tol_lambda"]


# ------------------------------------------------------------------ distances: FD in every branch
def voronoi_cases(rng):
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    pts = [(0.2, 0.2, 0.3), (-0.5, -0.4, 0.2), (1.6, -0.2, 0.1), (-0.3, 1.5, -0.2), (0.5, -0.5, 0.3), (-0.5, 0.5, -0.1), (0.8, 0.8, 0.2)]
    out = []
    for p in pts:
        R = np.linalg.qr(rng.standard_normal((3, 3)))[0]
        t = rng.standard_normal(3)
        out.append(np.concatenate([R @ np.array(p) + t] + [R @ v + t for v in tri]))
    return out


def test_point_triangle_gradient_fd_all_regions():
    rng = np.random.default_rng(1)
    for p12 in voronoi_cases(rng):
        d, g = ptd_grad(p12)
        gf = fd_grad(lambda q: ptd(q), p12)
        assert np.allclose(g, gf, atol=1e-6), (g, gf)
        assert np.allclose(g.reshape(4, 3).sum(0), 0, atol=1e-12)  # translation invariance


def test_segment_gradient_fd():
    rng = np.random.default_rng(2)
    for _ in range(20):
        p12 = rng.standard_normal(12)
        d, g = ssd_grad(p12)
        gf = fd_grad(lambda q: ssd_grad(q)[0], p12)
        assert np.allclose(g, gf, atol=1e-5), (g, gf)
        assert abs(d - ssd(p12)) < 1e-9  # witness distance equals the clamped closest pair
        assert np.allclose(g.reshape(4, 3).sum(0), 0, atol=1e-12)


def test_degenerate_triangle_no_nan():
    p12 = np.array([0.3, 0.2, 0.1, 0, 0, 0, 0, 0, 0, 1, 0, 0], float)  # collapsed edge a-b
    for fp64 in (0, 1):
        d, g = ptd_grad(p12, fp64)
        assert np.isfinite(d) and np.all(np.isfinite(g))


def test_edge_triangle_intersect_cases():
    tri = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    pierce = np.array([0.2, 0.2, -1, 0.2, 0.2, 1] + tri, float)
    miss = np.array([2.0, 2.0, -1, 2.0, 2.0, 1] + tri, float)
    touch = np.array([0.2, 0.2, 0, 0.2, 0.2, 1] + tri, float)  # endpoint on the plane -> not strict
    for fp64 in (0, 1):
        assert lib.oracle_edge_triangle_intersect(fp64, O.dp(pierce)) == 1
        assert lib.oracle_edge_triangle_intersect(fp64, O.dp(miss)) == 0
        assert lib.oracle_edge_triangle_intersect(fp64, O.dp(touch)) == 0


# ------------------------------------------------------------------ barriers, stiffness, friction
@pytest.mark.parametrize("kind", [0, 1])
def test_barrier_fd_and_c2_at_dhat(kind):
    dhat, kappa = 1e-3, 7.0
    out = np.zeros(3)

    def ev(d):
        o = np.zeros(3)
        return o if not lib.oracle_barrier(1, kind, d, dhat, kappa, O.dp(o)) else o

    for d in (1e-4, 3e-4, 6e-4, 9e-4):
        e, g, h = ev(d)
        hh = 1e-9
        assert abs(g - (ev(d + hh)[0] - ev(d - hh)[0]) / (2 * hh)) < 1e-5 * max(1, abs(g))
        assert abs(h - (ev(d + hh)[1] - ev(d - hh)[1]) / (2 * hh)) < 1e-4 * max(1, abs(h))
        assert g < 0 and h > 0  # repulsive, convex
    # inactive at and beyond dhat; value, slope and curvature vanish approaching it
    assert lib.oracle_barrier(1, kind, dhat, dhat, kappa, O.dp(out)) == 0
    e, g, h = ev(dhat * (1 - 1e-7))
    assert abs(e) < 1e-12 and abs(g) < 1e-9 * kappa
    if kind == 0:
        assert abs(h) < 1e-3 * kappa


def test_rayleigh_single_participant_closed_form():
    # N = 1 reduces exactly to n^T E n + m / d^2 (contact.rs:980-990 doc comment)
    rng = np.random.default_rng(3)
    for _ in range(10):
        a = rng.standard_normal((3, 3))
        E = a @ a.T
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        d, m = 0.3e-3 + rng.random() * 1e-3, 1e-4
        k = lib.oracle_rayleigh_stiffness(1, O.dp(E.ravel().copy()), O.dp(np.array([m])), O.dp(np.array([1.0])), 1,
                                          O.dp(n * d), 0.0)
        assert abs(k - (n @ E @ n + m / d**2)) < 1e-9 * k


def test_rayleigh_static_partner_uses_max_mass():
    # a zero mass is replaced by the stencil's largest (barrier.cu semantics)
    el = np.zeros((16, 9))
    w = np.array([1.0, -0.5, -0.3, -0.2])
    e = np.array([0.0, 0.0, 1e-3])
    k1 = lib.oracle_rayleigh_stiffness(1, O.dp(el.ravel()), O.dp(np.array([2e-4, 0, 0, 0.0])), O.dp(w), 4, O.dp(e), 0.0)
    k2 = lib.oracle_rayleigh_stiffness(1, O.dp(el.ravel()), O.dp(np.array([2e-4, 2e-4, 2e-4, 2e-4])), O.dp(w), 4, O.dp(e), 0.0)
    assert abs(k1 - k2) < 1e-9 * k2


def test_friction_coulomb_bound_and_psd():
    n = np.array([0.0, 0.0, 1.0])
    f = n * -5.0  # barrier force along -n (dE/dd < 0 convention: contact = -n.f_c = 5)
    mu = 0.3
    g, H = np.zeros(3), np.zeros(9)
    assert lib.oracle_friction(1, O.dp(f), O.dp(np.array([0.02, 0.01, 0.4])), O.dp(n), mu, 1e-5, O.dp(g), O.dp(H)) == 1
    assert abs(np.linalg.norm(g) - mu * 5.0) < 1e-12  # sliding: Coulomb bound
    assert np.all(np.linalg.eigvalsh(H.reshape(3, 3)) > -1e-12)
    assert abs(g @ n) < 1e-12  # tangential
    assert lib.oracle_friction(1, O.dp(f), O.dp(np.array([0.02, 0.01, 0.4])), O.dp(n), 0.0, 1e-5, O.dp(g), O.dp(H)) == 0


# ------------------------------------------------------------------ elements
def shell(p9, rest9, model=0, mu=19.2, lam=28.8, eps=1e-6, fp64=1):
    g, h = np.zeros(9), np.zeros(81)
    e = lib.oracle_shell_element(fp64, model, mu, lam, eps, O.dp(O.d64(rest9)), O.dp(O.d64(p9)), O.dp(g), O.dp(h))
    return e, g, h.reshape(9, 9)


@pytest.mark.parametrize("model", [0, 1])
def test_shell_gradient_fd_and_psd(model):
    rng = np.random.default_rng(4)
    rest = np.array([0, 0, 0, 1, 0, 0, 0.2, 0.9, 0], float)
    for _ in range(6):
        p = rest + 0.15 * rng.standard_normal(9)
        e, g, h = shell(p, rest, model)
        gf = fd_grad(lambda q: shell(q, rest, model)[0], p)
        assert np.allclose(g, gf, atol=1e-6 * max(1, np.abs(g).max())), (g, gf)
        assert np.linalg.eigvalsh(0.5 * (h + h.T)).min() > -1e-8 * np.abs(h).max()


def test_shell_hessian_exact_when_unclamped_and_rigid_invariant():
    # ARAP at F = diag(1.2, 1.0) (mu 19.2, lambda 28.8): twist 7.5, flip 13.4, out-of-plane
    # 8.0 / 6.9 and both stretch eigenvalues are positive, so the PSD filter is the identity
    # and the analytic element Hessian must equal the FD Hessian of the gradient.
    rest = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], float)
    p = rest.copy()
    p[3] = 1.2
    R = np.linalg.qr(np.random.default_rng(5).standard_normal((3, 3)))[0]
    p = (p.reshape(3, 3) @ R.T + np.array([0.3, -1, 2])).ravel()
    e, g, h = shell(p, rest, model=1)
    hf = np.stack([fd_grad(lambda q, i=i: shell(q, rest, 1)[1][i], p) for i in range(9)])
    assert np.allclose(h, hf, atol=5e-5 * np.abs(h).max()), np.abs(h - hf).max()
    e0 = shell(rest.copy(), rest, model=1)[0]
    assert abs(e0) < 1e-12  # rest state carries no energy
    p2 = (p.reshape(3, 3) @ R + np.array([-2.0, 0.5, 0.1])).ravel()  # another rigid motion
    assert abs(shell(p2, rest, model=1)[0] - e) < 1e-9 * abs(e)


def test_shell_stvk_mode_curvatures_are_clamped_eigenvalues():
    # StVK stretched by 1.3 x 1.1: the flip curvature (d0 - d1)/(s0 - s1) is negative. Along
    # each mode direction of the SVD frame the element Hessian must give max(lambda_k, 0)
    # where lambda_k is the true curvature (FD Hessian), and the gradient stays exact.
    rest = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], float)  # Dm = I: F = [x1 - x0, x2 - x0]
    F = np.array([[1.3, 0.05], [0.02, 1.1], [0.03, -0.04]])
    p = np.concatenate([[0, 0, 0], F[:, 0], F[:, 1]])
    e, g, h = shell(p, rest, model=0)
    hf = np.stack([fd_grad(lambda q, i=i: shell(q, rest, 0)[1][i], p, 1e-6) for i in range(9)])
    hf = 0.5 * (hf + hf.T)
    U, s, Vt = np.linalg.svd(F, full_matrices=False)
    u2 = np.cross(U[:, 0], U[:, 1])
    U3 = np.column_stack([U, u2])
    r = np.sqrt(0.5)
    modes = [np.array([[0, r], [-r, 0], [0, 0]]), np.array([[0, r], [r, 0], [0, 0]]),
             np.array([[0, 0], [0, 0], [1, 0]]), np.array([[0, 0], [0, 0], [0, 1]])]
    negative_seen = False
    for q in modes:
        Q = U3 @ q @ Vt
        dx = np.concatenate([[0, 0, 0], Q[:, 0], Q[:, 1]])
        lam_true = dx @ hf @ dx / (dx @ dx) * 2.0  # area 1/2 folds into the F-space curvature
        lam_elem = dx @ h @ dx / (dx @ dx) * 2.0
        negative_seen |= lam_true < 0
        assert abs(lam_elem - max(lam_true, 0.0)) < 2e-4 * max(1.0, abs(lam_true)), (lam_elem, lam_true)
    assert negative_seen


def test_dihedral_gradient_fd_and_flat_zero():
    import ctypes as C
    rng = np.random.default_rng(6)
    flat = np.array([0, 1, 0, 0, 0, 0, 1, 0, 0, 1, -1, 0], float)  # tri A=(v0,v1,v2), B=(v1,v2,v3) coplanar
    g = np.zeros(12)
    assert abs(lib.oracle_dihedral(1, O.dp(flat), O.dp(g))) < 1e-12
    for _ in range(8):
        p = flat + 0.2 * rng.standard_normal(12)
        th = lib.oracle_dihedral(1, O.dp(p), O.dp(g))
        gf = fd_grad(lambda q: lib.oracle_dihedral(1, O.dp(O.d64(q)), O.dp(np.zeros(12))), p)
        assert np.allclose(g, gf, atol=1e-6), (g, gf)


def test_strand_gradient_fd_psd_and_exact_when_unclamped():
    rng = np.random.default_rng(7)
    for _ in range(6):
        p = np.array([-1, 0.1, 0, 0, 0, 0, 1, 0.3, 0.1], float) + 0.1 * rng.standard_normal(9)
        rest = 2.5
        g, h = np.zeros(9), np.zeros(81)
        th = lib.oracle_strand(1, O.dp(p), rest, O.dp(g), O.dp(h))
        energy = lambda q: 0.5 * (lib.oracle_strand(1, O.dp(O.d64(q)), rest, O.dp(np.zeros(9)), O.dp(np.zeros(81))) - rest) ** 2
        assert np.allclose(g, fd_grad(energy, p), atol=1e-6)
        H = h.reshape(9, 9)
        assert np.linalg.eigvalsh(0.5 * (H + H.T)).min() > -1e-9
        # the true Hessian (FD of the gradient) minus the projection is negative semidefinite
        hf = np.stack([fd_grad(lambda q, i=i: (lambda gg: (lib.oracle_strand(1, O.dp(O.d64(q)), rest, O.dp(gg), O.dp(np.zeros(81))), gg)[1])(np.zeros(9))[i], p, h=1e-5) for i in range(9)])
        diff = 0.5 * (hf + hf.T) - H
        assert np.linalg.eigvalsh(diff).max() < 1e-4 * np.abs(H).max()


def test_rod_stretch_exact_when_stretched():
    d = np.array([0.3, 0.4, 1.2])
    rest, k = 1.0, 1e3
    h, ga = np.zeros(9), np.zeros(3)
    lib.oracle_rod_stretch(1, O.dp(d), rest, k, O.dp(h), O.dp(ga))
    energy = lambda q: 0.5 * k * (np.linalg.norm(q) - rest) ** 2
    assert np.allclose(ga, fd_grad(energy, d), atol=1e-6)
    hf = np.stack([fd_grad(lambda q, i=i: fd_grad(energy, q, 1e-5)[i], d, 1e-5) for i in range(3)])
    assert np.allclose(h.reshape(3, 3), hf, rtol=1e-3, atol=1e-3)


def test_spring_block_exact_when_stretched_and_projected_when_compressed():
    # implicit.rs:3240-3261: gradient k (l - L) d/l of 1/2 k (l - L)^2; stretched, the block
    # is the exact Hessian; compressed, the tangential term (1 - L/l < 0) is clamped, leaving
    # k n n^T, and the unprojected form keeps it (the exact Hessian again)
    rest, k = 1.0, 8e2
    energy = lambda q: 0.5 * k * (np.linalg.norm(q) - rest) ** 2
    for d in (np.array([0.3, 0.4, 1.2]), np.array([0.2, -0.3, 0.5])):
        exact = np.stack([fd_grad(lambda q, i=i: fd_grad(energy, q, 1e-5)[i], d, 1e-5) for i in range(3)])
        for project in (1, 0):
            h, ga = np.zeros(9), np.zeros(3)
            lib.oracle_spring(1, O.dp(d), rest, k, project, O.dp(h), O.dp(ga))
            assert np.allclose(ga, fd_grad(energy, d), atol=1e-6)
            l = np.linalg.norm(d)
            if l >= rest or not project:
                assert np.allclose(h.reshape(3, 3), exact, rtol=1e-3, atol=1e-3)
            else:
                n = d / l
                assert np.allclose(h.reshape(3, 3), k * np.outer(n, n), rtol=1e-12, atol=1e-9)
    # a degenerate edge: isotropic k, no force
    h, ga = np.zeros(9), np.zeros(3)
    lib.oracle_spring(0, O.dp(np.zeros(3)), rest, k, 1, O.dp(h), O.dp(ga))
    assert np.array_equal(h.reshape(3, 3), k * np.eye(3)) and not ga.any()


# ------------------------------------------------------------------ linear solve
def test_cg_manufactured_laplacian_and_zero_rhs():
    n = 3 * GOLD["pcg_manufactured"]["n"]
    rp, ci, va = [0], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                ci.append(j)
                va.append(v)
        rp.append(len(ci))
    xs = np.linspace(-1, 1, n).astype(np.float32)
    b = O.spmv(rp, ci, va, xs)
    x, it, res = O.cg(rp, ci, va, b, 1e-7, 100000)
    assert np.abs(x - xs).max() < 1e-3
    x0, it0, res0 = O.cg(rp, ci, va, np.zeros(n, np.float32), 1e-6, 10)
    assert np.all(x0 == 0)


# ------------------------------------------------------------------ LBVH stages
def test_morton_decode_roundtrip():
    s = nguo_hip.Scene(0, seed=0)
    pos, tris = O.f32(s.positions), O.u32(s.tris)
    codes = np.zeros(len(tris), np.uint32)
    lib.oracle_morton_codes(len(pos), O.fp(pos), len(tris), O.up(tris), O.up(codes))
    c = (pos[tris[:, 0]] + pos[tris[:, 1]] + pos[tris[:, 2]]) / np.float32(3.0)
    lo, hi = c.min(0), c.max(0)
    inv = np.where(hi - lo > 0, np.float32(1.0) / (hi - lo), np.float32(0.0)).astype(np.float32)
    q = np.clip(np.clip((c - lo) * inv, 0, 1) * np.float32(1023.0), 0, 1023).astype(np.uint32)

    def compact(x):
        x &= 0x09249249
        x = (x ^ (x >> 2)) & 0x030C30C3
        x = (x ^ (x >> 4)) & 0x0300F00F
        x = (x ^ (x >> 8)) & 0x030000FF
        x = (x ^ (x >> 16)) & 0x000003FF
        return x

    assert np.array_equal(compact(codes >> 2), q[:, 0])
    assert np.array_equal(compact(codes >> 1), q[:, 1])
    assert np.array_equal(compact(codes), q[:, 2])


def radix_tree_reference(codes):
    """Binary radix tree on (code, index) keys by recursive highest-differing-bit splits,
    numbered like Karras (node covering [f, l] split at g: children g and g + 1)."""
    n = len(codes)
    keys = [(int(c) << 32) | i for i, c in enumerate(codes)]
    left = np.zeros(n - 1, np.uint64)
    right = np.zeros(n - 1, np.uint64)
    LEAF = 1 << 31

    def split(f, l):
        x = keys[f] ^ keys[l]
        hb = x.bit_length() - 1
        g = f
        for i in range(f, l):
            if (keys[i + 1] >> hb) & 1 != (keys[f] >> hb) & 1:
                g = i
                break
        return g

    stack = [(0, n - 1, 0)]
    while stack:
        f, l, idx = stack.pop()
        g = split(f, l)
        left[idx] = (g | LEAF) if g == f else g
        right[idx] = ((g + 1) | LEAF) if g + 1 == l else g + 1
        if g != f:
            stack.append((f, g, g))
        if g + 1 != l:
            stack.append((g + 1, l, g + 1))
    return left.astype(np.uint32), right.astype(np.uint32)


def test_karras_matches_recursive_radix_tree_and_aabbs():
    rng = np.random.default_rng(8)
    for n in (2, 3, 17, 300):
        codes = np.sort(rng.integers(0, 1 << 12, n).astype(np.uint32))  # duplicates on purpose
        left, right = np.zeros(n - 1, np.uint32), np.zeros(n - 1, np.uint32)
        lib.oracle_karras(n, O.up(codes), O.up(left), O.up(right))
        rl, rr = radix_tree_reference(codes)
        assert np.array_equal(left, rl) and np.array_equal(right, rr)
        leaf = rng.standard_normal((n, 6)).astype(np.float32)
        leaf[:, 3:] = leaf[:, :3] + np.abs(leaf[:, 3:])
        node = np.zeros((n - 1, 6), np.float32)
        lib.oracle_propagate_aabb(n, O.up(left), O.up(right), O.fp(leaf), O.fp(node))

        def rng_of(c, is_leaf):  # leaf range of a subtree
            if is_leaf:
                return c, c
            a = rng_of(int(left[c] & 0x7FFFFFFF), bool(left[c] >> 31))
            b = rng_of(int(right[c] & 0x7FFFFFFF), bool(right[c] >> 31))
            return a[0], b[1]

        for i in range(n - 1):
            f, l = rng_of(i, False)
            assert np.array_equal(node[i, :3], leaf[f:l + 1, :3].min(0)) and np.array_equal(node[i, 3:], leaf[f:l + 1, 3:].max(0))


def test_sort_pairs_is_stable():
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 50, 2000).astype(np.uint32)
    vals = np.arange(2000, dtype=np.uint32)
    k, v = keys.copy(), vals.copy()
    lib.oracle_sort_pairs(len(k), O.up(k), O.up(v))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


# ------------------------------------------------------------------ whole-solver checks
def test_free_fall_is_exact_inertial_target():
    # a resting sheet in free fall: elastic and bending forces vanish for a translation, so
    # one implicit step lands exactly on x0 + h^2 g (v0 = 0)
    s = nguo_hip.Scene(100, seed=0)
    s.desc.params.self_collision = 0
    o = O.OracleSolver(s, fp64=True)
    x0 = o.positions_d()
    o.step(1)
    x1 = o.positions_d()
    h = s.params.dt
    free = s.inv_mass > 0
    want = x0.copy()
    want[free, 1] += h * h * -9.8
    assert np.abs(x1 - want).max() < 1e-7


def test_oracle_drape_invariants_small():
    s = nguo_hip.Scene(100, seed=0)
    o = O.OracleSolver(s)
    for _ in range(40):
        o.step(1)
        m = o.metrics()
        assert m["all_finite"] == 1 and m["penetration_free"] == 1
        assert m["max_reresid"] < 5e-6
    x = o.positions()
    free = s.inv_mass > 0
    # the sheet rests on the faceted sphere: in contact, but no free vertex inside the body
    assert np.linalg.norm(x[free], axis=1).min() < 0.3 + 2e-3
    assert np.all(o.collider_sides(x) == 1)
    assert len(o.stencils(x)[0]) > 0


def grid_system(nx, ny, mass=1e-2, k=1.0):
    """Block CSR of a mass + spring grid (3 dofs per vertex), SPD."""
    n = nx * ny
    rows = {}
    def add(i, j, blk):
        rows.setdefault(i, {}).setdefault(j, np.zeros((3, 3)))
        rows[i][j] += blk
    for v in range(n):
        add(v, v, mass * np.eye(3))
    for y in range(ny):
        for x in range(nx):
            v = y * nx + x
            for dx, dy in ((1, 0), (0, 1), (1, 1)):
                if x + dx < nx and y + dy < ny:
                    w = v + dy * nx + dx
                    K = k * np.eye(3)
                    add(v, v, K); add(w, w, K); add(v, w, -K); add(w, v, -K)
    rp, ci, va = [0], [], []
    for i in range(n):
        cols = sorted(rows[i])
        for r in range(3):
            for j in cols:
                for c in range(3):
                    ci.append(3 * j + c)
                    va.append(rows[i][j][r, c])
            rp.append(len(ci))
    return np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, np.float32)


def test_schwarz_beats_block_jacobi_on_a_grid():
    # ec/linsolve.rs:915-968: same solution, fewer iterations
    rp, ci, va = grid_system(24, 24)
    b = np.random.default_rng(12).standard_normal(len(rp) - 1).astype(np.float32)
    xj, itj, rj = O.cg(rp, ci, va, b, 1e-6, 100000)
    x1, it1, r1 = O.cg_schwarz(rp, ci, va, b, 1e-6, 100000, 1)
    assert it1 == itj  # one-vertex domains are the block-Jacobi preconditioner
    xs, its, rs = O.cg_schwarz(rp, ci, va, b, 1e-6, 100000, 16)
    assert its < 0.7 * itj, (its, itj)
    assert np.abs(xs - xj).max() < 1e-3 * np.abs(xj).max()


def test_schwarz_domains_partition_and_single_vertex_limit():
    rp, ci, va = grid_system(10, 7)
    doms, _ = O.schwarz(rp, ci, va, 16)
    allv = np.sort(np.concatenate(doms))
    assert np.array_equal(allv, np.arange(70))
    assert max(len(d) for d in doms) <= 16 and doms[0][0] == 0
    # domain of one vertex: the apply is the 3x3 diagonal-block inverse (up to the 1e-6 floor)
    r = np.random.default_rng(13).standard_normal(210)
    _, z = O.schwarz(rp, ci, va, 1, r)
    n = 210
    for v in range(70):
        blk = np.zeros((3, 3))
        for rr in range(3):
            row = 3 * v + rr
            for k in range(rp[row], rp[row + 1]):
                if ci[k] // 3 == v:
                    blk[rr, ci[k] % 3] = va[k]
        assert np.allclose(z[3 * v:3 * v + 3], np.linalg.solve(blk, r[3 * v:3 * v + 3]), rtol=1e-5)


def test_two_level_schwarz_fewer_iterations_on_a_stiff_chain():
    # MultilevelSchwarz (linsolve.rs:522-657): the coarse correction couples domains, so a
    # long stiff chain (rod-like) needs fewer iterations with it; weight 0 is one level
    n = 400
    rows = {}
    def add(i, j, blk):
        rows.setdefault(i, {}).setdefault(j, np.zeros((3, 3)))
        rows[i][j] += blk
    for v in range(n):
        add(v, v, 1e-6 * np.eye(3))
    for v in range(n - 1):
        K = np.eye(3)
        add(v, v, K); add(v + 1, v + 1, K); add(v, v + 1, -K); add(v + 1, v, -K)
    rp, ci, va = [0], [], []
    for i in range(n):
        for r in range(3):
            for j in sorted(rows[i]):
                for c in range(3):
                    ci.append(3 * j + c); va.append(rows[i][j][r, c])
            rp.append(len(ci))
    rp, ci, va = np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, np.float32)
    b = np.random.default_rng(14).standard_normal(3 * n).astype(np.float32)
    x1, it1, _ = O.cg_schwarz(rp, ci, va, b, 1e-6, 100000, 8, 0.0)
    x2, it2, _ = O.cg_schwarz(rp, ci, va, b, 1e-6, 100000, 8, 0.7)
    assert it2 < 0.8 * it1, (it2, it1)
    assert np.abs(x2 - x1).max() < 1e-3 * np.abs(x1).max()
