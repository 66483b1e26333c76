"""More pins of the oracle, each by an independent route:

- the contact stencil list (vertex-triangle, then edge-edge) equal, as a list, to a brute
  force over every pair under the reference's candidate rules (broadphase.rs:269-329 grid
  cells, broadphase.rs:756-775 one-ring exclusion, implicit.rs:678-753 activation), with
  the gaps and gradients bitwise the element routines';
- the inside side of a mesh collider on vertices placed behind a facet;
- the Rayleigh contact stiffness for 2-4 participants against a numpy Rayleigh quotient of
  the dense block matrix;
- the two-level additive Schwarz apply against dense numpy S0 + w C S1 C^T;
- the stencil friction proximity weights against the closest points they encode.
CPU only."""
import numpy as np
import pytest

from conftest import nguo_hip
import oracle_py as O
from cases import inside_state

lib = O.load()


# ------------------------------------------------------------------ brute-force stencils
def one_ring(nv, tris):
    """Dense exclusion matrix: ex[v, u] when u == v or u and v share a triangle."""
    ex = np.eye(nv, dtype=bool)
    for a in range(3):
        for b in range(3):
            ex[tris[:, a], tris[:, b]] = True
    return ex


def grid_cell(x, tris, dhat):
    """max(dhat, mean largest AABB side, 1e-6), the mean summed in FP32 in triangle order."""
    a, b, c = x[tris[:, 0]], x[tris[:, 1]], x[tris[:, 2]]
    d = np.maximum(np.maximum(a, b), c) - np.minimum(np.minimum(a, b), c)
    ext = np.maximum(np.maximum(d[:, 0], d[:, 1]), d[:, 2])
    s = np.float32(0.0)
    for e in ext:
        s = np.float32(s + e)
    mean = np.float32(s / np.float32(len(tris))) if len(tris) else np.float32(0.0)
    return np.float32(max(max(np.float32(dhat), mean), np.float32(1e-6)))


def elem_vt(p12):
    g = np.zeros(12)
    d = lib.oracle_point_tri_distance_grad(0, O.dp(O.d64(p12)), O.dp(g))
    return np.float32(d), g.astype(np.float32)


def elem_ee(p12):
    g = np.zeros(12)
    d = lib.oracle_segment_distance_grad(0, O.dp(O.d64(p12)), O.dp(g))
    return np.float32(d), g.astype(np.float32)


def brute_stencils(scene, x, side):
    """Every active stencil at x by brute force: (idx rows, gaps, gradient rows)."""
    x = np.asarray(x, np.float32)
    tris = scene.tris.astype(np.int64)
    nv = len(x)
    mov = scene.inv_mass > 0
    static = scene.static.astype(bool)
    dhat = np.float32(scene.params.contact_dhat)
    ex = one_ring(nv, tris)
    coll_tri = static[tris].all(axis=1)
    cell = grid_cell(x, tris, dhat)
    inv = np.float32(1.0) / cell
    key = np.floor(x * inv).astype(np.int64)                      # (V, 3)
    a, b, c = x[tris[:, 0]], x[tris[:, 1]], x[tris[:, 2]]
    lo = np.minimum(np.minimum(a, b), c) - dhat
    hi = np.maximum(np.maximum(a, b), c) + dhat
    klo, khi = np.floor(lo * inv).astype(np.int64), np.floor(hi * inv).astype(np.int64)
    tmov = mov[tris].any(axis=1)
    vt = []
    for v in range(nv):
        # grid candidates (the vertex's cell inside the padded triangle's cell range), the
        # movability and exclusion filters
        cand = np.all((klo <= key[v]) & (key[v] <= khi), axis=1) & (tmov | mov[v])
        cand &= ~(ex[v, tris[:, 0]] | ex[v, tris[:, 1]] | ex[v, tris[:, 2]])
        signed = coll_tri & (side[v] < 0)
        for t in np.nonzero(cand)[0]:
            d, g = elem_vt(np.concatenate([x[v], a[t], b[t], c[t]]))
            if signed[t]:
                d, g = -d, -g
            if d < dhat:
                vt.append(((v, *tris[t]), d, g))
    # edge-edge: sorted unique mesh edges; a pair with no shared vertex, one edge movable and
    # segment distance below dhat (its padded AABBs overlap then, which prefilters)
    e = np.unique(np.sort(np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]), axis=1), axis=0)
    emov = mov[e].any(axis=1)
    elo = np.minimum(x[e[:, 0]], x[e[:, 1]]).astype(np.float64)
    ehi = np.maximum(x[e[:, 0]], x[e[:, 1]]).astype(np.float64)
    pad = 1.01 * float(dhat)
    ee = []
    for i in range(len(e)):
        j = np.arange(i + 1, len(e))
        ok = (emov[i] | emov[j]) & np.all((elo[j] <= ehi[i] + pad) & (ehi[j] >= elo[i] - pad), axis=1)
        ok &= (e[j, 0] != e[i, 0]) & (e[j, 0] != e[i, 1]) & (e[j, 1] != e[i, 0]) & (e[j, 1] != e[i, 1])
        for jj in j[ok]:
            d, g = elem_ee(np.concatenate([x[e[i, 0]], x[e[i, 1]], x[e[jj, 0]], x[e[jj, 1]]]))
            if d < dhat:
                ee.append(((*e[i], *e[jj]), d, g))
    rows = vt + ee
    idx = np.array([r[0] for r in rows], np.uint32).reshape(-1, 4)
    return idx, np.array([r[1] for r in rows], np.float32), np.array([r[2] for r in rows], np.float32).reshape(-1, 12), len(vt)


def check_equal_to_brute(scene, x, side):
    o = O.OracleSolver(scene)
    idx, d, g = o.stencils(x)
    bidx, bd, bg, nvt = brute_stencils(scene, x, side)
    assert len(bidx) > 0
    # the vertex-triangle part in (v, triangle index) order: map brute rows to it
    tri_id = {tuple(r): k for k, r in enumerate(scene.tris.astype(np.int64).tolist())}
    order = sorted(range(nvt), key=lambda k: (int(bidx[k, 0]), tri_id[tuple(int(a) for a in bidx[k, 1:])]))
    perm = np.array(order + list(range(nvt, len(bidx))), np.int64)
    bidx, bd, bg = bidx[perm], bd[perm], bg[perm]
    assert idx.shape == bidx.shape, (len(idx), len(bidx))
    assert np.array_equal(idx, bidx)
    assert np.array_equal(d, bd)
    assert np.array_equal(g, bg)
    return idx, d, nvt


@pytest.mark.parametrize("cfg,steps", [(100, 12), (100, 30), (0, 16)])
def test_stencils_equal_brute_force(cfg, steps):
    s = nguo_hip.Scene(cfg, seed=cfg)
    o = O.OracleSolver(s)
    o.step(steps)
    x = o.positions()
    side = o.collider_sides(x)
    assert side is not None and np.all(side == 1)  # every vertex outside the body here
    idx, d, nvt = check_equal_to_brute(s, x, side)
    assert nvt > 0 and len(idx) > nvt, "both stencil kinds must be present"


def test_collider_inside_side_and_signed_stencils():
    s, x, v, vs, vd, ts, td = inside_state()
    o = O.OracleSolver(s)
    side = o.collider_sides(x)
    dhat = float(s.params.contact_dhat)
    # within dhat behind a facet: inside; deeper than dhat: the side is clamped to +1
    want = np.ones(len(x), np.float32)
    want[vs] = -1.0
    assert np.array_equal(side, want)
    idx, d, nvt = check_equal_to_brute(s, x, side)
    static = s.static.astype(bool)
    mine = [k for k in range(nvt) if idx[k, 0] == vs and static[idx[k, 1:]].all()]
    # the facet it sits behind at -0.4 dhat, and every other collider facet of its grid cells
    # at its (negative) signed gap
    facet = [k for k in mine if tuple(int(a) for a in idx[k, 1:]) == ts]
    assert len(facet) == 1 and abs(d[facet[0]] + 0.4 * dhat) < 1e-7
    assert len(mine) >= 2 and np.all(d[mine] < 0)
    # the deep vertex has no collider stencil: its unsigned gaps exceed dhat
    assert not any(idx[k, 0] == vd and static[idx[k, 1:]].all() for k in range(nvt))


# ------------------------------------------------------------------ Rayleigh stiffness
def rayleigh_dense(el, mass, w, e, offset):
    """w^T K w / w^T w with K = blocks + m_i / g^2 I, w = kron(weights, e) (numpy)."""
    n = len(mass)
    g = np.linalg.norm(e) - offset
    if g <= 0:
        return 0.0
    m = np.where(mass > 0, mass, mass.max())
    K = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for j in range(n):
            K[3 * i:3 * i + 3, 3 * j:3 * j + 3] = el[i * n + j].reshape(3, 3)
        K[3 * i:3 * i + 3, 3 * i:3 * i + 3] += m[i] / g**2 * np.eye(3)
    wv = np.kron(w, e)
    return wv @ K @ wv / (wv @ wv)


@pytest.mark.parametrize("n", [2, 3, 4])
def test_rayleigh_matches_dense_quotient(n):
    rng = np.random.default_rng(40 + n)
    for trial in range(20):
        a = rng.standard_normal((3 * n, 3 * n))
        K = a @ a.T + 0.3 * rng.standard_normal((3 * n, 3 * n))  # blocks need not be symmetric
        el = np.zeros((16, 9))
        for i in range(n):
            for j in range(n):
                el[i * n + j] = K[3 * i:3 * i + 3, 3 * j:3 * j + 3].ravel()
        mass = rng.uniform(1e-4, 5e-4, n)
        if trial % 3 == 0:
            mass[rng.integers(1, n)] = 0.0  # a pinned participant takes the stencil's largest mass
        w = np.concatenate([[1.0], -rng.dirichlet(np.ones(n - 1))]) if trial % 2 else rng.uniform(-1, 1, n)
        e = rng.standard_normal(3) * 1e-3
        off = 0.2 * np.linalg.norm(e) if trial % 4 == 1 else 0.0
        want = rayleigh_dense(el[: n * n], mass, w, e, off)
        got = lib.oracle_rayleigh_stiffness(1, O.dp(el.ravel().copy()), O.dp(mass.copy()), O.dp(w.copy()), n, O.dp(e.copy()), off)
        assert abs(got - want) <= 1e-10 * abs(want), (n, trial, got, want)
        got32 = lib.oracle_rayleigh_stiffness(0, O.dp(el.ravel().copy()), O.dp(mass.copy()), O.dp(w.copy()), n, O.dp(e.copy()), off)
        assert abs(got32 - want) <= 2e-5 * abs(want) + 1e-5 * np.abs(K).max(), (n, trial, got32, want)
    # a gap at or below the offset, or no direction, gives no stiffness
    z = np.zeros(16 * 9)
    assert lib.oracle_rayleigh_stiffness(1, O.dp(z), O.dp(np.ones(n)), O.dp(np.ones(n)), n, O.dp(np.array([1e-3, 0, 0])), 1e-3) == 0.0
    assert lib.oracle_rayleigh_stiffness(1, O.dp(z), O.dp(np.ones(n)), O.dp(np.zeros(n)), n, O.dp(np.array([1e-3, 0, 0])), 0.0) == 0.0


# ------------------------------------------------------------------ two-level Schwarz
def grid_csr(nx, ny, rng):
    """3x3-block CSR of a random SPD operator on an nx x ny grid graph (8-neighbour)."""
    n = nx * ny
    A = np.zeros((3 * n, 3 * n))
    for y in range(ny):
        for x in range(nx):
            v = y * nx + x
            A[3 * v:3 * v + 3, 3 * v:3 * v + 3] += 1e-2 * np.eye(3)
            for dx, dy in ((1, 0), (0, 1), (1, 1), (-1, 1)):
                if 0 <= x + dx < nx and y + dy < ny:
                    w = v + dy * nx + dx
                    q = rng.standard_normal((3, 3))
                    K = q @ q.T + np.eye(3)
                    for (i, j, s) in ((v, v, 1), (w, w, 1), (v, w, -1), (w, v, -1)):
                        A[3 * i:3 * i + 3, 3 * j:3 * j + 3] += s * K
    A = A.astype(np.float32).astype(np.float64)
    rp, ci, va = [0], [], []
    for r in range(3 * n):
        nz = np.nonzero(np.abs(A[r].reshape(-1, 3)).sum(axis=1))[0]
        for j in nz:
            for c in range(3):
                ci.append(3 * j + c)
                va.append(A[r, 3 * j + c])
        rp.append(len(ci))
    return A, np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, np.float32)


def floored_inv(Ad):
    """Inverse of a domain block with its diagonal raised by 1e-6 mean(diag) + 1e-8 (the
    reference's factorisation floor, linsolve.rs:217-224), read from its lower triangle as a
    Cholesky factorisation reads it (linsolve.rs:80-99)."""
    L = np.tril(Ad)
    Ad = L + np.tril(Ad, -1).T
    return np.linalg.inv(Ad + (1e-6 * np.trace(Ad) / len(Ad) + 1e-8) * np.eye(len(Ad)))


def one_level_dense(A, doms):
    S = np.zeros_like(A)
    for dv in doms:
        dofs = np.concatenate([[3 * v, 3 * v + 1, 3 * v + 2] for v in dv])
        S[np.ix_(dofs, dofs)] += floored_inv(A[np.ix_(dofs, dofs)])
    return S


@pytest.mark.parametrize("weight", [0.0, 0.7])
def test_two_level_schwarz_apply_matches_dense(weight):
    rng = np.random.default_rng(21)
    A, rp, ci, va = grid_csr(8, 8, rng)
    domain = 8  # 8 aggregates of 8 vertices: the coarse operator fits one coarse domain
    doms, z = O.schwarz(rp, ci, va, domain, rng.standard_normal(len(A)), weight)
    assert len(doms) == 8 and all(len(d) <= domain for d in doms)
    S = one_level_dense(A, doms)
    if weight:
        C = np.zeros((len(A), 3 * len(doms)))  # prolongation: vertex -> its aggregate
        for k, dv in enumerate(doms):
            for v in dv:
                C[3 * v:3 * v + 3, 3 * k:3 * k + 3] = np.eye(3)
        # A1 = C^T A C with its entries summed in FP32 in fine CSR order (linsolve.rs:613-625):
        # the aggregates cancel their springs, so the FP32 sums differ from FP64 ones by 1e-5
        agg = np.zeros(len(A) // 3, np.int64)
        for k, dv in enumerate(doms):
            agg[dv] = k
        A1 = np.zeros((3 * len(doms), 3 * len(doms)), np.float32)
        for r in range(len(A)):
            for k in range(rp[r], rp[r + 1]):
                cr, cc = 3 * agg[r // 3] + r % 3, 3 * agg[ci[k] // 3] + ci[k] % 3
                A1[cr, cc] = np.float32(A1[cr, cc] + va[k])
        A1 = A1.astype(np.float64)
        S = S + weight * C @ floored_inv(A1) @ C.T
    for r in (rng.standard_normal(len(A)) for _ in range(3)):
        _, z = O.schwarz(rp, ci, va, domain, r, weight)
        want = S @ r
        assert np.abs(z - want).max() <= 1e-6 * np.abs(want).max(), np.abs(z - want).max() / np.abs(want).max()
    # the preconditioner is symmetric: u^T S v == v^T S u through the oracle's apply
    u, v = rng.standard_normal(len(A)), rng.standard_normal(len(A))
    _, su = O.schwarz(rp, ci, va, domain, u, weight)
    _, sv = O.schwarz(rp, ci, va, domain, v, weight)
    assert abs(v @ su - u @ sv) <= 1e-9 * abs(v @ su)


# ------------------------------------------------------------------ friction proximity weights
@pytest.mark.parametrize("cfg,steps", [(100, 30), (0, 16)])
def test_friction_weights_are_closest_point_coordinates(cfg, steps):
    s = nguo_hip.Scene(cfg, seed=cfg)
    o = O.OracleSolver(s)
    o.step(steps)
    # one inertial step on: the friction of a stencil is active once its vertices have moved
    x1 = (o.positions() + s.params.dt * o.velocities()).astype(np.float32)
    x = x1.astype(np.float64)
    o.freeze(x, o.x_tilde_d(), 1.0 / s.params.dt**2)
    f = o.frozen_stencils()
    side = o.collider_sides(x1)  # a vertex may have moved inside the body: signed stencils
    bidx, _, _, nvt = brute_stencils(s, x1, side)
    assert len(bidx) == len(f["idx"]) and np.array_equal(np.sort(bidx[nvt:], axis=0), np.sort(f["idx"][nvt:], axis=0))
    fr = np.nonzero(f["fric"])[0]
    assert np.any(fr < nvt) and np.any(fr >= nvt), "friction on both stencil kinds"
    for k in fr:
        w, i = f["w"][k], f["idx"][k].astype(np.int64)
        P = x[i]
        sep = w @ P  # p - q for V-T, (closest on a) - (closest on b) for E-E
        d = f["d"][k]
        if k < nvt:
            # (1, -b0, -b1, -b2): b the barycentric coordinates of a point of the triangle
            assert abs(w[0] - 1) < 1e-6 and np.all(-w[1:] > -1e-6) and abs(-w[1:].sum() - 1) < 1e-6
        else:
            # (1 - s, s, -(1 - t), -t): a point of each segment
            assert np.all(w[:2] > -1e-6) and np.all(-w[2:] > -1e-6)
            assert abs(w[:2].sum() - 1) < 1e-6 and abs(w[2:].sum() + 1) < 1e-6
        # the points they weight are |d| apart, so they are the closest points (unique for
        # non-parallel primitives), and the friction normal is their direction (reversed for
        # a signed gap behind a collider facet)
        assert abs(np.linalg.norm(sep) - abs(d)) < 2e-7, (k, np.linalg.norm(sep), d)
        # (FP32 weights place the points to ~2e-8: a short separation's direction blurs)
        assert np.abs(f["n"][k] - np.sign(d) * sep / np.linalg.norm(sep)).max() < 1e-3 + 3e-8 / abs(d)
