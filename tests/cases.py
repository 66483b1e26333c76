"""Scenes and checks shared by the CPU and GPU tests beyond the generated configs: the
analytic-collider variant of the small sheet, parameter overrides, a sheet state with vertices
moved inside the body, and the device-vs-oracle stencil comparison."""
import ctypes as C

import numpy as np

from conftest import nguo_hip


def collider_scene():
    """The 10 x 10 sheet falling on an analytic sphere in place of the mesh icosphere (the
    analytic-collider branch, implicit.rs:2954-3021; contact.rs:1082-1125), self contact off."""
    s = nguo_hip.Scene(100, seed=0)
    st = s.static.astype(bool)
    ctr = s.positions[st].astype(np.float64).mean(axis=0)
    r = float(np.linalg.norm(s.positions[st] - ctr, axis=1).mean())
    arr = (nguo_hip.Collider * 1)()
    arr[0].kind = 1
    arr[0].a[:] = [float(c) for c in ctr]
    arr[0].radius = r
    s._colliders = arr  # keeps the array alive with the scene
    s.desc.colliders = C.cast(arr, C.POINTER(nguo_hip.Collider))
    s.desc.n_colliders = 1
    s.desc.params.self_collision = 0
    return s


def spring_scene():
    """The 10 x 10 sheet on the icosphere with edge springs (k = 800 over every unique
    triangle edge, implicit.rs:2487-2507) in place of the shells: the mass-spring element set
    of the reference's scene files (scene.rs:165-229 build_implicit)."""
    s = nguo_hip.Scene(100, seed=0)
    p = s.desc.params
    p.stretch_k = 800.0
    p.shell_mu = 0.0
    p.shell_lambda = 0.0
    return s


def scene(cfg, **params):
    """Generated config `cfg` (seed = cfg), "collider" or "springs", with params overridden."""
    s = collider_scene() if cfg == "collider" else spring_scene() if cfg == "springs" else nguo_hip.Scene(cfg, seed=cfg)
    for k, v in params.items():
        setattr(s.desc.params, k, v)
    return s


def inside_state(steps=30, shallow=0.4, deep=2.5):
    """The 10 x 10 sheet resting on the icosphere after `steps` oracle steps, with two free
    vertices moved inside the body: one `shallow` * dhat behind the centroid of a facet it
    touches (within dhat of the surface, so its side is -1 and its gaps to collider facets are
    signed), one `deep` * dhat behind another facet (deeper than dhat: side +1, unsigned).
    Returns (scene, x, v, shallow vertex, deep vertex, their facets)."""
    import oracle_py as O

    s = nguo_hip.Scene(100, seed=0)
    o = O.OracleSolver(s)
    o.step(steps)
    x, v = o.positions().copy(), o.velocities()
    static = s.static.astype(bool)
    tris = s.tris.astype(np.int64)
    facets = np.nonzero(static[tris].all(axis=1))[0]
    free = np.nonzero(s.inv_mass > 0)[0]
    # each free vertex's nearest collider facet, by its centroid distance
    cen = x[tris[facets]].astype(np.float64).mean(axis=1)
    dist = np.linalg.norm(x[free].astype(np.float64)[:, None, :] - cen[None, :, :], axis=2)
    near = dist.argmin(axis=1)
    order = np.argsort(dist[np.arange(len(free)), near])
    centre = s.positions[static].astype(np.float64).mean(axis=0)
    dhat = float(s.params.contact_dhat)
    picked = []
    for k in order:
        vert, tri = int(free[k]), tuple(int(a) for a in tris[facets[near[k]]])
        if any(vert == p[0] or tri == p[1] or set(tri) & set(p[1]) for p in picked):
            continue
        picked.append((vert, tri))
        if len(picked) == 2:
            break
    for (vert, tri), depth in zip(picked, (shallow, deep)):
        a, b, c = (x[k].astype(np.float64) for k in tri)
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        if n @ ((a + b + c) / 3 - centre) < 0:
            n = -n
        x[vert] = ((a + b + c) / 3 - depth * dhat * n).astype(np.float32)
    return s, x, v, picked[0][0], picked[1][0], picked[0][1], picked[1][1]


def check_stencils(ctx, s, o, x, v):
    """The device's sorted stencil list, gaps, gradients, Rayleigh stiffnesses and collider
    sides at x against the oracle's, bitwise. Returns the oracle's (idx, d)."""
    idx, d, g = o.stencils(x)
    assert len(idx) > 0, "the state must be in contact"
    st = nguo_hip.Stepper(ctx, s)
    st.set_state(x, v)
    gidx, gd, gg, gk = st.stencils(x)
    assert np.array_equal(gidx, idx)
    assert np.array_equal(gd, d)  # same FP32 operation order: bitwise equal gaps
    assert np.array_equal(gg, g)
    assert np.all(np.isfinite(gk)) and np.all(gk > 0)
    # Rayleigh stiffness against the elastic blocks at x (implicit.rs:755-795): the blocks are
    # summed in the reference's push order on both sides, so the stiffnesses agree bitwise
    ok = o.kappas(x, 1.0 / (s.params.dt * s.params.dt))[: len(idx)]
    assert np.array_equal(gk, ok)
    side = o.collider_sides(x)
    assert np.array_equal(st.collider_sides(x), side)
    st.close()
    return idx, d


def pinned_state(steps=14, npin=3):
    """The spring sheet resting on the icosphere (`steps` oracle steps of spring_scene), rebuilt
    from that state with the `npin` free vertices nearest the body pinned: springs and hinges
    then join pinned and free vertices, and contact stencils include pinned cloth vertices.
    Returns (scene, x, v, pinned vertex ids)."""
    import oracle_py as O

    s0 = spring_scene()
    o = O.OracleSolver(s0)
    o.step(steps)
    x, v = o.positions(), o.velocities()
    static = s0.static.astype(bool)
    idx, d, _ = o.stencils(x)
    free = s0.inv_mass > 0
    near = []
    for r, dd in sorted(zip(idx.tolist(), d.tolist()), key=lambda t: t[1]):
        for a in r:
            if free[a] and a not in near:
                near.append(a)
    pins = np.array(near[:npin], np.int64)
    assert len(pins) == npin, "the sheet must touch the body"
    im = s0.inv_mass.copy()
    im[pins] = 0.0
    v = v.copy()
    v[pins] = 0.0
    cols = [s0.desc.colliders[k] for k in range(s0.desc.n_colliders)]
    s = nguo_hip.ArrayScene(x, im, s0.tris, s0.desc.params, static=static.astype(np.uint8), colliders=cols)
    return s, x, v, pins


def max_stretch_ratio(x, rest_pos, tris):
    """implicit.rs:2193-2209 max_stretch_ratio over the mesh's unique edges (state.rs:39-65),
    in FP32 with glam's length order ((x*x + y*y) + z*z): the max_sigma channel."""
    t = np.asarray(tris, dtype=np.int64)
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    e = np.unique(e, axis=0)
    if len(e) == 0:
        return np.float32(0.0)

    def length(p):
        d = p[e[:, 0]].astype(np.float32) - p[e[:, 1]].astype(np.float32)
        return np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])

    rest, cur = length(rest_pos), length(x)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(rest > np.float32(1e-9), cur / rest, np.float32(1.0)).astype(np.float32)
    r = r[~np.isnan(r)]
    return np.float32(r.max()) if len(r) else np.float32(0.0)
