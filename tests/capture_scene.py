"""A synthetic scene in the reference capture's format (engine-io/src/golden.rs: input_bin/*.bin,
input_bin/param/*.bin, param.toml, info.toml, checkpoints/vert_<N>.bin, little-endian), written
from numpy, and its lowering onto the step restated here independently of mguso's C++
(engine-cli/src/gauge_driver.rs:318-1200 rules: lumped mass from the densities, isotropic and
Baraff-Witkin faces as shells, rod stretch k = young * density / length_factor with rest length
l * length_factor, rod bend stiffness = mean segment bend x lumped vertex mass at vertices of
two segments and no face, pins, walls and spheres), so the tests can run the oracle on the scene
`mguso gauge-scene` should have built. FP32 arithmetic in the reference's order where it
computes in f32."""
import os

import numpy as np

from conftest import nguo_hip

F32 = np.float32


def sheet_spec(n=10, seed=3):
    """A cloth sheet of n x n vertices 6 cm above a wall (y = -0.03) with a sphere under its
    middle, two corners pinned, alternate faces ARAP and Baraff-Witkin, per-face densities, and
    a free 8-segment rod beside it; vertices split into two displacement groups."""
    rng = np.random.default_rng(seed)
    h = F32(0.3 / (n - 1))
    g = np.array([[i * h, F32(0.0), j * h] for i in range(n) for j in range(n)], np.float64)
    tris = []
    for i in range(n - 1):
        for j in range(n - 1):
            v = i * n + j
            tris += [[v, v + 1, v + n], [v + n, v + 1, v + n + 1]]
    tris = np.array(tris, np.int64)
    nf = len(tris)
    rod0 = n * n
    rod = np.array([[0.33 + 0.02 * k, 0.01, 0.15] for k in range(9)], np.float64)
    template = np.concatenate([g, rod])
    nv = len(template)
    dmap = np.zeros(nv, np.uint32)
    dmap[rod0:] = 1
    disp = np.array([[0.0, 0.06, 0.0], [0.0, 0.05, 0.0]], np.float64)
    vel = np.zeros((nv, 3), np.float32)
    vel[:, 1] = -0.2
    vel[rod0:, 0] = 0.1
    rods = np.array([[rod0 + k, rod0 + k + 1] for k in range(8)], np.int64)
    return {
        "template": template, "dmap": dmap, "disp": disp, "vel": vel, "tris": tris, "rods": rods,
        "tri_model": np.array([0 if f % 2 == 0 else 2 for f in range(nf)], np.uint8),
        "tri_density": (0.2 + 0.05 * rng.random(nf)).astype(np.float32),
        "tri_young": np.full(nf, 2.0e3, np.float32), "tri_poiss": np.full(nf, 0.35, np.float32),
        "tri_bend": np.full(nf, 1.0e-4, np.float32), "tri_gap": np.full(nf, 4.0e-3, np.float32),
        "tri_friction": np.full(nf, 0.3, np.float32),
        "rod_young": np.full(8, 50.0, np.float32), "rod_density": np.full(8, 0.01, np.float32),
        "rod_lf": np.full(8, 0.97, np.float32), "rod_bend": np.full(8, 20.0, np.float32),
        "pins": [np.array([0, n - 1], np.uint64)],
        "walls": [{"pos": [[0.0, -0.03, 0.0]], "n": [0.0, 1.0, 0.0]}],
        "spheres": [{"pos": [[0.15, -0.046, 0.15]], "r": [0.1]}],
        "param": {"dt": 0.005, "fps": 100.0, "frames": 4, "gravity": "[0.0, -9.8, 0.0]", "barrier": '"cubic"',
                  "cg_tol": 1e-10, "friction_eps": 1e-5, "eiganalysis_eps": 1e-6, "max_dx": 1.0},
    }


def write_capture(root, spec, checkpoints=None):
    """Writes the capture files of spec under root; checkpoints maps frame -> positions [nv, 3]."""
    ib = os.path.join(root, "input_bin")
    pp = os.path.join(ib, "param")
    os.makedirs(pp, exist_ok=True)
    w = lambda name, arr, dt: open(os.path.join(ib, name), "wb").write(np.ascontiguousarray(arr, dt).tobytes())
    w("vert.bin", spec["template"], "<f8")
    if "dmap" in spec:
        w("vert_dmap.bin", spec["dmap"], "<u4")
        w("displacement.bin", spec["disp"], "<f8")
    w("vel.bin", spec["vel"], "<f4")
    if len(spec["tris"]):
        w("tri.bin", spec["tris"], "<u8")
    if len(spec.get("rods", [])):
        w("rod.bin", spec["rods"], "<u8")
    wp = lambda name, arr: open(os.path.join(pp, name), "wb").write(np.ascontiguousarray(arr).tobytes())
    for key, name in (("tri_density", "tri-density.bin"), ("tri_young", "tri-young-mod.bin"), ("tri_poiss", "tri-poiss-rat.bin"),
                      ("tri_bend", "tri-bend.bin"), ("tri_gap", "tri-contact-gap.bin"), ("tri_friction", "tri-friction.bin"),
                      ("rod_young", "rod-young-mod.bin"), ("rod_density", "rod-density.bin"), ("rod_lf", "rod-length-factor.bin"),
                      ("rod_bend", "rod-bend.bin")):
        if key in spec:
            wp(name, np.asarray(spec[key], "<f4"))
    if "tri_model" in spec:
        wp("tri-model.bin", np.asarray(spec["tri_model"], np.uint8))
    info = []
    for b, idx in enumerate(spec.get("pins", [])):
        w(f"pin-ind-{b}.bin", idx, "<u8")
        info += [f"[pin-{b}]", "operation_count = 0", "pull = 0.0", ""]
    for i, c in enumerate(spec.get("walls", [])):
        w(f"wall-pos-{i}.bin", np.array(c["pos"], np.float64), "<f8")
        info += [f"[wall-{i}]", f"nx = {c['n'][0]}", f"ny = {c['n'][1]}", f"nz = {c['n'][2]}", ""]
    for i, c in enumerate(spec.get("spheres", [])):
        w(f"sphere-pos-{i}.bin", np.array(c["pos"], np.float64), "<f8")
        w(f"sphere-radius-{i}.bin", np.array(c["r"], np.float32), "<f4")
        info += [f"[sphere-{i}]", ""]
    open(os.path.join(root, "info.toml"), "w").write("\n".join(info) + "\n")
    if "param" in spec:
        open(os.path.join(root, "param.toml"), "w").write("[param]\n" + "".join(f"{k} = {v}\n" for k, v in spec["param"].items()))
    if checkpoints:
        os.makedirs(os.path.join(root, "checkpoints"), exist_ok=True)
        for k, x in checkpoints.items():
            open(os.path.join(root, "checkpoints", f"vert_{k}.bin"), "wb").write(np.asarray(x, "<f4").tobytes())


def initial_positions(spec):
    x = spec["template"].astype(np.float32)
    if "dmap" in spec:
        x = x + spec["disp"].astype(np.float32)[spec["dmap"]]
    return x.astype(np.float32)


def _seq_sum(a):
    s = F32(0.0)
    for v in np.asarray(a, np.float32):
        s = F32(s + v)
    return s


def _len32(d):
    return F32(np.sqrt(F32(F32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])))


def lower(spec):
    """The ArrayScene gauge-scene should build from spec (no tets, no strain limits)."""
    x = initial_positions(spec)
    nv = len(x)
    tris = spec["tris"]
    mass = np.zeros(nv, np.float32)
    for f, (a, b, c) in enumerate(tris):
        e1, e2 = (x[b] - x[a]).astype(np.float32), (x[c] - x[a]).astype(np.float32)
        cr = np.array([F32(e1[1] * e2[2] - e1[2] * e2[1]), F32(e1[2] * e2[0] - e1[0] * e2[2]), F32(e1[0] * e2[1] - e1[1] * e2[0])], np.float32)
        area = F32(F32(0.5) * _len32(cr))
        m = F32(F32(spec["tri_density"][f] * area) / F32(3.0))
        for v in (a, b, c):
            mass[v] = F32(mass[v] + m)
    rods = spec.get("rods", np.zeros((0, 2), np.int64))
    for r, (a, b) in enumerate(rods):
        m = F32(F32(spec["rod_density"][r] * _len32((x[b] - x[a]).astype(np.float32))) / F32(2.0))
        mass[a] = F32(mass[a] + m)
        mass[b] = F32(mass[b] + m)
    lumped = mass.copy()
    pos = mass[mass > 0]
    fallback = F32(_seq_sum(pos) / F32(len(pos))) if len(pos) else F32(1.0)
    mass = np.where(mass > 0, mass, fallback).astype(np.float32)
    inv = (F32(1.0) / mass).astype(np.float32)
    for blk in spec.get("pins", []):
        inv[blk.astype(np.int64)] = 0.0
    shell = tris[np.isin(spec["tri_model"], [0, 1, 2])]
    # rods: rest l * lf, k = young * density / lf; hinges at vertices of two segments and no face
    rv, rr, rk = [], [], []
    for r, (a, b) in enumerate(rods):
        lf = spec["rod_lf"][r]
        rv.append([a, b])
        rr.append(max(F32(_len32((x[b] - x[a]).astype(np.float32)) * lf), F32(1e-9)))
        rk.append(F32(F32(spec["rod_young"][r] * spec["rod_density"][r]) / max(lf, F32(1e-9))))
    has_face = np.zeros(nv, bool)
    has_face[tris.reshape(-1)] = True
    nbr, deg, bend_sum = {}, np.zeros(nv, np.int64), np.zeros(nv, np.float32)
    for r, (a, b) in enumerate(rods):
        for v, o in ((a, b), (b, a)):
            nbr.setdefault(v, []).append(o)
            deg[v] += 1
        bend_sum[a] = F32(bend_sum[a] + spec["rod_bend"][r])
        bend_sum[b] = F32(bend_sum[b] + spec["rod_bend"][r])
    hv, hr, hs = [], [], []
    for v in range(nv):
        if deg[v] == 2 and not has_face[v]:
            hv.append([nbr[v][0], v, nbr[v][1]])
            hr.append(F32(np.pi))
            hs.append(F32(F32(F32(0.5) * bend_sum[v]) * lumped[v]))
    # material: mean young, representative poisson -> Lame (material.rs:31-44)
    young = F32(_seq_sum(spec["tri_young"]) / F32(len(spec["tri_young"])))
    nu = F32(min(max(F32(_seq_sum(spec["tri_poiss"]) / F32(len(spec["tri_poiss"]))), F32(1e-6)), F32(0.5 - 1e-4)))
    mu = F32(young / F32(F32(2.0) * F32(F32(1.0) + nu)))
    lam = F32(F32(young * nu) / F32(F32(F32(1.0) + nu) * F32(F32(1.0) - F32(F32(2.0) * nu))))
    mean = lambda a: F32(_seq_sum(a) / F32(len(a)))
    t = spec["param"]
    p = nguo_hip.Params()
    p.dt = t["dt"]
    for k, gk in enumerate([float(v) for v in t["gravity"].strip("[]").split(",")]):
        p.gravity[k] = gk
    p.newton_iters = 32
    p.bend_stiffness = mean(spec["tri_bend"])
    p.shell_mu, p.shell_lambda, p.shell_model = mu, lam, 1
    p.eiganalysis_eps = t.get("eiganalysis_eps", 1e-2)
    p.contact_dhat = mean(spec["tri_gap"])
    p.contact_kappa = 100.0
    p.friction = mean(spec["tri_friction"])
    p.friction_eps = t.get("friction_eps", 1e-5)
    p.barrier_kind = 1 if t.get("barrier") == '"cubic"' else 0
    p.self_collision = 1
    p.max_dx = t.get("max_dx", 1.0)
    p.target_toi = t.get("target_toi", 0.25)
    p.min_newton_steps = 0
    p.ccd_max_iter = 4096
    p.cg_tol = t.get("cg_tol", 1e-3)
    p.cg_max_iter = 10000
    p.floor_y = float("-inf")
    p.precond, p.schwarz_domain, p.schwarz_weight = 1, 32, 0.0
    colliders = []
    for c in spec.get("walls", []):
        k = nguo_hip.Collider()
        k.kind = 0
        k.a[:] = c["pos"][0]
        nn = np.asarray(c["n"], np.float32)
        k.b[:] = nn / _len32(nn)
        colliders.append(k)
    for c in spec.get("spheres", []):
        k = nguo_hip.Collider()
        k.kind = 1
        k.a[:] = c["pos"][0]
        k.radius = c["r"][0]
        colliders.append(k)
    return nguo_hip.ArrayScene(x, inv, tris, p, colliders=colliders, velocities=spec["vel"], shell_faces=shell,
                               rods=(np.array(rv), np.array(rr, np.float32), np.array(rk, np.float32)),
                               rod_hinges=(np.array(hv), np.array(hr, np.float32), np.array(hs, np.float32)))


def steps_to_frame(spec, frame):
    """Steps the runner takes to reach frame / fps: the clock adds the f32 dt in f64."""
    t, n, target = 0.0, 0, frame / spec["param"]["fps"]
    while t < target:
        t += float(F32(spec["param"]["dt"]))
        n += 1
    return n
