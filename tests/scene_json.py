"""The reference's scene-file lowering (engine-api/src/scene.rs:89-229 lower and
build_implicit, params.rs:41-122 SolverConfig defaults, material.rs:9-73 Material defaults)
restated in Python over numpy, independently of the mguso CLI's C++ lowering, so the tests can
run the oracle on the scene the CLI should have built. FP32 arithmetic where the reference
computes in f32 (total area, lumped mass)."""
import json

import numpy as np

from conftest import nguo_hip

CONFIG = {"dt": 1.0e-3, "fps": 60.0, "frames": 300, "gravity": [0.0, -9.8, 0.0], "cg-max-iter": 10000,
          "cg-tol": 1.0e-3, "friction-eps": 1.0e-5, "iterations": 4, "self-collision": False, "floor-y": None}
MATERIAL = {"young-mod": 1000.0, "density": 1.0, "bend": 10.0, "contact-gap": 1.0e-3, "friction": 0.0,
            "membrane-warp": 0.0, "membrane-weft": 0.0, "membrane-shear": 0.0, "strain-limit": 0.0,
            "strain-limit-weft": 0.0, "contact-dynamic": 0.0}


def config_of(scene):
    c = dict(CONFIG)
    c.update(scene.get("config", {}))
    return c


def steps_per_frame(c):
    return max(1, int(round((1.0 / c["fps"]) / float(np.float32(c["dt"])))))


def lower(path, precond=0):
    """(ArrayScene, config) of a scene file; precond 0 = block Jacobi, 1 = schwarz-ml."""
    scene = json.load(open(path))
    c = config_of(scene)
    pos, tris, pin, static = [], [], [], []
    material = dict(MATERIAL)
    for o in scene["objects"]:
        base = len(pos)
        st = bool(o.get("is_static", False))
        pos += [list(map(np.float32, p)) for p in o["positions"]]
        pin += [st] * len(o["positions"])
        static += [st] * len(o["positions"])
        tris += [[base + i for i in t] for t in o["tris"]]
        if not st:
            if o.get("preset") is not None:
                raise ValueError("preset: membrane model")
            material = dict(MATERIAL)
            material.update(o.get("material", {}))
        for p in o.get("pins", []):
            if base + p < len(pin):
                pin[base + p] = True
    x = np.array(pos, np.float32).reshape(-1, 3)
    t = np.array(tris, np.int64).reshape(-1, 3)
    area = np.float32(0.0)
    for a, b, cc in t:
        n = np.cross(x[b] - x[a], x[cc] - x[a]).astype(np.float32)
        area = np.float32(area + np.float32(0.5) * np.float32(np.sqrt(np.float32(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]))))
    nv = np.float32(max(len(x), 1))
    mass = max(np.float32(np.float32(area * np.float32(material["density"])) / nv), np.float32(1e-6))
    inv = np.full(len(x), np.float32(1.0) / mass, np.float32)
    inv[np.array(pin, bool)] = 0.0
    p = nguo_hip.Params()
    p.dt = c["dt"]
    for k in range(3):
        p.gravity[k] = c["gravity"][k]
    p.newton_iters = max(int(c["iterations"]), 1)
    p.stretch_k = max(material["young-mod"], 1.0)
    p.bend_stiffness = material["bend"]
    p.shell_model = 1  # ARAP, inert without shell faces
    p.eiganalysis_eps = 1e-6
    p.contact_dhat = max(np.float32(material["contact-gap"]), np.float32(1e-3))
    p.contact_kappa = max(material["young-mod"], 1.0)
    p.friction = material["friction"]
    p.friction_eps = c["friction-eps"]
    p.barrier_kind = 0  # log
    p.self_collision = 1 if c["self-collision"] else 0
    p.max_dx = float("inf")
    p.ccd_max_iter = 256
    p.cg_tol = c["cg-tol"]
    p.cg_max_iter = int(c["cg-max-iter"])
    p.floor_y = float("-inf") if c["floor-y"] is None else c["floor-y"]
    p.precond = precond
    p.schwarz_domain = 16 if precond == 1 else 32
    p.schwarz_weight = 0.7 if precond == 1 else 0.0
    colliders = []
    for cs in scene.get("colliders", []):
        k = nguo_hip.Collider()
        if cs["type"] == "plane":
            k.kind = 0
            k.a[:] = cs["point"]
            k.b[:] = cs["normal"]
        else:
            k.kind = 1
            k.a[:] = cs["center"]
            k.radius = cs["radius"]
        colliders.append(k)
    st = np.array(static, np.uint8) if any(static) else None
    return nguo_hip.ArrayScene(x, inv, t, p, static=st, colliders=colliders), c


def read_obj(path):
    """Positions and triangles of an OBJ the CLI wrote (v / f records, 1-based)."""
    v, f = [], []
    for line in open(path):
        w = line.split()
        if not w:
            continue
        if w[0] == "v":
            v.append([np.float32(a) for a in w[1:4]])
        elif w[0] == "f":
            f.append([int(a.split("/")[0]) - 1 for a in w[1:4]])
    return np.array(v, np.float32), np.array(f, np.int64)
