"""Scenes shared by the CPU and GPU tests beyond the generated configs: the analytic-collider
variant of the small sheet, and parameter overrides."""
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


def scene(cfg, **params):
    """Generated config `cfg` (seed = cfg) or "collider", with params overridden."""
    s = collider_scene() if cfg == "collider" else nguo_hip.Scene(cfg, seed=cfg)
    for k, v in params.items():
        setattr(s.desc.params, k, v)
    return s
