"""Contact queries and the CCD line search on trees of one to a few leaves (the frontier of a
one-leaf tree is the leaf itself; trees smaller than the frontier leave parts empty): a cloth
patch of 1 or 2 triangles falling fast onto a static patch of 1 or 2 triangles. Stencils
bitwise and the step (CCD-limited) against the oracle."""
import ctypes as C

import numpy as np
import pytest

from conftest import nguo_hip
import oracle_py as O


class TinyScene:
    """A scene description built from numpy arrays (kept alive with the descriptor)."""

    def __init__(self, n_cloth_tris, n_static_tris, gap=4e-4, speed=0.5, slide=0.0):
        quad = lambda z0, s: np.array([[0, 0, 0], [s, 0, 0], [0, 0, s], [s, 0, s]], np.float32) + np.float32([0, z0, 0])
        cloth = quad(gap, 0.1)[: 3 if n_cloth_tris == 1 else 4] + np.float32([0.02, 0.0, 0.02])
        body = quad(0.0, 0.5)[: 3 if n_static_tris == 1 else 4] - np.float32([0.1, 0.0, 0.1])
        self.positions = np.ascontiguousarray(np.concatenate([cloth, body]), np.float32)
        nc, nb = len(cloth), len(body)
        self.n_verts = nc + nb
        self.inv_mass = np.concatenate([np.full(nc, 1.0 / 2e-4), np.zeros(nb)]).astype(np.float32)
        self.static = np.concatenate([np.zeros(nc), np.ones(nb)]).astype(np.uint8)
        vel = np.zeros_like(self.positions)
        vel[:nc, 1] = -speed
        vel[:nc, 0] = slide
        self.velocities = vel
        tri = lambda k, o: [[o + 0, o + 1, o + 2], [o + 2, o + 1, o + 3]][:k]
        shell = np.array(tri(n_cloth_tris, 0), np.uint32)
        self.tris = np.ascontiguousarray(np.concatenate([shell, np.array(tri(n_static_tris, nc), np.uint32)]), np.uint32)
        self.shell = np.ascontiguousarray(shell, np.uint32)
        ref = nguo_hip.Scene(100, seed=0)
        d = nguo_hip.SceneDesc()
        d.n_verts = self.n_verts
        d.positions = self.positions.ctypes.data_as(C.POINTER(C.c_float))
        d.velocities = self.velocities.ctypes.data_as(C.POINTER(C.c_float))
        d.inv_mass = self.inv_mass.ctypes.data_as(C.POINTER(C.c_float))
        d.n_tris = len(self.tris)
        d.tris = self.tris.ctypes.data_as(C.POINTER(C.c_uint32))
        d.static_collider = self.static.ctypes.data_as(C.POINTER(C.c_uint8))
        d.n_shell_faces = len(self.shell)
        d.shell_faces = self.shell.ctypes.data_as(C.POINTER(C.c_uint32))
        d.params = ref.desc.params
        d.params.newton_iters = 4
        d.params.cg_tol = 1e-10
        self.desc = d
        self.params = d.params


@pytest.mark.gpu
@pytest.mark.parametrize("nc,ns", [(1, 1), (2, 1), (1, 2), (2, 2)])
def test_tiny_trees_stencils(ctx, nc, ns):
    sc = TinyScene(nc, ns, gap=4e-4)
    o = O.OracleSolver(sc)
    st = nguo_hip.Stepper(ctx, sc)
    x = sc.positions
    idx, d, g = o.stencils(x)
    assert len(idx) > 0, "the patches start within dhat"
    gidx, gd, gg, _ = st.stencils(x)
    assert np.array_equal(gidx, idx) and np.array_equal(gd, d) and np.array_equal(gg, g)
    st.close()


@pytest.mark.gpu
@pytest.mark.parametrize("nc,ns", [(1, 1), (2, 1), (1, 2), (2, 2)])
def test_tiny_trees_ccd_limited_step(ctx, nc, ns):
    # 2 mm above the static patch (outside dhat) at 0.5 m/s: the first Newton direction
    # crosses it and the CCD line search has to cut the step
    sc = TinyScene(nc, ns, gap=2e-3)
    o, o64 = O.OracleSolver(sc), O.OracleSolver(sc, fp64=True)
    st = nguo_hip.Stepper(ctx, sc)
    o.step(1)
    o64.step(1)
    st.advance(1)
    m = st.metrics()
    _, tois = st.newton_trace(positions=False)
    assert min(tois) < 1.0, "the fall is limited by the CCD step"
    spread = np.abs(o.positions() - o64.positions()).max()
    err = np.abs(st.positions() - o.positions()).max()
    print(f"nc={nc} ns={ns} toi={min(tois):.4f} err={err:.2e} spread={spread:.2e}")
    # against the FP32 mirror at the step bar (the FP64 twin drifts from both on the
    # two-triangle patches, where the cloth rests across the patch's diagonal edge)
    assert err <= 1e-5
    assert m.penetration_free == 1 and m.all_finite == 1
    st.close()


@pytest.mark.gpu
def test_sliding_contact_long_advancement(ctx):
    # sliding fast along the static patch while closing slowly: the advancement moves by
    # (d - target) / u_max with u_max dominated by the slide, so pairs need many iterations
    # (the ones past the inline budget go to the per-pair pass); same step as the oracle
    sc = TinyScene(2, 2, gap=1.5e-3, speed=0.3, slide=3.0)
    o = O.OracleSolver(sc)
    st = nguo_hip.Stepper(ctx, sc)
    o.step(1)
    st.advance(1)
    _, tois = st.newton_trace(positions=False)
    _, otois = o.newton_trace()
    err = np.abs(st.positions() - o.positions()).max()
    print(f"sliding: toi device {np.round(tois, 5)} oracle {np.round(otois, 5)} err={err:.2e}")
    assert min(tois) < 1.0
    assert np.allclose(tois, otois, rtol=1e-5, atol=0)
    assert err <= 1e-5
    st.close()
