"""LBVH stages on the device, bit for bit against the oracle (engine-gpu/src/morton.rs,
radix.rs, karras.rs, bvh_aabb.rs semantics)."""
import numpy as np
import pytest

from conftest import nguo_hip
import oracle_py as O


def oracle_morton(pos, tris):
    lib = O.load()
    pos, tris = O.f32(pos), O.u32(tris)
    codes = np.zeros(len(tris), np.uint32)
    lib.oracle_morton_codes(len(pos), O.fp(pos), len(tris), O.up(tris), O.up(codes))
    return codes


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", [0, 1])
def test_morton_sort_karras_aabb_bitexact(ctx, cfg):
    s = nguo_hip.Scene(cfg, seed=cfg)
    pos, tris = np.array(s.positions), np.array(s.tris)
    codes = nguo_hip.lbvh_morton(ctx, pos, tris)
    ref = oracle_morton(pos, tris)
    assert np.array_equal(codes, ref)
    keys, vals = nguo_hip.lbvh_sort_pairs(ctx, codes, np.arange(len(codes), dtype=np.uint32))
    rk, rv = O.u32(ref).copy(), np.arange(len(ref), dtype=np.uint32)
    O.load().oracle_sort_pairs(len(rk), O.up(rk), O.up(rv))
    assert np.array_equal(keys, rk) and np.array_equal(vals, rv)  # stable order, duplicates included
    left, right = nguo_hip.lbvh_karras(ctx, keys)
    ol, orr = np.zeros(len(keys) - 1, np.uint32), np.zeros(len(keys) - 1, np.uint32)
    O.load().oracle_karras(len(keys), O.up(keys), O.up(ol), O.up(orr))
    assert np.array_equal(left, ol) and np.array_equal(right, orr)
    t = tris[vals]
    p = pos[t]
    leaf = np.concatenate([p.min(1), p.max(1)], axis=1).astype(np.float32)
    node = nguo_hip.lbvh_aabb(ctx, left, right, leaf)
    on = np.zeros((len(keys) - 1, 6), np.float32)
    O.load().oracle_propagate_aabb(len(keys), O.up(left), O.up(right), O.fp(leaf), O.fp(on))
    assert np.array_equal(node, on)


@pytest.mark.gpu
def test_sort_pairs_stability_and_edges(ctx):
    rng = np.random.default_rng(11)
    # 1 .. 70 001 use the 1024-item tile (70 001: the multi-workgroup counter scan); from 2^20
    # the 4096-item tile; heavy duplication and full-range keys, ragged last tiles
    for n, full in ((1, False), (2, False), (1023, False), (1024, False), (1025, False), (70001, False), (70001, True),
                    (1 << 20, True), (3000001, False), (3000001, True)):
        if full:
            keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        else:
            keys = rng.integers(0, 37, n).astype(np.uint32) * np.uint32(0x01010101)
        vals = np.arange(n, dtype=np.uint32)
        k, v = nguo_hip.lbvh_sort_pairs(ctx, keys, vals)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order]), (n, full)
