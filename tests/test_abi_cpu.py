"""The C ABI library loads without a GPU, exports every entry point include/nguo_hip.h
declares, and its host-only parts (scene generators, argument checks) behave. CPU only."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, nguo_hip


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nguo_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|const char \*)\s*\**\s*(nguo_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = nguo_hip.load()
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(nguo_hip.EXPORTED) <= set(syms)


def test_no_device_is_reported_not_faked():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(nguo_hip.NguoError) as e:
        nguo_hip.Context(0)
    assert e.value.code in (nguo_hip.NGUO_E_DEVICE, nguo_hip.NGUO_E_INVALID)


@pytest.mark.parametrize("cfg,verts,tris", [(0, 1024 + 2562, 1922 + 5120), (100, 100 + 162, 162 + 320)])
def test_sheet_scene_sizes(cfg, verts, tris):
    s = nguo_hip.Scene(cfg, seed=cfg)
    assert s.n_verts == verts and len(s.tris) == tris
    assert (s.inv_mass > 0).sum() == (1024 if cfg == 0 else 100)
    assert np.all(s.static[s.inv_mass > 0] == 0) and np.all(s.static[s.inv_mass == 0] == 1)
    free = s.inv_mass > 0
    assert np.abs(s.positions[free, 1] - 0.5).max() <= 1e-3 + 1e-7  # +-1 mm jitter


def test_tube_scene_sizes_and_determinism():
    a = nguo_hip.Scene(1, seed=1)
    b = nguo_hip.Scene(1, seed=1)
    assert a.n_verts == 19950 + 3 * 201 + 12 * 2114
    assert len(a.tris) == 39600 + 12 * 4224
    assert a.desc.n_rods == 600 and a.desc.n_rod_hinges == 597
    assert np.array_equal(a.positions, b.positions)
    c = nguo_hip.Scene(1, seed=2)
    assert not np.array_equal(a.positions, c.positions)


def test_graded_sizes_are_uneven():
    counts = [int((nguo_hip.Scene(3, seed=3, param=f).inv_mass > 0).sum()) for f in (0.85, 0.9, 0.95, 1.0, 1.05, 1.1, 1.15, 1.2)]
    assert counts == sorted(counts) and counts[-1] > 2.5 * counts[0]


def test_unknown_scene_is_an_error():
    with pytest.raises(nguo_hip.NguoError) as e:
        nguo_hip.Scene(42)
    assert e.value.code == nguo_hip.NGUO_E_INVALID
    assert "unknown config" in str(e.value)
