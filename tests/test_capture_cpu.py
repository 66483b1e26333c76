"""mguso gauge-scene on synthetic captures, on the CPU (loading, lowering and reporting need no
device; --compare-frame 0 compares the loaded state without stepping). The reader's contract
follows engine-io/src/golden.rs's own tests (decodes_synthetic_buffers,
u64_index_decoders_refuse_to_truncate, rejects_misaligned_and_oversized,
reads_pin_list_from_synthetic, parses_param_toml_from_text, compare_identical_is_zero_and_passes,
compare_known_offset, loaded_positions_match_the_reference_frame_zero) and the report's lines
follow engine-cli/src/main.rs:264-360."""
import re

import numpy as np
import pytest

import capture_scene as CS
from test_cli_cpu import mguso


def report(r):
    assert r.returncode == 0, r.stderr
    return r.stdout


def write(tmp_path, spec, **kw):
    d = str(tmp_path / "scene")
    CS.write_capture(d, spec, **kw)
    return d


def test_report_counts_and_gaps(tmp_path):
    spec = CS.sheet_spec()
    out = report(mguso("gauge-scene", write(tmp_path, spec)))
    nf = len(spec["tris"])
    assert f"verts {len(spec['template'])}  tris {nf}  tets 0  rods 8 (bend hinges 7)" in out
    assert f"model split: baraff-witkin {nf // 2}  isotropic-shell {nf - nf // 2}" in out
    assert "pins 2 (0 kinematic)  pulls 0  colliders 2 (0 animated)" in out
    assert "NOT reproduced by this build:" in out
    assert f"{nf // 2} faces use the Baraff-Witkin membrane" in out
    assert "2 pins are held fixed" in out
    assert "per-hinge bending coefficients" in out
    # uniform arrays are reproduced exactly, so no averaging is reported for them
    assert "young-mod varies" not in out and "contact-gap varies" not in out and "friction varies" not in out


def test_param_toml_mapping(tmp_path):
    # parses_param_toml_from_text: [param] keys, quoted barrier, bracketed gravity; absent keys keep
    # the reference defaults
    spec = CS.sheet_spec()
    spec["param"] = {"dt": 0.001, "fps": 60.0, "frames": 200, "gravity": "[0.0, -9.8, 0.5]", "world_scaling": 2.0,
                     "barrier": '"cubic"'}
    out = report(mguso("gauge-scene", write(tmp_path, spec)))
    m = re.search(r"solver: dt (\S+)  gravity \[(\S+), (\S+), (\S+)\].*dhat (\S+)  friction (\S+)  barrier (\S+)  cg_tol (\S+)  newton cap (\d+)", out)
    assert m, out
    assert float(m.group(1)) == np.float32(0.001)
    assert [float(m.group(k)) for k in (2, 3, 4)] == [0.0, np.float32(-9.8), 0.5]
    # scene-wide values are the f32 running-sum means (gauge_driver.rs:254-261 mean_or)
    mean = lambda a: np.float32(CS._seq_sum(a) / np.float32(len(a)))
    assert float(m.group(5)) == mean(spec["tri_gap"]) and float(m.group(6)) == mean(spec["tri_friction"])
    assert m.group(7) == "cubic" and float(m.group(8)) == np.float32(1e-3) and m.group(9) == "32"
    assert "world_scaling != 1" in out


def test_varying_arrays_are_reported(tmp_path):
    spec = CS.sheet_spec()
    spec["tri_young"] = np.linspace(1e3, 3e3, len(spec["tris"])).astype(np.float32)
    spec["tri_gap"] = np.linspace(1e-3, 5e-3, len(spec["tris"])).astype(np.float32)
    out = report(mguso("gauge-scene", write(tmp_path, spec)))
    assert "per-face young-mod varies" in out and "per-face contact-gap varies" in out


@pytest.mark.parametrize("case,needle", [
    ("vert_misaligned", "not a multiple of 24"),
    ("tri_oversized", "exceeds u32::MAX"),
    ("pin_misaligned", "not a whole number of u64"),
    ("vel_count", "vel count"),
    ("tri_range", "references a vertex"),
    ("no_vert", "reading golden binary"),
])
def test_malformed_capture_fails(tmp_path, case, needle):
    spec = CS.sheet_spec()
    d = write(tmp_path, spec)
    ib = tmp_path / "scene" / "input_bin"
    if case == "vert_misaligned":
        (ib / "vert.bin").write_bytes(bytes(23))
    elif case == "tri_oversized":
        # u64_index_decoders_refuse_to_truncate: 0x1_0000_0005 would narrow to vertex 5
        t = spec["tris"].astype(np.uint64).copy()
        t[3, 1] = 0x100000005
        (ib / "tri.bin").write_bytes(t.astype("<u8").tobytes())
    elif case == "pin_misaligned":
        (ib / "pin-ind-1.bin").write_bytes(bytes(12))
    elif case == "vel_count":
        (ib / "vel.bin").write_bytes(spec["vel"][:-1].astype("<f4").tobytes())
    elif case == "tri_range":
        t = spec["tris"].copy()
        t[0, 0] = len(spec["template"])
        (ib / "tri.bin").write_bytes(t.astype("<u8").tobytes())
    else:
        (ib / "vert.bin").unlink()
    r = mguso("gauge-scene", d)
    assert r.returncode == 1 and r.stdout == "", (r.returncode, r.stdout)
    assert needle in r.stderr and "ERROR" in r.stderr, r.stderr


def gates(out):
    m = re.search(r"divergence vs reference: max_rel (\S+)  rms (\S+)  centroid_rel (\S+)  \(diagonal (\S+)\)", out)
    assert m, out
    ok = re.search(r"gates \(max_rel < 8e-2, rms < 2e-2, centroid_rel < 6e-3\): (PASS|FAIL)", out)
    return [float(m.group(k)) for k in (1, 2, 3, 4)], ok.group(1)


def test_frame_zero_is_the_displaced_template(tmp_path):
    # loaded_positions_match_the_reference_frame_zero: vert.bin + displacement[vert_dmap]
    # reproduces checkpoint 0; without the displacement the sheet and rod would sit 5-6 cm lower
    spec = CS.sheet_spec()
    x0 = CS.initial_positions(spec)
    out = report(mguso("gauge-scene", write(tmp_path, spec, checkpoints={0: x0}), "--compare-frame", "0"))
    (mx, rms, cen, diag), verdict = gates(out)
    assert mx == 0.0 and rms == 0.0 and cen == 0.0 and verdict == "PASS"
    assert "in 0 steps" in out


def test_compare_known_offset(tmp_path):
    # compare_known_offset: the unit triangle shifted by 0.1 in x: diagonal sqrt(2), every
    # distance 0.1, so max_rel = rms = centroid_rel = 0.1 / sqrt(2): max_rel passes, rms fails
    spec = {"template": np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64), "vel": np.zeros((3, 3), np.float32),
            "tris": np.array([[0, 1, 2]], np.int64), "tri_model": np.zeros(1, np.uint8), "tri_density": np.ones(1, np.float32),
            "param": {"dt": 0.01, "fps": 60.0}}
    ref = CS.initial_positions(spec) + np.array([0.1, 0, 0], np.float32)
    out = report(mguso("gauge-scene", write(tmp_path, spec, checkpoints={0: ref}), "--compare-frame", "0"))
    (mx, rms, cen, diag), verdict = gates(out)
    want = 0.1 / np.sqrt(2.0)
    assert abs(diag - np.sqrt(2.0)) < 1e-4
    for v in (mx, rms, cen):
        assert abs(v - want) < 1e-4
    assert verdict == "FAIL"


def test_compare_frame_needs_param_toml(tmp_path):
    spec = CS.sheet_spec()
    del spec["param"]
    r = mguso("gauge-scene", write(tmp_path, spec), "--compare-frame", "1")
    assert r.returncode == 1 and "param.toml is required" in r.stderr, r.stderr


def test_compare_frame_is_gauge_only():
    r = mguso("run", "a.json", "--compare-frame", "1")
    assert r.returncode == 2 and "--compare-frame" in r.stderr
