"""mguso gauge-scene on the device against the oracle: a synthetic capture (a pinned sheet of ARAP
and Baraff-Witkin faces with per-face densities settling onto a sphere over a wall, cubic
barrier, and a free rod with bending; two displacement groups) whose checkpoint at frame N is the
oracle's step of the same scene lowered independently (capture_scene.lower), so
--compare-frame N measures the reader, the mapping and the step together
(engine-cli/src/main.rs:264-450)."""
import re

import numpy as np
import pytest

import capture_scene as CS
import oracle_py as O
from test_capture_cpu import gates, write
from test_cli_cpu import mguso


@pytest.mark.gpu
def test_compare_frame_matches_the_oracle(tmp_path):
    spec = CS.sheet_spec()
    frame = 2
    n = CS.steps_to_frame(spec, frame)
    o = O.OracleSolver(CS.lower(spec))
    o.step(n)
    ref = o.positions()
    x0 = CS.initial_positions(spec)
    d = write(tmp_path, spec, checkpoints={0: x0, frame: ref})
    r = mguso("gauge-scene", d, "--compare-frame", str(frame))
    assert r.returncode == 0, r.stderr
    print(r.stdout)
    (mx, rms, cen, diag), verdict = gates(r.stdout)
    assert f"in {n} steps" in r.stdout
    # the sheet has settled onto the sphere (contact active from step 3): positions to FP32
    # trajectory agreement, far inside the gates
    assert verdict == "PASS"
    assert mx * diag <= 2e-6, (mx, diag)
    skill = re.search(r"SKILL vs null .*: max (\S+)  rms (\S+)", r.stdout)
    assert float(skill.group(1)) > 0.999 and float(skill.group(2)) > 0.999


@pytest.mark.gpu
def test_steps_report(tmp_path):
    spec = CS.sheet_spec()
    o = O.OracleSolver(CS.lower(spec))
    o.step(3)
    r = mguso("gauge-scene", write(tmp_path, spec), "--steps", "3")
    assert r.returncode == 0, r.stderr
    m = re.search(r"stepped 3 step\(s\) to t = (\S+): all finite, min y = (\S+)", r.stdout)
    assert m, r.stdout
    assert abs(float(m.group(1)) - 3 * np.float32(0.005)) < 1e-6
    assert abs(float(m.group(2)) - o.positions()[:, 1].min()) <= 2e-6
    assert re.search(r"max per step: newton \d+  cg \d+  contacts \d+", r.stdout)
