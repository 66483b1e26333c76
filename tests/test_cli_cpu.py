"""The mguso command-line runner's process contract on the CPU (no device needed): exit 2 and
a usage message on stderr for bad arguments (clap's behaviour, engine-cli/src/main.rs:62-91),
exit 1 with one ERROR log line on stderr and nothing on stdout for unreadable, malformed or
unsupported input and for a missing device (core-engine/docs/PROCESS_CONTRACT.md:10-27)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MGUSO = os.path.join(ROOT, "nguo-3d_amd", "bin", "mguso")
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")


def mguso(*args, env=None):
    assert os.path.exists(MGUSO), "build first: make -C nguo-3d_amd"
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([MGUSO, *args], capture_output=True, text=True, timeout=120, env=e)


def scene_file(tmp_path, text):
    p = tmp_path / "scene.json"
    p.write_text(text)
    return str(p)


def assert_failure(r, needle):
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert r.stdout == ""
    lines = [ln for ln in r.stderr.splitlines() if ln.strip()]
    assert len(lines) == 1 and "ERROR" in lines[0] and needle in lines[0], r.stderr


@pytest.mark.parametrize("args", [[], ["frobnicate"], ["bench"], ["bench", "a.json", "--steps", "x"],
                                  ["run", "a.json", "--steps", "3"], ["verify", "a.json", "--precond", "block"],
                                  ["bench", "a.json", "--solver", "gpu"], ["--log-format", "xml", "gpu-info"]])
def test_usage_errors_exit_2(args):
    r = mguso(*args)
    assert r.returncode == 2 and r.stdout == "" and r.stderr.startswith("error: ") and "Usage: mguso" in r.stderr


def test_help_and_version_exit_0():
    for a in ("--help", "-h"):
        r = mguso(a)
        assert r.returncode == 0 and "Commands:" in r.stdout and r.stderr == ""
    r = mguso("--version")
    assert r.returncode == 0 and r.stdout.startswith("mguso ")


def test_unreadable_scene_exits_1(tmp_path):
    assert_failure(mguso("bench", str(tmp_path / "missing.json")), "reading scene")


@pytest.mark.parametrize("text,needle", [
    ("{", "parsing scene JSON"),
    ('{"config": {}}', "missing field `objects`"),
    ('{"objects": [{"name": "a", "positions": [[0, 0, 0]]}]}', "missing field `tris`"),
    ('{"objects": [{"name": "a", "positions": [[0, 0]], "tris": []}]}', "expected 3 elements"),
    ('{"objects": [{"name": "a", "positions": [[0, 0, 0]], "tris": [[0, 0, 1.0]]}]}', "expected an integer"),
    ('{"objects": [{"name": "a", "positions": [[0, 0, 0]], "tris": [[0, 0, 5]]}]}', "out of range"),
    ('{"config": {"dt": "small"}, "objects": []}', "`config.dt`"),
    ('{"objects": [], "colliders": [{"type": "cone"}]}', "unknown variant `cone`"),
])
def test_malformed_scene_exits_1(tmp_path, text, needle):
    assert_failure(mguso("bench", scene_file(tmp_path, text)), needle)


def test_unsupported_inputs_exit_1(tmp_path):
    # a fabric preset selects the membrane model; XPBD and the CPU solvers are not in this build
    assert_failure(mguso("bench", os.path.join(SCENES, "fabric_drape.json")), "Baraff-Witkin membrane")
    sphere = os.path.join(SCENES, "sphere_drape.json")
    assert_failure(mguso("run", sphere), "XPBD preview tier")  # run defaults to --solver xpbd
    assert_failure(mguso("bench", sphere, "--cpu"), "--cpu")
    assert_failure(mguso("golden", str(tmp_path)), "capture corpus")
    text = json.dumps({"objects": [{"name": "c", "positions": [[0, 0, 0], [1, 0, 0], [0, 0, 1]], "tris": [[0, 1, 2]],
                                    "material": {"strain-limit": 0.1}}]})
    assert_failure(mguso("bench", scene_file(tmp_path, text)), "strain limiting")


@pytest.mark.parametrize("fmt", ["text", "json"])
def test_no_device_exits_1_with_one_log_line(fmt):
    # this container has no GPU: a valid scene fails at bring-up, logged in the chosen format
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present")
    r = mguso("--log-format", fmt, "bench", os.path.join(SCENES, "sphere_drape.json"), "--steps", "1")
    assert_failure(r, "GPU bring-up")
    if fmt == "json":
        rec = json.loads(r.stderr.strip())
        assert rec["level"] == "ERROR" and rec["target"] == "mguso" and "GPU bring-up" in rec["fields"]["error"]
