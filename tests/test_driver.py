"""The headless driver (nans_projekat_amd/driver/nans_headless.cpp), the
stand-in for the reference's platform loop (code/sdl_nans.cpp:395-415,
913-1000): scene files round-trip exactly, and on the GPU its trajectory equals
the library's own stepping and the oracle."""
import os
import subprocess

import numpy as np
import pytest

from nans_projekat_amd import build
from nans_projekat_amd import scenes as S


def _driver():
    if not os.path.exists(build.DRIVER):
        build.build_driver()
    return build.DRIVER


def _checksum(sc):
    # the driver's order: each array row-major, accumulated in binary64
    s = 0.0
    for k in ("pos", "vel", "angvel", "angles", "scale", "mass", "moi"):
        for x in np.asarray(getattr(sc, k), np.float32).reshape(-1):
            s += float(x)
    return s


@pytest.mark.parametrize("name", ["config1", "config4"])
def test_scene_file_round_trip(tmp_path, name):
    sc = S.config1() if name == "config1" else S.config4(8)
    path = str(tmp_path / "scene.txt")
    S.save_scene(sc, path)
    out = subprocess.run([_driver(), "--scene", path, "--steps", "0", "--parse-only"], capture_output=True, text=True,
                         check=True).stdout.split()
    assert out[1] == str(sc.nworlds) and out[3] == str(sc.nbodies)
    assert int(out[5]) == sc.params.solver_iterations
    assert float(out[9]) == _checksum(sc)


def test_bad_scene_is_rejected(tmp_path):
    path = str(tmp_path / "bad.txt")
    with open(path, "w") as f:
        f.write("nans-scene 1\nparams 1 2 3\n")
    res = subprocess.run([_driver(), "--scene", path, "--steps", "0", "--parse-only"], capture_output=True, text=True)
    assert res.returncode == 1 and "bad params line" in res.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_driver_trajectory_matches_oracle(tmp_path, oracle_mod, mode):
    sc = S.config1()
    scene, traj = str(tmp_path / "c1.txt"), str(tmp_path / "c1.bin")
    S.save_scene(sc, scene)
    from nans_projekat_amd import runtime
    res = subprocess.run([_driver(), "--lib", runtime.LIB_PATH, "--scene", scene, "--steps", "120", "--every", "40",
                          "--mode", mode, "--out", traj], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    t = S.load_trajectory(traj)
    assert list(t["steps"]) == [0, 40, 80, 120]
    np.testing.assert_array_equal(t["pos"][0], sc.pos)
    ws = oracle_mod.WorldState(sc, 0, "f32")
    for k in range(1, 4):
        ws.run(40)
        np.testing.assert_array_equal(t["pos"][k], ws.pos)
        np.testing.assert_array_equal(t["angles"][k], ws.angles)
