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


# --- the SimUpdateAndRender adapter (libnans_sim.so) under the platform loop ---

def _sim():
    if not os.path.exists(build.SIM):
        build.build_driver()
    return build.SIM


def test_sim_library_exports_the_plugin_symbol():
    import ctypes
    lib = ctypes.CDLL(_sim())
    assert hasattr(lib, "SimUpdateAndRender") and hasattr(lib, "nans_sim_release")


def test_sim_input_script_is_checked(tmp_path):
    path = str(tmp_path / "in.txt")
    with open(path, "w") as f:
        f.write("3 jump 1\n")
    res = subprocess.run([_driver(), "--sim", _sim(), "--frames", "1", "--input", path], capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 1 and "bad input line" in res.stderr


F32 = np.float32


def _front(yaw, pitch):
    """The camera's Front from Yaw/Pitch in binary32 (cos/sin rounded from binary64)."""
    rad = lambda d: F32(F32(d) * F32(0.017453292519943295))
    c = lambda x: F32(np.cos(np.float64(x)))
    s = lambda x: F32(np.sin(np.float64(x)))
    f = np.array([c(rad(yaw)) * c(rad(pitch)), s(rad(pitch)), s(rad(yaw)) * c(rad(pitch))], F32)
    d = F32(F32(f[0] * f[0] + f[1] * f[1]) + f[2] * f[2])
    return f * F32(F32(1.0) / F32(np.sqrt(d)))


@pytest.mark.gpu
def test_sim_platform_loop_matches_oracle(tmp_path, oracle_mod):
    """150 frames of the reference's Init scene through SimUpdateAndRender, with the
    library reloaded every 7 frames and input that nudges Cubes[3], resets the cubes,
    walks the camera and shoots the sphere twice; every 10th frame equals the oracle
    stepped with the same edits."""
    script = str(tmp_path / "in.txt")
    with open(script, "w") as f:
        f.write("20 debug_left 1\n30 debug_left 0\n40 debug_up 1\n43 debug_up 0\n60 debug_reset 1\n61 debug_reset 0\n"
                "90 shoot 1\n91 shoot 0\n100 forward 1\n105 forward 0\n120 shoot 1\n121 shoot 0\n")
    traj = str(tmp_path / "sim.bin")
    from nans_projekat_amd import runtime
    dt = F32(1.0) / F32(60.0)
    res = subprocess.run([_driver(), "--sim", _sim(), "--lib", runtime.LIB_PATH, "--frames", "150", "--dt",
                          "%.9g" % dt, "--reload-every", "7", "--every", "10", "--input", script, "--out", traj],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    assert "150 frames, 5 bodies, 21 reloads, 1 uploads" in res.stdout
    t = S.load_trajectory(traj)
    assert list(t["steps"]) == list(range(10, 151, 10))

    sc = S.demo_scene()
    ws = oracle_mod.WorldState(sc, 0, "f32")
    cam = np.array([0.0, 0.0, 3.0], F32)
    front = np.array([0.0, 0.0, -1.0], F32)
    speed = F32(0.05)
    held = set()
    events = {20: ("debug_left", 1), 30: ("debug_left", 0), 40: ("debug_up", 1), 43: ("debug_up", 0),
              60: ("debug_reset", 1), 61: ("debug_reset", 0), 90: ("shoot", 1), 91: ("shoot", 0),
              100: ("forward", 1), 105: ("forward", 0), 120: ("shoot", 1), 121: ("shoot", 0)}
    k = 0
    for frame in range(150):
        if frame in events:
            name, down = events[frame]
            (held.add if down else held.discard)(name)
        if "forward" in held:
            cam = (cam + speed * front).astype(F32)
        if "shoot" in held:
            s = ws.nb
            ws.pos[s] = cam
            ws.vel[s] = 0.0
            ws.angvel[s] = 0.0
            ws.torque[s] = 0.0
            ws.force[s] = (F32(0.0) + F32(4001.0) * front).astype(F32)
        for name, axis in (("debug_left", 0), ("debug_up", 1)):
            if name in held:
                ws.pos[3, axis] = F32(ws.pos[3, axis] + dt * F32(1.0))
        if "debug_reset" in held:
            ws.pos[:4] = np.array([(2, 3.5, 2), (2, 1, 2), (2, 4.5, 2), (0, 0, 0)], F32)
            ws.vel[:4] = 0.0
            ws.angvel[:4] = 0.0
        if cam[1] < 0.5:
            cam[1] = F32(0.5)
        ws.run(1)
        front = _front(-90.0, 0.0)
        if (frame + 1) % 10 == 0:
            np.testing.assert_array_equal(t["pos"][k], ws.pos, err_msg="frame %d" % (frame + 1))
            np.testing.assert_array_equal(t["angles"][k], ws.angles, err_msg="frame %d" % (frame + 1))
            k += 1
    assert k == 15
