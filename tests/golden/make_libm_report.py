"""How far the libm boundary moves the step: the binary32 oracle with the Model
rebuild's sin/cos correctly rounded (the default, and what the HIP kernels
compute) against the same oracle calling the host libm's sinf/cosf (what a
glibc-linked build of the reference calls through glm::rotate<float>,
code/nans.cpp:1915-1920).  Everything else is identical, so the difference is
the libm choice alone.

Writes libm_boundary.json: per config, the largest |position difference| after
the given number of steps, the worlds it covers, and how many of the boxes'
final angles (up to 256 worlds) have a binary32 sin or cos that differs between
the two.  The numbers belong to the
host libm named in the file (glibc's version).  CPU only.
Usage: python tests/golden/make_libm_report.py [--procs 8]
"""
import argparse
import ctypes
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libm_boundary.json")


def _world_diff(args):
    """max |pos_libm - pos_default| and the same for velocity after nsteps."""
    scene, w, nsteps = args
    import oracle as O

    a = O.WorldState(scene, w, "f32")
    b = O.WorldState(scene, w, "f32")
    a.run(nsteps)
    b.run(nsteps, opt="libm")
    dp = float(np.max(np.abs(a.pos.astype(np.float64) - b.pos.astype(np.float64)))) if len(a.pos) else 0.0
    same = bool(np.array_equal(a.pos, b.pos) and np.array_equal(a.vel, b.vel))
    return w, dp, same, a.angles[: a.nb].copy()


def angle_ulps(angles_deg):
    """How many of the angles' binary32 sin or cos differ between the host libm
    and the correctly rounded value (the Model rebuild's radians conversion)."""
    x = (np.asarray(angles_deg, np.float32) * np.float32(0.0174532925199432957692369)).astype(np.float32)
    libm = ctypes.CDLL("libm.so.6")
    for f in (libm.sinf, libm.cosf):
        f.restype = ctypes.c_float
        f.argtypes = [ctypes.c_float]
    diff = 0
    for v in x.ravel():
        v = float(v)
        if (libm.sinf(v) != np.float32(np.sin(v))) or (libm.cosf(v) != np.float32(np.cos(v))):
            diff += 1
    return diff, int(x.size)


def glibc_version():
    f = ctypes.CDLL(None).gnu_get_libc_version
    f.restype = ctypes.c_char_p
    return f().decode()


def report(procs):
    from nans_projekat_amd import scenes as S

    plan = {
        "config1": (S.config1(), [0], 240),
        "config3": (S.config3(), list(range(4096)), 240),
        "config4": (S.config4(), list(range(1024)), 240),
        "config5_block": (S.config5(nx=12, ny=4, nz=12, nsteps=240), [0], 240),
    }
    res = {"libm": "glibc " + glibc_version(), "generator": "tests/golden/make_libm_report.py"}
    with mp.get_context("spawn").Pool(procs) as pool:
        for name, (sc, worlds, nsteps) in plan.items():
            t = time.time()
            rows = pool.map(_world_diff, [(sc.world(w), 0, nsteps) for w in worlds], chunksize=16)
            rows = [(w, r[1], r[2], r[3]) for w, r in zip(worlds, rows)]
            worst = max(rows, key=lambda r: r[1])
            res[name] = {"steps": nsteps, "worlds": len(worlds), "bodies": int(sum(
                         int(sc.offsets[w + 1]) - int(sc.offsets[w]) for w in worlds)),
                         "max_abs_dpos_m": worst[1], "worst_world": worst[0],
                         "worlds_bit_identical": int(sum(r[2] for r in rows)),
                         "worlds_above_1e-3_m": int(sum(r[1] > 1e-3 for r in rows))}
            # the boxes' angles after the run: how often the two sin/cos differ
            ang = np.concatenate([r[3] for r in rows[:256]]) if rows else np.zeros((0, 3))
            diff, total = angle_ulps(ang)
            res[name]["final_box_angles_with_a_libm_difference"] = {"angles": total, "differing": diff}
            print(name, res[name], "%.0f s" % (time.time() - t))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    res = report(a.procs)
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
