"""Error of the binary32 step against the binary64 oracle, per config.

The HIP kernels equal the binary32 oracle bit for bit on every config (the
parity tests under tests/), in both arithmetic modes, so the error of the
binary32 oracle against the binary64 oracle is the error of the GPU step.  This
script measures it on the CPU and writes fp64_error.json:
  config2       1M narrowphase pairs: GJK intersect/separate disagreements;
                contact (EPA) disagreements, all and outside the 1e-5 touching
                band (the reporting side's depth above 1e-5); the angle between
                the contact normals and the
                relative depth error where both report a contact; in the
                reference arithmetic and in the FP64-islands mode;
  config1/3/4/5 body positions after 240 steps (config 5: a 12 x 4 x 12 block of
                its boxes, the full world being too slow for the binary64 CPU
                oracle): max, p99 and median of |pos32 - pos64| over the bodies,
                and the worlds whose largest difference exceeds 1e-3 m.
Usage: python tests/golden/make_fp64_report.py [--procs 8]
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fp64_error.json")
BAND = 1e-5


def _narrow(args):
    lo, hi, islands = args
    import oracle as O
    from nans_projekat_amd import scenes as S

    pairs = {k: v[lo:hi] for k, v in S.config2_pairs().items()}
    a = O.narrowphase(pairs, S.StepParams(fp64_islands=islands), "f32")
    b = O.narrowphase(pairs, S.StepParams(), "f64")
    keep = ("gjk", "contact", "normal", "depth")
    return {k: a[k] for k in keep}, {k: b[k] for k in keep}


def narrow_report(procs, islands):
    chunks = [(lo, lo + 62500, islands) for lo in range(0, 1_000_000, 62500)]
    with mp.get_context("spawn").Pool(procs) as pool:
        parts = pool.map(_narrow, chunks)
    cat = lambda side, k: np.concatenate([p[side][k] for p in parts])
    c32, c64 = cat(0, "contact"), cat(1, "contact")
    g32, g64 = cat(0, "gjk") == 1, cat(1, "gjk") == 1
    d32, d64 = cat(0, "depth"), cat(1, "depth")
    n32, n64 = cat(0, "normal").astype(np.float64), cat(1, "normal")
    dis = c32 != c64
    depth_of_reporting = np.where(c32 == 1, np.abs(d32), np.abs(d64))
    both = (c32 == 1) & (c64 == 1)
    u = n32[both] / np.linalg.norm(n32[both], axis=1, keepdims=True)
    v = n64[both] / np.linalg.norm(n64[both], axis=1, keepdims=True)
    ang = np.arccos(np.clip(np.sum(u * v, axis=1), -1.0, 1.0))
    rel = np.abs(d32[both] - d64[both]) / np.maximum(np.abs(d64[both]), 1e-3)
    q = lambda x, p: float(np.quantile(x, p)) if x.size else 0.0
    return {"pairs": int(c32.size), "gjk_intersecting_f32": int(g32.sum()), "gjk_intersecting_f64": int(g64.sum()),
            "gjk_classification_disagreements": int((g32 != g64).sum()),
            "contacts_f32": int(c32.sum()), "contacts_f64": int(c64.sum()),
            "classification_disagreements": int(dis.sum()),
            "disagreements_outside_1e-5_band": int((dis & (depth_of_reporting > BAND)).sum()),
            "normal_angle_rad": {"p50": q(ang, 0.5), "p99": q(ang, 0.99), "max": float(ang.max()) if ang.size else 0.0},
            "depth_rel_error": {"p50": q(rel, 0.5), "p99": q(rel, 0.99), "max": float(rel.max()) if rel.size else 0.0,
                                "note": "|d32 - d64| / max(|d64|, 1e-3)"}}


def _world(args):
    scene, nsteps = args
    import oracle as O

    a = O.WorldState(scene, 0, "f32")
    b = O.WorldState(scene, 0, "f64")
    a.run(nsteps)
    b.run(nsteps)
    return np.abs(a.pos.astype(np.float64) - b.pos).max(axis=1) if len(a.pos) else np.zeros(0)


def position_report(procs, scene, worlds, nsteps):
    with mp.get_context("spawn").Pool(procs) as pool:
        rows = pool.map(_world, [(scene.world(w), nsteps) for w in worlds], chunksize=8)
    allb = np.concatenate(rows)
    worst = np.array([r.max() if r.size else 0.0 for r in rows])
    return {"steps": nsteps, "worlds": len(worlds), "bodies": int(allb.size),
            "max_abs_dpos_m": float(allb.max()), "p99_abs_dpos_m": float(np.quantile(allb, 0.99)),
            "median_abs_dpos_m": float(np.median(allb)),
            "worlds_above_1e-3_m": int((worst > 1e-3).sum())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    from nans_projekat_amd import scenes as S

    res = {"generator": "tests/golden/make_fp64_report.py",
           "meaning": "binary32 step (HIP kernels = binary32 oracle bit for bit) against the binary64 oracle"}
    t = time.time()
    res["config2"] = narrow_report(a.procs, 0)
    res["config2_fp64_islands"] = narrow_report(a.procs, 1)
    print("config2 done %.0f s" % (time.time() - t), flush=True)
    plan = {"config1": (S.config1(), [0]),
            "config3": (S.config3(), list(range(4096))),
            "config4": (S.config4(), list(range(1024))),
            "config5_block": (S.config5(nx=12, ny=4, nz=12, nsteps=240), [0])}
    for name, (sc, worlds) in plan.items():
        res[name] = position_report(a.procs, sc, worlds, 240)
        print(name, res[name], "%.0f s" % (time.time() - t), flush=True)
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
