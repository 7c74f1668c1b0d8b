"""Golden final-state hashes of the full benchmark rollouts, from the CPU oracle.

  config 3: all 4096 worlds x 240 steps (worlds in parallel processes)
  config 5: the 8192-box world x 480 steps (one process; about half an hour)
The hash is SHA-256 over the binary32 bytes of pos, vel, angvel and angles
([N][3] each, worlds concatenated in order), so a GPU run of the same rollout
that is bit-exact with the oracle reproduces it.  Output: state_hashes.json.
Usage: python tests/golden/make_state_hashes.py [--procs 8] [--only config3|config5]
"""
import argparse
import hashlib
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "state_hashes.json")


def state_hash(pos, vel, angvel, angles):
    h = hashlib.sha256()
    for a in (pos, vel, angvel, angles):
        h.update(np.ascontiguousarray(a, dtype=np.float32).tobytes())
    return h.hexdigest()


def _c3_world(w):
    import oracle as O
    from nans_projekat_amd import scenes as S

    sc = S.config3()
    ws = O.WorldState(sc, w, "f32")
    ws.run(sc.nsteps)
    return w, ws.pos, ws.vel, ws.angvel, ws.angles


def config3(procs):
    from nans_projekat_amd import scenes as S

    sc = S.config3()
    parts = {}
    with mp.get_context("spawn").Pool(procs) as pool:
        for w, p, v, av, an in pool.imap_unordered(_c3_world, range(sc.nworlds), chunksize=8):
            parts[w] = (p, v, av, an)
    cat = lambda i: np.concatenate([parts[w][i] for w in range(sc.nworlds)])
    return {"steps": sc.nsteps, "bodies": sc.nbodies, "sha256": state_hash(cat(0), cat(1), cat(2), cat(3))}


def config5():
    import oracle as O
    from nans_projekat_amd import scenes as S

    sc = S.config5()
    ws = O.WorldState(sc, 0, "f32")
    ws.run(sc.nsteps)
    return {"steps": sc.nsteps, "bodies": sc.nbodies, "sha256": state_hash(ws.pos, ws.vel, ws.angvel, ws.angles)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--only", choices=("config3", "config5"))
    a = ap.parse_args()
    new = {}
    t = time.time()
    if a.only in (None, "config3"):
        new["config3_240_steps"] = config3(a.procs)
        print("config3 done in %.0f s" % (time.time() - t), flush=True)
    if a.only in (None, "config5"):
        new["config5_480_steps"] = config5()
        print("config5 done in %.0f s" % (time.time() - t), flush=True)
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    res.update(new)
    res["generator"] = "tests/golden/make_state_hashes.py"
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
