"""The first box-box contact of config 5 that launches a body (above 30 m/s), as a
small fixture: the state of its two boxes just before that step.  Runs the
binary32 oracle on all 8192 boxes with the HIP step's bounding-sphere cull
(results are the same as testing every pair) until a body first exceeds 30 m/s;
a few minutes on one core.  Writes config5_launch_pair.json.
Usage: python tests/golden/make_launch_pair.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config5_launch_pair.json")


def main():
    import oracle as O
    from nans_projekat_amd import scenes as S

    sc = S.config5(nsteps=1)
    ws = O.WorldState(sc, 0, "f32")
    O.counting(True, cull=True)
    for step in range(sc.nsteps if sc.nsteps > 1 else 480):
        before = {k: getattr(ws, k).copy() for k in ("pos", "vel", "angvel", "angles")}
        con = ws.run(1, contacts=True)
        sp = np.linalg.norm(ws.vel, axis=1)
        if sp.max() > 30.0:
            body = int(np.argmax(sp))
            cc = [c for c in con if c["type"] == O.CC and body in (int(c["index_a"]), int(c["index_b"]))]
            c = max(cc, key=lambda c: abs(float(c["dl_normal"])))
            pair = [int(c["index_a"]), int(c["index_b"])]
            res = {"generator": "tests/golden/make_launch_pair.py", "step": step, "launched_body": body,
                   "speed_after": float(sp[body]), "pair": pair, "dl_normal": float(c["dl_normal"])}
            for k, v in before.items():
                res[k] = v[pair].astype(float).tolist()
            for k in ("scale", "mass", "moi"):
                res[k] = np.asarray(getattr(sc, k))[pair].astype(float).tolist()
            with open(OUT, "w") as f:
                json.dump(res, f, indent=1)
            print("wrote", OUT, res["step"], pair)
            return
    raise SystemExit("no body exceeded 30 m/s")


if __name__ == "__main__":
    main()
