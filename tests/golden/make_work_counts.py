"""Freeze the algorithmic work of the benchmark workloads into work_counts.json.

Runs the CPU oracle with its event counters on (oracle/nans_oracle_impl.h) over
  config 3: all 4096 worlds x 240 steps, with the HIP step's bounding-sphere cull,
            so the events are those of the work the HIP step performs; totals
            and the counts of each step;
  config 2: the 1M narrowphase pairs (no cull: the narrowphase entry point has none).
bench.py turns the events into algorithmic flops with its per-event weights.
The counts do not depend on the machine: the oracle and the scene generator are
deterministic.  Usage: python tests/golden/make_work_counts.py [--procs 8]
"""
import argparse
import json
import multiprocessing as mp
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "work_counts.json")


def _c3_chunk(args):
    """Event counts of worlds [lo, hi), step by step: {event: [count of step t]}."""
    lo, hi, nsteps = args
    import oracle as O
    from nans_projekat_amd import scenes as S

    sc = S.config3()
    worlds = [O.WorldState(sc, w, "f32") for w in range(lo, hi)]
    O.counting(True, cull=True)
    per_step = {k: [] for k in O.COUNT_NAMES}
    prev = {k: 0 for k in O.COUNT_NAMES}
    for _ in range(nsteps):
        for ws in worlds:
            ws.run(1)
        c = O.counts()
        for k in O.COUNT_NAMES:
            per_step[k].append(c[k] - prev[k])
        prev = c
    O.counting(False)
    per_step["body_steps"] = [int(sc.offsets[hi]) - int(sc.offsets[lo])] * nsteps
    return per_step


def config3_counts(procs: int, nsteps: int = 240):
    """Totals over the run, and the same events per step (the bench times a window
    of steps from the initial state and prices exactly those steps)."""
    from nans_projekat_amd import scenes as S

    nw = S.config3().nworlds
    step = max(1, nw // (procs * 8))
    chunks = [(lo, min(nw, lo + step), nsteps) for lo in range(0, nw, step)]
    with mp.get_context("spawn").Pool(procs) as pool:
        parts = pool.map(_c3_chunk, chunks)
    per_step = {}
    for p in parts:
        for k, v in p.items():
            acc = per_step.setdefault(k, [0] * nsteps)
            for t in range(nsteps):
                acc[t] += v[t]
    total = {k: sum(v) for k, v in per_step.items()}
    total["steps"] = nsteps
    total["worlds"] = nw
    return total, per_step


def config2_counts():
    import oracle as O
    from nans_projekat_amd import scenes as S

    pairs = S.config2_pairs()
    O.counting(True, cull=False)
    O.narrowphase(pairs, S.StepParams(), "f32")
    c = O.counts()
    O.counting(False)
    c["n_pairs"] = len(pairs["type_a"])
    return c


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    t = time.time()
    c3, c3_steps = config3_counts(a.procs)
    res = {"config3_batched_240_steps": c3, "config3_batched_per_step": c3_steps,
           "config2_narrowphase": config2_counts()}
    res["generator"] = "tests/golden/make_work_counts.py"
    text = json.dumps(res, indent=1, sort_keys=True)
    # one line per per-step list
    text = re.sub(r"\[\s*(-?\d+(?:,\s*-?\d+)*)\s*\]", lambda m: "[" + ", ".join(m.group(1).split()) .replace(",,", ",") + "]",
                  text)
    with open(OUT, "w") as f:
        f.write(text)
    print("wrote %s in %.0f s" % (OUT, time.time() - t))


if __name__ == "__main__":
    main()
