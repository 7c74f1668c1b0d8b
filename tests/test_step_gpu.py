"""Parity of the HIP world step (plain and batched) with the CPU oracle's step
(code/nans.cpp:1758-1762 plus the Model rebuild): per-step ordered contact
lists and body state bit for bit, positions within 1e-3 m after 240 steps, and
the error against the FP64 oracle reported."""
import numpy as np
import torch
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S

from golden_util import load_state_hash, state_hash

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "angvel", "angles")


def _ctx(scene):
    ctx = runtime.Context(scene.params)
    ctx.upload(scene)
    ctx.record_contacts(True)
    return ctx


def _contacts_equal(got, ref, where):
    assert len(got) == len(ref), (where, len(got), len(ref))
    for f in ("type", "index_a", "index_b"):
        np.testing.assert_array_equal(got[f], ref[f], err_msg="%s %s" % (where, f))
    for f in ("point_a", "point_b", "normal", "dl_normal", "dl_tangent1", "dl_tangent2"):
        np.testing.assert_array_equal(got[f], ref[f].astype(np.float32), err_msg="%s %s" % (where, f))


def _state_equal(got, ws, a, b, where):
    for f in STATE:
        np.testing.assert_array_equal(got[f][a:b], getattr(ws, f), err_msg="%s %s" % (where, f))


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_config1_stepwise_bit_exact(oracle_mod, path):
    """Config 1 (27-cube stack + sphere), 240 steps: every step's contact list and
    the state after every step match the binary32 oracle exactly."""
    sc = S.config1()
    ctx = _ctx(sc)
    ws = oracle_mod.WorldState(sc, 0, "f32")
    step = ctx.step_plain if path == "plain" else ctx.step_batched
    for k in range(240):
        step(1)
        ref = ws.run(1, contacts=True)
        _contacts_equal(ctx.contacts(0), ref, "step %d" % k)
        _state_equal(ctx.download(), ws, 0, sc.nbodies, "step %d" % k)
    ctx.close()


def test_config1_600_steps_and_fp64_error(oracle_mod):
    """The full config-1 run: batched in one call (hipGraph of 600 steps) against
    the oracle; the deviation from the FP64 oracle is reported."""
    sc = S.config1()
    ctx = _ctx(sc)
    ctx.step_batched(sc.nsteps)
    got = ctx.download()
    ws = oracle_mod.WorldState(sc, 0, "f32")
    ws.run(sc.nsteps)
    _state_equal(got, ws, 0, sc.nbodies, "600 steps")
    w64 = oracle_mod.WorldState(sc, 0, "f64")
    w64.run(240)
    ctx2 = _ctx(sc)
    ctx2.step_batched(240)
    g2 = ctx2.download()
    err = np.abs(g2["pos"] - w64.pos).max()
    print("config 1, 240 steps: max |pos - pos_fp64| = %.3e m" % err)
    ctx.close()
    ctx2.close()


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_demo_scene_bit_exact(oracle_mod, path):
    """The reference's own Init scene (non-uniform box Cubes[3], r=0.25 sphere) with
    the reference's constants (70 iterations, Cr=Cf=0.1, dt=1/60)."""
    sc = S.demo_scene()
    ctx = _ctx(sc)
    ws = oracle_mod.WorldState(sc, 0, "f32")
    step = ctx.step_plain if path == "plain" else ctx.step_batched
    for k in range(120):
        step(1)
        ref = ws.run(1, contacts=True)
        _contacts_equal(ctx.contacts(0), ref, "step %d" % k)
    _state_equal(ctx.download(), ws, 0, sc.nbodies, "end")
    ctx.close()


def _c3_subset():
    full = S.config3(256, nsteps=1)
    sizes = np.diff(full.offsets)
    pick = [int(np.argmax(sizes)), int(np.argmin(sizes))] + list(range(0, 256, 19))
    return full.subset(sorted(set(pick)))


@pytest.mark.parametrize("path", ["plain", "batched", "batched_waves"])
def test_config3_worlds_bit_exact(oracle_mod, path):
    """A batch of config-3 worlds (8..256 bodies, 1-4 spheres) for 40 steps:
    every world's state and last contact list equal the oracle's (batched_waves:
    the worlds of at most 64 bodies one per wave)."""
    sc = _c3_subset()
    ctx = _ctx(sc)
    if path == "batched_waves":
        ctx.set_wave_worlds(64)
    (ctx.step_plain if path == "plain" else ctx.step_batched)(40)
    got = ctx.download()
    for w in range(sc.nworlds):
        ws = oracle_mod.WorldState(sc, w, "f32")
        ws.run(39)
        ref = ws.run(1, contacts=True)
        a, b = int(sc.offsets[w]), int(sc.offsets[w + 1])
        _state_equal(got, ws, a, b, "world %d" % w)
        _contacts_equal(ctx.contacts(w), ref, "world %d" % w)
    st = ctx.stats()
    assert st["epa_overflow"] == 0 and st["capacity_overflow"] == 0
    ctx.close()


def test_per_world_coefficients(oracle_mod):
    """Config-4 worlds carry their own friction and restitution."""
    sc = S.config4(12, nsteps=1)
    for path in ("plain", "batched"):
        ctx = _ctx(sc)
        (ctx.step_plain if path == "plain" else ctx.step_batched)(30)
        got = ctx.download()
        for w in range(sc.nworlds):
            ws = oracle_mod.WorldState(sc, w, "f32")
            ws.run(30)
            a, b = int(sc.offsets[w]), int(sc.offsets[w + 1])
            _state_equal(got, ws, a, b, "%s world %d" % (path, w))
        ctx.close()


def test_plain_equals_batched_graph_replay():
    """The batched graph replayed across calls equals the plain path."""
    sc = _c3_subset()
    a, b = _ctx(sc), _ctx(sc)
    for _ in range(3):
        a.step_batched(5)
    b.step_plain(15)
    ga, gb = a.download(), b.download()
    for f in STATE:
        np.testing.assert_array_equal(ga[f], gb[f])
    a.close()
    b.close()


# (threads per world, wave-world bound): every world by a workgroup (bound 0, the
# default -1), bounds that split the subset elsewhere, and (128, 256): every world
# by a wave
SPLITS = ((64, -1), (128, 0), (128, 64), (128, 256), (256, -1), (256, 16), (256, 32), (256, 64), (256, 128),
          (512, 0), (512, 64))


def test_workgroup_sizes_and_autotuner():
    """Every workgroup size and wave-world bound of the batched step gives the same
    bits; the autotuner leaves the state as it found it and its choice steps the
    same."""
    sc = _c3_subset()
    sizes = np.diff(sc.offsets)
    ref = None
    for tpw, bound in SPLITS:
        ctx = _ctx(sc)
        ctx.set_batched_block(tpw)
        ctx.set_wave_worlds(bound)
        print("threads %d, wave-world bound %d" % (tpw, bound), flush=True)
        ctx.step_batched(20)
        t = ctx.batched_timing()
        eff = 0 if tpw == 64 or bound < 0 else bound
        assert t["threads_per_world"] == tpw
        assert t["wave_world_bodies"] == eff and t["wave_worlds"] == int((sizes <= eff).sum()), t
        got = ctx.download()
        ctx.close()
        if ref is None:
            ref = got
        for f in STATE:
            np.testing.assert_array_equal(got[f], ref[f], err_msg="%d threads, bound %d: %s" % (tpw, bound, f))
    ctx = _ctx(sc)
    before = ctx.download()
    tune = ctx.autotune(3)
    print("autotune on the config-3 subset:", tune)
    assert tune["threads_per_world"] in (64, 128, 256, 512)
    assert tune["wave_world_bodies"] in runtime.TUNE_WAVE_BOUNDS
    after = ctx.download()
    for f in STATE:
        np.testing.assert_array_equal(after[f], before[f])
    ctx.step_batched(20)
    got = ctx.download()
    for f in STATE:
        np.testing.assert_array_equal(got[f], ref[f])
    ctx.close()


def test_full_config3_properties(oracle_mod):
    """All 4096 config-3 worlds, 240 steps, batched: finite state, no overflow, and
    three sampled worlds (largest, smallest, one median) equal to the oracle."""
    sc = S.config3()
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.step_batched(sc.nsteps)
    got = ctx.download()
    assert np.all(np.isfinite(got["pos"]))
    st = ctx.stats()
    assert st["epa_overflow"] == 0 and st["capacity_overflow"] == 0
    assert st["contacts"] > 0
    sizes = np.diff(sc.offsets)
    for w in (int(np.argmax(sizes)), int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2])):
        ws = oracle_mod.WorldState(sc, w, "f32")
        ws.run(sc.nsteps)
        a, b = int(sc.offsets[w]), int(sc.offsets[w + 1])
        assert np.abs(got["pos"][a:b] - ws.pos).max() <= 1e-3
        _state_equal(got, ws, a, b, "world %d" % w)
    # every world: the final state hashes to the oracle's (tests/golden/state_hashes.json)
    golden = load_state_hash("config3_240_steps")
    assert golden["bodies"] == sc.nbodies and golden["steps"] == sc.nsteps
    assert state_hash(got) == golden["sha256"]
    # the launch shape the step's speed rests on: 256 threads per world and three
    # workgroups per CU (a 256-body world's LDS within a third of the CU's 160 KB)
    t = ctx.batched_timing()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    assert t["threads_per_world"] == 256 and t["lds_bytes"] <= 160 * 1024 // 3, t
    assert t["workgroups"] == 3 * cus, (t, cus)
    ctx.close()
