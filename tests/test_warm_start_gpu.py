"""The opt-in warm start of the plain and the batched step on the GPU against
the oracle's warm start, bit for bit: config 1 step by step, config-4 worlds with
their own coefficients, the reference demo scene, size-stratified config-3 worlds
(batched, 8 to 249 bodies); the two steps alternating on one set of
contact lists; the lists emptied by set_state; the autotuner and the
differentiable rollout refusing the mode."""
import numpy as np
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "angvel", "angles")


def _warm_ctx(sc):
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.record_contacts(True)
    ctx.set_warm_start(True)
    return ctx


def _oracle(oracle_mod, sc, w):
    ws = oracle_mod.WorldState(sc, w, "f32")
    ws.enable_warm(16 * (int(sc.offsets[w + 1]) - int(sc.offsets[w])) + 64)  # the library's capacity
    return ws


def _equal(got, ws, a, b, where):
    for f in STATE:
        np.testing.assert_array_equal(got[f][a:b], getattr(ws, f), err_msg="%s %s" % (where, f))


def _stepper(ctx, path):
    return ctx.step_plain if path == "plain" else ctx.step_batched


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_config1_warm_stepwise(oracle_mod, path):
    sc = S.config1()
    ctx = _warm_ctx(sc)
    step_fn = _stepper(ctx, path)
    ws = _oracle(oracle_mod, sc, 0)
    carried = 0
    for step in range(240):
        step_fn(1)
        ref = ws.run(1, contacts=True)
        con = ctx.contacts(0)
        assert len(con) == len(ref), step
        for f in ("index_a", "index_b", "dl_normal", "dl_tangent1", "dl_tangent2"):
            np.testing.assert_array_equal(con[f], ref[f], err_msg="step %d %s" % (step, f))
        if step % 20 == 19:
            _equal(ctx.download(), ws, 0, sc.nbodies, "step %d" % step)
        carried = max(carried, int(ws.warm[2][0]))
    assert carried > 0
    # the mode changes results: the cold oracle differs
    cold = oracle_mod.WorldState(sc, 0, "f32")
    cold.run(240)
    assert not np.array_equal(cold.pos, ws.pos)
    ctx.close()


@pytest.mark.parametrize("path", ["plain", "batched"])
@pytest.mark.parametrize("scene", ["config4", "demo"])
def test_worlds_warm(oracle_mod, scene, path):
    sc = S.config4(8) if scene == "config4" else S.demo_scene()
    ctx = _warm_ctx(sc)
    _stepper(ctx, path)(90)
    got = ctx.download()
    for w in range(sc.nworlds):
        ws = _oracle(oracle_mod, sc, w)
        ws.run(90)
        _equal(got, ws, int(sc.offsets[w]), int(sc.offsets[w + 1]), "world %d" % w)
    ctx.close()


def test_set_state_empties_the_lists(oracle_mod):
    sc = S.config1()
    ctx = _warm_ctx(sc)
    ctx.step_plain(50)
    ctx.set_state(sc)          # back to the start: every contact starts from zero again
    ctx.step_plain(30)
    ws = _oracle(oracle_mod, sc, 0)
    ws.run(30)
    _equal(ctx.download(), ws, 0, sc.nbodies, "after set_state")
    ctx.close()


@pytest.mark.parametrize("wave_bound", [0, 64])
def test_config3_worlds_batched_warm(oracle_mod, wave_bound):
    """Six config-3 worlds (8 to 249 bodies), 60 warm-started batched steps; with
    wave_bound 64 the worlds of at most 64 bodies run one per wave."""
    sc = S.config3(128, nsteps=1).subset([0, 17, 40, 77, 102, 127])  # 8 to 249 bodies
    ctx = _warm_ctx(sc)
    ctx.set_wave_worlds(wave_bound)
    for k in range(3):
        ctx.step_batched(20)
    got = ctx.download()
    st = ctx.stats()
    assert st["epa_overflow"] == 0 and st["capacity_overflow"] == 0
    for w in range(sc.nworlds):
        ws = _oracle(oracle_mod, sc, w)
        ws.run(59)
        ref = ws.run(1, contacts=True)
        _equal(got, ws, int(sc.offsets[w]), int(sc.offsets[w + 1]), "world %d" % w)
        con = ctx.contacts(w)
        assert len(con) == len(ref), w
        for f in ("index_a", "index_b", "dl_normal", "dl_tangent1", "dl_tangent2"):
            np.testing.assert_array_equal(con[f], ref[f], err_msg="world %d %s" % (w, f))
    ctx.close()


@pytest.mark.parametrize("threads", [64, 512])
def test_config3_worlds_batched_warm_sizes(oracle_mod, threads):
    """The warm-started batched step at 64 and 512 threads per world (instantiations
    of their own) on six config-3 worlds, 40 steps, against the oracle's warm start."""
    sc = S.config3(128, nsteps=1).subset([0, 17, 40, 77, 102, 127])  # 8 to 249 bodies
    ctx = _warm_ctx(sc)
    ctx.set_batched_block(threads)
    ctx.step_batched(20)
    ctx.step_batched(20)
    assert ctx.batched_timing()["threads_per_world"] == threads
    got = ctx.download()
    st = ctx.stats()
    assert st["epa_overflow"] == 0 and st["capacity_overflow"] == 0
    for w in range(sc.nworlds):
        ws = _oracle(oracle_mod, sc, w)
        ws.run(39)
        ref = ws.run(1, contacts=True)
        _equal(got, ws, int(sc.offsets[w]), int(sc.offsets[w + 1]), "%d threads, world %d" % (threads, w))
        con = ctx.contacts(w)
        assert len(con) == len(ref), w
        for f in ("index_a", "index_b", "dl_normal", "dl_tangent1", "dl_tangent2"):
            np.testing.assert_array_equal(con[f], ref[f], err_msg="%d threads, world %d %s" % (threads, w, f))
    ctx.close()


def test_plain_and_batched_share_the_lists(oracle_mod):
    sc = S.config4(8)
    ctx = _warm_ctx(sc)
    ctx.step_plain(30)
    ctx.step_batched(30)   # picks up the accumulators the plain step left
    ctx.step_plain(15)
    ctx.step_batched(15)
    got = ctx.download()
    for w in range(sc.nworlds):
        ws = _oracle(oracle_mod, sc, w)
        ws.run(90)
        _equal(got, ws, int(sc.offsets[w]), int(sc.offsets[w + 1]), "world %d" % w)
    ctx.close()


def test_autotuner_and_rollout_refuse_the_mode():
    sc = S.config4(4)
    ctx = _warm_ctx(sc)
    with pytest.raises(runtime.NansError, match="warm"):
        ctx.autotune(1)
    with pytest.raises(runtime.NansError, match="warm"):
        ctx.rollout(8, 4)
    ctx.set_warm_start(False)
    ctx.step_batched(1)
    ctx.close()
