"""The world step in the FP64-islands mode (GJK's length test and EPA's final face
and depth in binary64), plain and batched, bit for bit against the oracle in the
same mode: config 1 for 240 steps and a size-stratified handful of config-3
worlds for 40 steps, with their last contact lists; the batched step under every
workgroup size and wave-world split, with the warm start, and in the
differentiable rollout's forward pass."""
import dataclasses

import numpy as np
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S

from test_step_gpu import _contacts_equal, _state_equal

pytestmark = pytest.mark.gpu


def _islands(sc):
    sc.params = dataclasses.replace(sc.params, fp64_islands=1)
    return sc


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_islands_mode_bit_exact(oracle_mod, path):
    for sc, nsteps in ((_islands(S.config1()), 240), (_islands(S.config3(128, nsteps=1).subset([0, 17, 40, 77, 101])), 40)):
        ctx = runtime.Context(sc.params)
        ctx.upload(sc)
        ctx.record_contacts(True)
        (ctx.step_plain if path == "plain" else ctx.step_batched)(nsteps)
        got = ctx.download()
        for w in range(sc.nworlds):
            ws = oracle_mod.WorldState(sc, w, "f32")
            ws.run(nsteps - 1)
            ref = ws.run(1, contacts=True)
            a, b = int(sc.offsets[w]), int(sc.offsets[w + 1])
            _state_equal(got, ws, a, b, "%s world %d" % (path, w))
            _contacts_equal(ctx.contacts(w), ref, "%s world %d" % (path, w))
        assert ctx.stats()["epa_overflow"] == 0
        ctx.close()


def _islands_subset():
    return _islands(S.config3(128, nsteps=1).subset([0, 17, 40, 77, 101]))  # 8 to 191 bodies


# (threads per world, wave-world bound): the FP64-islands instantiation of the
# batched step under every workgroup size, with every world by a workgroup and with
# the worlds of at most 16 or 64 bodies by one wave each
ISL_SPLITS = ((64, -1), (128, 64), (256, 16), (256, 64), (512, 0), (512, 64))


def test_islands_batched_shapes(oracle_mod):
    sc = _islands_subset()
    nsteps = 40
    refs = []
    for w in range(sc.nworlds):
        ws = oracle_mod.WorldState(sc, w, "f32")
        ws.run(nsteps - 1)
        refs.append((ws, ws.run(1, contacts=True)))
    for tpw, bound in ISL_SPLITS:
        print("islands: threads %d, wave-world bound %d" % (tpw, bound), flush=True)
        ctx = runtime.Context(sc.params)
        ctx.upload(sc)
        ctx.set_batched_block(tpw)
        ctx.set_wave_worlds(bound)
        ctx.record_contacts(True)
        ctx.step_batched(nsteps)
        assert ctx.batched_timing()["threads_per_world"] == tpw
        got = ctx.download()
        for w, (ws, ref) in enumerate(refs):
            a, b = int(sc.offsets[w]), int(sc.offsets[w + 1])
            where = "%d threads, bound %d, world %d" % (tpw, bound, w)
            _state_equal(got, ws, a, b, where)
            _contacts_equal(ctx.contacts(w), ref, where)
        assert ctx.stats()["epa_overflow"] == 0
        ctx.close()


def test_islands_warm_batched(oracle_mod):
    """The FP64-islands mode and the warm start together, batched, 60 steps."""
    sc = _islands_subset()
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.record_contacts(True)
    ctx.set_warm_start(True)
    ctx.set_wave_worlds(64)
    for k in range(3):
        ctx.step_batched(20)
    got = ctx.download()
    for w in range(sc.nworlds):
        ws = oracle_mod.WorldState(sc, w, "f32")
        ws.enable_warm(16 * (int(sc.offsets[w + 1]) - int(sc.offsets[w])) + 64)  # the library's capacity
        ws.run(59)
        ref = ws.run(1, contacts=True)
        _state_equal(got, ws, int(sc.offsets[w]), int(sc.offsets[w + 1]), "warm world %d" % w)
        con = ctx.contacts(w)
        assert len(con) == len(ref), w
        for f in ("index_a", "index_b", "dl_normal", "dl_tangent1", "dl_tangent2"):
            np.testing.assert_array_equal(con[f], ref[f], err_msg="warm world %d %s" % (w, f))
    ctx.close()


def test_islands_rollout_equals_step():
    """The differentiable rollout's forward steps in the FP64-islands mode land on
    the same state as the batched step's."""
    sc = _islands_subset()
    out = []
    for roll in (False, True):
        ctx = runtime.Context(sc.params)
        ctx.upload(sc)
        if roll:
            ctx.rollout(48, 16)
        else:
            ctx.step_batched(48)
        out.append(ctx.download())
        ctx.close()
    for f in ("pos", "vel", "angvel", "angles"):
        np.testing.assert_array_equal(out[1][f], out[0][f], err_msg=f)
