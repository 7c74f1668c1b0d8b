"""The opt-in quaternion orientation on the GPU (plain and batched steps, the latter
at several workgroup sizes and wave-world bounds) against the oracle's quaternion
mode, bit for bit; the Euler-to-quaternion conversion of
nans_set_orientation_mode; the differentiable rollout's refusal of the mode."""
import ctypes

import numpy as np
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S
from quat_util import euler_to_quat

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "angvel", "angles")


def _quat_ctx(sc, angles_seed=None):
    if angles_seed is not None:  # start from a random orientation
        sc.angles[:] = np.random.default_rng(angles_seed).uniform(-180, 180, sc.angles.shape).astype(np.float32)
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.record_contacts(True)
    ctx.set_orientation_mode("quaternion")
    return ctx


def _check_world(got, q, ctx, oracle_mod, sc, w, nsteps, q0):
    ws = oracle_mod.WorldState(sc, w, "f32")
    a, b = int(sc.offsets[w]), int(sc.offsets[w + 1])
    ws.quat = q0[a:b].copy()  # run() advances it in place
    ws.run(nsteps - 1)
    ref = ws.run(1, contacts=True)
    for f in STATE:
        np.testing.assert_array_equal(got[f][a:b], getattr(ws, f), err_msg="world %d %s" % (w, f))
    np.testing.assert_array_equal(q[a:b], ws.quat, err_msg="world %d quat" % w)
    con = ctx.contacts(w)
    assert len(con) == len(ref)
    for f in ("index_a", "index_b", "normal", "dl_normal"):
        np.testing.assert_array_equal(con[f], ref[f], err_msg="world %d %s" % (w, f))


def test_mode_switch_converts_the_angles():
    sc = S.config1()
    ctx = _quat_ctx(sc, angles_seed=1)
    q = ctx.quaternions()
    ref = euler_to_quat(sc.angles).astype(np.float32)
    assert np.abs(q - ref).max() <= 1.2e-7   # binary64 conversions, rounded once
    ctx.close()


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_config1_quaternion_mode_bit_exact(oracle_mod, path):
    sc = S.config1()
    ctx = _quat_ctx(sc, angles_seed=2)
    q0 = ctx.quaternions()
    (ctx.step_plain if path == "plain" else ctx.step_batched)(240)
    got, q = ctx.download(), ctx.quaternions()
    _check_world(got, q, ctx, oracle_mod, sc, 0, 240, q0)
    np.testing.assert_array_equal(got["angles"], sc.angles)   # left alone in this mode
    assert np.abs(q - q0).max() > 1e-2                        # and the bodies did turn
    ctx.close()


# (threads per world, wave-world bound): the default shape, the 64- and 512-thread
# instantiations, and the worlds of at most 16 bodies one per wave
QUAT_SHAPES = ((0, -1), (64, -1), (512, 0), (256, 16))


@pytest.mark.parametrize("threads,bound", QUAT_SHAPES)
def test_config3_worlds_quaternion_mode_batched(oracle_mod, threads, bound):
    full = S.config3(256, nsteps=1)
    sc = full.subset(list(range(0, 256, 23)))
    ctx = _quat_ctx(sc, angles_seed=3)
    ctx.set_batched_block(threads)
    ctx.set_wave_worlds(bound)
    q0 = ctx.quaternions()
    ctx.step_batched(60)
    got, q = ctx.download(), ctx.quaternions()
    for w in range(sc.nworlds):
        _check_world(got, q, ctx, oracle_mod, sc, w, 60, q0)
    ctx.close()


def test_switching_modes_recaptures_the_batched_graph(oracle_mod):
    """Euler steps, then quaternion steps on the same context: the captured graph
    follows the mode."""
    sc = S.config1()
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.record_contacts(True)
    ctx.step_batched(20)
    ws = oracle_mod.WorldState(sc, 0, "f32")
    ws.run(20)
    ctx.set_orientation_mode("quaternion")
    q0 = ctx.quaternions()
    ws.quat = q0.copy()
    ctx.step_batched(20)
    ws.run(20)
    got = ctx.download()
    for f in STATE:
        np.testing.assert_array_equal(got[f], getattr(ws, f), err_msg=f)
    np.testing.assert_array_equal(ctx.quaternions(), ws.quat)
    ctx.close()


def test_rollout_requires_euler():
    sc = S.config4(4)
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.set_orientation_mode("quaternion")
    with pytest.raises(runtime.NansError, match="Euler"):
        ctx.rollout(8, 4)
    ctx.close()
