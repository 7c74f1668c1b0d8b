"""The opt-in modes on a world large enough for the grid broadphase (1152 boxes
of config 5's grid, AUTO takes the grid from 1024 bodies): warm starting and
quaternion orientation through the plain step, bit for bit against the oracle's
same modes (which test every canonical pair), after the boxes have landed."""
import numpy as np
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "angvel", "angles")
NSTEPS = 150  # the lowest layer lands after about 64 steps


def _scene():
    sc = S.config5(nx=12, ny=8, nz=12, nsteps=NSTEPS)
    assert sc.nbodies == 1152
    return sc


def _equal(got, ws, where):
    for f in STATE:
        np.testing.assert_array_equal(got[f], getattr(ws, f), err_msg="%s %s" % (where, f))


def test_warm_start_with_the_grid(oracle_mod):
    sc = _scene()
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.record_contacts(True)
    ctx.set_warm_start(True)
    ws = oracle_mod.WorldState(sc, 0, "f32")
    ws.enable_warm(16 * sc.nbodies + 64)  # the library's capacity
    for k in range(3):
        ctx.step_plain(NSTEPS // 3)
        ref = ws.run(NSTEPS // 3, contacts=True)
        _equal(ctx.download(), ws, "after %d steps" % ((k + 1) * NSTEPS // 3))
        con = ctx.contacts(0)
        assert len(con) == len(ref)
        for f in ("index_a", "index_b", "dl_normal", "dl_tangent1", "dl_tangent2"):
            np.testing.assert_array_equal(con[f], ref[f], err_msg="%s after %d steps" % (f, (k + 1) * NSTEPS // 3))
    assert len(ref) > 0 and int(ws.warm[2][0]) > 0  # contacts, and accumulators carried
    st = ctx.stats()
    assert st["candidates"] > 0 and st["capacity_overflow"] == 0
    ctx.close()


def test_quaternions_with_the_grid(oracle_mod):
    sc = _scene()
    sc.angles[:] = np.random.default_rng(5).uniform(-30, 30, sc.angles.shape).astype(np.float32)
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.record_contacts(True)
    ctx.set_orientation_mode("quaternion")
    q0 = ctx.quaternions()
    ctx.step_plain(NSTEPS)
    got, q = ctx.download(), ctx.quaternions()
    ws = oracle_mod.WorldState(sc, 0, "f32")
    ws.quat = q0.copy()
    ws.run(NSTEPS - 1)
    ref = ws.run(1, contacts=True)
    _equal(got, ws, "after %d steps" % NSTEPS)
    np.testing.assert_array_equal(q, ws.quat)
    con = ctx.contacts(0)
    assert len(con) == len(ref) and len(ref) > 0
    for f in ("index_a", "index_b", "normal", "dl_normal"):
        np.testing.assert_array_equal(con[f], ref[f], err_msg=f)
    ctx.close()
