"""Ragged and degenerate worlds through the plain and batched steps, bit for bit
against the oracle: an empty world, single bodies, coincident boxes and spheres
(normalize(0) = NaN decides, code/nans.cpp:575), spheres only, a box inside a
larger box and a sphere inside a box (deep EPA), bodies far from the origin,
millimetre boxes, and a world at the batched path's largest size (512 bodies)."""
import numpy as np
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S
from nans_projekat_amd.scenes import _world

from test_step_gpu import _contacts_equal, _state_equal

pytestmark = pytest.mark.gpu

Z3 = np.zeros((0, 3), np.float32)
Z = np.zeros(0, np.float32)


def _boxes(pos, scale=None, mass=None, vel=None):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    scale = np.ones((n, 3), np.float32) if scale is None else np.asarray(scale, np.float32).reshape(n, 3)
    mass = np.ones(n, np.float32) if mass is None else np.asarray(mass, np.float32)
    return pos, scale, mass, vel


def _edge_scene():
    w = []
    w.append(_world(Z3, Z3, Z, Z3, Z3, Z, Z))                                           # empty
    w.append(_world(*_boxes([[1.2, 0.49, 1.0]])[:3], Z3, Z3, Z, Z))                     # one box
    w.append(_world(Z3, Z3, Z, [[0.3, 0.6, 0.2]], [[0.0, -3.0, 0.0]], [0.5], [2.0]))    # one sphere
    b = _boxes([[1.2, 3.0, 1.0], [1.2, 3.0, 1.0]])
    w.append(_world(*b[:3], Z3, Z3, Z, Z))                                              # coincident boxes
    w.append(_world(Z3, Z3, Z, [[0.0, 2.0, 0.0]] * 2, [[0.0, 0.0, 0.0]] * 2, [0.5, 0.5], [2.0, 2.0]))
    w.append(_world(Z3, Z3, Z, [[0.0, 0.5, 0.0], [0.0, 1.45, 0.0], [0.05, 2.4, 0.0]],  # spheres only
                    np.zeros((3, 3)), [0.5] * 3, [2.0] * 3))
    b = _boxes([[1.0, 2.0, 1.0], [1.3, 2.2, 0.9]], scale=[[4.0, 4.0, 4.0], [1.0, 1.0, 1.0]], mass=[8.0, 1.0])
    w.append(_world(*b[:3], Z3, Z3, Z, Z))                                              # box inside a box
    b = _boxes([[1.0, 1.5, 1.0]], scale=[[2.0, 3.0, 2.0]], mass=[3.0])
    w.append(_world(*b[:3], [[1.2, 1.9, 0.8]], [[0.0, 0.0, 0.0]], [0.4], [1.0]))       # sphere inside a box
    b = _boxes([[1e4, 1e4, -1e4], [1e4 + 0.999, 1e4, -1e4]])
    w.append(_world(*b[:3], [[1e4, 1e4 + 0.9, -1e4]], [[0.0, -1.0, 0.0]], [0.5], [2.0]))  # far away
    b = _boxes([[0.0, 0.0006, 0.0], [0.0009, 0.0016, 0.0003]], scale=np.full((2, 3), 1e-3), mass=[1e-6, 1e-6])
    w.append(_world(*b[:3], Z3, Z3, Z, Z))                                              # millimetre boxes
    g = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(7)
    pos = (g * np.array([1.02, 1.0, 1.02]) + np.array([-3.5, 0.5, -3.5]) + rng.uniform(-0.01, 0.01, g.shape))
    w.append(_world(*_boxes(pos.astype(np.float32))[:3], Z3, Z3, Z, Z))                # 512 bodies
    sc = S.concat(w)
    sc.params = S.StepParams(dt=1.0 / 120.0, solver_iterations=10, restitution=0.2, friction=0.5)
    sc.angles[:] = np.random.default_rng(3).uniform(-20.0, 20.0, sc.angles.shape).astype(np.float32)
    return sc


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_edge_worlds_bit_exact(oracle_mod, path):
    sc = _edge_scene()
    assert int(np.diff(sc.offsets).max()) == 512
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.record_contacts(True)
    nsteps = 20
    (ctx.step_plain if path == "plain" else ctx.step_batched)(nsteps)
    got = ctx.download()
    for w in range(sc.nworlds):
        a, b = int(sc.offsets[w]), int(sc.offsets[w + 1])
        ws = oracle_mod.WorldState(sc, w, "f32")
        ws.run(nsteps - 1)
        ref = ws.run(1, contacts=True)
        _state_equal(got, ws, a, b, "%s world %d" % (path, w))
        if b > a:
            _contacts_equal(ctx.contacts(w), ref, "%s world %d" % (path, w))
    st = ctx.stats()
    assert st["epa_overflow"] == 0 and st["capacity_overflow"] == 0
    ctx.close()
