"""The opt-in quaternion orientation in the oracle (no counterpart in the
reference): its integrator and Model against closed forms, and its Euler
conversion against the Euler Model.  GPU parity is in test_quaternion_gpu.py."""
import dataclasses

import numpy as np

from nans_projekat_amd import scenes as S
from quat_util import CORNERS, axis_angle, euler_matrix, euler_to_quat, qmul


def test_euler_conversion_matches_the_euler_model(oracle_mod):
    rng = np.random.default_rng(5)
    for _ in range(50):
        ang = rng.uniform(-400, 400, 3)
        pos, scale = rng.uniform(-5, 5, 3), rng.uniform(0.1, 3, 3)
        q = euler_to_quat(ang[None])[0]
        e = oracle_mod.box_vertices(pos, ang, scale, "f64")
        v = oracle_mod.box_vertices_quat(pos, q, scale, "f64")
        np.testing.assert_allclose(v, e, rtol=0, atol=1e-12)
        np.testing.assert_allclose(e, pos + (euler_matrix(ang) @ (CORNERS * scale).T).T, rtol=0, atol=1e-12)


def test_free_spin_closed_form(oracle_mod):
    """A box with constant W (no damping, no contact): each step rotates by
    2 atan(|W| dt / 2) about W in the world frame, so after n steps
    q_n = (cos n phi, sin n phi W/|W|) q_0 and the corners are
    P + Rot(W/|W|, 2 n phi) R(q_0) (S * corner)."""
    ang0 = np.array([[25.0, -40.0, 110.0]])
    scale = np.array([[1.0, 0.5, 2.0]])
    sc = S._world(np.array([[0.0, 1000.0, 0.0]]), scale, np.ones(1), np.zeros((0, 3)), np.zeros((0, 3)),
                  np.zeros(0), np.zeros(0))
    sc.angles[:] = ang0
    sc.angvel[0] = [0.3, -1.2, 0.7]
    W = sc.angvel[0].astype(np.float64)                 # the binary32 values the scene holds
    sc.params = dataclasses.replace(S.StepParams(), damping=0.0, dt=1.0 / 120.0)
    dt = float(np.float32(sc.params.dt))                # the step parameters are binary32
    ws = oracle_mod.WorldState(sc, 0, "f64")
    q0 = euler_to_quat(ang0)
    ws.quat = q0.copy()
    n = 200
    ws.run(n)
    np.testing.assert_array_equal(ws.angvel[0], W)      # no torque, no damping
    np.testing.assert_array_equal(ws.angles[0], ang0[0])  # angles are left alone
    w = np.linalg.norm(W)
    phi = np.arctan(w * dt / 2.0)
    r = np.concatenate([[np.cos(n * phi)], np.sin(n * phi) * W / w])
    np.testing.assert_allclose(ws.quat[0], qmul(r, q0[0]), rtol=0, atol=1e-12)
    v = oracle_mod.box_vertices_quat(ws.pos[0], ws.quat[0], scale[0], "f64")
    R = axis_angle(W / w, 2 * n * phi) @ euler_matrix(ang0[0])
    np.testing.assert_allclose(v, ws.pos[0] + (R @ (CORNERS * scale[0]).T).T, rtol=0, atol=1e-9)


def test_binary32_quaternions_stay_unit(oracle_mod):
    sc = S.config1()
    ws = oracle_mod.WorldState(sc, 0, "f32")
    ws.quat = euler_to_quat(ws.angles).astype(np.float32)
    ws.run(240)
    norm = np.linalg.norm(ws.quat.astype(np.float64), axis=1)
    assert np.all(np.abs(norm - 1) < 1e-6)
    assert np.abs(ws.quat[:, 1:]).max() > 1e-3   # something rotated
