"""CPU checks of the references the adjoint is tested against.

The adjoint differentiates the step with the GJK/EPA outcome held fixed: the
frozen-manifold rollout of oracle/nans_oracle.h.  Its binary32 replay must be
the step itself (bit for bit), and its two FP64 derivatives -- central finite
differences of the C replay and torch autograd of the independent
tests/adjoint_reference.py -- must agree with each other, which pins both."""
import numpy as np
import pytest
import torch

from nans_projekat_amd import scenes as S

from adjoint_reference import frozen_rollout


@pytest.mark.parametrize("scene,w,nsteps", [("config4", 1, 240), ("config1", 0, 240)])
def test_frozen_replay_is_the_step(oracle_mod, scene, w, nsteps):
    """Replaying the recorded contact frames in binary32 reproduces the binary32
    step bit for bit (so the frames capture everything the narrowphase fed the
    solver), and the binary64 replay stays close to it."""
    sc = S.config4(4) if scene == "config4" else S.config1()
    frames, count, fin = oracle_mod.frozen_record(sc, w, nsteps)
    ref = oracle_mod.WorldState(sc, w, "f32")
    ref.run(nsteps)
    for k in ("pos", "vel", "angvel", "angles"):
        np.testing.assert_array_equal(getattr(fin, k), getattr(ref, k))
    r32 = oracle_mod.frozen_run(sc, w, frames, count, nsteps, "f32")
    for k in ("pos", "vel", "angvel", "angles"):
        np.testing.assert_array_equal(r32[k], getattr(ref, k), err_msg=k)
    assert count.sum() == len(frames) > 0
    r64 = oracle_mod.frozen_run(sc, w, frames, count, nsteps, "f64")
    assert np.abs(r64["pos"] - ref.pos).max() < 1e-2


def _loss_c(oracle_mod, sc, w, frames, count, nsteps, target, **kw):
    st = oracle_mod.frozen_run(sc, w, frames, count, nsteps, "f64", **kw)
    return float(np.mean((st["pos"] - target) ** 2))


def test_fp64_autograd_matches_fp64_finite_differences(oracle_mod):
    """d MSE(final pos) / d(mass, v0, friction, restitution) of one config-4 world
    over 64 steps: torch autograd (float64) against central differences of the C
    replay (float64, h = 1e-6) agree to 1e-6 relative."""
    sc = S.config4(4)
    w, nsteps = 2, 64
    frames, count, fin = oracle_mod.frozen_record(sc, w, nsteps)
    ws = oracle_mod.WorldState(sc, w, "f64")
    target = fin.pos.astype(np.float64) + np.random.default_rng(5).normal(0, 0.05, fin.pos.shape)
    k_moi = ws.moi / ws.mass
    cf0, cr0 = float(ws.friction), float(ws.restitution)
    mass = torch.tensor(ws.mass, requires_grad=True)
    v0 = torch.tensor(ws.vel, requires_grad=True)
    cf = torch.tensor(cf0, dtype=torch.float64, requires_grad=True)
    cr = torch.tensor(cr0, dtype=torch.float64, requires_grad=True)
    P, _, _ = frozen_rollout(ws.params, torch.tensor(ws.pos), v0, torch.tensor(ws.angvel), mass,
                             mass * torch.tensor(k_moi), cf, cr, frames, count, nsteps)
    loss = torch.mean((P - torch.tensor(target)) ** 2)
    loss.backward()
    L = lambda **kw: _loss_c(oracle_mod, sc, w, frames, count, nsteps, target, **kw)
    assert abs(loss.item() - L()) <= 1e-12 * max(1.0, abs(L()))
    h = 1e-6
    fd_m = np.zeros(len(ws.mass))
    for i in range(len(ws.mass)):
        mp, mm = ws.mass.copy(), ws.mass.copy()
        mp[i] += h
        mm[i] -= h
        fd_m[i] = (L(mass=mp, moi=mp * k_moi) - L(mass=mm, moi=mm * k_moi)) / (2 * h)
    fd_v = np.zeros(ws.vel.size)
    for i in range(ws.vel.size):
        vp, vm = ws.vel.copy().reshape(-1), ws.vel.copy().reshape(-1)
        vp[i] += h
        vm[i] -= h
        fd_v[i] = (L(vel=vp.reshape(ws.vel.shape)) - L(vel=vm.reshape(ws.vel.shape))) / (2 * h)
    fd_cf = (L(friction=cf0 + h) - L(friction=cf0 - h)) / (2 * h)
    fd_cr = (L(restitution=cr0 + h) - L(restitution=cr0 - h)) / (2 * h)
    rel = lambda a, b: np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)
    assert rel(mass.grad.numpy(), fd_m) < 1e-6
    assert rel(v0.grad.numpy().reshape(-1), fd_v) < 1e-6
    assert rel([cf.grad.item()], [fd_cf]) < 1e-6
    assert rel([cr.grad.item()], [fd_cr]) < 1e-6
    assert np.linalg.norm(fd_m) > 0 and abs(fd_cf) > 0 and abs(fd_cr) > 0


def test_angular_velocity_does_not_reach_positions(oracle_mod):
    """Constraint's relative velocity uses W x N and W x T (code/nans.cpp:1149,
    1171-1173), which are orthogonal to N and T, so with frozen contact frames the
    initial angular velocity changes angles only: its position gradient is 0."""
    sc = S.config4(4)
    w, nsteps = 0, 64
    frames, count, fin = oracle_mod.frozen_record(sc, w, nsteps)
    ws = oracle_mod.WorldState(sc, w, "f64")
    base = oracle_mod.frozen_run(sc, w, frames, count, nsteps, "f64")
    bumped = oracle_mod.frozen_run(sc, w, frames, count, nsteps, "f64", angvel=ws.angvel + 0.5)
    # only rounding of the (zero) triple products W x N . N remains
    assert np.abs(base["pos"] - bumped["pos"]).max() < 1e-9
    assert np.abs(base["angles"] - bumped["angles"]).max() > 1e-2
