"""Regression tests of the batched step's host bookkeeping on the GPU:
- the autotuner restores the quaternion planes of the opt-in orientation mode;
- nans_set_state sets the quaternions from the given angles, as the upload does;
- a captured step graph is dropped when the per-workgroup scratch it points into
  is reallocated (a larger grid at another workgroup size), so a later step never
  replays kernel arguments that point at freed memory."""
import numpy as np
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "angvel", "angles")


def _subset(nworlds=64):
    full = S.config3(nworlds, nsteps=1)
    return full


def test_autotune_keeps_quaternions():
    sc = _subset()
    a = runtime.Context(sc.params)
    a.upload(sc)
    a.set_orientation_mode("quaternion")
    a.step_batched(3)
    q_before = a.quaternions()
    s_before = a.download()
    tune = a.autotune(3)
    print("autotune in quaternion mode:", tune)
    np.testing.assert_array_equal(a.quaternions(), q_before)
    s_after = a.download()
    for f in STATE:
        np.testing.assert_array_equal(s_after[f], s_before[f], err_msg=f)
    a.step_batched(10)
    b = runtime.Context(sc.params)
    b.upload(sc)
    b.set_orientation_mode("quaternion")
    b.step_batched(13)
    np.testing.assert_array_equal(a.quaternions(), b.quaternions())
    ga, gb = a.download(), b.download()
    for f in STATE:
        np.testing.assert_array_equal(ga[f], gb[f], err_msg=f)
    a.close()
    b.close()


def test_set_state_sets_quaternions_from_angles():
    sc = _subset(16)
    a = runtime.Context(sc.params)
    a.upload(sc)
    a.set_orientation_mode("quaternion")
    q0 = a.quaternions()
    a.step_batched(5)
    assert not np.array_equal(a.quaternions(), q0)
    a.set_state(sc)
    np.testing.assert_array_equal(a.quaternions(), q0)
    a.step_batched(5)
    b = runtime.Context(sc.params)
    b.upload(sc)
    b.set_orientation_mode("quaternion")
    b.step_batched(5)
    np.testing.assert_array_equal(a.quaternions(), b.quaternions())
    ga, gb = a.download(), b.download()
    for f in STATE:
        np.testing.assert_array_equal(ga[f], gb[f], err_msg=f)
    a.close()
    b.close()


def test_graph_dropped_when_scratch_grows():
    """Rollout at the default workgroup size (captures the step graph), the adjoint
    at 64 threads per world (a larger grid: the scratch is reallocated), then the
    default size again: the step equals an uninterrupted run bit for bit."""
    sc = S.config3(1536, nsteps=1)
    probe = runtime.Context(sc.params)
    probe.upload(sc)
    probe.step_batched(1)
    grid_default = probe.batched_timing()["workgroups"]
    probe.set_batched_block(64)
    probe.step_batched(1)
    grid64 = probe.batched_timing()["workgroups"]
    probe.close()
    print("workgroups: default %d, 64 threads %d" % (grid_default, grid64))
    assert grid64 > grid_default  # the sequence below reallocates the scratch

    a = runtime.Context(sc.params)
    a.upload(sc)
    a.rollout(8, 4)
    a.set_batched_block(64)
    a.adjoint(dl_pos=np.ones((sc.nbodies, 3), np.float32))
    a.set_batched_block(0)
    a.step_batched(4)
    b = runtime.Context(sc.params)
    b.upload(sc)
    b.rollout(8, 4)
    b.step_batched(4)
    ga, gb = a.download(), b.download()
    for f in STATE:
        np.testing.assert_array_equal(ga[f], gb[f], err_msg=f)
    a.close()
    b.close()


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_capacity_overflow_leaves_the_batch_unchanged(path):
    """A world of 512 coincident boxes has 130,816 candidate pairs, beyond both
    steps' candidate scratch (plain: 64 n + 4096, batched: 32 n + 1024): the call
    fails with NANS_ERR_CAPACITY and leaves the state and the statistics as they
    were; a normal world beside it is untouched as well."""
    n = 512
    big = S._world(np.tile(np.float32([[3.0, 5.0, 2.0]]), (n, 1)), np.ones((n, 3), np.float32),
                   np.ones(n, np.float32), np.zeros((0, 3)), np.zeros((0, 3)), [], [])
    sc = S.concat([S.config1(), big])
    sc.params = S.config1().params
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    before = ctx.download()
    stats = ctx.stats()
    with pytest.raises(runtime.NansError) as e:
        (ctx.step_plain if path == "plain" else ctx.step_batched)(2)
    assert e.value.code == runtime.ERR_CAPACITY
    after = ctx.download()
    for f in STATE:
        np.testing.assert_array_equal(after[f], before[f], err_msg=f)
    assert ctx.stats() == stats
    ctx.close()
