"""The grid broadphase of the plain step (nans_projekat_amd/csrc/nans_broadphase.hpp)
against the all-pairs cull it replaces: the same candidates (count, and hence the
same canonical list, since each emitted pair passes the same test and none repeats)
and bit-identical steps, on config 5 and on an adversarial world of mixed sizes,
touching neighbours, far-away clusters and non-finite bodies."""
import numpy as np
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S

pytestmark = pytest.mark.gpu


def _run(sc, mode, nsteps):
    ctx = runtime.Context(sc.params)
    ctx.upload(sc)
    ctx.set_broadphase(mode)
    ctx.record_contacts(True)
    ctx.step_plain(nsteps)
    out = ctx.download(), ctx.stats(), ctx.contacts(0)
    ctx.close()
    return out


def _same(a, b):
    (sa, ta, ca), (sb, tb, cb) = a, b
    for f in ("pos", "vel", "angvel", "angles"):
        np.testing.assert_array_equal(sa[f].view(np.uint32), sb[f].view(np.uint32), err_msg=f)
    for k in ("pair_tests", "candidates", "gjk_found", "contacts"):
        assert ta[k] == tb[k], k
    assert len(ca) == len(cb)
    for f in ca.dtype.names:
        np.testing.assert_array_equal(ca[f], cb[f], err_msg=f)


def test_grid_equals_all_pairs_config5():
    sc = S.config5()
    grid, allp = _run(sc, "grid", 120), _run(sc, "all_pairs", 120)
    _same(grid, allp)
    assert grid[1]["candidates"] > 0 and grid[1]["contacts"] > 0
    print("config 5, 120 steps: %d candidates, %d contacts" % (grid[1]["candidates"], grid[1]["contacts"]))


def adversarial_world(seed=7):
    """2400 boxes and 1200 spheres: touching rows at the cull's limit, sizes over
    three decades, a cluster 1e5 m away, and wild bodies (NaN, inf, 1e20 m, a
    negative radius)."""
    rng = np.random.default_rng(seed)
    nb, ns = 2400, 1200
    half = np.exp(rng.uniform(np.log(0.01), np.log(2.0), (nb, 3))).astype(np.float32)
    bpos = rng.uniform(-30, 30, (nb, 3)).astype(np.float32)
    bpos[:400] = (np.arange(400)[:, None] * np.float32(0.6) * np.array([1, 0, 0]) + [0, 5, 0]).astype(np.float32)
    half[:400] = 0.3 / np.sqrt(3.0)                    # bounding radius 0.3: neighbours 0.6 apart touch
    bpos[400:700] = (rng.uniform(-2, 2, (300, 3)) + 1e5).astype(np.float32)
    rad = np.exp(rng.uniform(np.log(0.01), np.log(3.0), ns)).astype(np.float32)
    spos = rng.uniform(-30, 30, (ns, 3)).astype(np.float32)
    spos[:200] = (bpos[:200] + [0, 0, 0.3]).astype(np.float32)
    bpos[700] = np.nan
    bpos[701] = [np.inf, 0, 0]
    bpos[702] = [1e20, 0, 0]
    spos[5] = [0, -np.inf, 0]
    rad[6] = -1.5
    scale = (2 * half).astype(np.float32)
    mass = np.prod(scale.astype(np.float64), axis=1).astype(np.float32) + np.float32(0.01)
    sc = S._world(bpos, scale, mass, spos, np.zeros((ns, 3)), rad, np.ones(ns, np.float32))
    sc.params = S.StepParams()
    return sc


def test_grid_equals_all_pairs_adversarial():
    sc = adversarial_world()
    grid, allp = _run(sc, "grid", 2), _run(sc, "all_pairs", 2)
    _same(grid, allp)
    # the five wild bodies pair with everything
    n = sc.nbodies
    assert grid[1]["candidates"] >= 2 * 5 * (n - 5)
    print("adversarial world, 2 steps: %d candidates, %d contacts" % (grid[1]["candidates"], grid[1]["contacts"]))


def test_auto_mode_picks_the_grid_for_large_worlds():
    # same results whichever mode AUTO picks; here just that it runs and matches
    sc = S.config5(nx=16, ny=4, nz=16)   # 1024 bodies: the grid
    _same(_run(sc, "auto", 8), _run(sc, "all_pairs", 8))
