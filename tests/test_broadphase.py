"""CPU check of the grid broadphase's soundness argument
(nans_projekat_amd/csrc/nans_broadphase.hpp): whenever the plain step's cull test
keeps a pair of tame bodies, their distance per axis is within the sum of their
reaches, and their cells at the coarser body's level differ by at most one per
axis, so the 27-cell search finds the pair.  Pairs are placed on the test's
boundary (just inside, on, and just outside the limit) over radii from 1e-3 to
1e14 and centre magnitudes from 1e-3 to 1e14, with the cull test in the device's
binary32 arithmetic (no contraction) and reaches, levels and cells in binary64."""
import numpy as np

F = np.float32


def cull_keep(pa, ra, pb, rb):
    """cull_keep of nans_projekat_amd/csrc/nans_world.hpp in binary32."""
    d = (pb - pa).astype(F)
    dist2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    a, b = np.abs(pa), np.abs(pb)
    mag = ((((a[:, 0] + a[:, 1]) + a[:, 2]) + b[:, 0]) + b[:, 1]) + b[:, 2]
    lim = (ra + rb) * F(1.0001) + F(1e-4) * (((F(1.0) + mag) + ra) + rb)
    return ~(dist2 > lim * lim), lim


def reach(p, r):
    """Per-body reach of the grid (k_bp_cells), binary64 from binary32 inputs."""
    mag = (np.abs(p[:, 0]) + np.abs(p[:, 1])) + np.abs(p[:, 2])
    return (np.abs(r).astype(np.float64) * 1.0002 + 1e-4 * (0.5 + mag.astype(np.float64))) * (1.0 + 4e-6)


def rho0_of(rmax):
    return (np.float64(rmax) * 1.0002 + 1e-4 * (0.5 + 128.0)) * (1.0 + 4e-6)


def level(rho, rho0):
    k = np.zeros(rho.shape, np.int64)
    lim = np.full(rho.shape, rho0, np.float64)
    for _ in range(63):
        up = rho > lim
        k += up
        lim = np.where(up, lim * 2.0, lim)
    return k


def coord(x, inv):
    u = np.floor(x.astype(np.float64) * inv)
    return np.clip(u, -2.0 ** 40, 2.0 ** 40).astype(np.int64)


def cells_at(p, k, rho0):
    inv = np.ldexp(1.0 / (2.02 * rho0), -k)
    return np.stack([coord(p[:, c], inv) for c in range(3)], axis=1)


def _boundary_pairs(rng, n, scale_r, scale_p):
    ra = (rng.uniform(0.0, 1.0, n) * scale_r).astype(F)
    rb = (rng.uniform(0.0, 1.0, n) * scale_r).astype(F)
    rb[::7] *= F(-1.0)                                  # a negative radius is tame too
    pa = (rng.uniform(-1.0, 1.0, (n, 3)) * scale_p).astype(F)
    _, lim = cull_keep(pa, ra, pa, rb)
    u = rng.normal(size=(n, 3))
    u[::5] = np.eye(3)[rng.integers(0, 3, n)[::5]]      # axis-aligned: the per-axis worst case
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    f = rng.uniform(0.999, 1.001, n)                    # straddle the limit
    pb = (pa.astype(np.float64) + u * (np.abs(lim) * f)[:, None]).astype(F)
    return pa, ra, pb, rb


def test_kept_pairs_lie_in_neighbouring_cells():
    rng = np.random.default_rng(11)
    checked = 0
    for scale_r in (1e-3, 0.1, 1.0, 30.0, 1e4, 1e8, 1e14):
        for scale_p in (1e-3, 1.0, 1e3, 1e6, 1e10, 1e14):
            pa, ra, pb, rb = _boundary_pairs(rng, 20000, scale_r, scale_p)
            keep, _ = cull_keep(pa, ra, pb, rb)
            mag = lambda p: (np.abs(p[:, 0]) + np.abs(p[:, 1])) + np.abs(p[:, 2])
            tame = (mag(pa) <= F(1e15)) & (mag(pb) <= F(1e15)) & (np.abs(ra) <= F(1e15)) & (np.abs(rb) <= F(1e15))
            sel = keep & tame
            pa, ra, pb, rb = pa[sel], ra[sel], pb[sel], rb[sel]
            rha, rhb = reach(pa, ra), reach(pb, rb)
            # the reach bound the proof rests on
            d = np.abs(pa.astype(np.float64) - pb.astype(np.float64)).max(axis=1)
            assert np.all(d <= rha + rhb), (scale_r, scale_p)
            # the tightest grid: rho0 from the pair's own radii, either body the finer one
            for rho0 in (rho0_of(np.maximum(np.abs(ra), np.abs(rb)).min()), rho0_of(0.0)):
                ka, kb = level(rha, rho0), level(rhb, rho0)
                k = np.maximum(ka, kb)
                assert np.all(rha <= rho0 * 2.0 ** ka) and np.all(rhb <= rho0 * 2.0 ** kb)
                assert np.all(np.abs(cells_at(pa, k, rho0) - cells_at(pb, k, rho0)) <= 1), (scale_r, scale_p)
            checked += int(sel.sum())
    assert checked > 300000


def test_the_boundary_is_exercised():
    # about half of the boundary pairs are kept, so the check above sits on the limit
    rng = np.random.default_rng(3)
    pa, ra, pb, rb = _boundary_pairs(rng, 20000, 1.0, 1.0)
    keep, _ = cull_keep(pa, ra, pb, rb)
    assert 0.3 < keep.mean() < 0.7


def test_far_cells_are_clamped_together():
    rho0 = rho0_of(1.0)
    k = np.array([0, 0, 0])
    x = np.array([[1e14, 0, 0], [1e14 * (1 + 1e-6), 0, 0], [-1e14, 0, 0]], F)
    c = cells_at(x, k, rho0)[:, 0]
    assert abs(int(c[0]) - int(c[1])) <= 1 and c[0] == 2 ** 40 and c[2] == -2 ** 40
