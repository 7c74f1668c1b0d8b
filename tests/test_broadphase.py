"""CPU check of the grid broadphase's soundness argument
(nans_projekat_amd/csrc/nans_broadphase.hpp): whenever the plain step's cull test
keeps a pair of tame bodies, their grid cells differ by at most one per axis, so
the 27-cell search finds the pair.  Pairs are placed on the test's boundary (just
inside, on, and just outside the limit) over radii and centre magnitudes from
1e-3 to 1e17, with the cull test and the cell coordinates evaluated in the
device's arithmetic (binary32 without contraction; cells in binary64)."""
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


def cell_inv(r, m):
    """1 / cell size of k_bp_bounds for the largest tame |radius| r and |x|+|y|+|z| m."""
    R, M = np.float64(r), np.float64(m)
    lim = 2.0 * R * 1.0001 + 1e-4 * (1.0 + 2.0 * M + 2.0 * R)
    return 1.0 / (lim * 1.01)


def coord(x, inv):
    u = np.floor(x.astype(np.float64) * inv)
    return np.clip(u, -2.0 ** 40, 2.0 ** 40).astype(np.int64)


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
        for scale_p in (1e-3, 1.0, 1e3, 1e6, 1e10, 1e15, 1e17):
            pa, ra, pb, rb = _boundary_pairs(rng, 20000, scale_r, scale_p)
            keep, _ = cull_keep(pa, ra, pb, rb)
            mag = lambda p: (np.abs(p[:, 0]) + np.abs(p[:, 1])) + np.abs(p[:, 2])
            tame = (mag(pa) <= F(1e18)) & (mag(pb) <= F(1e18))
            sel = keep & tame
            # the tightest grid: a world holding just the two bodies of each pair
            inv = cell_inv(np.maximum(np.abs(ra), np.abs(rb))[sel],
                           np.maximum(mag(pa), mag(pb))[sel])
            for c in range(3):
                da = coord(pa[sel, c], inv)
                db = coord(pb[sel, c], inv)
                assert np.all(np.abs(da - db) <= 1), (scale_r, scale_p, c)
            checked += int(sel.sum())
    assert checked > 300000


def test_the_boundary_is_exercised():
    # about half of the boundary pairs are kept, so the check above sits on the limit
    rng = np.random.default_rng(3)
    pa, ra, pb, rb = _boundary_pairs(rng, 20000, 1.0, 1.0)
    keep, _ = cull_keep(pa, ra, pb, rb)
    assert 0.3 < keep.mean() < 0.7


def test_far_cells_are_clamped_together():
    inv = cell_inv(1.0, 1e17)
    x = np.array([1e17, 1e17 * (1 + 1e-6), -1e17], F)
    c = coord(x, inv)
    assert abs(int(c[0]) - int(c[1])) <= 1 and c[2] < 0
