"""Invariants of the oracle's EvolveSimplex and ResolveCollision, call by call.

The oracle restates GJK (code/nans.cpp:571-769) and EPA (code/nans.cpp:787-904)
with the reference's quirks.  These tests pin that restatement with properties
the geometry fixes, checked on every recorded call of config-2 pairs (binary32)
through oracle.trace_pair:

GJK
  * supports: the pushed SupA maximises dot(v, d) over A's corners (first corner
    on a tie, code/nans.cpp:409-430), SupB does so along -d, a sphere's support
    lies at distance r from its centre along d, and P = SupA - SupB;
  * directions: the first is the unit vector between the centres, the second its
    exact negation whatever the first point was (code/nans.cpp:575-592);
  * line case: behind A the simplex keeps A and the direction is exactly -A,
    beyond B it keeps B and -B; in between it keeps both, the direction is
    perpendicular to AB, points at the origin, and has length |AB| times the
    distance of the origin from the line (the unit-edge projection,
    code/nans.cpp:539-562);
  * triangle case: the oldest vertex is never dropped (code/nans.cpp:636-706); a
    kept edge gets a direction perpendicular to it, towards the origin; a kept
    triangle gets its normal on the origin's side (with B, C swapped when the
    normal is flipped);
  * tetrahedron case: a found intersection holds the origin inside the
    tetrahedron; otherwise the newest vertex stays, one of the other three goes,
    and the direction is the kept face's normal towards the origin
    (code/nans.cpp:712-753);
  * ends: 'no' after the length test means |dir| <= 1e-4, after a push that the
    new point lies behind the direction.
EPA
  * the polytope is a closed, consistently oriented triangulated surface
    (every directed edge once, its reverse once; V - E + F = 2);
  * stored normals are unit, perpendicular to their face and point away from the
    origin (code/nans.cpp:291-322);
  * the chosen face has the least distance, and the new support is the
    Minkowski support along its stored normal;
  * convergence follows dot(N, P) - d < MAX_EPA_ERROR (code/nans.cpp:826);
  * the removed faces are exactly those that see the new point; the survivors
    keep their order and the new faces, all on the new point, follow
    (code/nans.cpp:869-903);
  * the contact: barycentric weights sum to one and rebuild N d on the face;
    the contact normal is -N (code/nans.cpp:829-866).
A sign error, a swapped operand, a dropped term or a transposed rule in either
routine breaks at least one of these on the pairs below."""
import numpy as np
import pytest

from nans_projekat_amd import scenes as S
import oracle as O

F32 = np.float32
BOX = 0


def _params():
    return S.StepParams()


def _corners(pos, quat, half):
    return O.box_vertices_quat(pos, quat, F32(2.0) * np.asarray(half, F32), "f32").astype(np.float64)


def _dot32(v, d):
    v = np.asarray(v, F32)
    d = np.asarray(d, F32)
    return (v[..., 0] * d[0] + v[..., 1] * d[1]) + v[..., 2] * d[2]


def _box_support(corners, d):
    """First corner of greatest binary32 dot(v, d) (the strict '>' scan from -FLT_MAX)."""
    k = _dot32(corners.astype(F32), d)
    best, idx = np.float32(-np.finfo(F32).max), -1
    for i in range(8):
        if k[i] > best:
            best, idx = k[i], i
    return corners[idx] if idx >= 0 else np.zeros(3)


class Shape:
    def __init__(self, kind, pos, quat, half):
        self.kind = int(kind)
        self.pos = np.asarray(pos, np.float64)
        self.r = float(half[0])
        self.corners = _corners(pos, quat, half) if self.kind == BOX else None

    def check_support(self, s, d, stats):
        d64 = np.asarray(d, np.float64)
        if self.kind == BOX:
            np.testing.assert_array_equal(s, _box_support(self.corners, d))
            # maximal over the corners (binary64, up to rounding of the dots)
            dots = self.corners @ d64
            scale = np.abs(self.corners).sum(axis=1).max() * np.abs(d64).max()
            assert s @ d64 >= dots.max() - 1e-5 * scale
            stats["box_support"] += 1
        else:
            off = s - self.pos
            n = np.linalg.norm(d64)
            assert abs(np.linalg.norm(off) - self.r) <= 1e-5 * (1.0 + self.r + np.abs(self.pos).max())
            assert off @ d64 >= (1.0 - 1e-5) * self.r * n
            stats["sphere_support"] += 1


def _unit(v):
    return v / np.linalg.norm(v)


def _perp_ok(dirv, e, tol=2e-4):
    return abs(dirv @ e) <= tol * np.linalg.norm(dirv) * np.linalg.norm(e)


def _check_gjk(recs, A, B, stats):
    d_first = None
    for k, r in enumerate(recs):
        n_in, n_out, pushed = int(r["n_in"]), int(r["n_out"]), int(r["pushed"])
        s_in, s_out, dirv = r["s_in"][:n_in], r["s_out"][:n_out], r["dir"]
        kept = n_out - pushed
        if r["result"] == 1:
            # found: only from the tetrahedron case, with the origin inside it
            assert n_in == 4 and n_out == 4 and not pushed
            T = np.array(s_in)
            M = np.stack([T[0] - T[3], T[1] - T[3], T[2] - T[3]], axis=1)
            if abs(np.linalg.det(M)) > 1e-9 * np.abs(M).max() ** 3:
                lam = np.linalg.solve(M, -T[3])
                bary = np.append(lam, 1.0 - lam.sum())
                assert bary.min() >= -1e-4, bary
                stats["found_inside"] += 1
            continue
        if pushed:
            d32 = dirv.astype(F32)
            A.check_support(r["sup_a"], d32, stats)
            B.check_support(r["sup_b"], (d32 * F32(-1.0)).astype(F32), stats)
            np.testing.assert_array_equal(s_out[-1], (r["sup_a"].astype(F32) - r["sup_b"].astype(F32)).astype(np.float64))
            if r["result"] == 2:
                assert _dot32(s_out[-1], d32) >= 0
            else:
                assert _dot32(s_out[-1], d32) < 0
        else:
            assert r["result"] == 0 and np.linalg.norm(dirv) <= 1e-4 * (1 + 1e-6)
            stats["length_exit"] += 1
        if n_in == 0:
            c = B.pos - A.pos
            np.testing.assert_allclose(dirv, _unit(c), rtol=0, atol=1e-6)
            d_first = dirv
            stats["first_dir"] += 1
        elif n_in == 1:
            np.testing.assert_array_equal(dirv, -d_first)
            np.testing.assert_array_equal(s_out[0], s_in[0])
            stats["second_dir"] += 1
        elif n_in == 2:
            a, b = s_in
            ab = b - a
            t = (-a) @ ab / (ab @ ab)
            if t < -1e-4:
                assert kept == 1
                np.testing.assert_array_equal(s_out[0], a)
                np.testing.assert_array_equal(dirv, -a)
                stats["line_a"] += 1
            elif t > 1 + 1e-4:
                assert kept == 1
                np.testing.assert_array_equal(s_out[0], b)
                np.testing.assert_array_equal(dirv, -b)
                stats["line_b"] += 1
            elif 1e-4 < t < 1 - 1e-4:
                assert kept == 2
                np.testing.assert_array_equal(s_out[:2], s_in)
                dist = np.linalg.norm(np.cross(a, ab)) / np.linalg.norm(ab)
                if dist > 1e-4 * (1 + np.abs(s_in).max()):
                    assert _perp_ok(dirv, ab)
                    assert dirv @ (-a) > 0
                    np.testing.assert_allclose(np.linalg.norm(dirv), np.linalg.norm(ab) * dist, rtol=2e-3)
                    stats["line_mid"] += 1
        elif n_in == 3:
            s0 = s_in[0]
            assert kept >= 1
            np.testing.assert_array_equal(s_out[0], s0)  # the oldest vertex stays
            n2o = -s0
            e1, e2 = s_in[1] - s0, s_in[2] - s0
            tn = np.cross(e1, e2)
            # where the origin lies seen from s0: inside the triangle's wedge at s0
            # (behind both edge planes), or outside it; ahead of s0 along each edge
            w2, w1 = np.cross(tn, e2) @ n2o, np.cross(e1, tn) @ n2o
            a1, a2 = e1 @ n2o, e2 @ n2o
            band = 1e-5 * (np.linalg.norm(tn) + 1e-30) * np.linalg.norm(n2o) * (np.linalg.norm(e1) + np.linalg.norm(e2))
            clear = min(abs(w1), abs(w2)) > band and min(abs(a1), abs(a2)) > 1e-5 * np.linalg.norm(n2o) * max(
                np.linalg.norm(e1), np.linalg.norm(e2))
            inside = w1 < 0 and w2 < 0
            if kept == 1:
                # the vertex region of s0: outside the wedge, behind s0 along both edges
                np.testing.assert_array_equal(dirv, n2o)
                if clear:
                    assert not inside and a1 < 0 and a2 < 0
                stats["tri_vertex"] += 1
            elif kept == 2:
                other = s_out[1]
                k = 1 if np.array_equal(other, s_in[1]) else 2
                assert np.array_equal(other, s_in[k])
                e = other - s0
                if clear:
                    # an edge region: outside the wedge, ahead of s0 along the kept edge
                    assert not inside and (a1 if k == 1 else a2) > 0
                if np.linalg.norm(np.cross(e, n2o)) > 1e-3 * np.linalg.norm(e) * np.linalg.norm(n2o):
                    assert _perp_ok(dirv, e)
                    assert dirv @ n2o > 0
                    stats["tri_edge"] += 1
            else:
                s1, s2 = s_out[1], s_out[2]
                swapped = np.array_equal(s1, s_in[2]) and np.array_equal(s2, s_in[1])
                assert swapped or (np.array_equal(s1, s_in[1]) and np.array_equal(s2, s_in[2]))
                if clear:
                    assert inside  # the face region
                nrm = np.cross(s1 - s0, s2 - s0)
                if np.linalg.norm(nrm) > 1e-6 * (1 + np.abs(s_in).max()) ** 2:
                    assert _perp_ok(dirv, s1 - s0) and _perp_ok(dirv, s2 - s0)
                    assert dirv @ n2o >= 0
                    # the kept winding's normal is the direction, towards the origin
                    assert nrm @ dirv > 0
                    stats["tri_face"] += 1
        elif n_in == 4:
            assert kept == 3
            D = s_in[3]
            np.testing.assert_array_equal(s_out[2], D)  # the newest vertex stays, last
            rest = [j for j in range(3) if not any(np.array_equal(s_in[j], s_out[m]) for m in range(2))]
            assert len(rest) == 1
            face = s_out[:3]
            if np.linalg.norm(np.cross(face[0] - D, face[1] - D)) > 1e-6 * (1 + np.abs(s_in).max()) ** 2:
                assert _perp_ok(dirv, face[0] - D) and _perp_ok(dirv, face[1] - D)
                assert dirv @ (-D) >= 0
                stats["tetra_face"] += 1


def _faces(rec):
    f = rec["face"][:int(rec["nfaces"])]
    return f[:, 0:3], f[:, 3:6], f[:, 6:9], f[:, 9:12]


def _check_epa(recs, A, B, stats, tol_epa):
    for k, r in enumerate(recs):
        Pa, Pb, Pc, N = _faces(r)
        nf = len(Pa)
        # closed, consistently oriented surface (vertices by exact equality)
        keys = {}
        vid = lambda p: keys.setdefault(tuple(p), len(keys))
        tri = [(vid(Pa[i]), vid(Pb[i]), vid(Pc[i])) for i in range(nf)]
        if len(set(keys)) < len(keys) or any(len(set(t)) < 3 for t in tri):
            continue
        edges = {}
        for a, b, c in tri:
            for e in ((a, b), (b, c), (c, a)):
                edges[e] = edges.get(e, 0) + 1
        assert all(v == 1 for v in edges.values())
        assert all((b, a) in edges for (a, b) in edges)
        assert len(keys) - len(edges) // 2 + nf == 2
        stats["epa_closed"] += 1
        # stored normals: unit, perpendicular, away from the origin
        for i in range(nf):
            np.testing.assert_allclose(np.linalg.norm(N[i]), 1.0, atol=1e-5)
            assert _perp_ok(N[i], Pb[i] - Pa[i], 1e-3) and _perp_ok(N[i], Pc[i] - Pa[i], 1e-3)
            assert _dot32(Pa[i], N[i]) >= 0
        # distances of the face planes (binary64) and the chosen face
        nrm = np.cross(Pb - Pa, Pc - Pa)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        dist = np.abs(np.einsum("ij,ij->i", nrm, Pa))
        best = int(r["best"])
        scale = 1e-5 * (1 + np.abs(Pa).max())
        assert r["cur"] <= dist.min() + scale and abs(r["cur"] - dist[best]) <= scale
        np.testing.assert_array_equal(r["dir"], N[best])
        stats["epa_closest"] += 1
        # the new support along the stored normal
        d32 = r["dir"].astype(F32)
        A.check_support(r["sa_new"], d32, stats)
        B.check_support(r["sb_new"], (d32 * F32(-1.0)).astype(F32), stats)
        gap = r["dir"] @ r["p_new"] - r["cur"]
        if abs(gap - tol_epa) > 1e-5:
            assert bool(r["converged"]) == (gap < tol_epa)
        if r["converged"]:
            U, V, W = r["bary"]
            a, b, c = Pa[best], Pb[best], Pc[best]
            if max(abs(U), abs(V), abs(W)) <= 1.0:
                np.testing.assert_allclose(U + V + W, 1.0, atol=1e-5)
                np.testing.assert_allclose(U * a + V * b + W * c, N[best] * r["cur"], atol=1e-4 * (1 + np.abs(a).max()))
                sa = U * r["face_sup"][0][:3] + V * r["face_sup"][1][:3] + W * r["face_sup"][2][:3]
                sb = U * r["face_sup"][0][3:] + V * r["face_sup"][1][3:] + W * r["face_sup"][2][3:]
                np.testing.assert_allclose(sa - sb, N[best] * r["cur"], atol=1e-4 * (1 + np.abs(a).max()))
                stats["epa_contact"] += 1
            continue
        # removed faces are those that see the new point
        vis = r["visible"][:nf].astype(bool)
        flip = np.where((np.einsum("ij,ij->i", nrm, Pa) < 0)[:, None], -nrm, nrm)
        side = np.einsum("ij,ij->i", flip, r["p_new"] - Pa)
        band = 1e-5 * (1 + np.abs(Pa).max() + np.abs(r["p_new"]).max())
        clear = np.abs(side) > band
        assert np.all(vis[clear] == (side[clear] > 0))
        assert vis.any()
        # the next polytope: survivors in order, then new faces on the new point
        if k + 1 < len(recs):
            nxt = recs[k + 1]
            qa, qb, qc, qn = _faces(nxt)
            ns = int((~vis).sum())
            assert int(nxt["nfaces"]) == int(r["nfaces_out"])
            np.testing.assert_array_equal(qa[:ns], Pa[~vis])
            np.testing.assert_array_equal(qb[:ns], Pb[~vis])
            np.testing.assert_array_equal(qc[:ns], Pc[~vis])
            np.testing.assert_array_equal(qa[ns:], np.broadcast_to(r["p_new"], qa[ns:].shape))
            stats["epa_expand"] += 1


def _pairs(n):
    pr = S.config2_pairs(4 * n)
    idx = np.concatenate([np.nonzero((pr["type_a"] == 0) & (pr["type_b"] == 0))[0][:n],
                          np.nonzero((pr["type_a"] == 0) & (pr["type_b"] == 1))[0][:n // 2],
                          np.nonzero((pr["type_a"] == 1) & (pr["type_b"] == 1))[0][:n // 8]])
    for i in idx:
        yield tuple((pr["type_" + s][i], pr["pos_" + s][i], pr["quat_" + s][i], pr["half_" + s][i]) for s in "ab")


@pytest.fixture(scope="module")
def traced():
    params = _params()
    stats = {k: 0 for k in ("box_support", "sphere_support", "first_dir", "second_dir", "line_a", "line_b",
                            "line_mid", "tri_vertex", "tri_edge", "tri_face", "tetra_face", "found_inside",
                            "length_exit", "epa_closed", "epa_closest", "epa_contact", "epa_expand")}
    outcomes = []
    tried = 0
    for sa, sb in _pairs(240):
        tried += 1
        tr = O.trace_pair(sa, sb, params)
        if tr is None:
            continue
        outcomes.append((sa, sb, tr))
    # nearly every pair's trace fits (a polytope that runs away does not)
    assert len(outcomes) >= 0.97 * tried
    return params, outcomes, stats


def test_trace_agrees_with_the_narrowphase(traced):
    # the traced call is the oracle's CheckCollision itself
    params, outcomes, _ = traced
    keys = ("type", "pos", "quat", "half")
    pr = {k + "_" + s: np.array([o[0 if s == "a" else 1][j] for o in outcomes])
          for j, k in enumerate(keys) for s in "ab"}
    ref = O.narrowphase(pr, params, stats=True)
    for i, (_, _, tr) in enumerate(outcomes):
        assert tr["gjk"] == ref["gjk"][i]
        assert len(tr["gjk_recs"]) == ref["stats"][i, O.STAT_GJK_CALLS]
        if tr["gjk"] == O.GJK_FOUND:
            assert len(tr["epa_recs"]) == ref["stats"][i, O.STAT_EPA_ITERS]
            np.testing.assert_array_equal(tr["normal"], ref["normal"][i])
            np.testing.assert_array_equal(tr["point_a"], ref["point_a"][i])


def test_gjk_invariants(traced):
    _, outcomes, stats = traced
    for sa, sb, tr in outcomes:
        _check_gjk(tr["gjk_recs"], Shape(*sa), Shape(*sb), stats)
    # every common case of EvolveSimplex was exercised (a line case that keeps one
    # end, or a triangle reduced to its oldest vertex, is rare in these pairs and
    # checked when it occurs)
    for k in ("first_dir", "second_dir", "line_mid", "tri_edge", "tri_face", "tetra_face", "found_inside",
              "box_support", "sphere_support", "length_exit"):
        assert stats[k] >= 10, (k, stats)


def test_epa_invariants(traced):
    params, outcomes, stats = traced
    n = 0
    for sa, sb, tr in outcomes:
        if tr["gjk"] != O.GJK_FOUND:
            continue
        n += 1
        _check_epa(tr["epa_recs"], Shape(*sa), Shape(*sb), stats, float(F32(params.epa_tolerance)))
        last = tr["epa_recs"][-1]
        if tr["epa"] == O.EPA_CONTACT:
            assert last["converged"]
            np.testing.assert_array_equal(tr["normal"], (-last["dir"].astype(F32)).astype(F32))
    assert n >= 100
    for k in ("epa_closed", "epa_closest", "epa_contact", "epa_expand"):
        assert stats[k] >= 50, (k, stats)
