"""The GJK/EPA invariant checks catch planted mistakes.

Each case copies the oracle's sources, plants one plausible transcription
mistake in EvolveSimplex, its helpers, the supports or ResolveCollision, builds
the copy with gcc and runs the call-by-call checks of test_gjk_epa_invariants on
config-2 pairs: at least one check must fail (or the polytopes must run away,
so that the traces no longer fit).  A mistake that reverses the new faces' winding
(PushTriangle(ns, B, A)) makes the polytope grow without bound and is left out:
it cannot finish in a test's time.  Mistakes in Constraint's quirks (Q11-Q13:
the subtracted body-B term, W x N, the unnormalised tangent) must fail one of the
closed-form checks of test_constraint_quirks."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
import test_constraint_quirks as Q
import test_gjk_epa_invariants as T

SRC = os.path.dirname(O.__file__)

MUTANTS = {
    "triangle_refers_to_newest": ("V3 n2o = vneg(S[0].P);", "V3 n2o = vneg(S[2].P);"),
    "line_u_v_swapped": ("*V = vdot(vneg(A), seg) / len;\n    *U = vdot(B, seg) / len;",
                         "*U = vdot(vneg(A), seg) / len;\n    *V = vdot(B, seg) / len;"),
    "second_direction_not_negated": ("dir = vmul(dir, (REAL)-1);\n        break;", "dir = vmul(dir, (REAL)1);\n        break;"),
    "support_tie_to_last_corner": ("if (dist > maxd) {", "if (dist >= maxd) {"),
    "epa_farthest_face": ("if (dist < cur) {", "if (dist > cur) {"),
    "epa_normal_not_flipped": ("if (vdot(t.A.P, t.N) < (REAL)0)\n        t.N = vmul(t.N, (REAL)-1);",
                               "if (0)\n        t.N = vmul(t.N, (REAL)-1);"),
    "horizon_keeps_reversed_edges": ("if (veq(p->e[i].A.P, B.P) && veq(p->e[i].B.P, A.P)) {",
                                     "if (veq(p->e[i].A.P, A.P) && veq(p->e[i].B.P, B.P)) {"),
    "visibility_sign": ("if (vdot(nrm, temp) > (REAL)0) {", "if (vdot(nrm, temp) < (REAL)0) {"),
    "tetra_erases_wrong_vertex": ("if (vdot(abd, D0) > (REAL)0) {\n            SUF(erase)(S, n, 2);",
                                  "if (vdot(abd, D0) > (REAL)0) {\n            SUF(erase)(S, n, 1);"),
    "barycentric_v_w_swapped": ("*V = (d11 * d20 - d01 * d21) / denom;\n    *W = (d00 * d21 - d01 * d20) / denom;",
                                "*W = (d11 * d20 - d01 * d21) / denom;\n    *V = (d00 * d21 - d01 * d20) / denom;"),
    "support_of_b_along_plus_d": ("v.SupB = SUF(shape_support)(b, vmul(d, (REAL)-1));",
                                  "v.SupB = SUF(shape_support)(b, d);"),
    "triple_cross_sign": ("return vsub(vmul(B, vdot(C, A)), vmul(A, vdot(C, B)));",
                          "return vsub(vmul(A, vdot(C, B)), vmul(B, vdot(C, A)));"),
    "epa_convergence_inverted": ("if (vdot(p.t[best].N, ns.P) - cur < (REAL)np->epa_tolerance) {",
                                 "if (vdot(p.t[best].N, ns.P) - cur > (REAL)np->epa_tolerance) {"),
    "triangle_edge_tests_swapped": ("if (vdot(e2n, n2o) > (REAL)0) {", "if (vdot(e1n, n2o) > (REAL)0) {"),
    "triangle_swap_skipped": ("Vert t = S[1];\n                S[1] = S[2];\n                S[2] = t;",
                              "Vert t = S[1];\n                (void)t;"),
}


# Constraint (code/nans.cpp:1020-1329), checked by test_constraint_quirks
CONSTRAINT_MUTANTS = {
    "effective_mass_adds_body_b": ("JMJn += InvI1 * vdot(RN1, RN1) - InvI2 * vdot(vneg(RN2), vneg(RN2));",
                                   "JMJn += InvI1 * vdot(RN1, RN1) + InvI2 * vdot(vneg(RN2), vneg(RN2));"),
    "relative_velocity_w_cross_r": ("V3 dVn = vsub(vsub(vadd(V1, vcross(W1, N)), V2), vcross(W2, N));",
                                    "V3 dVn = vsub(vsub(vadd(V1, vcross(W1, R1)), V2), vcross(W2, R2));"),
    "tangent_normalised": ("/* glm::normalize(T1) result discarded (nans.cpp:1133) */",
                           "T1 = vnormalize(T1);"),
}


def _build(tmp, name):
    old, new = MUTANTS[name] if name in MUTANTS else CONSTRAINT_MUTANTS[name]
    d = os.path.join(str(tmp), name)
    os.makedirs(d)
    for f in ("nans_oracle.h", "nans_oracle_impl.h", "nans_oracle_f32.c", "nans_oracle_f64.c"):
        shutil.copy(os.path.join(SRC, f), d)
    p = os.path.join(d, "nans_oracle_impl.h")
    s = open(p).read()
    assert s.count(old) == 1, name
    open(p, "w").write(s.replace(old, new))
    lib = os.path.join(d, "libnans_oracle_mutant.so")
    subprocess.run(["gcc", "-std=c11", "-fPIC", "-O2", "-ffp-contract=off", "-fno-fast-math", "-msse2",
                    "-mfpmath=sse", "-shared", "-o", lib, os.path.join(d, "nans_oracle_f32.c"),
                    os.path.join(d, "nans_oracle_f64.c"), "-lm"], check=True, capture_output=True)
    return lib


def _caught(params):
    stats = dict.fromkeys(("box_support", "sphere_support", "first_dir", "second_dir", "line_a", "line_b",
                           "line_mid", "tri_vertex", "tri_edge", "tri_face", "tetra_face", "found_inside",
                           "length_exit", "epa_closed", "epa_closest", "epa_contact", "epa_expand"), 0)
    tol = float(np.float32(params.epa_tolerance))
    tried = fit = 0
    for sa, sb in T._pairs(120):
        tried += 1
        tr = O.trace_pair(sa, sb, params)
        if tr is None:
            continue
        fit += 1
        A, B = T.Shape(*sa), T.Shape(*sb)
        try:
            T._check_gjk(tr["gjk_recs"], A, B, stats)
            if tr["gjk"] == O.GJK_FOUND:
                T._check_epa(tr["epa_recs"], A, B, stats, tol)
        except AssertionError:
            return True
    return fit < 0.97 * tried


def test_the_true_oracle_passes():
    assert not _caught(T._params())


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_is_caught(tmp_path, monkeypatch, name):
    monkeypatch.setattr(O, "LIB_PATH", _build(tmp_path, name))
    assert _caught(T._params()), name


@pytest.mark.parametrize("name", sorted(CONSTRAINT_MUTANTS))
def test_planted_constraint_mistake_is_caught(tmp_path, monkeypatch, name):
    monkeypatch.setattr(O, "LIB_PATH", _build(tmp_path, name))
    failed = 0
    for check in (Q.test_impulses_do_not_depend_on_spin, Q.test_effective_mass_signs, Q.test_unnormalised_tangents):
        try:
            check(O)
        except AssertionError:
            failed += 1
    assert failed >= 1, name
