"""The GJK/EPA invariant checks catch planted mistakes.

Each case copies the oracle's sources, plants one plausible transcription
mistake in EvolveSimplex, its helpers, the supports or ResolveCollision, builds
the copy with gcc and runs the call-by-call checks of test_gjk_epa_invariants on
config-2 pairs: at least one check must fail (or the polytopes must run away,
so that the traces no longer fit).  A mistake that reverses the new faces' winding
(PushTriangle(ns, B, A)) makes the polytope grow without bound and is left out:
it cannot finish in a test's time.  Mistakes in Constraint's quirks (Q11-Q13:
the subtracted body-B term, W x N, the unnormalised tangent) must fail one of the
closed-form checks of test_constraint_quirks.  Mistakes in SolveConstraints'
list order (reversed, or every contact reading the velocities of before the
solve), planted in all four of the oracle's solver loops at once, must fail the
impulse-momentum checks of test_solver_order."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
import test_constraint_quirks as Q
import test_gjk_epa_invariants as T
import test_epa_ties as ET
import test_solver_order as SO

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


# SolveConstraints' list order (code/nans.cpp:1538-1548), changed the same way in
# all four solver loops of the oracle (cold step, warm step, the frozen recording
# and the frozen replay), so that the cross-checks between those loops still
# agree; test_solver_order's closed forms must catch them
_LOOPS = ("for (int k = 0; k < cl.n; ++k)\n                SUF(constraint)(&w, &cl.c[k], p, NULL);",
          "for (int k = 0; k < cl.n; ++k)  /* in list order */",
          "for (int c = 0; c < cl.n; ++c)  /* SolveConstraints in list order */",
          "for (int c = 0; c < step_count[s]; ++c) {  /* SolveConstraints in list order */")
_SNAP = ("static void SUF(snap)(World *w)\n{\n"
         "    for (int i = 0; i < w->nb; ++i) { w->box[i].V0 = w->box[i].V; w->box[i].W0 = w->box[i].W; }\n"
         "    for (int i = 0; i < w->ns; ++i) { w->sph[i].V0 = w->sph[i].V; w->sph[i].W0 = w->sph[i].W; }\n"
         "    w->floor.V0 = w->floor.V; w->floor.W0 = w->floor.W;\n}\n")
SOLVER_MUTANTS = {
    "solver_reverse_order": [
        (_LOOPS[0], _LOOPS[0].replace("int k = 0; k < cl.n; ++k", "int k = cl.n - 1; k >= 0; --k")),
        (_LOOPS[1], _LOOPS[1].replace("int k = 0; k < cl.n; ++k", "int k = cl.n - 1; k >= 0; --k")),
        (_LOOPS[2], _LOOPS[2].replace("int c = 0; c < cl.n; ++c", "int c = cl.n - 1; c >= 0; --c")),
        (_LOOPS[3], _LOOPS[3].replace("int c = 0; c < step_count[s]; ++c", "int c = step_count[s] - 1; c >= 0; --c")),
    ],
    # every contact reads the velocities of before the solve (a Jacobi sweep)
    "solver_reads_presolve_velocities": [
        ("    V3 verts[8];\n} SUF(Body);", "    V3 verts[8];\n    V3 V0, W0;\n} SUF(Body);"),
        ("#define World SUF(World)\n", "#define World SUF(World)\n" + _SNAP),
        ("V3 V1 = A->V, V2 = B->V, W1 = A->W, W2 = B->W;", "V3 V1 = A->V0, V2 = B->V0, W1 = A->W0, W2 = B->W0;"),
    ] + [(loop, "SUF(snap)(&w);\n" + loop) for loop in _LOOPS],
}


def _build(tmp, name):
    if name in SOLVER_MUTANTS:
        edits = SOLVER_MUTANTS[name]
    else:
        edits = [MUTANTS[name] if name in MUTANTS else CONSTRAINT_MUTANTS[name]]
    d = os.path.join(str(tmp), name)
    os.makedirs(d)
    for f in ("nans_oracle.h", "nans_oracle_impl.h", "nans_oracle_f32.c", "nans_oracle_f64.c"):
        shutil.copy(os.path.join(SRC, f), d)
    p = os.path.join(d, "nans_oracle_impl.h")
    s = open(p).read()
    for old, new in edits:
        assert s.count(old) == 1, (name, old)
        s = s.replace(old, new)
    open(p, "w").write(s)
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


@pytest.mark.parametrize("name", sorted(SOLVER_MUTANTS))
def test_planted_solver_order_mistake_is_caught(tmp_path, monkeypatch, name):
    """The mutant keeps the oracle's loops consistent with each other (the frozen
    replay still equals the step), yet the closed forms of test_solver_order fail."""
    monkeypatch.setattr(O, "LIB_PATH", _build(tmp_path, name))
    failed = 0
    for check in (SO.test_box_then_sphere_serial_order, SO.test_stack_serial_order):
        try:
            check(O)
        except AssertionError:
            failed += 1
    assert failed == 2, name


def test_planted_tie_rule_mistake_is_caught(tmp_path, monkeypatch):
    """A '<=' closest-face scan (ties to the last face) fails test_epa_ties."""
    name = "epa_ties_to_last_face"
    MUTANTS_TIES = {name: ("if (dist < cur) {", "if (dist <= cur) {")}
    monkeypatch.setitem(MUTANTS, name, MUTANTS_TIES[name])
    monkeypatch.setattr(O, "LIB_PATH", _build(tmp_path, name))
    with pytest.raises(AssertionError):
        ET.test_closest_face_ties_go_to_the_first_face(O)
