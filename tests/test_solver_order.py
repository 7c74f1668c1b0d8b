"""SolveConstraints' serial list order (code/nans.cpp:1538-1548), pinned by
impulse-momentum algebra on two-contact worlds.

Each contact of the reference's solver reads the velocities the earlier contacts
of the list left behind.  For one contact with an approaching normal velocity,
the applied normal impulse brings the relative normal velocity to
    after = Cr * before - (beta / dt) * Depth                      (1)
(exact algebra of lambda = (-J dV + B) M_eff; pinned on its own by
test_oracle.test_post_impulse_normal_velocity).  Friction impulses lie along the
tangents and angular impulses enter J dV only through W x N (code/nans.cpp:1149),
both perpendicular to N, so (1) holds for any friction when the angular terms of
the effective mass vanish (huge moments of inertia here).

Two worlds, each with two contacts that share one body (the other body of the
pair being touched by that contact alone):
  floor, box, sphere on the box:   list [CF(box), CS(box, sphere)]
  floor, box, box on the box:      list [CC(bottom, top), CF(bottom)]
The body touched by one contact only tells, by conservation of linear momentum
in that contact, the shared body's velocity between the two contacts:
    v_shared_mid = v_shared_before_that_contact - (m_other / m_shared) * dv_other
With it, (1) must hold for each contact with "before" read at its turn in the
list.  Reversing the list, or reading every contact's velocities before the solve
(a Jacobi sweep), breaks (1) for at least one contact by a margin far above the
tolerance (checked as well)."""
import numpy as np

from nans_projekat_amd import scenes as S

HUGE_MOI = 1e12


def _params():
    p = S.StepParams(dt=1.0 / 120.0, restitution=0.2, friction=0.5, solver_iterations=10)
    p.floor_mass = 1e20  # the one-sided floor contact's effective mass is then the box's alone
    p.floor_moi = 1e30
    return p


def _run(oracle_mod, sc):
    """(contacts, post-solve velocities, pre-solve velocities): the latter from a
    contact-free twin (RK4 acts on each body alone)."""
    ws = oracle_mod.WorldState(sc, 0, "f64")
    con = ws.run(1, contacts=True)
    far = sc.copy()
    far.pos = far.pos + np.arange(sc.nbodies, dtype=np.float32)[:, None] * np.float32(10.0) + np.float32(200.0)
    free = oracle_mod.WorldState(far, 0, "f64")
    assert len(free.run(1, contacts=True)) == 0
    return con, ws.vel.copy(), free.vel.copy()


def _frame(c):
    n = np.asarray(c["normal"], float)
    n = n / np.linalg.norm(n)
    depth = float(np.dot(np.asarray(c["point_a"], float) - np.asarray(c["point_b"], float), n))
    return n, depth


def _law(before, depth, p):
    """(1): the relative normal velocity after the contact's impulse."""
    cr, k = float(np.float32(p.restitution)), float(np.float32(p.beta)) / float(np.float32(p.dt))
    return cr * before - k * depth


def box_and_sphere_world():
    sc = S._world([(3.0, 0.49, 2.0)], np.ones((1, 3), np.float32), [1.5], [(3.05, 1.475, 2.03)],
                  [(-0.4, -3.0, 0.25)], [0.5], [2.0], box_vel=[(0.3, -1.0, 0.0)])
    sc.moi[:] = HUGE_MOI
    sc.params = _params()
    return sc


def box_on_box_world():
    sc = S._world([(3.0, 0.49, 2.0), (3.1, 1.475, 2.05)], np.ones((2, 3), np.float32), [1.0, 0.75],
                  np.zeros((0, 3)), np.zeros((0, 3)), [], [], box_vel=[(0.3, -1.0, 0.0), (-0.2, -3.0, 0.1)])
    sc.moi[:] = HUGE_MOI
    sc.params = _params()
    return sc


def _residuals_box_and_sphere(oracle_mod):
    sc = box_and_sphere_world()
    p = sc.params
    con, vf, vp = _run(oracle_mod, sc)
    kinds = [(int(c["type"]), int(c["index_a"]), int(c["index_b"])) for c in con]
    assert kinds == [(oracle_mod.CF, 0, 0), (oracle_mod.CS, 0, 0)], kinds
    assert con[0]["dl_normal"] > 0 and con[1]["dl_normal"] > 0
    (n1, d1), (n2, d2) = _frame(con[0]), _frame(con[1])
    mb, ms = float(sc.mass[0]), float(sc.mass[1])
    box_mid = vf[0] + (ms / mb) * (vf[1] - vp[1])  # the box between CF and CS
    serial = [np.dot(box_mid, n1) - _law(np.dot(vp[0], n1), d1, p),
              np.dot(vf[0] - vf[1], n2) - _law(np.dot(box_mid - vp[1], n2), d2, p)]
    # the same laws if CS had read the pre-solve velocities (reversed or Jacobi)
    other = [np.dot(vf[0] - vf[1], n2) - _law(np.dot(vp[0] - vp[1], n2), d2, p)]
    return serial, other


def _residuals_box_on_box(oracle_mod):
    sc = box_on_box_world()
    p = sc.params
    con, vf, vp = _run(oracle_mod, sc)
    kinds = [(int(c["type"]), int(c["index_a"]), int(c["index_b"])) for c in con]
    assert kinds == [(oracle_mod.CC, 0, 1), (oracle_mod.CF, 0, 0)], kinds
    assert con[0]["dl_normal"] > 0 and con[1]["dl_normal"] > 0
    (n1, d1), (n2, d2) = _frame(con[0]), _frame(con[1])
    m0, m1 = float(sc.mass[0]), float(sc.mass[1])
    bottom_mid = vp[0] - (m1 / m0) * (vf[1] - vp[1])  # the bottom box between CC and CF
    serial = [np.dot(bottom_mid - vf[1], n1) - _law(np.dot(vp[0] - vp[1], n1), d1, p),
              np.dot(vf[0], n2) - _law(np.dot(bottom_mid, n2), d2, p)]
    other = [np.dot(vf[0], n2) - _law(np.dot(vp[0], n2), d2, p)]
    return serial, other


TOL = 1e-7  # m/s; binary64 with 1e-12 angular terms


def test_box_then_sphere_serial_order(oracle_mod):
    serial, other = _residuals_box_and_sphere(oracle_mod)
    assert max(abs(r) for r in serial) < TOL, serial
    assert min(abs(r) for r in other) > 1e4 * TOL, other  # the check has the power to tell


def test_stack_serial_order(oracle_mod):
    serial, other = _residuals_box_on_box(oracle_mod)
    assert max(abs(r) for r in serial) < TOL, serial
    assert min(abs(r) for r in other) > 1e4 * TOL, other


def test_first_step_of_the_warm_start_is_the_same_solve(oracle_mod):
    """The warm-started step with an empty carried list is the cold step: its
    solver loop obeys the same order (same residuals)."""
    for make in (box_and_sphere_world, box_on_box_world):
        sc = make()
        cold = oracle_mod.WorldState(sc, 0, "f64")
        cold.run(1)
        warm = oracle_mod.WorldState(sc, 0, "f64")
        warm.enable_warm(64)
        warm.run(1)
        np.testing.assert_array_equal(warm.vel, cold.vel)
