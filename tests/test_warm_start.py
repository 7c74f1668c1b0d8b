"""The opt-in warm start in the oracle (no counterpart in the reference, whose
DetectCollisions leaves "Warm starting?" as a TODO, code/nans.cpp:1374-1393):
with no room for carried contacts it is the cold step; for a sphere resting on the
ground the carried normal accumulator grows by (iterations x applied impulse)
every step, as the accumulate-and-clamp loop of code/nans.cpp:1176-1238 implies
when each step starts from the previous step's sum.  GPU parity is in
test_warm_start_gpu.py."""
import dataclasses

import numpy as np

from nans_projekat_amd import scenes as S


def test_no_room_is_the_cold_step(oracle_mod):
    sc = S.config1()
    cold = oracle_mod.WorldState(sc, 0, "f32")
    warm = oracle_mod.WorldState(sc, 0, "f32")
    warm.enable_warm(0)
    cold.run(60)
    warm.run(60)
    for f in ("pos", "vel", "angvel", "angles"):
        np.testing.assert_array_equal(getattr(warm, f), getattr(cold, f))


def _resting_sphere():
    # off the ground box's axis: a centre straight above the ground's centre gives GJK a
    # degenerate first direction and no contact
    sc = S._world(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), [(2.0, 0.499, 1.7)], [(0.0, 0.0, 0.0)],
                  [0.5], [2.0])
    sc.params = dataclasses.replace(S.StepParams(), solver_iterations=10, dt=1.0 / 120.0)
    return sc


def test_resting_contact_accumulates_across_steps(oracle_mod):
    sc = _resting_sphere()
    K = sc.params.solver_iterations
    for warm in (False, True):
        ws = oracle_mod.WorldState(sc, 0, "f64")
        if warm:
            ws.enable_warm(16)
        run, total, prev = 0, 0.0, 0.0
        for t in range(120):
            con = ws.run(1, contacts=True)
            if len(con) != 1:
                run, total, prev = 0, 0.0, 0.0
                continue
            dln = float(con["dl_normal"][0])
            assert con["type"][0] == 4 and dln > 0           # SF, pushing
            total += dln
            run += 1
            if warm:
                wp, wsum, wn, _ = ws.warm
                assert wn[0] == 1 and wp[0] == 0             # the one SF pair (canonical index 0)
                s = float(wsum[0, 0])
                # the sum carried in grows by K applied impulses
                assert abs(s - (prev + K * dln)) <= 1e-9 * max(1.0, s)
                prev = s
        assert run >= 60                                      # the contact persisted
        if warm:
            assert abs(prev - K * total) <= 1e-8 * prev       # the whole run's impulses, not one step's
