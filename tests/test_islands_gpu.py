"""The world step in the FP64-islands mode (GJK's length test and EPA's final face
and depth in binary64), plain and batched, bit for bit against the oracle in the
same mode: config 1 for 240 steps and a size-stratified handful of config-3
worlds for 40 steps, with their last contact lists."""
import dataclasses

import numpy as np
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S

from test_step_gpu import _contacts_equal, _state_equal

pytestmark = pytest.mark.gpu


def _islands(sc):
    sc.params = dataclasses.replace(sc.params, fp64_islands=1)
    return sc


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_islands_mode_bit_exact(oracle_mod, path):
    for sc, nsteps in ((_islands(S.config1()), 240), (_islands(S.config3(128, nsteps=1).subset([0, 17, 40, 77, 101])), 40)):
        ctx = runtime.Context(sc.params)
        ctx.upload(sc)
        ctx.record_contacts(True)
        (ctx.step_plain if path == "plain" else ctx.step_batched)(nsteps)
        got = ctx.download()
        for w in range(sc.nworlds):
            ws = oracle_mod.WorldState(sc, w, "f32")
            ws.run(nsteps - 1)
            ref = ws.run(1, contacts=True)
            a, b = int(sc.offsets[w]), int(sc.offsets[w + 1])
            _state_equal(got, ws, a, b, "%s world %d" % (path, w))
            _contacts_equal(ctx.contacts(w), ref, "%s world %d" % (path, w))
        assert ctx.stats()["epa_overflow"] == 0
        ctx.close()
