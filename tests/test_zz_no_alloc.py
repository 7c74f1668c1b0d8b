"""nans_narrowphase_noalloc (its EPA queue in a static device array, no device
allocation) against nans_narrowphase on the same pairs, bit for bit: two calls
in a row, the second on fewer pairs than the first, so that the second runs on
scratch the first left behind; sizes beyond the scratch are refused."""
import numpy as np
import pytest

from nans_projekat_amd import runtime
from nans_projekat_amd import scenes as S

pytestmark = pytest.mark.gpu

FIELDS = ("gjk", "contact", "status", "normal", "depth", "point_a", "point_b")


def _pairs(n, seed):
    return S.config2_pairs(n, seed=seed)


def test_noalloc_matches_narrowphase():
    ctx = runtime.Context(S.StepParams())
    for n, seed in ((runtime.NARROWPHASE_NOALLOC_MAX_PAIRS, 11), (3001, 12)):
        print("noalloc narrowphase: %d config-2 pairs, seed %d" % (n, seed), flush=True)
        dp = ctx.upload_pairs(_pairs(n, seed))
        ref = ctx.narrow_to_numpy(ctx.narrowphase(dp))
        got = ctx.narrow_to_numpy(ctx.narrowphase(dp, noalloc=True))
        for k in FIELDS:
            np.testing.assert_array_equal(got[k], ref[k], err_msg="%d pairs: %s" % (n, k))
        assert got["n_overflow"] == ref["n_overflow"]
        assert int(np.sum(got["contact"])) > 0
    ctx.close()


def test_noalloc_refuses_more_pairs_than_its_scratch():
    ctx = runtime.Context(S.StepParams())
    n = runtime.NARROWPHASE_NOALLOC_MAX_PAIRS + 1
    dp = ctx.upload_pairs(_pairs(n, 13))
    with pytest.raises(runtime.NansError) as e:
        ctx.narrowphase(dp, noalloc=True)
    assert e.value.code == runtime.ERR_CAPACITY
    ctx.close()
