"""GJK's termination test length(d) <= 0.0001f (code/nans.cpp:755-767) is evaluated
on the GPU as dot(d, d) <= T with T the largest binary32 whose correctly rounded
square root is <= 0.0001f.  Correctly rounded sqrt is monotone, so the two forms
agree for every input; this pins T."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_threshold_is_the_last_float_with_small_root():
    src = open(os.path.join(ROOT, "nans_projekat_amd", "csrc", "nans_narrow.hpp")).read()
    bits = int(re.search(r"GJK_LEN2_MAX_BITS = (0x[0-9a-fA-F]+)u", src).group(1), 16)
    T = np.array([bits], np.uint32).view(np.float32)[0]
    lim = np.float32(0.0001)
    assert np.sqrt(T) <= lim
    assert not (np.sqrt(np.nextafter(T, np.float32(np.inf))) <= lim)
    # and the two forms agree around it and far from it
    xs = np.array([0, T, np.nextafter(T, np.float32(0)), np.nextafter(T, np.float32(1)), 1e-9, 2e-8, 1.0, np.nan],
                  np.float32)
    assert np.array_equal(np.sqrt(xs) <= lim, xs <= T)
