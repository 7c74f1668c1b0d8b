"""Register budget of the step kernels, from the compiler's resource report for
gfx950 (no GPU needed).  The batched step (64-512 threads per world, cold and
warm-started, reference and FP64-islands arithmetic) is compiled for three waves
per SIMD at 256 threads (168 VGPRs, so that three worlds share a CU) and two at
the other sizes: the test holds each instantiation to its occupancy, the
three-wave ones to the few dozen spilled registers that budget costs, and the
two-wave ones to none (a warm-started variant that once spilled 67 registers at
the two-wave budget ran at half the speed of the cold one).
Every variant of the plain step's solve keeps its registers without spills."""
import os
import re
import subprocess

import pytest

from nans_projekat_amd import build as B

pytestmark = pytest.mark.skipif(not os.path.exists(B.HIPCC), reason="hipcc not installed")


def _resources(src):
    out = os.path.join(B.OBJ_DIR, "resource_probe.o")
    os.makedirs(B.OBJ_DIR, exist_ok=True)
    cmd = [B.HIPCC, *B.FLAGS, "--cuda-device-only", "-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    kern = {}
    name = None
    for line in res.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            kern[name] = {}
            continue
        m = re.search(r"remark:\s+([A-Za-z /\[\]]+?): (\d+)", line)
        if m and name:
            kern[name][m.group(1).strip()] = int(m.group(2))
    return kern


def test_batched_step_has_no_spills():
    kern = _resources(os.path.join(B.HERE, "csrc", "nans_step.hip"))
    batched = {k: v for k, v in kern.items() if "k_batched_step" in k}
    assert len(batched) == 16, sorted(kern)
    for k, v in batched.items():
        block = int(re.search(r"k_batched_stepILi(\d+)E", k).group(1))
        if block != 256:
            assert v.get("Occupancy [waves/SIMD]", 0) >= 2, (k, v)
            assert v.get("VGPRs Spill") == 0, (k, v)
            assert v.get("ScratchSize [bytes/lane]") == 0, (k, v)
        else:
            assert v.get("Occupancy [waves/SIMD]", 0) >= 3, (k, v)
            # the default step (cold, reference arithmetic) and the opt-in modes
            cold = "Lb0ELb0E" in k
            assert v.get("VGPRs Spill") <= (40 if cold else 64), (k, v)
    # the plain step's solve: 1024 threads (large worlds) and 256 with contact terms in LDS
    solve = {k: v for k, v in kern.items() if "k_plain_solve" in k}
    assert len(solve) == 4, sorted(kern)
    for k, v in solve.items():
        assert v.get("VGPRs Spill") == 0, (k, v)
        assert v.get("ScratchSize [bytes/lane]") == 0, (k, v)
