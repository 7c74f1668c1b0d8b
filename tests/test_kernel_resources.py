"""Register budget of the step kernels, from the compiler's resource report for
gfx950 (no GPU needed): every instantiation of the batched step (64-512 threads
per world, cold and warm-started, reference and FP64-islands arithmetic) keeps two waves per SIMD with no VGPR spills
and no scratch memory, and so does every variant of the plain step's solve.  A spill there costs the step its speed (a warm-started
variant that spilled 67 registers ran at half the speed of the cold one)."""
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
        assert v.get("VGPRs Spill") == 0, (k, v)
        assert v.get("ScratchSize [bytes/lane]") == 0, (k, v)
        assert v.get("Occupancy [waves/SIMD]", 0) >= 2, (k, v)
    # the plain step's solve: 1024 threads (large worlds) and 256 with contact terms in LDS
    solve = {k: v for k, v in kern.items() if "k_plain_solve" in k}
    assert len(solve) == 4, sorted(kern)
    for k, v in solve.items():
        assert v.get("VGPRs Spill") == 0, (k, v)
        assert v.get("ScratchSize [bytes/lane]") == 0, (k, v)
