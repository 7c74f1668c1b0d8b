"""Helpers for the golden fixtures under tests/golden/."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def state_hash(state: dict) -> str:
    """SHA-256 over the binary32 bytes of pos, vel, angvel, angles ([N][3] each),
    as tests/golden/make_state_hashes.py computes it from the oracle."""
    h = hashlib.sha256()
    for k in ("pos", "vel", "angvel", "angles"):
        h.update(np.ascontiguousarray(state[k], dtype=np.float32).tobytes())
    return h.hexdigest()


def load_state_hash(key: str) -> dict:
    path = os.path.join(GOLDEN, "state_hashes.json")
    if not os.path.exists(path):
        pytest.fail("tests/golden/state_hashes.json is missing (tests/golden/make_state_hashes.py writes it)")
    data = json.load(open(path))
    if key not in data:
        pytest.fail("%s is missing from tests/golden/state_hashes.json" % key)
    return data[key]
