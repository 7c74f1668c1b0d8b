"""MI355X-native (gfx950) simulation step of the nans_projekat rigid-body engine.

The package holds the HIP kernels and C ABI (csrc/, built into libnans_gpu.so by
build.py), the host binding of that ABI (runtime.py) and the seeded scene
generators shared by the GPU path and the checker (scenes.py).
"""
from . import scenes  # noqa: F401

__all__ = ["scenes"]
