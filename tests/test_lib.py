"""CPU checks of the C ABI: libnans_gpu.so and libnans_sim.so load and export
every function that include/nans_gpu.h and include/nans_sim.h declare, reports the reference's constants as defaults and
rejects bad arguments with an error code and message (no GPU work)."""
import ctypes
import os
import re

import pytest

from nans_projekat_amd import runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:nans_|Sim)\w+)\s*\(", text)))


def test_every_header_is_checked():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == \
        ["nans_gpu.h", "nans_sim.h"]


def test_library_exports_every_declared_function():
    lib = runtime.load()
    names = declared_functions("nans_gpu.h")
    assert len(names) >= 7
    for name in names:
        assert hasattr(lib, name), name


def test_plugin_exports_every_declared_function():
    from nans_projekat_amd import build
    if not os.path.exists(build.SIM):
        build.build_driver()
    lib = ctypes.CDLL(build.SIM)
    names = declared_functions("nans_sim.h")
    assert names == ["SimUpdateAndRender", "nans_sim_release"]
    for name in names:
        assert hasattr(lib, name), name


def test_default_params_are_reference_constants():
    p = runtime.Params()
    runtime.load().nans_default_params(ctypes.byref(p))
    f32 = lambda x: ctypes.c_float(x).value
    assert p.solver_iterations == 70                    # code/nans.cpp:1176
    assert p.beta == f32(0.3)                            # :1152
    assert p.restitution == f32(0.1)                     # :1154
    assert p.friction == f32(0.1)                        # :1159
    assert p.damping == f32(1.5)                         # code/nans.h:46
    assert p.gravity == f32(9.81)                        # code/nans.cpp:68
    assert p.max_gjk_iterations == 64 and p.max_epa_iterations == 64   # code/nans.h:54,56
    assert p.epa_tolerance == f32(0.001)                 # code/nans.h:55
    assert tuple(p.floor_pos) == (f32(1.2), f32(-0.5), f32(1.0))       # code/nans.cpp:1669
    assert tuple(p.floor_scale) == (100.0, 1.0, 100.0)   # code/nans.cpp:1878-1879
    assert p.floor_mass == 100000.0                      # code/nans.cpp:1676
    assert p.floor_moi == f32(f32(100000.0 / 12.0) * 20000.0)  # code/nans.cpp:1677-1678
    assert p.fp64_islands == 0


def test_bad_arguments_return_codes():
    lib = runtime.load()
    assert lib.nans_create(None, None) == runtime.ERR_INVALID
    assert b"NULL" in lib.nans_last_error()
    assert lib.nans_narrowphase(None, ctypes.c_int64(1), None, None, None, None) == runtime.ERR_INVALID
    assert lib.nans_set_params(None, None) == runtime.ERR_INVALID
    assert lib.nans_destroy(None) == runtime.OK
    assert lib.nans_rollout(None, 240, 16) == runtime.ERR_STATE
    assert lib.nans_adjoint(None, None) == runtime.ERR_STATE
