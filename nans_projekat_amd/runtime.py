"""Host-side binding of the C ABI (include/nans_gpu.h) over ctypes.

This is the Python mirror of the reference's plugin boundary: the platform
layer of the reference binds one C symbol with dlsym (code/sdl_nans.cpp:399-407)
and calls it once per frame; here a Context binds libnans_gpu.so and exposes
the same physics block (code/nans.cpp:1758-1762) as methods.  Device memory is
handed over as torch tensors (PyTorch is only the allocator and stream here).
The library is never replaced by a CPU path: if libnans_gpu.so is missing or
no GPU is present, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import scenes as _scenes

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnans_gpu.so")

OK, ERR_INVALID, ERR_HIP, ERR_CAPACITY, ERR_STATE = 0, -1, -2, -3, -4
SHAPE_BOX, SHAPE_SPHERE = 0, 1
CC, CS, CF, SS, SF = 0, 1, 2, 3, 4
GJK_NO, GJK_FOUND, GJK_CAP = 0, 1, 2
EPA_NONE, EPA_CONTACT, EPA_BARY, EPA_NONFINITE, EPA_ITER, EPA_EMPTY, EPA_OVERFLOW = range(7)
NARROWPHASE_NOALLOC_MAX_PAIRS = 262144


class NansError(RuntimeError):
    """A C-ABI call returned an error code."""

    def __init__(self, code: int, msg: str):
        super().__init__("nans error %d: %s" % (code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("solver_iterations", ctypes.c_int32), ("beta", ctypes.c_float), ("restitution", ctypes.c_float),
                ("friction", ctypes.c_float), ("damping", ctypes.c_float), ("gravity", ctypes.c_float),
                ("dt", ctypes.c_float), ("max_gjk_iterations", ctypes.c_int32), ("max_epa_iterations", ctypes.c_int32),
                ("epa_tolerance", ctypes.c_float), ("fp64_islands", ctypes.c_int32),
                ("floor_pos", ctypes.c_float * 3), ("floor_scale", ctypes.c_float * 3),
                ("floor_mass", ctypes.c_float), ("floor_moi", ctypes.c_float), ("device", ctypes.c_int32)]


class ShapesSoa(ctypes.Structure):
    _fields_ = [("type", ctypes.c_void_p), ("pos", ctypes.c_void_p), ("quat", ctypes.c_void_p),
                ("half", ctypes.c_void_p)]


class NarrowOut(ctypes.Structure):
    _fields_ = [("gjk", ctypes.c_void_p), ("contact", ctypes.c_void_p), ("normal", ctypes.c_void_p),
                ("depth", ctypes.c_void_p), ("point_a", ctypes.c_void_p), ("point_b", ctypes.c_void_p),
                ("status", ctypes.c_void_p)]


class BodiesSoa(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("pos", "vel", "angvel", "angles", "force", "torque", "scale", "mass",
                                               "moi", "kind")]


class Stats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("pair_tests", "candidates", "gjk_found", "contacts", "epa_overflow",
                                              "capacity_overflow")]


class Timing(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int32), ("elapsed_ms", ctypes.c_float), ("threads_per_world", ctypes.c_int32),
                ("workgroups", ctypes.c_int32), ("lds_bytes", ctypes.c_int32), ("wave_world_bodies", ctypes.c_int32),
                ("wave_worlds", ctypes.c_int32)]


TUNE_SIZES = (64, 128, 256, 512)
TUNE_WAVE_BOUNDS = (0, 16, 32, 64, 128)


class Tuning(ctypes.Structure):
    _fields_ = [("threads_per_world", ctypes.c_int32), ("steps", ctypes.c_int32), ("ms_per_step", ctypes.c_float * 4),
                ("wave_world_bodies", ctypes.c_int32), ("ms_per_step_wave", (ctypes.c_float * 5) * 4)]


class AdjointIo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("dl_pos", "dl_vel", "dl_angvel", "dl_angles", "dl_pos0", "dl_vel0",
                                               "dl_angvel0", "dl_angles0", "dl_mass", "dl_moi", "dl_friction",
                                               "dl_restitution")]


# nans_contact (include/nans_gpu.h) as a numpy record
CONTACT_DTYPE = np.dtype([("type", "<i4"), ("index_a", "<i4"), ("index_b", "<i4"), ("pad", "<i4"),
                          ("point_a", "<f4", 3), ("point_b", "<f4", 3), ("normal", "<f4", 3),
                          ("dl_normal", "<f4"), ("dl_tangent1", "<f4"), ("dl_tangent2", "<f4")])

_lib = None


def load() -> ctypes.CDLL:
    """Load libnans_gpu.so (built in-tree by nans_projekat_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("%s is missing: run `python -m nans_projekat_amd.build`" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        lib.nans_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        raise NansError(rc, load().nans_last_error().decode(errors="replace"))


def to_params(sp: "_scenes.StepParams | None" = None, device: int = 0, **overrides) -> Params:
    """nans_params from a StepParams (reference defaults when None)."""
    p = Params()
    load().nans_default_params(ctypes.byref(p))
    if sp is not None:
        for name in ("solver_iterations", "beta", "restitution", "friction", "damping", "gravity", "dt",
                     "max_gjk_iterations", "max_epa_iterations", "epa_tolerance", "fp64_islands",
                     "floor_mass", "floor_moi"):
            setattr(p, name, getattr(sp, name))
        p.floor_pos = (ctypes.c_float * 3)(*sp.floor_pos)
        p.floor_scale = (ctypes.c_float * 3)(*sp.floor_scale)
    for k, v in overrides.items():
        setattr(p, k, v)
    p.device = device
    return p


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _np(a) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if a is None else a.ctypes.data)


class Context:
    """Owns a nans_ctx.  Calls are ordered on torch's current stream of the
    context's device, so tensors produced by torch are ready when a call runs."""

    def __init__(self, params: "_scenes.StepParams | None" = None, device: int = 0, **overrides):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("no GPU: the HIP step has no CPU fallback")
        self.torch = torch
        self.device = device
        self.lib = load()
        self.params = to_params(params, device, **overrides)
        h = ctypes.c_void_p()
        check(self.lib.nans_create(ctypes.byref(h), ctypes.byref(self.params)))
        self.h = h
        self._sync_stream()

    def _sync_stream(self):
        s = self.torch.cuda.current_stream(self.device).cuda_stream
        check(self.lib.nans_set_stream(self.h, ctypes.c_void_p(s)))

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.nans_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: "_scenes.StepParams | None" = None, **overrides):
        self.params = to_params(params, self.device, **overrides)
        check(self.lib.nans_set_params(self.h, ctypes.byref(self.params)))

    # ------------------------------------------------------------------ narrowphase
    def upload_pairs(self, pairs: dict) -> dict:
        """Host pairs ([n][3]/[n][4] numpy, scenes.config2_pairs layout) -> device
        SoA planes ([3][n]/[4][n] torch tensors)."""
        torch = self.torch
        dev = "cuda:%d" % self.device
        out = {}
        for side in ("a", "b"):
            out["type_" + side] = torch.from_numpy(np.ascontiguousarray(pairs["type_" + side], np.int32)).to(dev)
            for k in ("pos", "quat", "half"):
                arr = np.ascontiguousarray(np.asarray(pairs["%s_%s" % (k, side)], np.float32).T)
                out["%s_%s" % (k, side)] = torch.from_numpy(arr).to(dev)
        out["n"] = len(pairs["type_a"])
        return out

    def alloc_narrow_out(self, n: int) -> dict:
        torch = self.torch
        dev = "cuda:%d" % self.device
        return dict(gjk=torch.empty(n, dtype=torch.int32, device=dev),
                    contact=torch.empty(n, dtype=torch.int32, device=dev),
                    normal=torch.empty((3, n), dtype=torch.float32, device=dev),
                    depth=torch.empty(n, dtype=torch.float64, device=dev),
                    point_a=torch.empty((3, n), dtype=torch.float32, device=dev),
                    point_b=torch.empty((3, n), dtype=torch.float32, device=dev),
                    status=torch.empty(n, dtype=torch.int32, device=dev))

    def narrowphase(self, dpairs: dict, out: dict | None = None, noalloc: bool = False) -> dict:
        """CheckCollision (code/nans.cpp:906-966) on device-resident pairs.
        noalloc: through nans_narrowphase_noalloc (no device allocation; at most
        NARROWPHASE_NOALLOC_MAX_PAIRS pairs per call)."""
        n = int(dpairs["n"])
        if out is None:
            out = self.alloc_narrow_out(n)
        a = ShapesSoa(*[_ptr(dpairs[k + "_a"]) for k in ("type", "pos", "quat", "half")])
        b = ShapesSoa(*[_ptr(dpairs[k + "_b"]) for k in ("type", "pos", "quat", "half")])
        o = NarrowOut(*[_ptr(out[k]) for k in ("gjk", "contact", "normal", "depth", "point_a", "point_b", "status")])
        nov = ctypes.c_int64(0)
        self._sync_stream()
        fn = self.lib.nans_narrowphase_noalloc if noalloc else self.lib.nans_narrowphase
        check(fn(self.h, ctypes.c_int64(n), ctypes.byref(a), ctypes.byref(b), ctypes.byref(o), ctypes.byref(nov)))
        out["n_overflow"] = nov.value
        return out

    @staticmethod
    def narrow_to_numpy(out: dict) -> dict:
        """Device outputs -> host arrays in the [n][3] layout of the oracle."""
        res = {}
        for k, v in out.items():
            if k == "n_overflow":
                res[k] = v
                continue
            a = v.cpu().numpy()
            res[k] = np.ascontiguousarray(a.T) if a.ndim == 2 else a
        return res

    # ------------------------------------------------------------------ worlds
    def upload(self, scene: "_scenes.Scene") -> None:
        """Copy a batch of worlds (scenes.Scene) to the device."""
        planes = lambda a: np.ascontiguousarray(np.asarray(a, np.float32).T)
        self._host = dict(pos=planes(scene.pos), vel=planes(scene.vel), angvel=planes(scene.angvel),
                          angles=planes(scene.angles), force=planes(scene.force), torque=planes(scene.torque),
                          scale=planes(scene.scale), mass=np.ascontiguousarray(scene.mass, np.float32),
                          moi=np.ascontiguousarray(scene.moi, np.float32),
                          kind=np.ascontiguousarray(scene.kind, np.int32))
        self._offsets = np.ascontiguousarray(scene.offsets, np.int32)
        fr = None if scene.world_friction is None else np.ascontiguousarray(scene.world_friction, np.float32)
        rs = None if scene.world_restitution is None else np.ascontiguousarray(scene.world_restitution, np.float32)
        self._coeffs = (fr, rs)
        b = BodiesSoa(*[_np(self._host[k]) for k in ("pos", "vel", "angvel", "angles", "force", "torque", "scale",
                                                     "mass", "moi", "kind")])
        self.torch.cuda.synchronize(self.device)
        check(self.lib.nans_upload_worlds(self.h, int(scene.nworlds), _np(self._offsets), ctypes.byref(b),
                                          _np(fr), _np(rs)))
        self.nworlds = int(scene.nworlds)
        self.nbodies = int(scene.nbodies)

    def download(self) -> dict:
        """pos, vel, angvel, angles, force, torque as [N][3] float32 arrays."""
        n = self.nbodies
        out = {k: np.zeros((3, n), np.float32) for k in ("pos", "vel", "angvel", "angles", "force", "torque")}
        b = BodiesSoa(*[_np(out.get(k)) for k in ("pos", "vel", "angvel", "angles", "force", "torque")],
                      None, None, None, None)
        check(self.lib.nans_download_state(self.h, ctypes.byref(b)))
        return {k: np.ascontiguousarray(v.T) for k, v in out.items()}

    def step_plain(self, nsteps: int = 1) -> None:
        self._sync_stream()
        check(self.lib.nans_step_plain(self.h, int(nsteps)))

    def step_batched(self, nsteps: int = 1) -> None:
        self._sync_stream()
        check(self.lib.nans_step_batched(self.h, int(nsteps)))

    def record_contacts(self, enable: bool = True) -> None:
        check(self.lib.nans_record_contacts(self.h, int(bool(enable))))

    def contacts(self, world: int = 0, max_contacts: int = 1 << 16) -> np.ndarray:
        buf = np.zeros(max_contacts, CONTACT_DTYPE)
        n = ctypes.c_int32(0)
        check(self.lib.nans_get_contacts(self.h, int(world), _np(buf), int(max_contacts), ctypes.byref(n)))
        return buf[: min(n.value, max_contacts)].copy()

    def stats(self) -> dict:
        st = Stats()
        check(self.lib.nans_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def reset_stats(self) -> None:
        check(self.lib.nans_reset_stats(self.h))

    def set_state(self, scene: "_scenes.Scene") -> None:
        """Overwrite the uploaded batch's state in place (same shape)."""
        planes = lambda a: np.ascontiguousarray(np.asarray(a, np.float32).T)
        st = {k: planes(getattr(scene, k)) for k in ("pos", "vel", "angvel", "angles", "force", "torque")}
        self._state_keep = st
        b = BodiesSoa(*[_np(st[k]) for k in ("pos", "vel", "angvel", "angles", "force", "torque")],
                      None, None, None, None)
        check(self.lib.nans_set_state(self.h, ctypes.byref(b)))

    def autotune(self, nsteps: int = 5) -> dict:
        """Time the batched step at 64/128/256/512 threads per world and wave-world
        bounds 0/16/32/64/128 bodies on the uploaded batch and keep the fastest,
        unless it beats the default by less than 2 % (nans_autotune_batched); the
        state is restored.  ms -1: skipped."""
        self._sync_stream()
        t = Tuning()
        check(self.lib.nans_autotune_batched(self.h, int(nsteps), ctypes.byref(t)))
        return {"threads_per_world": t.threads_per_world, "wave_world_bodies": t.wave_world_bodies,
                "steps": t.steps,
                "ms_per_step": dict(zip(TUNE_SIZES, [round(x, 4) for x in t.ms_per_step])),
                "ms_per_step_wave": {sz: dict(zip(TUNE_WAVE_BOUNDS, [round(x, 4) for x in t.ms_per_step_wave[i]]))
                                     for i, sz in enumerate(TUNE_SIZES)}}

    def set_batched_block(self, threads_per_world: int) -> None:
        check(self.lib.nans_set_batched_block(self.h, int(threads_per_world)))

    def set_wave_worlds(self, max_bodies: int) -> None:
        """Worlds of at most max_bodies bodies run one per wave (-1: the default, none; 0: none)."""
        check(self.lib.nans_set_wave_worlds(self.h, int(max_bodies)))

    def set_warm_start(self, enable: bool) -> None:
        """Opt-in warm start of the plain and batched steps (contact accumulators carried over)."""
        check(self.lib.nans_set_warm_start(self.h, 1 if enable else 0))

    ORIENTATION = {"euler": 0, "quaternion": 1}

    def set_orientation_mode(self, mode: str) -> None:
        """"euler" (the reference's degrees) or the opt-in "quaternion" mode."""
        check(self.lib.nans_set_orientation_mode(self.h, self.ORIENTATION[mode]))

    def quaternions(self) -> np.ndarray:
        """Orientation of the quaternion mode as [N][4] (w, x, y, z)."""
        q = np.zeros((4, self.nbodies), np.float32)
        check(self.lib.nans_get_quaternions(self.h, _np(q)))
        return np.ascontiguousarray(q.T)

    def set_quaternions(self, quat) -> None:
        q = np.ascontiguousarray(np.asarray(quat, np.float32).reshape(self.nbodies, 4).T)
        check(self.lib.nans_set_quaternions(self.h, _np(q)))

    BROADPHASE = {"auto": 0, "all_pairs": 1, "grid": 2}

    def set_broadphase(self, mode: str) -> None:
        """Broadphase of the plain step: "auto", "all_pairs" or "grid" (same
        candidates and results either way)."""
        check(self.lib.nans_set_broadphase(self.h, self.BROADPHASE[mode]))

    def batched_timing(self) -> dict:
        t = Timing()
        check(self.lib.nans_batched_timing(self.h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    # ------------------------------------------------------------------ differentiable rollout
    def rollout(self, nsteps: int, checkpoint_every: int = 16) -> None:
        """nsteps batched steps from the current state, keeping the state every
        checkpoint_every steps for adjoint()."""
        self._sync_stream()
        check(self.lib.nans_rollout(self.h, int(nsteps), int(checkpoint_every)))

    def adjoint(self, dl_pos=None, dl_vel=None, dl_angvel=None, dl_angles=None) -> dict:
        """Gradient of a loss of the final state through the last rollout(), given
        dL/d(final pos, vel, angvel, angles) as [N][3] arrays (None = zero).
        Returns pos0, vel0, angvel0, angles0 ([N][3]), mass, moi ([N]) and
        friction, restitution ([nworlds]); contact sets and frames are held at the
        rollout's (nans_adjoint in include/nans_gpu.h)."""
        n, nw = self.nbodies, self.nworlds
        planes = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, np.float32).reshape(n, 3).T)
        ins = [planes(a) for a in (dl_pos, dl_vel, dl_angvel, dl_angles)]
        outs = {k: np.zeros((3, n), np.float32) for k in ("pos0", "vel0", "angvel0", "angles0")}
        outs.update(mass=np.zeros(n, np.float32), moi=np.zeros(n, np.float32), friction=np.zeros(nw, np.float32),
                    restitution=np.zeros(nw, np.float32))
        io = AdjointIo(*[_np(a) for a in ins], *[_np(outs[k]) for k in ("pos0", "vel0", "angvel0", "angles0", "mass",
                                                                         "moi", "friction", "restitution")])
        self._sync_stream()
        check(self.lib.nans_adjoint(self.h, ctypes.byref(io)))
        return {k: (np.ascontiguousarray(v.T) if v.ndim == 2 else v) for k, v in outs.items()}
