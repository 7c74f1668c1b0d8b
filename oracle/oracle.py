"""ctypes front end of the CPU oracle (oracle/nans_oracle_impl.h).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker of the HIP path.  The product
package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnans_oracle.so")
LIB_O0_PATH = os.path.join(HERE, "lib", "libnans_oracle_O0.so")
# the binary32 oracle with the Model rebuild's sin/cos from the host libm's sinf/cosf
LIB_LIBM_PATH = os.path.join(HERE, "lib", "libnans_oracle_libm.so")

NSTAT = 8
STAT_GJK_CALLS, STAT_GJK_RESULT, STAT_EPA_ITERS, STAT_EPA_RESULT = 0, 1, 2, 3
STAT_MAX_FACES, STAT_MAX_EDGES, STAT_NVERTS, STAT_FACES_MADE = 4, 5, 6, 7
GJK_NO, GJK_FOUND, GJK_CAP = 0, 1, 2
EPA_NONE, EPA_CONTACT, EPA_BARY, EPA_NONFINITE, EPA_ITER, EPA_EMPTY = 0, 1, 2, 3, 4, 5
CC, CS, CF, SS, SF = 0, 1, 2, 3, 4


class NpParams(ctypes.Structure):
    _fields_ = [("max_gjk_iterations", ctypes.c_int32), ("max_epa_iterations", ctypes.c_int32),
                ("epa_tolerance", ctypes.c_float), ("fp64_islands", ctypes.c_int32)]


class StepParams(ctypes.Structure):
    _fields_ = [("np", NpParams), ("solver_iterations", ctypes.c_int32), ("beta", ctypes.c_float),
                ("restitution", ctypes.c_float), ("friction", ctypes.c_float), ("damping", ctypes.c_float),
                ("gravity", ctypes.c_float), ("dt", ctypes.c_float), ("floor_pos", ctypes.c_float * 3),
                ("floor_scale", ctypes.c_float * 3), ("floor_mass", ctypes.c_float), ("floor_moi", ctypes.c_float)]


class Contact32(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("index_a", ctypes.c_int32), ("index_b", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("point_a", ctypes.c_float * 3), ("point_b", ctypes.c_float * 3),
                ("normal", ctypes.c_float * 3), ("dl_normal", ctypes.c_float), ("dl_tangent1", ctypes.c_float),
                ("dl_tangent2", ctypes.c_float)]


class Contact64(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("index_a", ctypes.c_int32), ("index_b", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("point_a", ctypes.c_double * 3), ("point_b", ctypes.c_double * 3),
                ("normal", ctypes.c_double * 3), ("dl_normal", ctypes.c_double), ("dl_tangent1", ctypes.c_double),
                ("dl_tangent2", ctypes.c_double)]


CONTACT_DTYPE = {
    "f32": np.dtype([("type", "<i4"), ("index_a", "<i4"), ("index_b", "<i4"), ("pad", "<i4"),
                     ("point_a", "<f4", 3), ("point_b", "<f4", 3), ("normal", "<f4", 3),
                     ("dl_normal", "<f4"), ("dl_tangent1", "<f4"), ("dl_tangent2", "<f4")]),
    "f64": np.dtype([("type", "<i4"), ("index_a", "<i4"), ("index_b", "<i4"), ("pad", "<i4"),
                     ("point_a", "<f8", 3), ("point_b", "<f8", 3), ("normal", "<f8", 3),
                     ("dl_normal", "<f8"), ("dl_tangent1", "<f8"), ("dl_tangent2", "<f8")]),
}

_libs: dict = {}


def build() -> None:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(opt: str = "O2"):
    path = {"O2": LIB_PATH, "O0": LIB_O0_PATH, "libm": LIB_LIBM_PATH}[opt]
    if path not in _libs:
        if not os.path.exists(path):
            build()
        _libs[path] = ctypes.CDLL(path)
    return _libs[path]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def np_params(params) -> NpParams:
    return NpParams(int(params.max_gjk_iterations), int(params.max_epa_iterations),
                    float(params.epa_tolerance), int(params.fp64_islands))


def step_params(params, friction=None, restitution=None) -> StepParams:
    p = StepParams()
    p.np = np_params(params)
    p.solver_iterations = int(params.solver_iterations)
    p.beta = params.beta
    p.restitution = params.restitution if restitution is None else float(restitution)
    p.friction = params.friction if friction is None else float(friction)
    p.damping = params.damping
    p.gravity = params.gravity
    p.dt = params.dt
    p.floor_pos = (ctypes.c_float * 3)(*params.floor_pos)
    p.floor_scale = (ctypes.c_float * 3)(*params.floor_scale)
    p.floor_mass = params.floor_mass
    p.floor_moi = params.floor_moi
    return p


def narrowphase(pairs: dict, params, precision: str = "f32", stats: bool = False, opt: str = "O2"):
    """CheckCollision on explicit pairs (code/nans.cpp:906-966)."""
    ft = np.float32 if precision == "f32" else np.float64
    n = len(pairs["type_a"])
    c = lambda k: np.ascontiguousarray(pairs[k], dtype=ft)
    i32 = lambda k: np.ascontiguousarray(pairs[k], dtype=np.int32)
    out = dict(gjk=np.zeros(n, np.int32), contact=np.zeros(n, np.int32), normal=np.zeros((n, 3), ft),
               depth=np.zeros(n, np.float64), point_a=np.zeros((n, 3), ft), point_b=np.zeros((n, 3), ft))
    st = np.zeros((n, NSTAT), np.int32) if stats else None
    args = [i32("type_a"), c("pos_a"), c("quat_a"), c("half_a"), i32("type_b"), c("pos_b"), c("quat_b"), c("half_b")]
    fn = getattr(lib(opt), "ora_narrowphase_" + precision)
    prm = np_params(params)
    rc = fn(ctypes.c_int64(n), *[_ptr(a) for a in args], ctypes.byref(prm),
            _ptr(out["gjk"]), _ptr(out["contact"]), _ptr(out["normal"]), _ptr(out["depth"]),
            _ptr(out["point_a"]), _ptr(out["point_b"]), _ptr(st))
    if rc != 0:
        raise ValueError("oracle narrowphase rejected its arguments")
    if stats:
        out["stats"] = st
    return out


TRACE_MAXF = 128
GJK_REC = np.dtype([("n_in", "<i4"), ("n_out", "<i4"), ("result", "<i4"), ("pushed", "<i4"),
                    ("s_in", "<f8", (4, 3)), ("s_out", "<f8", (5, 3)), ("dir", "<f8", 3), ("sup_a", "<f8", 3),
                    ("sup_b", "<f8", 3)])
EPA_REC = np.dtype([("nfaces", "<i4"), ("best", "<i4"), ("converged", "<i4"), ("nfaces_out", "<i4"),
                    ("cur", "<f8"), ("dir", "<f8", 3), ("p_new", "<f8", 3), ("sa_new", "<f8", 3),
                    ("sb_new", "<f8", 3), ("bary", "<f8", 3), ("face_sup", "<f8", (3, 6)),
                    ("face", "<f8", (TRACE_MAXF, 12)), ("visible", "<i4", TRACE_MAXF), ("dist0", "<f8"),
                    ("dist", "<f8", TRACE_MAXF)])


def trace_pair(shape_a: tuple, shape_b: tuple, params, precision: str = "f32", gcap: int = 80, ecap: int = 80):
    """CheckCollision of one pair with its call-by-call trace (ora_trace_pair_*).
    A shape is (type, centre[3], quaternion w,x,y,z [4], half-extents or radius [3]).
    Returns None when the trace did not fit, else a dict with the GJK records
    (GJK_REC), the EPA records (EPA_REC) and the outcome."""
    ft = np.float32 if precision == "f32" else np.float64
    arr = lambda x: np.ascontiguousarray(x, dtype=ft)
    a = [arr(shape_a[k]) for k in (1, 2, 3)]
    b = [arr(shape_b[k]) for k in (1, 2, 3)]
    g = np.zeros(gcap, GJK_REC)
    e = np.zeros(ecap, EPA_REC)
    ng, ne, gj, ep = (ctypes.c_int32() for _ in range(4))
    nrm, pa, pb = np.zeros(3, ft), np.zeros(3, ft), np.zeros(3, ft)
    prm = np_params(params)
    rc = getattr(lib(), "ora_trace_pair_" + precision)(
        ctypes.c_int32(int(shape_a[0])), _ptr(a[0]), _ptr(a[1]), _ptr(a[2]),
        ctypes.c_int32(int(shape_b[0])), _ptr(b[0]), _ptr(b[1]), _ptr(b[2]), ctypes.byref(prm),
        _ptr(g), ctypes.c_int32(gcap), ctypes.byref(ng), _ptr(e), ctypes.c_int32(ecap), ctypes.byref(ne),
        ctypes.byref(gj), ctypes.byref(ep), _ptr(nrm), _ptr(pa), _ptr(pb))
    if rc < 0:
        raise ValueError("oracle trace rejected its arguments")
    if rc == 1:
        return None
    return dict(gjk_recs=g[:ng.value], epa_recs=e[:ne.value], gjk=gj.value, epa=ep.value, normal=nrm,
                point_a=pa, point_b=pb)


class WorldState:
    """One world's mutable state for the oracle (interleaved [n][3] arrays)."""

    def __init__(self, scene, w: int = 0, precision: str = "f32"):
        ft = np.float32 if precision == "f32" else np.float64
        a, b = int(scene.offsets[w]), int(scene.offsets[w + 1])
        self.precision = precision
        self.nb = int(scene.nboxes[w])
        self.ns = (b - a) - self.nb
        cp = lambda x: np.array(x[a:b], dtype=ft, order="C")  # a copy: run() updates it in place
        self.pos, self.vel, self.angvel, self.angles = cp(scene.pos), cp(scene.vel), cp(scene.angvel), cp(scene.angles)
        self.force, self.torque = cp(scene.force), cp(scene.torque)
        self.scale, self.mass, self.moi = cp(scene.scale), cp(scene.mass), cp(scene.moi)
        self.params = scene.params
        self.quat = None  # [n][4] (w, x, y, z) in the opt-in quaternion mode
        self.warm = None  # the opt-in warm start's contact list (enable_warm)
        self.friction = None if scene.world_friction is None else float(scene.world_friction[w])
        self.restitution = None if scene.world_restitution is None else float(scene.world_restitution[w])
        self.stats = np.zeros(8, np.int64)

    def enable_warm(self, cap: int):
        """Opt-in warm start: carry each contact's accumulators to the next step
        (at most cap contacts persist); the list starts empty."""
        ft = np.float32 if self.precision == "f32" else np.float64
        self.warm = (np.zeros(cap, np.int64), np.zeros((cap, 3), ft), np.zeros(1, np.int32), int(cap))

    def run(self, nsteps: int, contacts: bool = False, max_contacts: int = 1 << 16, opt: str = "O2"):
        """Advance nsteps; returns the contact list of the last step if asked."""
        prm = step_params(self.params, self.friction, self.restitution)
        cbuf = np.zeros(max_contacts, CONTACT_DTYPE[self.precision]) if contacts else None
        ncon = ctypes.c_int32(0)
        state = [_ptr(self.pos), _ptr(self.vel), _ptr(self.angvel), _ptr(self.angles)]
        extra = []
        if self.warm is not None:
            fn = getattr(lib(opt), "ora_world_run_warm_" + self.precision)
            state.append(_ptr(self.quat))
            wp, wsum, wn, cap = self.warm
            extra = [_ptr(wp), _ptr(wsum), _ptr(wn), cap]
        elif self.quat is None:
            fn = getattr(lib(opt), "ora_world_run_" + self.precision)
        else:
            fn = getattr(lib(opt), "ora_world_run_quat_" + self.precision)
            state.append(_ptr(self.quat))
        rc = fn(self.nb, self.ns, *state,
                _ptr(self.force), _ptr(self.torque), _ptr(self.scale), _ptr(self.mass), _ptr(self.moi),
                ctypes.byref(prm), int(nsteps), _ptr(cbuf), int(max_contacts), ctypes.byref(ncon),
                _ptr(self.stats), *extra)
        if rc != 0:
            raise ValueError("oracle world_run rejected its arguments")
        if contacts:
            return cbuf[: min(ncon.value, max_contacts)].copy()
        return None

    def pairs(self):
        """Every canonical pair of the current pose with its CheckCollision stats."""
        nb, ns = self.nb, self.ns
        npairs = nb * (nb - 1) // 2 + nb + ns + nb * ns + ns * (ns - 1) // 2
        t, ia, ib = (np.zeros(npairs, np.int32) for _ in range(3))
        st = np.zeros((npairs, NSTAT), np.int32)
        prm = step_params(self.params, self.friction, self.restitution)
        f32 = lambda x: np.ascontiguousarray(x, dtype=np.float32)
        rc = lib().ora_world_pairs_f32(nb, ns, _ptr(f32(self.pos)), _ptr(f32(self.angles)), _ptr(f32(self.scale)),
                                       ctypes.byref(prm), _ptr(t), _ptr(ia), _ptr(ib), _ptr(st))
        if rc != 0:
            raise ValueError("oracle world_pairs rejected its arguments")
        return t, ia, ib, st


def box_vertices_quat(pos, quat, scale, precision: str = "f32"):
    """Corners of the quaternion-mode Model [R(q) diag(scale) | pos], [8][3]."""
    ft = np.float32 if precision == "f32" else np.float64
    out = np.zeros((8, 3), ft)
    a = [np.ascontiguousarray(x, dtype=ft) for x in (pos, quat, scale)]
    getattr(lib(), "ora_box_vertices_quat_" + precision)(*[_ptr(x) for x in a], _ptr(out))
    return out


def box_vertices(pos, angles, scale, precision: str = "f32"):
    ft = np.float32 if precision == "f32" else np.float64
    out = np.zeros((8, 3), ft)
    a = [np.ascontiguousarray(x, dtype=ft) for x in (pos, angles, scale)]
    getattr(lib(), "ora_box_vertices_" + precision)(*[_ptr(x) for x in a], _ptr(out))
    return out


def rk4(v, w, force, torque, mass, moi, params, precision: str = "f32"):
    ft = np.float32 if precision == "f32" else np.float64
    cty = ctypes.c_float if precision == "f32" else ctypes.c_double
    v = np.array(v, ft)
    w = np.array(w, ft)
    f = np.ascontiguousarray(force, dtype=ft)
    t = np.ascontiguousarray(torque, dtype=ft)
    prm = step_params(params)
    getattr(lib(), "ora_rk4_" + precision)(_ptr(v), _ptr(w), _ptr(f), _ptr(t), cty(mass), cty(moi), ctypes.byref(prm))
    return v, w


MANIFOLD_DTYPE = np.dtype([("type", "<i4"), ("slot_a", "<i4"), ("slot_b", "<i4"), ("pad", "<i4"),
                           ("r1", "<f4", 3), ("r2", "<f4", 3), ("n", "<f4", 3)])


def frozen_record(scene, w: int, nsteps: int):
    """Binary32 rollout of world w recording every step's contact frames
    (ora_frozen_record_f32).  Returns (frames, step_count, final WorldState)."""
    ws = WorldState(scene, w, "f32")
    prm = step_params(ws.params, ws.friction, ws.restitution)
    counts_ = np.zeros(max(nsteps, 1), np.int32)
    total = ctypes.c_int64(0)
    cap = 64 * (ws.nb + ws.ns + 4) * max(nsteps, 1)
    for _ in range(2):
        frames = np.zeros(cap, MANIFOLD_DTYPE)
        trial = WorldState(scene, w, "f32")
        rc = lib().ora_frozen_record_f32(trial.nb, trial.ns, _ptr(trial.pos), _ptr(trial.vel), _ptr(trial.angvel),
                                         _ptr(trial.angles), _ptr(trial.force), _ptr(trial.torque),
                                         _ptr(trial.scale), _ptr(trial.mass), _ptr(trial.moi), ctypes.byref(prm),
                                         int(nsteps), _ptr(frames), ctypes.c_int64(cap), _ptr(counts_),
                                         ctypes.byref(total))
        if rc == 0:
            return frames[: total.value].copy(), counts_[:nsteps].copy(), trial
        if rc != -2:
            raise ValueError("oracle frozen_record rejected its arguments")
        cap = int(total.value)
    raise RuntimeError("frame capacity")


def frozen_run(scene, w: int, frames, step_count, nsteps: int, precision: str = "f64", pos=None, vel=None,
               angvel=None, angles=None, mass=None, moi=None, friction=None, restitution=None) -> dict:
    """Replay world w for nsteps with recorded contact frames (ora_frozen_run_*),
    from the scene's initial state or the given overrides; returns the final state."""
    ft = np.float32 if precision == "f32" else np.float64
    ws = WorldState(scene, w, precision)
    pick = lambda x, d: np.array(d if x is None else x, dtype=ft).reshape(np.shape(d)).copy()
    st = dict(pos=pick(pos, ws.pos), vel=pick(vel, ws.vel), angvel=pick(angvel, ws.angvel),
              angles=pick(angles, ws.angles), force=ws.force.copy(), torque=ws.torque.copy())
    m = pick(mass, ws.mass)
    I = pick(moi, ws.moi)
    prm = step_params(ws.params)
    cf = float(ws.params.friction if ws.friction is None else ws.friction) if friction is None else float(friction)
    cr = (float(ws.params.restitution if ws.restitution is None else ws.restitution) if restitution is None
          else float(restitution))
    fr = np.ascontiguousarray(frames, dtype=MANIFOLD_DTYPE)
    sc = np.ascontiguousarray(step_count, dtype=np.int32)
    fn = getattr(lib(), "ora_frozen_run_" + precision)
    rc = fn(ws.nb, ws.ns, _ptr(st["pos"]), _ptr(st["vel"]), _ptr(st["angvel"]), _ptr(st["angles"]),
            _ptr(st["force"]), _ptr(st["torque"]), _ptr(m), _ptr(I), ctypes.byref(prm), ctypes.c_double(cf),
            ctypes.c_double(cr), int(nsteps), _ptr(fr), _ptr(sc))
    if rc != 0:
        raise ValueError("oracle frozen_run rejected its arguments")
    return st


COUNT_NAMES = ("body_box", "body_sphere", "pairs", "candidates", "evolve0", "evolve1", "evolve2", "evolve3",
               "evolve4", "evolve_length", "support_box", "support_sphere", "epa_iter", "epa_faces", "epa_edges",
               "epa_final", "constraints", "solver_iterations")


def counting(enable: bool, cull: bool = False, precision: str = "f32") -> None:
    """Switch the oracle's work counters on (reset) or off; cull applies the HIP
    step's bounding-sphere cull so the counts describe the work it performs."""
    getattr(lib(), "ora_counting_" + precision)(int(enable), int(cull))


def counts(precision: str = "f32") -> dict:
    out = np.zeros(32, np.int64)
    getattr(lib(), "ora_get_counts_" + precision)(_ptr(out))
    return {k: int(out[i]) for i, k in enumerate(COUNT_NAMES)}
