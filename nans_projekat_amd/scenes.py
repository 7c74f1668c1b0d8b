"""Seeded synthetic scenes for the simulation step.

Every scene is generated on the host, once, so that the HIP path and the CPU
checker see identical bytes.  The random source is SplitMix64 (Steele, Lea &
Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014) with one
sub-stream per (config, world, purpose), seeded from 0x4E414E53 ("NANS").

World layout (shared with the C-ABI, see include/nans_gpu.h): bodies of all
worlds are concatenated; world w owns bodies [offsets[w], offsets[w+1]), its
boxes first and its spheres after them -- the reference keeps boxes and spheres
in the separate index spaces State->Cubes[] and State->Spheres[]
(code/nans.h:377-378), and pair indices below are relative to those spaces.

Body fields (float32):
  pos, vel, angvel, angles (Euler angles in degrees, code/nans.cpp:1915-1920),
  force, torque, scale (boxes: Model scale = full extents, code/nans.cpp:1921-1923
  and 1938-1940; spheres: scale[0] = radius), mass, moi.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SEED = 0x4E414E53

BOX = 0
SPHERE = 1

_M64 = (1 << 64) - 1


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 output function applied to counters x (uint64 array)."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based SplitMix64 stream: value k is splitmix64(key + k*golden)."""

    def __init__(self, *path: int, seed: int = SEED):
        key = seed & _M64
        for p in path:
            key = int(_splitmix64(np.array([(key ^ (p & _M64)) & _M64], dtype=np.uint64))[0])
        self.key = key
        self.pos = 0

    def u64(self, n: int) -> np.ndarray:
        idx = np.arange(self.pos, self.pos + n, dtype=np.uint64)
        self.pos += n
        with np.errstate(over="ignore"):
            ctr = np.uint64(self.key) + idx * np.uint64(0x9E3779B97F4A7C15)
        return _splitmix64(ctr)

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        """Doubles in [lo, hi) from the top 53 bits."""
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """Integers in [lo, hi] inclusive."""
        return lo + np.floor(self.uniform(n) * (hi - lo + 1)).astype(np.int64)


def unit_quaternions(st: Stream, n: int) -> np.ndarray:
    """Uniform unit quaternions (w, x, y, z), Shoemake's subgroup algorithm."""
    u1, u2, u3 = st.uniform(n), st.uniform(n), st.uniform(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([b * np.cos(2 * math.pi * u3), a * np.sin(2 * math.pi * u2),
                  a * np.cos(2 * math.pi * u2), b * np.sin(2 * math.pi * u3)], axis=1)
    return q.astype(np.float32)


def cube_moi(mass, size):
    """(m/12)*(2*s*s) in binary32, the cube formula of code/nans.cpp:1622-1623."""
    m = np.asarray(mass, dtype=np.float32)
    s = np.asarray(size, dtype=np.float32)
    return (m / np.float32(12.0)) * (np.float32(2.0) * s * s)


def box_moi(mass, scale):
    """Scalar inertia of a general box: (m/12)*(2/3)(a^2+b^2+c^2) with full
    extents a, b, c; reduces to cube_moi for a = b = c (the reference keeps a
    scalar isotropic MOI, code/nans.cpp:1609-1610)."""
    m = np.asarray(mass, dtype=np.float64)
    s = np.asarray(scale, dtype=np.float64)
    return ((m / 12.0) * (2.0 / 3.0) * np.sum(s * s, axis=-1)).astype(np.float32)


def sphere_moi(mass, radius):
    """(2/5) m r^2 (code/nans.cpp:1663-1664)."""
    m = np.asarray(mass, dtype=np.float32)
    r = np.asarray(radius, dtype=np.float64)
    return (np.float32(2.0 / 5.0) * m).astype(np.float64) * (r * r)


@dataclasses.dataclass
class StepParams:
    """Host mirror of nans_params (include/nans_gpu.h).  Defaults are the
    reference's hard-coded constants (code/nans.cpp:68,1152-1176, code/nans.h:46,54-56)."""
    solver_iterations: int = 70
    beta: float = 0.3
    restitution: float = 0.1
    friction: float = 0.1
    damping: float = 1.5
    gravity: float = 9.81
    dt: float = 1.0 / 60.0
    max_gjk_iterations: int = 64
    max_epa_iterations: int = 64
    epa_tolerance: float = 0.001
    fp64_islands: int = 0
    floor_pos: tuple = (1.2, -0.5, 1.0)
    floor_scale: tuple = (100.0, 1.0, 100.0)
    floor_mass: float = 100000.0
    floor_moi: float = float(cube_moi(100000.0, 100.0))


@dataclasses.dataclass
class Scene:
    """A batch of independent worlds (see module docstring for the layout)."""
    pos: np.ndarray
    vel: np.ndarray
    angvel: np.ndarray
    angles: np.ndarray
    force: np.ndarray
    torque: np.ndarray
    scale: np.ndarray
    mass: np.ndarray
    moi: np.ndarray
    kind: np.ndarray
    offsets: np.ndarray          # int32 [nworlds + 1]
    nboxes: np.ndarray           # int32 [nworlds]
    params: StepParams
    nsteps: int
    world_friction: np.ndarray | None = None     # float32 [nworlds] (config 4)
    world_restitution: np.ndarray | None = None

    @property
    def nworlds(self) -> int:
        return len(self.offsets) - 1

    @property
    def nbodies(self) -> int:
        return int(self.offsets[-1])

    def world(self, w: int) -> "Scene":
        a, b = int(self.offsets[w]), int(self.offsets[w + 1])
        f = None if self.world_friction is None else self.world_friction[w:w + 1].copy()
        r = None if self.world_restitution is None else self.world_restitution[w:w + 1].copy()
        return Scene(self.pos[a:b].copy(), self.vel[a:b].copy(), self.angvel[a:b].copy(),
                     self.angles[a:b].copy(), self.force[a:b].copy(), self.torque[a:b].copy(),
                     self.scale[a:b].copy(), self.mass[a:b].copy(), self.moi[a:b].copy(),
                     self.kind[a:b].copy(), np.array([0, b - a], dtype=np.int32),
                     self.nboxes[w:w + 1].copy(), dataclasses.replace(self.params), self.nsteps, f, r)

    def subset(self, worlds) -> "Scene":
        parts = [self.world(int(w)) for w in worlds]
        return concat(parts)

    def copy(self) -> "Scene":
        return dataclasses.replace(self, **{f.name: (getattr(self, f.name).copy() if isinstance(getattr(self, f.name), np.ndarray) else getattr(self, f.name))
                                            for f in dataclasses.fields(self)})


def concat(parts) -> Scene:
    offs = [0]
    for s in parts:
        offs.append(offs[-1] + s.nbodies)
    cat = lambda name: np.concatenate([getattr(s, name) for s in parts])
    fr = None if parts[0].world_friction is None else cat("world_friction")
    rs = None if parts[0].world_restitution is None else cat("world_restitution")
    return Scene(cat("pos"), cat("vel"), cat("angvel"), cat("angles"), cat("force"), cat("torque"),
                 cat("scale"), cat("mass"), cat("moi"), cat("kind"), np.array(offs, dtype=np.int32),
                 cat("nboxes"), parts[0].params, parts[0].nsteps, fr, rs)


def _world(box_pos, box_scale, box_mass, sph_pos, sph_vel, sph_r, sph_mass, box_vel=None):
    nb, ns = len(box_pos), len(sph_pos)
    n = nb + ns
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    pos = f32(np.concatenate([np.reshape(box_pos, (nb, 3)), np.reshape(sph_pos, (ns, 3))]))
    vel = np.zeros((n, 3), np.float32)
    if box_vel is not None:
        vel[:nb] = box_vel
    vel[nb:] = np.reshape(sph_vel, (ns, 3))
    scale = np.zeros((n, 3), np.float32)
    scale[:nb] = box_scale
    scale[nb:] = np.repeat(np.reshape(np.asarray(sph_r, np.float32), (ns, 1)), 3, axis=1)
    mass = f32(np.concatenate([np.reshape(box_mass, (nb,)), np.reshape(sph_mass, (ns,))]))
    moi = np.zeros(n, np.float32)
    cubes = np.all(box_scale == box_scale[:, :1], axis=1) if nb else np.zeros(0, bool)
    moi[:nb] = np.where(cubes, cube_moi(mass[:nb], box_scale[:, 0] if nb else 0), box_moi(mass[:nb], box_scale))
    moi[nb:] = sphere_moi(mass[nb:], scale[nb:, 0]).astype(np.float32)
    kind = np.array([BOX] * nb + [SPHERE] * ns, dtype=np.int32)
    z = lambda: np.zeros((n, 3), np.float32)
    return Scene(pos, vel, z(), z(), z(), z(), scale, mass, moi, kind,
                 np.array([0, n], np.int32), np.array([nb], np.int32), StepParams(), 0)


def config1(seed: int = SEED, nsteps: int = 600) -> Scene:
    """Plain reference scene: a 3x3x3 stack of unit cubes (half-extent 0.5, mass 1)
    at (i, 0.5+j, k) with U(+-0.01) jitter per axis, on the ground box; a sphere
    (r=0.5, m=2) at (-6, 1.5, 0) fired at (20, 0, 0).  dt=1/120, 10 iterations,
    restitution 0.2, friction 0.5."""
    st = Stream(1, 0, seed=seed)
    grid = np.array([(i, 0.5 + j, k) for i in (-1, 0, 1) for j in (0, 1, 2) for k in (-1, 0, 1)], np.float64)
    jit = st.uniform(27 * 3, -0.01, 0.01).reshape(27, 3)
    box_pos = (grid + jit).astype(np.float32)
    sc = _world(box_pos, np.ones((27, 3), np.float32), np.ones(27, np.float32),
                [(-6.0, 1.5, 0.0)], [(20.0, 0.0, 0.0)], [0.5], [2.0])
    sc.params = dataclasses.replace(StepParams(), solver_iterations=10, restitution=0.2, friction=0.5,
                                    dt=1.0 / 120.0)
    sc.nsteps = nsteps
    return sc


def demo_scene() -> Scene:
    """The reference's own Init scene, physics part (code/nans.cpp:1597-1678):
    cubes at (2,3.5,2), (2,1,2), (2,4.5,2) and the non-uniform box Cubes[3] at
    (1,1,1) with Model scale (0.5,1,0.5) (code/nans.cpp:1938-1940) keeping the
    unit-cube MOI; sphere r=0.25, m=2 at (0.1,1.1,1.1)."""
    box_pos = np.array([(2, 3.5, 2), (2, 1, 2), (2, 4.5, 2), (1, 1, 1)], np.float32)
    box_scale = np.array([(1, 1, 1), (1, 1, 1), (1, 1, 1), (0.5, 1, 0.5)], np.float32)
    sc = _world(box_pos, box_scale, np.ones(4, np.float32), [(0.1, 1.1, 1.1)], [(0, 0, 0)], [0.25], [2.0])
    sc.moi[:4] = cube_moi(1.0, 1.0)
    sc.moi[4] = np.float32((np.float32(2.0 / 5.0) * np.float32(2.0)) * math.pow(0.25, 2))
    sc.nsteps = 240
    return sc


def config2_pairs(n: int = 1_000_000, seed: int = SEED):
    """Narrowphase microbench pairs: 50% box-box, 30% box-sphere, 20% sphere-sphere.
    Box half-extents U[0.25,1]^3, radii U[0.25,1], uniform unit quaternions; A at
    the origin, B at distance U[0, 1.2 (R_A + R_B)] along a uniform direction, with
    R = |h| for boxes and r for spheres.  Returns a dict of float32/int32 arrays
    laid out [n][3] / [n][4]."""
    st = Stream(2, 0, seed=seed)
    kind = np.full(n, 0, np.int64)
    u = st.uniform(n)
    kind[u >= 0.5] = 1
    kind[u >= 0.8] = 2
    ta = np.where(kind == 2, SPHERE, BOX).astype(np.int32)
    tb = np.where(kind == 0, BOX, SPHERE).astype(np.int32)
    ha = st.uniform(3 * n, 0.25, 1.0).reshape(n, 3)
    hb = st.uniform(3 * n, 0.25, 1.0).reshape(n, 3)
    ha[ta == SPHERE, 1:] = 0.0
    hb[tb == SPHERE, 1:] = 0.0
    qa = unit_quaternions(st, n)
    qb = unit_quaternions(st, n)
    ra = np.where(ta == BOX, np.linalg.norm(ha, axis=1), ha[:, 0])
    rb = np.where(tb == BOX, np.linalg.norm(hb, axis=1), hb[:, 0])
    dist = st.uniform(n) * 1.2 * (ra + rb)
    d = st.normal(3 * n).reshape(n, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pa = np.zeros((n, 3), np.float32)
    pb = (d * dist[:, None]).astype(np.float32)
    return dict(type_a=ta, pos_a=pa, quat_a=qa, half_a=ha.astype(np.float32),
                type_b=tb, pos_b=pb, quat_b=qb, half_b=hb.astype(np.float32))


def _stack_world(st: Stream, n: int, nspheres: int) -> Scene:
    nc = n - nspheres
    side = max(1, int(round(nc ** (1.0 / 3.0))))
    idx = np.arange(nc)
    layer, rem = idx // (side * side), idx % (side * side)
    gx, gz = rem % side, rem // side
    off = (side - 1) / 2.0
    grid = np.stack([gx - off, 0.5 + layer, gz - off], axis=1).astype(np.float64)
    grid[:, 0] += 1.2
    grid[:, 2] += 1.0
    box_pos = grid + st.uniform(3 * nc, -0.01, 0.01).reshape(nc, 3)
    centre = np.array([1.2, 0.5 + (layer.max() if nc else 0) / 2.0, 1.0])
    az = st.uniform(nspheres, 0.0, 2.0 * math.pi)
    rad = st.uniform(nspheres, 6.0, 10.0) + side / 2.0
    hgt = st.uniform(nspheres, 0.5, 3.0)
    sph_pos = np.stack([centre[0] + rad * np.cos(az), hgt, centre[2] + rad * np.sin(az)], axis=1)
    speed = st.uniform(nspheres, 5.0, 25.0)
    dirn = centre[None, :] - sph_pos
    dirn /= np.linalg.norm(dirn, axis=1, keepdims=True)
    sph_vel = dirn * speed[:, None]
    return _world(box_pos.astype(np.float32), np.ones((nc, 3), np.float32), np.ones(nc, np.float32),
                  sph_pos, sph_vel, np.full(nspheres, 0.5), np.full(nspheres, 2.0))


def config3(nworlds: int = 4096, seed: int = SEED, nsteps: int = 240) -> Scene:
    """Batched worlds: body counts n = floor(exp(U[ln 8, ln 257))) (log-uniform in
    [8, 256]); s = U{1..4} spheres (r=0.5, m=2) aimed at the stack centre with
    speed U[5,25] m/s; n-s unit cubes in a jittered grid stack (side^2 per layer,
    side = round(cbrt(n-s))) on the ground.  dt=1/120, 10 iterations."""
    st = Stream(3, 0, seed=seed)
    sizes = np.floor(np.exp(st.uniform(nworlds, math.log(8.0), math.log(257.0)))).astype(np.int64)
    sizes = np.clip(sizes, 8, 256)
    nsph = st.integers(nworlds, 1, 4)
    parts = [_stack_world(Stream(3, 1 + w, seed=seed), int(sizes[w]), int(nsph[w])) for w in range(nworlds)]
    sc = concat(parts)
    sc.params = dataclasses.replace(StepParams(), solver_iterations=10, restitution=0.2, friction=0.5,
                                    dt=1.0 / 120.0)
    sc.nsteps = nsteps
    return sc


def config4(nworlds: int = 1024, seed: int = SEED, nsteps: int = 240, jitter: bool = True) -> Scene:
    """Differentiable rollout: 1024 worlds of 16 unit cubes in a 4x4 single layer
    (spacing 1.05) resting on the ground, each cube's position jittered by U(+-0.01)
    per axis as in configs 1 and 3 (a sub-stream per world); per-cube mass
    U[0.5,2]; per-world friction U[0.2,0.9] and restitution U[0,0.5]; initial
    velocity ~ N(0,1)^3.  jitter=False keeps the exact layout, whose cubes touch
    the ground with zero depth (the degenerate case the floor tests keep)."""
    st = Stream(4, 0, seed=seed)
    parts = []
    g = np.array([(1.2 + 1.05 * (i - 1.5), 0.5, 1.0 + 1.05 * (k - 1.5)) for i in range(4) for k in range(4)])
    for w in range(nworlds):
        m = st.uniform(16, 0.5, 2.0).astype(np.float32)
        v0 = st.normal(48).reshape(16, 3)
        pos = g + (Stream(4, 1 + w, seed=seed).uniform(48, -0.01, 0.01).reshape(16, 3) if jitter else 0.0)
        parts.append(_world(pos.astype(np.float32), np.ones((16, 3), np.float32), m,
                            np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros(0), box_vel=v0))
    sc = concat(parts)
    sc.world_friction = st.uniform(nworlds, 0.2, 0.9).astype(np.float32)
    sc.world_restitution = st.uniform(nworlds, 0.0, 0.5).astype(np.float32)
    sc.params = dataclasses.replace(StepParams(), solver_iterations=10, restitution=0.2, friction=0.5,
                                    dt=1.0 / 120.0)
    sc.nsteps = nsteps
    return sc


def config5(seed: int = SEED, nsteps: int = 480, nx: int = 32, ny: int = 8, nz: int = 32) -> Scene:
    """Large single world: nx*ny*nz (8192) boxes, half-extents U[0.2,0.6]^3, unit
    density (mass = volume), on a jittered grid at 1.4 m pitch centred over the
    ground box and dropped from y >= 1.  dt=1/240, 20 iterations."""
    st = Stream(5, 0, seed=seed)
    n = nx * ny * nz
    half = st.uniform(3 * n, 0.2, 0.6).reshape(n, 3)
    ii = np.arange(n)
    gx, gy, gz = ii % nx, (ii // nx) % ny, ii // (nx * ny)
    pos = np.stack([1.2 + 1.4 * (gx - (nx - 1) / 2.0), 1.0 + 1.4 * gy, 1.0 + 1.4 * (gz - (nz - 1) / 2.0)], axis=1)
    pos += st.uniform(3 * n, -0.05, 0.05).reshape(n, 3)
    scale = (2.0 * half).astype(np.float32)
    mass = np.prod(scale.astype(np.float64), axis=1).astype(np.float32)
    sc = _world(pos.astype(np.float32), scale, mass, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros(0))
    sc.params = dataclasses.replace(StepParams(), solver_iterations=20, restitution=0.2, friction=0.5,
                                    dt=1.0 / 240.0)
    sc.nsteps = nsteps
    return sc


def pair_counts(nb: int, ns: int):
    """Canonical pair counts per kind (code/nans.cpp:1355-1535): CC, CF, SF, CS, SS."""
    return nb * (nb - 1) // 2, nb, ns, nb * ns, ns * (ns - 1) // 2


# ---------------------------------------------------------------------- files of the headless driver
def save_scene(scene: Scene, path: str) -> None:
    """Write a scene in the text format nans_projekat_amd/driver/nans_headless.cpp
    reads ("nans-scene 1"); binary32 values are written with "%.9g", which reads back
    exactly.  Force and torque are not part of the format (zero)."""
    f9 = lambda x: "%.9g" % float(np.float32(x))
    p = scene.params
    with open(path, "w") as f:
        f.write("nans-scene 1\n")
        f.write("params %d %s %s %s %s %s %s %d %d %s %d %s %s %s %s\n" % (
            p.solver_iterations, f9(p.beta), f9(p.restitution), f9(p.friction), f9(p.damping), f9(p.gravity),
            f9(p.dt), p.max_gjk_iterations, p.max_epa_iterations, f9(p.epa_tolerance), p.fp64_islands,
            " ".join(f9(x) for x in p.floor_pos), " ".join(f9(x) for x in p.floor_scale), f9(p.floor_mass),
            f9(p.floor_moi)))
        f.write("worlds %d\n" % scene.nworlds)
        for w in range(scene.nworlds):
            a, b = int(scene.offsets[w]), int(scene.offsets[w + 1])
            fr = "-" if scene.world_friction is None else f9(scene.world_friction[w])
            rs = "-" if scene.world_restitution is None else f9(scene.world_restitution[w])
            f.write("world %d %s %s\n" % (b - a, fr, rs))
            for i in range(a, b):
                vals = [*scene.pos[i], *scene.vel[i], *scene.angvel[i], *scene.angles[i], *scene.scale[i],
                        scene.mass[i], scene.moi[i]]
                f.write("body %s %s\n" % ("box" if scene.kind[i] == BOX else "sphere", " ".join(f9(v) for v in vals)))


def load_trajectory(path: str) -> dict:
    """Read a trajectory written by nans_headless: {"steps": [nframes], "pos":
    [nframes][nbodies][3], "angles": [nframes][nbodies][3], "steps_per_frame": k}."""
    with open(path, "rb") as f:
        if f.read(8) != b"NANSTRJ1":
            raise ValueError("%s is not a NANSTRJ1 trajectory" % path)
        n, nframes, per = np.frombuffer(f.read(12), "<i4")
        rec = np.dtype([("step", "<i4"), ("pos", "<f4", (n, 3)), ("angles", "<f4", (n, 3))])
        fr = np.frombuffer(f.read(), rec, count=nframes)
    return {"steps": fr["step"].copy(), "pos": fr["pos"].copy(), "angles": fr["angles"].copy(),
            "steps_per_frame": int(per)}
