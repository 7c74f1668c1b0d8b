"""FP64 reference gradient of the frozen-manifold rollout, by torch autograd.

The function differentiated is the step with the GJK/EPA outcome held fixed
(oracle/nans_oracle.h, "Frozen-manifold rollouts"): per step, RK4 on every
body's V and W (code/nans.cpp:50-78, 974-1018), then Constraint for each
recorded contact frame in list order (code/nans.cpp:1125-1329, with
Depth = dot((PosA + R1) - (PosB + R2), N)), then IntegrateVelocities
(code/nans.cpp:1331-1349).  It is written here independently of the C replay
and of the HIP adjoint, in float64, with torch.where for the impulse clamps (so
the gradient of a clamped branch is that branch's), and serves as the second
reference of the adjoint beside central finite differences of the C replay.
"""
from __future__ import annotations

import math

import numpy as np
import torch

F64 = torch.float64


def _v(x):
    return torch.as_tensor(np.asarray(x, dtype=np.float64), dtype=F64)


def _cross(a, b):
    return torch.stack((a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]))


def _rk4(y0, force, inv, gm, damping, dt):
    f = lambda y: (force + gm - damping * y) * inv
    k1 = f(y0) * dt
    k2 = f(y0 + k1 / 2) * dt
    k3 = f(y0 + k2 / 2) * dt
    k4 = f(y0 + k3) * dt
    return y0 + (k1 + 2 * k2 + 2 * k3 + k4) / 6


def frozen_rollout(params, pos, vel, angvel, mass, moi, friction, restitution, frames, step_count, nsteps,
                   force=None, torque=None):
    """Final (pos, vel, angvel) of the frozen rollout; every argument that is a
    torch tensor with requires_grad takes part in autograd.  pos/vel/angvel are
    [n][3], mass/moi [n], friction/restitution scalars."""
    n = pos.shape[0]
    dt = float(np.float32(params.dt))
    damping = float(np.float32(params.damping))
    gravity = float(np.float32(params.gravity))
    beta = float(np.float32(params.beta))
    iters = int(params.solver_iterations)
    floor_c = _v(np.asarray(params.floor_pos, np.float32))
    floor_inv_m = 1.0 / float(np.float32(params.floor_mass))
    floor_inv_i = 1.0 / float(np.float32(params.floor_moi))
    zero3 = torch.zeros(3, dtype=F64)
    P = [pos[i] for i in range(n)]
    V = [vel[i] for i in range(n)]
    W = [angvel[i] for i in range(n)]
    Fo = [zero3 if force is None else _v(force[i]) for i in range(n)]
    To = [zero3 if torque is None else _v(torque[i]) for i in range(n)]
    down = _v([0.0, -1.0, 0.0])
    inv_m = [1.0 / mass[i] for i in range(n)]
    inv_i = [1.0 / moi[i] for i in range(n)]
    k = 0
    for s in range(nsteps):
        for i in range(n):
            V[i] = _rk4(V[i], Fo[i], inv_m[i], down * (mass[i] * gravity), damping, dt)
            W[i] = _rk4(W[i], To[i], inv_i[i], zero3, damping, dt)
            Fo[i] = zero3
            To[i] = zero3
        for _ in range(int(step_count[s])):
            fr = frames[k]
            k += 1
            a, b = int(fr["slot_a"]), int(fr["slot_b"])
            N = _v(fr["n"])
            R1 = _v(fr["r1"])
            R2 = _v(fr["r2"])
            if float(N[0]) >= float(np.float32(0.57735)):
                T1 = torch.stack((N[1], -N[0], torch.zeros((), dtype=F64)))
            else:
                T1 = torch.stack((torch.zeros((), dtype=F64), N[2], -N[1]))
            T2 = _cross(N, T1)
            im1, ii1 = inv_m[a], inv_i[a]
            if b >= 0:
                im2, ii2, PB, V2, W2 = inv_m[b], inv_i[b], P[b], V[b], W[b]
            else:
                im2, ii2, PB, V2, W2 = floor_inv_m, floor_inv_i, floor_c, zero3, zero3
            PA, V1, W1 = P[a], V[a], W[a]
            depth = torch.dot((PA + R1) - (PB + R2), N)
            RN1, RN2 = _cross(R1, N), _cross(R2, N)
            R1T1, R2T1, R1T2, R2T2 = _cross(R1, T1), _cross(R2, T1), _cross(R1, T2), _cross(R2, T2)
            jn = 1.0 / (im1 + im2 + ii1 * torch.dot(RN1, RN1) - ii2 * torch.dot(RN2, RN2))
            jt1 = 1.0 / (im1 + im2 + ii1 * torch.dot(R1T1, R1T1) - ii2 * torch.dot(R2T1, R2T1))
            jt2 = 1.0 / (im1 + im2 + ii1 * torch.dot(R1T2, R1T2) - ii2 * torch.dot(R2T2, R2T2))
            dvn = torch.dot(V1 + _cross(W1, N) - V2 - _cross(W2, N), N)
            dvt1 = torch.dot(V1 + _cross(W1, T1) - V2 - _cross(W2, T1), T1)
            dvt2 = torch.dot(V1 + _cross(W1, T2) - V2 - _cross(W2, T2), T2)
            bias = -beta / dt * depth + restitution * dvn
            ln = (-dvn + bias) * jn
            lt1 = -dvt1 * jt1
            lt2 = -dvt2 * jt2
            sn = st1 = st2 = torch.zeros((), dtype=F64)
            dln = dlt1 = dlt2 = None
            for _ in range(iters):
                old = sn
                sn = sn + ln
                sn = torch.where(sn < 0, torch.zeros((), dtype=F64), sn)
                dln = sn - old
                cap = math.sqrt(2.0) * friction * sn
                o1 = st1
                st1 = st1 + lt1
                st1 = torch.where(st1 < -cap, -cap, st1)
                st1 = torch.where(st1 > cap, cap, st1)
                dlt1 = st1 - o1
                o2 = st2
                st2 = st2 + lt2
                st2 = torch.where(st2 < -cap, -cap, st2)
                st2 = torch.where(st2 > cap, cap, st2)
                dlt2 = st2 - o2
            if dln is None:
                continue
            lin = N * dln + T1 * dlt1 + T2 * dlt2
            V[a] = V1 + lin * im1
            W[a] = W1 + (RN1 * dln + R1T1 * dlt1 + R1T2 * dlt2) * ii1
            if b >= 0:
                V[b] = V2 - lin * im2
                W[b] = W2 - (RN2 * dln + R2T1 * dlt1 + R2T2 * dlt2) * ii2
        for i in range(n):
            P[i] = P[i] + V[i] * dt
    return torch.stack(P), torch.stack(V), torch.stack(W)
