// nans_world.hpp -- per-body, per-pair and per-contact device routines of the
// world step (code/nans.cpp:974-1349, 1351-1548), shared by the plain
// (one world per launch sequence) and batched (one workgroup per world) paths.
//
// Body state is structure-of-arrays: a Planes view holds one plane per field
// with a common stride, pointing either into HBM (plain path) or into LDS
// (batched path).
#pragma once

#include "../../include/nans_gpu.h"
#include "nans_epa.hpp"

namespace nans {

// Field planes of a body view.
// The first F_NLDS fields are the ones the batched path keeps in LDS.
enum : int {
    F_PX = 0, F_PY, F_PZ, F_VX, F_VY, F_VZ, F_WX, F_WY, F_WZ, F_AX, F_AY, F_AZ,
    F_SX, F_SY, F_SZ, F_MASS, F_MOI,
    F_NLDS,
    F_FX = F_NLDS, F_FY, F_FZ, F_TX, F_TY, F_TZ,
    F_QW, F_QX, F_QY, F_QZ,  // orientation of the opt-in quaternion mode
    F_NFIELDS
};

struct Planes {
    float *base;
    int stride;
    __host__ __device__ float &at(int f, int i) const { return base[f * stride + i]; }
    __host__ __device__ V3 get3(int f, int i) const { return mk3(at(f, i), at(f + 1, i), at(f + 2, i)); }
    __host__ __device__ void set3(int f, int i, V3 v) const
    {
        at(f, i) = v.x;
        at(f + 1, i) = v.y;
        at(f + 2, i) = v.z;
    }
};

struct StepConsts {
    NarrowParams np;
    int solver_iterations;
    float beta, restitution, friction, damping, gravity, dt;
    int quat;  // 1: orientation is the quaternion planes (opt-in), 0: Euler degrees
};

// The ground box of code/nans.cpp:1666-1678: static, its V and W read as zero.
struct FloorDev {
    float c[3];
    float v[24];
    float brad;
    float mass, moi;
};

// ------------------------------------------------------------------ RK4 (code/nans.cpp:50-78)
// MovementFunction: (1/m)(F + m*g*(0,-1,0) - damping*v), with 1/m in binary64
__host__ __device__ inline V3 movement(V3 v, V3 F, float m, const StepConsts &c)
{
    float inv = (float)(1.0 / (double)m);
    V3 g = mul(mk3(0.0f, -1.0f, 0.0f), m * c.gravity);
    return mul(sub(add(F, g), mul(v, c.damping)), inv);
}

// RotationFunction: (1/I)(tau - damping*w)
__host__ __device__ inline V3 rotation(V3 w, V3 T, float I, const StepConsts &c)
{
    float inv = 1.0f / I;
    return mul(sub(T, mul(w, c.damping)), inv);
}

template <bool ROT>
__host__ __device__ inline V3 rk4(float dt, V3 y0, V3 F, float m, const StepConsts &c)
{
    auto f = [&](V3 y) { return ROT ? rotation(y, F, m, c) : movement(y, F, m, c); };
    V3 k1 = mul(f(y0), dt);
    V3 k2 = mul(f(add(y0, divs(k1, 2.0f))), dt);
    V3 k3 = mul(f(add(y0, divs(k2, 2.0f))), dt);
    V3 k4 = mul(f(add(y0, k3)), dt);
    V3 sum = add(add(add(k1, mul(k2, 2.0f)), mul(k3, 2.0f)), k4);
    return add(y0, mul(sum, 1.0f / 6.0f));
}

// IntegrateForces for one body (code/nans.cpp:974-1018): RK4 on V and W with the
// force and torque F, T (cleared by the caller).
__host__ __device__ inline void integrate_forces_body(const Planes &B, int i, V3 F, V3 T, const StepConsts &c)
{
    V3 V = B.get3(F_VX, i), W = B.get3(F_WX, i);
    float m = B.at(F_MASS, i), I = B.at(F_MOI, i);
    B.set3(F_VX, i, rk4<false>(c.dt, V, F, m, c));
    B.set3(F_WX, i, rk4<true>(c.dt, W, T, I, c));
}

// IntegrateVelocities for one body (code/nans.cpp:1331-1349)
__host__ __device__ inline void integrate_velocities_body(const Planes &B, int i, float dt)
{
    B.set3(F_PX, i, add(B.get3(F_PX, i), mul(B.get3(F_VX, i), dt)));
    B.set3(F_AX, i, add(B.get3(F_AX, i), mul(B.get3(F_WX, i), dt)));
}

// IntegrateVelocities of the opt-in quaternion mode: P += dt V, and
// q <- normalize(q + dt/2 (0, W) q) with W the world-frame angular velocity in
// rad/s (the quantity the Constraint terms W x N treat it as); in exact
// arithmetic one step rotates by 2 atan(|W| dt / 2) about W.  The quaternion
// planes may live in another view than the pose (Q: global, B: LDS).
__host__ __device__ inline void integrate_velocities_quat(const Planes &B, const Planes &Q, int i, float dt)
{
    B.set3(F_PX, i, add(B.get3(F_PX, i), mul(B.get3(F_VX, i), dt)));
    const V3 W = B.get3(F_WX, i);
    const float h = 0.5f * dt;
    const float w = Q.at(F_QW, i), x = Q.at(F_QX, i), y = Q.at(F_QY, i), z = Q.at(F_QZ, i);
    const float nw = w - h * ((W.x * x + W.y * y) + W.z * z);
    const float nx = x + h * (w * W.x + (W.y * z - W.z * y));
    const float ny = y + h * (w * W.y + (W.z * x - W.x * z));
    const float nz = z + h * (w * W.z + (W.x * y - W.y * x));
    const float inv = 1.0f / sqrtf(((nw * nw + nx * nx) + ny * ny) + nz * nz);
    Q.at(F_QW, i) = nw * inv;
    Q.at(F_QX, i) = nx * inv;
    Q.at(F_QY, i) = ny * inv;
    Q.at(F_QZ, i) = nz * inv;
}

// Corners of a box Model (UpdateVertices, code/nans.cpp:394-407) and the
// bounding radius max |v - P| used by the cull.
__device__ inline float box_vertices_model(const M4 &m, V3 P, float *out24)
{
    V3 v[8];
    model_vertices(m, v);
    float r2 = 0.0f;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        out24[3 * k] = v[k].x;
        out24[3 * k + 1] = v[k].y;
        out24[3 * k + 2] = v[k].z;
        V3 d = sub(v[k], P);
        float d2 = dot(d, d);
        bad |= d2 != d2;
        r2 = fmaxf(r2, d2);
    }
    // a NaN corner makes the radius NaN, which keeps every pair of the body
    return bad ? __int_as_float(0x7fc00000) : sqrtf(r2);
}

// Box corners of the Model rebuilt from the pose (code/nans.cpp:1913-1941), or in
// the quaternion mode from the quaternion planes of Q.
__device__ inline float box_vertices_body(const Planes &B, int i, float *out24)
{
    const V3 P = B.get3(F_PX, i);
    return box_vertices_model(model_from_pose(P, B.get3(F_AX, i), B.get3(F_SX, i)), P, out24);
}

__device__ inline float box_vertices_quat(const Planes &B, const Planes &Q, int i, float *out24)
{
    const V3 P = B.get3(F_PX, i);
    return box_vertices_model(
        model_from_quat(P, Q.at(F_QW, i), Q.at(F_QX, i), Q.at(F_QY, i), Q.at(F_QZ, i), B.get3(F_SX, i)), P, out24);
}

// ------------------------------------------------------------------ pairs (code/nans.cpp:1355-1535)
struct PairId {
    int type, i, j;  // i, j in the reference's index spaces (Cubes[] / Spheres[])
};

__host__ __device__ inline int64_t tri(int64_t n) { return n * (n - 1) / 2; }

__host__ __device__ inline int64_t pair_count(int nb, int ns)
{
    return tri(nb) + nb + ns + (int64_t)nb * ns + tri(ns);
}

// Row i of the upper triangle (i < j) holding linear index p; row i starts at i*(2n-i-1)/2.
__device__ inline void tri_decode(int64_t p, int n, int &i, int &j)
{
    // a binary32 estimate of the row; the two loops below make it exact
    const float b = 2.0f * (float)n - 1.0f;
    int r = (int)((b - __builtin_amdgcn_sqrtf(fmaxf(b * b - 8.0f * (float)p, 0.0f))) * 0.5f);
    if (r < 0) r = 0;
    auto start = [n](int64_t row) { return row * (2 * (int64_t)n - row - 1) / 2; };
    while (r + 1 < n && start(r + 1) <= p) ++r;
    while (r > 0 && start(r) > p) --r;
    i = r;
    j = (int)(p - start(r)) + r + 1;
}

// Canonical pair p of a world: CC(i<j), CF(i), SF(s), CS(cube-major), SS(s<t).
__device__ inline PairId decode_pair(int64_t p, int nb, int ns)
{
    PairId q;
    int64_t ncc = tri(nb);
    if (p < ncc) {
        q.type = NANS_CC;
        tri_decode(p, nb, q.i, q.j);
        return q;
    }
    p -= ncc;
    if (p < nb) {
        q.type = NANS_CF;
        q.i = q.j = (int)p;
        return q;
    }
    p -= nb;
    if (p < ns) {
        q.type = NANS_SF;
        q.i = q.j = (int)p;
        return q;
    }
    p -= ns;
    if (p < (int64_t)nb * ns) {
        q.type = NANS_CS;
        q.i = (int)(p / ns);
        q.j = (int)(p % ns);
        return q;
    }
    p -= (int64_t)nb * ns;
    q.type = NANS_SS;
    tri_decode(p, ns, q.i, q.j);
    return q;
}

// decode_pair for worlds whose pair count fits 24 bits (the batched path):
// binary32 estimates of row and quotient (hardware sqrt and reciprocal) made
// exact by integer fix-ups.  The terms that depend on the world's body counts
// alone are computed once per world (pair_decode32, where every lane of the
// wave runs) and made scalar, so that they cost no vector registers in the
// loops that decode pairs.
__device__ __forceinline__ float wave_uniform(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

struct PairDecode32 {
    int nb, ns;
    float bc, bbc;  // 2 nb - 1 and its square (box-box triangle)
    float bs, bbs;  // 2 ns - 1 and its square (sphere-sphere triangle)
    float rns;      // estimate of 1 / ns
};

// The body counts must be the same in every lane of the wave (one world per
// wave or workgroup), and every lane must run this.
__device__ inline PairDecode32 pair_decode32(int nb, int ns)
{
    PairDecode32 d;
    d.nb = nb;
    d.ns = ns;
    d.bc = wave_uniform(2.0f * (float)nb - 1.0f);
    d.bbc = wave_uniform(d.bc * d.bc);
    d.bs = wave_uniform(2.0f * (float)ns - 1.0f);
    d.bbs = wave_uniform(d.bs * d.bs);
    d.rns = wave_uniform(__builtin_amdgcn_rcpf((float)ns));
    return d;
}

// b = 2n - 1 and bb = b * b as binary32
__device__ inline void tri_decode32(int p, int n, float b, float bb, int &i, int &j)
{
    const float disc = bb - 8.0f * (float)p;
    int r = (int)((b - __builtin_amdgcn_sqrtf(disc > 0.0f ? disc : 0.0f)) * 0.5f);
    r = r < 0 ? 0 : (r > n - 1 ? n - 1 : r);
    auto start = [n](int row) { return row * (2 * n - row - 1) / 2; };
    while (r + 1 < n && start(r + 1) <= p) ++r;
    while (r > 0 && start(r) > p) --r;
    i = r;
    j = p - start(r) + r + 1;
}

__device__ inline PairId decode_pair32(int p, const PairDecode32 &d)
{
    const int nb = d.nb, ns = d.ns;
    PairId q;
    const int ncc = nb * (nb - 1) / 2;
    if (p < ncc) {
        q.type = NANS_CC;
        tri_decode32(p, nb, d.bc, d.bbc, q.i, q.j);
        return q;
    }
    p -= ncc;
    if (p < nb) {
        q.type = NANS_CF;
        q.i = q.j = p;
        return q;
    }
    p -= nb;
    if (p < ns) {
        q.type = NANS_SF;
        q.i = q.j = p;
        return q;
    }
    p -= ns;
    if (p < nb * ns) {
        q.type = NANS_CS;
        int i = (int)((float)p * d.rns);
        i = i < 0 ? 0 : i;
        while (i * ns > p) --i;
        while ((i + 1) * ns <= p) ++i;
        q.i = i;
        q.j = p - i * ns;
        return q;
    }
    p -= nb * ns;
    q.type = NANS_SS;
    tri_decode32(p, ns, d.bs, d.bbs, q.i, q.j);
    return q;
}

// World-local body slots of a pair (boxes first, then spheres); -1 = the floor.
__device__ inline void pair_slots(const PairId &q, int nb, int &a, int &b)
{
    switch (q.type) {
    case NANS_CC: a = q.i; b = q.j; break;
    case NANS_CS: a = q.i; b = nb + q.j; break;
    case NANS_CF: a = q.i; b = -1; break;
    case NANS_SS: a = nb + q.i; b = nb + q.j; break;
    default: a = nb + q.i; b = -1; break;
    }
}

// Conservative cull: the reference's first EvolveSimplex call searches along
// d = normalize(PB - PA) and stops with NoIntersection when dot(d, S_A(d) - S_B(-d)) < 0.
// Every support point lies within R (max vertex distance, or radius) of its centre,
// so that dot is at most -|PB - PA| + R_A + R_B; pairs whose centre distance exceeds
// R_A + R_B by a margin that covers binary32 rounding are therefore exactly the
// pairs that CheckCollision rejects on its first call, and are skipped.
__device__ inline bool cull_keep(V3 pa, float ra, V3 pb, float rb)
{
    V3 d = sub(pb, pa);
    float dist2 = dot(d, d);
    float mag = fabsf(pa.x) + fabsf(pa.y) + fabsf(pa.z) + fabsf(pb.x) + fabsf(pb.y) + fabsf(pb.z);
    float lim = (ra + rb) * 1.0001f + 1e-4f * (1.0f + mag + ra + rb);
    return !(dist2 > lim * lim);
}

// ------------------------------------------------------------------ Constraint (code/nans.cpp:1020-1329)
struct ContactRec {
    int type, ia, ib, slot_a;  // slot_a/slot_b: world-local body slots (-1 = floor)
    int slot_b, pad0, pad1, pad2;
    float pa[3], pb[3], n[3];
    float dln, dlt1, dlt2;
};

struct BodyRef {
    V3 P, V, W;
    float mass, moi;
};

__host__ __device__ inline BodyRef body_ref(const Planes &B, int slot, const FloorDev &fl)
{
    BodyRef r;
    if (slot < 0) {
        r.P = mk3(fl.c[0], fl.c[1], fl.c[2]);
        r.V = r.W = mk3(0.0f, 0.0f, 0.0f);
        r.mass = fl.mass;
        r.moi = fl.moi;
    } else {
        r.P = B.get3(F_PX, slot);
        r.V = B.get3(F_VX, slot);
        r.W = B.get3(F_WX, slot);
        r.mass = B.at(F_MASS, slot);
        r.moi = B.at(F_MOI, slot);
    }
    return r;
}

// Constraint's terms that depend only on positions, masses and the contact
// (code/nans.cpp:1115-1170): they do not change during SolveConstraints.
struct ContactTerms {
    V3 N, T1, T2, RN1, RN2, R1T1, R2T1, R1T2, R2T2;
    float JMJn, JMJt1, JMJt2;
    float bdepth;  // -beta / dt * Depth
    float InvM1, InvM2, InvI1, InvI2;
};

__host__ __device__ inline void contact_terms(const Planes &B, const ContactRec &c, const FloorDev &fl,
                                              const StepConsts &k, ContactTerms &t)
{
    BodyRef A = body_ref(B, c.slot_a, fl);
    BodyRef Bo = body_ref(B, c.slot_b, fl);
    V3 N = normalize(mk3(c.n[0], c.n[1], c.n[2]));
    const V3 PosA = A.P, PosB = Bo.P;
    t.InvI1 = 1.0f / A.moi;
    t.InvI2 = 1.0f / Bo.moi;
    t.InvM1 = 1.0f / A.mass;
    t.InvM2 = 1.0f / Bo.mass;
    if (veq(N, mk3(0.0f, 0.0f, 0.0f))) N = normalize(sub(PosB, PosA));
    const V3 R1 = sub(mk3(c.pa[0], c.pa[1], c.pa[2]), PosA);
    const V3 R2 = sub(mk3(c.pb[0], c.pb[1], c.pb[2]), PosB);
    const V3 T1 = N.x >= 0.57735f ? mk3(N.y, -N.x, 0.0f) : mk3(0.0f, N.z, -N.y);
    const V3 T2 = cross(N, T1);
    const float Depth = dot(sub(add(PosA, R1), add(PosB, R2)), N);
    t.N = N;
    t.T1 = T1;
    t.T2 = T2;
    t.RN1 = cross(R1, N);
    t.RN2 = cross(R2, N);
    float JMJn = t.InvM1 + t.InvM2;
    JMJn += t.InvI1 * dot(t.RN1, t.RN1) - t.InvI2 * dot(neg(t.RN2), neg(t.RN2));
    t.JMJn = 1.0f / JMJn;
    t.bdepth = -k.beta / k.dt * Depth;
    t.R1T1 = cross(R1, T1);
    t.R2T1 = cross(R2, T1);
    t.R1T2 = cross(R1, T2);
    t.R2T2 = cross(R2, T2);
    float JMJt1 = t.InvM1 + t.InvM2;
    JMJt1 += t.InvI1 * dot(t.R1T1, t.R1T1) - t.InvI2 * dot(neg(t.R2T1), neg(t.R2T1));
    t.JMJt1 = 1.0f / JMJt1;
    float JMJt2 = t.InvM1 + t.InvM2;
    JMJt2 += t.InvI1 * dot(t.R1T2, t.R1T2) - t.InvI2 * dot(neg(t.R2T2), neg(t.R2T2));
    t.JMJt2 = 1.0f / JMJt2;
}

// The rest of Constraint (code/nans.cpp:1148-1329): relative velocities from the
// current V, W of the bodies, the impulse iterations, and the application.
// friction / restitution are the pair's coefficients.
// sums (may be null): the accumulators DLNormalSum, DLTangent1Sum, DLTangent2Sum
// at entry (null: zero, as every contact_pair starts) and at exit, for the
// opt-in warm start.
__host__ __device__ inline void contact_apply(const Planes &B, ContactRec &c, const ContactTerms &t,
                                              const StepConsts &k, float Cf, float Cr, float *sums = nullptr)
{
    const V3 zero = mk3(0.0f, 0.0f, 0.0f);
    const V3 V1 = B.get3(F_VX, c.slot_a), W1 = B.get3(F_WX, c.slot_a);
    const V3 V2 = c.slot_b >= 0 ? B.get3(F_VX, c.slot_b) : zero;
    const V3 W2 = c.slot_b >= 0 ? B.get3(F_WX, c.slot_b) : zero;
    const V3 N = t.N, T1 = t.T1, T2 = t.T2;
    V3 dVn = sub(sub(add(V1, cross(W1, N)), V2), cross(W2, N));
    float JdVn = dot(dVn, N);
    float Bb = t.bdepth + Cr * JdVn;
    V3 dVt1 = sub(sub(add(V1, cross(W1, T1)), V2), cross(W2, T1));
    float JdVt1 = dot(dVt1, T1);
    V3 dVt2 = sub(sub(add(V1, cross(W1, T2)), V2), cross(W2, T2));
    float JdVt2 = dot(dVt2, T2);

    float sumN = 0.0f, sumT1 = 0.0f, sumT2 = 0.0f, dln = 0.0f, dlt1 = 0.0f, dlt2 = 0.0f;
    if (sums) {
        sumN = sums[0];
        sumT1 = sums[1];
        sumT2 = sums[2];
    }
    const double cf_sqrt2 = 1.4142135623730951 * (double)Cf;
    // The velocities are read once (code/nans.cpp:1176-1238), so every iteration
    // adds the same lambdas; only the last iteration's deltas are applied, so the
    // earlier ones update the sums alone.
    const float LambdaN = (-JdVn + Bb) * t.JMJn;
    const float LambdaT1 = (-JdVt1) * t.JMJt1;
    const float LambdaT2 = (-JdVt2) * t.JMJt2;
    auto iterate = [&]() {
        sumN += LambdaN;
        if (sumN < 0.0f) sumN = 0.0f;
        const float MaxLT = (float)(cf_sqrt2 * (double)sumN);
        sumT1 += LambdaT1;
        if (sumT1 < -MaxLT) sumT1 = -MaxLT;
        if (sumT1 > MaxLT) sumT1 = MaxLT;
        sumT2 += LambdaT2;
        if (sumT2 < -MaxLT) sumT2 = -MaxLT;
        if (sumT2 > MaxLT) sumT2 = MaxLT;
    };
    if (k.solver_iterations > 0) {
        // unrolled, so that an iteration's friction bound (from its normal sum
        // alone) is computed while the friction sums of earlier ones still run
        if (k.solver_iterations == 10) {
#pragma unroll
            for (int it = 1; it < 10; ++it) iterate();
        } else {
#pragma unroll 4
            for (int it = 1; it < k.solver_iterations; ++it) iterate();
        }
        const float oldN = sumN, oldT1 = sumT1, oldT2 = sumT2;
        iterate();
        dln = sumN - oldN;
        dlt1 = sumT1 - oldT1;
        dlt2 = sumT2 - oldT2;
    }
    c.dln = dln;
    c.dlt1 = dlt1;
    c.dlt2 = dlt2;
    if (sums) {
        sums[0] = sumN;
        sums[1] = sumT1;
        sums[2] = sumT2;
    }
    const float InvM1 = t.InvM1, InvM2 = t.InvM2, InvI1 = t.InvI1, InvI2 = t.InvI2;
    V3 LinN = mul(N, dln), LinT1 = mul(T1, dlt1), LinT2 = mul(T2, dlt2);
    V3 AngI1 = mul(t.RN1, dln), AngI2 = mul(t.RN2, dln);
    V3 AngI1T1 = mul(t.R1T1, dlt1), AngI2T1 = mul(t.R2T1, dlt1);
    V3 AngI1T2 = mul(t.R1T2, dlt2), AngI2T2 = mul(t.R2T2, dlt2);
    V3 va = V1, wa = W1, vb = V2, wb = W2;
    va = add(va, mul(LinN, InvM1));
    vb = sub(vb, mul(LinN, InvM2));
    wa = add(wa, mul(AngI1, InvI1));
    wb = sub(wb, mul(AngI2, InvI2));
    va = add(va, mul(LinT1, InvM1));
    vb = sub(vb, mul(LinT1, InvM2));
    wa = add(wa, mul(AngI1T1, InvI1));
    wb = sub(wb, mul(AngI2T1, InvI2));
    va = add(va, mul(LinT2, InvM1));
    vb = sub(vb, mul(LinT2, InvM2));
    wa = add(wa, mul(AngI1T2, InvI1));
    wb = sub(wb, mul(AngI2T2, InvI2));
    B.set3(F_VX, c.slot_a, va);
    B.set3(F_WX, c.slot_a, wa);
    if (c.slot_b >= 0) {  // the floor's velocities are never written (code/nans.cpp:1280-1289)
        B.set3(F_VX, c.slot_b, vb);
        B.set3(F_WX, c.slot_b, wb);
    }
}

// Solve one contact against the current V, W of its bodies and apply the impulses
// (code/nans.cpp:1020-1329).
__host__ __device__ inline void constraint(const Planes &B, ContactRec &c, const FloorDev &fl, const StepConsts &k,
                                           float Cf, float Cr, float *sums = nullptr)
{
    ContactTerms t;
    contact_terms(B, c, fl, k, t);
    contact_apply(B, c, t, k, Cf, Cr, sums);
}

}  // namespace nans
