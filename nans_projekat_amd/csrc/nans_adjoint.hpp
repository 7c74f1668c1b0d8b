// nans_adjoint.hpp -- reverse-mode derivatives of one world step with the
// GJK/EPA outcome held fixed.
//
// The forward step (nans_world.hpp) is RK4 on V and W with constant force
// (code/nans.cpp:50-78, 974-1018), Constraint for each contact in list order
// (code/nans.cpp:1020-1329) and IntegrateVelocities (code/nans.cpp:1331-1349).
// The adjoint treats the contact set and each contact's frame as constants: the
// unit normal N and the lever arms R1 = PointA - PosA, R2 = PointB - PosB of
// code/nans.cpp:1115-1124.  The bodies' positions still enter through
// Depth = dot((PosA + R1) - (PosB + R2), N) (code/nans.cpp:1139), their
// velocities through the relative velocities, and mass, moment of inertia,
// friction and restitution through every term that holds them.
//
// The impulse iterations are reversed as unrolled (code/nans.cpp:1176-1238):
// the forward sweep is replayed to record sumN after each iteration and which
// clamp fired, then the iterations are walked backwards; a clamped accumulator
// passes its adjoint to the clamp bound (and so to Cf and sumN) instead of to
// its input, as the derivative of the branch taken.
//
// Everything here is __host__ __device__ so that the same code can be checked
// on the CPU.
#pragma once

#include "nans_world.hpp"

namespace nans {

constexpr int ADJ_MAX_ITERS = 128;  // impulse iterations the reverse sweep records

// Reverse of rk4<ROT> (nans_world.hpp) for a fixed force F: returns the adjoint
// of y0 given the adjoint gy1 of the result, and adds the adjoint of the mass
// (ROT = false) or moment of inertia (ROT = true) to gm.
template <bool ROT>
__host__ __device__ inline V3 rk4_adjoint(float dt, V3 y0, V3 F, float m, const StepConsts &c, V3 gy1, float &gm)
{
    const float inv = ROT ? 1.0f / m : (float)(1.0 / (double)m);
    const V3 gvec = ROT ? mk3(0.0f, 0.0f, 0.0f) : mul(mk3(0.0f, -1.0f, 0.0f), m * c.gravity);
    // f(y) = (F + gvec - damping*y) * inv, k = f(y)*dt
    auto f = [&](V3 y) { return mul(sub(add(F, gvec), mul(y, c.damping)), inv); };
    const V3 k1 = mul(f(y0), dt);
    const V3 y1 = add(y0, divs(k1, 2.0f));
    const V3 k2 = mul(f(y1), dt);
    const V3 y2 = add(y0, divs(k2, 2.0f));
    const V3 k3 = mul(f(y2), dt);
    const V3 y3 = add(y0, k3);
    // y1' = y0 + (k1 + 2 k2 + 2 k3 + k4) / 6
    const V3 gs = mul(gy1, 1.0f / 6.0f);
    V3 gk1 = gs, gk2 = mul(gs, 2.0f), gk3 = mul(gs, 2.0f);
    const V3 gk4 = gs;
    V3 gy0 = gy1;
    float ginv = 0.0f, gdir = 0.0f;
    // one stage k = f(y)*dt: adjoint of y, and of inv and m (through gvec)
    auto stage = [&](V3 y, V3 gk) {
        const V3 gf = mul(gk, dt);
        ginv += dot(sub(add(F, gvec), mul(y, c.damping)), gf);
        if (!ROT) gdir += -c.gravity * gf.y * inv;
        return mul(gf, -c.damping * inv);
    };
    const V3 gy3 = stage(y3, gk4);
    gy0 = add(gy0, gy3);
    gk3 = add(gk3, gy3);
    const V3 gy2 = stage(y2, gk3);
    gy0 = add(gy0, gy2);
    gk2 = add(gk2, mul(gy2, 0.5f));
    const V3 gyb = stage(y1, gk2);
    gy0 = add(gy0, gyb);
    gk1 = add(gk1, mul(gyb, 0.5f));
    gy0 = add(gy0, stage(y0, gk1));
    gm += gdir - inv * inv * ginv;
    return gy0;
}

// Forward values Constraint read for one contact.
struct ContactBodies {
    V3 PA, PB;          // positions (PB: the floor's centre for floor pairs)
    V3 V1, W1, V2, W2;  // velocities when the contact was applied (floor: zero)
    float mA, IA, mB, IB;
    bool two_sided;     // B is a body (its velocities are written)
};

// Adjoints of one contact: gV*/gW* hold the adjoints of the velocities after the
// contact on entry and before it on return; the rest is output.
struct ContactGrad {
    V3 gV1, gW1, gV2, gW2;
    V3 gPA, gPB;
    float gmA, gIA, gmB, gIB, gCf, gCr;
};

__host__ __device__ inline void constraint_adjoint(const ContactRec &c, const ContactBodies &b, const StepConsts &k,
                                                   float Cf, float Cr, ContactGrad &g)
{
    // ---- forward values, as constraint() computes them
    V3 N = normalize(mk3(c.n[0], c.n[1], c.n[2]));
    if (veq(N, mk3(0.0f, 0.0f, 0.0f))) N = normalize(sub(b.PB, b.PA));
    const V3 R1 = sub(mk3(c.pa[0], c.pa[1], c.pa[2]), b.PA);
    const V3 R2 = sub(mk3(c.pb[0], c.pb[1], c.pb[2]), b.PB);
    const V3 T1 = N.x >= 0.57735f ? mk3(N.y, -N.x, 0.0f) : mk3(0.0f, N.z, -N.y);
    const V3 T2 = cross(N, T1);
    const float Depth = dot(sub(add(b.PA, R1), add(b.PB, R2)), N);
    const float InvI1 = 1.0f / b.IA, InvI2 = 1.0f / b.IB;
    const float InvM1 = 1.0f / b.mA, InvM2 = 1.0f / b.mB;
    const V3 RN1 = cross(R1, N), RN2 = cross(R2, N);
    const V3 R1T1 = cross(R1, T1), R2T1 = cross(R2, T1), R1T2 = cross(R1, T2), R2T2 = cross(R2, T2);
    const float qn1 = dot(RN1, RN1), qn2 = dot(RN2, RN2);
    const float q11 = dot(R1T1, R1T1), q21 = dot(R2T1, R2T1), q12 = dot(R1T2, R1T2), q22 = dot(R2T2, R2T2);
    const float JMJn = 1.0f / ((InvM1 + InvM2) + (InvI1 * qn1 - InvI2 * qn2));
    const float JMJt1 = 1.0f / ((InvM1 + InvM2) + (InvI1 * q11 - InvI2 * q21));
    const float JMJt2 = 1.0f / ((InvM1 + InvM2) + (InvI1 * q12 - InvI2 * q22));
    const float JdVn = dot(sub(sub(add(b.V1, cross(b.W1, N)), b.V2), cross(b.W2, N)), N);
    const float JdVt1 = dot(sub(sub(add(b.V1, cross(b.W1, T1)), b.V2), cross(b.W2, T1)), T1);
    const float JdVt2 = dot(sub(sub(add(b.V1, cross(b.W1, T2)), b.V2), cross(b.W2, T2)), T2);
    const float kb = -k.beta / k.dt;
    const float Bb = kb * Depth + Cr * JdVn;
    const float L = (-JdVn + Bb) * JMJn;
    const float LT1 = (-JdVt1) * JMJt1;
    const float LT2 = (-JdVt2) * JMJt2;

    // ---- replay of the iterations: sumN after each, and the clamps taken
    //      code bit 0: sumN clamped to 0; bits 1-2 / 3-4: tangent 1 / 2 set to
    //      -MaxLT (1) or +MaxLT (2)
    const int K = k.solver_iterations < ADJ_MAX_ITERS ? k.solver_iterations : ADJ_MAX_ITERS;
    float shist[ADJ_MAX_ITERS];
    unsigned char code[ADJ_MAX_ITERS];
    const double cf2 = 1.4142135623730951 * (double)Cf;
    float sumN = 0.0f, sumT1 = 0.0f, sumT2 = 0.0f, dln = 0.0f, dlt1 = 0.0f, dlt2 = 0.0f;
    for (int it = 0; it < K; ++it) {
        unsigned char cd = 0;
        const float old = sumN;
        sumN += L;
        if (sumN < 0.0f) {
            sumN = 0.0f;
            cd |= 1;
        }
        dln = sumN - old;
        const float M = (float)(cf2 * (double)sumN);
        float o1 = sumT1;
        sumT1 += LT1;
        int t1 = 0;
        if (sumT1 < -M) { sumT1 = -M; t1 = 1; }
        if (sumT1 > M) { sumT1 = M; t1 = 2; }
        dlt1 = sumT1 - o1;
        float o2 = sumT2;
        sumT2 += LT2;
        int t2 = 0;
        if (sumT2 < -M) { sumT2 = -M; t2 = 1; }
        if (sumT2 > M) { sumT2 = M; t2 = 2; }
        dlt2 = sumT2 - o2;
        shist[it] = sumN;
        code[it] = (unsigned char)(cd | (t1 << 1) | (t2 << 3));
    }

    // ---- reverse of the impulse application
    const V3 lin = add(add(mul(N, dln), mul(T1, dlt1)), mul(T2, dlt2));
    float gdln = InvM1 * dot(N, g.gV1) + InvI1 * dot(RN1, g.gW1);
    float gdlt1 = InvM1 * dot(T1, g.gV1) + InvI1 * dot(R1T1, g.gW1);
    float gdlt2 = InvM1 * dot(T2, g.gV1) + InvI1 * dot(R1T2, g.gW1);
    float gInvM1 = dot(lin, g.gV1);
    float gInvI1 = dot(add(add(mul(RN1, dln), mul(R1T1, dlt1)), mul(R1T2, dlt2)), g.gW1);
    float gInvM2 = 0.0f, gInvI2 = 0.0f;
    if (b.two_sided) {
        gdln -= InvM2 * dot(N, g.gV2) + InvI2 * dot(RN2, g.gW2);
        gdlt1 -= InvM2 * dot(T1, g.gV2) + InvI2 * dot(R2T1, g.gW2);
        gdlt2 -= InvM2 * dot(T2, g.gV2) + InvI2 * dot(R2T2, g.gW2);
        gInvM2 = -dot(lin, g.gV2);
        gInvI2 = -dot(add(add(mul(RN2, dln), mul(R2T1, dlt1)), mul(R2T2, dlt2)), g.gW2);
    }

    // ---- reverse of the iterations (only the last one's deltas are outputs)
    float gs = 0.0f, gt1 = 0.0f, gt2 = 0.0f;
    float gL = 0.0f, gLT1 = 0.0f, gLT2 = 0.0f, gCf = 0.0f;
    for (int it = K - 1; it >= 0; --it) {
        const bool last = it == K - 1;
        const int cd = code[it];
        const float sn = shist[it];
        float gM = 0.0f;
        // tangent 2: t' = clamp(t + LT2), dlt2 = t' - t
        {
            const float gnew = gt2 + (last ? gdlt2 : 0.0f);
            const int br = (cd >> 3) & 3;
            float gv = 0.0f;
            if (br == 0) gv = gnew;
            else gM += br == 1 ? -gnew : gnew;
            gt2 = gv - (last ? gdlt2 : 0.0f);
            gLT2 += gv;
        }
        // tangent 1
        {
            const float gnew = gt1 + (last ? gdlt1 : 0.0f);
            const int br = (cd >> 1) & 3;
            float gv = 0.0f;
            if (br == 0) gv = gnew;
            else gM += br == 1 ? -gnew : gnew;
            gt1 = gv - (last ? gdlt1 : 0.0f);
            gLT1 += gv;
        }
        // MaxLT = sqrt(2) Cf sumN
        const float c2 = (float)cf2;
        float gsn = gs + (last ? gdln : 0.0f) + c2 * gM;
        gCf += 1.41421356f * sn * gM;
        // sumN' = max(sumN + L, 0) (as the branch taken), dln = sumN' - sumN
        const float gu = (cd & 1) ? 0.0f : gsn;
        gs = gu - (last ? gdln : 0.0f);
        gL += gu;
    }

    // ---- reverse of the impulse terms
    const float gJMJn = (-JdVn + Bb) * gL;
    const float gBb = JMJn * gL;
    const float gJdVn = -JMJn * gL + Cr * gBb;
    const float gDepth = kb * gBb;
    const float gCr = JdVn * gBb;
    const float gJdVt1 = -JMJt1 * gLT1, gJMJt1 = -JdVt1 * gLT1;
    const float gJdVt2 = -JMJt2 * gLT2, gJMJt2 = -JdVt2 * gLT2;
    const float gDn = -JMJn * JMJn * gJMJn;
    const float gDt1 = -JMJt1 * JMJt1 * gJMJt1;
    const float gDt2 = -JMJt2 * JMJt2 * gJMJt2;
    gInvM1 += gDn + gDt1 + gDt2;
    gInvM2 += gDn + gDt1 + gDt2;
    gInvI1 += qn1 * gDn + q11 * gDt1 + q12 * gDt2;
    gInvI2 -= qn2 * gDn + q21 * gDt1 + q22 * gDt2;
    // relative velocities: d/dV1 dot(V1 + W1 x X - ..., X) = X, d/dW1 = X x X (zero)
    const V3 gdv = add(add(mul(N, gJdVn), mul(T1, gJdVt1)), mul(T2, gJdVt2));
    const V3 gdw = add(add(cross(N, mul(N, gJdVn)), cross(T1, mul(T1, gJdVt1))), cross(T2, mul(T2, gJdVt2)));
    g.gV1 = add(g.gV1, gdv);
    g.gW1 = add(g.gW1, gdw);
    if (b.two_sided) {
        g.gV2 = sub(g.gV2, gdv);
        g.gW2 = sub(g.gW2, gdw);
    }
    g.gPA = mul(N, gDepth);
    g.gPB = mul(N, -gDepth);
    g.gmA = -InvM1 * InvM1 * gInvM1;
    g.gIA = -InvI1 * InvI1 * gInvI1;
    g.gmB = b.two_sided ? -InvM2 * InvM2 * gInvM2 : 0.0f;
    g.gIB = b.two_sided ? -InvI2 * InvI2 * gInvI2 : 0.0f;
    g.gCf = gCf;
    g.gCr = gCr;
}

// Per-contact record of the adjoint kernel: the velocities Constraint read, the
// dependency round the contact was applied in, and its friction/restitution
// adjoints (summed per world in contact order afterwards).
struct AdjTape {
    float v1[3], w1[3], v2[3], w2[3];
    int round;
    float gcf, gcr;
    int pad;
};

}  // namespace nans
