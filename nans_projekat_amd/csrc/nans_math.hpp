// nans_math.hpp -- binary32 vector arithmetic with the operation order of glm,
// the (un-vendored) math library of the reference (code/nans.h:6-10).
//
// Every routine is written so that, compiled with -ffp-contract=off and
// IEEE-correct division and square root (hipcc's defaults for HIP), it yields
// the same bits as the reference's x86-64 SSE build (code/build.sh:5-14):
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)   = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   normalize(v) = v * (1 / sqrt(dot(v,v)))
//   length(v)    = sqrt(dot(v,v))
//   abs(x)       = x >= 0 ? x : -x
//   mat4 * vec4  = (c0*x + c1*y) + (c2*z + c3*w)
// sin/cos of the Model rebuild are the correctly rounded binary32 values (the
// binary64 results rounded once).  The reference calls the binary32 sinf/cosf
// of whatever libm it is linked with; glibc's are not correctly rounded and
// differ from these by one ulp for about 1% of the angles the steps produce,
// so results equal a glibc-linked reference's only where no such angle occurs.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

namespace nans {

struct V3 {
    float x, y, z;
};

__host__ __device__ inline V3 mk3(float x, float y, float z) { return V3{x, y, z}; }
__host__ __device__ inline V3 add(V3 a, V3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline V3 sub(V3 a, V3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline V3 mul(V3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline V3 divs(V3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__host__ __device__ inline V3 neg(V3 a) { return mk3(-a.x, -a.y, -a.z); }
__host__ __device__ inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__host__ __device__ inline V3 cross(V3 a, V3 b)
{
    return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__host__ __device__ inline float gsqrt(float x) { return __builtin_sqrtf(x); }
__host__ __device__ inline V3 normalize(V3 v) { return mul(v, 1.0f / gsqrt(dot(v, v))); }
__host__ __device__ inline float length(V3 v) { return gsqrt(dot(v, v)); }
__host__ __device__ inline float gabs(float x) { return x >= 0.0f ? x : -x; }
__host__ __device__ inline bool veq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
__host__ __device__ inline bool has_nan(V3 a) { return a.x != a.x || a.y != a.y || a.z != a.z; }

// Column-major glm::mat4, m[c*4 + r].
struct M4 {
    float m[16];
};

// sin and cos in binary64, rounded to binary32, with one argument reduction
__device__ inline void dsincosf(float a, float &s, float &c)
{
    double sd, cd;
    sincos((double)a, &sd, &cd);
    s = (float)sd;
    c = (float)cd;
}

// glm::translate(identity, v)
__device__ inline M4 m4_translate_identity(V3 v)
{
    M4 r;
#pragma unroll
    for (int i = 0; i < 16; ++i) r.m[i] = 0.0f;
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    M4 t = r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        t.m[12 + i] = ((r.m[0 + i] * v.x + r.m[4 + i] * v.y) + r.m[8 + i] * v.z) + r.m[12 + i];
    return t;
}

// glm::rotate(m, a, axis) (Rodrigues form of matrix_transform.inl)
__device__ inline M4 m4_rotate(const M4 &m, float a, V3 v)
{
    float s, c;
    dsincosf(a, s, c);
    V3 axis = normalize(v);
    V3 temp = mul(axis, 1.0f - c);
    float R[3][3];
    R[0][0] = c + temp.x * axis.x;
    R[0][1] = temp.x * axis.y + s * axis.z;
    R[0][2] = temp.x * axis.z - s * axis.y;
    R[1][0] = temp.y * axis.x - s * axis.z;
    R[1][1] = c + temp.y * axis.y;
    R[1][2] = temp.y * axis.z + s * axis.x;
    R[2][0] = temp.z * axis.x + s * axis.y;
    R[2][1] = temp.z * axis.y - s * axis.x;
    R[2][2] = c + temp.z * axis.z;
    M4 r;
#pragma unroll
    for (int col = 0; col < 3; ++col)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            r.m[col * 4 + i] = (m.m[0 + i] * R[col][0] + m.m[4 + i] * R[col][1]) + m.m[8 + i] * R[col][2];
#pragma unroll
    for (int i = 0; i < 4; ++i) r.m[12 + i] = m.m[12 + i];
    return r;
}

// glm::scale
__device__ inline M4 m4_scale(const M4 &m, V3 v)
{
    M4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        r.m[0 + i] = m.m[0 + i] * v.x;
        r.m[4 + i] = m.m[4 + i] * v.y;
        r.m[8 + i] = m.m[8 + i] * v.z;
        r.m[12 + i] = m.m[12 + i];
    }
    return r;
}

// xyz of glm mat4 * vec4(x, y, z, w)
__device__ inline V3 m4_apply(const M4 &m, float x, float y, float z, float w)
{
    float o[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        o[i] = (m.m[0 + i] * x + m.m[4 + i] * y) + (m.m[8 + i] * z + m.m[12 + i] * w);
    return mk3(o[0], o[1], o[2]);
}

__device__ inline float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

// Model = translate(P) * rotX(rad(A.x)) * rotY(rad(A.y)) * rotZ(rad(A.z)) * scale(S)
// (code/nans.cpp:1913-1924, 1930-1941, 1870-1879).
__device__ inline M4 model_from_pose(V3 pos, V3 ang, V3 scale)
{
    M4 m = m4_translate_identity(pos);
    m = m4_rotate(m, radians(ang.x), mk3(1.0f, 0.0f, 0.0f));
    m = m4_rotate(m, radians(ang.y), mk3(0.0f, 1.0f, 0.0f));
    m = m4_rotate(m, radians(ang.z), mk3(0.0f, 0.0f, 1.0f));
    return m4_scale(m, scale);
}

// Corner k of UpdateVertices (code/nans.cpp:399-406): (+,+,+),(+,+,-),(-,+,+),(-,+,-),
// (+,-,+),(+,-,-),(-,-,+),(-,-,-) times 0.5.
__device__ inline void box_corner(int k, float &x, float &y, float &z)
{
    x = (k & 2) ? -0.5f : 0.5f;
    y = (k & 4) ? -0.5f : 0.5f;
    z = (k & 1) ? -0.5f : 0.5f;
}

__device__ inline void model_vertices(const M4 &m, V3 *verts)
{
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float x, y, z;
        box_corner(k, x, y, z);
        verts[k] = m4_apply(m, x, y, z, 1.0f);
    }
}

// Opt-in quaternion orientation of the step (no counterpart in the reference,
// whose Angles are Euler degrees): q = (w, x, y, z), Model = [R(q) diag(S) | P].
__device__ inline M4 model_from_quat(V3 pos, float w, float x, float y, float z, V3 scale)
{
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, xz = x * z, yz = y * z;
    const float wx = w * x, wy = w * y, wz = w * z;
    M4 m;
    m.m[0] = (1.0f - 2.0f * (yy + zz)) * scale.x;
    m.m[1] = (2.0f * (xy + wz)) * scale.x;
    m.m[2] = (2.0f * (xz - wy)) * scale.x;
    m.m[3] = 0.0f;
    m.m[4] = (2.0f * (xy - wz)) * scale.y;
    m.m[5] = (1.0f - 2.0f * (xx + zz)) * scale.y;
    m.m[6] = (2.0f * (yz + wx)) * scale.y;
    m.m[7] = 0.0f;
    m.m[8] = (2.0f * (xz + wy)) * scale.z;
    m.m[9] = (2.0f * (yz - wx)) * scale.z;
    m.m[10] = (1.0f - 2.0f * (xx + yy)) * scale.z;
    m.m[11] = 0.0f;
    m.m[12] = pos.x;
    m.m[13] = pos.y;
    m.m[14] = pos.z;
    m.m[15] = 1.0f;
    return m;
}

// Box corners from (centre, unit quaternion (w,x,y,z), half-extents): Model columns
// R(q)[:,k] * (2 h_k) and (c, 1), applied with the glm mat4*vec4 form.  Used by the
// narrowphase entry point, whose shapes carry quaternions.
__device__ inline void box_from_quat(V3 c, float qw, float qx, float qy, float qz, V3 h, V3 *verts)
{
    float xx = qx * qx, yy = qy * qy, zz = qz * qz;
    float xy = qx * qy, xz = qx * qz, yz = qy * qz;
    float wx = qw * qx, wy = qw * qy, wz = qw * qz;
    float r00 = 1.0f - 2.0f * (yy + zz), r01 = 2.0f * (xy - wz), r02 = 2.0f * (xz + wy);
    float r10 = 2.0f * (xy + wz), r11 = 1.0f - 2.0f * (xx + zz), r12 = 2.0f * (yz - wx);
    float r20 = 2.0f * (xz - wy), r21 = 2.0f * (yz + wx), r22 = 1.0f - 2.0f * (xx + yy);
    float sx = 2.0f * h.x, sy = 2.0f * h.y, sz = 2.0f * h.z;
    M4 m;
    m.m[0] = r00 * sx; m.m[1] = r10 * sx; m.m[2] = r20 * sx; m.m[3] = 0.0f;
    m.m[4] = r01 * sy; m.m[5] = r11 * sy; m.m[6] = r21 * sy; m.m[7] = 0.0f;
    m.m[8] = r02 * sz; m.m[9] = r12 * sz; m.m[10] = r22 * sz; m.m[11] = 0.0f;
    m.m[12] = c.x; m.m[13] = c.y; m.m[14] = c.z; m.m[15] = 1.0f;
    model_vertices(m, verts);
}

}  // namespace nans
