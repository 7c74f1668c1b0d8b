// nans_narrow.hpp -- GJK (one lane per pair) and EPA (one wavefront per pair,
// polytope in LDS) with the decision rules of code/nans.cpp:409-966.
//
// GJK runs in registers.  For EPA the pair's shapes and final simplex are
// copied to a per-wave LDS block; the wave then keeps
//   * the vertex list  (P, SupA, SupB, equality class)       -- LDS
//   * the face list    (stored normal, distance, vertex ids) -- LDS, ordered
//   * the horizon edges                                       -- one per lane
// and evaluates each iteration of ResolveCollision (code/nans.cpp:804-902):
//   closest face      = wave argmin of the stored face distances (ties -> lowest
//                       index, the strict '<' scan of code/nans.cpp:809-822);
//   visible faces     = every lane tests its faces, ballot masks;
//   horizon edges     = the faces' edges pushed in face order with the exact
//                       reversed-edge cancellation of PushEdge (code/nans.cpp:
//                       232-266), each push one ballot over the lanes;
//   face removal      = ordered compaction (std::vector::erase keeps order);
//   new faces         = one per horizon edge, appended in edge order.
// A face's distance |dot(n, A)| never changes after it is made, so it is
// computed once; the result equals the reference's per-iteration recomputation
// bit for bit (the recomputed normal is the same expression).
#pragma once

#include "nans_math.hpp"

namespace nans {

enum : int { SHAPE_BOX = 0, SHAPE_SPHERE = 1 };
enum : int { GJK_NO = 0, GJK_FOUND = 1, GJK_CAP = 2, GJK_STILL = 3 };
enum : int {
    EPA_NONE = 0,       // EPA not run
    EPA_CONTACT = 1,    // converged; contact written
    EPA_BARY = 2,       // |barycentric| > 1 -> no contact (code/nans.cpp:836-840)
    EPA_NONFINITE = 3,  // NaN/inf barycentric -> no contact (code/nans.cpp:842-846)
    EPA_ITER = 4,       // 65 iterations without convergence (code/nans.cpp:903)
    EPA_EMPTY = 5,      // every face removed (the reference reads an empty vector)
    EPA_OVERFLOW = 6    // polytope exceeded the LDS capacity: flagged, no contact
};

struct NarrowParams {
    int max_gjk_iterations;   // 64 -> 65 EvolveSimplex calls
    int max_epa_iterations;   // 64 -> 65 iterations
    float epa_tolerance;      // MAX_EPA_ERROR
    int fp64_islands;         // GJK length test and EPA final face/depth in binary64
};

struct Shape {
    int kind;
    V3 c;
    float r;
    V3 v[8];
    const float *vm;  // the same corners in memory ([8][3]), or null
};

struct Vtx {
    V3 P, A, B;  // Minkowski point, support on A, support on B
};

// GetCubeSupport / GetFloorSupport (code/nans.cpp:409-461): strict '>' argmax from
// -FLT_MAX with a zero default (a NaN direction returns (0,0,0)).
__device__ inline V3 box_support(const V3 *v, V3 d)
{
    float maxd = -FLT_MAX;
    V3 r = mk3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float dist = dot(v[k], d);
        if (dist > maxd) {
            maxd = dist;
            r = v[k];
        }
    }
    return r;
}

// The same argmax when the corners are also in memory: the scan over the
// registers selects an index (two selects per corner instead of four) and the
// winner is read back from memory.
__device__ inline V3 box_support_shape(const Shape &s, V3 d)
{
    if (!s.vm) return box_support(s.v, d);
    float maxd = -FLT_MAX;
    int idx = -1;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float dist = dot(s.v[k], d);
        if (dist > maxd) {
            maxd = dist;
            idx = k;
        }
    }
    if (idx < 0) return mk3(0.0f, 0.0f, 0.0f);
    const float *w = s.vm + 3 * idx;
    return mk3(w[0], w[1], w[2]);
}

// GetSphereSupport (code/nans.cpp:432-438)
__device__ inline V3 sphere_support(V3 c, float r, V3 d) { return add(c, mul(normalize(d), r)); }

__device__ inline V3 support(const Shape &s, V3 d)
{
    return s.kind == SHAPE_BOX ? box_support(s.v, d) : sphere_support(s.c, s.r, d);
}

// CalculateSupport (code/nans.cpp:463-519).  A sphere's support along -d uses
// normalize(-d) = -normalize(d) and c + r (-n) = c - r n, both exact in IEEE
// arithmetic, so one normalize serves both shapes.
__device__ inline Vtx minkowski(const Shape &a, const Shape &b, V3 d)
{
    Vtx v;
    const bool sa = a.kind != SHAPE_BOX, sb = b.kind != SHAPE_BOX;
    V3 n = mk3(0.0f, 0.0f, 0.0f);
    if (sa || sb) n = normalize(d);
    v.A = sa ? add(a.c, mul(n, a.r)) : box_support_shape(a, d);
    v.B = sb ? sub(b.c, mul(n, b.r)) : box_support_shape(b, mul(d, -1.0f));
    v.P = sub(v.A, v.B);
    return v;
}

// TripleCross (code/nans.cpp:564-569)
__device__ inline V3 triple_cross(V3 a, V3 b, V3 c) { return sub(mul(b, dot(c, a)), mul(a, dot(c, b))); }

// ---------------------------------------------------------------------- GJK
// Simplex in four named register slots (no array, so nothing is placed in
// scratch); erase() shifts like std::vector::erase.
struct Simplex {
    Vtx s0, s1, s2, s3;
    int n;
};

// Per-component selects: a conditional on struct lvalues would select an address,
// which keeps the simplex out of registers.
__device__ __forceinline__ void fsel(V3 &d, bool c, const V3 &s)
{
    d.x = c ? s.x : d.x;
    d.y = c ? s.y : d.y;
    d.z = c ? s.z : d.z;
}

__device__ __forceinline__ void vsel(Vtx &d, bool c, const Vtx &s)
{
    fsel(d.P, c, s.P);
    fsel(d.A, c, s.A);
    fsel(d.B, c, s.B);
}

__device__ __forceinline__ void simplex_erase(Simplex &S, int idx)
{
    vsel(S.s0, idx <= 0, S.s1);
    vsel(S.s1, idx <= 1, S.s2);
    vsel(S.s2, idx <= 2, S.s3);
    S.n -= 1;
}

__device__ __forceinline__ void simplex_push(Simplex &S, const Vtx &v)
{
    vsel(S.s0, S.n == 0, v);
    vsel(S.s1, S.n == 1, v);
    vsel(S.s2, S.n == 2, v);
    vsel(S.s3, S.n >= 3, v);
    S.n += 1;
}

// GJK's termination test length(d) <= 0.0001f (code/nans.cpp:755-767) as one
// compare: with a correctly rounded (hence monotone) sqrt it holds exactly when
// dot(d, d) <= the largest binary32 whose square root rounds to <= 0.0001f, which
// is 0x322bcc77 (9.99999994e-09); NaN fails both forms.
constexpr unsigned GJK_LEN2_MAX_BITS = 0x322bcc77u;

// The first direction of a pair's GJK, normalize(PB - PA) (code/nans.cpp:575);
// the second is its negation.
__device__ __forceinline__ V3 gjk_first_dir(const Shape &a, const Shape &b) { return normalize(sub(b.c, a.c)); }

// EvolveSimplex (code/nans.cpp:571-769); d0 = gjk_first_dir(a, b), computed once
// per pair by the caller.
__device__ __forceinline__ int evolve_simplex(const Shape &a, const Shape &b, Simplex &S, int fp64, V3 d0)
{
    // The cases only choose the direction and which vertices leave; the simplex is
    // rearranged once after them (erase er1, then er2; swap s1 and s2), so that
    // lanes in different cases share one copy of the register moves.  Every case
    // reads the simplex before it changes, as the reference's does.
    V3 dir = d0;
    int er1 = -1, er2 = -1;
    bool swap12 = false;
    if (S.n == 1) {
        dir = mul(dir, -1.0f);
    } else if (S.n == 2) {
        // ClosestPointOnLine (code/nans.cpp:539-562)
        V3 A = S.s0.P, B = S.s1.P;
        V3 seg = normalize(sub(B, A));
        float len = length(seg);
        float V = dot(neg(A), seg) / len;
        float U = dot(B, seg) / len;
        V3 cp;
        if (U <= 0.0f) cp = B;
        else if (V <= 0.0f) cp = A;
        else cp = add(mul(A, U), mul(B, V));
        if (V <= 0.0f) er1 = 1;
        else if (U <= 0.0f) er1 = 0;
        dir = neg(cp);
    } else if (S.n == 3) {
        V3 n2o = neg(S.s0.P);
        V3 e1 = sub(S.s1.P, S.s0.P);
        V3 e2 = sub(S.s2.P, S.s0.P);
        V3 tn = cross(e1, e2);
        V3 e1n = cross(e1, tn);
        V3 e2n = cross(tn, e2);
        if (dot(e2n, n2o) > 0.0f) {
            if (dot(e2, n2o) > 0.0f) {
                dir = triple_cross(e2, n2o, e2);
                er1 = 1;
            } else if (dot(e1, n2o) > 0.0f) {
                dir = triple_cross(e1, n2o, e1);
                er1 = 2;
            } else {
                dir = n2o;
                er1 = 2;
                er2 = 1;
            }
        } else if (dot(e1n, n2o) > 0.0f) {
            if (dot(e1, n2o) > 0.0f) {
                dir = triple_cross(e1, n2o, e1);
                er1 = 2;
            } else {
                dir = n2o;
                er1 = 2;
                er2 = 1;
            }
        } else if (dot(tn, n2o) > 0.0f) {
            dir = tn;
        } else {
            dir = neg(tn);
            swap12 = true;
        }
    } else if (S.n == 4) {
        V3 DA = sub(S.s0.P, S.s3.P);
        V3 DB = sub(S.s1.P, S.s3.P);
        V3 DC = sub(S.s2.P, S.s3.P);
        V3 D0 = mul(S.s3.P, -1.0f);
        V3 abd = cross(DA, DB);
        V3 bcd = cross(DB, DC);
        V3 cad = cross(DC, DA);
        if (dot(abd, D0) > 0.0f) {
            er1 = 2;
            dir = abd;
        } else if (dot(bcd, D0) > 0.0f) {
            er1 = 0;
            dir = bcd;
        } else if (dot(cad, D0) > 0.0f) {
            er1 = 1;
            dir = cad;
        } else {
            return GJK_FOUND;
        }
    }
    if (swap12) {
        const Vtx t = S.s1;
        vsel(S.s1, true, S.s2);
        vsel(S.s2, true, t);
    }
    if (er1 >= 0) simplex_erase(S, er1);
    if (er2 >= 0) simplex_erase(S, er2);
    if (fp64) {
        double dx = (double)dir.x, dy = (double)dir.y, dz = (double)dir.z;
        if (__builtin_sqrt((dx * dx + dy * dy) + dz * dz) <= (double)0.0001f) return GJK_NO;
    } else if (dot(dir, dir) <= __uint_as_float(GJK_LEN2_MAX_BITS)) {
        return GJK_NO;
    }
    // AddSupport (code/nans.cpp:521-537): pushed whatever the test decides
    Vtx v = minkowski(a, b, dir);
    simplex_push(S, v);
    return dot(dir, v.P) >= 0.0f ? GJK_STILL : GJK_NO;
}

// The GJK loop of CheckCollision (code/nans.cpp:913-955).  Returns GJK_NO,
// GJK_FOUND or GJK_CAP; *calls receives the number of EvolveSimplex calls.
__device__ inline int gjk(const Shape &a, const Shape &b, const NarrowParams &p, Simplex &S, int *calls)
{
    S.n = 0;
    int res = GJK_STILL;
    unsigned it = 0;
    int c = 0;
    const V3 d0 = gjk_first_dir(a, b);
    while (res == GJK_STILL && it++ <= (unsigned)p.max_gjk_iterations) {
        res = evolve_simplex(a, b, S, p.fp64_islands, d0);
        ++c;
    }
    *calls = c;
    return res == GJK_STILL ? GJK_CAP : res;
}

}  // namespace nans
