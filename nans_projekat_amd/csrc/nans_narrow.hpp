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

// GetSphereSupport (code/nans.cpp:432-438)
__device__ inline V3 sphere_support(V3 c, float r, V3 d) { return add(c, mul(normalize(d), r)); }

__device__ inline V3 support(const Shape &s, V3 d)
{
    return s.kind == SHAPE_BOX ? box_support(s.v, d) : sphere_support(s.c, s.r, d);
}

// CalculateSupport (code/nans.cpp:463-519)
__device__ inline Vtx minkowski(const Shape &a, const Shape &b, V3 d)
{
    Vtx v;
    v.A = support(a, d);
    v.B = support(b, mul(d, -1.0f));
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

// EvolveSimplex (code/nans.cpp:571-769)
__device__ __forceinline__ int evolve_simplex(const Shape &a, const Shape &b, Simplex &S, int fp64)
{
    V3 dir = normalize(sub(b.c, a.c));
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
        if (V <= 0.0f) {
            simplex_erase(S, 1);
        } else if (U <= 0.0f) {
            simplex_erase(S, 0);
        }
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
                simplex_erase(S, 1);
            } else if (dot(e1, n2o) > 0.0f) {
                dir = triple_cross(e1, n2o, e1);
                simplex_erase(S, 2);
            } else {
                dir = n2o;
                simplex_erase(S, 2);
                simplex_erase(S, 1);
            }
        } else if (dot(e1n, n2o) > 0.0f) {
            if (dot(e1, n2o) > 0.0f) {
                dir = triple_cross(e1, n2o, e1);
                simplex_erase(S, 2);
            } else {
                dir = n2o;
                simplex_erase(S, 2);
                simplex_erase(S, 1);
            }
        } else if (dot(tn, n2o) > 0.0f) {
            dir = tn;
        } else {
            dir = neg(tn);
            Vtx t = S.s1;
            S.s1 = S.s2;
            S.s2 = t;
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
            simplex_erase(S, 2);
            dir = abd;
        } else if (dot(bcd, D0) > 0.0f) {
            simplex_erase(S, 0);
            dir = bcd;
        } else if (dot(cad, D0) > 0.0f) {
            simplex_erase(S, 1);
            dir = cad;
        } else {
            return GJK_FOUND;
        }
    }
    if (fp64) {
        double dx = (double)dir.x, dy = (double)dir.y, dz = (double)dir.z;
        if (__builtin_sqrt((dx * dx + dy * dy) + dz * dz) <= (double)0.0001f) return GJK_NO;
    } else if (length(dir) <= 0.0001f) {
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
    while (res == GJK_STILL && it++ <= (unsigned)p.max_gjk_iterations) {
        res = evolve_simplex(a, b, S, p.fp64_islands);
        ++c;
    }
    *calls = c;
    return res == GJK_STILL ? GJK_CAP : res;
}

// ---------------------------------------------------------------------- EPA
constexpr int EPA_MAXV = 72;   // 4 + 65 vertices at most (one per iteration)
constexpr int EPA_MAXF = 192;  // alive faces: a closed triangulation of V <= 69 points has 2V-4 <= 134
constexpr int EPA_FSLOTS = EPA_MAXF / 64;

struct EpaLds {
    float vp[EPA_MAXV][3];
    float va[EPA_MAXV][3];
    float vb[EPA_MAXV][3];
    int vcls[EPA_MAXV];    // equality class: first vertex with the same P, or -1-id if P has a NaN
    float fn[EPA_MAXF][3]; // stored normal (flipped away from the origin, PushTriangle)
    float fd[EPA_MAXF];    // |dot(unflipped normal, A)| with glm's abs
    int fid[EPA_MAXF];     // a | b << 8 | c << 16
    float sa[32];          // shape A: kind, c.xyz, r, 24 vertex floats
    float sb[32];
};

struct EpaResult {
    int status;  // EPA_*
    V3 pa, pb, n;
    double depth;
    int iters, max_faces, nverts;
};

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int lane_id() { return (int)(threadIdx.x & 63); }

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

__device__ inline int ffs64(uint64_t m) { return __builtin_ctzll(m); }

__device__ inline void shape_store(float *dst, const Shape &s)
{
    dst[0] = __int_as_float(s.kind);
    dst[1] = s.c.x; dst[2] = s.c.y; dst[3] = s.c.z;
    dst[4] = s.r;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        dst[5 + 3 * k] = s.v[k].x;
        dst[6 + 3 * k] = s.v[k].y;
        dst[7 + 3 * k] = s.v[k].z;
    }
}

// Support of an LDS-resident shape (all lanes read the same words: broadcast).
__device__ inline V3 support_lds(const float *s, V3 d)
{
    int kind = __float_as_int(s[0]);
    if (kind == SHAPE_BOX) {
        float maxd = -FLT_MAX;
        V3 r = mk3(0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            V3 v = mk3(s[5 + 3 * k], s[6 + 3 * k], s[7 + 3 * k]);
            float dist = dot(v, d);
            if (dist > maxd) {
                maxd = dist;
                r = v;
            }
        }
        return r;
    }
    return sphere_support(mk3(s[1], s[2], s[3]), s[4], d);
}

__device__ inline V3 ld3(const float (*a)[3], int i) { return mk3(a[i][0], a[i][1], a[i][2]); }
__device__ inline void st3(float (*a)[3], int i, V3 v)
{
    a[i][0] = v.x;
    a[i][1] = v.y;
    a[i][2] = v.z;
}

// PushTriangle (code/nans.cpp:291-322): normal, flip, and the face's distance.
__device__ inline void make_face(EpaLds &L, int slot, int a, int b, int c)
{
    V3 A = ld3(L.vp, a), B = ld3(L.vp, b), C = ld3(L.vp, c);
    V3 n = normalize(cross(sub(B, A), sub(C, A)));
    float d = gabs(dot(n, A));
    if (dot(A, n) < 0.0f) n = mul(n, -1.0f);
    st3(L.fn, slot, n);
    L.fd[slot] = d;
    L.fid[slot] = a | (b << 8) | (c << 16);
}

// Binary64 distance of face f's plane to the origin (FP64-islands final choice).
__device__ inline double face_dist64(const EpaLds &L, int f)
{
    int id = L.fid[f];
    int a = id & 255, b = (id >> 8) & 255, c = (id >> 16) & 255;
    double ax = L.vp[a][0], ay = L.vp[a][1], az = L.vp[a][2];
    double ux = (double)L.vp[b][0] - ax, uy = (double)L.vp[b][1] - ay, uz = (double)L.vp[b][2] - az;
    double vx = (double)L.vp[c][0] - ax, vy = (double)L.vp[c][1] - ay, vz = (double)L.vp[c][2] - az;
    double nx = uy * vz - vy * uz, ny = uz * vx - vz * ux, nz = ux * vy - vx * uy;
    double inv = 1.0 / __builtin_sqrt((nx * nx + ny * ny) + nz * nz);
    double d = ((nx * inv) * ax + (ny * inv) * ay) + (nz * inv) * az;
    return d >= 0.0 ? d : -d;
}

// DPP move of a 32-bit value (GFX9 encodings: quad_perm 0x00-0xff, row_mirror
// 0x140, row_half_mirror 0x141); every source lane of these patterns exists.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

__device__ __forceinline__ float readlane_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Minimum over the 64 lanes (no NaN inputs): quad and row reductions by DPP, then
// the four row results through readlane.
__device__ __forceinline__ float wave_min(float v)
{
    v = fminf(v, dpp_f<0xB1>(v));   // quad_perm [1,0,3,2]
    v = fminf(v, dpp_f<0x4E>(v));   // quad_perm [2,3,0,1]
    v = fminf(v, dpp_f<0x141>(v));  // row_half_mirror
    v = fminf(v, dpp_f<0x140>(v));  // row_mirror
    return fminf(fminf(readlane_f(v, 0), readlane_f(v, 16)), fminf(readlane_f(v, 32), readlane_f(v, 48)));
}

// Wave argmin over (value, index) among candidates; ties -> lowest index.
template <typename T>
__device__ inline void wave_argmin(bool &valid, T &v, int &i)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        bool ov = __shfl_xor((int)valid, off, 64) != 0;
        T ovv = __shfl_xor(v, off, 64);
        int oi = __shfl_xor(i, off, 64);
        bool take = ov && (!valid || ovv < v || (ovv == v && oi < i));
        if (take) {
            valid = true;
            v = ovv;
            i = oi;
        }
    }
}

// Class of vertex id (its P equals an earlier vertex's P -> that vertex's class).
__device__ inline int vertex_class(const EpaLds &L, int id, V3 P)
{
    if (has_nan(P)) return -1 - id;
    int lane = lane_id();
    for (int base = 0; base < id; base += 64) {
        int u = base + lane;
        bool m = u < id && veq(ld3(L.vp, u), P);
        uint64_t mm = ballot(m);
        if (mm) return base + ffs64(mm);
    }
    return id;
}

__device__ inline bool cls_eq(int x, int y) { return x >= 0 && x == y; }

// ResolveCollision (code/nans.cpp:787-904) executed by one full wavefront.  The
// caller has written the shapes to L.sa / L.sb and the GJK simplex to vertices
// 0..3, then called wave_sync().  Every lane returns the same result.
__device__ inline EpaResult epa_wave(EpaLds &L, const NarrowParams &p)
{
    const int lane = lane_id();
    EpaResult r;
    r.status = EPA_ITER;
    r.pa = r.pb = r.n = mk3(0.0f, 0.0f, 0.0f);
    r.depth = 0.0;
    r.iters = 0;
    r.max_faces = 4;
    r.nverts = 4;

    // vertex_class uses ballots, so every lane evaluates every vertex
    {
        int cls = 0;
        for (int v = 0; v < 4; ++v) {
            int c = vertex_class(L, v, ld3(L.vp, v));
            if (lane == v) cls = c;
        }
        if (lane < 4) L.vcls[lane] = cls;
    }
    if (lane < 4) {
        const int fa[4] = {0, 0, 0, 1}, fb[4] = {1, 2, 3, 3}, fc[4] = {2, 3, 1, 2};
        make_face(L, lane, fa[lane], fb[lane], fc[lane]);
    }
    wave_sync();
    int nf = 4, nv = 4;
    int iter = 0;
    while (iter++ <= p.max_epa_iterations) {
        r.iters += 1;
        if (nf == 0) {
            r.status = EPA_EMPTY;
            break;
        }
        // closest face (code/nans.cpp:807-822): the scan starts from face 0's
        // distance and takes strictly smaller ones, so the result is the first
        // face holding the minimum of the distances below d0 (or face 0)
        int id0 = L.fid[0];
        float d0 = gabs(dot(ld3(L.fn, 0), ld3(L.vp, id0 & 255)));
        float key[EPA_FSLOTS];
        float kmin = __builtin_inff();
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            int f = lane + 64 * k;
            float d = f < nf ? L.fd[f] : __builtin_inff();
            key[k] = d < d0 ? d : __builtin_inff();
            kmin = fminf(kmin, key[k]);
        }
        const float m = wave_min(kmin);
        int best = 0;
        if (m != __builtin_inff()) {
#pragma unroll
            for (int k = EPA_FSLOTS - 1; k >= 0; --k) {
                uint64_t b = ballot(key[k] == m);
                if (b) best = 64 * k + ffs64(b);
            }
        }
        float cur = m == __builtin_inff() ? d0 : L.fd[best];
        V3 N = ld3(L.fn, best);
        // new support along the stored normal
        Vtx ns;
        ns.A = support_lds(L.sa, N);
        ns.B = support_lds(L.sb, mul(N, -1.0f));
        ns.P = sub(ns.A, ns.B);
        if (dot(N, ns.P) - cur < p.epa_tolerance) {
            double depth = (double)cur;
            if (p.fp64_islands) {
                double b0 = face_dist64(L, 0);
                bool v64 = false;
                double bd = 0.0;
                int bi64 = 0x7fffffff;
#pragma unroll
                for (int k = 0; k < EPA_FSLOTS; ++k) {
                    int f = lane + 64 * k;
                    if (f < nf) {
                        double d = face_dist64(L, f);
                        if (d < b0 && (!v64 || d < bd)) {
                            v64 = true;
                            bd = d;
                            bi64 = f;
                        }
                    }
                }
                wave_argmin(v64, bd, bi64);
                int sel = v64 ? bi64 : 0;
                depth = v64 ? bd : b0;
                if (sel != best) {
                    best = sel;
                    cur = sel == 0 ? d0 : L.fd[sel];
                    N = ld3(L.fn, sel);
                }
            }
            int id = L.fid[best];
            int a = id & 255, b = (id >> 8) & 255, c = (id >> 16) & 255;
            V3 PA = ld3(L.vp, a), PB = ld3(L.vp, b), PC = ld3(L.vp, c);
            // Barycentric (code/nans.cpp:771-785) of N * d
            V3 P = mul(N, cur);
            V3 v0 = sub(PB, PA), v1 = sub(PC, PA), v2 = sub(P, PA);
            float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
            float d20 = dot(v2, v0), d21 = dot(v2, v1);
            float denom = d00 * d11 - d01 * d01;
            float V = (d11 * d20 - d01 * d21) / denom;
            float W = (d00 * d21 - d01 * d20) / denom;
            float U = 1.0f - V - W;
            if (fabsf(U) > 1.0f || fabsf(V) > 1.0f || fabsf(W) > 1.0f) {
                r.status = EPA_BARY;
            } else if (!isfinite(U) || !isfinite(V) || !isfinite(W)) {
                r.status = EPA_NONFINITE;
            } else {
                r.pa = add(add(mul(ld3(L.va, a), U), mul(ld3(L.va, b), V)), mul(ld3(L.va, c), W));
                r.n = mul(N, -1.0f);
                r.pb = add(add(mul(ld3(L.vb, a), U), mul(ld3(L.vb, b), V)), mul(ld3(L.vb, c), W));
                r.depth = depth;
                r.status = EPA_CONTACT;
            }
            break;
        }
        if (nv >= EPA_MAXV) {
            r.status = EPA_OVERFLOW;
            break;
        }
        // append the vertex
        const int vnew = nv;
        int ncls = vertex_class(L, vnew, ns.P);
        if (lane == 0) {
            st3(L.vp, vnew, ns.P);
            st3(L.va, vnew, ns.A);
            st3(L.vb, vnew, ns.B);
            L.vcls[vnew] = ncls;
        }
        nv += 1;
        r.nverts = nv;
        // visibility (code/nans.cpp:869-891): the recomputed flipped normal is the stored one
        uint64_t vis[EPA_FSLOTS];
        int fidk[EPA_FSLOTS];
        float fnk[EPA_FSLOTS][3], fdk[EPA_FSLOTS];
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            int f = lane + 64 * k;
            bool vv = false;
            fidk[k] = 0;
            fdk[k] = 0.0f;
            fnk[k][0] = fnk[k][1] = fnk[k][2] = 0.0f;
            if (f < nf) {
                int id = L.fid[f];
                V3 n = ld3(L.fn, f);
                V3 A = ld3(L.vp, id & 255);
                vv = dot(n, sub(ns.P, A)) > 0.0f;
                fidk[k] = id;
                fdk[k] = L.fd[f];
                fnk[k][0] = n.x; fnk[k][1] = n.y; fnk[k][2] = n.z;
            }
            vis[k] = ballot(vv);
        }
        wave_sync();
        // horizon edges in face order (PushEdge), one edge per lane
        int ne = 0, ea = 0, eb = 0, eca = 0, ecb = 0;
        bool overflow = false;
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            uint64_t m = vis[k];
            while (m) {
                int f = 64 * k + ffs64(m);
                m &= m - 1;
                int id = L.fid[f];
                int fv[3] = {id & 255, (id >> 8) & 255, (id >> 16) & 255};
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    int A = fv[e], B = fv[(e + 1) % 3];
                    int cA = L.vcls[A], cB = L.vcls[B];
                    bool match = lane < ne && cls_eq(eca, cB) && cls_eq(ecb, cA);
                    uint64_t mm = ballot(match);
                    if (mm) {
                        int j = ffs64(mm);
                        int na = __shfl_down(ea, 1, 64), nb = __shfl_down(eb, 1, 64);
                        int nca = __shfl_down(eca, 1, 64), ncb = __shfl_down(ecb, 1, 64);
                        if (lane >= j) {
                            ea = na; eb = nb; eca = nca; ecb = ncb;
                        }
                        ne -= 1;
                    } else if (ne < 64) {
                        if (lane == ne) {
                            ea = A; eb = B; eca = cA; ecb = cB;
                        }
                        ne += 1;
                    } else {
                        overflow = true;
                    }
                }
            }
        }
        if (overflow) {
            r.status = EPA_OVERFLOW;
            break;
        }
        // ordered removal of the visible faces
        int nsurv = 0;
        int dst[EPA_FSLOTS];
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            int f = lane + 64 * k;
            bool alive = f < nf && !((vis[k] >> lane) & 1ull);
            uint64_t am = ballot(alive);
            dst[k] = alive ? nsurv + __popcll(am & ((1ull << lane) - 1ull)) : -1;
            nsurv += __popcll(am);
        }
        if (nsurv + ne > EPA_MAXF) {
            r.status = EPA_OVERFLOW;
            break;
        }
        wave_sync();
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            if (dst[k] >= 0) {
                L.fn[dst[k]][0] = fnk[k][0];
                L.fn[dst[k]][1] = fnk[k][1];
                L.fn[dst[k]][2] = fnk[k][2];
                L.fd[dst[k]] = fdk[k];
                L.fid[dst[k]] = fidk[k];
            }
        }
        wave_sync();
        // new faces (code/nans.cpp:894-899): PushTriangle(NewSupport, e.A, e.B)
        if (lane < ne) make_face(L, nsurv + lane, vnew, ea, eb);
        nf = nsurv + ne;
        if (nf > r.max_faces) r.max_faces = nf;
        wave_sync();
    }
    return r;
}

// ---------------------------------------------------------------------- block driver
// CheckCollision (code/nans.cpp:906-966) over the candidates [0, ncand) of one
// workgroup, in two phases:
//   GJK  every lane runs one candidate's EvolveSimplex loop; a lane whose loop
//        ends takes the next candidate at once (work is handed out per wave by
//        one LDS atomic), so a wave is not held by its slowest lane;
//   EPA  the GJK hits are queued with their simplex; each wave takes queued hits
//        one at a time (LDS atomic) and runs ResolveCollision on them with the
//        polytope in its LDS block.
// Results go through the source object, per candidate, so their order is that
// of the candidates whatever the order of the work.
struct NarrowShared {
    int next;      // next candidate to hand out
    int nepa;      // GJK hits queued
    int epa_next;  // next queued hit to hand out
};

// A queued GJK hit's simplex: the supports on A and B of its four vertices; the
// Minkowski point is recomputed as SupA - SupB, the expression that made it.
struct SimplexRec {
    float a[12], b[12];
};

__device__ __forceinline__ void simplex_rec_put(SimplexRec &q, const Vtx &v, int k)
{
    q.a[3 * k] = v.A.x; q.a[3 * k + 1] = v.A.y; q.a[3 * k + 2] = v.A.z;
    q.b[3 * k] = v.B.x; q.b[3 * k + 1] = v.B.y; q.b[3 * k + 2] = v.B.z;
}

// Source interface (duck-typed):
//   void shapes(int c, Shape &A, Shape &B) const       both shapes of candidate c
//   void gjk_result(int c, int gjk) const               every candidate
//   void epa_result(int c, const EpaResult &r) const    every GJK hit (lane 0 calls)
template <int BLOCK, class Src>
__device__ void narrow_block(const Src &src, int ncand, const NarrowParams &np, EpaLds *epa, NarrowShared &sh,
                             int *qidx, SimplexRec *qsim)
{
    const int lane = lane_id();
    if (threadIdx.x == 0) {
        sh.next = 0;
        sh.nepa = 0;
        sh.epa_next = 0;
    }
    __syncthreads();
    {
        Shape A, B;
        Simplex S;
        S.n = 0;
        int c = -1, calls = 0;
        bool retired = false;
        for (;;) {
            const bool need = c < 0 && !retired;
            const uint64_t m = ballot(need);
            if (m) {
                int base = 0;
                if (lane == 0) base = atomicAdd(&sh.next, __popcll(m));
                base = __builtin_amdgcn_readfirstlane(base);
                if (need) {
                    const int my = base + __popcll(m & ((1ull << lane) - 1ull));
                    if (my < ncand) {
                        c = my;
                        src.shapes(c, A, B);
                        S.n = 0;
                        calls = 0;
                    } else {
                        retired = true;
                    }
                }
            }
            if (!ballot(c >= 0)) break;
            if (c >= 0) {
                const int res = evolve_simplex(A, B, S, np.fp64_islands);
                ++calls;
                if (res != GJK_STILL || calls > np.max_gjk_iterations) {
                    const int g = res == GJK_STILL ? GJK_CAP : res;
                    src.gjk_result(c, g);
                    if (g == GJK_FOUND) {
                        const int q = atomicAdd(&sh.nepa, 1);
                        qidx[q] = c;
                        simplex_rec_put(qsim[q], S.s0, 0);
                        simplex_rec_put(qsim[q], S.s1, 1);
                        simplex_rec_put(qsim[q], S.s2, 2);
                        simplex_rec_put(qsim[q], S.s3, 3);
                    }
                    c = -1;
                }
            }
        }
    }
    __syncthreads();
    const int nepa = sh.nepa;
    EpaLds &L = epa[threadIdx.x >> 6];
    for (;;) {
        int job = 0;
        if (lane == 0) job = atomicAdd(&sh.epa_next, 1);
        job = __builtin_amdgcn_readfirstlane(job);
        if (job >= nepa) break;
        const int c = qidx[job];
        Shape A, B;
        src.shapes(c, A, B);
        if (lane == 0) {
            shape_store(L.sa, A);
            shape_store(L.sb, B);
        }
        if (lane < 4) {
            const SimplexRec &q = qsim[job];
            V3 a = mk3(q.a[3 * lane], q.a[3 * lane + 1], q.a[3 * lane + 2]);
            V3 b = mk3(q.b[3 * lane], q.b[3 * lane + 1], q.b[3 * lane + 2]);
            st3(L.va, lane, a);
            st3(L.vb, lane, b);
            st3(L.vp, lane, sub(a, b));
        }
        wave_sync();
        const EpaResult r = epa_wave(L, np);
        if (lane == 0) src.epa_result(c, r);
        wave_sync();
    }
    __syncthreads();
}

}  // namespace nans
