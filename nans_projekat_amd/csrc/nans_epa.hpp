// nans_epa.hpp -- EPA (ResolveCollision, code/nans.cpp:787-904) by one wavefront
// per pair, and the workgroup narrowphase driver (GJK per lane, EPA per wave).
//
// Layout of the polytope:
//   vertices   registers: lane l holds vertex l (and lane l < 8 vertex 64 + l): its
//              P and its equality class (the first vertex with the same P, 255 if
//              P has a NaN: never equal); LDS keeps P, SupA, SupB for gathers;
//   faces      registers: lane l owns faces l, l+64, l+128 and keeps for each its
//              stored normal N, first vertex A, distance |dot(n, A)| and the packed
//              ids and classes of its three vertices;
//   edges      registers: horizon edge e lives in lane e;
//   shapes     registers: lanes 0-7 hold corner (lane) of box A, lanes 8-15 corner
//              (lane - 8) of box B; sphere centres and radii in every lane.
// One iteration of the reference's loop (code/nans.cpp:804-902):
//   closest face   the first face holding the minimum of the distances below face
//                  0's (the strict '<' scan from face 0, code/nans.cpp:807-822):
//                  DPP row minimum + ballots, the winner's data by readlane;
//   new support    along the winner's stored normal (code/nans.cpp:824-825): one
//                  corner dot product per lane and a DPP argmax over each group of
//                  8 lanes with the first-index tie rule of the strict '>' scan;
//   visibility     each lane tests its faces against the new point, ballots;
//   horizon        the visible faces' edges in face order with the exact
//                  reversed-edge cancellation of PushEdge (code/nans.cpp:232-266):
//                  when no edge key occurs twice, an edge survives iff its reverse
//                  is not among the pushed edges, which every lane checks at once;
//                  otherwise (or beyond 64 pushes) the pushes run one at a time;
//   compaction     survivors keep their order (std::vector::erase), new faces
//                  follow in edge order: one pass through an LDS staging area.
// A face's distance and flipped normal never change once it is made, so they are
// computed once, with the expression the reference re-evaluates every iteration.
#pragma once

#include "nans_narrow.hpp"

namespace nans {

constexpr int EPA_MAXV = 72;   // 4 + 65 vertices at most (one per iteration)
constexpr int EPA_MAXF = 192;  // alive faces: a closed triangulation of V <= 69 points has 2V-4 <= 134
constexpr int EPA_FSLOTS = EPA_MAXF / 64;
constexpr int CLS_NAN = 255;

struct EpaLds {
    float vp[EPA_MAXV][3];
    float va[EPA_MAXV][3];
    float vb[EPA_MAXV][3];
    float sfn[EPA_MAXF][3];  // compaction staging: stored normal
    float sfd[EPA_MAXF];     //                     distance
    int sfid[EPA_MAXF];      //                     a | b << 8 | c << 16
    int sfcl[EPA_MAXF];      //                     classes of a, b, c
    int ekey[128];           // horizon: pushed edges' classes (0..63), survivors (64..127)
    int eids[128];           //          and vertex ids, a | b << 8
};

struct EpaResult {
    int status;  // EPA_*
    V3 pa, pb, n;
    double depth;
    int iters, max_faces, nverts;
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// The threads that run one world's step or one narrowphase: the whole workgroup,
// or one wave of it.  whole must be uniform over the workgroup (made from
// constants or readfirstlane values), so that sync()'s barrier is reached by
// every wave or by none.  A wave team's sync orders its global and LDS accesses
// at workgroup scope (the memory it shares is the workgroup's), with no barrier.
struct Team {
    int rank;    // the thread's index in the team
    int size;    // threads in the team: the workgroup's, or 64
    bool whole;  // the whole workgroup

    __device__ static Team block(int threads) { return Team{(int)threadIdx.x, threads, true}; }
    __device__ static Team wave() { return Team{(int)(threadIdx.x & 63), 64, false}; }
    __device__ int waves() const { return size >> 6; }
    __device__ void sync() const
    {
        if (whole) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
    }
    // sync() that also returns the number of threads with p set
    __device__ int sync_count(bool p) const
    {
        if (whole) return __syncthreads_count(p);
        const int c = __popcll(__ballot(p));
        sync();
        return c;
    }
};
__device__ __forceinline__ int ffs64(uint64_t m) { return __builtin_ctzll(m); }

__device__ __forceinline__ V3 ld3(const float (*a)[3], int i) { return mk3(a[i][0], a[i][1], a[i][2]); }
__device__ __forceinline__ void st3(float (*a)[3], int i, V3 v)
{
    a[i][0] = v.x;
    a[i][1] = v.y;
    a[i][2] = v.z;
}

// DPP move of a 32-bit value (GFX9 encodings: quad_perm 0x00-0xff, row_mirror
// 0x140, row_half_mirror 0x141; gfx90a and later: row_newbcast 0x150-0x15f,
// lane n of each 16-lane row to the whole row); every source lane of these
// patterns exists.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

__device__ __forceinline__ float readlane_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

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


template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}

// The two shapes of a pair as one wave holds them for EPA.
struct EpaShapes {
    int kind_a, kind_b;  // uniform
    V3 ca, cb;           // centres (spheres)
    float ra, rb;        // radii (spheres)
    V3 k;                // lanes 0-7: corner (lane) of A; lanes 8-15: corner (lane - 8) of B
};

__device__ inline V3 pick_corner(const Shape &s, int k)
{
    V3 v = s.v[0];
#pragma unroll
    for (int j = 1; j < 8; ++j) {
        const bool t = k == j;
        v.x = t ? s.v[j].x : v.x;
        v.y = t ? s.v[j].y : v.y;
        v.z = t ? s.v[j].z : v.z;
    }
    return v;
}

__device__ inline EpaShapes epa_shapes(const Shape &A, const Shape &B)
{
    const int lane = lane_id();
    EpaShapes e;
    e.kind_a = A.kind;
    e.kind_b = B.kind;
    e.ca = A.c;
    e.cb = B.c;
    e.ra = A.r;
    e.rb = B.r;
    const V3 ka = pick_corner(A, lane & 7), kb = pick_corner(B, lane & 7);
    const bool onb = lane >= 8;
    e.k = mk3(onb ? kb.x : ka.x, onb ? kb.y : ka.y, onb ? kb.z : ka.z);
    return e;
}

// Supports of the boxes held as corners in lanes 0-7 (A, along d) and 8-15 (B,
// along -d), both at once: in each group of 8 lanes, the first corner of greatest
// dot(v, dir) among those above -FLT_MAX, else (0,0,0) -- GetCubeSupport's strict
// '>' scan from -FLT_MAX (code/nans.cpp:409-430).
__device__ __forceinline__ void box_supports(V3 kv, V3 d, V3 nd, V3 &sa, V3 &sb)
{
    const int lane = lane_id();
    const float dist = dot(kv, lane >= 8 ? nd : d);
    float key = dist > -FLT_MAX ? dist : -__builtin_inff();
    int idx = lane & 7;
    auto step = [&](float ok, int oi) {
        const bool take = ok > key || (ok == key && oi < idx);
        key = take ? ok : key;
        idx = take ? oi : idx;
    };
    step(dpp_f<0xB1>(key), dpp_i<0xB1>(idx));    // quad_perm [1,0,3,2]
    step(dpp_f<0x4E>(key), dpp_i<0x4E>(idx));    // quad_perm [2,3,0,1]
    step(dpp_f<0x141>(key), dpp_i<0x141>(idx));  // row_half_mirror: the other quad of the 8
    const float ka = readlane_f(key, 0), kb = readlane_f(key, 8);
    const int wa = readlane_i(idx, 0), wb = 8 + readlane_i(idx, 8);
    sa = ka == -__builtin_inff() ? mk3(0.0f, 0.0f, 0.0f)
                                 : mk3(readlane_f(kv.x, wa), readlane_f(kv.y, wa), readlane_f(kv.z, wa));
    sb = kb == -__builtin_inff() ? mk3(0.0f, 0.0f, 0.0f)
                                 : mk3(readlane_f(kv.x, wb), readlane_f(kv.y, wb), readlane_f(kv.z, wb));
}

// CalculateSupport (code/nans.cpp:463-519): S_A(d) and S_B(-d); every lane calls.
__device__ inline Vtx wave_support(const EpaShapes &s, V3 d)
{
    const V3 nd = mul(d, -1.0f);
    Vtx v;
    V3 ga = mk3(0.0f, 0.0f, 0.0f), gb = mk3(0.0f, 0.0f, 0.0f);
    if (s.kind_a == SHAPE_BOX || s.kind_b == SHAPE_BOX) box_supports(s.k, d, nd, ga, gb);
    v.A = s.kind_a == SHAPE_BOX ? ga : sphere_support(s.ca, s.ra, d);
    v.B = s.kind_b == SHAPE_BOX ? gb : sphere_support(s.cb, s.rb, nd);
    v.P = sub(v.A, v.B);
    return v;
}

// Position of the n-th set bit of m (n < popcount(m)).
__device__ __forceinline__ int nth_set_bit(uint64_t m, int n)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint64_t low = m & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        if (n >= c) {
            n -= c;
            pos += w;
            m >>= w;
        } else {
            m = low;
        }
    }
    return pos;
}

// Equality class of a new vertex with id nv and point P against the vertices in
// registers (vertex == compares P, code/nans.h:251-254): the first earlier vertex
// with the same P, CLS_NAN if P has a NaN.  Every lane calls.
__device__ __forceinline__ int vertex_class_reg(V3 v0, V3 v1, int nv, V3 P)
{
    if (has_nan(P)) return CLS_NAN;
    const int lane = lane_id();
    const uint64_t m0 = ballot(lane < nv && veq(v0, P));
    if (m0) return ffs64(m0);
    const uint64_t m1 = ballot(lane + 64 < nv && veq(v1, P));
    return m1 ? 64 + ffs64(m1) : nv;
}

__device__ __forceinline__ float pick3f(int k, float a, float b, float c) { return k == 0 ? a : (k == 1 ? b : c); }
__device__ __forceinline__ int pick3i(int k, int a, int b, int c) { return k == 0 ? a : (k == 1 ? b : c); }

// A face in registers.
struct FaceR {
    float nx, ny, nz;  // stored normal (flipped away from the origin)
    float ax, ay, az;  // first vertex A
    float d;           // |dot(unflipped normal, A)| with glm's abs
    int id;            // a | b << 8 | c << 16
    int cl;            // classes of a, b, c
};

// PushTriangle (code/nans.cpp:291-322) of vertices (A, B, C) = (a, b, c).
__device__ __forceinline__ FaceR make_face(V3 A, V3 B, V3 C, int a, int b, int c, int ca, int cb, int cc)
{
    V3 n = normalize(cross(sub(B, A), sub(C, A)));
    FaceR f;
    f.d = gabs(dot(n, A));
    if (dot(A, n) < 0.0f) n = mul(n, -1.0f);
    f.nx = n.x; f.ny = n.y; f.nz = n.z;
    f.ax = A.x; f.ay = A.y; f.az = A.z;
    f.id = a | (b << 8) | (c << 16);
    f.cl = ca | (cb << 8) | (cc << 16);
    return f;
}

// Binary64 distance of a face's plane to the origin (FP64-islands final choice).
__device__ inline double face_dist64(const EpaLds &L, const FaceR &f)
{
    int b = (f.id >> 8) & 255, c = (f.id >> 16) & 255;
    double ax = f.ax, ay = f.ay, az = f.az;
    double ux = (double)L.vp[b][0] - ax, uy = (double)L.vp[b][1] - ay, uz = (double)L.vp[b][2] - az;
    double vx = (double)L.vp[c][0] - ax, vy = (double)L.vp[c][1] - ay, vz = (double)L.vp[c][2] - az;
    double nx = uy * vz - vy * uz, ny = uz * vx - vz * ux, nz = ux * vy - vx * uy;
    double inv = 1.0 / __builtin_sqrt((nx * nx + ny * ny) + nz * nz);
    double d = ((nx * inv) * ax + (ny * inv) * ay) + (nz * inv) * az;
    return d >= 0.0 ? d : -d;
}

// ResolveCollision (code/nans.cpp:787-904) executed by one full wavefront.  The
// caller has written the GJK simplex to L.vp / L.va / L.vb[0..3], passes vertex
// (lane)'s P in p0 for lanes 0-3, and has called wave_sync().  Every lane returns
// the same result.
__device__ inline EpaResult epa_wave(EpaLds &L, const EpaShapes &sh, V3 p0, const NarrowParams &p)
{
    const int lane = lane_id();
    EpaResult r;
    r.status = EPA_ITER;
    r.pa = r.pb = r.n = mk3(0.0f, 0.0f, 0.0f);
    r.depth = 0.0;
    r.iters = 0;
    r.max_faces = 4;
    r.nverts = 4;

    // vertices in registers: v0 = vertex (lane), v1 = vertex (64 + lane)
    V3 v0 = lane < 4 ? p0 : mk3(0.0f, 0.0f, 0.0f);
    V3 v1 = mk3(0.0f, 0.0f, 0.0f);
    int c0 = 0;
    {
        // classes of the simplex vertices: the first earlier vertex with the same P
        int cls = lane;
        bool found = false;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const V3 pu = mk3(readlane_f(v0.x, u), readlane_f(v0.y, u), readlane_f(v0.z, u));
            if (!found && u < lane && veq(pu, v0)) {
                cls = u;
                found = true;
            }
        }
        c0 = has_nan(v0) ? CLS_NAN : cls;
    }
    FaceR F[EPA_FSLOTS];
#pragma unroll
    for (int k = 0; k < EPA_FSLOTS; ++k) F[k] = FaceR{0, 0, 0, 0, 0, 0, 0, 0, 0};
    {
        // ABC, ACD, ADB, BDC (code/nans.cpp:799-802)
        const int a = lane == 3 ? 1 : 0;
        const int b = lane == 0 ? 1 : (lane == 1 ? 2 : 3);
        const int c = lane == 0 ? 2 : (lane == 1 ? 3 : (lane == 2 ? 1 : 2));
        const int ca = __shfl(c0, a, 64), cb = __shfl(c0, b, 64), cc = __shfl(c0, c, 64);
        if (lane < 4) F[0] = make_face(ld3(L.vp, a), ld3(L.vp, b), ld3(L.vp, c), a, b, c, ca, cb, cc);
    }
    int nf = 4, nv = 4;
    int iter = 0;
    while (iter++ <= p.max_epa_iterations) {
        r.iters += 1;
        if (nf == 0) {
            r.status = EPA_EMPTY;
            break;
        }
        // closest face
        const float d0 = readlane_f(gabs(dot(mk3(F[0].nx, F[0].ny, F[0].nz), mk3(F[0].ax, F[0].ay, F[0].az))), 0);
        float key[EPA_FSLOTS];
        float kmin = __builtin_inff();
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            const int f = lane + 64 * k;
            key[k] = (f < nf && F[k].d < d0) ? F[k].d : __builtin_inff();
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
        int owner = best & 63, ks = best >> 6;
        V3 N = mk3(readlane_f(pick3f(ks, F[0].nx, F[1].nx, F[2].nx), owner),
                   readlane_f(pick3f(ks, F[0].ny, F[1].ny, F[2].ny), owner),
                   readlane_f(pick3f(ks, F[0].nz, F[1].nz, F[2].nz), owner));
        float cur = m == __builtin_inff() ? d0 : readlane_f(pick3f(ks, F[0].d, F[1].d, F[2].d), owner);
        // new support along the stored normal
        const Vtx ns = wave_support(sh, N);
        if (dot(N, ns.P) - cur < p.epa_tolerance) {
            double depth = (double)cur;
            if (p.fp64_islands) {
                wave_sync();
                double d64[EPA_FSLOTS];
#pragma unroll
                for (int k = 0; k < EPA_FSLOTS; ++k) d64[k] = (lane + 64 * k < nf) ? face_dist64(L, F[k]) : 0.0;
                const double f0 = __shfl(d64[0], 0, 64);
                bool v64 = false;
                double bd = 0.0;
                int bi64 = 0x7fffffff;
#pragma unroll
                for (int k = 0; k < EPA_FSLOTS; ++k) {
                    const int f = lane + 64 * k;
                    if (f < nf && d64[k] < f0 && (!v64 || d64[k] < bd)) {
                        v64 = true;
                        bd = d64[k];
                        bi64 = f;
                    }
                }
                wave_argmin(v64, bd, bi64);
                const int sel = v64 ? bi64 : 0;
                depth = v64 ? bd : f0;
                if (sel != best) {
                    best = sel;
                    owner = best & 63;
                    ks = best >> 6;
                    N = mk3(readlane_f(pick3f(ks, F[0].nx, F[1].nx, F[2].nx), owner),
                            readlane_f(pick3f(ks, F[0].ny, F[1].ny, F[2].ny), owner),
                            readlane_f(pick3f(ks, F[0].nz, F[1].nz, F[2].nz), owner));
                    cur = sel == 0 ? d0 : readlane_f(pick3f(ks, F[0].d, F[1].d, F[2].d), owner);
                }
            }
            const int id = readlane_i(pick3i(ks, F[0].id, F[1].id, F[2].id), owner);
            const int a = id & 255, b = (id >> 8) & 255, c = (id >> 16) & 255;
            wave_sync();
            V3 PA = ld3(L.vp, a), PB = ld3(L.vp, b), PC = ld3(L.vp, c);
            // Barycentric (code/nans.cpp:771-785) of N * d
            V3 P = mul(N, cur);
            V3 e0 = sub(PB, PA), e1 = sub(PC, PA), e2 = sub(P, PA);
            float d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1);
            float d20 = dot(e2, e0), d21 = dot(e2, e1);
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
        const int ncls = vertex_class_reg(v0, v1, vnew, ns.P);
        if (lane == (vnew & 63)) {
            if (vnew < 64) v0 = ns.P;
            else v1 = ns.P;
        }
        if (lane == 0) {
            st3(L.vp, vnew, ns.P);
            st3(L.va, vnew, ns.A);
            st3(L.vb, vnew, ns.B);
        }
        nv += 1;
        r.nverts = nv;
        // visibility (code/nans.cpp:869-891): the recomputed flipped normal is the stored one
        uint64_t vis[EPA_FSLOTS];
        int nvis = 0;
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            const int f = lane + 64 * k;
            const bool vv = f < nf && dot(mk3(F[k].nx, F[k].ny, F[k].nz),
                                          sub(ns.P, mk3(F[k].ax, F[k].ay, F[k].az))) > 0.0f;
            vis[k] = ballot(vv);
            nvis += __popcll(vis[k]);
        }
        // horizon edges in face order (PushEdge), edge e in lane e
        int ne = 0, ea = 0, eb = 0, eca = 0, ecb = 0;
        // fast path (all faces in one slot, at most 64 pushes): an edge survives iff
        // its reverse is not among the pushed edges, provided no key occurs twice
        bool sequential = nf > 64 || 3 * nvis > 64;
        if (!sequential) {
            const bool mine = (vis[0] >> lane) & 1ull;
            const int cl = F[0].cl;
            int key[3], rev[3], nsame[3] = {0, 0, 0}, nrev[3] = {0, 0, 0};
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const int ca = (cl >> (8 * e)) & 255, cb = (cl >> (8 * ((e + 1) % 3))) & 255;
                key[e] = ca | (cb << 8);
                rev[e] = (ca == CLS_NAN || cb == CLS_NAN) ? -1 : (cb | (ca << 8));
            }
            uint64_t mk = vis[0];
            while (mk) {
                const int o = ffs64(mk);
                mk &= mk - 1;
                const int co = readlane_i(cl, o);
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int kt = ((co >> (8 * t)) & 255) | (((co >> (8 * ((t + 1) % 3))) & 255) << 8);
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        nsame[e] += kt == key[e];
                        nrev[e] += kt == rev[e];
                    }
                }
            }
            bool dup = false;
#pragma unroll
            for (int e = 0; e < 3; ++e) dup |= rev[e] >= 0 && (nsame[e] > 1 || nrev[e] > 1);
            sequential = ballot(mine && dup) != 0ull;
            if (!sequential) {
                // kept edges of this lane's face, placed after those of earlier visible faces
                int keep = 0;
#pragma unroll
                for (int e = 0; e < 3; ++e) keep |= (rev[e] < 0 || nrev[e] == 0) ? (1 << e) : 0;
                const int cnt = mine ? __popc(keep) : 0;
                const uint64_t lower = (1ull << lane) - 1ull;
                const uint64_t b0 = ballot(cnt & 1), b1 = ballot(cnt & 2);
                int o = __popcll(b0 & lower) + 2 * __popcll(b1 & lower);
                ne = __popcll(b0) + 2 * __popcll(b1);
                if (mine) {
                    const int id = F[0].id;
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        if ((keep >> e) & 1) {
                            const int s0 = 8 * e, s1 = 8 * ((e + 1) % 3);
                            L.eids[o] = ((id >> s0) & 255) | (((id >> s1) & 255) << 8);
                            L.ekey[o] = key[e];
                            ++o;
                        }
                    }
                }
                wave_sync();
                if (lane < ne) {
                    const int k2 = L.ekey[lane], i2 = L.eids[lane];
                    ea = i2 & 255;
                    eb = (i2 >> 8) & 255;
                    eca = k2 & 255;
                    ecb = (k2 >> 8) & 255;
                }
                wave_sync();
            }
        }
        bool overflow = false;
        if (sequential) {
            ne = 0;
#pragma unroll
            for (int k = 0; k < EPA_FSLOTS; ++k) {
                uint64_t mk = vis[k];
                while (mk) {
                    const int o = ffs64(mk);
                    mk &= mk - 1;
                    const int id = readlane_i(F[k].id, o), cl = readlane_i(F[k].cl, o);
                    const int fv[3] = {id & 255, (id >> 8) & 255, (id >> 16) & 255};
                    const int fc[3] = {cl & 255, (cl >> 8) & 255, (cl >> 16) & 255};
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const int A = fv[e], B = fv[(e + 1) % 3];
                        const int cA = fc[e], cB = fc[(e + 1) % 3];
                        const bool match = lane < ne && eca != CLS_NAN && eca == cB && ecb != CLS_NAN && ecb == cA;
                        const uint64_t mm = ballot(match);
                        if (mm) {
                            const int j = ffs64(mm);
                            const int na = __shfl_down(ea, 1, 64), nb = __shfl_down(eb, 1, 64);
                            const int nca = __shfl_down(eca, 1, 64), ncb = __shfl_down(ecb, 1, 64);
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
        }
        if (overflow) {
            r.status = EPA_OVERFLOW;
            break;
        }
        // ordered removal of the visible faces, then the new ones
        int nsurv = 0;
        int dst[EPA_FSLOTS];
        uint64_t am0 = 0ull;
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            const int f = lane + 64 * k;
            const bool alive = f < nf && !((vis[k] >> lane) & 1ull);
            const uint64_t am = ballot(alive);
            if (k == 0) am0 = am;
            dst[k] = alive ? nsurv + __popcll(am & ((1ull << lane) - 1ull)) : -1;
            nsurv += __popcll(am);
        }
        if (nsurv + ne > EPA_MAXF) {
            r.status = EPA_OVERFLOW;
            break;
        }
        // PushTriangle(NewSupport, e.A, e.B) (code/nans.cpp:894-899) in lane e
        FaceR g = FaceR{0, 0, 0, 0, 0, 0, 0, 0, 0};
        wave_sync();  // lane 0's vertex stores before other lanes read L.vp
        if (lane < ne) g = make_face(ns.P, ld3(L.vp, ea), ld3(L.vp, eb), vnew, ea, eb, ncls, eca, ecb);
        const int nfn = nsurv + ne;
        if (nf <= 64 && nfn <= 64) {
            // lane f takes survivor f from its lane, or new face f - nsurv from its edge lane
            const bool old = lane < nsurv;
            const int src = old ? nth_set_bit(am0, lane) : lane - nsurv;
            const int adr = 4 * (src & 63);
            auto pull = [&](float a_, float b_) {
                const float x = __int_as_float(__builtin_amdgcn_ds_bpermute(adr, __float_as_int(a_)));
                const float y = __int_as_float(__builtin_amdgcn_ds_bpermute(adr, __float_as_int(b_)));
                return old ? x : y;
            };
            auto pulli = [&](int a_, int b_) {
                const int x = __builtin_amdgcn_ds_bpermute(adr, a_);
                const int y = __builtin_amdgcn_ds_bpermute(adr, b_);
                return old ? x : y;
            };
            FaceR h;
            h.nx = pull(F[0].nx, g.nx);
            h.ny = pull(F[0].ny, g.ny);
            h.nz = pull(F[0].nz, g.nz);
            h.ax = pull(F[0].ax, g.ax);
            h.ay = pull(F[0].ay, g.ay);
            h.az = pull(F[0].az, g.az);
            h.d = pull(F[0].d, g.d);
            h.id = pulli(F[0].id, g.id);
            h.cl = pulli(F[0].cl, g.cl);
            F[0] = h;
            nf = nfn;
            if (nf > r.max_faces) r.max_faces = nf;
            continue;
        }
        // otherwise through the LDS staging area
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            if (dst[k] >= 0) {
                L.sfn[dst[k]][0] = F[k].nx;
                L.sfn[dst[k]][1] = F[k].ny;
                L.sfn[dst[k]][2] = F[k].nz;
                L.sfd[dst[k]] = F[k].d;
                L.sfid[dst[k]] = F[k].id;
                L.sfcl[dst[k]] = F[k].cl;
            }
        }
        if (lane < ne) {
            const int s = nsurv + lane;
            L.sfn[s][0] = g.nx;
            L.sfn[s][1] = g.ny;
            L.sfn[s][2] = g.nz;
            L.sfd[s] = g.d;
            L.sfid[s] = g.id;
            L.sfcl[s] = g.cl;
        }
        nf = nfn;
        if (nf > r.max_faces) r.max_faces = nf;
        wave_sync();
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            const int f = lane + 64 * k;
            if (f < nf) {
                F[k].nx = L.sfn[f][0];
                F[k].ny = L.sfn[f][1];
                F[k].nz = L.sfn[f][2];
                F[k].d = L.sfd[f];
                F[k].id = L.sfid[f];
                F[k].cl = L.sfcl[f];
                const int a = F[k].id & 255;
                F[k].ax = L.vp[a][0];
                F[k].ay = L.vp[a][1];
                F[k].az = L.vp[a][2];
            }
        }
        wave_sync();
    }
    return r;
}

// ---------------------------------------------------------------------- four EPAs per wave
// ResolveCollision for 64/G pairs at once, one per G-lane group (G = 16: a row
// of the wave; G = 32: half a wave): the same steps as epa_wave, with the
// polytope of each pair held by its group -- face f and vertex v in lane f / v
// (P, SupA, SupB; P also in LDS for gathers) -- group reductions by row-local DPP
// (and one ds_bpermute across the two rows of a 32-lane group) and group
// broadcasts by ds_bpermute.  A polytope that outgrows its group (G+1 vertices,
// G+1 faces, G+1 horizon edges, or an edge key pushed twice) stops with EPA_SPILL
// and is redone from its simplex by the next wider tier (16 -> 32 -> epa_wave),
// so the result is always epa_wave's.
constexpr int EPA_SPILL = 99;

template <int G>
struct EpgLds {
    static_assert(G == 16 || G == 32, "groups of 16 or 32 lanes");
    float vp[64 / G][G][3];   // vertex P per group
    float sf[64 / G][G][9];   // compaction staging: n, a, d, id, cl
    int ekey[64];             // horizon edges per group (G each)
    int eids[64];
    // horizon: how often the visible faces push each directed edge (ca, cb),
    // one byte per key ca * G + cb, four keys per word
    uint32_t ecnt[64 / G][G * G / 4];
};
static_assert(sizeof(EpgLds<16>) <= sizeof(EpaLds), "the grouped EPA reuses the wave's EPA LDS");
static_assert(sizeof(EpgLds<32>) <= sizeof(EpaLds), "the grouped EPA reuses the wave's EPA LDS");
static_assert(16 * 4 * 64 <= sizeof(EpaLds), "GJK keeps 16 words per lane in the EPA LDS");

__device__ __forceinline__ float gpull_f(float v, int src_lane)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_lane << 2, __float_as_int(v)));
}
__device__ __forceinline__ int gpull_i(int v, int src_lane) { return __builtin_amdgcn_ds_bpermute(src_lane << 2, v); }
__device__ __forceinline__ V3 gpull3(V3 v, int src_lane)
{
    return mk3(gpull_f(v.x, src_lane), gpull_f(v.y, src_lane), gpull_f(v.z, src_lane));
}
// lane U of this lane's G-lane group, in every lane (G = 16: one row broadcast,
// no LDS round trip); the whole wave must be active
template <int G, int U>
__device__ __forceinline__ float gbcast_f(float v)
{
    if constexpr (G == 16) return dpp_f<0x150 + U>(v);
    else return gpull_f(v, (lane_id() & (64 - G)) + U);
}
template <int G, int U>
__device__ __forceinline__ int gbcast_i(int v)
{
    if constexpr (G == 16) return dpp_i<0x150 + U>(v);
    else return gpull_i(v, (lane_id() & (64 - G)) + U);
}
template <int G, int U>
__device__ __forceinline__ V3 gbcast3(V3 v)
{
    return mk3(gbcast_f<G, U>(v.x), gbcast_f<G, U>(v.y), gbcast_f<G, U>(v.z));
}

// this lane's G-lane group bits of a wave ballot
template <int G>
__device__ __forceinline__ uint32_t gbits(uint64_t m)
{
    return (uint32_t)(m >> (lane_id() & (64 - G))) & (uint32_t)((1ull << G) - 1ull);
}
// minimum over the lane's group of G (no NaN inputs), in every lane of the group:
// a row of 16 by DPP, then the two rows of a 32-lane group through ds_bpermute
template <int G>
__device__ __forceinline__ float group_min(float v)
{
    v = fminf(v, dpp_f<0xB1>(v));
    v = fminf(v, dpp_f<0x4E>(v));
    v = fminf(v, dpp_f<0x141>(v));
    v = fminf(v, dpp_f<0x140>(v));
    if (G == 32) v = fminf(v, gpull_f(v, lane_id() ^ 16));
    return v;
}

// The group's shapes: lanes 0-7 of each row hold the corners of A, 8-15 of B.
// MEM: the shapes carry only Shape::vm (corners in memory, null for a sphere),
// and each lane reads its one corner from there.
template <bool MEM = false>
__device__ inline EpaShapes epa_shapes_group(const Shape &A, const Shape &B)
{
    const int j = lane_id() & 15;
    EpaShapes e;
    e.kind_a = A.kind;
    e.kind_b = B.kind;
    e.ca = A.c;
    e.cb = B.c;
    e.ra = A.r;
    e.rb = B.r;
    const bool onb = j >= 8;
    if (MEM) {
        const float *vm = onb ? B.vm : A.vm;
        e.k = vm ? mk3(vm[3 * (j & 7)], vm[3 * (j & 7) + 1], vm[3 * (j & 7) + 2]) : mk3(0.0f, 0.0f, 0.0f);
    } else {
        const V3 ka = pick_corner(A, j & 7), kb = pick_corner(B, j & 7);
        e.k = mk3(onb ? kb.x : ka.x, onb ? kb.y : ka.y, onb ? kb.z : ka.z);
    }
    return e;
}

// CalculateSupport for each group's pair along its d (group-uniform).  Every row
// of 16 lanes holds the corners of A (lanes 0-7) and B (8-15); a butterfly of DPP
// moves inside each 8-lane half carries (dot, corner number, corner) to the first
// corner of greatest dot product -- GetCubeSupport's strict '>' scan in corner
// order -- and a row mirror hands each half the other half's winner.  No
// cross-lane traffic leaves the row, so a 32-lane group's two rows agree.
template <int G>
__device__ inline Vtx group_support(const EpaShapes &s, V3 d)
{
    const int lane = lane_id(), j = lane & 15;
    const V3 nd = mul(d, -1.0f);
    const float dist = dot(s.k, j >= 8 ? nd : d);
    float key = dist > -FLT_MAX ? dist : -__builtin_inff();
    int idx = j & 7;
    V3 cw = s.k;
    auto step = [&](float ok, int oi, float ox, float oy, float oz) {
        const bool take = ok > key || (ok == key && oi < idx);
        key = take ? ok : key;
        idx = take ? oi : idx;
        cw.x = take ? ox : cw.x;
        cw.y = take ? oy : cw.y;
        cw.z = take ? oz : cw.z;
    };
    step(dpp_f<0xB1>(key), dpp_i<0xB1>(idx), dpp_f<0xB1>(cw.x), dpp_f<0xB1>(cw.y), dpp_f<0xB1>(cw.z));
    step(dpp_f<0x4E>(key), dpp_i<0x4E>(idx), dpp_f<0x4E>(cw.x), dpp_f<0x4E>(cw.y), dpp_f<0x4E>(cw.z));
    step(dpp_f<0x141>(key), dpp_i<0x141>(idx), dpp_f<0x141>(cw.x), dpp_f<0x141>(cw.y), dpp_f<0x141>(cw.z));
    // this half's winner and, through the row mirror, the other half's
    const float ko = dpp_f<0x140>(key);
    const V3 co = mk3(dpp_f<0x140>(cw.x), dpp_f<0x140>(cw.y), dpp_f<0x140>(cw.z));
    const bool onb = j >= 8;
    const float ka = onb ? ko : key, kb = onb ? key : ko;
    const V3 ca = mk3(onb ? co.x : cw.x, onb ? co.y : cw.y, onb ? co.z : cw.z);
    const V3 cb = mk3(onb ? cw.x : co.x, onb ? cw.y : co.y, onb ? cw.z : co.z);
    Vtx v;
    v.A = s.kind_a == SHAPE_BOX ? (ka == -__builtin_inff() ? mk3(0.0f, 0.0f, 0.0f) : ca)
                                : sphere_support(s.ca, s.ra, d);
    v.B = s.kind_b == SHAPE_BOX ? (kb == -__builtin_inff() ? mk3(0.0f, 0.0f, 0.0f) : cb)
                                : sphere_support(s.cb, s.rb, nd);
    v.P = sub(v.A, v.B);
    return v;
}

// classes (ca | cb << 8) of edge e of a face with packed classes cl
__device__ __forceinline__ int edge_key(int cl, int e)
{
    return ((cl >> (8 * e)) & 255) | (((cl >> (8 * ((e + 1) % 3))) & 255) << 8);
}

// Every lane calls; lanes 0-3 of a group pass the supports of simplex vertex j in
// sa0 / sb0.  act: the group has a pair.  Returns the group's result (EPA_NONE for
// an idle group, EPA_SPILL to redo with the next wider tier).
// Binary64 distance of a group face's plane to the origin (FP64-islands final
// choice), its vertices b and c from the group's vertex LDS.
__device__ inline double group_face_dist64(const float (*vp)[3], const FaceR &f)
{
    const int b = (f.id >> 8) & 255, c = (f.id >> 16) & 255;
    const double ax = f.ax, ay = f.ay, az = f.az;
    const double ux = (double)vp[b][0] - ax, uy = (double)vp[b][1] - ay, uz = (double)vp[b][2] - az;
    const double vx = (double)vp[c][0] - ax, vy = (double)vp[c][1] - ay, vz = (double)vp[c][2] - az;
    const double nx = uy * vz - vy * uz, ny = uz * vx - vz * ux, nz = ux * vy - vx * uy;
    const double inv = 1.0 / __builtin_sqrt((nx * nx + ny * ny) + nz * nz);
    const double d = ((nx * inv) * ax + (ny * inv) * ay) + (nz * inv) * az;
    return d >= 0.0 ? d : -d;
}

// Argmin over (value, index) in each G-lane group, ties to the lowest index;
// every lane of the group receives the result.
template <int G>
__device__ inline void group_argmin64(bool &valid, double &v, int &i)
{
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(i, off, 64);
        const int ok = __shfl_xor(valid ? 1 : 0, off, 64);
        const bool take = ok && (!valid || ov < v || (ov == v && oi < i));
        v = take ? ov : v;
        i = take ? oi : i;
        valid = valid || ok;
    }
}

// ISL: the FP64-islands mode, whose final face choice and depth are binary64
// (as epa_wave's).
template <int G, bool ISL = false>
__device__ inline EpaResult epa_group(EpgLds<G> &L, const EpaShapes &sh, V3 sa0, V3 sb0, bool act,
                                      const NarrowParams &p)
{
    const int lane = lane_id(), j = lane & (G - 1), gb = lane & (64 - G), g = lane / G;
    auto gbits = [](uint64_t m) { return nans::gbits<G>(m); };
    EpaResult r;
    r.status = act ? EPA_ITER : EPA_NONE;
    r.pa = r.pb = r.n = mk3(0.0f, 0.0f, 0.0f);
    r.depth = 0.0;
    r.iters = 0;
    r.max_faces = 4;
    r.nverts = 4;
    bool run = act;
    const V3 z = mk3(0.0f, 0.0f, 0.0f);
    // vertex j of the group: P, SupA, SupB
    V3 VA = j < 4 ? sa0 : z, VB = j < 4 ? sb0 : z;
    V3 VP = j < 4 ? sub(sa0, sb0) : z;
    if (j < 4) st3(L.vp[g], j, VP);
    // the group's four simplex vertices in every lane of the group
    const V3 P0 = gbcast3<G, 0>(VP), P1 = gbcast3<G, 1>(VP), P2 = gbcast3<G, 2>(VP), P3 = gbcast3<G, 3>(VP);
    int c0;
    {
        int cls = j;
        bool found = false;
        const V3 pu[3] = {P0, P1, P2};
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            if (!found && u < j && veq(pu[u], VP)) {
                cls = u;
                found = true;
            }
        }
        c0 = has_nan(VP) ? CLS_NAN : cls;
    }
    FaceR F = FaceR{0, 0, 0, 0, 0, 0, 0, 0, 0};
    {
        // ABC, ACD, ADB, BDC (code/nans.cpp:799-802)
        const int a = j == 3 ? 1 : 0;
        const int b = j == 0 ? 1 : (j == 1 ? 2 : 3);
        const int c = j == 0 ? 2 : (j == 1 ? 3 : (j == 2 ? 1 : 2));
        const int C0 = gbcast_i<G, 0>(c0), C1 = gbcast_i<G, 1>(c0), C2 = gbcast_i<G, 2>(c0), C3 = gbcast_i<G, 3>(c0);
        V3 Pa = P0, Pb = P1, Pc = P1;
        fsel(Pa, a == 1, P1);
        fsel(Pb, b == 2, P2);
        fsel(Pb, b == 3, P3);
        fsel(Pc, c == 2, P2);
        fsel(Pc, c == 3, P3);
        const int ca = a == 1 ? C1 : C0;
        const int cb = b == 1 ? C1 : (b == 2 ? C2 : C3);
        const int cc = c == 1 ? C1 : (c == 2 ? C2 : C3);
        if (j < 4) F = make_face(Pa, Pb, Pc, a, b, c, ca, cb, cc);
    }
    int nf = 4, nv = 4, iter = 0;
    // A group that converges keeps its closest face (and, for the islands choice,
    // face 0's distance and the winner's index) and stops: its polytope is frozen
    // from then on, so the contact points are computed once for every group
    // after the loop rather than on each iteration where some group converges.
    bool cv = false;
    V3 cN = z;
    float ccur = 0.0f, cd0 = 0.0f;
    int cfid = 0, cbest = 0;
    // the horizon's edge counts (below), zero at the start of every trip
    constexpr int CW = G * G / 4;  // words of counts per group, G / 4 per lane
    auto clear_counts = [&]() {
#pragma unroll
        for (int k = 0; k < CW / G; ++k) L.ecnt[g][j * (CW / G) + k] = 0u;
    };
    clear_counts();
    wave_sync();
    for (;;) {
        // while (iter++ <= MAX_EPA_ITERATIONS), per group
        if (run && !(iter <= p.max_epa_iterations)) run = false;  // status stays EPA_ITER
        iter += 1;
        if (run) r.iters += 1;
        if (run && nf == 0) {
            r.status = EPA_EMPTY;
            run = false;
        }
        if (!ballot(run)) break;
        // closest face: the first face holding the minimum below face 0's distance
        const float d0 = gbcast_f<G, 0>(gabs(dot(mk3(F.nx, F.ny, F.nz), mk3(F.ax, F.ay, F.az))));
        const float k0 = (j < nf && F.d < d0) ? F.d : __builtin_inff();
        const float m = group_min<G>(k0);
        const uint32_t b0 = gbits(ballot(k0 == m));
        const int best = (m != __builtin_inff() && b0) ? ffs64(b0) : 0;
        const int owner = gb + best;
        const V3 N = mk3(gpull_f(F.nx, owner), gpull_f(F.ny, owner), gpull_f(F.nz, owner));
        const float dsel = gpull_f(F.d, owner);
        const float cur = m == __builtin_inff() ? d0 : dsel;
        int fid = gpull_i(F.id, owner);
        // new support along the stored normal
        const Vtx ns = group_support<G>(sh, N);
        const bool conv = run && dot(N, ns.P) - cur < p.epa_tolerance;
        if (conv) {
            cv = true;
            cN = N;
            ccur = cur;
            cd0 = d0;
            cfid = fid;
            cbest = best;
            run = false;
        }
        if (run && nv >= G) {
            r.status = EPA_SPILL;
            run = false;
        }
        // append the vertex (its class: the first earlier vertex with the same P)
        const int vnew = nv;
        const uint32_t meq = gbits(ballot(j < nv && veq(VP, ns.P)));
        const int ncls = has_nan(ns.P) ? CLS_NAN : (meq ? ffs64(meq) : nv);
        if (run) {
            if (j == vnew) {
                VP = ns.P;
                VA = ns.A;
                VB = ns.B;
            }
            if (j == 0) st3(L.vp[g], vnew, ns.P);
            nv += 1;
            r.nverts = nv;
        }
        // visibility
        const bool w0 = run && j < nf && dot(mk3(F.nx, F.ny, F.nz), sub(ns.P, mk3(F.ax, F.ay, F.az))) > 0.0f;
        const uint32_t vis = gbits(ballot(w0));
        const int nvis = __popc(vis);
        // horizon: edges of the visible faces in face order; an edge survives iff its
        // reverse is not among them, when no key is pushed twice (else spill).  The
        // visible faces' directed edges between non-NaN classes (< G) are counted
        // per key (at most 3G pushes, so a byte never wraps); bits of this lane's
        // three edges: same key pushed twice (s2), reverse key pushed (r1) and
        // pushed twice (r2).  Edges with a NaN class never match, as vertex ==
        // never holds for them.  The counts are zero here: cleared before the loop
        // and after each trip's last read of them.
        const int cl = F.cl;
        if (w0) {
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const int ke = edge_key(cl, e);
                const int ca = ke & 255, cb = ke >> 8;
                if (ca < G && cb < G) {  // classes are < G, or CLS_NAN
                    const int key = ca * G + cb;
                    atomicAdd(&L.ecnt[g][key >> 2], 1u << (8 * (key & 3)));
                }
            }
        }
        wave_sync();
        uint32_t s2 = 0, r1 = 0, r2 = 0;
        if (w0) {
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const int ke = edge_key(cl, e);
                const int ca = ke & 255, cb = ke >> 8;
                if (ca < G && cb < G) {
                    const int key = ca * G + cb, rkey = cb * G + ca;
                    const uint32_t nk = (L.ecnt[g][key >> 2] >> (8 * (key & 3))) & 255u;
                    const uint32_t nr = (L.ecnt[g][rkey >> 2] >> (8 * (rkey & 3))) & 255u;
                    s2 |= (nk >= 2u ? 1u : 0u) << e;
                    r1 |= (nr >= 1u ? 1u : 0u) << e;
                    r2 |= (nr >= 2u ? 1u : 0u) << e;
                }
            }
        }
        uint32_t nanbits = 0;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const int ke = edge_key(cl, e);
            nanbits |= ((ke & 255) == CLS_NAN || (ke >> 8) == CLS_NAN) ? (1u << e) : 0u;
        }
        const bool dup = w0 && ((s2 | r2) & ~nanbits & 7u) != 0u;
        const int keep = w0 ? (int)(~r1 & 7u) : 0;   // NaN edges never match: r1 has no NaN bits
        const int cnt = __popc(keep);
        const uint32_t lower = (1u << j) - 1u;
        const uint32_t p0 = gbits(ballot(cnt & 1)), p1 = gbits(ballot(cnt & 2));
        const int ne = __popc(p0) + 2 * __popc(p1);
        if (run && (gbits(ballot(dup)) != 0u || ne > G)) {
            r.status = EPA_SPILL;
            run = false;
        }
        if (run && keep) {
            int o = __popc(p0 & lower) + 2 * __popc(p1 & lower);
            const int id = F.id;
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                if ((keep >> e) & 1) {
                    L.eids[gb + o] = edge_key(id, e);
                    L.ekey[gb + o] = edge_key(cl, e);
                    ++o;
                }
            }
        }
        wave_sync();
        clear_counts();  // every read of the counts is behind the sync above
        int ea = 0, eb = 0, eca = 0, ecb = 0;
        if (run && j < ne) {
            const int k2 = L.ekey[gb + j], i2 = L.eids[gb + j];
            ea = i2 & 255;
            eb = i2 >> 8;
            eca = k2 & 255;
            ecb = k2 >> 8;
        }
        // ordered removal of the visible faces, then the new ones, through LDS
        const bool al = j < nf && !w0;
        const uint32_t alive = gbits(ballot(al));
        const int nsurv = __popc(alive);
        if (run && nsurv + ne > G) {
            r.status = EPA_SPILL;
            run = false;
        }
        FaceR gn = FaceR{0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (run && j < ne) gn = make_face(ns.P, ld3(L.vp[g], ea), ld3(L.vp[g], eb), vnew, ea, eb, ncls, eca, ecb);
        auto stage = [&](int at, const FaceR &f) {
            float *d = L.sf[g][at];
            d[0] = f.nx; d[1] = f.ny; d[2] = f.nz;
            d[3] = f.ax; d[4] = f.ay; d[5] = f.az;
            d[6] = f.d;
            d[7] = __int_as_float(f.id);
            d[8] = __int_as_float(f.cl);
        };
        if (run) {
            if (al) stage(__popc(alive & lower), F);
            if (j < ne) stage(nsurv + j, gn);
        }
        wave_sync();
        if (run) {
            const int nfn = nsurv + ne;
            if (j < nfn) {
                const float *d = L.sf[g][j];
                F.nx = d[0]; F.ny = d[1]; F.nz = d[2];
                F.ax = d[3]; F.ay = d[4]; F.az = d[5];
                F.d = d[6];
                F.id = __float_as_int(d[7]);
                F.cl = __float_as_int(d[8]);
            }
            nf = nfn;
            if (nf > r.max_faces) r.max_faces = nf;
        }
        wave_sync();
    }
    if (ballot(cv)) {
        V3 Nc = cN;
        float cc = ccur;
        double depth = (double)ccur;
        int fid = cfid;
        if constexpr (ISL) {
            // the final face by binary64 plane distance: the first of the least
            // below face 0's, else face 0 (epa_wave's islands choice per group)
            wave_sync();
            const double d64 = j < nf ? group_face_dist64(L.vp[g], F) : 0.0;
            const double f0 = __shfl(d64, gb, 64);
            bool v64 = cv && j < nf && d64 < f0;
            double bd = v64 ? d64 : 0.0;
            int bi = v64 ? j : 0x7fffffff;
            group_argmin64<G>(v64, bd, bi);
            const int sel = v64 ? bi : 0;
            depth = v64 ? bd : f0;
            const int so = gb + sel;
            const V3 Ns = mk3(gpull_f(F.nx, so), gpull_f(F.ny, so), gpull_f(F.nz, so));
            const float ds = gpull_f(F.d, so);
            const int fs = gpull_i(F.id, so);
            if (sel != cbest) {
                Nc = Ns;
                cc = sel == 0 ? cd0 : ds;
                fid = fs;
            }
        }
        const int a = fid & 255, b = (fid >> 8) & 255, c = (fid >> 16) & 255;
        const int am = gb + (a & (G - 1)), bm = gb + (b & (G - 1)), cm = gb + (c & (G - 1));
        const V3 PA = gpull3(VP, am), PB = gpull3(VP, bm), PC = gpull3(VP, cm);
        const V3 AA = gpull3(VA, am), AB = gpull3(VA, bm), AC = gpull3(VA, cm);
        const V3 BA = gpull3(VB, am), BB = gpull3(VB, bm), BC = gpull3(VB, cm);
        if (cv) {
            const V3 P = mul(Nc, cc);
            const V3 e0 = sub(PB, PA), e1 = sub(PC, PA), e2 = sub(P, PA);
            const float d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1);
            const float d20 = dot(e2, e0), d21 = dot(e2, e1);
            const float denom = d00 * d11 - d01 * d01;
            const float V = (d11 * d20 - d01 * d21) / denom;
            const float W = (d00 * d21 - d01 * d20) / denom;
            const float U = 1.0f - V - W;
            if (fabsf(U) > 1.0f || fabsf(V) > 1.0f || fabsf(W) > 1.0f) {
                r.status = EPA_BARY;
            } else if (!isfinite(U) || !isfinite(V) || !isfinite(W)) {
                r.status = EPA_NONFINITE;
            } else {
                r.pa = add(add(mul(AA, U), mul(AB, V)), mul(AC, W));
                r.n = mul(Nc, -1.0f);
                r.pb = add(add(mul(BA, U), mul(BB, V)), mul(BC, W));
                r.depth = depth;
                r.status = EPA_CONTACT;
            }
        }
    }
    return r;
}

// ---------------------------------------------------------------------- block driver
// CheckCollision (code/nans.cpp:906-966) over the candidates [0, ncand) of one
// workgroup, in two phases:
//   GJK  every lane runs one candidate's EvolveSimplex loop; lanes whose loops
//        have ended take the next candidates together once GJK_REFILL_MIN of
//        them wait (work is handed out per wave by one LDS atomic), so a wave is
//        not held by its slowest lane;
//   EPA  the GJK hits are queued with their simplex; each wave takes queued hits
//        (LDS atomic) and runs ResolveCollision on them with the polytope in its
//        LDS block.  In the reference's arithmetic a wave starts on the queue as
//        soon as its own GJK lanes are done, while other waves still run GJK: a
//        hit's queue slot is published with a release store of its candidate
//        index, and a wave that takes a slot not yet published waits for it, or
//        for every wave to leave GJK (then the slot will never be filled).
// Results go through the source object, per candidate, so their order is that
// of the candidates whatever the order of the work.
struct NarrowShared {
    int next;      // next candidate to hand out
    int nepa;      // GJK hits queued
    int epa_next;  // next queued hit to hand out
    int gjk_done;  // waves that have left the GJK phase
    int stalled;   // queue waits that hit their bound (counted as EPA overflow)
};

// Bound of a wave's wait for a queue slot (each round sleeps ~128 cycles): far
// beyond any GJK phase, so reaching it means a fault, which is counted and the
// slot skipped rather than waited on for ever.
constexpr int EPA_WAIT_ROUNDS = 1 << 20;

// First EvolveSimplex call whose state enters the cycle check (any start finds
// a cycle; a later one spares the loops that end early).
constexpr int GJK_CYCLE_START = 8;

// Idle GJK lanes take new candidates once this many of them wait (or no lane
// is busy), so that the refill code runs once for a batch of lanes (config 3 on
// the MI355X: 1.429 ms per step at 16, 1.413 at 32-56, 1.484 at 64).
constexpr int GJK_REFILL_MIN = 48;

// A queued GJK hit's simplex: the supports on A and B of its four vertices (the
// Minkowski point is recomputed as SupA - SupB, the expression that made it),
// and the pair's key (the source's key(c)), so that EPA looks nothing else up.
struct SimplexRec {
    float a[12], b[12];
    int key;
};

__device__ __forceinline__ void simplex_rec_put(SimplexRec &q, const Vtx &v, int k)
{
    q.a[3 * k] = v.A.x; q.a[3 * k + 1] = v.A.y; q.a[3 * k + 2] = v.A.z;
    q.b[3 * k] = v.B.x; q.b[3 * k + 1] = v.B.y; q.b[3 * k + 2] = v.B.z;
}

__device__ __forceinline__ void simplex_rec_store(SimplexRec &q, const Simplex &S)
{
    simplex_rec_put(q, S.s0, 0);
    simplex_rec_put(q, S.s1, 1);
    simplex_rec_put(q, S.s2, 2);
    simplex_rec_put(q, S.s3, 3);
}

// One lane's GJK loop (CheckCollision's EvolveSimplex calls, code/nans.cpp:906-966)
// over the candidates it is handed.
//
// GJK cycles: EvolveSimplex's next call depends only on the simplex's Minkowski
// points (SupA/SupB ride along), so once the points repeat bit for bit the calls
// repeat for ever and the loop can only end at the cap.  Each lane keeps one
// earlier state (Brent's scheme, started at call GJK_CYCLE_START: saved after that
// call and then 1, 2, 4, ... calls later) in cyc[], its wave's [16][64] floats
// (P of s0..s3, n, then the pair's first direction).  Most loops end before the
// start, and pay nothing for the check.
struct GjkLane {
    Shape A, B;
    Simplex S;
    int c = -1, key = 0, calls = 0, power = 1, lam = 0;
    bool saved = false, retired = false;
};

// Candidate c for the lane (c >= ncand: none left, the lane retires).
template <class Src>
__device__ inline void gjk_lane_take(GjkLane &G, const Src &src, int c, int ncand, float *cyc)
{
    if (c >= ncand) {
        G.retired = true;
        return;
    }
    const int tid = lane_id();
    G.c = c;
    G.key = src.key(c);
    src.shapes(G.key, G.A, G.B);
    const V3 d0 = gjk_first_dir(G.A, G.B);  // kept beside the cycle state
    cyc[13 * 64 + tid] = d0.x;
    cyc[14 * 64 + tid] = d0.y;
    cyc[15 * 64 + tid] = d0.z;
    G.S.n = 0;
    G.calls = 0;
    G.saved = false;
}

// One EvolveSimplex call of the lane's candidate, with the cycle check: the
// outcome once the loop has ended (GJK_NO, GJK_FOUND, or GJK_CAP at the 65-call
// cap), else GJK_STILL.
__device__ inline int gjk_lane_call(GjkLane &G, float *cyc, const NarrowParams &np)
{
    const int tid = lane_id();
    Simplex &S = G.S;
    const V3 d0 = mk3(cyc[13 * 64 + tid], cyc[14 * 64 + tid], cyc[15 * 64 + tid]);
    int res = evolve_simplex(G.A, G.B, S, np.fp64_islands, d0);
    ++G.calls;
    if (res == GJK_STILL && G.calls <= np.max_gjk_iterations && G.calls >= GJK_CYCLE_START) {
        const float st[13] = {S.s0.P.x, S.s0.P.y, S.s0.P.z, S.s1.P.x, S.s1.P.y, S.s1.P.z,
                              S.s2.P.x, S.s2.P.y, S.s2.P.z, S.s3.P.x, S.s3.P.y, S.s3.P.z,
                              __int_as_float(S.n)};
        if (G.saved) {
            bool same = __float_as_int(cyc[12 * 64 + tid]) == S.n;
#pragma unroll
            for (int k = 0; k < 12; ++k)
                same = same && (k >= 3 * S.n || __float_as_int(cyc[k * 64 + tid]) == __float_as_int(st[k]));
            if (same) res = GJK_CAP;  // a repeated state: the cap is certain
        }
        ++G.lam;
        if (res == GJK_STILL && (!G.saved || G.lam == G.power)) {
#pragma unroll
            for (int k = 0; k < 13; ++k) cyc[k * 64 + tid] = st[k];
            G.power = G.saved ? 2 * G.power : 1;
            G.saved = true;
            G.lam = 0;
        }
    }
    if (res != GJK_STILL || G.calls > np.max_gjk_iterations) return res == GJK_STILL ? GJK_CAP : res;
    return GJK_STILL;
}

// Source interface (duck-typed):
//   int key(int c) const                                 the pair of candidate c, as an int
//   static constexpr bool kMemCorners                    box corners readable in memory:
//   void shapes_ref(int key, Shape &A, Shape &B) const   then both shapes of the pair
//                                                        with Shape::vm and no v[]
//   void shapes(int key, Shape &A, Shape &B) const       both shapes of the pair
//   void gjk_result(int c, int gjk) const                 every candidate
//   void epa_result(int c, int key, const EpaResult &r) const   every GJK hit (lane 0 calls)
// WIDE: polytopes that outgrow a 16-lane group go to a 32-lane group before the
// whole wave (worth its registers where many do, as among random deep overlaps).
// ISL: the FP64-islands mode runs the grouped tiers too (their final face choice
// in binary64); without it, that mode runs every polytope by the whole wave.
// The team's wave k uses epa[k]; sh, qidx and qsim are the team's own.
template <bool WIDE, bool ISL = false, class Src>
__device__ void narrow_team(const Team &team, const Src &src, int ncand, const NarrowParams &np, EpaLds *epa,
                            NarrowShared &sh, int *qidx, SimplexRec *qsim)
{
    const bool grouped = ISL || !np.fp64_islands;
    const int lane = lane_id();
    if (team.rank == 0) {
        sh.next = 0;
        sh.nepa = 0;
        sh.epa_next = 0;
        sh.gjk_done = 0;
        sh.stalled = 0;
    }
    if (grouped)
        for (int k = team.rank; k < ncand; k += team.size) qidx[k] = -1;  // queue slots: unpublished
    team.sync();
    {
        // GJK lanes (gjk_lane_call); the cycle state lives in the wave's EPA LDS,
        // idle until the wave's GJK ends
        float *cyc = reinterpret_cast<float *>(&epa[team.rank >> 6]);
        GjkLane G;
        for (;;) {
            const bool need = G.c < 0 && !G.retired;
            const uint64_t m = ballot(need);
            if (m && (__popcll(m) >= GJK_REFILL_MIN || !ballot(G.c >= 0))) {
                int base = 0;
                if (lane == 0) base = atomicAdd(&sh.next, __popcll(m));
                base = __builtin_amdgcn_readfirstlane(base);
                if (need) gjk_lane_take(G, src, base + __popcll(m & ((1ull << lane) - 1ull)), ncand, cyc);
            }
            if (!ballot(G.c >= 0)) break;
            if (G.c >= 0) {
                const int g = gjk_lane_call(G, cyc, np);
                if (g != GJK_STILL) {
                    src.gjk_result(G.c, g);
                    if (g == GJK_FOUND) {
                        const int q = atomicAdd(&sh.nepa, 1);
                        simplex_rec_store(qsim[q], G.S);
                        qsim[q].key = G.key;
                        // publish the slot after its simplex
                        __hip_atomic_store(&qidx[q], G.c, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    G.c = -1;
                }
            }
        }
        // this wave's queue slots are published before it counts itself out
        if (lane == 0) __hip_atomic_fetch_add(&sh.gjk_done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    EpaLds &L = epa[team.rank >> 6];
    if (grouped) {
        // four pairs per wave, one per 16-lane group, started while other waves
        // may still be in GJK; a spilled polytope is redone by the whole wave, or
        // with WIDE by a 32-lane group (two at a time) and, if it spills there
        // too, by the whole wave
        const int NW = team.waves();
        EpgLds<16> &G = *reinterpret_cast<EpgLds<16> *>(&L);
        EpgLds<32> &G2 = *reinterpret_cast<EpgLds<32> *>(&L);
        const int j = lane & 15;
        // A pass takes four queue slots, one per group.  The next pass's slots are
        // claimed while this one starts, their contents loaded (relaxed) under its
        // work, and then confirmed by an acquire fence, or waited for.
        auto claim = [&]() {
            int b = 0;
            if (lane == 0) b = atomicAdd(&sh.epa_next, 4);
            return __builtin_amdgcn_readfirstlane(b);
        };
        auto peek = [&](int b) {
            const int job = b + (lane >> 4);
            return job < ncand ? __hip_atomic_load(&qidx[job], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : -1;
        };
        // the candidate of this group's slot, or -1 once it is certain the slot
        // stays empty (every wave has left GJK and it was never published)
        auto resolve = [&](int b, int seen) {
            const int job = b + (lane >> 4);
            int c = -1;
            if (job < ncand) {
                if (seen >= 0 && seen < ncand) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    return seen;
                }
                for (int round = 0;; ++round) {
                    c = __hip_atomic_load(&qidx[job], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (c >= ncand) {  // not a candidate of this block: never published here
                        if (j == 0) atomicAdd(&sh.stalled, 1);
                        c = -1;
                        break;
                    }
                    if (c >= 0) break;
                    if (__hip_atomic_load(&sh.gjk_done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == NW) {
                        c = __hip_atomic_load(&qidx[job], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (c >= ncand) {
                            if (j == 0) atomicAdd(&sh.stalled, 1);
                            c = -1;
                        }
                        break;
                    }
                    if (round >= EPA_WAIT_ROUNDS) {
                        if (j == 0) atomicAdd(&sh.stalled, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            return c;
        };
        int base = claim();
        int c = resolve(base, peek(base));
        for (;;) {
            // slots are filled in order, so an empty first slot ends the queue
            if (__shfl(c, 0, 64) < 0) break;
            const int job = base + (lane >> 4);
            const bool act = c >= 0;
            const int jq = act ? job : base;
            if (!act) c = __shfl(c, 0, 64);
            const int key = qsim[jq].key;
            Shape A, B;
            if constexpr (Src::kMemCorners) src.shapes_ref(key, A, B);
            else src.shapes(key, A, B);
            const EpaShapes es = epa_shapes_group<Src::kMemCorners>(A, B);
            V3 sa0 = mk3(0.0f, 0.0f, 0.0f), sb0 = sa0;
            if (j < 4) {
                const SimplexRec &q = qsim[jq];
                sa0 = mk3(q.a[3 * j], q.a[3 * j + 1], q.a[3 * j + 2]);
                sb0 = mk3(q.b[3 * j], q.b[3 * j + 1], q.b[3 * j + 2]);
            }
            wave_sync();
            const int nbase = claim();
            const int seen = peek(nbase);
            // WIDE: a pair with a sphere goes to the 32-lane tier at once (a sphere's
            // support gives a new vertex at every iteration, so its polytopes grow)
            const bool sph = WIDE && (A.kind != SHAPE_BOX || B.kind != SHAPE_BOX);
            const EpaResult r = epa_group<16, ISL>(G, es, sa0, sb0, act && !sph, np);
            if (act && !sph && j == 0 && r.status != EPA_SPILL) src.epa_result(c, key, r);
            const uint64_t spill = ballot(act && j == 0 && (sph || r.status == EPA_SPILL));
            wave_sync();
            const int nsp = __popcll(spill);
            for (int t = 0; WIDE && 2 * t < nsp; ++t) {
                // spilled group k = 2t + h goes to 32-lane group h
                const int h = lane >> 5, j2 = lane & 31;
                const int k = 2 * t + h;
                const bool act2 = k < nsp;
                const int sg = nth_set_bit(spill, act2 ? k : 2 * t) >> 4;
                const int cs = __shfl(c, 16 * sg, 64);
                const int ks = __shfl(key, 16 * sg, 64);
                const int sj = base + sg;
                Shape SA, SB;
                src.shapes(ks, SA, SB);
                const EpaShapes e2 = epa_shapes_group(SA, SB);
                V3 a2 = mk3(0.0f, 0.0f, 0.0f), b2 = a2;
                if (j2 < 4) {
                    const SimplexRec &q = qsim[sj];
                    a2 = mk3(q.a[3 * j2], q.a[3 * j2 + 1], q.a[3 * j2 + 2]);
                    b2 = mk3(q.b[3 * j2], q.b[3 * j2 + 1], q.b[3 * j2 + 2]);
                }
                wave_sync();
                const EpaResult r2 = epa_group<32, ISL>(G2, e2, a2, b2, act2, np);
                if (act2 && j2 == 0 && r2.status != EPA_SPILL) src.epa_result(cs, ks, r2);
                uint64_t spill2 = ballot(act2 && j2 == 0 && r2.status == EPA_SPILL);
                wave_sync();
                while (spill2) {
                    const int src_lane = ffs64(spill2);  // 0 or 32
                    spill2 &= spill2 - 1;
                    const int cw = __shfl(cs, src_lane, 64);
                    const int kw = __shfl(ks, src_lane, 64);
                    const int sw = __shfl(sj, src_lane, 64);
                    Shape WA, WB;
                    src.shapes(kw, WA, WB);
                    const EpaShapes ew = epa_shapes(WA, WB);
                    V3 p0 = mk3(0.0f, 0.0f, 0.0f);
                    if (lane < 4) {
                        const SimplexRec &q = qsim[sw];
                        V3 a = mk3(q.a[3 * lane], q.a[3 * lane + 1], q.a[3 * lane + 2]);
                        V3 b = mk3(q.b[3 * lane], q.b[3 * lane + 1], q.b[3 * lane + 2]);
                        p0 = sub(a, b);
                        st3(L.va, lane, a);
                        st3(L.vb, lane, b);
                        st3(L.vp, lane, p0);
                    }
                    wave_sync();
                    const EpaResult rw = epa_wave(L, ew, p0, np);
                    if (lane == 0) src.epa_result(cw, kw, rw);
                    wave_sync();
                }
            }
            for (uint64_t sp = WIDE ? 0ull : spill; sp; sp &= sp - 1) {
                const int sg = ffs64(sp) >> 4;  // the spilled group
                const int cs = __shfl(c, 16 * sg, 64);
                const int ks = __shfl(key, 16 * sg, 64);
                Shape SA, SB;
                src.shapes(ks, SA, SB);
                const EpaShapes ew = epa_shapes(SA, SB);
                V3 p0 = mk3(0.0f, 0.0f, 0.0f);
                if (lane < 4) {
                    const SimplexRec &q = qsim[base + sg];
                    V3 a = mk3(q.a[3 * lane], q.a[3 * lane + 1], q.a[3 * lane + 2]);
                    V3 b = mk3(q.b[3 * lane], q.b[3 * lane + 1], q.b[3 * lane + 2]);
                    p0 = sub(a, b);
                    st3(L.va, lane, a);
                    st3(L.vb, lane, b);
                    st3(L.vp, lane, p0);
                }
                wave_sync();
                const EpaResult rw = epa_wave(L, ew, p0, np);
                if (lane == 0) src.epa_result(cs, ks, rw);
                wave_sync();
            }
            base = nbase;
            c = resolve(base, seen);
        }
        team.sync();
        return;
    }
    team.sync();
    const int nepa = sh.nepa;
    for (;;) {
        int job = 0;
        if (lane == 0) job = atomicAdd(&sh.epa_next, 1);
        job = __builtin_amdgcn_readfirstlane(job);
        if (job >= nepa) break;
        const int c = qidx[job];
        const int key = qsim[job].key;
        Shape A, B;
        src.shapes(key, A, B);
        const EpaShapes es = epa_shapes(A, B);
        V3 p0 = mk3(0.0f, 0.0f, 0.0f);
        if (lane < 4) {
            const SimplexRec &q = qsim[job];
            V3 a = mk3(q.a[3 * lane], q.a[3 * lane + 1], q.a[3 * lane + 2]);
            V3 b = mk3(q.b[3 * lane], q.b[3 * lane + 1], q.b[3 * lane + 2]);
            p0 = sub(a, b);
            st3(L.va, lane, a);
            st3(L.vb, lane, b);
            st3(L.vp, lane, p0);
        }
        wave_sync();
        const EpaResult r = epa_wave(L, es, p0, np);
        if (lane == 0) src.epa_result(c, key, r);
        wave_sync();
    }
    team.sync();
}

// narrow_team by the whole workgroup of BLOCK threads
template <int BLOCK, bool WIDE, bool ISL = false, class Src>
__device__ inline void narrow_block(const Src &src, int ncand, const NarrowParams &np, EpaLds *epa,
                                    NarrowShared &sh, int *qidx, SimplexRec *qsim)
{
    narrow_team<WIDE, ISL>(Team::block(BLOCK), src, ncand, np, epa, sh, qidx, qsim);
}

}  // namespace nans
