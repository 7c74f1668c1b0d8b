// nans_epa.hpp -- EPA (ResolveCollision, code/nans.cpp:787-904) by one wavefront
// per pair, and the workgroup narrowphase driver (GJK per lane, EPA per wave).
//
// Layout of the polytope:
//   vertices   LDS: P, SupA, SupB and an equality class per vertex (the first
//              vertex with the same P, 255 if P has a NaN: never equal);
//   faces      registers: lane l owns faces l, l+64, l+128 and keeps for each its
//              stored normal N, first vertex A, distance |dot(n, A)| and the packed
//              ids and classes of its three vertices;
//   edges      registers: horizon edge e lives in lane e.
// One iteration of the reference's loop (code/nans.cpp:804-902):
//   closest face   the first face holding the minimum of the distances below face
//                  0's (the strict '<' scan from face 0, code/nans.cpp:807-822):
//                  DPP row minimum + ballots, the winner's data by readlane;
//   new support    along the winner's stored normal (code/nans.cpp:824-825);
//   visibility     each lane tests its faces against the new point, ballots;
//   horizon        the visible faces' edges pushed in face order with the exact
//                  reversed-edge cancellation of PushEdge (code/nans.cpp:232-266),
//                  one ballot per push;
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
    int vcls[EPA_MAXV];
    float sfn[EPA_MAXF][3];  // compaction staging: stored normal
    float sfd[EPA_MAXF];     //                     distance
    int sfid[EPA_MAXF];      //                     a | b << 8 | c << 16
    int sfcl[EPA_MAXF];      //                     classes of a, b, c
    float sa[32];            // shape A: kind, c.xyz, r, 24 corner floats
    float sb[32];
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
__device__ __forceinline__ int ffs64(uint64_t m) { return __builtin_ctzll(m); }

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

// Support of an LDS-resident shape (every lane reads the same words: broadcast).
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

__device__ __forceinline__ V3 ld3(const float (*a)[3], int i) { return mk3(a[i][0], a[i][1], a[i][2]); }
__device__ __forceinline__ void st3(float (*a)[3], int i, V3 v)
{
    a[i][0] = v.x;
    a[i][1] = v.y;
    a[i][2] = v.z;
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

// Equality class of vertex id: the first earlier vertex with the same P (vertex
// == compares P, code/nans.h:251-254), CLS_NAN if P has a NaN (never equal).
// Every lane must call it.
__device__ inline int vertex_class(const EpaLds &L, int id, V3 P)
{
    if (has_nan(P)) return CLS_NAN;
    int lane = lane_id();
    for (int base = 0; base < id; base += 64) {
        int u = base + lane;
        bool m = u < id && veq(ld3(L.vp, u), P);
        uint64_t mm = ballot(m);
        if (mm) return base + ffs64(mm);
    }
    return id;
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

    {
        int cls = 0;
        for (int v = 0; v < 4; ++v) {
            int c = vertex_class(L, v, ld3(L.vp, v));
            if (lane == v) cls = c;
        }
        if (lane < 4) L.vcls[lane] = cls;
    }
    wave_sync();
    FaceR F[EPA_FSLOTS];
#pragma unroll
    for (int k = 0; k < EPA_FSLOTS; ++k) F[k] = FaceR{0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (lane < 4) {
        // ABC, ACD, ADB, BDC (code/nans.cpp:799-802)
        const int a = lane == 3 ? 1 : 0;
        const int b = lane == 0 ? 1 : (lane == 1 ? 2 : 3);
        const int c = lane == 0 ? 2 : (lane == 1 ? 3 : (lane == 2 ? 1 : 2));
        F[0] = make_face(ld3(L.vp, a), ld3(L.vp, b), ld3(L.vp, c), a, b, c, L.vcls[a], L.vcls[b], L.vcls[c]);
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
        Vtx ns;
        ns.A = support_lds(L.sa, N);
        ns.B = support_lds(L.sb, mul(N, -1.0f));
        ns.P = sub(ns.A, ns.B);
        if (dot(N, ns.P) - cur < p.epa_tolerance) {
            double depth = (double)cur;
            if (p.fp64_islands) {
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
        const int ncls = vertex_class(L, vnew, ns.P);
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
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            const int f = lane + 64 * k;
            const bool vv = f < nf && dot(mk3(F[k].nx, F[k].ny, F[k].nz),
                                          sub(ns.P, mk3(F[k].ax, F[k].ay, F[k].az))) > 0.0f;
            vis[k] = ballot(vv);
        }
        // horizon edges in face order (PushEdge), one edge per lane
        int ne = 0, ea = 0, eb = 0, eca = 0, ecb = 0;
        bool overflow = false;
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
        if (overflow) {
            r.status = EPA_OVERFLOW;
            break;
        }
        // ordered removal of the visible faces, then the new ones, through LDS
        int nsurv = 0;
        int dst[EPA_FSLOTS];
#pragma unroll
        for (int k = 0; k < EPA_FSLOTS; ++k) {
            const int f = lane + 64 * k;
            const bool alive = f < nf && !((vis[k] >> lane) & 1ull);
            const uint64_t am = ballot(alive);
            dst[k] = alive ? nsurv + __popcll(am & ((1ull << lane) - 1ull)) : -1;
            nsurv += __popcll(am);
        }
        if (nsurv + ne > EPA_MAXF) {
            r.status = EPA_OVERFLOW;
            break;
        }
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
            // PushTriangle(NewSupport, e.A, e.B) (code/nans.cpp:894-899)
            FaceR g = make_face(ns.P, ld3(L.vp, ea), ld3(L.vp, eb), vnew, ea, eb, ncls, eca, ecb);
            const int s = nsurv + lane;
            L.sfn[s][0] = g.nx;
            L.sfn[s][1] = g.ny;
            L.sfn[s][2] = g.nz;
            L.sfd[s] = g.d;
            L.sfid[s] = g.id;
            L.sfcl[s] = g.cl;
        }
        nf = nsurv + ne;
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
