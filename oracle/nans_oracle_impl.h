/*
 * nans_oracle_impl.h -- body of the CPU oracle, compiled once per precision.
 *
 * TEST INFRASTRUCTURE ONLY (see nans_oracle.h).  Included by
 * nans_oracle_f32.c / nans_oracle_f64.c with
 *   REAL      float | double
 *   SUF(x)    x##_f32 | x##_f64
 *   ORA_F32   1 | 0
 * Every function cites the code/nans.cpp lines it restates.  The restatement is
 * deliberately literal (same operation order, same std::vector erase semantics,
 * recomputed normals) so that it can serve as the checker of the HIP kernels,
 * which use an equivalent but reorganised formulation.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#if ORA_F32
#define SQRT sqrtf
/* glm::rotate calls the binary32 cos/sin of the linked libm (code/nans.cpp:1915-1920
 * through matrix_transform.inl).  Those are implementation-defined to within an ulp
 * (glibc's sinf/cosf are not correctly rounded for about 1% of arguments), so the
 * restatement uses the correctly rounded binary32 value, obtained by rounding the
 * binary64 result.  Built with ORA_LIBM_SINF (libnans_oracle_libm.so) it calls the
 * host libm's sinf/cosf instead, as a glibc-linked build of the reference does;
 * that variant measures how far the choice moves the results. */
#ifdef ORA_LIBM_SINF
static inline float ora_sinf(float x) { return sinf(x); }
static inline float ora_cosf(float x) { return cosf(x); }
#else
static inline float ora_sinf(float x) { return (float)sin((double)x); }
static inline float ora_cosf(float x) { return (float)cos((double)x); }
#endif
#define SIN ora_sinf
#define COS ora_cosf
#else
#define SQRT sqrt
#define SIN sin
#define COS cos
#endif

/* Work counters: events of the step (bodies integrated, pairs tested, GJK calls
 * by simplex size, support evaluations, EPA iterations with their face and edge
 * counts, contacts solved), from which the measurement harness derives the
 * algorithmic flop count of a workload.  Off unless ora_counting_* enables them;
 * the same call can switch on the product's bounding-sphere cull so that the
 * counted work is the work the HIP step performs. */
enum {
    OC_BODY_BOX, OC_BODY_SPH, OC_PAIRS, OC_CAND, OC_EV0, OC_EV1, OC_EV2, OC_EV3, OC_EV4, OC_EVLEN,
    OC_SUP_BOX, OC_SUP_SPH, OC_EPA_ITER, OC_EPA_F, OC_EPA_E, OC_EPA_FINAL, OC_CONSTRAINT, OC_SOLVER_IT,
    OC_N
};
static long long SUF(g_counts)[32];
static int SUF(g_counting);
static int SUF(g_cull);
#define OC(k, n)                                   \
    do {                                           \
        if (SUF(g_counting)) SUF(g_counts)[k] += (n); \
    } while (0)

void SUF(ora_counting)(int32_t enable, int32_t cull)
{
    SUF(g_counting) = enable;
    SUF(g_cull) = cull;
    memset(SUF(g_counts), 0, sizeof(SUF(g_counts)));
}

void SUF(ora_get_counts)(int64_t *out)
{
    for (int i = 0; i < 32; ++i) out[i] = SUF(g_counts)[i];
}

typedef struct { REAL x, y, z; } SUF(V3);
#define V3 SUF(V3)
typedef struct { V3 P, SupA, SupB; } SUF(Vert);  /* vertex, code/nans.h:245-255 */
#define Vert SUF(Vert)
typedef struct { REAL m[16]; } SUF(M4);            /* column-major glm::mat4, m[c*4+r] */
#define M4 SUF(M4)

static inline V3 SUF(v3)(REAL x, REAL y, REAL z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
#define mk3 SUF(v3)
static inline V3 SUF(vadd)(V3 a, V3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 SUF(vsub)(V3 a, V3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 SUF(vmul)(V3 a, REAL s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline V3 SUF(vdiv)(V3 a, REAL s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline V3 SUF(vneg)(V3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline REAL SUF(vdot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 SUF(vcross)(V3 a, V3 b)
{
    return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline V3 SUF(vnormalize)(V3 v) { return SUF(vmul)(v, (REAL)1 / SQRT(SUF(vdot)(v, v))); }
static inline REAL SUF(vlength)(V3 v) { return SQRT(SUF(vdot)(v, v)); }
static inline REAL SUF(gabs)(REAL x) { return x >= (REAL)0 ? x : -x; }
static inline int SUF(veq)(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
#define vadd SUF(vadd)
#define vsub SUF(vsub)
#define vmul SUF(vmul)
#define vdiv SUF(vdiv)
#define vneg SUF(vneg)
#define vdot SUF(vdot)
#define vcross SUF(vcross)
#define vnormalize SUF(vnormalize)
#define vlength SUF(vlength)
#define gabs SUF(gabs)
#define veq SUF(veq)

/* ---------------------------------------------------------------- glm matrix ops */
static M4 SUF(m4_identity)(void)
{
    M4 r;
    memset(&r, 0, sizeof(r));
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = (REAL)1;
    return r;
}

/* glm::translate: Result[3] = m[0]*v[0] + m[1]*v[1] + m[2]*v[2] + m[3] */
static M4 SUF(m4_translate)(M4 m, V3 v)
{
    M4 r = m;
    for (int i = 0; i < 4; ++i)
        r.m[12 + i] = ((m.m[0 + i] * v.x + m.m[4 + i] * v.y) + m.m[8 + i] * v.z) + m.m[12 + i];
    return r;
}

/* glm::rotate(m, angle, axis), Rodrigues form of matrix_transform.inl */
static M4 SUF(m4_rotate)(M4 m, REAL a, V3 v)
{
    REAL c = COS(a);
    REAL s = SIN(a);
    V3 axis = vnormalize(v);
    V3 temp = vmul(axis, (REAL)1 - c);
    REAL R[3][3];
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
    for (int col = 0; col < 3; ++col)
        for (int i = 0; i < 4; ++i)
            r.m[col * 4 + i] = (m.m[0 + i] * R[col][0] + m.m[4 + i] * R[col][1]) + m.m[8 + i] * R[col][2];
    for (int i = 0; i < 4; ++i)
        r.m[12 + i] = m.m[12 + i];
    return r;
}

/* glm::scale */
static M4 SUF(m4_scale)(M4 m, V3 v)
{
    M4 r;
    for (int i = 0; i < 4; ++i) {
        r.m[0 + i] = m.m[0 + i] * v.x;
        r.m[4 + i] = m.m[4 + i] * v.y;
        r.m[8 + i] = m.m[8 + i] * v.z;
        r.m[12 + i] = m.m[12 + i];
    }
    return r;
}

/* glm mat4 * vec4, xyz of the result: (c0*x + c1*y) + (c2*z + c3*w) */
static V3 SUF(m4_apply)(const M4 *m, REAL x, REAL y, REAL z, REAL w)
{
    REAL o[3];
    for (int i = 0; i < 3; ++i)
        o[i] = (m->m[0 + i] * x + m->m[4 + i] * y) + (m->m[8 + i] * z + m->m[12 + i] * w);
    return mk3(o[0], o[1], o[2]);
}

/* glm::radians */
static inline REAL SUF(radians)(REAL deg) { return deg * (REAL)0.01745329251994329576923690768489; }

/* Model rebuild of the render section: translate * rotX * rotY * rotZ * scale with
 * the Euler angles read as degrees (code/nans.cpp:1913-1924, 1930-1941, 1870-1879). */
static M4 SUF(model_from_pose)(V3 pos, V3 ang, V3 scale)
{
    M4 m = SUF(m4_identity)();
    m = SUF(m4_translate)(m, pos);
    m = SUF(m4_rotate)(m, SUF(radians)(ang.x), mk3(1, 0, 0));
    m = SUF(m4_rotate)(m, SUF(radians)(ang.y), mk3(0, 1, 0));
    m = SUF(m4_rotate)(m, SUF(radians)(ang.z), mk3(0, 0, 1));
    m = SUF(m4_scale)(m, scale);
    return m;
}

/* UpdateVertices / FloorUpdateVertices (code/nans.cpp:379-407): corner order */
static const REAL SUF(kCorner)[8][3] = {
    {0.5, 0.5, 0.5}, {0.5, 0.5, -0.5}, {-0.5, 0.5, 0.5}, {-0.5, 0.5, -0.5},
    {0.5, -0.5, 0.5}, {0.5, -0.5, -0.5}, {-0.5, -0.5, 0.5}, {-0.5, -0.5, -0.5}};

static void SUF(update_vertices)(const M4 *m, V3 *verts)
{
    for (int k = 0; k < 8; ++k)
        verts[k] = SUF(m4_apply)(m, SUF(kCorner)[k][0], SUF(kCorner)[k][1], SUF(kCorner)[k][2], (REAL)1);
}

/* Opt-in quaternion orientation (no counterpart in the reference, whose Angles
 * are Euler degrees, code/nans.cpp:1339, 1913-1924).  q = (w, x, y, z), unit.
 * Model = [R(q) * diag(S) | P]: the columns of the rotation matrix of q scaled by
 * the Model scale, applied to the corners like the Euler Model. */
static M4 SUF(model_from_quat)(V3 pos, const REAL *q, V3 scale)
{
    const REAL w = q[0], x = q[1], y = q[2], z = q[3];
    const REAL xx = x * x, yy = y * y, zz = z * z;
    const REAL xy = x * y, xz = x * z, yz = y * z;
    const REAL wx = w * x, wy = w * y, wz = w * z;
    M4 m;
    m.m[0] = ((REAL)1 - (REAL)2 * (yy + zz)) * scale.x;
    m.m[1] = ((REAL)2 * (xy + wz)) * scale.x;
    m.m[2] = ((REAL)2 * (xz - wy)) * scale.x;
    m.m[3] = 0;
    m.m[4] = ((REAL)2 * (xy - wz)) * scale.y;
    m.m[5] = ((REAL)1 - (REAL)2 * (xx + zz)) * scale.y;
    m.m[6] = ((REAL)2 * (yz + wx)) * scale.y;
    m.m[7] = 0;
    m.m[8] = ((REAL)2 * (xz + wy)) * scale.z;
    m.m[9] = ((REAL)2 * (yz - wx)) * scale.z;
    m.m[10] = ((REAL)1 - (REAL)2 * (xx + yy)) * scale.z;
    m.m[11] = 0;
    m.m[12] = pos.x;
    m.m[13] = pos.y;
    m.m[14] = pos.z;
    m.m[15] = 1;
    return m;
}

/* q <- normalize(q + dt/2 (0, W) q): W is the world-frame angular velocity in
 * rad/s, the quantity the Constraint terms (W x N, code/nans.cpp:1149) treat it
 * as.  In exact arithmetic one step rotates by 2 atan(|W| dt / 2) about W. */
static void SUF(quat_step)(REAL *q, V3 W, REAL dt)
{
    const REAL h = (REAL)0.5 * dt;
    const REAL w = q[0], x = q[1], y = q[2], z = q[3];
    const REAL nw = w - h * ((W.x * x + W.y * y) + W.z * z);
    const REAL nx = x + h * (w * W.x + (W.y * z - W.z * y));
    const REAL ny = y + h * (w * W.y + (W.z * x - W.x * z));
    const REAL nz = z + h * (w * W.z + (W.x * y - W.y * x));
    const REAL inv = (REAL)1 / SQRT(((nw * nw + nx * nx) + ny * ny) + nz * nz);
    q[0] = nw * inv;
    q[1] = nx * inv;
    q[2] = ny * inv;
    q[3] = nz * inv;
}

/* ---------------------------------------------------------------- shapes, support */
typedef struct {
    int kind;       /* ORA_SHAPE_BOX / ORA_SHAPE_SPHERE */
    V3 center;      /* Position used for the first GJK direction */
    V3 verts[8];    /* world-space corners (boxes) */
    REAL radius;    /* spheres */
} SUF(Shape);
#define Shape SUF(Shape)

/* GetCubeSupport / GetFloorSupport (nans.cpp:409-430, 440-461): strict '>' argmax,
 * starting from -FLT_MAX and the zero vector (a NaN direction yields (0,0,0)). */
static V3 SUF(box_support)(const V3 *verts, V3 d)
{
    REAL maxd = (REAL)(-FLT_MAX);
    V3 r = mk3(0, 0, 0);
    for (int k = 0; k < 8; ++k) {
        REAL dist = vdot(verts[k], d);
        if (dist > maxd) {
            maxd = dist;
            r = verts[k];
        }
    }
    return r;
}

/* GetSphereSupport (nans.cpp:432-438): P + r * normalize(d) */
static V3 SUF(sphere_support)(V3 c, REAL r, V3 d) { return vadd(c, vmul(vnormalize(d), r)); }

static V3 SUF(shape_support)(const Shape *s, V3 d)
{
    OC(s->kind == ORA_SHAPE_BOX ? OC_SUP_BOX : OC_SUP_SPH, 1);
    return s->kind == ORA_SHAPE_BOX ? SUF(box_support)(s->verts, d) : SUF(sphere_support)(s->center, s->radius, d);
}

/* CalculateSupport (nans.cpp:463-519): S_A(d) - S_B(-1*d) */
static Vert SUF(calc_support)(const Shape *a, const Shape *b, V3 d)
{
    Vert v;
    v.SupA = SUF(shape_support)(a, d);
    v.SupB = SUF(shape_support)(b, vmul(d, (REAL)-1));
    v.P = vsub(v.SupA, v.SupB);
    return v;
}

/* ---------------------------------------------------------------- GJK */
enum { SUF(kNo) = 0, SUF(kFound) = 1, SUF(kStill) = 2 };

/* Trace of ora_trace_pair (NULL buffers: off). */
static ora_gjk_rec *SUF(g_tg);
static int SUF(g_tgcap), SUF(g_tgn);
static ora_epa_rec *SUF(g_te);
static int SUF(g_tecap), SUF(g_ten), SUF(g_tovf);
static V3 SUF(g_tdir);
static int SUF(g_tdir_set), SUF(g_tpushed);

static void SUF(put3)(double *d, V3 v)
{
    d[0] = (double)v.x;
    d[1] = (double)v.y;
    d[2] = (double)v.z;
}

static void SUF(erase)(Vert *s, int *n, int idx)
{
    for (int i = idx; i < *n - 1; ++i)
        s[i] = s[i + 1];
    --*n;
}

/* ClosestPointOnLine (nans.cpp:539-562) */
static V3 SUF(closest_point_on_line)(V3 A, V3 B, REAL *U, REAL *V)
{
    V3 seg = vnormalize(vsub(B, A));
    REAL len = vlength(seg);
    *V = vdot(vneg(A), seg) / len;
    *U = vdot(B, seg) / len;
    if (*U <= (REAL)0)
        return B;
    else if (*V <= (REAL)0)
        return A;
    return vadd(vmul(A, *U), vmul(B, *V));
}

/* TripleCross (nans.cpp:564-569): B*(C.A) - A*(C.B) */
static V3 SUF(triple_cross)(V3 A, V3 B, V3 C) { return vsub(vmul(B, vdot(C, A)), vmul(A, vdot(C, B))); }

/* EvolveSimplex (nans.cpp:571-769). */
static int SUF(evolve_simplex)(const Shape *a, const Shape *b, Vert *S, int *n, int fp64)
{
    V3 dir = vnormalize(vsub(b->center, a->center));
    if (*n >= 0 && *n <= 4)
        OC(OC_EV0 + *n, 1);
    switch (*n) {
    case 0:
        break;
    case 1:
        dir = vmul(dir, (REAL)-1);
        break;
    case 2: {
        REAL U = 0, V = 0;
        V3 cp = SUF(closest_point_on_line)(S[0].P, S[1].P, &U, &V);
        if (V <= (REAL)0) {
            SUF(erase)(S, n, 1);
            dir = vneg(cp);
        } else if (U <= (REAL)0) {
            SUF(erase)(S, n, 0);
            dir = vneg(cp);
        } else {
            dir = vneg(cp);
        }
    } break;
    case 3: {
        V3 n2o = vneg(S[0].P);
        V3 e1 = vsub(S[1].P, S[0].P);
        V3 e2 = vsub(S[2].P, S[0].P);
        V3 tn = vcross(e1, e2);
        V3 e1n = vcross(e1, tn);
        V3 e2n = vcross(tn, e2);
        if (vdot(e2n, n2o) > (REAL)0) {
            if (vdot(e2, n2o) > (REAL)0) {
                dir = SUF(triple_cross)(e2, n2o, e2);
                SUF(erase)(S, n, 1);
            } else if (vdot(e1, n2o) > (REAL)0) {
                dir = SUF(triple_cross)(e1, n2o, e1);
                SUF(erase)(S, n, 2);
            } else {
                dir = n2o;
                SUF(erase)(S, n, 2);
                SUF(erase)(S, n, 1);
            }
        } else {
            if (vdot(e1n, n2o) > (REAL)0) {
                if (vdot(e1, n2o) > (REAL)0) {
                    dir = SUF(triple_cross)(e1, n2o, e1);
                    SUF(erase)(S, n, 2);
                } else {
                    dir = n2o;
                    SUF(erase)(S, n, 2);
                    SUF(erase)(S, n, 1);
                }
            } else if (vdot(tn, n2o) > (REAL)0) {
                dir = tn;
            } else {
                dir = vneg(tn);
                Vert t = S[1];
                S[1] = S[2];
                S[2] = t;
            }
        }
    } break;
    case 4: {
        V3 DA = vsub(S[0].P, S[3].P);
        V3 DB = vsub(S[1].P, S[3].P);
        V3 DC = vsub(S[2].P, S[3].P);
        V3 D0 = vmul(S[3].P, (REAL)-1);
        V3 abd = vcross(DA, DB);
        V3 bcd = vcross(DB, DC);
        V3 cad = vcross(DC, DA);
        if (vdot(abd, D0) > (REAL)0) {
            SUF(erase)(S, n, 2);
            dir = abd;
        } else if (vdot(bcd, D0) > (REAL)0) {
            SUF(erase)(S, n, 0);
            dir = bcd;
        } else if (vdot(cad, D0) > (REAL)0) {
            SUF(erase)(S, n, 1);
            dir = cad;
        } else {
            return SUF(kFound);
        }
    } break;
    default:
        fprintf(stderr, "oracle: simplex with %d vertices\n", *n);
        return SUF(kNo);
    }
    /* termination test (nans.cpp:755-759); binary64 in the FP64-islands mode */
    OC(OC_EVLEN, 1);
    if (SUF(g_tg)) {
        SUF(g_tdir) = dir;
        SUF(g_tdir_set) = 1;
    }
    if (fp64) {
        double dx = (double)dir.x, dy = (double)dir.y, dz = (double)dir.z;
        if (sqrt((dx * dx + dy * dy) + dz * dz) <= (double)0.0001f)
            return SUF(kNo);
    } else if (vlength(dir) <= (REAL)0.0001f) {
        return SUF(kNo);
    }
    /* AddSupport (nans.cpp:521-537): the vertex is pushed whatever the test says */
    Vert v = SUF(calc_support)(a, b, dir);
    S[(*n)++] = v;
    if (SUF(g_tg))
        SUF(g_tpushed) = 1;
    return vdot(dir, v.P) >= (REAL)0 ? SUF(kStill) : SUF(kNo);
}

/* evolve_simplex, recorded when a trace is on */
static int SUF(evolve_call)(const Shape *a, const Shape *b, Vert *S, int *n, int fp64)
{
    if (!SUF(g_tg))
        return SUF(evolve_simplex)(a, b, S, n, fp64);
    ora_gjk_rec r;
    memset(&r, 0, sizeof(r));
    r.n_in = *n;
    for (int i = 0; i < *n && i < 4; ++i)
        SUF(put3)(r.s_in[i], S[i].P);
    SUF(g_tdir_set) = 0;
    SUF(g_tpushed) = 0;
    int res = SUF(evolve_simplex)(a, b, S, n, fp64);
    r.n_out = *n;
    for (int i = 0; i < *n && i < 5; ++i)
        SUF(put3)(r.s_out[i], S[i].P);
    r.result = res == SUF(kFound) ? 1 : (res == SUF(kStill) ? 2 : 0);
    if (SUF(g_tdir_set))
        SUF(put3)(r.dir, SUF(g_tdir));
    r.pushed = SUF(g_tpushed);
    if (r.pushed) {
        SUF(put3)(r.sup_a, S[*n - 1].SupA);
        SUF(put3)(r.sup_b, S[*n - 1].SupB);
    }
    if (SUF(g_tgn) < SUF(g_tgcap))
        SUF(g_tg)[SUF(g_tgn)++] = r;
    else
        SUF(g_tovf) = 1;
    return res;
}

/* ---------------------------------------------------------------- EPA */
typedef struct { Vert A, B, C; V3 N; } SUF(Tri);   /* triangle, nans.h:270-277 */
#define Tri SUF(Tri)
typedef struct { Vert A, B; } SUF(Edge);           /* edge, nans.h:263-268 */
#define Edge SUF(Edge)

typedef struct {
    Tri *t;
    int nt, ct;
    Edge *e;
    int ne, ce;
} SUF(Poly);
#define Poly SUF(Poly)

/* PushTriangle (nans.cpp:291-322) */
static void SUF(push_triangle)(Poly *p, Vert A, Vert B, Vert C)
{
    if (p->nt == p->ct) {
        p->ct = p->ct ? 2 * p->ct : 64;
        p->t = (Tri *)realloc(p->t, sizeof(Tri) * (size_t)p->ct);
    }
    Tri t;
    t.A = A;
    t.B = B;
    t.C = C;
    t.N = vnormalize(vcross(vsub(B.P, A.P), vsub(C.P, A.P)));
    if (vdot(t.A.P, t.N) < (REAL)0)
        t.N = vmul(t.N, (REAL)-1);
    p->t[p->nt++] = t;
}

/* PushEdge (nans.cpp:232-266): cancel the first reversed edge (vertex == compares P) */
static void SUF(push_edge)(Poly *p, Vert A, Vert B)
{
    for (int i = 0; i < p->ne; ++i) {
        if (veq(p->e[i].A.P, B.P) && veq(p->e[i].B.P, A.P)) {
            memmove(&p->e[i], &p->e[i + 1], sizeof(Edge) * (size_t)(p->ne - i - 1));
            --p->ne;
            return;
        }
    }
    if (p->ne == p->ce) {
        p->ce = p->ce ? 2 * p->ce : 64;
        p->e = (Edge *)realloc(p->e, sizeof(Edge) * (size_t)p->ce);
    }
    p->e[p->ne].A = A;
    p->e[p->ne].B = B;
    ++p->ne;
}

/* Barycentric (nans.cpp:771-785) */
static void SUF(barycentric)(V3 P, V3 A, V3 B, V3 C, REAL *U, REAL *V, REAL *W)
{
    V3 v0 = vsub(B, A), v1 = vsub(C, A), v2 = vsub(P, A);
    REAL d00 = vdot(v0, v0);
    REAL d01 = vdot(v0, v1);
    REAL d11 = vdot(v1, v1);
    REAL d20 = vdot(v2, v0);
    REAL d21 = vdot(v2, v1);
    REAL denom = d00 * d11 - d01 * d01;
    *V = (d11 * d20 - d01 * d21) / denom;
    *W = (d00 * d21 - d01 * d20) / denom;
    *U = (REAL)1 - *V - *W;
}

static int SUF(is_valid)(REAL f) { return !(isnan(f) || isinf(f)); } /* IsValid, nans.cpp:4-17 */

typedef struct {
    V3 point_a, point_b, normal;
    double depth;
} SUF(EpaOut);
#define EpaOut SUF(EpaOut)

/* Binary64 distance of a face plane to the origin, for the FP64-islands final
 * face choice. */
static double SUF(face_dist64)(const Tri *t)
{
    double ax = t->A.P.x, ay = t->A.P.y, az = t->A.P.z;
    double ux = (double)t->B.P.x - ax, uy = (double)t->B.P.y - ay, uz = (double)t->B.P.z - az;
    double vx = (double)t->C.P.x - ax, vy = (double)t->C.P.y - ay, vz = (double)t->C.P.z - az;
    double nx = uy * vz - vy * uz, ny = uz * vx - vz * ux, nz = ux * vy - vx * uy;
    double inv = 1.0 / sqrt((nx * nx + ny * ny) + nz * nz);
    double d = ((nx * inv) * ax + (ny * inv) * ay) + (nz * inv) * az;
    return d >= 0.0 ? d : -d;
}

/* A face's distance in ResolveCollision's closest-face scan (nans.cpp:807-822):
 * the plane through A with the recomputed, unflipped normal, as an absolute value. */
static REAL SUF(scan_dist)(const Tri *t)
{
    V3 ab = vsub(t->B.P, t->A.P);
    V3 ac = vsub(t->C.P, t->A.P);
    V3 nrm = vnormalize(vcross(ab, ac));
    return gabs(vdot(nrm, t->A.P));
}

/* ResolveCollision (nans.cpp:787-904), literal: std::vector erase semantics,
 * closest face by recomputed unflipped normals, visible faces by recomputed
 * flipped normals. */
static int SUF(epa)(const Shape *a, const Shape *b, const Vert *S, const ora_np_params *np,
                    EpaOut *out, int32_t *st)
{
    Poly p;
    memset(&p, 0, sizeof(p));
    Vert A = S[0], B = S[1], C = S[2], D = S[3];
    SUF(push_triangle)(&p, A, B, C);
    SUF(push_triangle)(&p, A, C, D);
    SUF(push_triangle)(&p, A, D, B);
    SUF(push_triangle)(&p, B, D, C);
    int nverts = 4, made = 4, maxf = 4, maxe = 0, iters = 0;
    int result = ORA_EPA_ITER;
    int cur_iter = 0;
    OC(OC_EPA_E, 4);
    while (cur_iter++ <= np->max_epa_iterations) {
        ++iters;
        OC(OC_EPA_ITER, 1);
        if (p.nt == 0) {
            result = ORA_EPA_EMPTY;
            break;
        }
        REAL cur = gabs(vdot(p.t[0].N, p.t[0].A.P));
        int best = 0;
        for (int i = 0; i < p.nt; ++i) {
            REAL dist = SUF(scan_dist)(&p.t[i]);
            if (dist < cur) {
                cur = dist;
                best = i;
            }
        }
        V3 dir = p.t[best].N;
        Vert ns = SUF(calc_support)(a, b, dir);
        ++nverts;
        ora_epa_rec *tr = NULL;
        if (SUF(g_te)) {
            if (SUF(g_ten) < SUF(g_tecap) && p.nt <= ORA_TRACE_MAXF) {
                tr = &SUF(g_te)[SUF(g_ten)++];
                memset(tr, 0, sizeof(*tr));
                tr->nfaces = p.nt;
                for (int i = 0; i < p.nt; ++i) {
                    SUF(put3)(&tr->face[i][0], p.t[i].A.P);
                    SUF(put3)(&tr->face[i][3], p.t[i].B.P);
                    SUF(put3)(&tr->face[i][6], p.t[i].C.P);
                    SUF(put3)(&tr->face[i][9], p.t[i].N);
                }
                tr->best = best;
                tr->cur = (double)cur;
                tr->dist0 = (double)gabs(vdot(p.t[0].N, p.t[0].A.P));
                for (int i = 0; i < p.nt; ++i) tr->dist[i] = (double)SUF(scan_dist)(&p.t[i]);
                SUF(put3)(tr->dir, dir);
                SUF(put3)(tr->p_new, ns.P);
                SUF(put3)(tr->sa_new, ns.SupA);
                SUF(put3)(tr->sb_new, ns.SupB);
            } else {
                SUF(g_tovf) = 1;
            }
        }
        if (vdot(p.t[best].N, ns.P) - cur < (REAL)np->epa_tolerance) {
            double depth = (double)cur;
            if (np->fp64_islands) {
                /* final face choice and depth in binary64 */
                double best64 = SUF(face_dist64)(&p.t[0]);
                int bi = 0;
                for (int i = 0; i < p.nt; ++i) {
                    double d64 = SUF(face_dist64)(&p.t[i]);
                    if (d64 < best64) {
                        best64 = d64;
                        bi = i;
                    }
                }
                if (bi != best) {
                    best = bi;
                    if (bi == 0) {
                        cur = gabs(vdot(p.t[0].N, p.t[0].A.P));
                    } else {
                        V3 ab = vsub(p.t[bi].B.P, p.t[bi].A.P);
                        V3 ac = vsub(p.t[bi].C.P, p.t[bi].A.P);
                        cur = gabs(vdot(vnormalize(vcross(ab, ac)), p.t[bi].A.P));
                    }
                }
                depth = best64;
            }
            Tri *t = &p.t[best];
            REAL U, V, W;
            SUF(barycentric)(vmul(t->N, cur), t->A.P, t->B.P, t->C.P, &U, &V, &W);
            if (tr) {
                tr->converged = 1;
                tr->best = best;
                tr->cur = (double)cur;
                tr->bary[0] = (double)U;
                tr->bary[1] = (double)V;
                tr->bary[2] = (double)W;
                SUF(put3)(&tr->face_sup[0][0], t->A.SupA);
                SUF(put3)(&tr->face_sup[0][3], t->A.SupB);
                SUF(put3)(&tr->face_sup[1][0], t->B.SupA);
                SUF(put3)(&tr->face_sup[1][3], t->B.SupB);
                SUF(put3)(&tr->face_sup[2][0], t->C.SupA);
                SUF(put3)(&tr->face_sup[2][3], t->C.SupB);
                tr->nfaces_out = p.nt;
            }
            if (fabs(U) > 1.0 || fabs(V) > 1.0 || fabs(W) > 1.0) {
                result = ORA_EPA_BARY;
            } else if (!SUF(is_valid)(U) || !SUF(is_valid)(V) || !SUF(is_valid)(W)) {
                result = ORA_EPA_NONFINITE;
            } else {
                out->point_a = vadd(vadd(vmul(t->A.SupA, U), vmul(t->B.SupA, V)), vmul(t->C.SupA, W));
                out->normal = vmul(t->N, (REAL)-1);
                out->point_b = vadd(vadd(vmul(t->A.SupB, U), vmul(t->B.SupB, V)), vmul(t->C.SupB, W));
                out->depth = depth;
                result = ORA_EPA_CONTACT;
            }
            OC(OC_EPA_FINAL, 1);
            break;
        }
        OC(OC_EPA_F, p.nt);
        int orig = 0;
        for (int i = 0; i < p.nt; ++orig) {
            Tri *t = &p.t[i];
            V3 temp = vsub(ns.P, t->A.P);
            V3 ab = vsub(t->B.P, t->A.P);
            V3 ac = vsub(t->C.P, t->A.P);
            V3 nrm = vnormalize(vcross(ab, ac));
            if (vdot(nrm, t->A.P) < (REAL)0)
                nrm = vmul(nrm, (REAL)-1);
            if (vdot(nrm, temp) > (REAL)0) {
                Vert ta = t->A, tb = t->B, tc = t->C;
                SUF(push_edge)(&p, ta, tb);
                if (p.ne > maxe) maxe = p.ne;
                SUF(push_edge)(&p, tb, tc);
                if (p.ne > maxe) maxe = p.ne;
                SUF(push_edge)(&p, tc, ta);
                if (p.ne > maxe) maxe = p.ne;
                memmove(&p.t[i], &p.t[i + 1], sizeof(Tri) * (size_t)(p.nt - i - 1));
                --p.nt;
                if (tr)
                    tr->visible[orig] = 1;
                continue;
            }
            ++i;
        }
        OC(OC_EPA_E, p.ne);
        for (int i = 0; i < p.ne; ++i) {
            SUF(push_triangle)(&p, ns, p.e[i].A, p.e[i].B);
            ++made;
        }
        p.ne = 0;
        if (p.nt > maxf) maxf = p.nt;
        if (tr)
            tr->nfaces_out = p.nt;
    }
    if (st) {
        st[ORA_STAT_EPA_ITERS] = iters;
        st[ORA_STAT_EPA_RESULT] = result;
        st[ORA_STAT_MAX_FACES] = maxf;
        st[ORA_STAT_MAX_EDGES] = maxe;
        st[ORA_STAT_NVERTS] = nverts;
        st[ORA_STAT_FACES_MADE] = made;
    }
    free(p.t);
    free(p.e);
    return result;
}

/* CheckCollision (nans.cpp:906-966): GJK then EPA.  Returns 1 on contact. */
static int SUF(check_collision)(const Shape *a, const Shape *b, const ora_np_params *np,
                                int *gjk_result, EpaOut *out, int32_t *st)
{
    Vert S[5];
    int n = 0, res = SUF(kStill);
    unsigned it = 0, calls = 0;
    while (res == SUF(kStill) && it++ <= (unsigned)np->max_gjk_iterations) {
        res = SUF(evolve_call)(a, b, S, &n, np->fp64_islands);
        ++calls;
    }
    int g = res == SUF(kFound) ? ORA_GJK_FOUND : (res == SUF(kNo) ? ORA_GJK_NO : ORA_GJK_CAP);
    if (st) {
        memset(st, 0, sizeof(int32_t) * ORA_NSTAT);
        st[ORA_STAT_GJK_CALLS] = (int32_t)calls;
        st[ORA_STAT_GJK_RESULT] = g;
    }
    *gjk_result = g;
    if (res == SUF(kFound))
        return SUF(epa)(a, b, S, np, out, st) == ORA_EPA_CONTACT;
    return 0;
}

/* ---------------------------------------------------------------- narrowphase API */
/* Box corners from (center, unit quaternion w,x,y,z, half-extents): Model columns
 * R(q)[:,k] * (2 h_k) and (c, 1), applied with the glm mat4*vec4 form. */
static void SUF(box_from_quat)(const REAL *c, const REAL *q, const REAL *h, V3 *verts)
{
    REAL w = q[0], x = q[1], y = q[2], z = q[3];
    REAL xx = x * x, yy = y * y, zz = z * z;
    REAL xy = x * y, xz = x * z, yz = y * z;
    REAL wx = w * x, wy = w * y, wz = w * z;
    REAL r00 = (REAL)1 - (REAL)2 * (yy + zz), r01 = (REAL)2 * (xy - wz), r02 = (REAL)2 * (xz + wy);
    REAL r10 = (REAL)2 * (xy + wz), r11 = (REAL)1 - (REAL)2 * (xx + zz), r12 = (REAL)2 * (yz - wx);
    REAL r20 = (REAL)2 * (xz - wy), r21 = (REAL)2 * (yz + wx), r22 = (REAL)1 - (REAL)2 * (xx + yy);
    REAL sx = (REAL)2 * h[0], sy = (REAL)2 * h[1], sz = (REAL)2 * h[2];
    M4 m;
    m.m[0] = r00 * sx; m.m[1] = r10 * sx; m.m[2] = r20 * sx; m.m[3] = 0;
    m.m[4] = r01 * sy; m.m[5] = r11 * sy; m.m[6] = r21 * sy; m.m[7] = 0;
    m.m[8] = r02 * sz; m.m[9] = r12 * sz; m.m[10] = r22 * sz; m.m[11] = 0;
    m.m[12] = c[0]; m.m[13] = c[1]; m.m[14] = c[2]; m.m[15] = 1;
    SUF(update_vertices)(&m, verts);
}

static void SUF(make_shape)(Shape *s, int32_t type, const REAL *c, const REAL *q, const REAL *h)
{
    s->kind = type;
    s->center = mk3(c[0], c[1], c[2]);
    if (type == ORA_SHAPE_BOX) {
        SUF(box_from_quat)(c, q, h, s->verts);
        s->radius = 0;
    } else {
        s->radius = h[0];
    }
}

int SUF(ora_narrowphase)(int64_t n,
                         const int32_t *type_a, const REAL *pos_a, const REAL *quat_a, const REAL *half_a,
                         const int32_t *type_b, const REAL *pos_b, const REAL *quat_b, const REAL *half_b,
                         const ora_np_params *p,
                         int32_t *gjk, int32_t *contact, REAL *normal, double *depth,
                         REAL *point_a, REAL *point_b, int32_t *stats)
{
    if (n < 0 || !p || !type_a || !type_b || !pos_a || !pos_b || !quat_a || !quat_b || !half_a || !half_b)
        return -1;
    for (int64_t i = 0; i < n; ++i) {
        Shape a, b;
        if ((type_a[i] != ORA_SHAPE_BOX && type_a[i] != ORA_SHAPE_SPHERE) ||
            (type_b[i] != ORA_SHAPE_BOX && type_b[i] != ORA_SHAPE_SPHERE))
            return -1;
        SUF(make_shape)(&a, type_a[i], pos_a + 3 * i, quat_a + 4 * i, half_a + 3 * i);
        SUF(make_shape)(&b, type_b[i], pos_b + 3 * i, quat_b + 4 * i, half_b + 3 * i);
        EpaOut o;
        memset(&o, 0, sizeof(o));
        int g = 0;
        int c = SUF(check_collision)(&a, &b, p, &g, &o, stats ? stats + ORA_NSTAT * i : NULL);
        if (gjk) gjk[i] = g;
        if (contact) contact[i] = c;
        if (normal) { normal[3 * i] = o.normal.x; normal[3 * i + 1] = o.normal.y; normal[3 * i + 2] = o.normal.z; }
        if (depth) depth[i] = o.depth;
        if (point_a) { point_a[3 * i] = o.point_a.x; point_a[3 * i + 1] = o.point_a.y; point_a[3 * i + 2] = o.point_a.z; }
        if (point_b) { point_b[3 * i] = o.point_b.x; point_b[3 * i + 1] = o.point_b.y; point_b[3 * i + 2] = o.point_b.z; }
    }
    return 0;
}

int SUF(ora_trace_pair)(int32_t type_a, const REAL *pos_a, const REAL *quat_a, const REAL *half_a,
                        int32_t type_b, const REAL *pos_b, const REAL *quat_b, const REAL *half_b,
                        const ora_np_params *p, ora_gjk_rec *g, int32_t gcap, int32_t *ng,
                        ora_epa_rec *e, int32_t ecap, int32_t *ne, int32_t *gjk, int32_t *epa,
                        REAL *normal, REAL *point_a, REAL *point_b)
{
    if (!p || !pos_a || !pos_b || !quat_a || !quat_b || !half_a || !half_b || !g || !e || gcap < 0 || ecap < 0 ||
        !ng || !ne || !gjk || !epa || !normal || !point_a || !point_b)
        return -1;
    if ((type_a != ORA_SHAPE_BOX && type_a != ORA_SHAPE_SPHERE) ||
        (type_b != ORA_SHAPE_BOX && type_b != ORA_SHAPE_SPHERE))
        return -1;
    Shape a, b;
    SUF(make_shape)(&a, type_a, pos_a, quat_a, half_a);
    SUF(make_shape)(&b, type_b, pos_b, quat_b, half_b);
    SUF(g_tg) = g;
    SUF(g_tgcap) = gcap;
    SUF(g_tgn) = 0;
    SUF(g_te) = e;
    SUF(g_tecap) = ecap;
    SUF(g_ten) = 0;
    SUF(g_tovf) = 0;
    EpaOut o;
    memset(&o, 0, sizeof(o));
    int32_t st[ORA_NSTAT];
    int gr = 0;
    SUF(check_collision)(&a, &b, p, &gr, &o, st);
    SUF(g_tg) = NULL;
    SUF(g_te) = NULL;
    *ng = SUF(g_tgn);
    *ne = SUF(g_ten);
    *gjk = gr;
    *epa = gr == ORA_GJK_FOUND ? st[ORA_STAT_EPA_RESULT] : ORA_EPA_NONE;
    normal[0] = o.normal.x; normal[1] = o.normal.y; normal[2] = o.normal.z;
    point_a[0] = o.point_a.x; point_a[1] = o.point_a.y; point_a[2] = o.point_a.z;
    point_b[0] = o.point_b.x; point_b[1] = o.point_b.y; point_b[2] = o.point_b.z;
    return SUF(g_tovf) ? 1 : 0;
}

/* ---------------------------------------------------------------- world step */
typedef struct {
    V3 P, V, F, A, W, T, S;
    REAL q[4];      /* orientation in quaternion mode (w, x, y, z) */
    REAL mass, moi;
    M4 model;
    V3 verts[8];
} SUF(Body);
#define Body SUF(Body)

typedef struct {
    int nb, ns;
    Body *box;      /* Cubes[] */
    Body *sph;      /* Spheres[] */
    Body floor;     /* Floor (a box, nans.cpp:1666-1678) */
    int quat;       /* 1: bodies carry quaternions (opt-in), 0: Euler degrees */
} SUF(World);
#define World SUF(World)

typedef struct {
    int32_t type, ia, ib;
    V3 PA, PB, N;
    REAL dln, dlt1, dlt2;
} SUF(Contact);
#define Contact SUF(Contact)

static void SUF(rebuild_model)(Body *b, int quat)
{
    b->model = quat ? SUF(model_from_quat)(b->P, b->q, b->S) : SUF(model_from_pose)(b->P, b->A, b->S);
    SUF(update_vertices)(&b->model, b->verts);
}

/* MovementFunction (nans.cpp:63-70): (1/m)(F + m*g*(0,-1,0) - damping*v), 1/m in double */
static V3 SUF(movement)(V3 v, V3 F, REAL m, const ora_step_params *p)
{
#if ORA_F32
    REAL inv = (REAL)(1.0 / (double)m);
#else
    REAL inv = (REAL)1 / m;
#endif
    V3 g = vmul(mk3(0, -1, 0), m * (REAL)p->gravity);
    return vmul(vsub(vadd(F, g), vmul(v, (REAL)p->damping)), inv);
}

/* RotationFunction (nans.cpp:72-78): (1/I)(tau - damping*w) */
static V3 SUF(rotation)(V3 w, V3 T, REAL I, const ora_step_params *p)
{
    REAL inv = (REAL)1 / I;
    return vmul(vsub(T, vmul(w, (REAL)p->damping)), inv);
}

/* RK4 (nans.cpp:50-61) with the force held constant over the step */
static V3 SUF(rk4)(int rot, REAL dt, V3 y0, V3 F, REAL m, const ora_step_params *p)
{
#define ORA_F(y) (rot ? SUF(rotation)((y), F, m, p) : SUF(movement)((y), F, m, p))
    V3 k1 = vmul(ORA_F(y0), dt);
    V3 k2 = vmul(ORA_F(vadd(y0, vdiv(k1, (REAL)2))), dt);
    V3 k3 = vmul(ORA_F(vadd(y0, vdiv(k2, (REAL)2))), dt);
    V3 k4 = vmul(ORA_F(vadd(y0, k3)), dt);
#undef ORA_F
#if ORA_F32
    const REAL sixth = 1.0f / 6.0f;
#else
    const REAL sixth = 1.0 / 6.0;
#endif
    V3 sum = vadd(vadd(vadd(k1, vmul(k2, (REAL)2)), vmul(k3, (REAL)2)), k4);
    return vadd(y0, vmul(sum, sixth));
}

/* IntegrateForces (nans.cpp:974-1018): cubes then spheres; forces cleared */
static void SUF(integrate_forces)(World *w, const ora_step_params *p)
{
    REAL dt = (REAL)p->dt;
    OC(OC_BODY_BOX, w->nb);
    OC(OC_BODY_SPH, w->ns);
    for (int i = 0; i < w->nb; ++i) {
        Body *b = &w->box[i];
        b->V = SUF(rk4)(0, dt, b->V, b->F, b->mass, p);
        b->W = SUF(rk4)(1, dt, b->W, b->T, b->moi, p);
        SUF(update_vertices)(&b->model, b->verts);
        b->F = mk3(0, 0, 0);
        b->T = mk3(0, 0, 0);
    }
    for (int i = 0; i < w->ns; ++i) {
        Body *b = &w->sph[i];
        b->V = SUF(rk4)(0, dt, b->V, b->F, b->mass, p);
        b->W = SUF(rk4)(1, dt, b->W, b->T, b->moi, p);
        b->F = mk3(0, 0, 0);
        b->T = mk3(0, 0, 0);
    }
    SUF(update_vertices)(&w->floor.model, w->floor.verts);
}

static void SUF(body_shape)(Shape *s, const Body *b, int kind)
{
    s->kind = kind;
    s->center = b->P;
    if (kind == ORA_SHAPE_BOX) {
        memcpy(s->verts, b->verts, sizeof(s->verts));
        s->radius = 0;
    } else {
        s->radius = b->S.x;
    }
}

/* The HIP step's bounding-sphere cull (counting mode only): keep the pair unless
 * the centre distance exceeds R_A + R_B plus a rounding margin (binary32 only). */
static float SUF(shape_radius)(const Shape *s)
{
    if (s->kind != ORA_SHAPE_BOX)
        return (float)s->radius;
    float r2 = 0.0f;
    for (int k = 0; k < 8; ++k) {
        float dx = (float)s->verts[k].x - (float)s->center.x, dy = (float)s->verts[k].y - (float)s->center.y;
        float dz = (float)s->verts[k].z - (float)s->center.z;
        float d2 = (dx * dx + dy * dy) + dz * dz;
        r2 = d2 > r2 ? d2 : r2;
    }
    return sqrtf(r2);
}

static int SUF(cull_keep)(const Shape *a, const Shape *b)
{
    float ra = SUF(shape_radius)(a), rb = SUF(shape_radius)(b);
    float ax = (float)a->center.x, ay = (float)a->center.y, az = (float)a->center.z;
    float bx = (float)b->center.x, by = (float)b->center.y, bz = (float)b->center.z;
    float dx = bx - ax, dy = by - ay, dz = bz - az;
    float dist2 = (dx * dx + dy * dy) + dz * dz;
    float mag = fabsf(ax) + fabsf(ay) + fabsf(az) + fabsf(bx) + fabsf(by) + fabsf(bz);
    float lim = (ra + rb) * 1.0001f + 1e-4f * (1.0f + mag + ra + rb);
    return !(dist2 > lim * lim);
}

/* Body pair of a contact type: (A, kindA, B, kindB). */
static void SUF(pair_bodies)(World *w, int type, int ia, int ib, Body **A, int *ka, Body **B, int *kb)
{
    switch (type) {
    case ORA_CC: *A = &w->box[ia]; *ka = ORA_SHAPE_BOX; *B = &w->box[ib]; *kb = ORA_SHAPE_BOX; break;
    case ORA_CS: *A = &w->box[ia]; *ka = ORA_SHAPE_BOX; *B = &w->sph[ib]; *kb = ORA_SHAPE_SPHERE; break;
    case ORA_CF: *A = &w->box[ia]; *ka = ORA_SHAPE_BOX; *B = &w->floor; *kb = ORA_SHAPE_BOX; break;
    case ORA_SS: *A = &w->sph[ia]; *ka = ORA_SHAPE_SPHERE; *B = &w->sph[ib]; *kb = ORA_SHAPE_SPHERE; break;
    default: *A = &w->sph[ia]; *ka = ORA_SHAPE_SPHERE; *B = &w->floor; *kb = ORA_SHAPE_BOX; break;
    }
}

typedef struct {
    Contact *c;
    int n, cap;
    int64_t tests, found;
    int maxf, maxe, maxv;
    int64_t itercap, empty;
} SUF(ContactList);
#define ContactList SUF(ContactList)

static void SUF(test_pair)(World *w, ContactList *cl, int type, int ia, int ib, const ora_step_params *p,
                           int32_t *stat_out)
{
    Body *A, *B;
    int ka, kb;
    SUF(pair_bodies)(w, type, ia, ib, &A, &ka, &B, &kb);
    Shape sa, sb;
    SUF(body_shape)(&sa, A, ka);
    SUF(body_shape)(&sb, B, kb);
    OC(OC_PAIRS, 1);
    if (SUF(g_cull) && !SUF(cull_keep)(&sa, &sb))
        return;
    OC(OC_CAND, 1);
    EpaOut o;
    int32_t st[ORA_NSTAT];
    int g = 0;
    int hit = SUF(check_collision)(&sa, &sb, &p->np, &g, &o, st);
    if (stat_out)
        memcpy(stat_out, st, sizeof(st));
    cl->tests++;
    if (g == ORA_GJK_FOUND) {
        cl->found++;
        if (st[ORA_STAT_MAX_FACES] > cl->maxf) cl->maxf = st[ORA_STAT_MAX_FACES];
        if (st[ORA_STAT_MAX_EDGES] > cl->maxe) cl->maxe = st[ORA_STAT_MAX_EDGES];
        if (st[ORA_STAT_NVERTS] > cl->maxv) cl->maxv = st[ORA_STAT_NVERTS];
        if (st[ORA_STAT_EPA_RESULT] == ORA_EPA_ITER) cl->itercap++;
        if (st[ORA_STAT_EPA_RESULT] == ORA_EPA_EMPTY) cl->empty++;
    }
    if (!hit)
        return;
    /* the duplicate scan of nans.cpp:1374-1393 never matches: every (type, i, j)
     * is tested once per step, so the pair is appended unconditionally */
    if (cl->n == cl->cap) {
        cl->cap = cl->cap ? 2 * cl->cap : 256;
        cl->c = (Contact *)realloc(cl->c, sizeof(Contact) * (size_t)cl->cap);
    }
    Contact *c = &cl->c[cl->n++];
    memset(c, 0, sizeof(*c));
    c->type = type;
    c->ia = ia;
    c->ib = ib;
    c->PA = o.point_a;
    c->PB = o.point_b;
    c->N = o.normal;
}

/* DetectCollisions (nans.cpp:1351-1536): CC(i<j), CF, SF, CS (cube-major), SS(i<j) */
static void SUF(detect_collisions)(World *w, ContactList *cl, const ora_step_params *p)
{
    for (int i = 0; i < w->nb; ++i)
        for (int j = i + 1; j < w->nb; ++j)
            SUF(test_pair)(w, cl, ORA_CC, i, j, p, NULL);
    for (int i = 0; i < w->nb; ++i)
        SUF(test_pair)(w, cl, ORA_CF, i, i, p, NULL);
    for (int s = 0; s < w->ns; ++s)
        SUF(test_pair)(w, cl, ORA_SF, s, s, p, NULL);
    for (int i = 0; i < w->nb; ++i)
        for (int s = 0; s < w->ns; ++s)
            SUF(test_pair)(w, cl, ORA_CS, i, s, p, NULL);
    for (int s = 0; s < w->ns; ++s)
        for (int t = s + 1; t < w->ns; ++t)
            SUF(test_pair)(w, cl, ORA_SS, s, t, p, NULL);
}

/* Constraint (nans.cpp:1020-1329) on a copy of the contact */
/* The contact frame of Constraint (nans.cpp:1115-1124): unit normal (with the
 * zero-normal fallback) and lever arms from the bodies' positions. */
static void SUF(contact_frame)(World *w, const Contact *pair, V3 *N, V3 *R1, V3 *R2)
{
    Body *A, *B;
    int ka, kb;
    SUF(pair_bodies)(w, pair->type, pair->ia, pair->ib, &A, &ka, &B, &kb);
    *N = vnormalize(pair->N);
    if (veq(*N, mk3(0, 0, 0)))
        *N = vnormalize(vsub(B->P, A->P)); /* nans.cpp:1115-1119 */
    *R1 = vsub(pair->PA, A->P);
    *R2 = vsub(pair->PB, B->P);
}

/* Constraint's solve and apply (nans.cpp:1125-1329) for a contact frame; A and
 * B's velocities are updated (B's only when two_sided).  dl[3] receives the
 * applied impulses. */
static void SUF(constraint_solve)(Body *A, Body *B, int two_sided, V3 N, V3 R1, V3 R2,
                                  const ora_step_params *p, REAL Cf, REAL Cr, REAL *dl, REAL *sums);

/* sums (may be NULL): the accumulators DLNormalSum, DLTangent1Sum, DLTangent2Sum
 * at entry (NULL: zero, as every contact_pair starts, nans.cpp:1365) and at exit;
 * the opt-in warm start carries them from one step to the next. */
static void SUF(constraint)(World *w, Contact *pair, const ora_step_params *p, REAL *sums)
{
    Body *A, *B;
    int ka, kb;
    SUF(pair_bodies)(w, pair->type, pair->ia, pair->ib, &A, &ka, &B, &kb);
    V3 N, R1, R2;
    SUF(contact_frame)(w, pair, &N, &R1, &R2);
    int two_sided = pair->type == ORA_CC || pair->type == ORA_CS || pair->type == ORA_SS;
    REAL dl[3];
    SUF(constraint_solve)(A, B, two_sided, N, R1, R2, p, (REAL)p->friction, (REAL)p->restitution, dl, sums);
    pair->dln = dl[0];
    pair->dlt1 = dl[1];
    pair->dlt2 = dl[2];
}

static void SUF(constraint_solve)(Body *A, Body *B, int two_sided, V3 N, V3 R1, V3 R2,
                                  const ora_step_params *p, REAL Cf, REAL Cr, REAL *dl, REAL *sums)
{
    V3 PosA = A->P, PosB = B->P;
    REAL InvI1 = (REAL)1 / A->moi, InvI2 = (REAL)1 / B->moi;
    REAL InvM1 = (REAL)1 / A->mass, InvM2 = (REAL)1 / B->mass;
    V3 V1 = A->V, V2 = B->V, W1 = A->W, W2 = B->W;
    V3 T1;
    if (N.x >= (REAL)0.57735f)
        T1 = mk3(N.y, -N.x, 0);
    else
        T1 = mk3(0, N.z, -N.y);
    /* glm::normalize(T1) result discarded (nans.cpp:1133) */
    V3 T2 = vcross(N, T1);
    REAL Depth = vdot(vsub(vadd(PosA, R1), vadd(PosB, R2)), N);
    V3 RN1 = vcross(R1, N);
    V3 RN2 = vcross(R2, N);
    REAL JMJn = InvM1 + InvM2;
    JMJn += InvI1 * vdot(RN1, RN1) - InvI2 * vdot(vneg(RN2), vneg(RN2));
    JMJn = (REAL)1 / JMJn;
    V3 dVn = vsub(vsub(vadd(V1, vcross(W1, N)), V2), vcross(W2, N));
    REAL JdVn = vdot(dVn, N);
    REAL Beta = (REAL)p->beta;
    REAL Bb = -Beta / (REAL)p->dt * Depth + Cr * JdVn;
    V3 R1T1 = vcross(R1, T1);
    V3 R2T1 = vcross(R2, T1);
    V3 R1T2 = vcross(R1, T2);
    V3 R2T2 = vcross(R2, T2);
    REAL JMJt1 = InvM1 + InvM2;
    JMJt1 += InvI1 * vdot(R1T1, R1T1) - InvI2 * vdot(vneg(R2T1), vneg(R2T1));
    JMJt1 = (REAL)1 / JMJt1;
    REAL JMJt2 = InvM1 + InvM2;
    JMJt2 += InvI1 * vdot(R1T2, R1T2) - InvI2 * vdot(vneg(R2T2), vneg(R2T2));
    JMJt2 = (REAL)1 / JMJt2;
    V3 dVt1 = vsub(vsub(vadd(V1, vcross(W1, T1)), V2), vcross(W2, T1));
    REAL JdVt1 = vdot(dVt1, T1);
    V3 dVt2 = vsub(vsub(vadd(V1, vcross(W1, T2)), V2), vcross(W2, T2));
    REAL JdVt2 = vdot(dVt2, T2);

    REAL sumN = 0, sumT1 = 0, sumT2 = 0, dln = 0, dlt1 = 0, dlt2 = 0;
    if (sums) {
        sumN = sums[0];
        sumT1 = sums[1];
        sumT2 = sums[2];
    }
    int iter = p->solver_iterations;
    OC(OC_CONSTRAINT, 1);
    OC(OC_SOLVER_IT, iter);
    while (iter--) {
        REAL LambdaN = (-JdVn + Bb) * JMJn;
        REAL old = sumN;
        sumN += LambdaN;
        if (sumN < (REAL)0)
            sumN = 0;
        dln = sumN - old;

        REAL LambdaT1 = (-JdVt1) * JMJt1;
        REAL oldT1 = sumT1;
        sumT1 += LambdaT1;
#if ORA_F32
        REAL MaxLT1 = (REAL)(sqrt(2.0) * (double)Cf * (double)sumN);
#else
        REAL MaxLT1 = sqrt(2.0) * Cf * sumN;
#endif
        if (sumT1 < -MaxLT1) sumT1 = -MaxLT1;
        if (sumT1 > MaxLT1) sumT1 = MaxLT1;
        dlt1 = sumT1 - oldT1;

        REAL LambdaT2 = (-JdVt2) * JMJt2;
        REAL oldT2 = sumT2;
        sumT2 += LambdaT2;
#if ORA_F32
        REAL MaxLT2 = (REAL)(sqrt(2.0) * (double)Cf * (double)sumN);
#else
        REAL MaxLT2 = sqrt(2.0) * Cf * sumN;
#endif
        if (sumT2 < -MaxLT2) sumT2 = -MaxLT2;
        if (sumT2 > MaxLT2) sumT2 = MaxLT2;
        dlt2 = sumT2 - oldT2;
    }
    dl[0] = dln;
    dl[1] = dlt1;
    dl[2] = dlt2;
    if (sums) {
        sums[0] = sumN;
        sums[1] = sumT1;
        sums[2] = sumT2;
    }
    V3 LinearImpulse = vmul(N, dln);
    V3 LinearImpulseT1 = vmul(T1, dlt1);
    V3 LinearImpulseT2 = vmul(T2, dlt2);
    V3 AngularI1 = vmul(RN1, dln);
    V3 AngularI2 = vmul(RN2, dln);
    V3 AngularI1T1 = vmul(R1T1, dlt1);
    V3 AngularI2T1 = vmul(R2T1, dlt1);
    V3 AngularI1T2 = vmul(R1T2, dlt2);
    V3 AngularI2T2 = vmul(R2T2, dlt2);
    A->V = vadd(A->V, vmul(LinearImpulse, InvM1));
    if (two_sided) B->V = vsub(B->V, vmul(LinearImpulse, InvM2));
    A->W = vadd(A->W, vmul(AngularI1, InvI1));
    if (two_sided) B->W = vsub(B->W, vmul(AngularI2, InvI2));
    A->V = vadd(A->V, vmul(LinearImpulseT1, InvM1));
    if (two_sided) B->V = vsub(B->V, vmul(LinearImpulseT1, InvM2));
    A->W = vadd(A->W, vmul(AngularI1T1, InvI1));
    if (two_sided) B->W = vsub(B->W, vmul(AngularI2T1, InvI2));
    A->V = vadd(A->V, vmul(LinearImpulseT2, InvM1));
    if (two_sided) B->V = vsub(B->V, vmul(LinearImpulseT2, InvM2));
    A->W = vadd(A->W, vmul(AngularI1T2, InvI1));
    if (two_sided) B->W = vsub(B->W, vmul(AngularI2T2, InvI2));
}

/* IntegrateVelocities (nans.cpp:1331-1349) followed by the Model rebuild of the
 * render section (nans.cpp:1870-1881, 1913-1941) that feeds the next step. */
static void SUF(integrate_velocities)(World *w, const ora_step_params *p)
{
    REAL dt = (REAL)p->dt;
    for (int i = 0; i < w->nb; ++i) {
        Body *b = &w->box[i];
        b->P = vadd(b->P, vmul(b->V, dt));
        if (w->quat) SUF(quat_step)(b->q, b->W, dt);
        else b->A = vadd(b->A, vmul(b->W, dt));
        SUF(update_vertices)(&b->model, b->verts);
    }
    for (int i = 0; i < w->ns; ++i) {
        Body *b = &w->sph[i];
        b->P = vadd(b->P, vmul(b->V, dt));
        if (w->quat) SUF(quat_step)(b->q, b->W, dt);
        else b->A = vadd(b->A, vmul(b->W, dt));
    }
}

static void SUF(rebuild_models)(World *w)
{
    SUF(rebuild_model)(&w->floor, 0);
    for (int i = 0; i < w->nb; ++i)
        SUF(rebuild_model)(&w->box[i], w->quat);
}

static void SUF(world_load)(World *w, int32_t nb, int32_t ns, const REAL *pos, const REAL *vel,
                            const REAL *angvel, const REAL *angles, const REAL *force, const REAL *torque,
                            const REAL *scale, const REAL *mass, const REAL *moi, const ora_step_params *p,
                            const REAL *quat)
{
    w->nb = nb;
    w->ns = ns;
    w->quat = quat != NULL;
    w->box = (Body *)calloc((size_t)(nb > 0 ? nb : 1), sizeof(Body));
    w->sph = (Body *)calloc((size_t)(ns > 0 ? ns : 1), sizeof(Body));
    for (int k = 0; k < nb + ns; ++k) {
        Body *b = k < nb ? &w->box[k] : &w->sph[k - nb];
        b->P = mk3(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]);
        b->V = mk3(vel[3 * k], vel[3 * k + 1], vel[3 * k + 2]);
        b->W = mk3(angvel[3 * k], angvel[3 * k + 1], angvel[3 * k + 2]);
        b->A = mk3(angles[3 * k], angles[3 * k + 1], angles[3 * k + 2]);
        b->F = force ? mk3(force[3 * k], force[3 * k + 1], force[3 * k + 2]) : mk3(0, 0, 0);
        b->T = torque ? mk3(torque[3 * k], torque[3 * k + 1], torque[3 * k + 2]) : mk3(0, 0, 0);
        b->S = mk3(scale[3 * k], scale[3 * k + 1], scale[3 * k + 2]);
        b->mass = mass[k];
        b->moi = moi[k];
        for (int c = 0; c < 4; ++c) b->q[c] = quat ? quat[4 * k + c] : (REAL)(c == 0);
    }
    memset(&w->floor, 0, sizeof(w->floor));
    w->floor.P = mk3((REAL)p->floor_pos[0], (REAL)p->floor_pos[1], (REAL)p->floor_pos[2]);
    w->floor.S = mk3((REAL)p->floor_scale[0], (REAL)p->floor_scale[1], (REAL)p->floor_scale[2]);
    w->floor.mass = (REAL)p->floor_mass;
    w->floor.moi = (REAL)p->floor_moi;
    SUF(rebuild_models)(w);
}

static void SUF(world_store)(const World *w, REAL *pos, REAL *vel, REAL *angvel, REAL *angles,
                             REAL *force, REAL *torque, REAL *quat)
{
    for (int k = 0; k < w->nb + w->ns; ++k) {
        const Body *b = k < w->nb ? &w->box[k] : &w->sph[k - w->nb];
        if (quat)
            for (int c = 0; c < 4; ++c) quat[4 * k + c] = b->q[c];
        pos[3 * k] = b->P.x; pos[3 * k + 1] = b->P.y; pos[3 * k + 2] = b->P.z;
        vel[3 * k] = b->V.x; vel[3 * k + 1] = b->V.y; vel[3 * k + 2] = b->V.z;
        angvel[3 * k] = b->W.x; angvel[3 * k + 1] = b->W.y; angvel[3 * k + 2] = b->W.z;
        angles[3 * k] = b->A.x; angles[3 * k + 1] = b->A.y; angles[3 * k + 2] = b->A.z;
        if (force) { force[3 * k] = b->F.x; force[3 * k + 1] = b->F.y; force[3 * k + 2] = b->F.z; }
        if (torque) { torque[3 * k] = b->T.x; torque[3 * k + 1] = b->T.y; torque[3 * k + 2] = b->T.z; }
    }
}

static int SUF(params_ok)(const ora_step_params *p)
{
    return p && p->solver_iterations >= 0 && p->np.max_gjk_iterations >= 0 && p->np.max_epa_iterations >= 0;
}

#if ORA_F32
#define ORA_CONTACT_T ora_contact_f32
#else
#define ORA_CONTACT_T ora_contact_f64
#endif

/* Canonical index of a pair (DetectCollisions order, nans.cpp:1355-1535). */
static int64_t SUF(pair_index)(int type, int ia, int ib, int64_t nb, int64_t ns)
{
    const int64_t ncc = nb * (nb - 1) / 2;
    switch (type) {
    case ORA_CC: return (int64_t)ia * (2 * nb - ia - 1) / 2 + (ib - ia - 1);
    case ORA_CF: return ncc + ia;
    case ORA_SF: return ncc + nb + ia;
    case ORA_CS: return ncc + nb + ns + (int64_t)ia * ns + ib;
    default: return ncc + nb + ns + nb * ns + (int64_t)ia * (2 * ns - ia - 1) / 2 + (ib - ia - 1);
    }
}

/* Opt-in warm start (no counterpart in the reference, whose DetectCollisions
 * leaves "Warm starting?" as a TODO on a cache hit, nans.cpp:1374-1393): the
 * previous step's contacts as (canonical pair index, three accumulators), sorted
 * by index, at most cap of them; a contact whose pair is listed starts from those
 * accumulators. */
typedef struct {
    int64_t *p;
    REAL *sums;
    int32_t *n;
    int32_t cap;
} SUF(Warm);
#define Warm SUF(Warm)

static int SUF(world_run)(int32_t nb, int32_t ns,
                          REAL *pos, REAL *vel, REAL *angvel, REAL *angles, REAL *quat,
                          REAL *force, REAL *torque,
                          const REAL *scale, const REAL *mass, const REAL *moi,
                          const ora_step_params *p, int32_t nsteps,
                          ORA_CONTACT_T *contacts, int32_t max_contacts, int32_t *ncontacts,
                          int64_t *stats, const Warm *warm)
{
    if (nb < 0 || ns < 0 || nsteps < 0 || !SUF(params_ok)(p) || !pos || !vel || !angvel || !angles ||
        !scale || !mass || !moi)
        return -1;
    World w;
    SUF(world_load)(&w, nb, ns, pos, vel, angvel, angles, force, torque, scale, mass, moi, p, quat);
    ContactList cl;
    memset(&cl, 0, sizeof(cl));
    int64_t st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < nsteps; ++s) {
        cl.n = 0;
        SUF(integrate_forces)(&w, p);
        SUF(detect_collisions)(&w, &cl, p);
        /* SolveConstraints (nans.cpp:1538-1548): in list order, on copies */
        if (!warm) {
            for (int k = 0; k < cl.n; ++k)
                SUF(constraint)(&w, &cl.c[k], p, NULL);
        } else {
            const int m = cl.n < warm->cap ? cl.n : warm->cap;
            const size_t nc = (size_t)(cl.n > 0 ? cl.n : 1);
            int64_t *pk_ = (int64_t *)malloc(sizeof(int64_t) * nc);
            REAL *sums = (REAL *)calloc(3 * nc, sizeof(REAL));  /* each contact's accumulators */
            int j = 0;  /* both lists ascend: one merge pass finds the matches */
            for (int k = 0; k < cl.n; ++k) {
                const Contact *c = &cl.c[k];
                pk_[k] = SUF(pair_index)(c->type, c->ia, c->ib, nb, ns);
                while (j < *warm->n && warm->p[j] < pk_[k]) ++j;
                if (j < *warm->n && warm->p[j] == pk_[k])
                    for (int q = 0; q < 3; ++q) sums[3 * k + q] = warm->sums[3 * j + q];
            }
            for (int k = 0; k < cl.n; ++k)  /* in list order */
                SUF(constraint)(&w, &cl.c[k], p, sums + 3 * k);
            memcpy(warm->p, pk_, sizeof(int64_t) * (size_t)m);
            memcpy(warm->sums, sums, sizeof(REAL) * 3 * (size_t)m);
            *warm->n = m;
            free(pk_);
            free(sums);
        }
        SUF(integrate_velocities)(&w, p);
        SUF(rebuild_models)(&w);
        st[2] += cl.n;
    }
    st[0] = cl.tests;
    st[1] = cl.found;
    st[3] = cl.maxf;
    st[4] = cl.maxe;
    st[5] = cl.maxv;
    st[6] = cl.itercap;
    st[7] = cl.empty;
    if (stats)
        memcpy(stats, st, sizeof(st));
    if (contacts && ncontacts) {
        int nout = cl.n < max_contacts ? cl.n : max_contacts;
        for (int k = 0; k < nout; ++k) {
            ORA_CONTACT_T *o = &contacts[k];
            const Contact *c = &cl.c[k];
            o->type = c->type;
            o->index_a = c->ia;
            o->index_b = c->ib;
            o->pad = 0;
            o->point_a[0] = c->PA.x; o->point_a[1] = c->PA.y; o->point_a[2] = c->PA.z;
            o->point_b[0] = c->PB.x; o->point_b[1] = c->PB.y; o->point_b[2] = c->PB.z;
            o->normal[0] = c->N.x; o->normal[1] = c->N.y; o->normal[2] = c->N.z;
            o->dl_normal = c->dln;
            o->dl_tangent1 = c->dlt1;
            o->dl_tangent2 = c->dlt2;
        }
        *ncontacts = cl.n;
    } else if (ncontacts) {
        *ncontacts = cl.n;
    }
    SUF(world_store)(&w, pos, vel, angvel, angles, force, torque, quat);
    free(cl.c);
    free(w.box);
    free(w.sph);
    return 0;
}

int SUF(ora_world_run)(int32_t nb, int32_t ns,
                       REAL *pos, REAL *vel, REAL *angvel, REAL *angles,
                       REAL *force, REAL *torque,
                       const REAL *scale, const REAL *mass, const REAL *moi,
                       const ora_step_params *p, int32_t nsteps,
                       ORA_CONTACT_T *contacts, int32_t max_contacts, int32_t *ncontacts,
                       int64_t *stats)
{
    return SUF(world_run)(nb, ns, pos, vel, angvel, angles, NULL, force, torque, scale, mass, moi, p, nsteps,
                          contacts, max_contacts, ncontacts, stats, NULL);
}

int SUF(ora_world_run_quat)(int32_t nb, int32_t ns,
                            REAL *pos, REAL *vel, REAL *angvel, REAL *angles, REAL *quat,
                            REAL *force, REAL *torque,
                            const REAL *scale, const REAL *mass, const REAL *moi,
                            const ora_step_params *p, int32_t nsteps,
                            ORA_CONTACT_T *contacts, int32_t max_contacts, int32_t *ncontacts,
                            int64_t *stats)
{
    if (!quat) return -1;
    return SUF(world_run)(nb, ns, pos, vel, angvel, angles, quat, force, torque, scale, mass, moi, p, nsteps,
                          contacts, max_contacts, ncontacts, stats, NULL);
}

int SUF(ora_world_run_warm)(int32_t nb, int32_t ns,
                            REAL *pos, REAL *vel, REAL *angvel, REAL *angles, REAL *quat,
                            REAL *force, REAL *torque,
                            const REAL *scale, const REAL *mass, const REAL *moi,
                            const ora_step_params *p, int32_t nsteps,
                            ORA_CONTACT_T *contacts, int32_t max_contacts, int32_t *ncontacts,
                            int64_t *stats, int64_t *warm_pair, REAL *warm_sums, int32_t *warm_n, int32_t warm_cap)
{
    if (!warm_pair || !warm_sums || !warm_n || warm_cap < 0 || *warm_n < 0 || *warm_n > warm_cap) return -1;
    Warm wm;
    wm.p = warm_pair;
    wm.sums = warm_sums;
    wm.n = warm_n;
    wm.cap = warm_cap;
    return SUF(world_run)(nb, ns, pos, vel, angvel, angles, quat, force, torque, scale, mass, moi, p, nsteps,
                          contacts, max_contacts, ncontacts, stats, &wm);
}

/* ---------------------------------------------------------------- frozen-manifold rollouts */
static Body *SUF(slot_body)(World *w, int slot)
{
    if (slot < 0) return &w->floor;
    return slot < w->nb ? &w->box[slot] : &w->sph[slot - w->nb];
}

#if ORA_F32
int SUF(ora_frozen_record)(int32_t nb, int32_t ns,
                           REAL *pos, REAL *vel, REAL *angvel, REAL *angles,
                           REAL *force, REAL *torque,
                           const REAL *scale, const REAL *mass, const REAL *moi,
                           const ora_step_params *p, int32_t nsteps,
                           ora_manifold *out, int64_t max_out, int32_t *step_count, int64_t *total)
{
    if (nb < 0 || ns < 0 || nsteps < 0 || !SUF(params_ok)(p) || !pos || !vel || !angvel || !angles ||
        !scale || !mass || !moi || !step_count || max_out < 0 || (max_out > 0 && !out))
        return -1;
    World w;
    SUF(world_load)(&w, nb, ns, pos, vel, angvel, angles, force, torque, scale, mass, moi, p, NULL);
    ContactList cl;
    memset(&cl, 0, sizeof(cl));
    int64_t k = 0;
    for (int s = 0; s < nsteps; ++s) {
        cl.n = 0;
        SUF(integrate_forces)(&w, p);
        SUF(detect_collisions)(&w, &cl, p);
        /* frames from the positions the solve reads (they do not change during it) */
        for (int c = 0; c < cl.n; ++c, ++k) {
            if (k >= max_out) continue;
            const Contact *ct = &cl.c[c];
            V3 N, R1, R2;
            SUF(contact_frame)(&w, ct, &N, &R1, &R2);
            ora_manifold *m = &out[k];
            m->type = ct->type;
            switch (ct->type) {
            case ORA_CC: m->slot_a = ct->ia; m->slot_b = ct->ib; break;
            case ORA_CS: m->slot_a = ct->ia; m->slot_b = nb + ct->ib; break;
            case ORA_CF: m->slot_a = ct->ia; m->slot_b = -1; break;
            case ORA_SS: m->slot_a = nb + ct->ia; m->slot_b = nb + ct->ib; break;
            default: m->slot_a = nb + ct->ia; m->slot_b = -1; break;
            }
            m->pad = 0;
            m->r1[0] = R1.x; m->r1[1] = R1.y; m->r1[2] = R1.z;
            m->r2[0] = R2.x; m->r2[1] = R2.y; m->r2[2] = R2.z;
            m->n[0] = N.x; m->n[1] = N.y; m->n[2] = N.z;
        }
        step_count[s] = cl.n;
        for (int c = 0; c < cl.n; ++c)  /* SolveConstraints in list order */
            SUF(constraint)(&w, &cl.c[c], p, NULL);
        SUF(integrate_velocities)(&w, p);
        SUF(rebuild_models)(&w);
    }
    if (total) *total = k;
    SUF(world_store)(&w, pos, vel, angvel, angles, force, torque, NULL);
    free(cl.c);
    free(w.box);
    free(w.sph);
    return k > max_out ? -2 : 0;
}
#endif

int SUF(ora_frozen_run)(int32_t nb, int32_t ns,
                        REAL *pos, REAL *vel, REAL *angvel, REAL *angles,
                        REAL *force, REAL *torque, const REAL *mass, const REAL *moi,
                        const ora_step_params *p, double friction, double restitution, int32_t nsteps,
                        const ora_manifold *m, const int32_t *step_count)
{
    if (nb < 0 || ns < 0 || nsteps < 0 || !SUF(params_ok)(p) || !pos || !vel || !angvel || !angles ||
        !mass || !moi || (nsteps > 0 && !step_count))
        return -1;
    World w;
    memset(&w, 0, sizeof(w));
    w.nb = nb;
    w.ns = ns;
    w.box = (Body *)calloc((size_t)(nb > 0 ? nb : 1), sizeof(Body));
    w.sph = (Body *)calloc((size_t)(ns > 0 ? ns : 1), sizeof(Body));
    for (int k = 0; k < nb + ns; ++k) {
        Body *b = SUF(slot_body)(&w, k);
        b->P = mk3(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]);
        b->V = mk3(vel[3 * k], vel[3 * k + 1], vel[3 * k + 2]);
        b->W = mk3(angvel[3 * k], angvel[3 * k + 1], angvel[3 * k + 2]);
        b->A = mk3(angles[3 * k], angles[3 * k + 1], angles[3 * k + 2]);
        b->F = force ? mk3(force[3 * k], force[3 * k + 1], force[3 * k + 2]) : mk3(0, 0, 0);
        b->T = torque ? mk3(torque[3 * k], torque[3 * k + 1], torque[3 * k + 2]) : mk3(0, 0, 0);
        b->mass = mass[k];
        b->moi = moi[k];
    }
    w.floor.P = mk3((REAL)p->floor_pos[0], (REAL)p->floor_pos[1], (REAL)p->floor_pos[2]);
    w.floor.mass = (REAL)p->floor_mass;
    w.floor.moi = (REAL)p->floor_moi;
    const REAL dt = (REAL)p->dt;
    int64_t k = 0;
    int rc = 0;
    for (int s = 0; s < nsteps && rc == 0; ++s) {
        for (int i = 0; i < nb + ns; ++i) {  /* IntegrateForces without the corner refresh */
            Body *b = SUF(slot_body)(&w, i);
            b->V = SUF(rk4)(0, dt, b->V, b->F, b->mass, p);
            b->W = SUF(rk4)(1, dt, b->W, b->T, b->moi, p);
            b->F = mk3(0, 0, 0);
            b->T = mk3(0, 0, 0);
        }
        const int64_t k0 = k;
        k += step_count[s];
        for (int c = 0; c < step_count[s]; ++c) {  /* SolveConstraints in list order */
            const ora_manifold *f = &m[k0 + c];
            if (f->slot_a < 0 || f->slot_a >= nb + ns || f->slot_b < -1 || f->slot_b >= nb + ns) {
                rc = -1;
                break;
            }
            REAL dl[3];
            SUF(constraint_solve)(SUF(slot_body)(&w, f->slot_a), SUF(slot_body)(&w, f->slot_b), f->slot_b >= 0,
                                  mk3(f->n[0], f->n[1], f->n[2]), mk3(f->r1[0], f->r1[1], f->r1[2]),
                                  mk3(f->r2[0], f->r2[1], f->r2[2]), p, (REAL)friction, (REAL)restitution, dl, NULL);
        }
        for (int i = 0; i < nb + ns; ++i) {  /* IntegrateVelocities */
            Body *b = SUF(slot_body)(&w, i);
            b->P = vadd(b->P, vmul(b->V, dt));
            b->A = vadd(b->A, vmul(b->W, dt));
        }
    }
    for (int k2 = 0; k2 < nb + ns; ++k2) {
        const Body *b = SUF(slot_body)(&w, k2);
        pos[3 * k2] = b->P.x; pos[3 * k2 + 1] = b->P.y; pos[3 * k2 + 2] = b->P.z;
        vel[3 * k2] = b->V.x; vel[3 * k2 + 1] = b->V.y; vel[3 * k2 + 2] = b->V.z;
        angvel[3 * k2] = b->W.x; angvel[3 * k2 + 1] = b->W.y; angvel[3 * k2 + 2] = b->W.z;
        angles[3 * k2] = b->A.x; angles[3 * k2 + 1] = b->A.y; angles[3 * k2 + 2] = b->A.z;
        if (force) { force[3 * k2] = b->F.x; force[3 * k2 + 1] = b->F.y; force[3 * k2 + 2] = b->F.z; }
        if (torque) { torque[3 * k2] = b->T.x; torque[3 * k2 + 1] = b->T.y; torque[3 * k2 + 2] = b->T.z; }
    }
    free(w.box);
    free(w.sph);
    return rc;
}

#if ORA_F32
int SUF(ora_world_pairs)(int32_t nb, int32_t ns,
                         const REAL *pos, const REAL *angles, const REAL *scale,
                         const ora_step_params *p,
                         int32_t *type, int32_t *index_a, int32_t *index_b, int32_t *stats)
{
    if (nb < 0 || ns < 0 || !SUF(params_ok)(p) || !pos || !angles || !scale)
        return -1;
    int n = nb + ns;
    REAL *zero = (REAL *)calloc((size_t)(3 * (n > 0 ? n : 1)), sizeof(REAL));
    REAL *ones = (REAL *)malloc(sizeof(REAL) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; ++k)
        ones[k] = 1;
    World w;
    SUF(world_load)(&w, nb, ns, pos, zero, zero, angles, NULL, NULL, scale, ones, ones, p, NULL);
    ContactList cl;
    memset(&cl, 0, sizeof(cl));
    int64_t k = 0;
#define ORA_PAIR(T, I, J)                                                          \
    do {                                                                           \
        type[k] = (T);                                                             \
        index_a[k] = (I);                                                          \
        index_b[k] = (J);                                                          \
        SUF(test_pair)(&w, &cl, (T), (I), (J), p, stats ? stats + ORA_NSTAT * k : NULL); \
        ++k;                                                                       \
    } while (0)
    for (int i = 0; i < nb; ++i)
        for (int j = i + 1; j < nb; ++j)
            ORA_PAIR(ORA_CC, i, j);
    for (int i = 0; i < nb; ++i)
        ORA_PAIR(ORA_CF, i, i);
    for (int s = 0; s < ns; ++s)
        ORA_PAIR(ORA_SF, s, s);
    for (int i = 0; i < nb; ++i)
        for (int s = 0; s < ns; ++s)
            ORA_PAIR(ORA_CS, i, s);
    for (int s = 0; s < ns; ++s)
        for (int t = s + 1; t < ns; ++t)
            ORA_PAIR(ORA_SS, s, t);
#undef ORA_PAIR
    free(cl.c);
    free(w.box);
    free(w.sph);
    free(zero);
    free(ones);
    return 0;
}
#endif

void SUF(ora_box_vertices_quat)(const REAL *pos, const REAL *quat, const REAL *scale, REAL *out)
{
    M4 m = SUF(model_from_quat)(mk3(pos[0], pos[1], pos[2]), quat, mk3(scale[0], scale[1], scale[2]));
    V3 v[8];
    SUF(update_vertices)(&m, v);
    for (int k = 0; k < 8; ++k) {
        out[3 * k] = v[k].x;
        out[3 * k + 1] = v[k].y;
        out[3 * k + 2] = v[k].z;
    }
}

void SUF(ora_box_vertices)(const REAL *pos, const REAL *angles, const REAL *scale, REAL *out)
{
    M4 m = SUF(model_from_pose)(mk3(pos[0], pos[1], pos[2]), mk3(angles[0], angles[1], angles[2]),
                                mk3(scale[0], scale[1], scale[2]));
    V3 v[8];
    SUF(update_vertices)(&m, v);
    for (int k = 0; k < 8; ++k) {
        out[3 * k] = v[k].x;
        out[3 * k + 1] = v[k].y;
        out[3 * k + 2] = v[k].z;
    }
}

void SUF(ora_rk4)(REAL *v, REAL *w, const REAL *force, const REAL *torque, REAL mass, REAL moi,
                  const ora_step_params *p)
{
    V3 V = mk3(v[0], v[1], v[2]), W = mk3(w[0], w[1], w[2]);
    V3 F = mk3(force[0], force[1], force[2]), T = mk3(torque[0], torque[1], torque[2]);
    V = SUF(rk4)(0, (REAL)p->dt, V, F, mass, p);
    W = SUF(rk4)(1, (REAL)p->dt, W, T, moi, p);
    v[0] = V.x; v[1] = V.y; v[2] = V.z;
    w[0] = W.x; w[1] = W.y; w[2] = W.z;
}
