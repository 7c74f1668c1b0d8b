/*
 * nans_oracle.h -- CPU oracle for the nans_projekat simulation step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a literal C restatement of the
 * reference's physics block (code/nans.cpp) used as the checker for the
 * HIP implementation.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never
 * links, loads or calls it.
 *
 * The same source (nans_oracle_impl.h) is compiled twice:
 *   *_f32 : binary32 arithmetic with the reference's double islands
 *           (code/nans.cpp:68 (real32)(1.0/Mass), :1199/:1214 sqrt(2)*Cf*Sum),
 *           built -O2 -ffp-contract=off -fno-fast-math on x86-64 SSE, which
 *           reproduces the reference's g++ -O0 SSE scalar arithmetic
 *           (code/build.sh:5-14: no -O flag, no FMA).
 *   *_f64 : the same algorithm in binary64 (the "FP64 oracle").
 *
 * Vector arithmetic follows glm's published source semantics (glm is an
 * un-vendored dependency of the reference, code/nans.h:6-10, version not
 * pinned; the forms used are those of glm 0.9.8/0.9.9):
 *   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   cross(a,b)    = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
 *   normalize(v)  = v * (1 / sqrt(dot(v,v)))
 *   length(v)     = sqrt(dot(v,v))
 *   abs(x)        = x >= 0 ? x : -x
 *   mat4 * vec4   = (c0*x + c1*y) + (c2*z + c3*w)
 *   translate, rotate (axis-angle), scale, radians as in matrix_transform.inl.
 * The binary32 sin/cos of rotate are the correctly rounded values (the binary64
 * results rounded to binary32); glibc's sinf/cosf agree with them except for
 * about 1% of arguments, where they are one ulp off.  Built with ORA_LIBM_SINF
 * the oracle calls the host libm instead (lib/libnans_oracle_libm.so), to
 * measure what that choice moves.
 *
 * EvolveSimplex and ResolveCollision are checked call by call through
 * ora_trace_pair_* against properties the geometry fixes (tests/
 * test_gjk_epa_invariants.py), the closest-face tie rule on exact distance ties
 * (tests/test_epa_ties.py) and SolveConstraints' list order by impulse-momentum
 * algebra (tests/test_solver_order.py), each with planted mistakes in tests/
 * test_oracle_mutants.py.
 */
#ifndef NANS_ORACLE_H
#define NANS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Shape kinds of the narrowphase entry point. */
#define ORA_SHAPE_BOX 0
#define ORA_SHAPE_SPHERE 1

/* Pair kinds, numbered as collision_type in code/nans.h:73-89. */
#define ORA_CC 0
#define ORA_CS 1
#define ORA_CF 2
#define ORA_SS 3
#define ORA_SF 4

/* GJK outcome (evolve_result, code/nans.h:91-96) plus the iteration cap. */
#define ORA_GJK_NO 0
#define ORA_GJK_FOUND 1
#define ORA_GJK_CAP 2

/* EPA outcome codes. */
#define ORA_EPA_NONE 0        /* EPA not run (GJK did not find an intersection) */
#define ORA_EPA_CONTACT 1     /* converged, contact written (nans.cpp:828-866) */
#define ORA_EPA_BARY 2        /* |barycentric| > 1 -> no contact (nans.cpp:836-840) */
#define ORA_EPA_NONFINITE 3   /* NaN/inf barycentric -> no contact (nans.cpp:842-846) */
#define ORA_EPA_ITER 4        /* 65 iterations without convergence (nans.cpp:903) */
#define ORA_EPA_EMPTY 5       /* every face removed: the reference reads Triangle[0]
                                 of an empty vector here (undefined behaviour); the
                                 oracle reports no contact */

/* Per-pair statistics written by the narrowphase entry points. */
#define ORA_STAT_GJK_CALLS 0   /* EvolveSimplex calls */
#define ORA_STAT_GJK_RESULT 1  /* ORA_GJK_* */
#define ORA_STAT_EPA_ITERS 2   /* EPA loop iterations entered */
#define ORA_STAT_EPA_RESULT 3  /* ORA_EPA_* */
#define ORA_STAT_MAX_FACES 4   /* largest face count seen */
#define ORA_STAT_MAX_EDGES 5   /* largest horizon edge list */
#define ORA_STAT_NVERTS 6      /* polytope vertices at exit */
#define ORA_STAT_FACES_MADE 7  /* faces created in total */
#define ORA_NSTAT 8

typedef struct ora_np_params {
    int32_t max_gjk_iterations; /* MAX_GJK_ITERATIONS (64): the loop runs 65 calls */
    int32_t max_epa_iterations; /* MAX_EPA_ITERATIONS (64): the loop runs 65 times */
    float epa_tolerance;        /* MAX_EPA_ERROR (0.001f) */
    int32_t fp64_islands;       /* 0: reference arithmetic; 1: GJK length test and
                                   EPA final face choice / depth in binary64 */
} ora_np_params;

typedef struct ora_step_params {
    ora_np_params np;
    int32_t solver_iterations;  /* 70 in the reference (nans.cpp:1176) */
    float beta;                 /* 0.3f (nans.cpp:1152) */
    float restitution;          /* Cr, 0.1 (nans.cpp:1154) */
    float friction;             /* Cf, 0.1f (nans.cpp:1159) */
    float damping;              /* GLOBAL_FRICTION 1.5f (nans.h:46) */
    float gravity;              /* 9.81f (nans.cpp:68) */
    float dt;
    float floor_pos[3];         /* (1.2, -0.5, 1.0) (nans.cpp:1669) */
    float floor_scale[3];       /* (Size, 1, Size) = (100, 1, 100) (nans.cpp:1878) */
    float floor_mass;           /* 1e5 (nans.cpp:1676) */
    float floor_moi;            /* (m/12)*(2*Size*Size) (nans.cpp:1677) */
} ora_step_params;

/* One contact of the ordered contact list (contact_pair, code/nans.h:339-372). */
typedef struct ora_contact_f32 {
    int32_t type, index_a, index_b, pad;
    float point_a[3], point_b[3], normal[3];
    float dl_normal, dl_tangent1, dl_tangent2; /* impulses applied by Constraint */
} ora_contact_f32;

typedef struct ora_contact_f64 {
    int32_t type, index_a, index_b, pad;
    double point_a[3], point_b[3], normal[3];
    double dl_normal, dl_tangent1, dl_tangent2;
} ora_contact_f64;

/* Narrowphase over n explicit pairs.  Arrays are interleaved per pair:
 * pos [n][3], quat [n][4] (w,x,y,z), half [n][3] (spheres: half[0] = radius).
 * Box vertices are built as Model = [R(q)*2h | c] applied to the corners
 * (+-0.5,+-0.5,+-0.5,1) in the order of code/nans.cpp:399-406.
 * Outputs: gjk (ORA_GJK_*), contact (0/1), normal [n][3], depth (double),
 * point_a/point_b [n][3]; stats [n][ORA_NSTAT] may be NULL.
 * Returns 0, or -1 on bad arguments. */
int ora_narrowphase_f32(int64_t n,
                        const int32_t *type_a, const float *pos_a, const float *quat_a, const float *half_a,
                        const int32_t *type_b, const float *pos_b, const float *quat_b, const float *half_b,
                        const ora_np_params *p,
                        int32_t *gjk, int32_t *contact, float *normal, double *depth,
                        float *point_a, float *point_b, int32_t *stats);
int ora_narrowphase_f64(int64_t n,
                        const int32_t *type_a, const double *pos_a, const double *quat_a, const double *half_a,
                        const int32_t *type_b, const double *pos_b, const double *quat_b, const double *half_b,
                        const ora_np_params *p,
                        int32_t *gjk, int32_t *contact, double *normal, double *depth,
                        double *point_a, double *point_b, int32_t *stats);

/* Advance one world by nsteps.  The world holds nb boxes followed by ns spheres;
 * per-body arrays are interleaved [n][3] (n = nb + ns):
 *   pos, vel, angvel, angles (degrees, as code/nans.cpp:1915-1920 reads them),
 *   force, torque  (in/out; cleared by IntegrateForces, nans.cpp:92-110),
 *   scale          (boxes: Model scale = full extents; spheres: scale[0] = radius),
 *   mass, moi      [n].
 * The box Model is rebuilt from the pose at entry (the initial Model is built from
 * the initial pose) and after every step (nans.cpp:1913-1941).
 * If contacts != NULL the contact list of the LAST step is written there (up to
 * max_contacts entries) and its length to *ncontacts.  stats (may be NULL):
 *   [0] pair tests, [1] GJK found, [2] contacts, [3] max EPA faces, [4] max EPA edges,
 *   [5] max EPA vertices, [6] EPA iteration-cap exits, [7] EPA empty-polytope exits
 * accumulated over all steps.  Returns 0, or -1 on bad arguments. */
int ora_world_run_f32(int32_t nb, int32_t ns,
                      float *pos, float *vel, float *angvel, float *angles,
                      float *force, float *torque,
                      const float *scale, const float *mass, const float *moi,
                      const ora_step_params *p, int32_t nsteps,
                      ora_contact_f32 *contacts, int32_t max_contacts, int32_t *ncontacts,
                      int64_t *stats);
int ora_world_run_f64(int32_t nb, int32_t ns,
                      double *pos, double *vel, double *angvel, double *angles,
                      double *force, double *torque,
                      const double *scale, const double *mass, const double *moi,
                      const ora_step_params *p, int32_t nsteps,
                      ora_contact_f64 *contacts, int32_t max_contacts, int32_t *ncontacts,
                      int64_t *stats);

/* ora_world_run with the opt-in quaternion orientation (no counterpart in the
 * reference): quat [n][4] (w, x, y, z, unit) is the orientation in and out; the box
 * Model is [R(q) diag(scale) | pos] and IntegrateVelocities advances
 * q <- normalize(q + dt/2 (0, angvel) q) instead of the angles (left unchanged). */
int ora_world_run_quat_f32(int32_t nb, int32_t ns,
                           float *pos, float *vel, float *angvel, float *angles, float *quat,
                           float *force, float *torque,
                           const float *scale, const float *mass, const float *moi,
                           const ora_step_params *p, int32_t nsteps,
                           ora_contact_f32 *contacts, int32_t max_contacts, int32_t *ncontacts,
                           int64_t *stats);
int ora_world_run_quat_f64(int32_t nb, int32_t ns,
                           double *pos, double *vel, double *angvel, double *angles, double *quat,
                           double *force, double *torque,
                           const double *scale, const double *mass, const double *moi,
                           const ora_step_params *p, int32_t nsteps,
                           ora_contact_f64 *contacts, int32_t max_contacts, int32_t *ncontacts,
                           int64_t *stats);

/* ora_world_run with the opt-in warm start (no counterpart in the reference; its
 * DetectCollisions marks "Warm starting?" as a TODO on a cache hit,
 * code/nans.cpp:1374-1393).  warm_pair / warm_sums / *warm_n (in/out, capacity
 * warm_cap) list the previous step's contacts as canonical pair indices
 * (ascending) with their final accumulators DLNormalSum, DLTangent1Sum,
 * DLTangent2Sum; a contact of the same pair starts from those accumulators
 * instead of zero, everything else as ora_world_run (or, with quat != NULL, as
 * ora_world_run_quat).  After each step the list holds that step's first
 * warm_cap contacts. */
int ora_world_run_warm_f32(int32_t nb, int32_t ns,
                           float *pos, float *vel, float *angvel, float *angles, float *quat,
                           float *force, float *torque,
                           const float *scale, const float *mass, const float *moi,
                           const ora_step_params *p, int32_t nsteps,
                           ora_contact_f32 *contacts, int32_t max_contacts, int32_t *ncontacts,
                           int64_t *stats, int64_t *warm_pair, float *warm_sums, int32_t *warm_n, int32_t warm_cap);
int ora_world_run_warm_f64(int32_t nb, int32_t ns,
                           double *pos, double *vel, double *angvel, double *angles, double *quat,
                           double *force, double *torque,
                           const double *scale, const double *mass, const double *moi,
                           const ora_step_params *p, int32_t nsteps,
                           ora_contact_f64 *contacts, int32_t max_contacts, int32_t *ncontacts,
                           int64_t *stats, int64_t *warm_pair, double *warm_sums, int32_t *warm_n,
                           int32_t warm_cap);

/* Every canonical pair of the world's current pose (DetectCollisions order,
 * nans.cpp:1355-1535): type/index_a/index_b and the per-pair stats of
 * CheckCollision, without solving or stepping.  Arrays have room for
 * npairs = nb(nb-1)/2 + nb + ns + nb*ns + ns(ns-1)/2 entries. */
int ora_world_pairs_f32(int32_t nb, int32_t ns,
                        const float *pos, const float *angles, const float *scale,
                        const ora_step_params *p,
                        int32_t *type, int32_t *index_a, int32_t *index_b, int32_t *stats);

/* Trace of one pair's CheckCollision, call by call, for the invariant tests of
 * EvolveSimplex (code/nans.cpp:571-769) and ResolveCollision (code/nans.cpp:
 * 787-904).  Values are widened to binary64 (exactly, from binary32). */
#define ORA_TRACE_MAXF 128
typedef struct ora_gjk_rec {
    int32_t n_in, n_out, result, pushed;   /* simplex sizes; result 0 no, 1 found, 2 still */
    double s_in[4][3];                     /* Minkowski points before the call */
    double s_out[5][3];                    /* and after it (a pushed vertex last) */
    double dir[3];                         /* direction of the length test and the support */
    double sup_a[3], sup_b[3];             /* supports of the pushed vertex */
} ora_gjk_rec;

typedef struct ora_epa_rec {
    int32_t nfaces, best, converged, nfaces_out;
    double cur;                            /* distance of the chosen face */
    double dir[3];                         /* its stored normal: the search direction */
    double p_new[3], sa_new[3], sb_new[3]; /* the new support */
    double bary[3];                        /* converged: U, V, W of the contact */
    double face_sup[3][6];                 /* converged: SupA, SupB of the face's A, B, C */
    double face[ORA_TRACE_MAXF][12];       /* faces before the iteration: A, B, C, stored N */
    int32_t visible[ORA_TRACE_MAXF];       /* faces the iteration removed */
    double dist0;                          /* face 0's distance from its stored normal (the scan's start) */
    double dist[ORA_TRACE_MAXF];           /* each face's distance as the closest-face scan computes it */
} ora_epa_rec;

/* CheckCollision of one pair (shapes as in ora_narrowphase_*) with its trace:
 * up to gcap GJK calls into g and ecap EPA iterations into e (counts in *ng,
 * *ne); the outcome in *gjk, *epa and the contact in normal, point_a, point_b.
 * Returns 0, 1 if a trace did not fit (a cap or ORA_TRACE_MAXF faces), -1 on bad
 * arguments. */
int ora_trace_pair_f32(int32_t type_a, const float *pos_a, const float *quat_a, const float *half_a,
                       int32_t type_b, const float *pos_b, const float *quat_b, const float *half_b,
                       const ora_np_params *p, ora_gjk_rec *g, int32_t gcap, int32_t *ng,
                       ora_epa_rec *e, int32_t ecap, int32_t *ne, int32_t *gjk, int32_t *epa,
                       float *normal, float *point_a, float *point_b);
int ora_trace_pair_f64(int32_t type_a, const double *pos_a, const double *quat_a, const double *half_a,
                       int32_t type_b, const double *pos_b, const double *quat_b, const double *half_b,
                       const ora_np_params *p, ora_gjk_rec *g, int32_t gcap, int32_t *ng,
                       ora_epa_rec *e, int32_t ecap, int32_t *ne, int32_t *gjk, int32_t *epa,
                       double *normal, double *point_a, double *point_b);

/* Box corners of the Model built from a pose (translate * rotX * rotY * rotZ * scale,
 * degrees), in the order of code/nans.cpp:399-406: out [8][3]. */
void ora_box_vertices_f32(const float *pos, const float *angles, const float *scale, float *out);
void ora_box_vertices_f64(const double *pos, const double *angles, const double *scale, double *out);
/* Box corners of the quaternion-mode Model [R(q) diag(scale) | pos], quat (w, x, y, z). */
void ora_box_vertices_quat_f32(const float *pos, const float *quat, const float *scale, float *out);
void ora_box_vertices_quat_f64(const double *pos, const double *quat, const double *scale, double *out);

/* Work counters of the step (see nans_oracle_impl.h): enable/disable (and reset),
 * optionally applying the HIP step's bounding-sphere cull; read 32 counters. */
#define ORA_NCOUNT 32
void ora_counting_f32(int32_t enable, int32_t cull);
void ora_counting_f64(int32_t enable, int32_t cull);
void ora_get_counts_f32(int64_t *out);
void ora_get_counts_f64(int64_t *out);

/* RK4 of code/nans.cpp:50-78 on one body: v, w in/out [3]. */
void ora_rk4_f32(float *v, float *w, const float *force, const float *torque,
                 float mass, float moi, const ora_step_params *p);
void ora_rk4_f64(double *v, double *w, const double *force, const double *torque,
                 double mass, double moi, const ora_step_params *p);

/* ---- Frozen-manifold rollouts: the function the step's adjoint differentiates.
 *
 * The adjoint holds the GJK/EPA outcome of every step fixed: which pairs are in
 * contact, in which order, and each contact's frame -- the unit normal N and the
 * lever arms R1 = PointA - PosA, R2 = PointB - PosB that Constraint derives from
 * the EPA points (nans.cpp:1115-1124).  Everything else of the step is a smooth
 * (piecewise, through the impulse clamps) function of the state and parameters:
 * RK4 (nans.cpp:50-78), Constraint with Depth = dot((PosA + R1) - (PosB + R2), N)
 * (nans.cpp:1139) and the impulse iterations, IntegrateVelocities.
 *
 * ora_frozen_record_f32 runs the binary32 step (as ora_world_run_f32) and writes
 * the frame of every contact of every step: out[] in step order, step_count[s]
 * contacts in step s.  Slots number the world's bodies boxes first (0..nb-1),
 * then spheres (nb..nb+ns-1); the floor is -1.  Returns 0, -1 on bad arguments,
 * -2 when more than max_out frames occurred (*total holds the count needed).
 *
 * ora_frozen_run_* replays nsteps steps with those frames instead of the
 * narrowphase.  With the recorded binary32 inputs, ora_frozen_run_f32 reproduces
 * ora_world_run_f32 bit for bit; ora_frozen_run_f64 is the same function in
 * binary64, on which finite differences of the adjoint are taken.  friction and
 * restitution are arguments (in binary64, so that they can be perturbed finely);
 * the binary32 replay rounds them to binary32 as the step does. */
typedef struct ora_manifold {
    int32_t type, slot_a, slot_b, pad;
    float r1[3], r2[3], n[3];
} ora_manifold;

int ora_frozen_record_f32(int32_t nb, int32_t ns,
                          float *pos, float *vel, float *angvel, float *angles,
                          float *force, float *torque,
                          const float *scale, const float *mass, const float *moi,
                          const ora_step_params *p, int32_t nsteps,
                          ora_manifold *out, int64_t max_out, int32_t *step_count, int64_t *total);
int ora_frozen_run_f32(int32_t nb, int32_t ns,
                       float *pos, float *vel, float *angvel, float *angles,
                       float *force, float *torque, const float *mass, const float *moi,
                       const ora_step_params *p, double friction, double restitution, int32_t nsteps,
                       const ora_manifold *m, const int32_t *step_count);
int ora_frozen_run_f64(int32_t nb, int32_t ns,
                       double *pos, double *vel, double *angvel, double *angles,
                       double *force, double *torque, const double *mass, const double *moi,
                       const ora_step_params *p, double friction, double restitution, int32_t nsteps,
                       const ora_manifold *m, const int32_t *step_count);

#ifdef __cplusplus
}
#endif
#endif
