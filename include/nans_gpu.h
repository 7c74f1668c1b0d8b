/*
 * nans_gpu.h -- C ABI of the MI355X (gfx950) simulation step.
 *
 * Drop-in for the physics block of the reference's plugin entry point
 * SimUpdateAndRender (code/nans.h:396-397, code/nans.cpp:1719-1762), which the
 * platform layer binds with dlsym (code/sdl_nans.cpp:399-407).  The step is the
 * reference's IntegrateForces -> DetectCollisions -> SolveConstraints ->
 * IntegrateVelocities (code/nans.cpp:1758-1762) followed by the Model rebuild
 * that the reference performs in its render section (code/nans.cpp:1870-1881,
 * 1913-1941) and that feeds the next step's box vertices.
 *
 * Conventions
 *   - Every call returns int: NANS_OK (0) or a negative NANS_ERR_* code; the
 *     message of the last failure on the calling thread is nans_last_error().
 *     Nothing prints and nothing aborts (the reference only printf's,
 *     code/nans.cpp:752,1109,1324).
 *   - Vector arrays are structure-of-arrays component planes: a [3][n] array
 *     holds all x, then all y, then all z.
 *   - Calls are ordered on the context's stream and synchronise before they
 *     return.
 *   - One thread drives a context; a context owns one device.
 */
#ifndef NANS_GPU_H
#define NANS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NANS_OK 0
#define NANS_ERR_INVALID (-1)   /* bad argument */
#define NANS_ERR_HIP (-2)       /* HIP runtime error */
#define NANS_ERR_CAPACITY (-3)  /* input exceeds a documented capacity */
#define NANS_ERR_STATE (-4)     /* call not valid in the context's state */

/* Shape kinds of the narrowphase entry point and of bodies. */
#define NANS_SHAPE_BOX 0
#define NANS_SHAPE_SPHERE 1

/* Pair kinds, numbered as collision_type (code/nans.h:73-89). */
#define NANS_CC 0
#define NANS_CS 1
#define NANS_CF 2
#define NANS_SS 3
#define NANS_SF 4

/* GJK outcome (evolve_result, code/nans.h:91-96, plus the 65-call cap). */
#define NANS_GJK_NO 0
#define NANS_GJK_FOUND 1
#define NANS_GJK_CAP 2

/* EPA outcome per pair. */
#define NANS_EPA_NONE 0       /* EPA not run */
#define NANS_EPA_CONTACT 1    /* contact (code/nans.cpp:828-866) */
#define NANS_EPA_BARY 2       /* |barycentric| > 1 -> no contact (code/nans.cpp:836-840) */
#define NANS_EPA_NONFINITE 3  /* NaN/inf barycentric -> no contact (code/nans.cpp:842-846) */
#define NANS_EPA_ITER 4       /* 65 iterations without convergence (code/nans.cpp:903) */
#define NANS_EPA_EMPTY 5      /* every face removed; the reference indexes an empty vector */
#define NANS_EPA_OVERFLOW 6   /* polytope beyond 72 vertices / 192 faces / 64 horizon
                                 edges: flagged and counted, never silently wrong */

/* Parameters.  nans_default_params() fills the reference's constants:
 * 70 solver iterations (code/nans.cpp:1176), beta 0.3 (:1152), restitution 0.1
 * (:1154), friction 0.1 (:1159), damping GLOBAL_FRICTION 1.5 (code/nans.h:46),
 * gravity 9.81 (code/nans.cpp:68), 64 GJK / EPA iterations and EPA tolerance
 * 0.001 (code/nans.h:54-56), the floor box of code/nans.cpp:1666-1678 with
 * Model scale (100, 1, 100) (code/nans.cpp:1878-1879), dt 1/60. */
typedef struct nans_params {
    int32_t solver_iterations;
    float beta;
    float restitution;
    float friction;
    float damping;
    float gravity;
    float dt;
    int32_t max_gjk_iterations;
    int32_t max_epa_iterations;
    float epa_tolerance;
    int32_t fp64_islands;     /* 0: reference arithmetic (default).  1: GJK's
                                 termination length and EPA's final face choice
                                 and depth in binary64. */
    float floor_pos[3];
    float floor_scale[3];
    float floor_mass;
    float floor_moi;
    int32_t device;           /* HIP device ordinal */
} nans_params;

typedef struct nans_ctx nans_ctx;

/* Shapes of the narrowphase entry point (device pointers, n entries each). */
typedef struct nans_shapes_soa {
    const int32_t *type;   /* NANS_SHAPE_* */
    const float *pos;      /* [3][n] centre; also the GJK start direction */
    const float *quat;     /* [4][n] unit quaternion w,x,y,z (boxes) */
    const float *half;     /* [3][n] box half-extents; spheres: plane 0 = radius */
} nans_shapes_soa;

/* Narrowphase outputs (device pointers; any may be NULL). */
typedef struct nans_narrow_out {
    int32_t *gjk;          /* NANS_GJK_* */
    int32_t *contact;      /* 1 if EPA produced a contact */
    float *normal;         /* [3][n] contact normal (-face normal, code/nans.cpp:851) */
    double *depth;         /* EPA distance of the chosen face (binary64 in fp64_islands mode) */
    float *point_a;        /* [3][n] contact point on A (code/nans.cpp:848-855) */
    float *point_b;        /* [3][n] contact point on B (code/nans.cpp:862-865) */
    int32_t *status;       /* NANS_EPA_* */
} nans_narrow_out;

/* Fill p with the reference's constants (see nans_params). */
void nans_default_params(nans_params *p);

/* Message of the last failed call on this thread ("" if none). */
const char *nans_last_error(void);

/* Create / destroy a context on p->device (p may be NULL for the defaults).
 * Replaces the reference's one-time Init guarded by Memory->IsInitialized
 * (code/nans.cpp:1721-1729). */
int nans_create(nans_ctx **out, const nans_params *p);
int nans_destroy(nans_ctx *ctx);

/* Replace the parameters (dt, solver constants ...) used by later calls. */
int nans_set_params(nans_ctx *ctx, const nans_params *p);

/* Order later calls on this hipStream_t (NULL: the null stream). */
int nans_set_stream(nans_ctx *ctx, void *hip_stream);

/* CheckCollision (code/nans.cpp:906-966) on n independent pairs: GJK then EPA.
 * Box corners come from Model = [R(q) * 2h | centre] applied to (+-0.5, +-0.5,
 * +-0.5, 1) in the corner order of code/nans.cpp:399-406.  If n_overflow is not
 * NULL it receives the number of NANS_EPA_OVERFLOW pairs. */
int nans_narrowphase(nans_ctx *ctx, int64_t n, const nans_shapes_soa *a, const nans_shapes_soa *b,
                     const nans_narrow_out *out, int64_t *n_overflow);

/* nans_narrowphase without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as nans_narrowphase, except that at most
 * NANS_NARROWPHASE_NOALLOC_MAX_PAIRS pairs are accepted per call (more:
 * NANS_ERR_CAPACITY; split the call).  Its EPA queue lives in a static device
 * array of 26 MiB per device, held until the process ends and shared by every
 * call on that device: calls that overlap on one device (several threads or
 * streams at once) race on it. */
#define NANS_NARROWPHASE_NOALLOC_MAX_PAIRS 262144
int nans_narrowphase_noalloc(nans_ctx *ctx, int64_t n, const nans_shapes_soa *a, const nans_shapes_soa *b,
                             const nans_narrow_out *out, int64_t *n_overflow);

/* ------------------------------------------------------------------ worlds */

/* Bodies of a batch of worlds, host pointers.  World w owns bodies
 * [world_offsets[w], world_offsets[w+1]): its boxes first, then its spheres (the
 * reference's State->Cubes[] and State->Spheres[], code/nans.h:377-378; pair
 * indices are relative to those two spaces).  Each world has its own copy of the
 * static floor box of nans_params.  Vector arrays are [3][N] planes. */
typedef struct nans_bodies_soa {
    float *pos;            /* Position */
    float *vel;            /* V */
    float *angvel;         /* W */
    float *angles;         /* Angles, in degrees as the Model rebuild reads them */
    float *force;          /* Forces (may be NULL: zero); cleared by each step */
    float *torque;         /* Torque (may be NULL: zero); cleared by each step */
    const float *scale;    /* boxes: Model scale (full extents); spheres: plane 0 = Radius */
    const float *mass;     /* [N] */
    const float *moi;      /* [N] scalar moment of inertia */
    const int32_t *kind;   /* [N] NANS_SHAPE_* */
} nans_bodies_soa;

/* One contact of a world's ordered contact list (contact_pair, code/nans.h:339-372):
 * the pair as DetectCollisions lists it and the impulses Constraint applied. */
typedef struct nans_contact {
    int32_t type;          /* NANS_CC ... NANS_SF */
    int32_t index_a, index_b;  /* Cubes[]/Spheres[] indices; floor pairs repeat index_a */
    int32_t pad;
    float point_a[3], point_b[3], normal[3];
    float dl_normal, dl_tangent1, dl_tangent2;
} nans_contact;

/* Counters accumulated since the upload or nans_reset_stats. */
typedef struct nans_stats {
    int64_t pair_tests;        /* canonical pairs considered */
    int64_t candidates;        /* pairs that reached GJK */
    int64_t gjk_found;         /* GJK intersections */
    int64_t contacts;          /* EPA contacts */
    int64_t epa_overflow;      /* NANS_EPA_OVERFLOW pairs */
    int64_t capacity_overflow; /* steps whose candidates or contacts exceeded scratch */
} nans_stats;

/* Copy a batch of worlds to the device (replacing any previous batch).  Per-world
 * friction / restitution arrays are optional (NULL: the params' values); the
 * reference has one global Cf and Cr (code/nans.cpp:1154,1159).  At most 16384
 * bodies per world. */
int nans_upload_worlds(nans_ctx *ctx, int32_t nworlds, const int32_t *world_offsets, const nans_bodies_soa *bodies,
                       const float *world_friction, const float *world_restitution);

/* Copy pos, vel, angvel, angles, force, torque of every body back (non-NULL fields). */
int nans_download_state(nans_ctx *ctx, nans_bodies_soa *bodies);

/* Advance every world by nsteps with the plain path: one world per launch
 * sequence (the drop-in for the reference's physics block).
 * Both stepping calls (and nans_rollout) return NANS_ERR_CAPACITY when a step
 * had more pair candidates or contacts than its scratch holds; the batch is
 * then left as it was before the call (state, force, torque, quaternions and
 * statistics), with the warm-start lists emptied. */
int nans_step_plain(nans_ctx *ctx, int32_t nsteps);

/* Advance every world by nsteps with the batched path: one persistent launch per
 * step over all worlds (at most 512 bodies per world).  The step (work-counter
 * reset + step kernel) is captured once in a hipGraph and replayed nsteps times;
 * the graph is kept until the batch, the parameters or the tuning change.  With
 * the warm start on (nans_set_warm_start), the warm-started step kernel is
 * launched step by step instead. */
int nans_step_batched(nans_ctx *ctx, int32_t nsteps);

/* Device time of the last nans_step_batched call, from HIP events recorded on
 * the stream its kernels run on, and the launch shape it used. */
typedef struct nans_timing {
    int32_t steps;
    float elapsed_ms;
    int32_t threads_per_world;
    int32_t workgroups;
    int32_t lds_bytes;
    int32_t wave_world_bodies;  /* worlds of at most this many bodies ran one per wave */
    int32_t wave_worlds;        /* and there were this many */
} nans_timing;
int nans_batched_timing(nans_ctx *ctx, nans_timing *timing);

/* Overwrite pos, vel, angvel, angles (and force, torque; NULL = zero) of the
 * uploaded batch in place, keeping buffers and the captured graph.  The
 * quaternions of the opt-in orientation mode are set from the given angles, as
 * nans_upload_worlds sets them. */
int nans_set_state(nans_ctx *ctx, const nans_bodies_soa *bodies);

/* Orientation of both steps.  EULER (the default) is the reference's: Angles in
 * degrees advanced by Angles += dt * W (code/nans.cpp:1339) and read by the Model
 * rebuild as translate * rotX * rotY * rotZ * scale (code/nans.cpp:1913-1924).
 * QUATERNION is an opt-in mode with no counterpart in the reference (its README
 * lists quaternions as a TODO): each body carries a unit quaternion q = (w, x, y,
 * z), the box Model is [R(q) diag(scale) | pos], and IntegrateVelocities advances
 * q <- normalize(q + dt/2 (0, W) q), W read as the world-frame angular velocity in
 * rad/s (as the Constraint terms W x N read it).  Results differ from the
 * reference's by design.  Switching to QUATERNION takes each body's quaternion
 * from its current angles (R(q) = Rx Ry Rz); the angles are then left as they are.
 * nans_upload_worlds sets the quaternions from the uploaded angles;
 * nans_set_quaternions / nans_get_quaternions move them as [4][N] planes (w, x, y,
 * z).  The differentiable rollout requires EULER. */
#define NANS_ORIENT_EULER 0
#define NANS_ORIENT_QUATERNION 1
int nans_set_orientation_mode(nans_ctx *ctx, int32_t mode);
int nans_set_quaternions(nans_ctx *ctx, const float *quat);
int nans_get_quaternions(nans_ctx *ctx, float *quat);

/* Warm start of the plain step (opt-in; no counterpart in the reference, whose
 * DetectCollisions marks "Warm starting?" as a TODO where a contact pair of the
 * previous frame is found again, code/nans.cpp:1374-1393).  When enabled, each
 * contact whose pair (same kind and indices) was a contact in the previous step
 * starts its accumulators DLNormalSum, DLTangent1Sum, DLTangent2Sum from that
 * step's final values instead of zero; everything else, including the applied
 * impulse being the last iteration's delta (code/nans.cpp:1176-1238), is
 * unchanged.  At most 16 n + 64 contacts of a world carry over; the others start
 * from zero.  Enabling it, nans_set_state and nans_upload_worlds empty the
 * lists.  Results differ from the reference's by design.  The plain and the
 * batched step share the lists, so the two can alternate; the batched step then
 * launches step by step instead of replaying its graph.  The autotuner and the
 * differentiable rollout refuse the mode. */
int nans_set_warm_start(nans_ctx *ctx, int32_t enable);

/* Broadphase of the plain step.  The reference tests every pair of a world
 * (code/nans.cpp:1355-1535); the plain step keeps the pairs whose bounding
 * spheres can touch, found either by testing every canonical pair or with a
 * uniform grid.  Both give the same candidates in the same canonical order, so
 * results are identical; AUTO (the default) takes the grid for worlds of 1024
 * bodies or more.  The batched step always tests every pair of its (small)
 * worlds. */
#define NANS_BROADPHASE_AUTO 0
#define NANS_BROADPHASE_ALL_PAIRS 1
#define NANS_BROADPHASE_GRID 2
int nans_set_broadphase(nans_ctx *ctx, int32_t mode);

/* Autotuner of the batched step (no counterpart in the reference, which steps one
 * scene on one thread).  The batched step runs the larger worlds one per
 * workgroup and, after them, the worlds of at most `wave-world bound` bodies one
 * per wave (each wave of a workgroup stepping its own world).  The workgroup size
 * sets how many canonical pairs of a world are culled and GJK-tested at once, how
 * many EPA polytopes are in flight (one per wave) and the width of the solver's
 * dependency rounds; the bound trades that width for more worlds in flight.
 * nans_autotune_batched times nsteps steps of the uploaded batch for each
 * workgroup size (64, 128, 256, 512 threads) and bound (0, 16, 32, 64, 128
 * bodies; at 64 threads only 0), skipping combinations whose LDS does not fit or
 * that split the worlds as the previous bound did (ms = -1); it restores the
 * state, force/torque, quaternions and statistics, and keeps the fastest
 * combination for later nans_step_batched / nans_rollout calls, unless it is
 * less than 2 % faster than the default (which then stays: a few timed steps do
 * not rank shapes closer than that).  Results do not depend on either.  nans_set_batched_block forces a size (0: the default, 128
 * threads when no world holds more than 64 bodies, else 256);
 * nans_set_wave_worlds forces a bound (-1: the default, 0 = every world by a
 * whole workgroup; at most 512). */
typedef struct nans_tuning {
    int32_t threads_per_world;      /* chosen size */
    int32_t steps;                  /* steps timed per candidate */
    float ms_per_step[4];           /* 64, 128, 256, 512 threads per world, each at its best bound */
    int32_t wave_world_bodies;      /* chosen bound */
    float ms_per_step_wave[4][5];   /* [size][bound]: bounds 0, 16, 32, 64, 128 bodies */
} nans_tuning;
int nans_autotune_batched(nans_ctx *ctx, int32_t nsteps, nans_tuning *out);
int nans_set_batched_block(nans_ctx *ctx, int32_t threads_per_world);
int nans_set_wave_worlds(nans_ctx *ctx, int32_t max_bodies);

/* Keep the contact list of each world's last step (off by default); at most
 * 16 n + 64 contacts per world of n bodies are kept. */
int nans_record_contacts(nans_ctx *ctx, int32_t enable);

/* The recorded contact list of a world: *n receives its length, up to
 * max_contacts entries are copied to out (host memory). */
int nans_get_contacts(nans_ctx *ctx, int32_t world, nans_contact *out, int32_t max_contacts, int32_t *n);

int nans_get_stats(nans_ctx *ctx, nans_stats *stats);
int nans_reset_stats(nans_ctx *ctx);

/* ---- Differentiable rollout (no counterpart in the reference, which has no
 * gradients; it differentiates the reference's physics block, code/nans.cpp:
 * 1758-1762, through the batched step).
 *
 * nans_rollout advances every world by nsteps batched steps from the current
 * state, as nans_step_batched does, and keeps the state (pos, vel, angvel,
 * angles, force, torque) every checkpoint_every steps on the device.
 *
 * nans_adjoint returns the gradient of a loss L(final state) with respect to the
 * initial state and the per-body mass and moment of inertia and the per-world
 * friction and restitution, given dL/d(final state).  It is a hand-written
 * reverse pass of RK4 (code/nans.cpp:50-78), of Constraint with its unrolled
 * impulse iterations (code/nans.cpp:1020-1329) and of IntegrateVelocities
 * (code/nans.cpp:1331-1349); the GJK/EPA outcome -- which pairs touch, in which
 * order, with which normal and lever arms PointA - PosA, PointB - PosB -- is held
 * fixed at the rollout's.  The segments between checkpoints are recomputed on the
 * device.  The mass gradient counts mass only: when the moment of inertia is a
 * function of the mass, add dl_moi * d(moi)/d(mass).  The batch is left at the
 * rollout's final state.  Arrays are host memory: vectors [3][N] planes, dl_mass
 * and dl_moi [N], dl_friction and dl_restitution [nworlds]; NULL inputs count as
 * zero, NULL outputs are skipped.  At most 128 impulse iterations. */
int nans_rollout(nans_ctx *ctx, int32_t nsteps, int32_t checkpoint_every);

typedef struct nans_adjoint_io {
    const float *dl_pos, *dl_vel, *dl_angvel, *dl_angles;  /* in: dL/d(final state) */
    float *dl_pos0, *dl_vel0, *dl_angvel0, *dl_angles0;    /* out: dL/d(initial state) */
    float *dl_mass, *dl_moi;                                /* out: per body */
    float *dl_friction, *dl_restitution;                    /* out: per world */
} nans_adjoint_io;
int nans_adjoint(nans_ctx *ctx, const nans_adjoint_io *io);

#ifdef __cplusplus
}
#endif
#endif
