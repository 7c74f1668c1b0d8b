// nans_batch.hpp -- the device-resident world batch shared by the step
// (nans_step.hip) and its adjoint (nans_adjoint.hip).
#pragma once

#include "nans_common.hpp"
#include "nans_world.hpp"

#include <vector>

struct nans_worlds;

namespace nans {

constexpr int PLAIN_MAX_BODIES = 16384;  // solver head[] lives in LDS (64 KiB)
constexpr int BATCH_MAX_BODIES = 512;    // bodies of one world kept in LDS by the batched path
constexpr int SCAN_THREADS = 1024;
constexpr int SOLVE_THREADS = 1024;
// The plain step's solve for worlds of at most PLAIN_SMALL_BODIES bodies: 256
// threads, and the terms of the first PLAIN_SMALL_TERMS contacts kept in LDS.
constexpr int PLAIN_SMALL_THREADS = 256;
constexpr int PLAIN_SMALL_BODIES = 512;
constexpr int PLAIN_SMALL_TERMS = 320;
constexpr int CULL_CHUNK = 4096;          // canonical pairs per block of the plain cull

enum : int {
    CNT_EPA_OVERFLOW = 0,
    CNT_CAND_OVERFLOW,
    CNT_CON_OVERFLOW,
    CNT_CANDIDATES,
    CNT_CONTACTS,
    CNT_GJK_FOUND,
    CNT_PAIR_TESTS,
    CNT_N
};

// Arguments of k_batched_step (one persistent launch = one step of every world).
struct BatchArgs {
    float *planes;          // [F_NFIELDS][N]
    int64_t N;
    const int *offsets;     // [nw+1]
    const int *nboxes;      // [nw]
    const int *order;       // largest world first
    int nworlds;
    int nbig;               // order[0, nbig) run by a whole workgroup, the rest one per wave
    int nmax;               // LDS stride of a workgroup's world (>= largest world)
    int wrows;              // LDS stride of a wave's world (>= largest of order[nbig, nworlds))
    int lds_floats;         // floats of world LDS before the polytope LDS (a multiple of 4)
    int *queue;             // this step's work counters: [0] workgroup worlds, [1] wave worlds
    int *queue_next;        // the next step's counters, cleared by this launch (or null)
    StepConsts k;
    const FloorDev *fl;
    const float *wfric, *wrest;  // per-world coefficients (may be null)
    int *cand;              // per-workgroup scratch, cap entries each
    ContactRec *rec;
    int *flag;
    int *qidx;
    SimplexRec *qsim;
    int cap;                // entries per workgroup
    int wcap;               // entries per wave in the wave phase (cap / waves)
    float *corners;         // per-workgroup box corners, corner_stride floats each
    int64_t corner_stride;
    ContactRec *rec_out;    // recording (may be null)
    const int64_t *rec_off;
    int *rec_count;
    const int *rec_cap;
    unsigned long long *cnt;
};

// The batched step's arguments; the warm-started instantiation adds, per world w,
// the previous and this step's contact lists (canonical pair index, three
// accumulators; ascending) at warm_off[w], warm_cap[w] entries each.  The cold
// step's arguments stay BatchArgs exactly.
template <bool WARM>
struct BatchArgsT : BatchArgs {};
template <>
struct BatchArgsT<true> : BatchArgs {
    const int *warm_p_prev = nullptr;
    const float *warm_s_prev = nullptr;
    const int *warm_n_prev = nullptr;
    int *warm_p_next = nullptr;
    float *warm_s_next = nullptr;
    int *warm_n_next = nullptr;
    const int64_t *warm_off = nullptr;
    const int *warm_cap = nullptr;
};

struct AdjointTape;  // nans_adjoint.hip
void adjoint_release(nans_worlds *W);

template <typename T>
inline int dmalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    NANS_HIP(hipMalloc((void **)p, n * sizeof(T)));
    return NANS_OK;
}

void wfree(void *p);
StepConsts step_consts(const nans_params &p, int orientation);
int warm_reset(nans_ctx *ctx);
int setup_floor(nans_ctx *ctx, hipStream_t s);
int check_overflow(nans_ctx *ctx);
// Saves the state a stepping call starts from; step_finish checks the call for a
// capacity overflow and, if it overflowed, restores that state and the counters
// (and empties the warm-start lists) before returning NANS_ERR_CAPACITY.
int snapshot_take(nans_ctx *ctx, hipStream_t s);
int step_finish(nans_ctx *ctx);
// nsteps batched steps replayed from the captured one-step graph
int batched_run(nans_ctx *ctx, int nsteps, int block);
// threads per world of the batched step (the autotuner's choice or the default)
int batched_block(nans_ctx *ctx);
// One batched step launched directly on stream s, recording every world's
// contact list into rec_out (per-world offsets rec_off, capacities rec_cap,
// counts rec_count).
int batched_step_recorded(nans_ctx *ctx, hipStream_t s, ContactRec *rec_out, const int64_t *rec_off,
                          int *rec_count, const int *rec_cap);


}  // namespace nans

struct nans_worlds {
    int nworlds = 0;
    int64_t N = 0;
    int nmax = 0;
    int64_t max_pairs = 0;
    std::vector<int> h_off, h_nb;
    std::vector<float> h_fric, h_rest;
    float *planes = nullptr;
    float *verts = nullptr;
    float *brad = nullptr;
    int *offsets = nullptr;
    int *nboxes = nullptr;
    int *order = nullptr;
    float *wfric = nullptr;
    float *wrest = nullptr;
    nans::FloorDev *floor = nullptr;
    unsigned long long *cnt = nullptr;
    // plain scratch
    int cand_cap = 0;
    int *cand = nullptr;
    nans::ContactRec *rec = nullptr;
    int *flag = nullptr;
    nans::ContactRec *con = nullptr;
    int *done = nullptr;
    int *qidx = nullptr;
    nans::SimplexRec *qsim = nullptr;
    int *blk_counts = nullptr;
    int *blk_offsets = nullptr;
    int *ncand = nullptr;
    int max_blk = 0;
    // Further lanes of the plain step for batches of several worlds: world w runs on
    // lane w % lanes (lane 0 is the context's stream with the scratch above; a world
    // that takes the grid broadphase always runs on lane 0, the grid's only user).
    struct PlainLane {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        int *cand = nullptr;
        nans::ContactRec *rec = nullptr;
        int *flag = nullptr;
        nans::ContactRec *con = nullptr;
        int *done = nullptr;
        int *qidx = nullptr;
        nans::SimplexRec *qsim = nullptr;
        int *blk_counts = nullptr;
        int *blk_offsets = nullptr;
        int *ncand = nullptr;
        int *warm_cp = nullptr;
        float *warm_sums = nullptr;
    };
    std::vector<PlainLane> plain_lanes;  // lanes 1.. (lane 0 is not stored)
    hipEvent_t plain_ev0 = nullptr;
    // hierarchical-grid broadphase of the plain step (sized for the largest world)
    int bp_buckets = 0;          // power of two >= 2 * nmax
    long long *bp_cell = nullptr;  // [3][nmax] cell coordinates
    int *bp_hslot = nullptr;     // [3][nmax] bucket, slot in bucket, level
    int *bp_sorted = nullptr;    // [nmax] bodies grouped by bucket
    int *bp_bcnt = nullptr;      // [bp_buckets]
    int *bp_boff = nullptr;      // [bp_buckets]
    int *bp_counts = nullptr;    // [3 nmax] row counts
    int *bp_offsets = nullptr;   // [3 nmax]
    int *bp_tmp = nullptr;       // [cand_cap] unsorted rows
    int *bp_stash = nullptr;     // [2 nmax][32] first pairs of each row
    int *bp_ovf = nullptr;       // a row overflowed its stash
    // opt-in warm start of the plain and batched steps: two contact lists per world
    int warm_ready = 0, warm_flip = 0;
    std::vector<int64_t> h_warm_off;
    std::vector<int> h_warm_cap;
    int64_t *warm_offd = nullptr;           // [nworlds] h_warm_off on the device
    int *warm_capd = nullptr;               // [nworlds] h_warm_cap on the device
    int *warm_p[2] = {nullptr, nullptr};    // canonical pair indices
    float *warm_s[2] = {nullptr, nullptr};  // [3] accumulators per contact
    int *warm_n[2] = {nullptr, nullptr};    // [nworlds] list lengths
    int *warm_cp = nullptr;                 // [cand_cap] pair index per contact of the world being solved
    float *warm_sums = nullptr;             // [3 cand_cap] its accumulators
    void *bp_grid = nullptr;     // nans::BpGrid
    // batched scratch
    int b_grid = 0;
    int b_cap = 0;
    int *b_cand = nullptr;
    nans::ContactRec *b_rec = nullptr;
    int *b_flag = nullptr;
    int *b_qidx = nullptr;
    nans::SimplexRec *b_qsim = nullptr;
    float *b_corners = nullptr;    // box corners of the batched step's worlds in flight
    int64_t b_corner_floats = 0;   // its size
    int *queue = nullptr;
    int queue_len = 0;
    // recording
    int record = 0;
    nans::ContactRec *rec_buf = nullptr;
    int64_t *rec_off = nullptr;
    int *rec_count = nullptr;
    int *rec_cap = nullptr;
    std::vector<int64_t> h_rec_off;
    std::vector<int> h_rec_cap;
    // graph cache of the batched step
    hipGraphExec_t graph = nullptr;   // two steps, on counter buffers A then B
    hipGraphExec_t graph1 = nullptr;  // one step, on buffer A
    int graph_steps = -1;
    int graph_block = 0;
    nans_params graph_params;
    int graph_orientation = 0;
    hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
    int graph_grid = 0, graph_lds = 0;
    int graph_wave = -1;     // worlds of at most this many bodies ran one per wave
    int graph_nwave = 0;     // and there were this many of them
    int timed_steps = 0;
    float timed_ms = 0.0f;
    int tuned_block = 0;   // autotuner's choice of threads per world (0: not tuned)
    int tuned_wave = -1;   // autotuner's or caller's wave-world bound in bodies (-1: default)
    nans::AdjointTape *adj = nullptr;  // checkpoints and scratch of the differentiable rollout
    float *snap = nullptr;             // state at the start of a stepping call (snapshot_take)
};
