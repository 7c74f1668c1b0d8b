// nans_adjoint.hip -- differentiable rollout of the batched step: forward with
// checkpoints, and the hand-written adjoint (nans_adjoint.hpp) run backwards
// step by step.
//
// nans_rollout keeps the state (P, V, W, angles, force, torque) every
// checkpoint_every steps.  nans_adjoint walks the checkpoint segments from the
// last one: it restores the segment's checkpoint, recomputes the segment's
// steps with the batched step kernel (bit-identical to the rollout) while
// keeping each step's state and ordered contact list, then runs k_adjoint_step
// for the segment's steps in reverse.  Each launch of k_adjoint_step takes one
// world per workgroup: it replays the step's RK4 and SolveConstraints from the
// kept state and contact list (the same routines as the step, so the velocities
// after the solve equal the next kept state's, which is checked), recording the
// velocities each contact read and its dependency round, then propagates the
// adjoint back through IntegrateVelocities, the contacts in reverse round order
// (contacts of one round touch disjoint bodies), and RK4.  Per-body mass and
// moment-of-inertia adjoints accumulate in place; per-world friction and
// restitution adjoints are summed over the contacts in a fixed order, so a run
// is deterministic and independent of checkpoint_every.
#include "nans_adjoint.hpp"
#include "nans_batch.hpp"

#include <algorithm>
#include <climits>
#include <vector>

namespace nans {

enum : int { ADJ_CNT_OVERFLOW = 0, ADJ_CNT_REPLAY, ADJ_CNT_N };
constexpr int ST_FIELDS = 18;  // kept state: P V W A (12 planes) then F T (6 planes)

struct AdjArgs {
    const float *state;   // [ST_FIELDS][N] at the start of the step
    const float *next;    // [ST_FIELDS][N] after it
    const float *planes;  // the batch's planes (scale, mass, moi)
    int64_t N;
    const int *offsets;
    int nmax;
    StepConsts k;
    const FloorDev *fl;
    const float *wfric, *wrest;
    const ContactRec *rec;  // the step's contact lists
    const int64_t *rec_off;
    const int *rec_count;
    const int *rec_cap;
    AdjTape *tape;          // per contact, same offsets as rec
    float *adj;             // [12][N]: adjoints of P V W A, in: after the step, out: before
    float *g_mass, *g_moi;  // [N], accumulated
    float *g_fric, *g_rest; // [nworlds], accumulated
    unsigned long long *cnt;
};

// LDS: body planes (F_NLDS fields), adjoint planes gP gV gW (9) + gm gI (2),
// head[] of the dependency rounds.
constexpr int ADJ_LDS_PLANES = 11;

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_adjoint_step(AdjArgs a)
{
    extern __shared__ float smem[];
    __shared__ FloorDev fl;
    __shared__ float red[2][BLOCK];
    const int w = blockIdx.x, tid = threadIdx.x, nmax = a.nmax;
    const int b0 = a.offsets[w];
    const int n = a.offsets[w + 1] - b0;
    const int ncon = a.rec_count[w];
    if (ncon > a.rec_cap[w]) {  // uniform over the workgroup
        if (tid == 0) atomicAdd(&a.cnt[ADJ_CNT_OVERFLOW], 1ull);
        return;
    }
    if (tid == 0) fl = *a.fl;
    const Planes Lb{smem, nmax};
    float *ga = smem + F_NLDS * nmax;
    const Planes Gd{ga, nmax};  // 0..2 gP, 3..5 gV, 6..8 gW, 9 gm, 10 gI
    int *head = (int *)(ga + ADJ_LDS_PLANES * nmax);
    const Planes S{const_cast<float *>(a.state) + b0, (int)a.N};
    const Planes Nx{const_cast<float *>(a.next) + b0, (int)a.N};
    const Planes G{const_cast<float *>(a.planes) + b0, (int)a.N};
    const Planes Ad{a.adj + b0, (int)a.N};
    const ContactRec *con = a.rec + a.rec_off[w];
    AdjTape *tape = a.tape + a.rec_off[w];
    const float dt = a.k.dt;

    // state at the start of the step, RK4, and the adjoint through IntegrateVelocities
    for (int i = tid; i < n; i += BLOCK) {
#pragma unroll
        for (int f = 0; f < F_SX; ++f) Lb.at(f, i) = S.at(f, i);
#pragma unroll
        for (int f = F_SX; f < F_NLDS; ++f) Lb.at(f, i) = G.at(f, i);
        integrate_forces_body(Lb, i, S.get3(12, i), S.get3(15, i), a.k);
        // P' = P + dt V, A' = A + dt W (V, W after the solve)
        const V3 gP = Ad.get3(0, i), gV = Ad.get3(3, i), gW = Ad.get3(6, i), gA = Ad.get3(9, i);
        Gd.set3(0, i, gP);
        Gd.set3(3, i, add(gV, mul(gP, dt)));
        Gd.set3(6, i, add(gW, mul(gA, dt)));
        Gd.at(9, i) = 0.0f;
        Gd.at(10, i) = 0.0f;
    }
    for (int c = tid; c < ncon; c += BLOCK) tape[c].round = -1;
    __syncthreads();
    const float Cf = a.wfric ? a.wfric[w] : a.k.friction;
    const float Cr = a.wrest ? a.wrest[w] : a.k.restitution;

    // forward SolveConstraints by dependency rounds, keeping what each contact read
    int rounds = 0;
    int pending = ncon;
    while (pending > 0) {
        for (int i = tid; i < n; i += BLOCK) head[i] = INT_MAX;
        __syncthreads();
        for (int c = tid; c < ncon; c += BLOCK) {
            if (tape[c].round < 0) {
                atomicMin(&head[con[c].slot_a], c);
                const int sb = con[c].slot_b;
                if (sb >= 0) atomicMin(&head[sb], c);
            }
        }
        __syncthreads();
        int left = 0;
        for (int c = tid; c < ncon; c += BLOCK) {
            if (tape[c].round < 0) {
                const int sa = con[c].slot_a, sb = con[c].slot_b;
                if (head[sa] == c && (sb < 0 || head[sb] == c)) {
                    AdjTape &t = tape[c];
                    const V3 v1 = Lb.get3(F_VX, sa), w1 = Lb.get3(F_WX, sa);
                    const V3 v2 = sb >= 0 ? Lb.get3(F_VX, sb) : mk3(0.0f, 0.0f, 0.0f);
                    const V3 w2 = sb >= 0 ? Lb.get3(F_WX, sb) : mk3(0.0f, 0.0f, 0.0f);
                    t.v1[0] = v1.x; t.v1[1] = v1.y; t.v1[2] = v1.z;
                    t.w1[0] = w1.x; t.w1[1] = w1.y; t.w1[2] = w1.z;
                    t.v2[0] = v2.x; t.v2[1] = v2.y; t.v2[2] = v2.z;
                    t.w2[0] = w2.x; t.w2[1] = w2.y; t.w2[2] = w2.z;
                    t.round = rounds;
                    ContactRec r = con[c];
                    constraint(Lb, r, fl, a.k, Cf, Cr);
                } else {
                    left = 1;
                }
            }
        }
        pending = __syncthreads_count(left);
        ++rounds;
    }

    // the replay must land on the kept next state
    int bad = 0;
    for (int i = tid; i < n; i += BLOCK) {
        const V3 v = Lb.get3(F_VX, i), wv = Lb.get3(F_WX, i);
        const V3 p = add(Lb.get3(F_PX, i), mul(v, dt));
        bad |= !veq(v, Nx.get3(3, i)) || !veq(wv, Nx.get3(6, i)) || !veq(p, Nx.get3(0, i));
    }
    if (__syncthreads_count(bad) && tid == 0) atomicAdd(&a.cnt[ADJ_CNT_REPLAY], 1ull);

    // reverse SolveConstraints, latest round first
    for (int r = rounds - 1; r >= 0; --r) {
        for (int c = tid; c < ncon; c += BLOCK) {
            AdjTape &t = tape[c];
            if (t.round != r) continue;
            const ContactRec rc = con[c];
            const int sa = rc.slot_a, sb = rc.slot_b;
            ContactBodies cb;
            cb.PA = Lb.get3(F_PX, sa);
            cb.V1 = mk3(t.v1[0], t.v1[1], t.v1[2]);
            cb.W1 = mk3(t.w1[0], t.w1[1], t.w1[2]);
            cb.V2 = mk3(t.v2[0], t.v2[1], t.v2[2]);
            cb.W2 = mk3(t.w2[0], t.w2[1], t.w2[2]);
            cb.mA = Lb.at(F_MASS, sa);
            cb.IA = Lb.at(F_MOI, sa);
            cb.two_sided = sb >= 0;
            ContactGrad g;
            g.gV1 = Gd.get3(3, sa);
            g.gW1 = Gd.get3(6, sa);
            if (sb >= 0) {
                cb.PB = Lb.get3(F_PX, sb);
                cb.mB = Lb.at(F_MASS, sb);
                cb.IB = Lb.at(F_MOI, sb);
                g.gV2 = Gd.get3(3, sb);
                g.gW2 = Gd.get3(6, sb);
            } else {
                cb.PB = mk3(fl.c[0], fl.c[1], fl.c[2]);
                cb.mB = fl.mass;
                cb.IB = fl.moi;
                g.gV2 = g.gW2 = mk3(0.0f, 0.0f, 0.0f);
            }
            constraint_adjoint(rc, cb, a.k, Cf, Cr, g);
            Gd.set3(3, sa, g.gV1);
            Gd.set3(6, sa, g.gW1);
            Gd.set3(0, sa, add(Gd.get3(0, sa), g.gPA));
            Gd.at(9, sa) += g.gmA;
            Gd.at(10, sa) += g.gIA;
            if (sb >= 0) {
                Gd.set3(3, sb, g.gV2);
                Gd.set3(6, sb, g.gW2);
                Gd.set3(0, sb, add(Gd.get3(0, sb), g.gPB));
                Gd.at(9, sb) += g.gmB;
                Gd.at(10, sb) += g.gIB;
            }
            t.gcf = g.gCf;
            t.gcr = g.gCr;
        }
        __syncthreads();
    }

    // reverse RK4 and write back
    for (int i = tid; i < n; i += BLOCK) {
        float gm = Gd.at(9, i), gI = Gd.at(10, i);
        const float m = Lb.at(F_MASS, i), I = Lb.at(F_MOI, i);
        const V3 gV0 = rk4_adjoint<false>(dt, S.get3(3, i), S.get3(12, i), m, a.k, Gd.get3(3, i), gm);
        const V3 gW0 = rk4_adjoint<true>(dt, S.get3(6, i), S.get3(15, i), I, a.k, Gd.get3(6, i), gI);
        Ad.set3(0, i, Gd.get3(0, i));
        Ad.set3(3, i, gV0);
        Ad.set3(6, i, gW0);
        a.g_mass[b0 + i] += gm;
        a.g_moi[b0 + i] += gI;
    }
    // friction / restitution adjoints of the world: strided partial sums, fixed tree
    float sf = 0.0f, sr = 0.0f;
    for (int c = tid; c < ncon; c += BLOCK) {
        sf += tape[c].gcf;
        sr += tape[c].gcr;
    }
    red[0][tid] = sf;
    red[1][tid] = sr;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] += red[0][tid + s];
            red[1][tid] += red[1][tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        a.g_fric[w] += red[0][0];
        a.g_rest[w] += red[1][0];
    }
}

__global__ void k_fill_f32(float *p, int64_t n, float v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_adj_zero_u64(unsigned long long *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

struct AdjointTape {
    int steps = 0, every = 0, ncp = 0;
    nans_params params;
    float *cp = nullptr;        // [ncp][ST_FIELDS][N]
    float *last = nullptr;      // [ST_FIELDS][N] after the rollout
    // backward scratch
    int seg_len = 0;
    float *seg = nullptr;       // [seg_len + 1][ST_FIELDS][N]
    ContactRec *seg_rec = nullptr;  // [seg_len][total]
    int *seg_cnt = nullptr;     // [seg_len][nworlds]
    AdjTape *tape = nullptr;    // [total]
    int64_t total = 0;          // contact capacity summed over worlds
    int64_t *rec_off = nullptr;
    int *rec_cap = nullptr;
    float *adj = nullptr;       // [12][N]
    float *g_mass = nullptr, *g_moi = nullptr, *g_fric = nullptr, *g_rest = nullptr;
    unsigned long long *cnt = nullptr;
};

static void free_segment(AdjointTape *T)
{
    for (void *p : {(void *)T->seg, (void *)T->seg_rec, (void *)T->seg_cnt}) wfree(p);
    T->seg = nullptr;
    T->seg_rec = nullptr;
    T->seg_cnt = nullptr;
    T->seg_len = 0;
}

void adjoint_release(nans_worlds *W)
{
    AdjointTape *T = W->adj;
    if (!T) return;
    free_segment(T);
    for (void *p : {(void *)T->cp, (void *)T->last, (void *)T->tape, (void *)T->rec_off, (void *)T->rec_cap,
                    (void *)T->adj, (void *)T->g_mass, (void *)T->g_moi, (void *)T->g_fric, (void *)T->g_rest,
                    (void *)T->cnt})
        wfree(p);
    delete T;
    W->adj = nullptr;
}

// state planes <-> a kept slot [ST_FIELDS][N]
static int save_state(nans_worlds *W, float *dst, hipStream_t s)
{
    const size_t N = (size_t)W->N;
    NANS_HIP(hipMemcpyAsync(dst, W->planes, sizeof(float) * 12 * N, hipMemcpyDeviceToDevice, s));
    NANS_HIP(hipMemcpyAsync(dst + 12 * N, W->planes + (size_t)F_FX * N, sizeof(float) * 6 * N,
                            hipMemcpyDeviceToDevice, s));
    return NANS_OK;
}

static int load_state(nans_worlds *W, const float *src, hipStream_t s)
{
    const size_t N = (size_t)W->N;
    NANS_HIP(hipMemcpyAsync(W->planes, src, sizeof(float) * 12 * N, hipMemcpyDeviceToDevice, s));
    NANS_HIP(hipMemcpyAsync(W->planes + (size_t)F_FX * N, src + 12 * N, sizeof(float) * 6 * N,
                            hipMemcpyDeviceToDevice, s));
    return NANS_OK;
}

template <int BLOCK>
static int adjoint_launch(nans_ctx *ctx, AdjArgs &a, hipStream_t s)
{
    nans_worlds *W = ctx->worlds;
    const size_t lds = sizeof(float) * (size_t)(F_NLDS + ADJ_LDS_PLANES) * W->nmax + sizeof(int) * (size_t)W->nmax;
    NANS_HIP(hipFuncSetAttribute((const void *)k_adjoint_step<BLOCK>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
    k_adjoint_step<BLOCK><<<W->nworlds, BLOCK, lds, s>>>(a);
    NANS_HIP(hipGetLastError());
    return NANS_OK;
}

}  // namespace nans

using namespace nans;

extern "C" {

int nans_rollout(nans_ctx *ctx, int32_t nsteps, int32_t checkpoint_every)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(ctx->orientation == NANS_ORIENT_EULER, NANS_ERR_STATE,
                 "the differentiable rollout follows the reference's Euler orientation only");
    NANS_REQUIRE(!ctx->warm, NANS_ERR_STATE, "the differentiable rollout does not warm-start contacts");
    NANS_REQUIRE(nsteps >= 1 && checkpoint_every >= 1, NANS_ERR_INVALID,
                 "nsteps and checkpoint_every must be >= 1 (got %d, %d)", nsteps, checkpoint_every);
    NANS_REQUIRE(ctx->p.solver_iterations <= ADJ_MAX_ITERS, NANS_ERR_INVALID,
                 "the adjoint records at most %d impulse iterations (solver_iterations = %d)", ADJ_MAX_ITERS,
                 ctx->p.solver_iterations);
    nans_worlds *W = ctx->worlds;
    NANS_REQUIRE(W->nmax <= BATCH_MAX_BODIES, NANS_ERR_CAPACITY,
                 "the differentiable rollout uses the batched step: at most %d bodies per world (largest: %d)",
                 BATCH_MAX_BODIES, W->nmax);
    NANS_HIP(hipSetDevice(ctx->p.device));
    hipStream_t s = ctx->stream;
    NANS_HIP(hipStreamSynchronize(s));
    if (!W->adj) W->adj = new AdjointTape();
    AdjointTape *T = W->adj;
    const size_t slot = (size_t)ST_FIELDS * W->N;
    const int ncp = (nsteps + checkpoint_every - 1) / checkpoint_every;
    if (ncp != T->ncp || !T->cp) {
        wfree(T->cp);
        T->cp = nullptr;
        T->ncp = 0;
        NANS_TRY(dmalloc(&T->cp, slot * ncp));
        T->ncp = ncp;
    }
    if (!T->last) NANS_TRY(dmalloc(&T->last, slot));
    T->steps = 0;
    NANS_TRY(snapshot_take(ctx, s));
    const int block = batched_block(ctx);
    for (int j = 0; j < ncp; ++j) {
        const int len = std::min(checkpoint_every, nsteps - j * checkpoint_every);
        NANS_TRY(save_state(W, T->cp + slot * j, s));
        NANS_TRY(batched_run(ctx, len, block));
    }
    NANS_TRY(save_state(W, T->last, s));
    NANS_HIP(hipStreamSynchronize(s));
    NANS_TRY(step_finish(ctx));
    T->steps = nsteps;
    T->every = checkpoint_every;
    T->params = ctx->p;
    return NANS_OK;
}

int nans_adjoint(nans_ctx *ctx, const nans_adjoint_io *io)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(io, NANS_ERR_INVALID, "io is NULL");
    nans_worlds *W = ctx->worlds;
    AdjointTape *T = W->adj;
    NANS_REQUIRE(T && T->steps > 0, NANS_ERR_STATE, "no rollout to differentiate (call nans_rollout first)");
    NANS_REQUIRE(memcmp(&T->params, &ctx->p, sizeof(nans_params)) == 0, NANS_ERR_STATE,
                 "the parameters changed since nans_rollout");
    NANS_HIP(hipSetDevice(ctx->p.device));
    hipStream_t s = ctx->stream;
    NANS_HIP(hipStreamSynchronize(s));
    const int64_t N = W->N;
    const int nw = W->nworlds;
    const size_t slot = (size_t)ST_FIELDS * N;
    // contact capacity per world and step, as the contact recording of the step
    if (!T->rec_off) {
        std::vector<int64_t> off(nw + 1, 0);
        std::vector<int> cap(nw);
        for (int w = 0; w < nw; ++w) {
            cap[w] = 16 * (W->h_off[w + 1] - W->h_off[w]) + 64;
            off[w + 1] = off[w] + cap[w];
        }
        T->total = off[nw];
        NANS_TRY(dmalloc(&T->rec_off, (size_t)nw));
        NANS_TRY(dmalloc(&T->rec_cap, (size_t)nw));
        NANS_HIP(hipMemcpy(T->rec_off, off.data(), sizeof(int64_t) * nw, hipMemcpyHostToDevice));
        NANS_HIP(hipMemcpy(T->rec_cap, cap.data(), sizeof(int) * nw, hipMemcpyHostToDevice));
        NANS_TRY(dmalloc(&T->tape, (size_t)T->total));
        NANS_TRY(dmalloc(&T->adj, (size_t)12 * N));
        NANS_TRY(dmalloc(&T->g_mass, (size_t)N));
        NANS_TRY(dmalloc(&T->g_moi, (size_t)N));
        NANS_TRY(dmalloc(&T->g_fric, (size_t)nw));
        NANS_TRY(dmalloc(&T->g_rest, (size_t)nw));
        NANS_TRY(dmalloc(&T->cnt, ADJ_CNT_N));
    }
    const int L = std::min(T->every, T->steps);
    if (T->seg_len != L) {
        free_segment(T);
        NANS_TRY(dmalloc(&T->seg, slot * (L + 1)));
        NANS_TRY(dmalloc(&T->seg_rec, (size_t)T->total * L));
        NANS_TRY(dmalloc(&T->seg_cnt, (size_t)nw * L));
        T->seg_len = L;
    }
    // initial adjoint: dL/d(final state)
    {
        std::vector<float> h((size_t)12 * N, 0.0f);
        const float *src[4] = {io->dl_pos, io->dl_vel, io->dl_angvel, io->dl_angles};
        for (int f = 0; f < 4; ++f)
            if (src[f]) std::copy(src[f], src[f] + 3 * N, h.begin() + (size_t)3 * f * N);
        NANS_HIP(hipMemcpy(T->adj, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    }
    auto fill = [&](float *p, int64_t n) { k_fill_f32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, 0.0f); };
    fill(T->g_mass, N);
    fill(T->g_moi, N);
    fill(T->g_fric, nw);
    fill(T->g_rest, nw);
    k_adj_zero_u64<<<1, 64, 0, s>>>(T->cnt, ADJ_CNT_N);
    NANS_HIP(hipGetLastError());
    // the recomputation must not count in the step statistics
    unsigned long long saved[CNT_N];
    NANS_HIP(hipMemcpy(saved, W->cnt, sizeof(saved), hipMemcpyDeviceToHost));

    AdjArgs a;
    a.planes = W->planes;
    a.N = N;
    a.offsets = W->offsets;
    a.nmax = W->nmax;
    a.k = step_consts(ctx->p, NANS_ORIENT_EULER);
    a.fl = W->floor;
    a.wfric = W->wfric;
    a.wrest = W->wrest;
    a.rec_off = T->rec_off;
    a.rec_cap = T->rec_cap;
    a.tape = T->tape;
    a.adj = T->adj;
    a.g_mass = T->g_mass;
    a.g_moi = T->g_moi;
    a.g_fric = T->g_fric;
    a.g_rest = T->g_rest;
    a.cnt = T->cnt;
    NANS_TRY(setup_floor(ctx, s));
    int rc = NANS_OK;
    for (int j = T->ncp - 1; j >= 0 && rc == NANS_OK; --j) {
        const int t0 = j * T->every;
        const int len = std::min(T->every, T->steps - t0);
        rc = load_state(W, T->cp + slot * j, s);
        for (int st = 0; st < len && rc == NANS_OK; ++st) {
            rc = save_state(W, T->seg + slot * st, s);
            if (rc == NANS_OK)
                rc = batched_step_recorded(ctx, s, T->seg_rec + (size_t)T->total * st, T->rec_off,
                                           T->seg_cnt + (size_t)nw * st, T->rec_cap);
        }
        if (rc == NANS_OK) rc = save_state(W, T->seg + slot * len, s);
        for (int st = len - 1; st >= 0 && rc == NANS_OK; --st) {
            a.state = T->seg + slot * st;
            a.next = T->seg + slot * (st + 1);
            a.rec = T->seg_rec + (size_t)T->total * st;
            a.rec_count = T->seg_cnt + (size_t)nw * st;
            rc = W->nmax <= 64 ? adjoint_launch<64>(ctx, a, s) : adjoint_launch<256>(ctx, a, s);
        }
    }
    // back to the state after the rollout, and the statistics before this call
    if (rc == NANS_OK) rc = load_state(W, T->last, s);
    NANS_HIP(hipStreamSynchronize(s));
    NANS_HIP(hipMemcpy(W->cnt, saved, sizeof(saved), hipMemcpyHostToDevice));
    if (rc != NANS_OK) return rc;
    unsigned long long c[ADJ_CNT_N];
    NANS_HIP(hipMemcpy(c, T->cnt, sizeof(c), hipMemcpyDeviceToHost));
    NANS_REQUIRE(c[ADJ_CNT_OVERFLOW] == 0, NANS_ERR_CAPACITY,
                 "a world had more contacts in one step than the adjoint keeps (16 per body + 64)");
    NANS_REQUIRE(c[ADJ_CNT_REPLAY] == 0, NANS_ERR_STATE,
                 "the adjoint's replay of a step did not reproduce the rollout (%llu world-steps)",
                 (unsigned long long)c[ADJ_CNT_REPLAY]);
    // outputs
    std::vector<float> h((size_t)12 * N);
    NANS_HIP(hipMemcpy(h.data(), T->adj, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    float *dst[4] = {io->dl_pos0, io->dl_vel0, io->dl_angvel0, io->dl_angles0};
    for (int f = 0; f < 4; ++f)
        if (dst[f]) std::copy(h.begin() + (size_t)3 * f * N, h.begin() + (size_t)3 * (f + 1) * N, dst[f]);
    if (io->dl_mass) NANS_HIP(hipMemcpy(io->dl_mass, T->g_mass, sizeof(float) * N, hipMemcpyDeviceToHost));
    if (io->dl_moi) NANS_HIP(hipMemcpy(io->dl_moi, T->g_moi, sizeof(float) * N, hipMemcpyDeviceToHost));
    if (io->dl_friction)
        NANS_HIP(hipMemcpy(io->dl_friction, T->g_fric, sizeof(float) * nw, hipMemcpyDeviceToHost));
    if (io->dl_restitution)
        NANS_HIP(hipMemcpy(io->dl_restitution, T->g_rest, sizeof(float) * nw, hipMemcpyDeviceToHost));
    return check_overflow(ctx);
}

}  // extern "C"
