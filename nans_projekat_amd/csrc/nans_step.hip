// nans_step.hip -- world batches and the simulation step.
//
// One step is the reference's physics block (code/nans.cpp:1758-1762):
//   IntegrateForces -> DetectCollisions -> SolveConstraints -> IntegrateVelocities
// followed by the Model rebuild of its render section (code/nans.cpp:1870-1881,
// 1913-1941), whose box corners the next step's narrowphase reads.  Here the
// Model and corners are rebuilt from the pose at the start of each step, which is
// the same value.
//
// Two launch structures share the device routines of nans_world.hpp:
//   plain    one world per launch sequence, grid-wide kernels:
//            forces+corners | cull count | scan | cull write | narrowphase | solve
//   batched  one persistent launch per step; each workgroup takes worlds from a
//            queue (largest first) and runs the whole step of a world with its
//            bodies in LDS; the steps of a call are captured in a hipGraph.
//
// SolveConstraints applies the contacts one after another in list order
// (code/nans.cpp:1541-1547), each reading the velocities left by the previous
// ones.  Both paths reproduce that order exactly with dependency rounds: in each
// round a contact is applied iff it is the earliest pending contact of each of its
// dynamic bodies.  Contacts applied in one round touch disjoint bodies, and every
// body sees its contacts in list order, so the result equals the serial loop bit
// for bit.  The floor's velocities are never written, so it orders nothing.
#include "nans_batch.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <numeric>
#include <type_traits>
#include <vector>

using namespace nans;

namespace nans {


// ------------------------------------------------------------------ small block helpers
// The thread's rank among the team's threads with pred set, and their number
// (wave_tot: one int per wave of the team).
__device__ inline int team_rank(const Team &t, bool pred, int *wave_tot, int &total)
{
    const int lane = lane_id(), w = t.rank >> 6;
    uint64_t m = ballot(pred);
    int r = __popcll(m & ((1ull << lane) - 1ull));
    if (!t.whole) {
        total = __popcll(m);
        return r;
    }
    if (lane == 0) wave_tot[w] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < t.waves(); ++i) {
        int c = wave_tot[i];
        off += i < w ? c : 0;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + r;
}

template <int BLOCK>
__device__ inline int block_rank(bool pred, int *wave_tot, int &total)
{
    return team_rank(Team::block(BLOCK), pred, wave_tot, total);
}

// A body's shape without its corner registers: kind, centre, radius and vm.
__device__ inline Shape body_shape_ref(const Planes &B, const float *verts, int slot, int nb, const FloorDev &fl)
{
    Shape s;
    if (slot < 0) {
        s.kind = SHAPE_BOX;
        s.c = mk3(fl.c[0], fl.c[1], fl.c[2]);
        s.r = 0.0f;
        s.vm = fl.v;
    } else if (slot < nb) {
        s.kind = SHAPE_BOX;
        s.c = B.get3(F_PX, slot);
        s.r = 0.0f;
        s.vm = verts + 24 * slot;
    } else {
        s.kind = SHAPE_SPHERE;
        s.c = B.get3(F_PX, slot);
        s.r = B.at(F_SX, slot);
        s.vm = nullptr;
    }
    return s;
}

__device__ inline Shape body_shape(const Planes &B, const float *verts, int slot, int nb, const FloorDev &fl)
{
    Shape s;
    if (slot < 0) {
        s.kind = SHAPE_BOX;
        s.c = mk3(fl.c[0], fl.c[1], fl.c[2]);
        s.r = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) s.v[k] = mk3(fl.v[3 * k], fl.v[3 * k + 1], fl.v[3 * k + 2]);
        s.vm = fl.v;
    } else if (slot < nb) {
        s.kind = SHAPE_BOX;
        s.c = B.get3(F_PX, slot);
        s.r = 0.0f;
        const float *v = verts + 24 * slot;
#pragma unroll
        for (int k = 0; k < 8; ++k) s.v[k] = mk3(v[3 * k], v[3 * k + 1], v[3 * k + 2]);
        s.vm = v;
    } else {
        s.kind = SHAPE_SPHERE;
        s.c = B.get3(F_PX, slot);
        s.r = B.at(F_SX, slot);
#pragma unroll
        for (int k = 0; k < 8; ++k) s.v[k] = mk3(0.0f, 0.0f, 0.0f);
        s.vm = nullptr;
    }
    return s;
}

__device__ inline void slot_sphere(const Planes &B, const float *brad, int slot, const FloorDev &fl, V3 &c, float &r)
{
    if (slot < 0) {
        c = mk3(fl.c[0], fl.c[1], fl.c[2]);
        r = fl.brad;
    } else {
        c = B.get3(F_PX, slot);
        r = brad[slot];
    }
}

// Candidates of a world for the narrowphase driver: canonical pair indices
// cand[0..ncand); results as a contact flag and record per candidate.  SMALL:
// the world's pair count fits 24 bits (the batched path), so pairs decode in
// 32-bit arithmetic.
static_assert((long long)BATCH_MAX_BODIES * BATCH_MAX_BODIES < (1LL << 24),
              "every pair index of a batched world fits the 32-bit decode");
template <bool SMALL>
struct WorldSrc {
    Planes B;
    const float *verts;
    int nb, ns;
    const FloorDev *fl;
    const int *cand;
    ContactRec *rec;
    int *flag;
    unsigned long long *cnt;
    PairDecode32 dec;  // SMALL: the world's decode terms
    __device__ int key(int c) const { return cand[c]; }
    __device__ void shapes(int key, Shape &A, Shape &Bs) const
    {
        PairId q = SMALL ? decode_pair32(key, dec) : decode_pair(key, nb, ns);
        int sa, sb;
        pair_slots(q, nb, sa, sb);
        A = body_shape(B, verts, sa, nb, *fl);
        Bs = body_shape(B, verts, sb, nb, *fl);
    }
    static constexpr bool kMemCorners = true;
    __device__ void shapes_ref(int key, Shape &A, Shape &Bs) const
    {
        PairId q = SMALL ? decode_pair32(key, dec) : decode_pair(key, nb, ns);
        int sa, sb;
        pair_slots(q, nb, sa, sb);
        A = body_shape_ref(B, verts, sa, nb, *fl);
        Bs = body_shape_ref(B, verts, sb, nb, *fl);
    }
    __device__ void gjk_result(int c, int g) const { flag[c] = 0; }
    __device__ void epa_result(int c, int key, const EpaResult &r) const
    {
        if (r.status == EPA_OVERFLOW) atomicAdd(&cnt[CNT_EPA_OVERFLOW], 1ull);
        if (r.status != EPA_CONTACT) return;
        PairId q = SMALL ? decode_pair32(key, dec) : decode_pair(key, nb, ns);
        ContactRec &o = rec[c];
        int sa, sb;
        pair_slots(q, nb, sa, sb);
        o.type = q.type;
        o.ia = q.i;
        o.ib = q.j;
        o.slot_a = sa;
        o.slot_b = sb;
        o.pad0 = o.pad1 = o.pad2 = 0;
        o.pa[0] = r.pa.x; o.pa[1] = r.pa.y; o.pa[2] = r.pa.z;
        o.pb[0] = r.pb.x; o.pb[1] = r.pb.y; o.pb[2] = r.pb.z;
        o.n[0] = r.n.x; o.n[1] = r.n.y; o.n[2] = r.n.z;
        o.dln = o.dlt1 = o.dlt2 = 0.0f;
        flag[c] = 1;
    }
};

// Dependency rounds of SolveConstraints over con[0..ncon) (see the file comment).
// head[] has room for n ints; done[] for ncon.  The body slots of each thread's
// first REG contacts stay in registers for all rounds; contacts beyond REG * BLOCK
// are read from con[] each round.  The plain step's solve keeps SOLVE_REG; the
// batched step, compiled for three waves per SIMD, keeps one (two cost it 2.5 %
// of its step in spilled registers).
constexpr int SOLVE_REG = 2;
constexpr int BATCH_SOLVE_REG = 1;

// Contact c of a world's ordered list.
struct DirectCon {
    ContactRec *p;
    __device__ ContactRec &operator()(int c) const { return p[c]; }
};

// Room for Constraint's position-only terms of the first contacts: two
// segments of LDS that are idle during SolveConstraints (terms of contact c in
// a[c] for c < na, else in b[c - na] for c < na + nb).
struct PreTerms {
    ContactTerms *a = nullptr;
    int na = 0;
    ContactTerms *b = nullptr;
    int nb = 0;
    __device__ ContactTerms &operator()(int c) const { return c < na ? a[c] : b[c - na]; }
};

// pre receives the terms of the first contacts in one parallel pass (as many as
// it holds), so that the rounds only evaluate the velocity-dependent part for
// them.  con(c) is contact c's record (DirectCon).
template <bool WARM = false, int REG = SOLVE_REG, class Con>
__device__ inline void solve_team(const Team &t, const Planes &B, int n, const Con &con, int ncon, int *done,
                                  int *head, const FloorDev &fl, const StepConsts &k, float Cf, float Cr,
                                  const PreTerms &pre = PreTerms{}, float *warm = nullptr)
{
    const int tid = t.rank, BLOCK = t.size;
    const int npre = min(pre.na + pre.nb, ncon);
    for (int c = tid; c < npre; c += BLOCK) contact_terms(B, con(c), fl, k, pre(c));
    // a thread's first REG contacts live in registers: their bodies, and
    // the impulses they apply, written back once after the rounds (no global
    // memory access on a round's critical path when their terms are in pre[])
    int sa[REG], sb[REG];
    bool pend[REG];
    float dl[REG][3];
#pragma unroll
    for (int r = 0; r < REG; ++r) {
        const int c = tid + r * BLOCK;
        pend[r] = c < ncon;
        sa[r] = pend[r] ? con(c).slot_a : 0;
        sb[r] = pend[r] ? con(c).slot_b : -1;
        dl[r][0] = dl[r][1] = dl[r][2] = 0.0f;
    }
    for (int c = tid + REG * BLOCK; c < ncon; c += BLOCK) done[c] = 0;
    // Round k's key of contact c is (k << 21) | (2^21 - 1 - c) (contacts number
    // below 2^21 on both paths): the largest key a body receives in round k is its
    // earliest pending contact, and keys of earlier rounds are smaller, so head[]
    // needs clearing only when the round number wraps (every 2047 rounds).
    unsigned *hk = reinterpret_cast<unsigned *>(head);
    for (int i = tid; i < n; i += BLOCK) hk[i] = 0u;
    t.sync();
    int pending = ncon;
    unsigned round = 0;
    while (pending > 0) {
        if (++round == 2048u) {
            for (int i = tid; i < n; i += BLOCK) hk[i] = 0u;
            t.sync();
            round = 1;
        }
        const unsigned rk = (round << 21) | 0x1FFFFFu;
        auto key = [rk](int c) { return rk - (unsigned)c; };
#pragma unroll
        for (int r = 0; r < REG; ++r) {
            if (pend[r]) {
                const unsigned kc = key(tid + r * BLOCK);
                atomicMax(&hk[sa[r]], kc);
                if (sb[r] >= 0) atomicMax(&hk[sb[r]], kc);
            }
        }
        for (int c = tid + REG * BLOCK; c < ncon; c += BLOCK) {
            if (!done[c]) {
                atomicMax(&hk[con(c).slot_a], key(c));
                const int b = con(c).slot_b;
                if (b >= 0) atomicMax(&hk[b], key(c));
            }
        }
        t.sync();
        int left = 0;
        auto apply = [&](int c) {
            ContactRec r = con(c);
            float *sums = WARM ? warm + 3 * c : nullptr;
            if (c < npre) contact_apply(B, r, pre(c), k, Cf, Cr, sums);
            else constraint(B, r, fl, k, Cf, Cr, sums);
            con(c).dln = r.dln;
            con(c).dlt1 = r.dlt1;
            con(c).dlt2 = r.dlt2;
        };
#pragma unroll
        for (int r = 0; r < REG; ++r) {
            if (pend[r]) {
                const int c = tid + r * BLOCK;
                const unsigned kc = key(c);
                if (hk[sa[r]] == kc && (sb[r] < 0 || hk[sb[r]] == kc)) {
                    ContactRec rr;
                    float *sums = WARM ? warm + 3 * c : nullptr;
                    if (c < npre) {
                        rr.slot_a = sa[r];
                        rr.slot_b = sb[r];
                        contact_apply(B, rr, pre(c), k, Cf, Cr, sums);
                    } else {
                        rr = con(c);
                        constraint(B, rr, fl, k, Cf, Cr, sums);
                    }
                    dl[r][0] = rr.dln;
                    dl[r][1] = rr.dlt1;
                    dl[r][2] = rr.dlt2;
                    pend[r] = false;
                } else {
                    left = 1;
                }
            }
        }
        for (int c = tid + REG * BLOCK; c < ncon; c += BLOCK) {
            if (!done[c]) {
                const int a = con(c).slot_a, b = con(c).slot_b;
                if (hk[a] == key(c) && (b < 0 || hk[b] == key(c))) {
                    apply(c);
                    done[c] = 1;
                } else {
                    left = 1;
                }
            }
        }
        pending = t.sync_count(left);
    }
#pragma unroll
    for (int r = 0; r < REG; ++r) {
        const int c = tid + r * BLOCK;
        if (c < ncon) {
            con(c).dln = dl[r][0];
            con(c).dlt1 = dl[r][1];
            con(c).dlt2 = dl[r][2];
        }
    }
}

template <int BLOCK, bool WARM = false, class Con>
__device__ inline void solve_rounds(const Planes &B, int n, const Con &con, int ncon, int *done, int *head,
                                    const FloorDev &fl, const StepConsts &k, float Cf, float Cr,
                                    const PreTerms &pre = PreTerms{}, float *warm = nullptr)
{
    solve_team<WARM>(Team::block(BLOCK), B, n, con, ncon, done, head, fl, k, Cf, Cr, pre, warm);
}

// ------------------------------------------------------------------ plain path kernels
__global__ void k_floor_setup(FloorDev *fl, float px, float py, float pz, float sx, float sy, float sz, float mass,
                              float moi)
{
    if (threadIdx.x != 0) return;
    M4 m = model_from_pose(mk3(px, py, pz), mk3(0.0f, 0.0f, 0.0f), mk3(sx, sy, sz));
    V3 v[8];
    model_vertices(m, v);
    float r2 = 0.0f;
    for (int k = 0; k < 8; ++k) {
        fl->v[3 * k] = v[k].x;
        fl->v[3 * k + 1] = v[k].y;
        fl->v[3 * k + 2] = v[k].z;
        V3 d = sub(v[k], mk3(px, py, pz));
        r2 = fmaxf(r2, dot(d, d));
    }
    fl->c[0] = px;
    fl->c[1] = py;
    fl->c[2] = pz;
    fl->brad = sqrtf(r2);
    fl->mass = mass;
    fl->moi = moi;
}

__global__ __launch_bounds__(256) void k_plain_forces(Planes B, float *verts, float *brad, int n, int nb, StepConsts c)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    V3 F = B.get3(F_FX, i), T = B.get3(F_TX, i);
    integrate_forces_body(B, i, F, T, c);
    B.set3(F_FX, i, mk3(0.0f, 0.0f, 0.0f));
    B.set3(F_TX, i, mk3(0.0f, 0.0f, 0.0f));
    brad[i] = i >= nb ? B.at(F_SX, i)
                      : (c.quat ? box_vertices_quat(B, B, i, verts + 24 * i) : box_vertices_body(B, i, verts + 24 * i));
}

template <bool WRITE>
__global__ __launch_bounds__(256) void k_plain_cull(Planes B, const float *brad, int nb, int ns, int64_t npairs,
                                                    const FloorDev *flp, int *blk_counts, const int *blk_offsets,
                                                    int *cand, int cap, unsigned long long *cnt)
{
    __shared__ int wave_tot[4];
    __shared__ FloorDev fl;
    if (threadIdx.x == 0) fl = *flp;
    __syncthreads();
    const int64_t start = (int64_t)blockIdx.x * CULL_CHUNK;
    const int64_t end = min(start + CULL_CHUNK, npairs);
    int count = 0;
    int base_out = WRITE ? blk_offsets[blockIdx.x] : 0;
    for (int64_t base = start; base < end; base += 256) {
        int64_t p = base + threadIdx.x;
        bool keep = false;
        if (p < end) {
            PairId q = decode_pair(p, nb, ns);
            int sa, sb;
            pair_slots(q, nb, sa, sb);
            V3 ca, cb;
            float ra, rb;
            slot_sphere(B, brad, sa, fl, ca, ra);
            slot_sphere(B, brad, sb, fl, cb, rb);
            keep = cull_keep(ca, ra, cb, rb);
        }
        int total;
        int r = block_rank<256>(keep, wave_tot, total);
        if (WRITE && keep) {
            int dst = base_out + count + r;
            if (dst < cap) cand[dst] = (int)p;
        }
        count += total;
    }
    if (!WRITE && threadIdx.x == 0) blk_counts[blockIdx.x] = count;
    if (WRITE && threadIdx.x == 0) atomicAdd(&cnt[CNT_PAIR_TESTS], (unsigned long long)(end - start));
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan(const int *counts, int nblk, int *offsets, int *total, int cap,
                                                       unsigned long long *cnt)
{
    __shared__ int wave_tot[SCAN_THREADS / 64];
    int running = 0;
    for (int base = 0; base < nblk; base += SCAN_THREADS) {
        int i = base + threadIdx.x;
        int v = i < nblk ? counts[i] : 0;
        // inclusive scan within the block: wave scan then wave offsets
        const int lane = lane_id(), w = threadIdx.x >> 6;
        int x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wave_tot[w] = x;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int k = 0; k < SCAN_THREADS / 64; ++k) {
            int t = wave_tot[k];
            woff += k < w ? t : 0;
            tot += t;
        }
        if (i < nblk) offsets[i] = running + woff + x - v;
        running += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) {
        *total = min(running, cap);
        atomicAdd(&cnt[CNT_CANDIDATES], (unsigned long long)running);
        if (running > cap) atomicAdd(&cnt[CNT_CAND_OVERFLOW], 1ull);
    }
}

}  // namespace nans
#include "nans_broadphase.hpp"
namespace nans {

constexpr int NARROW_RANGE = 512;  // largest share of candidates per workgroup of the plain narrowphase

__global__ __launch_bounds__(256) void k_plain_narrow(Planes B, const float *verts, int nb, int ns,
                                                      const FloorDev *flp, const int *cand, const int *ncand,
                                                      NarrowParams np, ContactRec *rec, int *flag, int *qidx,
                                                      SimplexRec *qsim, unsigned long long *cnt)
{
    __shared__ EpaLds lds[4];
    __shared__ NarrowShared sh;
    __shared__ FloorDev fl;
    if (threadIdx.x == 0) fl = *flp;
    __syncthreads();
    const int total = *ncand;
    // shares of 64..NARROW_RANGE candidates, small enough to give every workgroup one
    const int range = min(NARROW_RANGE, max(64, ((total + (int)gridDim.x - 1) / (int)gridDim.x + 63) & ~63));
    for (int base = blockIdx.x * range; base < total; base += gridDim.x * range) {
        const int count = min(range, total - base);
        WorldSrc<false> src{B, verts, nb, ns, &fl, cand + base, rec + base, flag + base, cnt, PairDecode32{}};
        narrow_block<256, false>(src, count, np, lds, sh, qidx + base, qsim + base);
        if (threadIdx.x == 0) {
            atomicAdd(&cnt[CNT_GJK_FOUND], (unsigned long long)sh.nepa);
            if (sh.stalled) atomicAdd(&cnt[CNT_EPA_OVERFLOW], (unsigned long long)sh.stalled);
        }
    }
}

struct SolveArgs {
    Planes B;
    int n;
    const int *flag;
    const ContactRec *rec;
    const int *ncand;
    ContactRec *con;
    int *done;
    int con_cap;
    StepConsts k;
    const FloorDev *fl;
    float Cf, Cr;
    int npre;             // contact terms kept in the dynamic LDS after head[]
    ContactRec *rec_out;  // recording (may be null)
    int rec_cap;
    int *rec_count;
    unsigned long long *cnt;
    // opt-in warm start: the world's contact lists of the previous and this step
    // (canonical pair index, three accumulators; ascending), scratch per contact
    const int *cand;
    const int *warm_p_prev;
    const float *warm_s_prev;
    const int *warm_n_prev;
    int *warm_p_next;
    float *warm_s_next;
    int *warm_n_next;
    int warm_cap;
    int *warm_cp;
    float *warm_sums;
};

// Ordered contact compaction, dependency-round solve and IntegrateVelocities of
// one world, by one workgroup of T threads; head[] in dynamic LDS, followed by
// npre contact terms (T = PLAIN_SMALL_THREADS for small worlds, whose 256 threads
// have the registers to keep the terms' application free of spills; the
// 1024-thread variant for large worlds keeps none).
template <bool WARM, int T = SOLVE_THREADS>
__global__ __launch_bounds__(T) void k_plain_solve(SolveArgs a)
{
    extern __shared__ int head[];
    __shared__ int wave_tot[T / 64];
    __shared__ FloorDev fl;
    if (threadIdx.x == 0) fl = *a.fl;
    __syncthreads();
    const int total = *a.ncand;
    int ncon = 0;
    for (int base = 0; base < total; base += T) {
        int c = base + threadIdx.x;
        bool hit = c < total && a.flag[c];
        int tot;
        int r = block_rank<T>(hit, wave_tot, tot);
        if (hit && ncon + r < a.con_cap) {
            a.con[ncon + r] = a.rec[c];
            if (WARM) a.warm_cp[ncon + r] = a.cand[c];
        }
        ncon += tot;
    }
    if (ncon > a.con_cap) {
        if (threadIdx.x == 0) atomicAdd(&a.cnt[CNT_CON_OVERFLOW], 1ull);
        ncon = a.con_cap;
    }
    if (threadIdx.x == 0) atomicAdd(&a.cnt[CNT_CONTACTS], (unsigned long long)ncon);
    __syncthreads();
    if (WARM) {
        // accumulators of the same pair in the previous step's list, else zero
        const int np = *a.warm_n_prev;
        for (int c = threadIdx.x; c < ncon; c += T) {
            const int key = a.warm_cp[c];
            int lo = 0, hi = np;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (a.warm_p_prev[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            const bool hit = lo < np && a.warm_p_prev[lo] == key;
            for (int q = 0; q < 3; ++q) a.warm_sums[3 * c + q] = hit ? a.warm_s_prev[3 * lo + q] : 0.0f;
        }
        __syncthreads();
    }
    if constexpr (T == SOLVE_THREADS) {  // no terms kept: the path that applies them is not compiled
        solve_rounds<T, WARM>(a.B, a.n, DirectCon{a.con}, ncon, a.done, head, fl, a.k, a.Cf, a.Cr, PreTerms{},
                              a.warm_sums);
    } else {
        PreTerms pre;
        pre.a = reinterpret_cast<ContactTerms *>(head + ((a.n + 3) & ~3));
        pre.na = a.npre;
        solve_rounds<T, WARM>(a.B, a.n, DirectCon{a.con}, ncon, a.done, head, fl, a.k, a.Cf, a.Cr, pre, a.warm_sums);
    }
    if (WARM) {
        __syncthreads();
        const int m = min(ncon, a.warm_cap);
        for (int c = threadIdx.x; c < m; c += T) {
            a.warm_p_next[c] = a.warm_cp[c];
            for (int q = 0; q < 3; ++q) a.warm_s_next[3 * c + q] = a.warm_sums[3 * c + q];
        }
        if (threadIdx.x == 0) *a.warm_n_next = m;
    }
    for (int i = threadIdx.x; i < a.n; i += T) {
        if (a.k.quat) integrate_velocities_quat(a.B, a.B, i, a.k.dt);
        else integrate_velocities_body(a.B, i, a.k.dt);
    }
    if (a.rec_out) {
        __syncthreads();
        for (int c = threadIdx.x; c < ncon && c < a.rec_cap; c += T) a.rec_out[c] = a.con[c];
        if (threadIdx.x == 0) *a.rec_count = ncon;
    }
}

// ------------------------------------------------------------------ batched kernel

// Floats of one world's LDS for up to rows bodies (F_NLDS planes, the bounding
// radius and the solver's head per body).
__host__ __device__ constexpr int world_lds_floats(int rows) { return (F_NLDS + 2) * rows; }

// Floats of one world's box corners (global memory) for up to rows bodies.
__host__ __device__ constexpr int64_t world_corner_floats(int rows) { return 24LL * rows; }

// One world's memory in the batched step: body planes, the solver's heads and
// bounding radii in LDS for up to rows bodies, and the box corners in the
// workgroup's global scratch.  The corners stay out of LDS so that a world of
// 256 bodies fits three workgroups to a CU (three waves per SIMD); the GJK
// lanes hold a pair's corners in registers, and the rest of their reads hit
// the L1/L2.  The radii are idle once the narrowphase ends: the solver keeps
// contact terms there.
struct WorldLds {
    float *body;   // [F_NLDS][rows] (LDS)
    int *head;     // [rows] (LDS)
    float *brad;   // [rows] (LDS)
    float *verts;  // [rows][24] (global)
    int rows;
};

__device__ inline WorldLds world_lds(float *base, int rows, float *corners)
{
    WorldLds L;
    L.body = base;
    L.head = reinterpret_cast<int *>(base + F_NLDS * rows);
    L.brad = base + (F_NLDS + 1) * rows;
    L.verts = corners;
    L.rows = rows;
    return L;
}

// One world's global scratch, cap entries each: candidates, their contact records
// and flags (the flags become the solver's done marks), the EPA queue (whose
// simplex memory then holds the ordered contact list).
struct WorldScratch {
    int *cand;
    ContactRec *rec;
    int *flag;
    int *qidx;
    SimplexRec *qsim;
    int cap;
};

// The whole step of world w by team t (the workgroup, or one wave of it): epa[k]
// is the polytope LDS of the team's wave k, wave_tot one int per wave.
template <int BLOCK, bool WARM, bool ISL>
__device__ __forceinline__ void batched_world(const BatchArgsT<WARM> &a, const Team &t, int w, const WorldLds &L,
                                              const WorldScratch &S, EpaLds *epa, NarrowShared &nsh,
                                              int *wave_tot, const FloorDev &fl)
{
    constexpr int NWMAX = BLOCK / 64;
    const int tid = t.rank, T = t.size, NW = t.waves();
    const int b0 = a.offsets[w];
    const int n = a.offsets[w + 1] - b0;
    const int nb = a.nboxes[w];
    const int ns = n - nb;
    const Planes Lb{L.body, L.rows};
    const Planes Gw{a.planes + b0, (int)a.N};
    float *const vertsL = L.verts, *const bradL = L.brad;
    int *cand = S.cand;
    ContactRec *rec = S.rec;
    int *flag = S.flag;
    int *done = flag;
    ContactRec *con = (ContactRec *)S.qsim;
    // load the world, IntegrateForces, corners
    for (int i = tid; i < n; i += T) {
#pragma unroll
        for (int f = 0; f < F_NLDS; ++f) Lb.at(f, i) = Gw.at(f, i);
        V3 F = Gw.get3(F_FX, i), Tq = Gw.get3(F_TX, i);
        integrate_forces_body(Lb, i, F, Tq, a.k);
        Gw.set3(F_FX, i, mk3(0.0f, 0.0f, 0.0f));
        Gw.set3(F_TX, i, mk3(0.0f, 0.0f, 0.0f));
        bradL[i] = i >= nb ? Lb.at(F_SX, i)
                           : (a.k.quat ? box_vertices_quat(Lb, Gw, i, vertsL + 24 * i)
                                       : box_vertices_body(Lb, i, vertsL + 24 * i));
    }
    t.sync();
    // cull over the canonical pairs, in canonical order: wave w takes the w-th
    // contiguous share of the list, with no workgroup barrier, and appends the
    // survivors to its own region of the contact-record scratch (idle until the
    // narrowphase); the regions are then concatenated into cand[] in wave order
    // (a one-wave team writes cand[] directly)
    const int64_t npairs = pair_count(nb, ns);
    const PairDecode32 dec = pair_decode32(nb, ns);
    const int wv = tid >> 6, lane = tid & 63;
    const int64_t share = (npairs + NW - 1) / NW;
    const int64_t q0 = min(npairs, (int64_t)wv * share), q1 = min(npairs, q0 + share);
    int *wbuf = NW == 1 ? cand : reinterpret_cast<int *>(rec) + (int64_t)wv * S.cap;  // cap ints per wave
    int mine = 0;
    // A 64-pair chunk of box-box pairs covers the end of row ci from column cj
    // and the rows after it: a uniform walk over those rows (one or two while
    // rows are long) gives each lane its pair and the next chunk's start; the
    // other kinds (after the box-box triangle) are decoded per lane.
    const int ncc = nb * (nb - 1) / 2;
    int ci = 0, cj = 0;
    if (q0 < min(q1, (int64_t)ncc)) {
        const PairId s0 = decode_pair32((int)q0, dec);
        ci = __builtin_amdgcn_readfirstlane(s0.i);  // the same in every lane: kept scalar
        cj = __builtin_amdgcn_readfirstlane(s0.j);
    }
    // this lane's pair of the chunk starting at row ci, column cj; advances (ci, cj)
    auto walk = [&](int &sa, int &sb) {
        int acc = 0, row = ci, col = cj;
        sa = sb = 0;
        for (;;) {
            const int len = nb - col;  // pairs (row, col..nb-1)
            if (lane >= acc && lane < acc + len) {
                sa = row;
                sb = col + lane - acc;
            }
            if (acc + len >= 64 || row >= nb - 2) {  // the chunk ends in this row, or the triangle does
                if (acc + len > 64) {
                    ci = row;
                    cj = col + 64 - acc;
                } else {
                    ci = row + 1;
                    cj = row + 2;
                }
                break;
            }
            acc += len;
            row += 1;
            col = row + 1;
        }
    };
    auto append = [&](bool k, int p) {
        const uint64_t m = ballot(k);
        const int r = mine + __popcll(m & ((1ull << lane) - 1ull));
        if (k && r < S.cap) wbuf[r] = p;
        mine += __popcll(m);
    };
    for (int64_t b = q0; b < q1; b += 64) {
        const int p = (int)(b + lane);
        bool k = false;
        if (b < ncc) {
            int sa, sb;
            walk(sa, sb);
            if (p < q1 && p < ncc) k = cull_keep(Lb.get3(F_PX, sa), bradL[sa], Lb.get3(F_PX, sb), bradL[sb]);
        }
        if (p >= ncc && p < q1) {
            const PairId q = decode_pair32(p, dec);
            int sa, sb;
            pair_slots(q, nb, sa, sb);
            V3 ca, cb;
            float ra, rb;
            slot_sphere(Lb, bradL, sa, fl, ca, ra);
            slot_sphere(Lb, bradL, sb, fl, cb, rb);
            k = cull_keep(ca, ra, cb, rb);
        }
        append(k, p);
    }
    int ncand = mine;
    if (NW > 1) {
        if (lane == 0) wave_tot[wv] = mine;
        t.sync();
        ncand = 0;
        int woff[NWMAX];
#pragma unroll
        for (int i = 0; i < NWMAX; ++i) {
            woff[i] = ncand;
            ncand += i < NW ? wave_tot[i] : 0;
        }
        if constexpr (WARM || BLOCK > 256) {
            // each wave's region in turn, one load at a time (the warm-started and the
            // 512-thread kernels spill with the two-load copy below)
#pragma unroll
            for (int i = 0; i < NWMAX; ++i) {
                if (i >= NW) break;
                const int cnt = min(wave_tot[i], S.cap);
                const int *src = reinterpret_cast<const int *>(rec) + (int64_t)i * S.cap;
                for (int k = tid; k < cnt && woff[i] + k < S.cap; k += T) cand[woff[i] + k] = src[k];
            }
        } else {
            // cand[c] for c < min(ncand, cap) from the region of the wave whose share
            // holds c; a thread's copies are independent, so it issues two loads
            // before their stores
            const int nc = min(ncand, S.cap);
            const int *wbase = reinterpret_cast<const int *>(rec);
            auto src = [&](int c) {
                int wi = 0, off = 0;
#pragma unroll
                for (int q = 1; q < NWMAX; ++q) {
                    if (q < NW && c >= woff[q]) {
                        wi = q;
                        off = woff[q];
                    }
                }
                return wbase + (int64_t)wi * S.cap + (c - off);
            };
            for (int c = tid; c < nc; c += 2 * T) {
                const int c2 = c + T;
                const int v = *src(c);
                const int v2 = c2 < nc ? *src(c2) : 0;
                cand[c] = v;
                if (c2 < nc) cand[c2] = v2;
            }
        }
    }
    t.sync();
    if (ncand > S.cap) {
        if (tid == 0) atomicAdd(&a.cnt[CNT_CAND_OVERFLOW], 1ull);
        ncand = S.cap;
    }
    if (tid == 0) {
        atomicAdd(&a.cnt[CNT_PAIR_TESTS], (unsigned long long)npairs);
        atomicAdd(&a.cnt[CNT_CANDIDATES], (unsigned long long)ncand);
    }
    t.sync();
    // narrowphase over the candidates (GJK per lane, EPA per wave)
    {
        WorldSrc<true> src{Lb, vertsL, nb, ns, &fl, cand, rec, flag, a.cnt, dec};
        narrow_team<false, ISL>(t, src, ncand, a.k.np, epa, nsh, S.qidx, S.qsim);
    }
    if (tid == 0) {
        atomicAdd(&a.cnt[CNT_GJK_FOUND], (unsigned long long)nsh.nepa);
        if (nsh.stalled) atomicAdd(&a.cnt[CNT_EPA_OVERFLOW], (unsigned long long)nsh.stalled);
    }
    // the contacts in candidate (= canonical) order; con reuses the queue's memory
    // (warm start: cand[] is compacted alongside, in place, to the contacts'
    // canonical pair indices; team_rank's barriers order the reads before the
    // writes, and a one-wave team reads before it writes in each pass)
    int ncon = 0;
    for (int base = 0; base < ncand; base += T) {
        int c = base + tid;
        bool hit = c < ncand && flag[c];
        const int key = WARM && hit ? cand[c] : 0;
        int tot;
        int rk = team_rank(t, hit, wave_tot, tot);
        if (hit) {
            con[ncon + rk] = rec[c];
            if (WARM) cand[ncon + rk] = key;
        }
        ncon += tot;
    }
    if (tid == 0) atomicAdd(&a.cnt[CNT_CONTACTS], (unsigned long long)ncon);
    t.sync();
    // warm start: each contact's accumulators from the previous step's list (zero
    // when its pair was no contact there), held in rec[]'s memory, idle from here
    float *wsum = reinterpret_cast<float *>(rec);  // [3 ncon]
    if constexpr (WARM) {
        const int64_t off = a.warm_off[w];
        const int *pp = a.warm_p_prev + off;
        const float *sp = a.warm_s_prev + 3 * off;
        const int np = a.warm_n_prev[w];
        for (int c = tid; c < ncon; c += T) {
            const int key = cand[c];
            int lo = 0, hi = np;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (pp[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            const bool hit = lo < np && pp[lo] == key;
            for (int q = 0; q < 3; ++q) wsum[3 * c + q] = hit ? sp[3 * lo + q] : 0.0f;
        }
        t.sync();
    }
    const float Cf = a.wfric ? a.wfric[w] : a.k.friction;
    const float Cr = a.wrest ? a.wrest[w] : a.k.restitution;
    // contact terms in the radii, then in the team's polytope LDS
    PreTerms pre;
    pre.a = reinterpret_cast<ContactTerms *>(L.brad);
    pre.na = (int)(sizeof(float) * L.rows / sizeof(ContactTerms));
    pre.b = reinterpret_cast<ContactTerms *>(epa);
    pre.nb = (int)(NW * sizeof(EpaLds) / sizeof(ContactTerms));
    solve_team<WARM, BATCH_SOLVE_REG>(t, Lb, n, DirectCon{con}, ncon, done, L.head, fl, a.k, Cf, Cr, pre,
                                      WARM ? wsum : nullptr);
    if constexpr (WARM) {  // this step's list: the first warm_cap[w] contacts (contact c's sums are team thread c % T's)
        const int64_t off = a.warm_off[w];
        const int m = min(ncon, a.warm_cap[w]);
        for (int c = tid; c < m; c += T) {
            a.warm_p_next[off + c] = cand[c];
            for (int q = 0; q < 3; ++q) a.warm_s_next[3 * (off + c) + q] = wsum[3 * c + q];
        }
        if (tid == 0) a.warm_n_next[w] = m;
    }
    // IntegrateVelocities and write-back
    for (int i = tid; i < n; i += T) {
        if (a.k.quat) integrate_velocities_quat(Lb, Gw, i, a.k.dt);
        else integrate_velocities_body(Lb, i, a.k.dt);
#pragma unroll
        for (int f = 0; f < F_SX; ++f) Gw.at(f, i) = Lb.at(f, i);
    }
    if (a.rec_out) {
        ContactRec *dst = a.rec_out + a.rec_off[w];
        int rc = a.rec_cap[w];
        for (int c = tid; c < ncon && c < rc; c += T) dst[c] = con[c];
        if (tid == 0) a.rec_count[w] = ncon;
    }
    t.sync();
}

// The persistent step: the worlds order[0, nbig) (the larger ones, largest
// first) each by a whole workgroup, then order[nbig, nworlds) each by one wave,
// every wave of a workgroup taking its own.  Each phase hands out its worlds
// from its own counter (queue[0], queue[1]).
// WARM: the opt-in warm start, each contact's accumulators carried over from the
// world's previous step (a separate instantiation; the default step has none of it).
// ISL: the opt-in FP64-islands mode, EPA in the grouped tiers with a binary64 final
// face (a separate instantiation, so that the default step carries no binary64 code).
// Registers: at 256 threads three waves per SIMD (168 VGPRs; the compiler
// spills a few dozen, which a third world per CU more than repays); two at 64
// and 128 threads, the sizes for batches of small worlds that fit the chip at
// once (a step then lasts one world's step, which the spills would lengthen),
// and at 512, whose workgroup fits a CU once.
__host__ __device__ constexpr int batch_waves_per_simd(int block) { return block == 256 ? 3 : 2; }

template <int BLOCK, bool WARM = false, bool ISL = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(batch_waves_per_simd(BLOCK)))) void
k_batched_step(BatchArgsT<WARM> a)
{
    extern __shared__ float smem[];
    constexpr int NW = BLOCK / 64;
    __shared__ int wave_tot[NW];
    __shared__ int s_job;
    const FloorDev &fl = *a.fl;  // in global memory, as the box corners
    __shared__ NarrowShared nsh[NW];
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // the next launch's counters: no launch of this step reads them
    if (a.queue_next && blockIdx.x == 0 && threadIdx.x == 0) {
        a.queue_next[0] = 0;
        a.queue_next[1] = 0;
    }
    EpaLds *const epa = reinterpret_cast<EpaLds *>(smem + a.lds_floats);
    const int64_t wg = (int64_t)blockIdx.x * a.cap;
    // the workgroup's team, LDS and scratch until the first phase ends
    Team t = Team::block(BLOCK);
    float *const corners = a.corners + (int64_t)blockIdx.x * a.corner_stride;
    WorldLds L = world_lds(smem, a.nmax, corners);
    WorldScratch S{a.cand + wg, a.rec + wg, a.flag + wg, a.qidx + wg, a.qsim + wg, a.cap};
    EpaLds *tepa = epa;
    int tsh = 0;
    bool whole = true;
    for (;;) {
        int job = 0;
        if (whole) {
            if (threadIdx.x == 0) s_job = atomicAdd(&a.queue[0], 1);
            __syncthreads();
            job = __builtin_amdgcn_readfirstlane(s_job);
            __syncthreads();
            if (job >= a.nbig) {  // the wave phase: each wave its own LDS part and scratch
                whole = false;
                t = Team::wave();
                L = world_lds(smem + wv * world_lds_floats(a.wrows), a.wrows,
                              corners + wv * world_corner_floats(a.wrows));
                const int64_t o = wg + (int64_t)wv * a.wcap;
                S = WorldScratch{a.cand + o, a.rec + o, a.flag + o, a.qidx + o, a.qsim + o, a.wcap};
                tepa = epa + wv;
                tsh = wv;
            }
        }
        if (!whole) {
            int j = 0;
            if ((threadIdx.x & 63) == 0) j = atomicAdd(&a.queue[1], 1);
            job = a.nbig + __builtin_amdgcn_readfirstlane(j);
        }
        if (job >= a.nworlds) break;
        batched_world<BLOCK, WARM, ISL>(a, t, a.order[job], L, S, tepa, nsh[tsh], wave_tot, fl);
    }
}

__global__ void k_zero_i32(int *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace nans

// ====================================================================== host side

namespace nans {

void wfree(void *p)
{
    if (p) (void)hipFree(p);
}

static void drop_graph(nans_worlds *W)
{
    if (W->graph) (void)hipGraphExecDestroy(W->graph);
    if (W->graph1) (void)hipGraphExecDestroy(W->graph1);
    W->graph = nullptr;
    W->graph1 = nullptr;
    W->graph_steps = -1;
}

// Frees the batched step's per-workgroup scratch, and with it the captured graph,
// whose kernel arguments point into that scratch.
static void release_batch_scratch(nans_worlds *W)
{
    drop_graph(W);
    wfree(W->b_cand);
    wfree(W->b_rec);
    wfree(W->b_flag);
    wfree(W->b_qidx);
    wfree(W->b_qsim);
    wfree(W->b_corners);
    W->b_cand = nullptr;
    W->b_rec = nullptr;
    W->b_flag = nullptr;
    W->b_qidx = nullptr;
    W->b_qsim = nullptr;
    W->b_corners = nullptr;
    W->b_corner_floats = 0;
    W->b_grid = 0;
}

static void release_recording(nans_worlds *W)
{
    wfree(W->rec_buf);
    wfree(W->rec_off);
    wfree(W->rec_count);
    wfree(W->rec_cap);
    W->rec_buf = nullptr;
    W->rec_off = nullptr;
    W->rec_count = nullptr;
    W->rec_cap = nullptr;
}

static void release_plain_lanes(nans_worlds *W)
{
    for (auto &L : W->plain_lanes) {
        for (void *q : {(void *)L.cand, (void *)L.rec, (void *)L.flag, (void *)L.con, (void *)L.done,
                        (void *)L.qidx, (void *)L.qsim, (void *)L.blk_counts, (void *)L.blk_offsets,
                        (void *)L.ncand, (void *)L.warm_cp, (void *)L.warm_sums})
            wfree(q);
        if (L.ev) (void)hipEventDestroy(L.ev);
        if (L.s) (void)hipStreamDestroy(L.s);
    }
    W->plain_lanes.clear();
    if (W->plain_ev0) (void)hipEventDestroy(W->plain_ev0);
    W->plain_ev0 = nullptr;
}

int worlds_release(nans_ctx *ctx)
{
    if (!ctx || !ctx->worlds) return NANS_OK;
    nans_worlds *W = ctx->worlds;
    drop_graph(W);
    for (void *p : {(void *)W->planes, (void *)W->verts, (void *)W->brad, (void *)W->offsets, (void *)W->nboxes,
                    (void *)W->order, (void *)W->wfric, (void *)W->wrest, (void *)W->floor, (void *)W->cnt,
                    (void *)W->cand, (void *)W->rec, (void *)W->flag, (void *)W->con, (void *)W->done,
                    (void *)W->qidx, (void *)W->qsim,
                    (void *)W->blk_counts, (void *)W->blk_offsets, (void *)W->ncand, (void *)W->queue,
                    (void *)W->bp_cell, (void *)W->bp_hslot, (void *)W->bp_sorted,
                    (void *)W->bp_bcnt, (void *)W->bp_boff, (void *)W->bp_counts, (void *)W->bp_offsets,
                    (void *)W->bp_tmp, (void *)W->bp_stash, (void *)W->bp_ovf, W->bp_grid, (void *)W->warm_p[0], (void *)W->warm_p[1], (void *)W->warm_s[0],
                    (void *)W->warm_s[1], (void *)W->warm_n[0], (void *)W->warm_n[1], (void *)W->warm_cp,
                    (void *)W->warm_sums, (void *)W->warm_offd, (void *)W->warm_capd, (void *)W->snap})
        wfree(p);
    for (hipEvent_t ev : {W->ev_in, W->ev_out, W->ev_t0, W->ev_t1})
        if (ev) (void)hipEventDestroy(ev);
    release_batch_scratch(W);
    release_recording(W);
    release_plain_lanes(W);
    adjoint_release(W);
    delete W;
    ctx->worlds = nullptr;
    return NANS_OK;
}

// Quaternions (w, x, y, z planes) of Euler angles in degrees ([3][n] planes),
// composed as the Model rebuild composes them: q = qx(ax) qy(ay) qz(az), so that
// R(q) = Rx Ry Rz (code/nans.cpp:1913-1924).  Binary64, rounded once.
static void euler_to_quat_planes(const float *ang, float *q, int64_t n)
{
    const double k = 3.14159265358979323846 / 360.0;  // degrees to half-angle radians
    for (int64_t i = 0; i < n; ++i) {
        const double hx = ang[i] * k, hy = ang[n + i] * k, hz = ang[2 * n + i] * k;
        const double cx = cos(hx), sx = sin(hx), cy = cos(hy), sy = sin(hy), cz = cos(hz), sz = sin(hz);
        q[i] = (float)(cx * cy * cz - sx * sy * sz);
        q[n + i] = (float)(sx * cy * cz + cx * sy * sz);
        q[2 * n + i] = (float)(cx * sy * cz - sx * cy * sz);
        q[3 * n + i] = (float)(cx * cy * sz + sx * sy * cz);
    }
}

StepConsts step_consts(const nans_params &p, int orientation)
{
    StepConsts k;
    k.quat = orientation == NANS_ORIENT_QUATERNION;
    k.np = narrow_params(p);
    k.solver_iterations = p.solver_iterations;
    k.beta = p.beta;
    k.restitution = p.restitution;
    k.friction = p.friction;
    k.damping = p.damping;
    k.gravity = p.gravity;
    k.dt = p.dt;
    return k;
}

int setup_floor(nans_ctx *ctx, hipStream_t s)
{
    const nans_params &p = ctx->p;
    k_floor_setup<<<1, 64, 0, s>>>(ctx->worlds->floor, p.floor_pos[0], p.floor_pos[1], p.floor_pos[2],
                                   p.floor_scale[0], p.floor_scale[1], p.floor_scale[2], p.floor_mass, p.floor_moi);
    NANS_HIP(hipGetLastError());
    return NANS_OK;
}


static int alloc_recording(nans_ctx *ctx)
{
    nans_worlds *W = ctx->worlds;
    release_recording(W);
    W->h_rec_off.assign(W->nworlds + 1, 0);
    W->h_rec_cap.assign(W->nworlds, 0);
    for (int w = 0; w < W->nworlds; ++w) {
        int n = W->h_off[w + 1] - W->h_off[w];
        W->h_rec_cap[w] = 16 * n + 64;
        W->h_rec_off[w + 1] = W->h_rec_off[w] + W->h_rec_cap[w];
    }
    NANS_TRY(dmalloc(&W->rec_buf, (size_t)W->h_rec_off[W->nworlds]));
    NANS_TRY(dmalloc(&W->rec_off, (size_t)W->nworlds));
    NANS_TRY(dmalloc(&W->rec_count, (size_t)W->nworlds));
    NANS_TRY(dmalloc(&W->rec_cap, (size_t)W->nworlds));
    NANS_HIP(hipMemcpy(W->rec_off, W->h_rec_off.data(), sizeof(int64_t) * W->nworlds, hipMemcpyHostToDevice));
    NANS_HIP(hipMemcpy(W->rec_cap, W->h_rec_cap.data(), sizeof(int) * W->nworlds, hipMemcpyHostToDevice));
    std::vector<int> z(W->nworlds, 0);
    NANS_HIP(hipMemcpy(W->rec_count, z.data(), sizeof(int) * W->nworlds, hipMemcpyHostToDevice));
    return NANS_OK;
}

// How the batched step splits the worlds between whole workgroups and single
// waves when worlds of at most `bodies` bodies run one per wave.
struct WaveSplit {
    int bodies;  // the bound (0: every world by a workgroup)
    int nbig;    // worlds run by a workgroup: the first nbig of the order (largest first)
    int nwave;   // worlds run one per wave
    int wrows;   // largest of those
    int64_t wpairs;  // their largest pair count
};

static WaveSplit wave_split(const nans_worlds *W, int bodies)
{
    WaveSplit sp{bodies, 0, 0, 0, 0};
    for (int w = 0; w < W->nworlds; ++w) {
        const int n = W->h_off[w + 1] - W->h_off[w];
        if (n > bodies) {
            ++sp.nbig;
        } else {
            ++sp.nwave;
            sp.wrows = std::max(sp.wrows, n);
            sp.wpairs = std::max(sp.wpairs, pair_count(W->h_nb[w], n - W->h_nb[w]));
        }
    }
    return sp;
}

// Floats of world LDS ahead of the polytope LDS: a workgroup's world, or one world
// per wave, whichever is larger, rounded up to 16 bytes.
static int batch_lds_floats(int nmax, const WaveSplit &sp, int block)
{
    int f = std::max(world_lds_floats(nmax), sp.nwave ? (block / 64) * world_lds_floats(sp.wrows) : 0);
    return (f + 3) & ~3;
}

static size_t batch_lds_bytes(int nmax, const WaveSplit &sp, int block)
{
    return sizeof(float) * (size_t)batch_lds_floats(nmax, sp, block) + sizeof(EpaLds) * (size_t)(block / 64);
}

// Default bound of the worlds run one per wave: none.  A world on one wave takes
// 1.5x (8-15 bodies) to 2.9x (32-63 bodies) as long as on a 256-thread
// workgroup, so the wave phase, which runs last, lengthens the step's tail more
// than it saves (config 3 on the MI355X: 1.93 ms per step at bound 0, 1.92 at
// 16, 2.12 at 64); the autotuner tries the bounds per batch.
constexpr int WAVE_WORLD_BODIES = 0;

// The wave-world bound at a workgroup size: none at one wave per workgroup.
static int wave_bodies(const nans_worlds *W, int block)
{
    if (block <= 64) return 0;
    return W->tuned_wave >= 0 ? W->tuned_wave : WAVE_WORLD_BODIES;
}

}  // namespace nans

namespace nans {

int check_overflow(nans_ctx *ctx)
{
    unsigned long long c[CNT_N];
    NANS_HIP(hipMemcpy(c, ctx->worlds->cnt, sizeof(c), hipMemcpyDeviceToHost));
    NANS_REQUIRE(c[CNT_CAND_OVERFLOW] == 0 && c[CNT_CON_OVERFLOW] == 0, NANS_ERR_CAPACITY,
                 "pair candidates or contacts exceeded the scratch capacity; the batch is left as it was "
                 "before the call");
    return NANS_OK;
}

// A stepping call that drops candidates or contacts beyond its scratch has
// advanced the state without them.  The call saves the state it starts from --
// the planes [0, F_SX) (pos, vel, angvel, angles) and [F_FX, F_NFIELDS) (force,
// torque, quaternions) and the counters -- and restores it when check_overflow
// reports the overflow, so the error leaves the batch as the call found it.
constexpr int SNAP_HEAD = F_SX, SNAP_TAIL = F_NFIELDS - F_FX;

int snapshot_take(nans_ctx *ctx, hipStream_t s)
{
    nans_worlds *W = ctx->worlds;
    const size_t N = (size_t)W->N;
    if (!W->snap) NANS_TRY(dmalloc(&W->snap, (SNAP_HEAD + SNAP_TAIL) * N + 2 * CNT_N));
    NANS_HIP(hipMemcpyAsync(W->snap, W->planes, sizeof(float) * SNAP_HEAD * N, hipMemcpyDeviceToDevice, s));
    NANS_HIP(hipMemcpyAsync(W->snap + SNAP_HEAD * N, W->planes + (size_t)F_FX * N, sizeof(float) * SNAP_TAIL * N,
                            hipMemcpyDeviceToDevice, s));
    NANS_HIP(hipMemcpyAsync(W->snap + (SNAP_HEAD + SNAP_TAIL) * N, W->cnt, sizeof(unsigned long long) * CNT_N,
                            hipMemcpyDeviceToDevice, s));
    return NANS_OK;
}

// check_overflow, restoring the call's starting state when it fails.
int step_finish(nans_ctx *ctx)
{
    const int rc = check_overflow(ctx);
    if (rc != NANS_ERR_CAPACITY) return rc;
    nans_worlds *W = ctx->worlds;
    const size_t N = (size_t)W->N;
    hipStream_t s = ctx->stream;
    NANS_HIP(hipMemcpyAsync(W->planes, W->snap, sizeof(float) * SNAP_HEAD * N, hipMemcpyDeviceToDevice, s));
    NANS_HIP(hipMemcpyAsync(W->planes + (size_t)F_FX * N, W->snap + SNAP_HEAD * N, sizeof(float) * SNAP_TAIL * N,
                            hipMemcpyDeviceToDevice, s));
    NANS_HIP(hipMemcpyAsync(W->cnt, W->snap + (SNAP_HEAD + SNAP_TAIL) * N, sizeof(unsigned long long) * CNT_N,
                            hipMemcpyDeviceToDevice, s));
    NANS_TRY(warm_reset(ctx));
    NANS_HIP(hipStreamSynchronize(s));
    return rc;
}

}  // namespace nans

namespace nans {

// The grid broadphase replaces the all-pairs cull for worlds of at least this
// many bodies (both give the same candidates).
constexpr int GRID_MIN_BODIES = 1024;

static bool use_grid(const nans_ctx *ctx, int n)
{
    if (ctx->broadphase == NANS_BROADPHASE_GRID) return true;
    if (ctx->broadphase == NANS_BROADPHASE_ALL_PAIRS) return false;
    return n >= GRID_MIN_BODIES;
}

// Contact lists of the opt-in warm start, allocated on its first step: per
// world, two lists (the previous step's and this step's) of at most 16 n + 64
// contacts, and per-contact scratch of the size of the candidate scratch (plain
// step; the batched step keeps its per-contact scratch in its workgroup's).
// The plain and the batched step share the lists, so the two can be mixed.
static int warm_prepare(nans_ctx *ctx)
{
    nans_worlds *W = ctx->worlds;
    if (W->warm_ready) return NANS_OK;
    W->h_warm_off.assign(W->nworlds + 1, 0);
    W->h_warm_cap.assign(W->nworlds, 0);
    for (int w = 0; w < W->nworlds; ++w) {
        W->h_warm_cap[w] = 16 * (W->h_off[w + 1] - W->h_off[w]) + 64;
        W->h_warm_off[w + 1] = W->h_warm_off[w] + W->h_warm_cap[w];
    }
    const size_t total = (size_t)W->h_warm_off[W->nworlds];
    for (int b = 0; b < 2; ++b) {
        NANS_TRY(dmalloc(&W->warm_p[b], total));
        NANS_TRY(dmalloc(&W->warm_s[b], 3 * total));
        NANS_TRY(dmalloc(&W->warm_n[b], (size_t)W->nworlds));
    }
    NANS_TRY(dmalloc(&W->warm_cp, (size_t)W->cand_cap));
    NANS_TRY(dmalloc(&W->warm_sums, (size_t)3 * W->cand_cap));
    NANS_TRY(dmalloc(&W->warm_offd, (size_t)W->nworlds));
    NANS_TRY(dmalloc(&W->warm_capd, (size_t)W->nworlds));
    NANS_HIP(hipMemcpy(W->warm_offd, W->h_warm_off.data(), sizeof(int64_t) * W->nworlds, hipMemcpyHostToDevice));
    NANS_HIP(hipMemcpy(W->warm_capd, W->h_warm_cap.data(), sizeof(int) * W->nworlds, hipMemcpyHostToDevice));
    W->warm_ready = 1;
    return warm_reset(ctx);
}

// Empties the warm-start lists (the next step starts every contact from zero).
int warm_reset(nans_ctx *ctx)
{
    nans_worlds *W = ctx->worlds;
    if (!W || !W->warm_ready) return NANS_OK;
    for (int b = 0; b < 2; ++b) k_zero_i32<<<(W->nworlds + 255) / 256, 256, 0, ctx->stream>>>(W->warm_n[b], W->nworlds);
    NANS_HIP(hipGetLastError());
    W->warm_flip = 0;
    return NANS_OK;
}

// Candidates of one world from the hierarchical grid (nans_broadphase.hpp), in
// canonical order, into W->cand / W->ncand.
static void grid_broadphase(nans_worlds *W, const Planes &B, const float *brad, int nb, int ns, int64_t npairs,
                            hipStream_t s)
{
    const int n = nb + ns, blocks = (n + 255) / 256, L = 3 * nb + 2 * ns;
    BpGrid *g = (BpGrid *)W->bp_grid;
    const unsigned mask = (unsigned)W->bp_buckets - 1;
    k_bp_bounds<<<1, BP_THREADS, 0, s>>>(B, brad, n, W->bp_bcnt, W->bp_buckets, W->bp_counts, L, g, W->bp_ovf,
                                          W->cnt, (long long)npairs);
    k_bp_cells<<<blocks, 256, 0, s>>>(B, brad, n, g, mask, W->bp_cell, W->nmax, W->bp_hslot, W->bp_bcnt);
    scan_counts(W->bp_bcnt, W->bp_buckets, W->bp_boff, nullptr, 0, nullptr, s);
    k_bp_scatter<<<blocks, 256, 0, s>>>(n, W->bp_hslot, W->nmax, W->bp_boff, W->bp_sorted);
    BpFind f;
    f.B = B;
    f.brad = brad;
    f.nb = nb;
    f.ns = ns;
    f.fl = W->floor;
    f.g = g;
    f.mask = mask;
    f.cell = W->bp_cell;
    f.stride = W->nmax;
    f.hslot = W->bp_hslot;
    f.bucket_cnt = W->bp_bcnt;
    f.bucket_off = W->bp_boff;
    f.sorted = W->bp_sorted;
    f.counts = W->bp_counts;
    f.offsets = W->bp_offsets;
    f.tmp = W->bp_tmp;
    f.cand = W->cand;
    f.cap = W->cand_cap;
    f.stash = W->bp_stash;
    f.ovf = W->bp_ovf;
    const int wave_blocks = (n + 3) / 4;  // one wave per body
    k_bp_find<false><<<wave_blocks, 256, 0, s>>>(f);
    scan_counts(W->bp_counts, L, W->bp_offsets, W->ncand, W->cand_cap, W->cnt, s);
    k_bp_find<true><<<wave_blocks, 256, 0, s>>>(f);
    const int sort_waves = 2 * nb + ns + (nb + ns + 63) / 64;  // rows, then flags 64 per wave
    k_bp_sort<<<(sort_waves + 3) / 4, 256, 0, s>>>(nb, ns, W->bp_offsets, W->ncand, W->bp_tmp, W->bp_stash,
                                                   W->bp_ovf, W->cand, W->cand_cap);
}

}  // namespace nans

extern "C" {

int nans_upload_worlds(nans_ctx *ctx, int32_t nworlds, const int32_t *world_offsets, const nans_bodies_soa *b,
                       const float *world_friction, const float *world_restitution)
{
    NANS_REQUIRE(ctx, NANS_ERR_INVALID, "ctx is NULL");
    NANS_REQUIRE(nworlds >= 1, NANS_ERR_INVALID, "nworlds must be >= 1");
    NANS_REQUIRE(world_offsets && b, NANS_ERR_INVALID, "world_offsets or bodies is NULL");
    NANS_REQUIRE(world_offsets[0] == 0, NANS_ERR_INVALID, "world_offsets[0] must be 0");
    int64_t N = world_offsets[nworlds];
    NANS_REQUIRE(N >= 0 && N < (1LL << 30), NANS_ERR_CAPACITY, "body count out of range");
    NANS_REQUIRE(N == 0 || (b->pos && b->vel && b->angvel && b->angles && b->scale && b->mass && b->moi && b->kind),
                 NANS_ERR_INVALID, "a body array is NULL");
    std::vector<int> off(world_offsets, world_offsets + nworlds + 1), nbx(nworlds);
    int nmax = 0;
    int64_t maxp = 0;
    for (int w = 0; w < nworlds; ++w) {
        NANS_REQUIRE(off[w + 1] >= off[w], NANS_ERR_INVALID, "world_offsets must be non-decreasing");
        int n = off[w + 1] - off[w];
        int nb = 0;
        while (nb < n && b->kind[off[w] + nb] == NANS_SHAPE_BOX) ++nb;
        for (int i = nb; i < n; ++i)
            NANS_REQUIRE(b->kind[off[w] + i] == NANS_SHAPE_SPHERE, NANS_ERR_INVALID,
                         "world %d: boxes must precede spheres and kinds must be box or sphere", w);
        for (int i = 0; i < n; ++i)
            NANS_REQUIRE(b->mass[off[w] + i] > 0.0f && b->moi[off[w] + i] > 0.0f, NANS_ERR_INVALID,
                         "world %d body %d: mass and moi must be positive", w, i);
        nbx[w] = nb;
        nmax = std::max(nmax, n);
        maxp = std::max(maxp, pair_count(nb, n - nb));
    }
    NANS_REQUIRE(nmax <= PLAIN_MAX_BODIES, NANS_ERR_CAPACITY, "a world holds %d bodies; the limit is %d", nmax,
                 PLAIN_MAX_BODIES);
    NANS_HIP(hipSetDevice(ctx->p.device));
    NANS_HIP(hipDeviceSynchronize());
    worlds_release(ctx);
    nans_worlds *W = new nans_worlds();
    ctx->worlds = W;
    W->nworlds = nworlds;
    W->N = N;
    W->nmax = nmax;
    W->max_pairs = maxp;
    W->h_off = off;
    W->h_nb = nbx;
    NANS_TRY(dmalloc(&W->planes, (size_t)F_NFIELDS * N));
    NANS_TRY(dmalloc(&W->verts, (size_t)24 * N));
    NANS_TRY(dmalloc(&W->brad, (size_t)N));
    NANS_TRY(dmalloc(&W->offsets, (size_t)nworlds + 1));
    NANS_TRY(dmalloc(&W->nboxes, (size_t)nworlds));
    NANS_TRY(dmalloc(&W->order, (size_t)nworlds));
    NANS_TRY(dmalloc(&W->floor, 1));
    NANS_TRY(dmalloc(&W->cnt, CNT_N));
    // host staging in plane layout
    std::vector<float> h((size_t)F_NFIELDS * N, 0.0f);
    auto put3 = [&](int f, const float *src) {
        if (!src) return;
        for (int c = 0; c < 3; ++c) std::copy(src + c * N, src + (c + 1) * N, h.begin() + (size_t)(f + c) * N);
    };
    put3(F_PX, b->pos);
    put3(F_VX, b->vel);
    put3(F_WX, b->angvel);
    put3(F_AX, b->angles);
    put3(F_SX, b->scale);
    put3(F_FX, b->force);
    put3(F_TX, b->torque);
    std::copy(b->mass, b->mass + N, h.begin() + (size_t)F_MASS * N);
    std::copy(b->moi, b->moi + N, h.begin() + (size_t)F_MOI * N);
    euler_to_quat_planes(h.data() + (size_t)F_AX * N, h.data() + (size_t)F_QW * N, N);
    NANS_HIP(hipMemcpy(W->planes, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    NANS_HIP(hipMemcpy(W->offsets, off.data(), sizeof(int) * (nworlds + 1), hipMemcpyHostToDevice));
    NANS_HIP(hipMemcpy(W->nboxes, nbx.data(), sizeof(int) * nworlds, hipMemcpyHostToDevice));
    // largest world first (cost ~ pairs + bodies)
    std::vector<int> ord(nworlds);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        int nx = off[x + 1] - off[x], ny = off[y + 1] - off[y];
        return nx > ny;
    });
    NANS_HIP(hipMemcpy(W->order, ord.data(), sizeof(int) * nworlds, hipMemcpyHostToDevice));
    if (world_friction) {
        W->h_fric.assign(world_friction, world_friction + nworlds);
        NANS_TRY(dmalloc(&W->wfric, (size_t)nworlds));
        NANS_HIP(hipMemcpy(W->wfric, world_friction, sizeof(float) * nworlds, hipMemcpyHostToDevice));
    }
    if (world_restitution) {
        W->h_rest.assign(world_restitution, world_restitution + nworlds);
        NANS_TRY(dmalloc(&W->wrest, (size_t)nworlds));
        NANS_HIP(hipMemcpy(W->wrest, world_restitution, sizeof(float) * nworlds, hipMemcpyHostToDevice));
    }
    // plain scratch: candidates of the largest world
    static_assert(64LL * PLAIN_MAX_BODIES + 4096 < (1LL << 21), "contact numbers fit the solver's round keys");
    int64_t cap = std::min<int64_t>(maxp, 64LL * nmax + 4096);
    W->cand_cap = (int)cap;
    W->max_blk = (int)((maxp + CULL_CHUNK - 1) / CULL_CHUNK) + 1;
    NANS_TRY(dmalloc(&W->cand, (size_t)cap));
    NANS_TRY(dmalloc(&W->rec, (size_t)cap));
    NANS_TRY(dmalloc(&W->flag, (size_t)cap));
    NANS_TRY(dmalloc(&W->con, (size_t)cap));
    NANS_TRY(dmalloc(&W->done, (size_t)cap));
    NANS_TRY(dmalloc(&W->qidx, (size_t)cap));
    NANS_TRY(dmalloc(&W->qsim, (size_t)cap));
    NANS_TRY(dmalloc(&W->blk_counts, (size_t)W->max_blk));
    NANS_TRY(dmalloc(&W->blk_offsets, (size_t)W->max_blk));
    NANS_TRY(dmalloc(&W->ncand, 1));
    W->bp_buckets = 64;
    while (W->bp_buckets < 2 * nmax) W->bp_buckets *= 2;
    NANS_TRY(dmalloc(&W->bp_cell, (size_t)3 * nmax));
    NANS_TRY(dmalloc(&W->bp_hslot, (size_t)3 * nmax));
    NANS_TRY(dmalloc(&W->bp_sorted, (size_t)nmax));
    NANS_TRY(dmalloc(&W->bp_bcnt, (size_t)W->bp_buckets));
    NANS_TRY(dmalloc(&W->bp_boff, (size_t)W->bp_buckets));
    NANS_TRY(dmalloc(&W->bp_counts, (size_t)3 * nmax));
    NANS_TRY(dmalloc(&W->bp_offsets, (size_t)3 * nmax));
    NANS_TRY(dmalloc(&W->bp_tmp, (size_t)cap));
    NANS_TRY(dmalloc(&W->bp_stash, (size_t)2 * nmax * BP_STASH));
    NANS_TRY(dmalloc(&W->bp_ovf, 1));
    NANS_TRY(dmalloc((BpGrid **)&W->bp_grid, 1));
    for (const void *f : {(const void *)k_plain_solve<false>, (const void *)k_plain_solve<true>})
        NANS_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(sizeof(int) * PLAIN_MAX_BODIES)));
    for (const void *f : {(const void *)k_plain_solve<false, PLAIN_SMALL_THREADS>,
                          (const void *)k_plain_solve<true, PLAIN_SMALL_THREADS>})
        NANS_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(sizeof(int) * PLAIN_SMALL_BODIES + sizeof(ContactTerms) * PLAIN_SMALL_TERMS)));
    k_zero_u64<<<1, 64>>>(W->cnt, CNT_N);
    NANS_HIP(hipGetLastError());
    NANS_TRY(setup_floor(ctx, nullptr));
    if (W->record) NANS_TRY(alloc_recording(ctx));
    NANS_HIP(hipDeviceSynchronize());
    return NANS_OK;
}

int nans_download_state(nans_ctx *ctx, nans_bodies_soa *b)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(b, NANS_ERR_INVALID, "bodies is NULL");
    nans_worlds *W = ctx->worlds;
    int64_t N = W->N;
    NANS_HIP(hipSetDevice(ctx->p.device));
    NANS_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<float> h((size_t)F_NFIELDS * N);
    NANS_HIP(hipMemcpy(h.data(), W->planes, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    auto get3 = [&](int f, float *dst) {
        if (!dst) return;
        for (int c = 0; c < 3; ++c)
            std::copy(h.begin() + (size_t)(f + c) * N, h.begin() + (size_t)(f + c + 1) * N, dst + c * N);
    };
    get3(F_PX, b->pos);
    get3(F_VX, b->vel);
    get3(F_WX, b->angvel);
    get3(F_AX, b->angles);
    get3(F_FX, b->force);
    get3(F_TX, b->torque);
    return NANS_OK;
}

int nans_record_contacts(nans_ctx *ctx, int32_t enable)
{
    NANS_REQUIRE(ctx, NANS_ERR_INVALID, "ctx is NULL");
    if (!ctx->worlds) {
        NANS_REQUIRE(!enable, NANS_ERR_STATE, "upload worlds before enabling contact recording");
        return NANS_OK;
    }
    NANS_HIP(hipSetDevice(ctx->p.device));
    NANS_HIP(hipDeviceSynchronize());
    nans_worlds *W = ctx->worlds;
    W->record = enable ? 1 : 0;
    drop_graph(W);
    if (W->record) return alloc_recording(ctx);
    release_recording(W);
    return NANS_OK;
}

int nans_get_contacts(nans_ctx *ctx, int32_t world, nans_contact *out, int32_t max_contacts, int32_t *n)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    nans_worlds *W = ctx->worlds;
    NANS_REQUIRE(W->record, NANS_ERR_STATE, "contact recording is off (nans_record_contacts)");
    NANS_REQUIRE(world >= 0 && world < W->nworlds, NANS_ERR_INVALID, "world %d out of range", world);
    NANS_REQUIRE(n, NANS_ERR_INVALID, "n is NULL");
    NANS_HIP(hipSetDevice(ctx->p.device));
    NANS_HIP(hipStreamSynchronize(ctx->stream));
    int cnt = 0;
    NANS_HIP(hipMemcpy(&cnt, W->rec_count + world, sizeof(int), hipMemcpyDeviceToHost));
    NANS_REQUIRE(cnt <= W->h_rec_cap[world], NANS_ERR_CAPACITY, "world %d had %d contacts; %d were recorded", world,
                 cnt, W->h_rec_cap[world]);
    *n = cnt;
    int m = std::min(cnt, max_contacts);
    if (m > 0 && out) {
        std::vector<ContactRec> tmp(m);
        NANS_HIP(hipMemcpy(tmp.data(), W->rec_buf + W->h_rec_off[world], sizeof(ContactRec) * m,
                           hipMemcpyDeviceToHost));
        for (int i = 0; i < m; ++i) {
            out[i].type = tmp[i].type;
            out[i].index_a = tmp[i].ia;
            out[i].index_b = tmp[i].ib;
            out[i].pad = 0;
            for (int c = 0; c < 3; ++c) {
                out[i].point_a[c] = tmp[i].pa[c];
                out[i].point_b[c] = tmp[i].pb[c];
                out[i].normal[c] = tmp[i].n[c];
            }
            out[i].dl_normal = tmp[i].dln;
            out[i].dl_tangent1 = tmp[i].dlt1;
            out[i].dl_tangent2 = tmp[i].dlt2;
        }
    }
    return NANS_OK;
}

int nans_get_stats(nans_ctx *ctx, nans_stats *st)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(st, NANS_ERR_INVALID, "stats is NULL");
    NANS_HIP(hipSetDevice(ctx->p.device));
    NANS_HIP(hipStreamSynchronize(ctx->stream));
    unsigned long long c[CNT_N];
    NANS_HIP(hipMemcpy(c, ctx->worlds->cnt, sizeof(c), hipMemcpyDeviceToHost));
    st->pair_tests = (int64_t)c[CNT_PAIR_TESTS];
    st->candidates = (int64_t)c[CNT_CANDIDATES];
    st->gjk_found = (int64_t)c[CNT_GJK_FOUND];
    st->contacts = (int64_t)c[CNT_CONTACTS];
    st->epa_overflow = (int64_t)c[CNT_EPA_OVERFLOW];
    st->capacity_overflow = (int64_t)(c[CNT_CAND_OVERFLOW] + c[CNT_CON_OVERFLOW]);
    return NANS_OK;
}

int nans_reset_stats(nans_ctx *ctx)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_HIP(hipSetDevice(ctx->p.device));
    k_zero_u64<<<1, 64, 0, ctx->stream>>>(ctx->worlds->cnt, CNT_N);
    NANS_HIP(hipGetLastError());
    NANS_HIP(hipStreamSynchronize(ctx->stream));
    return NANS_OK;
}


// Lanes of the plain step: as many as worlds, at most PLAIN_LANES (each a stream
// with its own scratch).  The worlds of a batch are independent, and one world's
// launches are short, so several in flight keep the GPU busy.
constexpr int PLAIN_LANES = 8;

static int plain_lanes_prepare(nans_ctx *ctx, int lanes)
{
    nans_worlds *W = ctx->worlds;
    const size_t cap = (size_t)W->cand_cap;
    if (!W->plain_ev0) NANS_HIP(hipEventCreateWithFlags(&W->plain_ev0, hipEventDisableTiming));
    while ((int)W->plain_lanes.size() < lanes - 1) {
        W->plain_lanes.emplace_back();
        nans_worlds::PlainLane &L = W->plain_lanes.back();
        NANS_HIP(hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking));
        NANS_HIP(hipEventCreateWithFlags(&L.ev, hipEventDisableTiming));
        NANS_TRY(dmalloc(&L.cand, cap));
        NANS_TRY(dmalloc(&L.rec, cap));
        NANS_TRY(dmalloc(&L.flag, cap));
        NANS_TRY(dmalloc(&L.con, cap));
        NANS_TRY(dmalloc(&L.done, cap));
        NANS_TRY(dmalloc(&L.qidx, cap));
        NANS_TRY(dmalloc(&L.qsim, cap));
        NANS_TRY(dmalloc(&L.blk_counts, (size_t)W->max_blk));
        NANS_TRY(dmalloc(&L.blk_offsets, (size_t)W->max_blk));
        NANS_TRY(dmalloc(&L.ncand, 1));
    }
    if (ctx->warm) {
        for (auto &L : W->plain_lanes) {
            if (!L.warm_cp) NANS_TRY(dmalloc(&L.warm_cp, cap));
            if (!L.warm_sums) NANS_TRY(dmalloc(&L.warm_sums, 3 * cap));
        }
    }
    return NANS_OK;
}

int nans_step_plain(nans_ctx *ctx, int32_t nsteps)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(nsteps >= 0, NANS_ERR_INVALID, "nsteps < 0");
    nans_worlds *W = ctx->worlds;
    NANS_HIP(hipSetDevice(ctx->p.device));
    hipStream_t s = ctx->stream;
    if (ctx->warm) NANS_TRY(warm_prepare(ctx));
    NANS_TRY(setup_floor(ctx, s));
    NANS_TRY(snapshot_take(ctx, s));
    const StepConsts k = step_consts(ctx->p, ctx->orientation);
    const int lanes = std::min(PLAIN_LANES, W->nworlds);
    NANS_TRY(plain_lanes_prepare(ctx, lanes));
    // the lanes start after the set-up above on the context's stream
    NANS_HIP(hipEventRecord(W->plain_ev0, s));
    for (int l = 1; l < lanes; ++l) NANS_HIP(hipStreamWaitEvent(W->plain_lanes[l - 1].s, W->plain_ev0, 0));
    for (int step = 0; step < nsteps; ++step) {
        for (int w = 0; w < W->nworlds; ++w) {
            const int b0 = W->h_off[w], n = W->h_off[w + 1] - b0, nb = W->h_nb[w], ns = n - nb;
            if (n == 0) continue;
            // this world's lane: its stream and scratch (always the same lane for a
            // world, so its steps stay in order)
            const int l = use_grid(ctx, n) ? 0 : w % lanes;
            const nans_worlds::PlainLane *L = l ? &W->plain_lanes[l - 1] : nullptr;
            hipStream_t ls = L ? L->s : s;
            int *cand = L ? L->cand : W->cand, *flag = L ? L->flag : W->flag, *done = L ? L->done : W->done;
            int *qidx = L ? L->qidx : W->qidx, *ncand = L ? L->ncand : W->ncand;
            int *blk_counts = L ? L->blk_counts : W->blk_counts, *blk_offsets = L ? L->blk_offsets : W->blk_offsets;
            ContactRec *rec = L ? L->rec : W->rec, *con = L ? L->con : W->con;
            SimplexRec *qsim = L ? L->qsim : W->qsim;
            Planes B{W->planes + b0, (int)W->N};
            float *verts = W->verts + (int64_t)24 * b0;
            float *brad = W->brad + b0;
            const int64_t npairs = pair_count(nb, ns);
            const int nblk = (int)((npairs + CULL_CHUNK - 1) / CULL_CHUNK);
            k_plain_forces<<<(n + 255) / 256, 256, 0, ls>>>(B, verts, brad, n, nb, k);
            if (nblk > 0) {
                if (use_grid(ctx, n)) {
                    grid_broadphase(W, B, brad, nb, ns, npairs, s);
                } else {
                    k_plain_cull<false><<<nblk, 256, 0, ls>>>(B, brad, nb, ns, npairs, W->floor, blk_counts,
                                                              nullptr, nullptr, 0, W->cnt);
                    scan_counts(blk_counts, nblk, blk_offsets, ncand, W->cand_cap, W->cnt, ls);
                    k_plain_cull<true><<<nblk, 256, 0, ls>>>(B, brad, nb, ns, npairs, W->floor, nullptr,
                                                             blk_offsets, cand, W->cand_cap, W->cnt);
                }
                int grid = std::max(1, std::min((W->cand_cap + NARROW_RANGE - 1) / NARROW_RANGE, 2048));
                k_plain_narrow<<<grid, 256, 0, ls>>>(B, verts, nb, ns, W->floor, cand, ncand, k.np, rec, flag,
                                                     qidx, qsim, W->cnt);
            } else {
                k_zero_i32<<<1, 64, 0, ls>>>(ncand, 1);
            }
            SolveArgs a;
            a.B = B;
            a.n = n;
            a.flag = flag;
            a.rec = rec;
            a.ncand = ncand;
            a.con = con;
            a.done = done;
            a.con_cap = W->cand_cap;
            a.k = k;
            a.fl = W->floor;
            a.Cf = W->h_fric.empty() ? k.friction : W->h_fric[w];
            a.Cr = W->h_rest.empty() ? k.restitution : W->h_rest[w];
            a.rec_out = W->record ? W->rec_buf + W->h_rec_off[w] : nullptr;
            a.rec_cap = W->record ? W->h_rec_cap[w] : 0;
            a.rec_count = W->record ? W->rec_count + w : nullptr;
            a.cnt = W->cnt;
            a.cand = cand;
            // worlds of at most PLAIN_SMALL_BODIES bodies: 256 threads, contact terms in LDS
            const bool small = n <= PLAIN_SMALL_BODIES;
            a.npre = small ? PLAIN_SMALL_TERMS : 0;
            const size_t lds = sizeof(int) * (size_t)((n + 3) & ~3) + sizeof(ContactTerms) * (size_t)a.npre;
            if (ctx->warm) {
                const int cur = W->warm_flip, nxt = cur ^ 1;
                const int64_t off = W->h_warm_off[w];
                a.warm_p_prev = W->warm_p[cur] + off;
                a.warm_s_prev = W->warm_s[cur] + 3 * off;
                a.warm_n_prev = W->warm_n[cur] + w;
                a.warm_p_next = W->warm_p[nxt] + off;
                a.warm_s_next = W->warm_s[nxt] + 3 * off;
                a.warm_n_next = W->warm_n[nxt] + w;
                a.warm_cap = W->h_warm_cap[w];
                a.warm_cp = L ? L->warm_cp : W->warm_cp;
                a.warm_sums = L ? L->warm_sums : W->warm_sums;
                if (small) k_plain_solve<true, PLAIN_SMALL_THREADS><<<1, PLAIN_SMALL_THREADS, lds, ls>>>(a);
                else k_plain_solve<true><<<1, SOLVE_THREADS, lds, ls>>>(a);
            } else {
                if (small) k_plain_solve<false, PLAIN_SMALL_THREADS><<<1, PLAIN_SMALL_THREADS, lds, ls>>>(a);
                else k_plain_solve<false><<<1, SOLVE_THREADS, lds, ls>>>(a);
            }
            NANS_HIP(hipGetLastError());
        }
        if (ctx->warm) W->warm_flip ^= 1;
    }
    // the context's stream waits for every lane
    for (int l = 1; l < lanes; ++l) {
        NANS_HIP(hipEventRecord(W->plain_lanes[l - 1].ev, W->plain_lanes[l - 1].s));
        NANS_HIP(hipStreamWaitEvent(s, W->plain_lanes[l - 1].ev, 0));
    }
    NANS_HIP(hipStreamSynchronize(s));
    return step_finish(ctx);
}

}  // extern "C"

namespace nans {

// f(B, I) with B the workgroup size (64, 128, 256 or 512) and I the FP64-islands
// flag as compile-time constants: picks the batched kernel's instantiation.
template <class F>
static int by_shape(int block, bool isl, F &&f)
{
    auto go = [&](auto B) { return isl ? f(B, std::true_type{}) : f(B, std::false_type{}); };
    switch (block) {
    case 64: return go(std::integral_constant<int, 64>{});
    case 128: return go(std::integral_constant<int, 128>{});
    case 512: return go(std::integral_constant<int, 512>{});
    default: return go(std::integral_constant<int, 256>{});
    }
}

template <int BLOCK, bool WARM, bool ISL>
static int batched_config(nans_ctx *ctx, const WaveSplit &sp, int *grid, size_t *lds)
{
    nans_worlds *W = ctx->worlds;
    *lds = batch_lds_bytes(W->nmax, sp, BLOCK);
    NANS_HIP(hipFuncSetAttribute((const void *)k_batched_step<BLOCK, WARM, ISL>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)*lds));
    int per_cu = 0;
    NANS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batched_step<BLOCK, WARM, ISL>, BLOCK, *lds));
    NANS_REQUIRE(per_cu > 0, NANS_ERR_CAPACITY, "a world of %d bodies does not fit the LDS of one workgroup",
                 W->nmax);
    int cus = 0;
    NANS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->p.device));
    // no more workgroups than jobs: one per workgroup world, one per NW wave worlds
    const int jobs = sp.nbig + (sp.nwave + BLOCK / 64 - 1) / (BLOCK / 64);
    *grid = std::max(1, std::min(jobs, per_cu * cus));
    return NANS_OK;
}

// Launch shape, per-workgroup scratch and kernel arguments of the batched step.
template <int BLOCK, bool WARM, bool ISL>
static int batched_prepare(nans_ctx *ctx, BatchArgs &a, int *grid_out, size_t *lds_out,
                           WaveSplit *split_out = nullptr)
{
    nans_worlds *W = ctx->worlds;
    int grid = 0;
    size_t lds = 0;
    const WaveSplit sp = wave_split(W, wave_bodies(W, BLOCK));
    NANS_TRY((batched_config<BLOCK, WARM, ISL>(ctx, sp, &grid, &lds)));
    constexpr int NW = BLOCK / 64;
    int64_t cap = std::min<int64_t>(W->max_pairs, 32LL * W->nmax + 1024);
    cap = std::max<int64_t>(cap, 64);
    // in the wave phase each wave has cap / NW entries: room for every pair of its world
    if (sp.nwave) cap = std::max<int64_t>(cap, NW * std::max<int64_t>(sp.wpairs, 64));
    if (grid > W->b_grid || cap != W->b_cap) {
        release_batch_scratch(W);
        NANS_TRY(dmalloc(&W->b_cand, (size_t)grid * cap));
        NANS_TRY(dmalloc(&W->b_rec, (size_t)grid * cap));
        NANS_TRY(dmalloc(&W->b_flag, (size_t)grid * cap));
        NANS_TRY(dmalloc(&W->b_qidx, (size_t)grid * cap));
        NANS_TRY(dmalloc(&W->b_qsim, (size_t)grid * cap));
        W->b_grid = grid;
        W->b_cap = (int)cap;
    }
    // two buffers of work counters, A = queue[0, 2) and B = queue[2, 4), zero here;
    // every call leaves A zero (B is cleared by the launch before the one using it)
    if (W->queue_len < 4) {
        drop_graph(W);
        wfree(W->queue);
        W->queue = nullptr;
        NANS_TRY(dmalloc(&W->queue, 4));
        NANS_HIP(hipMemset(W->queue, 0, sizeof(int) * 4));
        W->queue_len = 4;
    }
    a.planes = W->planes;
    a.N = W->N;
    a.offsets = W->offsets;
    a.nboxes = W->nboxes;
    a.order = W->order;
    a.nworlds = W->nworlds;
    a.nbig = sp.nbig;
    a.nmax = W->nmax;
    a.wrows = sp.wrows;
    a.lds_floats = batch_lds_floats(W->nmax, sp, BLOCK);
    a.k = step_consts(ctx->p, ctx->orientation);
    a.fl = W->floor;
    a.wfric = W->wfric;
    a.wrest = W->wrest;
    a.cand = W->b_cand;
    a.rec = W->b_rec;
    a.flag = W->b_flag;
    a.qidx = W->b_qidx;
    a.qsim = W->b_qsim;
    a.cap = W->b_cap;
    a.wcap = W->b_cap / NW;
    // box corners: a workgroup world's, or its waves' worlds side by side
    a.corner_stride = std::max(world_corner_floats(W->nmax), NW * world_corner_floats(sp.wrows));
    if (W->b_corner_floats < a.corner_stride * grid) {
        drop_graph(W);
        wfree(W->b_corners);
        W->b_corners = nullptr;
        W->b_corner_floats = 0;
        NANS_TRY(dmalloc(&W->b_corners, (size_t)(a.corner_stride * grid)));
        W->b_corner_floats = a.corner_stride * grid;
    }
    a.corners = W->b_corners;
    a.rec_out = W->record ? W->rec_buf : nullptr;
    a.rec_off = W->rec_off;
    a.rec_count = W->rec_count;
    a.rec_cap = W->rec_cap;
    a.cnt = W->cnt;
    a.queue = W->queue;
    a.queue_next = nullptr;
    if (split_out) *split_out = sp;
    *grid_out = grid;
    *lds_out = lds;
    return NANS_OK;
}

template <int BLOCK, bool ISL>
static int batched_capture(nans_ctx *ctx, int nsteps, hipStream_t cs)
{
    nans_worlds *W = ctx->worlds;
    BatchArgsT<false> a;
    int grid = 0;
    size_t lds = 0;
    WaveSplit sp;
    NANS_TRY((batched_prepare<BLOCK, false, ISL>(ctx, a, &grid, &lds, &sp)));
    // two steps: the step kernel on counter buffer A, clearing B, then on B, clearing
    // A (no clearing kernel or memset node per step); and one step on A, after which
    // the caller clears A.  A call of nsteps replays the first graph nsteps / 2 times
    // and the second once if nsteps is odd.
    int *const qa = W->queue, *const qb = W->queue + 2;
    auto capture = [&](hipGraphExec_t *out, int steps) -> int {
        hipGraph_t g = nullptr;
        NANS_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < steps; ++k) {
            a.queue = k ? qb : qa;
            a.queue_next = steps == 1 ? nullptr : (k ? qa : qb);
            k_batched_step<BLOCK, false, ISL><<<grid, BLOCK, lds, cs>>>(a);
        }
        hipError_t e = hipStreamEndCapture(cs, &g);
        NANS_REQUIRE(e == hipSuccess, NANS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        NANS_REQUIRE(e == hipSuccess, NANS_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
        return NANS_OK;
    };
    drop_graph(W);
    NANS_TRY(capture(&W->graph, 2));
    NANS_TRY(capture(&W->graph1, 1));
    W->graph_steps = nsteps;
    W->graph_block = BLOCK;
    W->graph_params = ctx->p;
    W->graph_orientation = ctx->orientation;
    W->graph_grid = grid;
    W->graph_lds = (int)lds;
    W->graph_wave = sp.bodies;
    W->graph_nwave = sp.nwave;
    return NANS_OK;
}

int batched_run(nans_ctx *ctx, int nsteps, int block)
{
    nans_worlds *W = ctx->worlds;
    NANS_REQUIRE(W->nmax <= BATCH_MAX_BODIES, NANS_ERR_CAPACITY,
                 "the batched step keeps a world in LDS: at most %d bodies per world (largest: %d)",
                 BATCH_MAX_BODIES, W->nmax);
    hipStream_t cs = ctx->own_stream;
    if (!W->ev_in) {
        NANS_HIP(hipEventCreateWithFlags(&W->ev_in, hipEventDisableTiming));
        NANS_HIP(hipEventCreateWithFlags(&W->ev_out, hipEventDisableTiming));
        NANS_HIP(hipEventCreate(&W->ev_t0));
        NANS_HIP(hipEventCreate(&W->ev_t1));
    }
    bool reuse = W->graph && W->graph_block == block && W->graph_wave == wave_bodies(W, block) &&
                 W->graph_orientation == ctx->orientation &&
                 memcmp(&W->graph_params, &ctx->p, sizeof(nans_params)) == 0;
    if (!reuse) {
        NANS_TRY(setup_floor(ctx, ctx->stream));
        NANS_HIP(hipStreamSynchronize(ctx->stream));
        NANS_TRY(by_shape(block, ctx->p.fp64_islands != 0, [&](auto B, auto I) {
            return batched_capture<decltype(B)::value, decltype(I)::value>(ctx, 1, cs);
        }));
    }
    NANS_HIP(hipEventRecord(W->ev_in, ctx->stream));
    NANS_HIP(hipStreamWaitEvent(cs, W->ev_in, 0));
    // events on the stream the kernels run on bracket the nsteps graph launches
    NANS_HIP(hipEventRecord(W->ev_t0, cs));
    int s = 0;
    for (; s + 2 <= nsteps; s += 2) NANS_HIP(hipGraphLaunch(W->graph, cs));
    if (s < nsteps) {
        NANS_HIP(hipGraphLaunch(W->graph1, cs));
        k_zero_i32<<<1, 64, 0, cs>>>(W->queue, 2);  // A zero again for the next call
    }
    NANS_HIP(hipEventRecord(W->ev_t1, cs));
    NANS_HIP(hipEventRecord(W->ev_out, cs));
    NANS_HIP(hipStreamWaitEvent(ctx->stream, W->ev_out, 0));
    W->timed_steps = nsteps;
    return NANS_OK;
}

// The warm-started batched step: launched step by step (no graph), since each
// step swaps the previous and next contact lists.
template <int BLOCK, bool ISL>
static int batched_warm_launches(nans_ctx *ctx, int nsteps, hipStream_t cs)
{
    nans_worlds *W = ctx->worlds;
    BatchArgsT<true> a;
    int grid = 0;
    size_t lds = 0;
    WaveSplit sp;
    NANS_TRY((batched_prepare<BLOCK, true, ISL>(ctx, a, &grid, &lds, &sp)));
    a.warm_off = W->warm_offd;
    a.warm_cap = W->warm_capd;
    for (int s = 0; s < nsteps; ++s) {
        const int cur = W->warm_flip, nxt = cur ^ 1;
        a.warm_p_prev = W->warm_p[cur];
        a.warm_s_prev = W->warm_s[cur];
        a.warm_n_prev = W->warm_n[cur];
        a.warm_p_next = W->warm_p[nxt];
        a.warm_s_next = W->warm_s[nxt];
        a.warm_n_next = W->warm_n[nxt];
        k_batched_step<BLOCK, true, ISL><<<grid, BLOCK, lds, cs>>>(a);
        k_zero_i32<<<1, 64, 0, cs>>>(a.queue, 2);  // counter buffer A is zero between launches
        NANS_HIP(hipGetLastError());
        W->warm_flip = nxt;
    }
    W->graph_block = BLOCK;
    W->graph_grid = grid;
    W->graph_lds = (int)lds;
    W->graph_wave = sp.bodies;
    W->graph_nwave = sp.nwave;
    return NANS_OK;
}

int batched_run_warm(nans_ctx *ctx, int nsteps, int block)
{
    nans_worlds *W = ctx->worlds;
    NANS_REQUIRE(W->nmax <= BATCH_MAX_BODIES, NANS_ERR_CAPACITY,
                 "the batched step keeps a world in LDS: at most %d bodies per world (largest: %d)",
                 BATCH_MAX_BODIES, W->nmax);
    hipStream_t cs = ctx->own_stream;
    if (!W->ev_in) {
        NANS_HIP(hipEventCreateWithFlags(&W->ev_in, hipEventDisableTiming));
        NANS_HIP(hipEventCreateWithFlags(&W->ev_out, hipEventDisableTiming));
        NANS_HIP(hipEventCreate(&W->ev_t0));
        NANS_HIP(hipEventCreate(&W->ev_t1));
    }
    NANS_TRY(warm_prepare(ctx));
    NANS_TRY(setup_floor(ctx, ctx->stream));
    drop_graph(W);  // the graph_* launch shape below describes these launches
    NANS_HIP(hipEventRecord(W->ev_in, ctx->stream));
    NANS_HIP(hipStreamWaitEvent(cs, W->ev_in, 0));
    NANS_HIP(hipEventRecord(W->ev_t0, cs));
    NANS_TRY(by_shape(block, ctx->p.fp64_islands != 0, [&](auto B, auto I) {
        return batched_warm_launches<decltype(B)::value, decltype(I)::value>(ctx, nsteps, cs);
    }));
    NANS_HIP(hipEventRecord(W->ev_t1, cs));
    NANS_HIP(hipEventRecord(W->ev_out, cs));
    NANS_HIP(hipStreamWaitEvent(ctx->stream, W->ev_out, 0));
    W->timed_steps = nsteps;
    return NANS_OK;
}

// Threads per world: the autotuner's or the caller's choice, else 128 for worlds
// of at most 64 bodies (all of a batch of such worlds fit on the chip at once, and
// a step's latency is one world's) and 256 for larger ones.
int batched_block(nans_ctx *ctx)
{
    const nans_worlds *W = ctx->worlds;
    if (W->tuned_block) return W->tuned_block;
    return W->nmax <= 64 ? 128 : 256;
}

template <int BLOCK, bool ISL>
static int batched_recorded_launch(nans_ctx *ctx, hipStream_t s, ContactRec *rec_out, const int64_t *rec_off,
                                   int *rec_count, const int *rec_cap)
{
    BatchArgsT<false> a;
    int grid = 0;
    size_t lds = 0;
    NANS_TRY((batched_prepare<BLOCK, false, ISL>(ctx, a, &grid, &lds)));
    a.rec_out = rec_out;
    a.rec_off = rec_off;
    a.rec_count = rec_count;
    a.rec_cap = rec_cap;
    k_batched_step<BLOCK, false, ISL><<<grid, BLOCK, lds, s>>>(a);
    k_zero_i32<<<1, 64, 0, s>>>(a.queue, 2);  // counter buffer A is zero between launches
    NANS_HIP(hipGetLastError());
    return NANS_OK;
}

int batched_step_recorded(nans_ctx *ctx, hipStream_t s, ContactRec *rec_out, const int64_t *rec_off, int *rec_count,
                          const int *rec_cap)
{
    nans_worlds *W = ctx->worlds;
    NANS_REQUIRE(W->nmax <= BATCH_MAX_BODIES, NANS_ERR_CAPACITY,
                 "the batched step keeps a world in LDS: at most %d bodies per world (largest: %d)",
                 BATCH_MAX_BODIES, W->nmax);
    return by_shape(batched_block(ctx), ctx->p.fp64_islands != 0, [&](auto B, auto I) {
        return batched_recorded_launch<decltype(B)::value, decltype(I)::value>(ctx, s, rec_out, rec_off, rec_count,
                                                                              rec_cap);
    });
}

}  // namespace nans

extern "C" {

int nans_step_batched(nans_ctx *ctx, int32_t nsteps)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(nsteps >= 0, NANS_ERR_INVALID, "nsteps < 0");
    if (nsteps == 0) return NANS_OK;
    NANS_HIP(hipSetDevice(ctx->p.device));
    nans_worlds *W = ctx->worlds;
    int block = batched_block(ctx);
    NANS_TRY(snapshot_take(ctx, ctx->stream));
    if (ctx->warm) NANS_TRY(batched_run_warm(ctx, nsteps, block));
    else NANS_TRY(batched_run(ctx, nsteps, block));
    NANS_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    NANS_HIP(hipEventElapsedTime(&ms, W->ev_t0, W->ev_t1));
    W->timed_ms = ms;
    return step_finish(ctx);
}

int nans_batched_timing(nans_ctx *ctx, nans_timing *t)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(t, NANS_ERR_INVALID, "timing is NULL");
    nans_worlds *W = ctx->worlds;
    NANS_REQUIRE(W->timed_steps > 0, NANS_ERR_STATE, "no batched step has run");
    t->steps = W->timed_steps;
    t->elapsed_ms = W->timed_ms;
    t->threads_per_world = W->graph_block;
    t->workgroups = W->graph_grid;
    t->lds_bytes = W->graph_lds;
    t->wave_world_bodies = W->graph_wave;
    t->wave_worlds = W->graph_nwave;
    return NANS_OK;
}

int nans_set_batched_block(nans_ctx *ctx, int32_t threads_per_world)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(threads_per_world == 0 || threads_per_world == 64 || threads_per_world == 128 ||
                     threads_per_world == 256 || threads_per_world == 512,
                 NANS_ERR_INVALID, "threads_per_world must be 0, 64, 128, 256 or 512 (got %d)", threads_per_world);
    ctx->worlds->tuned_block = threads_per_world;
    return NANS_OK;
}

int nans_set_wave_worlds(nans_ctx *ctx, int32_t max_bodies)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(max_bodies >= -1 && max_bodies <= BATCH_MAX_BODIES, NANS_ERR_INVALID,
                 "max_bodies must be -1 (default), 0 (none) or at most %d (got %d)", BATCH_MAX_BODIES, max_bodies);
    ctx->worlds->tuned_wave = max_bodies;
    return NANS_OK;
}

// How much faster than the default shape another shape must time to be chosen.
constexpr float TUNE_MARGIN = 0.02f;

int nans_autotune_batched(nans_ctx *ctx, int32_t nsteps, nans_tuning *out)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(!ctx->warm, NANS_ERR_STATE, "the autotuner times the cold batched step: switch warm starting off");
    NANS_REQUIRE(nsteps >= 1, NANS_ERR_INVALID, "nsteps must be >= 1");
    nans_worlds *W = ctx->worlds;
    NANS_REQUIRE(W->nmax <= BATCH_MAX_BODIES, NANS_ERR_CAPACITY,
                 "the batched step keeps a world in LDS: at most %d bodies per world (largest: %d)",
                 BATCH_MAX_BODIES, W->nmax);
    NANS_HIP(hipSetDevice(ctx->p.device));
    hipStream_t s = ctx->stream;
    NANS_HIP(hipStreamSynchronize(s));
    const size_t N = (size_t)W->N;
    // the trials must leave no trace: keep state (pos, vel, angvel, angles), force,
    // torque, the quaternions of the opt-in orientation mode and the counters
    constexpr size_t NSTATE = F_SX, NTAIL = F_NFIELDS - F_FX;  // planes [0, F_SX) and [F_FX, F_NFIELDS)
    float *keep = nullptr;
    NANS_TRY(dmalloc(&keep, (NSTATE + NTAIL) * N));
    unsigned long long cnt[CNT_N];
    NANS_HIP(hipMemcpy(cnt, W->cnt, sizeof(cnt), hipMemcpyDeviceToHost));
    NANS_HIP(hipMemcpyAsync(keep, W->planes, sizeof(float) * NSTATE * N, hipMemcpyDeviceToDevice, s));
    NANS_HIP(hipMemcpyAsync(keep + NSTATE * N, W->planes + (size_t)F_FX * N, sizeof(float) * NTAIL * N,
                            hipMemcpyDeviceToDevice, s));
    const int cand[4] = {64, 128, 256, 512};
    const int bounds[5] = {0, 16, 32, 64, 128};
    int best = 0, best_wave = 0;
    float best_ms = 0.0f;
    // the default shape (batched_block's size, no wave phase) and its time
    const int def_block = W->nmax <= 64 ? 128 : 256;
    float def_ms = -1.0f;
    int rc = NANS_OK;
    int max_lds = 0;
    NANS_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->p.device));
    const int keep_wave = W->tuned_wave;
    if (out) {
        for (int c = 0; c < 4; ++c) {
            out->ms_per_step[c] = -1.0f;
            for (int b = 0; b < 5; ++b) out->ms_per_step_wave[c][b] = -1.0f;
        }
    }
    for (int c = 0; c < 4 && rc == NANS_OK; ++c) {
        int prev_nbig = -1;
        for (int b = 0; b < 5 && rc == NANS_OK; ++b) {
            if (cand[c] == 64 && b > 0) break;  // one wave per workgroup: no wave phase
            const WaveSplit sp = wave_split(W, bounds[b]);
            // the same split as the previous bound, or LDS that does not fit: skipped
            if (sp.nbig == prev_nbig || batch_lds_bytes(W->nmax, sp, cand[c]) > (size_t)max_lds) continue;
            prev_nbig = sp.nbig;
            W->tuned_wave = bounds[b];
            float ms = -1.0f;
            // one untimed step captures the graph; then nsteps timed steps
            int r = batched_run(ctx, 1, cand[c]);
            if (r == NANS_OK) r = batched_run(ctx, nsteps, cand[c]);
            if (r == NANS_OK) {
                NANS_HIP(hipStreamSynchronize(s));
                NANS_HIP(hipEventElapsedTime(&ms, W->ev_t0, W->ev_t1));
                ms /= (float)nsteps;
                if (cand[c] == def_block && bounds[b] == WAVE_WORLD_BODIES) def_ms = ms;
                if (!best || ms < best_ms) {
                    best = cand[c];
                    best_wave = bounds[b];
                    best_ms = ms;
                }
                if (out) {
                    out->ms_per_step_wave[c][b] = ms;
                    if (out->ms_per_step[c] < 0.0f || ms < out->ms_per_step[c]) out->ms_per_step[c] = ms;
                }
            } else {
                rc = r;
            }
            NANS_HIP(hipMemcpyAsync(W->planes, keep, sizeof(float) * NSTATE * N, hipMemcpyDeviceToDevice, s));
            NANS_HIP(hipMemcpyAsync(W->planes + (size_t)F_FX * N, keep + NSTATE * N, sizeof(float) * NTAIL * N,
                                    hipMemcpyDeviceToDevice, s));
        }
    }
    W->tuned_wave = keep_wave;
    NANS_HIP(hipStreamSynchronize(s));
    wfree(keep);
    NANS_HIP(hipMemcpy(W->cnt, cnt, sizeof(cnt), hipMemcpyHostToDevice));
    if (rc != NANS_OK) return rc;
    NANS_REQUIRE(best > 0, NANS_ERR_CAPACITY, "no workgroup size fits a world of %d bodies", W->nmax);
    // a few steps from one state rank shapes within a percent or two of each other
    // by chance (config 3: 256 threads at bound 16 timed 0.1 % ahead of bound 0
    // over its first 5 steps and ran 1.5 % behind it over the next 20), so the
    // default stays unless another shape beats it by TUNE_MARGIN
    if (def_ms > 0.0f && best_ms > def_ms * (1.0f - TUNE_MARGIN)) {
        best = def_block;
        best_wave = WAVE_WORLD_BODIES;
    }
    W->tuned_block = best;
    W->tuned_wave = best_wave;
    W->timed_steps = 0;
    if (out) {
        out->threads_per_world = best;
        out->wave_world_bodies = best_wave;
        out->steps = nsteps;
    }
    return NANS_OK;
}

int nans_set_orientation_mode(nans_ctx *ctx, int32_t mode)
{
    NANS_REQUIRE(ctx, NANS_ERR_INVALID, "ctx is NULL");
    NANS_REQUIRE(mode == NANS_ORIENT_EULER || mode == NANS_ORIENT_QUATERNION, NANS_ERR_INVALID,
                 "orientation mode %d is not one of NANS_ORIENT_*", (int)mode);
    nans_worlds *W = ctx->worlds;
    if (W && mode == NANS_ORIENT_QUATERNION && ctx->orientation != mode) {
        // the quaternions of the current angles
        const int64_t N = W->N;
        NANS_HIP(hipSetDevice(ctx->p.device));
        NANS_HIP(hipStreamSynchronize(ctx->stream));
        std::vector<float> ang((size_t)3 * N), q((size_t)4 * N);
        NANS_HIP(hipMemcpy(ang.data(), W->planes + (size_t)F_AX * N, sizeof(float) * ang.size(),
                           hipMemcpyDeviceToHost));
        euler_to_quat_planes(ang.data(), q.data(), N);
        NANS_HIP(hipMemcpy(W->planes + (size_t)F_QW * N, q.data(), sizeof(float) * q.size(), hipMemcpyHostToDevice));
    }
    ctx->orientation = mode;
    return NANS_OK;
}

int nans_set_quaternions(nans_ctx *ctx, const float *quat)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(quat, NANS_ERR_INVALID, "quat is NULL");
    nans_worlds *W = ctx->worlds;
    NANS_HIP(hipSetDevice(ctx->p.device));
    NANS_HIP(hipStreamSynchronize(ctx->stream));
    NANS_HIP(hipMemcpy(W->planes + (size_t)F_QW * W->N, quat, sizeof(float) * 4 * W->N, hipMemcpyHostToDevice));
    return NANS_OK;
}

int nans_get_quaternions(nans_ctx *ctx, float *quat)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(quat, NANS_ERR_INVALID, "quat is NULL");
    nans_worlds *W = ctx->worlds;
    NANS_HIP(hipSetDevice(ctx->p.device));
    NANS_HIP(hipStreamSynchronize(ctx->stream));
    NANS_HIP(hipMemcpy(quat, W->planes + (size_t)F_QW * W->N, sizeof(float) * 4 * W->N, hipMemcpyDeviceToHost));
    return NANS_OK;
}

int nans_set_state(nans_ctx *ctx, const nans_bodies_soa *b)
{
    NANS_REQUIRE(ctx && ctx->worlds, NANS_ERR_STATE, "no worlds uploaded");
    NANS_REQUIRE(b && b->pos && b->vel && b->angvel && b->angles, NANS_ERR_INVALID, "a state array is NULL");
    nans_worlds *W = ctx->worlds;
    const int64_t N = W->N;
    NANS_HIP(hipSetDevice(ctx->p.device));
    NANS_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<float> h((size_t)12 * N);
    const float *src[4] = {b->pos, b->vel, b->angvel, b->angles};
    for (int f = 0; f < 4; ++f) std::copy(src[f], src[f] + 3 * N, h.begin() + (size_t)3 * f * N);
    NANS_HIP(hipMemcpy(W->planes, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    std::vector<float> ft((size_t)6 * N, 0.0f);
    if (b->force) std::copy(b->force, b->force + 3 * N, ft.begin());
    if (b->torque) std::copy(b->torque, b->torque + 3 * N, ft.begin() + 3 * N);
    NANS_HIP(hipMemcpy(W->planes + (size_t)F_FX * N, ft.data(), sizeof(float) * ft.size(), hipMemcpyHostToDevice));
    // the quaternions of the given angles, as at upload (used by the opt-in quaternion mode)
    std::vector<float> q((size_t)4 * N);
    euler_to_quat_planes(h.data() + (size_t)F_AX * N, q.data(), N);
    NANS_HIP(hipMemcpy(W->planes + (size_t)F_QW * N, q.data(), sizeof(float) * q.size(), hipMemcpyHostToDevice));
    return warm_reset(ctx);
}

int nans_set_warm_start(nans_ctx *ctx, int32_t enable)
{
    NANS_REQUIRE(ctx, NANS_ERR_INVALID, "ctx is NULL");
    NANS_REQUIRE(enable == 0 || enable == 1, NANS_ERR_INVALID, "enable must be 0 or 1");
    if (enable && !ctx->warm) NANS_TRY(warm_reset(ctx));
    ctx->warm = enable;
    return NANS_OK;
}

}  // extern "C"
