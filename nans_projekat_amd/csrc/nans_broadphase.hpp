// nans_broadphase.hpp -- hierarchical-grid broadphase of the plain step.
//
// The reference tests every pair of a world (code/nans.cpp:1355-1535).  The
// plain step's all-pairs cull (k_plain_cull) keeps the canonical pairs for which
// cull_keep holds and hands them to the narrowphase in canonical order.  The grid
// below finds exactly the same pairs without testing every one and writes them
// in the same order, so the narrowphase, the solver and every result are
// unchanged.
//
// Reach.  cull_keep's limit for a pair grows with the pair's radii and with the
// magnitude of its centres.  Each body gets a reach
//     rho = (|r| * 1.0002 + 1e-4 * (0.5 + |x|+|y|+|z|)) * (1 + 4e-6)      (binary64)
// such that for two tame bodies (below) cull_keep(a, b) implies, per axis,
// |x_a - x_b| <= rho_a + rho_b: the limit is at most
// (|r_a|+|r_b|) * 1.0001 + 1e-4 (1 + mag_a + mag_b + |r_a| + |r_b|) up to binary32
// rounding of its evaluation and of the magnitude sums (a relative 2e-6 at most),
// and the distance test loses at most a further 4 * 2^-24; the factor 1 + 4e-6
// covers both.
//
// Levels.  Level k has cells of size h_k = 2.02 * rho0 * 2^k, where rho0 is the
// reach of the largest body at magnitude 128; a body lives in the lowest level
// with rho <= rho0 * 2^k.  For a kept pair with levels k_a <= k_b,
// |x_a - x_b| <= rho0 (2^k_a + 2^k_b) <= h_k_b / 1.01.  Cell coordinates are
// floor(x * (1 / h_k)) in binary64, clamped to +-2^40: below 2^41 cells the
// product is within 2^-11 of x / h_k, so the two coordinates of a kept pair
// differ by less than 1/1.01 + 2^-10 < 1 and their floors by at most one, which
// clamping (1-Lipschitz) keeps; a pair with a centre beyond 2^41 cells has both
// centres beyond 2^40 on the same side, and both clamp to the same cell.  So a
// body that searches the 27 cells around its centre at its own level and at
// every coarser level finds every partner of the same or a coarser level, and a
// pair is found once: by the finer body, or at equal levels by the body with
// the smaller index.  Cells of all levels share one hash table; a bucket entry
// counts only when its level and cell equal the searched ones.
//
// Wild bodies.  A body is tame when |x|+|y|+|z| and |r| are both <= 1e15 (which
// also bounds the level below 58 and keeps every term of cull_keep finite); the
// others (NaN, inf, huge) are wild: each sweeps every body, taking every tame
// partner and the wild partners of larger index.
//
// Output.  A found pair is credited to the canonical row of its owner (CC: the
// smaller box, CS: the box, SS: the smaller sphere).  Rows are laid out as
// [CC rows | CF flags | SF flags | CS rows | SS rows]; one exclusive scan of the
// per-row counts gives every row's place in canonical order, and each row is
// finally sorted into place.  The counting pass also keeps the first BP_STASH
// pairs of every row; only when a row holds more does a second pass drop every
// pair into its row's place.
#pragma once

namespace nans {

constexpr int BP_THREADS = 1024;
constexpr double BP_CLAMP = 1099511627776.0;  // 2^40 cells per axis
constexpr float BP_TAME = 1e15f;
constexpr int BP_LEVELS = 64;
constexpr int BP_STASH = 32;  // pairs per row kept by the counting pass

struct BpGrid {
    double rho0;                // reach covered by level 0
    double inv0;                // 1 / cell size of level 0; level k: inv0 / 2^k
    unsigned long long levels;  // bit k: level k holds a body
};

__device__ inline float bp_mag(V3 p) { return fabsf(p.x) + fabsf(p.y) + fabsf(p.z); }

__device__ inline bool bp_tame(V3 p, float r) { return bp_mag(p) <= BP_TAME && fabsf(r) <= BP_TAME; }

__device__ inline double bp_reach(V3 p, float r)
{
    return ((double)fabsf(r) * 1.0002 + 1e-4 * (0.5 + (double)bp_mag(p))) * (1.0 + 4e-6);
}

__device__ inline int bp_level(double rho, double rho0)
{
    int k = 0;
    double lim = rho0;
    while (rho > lim && k < BP_LEVELS - 1) {
        lim *= 2.0;
        ++k;
    }
    return k;
}

__device__ inline long long bp_coord(float x, double inv)
{
    double u = floor((double)x * inv);
    u = fmin(fmax(u, -BP_CLAMP), BP_CLAMP);
    return (long long)u;
}

__device__ inline unsigned bp_hash(int level, long long x, long long y, long long z, unsigned mask)
{
    unsigned long long h = (unsigned long long)x * 0x9E3779B97F4A7C15ull;
    h ^= (unsigned long long)y * 0xC2B2AE3D27D4EB4Full;
    h ^= (unsigned long long)z * 0x165667B19E3779F9ull;
    h ^= (unsigned long long)(level + 1) * 0xD6E8FEB86659FD93ull;
    h ^= h >> 29;
    return (unsigned)h & mask;
}

__device__ inline float bp_wave_max(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// Inclusive sum over the 64 lanes of a wave with DPP moves (VALU, no LDS
// crossbar): shifts by 1, 2, 4, 8 within each 16-lane row, then the last lane of
// a row broadcast into the next row (row_bcast:15) and lane 31 into the upper half
// (row_bcast:31).  Lanes a move does not reach add 0 (bound_ctrl off, old = 0).
__device__ inline int wave_incl_sum(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return x;
}

// One workgroup: clears the bucket and row counters and sets the level-0 reach
// from the tame bodies' largest radius.
__global__ __launch_bounds__(BP_THREADS) void k_bp_bounds(Planes B, const float *brad, int n, int *bucket_cnt,
                                                          int nbuckets, int *counts, int ncounts, BpGrid *g,
                                                          int *ovf, unsigned long long *cnt, long long npairs)
{
    if (threadIdx.x == 0) *ovf = 0;
    __shared__ float s_r[BP_THREADS / 64];
    for (int k = threadIdx.x; k < nbuckets; k += BP_THREADS) bucket_cnt[k] = 0;
    for (int k = threadIdx.x; k < ncounts; k += BP_THREADS) counts[k] = 0;
    __syncthreads();
    float rmax = 0.0f;
    for (int i = threadIdx.x; i < n; i += BP_THREADS) {
        const V3 p = B.get3(F_PX, i);
        const float r = brad[i];
        if (bp_tame(p, r)) rmax = fmaxf(rmax, fabsf(r));
    }
    rmax = bp_wave_max(rmax);
    if ((threadIdx.x & 63) == 0) s_r[threadIdx.x >> 6] = rmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < BP_THREADS / 64; ++k) rmax = fmaxf(rmax, s_r[k]);
        const double rho0 = ((double)rmax * 1.0002 + 1e-4 * (0.5 + 128.0)) * (1.0 + 4e-6);
        g->rho0 = rho0;
        g->inv0 = 1.0 / (2.02 * rho0);
        g->levels = 0ull;
        atomicAdd(&cnt[CNT_PAIR_TESTS], (unsigned long long)npairs);
    }
}

// Level, cell and bucket of every tame body: hslot = [bucket][slot][level]
// planes (bucket -1: wild).
__global__ __launch_bounds__(256) void k_bp_cells(Planes B, const float *brad, int n, BpGrid *g, unsigned mask,
                                                  long long *cell, int stride, int *hslot, int *bucket_cnt)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    unsigned long long bit = 0ull;
    if (i < n) {
        const V3 p = B.get3(F_PX, i);
        const float r = brad[i];
        if (!bp_tame(p, r)) {
            hslot[i] = -1;
        } else {
            const int k = bp_level(bp_reach(p, r), g->rho0);
            const double inv = ldexp(g->inv0, -k);
            const long long cx = bp_coord(p.x, inv), cy = bp_coord(p.y, inv), cz = bp_coord(p.z, inv);
            cell[i] = cx;
            cell[stride + i] = cy;
            cell[2 * stride + i] = cz;
            const unsigned h = bp_hash(k, cx, cy, cz, mask);
            hslot[i] = (int)h;
            hslot[stride + i] = atomicAdd(&bucket_cnt[h], 1);
            hslot[2 * stride + i] = k;
            bit = 1ull << k;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) bit |= __shfl_xor(bit, off, 64);
    if ((threadIdx.x & 63) == 0 && bit) atomicOr(&g->levels, bit);
}

__global__ __launch_bounds__(256) void k_bp_scatter(int n, const int *hslot, int stride, const int *bucket_off,
                                                    int *sorted)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int h = hslot[i];
    if (h >= 0) sorted[bucket_off[h] + hslot[stride + i]] = i;
}

struct BpFind {
    Planes B;
    const float *brad;
    int nb, ns;
    const FloorDev *fl;
    const BpGrid *g;
    unsigned mask;
    const long long *cell;
    int stride;
    const int *hslot;
    const int *bucket_cnt, *bucket_off, *sorted;
    int *counts;           // [3 nb + 2 ns] row counts
    const int *offsets;    // their exclusive scan (WRITE)
    int *tmp;              // rows before sorting (same offsets as cand)
    int *cand;
    int cap;
    int *stash;            // [2 nb + ns][BP_STASH] the first pairs of every row
    int *ovf;              // a row holds more than BP_STASH pairs
};

// Stash row of a counts index (CC rows, CS rows, SS rows; flags have none).
__device__ inline int bp_stash_row(int idx, int nb, int ns)
{
    return idx < nb ? idx : (idx < 3 * nb + ns ? nb + (idx - 2 * nb - ns) : 2 * nb + (idx - 3 * nb - ns));
}

// Canonical row (index into the counts layout) and index of the pair {i, j}, and
// its owner a and partner b in cull_keep's argument order.
struct BpPair {
    int row, a, b;
    int64_t p;
};

__device__ inline BpPair bp_pair(int i, int j, int nb, int ns)
{
    BpPair q;
    const bool ci = i < nb, cj = j < nb;
    const int64_t base_cs = tri(nb) + nb + ns;
    if (ci && cj) {
        q.a = min(i, j);
        q.b = max(i, j);
        q.row = q.a;
        q.p = (int64_t)q.a * (2 * (int64_t)nb - q.a - 1) / 2 + (q.b - q.a - 1);
    } else if (ci != cj) {
        q.a = ci ? i : j;
        q.b = ci ? j : i;
        q.row = 2 * nb + ns + q.a;
        q.p = base_cs + (int64_t)q.a * ns + (q.b - nb);
    } else {
        q.a = min(i, j);
        q.b = max(i, j);
        const int sa = q.a - nb, sb = q.b - nb;
        q.row = 3 * nb + ns + sa;
        q.p = base_cs + (int64_t)nb * ns + (int64_t)sa * (2 * (int64_t)ns - sa - 1) / 2 + (sb - sa - 1);
    }
    return q;
}

// Pairs of every body, one wave per body: counts them into their rows
// (WRITE = false) or drops their canonical indices into tmp (WRITE = true).
template <bool WRITE>
__global__ __launch_bounds__(256) void k_bp_find(BpFind a)
{
    if (WRITE && *a.ovf == 0) return;  // every row fitted the stash
    __shared__ FloorDev fl;
    if (threadIdx.x == 0) fl = *a.fl;
    __syncthreads();
    const int nb = a.nb, ns = a.ns, n = nb + ns;
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    // one candidate partner per lane (j < 0: none)
    auto take = [&](int j) {
        if (j < 0) return;
        const BpPair q = bp_pair(i, j, nb, ns);
        if (!cull_keep(a.B.get3(F_PX, q.a), a.brad[q.a], a.B.get3(F_PX, q.b), a.brad[q.b])) return;
        if (!WRITE) {
            const int k = atomicAdd(&a.counts[q.row], 1);
            if (k < BP_STASH) a.stash[bp_stash_row(q.row, nb, ns) * BP_STASH + k] = (int)q.p;
            else *a.ovf = 1;
        } else {
            const int dst = a.offsets[q.row] + atomicSub(&a.counts[q.row], 1) - 1;
            if (dst < a.cap) a.tmp[dst] = (int)q.p;
        }
    };
    const int s = a.stride;
    if (a.hslot[i] < 0) {
        for (int base = 0; base < n; base += 64) {
            const int j = base + lane;
            take(j < n && j != i && (a.hslot[j] >= 0 || j > i) ? j : -1);
        }
    } else {
        const V3 pi = a.B.get3(F_PX, i);
        const int ki = a.hslot[2 * s + i];
        unsigned long long lv = (a.g->levels >> ki) << ki;
        while (lv) {
            const int k = __ffsll((long long)lv) - 1;
            lv &= lv - 1ull;
            const double inv = ldexp(a.g->inv0, -k);
            long long x = 0, y = 0, z = 0;
            int rs = 0, rc = 0;
            if (lane < 27) {
                x = bp_coord(pi.x, inv) + (lane % 3) - 1;
                y = bp_coord(pi.y, inv) + (lane / 3) % 3 - 1;
                z = bp_coord(pi.z, inv) + lane / 9 - 1;
                const unsigned h = bp_hash(k, x, y, z, a.mask);
                rs = a.bucket_off[h];
                rc = a.bucket_cnt[h];
            }
            const int incl = wave_incl_sum(rc);  // inclusive scan of the cells' entry counts
            const int total = __shfl(incl, 63, 64);
            for (int base = 0; base < total; base += 64) {
                const int t = base + lane;
                int owner = 0;  // the cell lane holding entry t
                for (int q = 0; q < 27; ++q) owner += __shfl(incl, q, 64) <= t ? 1 : 0;
                owner = min(owner, 26);
                const int ostart = __shfl(incl - rc, owner, 64), ors = __shfl(rs, owner, 64);
                const long long ox = __shfl(x, owner, 64), oy = __shfl(y, owner, 64), oz = __shfl(z, owner, 64);
                int j = -1;
                if (t < total) {
                    const int c = a.sorted[ors + (t - ostart)];
                    if (a.hslot[2 * s + c] == k && (k > ki || c > i) && a.cell[c] == ox && a.cell[s + c] == oy &&
                        a.cell[2 * s + c] == oz)
                        j = c;
                }
                take(j);
            }
        }
    }
    if (WRITE || lane != 0) return;
    const bool cube = i < nb;
    const int idx_f = cube ? nb + i : 2 * nb + (i - nb);  // CF / SF flag (written by k_bp_sort)
    a.counts[idx_f] = cull_keep(a.B.get3(F_PX, i), a.brad[i], mk3(fl.c[0], fl.c[1], fl.c[2]), fl.brad) ? 1 : 0;
}

// Sorts every CC / CS / SS row into cand, from the stash or, after an overflow,
// from tmp, one wave per row (rank sort: the entries of a row are distinct); the
// waves after the rows write the CF / SF entries, one flag per lane.
__global__ __launch_bounds__(256) void k_bp_sort(int nb, int ns, const int *offsets, const int *ncand, const int *tmp,
                                                 const int *stash, const int *ovf, int *cand, int cap)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nrows = 2 * nb + ns;
    if (r >= nrows) {
        const int f = (r - nrows) * 64 + lane;  // flag f: counts index nb + f, canonical index tri(nb) + f
        if (f < nb + ns) {
            const int idx = nb + f;
            const int o = offsets[idx];
            const int end = idx + 1 < 3 * nb + 2 * ns ? offsets[idx + 1] : *ncand;
            if (end > o && o < cap) cand[o] = (int)(tri(nb) + f);
        }
        return;
    }
    const int idx = r < nb ? r : (r < 2 * nb ? 2 * nb + ns + (r - nb) : 3 * nb + ns + (r - 2 * nb));
    const int L = 3 * nb + 2 * ns;
    const int o = offsets[idx];
    const int end = min(idx + 1 < L ? offsets[idx + 1] : *ncand, cap);
    const int e = end - o;
    // the stash row r holds this row's entries when no row overflowed
    const int *src = *ovf ? tmp + o : stash + (size_t)r * BP_STASH;
    for (int c = 0; c < e; c += 64) {
        const int k = c + lane;
        const int v = k < e ? src[k] : INT_MAX;
        int rank = 0;
        for (int d = 0; d < e; d += 64) {
            const int w = d + lane < e ? src[d + lane] : INT_MAX;
            for (int q = 0; q < 64; ++q) rank += __shfl(w, q, 64) < v ? 1 : 0;
        }
        if (k < e) cand[o + rank] = v;
    }
}

// Exclusive scan of n <= PER * SCAN_THREADS counts by one workgroup.  Element
// e = k * SCAN_THREADS + t is loaded by thread t (coalesced); the wave-level scans of
// all PER columns run in registers, their (column, wave) totals are scanned by one
// wave, and every element gets its offset after two barriers.  With total !=
// nullptr it records the candidate count and its overflow like k_scan.
template <int PER>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_cols(const int *counts, int n, int *offsets, int *total,
                                                            int cap, unsigned long long *cnt)
{
    constexpr int NWV = SCAN_THREADS / 64;
    __shared__ int wt[PER * NWV];
    __shared__ int s_tot;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int v[PER], x[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = k * SCAN_THREADS + t;
        v[k] = e < n ? counts[e] : 0;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int y = wave_incl_sum(v[k]);
        x[k] = y;
        if (lane == 63) wt[k * NWV + w] = y;
    }
    __syncthreads();
    if (w == 0) {
        int run = 0;
        for (int base = 0; base < PER * NWV; base += 64) {
            const int i = base + lane;
            const int val = i < PER * NWV ? wt[i] : 0;
            const int y = wave_incl_sum(val);
            if (i < PER * NWV) wt[i] = run + y - val;
            run += __shfl(y, 63, 64);
        }
        if (lane == 0) s_tot = run;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = k * SCAN_THREADS + t;
        if (e < n) offsets[e] = wt[k * NWV + w] + x[k] - v[k];
    }
    if (t == 0 && total) {
        const int tot = s_tot;
        *total = min(tot, cap);
        atomicAdd(&cnt[CNT_CANDIDATES], (unsigned long long)tot);
        if (tot > cap) atomicAdd(&cnt[CNT_CAND_OVERFLOW], 1ull);
    }
}

// k_scan_cols for n <= 48 * SCAN_THREADS, k_scan beyond.
inline void scan_counts(const int *counts, int n, int *offsets, int *total, int cap, unsigned long long *cnt,
                        hipStream_t s)
{
    if (n <= 8 * SCAN_THREADS)
        k_scan_cols<8><<<1, SCAN_THREADS, 0, s>>>(counts, n, offsets, total, cap, cnt);
    else if (n <= 16 * SCAN_THREADS)
        k_scan_cols<16><<<1, SCAN_THREADS, 0, s>>>(counts, n, offsets, total, cap, cnt);
    else if (n <= 48 * SCAN_THREADS)
        k_scan_cols<48><<<1, SCAN_THREADS, 0, s>>>(counts, n, offsets, total, cap, cnt);
    else
        k_scan<<<1, SCAN_THREADS, 0, s>>>(counts, n, offsets, total, cap, cnt);
}

}  // namespace nans
