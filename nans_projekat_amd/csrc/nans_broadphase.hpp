// nans_broadphase.hpp -- uniform-grid broadphase of the plain step.
//
// The reference tests every pair of a world (code/nans.cpp:1355-1535).  The
// plain step's all-pairs cull (k_plain_cull) keeps the canonical pairs for which
// cull_keep holds and hands them to the narrowphase in canonical order.  The grid
// below finds exactly the same pairs with O(n) expected work and writes them in
// the same order, so the narrowphase, the solver and every result are unchanged.
//
// Why no pair is lost.  A body is tame when |x|+|y|+|z| of its centre and the
// magnitude of its bounding radius are both <= 1e18; the others (NaN, inf, huge)
// are wild and are paired with every body by an exhaustive scan, so only pairs of
// two tame bodies rely on the grid.  For two tame bodies no term of cull_keep
// overflows, and cull_keep(a, b) implies, per axis,
//     |x_a - x_b| <= |lim| (1 + 4u),   |lim| <= (2R * 1.0001 + 1e-4 (1 + 2M + 2R)) (1 + 8u)
// with u = 2^-24, R the largest tame |radius| and M the largest tame
// |x|+|y|+|z|.  The cell size h is that bound's binary64 value times 1.01, so
// |x_a - x_b| / h < 0.991.  Cell coordinates are floor(x * (1/h)) in binary64,
// clamped to +-2^40: below 2^41 cells the product is within 2^-11 of x / h, so
// the two products of a kept pair differ by less than 0.991 + 2^-10 < 1 and
// their floors by at most one, which clamping (1-Lipschitz) keeps; a pair with a
// centre beyond 2^41 cells has both centres beyond 2^40 on the same side, and
// both clamp to the same cell.  Hence the 27 cells around a body's cell hold
// every tame partner that cull_keep accepts.  Cells are hashed into buckets; a bucket entry counts only
// when its own cell equals the searched cell, so each pair is found once.
//
// Output.  Each pair is written by the body that owns its canonical row: a box
// writes its CC row (boxes j > i), its CF flag and its CS row (every sphere); a
// sphere writes its SF flag and its SS row (spheres t > s).  One exclusive scan
// over the per-body counts laid out as [CC rows | CF | SF | CS rows | SS rows]
// gives every write offset in canonical order; each row is then sorted in place.
#pragma once

namespace nans {

constexpr int BP_THREADS = 1024;
constexpr double BP_CLAMP = 1099511627776.0;  // 2^40 cells per axis

struct BpGrid {
    double inv;  // 1 / cell size
    int nwild;
    int pad;
};

__device__ inline float bp_mag(V3 p) { return fabsf(p.x) + fabsf(p.y) + fabsf(p.z); }

__device__ inline bool bp_tame(V3 p, float r) { return bp_mag(p) <= 1e18f && fabsf(r) <= 1e18f; }

__device__ inline long long bp_coord(float x, double inv)
{
    double u = floor((double)x * inv);
    u = fmin(fmax(u, -BP_CLAMP), BP_CLAMP);
    return (long long)u;
}

__device__ inline unsigned bp_hash(long long x, long long y, long long z, unsigned mask)
{
    unsigned long long h = (unsigned long long)x * 0x9E3779B97F4A7C15ull;
    h ^= (unsigned long long)y * 0xC2B2AE3D27D4EB4Full;
    h ^= (unsigned long long)z * 0x165667B19E3779F9ull;
    h ^= h >> 29;
    return (unsigned)h & mask;
}

__device__ inline float bp_wave_max(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// One workgroup: clears the bucket counters, lists the wild bodies and sets the
// cell size from the tame bodies' largest radius and centre magnitude.
__global__ __launch_bounds__(BP_THREADS) void k_bp_bounds(Planes B, const float *brad, int n, int *bucket_cnt,
                                                          int nbuckets, BpGrid *g, int *wild,
                                                          unsigned long long *cnt, long long npairs)
{
    __shared__ float s_r[BP_THREADS / 64], s_m[BP_THREADS / 64];
    __shared__ int s_nw;
    if (threadIdx.x == 0) s_nw = 0;
    for (int k = threadIdx.x; k < nbuckets; k += BP_THREADS) bucket_cnt[k] = 0;
    __syncthreads();
    float rmax = 0.0f, mmax = 0.0f;
    for (int i = threadIdx.x; i < n; i += BP_THREADS) {
        const V3 p = B.get3(F_PX, i);
        const float r = brad[i];
        if (bp_tame(p, r)) {
            rmax = fmaxf(rmax, fabsf(r));
            mmax = fmaxf(mmax, bp_mag(p));
        } else {
            wild[atomicAdd(&s_nw, 1)] = i;
        }
    }
    rmax = bp_wave_max(rmax);
    mmax = bp_wave_max(mmax);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_r[w] = rmax, s_m[w] = mmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < BP_THREADS / 64; ++k) rmax = fmaxf(rmax, s_r[k]), mmax = fmaxf(mmax, s_m[k]);
        const double R = rmax, M = mmax;
        const double lim = 2.0 * R * 1.0001 + 1e-4 * (1.0 + 2.0 * M + 2.0 * R);
        g->inv = 1.0 / (lim * 1.01);
        g->nwild = s_nw;
        atomicAdd(&cnt[CNT_PAIR_TESTS], (unsigned long long)npairs);
    }
}

// Cell and bucket of every tame body (hslot = [bucket][slot in bucket], -1: wild).
__global__ __launch_bounds__(256) void k_bp_cells(Planes B, const float *brad, int n, const BpGrid *g,
                                                  unsigned mask, long long *cell, int stride, int *hslot,
                                                  int *bucket_cnt)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const V3 p = B.get3(F_PX, i);
    if (!bp_tame(p, brad[i])) {
        hslot[i] = -1;
        return;
    }
    const double inv = g->inv;
    const long long cx = bp_coord(p.x, inv), cy = bp_coord(p.y, inv), cz = bp_coord(p.z, inv);
    cell[i] = cx;
    cell[stride + i] = cy;
    cell[2 * stride + i] = cz;
    const unsigned h = bp_hash(cx, cy, cz, mask);
    hslot[i] = (int)h;
    hslot[stride + i] = atomicAdd(&bucket_cnt[h], 1);
}

__global__ __launch_bounds__(256) void k_bp_scatter(int n, const int *hslot, int stride, const int *bucket_off,
                                                    int *sorted)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int h = hslot[i];
    if (h >= 0) sorted[bucket_off[h] + hslot[stride + i]] = i;
}

struct BpRows {
    Planes B;
    const float *brad;
    int nb, ns;
    const FloorDev *fl;
    const BpGrid *g;
    unsigned mask;
    const long long *cell;
    int stride;
    const int *hslot;
    const int *bucket_cnt, *bucket_off, *sorted, *wild;
    int *counts;           // [3 nb + 2 ns] counts (WRITE = false)
    const int *offsets;    // their exclusive scan (WRITE = true)
    int *cand;
    int cap;
};

// Rows of every body: counts them (WRITE = false) or writes the canonical pair
// indices and sorts each row (WRITE = true).
template <bool WRITE>
__global__ __launch_bounds__(256) void k_bp_rows(BpRows a)
{
    __shared__ FloorDev fl;
    if (threadIdx.x == 0) fl = *a.fl;
    __syncthreads();
    const int nb = a.nb, ns = a.ns, n = nb + ns;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const bool cube = i < nb;
    const V3 pi = a.B.get3(F_PX, i);
    const float ri = a.brad[i];
    const int64_t ncc = tri(nb);
    const int64_t base_cs = ncc + nb + ns, base_ss = base_cs + (int64_t)nb * ns;
    const int idx_a = cube ? i : 3 * nb + ns + (i - nb);      // CC row / SS row
    const int idx_b = 2 * nb + ns + i;                        // CS row (boxes)
    const int idx_f = cube ? nb + i : 2 * nb + (i - nb);      // CF / SF flag
    const int oa = WRITE ? a.offsets[idx_a] : 0;
    const int ob = WRITE && cube ? a.offsets[idx_b] : 0;
    // start of this body's row in the canonical CC / SS triangle (the layout tri_decode reads)
    const int64_t row_a = cube ? (int64_t)i * (2 * (int64_t)nb - i - 1) / 2
                               : base_ss + (int64_t)(i - nb) * (2 * (int64_t)ns - (i - nb) - 1) / 2;
    int na = 0, nbr = 0;
    auto emit = [&](int dst, int64_t p) {
        if (WRITE && dst < a.cap) a.cand[dst] = (int)p;
    };
    auto visit = [&](int j) {
        if (j <= i && !(cube && j >= nb)) return;  // rows hold j > i, except a box's CS row
        const V3 pj = a.B.get3(F_PX, j);
        if (!cull_keep(pi, ri, pj, a.brad[j])) return;
        if (cube && j >= nb) {
            emit(ob + nbr, base_cs + (int64_t)i * ns + (j - nb));
            ++nbr;
        } else {
            emit(oa + na, row_a + (j - i - 1));
            ++na;
        }
    };
    if (a.hslot[i] < 0) {
        for (int j = 0; j < n; ++j) visit(j);
    } else {
        const int s = a.stride;
        const long long c0 = a.cell[i], c1 = a.cell[s + i], c2 = a.cell[2 * s + i];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const long long x = c0 + dx, y = c1 + dy, z = c2 + dz;
                    const unsigned h = bp_hash(x, y, z, a.mask);
                    const int e = a.bucket_off[h] + a.bucket_cnt[h];
                    for (int k = a.bucket_off[h]; k < e; ++k) {
                        const int j = a.sorted[k];
                        if (a.cell[j] == x && a.cell[s + j] == y && a.cell[2 * s + j] == z) visit(j);
                    }
                }
        const int nw = a.g->nwild;
        for (int k = 0; k < nw; ++k) visit(a.wild[k]);
    }
    const V3 fc = mk3(fl.c[0], fl.c[1], fl.c[2]);
    const bool fkeep = cull_keep(pi, ri, fc, fl.brad);
    if (!WRITE) {
        a.counts[idx_a] = na;
        if (cube) a.counts[idx_b] = nbr;
        a.counts[idx_f] = fkeep ? 1 : 0;
        return;
    }
    if (fkeep) emit(a.offsets[idx_f], cube ? ncc + i : ncc + nb + (i - nb));
    auto sort_row = [&](int o, int m) {
        const int e = min(o + m, a.cap);
        for (int k = o + 1; k < e; ++k) {
            const int v = a.cand[k];
            int q = k - 1;
            while (q >= o && a.cand[q] > v) {
                a.cand[q + 1] = a.cand[q];
                --q;
            }
            a.cand[q + 1] = v;
        }
    };
    sort_row(oa, na);
    if (cube) sort_row(ob, nbr);
}

}  // namespace nans
