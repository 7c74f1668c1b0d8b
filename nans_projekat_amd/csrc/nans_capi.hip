// nans_capi.hip -- context management and the narrowphase entry point of the
// C ABI declared in include/nans_gpu.h.
#include "nans_common.hpp"

#include <mutex>
#include <new>

namespace nans {

static thread_local char g_error[1024];

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

struct NarrowIn {
    const int32_t *type;
    const float *pos, *quat, *half;
};

struct NarrowOutDev {
    int32_t *gjk, *contact, *status;
    float *normal, *pa, *pb;
    double *depth;
};

__device__ inline Shape load_shape(const NarrowIn &s, int64_t i, int64_t n)
{
    Shape sh;
    sh.vm = nullptr;
    sh.kind = s.type[i];
    sh.c = mk3(s.pos[i], s.pos[n + i], s.pos[2 * n + i]);
    V3 h = mk3(s.half[i], s.half[n + i], s.half[2 * n + i]);
    sh.r = h.x;
    if (sh.kind == SHAPE_BOX) {
        box_from_quat(sh.c, s.quat[i], s.quat[n + i], s.quat[2 * n + i], s.quat[3 * n + i], h, sh.v);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) sh.v[k] = mk3(0.0f, 0.0f, 0.0f);
    }
    return sh;
}

__device__ inline void write_pair(const NarrowOutDev &o, int64_t i, int64_t n, int g, int status, V3 nrm, V3 pa,
                                  V3 pb, double depth)
{
    if (o.gjk) o.gjk[i] = g;
    if (o.contact) o.contact[i] = status == EPA_CONTACT ? 1 : 0;
    if (o.status) o.status[i] = status;
    if (o.normal) { o.normal[i] = nrm.x; o.normal[n + i] = nrm.y; o.normal[2 * n + i] = nrm.z; }
    if (o.pa) { o.pa[i] = pa.x; o.pa[n + i] = pa.y; o.pa[2 * n + i] = pa.z; }
    if (o.pb) { o.pb[i] = pb.x; o.pb[n + i] = pb.y; o.pb[2 * n + i] = pb.z; }
    if (o.depth) o.depth[i] = depth;
}

// Candidates of one workgroup: pairs [base, base + count) of the input arrays.
struct PairSrc {
    NarrowIn a, b;
    int64_t base, n;
    NarrowOutDev o;
    unsigned long long *overflow;
    static constexpr bool kMemCorners = false;
    __device__ int key(int c) const { return c; }
    __device__ void shapes_ref(int key, Shape &A, Shape &B) const { shapes(key, A, B); }
    __device__ void shapes(int key, Shape &A, Shape &B) const
    {
        A = load_shape(a, base + key, n);
        B = load_shape(b, base + key, n);
    }
    __device__ void gjk_result(int c, int g) const
    {
        const V3 z = mk3(0.0f, 0.0f, 0.0f);
        write_pair(o, base + c, n, g, EPA_NONE, z, z, z, 0.0);
    }
    __device__ void epa_result(int c, int key, const EpaResult &r) const
    {
        write_pair(o, base + c, n, GJK_FOUND, r.status, r.n, r.pa, r.pb, r.depth);
        if (r.status == EPA_OVERFLOW) atomicAdd(overflow, 1ull);
    }
};

// Pairs per workgroup of the narrowphase entry point, and the kernel's occupancy
// floor: three waves per SIMD (168 VGPRs, a few spilled) and short ranges, so
// that slow EPA ranges are spread over many workgroups; on 1M random pairs this
// runs 5.1 ms against 6.8 ms at two waves per SIMD and 1024-pair ranges.
constexpr int NP_RANGE = 256;

// ISL: the FP64-islands mode (grouped EPA tiers with a binary64 final face).
template <bool ISL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_narrowphase(
    int64_t n, NarrowIn a, NarrowIn b, NarrowParams p, NarrowOutDev o, unsigned long long *overflow, int *qidx,
    SimplexRec *qsim)
{
    __shared__ EpaLds lds[4];
    __shared__ NarrowShared sh;
    const int64_t base = (int64_t)blockIdx.x * NP_RANGE;
    const int count = (int)min((int64_t)NP_RANGE, n - base);
    PairSrc src{a, b, base, n, o, overflow};
    narrow_block<256, true, ISL>(src, count, p, lds, sh, qidx + base, qsim + base);
    if (threadIdx.x == 0 && sh.stalled) atomicAdd(overflow, (unsigned long long)sh.stalled);
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Scratch of nans_narrowphase_noalloc, one copy per device for the life of the
// process: the EPA overflow counter (in its own 256-byte slot), then the EPA queue
// of k_narrowphase -- a candidate index and a GJK simplex per pair -- for up to
// NANS_NARROWPHASE_NOALLOC_MAX_PAIRS pairs (26 MiB in all).
constexpr size_t NA_CNT_BYTES = 256;
constexpr size_t NA_QIDX_BYTES = sizeof(int) * (size_t)NANS_NARROWPHASE_NOALLOC_MAX_PAIRS;
constexpr size_t NA_QSIM_BYTES = sizeof(SimplexRec) * (size_t)NANS_NARROWPHASE_NOALLOC_MAX_PAIRS;
constexpr size_t NA_BYTES = NA_CNT_BYTES + NA_QIDX_BYTES + NA_QSIM_BYTES;
static_assert(NA_QIDX_BYTES % 256 == 0, "the simplex records start 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// Device address of each device's copy of scratch[], looked up once per device
// (with that device current) by nans_create.  The name scratch is a host-side
// stand-in in host code; only these addresses are passed to kernels.
constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[NA_MAX_DEVICES];

static int noalloc_scratch_init(int device, void **out)
{
    NANS_REQUIRE(device >= 0 && device < NA_MAX_DEVICES, NANS_ERR_INVALID, "device %d beyond %d", device,
                 NA_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        NANS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        NANS_REQUIRE(ptr, NANS_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[device] = ptr;
    }
    *out = g_scratch_dev[device];
    return NANS_OK;
}

}  // namespace nans

using namespace nans;

extern "C" {

const char *nans_last_error(void) { return g_error; }

void nans_default_params(nans_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->solver_iterations = 70;
    p->beta = 0.3f;
    p->restitution = (float)0.1;
    p->friction = 0.1f;
    p->damping = 1.5f;
    p->gravity = 9.81f;
    p->dt = 1.0f / 60.0f;
    p->max_gjk_iterations = 64;
    p->max_epa_iterations = 64;
    p->epa_tolerance = 0.001f;
    p->fp64_islands = 0;
    p->floor_pos[0] = 1.2f;
    p->floor_pos[1] = -0.5f;
    p->floor_pos[2] = 1.0f;
    p->floor_scale[0] = 100.0f;
    p->floor_scale[1] = 1.0f;
    p->floor_scale[2] = 100.0f;
    p->floor_mass = 100000.0f;
    p->floor_moi = (p->floor_mass / 12.0f) * (2.0f * 100.0f * 100.0f);
    p->device = 0;
}

static int check_params(const nans_params *p)
{
    NANS_REQUIRE(p, NANS_ERR_INVALID, "params is NULL");
    NANS_REQUIRE(p->solver_iterations >= 0 && p->solver_iterations <= 100000, NANS_ERR_INVALID,
                 "solver_iterations out of range");
    NANS_REQUIRE(p->max_gjk_iterations >= 0 && p->max_gjk_iterations <= 100000, NANS_ERR_INVALID,
                 "max_gjk_iterations out of range");
    NANS_REQUIRE(p->max_epa_iterations >= 0 && p->max_epa_iterations <= 100000, NANS_ERR_INVALID,
                 "max_epa_iterations out of range");
    NANS_REQUIRE(p->fp64_islands == 0 || p->fp64_islands == 1, NANS_ERR_INVALID, "fp64_islands must be 0 or 1");
    NANS_REQUIRE(p->floor_mass > 0.0f && p->floor_moi > 0.0f, NANS_ERR_INVALID, "floor mass/moi must be positive");
    return NANS_OK;
}

int nans_create(nans_ctx **out, const nans_params *p)
{
    NANS_REQUIRE(out, NANS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    nans_params q;
    if (p) q = *p;
    else nans_default_params(&q);
    int rc = check_params(&q);
    if (rc) return rc;
    int ndev = 0;
    NANS_HIP(hipGetDeviceCount(&ndev));
    NANS_REQUIRE(q.device >= 0 && q.device < ndev, NANS_ERR_INVALID, "device %d not present (%d devices)", q.device, ndev);
    NANS_HIP(hipSetDevice(q.device));
    nans_ctx *c = new (std::nothrow) nans_ctx();
    NANS_REQUIRE(c, NANS_ERR_CAPACITY, "out of host memory");
    c->p = q;
    c->stream = nullptr;
    c->worlds = nullptr;
    c->np_cap = 0;
    c->np_qidx = nullptr;
    c->np_qsim = nullptr;
    c->na_scratch = nullptr;
    rc = noalloc_scratch_init(q.device, &c->na_scratch);
    if (rc) {
        delete c;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error("hipStreamCreateWithFlags: %s", hipGetErrorString(e));
        return NANS_ERR_HIP;
    }
    e = hipMalloc((void **)&c->d_counters, 64 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        hipStreamDestroy(c->own_stream);
        delete c;
        set_error("hipMalloc: %s", hipGetErrorString(e));
        return NANS_ERR_HIP;
    }
    *out = c;
    return NANS_OK;
}

int nans_destroy(nans_ctx *ctx)
{
    if (!ctx) return NANS_OK;
    NANS_HIP(hipSetDevice(ctx->p.device));
    NANS_HIP(hipDeviceSynchronize());
    worlds_release(ctx);
    hipFree(ctx->d_counters);
    if (ctx->np_qidx) hipFree(ctx->np_qidx);
    if (ctx->np_qsim) hipFree(ctx->np_qsim);
    hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return NANS_OK;
}

int nans_set_params(nans_ctx *ctx, const nans_params *p)
{
    NANS_REQUIRE(ctx, NANS_ERR_INVALID, "ctx is NULL");
    int rc = check_params(p);
    if (rc) return rc;
    NANS_REQUIRE(p->device == ctx->p.device, NANS_ERR_INVALID, "the device of a context cannot change");
    ctx->p = *p;
    return NANS_OK;
}

int nans_set_stream(nans_ctx *ctx, void *hip_stream)
{
    NANS_REQUIRE(ctx, NANS_ERR_INVALID, "ctx is NULL");
    ctx->stream = (hipStream_t)hip_stream;
    return NANS_OK;
}

int nans_set_broadphase(nans_ctx *ctx, int32_t mode)
{
    NANS_REQUIRE(ctx, NANS_ERR_INVALID, "ctx is NULL");
    NANS_REQUIRE(mode == NANS_BROADPHASE_AUTO || mode == NANS_BROADPHASE_ALL_PAIRS || mode == NANS_BROADPHASE_GRID,
                 NANS_ERR_INVALID, "broadphase mode %d is not one of NANS_BROADPHASE_*", (int)mode);
    ctx->broadphase = mode;
    return NANS_OK;
}

}  // extern "C"

namespace nans {

static int narrow_check(const nans_ctx *ctx, int64_t n, const nans_shapes_soa *a, const nans_shapes_soa *b,
                        const nans_narrow_out *out)
{
    NANS_REQUIRE(ctx, NANS_ERR_INVALID, "ctx is NULL");
    NANS_REQUIRE(n >= 0, NANS_ERR_INVALID, "n < 0");
    NANS_REQUIRE(n <= (int64_t)1 << 31, NANS_ERR_CAPACITY, "n = %lld exceeds 2^31 pairs", (long long)n);
    NANS_REQUIRE(a && b && out, NANS_ERR_INVALID, "shape or output descriptor is NULL");
    NANS_REQUIRE(n == 0 || (a->type && a->pos && a->quat && a->half && b->type && b->pos && b->quat && b->half),
                 NANS_ERR_INVALID, "a shape array is NULL");
    return NANS_OK;
}

// k_narrowphase over n pairs on the context's stream with the given overflow
// counter and EPA queue (n entries), then the counter to the host.
static int narrow_launch(nans_ctx *ctx, int64_t n, const nans_shapes_soa *a, const nans_shapes_soa *b,
                         const nans_narrow_out *out, unsigned long long *cnt, int *qidx, SimplexRec *qsim,
                         int64_t *n_overflow)
{
    hipStream_t s = ctx->stream;
    k_fill_u64<<<1, 64, 0, s>>>(cnt, 1, 0ull);
    NANS_HIP(hipGetLastError());
    if (n > 0) {
        NarrowIn ia{a->type, a->pos, a->quat, a->half};
        NarrowIn ib{b->type, b->pos, b->quat, b->half};
        NarrowOutDev o{out->gjk, out->contact, out->status, out->normal, out->point_a, out->point_b, out->depth};
        unsigned grid = (unsigned)((n + NP_RANGE - 1) / NP_RANGE);
        if (ctx->p.fp64_islands)
            k_narrowphase<true><<<grid, 256, 0, s>>>(n, ia, ib, narrow_params(ctx->p), o, cnt, qidx, qsim);
        else
            k_narrowphase<false><<<grid, 256, 0, s>>>(n, ia, ib, narrow_params(ctx->p), o, cnt, qidx, qsim);
        NANS_HIP(hipGetLastError());
    }
    unsigned long long ov = 0;
    NANS_HIP(hipMemcpyAsync(&ov, cnt, sizeof(ov), hipMemcpyDeviceToHost, s));
    NANS_HIP(hipStreamSynchronize(s));
    if (n_overflow) *n_overflow = (int64_t)ov;
    return NANS_OK;
}

}  // namespace nans

extern "C" {

int nans_narrowphase(nans_ctx *ctx, int64_t n, const nans_shapes_soa *a, const nans_shapes_soa *b,
                     const nans_narrow_out *out, int64_t *n_overflow)
{
    NANS_TRY(narrow_check(ctx, n, a, b, out));
    NANS_HIP(hipSetDevice(ctx->p.device));
    if (n > ctx->np_cap) {
        hipFree(ctx->np_qidx);
        hipFree(ctx->np_qsim);
        ctx->np_qidx = nullptr;
        ctx->np_qsim = nullptr;
        ctx->np_cap = 0;
        NANS_HIP(hipMalloc((void **)&ctx->np_qidx, sizeof(int) * (size_t)n));
        NANS_HIP(hipMalloc((void **)&ctx->np_qsim, sizeof(SimplexRec) * (size_t)n));
        ctx->np_cap = n;
    }
    return narrow_launch(ctx, n, a, b, out, ctx->d_counters, ctx->np_qidx, (SimplexRec *)ctx->np_qsim, n_overflow);
}

int nans_narrowphase_noalloc(nans_ctx *ctx, int64_t n, const nans_shapes_soa *a, const nans_shapes_soa *b,
                             const nans_narrow_out *out, int64_t *n_overflow)
{
    NANS_TRY(narrow_check(ctx, n, a, b, out));
    NANS_REQUIRE(n <= NANS_NARROWPHASE_NOALLOC_MAX_PAIRS, NANS_ERR_CAPACITY,
                 "n = %lld exceeds the %d pairs of the no-allocation scratch; split the call",
                 (long long)n, NANS_NARROWPHASE_NOALLOC_MAX_PAIRS);
    NANS_REQUIRE(ctx->na_scratch, NANS_ERR_STATE, "the context has no scratch address");
    NANS_HIP(hipSetDevice(ctx->p.device));
    unsigned char *base = static_cast<unsigned char *>(ctx->na_scratch);
    return narrow_launch(ctx, n, a, b, out, reinterpret_cast<unsigned long long *>(base),
                         reinterpret_cast<int *>(base + NA_CNT_BYTES),
                         reinterpret_cast<SimplexRec *>(base + NA_CNT_BYTES + NA_QIDX_BYTES), n_overflow);
}

}  // extern "C"
