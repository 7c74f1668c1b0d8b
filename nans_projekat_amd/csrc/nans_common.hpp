// nans_common.hpp -- context, error reporting and small helpers shared by the
// C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/nans_gpu.h"
#include "nans_epa.hpp"

namespace nans {

void set_error(const char *fmt, ...);


#define NANS_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            ::nans::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                              __LINE__);                                                     \
            return NANS_ERR_HIP;                                                             \
        }                                                                                    \
    } while (0)

#define NANS_REQUIRE(cond, code, ...)   \
    do {                                \
        if (!(cond)) {                  \
            ::nans::set_error(__VA_ARGS__); \
            return (code);              \
        }                               \
    } while (0)

#define NANS_TRY(x)             \
    do {                        \
        int rc_ = (x);          \
        if (rc_) return rc_;    \
    } while (0)

inline NarrowParams narrow_params(const nans_params &p)
{
    NarrowParams q;
    q.max_gjk_iterations = p.max_gjk_iterations;
    q.max_epa_iterations = p.max_epa_iterations;
    q.epa_tolerance = p.epa_tolerance;
    q.fp64_islands = p.fp64_islands;
    return q;
}

}  // namespace nans

struct nans_worlds;   // device-resident world batch (nans_step.hip)

namespace nans {
int worlds_release(nans_ctx *ctx);  // frees the world batch of a context
}

struct nans_ctx {
    nans_params p;
    hipStream_t stream;          // stream the calls are ordered on
    hipStream_t own_stream;      // private stream for graph capture / replay
    unsigned long long *d_counters;  // small device scratch counters
    nans_worlds *worlds;
    int64_t np_cap;                  // narrowphase entry point: queue capacity (pairs)
    int *np_qidx;                    // queued GJK hits
    void *np_qsim;                   // their simplices (SimplexRec)
    void *na_scratch;                // this device's scratch of nans_narrowphase_noalloc
    int broadphase;                  // NANS_BROADPHASE_* of the plain step
    int orientation;                 // NANS_ORIENT_* of both steps
    int warm;                        // opt-in warm start of the plain and batched steps
};
