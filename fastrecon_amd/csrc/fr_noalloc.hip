// Gridding without device allocation: fr_grid_execute_noalloc.
//
// fr_grid_execute keeps the partial sums of heavy units' slices (the only
// device memory a gridding pass needs beyond its inputs and outputs) in a
// buffer the plan allocates on first use and grows with the coil count.
// Callers that cannot allocate (code under HIP graph capture, tight serving
// loops) use this entry point instead: the same kernels, with those partial
// sums in `scratch`, a file-scope device array of FR_SCRATCH_BYTES that each
// GPU holds for the life of the process.
//
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU (several contexts or streams at once) race.  Each pass writes every
// partial it reads before reading it, so nothing needs clearing between
// calls.
#include "fr_internal.h"

#include <mutex>

namespace {

constexpr size_t FR_SCRATCH_BYTES = size_t(32) << 20;     // 32 MiB per GPU
constexpr int FR_SCRATCH_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[FR_SCRATCH_BYTES];

std::mutex g_scratch_mutex;
void *g_scratch_ptr[FR_SCRATCH_DEVICES] = {};

}  // namespace

// Called by fr_ctx_create with the context's device current: the device
// address of this device's copy of `scratch`, looked up once per device
// (the host-side name is not a device address).
fr_status fr_scratch_init(fr_ctx *ctx)
{
    FR_CHECK(ctx, ctx->device >= 0 && ctx->device < FR_SCRATCH_DEVICES, FR_ERR_UNSUPPORTED,
             "device %d beyond the %d the scratch table holds", ctx->device, FR_SCRATCH_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[ctx->device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        FR_CHECK(ctx, e == hipSuccess && ptr != nullptr, FR_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                 hipGetErrorString(e));
        g_scratch_ptr[ctx->device] = ptr;
    }
    ctx->scratch = g_scratch_ptr[ctx->device];
    return FR_OK;
}

extern "C" fr_status fr_grid_execute_noalloc(fr_plan *p, const fr_samples *s, fr_grids *out)
{
    if (!p || !s || !out) return FR_ERR_INVALID;
    fr_ctx *ctx = p->ctx;
    fr_status e = fr_grid_check(p, s, out);
    if (e) return e;
    FR_CHECK(ctx, ctx->scratch != nullptr, FR_ERR_STATE, "the context has no scratch (fr_ctx_create did not set it)");
    const size_t need = fr_grid_scratch_bytes(p, s->C);
    FR_CHECK(ctx, need <= FR_SCRATCH_BYTES, FR_ERR_UNSUPPORTED,
             "fr_grid_execute_noalloc: the plan's heavy-unit slices need %zu bytes of scratch for %d coils, more "
             "than the %zu bytes of static scratch (use fr_grid_execute, or a plan with fewer slices)",
             need, s->C, FR_SCRATCH_BYTES);
    return fr_launch_grid(p, s, out->data, 0, (float2 *)ctx->scratch, FR_SCRATCH_BYTES);
}
