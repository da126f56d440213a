// Launch-level machinery on top of the kernels: one reconstruction pass
// (gridding, then the cropped transforms with the coil sum of squares in their
// last pass) captured into a HIP graph and
// replayed, and an autotuner that times candidate unit shapes / slice sizes
// on the actual trajectory and keeps the fastest plan settings.
//
// Neither has a counterpart in the reference (its CUDA path launches one
// kernel per coil and partition with a device synchronisation after each,
// cuGridLut.cu:261-309).  With batched frames (fr_samples_create_frames) one
// captured pass grids every frame of a dynamic series in one launch.
#include "fr_internal.h"

#include <algorithm>
#include <vector>

struct fr_graph {
    fr_ctx *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// FFT only, or FFT + shift + crop + SOS through the cropped transforms
// (fr_fft_sos: the grids are consumed, which a replayed pass does not mind
// since gridding rewrites them).
static fr_status image_stage(fr_fft *f, fr_grids *g, unsigned rcx, unsigned rcy, unsigned rcz, int apply_shift,
                             const float *deapod_dev, float *out_dev)
{
    if (!f) return FR_OK;
    if (!out_dev) return fr_fft_execute(f, g);
    return fr_fft_sos(f, g, rcx, rcy, rcz, apply_shift, deapod_dev, out_dev);
}

extern "C" fr_status fr_graph_create(fr_plan *p, const fr_samples *s, fr_grids *g, fr_fft *f, unsigned rcx,
                                     unsigned rcy, unsigned rcz, int apply_shift, const float *deapod_dev,
                                     float *out_dev, fr_graph **out)
{
    if (!p || !s || !g || !out) return FR_ERR_INVALID;
    fr_ctx *ctx = p->ctx;
    *out = nullptr;
    FR_CHECK(ctx, out_dev == nullptr || f != nullptr, FR_ERR_INVALID, "the SOS stage needs the FFT stage");
    // One eager pass first: it performs every lazy allocation (slice
    // partials) and validates the arguments, so the captured pass only
    // enqueues kernels.
    fr_status st = fr_grid_execute(p, const_cast<fr_samples *>(s), g);
    if (!st) st = image_stage(f, g, rcx, rcy, rcz, apply_shift, deapod_dev, out_dev);
    if (st) return st;
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));

    fr_graph *gr = new (std::nothrow) fr_graph;
    if (!gr) return FR_ERR_NOMEM;
    gr->ctx = ctx;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete gr;
        return fr_fail(ctx, FR_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    st = fr_grid_execute(p, const_cast<fr_samples *>(s), g);
    if (!st) st = image_stage(f, g, rcx, rcy, rcz, apply_shift, deapod_dev, out_dev);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(ctx->stream, &graph);
    if (st || e != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        delete gr;
        return st ? st : fr_fail(ctx, FR_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    gr->graph = graph;
    e = hipGraphInstantiate(&gr->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(graph);
        delete gr;
        return fr_fail(ctx, FR_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    *out = gr;
    return FR_OK;
}

extern "C" fr_status fr_graph_launch(fr_graph *gr)
{
    if (!gr) return FR_ERR_INVALID;
    FR_HIP(gr->ctx, hipGraphLaunch(gr->exec, gr->ctx->stream));
    return FR_OK;
}

extern "C" void fr_graph_destroy(fr_graph *gr)
{
    if (!gr) return;
    hipStreamSynchronize(gr->ctx->stream);
    if (gr->exec) hipGraphExecDestroy(gr->exec);
    if (gr->graph) hipGraphDestroy(gr->graph);
    delete gr;
}

extern "C" fr_status fr_autotune(fr_ctx *ctx, const fr_samples *s, unsigned N, float kWidth, const float *lut, int L,
                                 const fr_tune *candidates, int ncand, int reps, fr_tune *best, float *best_ms,
                                 float *cand_ms)
{
    if (!ctx || !s || !lut || !best) return FR_ERR_INVALID;
    FR_CHECK(ctx, reps >= 1 && reps <= 1000, FR_ERR_INVALID, "reps must be in [1, 1000]");
    std::vector<fr_tune> cand;
    if (candidates && ncand > 0) cand.assign(candidates, candidates + ncand);
    else cand = fr_tune_structural_candidates(s->dim);
    fr_grids *g = nullptr;
    fr_status st = s->frames > 0 ? fr_grids_create_frames(ctx, N, s->C, s->frames, &g)
                                 : fr_grids_create(ctx, s->dim, N, s->C, &g);
    if (st) return st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        fr_grids_destroy(g);
        return fr_fail(ctx, FR_ERR_HIP, "hipEventCreate failed");
    }
    float best_t = -1.f;
    for (size_t i = 0; i < cand.size() && !st; ++i) {
        fr_plan *p = nullptr;
        st = fr_plan_create(ctx, s, N, kWidth, lut, L, &cand[i], &p);
        if (st) {
            // a candidate the plan refuses (too many cells, ...) is skipped
            if (cand_ms) cand_ms[i] = -1.f;
            st = FR_OK;
            continue;
        }
        st = fr_grid_execute(p, const_cast<fr_samples *>(s), g);           // warm-up (lazy allocations)
        float ms = 0.f;
        if (!st) {
            hipEventRecord(e0, ctx->stream);
            for (int r = 0; r < reps && !st; ++r) st = fr_grid_execute(p, const_cast<fr_samples *>(s), g);
            hipEventRecord(e1, ctx->stream);
            hipEventSynchronize(e1);
            hipEventElapsedTime(&ms, e0, e1);
            ms /= (float)reps;
        }
        fr_plan_destroy(p);
        if (st) break;
        if (cand_ms) cand_ms[i] = ms;
        if (best_t < 0.f || ms < best_t) {
            best_t = ms;
            *best = cand[i];
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    fr_grids_destroy(g);
    if (st) return st;
    FR_CHECK(ctx, best_t >= 0.f, FR_ERR_INVALID, "no candidate produced a plan");
    if (best_ms) *best_ms = best_t;
    return FR_OK;
}
