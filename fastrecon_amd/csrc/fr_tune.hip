// Autotuning search and its persistent cache (net-new: the reference's
// gridding has no tunable plan).
//
// fr_autotune_search times, on the caller's samples:
//   A. the default plan and the structural candidates (unit box and
//      accumulation mode, fr_tune_structural_candidates), one plan each;
//   B. for the fastest so far, the heavy-unit slice size (split_taps) and,
//      for position-weight plans, the write-out grouping (tiles of four,
//      single units; pairs are the default): one plan each;
//   C. for a position-weight winner, the launch-only knobs on that one plan
//      (items per wave, XCD run length, waves per SIMD, store kind), one knob
//      at a time in that order, each kept at its fastest value;
//   D. the default and the winner again, in turn, three rounds each: the
//      winner is kept only if its best time beats the default's by 1%.
// The winner is returned as an explicit fr_tune (a default winner is spelled
// out from its plan), so a plan built from it reproduces what was timed.
//
// The cache is a text file, one line per key:
//   <key> | <12 fr_tune fields> | best_ms=<ms> default_ms=<ms>
// keyed by device (architecture, CU count) and workload shape (dim, N, M,
// C, sample type, frames, kernel width, LUT length); the trajectory itself
// is not part of the key, so workloads of one shape share their tune.  A hit
// returns the stored tune without building or timing anything.
#include "fr_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

std::vector<fr_tune> fr_tune_structural_candidates(int dim)
{
    // {tile_x, tile_y, tile_z, accumulate}: position-weight boxes of 16
    // cells in the four aspect ratios, matrix-core tap boxes of 16 cells,
    // and 32-cell boxes with LDS read-add-write
    static const int s3[][4] = {{4, 2, 2, 6}, {4, 4, 1, 6}, {2, 2, 4, 6}, {8, 2, 1, 6},
                                {2, 4, 2, 6}, {4, 2, 2, 5}, {8, 2, 1, 5}, {4, 4, 2, 2}};
    static const int s2[][4] = {{4, 4, 1, 6}, {8, 2, 1, 6}, {16, 1, 1, 6}, {2, 8, 1, 6},
                                {8, 2, 1, 5}, {16, 2, 1, 5}, {8, 4, 1, 2}, {4, 4, 1, 2}};
    const int(*s)[4] = dim == 3 ? s3 : s2;
    std::vector<fr_tune> c;
    for (int i = 0; i < 8; ++i) {
        fr_tune t{};
        t.tile_x = s[i][0];
        t.tile_y = s[i][1];
        t.tile_z = s[i][2];
        t.accumulate = s[i][3];
        c.push_back(t);
    }
    return c;
}

namespace {

constexpr int NFIELDS = sizeof(fr_tune) / sizeof(int);
static_assert(sizeof(fr_tune) == NFIELDS * sizeof(int), "fr_tune is a plain array of ints");

std::string key_string(const fr_tune_key &k)
{
    char buf[256];
    snprintf(buf, sizeof buf, "arch=%.31s cus=%d dim=%d N=%u M=%lld C=%d type=%d frames=%d kw=%.6g L=%d",
             k.arch[0] ? k.arch : "?", k.cus, k.dim, k.N, (long long)k.M, k.C, k.type, k.frames, (double)k.kwidth,
             k.L);
    return buf;
}

std::string tune_string(const fr_tune &t)
{
    const int *f = reinterpret_cast<const int *>(&t);
    std::string out;
    for (int i = 0; i < NFIELDS; ++i) out += (i ? " " : "") + std::to_string(f[i]);
    return out;
}

bool parse_line(const std::string &line, std::string &key, fr_tune &t, float &best_ms, float &default_ms)
{
    const size_t a = line.find(" | ");
    if (a == std::string::npos) return false;
    const size_t b = line.find(" | ", a + 3);
    if (b == std::string::npos) return false;
    key = line.substr(0, a);
    std::istringstream fs(line.substr(a + 3, b - a - 3));
    int *f = reinterpret_cast<int *>(&t);
    for (int i = 0; i < NFIELDS; ++i)
        if (!(fs >> f[i])) return false;
    double bm = 0.0, dm = 0.0;
    if (sscanf(line.c_str() + b + 3, "best_ms=%lf default_ms=%lf", &bm, &dm) != 2) return false;
    best_ms = (float)bm;
    default_ms = (float)dm;
    return true;
}

// the plan's effective settings as an explicit tune
fr_tune explicit_tune(const fr_plan *p, const fr_tune *asked)
{
    fr_tune t = asked ? *asked : fr_tune{};
    t.tile_x = p->geom.bx;
    t.tile_y = p->geom.by;
    t.tile_z = p->geom.bz;
    t.accumulate = p->accumulate;
    if (p->accumulate == FR_ACC_POS) {
        t.items_per_wave = p->pos_ipw;
        t.xcd_run = p->xcd_run;
        t.stores = p->pos_stores;
        t.waves_per_simd = p->pos_waves;
    }
    return t;
}

struct timer {
    fr_ctx *ctx;
    const fr_samples *s;
    fr_grids *g;
    int reps;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int timed = 0;

    fr_status time(fr_plan *p, float *ms)
    {
        fr_status st = fr_grid_execute(p, const_cast<fr_samples *>(s), g);   // warm-up (lazy allocations)
        if (st) return st;
        FR_HIP(ctx, hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps && !st; ++r) st = fr_grid_execute(p, const_cast<fr_samples *>(s), g);
        FR_HIP(ctx, hipEventRecord(e1, ctx->stream));
        FR_HIP(ctx, hipEventSynchronize(e1));
        FR_HIP(ctx, hipEventElapsedTime(ms, e0, e1));
        *ms /= (float)reps;
        ++timed;
        return st;
    }

    // builds the plan of `t` (nullptr: default), times it; a plan the
    // builder refuses gives *ms = -1 and no error
    fr_status time_tune(const fr_tune *t, const float *lut, int L, unsigned N, float kWidth, float *ms,
                        fr_tune *effective, fr_plan **keep)
    {
        fr_plan *p = nullptr;
        *ms = -1.f;
        if (fr_plan_create(ctx, s, N, kWidth, lut, L, t, &p) != FR_OK) return FR_OK;
        fr_status st = time(p, ms);
        if (!st && effective) *effective = explicit_tune(p, t);
        if (!st && keep)
            *keep = p;
        else
            fr_plan_destroy(p);
        return st;
    }
};

}  // namespace

extern "C" fr_status fr_tune_key_of(fr_ctx *ctx, const fr_samples *s, unsigned N, float kWidth, int L,
                                    fr_tune_key *key)
{
    if (!ctx || !s || !key) return FR_ERR_INVALID;
    *key = fr_tune_key{};
    snprintf(key->arch, sizeof key->arch, "%s", ctx->arch.c_str());
    key->cus = ctx->cu_count;
    key->dim = s->dim;
    key->N = N;
    key->M = s->M;
    key->C = s->C;
    key->type = s->type;
    key->frames = s->frames;
    key->kwidth = kWidth;
    key->L = L;
    return FR_OK;
}

extern "C" fr_status fr_tune_cache_lookup(const char *path, const fr_tune_key *key, fr_tune *tune, float *best_ms,
                                          float *default_ms, int *found)
{
    if (!path || !key || !tune || !found) return FR_ERR_INVALID;
    *found = 0;
    std::ifstream in(path);
    if (!in) return FR_OK;                 // no cache yet: a miss
    const std::string want = key_string(*key);
    std::string line;
    while (std::getline(in, line)) {
        std::string k;
        fr_tune t{};
        float bm = 0.f, dm = 0.f;
        if (line.empty() || line[0] == '#' || !parse_line(line, k, t, bm, dm) || k != want) continue;
        *tune = t;
        if (best_ms) *best_ms = bm;
        if (default_ms) *default_ms = dm;
        *found = 1;                        // the last line of a key wins
    }
    return FR_OK;
}

extern "C" fr_status fr_tune_cache_store(const char *path, const fr_tune_key *key, const fr_tune *tune,
                                         float best_ms, float default_ms)
{
    if (!path || !key || !tune) return FR_ERR_INVALID;
    const std::string want = key_string(*key);
    std::vector<std::string> keep;
    {
        std::ifstream in(path);
        std::string line;
        while (in && std::getline(in, line)) {
            std::string k;
            fr_tune t{};
            float bm = 0.f, dm = 0.f;
            if (parse_line(line, k, t, bm, dm) && k == want) continue;   // replaced below
            if (!line.empty()) keep.push_back(line);
        }
    }
    if (keep.empty()) keep.push_back("# fastrecon_amd autotune cache: <key> | <fr_tune fields> | <times>");
    char times[96];
    snprintf(times, sizeof times, "best_ms=%.9g default_ms=%.9g", (double)best_ms, (double)default_ms);
    keep.push_back(want + " | " + tune_string(*tune) + " | " + times);
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream out(tmp, std::ios::trunc);
        if (!out) return FR_ERR_INVALID;
        for (const std::string &l : keep) out << l << "\n";
        if (!out) return FR_ERR_INVALID;
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return FR_ERR_INVALID;
    }
    return FR_OK;
}

extern "C" fr_status fr_autotune_search(fr_ctx *ctx, const fr_samples *s, unsigned N, float kWidth, const float *lut,
                                        int L, int reps, const char *cache_path, fr_tune *best, float *best_ms,
                                        float *default_ms, int *timed)
{
    if (!ctx || !s || !lut || !best) return FR_ERR_INVALID;
    FR_CHECK(ctx, reps >= 1 && reps <= 1000, FR_ERR_INVALID, "reps must be in [1, 1000]");
    fr_tune_key key;
    fr_status st = fr_tune_key_of(ctx, s, N, kWidth, L, &key);
    if (st) return st;
    if (timed) *timed = 0;
    if (cache_path) {
        int found = 0;
        float bm = 0.f, dm = 0.f;
        fr_tune t{};
        FR_CHECK(ctx, fr_tune_cache_lookup(cache_path, &key, &t, &bm, &dm, &found) == FR_OK, FR_ERR_INVALID,
                 "cannot read the tune cache %s", cache_path);
        if (found) {
            *best = t;
            if (best_ms) *best_ms = bm;
            if (default_ms) *default_ms = dm;
            return FR_OK;
        }
    }
    timer T{ctx, s, nullptr, reps};
    st = s->frames > 0 ? fr_grids_create_frames(ctx, N, s->C, s->frames, &T.g)
                       : fr_grids_create(ctx, s->dim, N, s->C, &T.g);
    if (st) return st;
    if (hipEventCreate(&T.e0) != hipSuccess || hipEventCreate(&T.e1) != hipSuccess) {
        if (T.e0) hipEventDestroy(T.e0);
        fr_grids_destroy(T.g);
        return fr_fail(ctx, FR_ERR_HIP, "hipEventCreate failed");
    }
    fr_tune win{};
    float win_ms = -1.f, def_ms = -1.f;
    auto consider = [&](const fr_tune *t) -> fr_status {
        float ms;
        fr_tune eff{};
        fr_status e = T.time_tune(t, lut, L, N, kWidth, &ms, &eff, nullptr);
        if (!e && ms >= 0.f && (win_ms < 0.f || ms < win_ms)) {
            win_ms = ms;
            win = eff;
        }
        if (!t) def_ms = ms;
        return e;
    };
    // A: default (twice: the first pass also brings the clocks up) and the
    // structural candidates
    st = consider(nullptr);
    if (!st) {
        win_ms = -1.f;
        st = consider(nullptr);
    }
    const fr_tune def_tune = win;
    for (const fr_tune &c : fr_tune_structural_candidates(s->dim))
        if (!st) st = consider(&c);
    // B: slice size and (position-weight) write-out grouping around the winner
    if (!st && win_ms >= 0.f) {
        const fr_tune base = win;
        for (int split : {256, 1024, 4096, 16384, 65536}) {
            fr_tune t = base;
            t.split_taps = split;
            if (!st) st = consider(&t);
        }
        if (base.accumulate == FR_ACC_POS)
            for (int tw : {1, 2}) {
                fr_tune t = win;
                t.tile_writes = tw;
                if (!st) st = consider(&t);
            }
    }
    // C: launch-only knobs of a position-weight winner, on its one plan
    if (!st && win_ms >= 0.f && win.accumulate == FR_ACC_POS) {
        fr_plan *p = nullptr;
        float ms = -1.f;
        st = T.time_tune(&win, lut, L, N, kWidth, &ms, nullptr, &p);
        if (!st && p) {
            win_ms = ms;     // the knobs below are compared on this plan instance
            struct knob {
                int *plan_field;
                int fr_tune::*tune_field;
                std::vector<int> values;
            };
            const knob knobs[] = {{&p->pos_ipw, &fr_tune::items_per_wave, {1, 2, 4}},
                                  {&p->xcd_run, &fr_tune::xcd_run, {64, 256, 1024, 4096}},
                                  {&p->pos_waves, &fr_tune::waves_per_simd, {3, 4}},
                                  {&p->pos_stores, &fr_tune::stores, {1, 2}}};
            for (const knob &k : knobs) {
                int keep = *k.plan_field;
                for (int v : k.values) {
                    if (st || v == win.*k.tune_field) continue;
                    *k.plan_field = v;
                    float m = -1.f;
                    st = T.time(p, &m);
                    if (!st && m < win_ms) {
                        win_ms = m;
                        win.*k.tune_field = v;
                        keep = v;
                    }
                }
                *k.plan_field = keep;
            }
            fr_plan_destroy(p);
        }
    }
    // confirmation: the default and the winner timed in turn, three rounds,
    // best of each; the winner must beat the default by 1% or the default
    // (spelled out) is returned
    if (!st && win_ms >= 0.f && def_ms >= 0.f && memcmp(&win, &def_tune, sizeof(fr_tune)) != 0) {
        fr_plan *pd = nullptr, *pw = nullptr;
        float md = -1.f, mw = -1.f, m = -1.f;
        st = T.time_tune(nullptr, lut, L, N, kWidth, &md, nullptr, &pd);
        if (!st) st = T.time_tune(&win, lut, L, N, kWidth, &mw, nullptr, &pw);
        for (int r = 0; r < 2 && !st && pd && pw; ++r) {
            st = T.time(pd, &m);
            if (!st) md = std::min(md, m);
            if (!st) st = T.time(pw, &m);
            if (!st) mw = std::min(mw, m);
        }
        if (pd) fr_plan_destroy(pd);
        if (pw) fr_plan_destroy(pw);
        if (!st && md >= 0.f && mw >= 0.f) {
            def_ms = md;
            if (mw < 0.99f * md) {
                win_ms = mw;
            } else {
                win = def_tune;
                win_ms = md;
            }
        }
    }
    hipEventDestroy(T.e0);
    hipEventDestroy(T.e1);
    fr_grids_destroy(T.g);
    if (st) return st;
    FR_CHECK(ctx, win_ms >= 0.f, FR_ERR_INVALID, "no candidate produced a plan");
    *best = win;
    if (best_ms) *best_ms = win_ms;
    if (default_ms) *default_ms = def_ms;
    if (timed) *timed = T.timed;
    if (cache_path)
        FR_CHECK(ctx, fr_tune_cache_store(cache_path, &key, &win, win_ms, def_ms) == FR_OK, FR_ERR_INVALID,
                 "cannot write the tune cache %s", cache_path);
    return FR_OK;
}
