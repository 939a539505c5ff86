// st_tune.cpp -- autotuner of the batched parcel step.
//
// Candidates: parcels per brick (which sets each cloud's brick dims), threads per
// workgroup, and the cap on persistent workgroups per CU.  Each candidate is timed the
// way a batch runs in production: its batch is built, one fluid step of kTuneSubsteps
// substeps is captured as a hipGraph, and after one warm-up replay `trials` (at least
// kMinTrials) replays are timed with HIP events on the context stream; the candidate's
// figure is the median replay divided by the passes per step.  The clouds and sources
// are restored from a snapshot afterwards, so tuning does not advance the caller's
// simulation.  The winner and the runner-up's figure are cached per device and
// workload signature in a JSON-lines file.
#include "st_handles.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace st::host;

namespace {

constexpr int kTuneSubsteps = 10;  // one fluid step of the reference's nSteps substeps
constexpr int kMinTrials = 5;

struct Snapshot {
    std::vector<void *> bufs;
    std::vector<size_t> sizes;
    std::vector<void *> srcs;
};

// FNV-1a over the workload description.
uint64_t fnv(const std::string &s) {
    uint64_t h = 1469598103934665603ull;
    for (unsigned char ch : s) {
        h ^= ch;
        h *= 1099511628211ull;
    }
    return h;
}

int snap_take(st_ctx *ctx, int n, st_cloud *const *clouds, st_field *const *fields, Snapshot &sn) {
    hipStream_t s = ctx->stream;
    auto keep = [&](void *src, size_t bytes) -> int {
        void *d = nullptr;
        ST_HIP(hipMalloc(&d, bytes > 0 ? bytes : 1));
        sn.bufs.push_back(d);
        sn.sizes.push_back(bytes);
        sn.srcs.push_back(src);
        if (bytes) ST_HIP(hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToDevice, s));
        return ST_OK;
    };
    for (int i = 0; i < n; ++i) {
        st_cloud *c = clouds[i];
        for (int f = 0; f < kNF; ++f)
            if (c->buf[c->cur][f]) ST_TRY(keep(c->buf[c->cur][f], (size_t)c->n * c->es));
        ST_TRY(keep(c->pid[c->cur], (size_t)c->n * 4));
        st_field *fl = fields[i];
        ST_TRY(merge_source(fl));
        ST_TRY(keep(fl->src, (size_t)fl->ncell * 3 * 8));
    }
    ST_HIP(hipStreamSynchronize(s));
    return ST_OK;
}

int snap_restore(st_ctx *ctx, int n, st_cloud *const *clouds, st_field *const *fields, const Snapshot &sn) {
    hipStream_t s = ctx->stream;
    for (size_t k = 0; k < sn.bufs.size(); ++k)
        if (sn.sizes[k]) ST_HIP(hipMemcpyAsync(sn.srcs[k], sn.bufs[k], sn.sizes[k], hipMemcpyDeviceToDevice, s));
    for (int i = 0; i < n; ++i) {
        // the trials' stray sums are dropped; src was merged before the snapshot
        ST_HIP(st::launch_fill_f64(fields[i]->stray, 0.0, (size_t)fields[i]->ncell * 3, s));
        fields[i]->stray_dirty = false;
        clouds[i]->sorted = false;  // the restored order is the pre-tuning one
    }
    ST_HIP(hipStreamSynchronize(s));
    return ST_OK;
}

void snap_free(Snapshot &sn) {
    for (void *b : sn.bufs) (void)hipFree(b);
    sn.bufs.clear();
}

bool cache_lookup(const char *path, const std::string &key, st_tune_result *r) {
    if (!path || !*path) return false;
    FILE *fp = std::fopen(path, "r");
    if (!fp) return false;
    char line[1024];
    bool found = false;
    while (std::fgets(line, sizeof line, fp)) {
        char k[128];
        st_tune_result t{};
        if (std::sscanf(line, "{\"key\": \"%127[^\"]\", \"target_parcels_per_brick\": %d, \"block\": %d, "
                              "\"blocks_per_cu\": %d, \"ms_per_pass\": %lf, \"runner_up_ms_per_pass\": %lf",
                        k, &t.target_parcels_per_brick, &t.block, &t.blocks_per_cu, &t.ms_per_pass,
                        &t.runner_up_ms_per_pass) == 6 &&
            key == k) {
            *r = t;  // the last matching line wins
            found = true;
        }
    }
    std::fclose(fp);
    return found;
}

void cache_store(const char *path, const std::string &key, const std::string &desc, const st_tune_result &r) {
    if (!path || !*path) return;
    FILE *fp = std::fopen(path, "a");
    if (!fp) return;
    std::fprintf(fp,
                 "{\"key\": \"%s\", \"target_parcels_per_brick\": %d, \"block\": %d, \"blocks_per_cu\": %d, "
                 "\"ms_per_pass\": %.6f, \"runner_up_ms_per_pass\": %.6f, \"workload\": \"%s\"}\n",
                 key.c_str(), r.target_parcels_per_brick, r.block, r.blocks_per_cu, r.ms_per_pass,
                 r.runner_up_ms_per_pass, desc.c_str());
    std::fclose(fp);
}

}  // namespace

extern "C" {

int st_autotune_batch(st_ctx *ctx, int32_t n, st_cloud *const *clouds, st_field *const *fields,
                      const st_params *base, double dt_fluid, int32_t trials, const char *cache_path,
                      st_tune_result *out) {
    if (!out) return fail(ST_E_INVALID, "null output");
    ST_TRY(check_ctx(ctx));
    if (n < 1 || !clouds || !fields) return fail(ST_E_INVALID, "need n >= 1 clouds and fields");
    if (trials < 1) return fail(ST_E_INVALID, "trials must be >= 1");
    if (!std::isfinite(dt_fluid) || !(dt_fluid > 0)) return fail(ST_E_INVALID, "dt_fluid must be finite and > 0");
    st_params p;
    if (base)
        p = *base;
    else
        st_params_default(&p);
    hipDeviceProp_t prop;
    ST_HIP(hipGetDeviceProperties(&prop, ctx->device));
    std::string desc = std::string(prop.gcnArchName) + " cus=" + std::to_string(prop.multiProcessorCount) +
                       " prec=" + std::to_string(clouds[0]->prec) + " bnd=" + std::to_string(p.boundary) +
                       " flush=" + std::to_string(p.flush) + " spp=" + std::to_string(p.substeps_per_pass);
    for (int i = 0; i < n; ++i) {
        if (!clouds[i] || !fields[i]) return fail(ST_E_INVALID, "null cloud or field %d", i);
        desc += " " + std::to_string(clouds[i]->n) + "@" + std::to_string(fields[i]->mesh.n[0]) + "x" +
                std::to_string(fields[i]->mesh.n[1]) + "x" + std::to_string(fields[i]->mesh.n[2]);
    }
    char keybuf[32];
    std::snprintf(keybuf, sizeof keybuf, "%016llx", (unsigned long long)fnv(desc));
    const std::string key = keybuf;
    if (cache_lookup(cache_path, key, out)) {
        out->from_cache = 1;
        return ST_OK;
    }
    Snapshot sn;
    int r = snap_take(ctx, n, clouds, fields, sn);
    if (r != ST_OK) {
        snap_free(sn);
        return r;
    }
    const int targets[] = {1024, 2048, 4096, 8192, 16384};
    const int blocks[] = {256, 512, 1024};
    const int caps[] = {0, 4, 3, 2};
    st_tune_result best{};
    best.ms_per_pass = INFINITY;
    best.runner_up_ms_per_pass = INFINITY;
    const int per = p.substeps_per_pass > 0 ? p.substeps_per_pass : kTuneSubsteps;
    const int passes = (kTuneSubsteps + per - 1) / per;
    const int ntrials = std::max(trials, kMinTrials);
    hipEvent_t ev[2] = {nullptr, nullptr};  // timing only: no system-scope fence (as the pass timer's)
    if (hipEventCreateWithFlags(&ev[0], hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&ev[1], hipEventDisableSystemFence) != hipSuccess) {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        snap_free(sn);
        return fail(ST_E_DEVICE, "autotuner events");
    }
    hipStream_t s = ctx->stream;
    for (int tgt : targets)
        for (int blk : blocks)
            for (int cap : caps) {
                st_params q = p;
                q.block = blk;
                st_batch *b = nullptr;
                r = st_batch_create(ctx, n, clouds, fields, &q, tgt, cap, &b);
                if (r != ST_OK) {  // a candidate that does not fit is skipped
                    st_batch_destroy(b);
                    r = snap_restore(ctx, n, clouds, fields, sn);
                    if (r != ST_OK) break;
                    continue;
                }
                st_graph *g = nullptr;
                r = st_batch_graph_capture(b, dt_fluid, kTuneSubsteps, 0, &g);
                std::vector<double> times;
                if (r == ST_OK) r = st_graph_launch(g);  // warm-up
                for (int t = 0; t < ntrials && r == ST_OK; ++t) {
                    float ms = 0.f;
                    hipError_t e = hipEventRecord(ev[0], s);
                    if (e == hipSuccess) r = st_graph_launch(g);
                    if (e == hipSuccess && r == ST_OK) e = hipEventRecord(ev[1], s);
                    if (e == hipSuccess && r == ST_OK) e = hipEventSynchronize(ev[1]);
                    if (e == hipSuccess && r == ST_OK) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
                    if (e != hipSuccess) r = fail(ST_E_DEVICE, "autotuner timing: %s", hipGetErrorString(e));
                    times.push_back((double)ms);
                }
                st_graph_destroy(g);
                st_batch_destroy(b);
                const int rr = snap_restore(ctx, n, clouds, fields, sn);
                if (r != ST_OK || rr != ST_OK) {
                    if (r == ST_OK) r = rr;
                    break;
                }
                std::sort(times.begin(), times.end());
                const size_t h = times.size() / 2;
                const double med = times.size() % 2 ? times[h] : 0.5 * (times[h - 1] + times[h]);
                const double ms = med / passes;
                if (ms < best.ms_per_pass) {
                    best.runner_up_ms_per_pass = best.ms_per_pass;
                    best.ms_per_pass = ms;
                    best.target_parcels_per_brick = tgt;
                    best.block = blk;
                    best.blocks_per_cu = cap;
                } else if (ms < best.runner_up_ms_per_pass) {
                    best.runner_up_ms_per_pass = ms;
                }
            }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    snap_free(sn);
    if (r != ST_OK) return r;
    if (!std::isfinite(best.ms_per_pass)) return fail(ST_E_DEVICE, "no autotuner candidate ran");
    best.from_cache = 0;
    cache_store(cache_path, key, desc, best);
    *out = best;
    return ST_OK;
}

}  // extern "C"
