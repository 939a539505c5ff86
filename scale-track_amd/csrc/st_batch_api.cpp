// st_batch_api.cpp -- many independent clouds advanced by one persistent launch per pass.
//
// Each cloud keeps its own mesh, field and brick sort.  Work items are (cloud, brick)
// pairs, handed out by an atomic counter to a grid sized to fill the GPU once; clouds
// are ordered densest first so the longest items start first.  Brick dimensions are
// chosen per cloud so a brick holds about `target_parcels_per_brick` parcels.
#include "st_handles.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

using namespace st::host;

struct st_batch {
    st_ctx *ctx = nullptr;
    int prec = ST_FP64;
    st_params p{};
    int32_t target = 0;
    int32_t blocks_per_cu = 0;
    std::vector<st_cloud *> clouds;  // in work order: no per-parcel density first, then
                                     // per-parcel density; densest first in each group
    std::vector<st_field *> fields;
    std::vector<st_params> cloud_params;
    std::vector<int> caller_index;   // slot -> position in the caller's arrays
    void *descs = nullptr;           // device PassArgs<T>[n]
    std::vector<unsigned char> host_descs;
    int32_t *item0 = nullptr;        // device [n + 1]
    st::StatsTarget *targets = nullptr;  // device [n]: statistics reduction of each cloud
    unsigned int *counter = nullptr; // device
    int32_t total_items = 0;
    int32_t nplain = 0;              // leading clouds without a per-parcel density
    std::vector<int32_t> host_item0; // [n + 1]
    int grid[2] = {0, 0};            // persistent grid of each group (0: group empty)
    size_t lds = 0;
    double dt_sub = -1.0;            // substep the device table holds
    int pins = 0;                    // live graphs of this batch
    st::FillItem *fills = nullptr;   // device: (owner flush) every field's stray, then every src
    uint32_t *fill_chunk0 = nullptr; // device [nfill + 1]
    int32_t nfill = 0;
    uint32_t fill_chunks = 0;
    int32_t nfill_stray = 0;         // the stray items, a prefix of the table
    uint32_t fill_chunks_stray = 0;
    bool pinned_members = false;     // the clouds and fields are pinned by this batch
};

namespace st {
namespace host {
void batch_pin(st_batch *b) { ++b->pins; }
void batch_unpin(st_batch *b) { --b->pins; }
}  // namespace host
}  // namespace st

namespace {

// Static LDS of the batch kernel (its statistics reduction and item slot), kept free
// next to the tile when bricks are sized for a number of workgroups per CU.
constexpr size_t kBatchStaticLds = st::kPassStaticLds;

// Brick of about `cells` cells: powers of two, x extent largest (rows stay contiguous),
// at most `max_cells` including the one-cell halo.
void auto_brick(int64_t cells, const int32_t n[3], int64_t max_cells, int32_t b[3]) {
    int64_t c = 1;
    while (c * 2 <= cells && c < 2048) c *= 2;
    while (c > 1) {  // largest halo tile of c cells: a cube-ish brick of side (c)^(1/3)
        const double side = std::cbrt((double)c) + 2.0;
        if (side * side * side * 1.2 <= (double)max_cells) break;
        c /= 2;
    }
    b[0] = b[1] = b[2] = 1;
    int axis = 0;
    while (c > 1) {
        if (b[axis] < n[axis]) {
            b[axis] *= 2;
            c /= 2;
        } else if (b[0] >= n[0] && b[1] >= n[1] && b[2] >= n[2]) {
            break;
        }
        axis = (axis + 1) % 3;
    }
}

template <class T>
int build_table(st_batch *b, double dt_sub) {
    const size_t n = b->clouds.size();
    b->host_descs.assign(n * sizeof(st::PassArgs<T>), 0);
    st::PassArgs<T> *hd = (st::PassArgs<T> *)b->host_descs.data();
    std::vector<int32_t> item0(n + 1, 0);
    std::vector<st::StatsTarget> targets(n);
    int32_t maxcells = 1;
    for (size_t i = 0; i < n; ++i) {
        st_cloud *c = b->clouds[i];
        st_field *f = b->fields[i];
        const st_params &p = b->cloud_params[i];
        st::PassArgs<T> a{};
        a.s = soa<T>(c, c->cur);
        a.U = (const T *)f->U;
        a.src = f->src;
        a.stray = f->stray;
        a.brick_start = c->brick_start;
        a.g = f->geo;
        a.ph = make_phys<T>(c, f, dt_sub, p.cd);
        ST_TRY(ensure_part(c, c->nbricks));
        a.part = c->part;
        a.nbricks = c->nbricks;
        targets[i] = st::StatsTarget{c->part, c->nbricks, c->bbox, c->resort + 1};
        for (int k = 0; k < 3; ++k) {
            a.nb[k] = c->nb[k];
            a.bd[k] = p.brick[k];
        }
        a.tile_cells = (p.brick[0] + 2) * (p.brick[1] + 2) * (p.brick[2] + 2);
        a.nsteps = 1;
        a.flush = p.flush;
        a.np = (uint32_t)c->n;
        hd[i] = a;
        item0[i + 1] = item0[i] + c->nbricks;
        maxcells = std::max(maxcells, a.tile_cells);
    }
    b->total_items = item0[n];
    b->host_item0 = item0;
    b->lds = (size_t)maxcells * 3 * (sizeof(double) + sizeof(T));
    hipStream_t s = b->ctx->stream;
    ST_HIP(hipMemcpyAsync(b->descs, hd, b->host_descs.size(), hipMemcpyHostToDevice, s));
    ST_HIP(hipMemcpyAsync(b->item0, item0.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
    ST_HIP(hipMemcpyAsync(b->targets, targets.data(), n * sizeof(st::StatsTarget), hipMemcpyHostToDevice, s));
    ST_HIP(hipStreamSynchronize(s));
    int dev = b->ctx->device;
    int cus = 0;
    ST_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    for (int g = 0; g < 2; ++g) {
        const int32_t c0 = g == 0 ? 0 : b->nplain, c1 = g == 0 ? b->nplain : (int32_t)n;
        const int32_t items = item0[c1] - item0[c0];
        b->grid[g] = 0;
        if (c1 == c0) continue;
        int per = st::batch_max_blocks_per_cu<T>(b->p.boundary == ST_PERIODIC, g, b->p.block, b->lds);
        if (per < 1) return fail(ST_E_RANGE, "batch kernel does not fit: block %d, %zu B LDS", b->p.block, b->lds);
        if (b->blocks_per_cu > 0) per = std::min(per, b->blocks_per_cu);
        b->grid[g] = std::max(1, std::min(cus * per, items));
    }
    b->dt_sub = dt_sub;
    return ST_OK;
}

int table_for(st_batch *b, double dt_sub) {
    if (b->dt_sub == dt_sub && b->descs) return ST_OK;
    return b->prec == ST_FP64 ? build_table<double>(b, dt_sub) : build_table<float>(b, dt_sub);
}

int sort_all(st_batch *b) {
    for (size_t i = 0; i < b->clouds.size(); ++i) ST_TRY(sort_any(b->clouds[i], b->fields[i], b->cloud_params[i]));
    b->dt_sub = -1.0;  // brick counts may have changed: rebuild the table
    return ST_OK;
}

// take_zero: every field's source counts as zero in the first pass, whose bricks then
// store all their own cells (owner flush only); otherwise `src` holds what it holds.
int queue_passes(st_batch *b, double dt_sub, int nsub, bool capturing, bool take_zero) {
    ST_TRY(table_for(b, dt_sub));
    hipStream_t s = b->ctx->stream;
    const int per = b->p.substeps_per_pass > 0 ? b->p.substeps_per_pass : nsub;
    if (!capturing) ST_HIP(timer_start(b->ctx, s));
    int launches = 0;
    const int32_t n = (int32_t)b->clouds.size();
    const int periodic = b->p.boundary == ST_PERIODIC;
    for (int done = 0; done < nsub; done += per) {
        const int ns = std::min(per, nsub - done);
        const int stats = done + ns >= nsub;
        // one persistent launch per group of clouds sharing the density form, so each
        // launch runs the kernel instance of that form
        for (int g = 0; g < 2; ++g) {
            if (!b->grid[g]) continue;
            const int32_t c0 = g == 0 ? 0 : b->nplain, nc = g == 0 ? b->nplain : n - b->nplain;
            ST_HIP(st::launch_fill_u32(b->counter, (unsigned)b->host_item0[c0], 1, s));
            if (b->prec == ST_FP64)
                ST_HIP(st::launch_pass_batch<double>((const st::PassArgs<double> *)b->descs + c0, b->item0 + c0, nc,
                                                     ns, stats, b->counter, periodic, g, b->p.block, b->grid[g],
                                                     b->lds, take_zero && done == 0, s));
            else
                ST_HIP(st::launch_pass_batch<float>((const st::PassArgs<float> *)b->descs + c0, b->item0 + c0, nc,
                                                    ns, stats, b->counter, periodic, g, b->p.block, b->grid[g],
                                                    b->lds, take_zero && done == 0, s));
        }
        ++launches;  // passes: the kernel time per pass spans both launches
    }
    if (!capturing) ST_HIP(timer_stop(b->ctx, s, launches));
    ST_HIP(st::launch_reduce_stats_many(b->targets, (int32_t)b->clouds.size(), s));
    if (b->p.flush == ST_FLUSH_OWNER)
        for (st_field *f : b->fields) f->stray_dirty = true;
    return ST_OK;
}

int destroy(st_batch *b) {
    if (!b) return ST_OK;
    if (b->pins > 0) return fail(ST_E_STATE, "batch is used by %d live graph(s); destroy them first", b->pins);
    quiesce_ctx(b->ctx);
    if (b->pinned_members) {
        for (st_cloud *c : b->clouds) unpin(c);
        for (st_field *f : b->fields) unpin(f);
    }
    if (b->descs) (void)hipFree(b->descs);
    if (b->item0) (void)hipFree(b->item0);
    if (b->targets) (void)hipFree(b->targets);
    if (b->counter) (void)hipFree(b->counter);
    if (b->fills) (void)hipFree(b->fills);
    if (b->fill_chunk0) (void)hipFree(b->fill_chunk0);
    st_ctx *ctx = b->ctx;
    delete b;
    release_ctx(ctx);
    return ST_OK;
}

}  // namespace

extern "C" {

int st_batch_create(st_ctx *ctx, int32_t n, st_cloud *const *clouds, st_field *const *fields,
                    const st_params *params, int32_t target_parcels_per_brick, int32_t blocks_per_cu,
                    st_batch **out) {
    if (!out) return fail(ST_E_INVALID, "null output handle");
    *out = nullptr;
    ST_TRY(check_ctx(ctx));
    if (n < 1 || !clouds || !fields) return fail(ST_E_INVALID, "need n >= 1 clouds and fields");
    if (target_parcels_per_brick < 0 || blocks_per_cu < 0)
        return fail(ST_E_INVALID, "target_parcels_per_brick and blocks_per_cu must be >= 0");
    st_params p;
    if (params)
        p = *params;
    else
        st_params_default(&p);
    p.method = ST_METHOD_SORTED;
    ST_TRY(validate_params(p));
    if (p.boundary == ST_CHANNEL) return fail(ST_E_UNSUPPORTED, "batches take ST_WALL or ST_PERIODIC clouds");
    std::vector<int> idx(n);
    std::vector<double> density(n);
    for (int i = 0; i < n; ++i) {
        st_params tmp;
        ST_TRY(prepare(ctx, clouds[i], fields[i], 1.0, 1, &p, tmp));
        if (clouds[i]->prec != clouds[0]->prec) return fail(ST_E_INVALID, "all clouds of a batch share one precision");
        for (int j = 0; j < i; ++j)
            if (clouds[j] == clouds[i] || fields[j] == fields[i])
                return fail(ST_E_INVALID, "a cloud or field appears twice in the batch");
        idx[i] = i;
        density[i] = (double)clouds[i]->n / (double)fields[i]->ncell;
    }
    // clouds without a per-parcel density first, then those with one; densest first in
    // each group (the longest work items are handed out first)
    auto has_rho = [&](int i) { return clouds[i]->has_rho != 0; };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) {
        if (has_rho(a) != has_rho(c)) return !has_rho(a);
        return density[a] > density[c];
    });
    st_batch *b = new (std::nothrow) st_batch;
    if (!b) return fail(ST_E_NOMEM, "host allocation failed");
    b->ctx = ctx;
    retain_ctx(ctx);
    b->prec = clouds[0]->prec;
    b->p = p;
    b->target = target_parcels_per_brick;
    b->blocks_per_cu = blocks_per_cu;
    for (int k = 0; k < n; ++k) {
        const int i = idx[k];
        if (!has_rho(i)) b->nplain = k + 1;
        b->clouds.push_back(clouds[i]);
        b->fields.push_back(fields[i]);
        b->caller_index.push_back(i);
        st_params pc = p;
        if (target_parcels_per_brick > 0) {
            const double want = target_parcels_per_brick / std::max(density[i], 1e-9);
            const size_t per_cell = 3 * (sizeof(double) + (clouds[i]->prec == ST_FP64 ? 8 : 4));
            // With a cap of k workgroups per CU, every cloud's tile fits a k-th of the
            // CU's LDS, so k workgroups can be resident whichever clouds they hold
            // (one LDS size serves the whole launch: the largest tile's).
            size_t budget = ctx->max_lds;
            if (blocks_per_cu > 0) {
                const size_t share = ctx->lds_per_cu / (size_t)blocks_per_cu;
                budget = std::min(budget, share > kBatchStaticLds ? share - kBatchStaticLds : 0);
            }
            auto_brick((int64_t)std::max(1.0, want), fields[i]->mesh.n, (int64_t)(budget / per_cell), pc.brick);
            auto tile = [&] {
                return clouds[i]->prec == ST_FP64 ? st::tile_lds_bytes<double>(pc.brick) : st::tile_lds_bytes<float>(pc.brick);
            };
            while ((tile() > budget || check_tile(ctx, clouds[i]->prec, pc.brick) != ST_OK) &&
                   pc.brick[0] * pc.brick[1] * pc.brick[2] > 1) {
                int ax = pc.brick[0] >= pc.brick[1] && pc.brick[0] >= pc.brick[2] ? 0 : (pc.brick[1] >= pc.brick[2] ? 1 : 2);
                pc.brick[ax] = std::max(1, pc.brick[ax] / 2);
            }
        }
        int r = check_tile(ctx, clouds[i]->prec, pc.brick);
        if (r != ST_OK) {
            delete b;
            release_ctx(ctx);
            return r;
        }
        b->cloud_params.push_back(pc);
    }
    hipError_t e = hipMalloc(&b->descs, (size_t)n * std::max(sizeof(st::PassArgs<double>), sizeof(st::PassArgs<float>)));
    if (e == hipSuccess) e = hipMalloc((void **)&b->item0, (size_t)(n + 1) * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&b->targets, (size_t)n * sizeof(st::StatsTarget));
    if (e == hipSuccess) e = hipMalloc((void **)&b->counter, sizeof(unsigned int));
    // the table that zeroes every field's stray sums and source in one launch; its
    // prefix of stray items alone serves a step whose first pass takes the source's zero
    std::vector<st::FillItem> fills;
    if (p.flush == ST_FLUSH_OWNER)
        for (st_field *f : b->fields) fills.push_back({f->stray, (size_t)f->ncell * 3});
    b->nfill_stray = (int32_t)fills.size();
    for (st_field *f : b->fields) fills.push_back({f->src, (size_t)f->ncell * 3});
    std::vector<uint32_t> chunk0(fills.size() + 1, 0);
    for (size_t i = 0; i < fills.size(); ++i) chunk0[i + 1] = chunk0[i] + st::fill_many_chunks(fills[i].n);
    b->nfill = (int32_t)fills.size();
    b->fill_chunks = chunk0.back();
    b->fill_chunks_stray = chunk0[(size_t)b->nfill_stray];
    if (e == hipSuccess) e = hipMalloc((void **)&b->fills, fills.size() * sizeof(st::FillItem));
    if (e == hipSuccess) e = hipMalloc((void **)&b->fill_chunk0, chunk0.size() * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->fills, fills.data(), fills.size() * sizeof(st::FillItem), hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->fill_chunk0, chunk0.data(), chunk0.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        destroy(b);
        return fail(ST_E_NOMEM, "batch allocation: %s", hipGetErrorString(e));
    }
    int r = sort_all(b);
    if (r == ST_OK) r = table_for(b, 0.0);  // sizes the grid; the first evolve sets dt
    if (r != ST_OK) {
        destroy(b);
        return r;
    }
    for (st_cloud *c : b->clouds) pin(c);
    for (st_field *f : b->fields) pin(f);
    b->pinned_members = true;
    *out = b;
    return ST_OK;
}

int st_batch_destroy(st_batch *b) { return destroy(b); }

int st_batch_sort(st_batch *b) {
    if (!b) return fail(ST_E_INVALID, "null batch");
    ST_TRY(check_ctx(b->ctx));
    return sort_all(b);
}

int st_batch_evolve(st_batch *b, double dt_fluid, int32_t nsub) {
    if (!b) return fail(ST_E_INVALID, "null batch");
    ST_TRY(check_ctx(b->ctx));
    if (!std::isfinite(dt_fluid) || !(dt_fluid > 0)) return fail(ST_E_INVALID, "dt_fluid must be finite and > 0");
    if (nsub < 1) return fail(ST_E_INVALID, "nsub must be >= 1");
    // zeros pending on every field (st_field_zero_source) are taken by the first pass;
    // otherwise the kernel reads `src`, so pending ones are written first
    bool take = b->p.flush == ST_FLUSH_OWNER;
    for (st_field *f : b->fields) take = take && f->src_zero;
    if (!take)
        for (st_field *f : b->fields) ST_TRY(materialize_zero(f));
    ST_TRY(queue_passes(b, substep_dt(b->prec, dt_fluid, nsub), nsub, false, take));
    if (take)
        for (st_field *f : b->fields) f->src_zero = false;
    return ST_OK;
}

int st_batch_graph_capture(st_batch *b, double dt_fluid, int32_t nsub, int32_t zero_source, st_graph **out) {
    if (!out) return fail(ST_E_INVALID, "null output handle");
    *out = nullptr;
    if (!b) return fail(ST_E_INVALID, "null batch");
    ST_TRY(check_ctx(b->ctx));
    if (!std::isfinite(dt_fluid) || !(dt_fluid > 0)) return fail(ST_E_INVALID, "dt_fluid must be finite and > 0");
    if (nsub < 1) return fail(ST_E_INVALID, "nsub must be >= 1");
    const double dt_sub = substep_dt(b->prec, dt_fluid, nsub);
    ST_TRY(table_for(b, dt_sub));  // table upload happens before the capture
    for (st_field *f : b->fields) ST_TRY(materialize_zero(f));  // a zero pending now belongs to no replay
    st_graph *g = new (std::nothrow) st_graph;
    if (!g) return fail(ST_E_NOMEM, "host allocation failed");
    g->ctx = b->ctx;
    g->more_fields = b->fields;
    g->more_clouds = b->clouds;
    g->owner_flush = b->p.flush == ST_FLUSH_OWNER;
    hipStream_t s = b->ctx->stream;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete g;
        return fail(ST_E_DEVICE, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    int r = ST_OK;
    // with the owner flush the first captured pass takes the sources' zero and only the
    // stray sums are filled; otherwise every source is filled too, in one launch
    const bool take = zero_source && b->p.flush == ST_FLUSH_OWNER;
    if (zero_source && (take ? b->nfill_stray : b->nfill) > 0) {
        e = take ? st::launch_fill_many(b->fills, b->fill_chunk0, b->nfill_stray, b->fill_chunks_stray, s)
                 : st::launch_fill_many(b->fills, b->fill_chunk0, b->nfill, b->fill_chunks, s);
        if (e != hipSuccess) r = fail(ST_E_DEVICE, "capture fill: %s", hipGetErrorString(e));
    }
    if (r == ST_OK) r = queue_passes(b, dt_sub, nsub, true, take);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s, &graph);
    if (r != ST_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        delete g;
        return r != ST_OK ? r : fail(ST_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        delete g;
        return fail(ST_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    g->graph = graph;
    retain_ctx(b->ctx);
    g->batch = b;
    batch_pin(b);
    for (st_cloud *c : b->clouds) pin(c);
    for (st_field *f : b->fields) pin(f);
    *out = g;
    return ST_OK;
}

int st_batch_info(const st_batch *b, int32_t *total_items, int32_t *grid, int32_t *brick_of_cloud) {
    if (!b) return fail(ST_E_INVALID, "null batch");
    if (total_items) *total_items = b->total_items;
    if (grid) *grid = b->grid[0] + b->grid[1];
    if (brick_of_cloud)  // in the caller's cloud order
        for (size_t i = 0; i < b->clouds.size(); ++i)
            for (int k = 0; k < 3; ++k) brick_of_cloud[3 * b->caller_index[i] + k] = b->cloud_params[i].brick[k];
    return ST_OK;
}

}  // extern "C"
