// st_api.cpp -- C-ABI of the gfx950 parcel step (declared in include/scaletrack.h).
//
// Replaces the device-side calls of the reference's tracker: the per-chunk launch of
// evolve_on_device! (run/icoJuliaParcelFoam/cavity3D/asyncParallelTracking.jl:1098-1125),
// the UTrans fill! (:1064-1067), the U/UTrans copies (:1076-1083, :1148-1152) and the
// bounding-box readback (:1106, :1113).  No status escapes as an exception.
#include "st_handles.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <new>
#include <thread>

namespace st {
namespace host {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char *last_error_text() { return g_err.c_str(); }

int check_ctx(st_ctx *ctx) {
    if (!ctx) return fail(ST_E_INVALID, "null context");
    if (ctx->closed) return fail(ST_E_STATE, "context was destroyed");
    ST_HIP(hipSetDevice(ctx->device));
    return ST_OK;
}

void retain_ctx(st_ctx *ctx) {
    if (ctx) ++ctx->refs;
}

hipError_t timer_start(st_ctx *ctx, hipStream_t s) {
    return hipEventRecord(ctx->tev[ctx->tcount % kTimerRing][0], s);
}

hipError_t timer_stop(st_ctx *ctx, hipStream_t s, int launches) {
    const int slot = (int)(ctx->tcount % kTimerRing);
    const hipError_t e = hipEventRecord(ctx->tev[slot][1], s);
    if (e == hipSuccess) {
        ctx->tlaunches[slot] = launches;
        ++ctx->tcount;
    }
    return e;
}

static void free_ctx(st_ctx *ctx) {
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int k = 0; k < kTimerRing; ++k)
        for (int j = 0; j < 2; ++j)
            if (ctx->tev[k][j]) (void)hipEventDestroy(ctx->tev[k][j]);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void release_ctx(st_ctx *ctx) {
    if (!ctx) return;
    if (--ctx->refs <= 0 && ctx->closed) free_ctx(ctx);
}

void quiesce_ctx(st_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
}

// Runs body(lo, hi) over [0, n), split over up to kHostThreads host threads when n is
// large: the argument checks and staging copies of big uploads and injections run on
// the caller's thread, often while the device waits for them.  Each call of body
// touches only its own range.  If a thread cannot be started its range runs here.
constexpr int64_t kHostThreads = 8;
constexpr int64_t kHostChunk = int64_t(1) << 16;

template <class F>
void host_parallel(int64_t n, const F &body) {
    const int64_t hw = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
    const int64_t nt = std::min<int64_t>(std::min<int64_t>(kHostThreads, hw), n / kHostChunk);
    if (nt < 2) {
        body((int64_t)0, n);
        return;
    }
    const int64_t step = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    th.reserve((size_t)nt);
    for (int64_t t = 1; t < nt; ++t) {
        const int64_t lo = t * step, hi = std::min(n, lo + step);
        if (lo >= hi) continue;
        try {
            th.emplace_back([&body, lo, hi] { body(lo, hi); });
        } catch (...) {
            body(lo, hi);
        }
    }
    body((int64_t)0, std::min(n, step));
    for (std::thread &x : th) x.join();
}

bool finite_all(const void *p, int64_t n, size_t es) {
    std::atomic<bool> ok{true};
    host_parallel(n, [&](int64_t lo, int64_t hi) {
        bool good = true;  // no early exit: the loops vectorise
        if (es == 8) {
            const double *a = (const double *)p;
            for (int64_t i = lo; i < hi; ++i) good &= std::isfinite(a[i]);
        } else {
            const float *a = (const float *)p;
            for (int64_t i = lo; i < hi; ++i) good &= std::isfinite(a[i]);
        }
        if (!good) ok.store(false, std::memory_order_relaxed);
    });
    return ok.load();
}

bool positive_all(const void *p, int64_t n, size_t es) {
    std::atomic<bool> ok{true};
    host_parallel(n, [&](int64_t lo, int64_t hi) {
        bool good = true;
        if (es == 8) {
            const double *a = (const double *)p;
            for (int64_t i = lo; i < hi; ++i) good &= a[i] > 0;
        } else {
            const float *a = (const float *)p;
            for (int64_t i = lo; i < hi; ++i) good &= a[i] > 0;
        }
        if (!good) ok.store(false, std::memory_order_relaxed);
    });
    return ok.load();
}

Geo make_geo(const st_mesh &m) {
    Geo g{};
    for (int c = 0; c < 3; ++c) {
        g.n[c] = m.n[c];
        g.o[c] = m.origin[c];
        const double L = m.ending[c] - m.origin[c];
        const double delta = L / (double)m.n[c];
        g.rd[c] = 1.0 / delta;
    }
    g.nxny = (uint32_t)m.n[0] * (uint32_t)m.n[1];
    return g;
}

int alloc_set(st_cloud *c, int which) {
    const size_t bytes = (size_t)(c->cap > 0 ? c->cap : 1) * c->es;
    for (int f = 0; f < kNF; ++f) {
        if (f == 7 && !c->has_rho) continue;
        if (!c->buf[which][f]) ST_HIP(hipMalloc(&c->buf[which][f], bytes));
    }
    if (!c->pid[which]) ST_HIP(hipMalloc((void **)&c->pid[which], (size_t)(c->cap > 0 ? c->cap : 1) * 4));
    return ST_OK;
}

int validate_params(const st_params &p) {
    if (p.boundary != ST_WALL && p.boundary != ST_PERIODIC && p.boundary != ST_CHANNEL)
        return fail(ST_E_INVALID, "boundary must be ST_WALL, ST_PERIODIC or ST_CHANNEL (got %d)", p.boundary);
    if (p.method != ST_METHOD_PLAIN && p.method != ST_METHOD_SORTED)
        return fail(ST_E_INVALID, "method must be ST_METHOD_PLAIN or ST_METHOD_SORTED (got %d)", p.method);
    if (p.flush != ST_FLUSH_OWNER && p.flush != ST_FLUSH_ATOMIC)
        return fail(ST_E_INVALID, "flush must be ST_FLUSH_OWNER or ST_FLUSH_ATOMIC (got %d)", p.flush);
    if (p.substeps_per_pass < 0) return fail(ST_E_INVALID, "substeps_per_pass must be >= 0");
    if (!std::isfinite(p.cd) || p.cd < 0) return fail(ST_E_INVALID, "cd must be finite and >= 0");
    for (int c = 0; c < 3; ++c)
        if (p.brick[c] < 1 || p.brick[c] > 64) return fail(ST_E_INVALID, "brick dims must be in [1, 64]");
    const int64_t tc = (int64_t)p.brick[0] * p.brick[1] * p.brick[2];
    if (tc > 2048) return fail(ST_E_INVALID, "brick holds %lld cells; at most 2048", (long long)tc);
    if (p.block != 64 && p.block != 128 && p.block != 256 && p.block != 512 && p.block != 1024)
        return fail(ST_E_INVALID, "block must be 64, 128, 256, 512 or 1024 (got %d)", p.block);
    if (!(p.resort_fraction >= 0)) return fail(ST_E_INVALID, "resort_fraction must be >= 0");
    if (p.passes_per_launch < 1 || p.passes_per_launch > 1024)
        return fail(ST_E_INVALID, "passes_per_launch must be in [1, 1024] (got %d)", p.passes_per_launch);
    return ST_OK;
}

// Brick counts up to this sort through per-workgroup LDS histograms (32 KB of LDS);
// larger ones through one global counter per brick.
constexpr int64_t kSortLdsBins = 8192;

void pin(st_cloud *c) { ++c->pins; }
void unpin(st_cloud *c) { --c->pins; }
void pin(st_field *f) { ++f->pins; }
void unpin(st_field *f) { --f->pins; }

// Histogram rows of the chunked sort (one per workgroup of at least 4096 parcels, at
// most 512), or 0 when the brick count takes the global-counter path.
static int32_t sort_rows(int64_t n, int64_t nbr) {
    if (nbr > kSortLdsBins) return 0;
    return (int32_t)std::max<int64_t>(1, std::min<int64_t>(512, (n + 4095) / 4096));
}

// Device buffer `*p` of at least `need` elements of `es` bytes.  Growing one that exists
// frees it, which a pinned cloud refuses (captured graphs hold its address).
template <class U>
static int grow(st_cloud *c, U **p, size_t &cap, size_t need, const char *what) {
    if (*p && need <= cap) return ST_OK;
    if (*p) {
        if (c->pins > 0)
            return fail(ST_E_STATE,
                        "the cloud's %s would be reallocated, but %d graph(s), batch(es) or adjoint(s) "
                        "hold its device buffers; destroy them first",
                        what, c->pins);
        ST_HIP(hipFree(*p));
        *p = nullptr;
        cap = 0;
    }
    ST_HIP(hipMalloc((void **)p, need * sizeof(U)));
    cap = need;
    return ST_OK;
}

// Queues the sort of the cloud by brick (nb bricks per direction, nbr in all, rows as
// sort_rows gives) into the second buffer set and its copy back over the first, all
// gated on gate[0] when gate is set (a gated sort that runs counts itself in count).
// Every buffer must exist.
template <class T>
static int enqueue_sort(st_cloud *c, st_field *f, const st_params &p, const int32_t nb[3], int64_t nbr,
                        int32_t rows, const unsigned int *gate, unsigned int *count) {
    hipStream_t s = c->ctx->stream;
    const uint32_t np = (uint32_t)c->n;
    if (rows > 0) {
        ST_HIP(launch_sort_chunked<T>(soa<T>(c, c->cur), c->pid[c->cur], soa<T>(c, 1 - c->cur), c->pid[1 - c->cur],
                                      np, f->geo, nb, p.brick, p.boundary == ST_PERIODIC, rows, c->keys, c->hist,
                                      c->counts, c->brick_start, c->cursor, gate, s));
    } else {
        ST_HIP(launch_fill_u32(c->counts, 0u, (size_t)nbr, s, gate));
        ST_HIP(launch_brick_keys<T>(soa<T>(c, c->cur), np, f->geo, nb, p.brick, p.boundary == ST_PERIODIC,
                                    c->keys, c->counts, gate, s));
        ST_HIP(launch_scan(c->counts, (int32_t)nbr, c->brick_start, c->cursor, gate, s));
        ST_HIP(launch_scatter<T>(soa<T>(c, c->cur), c->pid[c->cur], soa<T>(c, 1 - c->cur), c->pid[1 - c->cur],
                                 c->keys, c->cursor, np, gate, s));
    }
    // The sorted order is copied back so the cloud's device pointers never change:
    // graphs and batch tables captured earlier stay valid across re-sorts.
    ST_HIP(launch_copy_back<T>(soa<T>(c, 1 - c->cur), c->pid[1 - c->cur], soa<T>(c, c->cur), c->pid[c->cur], np,
                               gate, count, s));
    return ST_OK;
}

static bool same_geometry(const st_cloud *c, const st_field *f, const st_params &p) {
    if (c->sort_periodic != p.boundary) return false;
    for (int a = 0; a < 3; ++a)
        if (c->sort_n[a] != f->mesh.n[a] || c->sort_bd[a] != p.brick[a] || c->sort_o[a] != f->mesh.origin[a] ||
            c->sort_e[a] != f->mesh.ending[a])
            return false;
    return true;
}

// The sort buffers for nbr bricks and the cloud's capacity: grown (when may_alloc) or
// only checked (ST_E_STATE when too small).
static int sort_buffers(st_cloud *c, int64_t nbr, bool may_alloc) {
    const int32_t rows = sort_rows(c->n, nbr);
    const size_t n1 = (size_t)(c->cap > 0 ? c->cap : 1);
    if (!may_alloc) {
        const bool ok = c->counts && c->cursor && c->brick_start && c->bins_cap >= nbr && c->keys &&
                        c->keys_cap >= n1 && (rows == 0 || (c->hist && c->hist_cap >= (size_t)rows * (size_t)nbr)) &&
                        c->buf[1 - c->cur][0] && c->pid[1 - c->cur];
        return ok ? ST_OK
                  : fail(ST_E_STATE, "the cloud's sort buffers do not hold its %lld parcels; a sort (st_cloud_sort or "
                                     "st_evolve) sizes them",
                         (long long)c->n);
    }
    size_t bins = (size_t)c->bins_cap, bins1 = (size_t)c->bins_cap + 1;
    size_t bins_c = bins;
    ST_TRY(grow(c, &c->counts, bins, (size_t)nbr, "sort counters"));
    ST_TRY(grow(c, &c->cursor, bins_c, (size_t)nbr, "sort cursors"));
    ST_TRY(grow(c, &c->brick_start, bins1, (size_t)nbr + 1, "brick ranges"));
    c->bins_cap = (int32_t)std::max<int64_t>(c->bins_cap, nbr);
    ST_TRY(grow(c, &c->keys, c->keys_cap, n1, "sort keys"));
    if (rows > 0) ST_TRY(grow(c, &c->hist, c->hist_cap, (size_t)rows * (size_t)nbr, "sort histograms"));
    return alloc_set(c, 1 - c->cur);
}

bool sort_buffers_fit(const st_cloud *c) {
    const int32_t rows = sort_rows(c->n, c->nbricks);
    const size_t n1 = (size_t)(c->cap > 0 ? c->cap : 1);
    return c->counts && c->cursor && c->brick_start && c->bins_cap >= c->nbricks && c->keys && c->keys_cap >= n1 &&
           (rows == 0 || (c->hist && c->hist_cap >= (size_t)rows * (size_t)c->nbricks)) && c->buf[1 - c->cur][0] &&
           c->pid[1 - c->cur];
}

template <class T>
int sort_cloud(st_cloud *c, st_field *f, const st_params &p) {
    int32_t nb[3];
    int64_t nbr = 1;
    for (int a = 0; a < 3; ++a) {
        nb[a] = (f->mesh.n[a] + p.brick[a] - 1) / p.brick[a];
        nbr *= nb[a];
    }
    if (nbr > (1 << 26)) return fail(ST_E_RANGE, "too many bricks (%lld)", (long long)nbr);
    if (c->pins > 0 && c->nbricks > 0 && !same_geometry(c, f, p))
        return fail(ST_E_STATE,
                    "%d graph(s), batch(es) or adjoint(s) captured this cloud sorted by brick %dx%dx%d of its "
                    "mesh; a sort by another brick, mesh or boundary form would change what they step: destroy "
                    "them first",
                    c->pins, c->sort_bd[0], c->sort_bd[1], c->sort_bd[2]);
    const int32_t rows = sort_rows(c->n, nbr);
    ST_TRY(sort_buffers(c, nbr, true));
    ST_TRY(enqueue_sort<T>(c, f, p, nb, nbr, rows, nullptr, nullptr));
    c->permuted = true;
    c->sorted = true;
    c->sorted_n = c->n;
    for (int a = 0; a < 3; ++a) {
        c->sort_n[a] = f->mesh.n[a];
        c->sort_bd[a] = p.brick[a];
        c->nb[a] = nb[a];
        c->sort_o[a] = f->mesh.origin[a];
        c->sort_e[a] = f->mesh.ending[a];
    }
    c->sort_periodic = p.boundary;
    c->nbricks = (int32_t)nbr;
    c->strays_pending = false;
    c->last_stray_fraction = -1.0;
    ++c->epoch;
    return ST_OK;
}

// The re-sort queued behind a step of a sorted cloud, run only if the step's statistics
// reduction set gate[0] (its buffers exist: the cloud was sorted with this geometry).
template <class T>
static int queue_resort(st_cloud *c, st_field *f, const st_params &p, const unsigned int *gate) {
    const int64_t nbr = c->nbricks;
    return enqueue_sort<T>(c, f, p, c->nb, nbr, sort_rows(c->n, nbr), gate, c->resort + 2);
}

bool sort_matches(const st_cloud *c, const st_field *f, const st_params &p) {
    if (!c->sorted || c->sort_periodic != p.boundary) return false;
    for (int a = 0; a < 3; ++a)
        if (c->sort_n[a] != f->mesh.n[a] || c->sort_bd[a] != p.brick[a] ||
            c->sort_o[a] != f->mesh.origin[a] || c->sort_e[a] != f->mesh.ending[a])
            return false;
    return true;
}

// Parcels among (x, y, z)[0 .. k) outside the mesh the cloud is sorted for, by the
// device's range check (k_range_check).
template <class T>
int64_t cells_outside(const st_cloud *c, const void *x, const void *y, const void *z, int64_t k) {
    st_mesh m;
    for (int a = 0; a < 3; ++a) {
        m.n[a] = c->sort_n[a];
        m.origin[a] = c->sort_o[a];
        m.ending[a] = c->sort_e[a];
    }
    const Geo g = make_geo(m);
    const T *p[3] = {(const T *)x, (const T *)y, (const T *)z};
    std::atomic<int64_t> bad{0};
    host_parallel(k, [&](int64_t lo, int64_t hi) {
        int64_t b = 0;
        for (int64_t i = lo; i < hi; ++i) {
            bool out = false;
            for (int a = 0; a < 3; ++a) {
                const double v = std::floor(((double)p[a][i] - g.o[a]) * g.rd[a]);
                out |= !(v >= 0.0 && v < (double)g.n[a]);
            }
            b += out;
        }
        bad.fetch_add(b, std::memory_order_relaxed);
    });
    return bad.load();
}

int range_check(st_cloud *c, st_field *f) {
    hipStream_t s = c->ctx->stream;
    ST_HIP(launch_fill_u32(c->bad, 0u, 1, s));
    if (c->prec == ST_FP64)
        ST_HIP(launch_range_check<double>(soa<double>(c, c->cur), (uint32_t)c->fresh_from, (uint32_t)c->n, f->geo,
                                          c->bad, s));
    else
        ST_HIP(launch_range_check<float>(soa<float>(c, c->cur), (uint32_t)c->fresh_from, (uint32_t)c->n, f->geo,
                                         c->bad, s));
    unsigned int bad = 0;
    ST_HIP(hipMemcpyAsync(&bad, c->bad, 4, hipMemcpyDeviceToHost, s));
    ST_HIP(hipStreamSynchronize(s));
    if (bad)
        return fail(ST_E_RANGE,
                    "%u parcels lie outside the mesh [origin, ending): their cell index "
                    "floor((x - origin) / delta) is outside [0, N) on some axis",
                    bad);
    c->fresh = false;
    c->fresh_from = c->n;
    return ST_OK;
}

// Persistent workgroups of the multi-pass launch: the most that fill every CU equally
// (a multiple of the CU count, at most the kernel's occupancy per CU) and split the
// bricks evenly, so no workgroup and no CU owns more bricks than another; the most that
// fit when no such count exists.
template <class T>
static int multi_grid(const st_ctx *ctx, const st_cloud *c, const st_params &p, int32_t tile_cells) {
    const size_t lds = (size_t)tile_cells * 3 * (sizeof(double) + sizeof(T));
    const int per_cu = pass_multi_blocks_per_cu<T>(p.boundary, c->has_rho, p.block, lds);
    if (per_cu < 1) return 0;
    for (int m = per_cu; m >= 1; --m) {
        const int64_t g = (int64_t)m * ctx->cus;
        if (g <= c->nbricks && c->nbricks % g == 0) return (int)g;
    }
    return (int)std::min<int64_t>((int64_t)per_cu * ctx->cus, std::max<int32_t>(1, c->nbricks));
}

int32_t tail_blocks(const st_cloud *c, const st_params &p) {
    if (p.method != ST_METHOD_SORTED || c->n <= c->sorted_n) return 0;
    return (int32_t)((c->n - c->sorted_n + p.block - 1) / p.block);
}

int32_t stats_slots(const st_cloud *c, const st_params &p) {
    return p.method == ST_METHOD_SORTED ? c->nbricks + tail_blocks(c, p) : (int32_t)((c->n + p.block - 1) / p.block);
}

// Takes in the last step's statistics once they have arrived, without waiting: the
// stray share, and -- when that step's re-sort ran -- that every parcel up to the count
// it sorted is in brick order again (sorted_n only ever grows to a count a re-sort is
// known to have covered, so the tail bound never falls short).
static void observe_step(st_cloud *c) {
    if (!c->strays_pending || hipEventQuery(c->strays_ev) != hipSuccess) return;
    c->strays_pending = false;
    c->last_stray_fraction = c->strays_den > 0 ? (double)c->strays_host[1] / (double)c->strays_den : 0.0;
    if (c->strays_host[0] && c->sorted) c->sorted_n = std::max(c->sorted_n, c->strays_den);
}

int cloud_stats(st_cloud *c, int32_t nslots, StatsBuf &sb) {
    ST_TRY(ensure_part(c, nslots));
    sb.part = c->part;
    sb.cap = c->part_cap;
    sb.bbox = c->bbox;
    sb.gate = c->resort;
    return ST_OK;
}

template <class T>
int run_passes(st_ctx *ctx, st_cloud *c, st_field *f, double dt_sub, int nsub, const st_params &p,
               bool capturing, const StatsBuf &sb, bool device_resort, T *ck, const Soa<T> *in,
               const Soa<T> *out) {
    hipStream_t s = ctx->stream;
    if (!capturing) observe_step(c);
    // in/out: the adjoint's forward steps a segment from one set of state arrays into
    // another (one sorted launch, FP64); every other step works in place
    if ((in || out) && (!ck && !out)) return fail(ST_E_INVALID, "state arrays given without a checkpointing pass");
    if ((in || out) && (p.method != ST_METHOD_SORTED || (p.substeps_per_pass > 0 && p.substeps_per_pass < nsub)))
        return fail(ST_E_INVALID, "separate state arrays need one sorted launch for the segment");
    PassArgs<T> a{};
    a.s = in ? *in : soa<T>(c, c->cur);
    a.out = out ? *out : a.s;
    a.U = (const T *)f->U;
    a.src = f->src;
    a.stray = f->stray;
    a.brick_start = c->brick_start;
    a.g = f->geo;
    a.ph = make_phys<T>(c, f, dt_sub, p.cd);
    a.part = nullptr;
    a.nbricks = c->nbricks;
    for (int k = 0; k < 3; ++k) {
        a.nb[k] = c->nb[k];
        a.bd[k] = p.brick[k];
    }
    a.tile_cells = (p.brick[0] + 2) * (p.brick[1] + 2) * (p.brick[2] + 2);
    a.flush = p.flush;
    a.np = (uint32_t)c->n;
    const int per = p.substeps_per_pass > 0 ? p.substeps_per_pass : nsub;
    const int bnd = p.boundary;
    // the unsorted tail of a sorted cloud (parcels injected since its last sort)
    const int32_t tblocks = tail_blocks(c, p);
    PassArgs<T> ta = a;
    ta.tail0 = c->brick_start + c->nbricks;
    // passes of one persistent launch (sorted method, no checkpoint, no tail)
    const int group = p.method == ST_METHOD_SORTED && !ck && tblocks == 0 ? std::max(1, p.passes_per_launch) : 1;
    const int grid = group > 1 ? multi_grid<T>(ctx, c, p, a.tile_cells) : 0;
    if (group > 1 && grid < 1) return fail(ST_E_RANGE, "the multi-pass kernel does not fit: block %d", p.block);
    // A pending zero of the source (st_field_zero_source) is taken by the first launch
    // when that is one sorted owner-flush pass over the bricks: every brick starts its
    // own cells' sums from zero and stores all of them.  Every other launch reads `src`,
    // so the zeros are written first.  Under capture a pending zero is the captured
    // step's own (st_graph_capture writes any earlier one before it begins), and the
    // captured first pass then takes it on every replay.
    const bool lazy = f->src_zero && p.method == ST_METHOD_SORTED && p.flush == ST_FLUSH_OWNER && group == 1;
    if (f->src_zero && !lazy) {
        if (capturing) return fail(ST_E_STATE, "a pending zero of the source inside a capture");
        ST_TRY(materialize_zero(f));
    }

    // Statistics (bounding box, strays) are taken in the last pass only.
    const int32_t nslots = stats_slots(c, p);
    if (nslots > sb.cap)
        return fail(ST_E_RANGE, "the step needs %d statistics slots; %d are available", nslots, sb.cap);
    if (!capturing) ST_HIP(timer_start(ctx, s));
    int passes = 0;
    for (int done = 0; done < nsub;) {
        a.nsteps = per < nsub - done ? per : nsub - done;
        // equal passes that follow are run by the same launch
        int k = 1;
        while (k < group && done + (k + 1) * a.nsteps <= nsub &&
               (per < nsub - done - k * a.nsteps ? per : nsub - done - k * a.nsteps) == a.nsteps)
            ++k;
        const int next = done + k * a.nsteps;
        a.part = next >= nsub ? sb.part : nullptr;
        a.ck = done == 0 ? ck : nullptr;
        a.zero_src = lazy && done == 0 ? 1 : 0;
        ta.zero_src = a.zero_src;
        if (k > 1) {
            ST_HIP(launch_pass_multi<T>(a, k, bnd, p.block, grid, s));
        } else if (tblocks > 0) {  // the bricks and the unsorted tail in one launch
            ta.nsteps = a.nsteps;
            ta.part = a.part;
            ta.ck = a.ck;
            ST_HIP(launch_pass_sorted_tail<T>(ta, (uint32_t)tblocks, bnd, p.block, s));
        } else if (p.method == ST_METHOD_SORTED) {
            ST_HIP(launch_pass_sorted<T>(a, bnd, p.block, s));
        } else {
            ST_HIP(launch_pass_plain<T>(a, bnd, p.block, s));
        }
        passes += k;
        done = next;
    }
    if (lazy) f->src_zero = false;
    // the recorded time covers every pass; per-pass figures divide by the pass count
    if (!capturing) ST_HIP(timer_stop(ctx, s, passes));
    bool resort = device_resort && p.method == ST_METHOD_SORTED && c->sorted && c->n > 0;
    if (resort) {
        const int r = sort_buffers(c, c->nbricks, sb.may_alloc);
        if (r != ST_OK) {
            if (sb.may_alloc) return r;
            resort = false;  // no room to re-sort without allocating: the tail stays
        }
    }
    // The gate is written every step: 0 unless a re-sort was queued and is due.  Outside
    // graph capture the reduction also writes the gate and the strays to the cloud's
    // pinned words (the host never writes them: an earlier step's may still be in flight).
    const bool to_host = p.method == ST_METHOD_SORTED && !capturing;
    StatsTarget t{sb.part, nslots, sb.bbox, sb.gate + 1, sb.gate, 0xffffffffu, to_host ? c->strays_host_dev : nullptr};
    if (resort) {
        // share > resort_fraction  <=>  strays > floor(resort_fraction * n)
        const double th = std::floor(p.resort_fraction * (double)c->n);
        t.threshold = th >= 4294967295.0 ? 0xffffffffu : (uint32_t)th;
    }
    ST_HIP(launch_reduce_stats(t, s));
    if (p.method == ST_METHOD_SORTED && p.flush == ST_FLUSH_OWNER) f->stray_dirty = true;
    if (resort) ST_TRY(queue_resort<T>(c, f, p, sb.gate));
    if (to_host) {
        // the pinned words are read once the event behind this step's re-sort completes
        ST_HIP(hipEventRecord(c->strays_ev, s));
        c->strays_pending = true;
        c->strays_den = c->n;
    }
    ++c->epoch;
    return ST_OK;
}

int prepare(st_ctx *ctx, st_cloud *c, st_field *f, double dt_fluid, int nsub, const st_params *pin,
            st_params &p) {
    ST_TRY(check_ctx(ctx));
    if (!c || !f) return fail(ST_E_INVALID, "null cloud or field");
    if (c->ctx != ctx || f->ctx != ctx) return fail(ST_E_INVALID, "cloud/field belong to another context");
    if (!std::isfinite(dt_fluid) || !(dt_fluid > 0)) return fail(ST_E_INVALID, "dt_fluid must be finite and > 0");
    if (nsub < 1) return fail(ST_E_INVALID, "nsub must be >= 1 (got %d)", nsub);
    if (pin)
        p = *pin;
    else
        st_params_default(&p);
    ST_TRY(validate_params(p));
    if (c->prec != f->prec) return fail(ST_E_INVALID, "cloud and field precision differ");
    if (!c->uploaded) return fail(ST_E_STATE, "cloud has no parcels uploaded");
    if (!f->U_set) return fail(ST_E_STATE, "field velocity U was never uploaded");
    if (c->fresh) ST_TRY(range_check(c, f));
    return ST_OK;
}

// A host-side sort when the cloud is not sorted for this mesh and brick; re-sorts of a
// sorted cloud are decided on the device (run_passes).
int maybe_sort(st_cloud *c, st_field *f, const st_params &p) {
    if (sort_matches(c, f, p)) return ST_OK;
    return c->prec == ST_FP64 ? sort_cloud<double>(c, f, p) : sort_cloud<float>(c, f, p);
}


int sort_any(st_cloud *c, st_field *f, const st_params &p) {
    return c->prec == ST_FP64 ? sort_cloud<double>(c, f, p) : sort_cloud<float>(c, f, p);
}

int ensure_part(st_cloud *c, int32_t nslots) {
    if (nslots <= c->part_cap) return ST_OK;
    if (c->part && c->pins > 0)
        return fail(ST_E_STATE,
                    "the cloud's statistics slots would be reallocated (%d -> %d), but %d graph(s), batch(es) "
                    "or adjoint(s) hold them; destroy them first",
                    c->part_cap, nslots, c->pins);
    // grown by half at least: a cloud whose unsorted tail grows step by step (injection)
    // needs a few more slots every step, and each reallocation waits for the device
    const int32_t cap = (int32_t)std::min<int64_t>(INT32_MAX, std::max<int64_t>(nslots, (int64_t)c->part_cap * 3 / 2));
    if (c->part) ST_HIP(hipFree(c->part));
    c->part = nullptr;
    c->part_cap = 0;
    ST_HIP(hipMalloc((void **)&c->part, (size_t)cap * kStatSlot * sizeof(double)));
    c->part_cap = cap;
    return ST_OK;
}

int check_tile(const st_ctx *ctx, int prec, const int32_t brick[3]) {
    // the halo tile plus the kernels' static LDS (statistics reduction, batch item slot)
    const size_t need =
        (prec == ST_FP64 ? tile_lds_bytes<double>(brick) : tile_lds_bytes<float>(brick)) + kPassStaticLds;
    if (need > ctx->max_lds)
        return fail(ST_E_INVALID,
                    "brick %dx%dx%d needs %zu B of LDS with its halo (%zu B of it the kernel's own); this device "
                    "allows %zu",
                    brick[0], brick[1], brick[2], need, kPassStaticLds, ctx->max_lds);
    return ST_OK;
}

int materialize_zero(st_field *f) {
    if (f->src_zero) {
        ST_HIP(launch_fill_f64(f->src, 0.0, (size_t)f->ncell * 3, f->ctx->stream));
        f->src_zero = false;
    }
    return ST_OK;
}

int merge_source(st_field *f) {
    ST_TRY(materialize_zero(f));
    if (f->stray_dirty) {
        ST_HIP(launch_merge(f->src, f->stray, (size_t)f->ncell * 3, f->ctx->stream));
        f->stray_dirty = false;
    }
    return ST_OK;
}

template <class T>
Phys<T> make_phys(const st_cloud *c, const st_field *f, double dt_sub, double cd) {
    Phys<T> ph;
    ph.dt = (T)dt_sub;
    ph.mu = (T)c->mu_c;
    ph.rho = (T)c->rho_d;
    ph.cd = (T)cd;
    for (int k = 0; k < 3; ++k) ph.L[k] = (T)(f->mesh.ending[k] - f->mesh.origin[k]);
    ph.rdr = c->rdr;
    ph.rho_c = c->rho_c;
    return ph;
}
template Phys<float> make_phys<float>(const st_cloud *, const st_field *, double, double);
template Phys<double> make_phys<double>(const st_cloud *, const st_field *, double, double);
template int run_passes<float>(st_ctx *, st_cloud *, st_field *, double, int, const st_params &, bool,
                               const StatsBuf &, bool, float *, const Soa<float> *, const Soa<float> *);
template int run_passes<double>(st_ctx *, st_cloud *, st_field *, double, int, const st_params &, bool,
                                const StatsBuf &, bool, double *, const Soa<double> *, const Soa<double> *);

}  // namespace host
}  // namespace st

using namespace st::host;

extern "C" {

int st_version(void) { return ST_ABI_VERSION; }

const char *st_last_error(void) { return last_error_text(); }

void st_params_default(st_params *p) {
    if (!p) return;
    p->boundary = ST_WALL;
    p->method = ST_METHOD_SORTED;
    p->flush = ST_FLUSH_OWNER;
    p->substeps_per_pass = 0;
    p->cd = 1.0;
    p->brick[0] = 32;
    p->brick[1] = 8;
    p->brick[2] = 2;
    p->block = 512;
    p->resort_fraction = 0.08;
    p->passes_per_launch = 1;
}

int st_device_count(int32_t *count) {
    if (!count) return fail(ST_E_INVALID, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(ST_E_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *count = n;
    return ST_OK;
}

int st_ctx_create(int32_t device, st_ctx **out) {
    if (!out) return fail(ST_E_INVALID, "null output handle");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(ST_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(ST_E_INVALID, "device %d out of range [0, %d)", device, n);
    st_ctx *c = new (std::nothrow) st_ctx;
    if (!c) return fail(ST_E_NOMEM, "host allocation failed");
    c->device = device;
    e = hipSetDevice(device);
    int lds = 0;
    if (e == hipSuccess && hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
        lds > 0)
        c->max_lds = (size_t)lds;
    if (e == hipSuccess &&
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) == hipSuccess &&
        lds > 0)
        c->lds_per_cu = (size_t)lds;
    if (e == hipSuccess && hipDeviceGetAttribute(&lds, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        lds > 0)
        c->cus = lds;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // st_evolve_noalloc's scratch address on this device (a failed lookup only disables
    // that entry point, which then says so)
    if (e == hipSuccess) (void)noalloc_setup(device);
    // The pass timer's events give timestamps only, and nothing waits on them for data:
    // without the system-scope fence a record costs no L2 writeback and invalidate
    // (≈ 6 µs of idle device between the launches around it, measured).
    for (int k = 0; k < kTimerRing && e == hipSuccess; ++k) {
        e = hipEventCreateWithFlags(&c->tev[k][0], hipEventDisableSystemFence);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->tev[k][1], hipEventDisableSystemFence);
    }
    if (e != hipSuccess) {
        st_ctx_destroy(c);
        return fail(ST_E_DEVICE, "context creation: %s", hipGetErrorString(e));
    }
    *out = c;
    return ST_OK;
}

int st_ctx_destroy(st_ctx *ctx) {
    if (!ctx || ctx->closed) return ST_OK;
    quiesce_ctx(ctx);
    ctx->closed = true;
    if (ctx->refs <= 0) free_ctx(ctx);
    return ST_OK;
}

int st_ctx_synchronize(st_ctx *ctx) {
    ST_TRY(check_ctx(ctx));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    return ST_OK;
}

int st_cloud_create(st_ctx *ctx, int32_t precision, int64_t n, double mu_c, double rho_d,
                    double rho_d_by_rho_c, double rho_c, int32_t per_parcel_rho, st_cloud **out) {
    if (!out) return fail(ST_E_INVALID, "null output handle");
    *out = nullptr;
    ST_TRY(check_ctx(ctx));
    if (precision != ST_FP64 && precision != ST_FP32) return fail(ST_E_INVALID, "precision must be ST_FP64 or ST_FP32");
    if (n < 0 || n > 0x7fffffffLL) return fail(ST_E_RANGE, "parcel count %lld outside [0, 2^31)", (long long)n);
    // each parcel array is addressed with 32-bit byte offsets on the device
    if ((uint64_t)n * (precision == ST_FP64 ? 8u : 4u) >= (1ull << 32))
        return fail(ST_E_RANGE, "parcel arrays of %lld parcels exceed 4 GiB; split the cloud", (long long)n);
    if (!std::isfinite(mu_c) || mu_c < 0) return fail(ST_E_INVALID, "mu_c must be finite and >= 0");
    if (!per_parcel_rho && (!std::isfinite(rho_d) || !(rho_d > 0)))
        return fail(ST_E_INVALID, "rho_d must be finite and > 0");
    if (!per_parcel_rho && (!std::isfinite(rho_d_by_rho_c) || rho_d_by_rho_c < 0))
        return fail(ST_E_INVALID, "rho_d_by_rho_c must be finite and >= 0");
    if (per_parcel_rho && (!std::isfinite(rho_c) || !(rho_c > 0)))
        return fail(ST_E_INVALID, "rho_c must be finite and > 0 with per-parcel density");
    st_cloud *c = new (std::nothrow) st_cloud;
    if (!c) return fail(ST_E_NOMEM, "host allocation failed");
    c->ctx = ctx;
    retain_ctx(ctx);
    c->prec = precision;
    c->es = precision == ST_FP64 ? 8 : 4;
    c->n = n;
    c->cap = n;
    c->mu_c = mu_c;
    c->rho_d = per_parcel_rho ? 1.0 : rho_d;
    c->rdr = per_parcel_rho ? 0.0 : rho_d_by_rho_c;
    c->rho_c = per_parcel_rho ? rho_c : 0.0;
    c->has_rho = per_parcel_rho ? 1 : 0;
    int r = alloc_set(c, 0);
    hipError_t e = hipSuccess;
    if (r == ST_OK) e = hipMalloc((void **)&c->bbox, 6 * sizeof(double));
    if (r == ST_OK && e == hipSuccess) e = hipMalloc((void **)&c->bad, 4);
    if (r == ST_OK && e == hipSuccess) e = hipMalloc((void **)&c->resort, 4 * sizeof(unsigned int));
    if (r == ST_OK && e == hipSuccess) e = st::launch_fill_u32(c->resort, 0u, 4, ctx->stream);
    if (r == ST_OK && e == hipSuccess) e = hipHostMalloc((void **)&c->strays_host, 8, hipHostMallocDefault);
    if (r == ST_OK && e == hipSuccess) e = hipHostGetDevicePointer((void **)&c->strays_host_dev, c->strays_host, 0);
    if (r == ST_OK && e == hipSuccess) e = hipEventCreateWithFlags(&c->strays_ev, hipEventDisableTiming);
    if (r == ST_OK && e == hipSuccess) e = st::launch_iota(c->pid[0], (uint32_t)n, ctx->stream);
    if (r != ST_OK || e != hipSuccess) {
        std::string msg = r != ST_OK ? g_err : std::string(hipGetErrorString(e));
        st_cloud_destroy(c);
        return fail(r != ST_OK ? r : ST_E_DEVICE, "cloud allocation: %s", msg.c_str());
    }
    *out = c;
    return ST_OK;
}

int st_cloud_destroy(st_cloud *c) {
    if (!c) return ST_OK;
    if (c->pins > 0)
        return fail(ST_E_STATE, "cloud is used by %d live graph(s), batch(es) or adjoint(s); destroy them first",
                    c->pins);
    quiesce_ctx(c->ctx);
    for (int w = 0; w < 2; ++w) {
        for (int f = 0; f < kNF; ++f)
            if (c->buf[w][f]) hipFree(c->buf[w][f]);
        if (c->pid[w]) hipFree(c->pid[w]);
    }
    if (c->keys) hipFree(c->keys);
    if (c->counts) hipFree(c->counts);
    if (c->cursor) hipFree(c->cursor);
    if (c->hist) hipFree(c->hist);
    if (c->brick_start) hipFree(c->brick_start);
    if (c->inj_buf) hipFree(c->inj_buf);
    if (c->stage) hipHostFree(c->stage);
    if (c->stage_ev) hipEventDestroy(c->stage_ev);
    if (c->bbox) hipFree(c->bbox);
    if (c->part) hipFree(c->part);
    if (c->bad) hipFree(c->bad);
    if (c->resort) hipFree(c->resort);
    if (c->strays_host) hipHostFree(c->strays_host);
    if (c->strays_ev) hipEventDestroy(c->strays_ev);
    st_ctx *ctx = c->ctx;
    delete c;
    release_ctx(ctx);
    return ST_OK;
}

int st_cloud_upload(st_cloud *c, const void *d, const void *x, const void *y, const void *z,
                    const void *u, const void *v, const void *w, const void *rho) {
    if (!c) return fail(ST_E_INVALID, "null cloud");
    ST_TRY(check_ctx(c->ctx));
    const void *src[kNF] = {d, x, y, z, u, v, w, rho};
    for (int f = 0; f < 7; ++f)
        if (!src[f] && c->n) return fail(ST_E_INVALID, "null parcel array %d", f);
    if (c->has_rho && !rho && c->n) return fail(ST_E_INVALID, "cloud has per-parcel density but rho is NULL");
    for (int f = 0; f < kNF; ++f) {
        if (f == 7 && !c->has_rho) continue;
        if (c->n && !finite_all(src[f], c->n, c->es)) return fail(ST_E_INVALID, "parcel array %d holds a non-finite value", f);
    }
    if (c->n && !positive_all(d, c->n, c->es)) return fail(ST_E_INVALID, "parcel diameters must be > 0");
    if (c->has_rho && c->n && !positive_all(rho, c->n, c->es)) return fail(ST_E_INVALID, "parcel densities must be > 0");
    hipStream_t s = c->ctx->stream;
    c->cur = 0;
    const size_t bytes = (size_t)c->n * c->es;
    for (int f = 0; f < kNF; ++f) {
        if (f == 7 && !c->has_rho) continue;
        if (bytes) ST_HIP(hipMemcpyAsync(c->buf[0][f], src[f], bytes, hipMemcpyHostToDevice, s));
    }
    ST_HIP(st::launch_iota(c->pid[0], (uint32_t)c->n, s));
    ST_HIP(hipStreamSynchronize(s));
    c->uploaded = true;
    c->fresh = true;
    c->fresh_from = 0;
    c->permuted = false;
    c->sorted = false;
    c->strays_pending = false;
    ++c->epoch;
    return ST_OK;
}

int st_cloud_download(st_cloud *c, void *d, void *x, void *y, void *z, void *u, void *v, void *w) {
    if (!c) return fail(ST_E_INVALID, "null cloud");
    ST_TRY(check_ctx(c->ctx));
    void *dst[7] = {d, x, y, z, u, v, w};
    for (int f = 0; f < 7; ++f)
        if (!dst[f] && c->n) return fail(ST_E_INVALID, "null output array %d", f);
    if (!c->uploaded) return fail(ST_E_STATE, "cloud has no parcels uploaded");
    hipStream_t s = c->ctx->stream;
    int from = c->cur;
    if (c->permuted) {
        ST_TRY(alloc_set(c, 1 - c->cur));
        if (c->prec == ST_FP64)
            ST_HIP(st::launch_unpermute<double>(soa<double>(c, c->cur), c->pid[c->cur], soa<double>(c, 1 - c->cur), (uint32_t)c->n, s));
        else
            ST_HIP(st::launch_unpermute<float>(soa<float>(c, c->cur), c->pid[c->cur], soa<float>(c, 1 - c->cur), (uint32_t)c->n, s));
        from = 1 - c->cur;
    }
    const size_t bytes = (size_t)c->n * c->es;
    for (int f = 0; f < 7; ++f)
        if (bytes) ST_HIP(hipMemcpyAsync(dst[f], c->buf[from][f], bytes, hipMemcpyDeviceToHost, s));
    ST_HIP(hipStreamSynchronize(s));
    return ST_OK;
}

int st_cloud_reserve(st_cloud *c, int64_t capacity) {
    if (!c) return fail(ST_E_INVALID, "null cloud");
    ST_TRY(check_ctx(c->ctx));
    if (capacity <= c->cap) return ST_OK;
    if (capacity > 0x7fffffffLL || (uint64_t)capacity * c->es >= (1ull << 32))
        return fail(ST_E_RANGE, "a capacity of %lld parcels exceeds the 4 GiB per-array limit", (long long)capacity);
    if (c->pins > 0)
        return fail(ST_E_STATE, "cloud is used by %d live graph(s), batch(es) or adjoint(s); its arrays cannot move",
                    c->pins);
    hipStream_t s = c->ctx->stream;
    ST_HIP(hipStreamSynchronize(s));
    // the current set moves into arrays of the new capacity; the other set and the sort
    // keys are dropped and come back at the new size when a sort needs them
    void *nb[kNF] = {};
    int32_t *npid = nullptr;
    int r = ST_OK;
    for (int f = 0; f < kNF && r == ST_OK; ++f) {
        if (f == 7 && !c->has_rho) continue;
        if (hipMalloc(&nb[f], (size_t)capacity * c->es) != hipSuccess) r = ST_E_NOMEM;
    }
    if (r == ST_OK && hipMalloc((void **)&npid, (size_t)capacity * 4) != hipSuccess) r = ST_E_NOMEM;
    if (r != ST_OK) {
        for (void *q : nb)
            if (q) (void)hipFree(q);
        if (npid) (void)hipFree(npid);
        return fail(ST_E_NOMEM, "cannot grow the cloud to %lld parcels", (long long)capacity);
    }
    for (int f = 0; f < kNF; ++f) {
        if (f == 7 && !c->has_rho) continue;
        if (c->n) ST_HIP(hipMemcpyAsync(nb[f], c->buf[c->cur][f], (size_t)c->n * c->es, hipMemcpyDeviceToDevice, s));
    }
    if (c->n) ST_HIP(hipMemcpyAsync(npid, c->pid[c->cur], (size_t)c->n * 4, hipMemcpyDeviceToDevice, s));
    ST_HIP(hipStreamSynchronize(s));
    for (int w = 0; w < 2; ++w) {
        for (int f = 0; f < kNF; ++f)
            if (c->buf[w][f]) ST_HIP(hipFree(c->buf[w][f]));
        if (c->pid[w]) ST_HIP(hipFree(c->pid[w]));
        for (int f = 0; f < kNF; ++f) c->buf[w][f] = nullptr;
        c->pid[w] = nullptr;
    }
    for (int f = 0; f < kNF; ++f) c->buf[0][f] = nb[f];
    c->pid[0] = npid;
    c->cur = 0;
    if (c->keys) ST_HIP(hipFree(c->keys));
    c->keys = nullptr;
    c->keys_cap = 0;
    c->cap = capacity;
    ++c->epoch;
    return ST_OK;
}

int st_cloud_inject(st_cloud *c, int64_t k, const void *d, const void *x, const void *y, const void *z,
                    const void *u, const void *v, const void *w, const void *rho) {
    if (!c) return fail(ST_E_INVALID, "null cloud");
    ST_TRY(check_ctx(c->ctx));
    if (k < 0) return fail(ST_E_INVALID, "negative parcel count");
    if (k == 0) return ST_OK;
    if (c->pins > 0)
        return fail(ST_E_STATE,
                    "cloud is used by %d live graph(s), batch(es) or adjoint(s), which step a fixed parcel count",
                    c->pins);
    if (!c->uploaded && c->n > 0) return fail(ST_E_STATE, "upload the cloud's parcels before injecting more");
    const void *src[kNF] = {d, x, y, z, u, v, w, rho};
    for (int f = 0; f < 7; ++f)
        if (!src[f]) return fail(ST_E_INVALID, "null parcel array %d", f);
    if (c->has_rho && !rho) return fail(ST_E_INVALID, "cloud has per-parcel density but rho is NULL");
    for (int f = 0; f < kNF; ++f) {
        if (f == 7 && !c->has_rho) continue;
        if (!finite_all(src[f], k, c->es)) return fail(ST_E_INVALID, "injected array %d holds a non-finite value", f);
    }
    if (!positive_all(d, k, c->es)) return fail(ST_E_INVALID, "injected diameters must be > 0");
    if (c->has_rho && !positive_all(rho, k, c->es)) return fail(ST_E_INVALID, "injected densities must be > 0");
    // A sorted cloud knows its mesh: the injected parcels are checked against it here, on
    // the host, as the first step after an upload checks on the device (floor((x - origin)
    // * rDelta) in [0, N) per axis, in the same double arithmetic), so that no step has
    // to wait for a check.
    const bool checked = c->sorted;
    if (checked) {
        const int64_t bad = c->es == 8 ? cells_outside<double>(c, x, y, z, k) : cells_outside<float>(c, x, y, z, k);
        if (bad)
            return fail(ST_E_RANGE,
                        "%lld injected parcels lie outside the mesh [origin, ending): their cell index "
                        "floor((x - origin) / delta) is outside [0, N) on some axis",
                        (long long)bad);
    }
    const int64_t n1 = c->n + k;
    if (n1 > c->cap) ST_TRY(st_cloud_reserve(c, std::max<int64_t>(n1, c->cap + c->cap / 2)));
    hipStream_t s = c->ctx->stream;
    // The host arrays go through a pinned staging buffer of the cloud: copied into it on
    // the host, while the device may still be stepping earlier work, then read by a
    // kernel queued on the stream; the call returns without waiting for the stream.
    const int nf = c->has_rho ? kNF : 7;
    const size_t part = (size_t)k * c->es;
    if (!c->stage_ev) ST_HIP(hipEventCreateWithFlags(&c->stage_ev, hipEventDisableTiming));
    else ST_HIP(hipEventSynchronize(c->stage_ev));  // the last copy out of the staging buffer is done
    if (part * nf > c->stage_cap) {
        if (c->stage) ST_HIP(hipHostFree(c->stage));
        c->stage = nullptr;
        c->stage_cap = 0;
        if (hipHostMalloc(&c->stage, part * nf, hipHostMallocDefault) != hipSuccess)
            return fail(ST_E_NOMEM, "cannot stage %lld injected parcels in pinned host memory", (long long)k);
        c->stage_cap = part * nf;
    }
    const size_t es = c->es;
    host_parallel(k, [&](int64_t lo, int64_t hi) {
        for (int f = 0; f < nf; ++f)
            std::memcpy((char *)c->stage + (size_t)f * part + (size_t)lo * es, (const char *)src[f] + (size_t)lo * es,
                        (size_t)(hi - lo) * es);
    });
    void *dstage = nullptr;
    ST_HIP(hipHostGetDevicePointer(&dstage, c->stage, 0));
    ST_TRY(copy_injected(c, dstage, c->n, k));
    // in a sorted cloud the new parcels are its unsorted tail: ordered by cell, the tail's
    // waves can sum their deposits before the atomics
    ST_TRY(sort_injected(c, c->n, k));
    ST_HIP(hipEventRecord(c->stage_ev, s));
    if (!checked) {
        if (!c->fresh) c->fresh_from = c->n;
        c->fresh = true;
    }
    c->n = n1;
    c->uploaded = true;
    // the new parcels sit at the end: sorted steps step them as the unsorted tail until
    // a re-sort takes them in
    ++c->epoch;
    return ST_OK;
}

int st_cloud_compact(st_cloud *c, int64_t *n_live) {
    if (!c) return fail(ST_E_INVALID, "null cloud");
    ST_TRY(check_ctx(c->ctx));
    if (c->pins > 0)
        return fail(ST_E_STATE,
                    "cloud is used by %d live graph(s), batch(es) or adjoint(s), which step a fixed parcel count",
                    c->pins);
    if (!c->uploaded || c->n == 0) {
        if (n_live) *n_live = c->n;
        return ST_OK;
    }
    hipStream_t s = c->ctx->stream;
    const uint32_t np = (uint32_t)c->n;
    const uint32_t tiles = st::scan_tiles(np);
    ST_TRY(alloc_set(c, 1 - c->cur));
    uint32_t *tmp = nullptr;
    const size_t words = 4 * (size_t)np + 4 * ((size_t)tiles + 1);
    ST_HIP(hipMalloc((void **)&tmp, words * 4));
    uint32_t *live_dev = tmp, *live_up = tmp + np, *pos_dev = tmp + 2 * (size_t)np, *new_up = tmp + 3 * (size_t)np;
    uint32_t *tile_sums = tmp + 4 * (size_t)np;
    int32_t *off_dev = (int32_t *)(tile_sums + tiles + 1);
    int32_t *off_up = off_dev + tiles + 1;
    uint32_t *cursor = (uint32_t *)(off_up + tiles + 1);
    hipError_t e = c->prec == ST_FP64
                       ? st::launch_compact<double>(soa<double>(c, c->cur), c->pid[c->cur], soa<double>(c, 1 - c->cur),
                                                    c->pid[1 - c->cur], np, live_dev, live_up, pos_dev, new_up,
                                                    tile_sums, off_dev, off_up, cursor, s)
                       : st::launch_compact<float>(soa<float>(c, c->cur), c->pid[c->cur], soa<float>(c, 1 - c->cur),
                                                   c->pid[1 - c->cur], np, live_dev, live_up, pos_dev, new_up,
                                                   tile_sums, off_dev, off_up, cursor, s);
    int32_t live = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&live, off_dev + tiles, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && live > 0) {
        if (c->prec == ST_FP64)
            e = st::launch_copy_back<double>(soa<double>(c, 1 - c->cur), c->pid[1 - c->cur], soa<double>(c, c->cur),
                                             c->pid[c->cur], (uint32_t)live, nullptr, nullptr, s);
        else
            e = st::launch_copy_back<float>(soa<float>(c, 1 - c->cur), c->pid[1 - c->cur], soa<float>(c, c->cur),
                                            c->pid[c->cur], (uint32_t)live, nullptr, nullptr, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(ST_E_DEVICE, "compaction: %s", hipGetErrorString(e));
    c->n = live;
    c->permuted = true;
    c->sorted = false;
    c->fresh_from = c->fresh ? 0 : c->n;  // unchecked parcels moved: check them all
    ++c->epoch;
    if (n_live) *n_live = live;
    return ST_OK;
}

int st_cloud_count(const st_cloud *c, int64_t *n) {
    if (!c || !n) return fail(ST_E_INVALID, "null argument");
    *n = c->n;
    return ST_OK;
}

int st_field_create(st_ctx *ctx, const st_mesh *mesh, int32_t precision, st_field **out) {
    if (!out) return fail(ST_E_INVALID, "null output handle");
    *out = nullptr;
    ST_TRY(check_ctx(ctx));
    if (!mesh) return fail(ST_E_INVALID, "null mesh");
    if (precision != ST_FP64 && precision != ST_FP32) return fail(ST_E_INVALID, "precision must be ST_FP64 or ST_FP32");
    int64_t nc = 1;
    for (int a = 0; a < 3; ++a) {
        if (mesh->n[a] < 1) return fail(ST_E_INVALID, "mesh n[%d] = %d must be >= 1", a, mesh->n[a]);
        if (!std::isfinite(mesh->origin[a]) || !std::isfinite(mesh->ending[a]) || !(mesh->ending[a] > mesh->origin[a]))
            return fail(ST_E_INVALID, "mesh axis %d needs finite origin < ending", a);
        nc *= mesh->n[a];
    }
    if (nc * 3 >= 0x7fffffffLL) return fail(ST_E_RANGE, "mesh of %lld cells is above the 32-bit index limit", (long long)nc);
    st_field *f = new (std::nothrow) st_field;
    if (!f) return fail(ST_E_NOMEM, "host allocation failed");
    f->ctx = ctx;
    retain_ctx(ctx);
    f->mesh = *mesh;
    f->geo = make_geo(*mesh);
    f->prec = precision;
    f->es = precision == ST_FP64 ? 8 : 4;
    f->ncell = nc;
    hipError_t e = hipMalloc(&f->U, (size_t)nc * 3 * f->es);
    if (e == hipSuccess) e = hipMalloc((void **)&f->src, (size_t)nc * 3 * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&f->stray, (size_t)nc * 3 * 8);
    if (e == hipSuccess) e = st::launch_fill_f64(f->src, 0.0, (size_t)nc * 3, ctx->stream);
    if (e == hipSuccess) e = st::launch_fill_f64(f->stray, 0.0, (size_t)nc * 3, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        st_field_destroy(f);
        return fail(e == hipErrorOutOfMemory ? ST_E_NOMEM : ST_E_DEVICE, "field allocation: %s", hipGetErrorString(e));
    }
    *out = f;
    return ST_OK;
}

int st_field_destroy(st_field *f) {
    if (!f) return ST_OK;
    if (f->pins > 0)
        return fail(ST_E_STATE, "field is used by %d live graph(s), batch(es) or adjoint(s); destroy them first",
                    f->pins);
    quiesce_ctx(f->ctx);
    if (f->U) hipFree(f->U);
    if (f->src) hipFree(f->src);
    if (f->stray) hipFree(f->stray);
    st_ctx *ctx = f->ctx;
    delete f;
    release_ctx(ctx);
    return ST_OK;
}

int st_field_upload_U(st_field *f, const void *U, int32_t elem_bytes) {
    if (!f) return fail(ST_E_INVALID, "null field");
    ST_TRY(check_ctx(f->ctx));
    if (!U) return fail(ST_E_INVALID, "null U");
    if (elem_bytes != 4 && elem_bytes != 8) return fail(ST_E_INVALID, "elem_bytes must be 4 or 8 (got %d)", elem_bytes);
    const int64_t n = f->ncell * 3;
    if (!finite_all(U, n, (size_t)elem_bytes)) return fail(ST_E_INVALID, "U holds a non-finite value");
    hipStream_t s = f->ctx->stream;
    if ((size_t)elem_bytes == f->es) {
        ST_HIP(hipMemcpyAsync(f->U, U, (size_t)n * f->es, hipMemcpyHostToDevice, s));
        ST_HIP(hipStreamSynchronize(s));
    } else {
        std::vector<unsigned char> tmp((size_t)n * f->es);
        if (f->es == 8) {
            const float *a = (const float *)U;
            double *b = (double *)tmp.data();
            for (int64_t i = 0; i < n; ++i) b[i] = (double)a[i];
        } else {
            const double *a = (const double *)U;
            float *b = (float *)tmp.data();
            for (int64_t i = 0; i < n; ++i) b[i] = (float)a[i];
        }
        ST_HIP(hipMemcpyAsync(f->U, tmp.data(), tmp.size(), hipMemcpyHostToDevice, s));
        ST_HIP(hipStreamSynchronize(s));
    }
    f->U_set = true;
    return ST_OK;
}

int st_field_zero_source(st_field *f) {
    if (!f) return fail(ST_E_INVALID, "null field");
    ST_TRY(check_ctx(f->ctx));
    // The zeros of `src` are written by the next step's first sorted owner-flush pass,
    // which then stores every cell (no fill launch in front of it); any other use of the
    // source writes them first (materialize_zero).
    f->src_zero = true;
    if (f->stray_dirty) ST_HIP(st::launch_fill_f64(f->stray, 0.0, (size_t)f->ncell * 3, f->ctx->stream));
    f->stray_dirty = false;
    return ST_OK;
}

int st_field_merge_source(st_field *f) {
    if (!f) return fail(ST_E_INVALID, "null field");
    ST_TRY(check_ctx(f->ctx));
    return merge_source(f);
}

int st_field_download_source(st_field *f, void *out, int32_t elem_bytes) {
    if (!f) return fail(ST_E_INVALID, "null field");
    ST_TRY(check_ctx(f->ctx));
    if (!out) return fail(ST_E_INVALID, "null output");
    if (elem_bytes != 4 && elem_bytes != 8) return fail(ST_E_INVALID, "elem_bytes must be 4 or 8 (got %d)", elem_bytes);
    hipStream_t s = f->ctx->stream;
    const size_t n = (size_t)f->ncell * 3;
    ST_TRY(merge_source(f));
    if (elem_bytes == 8) {
        ST_HIP(hipMemcpyAsync(out, f->src, n * 8, hipMemcpyDeviceToHost, s));
        ST_HIP(hipStreamSynchronize(s));
    } else {
        std::vector<double> tmp(n);
        ST_HIP(hipMemcpyAsync(tmp.data(), f->src, n * 8, hipMemcpyDeviceToHost, s));
        ST_HIP(hipStreamSynchronize(s));
        float *o = (float *)out;
        for (size_t i = 0; i < n; ++i) o[i] = (float)tmp[i];
    }
    return ST_OK;
}

int st_evolve(st_ctx *ctx, st_cloud *c, st_field *f, double dt_fluid, int32_t nsub, const st_params *params,
              double bbox_out[6]) {
    st_params p;
    ST_TRY(prepare(ctx, c, f, dt_fluid, nsub, params, p));
    if (c->n == 0) {
        if (bbox_out)
            for (int k = 0; k < 6; ++k) bbox_out[k] = k < 3 ? INFINITY : -INFINITY;
        return ST_OK;
    }
    if (p.method == ST_METHOD_SORTED) {
        ST_TRY(check_tile(ctx, c->prec, p.brick));
        ST_TRY(maybe_sort(c, f, p));
    }
    StatsBuf sb;
    ST_TRY(cloud_stats(c, stats_slots(c, p), sb));
    if (c->prec == ST_FP64)
        ST_TRY(run_passes<double>(ctx, c, f, substep_dt(c->prec, dt_fluid, nsub), nsub, p, false, sb, true));
    else
        ST_TRY(run_passes<float>(ctx, c, f, substep_dt(c->prec, dt_fluid, nsub), nsub, p, false, sb, true));
    if (bbox_out) {
        ST_HIP(hipMemcpyAsync(bbox_out, sb.bbox, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
    }
    return ST_OK;
}

int st_cloud_stray_fraction(st_cloud *c, double *fraction) {
    if (!c || !fraction) return fail(ST_E_INVALID, "null argument");
    ST_TRY(check_ctx(c->ctx));
    if (c->strays_pending) {
        ST_HIP(hipEventSynchronize(c->strays_ev));
        observe_step(c);
    }
    *fraction = c->last_stray_fraction;
    return ST_OK;
}

int st_cloud_resort_count(st_cloud *c, int64_t *count) {
    if (!c || !count) return fail(ST_E_INVALID, "null argument");
    ST_TRY(check_ctx(c->ctx));
    unsigned int v = 0;
    ST_HIP(hipMemcpyAsync(&v, c->resort + 2, sizeof v, hipMemcpyDeviceToHost, c->ctx->stream));
    ST_HIP(hipStreamSynchronize(c->ctx->stream));
    *count = (int64_t)v;
    return ST_OK;
}

int st_cloud_sort(st_ctx *ctx, st_cloud *c, st_field *f, const st_params *params) {
    st_params p;
    ST_TRY(prepare(ctx, c, f, 1.0, 1, params, p));
    if (c->n == 0) return ST_OK;
    return c->prec == ST_FP64 ? sort_cloud<double>(c, f, p) : sort_cloud<float>(c, f, p);
}

int st_graph_capture(st_ctx *ctx, st_cloud *c, st_field *f, double dt_fluid, int32_t nsub,
                     const st_params *params, int32_t zero_source, st_graph **out) {
    if (!out) return fail(ST_E_INVALID, "null output handle");
    *out = nullptr;
    st_params p;
    ST_TRY(prepare(ctx, c, f, dt_fluid, nsub, params, p));
    if (c->n == 0) return fail(ST_E_INVALID, "cannot capture a step of an empty cloud");
    if (p.method == ST_METHOD_SORTED) ST_TRY(check_tile(ctx, c->prec, p.brick));
    if (p.method == ST_METHOD_SORTED && !sort_matches(c, f, p))
        ST_TRY(c->prec == ST_FP64 ? sort_cloud<double>(c, f, p) : sort_cloud<float>(c, f, p));
    // every allocation happens before the capture: none is permitted inside it
    if (p.method == ST_METHOD_SORTED) ST_TRY(sort_buffers(c, c->nbricks, true));
    StatsBuf sb;
    ST_TRY(cloud_stats(c, stats_slots(c, p), sb));
    ST_TRY(materialize_zero(f));  // a zero pending now belongs to no replay
    ST_HIP(hipStreamSynchronize(ctx->stream));
    st_graph *g = new (std::nothrow) st_graph;
    if (!g) return fail(ST_E_NOMEM, "host allocation failed");
    g->ctx = ctx;
    g->field = f;
    g->cloud = c;
    g->owner_flush = p.method == ST_METHOD_SORTED && p.flush == ST_FLUSH_OWNER;
    hipStream_t s = ctx->stream;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete g;
        return fail(ST_E_DEVICE, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    int r = ST_OK;
    if (zero_source) {
        // with the owner flush and one pass per launch the captured first pass takes the
        // zero (no fill node); the stray sums are cleared by a fill kernel either way
        if (g->owner_flush && p.passes_per_launch <= 1)
            f->src_zero = true;
        else
            e = st::launch_fill_f64(f->src, 0.0, (size_t)f->ncell * 3, s);
        if (e == hipSuccess && g->owner_flush) e = st::launch_fill_f64(f->stray, 0.0, (size_t)f->ncell * 3, s);
        if (e != hipSuccess) r = fail(ST_E_DEVICE, "capture fill: %s", hipGetErrorString(e));
    }
    if (r == ST_OK)
        r = c->prec == ST_FP64
                ? run_passes<double>(ctx, c, f, substep_dt(c->prec, dt_fluid, nsub), nsub, p, true, sb, true)
                : run_passes<float>(ctx, c, f, substep_dt(c->prec, dt_fluid, nsub), nsub, p, true, sb, true);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s, &graph);
    f->src_zero = false;  // a zero the capture set belongs to the replays, not to the field
    if (r != ST_OK || e != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        delete g;
        return r != ST_OK ? r : fail(ST_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(graph);
        delete g;
        return fail(ST_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    g->graph = graph;
    retain_ctx(ctx);
    pin(c);
    pin(f);
    *out = g;
    return ST_OK;
}

int st_graph_launch(st_graph *g) {
    if (!g) return fail(ST_E_INVALID, "null graph");
    ST_TRY(check_ctx(g->ctx));
    // the replay's kernels read and write `src` as it is in memory
    if (g->field) ST_TRY(materialize_zero(g->field));
    for (st_field *f : g->more_fields) ST_TRY(materialize_zero(f));
    ST_HIP(hipGraphLaunch(g->exec, g->ctx->stream));
    if (g->owner_flush) {
        if (g->field) g->field->stray_dirty = true;
        for (st_field *f : g->more_fields) f->stray_dirty = true;
    }
    if (g->cloud) ++g->cloud->epoch;
    for (st_cloud *c : g->more_clouds) ++c->epoch;
    return ST_OK;
}

int st_graph_destroy(st_graph *g) {
    if (!g) return ST_OK;
    quiesce_ctx(g->ctx);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->cloud) unpin(g->cloud);
    if (g->field) unpin(g->field);
    for (st_cloud *c : g->more_clouds) unpin(c);
    for (st_field *f : g->more_fields) unpin(f);
    if (g->batch) batch_unpin(g->batch);
    st_ctx *ctx = g->ctx;
    delete g;
    release_ctx(ctx);
    return ST_OK;
}

int st_last_kernel_ms(st_ctx *ctx, double *ms, int32_t *launches) {
    ST_TRY(check_ctx(ctx));
    if (!ms || !launches) return fail(ST_E_INVALID, "null argument");
    *ms = -1.0;
    *launches = 0;
    if (ctx->tcount == 0) return ST_OK;
    const int slot = (int)((ctx->tcount - 1) % kTimerRing);
    *launches = ctx->tlaunches[slot];
    if (*launches <= 0) return ST_OK;
    ST_HIP(hipEventSynchronize(ctx->tev[slot][1]));
    float t = 0.f;
    ST_HIP(hipEventElapsedTime(&t, ctx->tev[slot][0], ctx->tev[slot][1]));
    *ms = (double)t / *launches;
    return ST_OK;
}

int st_diag_kernel_times(st_ctx *ctx, int32_t max, double *ms, int32_t *launches, int32_t *count) {
    ST_TRY(check_ctx(ctx));
    if (!ms || !launches || !count || max < 1) return fail(ST_E_INVALID, "null argument or max < 1");
    const int64_t have = std::min<int64_t>(ctx->tcount, kTimerRing);
    const int32_t n = (int32_t)std::min<int64_t>(have, max);
    *count = n;
    if (n == 0) return ST_OK;
    ST_HIP(hipEventSynchronize(ctx->tev[(ctx->tcount - 1) % kTimerRing][1]));
    for (int32_t i = 0; i < n; ++i) {  // oldest first
        const int slot = (int)((ctx->tcount - n + i) % kTimerRing);
        launches[i] = ctx->tlaunches[slot];
        ms[i] = -1.0;
        if (launches[i] <= 0) continue;
        float t = 0.f;
        ST_HIP(hipEventElapsedTime(&t, ctx->tev[slot][0], ctx->tev[slot][1]));
        ms[i] = (double)t / launches[i];
    }
    return ST_OK;
}

int st_host_alloc(int64_t bytes, void **out) {
    if (out) *out = nullptr;
    if (!out || bytes < 1) return fail(ST_E_INVALID, "need bytes >= 1 and an output");
    const hipError_t e = hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(e == hipErrorOutOfMemory ? ST_E_NOMEM : ST_E_DEVICE, "page-locked host allocation of %lld B: %s",
                    (long long)bytes, hipGetErrorString(e));
    }
    return ST_OK;
}

int st_host_free(void *ptr) {
    if (ptr) ST_HIP(hipHostFree(ptr));
    return ST_OK;
}

int st_diag_copy_bandwidth(st_ctx *ctx, int64_t bytes, int32_t reps, double *gbs) {
    ST_TRY(check_ctx(ctx));
    if (!gbs || bytes < 16 || reps < 1) return fail(ST_E_INVALID, "need bytes >= 16, reps >= 1 and an output");
    bytes &= ~(int64_t)15;
    void *a = nullptr, *b = nullptr;
    ST_HIP(hipMalloc(&a, (size_t)bytes));
    hipError_t e = hipMalloc(&b, (size_t)bytes);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(ST_E_NOMEM, "copy buffers: %s", hipGetErrorString(e));
    }
    hipStream_t s = ctx->stream;
    e = st::launch_fill_u32((uint32_t *)a, 0x3f800000u, (size_t)bytes / 4, s);
    if (e == hipSuccess) e = st::launch_copy(a, b, (size_t)bytes, s);  // warm-up
    const int slot = (int)(ctx->tcount % kTimerRing);
    if (e == hipSuccess) e = timer_start(ctx, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = st::launch_copy(a, b, (size_t)bytes, s);
    if (e == hipSuccess) e = timer_stop(ctx, s, reps);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->tev[slot][1]);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->tev[slot][0], ctx->tev[slot][1]);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(ST_E_DEVICE, "copy: %s", hipGetErrorString(e));
    *gbs = 2.0 * (double)bytes * reps / ((double)ms * 1e-3) / 1e9;
    return ST_OK;
}

}  // extern "C"
