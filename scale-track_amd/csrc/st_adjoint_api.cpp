// st_adjoint_api.cpp -- C-ABI of the checkpointed adjoint of the parcel step (FP64).
//
// Forward: the parcel step over nsub substeps, run as passes of K substeps with the
// parcel state (x, v) saved before each pass.  Backward: one reverse launch per pass,
// last first, from the saved state (st_adjoint.hip).  Gradients of a loss of the
// deposited source (and optionally of the final state) with respect to the per-parcel
// diameter and density and to the drag-coefficient multiplier.
#include "st_handles.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

using namespace st::host;

struct st_adjoint {
    st_ctx *ctx = nullptr;
    st_cloud *cloud = nullptr;
    st_field *field = nullptr;
    st_params p{};
    int32_t nsub = 0, K = 0, nseg = 0;
    double dt_sub = 0.0;
    uint32_t n = 0;
    bool forward_done = false;
    bool backward_done = false;
    // seeds given since the forward; one not given is zero, and the sweep reads nullptr
    // in its place (no fill in the forward, no read in every segment)
    bool xb_set = false, vseed_set = false, G_set = false;
    int64_t epoch = -1;  // the cloud's epoch right after the forward
    double *ck = nullptr;    // nseg * 6 * n
    double *xb = nullptr;    // 3n, the final-position cotangent; the sweep only reads it (the
                             // adjoint of x is constant through the substeps, so it is dL/dx0)
    double *vb = nullptr;    // 3n, propagated by the backward sweep (then dL/dv0)
    double *vseed = nullptr; // 3n, the final-velocity cotangent each sweep starts from
    double *fbar = nullptr;  // n
    double *mbar = nullptr;  // n
    double *G = nullptr;     // 3 * ncell
    double *UG = nullptr;    // 6 * ncell: U and G of each cell side by side (reverse sweep)
    double *T = nullptr;     // 3 * ncell: target source of st_adjoint_source_misfit
    double *tmp = nullptr;   // 6n staging
    double *scal = nullptr;  // [0] loss, [1] d L / d cd, [2, 2 + kSumBlocks) partial sums
};

namespace {

int destroy(st_adjoint *a) {
    if (!a) return ST_OK;
    quiesce_ctx(a->ctx);
    if (a->cloud) unpin(a->cloud);
    if (a->field) unpin(a->field);
    for (double *q : {a->ck, a->xb, a->vb, a->vseed, a->fbar, a->mbar, a->G, a->UG, a->T, a->tmp, a->scal})
        if (q) (void)hipFree(q);
    st_ctx *ctx = a->ctx;
    delete a;
    release_ctx(ctx);
    return ST_OK;
}

int check(st_adjoint *a) {
    if (!a) return fail(ST_E_INVALID, "null adjoint handle");
    return check_ctx(a->ctx);
}

// The reverse sweep replays the forward from its checkpoints in the device order of
// the forward: a cloud stepped, sorted or uploaded since then has another state.
int check_unchanged(st_adjoint *a) {
    if (a->forward_done && a->cloud->epoch != a->epoch)
        return fail(ST_E_STATE,
                    "the cloud was stepped, sorted or uploaded after st_adjoint_forward; run the forward again");
    return ST_OK;
}

}  // namespace

extern "C" {

int st_adjoint_create(st_ctx *ctx, st_cloud *c, st_field *f, int32_t nsub, int32_t checkpoint_every,
                      const st_params *params, st_adjoint **out) {
    if (!out) return fail(ST_E_INVALID, "null output handle");
    *out = nullptr;
    st_params p;
    ST_TRY(prepare(ctx, c, f, 1.0, nsub, params, p));
    if (c->prec != ST_FP64) return fail(ST_E_UNSUPPORTED, "the adjoint runs in FP64; create the cloud with ST_FP64");
    if (p.boundary == ST_CHANNEL)
        return fail(ST_E_UNSUPPORTED, "the adjoint takes ST_WALL or ST_PERIODIC (removals are not differentiable)");
    if (checkpoint_every < 1 || checkpoint_every > nsub)
        return fail(ST_E_INVALID, "checkpoint_every must be in [1, nsub] (got %d)", checkpoint_every);
    st_adjoint *a = new (std::nothrow) st_adjoint;
    if (!a) return fail(ST_E_NOMEM, "host allocation failed");
    a->ctx = ctx;
    retain_ctx(ctx);
    a->cloud = c;
    a->field = f;
    pin(c);
    pin(f);
    a->p = p;
    a->p.substeps_per_pass = checkpoint_every;
    a->nsub = nsub;
    a->K = checkpoint_every;
    a->nseg = (nsub + checkpoint_every - 1) / checkpoint_every;
    a->n = (uint32_t)c->n;
    const size_t n = a->n > 0 ? a->n : 1;
    const size_t nc3 = (size_t)f->ncell * 3;
    hipError_t e = hipMalloc((void **)&a->ck, (size_t)a->nseg * 6 * n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->xb, 3 * n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->vb, 3 * n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->vseed, 3 * n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->fbar, n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->mbar, n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->G, nc3 * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->UG, 2 * nc3 * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->T, nc3 * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->tmp, 6 * n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&a->scal, (2 + st::kSumBlocks) * 8);
    if (e == hipSuccess) e = st::launch_fill_f64(a->G, 0.0, nc3, ctx->stream);
    if (e != hipSuccess) {
        destroy(a);
        return fail(e == hipErrorOutOfMemory ? ST_E_NOMEM : ST_E_DEVICE, "adjoint allocation: %s",
                    hipGetErrorString(e));
    }
    *out = a;
    return ST_OK;
}

int st_adjoint_destroy(st_adjoint *a) { return destroy(a); }

int st_adjoint_forward(st_adjoint *a, double dt_fluid) {
    ST_TRY(check(a));
    st_cloud *c = a->cloud;
    st_field *f = a->field;
    st_params p;
    ST_TRY(prepare(a->ctx, c, f, dt_fluid, a->nsub, &a->p, p));
    hipStream_t s = a->ctx->stream;
    // sorted, and with no unsorted tail: the checkpoints follow the sorted passes.  The
    // adjoint's own hold on the cloud does not stop its forward from sorting the cloud
    // for its parameters or growing the statistics slots: the forward defines what the
    // adjoint holds.  Holds of graphs, batches and other adjoints still do.
    unpin(c);
    int r = ST_OK;
    if (p.method == ST_METHOD_SORTED && (!sort_matches(c, f, p) || c->sorted_n != c->n)) r = sort_any(c, f, p);
    StatsBuf sb;
    if (r == ST_OK) r = cloud_stats(c, stats_slots(c, p), sb);
    pin(c);
    ST_TRY(r);
    ST_TRY(st_field_zero_source(f));
    a->dt_sub = substep_dt(ST_FP64, dt_fluid, a->nsub);
    const uint32_t n = a->n;
    // Sorted method: segment s steps its state from checkpoint s into checkpoint s + 1
    // (one launch each), so no segment copies a state it can write where it is needed;
    // the first reads the cloud's arrays and also writes checkpoint 0, the last writes
    // the cloud's arrays.
    const st::Soa<double> cloud = soa<double>(c, c->cur);
    auto ck_soa = [&](int32_t k) {
        st::Soa<double> q = cloud;  // d and rho stay the cloud's
        double *b = a->ck + (size_t)k * 6 * n;
        q.x = b;
        q.y = b + n;
        q.z = b + 2 * (size_t)n;
        q.u = b + 3 * (size_t)n;
        q.v = b + 4 * (size_t)n;
        q.w = b + 5 * (size_t)n;
        return q;
    };
    for (int32_t seg = 0; seg < a->nseg; ++seg) {
        double *ck = a->ck + (size_t)seg * 6 * n;
        const int32_t ks = std::min(a->K, a->nsub - seg * a->K);
        if (p.method == ST_METHOD_SORTED) {
            const st::Soa<double> in = seg == 0 ? cloud : ck_soa(seg);
            const st::Soa<double> out = seg == a->nseg - 1 ? cloud : ck_soa(seg + 1);
            ST_TRY(run_passes<double>(a->ctx, c, f, a->dt_sub, ks, p, false, sb, false, seg == 0 ? ck : nullptr, &in,
                                      &out));
        } else {
            st::Soa<double> cur = soa<double>(c, c->cur);
            const double *fields[6] = {cur.x, cur.y, cur.z, cur.u, cur.v, cur.w};
            ST_HIP(st::launch_save6(fields, ck, n, s));
            ST_TRY(run_passes<double>(a->ctx, c, f, a->dt_sub, ks, p, false, sb, false));
        }
    }
    a->xb_set = a->vseed_set = a->G_set = false;
    a->forward_done = true;
    a->backward_done = false;
    a->epoch = c->epoch;
    return ST_OK;
}

int st_adjoint_source_misfit(st_adjoint *a, const double *target, double *loss) {
    ST_TRY(check(a));
    if (!target || !loss) return fail(ST_E_INVALID, "null target or loss");
    if (!a->forward_done) return fail(ST_E_STATE, "run st_adjoint_forward first");
    st_field *f = a->field;
    hipStream_t s = a->ctx->stream;
    const size_t nc3 = (size_t)f->ncell * 3;
    ST_TRY(merge_source(f));
    hipError_t e = hipMemcpyAsync(a->T, target, nc3 * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = st::launch_misfit(f->src, a->T, a->G, a->scal, a->scal + 2, nc3, s);
    if (e == hipSuccess) a->G_set = true;
    if (e == hipSuccess) e = hipMemcpyAsync(loss, a->scal, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(ST_E_DEVICE, "source misfit: %s", hipGetErrorString(e));
    return ST_OK;
}

int st_adjoint_set_source_cotangent(st_adjoint *a, const double *g) {
    ST_TRY(check(a));
    if (!g) return fail(ST_E_INVALID, "null cotangent");
    if (!a->forward_done) return fail(ST_E_STATE, "run st_adjoint_forward first");
    const size_t nc3 = (size_t)a->field->ncell * 3;
    ST_HIP(hipMemcpyAsync(a->G, g, nc3 * 8, hipMemcpyHostToDevice, a->ctx->stream));
    ST_HIP(hipStreamSynchronize(a->ctx->stream));
    a->G_set = true;
    return ST_OK;
}

int st_adjoint_set_state_cotangent(st_adjoint *a, const double *xbar, const double *vbar) {
    ST_TRY(check(a));
    if (!a->forward_done) return fail(ST_E_STATE, "run st_adjoint_forward first");
    ST_TRY(check_unchanged(a));
    hipStream_t s = a->ctx->stream;
    const uint32_t n = a->n;
    const int32_t *pid = a->cloud->pid[a->cloud->cur];
    if (xbar) {
        ST_HIP(hipMemcpyAsync(a->tmp, xbar, 3 * (size_t)n * 8, hipMemcpyHostToDevice, s));
        ST_HIP(st::launch_gather_pid(a->tmp, pid, a->xb, n, 3, s));
        a->xb_set = true;
    }
    if (vbar) {
        ST_HIP(hipMemcpyAsync(a->tmp + 3 * (size_t)n, vbar, 3 * (size_t)n * 8, hipMemcpyHostToDevice, s));
        ST_HIP(st::launch_gather_pid(a->tmp + 3 * (size_t)n, pid, a->vseed, n, 3, s));
        a->vseed_set = true;
    }
    ST_HIP(hipStreamSynchronize(s));
    return ST_OK;
}

int st_adjoint_backward(st_adjoint *a) {
    ST_TRY(check(a));
    if (!a->forward_done) return fail(ST_E_STATE, "run st_adjoint_forward first");
    ST_TRY(check_unchanged(a));
    st_cloud *c = a->cloud;
    st_field *f = a->field;
    hipStream_t s = a->ctx->stream;
    const uint32_t n = a->n;
    // The first segment swept (the last in time) reads the velocity cotangent from the
    // seed and starts fbar and mbar at zero; the later ones carry vb, fbar and mbar on.
    // So each backward starts again from the seeds (one forward serves several losses),
    // with no copy or fill in front of it.
    st::Soa<double> cur = soa<double>(c, c->cur);
    st::AdjArgsHost h{};
    h.d = cur.d;
    h.rho = cur.rho;
    h.U = (const double *)f->U;
    h.G = a->G;
    h.UG = a->UG;
    ST_HIP(st::launch_interleave3(h.U, a->G_set ? a->G : nullptr, a->UG, (size_t)f->ncell, s));
    h.xb = a->xb_set ? a->xb : nullptr;
    h.vb = a->vb;
    h.fbar = a->fbar;
    h.mbar = a->mbar;
    h.g = f->geo;
    h.ph = make_phys<double>(c, f, a->dt_sub, a->p.cd);
    h.np = n;
    h.periodic = a->p.boundary == ST_PERIODIC;
    ST_HIP(timer_start(a->ctx, s));
    for (int32_t seg = a->nseg - 1; seg >= 0; --seg) {
        h.ck = a->ck + (size_t)seg * 6 * n;
        h.K = std::min(a->K, a->nsub - seg * a->K);
        h.first = seg == a->nseg - 1;
        h.vb_in = !h.first ? a->vb : a->vseed_set ? a->vseed : nullptr;
        ST_HIP(st::launch_adjoint_segment(h, s));
    }
    ST_HIP(timer_stop(a->ctx, s, a->nseg));
    // gd -> tmp[0, n), grho -> tmp[n, 2n), device order; dL/dcd -> scal[1]
    ST_HIP(st::launch_adjoint_grads(cur.d, cur.rho, a->fbar, a->mbar, a->tmp, cur.rho ? a->tmp + n : nullptr,
                                    a->scal + 1, a->scal + 2, h.ph, n, s));
    a->backward_done = true;
    return ST_OK;
}

int st_adjoint_gradients(st_adjoint *a, double *grad_d, double *grad_rho, double *grad_cd, double *grad_x0,
                         double *grad_v0) {
    ST_TRY(check(a));
    if (!a->backward_done) return fail(ST_E_STATE, "run st_adjoint_forward and st_adjoint_backward first");
    ST_TRY(check_unchanged(a));
    st_cloud *c = a->cloud;
    if (grad_rho && !c->has_rho) return fail(ST_E_INVALID, "grad_rho needs a cloud with per-parcel density");
    hipStream_t s = a->ctx->stream;
    const uint32_t n = a->n;
    const int32_t *pid = c->pid[c->cur];
    double *stage = a->tmp + 2 * (size_t)n;  // [2n, 3n) of the staging buffer
    struct Item {
        const double *src;
        double *dst;
    } items[8] = {{a->tmp, grad_d}, {a->tmp + n, grad_rho}};
    int ni = 2;
    if (grad_x0 && a->xb_set)
        for (int k = 0; k < 3; ++k) items[ni++] = {a->xb + (size_t)k * n, grad_x0 + (size_t)k * n};
    else if (grad_x0)  // dL/dx0 is the final-position cotangent, and none was given
        std::fill(grad_x0, grad_x0 + 3 * (size_t)n, 0.0);
    if (grad_v0)
        for (int k = 0; k < 3; ++k) items[ni++] = {a->vb + (size_t)k * n, grad_v0 + (size_t)k * n};
    // one staging array reused in stream order: each scatter follows the previous copy
    for (int i = 0; i < ni; ++i) {
        if (!items[i].dst || !n) continue;
        ST_HIP(st::launch_scatter_pid(items[i].src, pid, stage, n, s));
        ST_HIP(hipMemcpyAsync(items[i].dst, stage, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    }
    if (grad_cd) ST_HIP(hipMemcpyAsync(grad_cd, a->scal + 1, 8, hipMemcpyDeviceToHost, s));
    ST_HIP(hipStreamSynchronize(s));
    return ST_OK;
}

}  // extern "C"
