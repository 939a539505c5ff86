// st_noalloc.hip -- st_evolve for callers that cannot allocate device memory.
//
// st_evolve grows a cloud's per-workgroup statistics slots on demand (hipMalloc) and
// sorts a cloud that is not yet sorted for the mesh and brick (which allocates the
// sort buffers).  This variant does neither: the statistics slots, the reduced
// bounding box, the stray count and the re-sort gate live in one file-scope device
// array, `scratch`, of fixed size, and a cloud that would need a first sort is refused.
// Re-sorts of an already sorted cloud reuse the buffers its first sort allocated.
//
// Each GPU the library runs on holds its own copy of `scratch` until the process ends.
// Its address is looked up once per device, with that device current, when a context
// is created on it (st_ctx_create calls noalloc_setup); host code only ever uses that
// address, never the array's name.  Every call on one GPU shares that GPU's array and
// it keeps the last call's contents; nothing in it needs to start at zero (the last
// pass writes every slot the reduction reads, and the reduction writes the rest).
// Calls that overlap on one GPU (two contexts, or two host threads) race on it.
#include "st_handles.h"

#include <cmath>
#include <mutex>

namespace {

constexpr int kMaxDevices = 64;
constexpr size_t kHeaderBytes = 256;  // bbox (6 doubles), re-sort gate, stray count
constexpr int32_t kMaxSlots = 65536;  // statistics slots: bricks (sorted) or workgroups (plain)
constexpr size_t kScratchBytes = kHeaderBytes + (size_t)kMaxSlots * st::kStatSlot * sizeof(double);  // 4 MiB + 256 B

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};

unsigned char *device_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    return device >= 0 && device < kMaxDevices ? g_ptr[device] : nullptr;
}

}  // namespace

namespace st {
namespace host {

// Looks up this device's copy of `scratch`; the device must be current.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return ST_OK;  // the entry point refuses such devices
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return ST_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(ST_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_ptr[device] = (unsigned char *)p;
    return ST_OK;
}

}  // namespace host
}  // namespace st

using namespace st::host;

extern "C" int st_evolve_noalloc(st_ctx *ctx, st_cloud *c, st_field *f, double dt_fluid, int32_t nsub,
                                 const st_params *params, double bbox_out[6]) {
    st_params p;
    ST_TRY(prepare(ctx, c, f, dt_fluid, nsub, params, p));
    if (c->n == 0) {
        if (bbox_out)
            for (int k = 0; k < 6; ++k) bbox_out[k] = k < 3 ? INFINITY : -INFINITY;
        return ST_OK;
    }
    unsigned char *base = device_scratch(ctx->device);
    if (!base)
        return fail(ST_E_STATE, "no statistics scratch on device %d (device index above %d, or its lookup failed)",
                    ctx->device, kMaxDevices - 1);
    if (p.method == ST_METHOD_SORTED) {
        ST_TRY(check_tile(ctx, c->prec, p.brick));
        if (!sort_matches(c, f, p))
            return fail(ST_E_STATE,
                        "st_evolve_noalloc steps clouds already sorted for this mesh and brick; call "
                        "st_cloud_sort (which allocates the sort buffers) first");
    }
    const int32_t nslots = stats_slots(c, p);
    if (nslots > kMaxSlots)
        return fail(ST_E_RANGE,
                    "the step needs %d statistics slots (%s); st_evolve_noalloc holds at most %d: use larger "
                    "bricks or workgroups, or st_evolve",
                    nslots, p.method == ST_METHOD_SORTED ? "one per brick" : "one per workgroup", kMaxSlots);
    StatsBuf sb;
    sb.bbox = (double *)base;
    sb.gate = (unsigned int *)(base + 6 * sizeof(double));  // [0] re-sort gate, [1] stray count
    sb.part = (double *)(base + kHeaderBytes);
    sb.cap = kMaxSlots;
    sb.may_alloc = false;  // a re-sort runs only if the cloud's sort buffers already fit
    const double dt_sub = substep_dt(c->prec, dt_fluid, nsub);
    if (c->prec == ST_FP64)
        ST_TRY(run_passes<double>(ctx, c, f, dt_sub, nsub, p, false, sb, true));
    else
        ST_TRY(run_passes<float>(ctx, c, f, dt_sub, nsub, p, false, sb, true));
    if (bbox_out) {
        ST_HIP(hipMemcpyAsync(bbox_out, sb.bbox, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
    }
    return ST_OK;
}
