// st_handles.h -- host-side handle types and helpers shared by the C-ABI sources.
#pragma once

#include "../../include/scaletrack.h"
#include "st_internal.h"

#include <string>
#include <vector>

namespace st {
namespace host {
int fail(int code, const char *fmt, ...);
}  // namespace host
}  // namespace st

#define ST_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return ::st::host::fail(e_ == hipErrorOutOfMemory ? ST_E_NOMEM : ST_E_DEVICE,         \
                                    "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

#define ST_TRY(call)                \
    do {                            \
        int r_ = (call);            \
        if (r_ != ST_OK) return r_; \
    } while (0)

namespace st {
namespace host {
constexpr int kNF = 8;  // d, x, y, z, u, v, w, rho
}  // namespace host
}  // namespace st

// Event pairs around the dominant kernel launches of the last kTimerRing timed calls
// (st_evolve, batch and adjoint passes), so that timings of consecutive calls can be
// read after one synchronisation.
constexpr int kTimerRing = 64;

struct st_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t tev[kTimerRing][2] = {};
    int32_t tlaunches[kTimerRing] = {};  // launches between the pair of each slot
    int64_t tcount = 0;                  // timed calls so far (slot = index % kTimerRing)
    size_t max_lds = 65536;      // LDS bytes one workgroup may use on this device
    size_t lds_per_cu = 163840;  // LDS bytes of one CU (all resident workgroups together)
    int cus = 256;               // compute units of the device
    // Clouds, fields, graphs, batches and adjoints hold a reference: st_ctx_destroy only
    // marks the context closed while any of them is alive; the last release frees it.
    int refs = 0;
    bool closed = false;
};

struct st_cloud {
    st_ctx *ctx = nullptr;
    int prec = ST_FP64;
    int64_t n = 0;
    int64_t cap = 0;  // parcels each device array holds (n <= cap; grown by inject/reserve)
    size_t es = 8;
    double mu_c = 0, rho_d = 0, rdr = 0, rho_c = 0;
    int has_rho = 0;
    void *buf[2][st::host::kNF] = {};
    int32_t *pid[2] = {nullptr, nullptr};
    int cur = 0;
    bool uploaded = false;
    bool fresh = false;     // parcels from fresh_from on were not yet range-checked
    int64_t fresh_from = 0;
    bool permuted = false;  // device order differs from upload order
    // sort state
    uint32_t *keys = nullptr;
    size_t keys_cap = 0;
    uint32_t *counts = nullptr;
    uint32_t *cursor = nullptr;
    uint32_t *hist = nullptr;            // per-chunk brick histograms of the chunked sort
    size_t hist_cap = 0;                 // elements
    int32_t *brick_start = nullptr;
    int32_t bins_cap = 0;
    void *inj_buf = nullptr;             // keys, indices and radix-sort scratch of sort_injected
    size_t inj_cap = 0;                  // bytes
    void *stage = nullptr;               // pinned host staging of injected parcels
    size_t stage_cap = 0;                // bytes
    hipEvent_t stage_ev = nullptr;       // recorded behind the last copy out of `stage`
    bool sorted = false;
    int32_t sort_n[3] = {0, 0, 0}, sort_bd[3] = {0, 0, 0}, sort_periodic = -1;
    int32_t nb[3] = {0, 0, 0};
    int32_t nbricks = 0;
    double sort_o[3] = {0, 0, 0}, sort_e[3] = {0, 0, 0};
    // statistics of the last launch sequence
    double *bbox = nullptr;              // device, 6: min xyz, max xyz of the last call
    double *part = nullptr;              // device, per-workgroup statistics slots
    int32_t part_cap = 0;
    unsigned int *bad = nullptr;         // device: parcels outside the mesh (the range check's count)
    unsigned int *strays_host = nullptr; // pinned: [0] the last step's re-sort gate, [1] its strays,
                                         // written by the statistics reduction
    unsigned int *strays_host_dev = nullptr;  // the same words' device address
    hipEvent_t strays_ev = nullptr;
    bool strays_pending = false;
    int64_t strays_den = 0;  // parcels behind strays_host
    // Parcels [0, sorted_n) were in brick order after the last sort the host knows of;
    // parcels injected since then form an unsorted tail [brick_start[nbricks], n) that
    // each pass steps in workgroups after the bricks.  The device re-sorts fold the tail
    // in; the host learns of one when it reads the gate back, so sorted_n may lag (the
    // tail's workgroups read its start from the device and cover n - sorted_n parcels at
    // most).
    int64_t sorted_n = 0;
    double last_stray_fraction = -1.0;
    // device: [0] gate of the re-sort queued behind the last step and [1] that step's
    // stray count (both written by the statistics reduction, read back as one pair),
    // [2] re-sorts the device decided on so far, [3] unused
    unsigned int *resort = nullptr;
    // Graphs, batches and adjoints holding raw device pointers of this cloud: while any
    // is alive the cloud refuses to be destroyed, re-sorted with another brick geometry,
    // or to grow the buffers they captured.
    int pins = 0;
    // Bumped whenever the device order or state of the parcels may have changed (upload,
    // sort, step, queued re-sort): an adjoint checks it between its forward and backward.
    int64_t epoch = 0;
};

struct st_field {
    st_ctx *ctx = nullptr;
    st_mesh mesh{};
    st::Geo geo{};
    int prec = ST_FP64;
    size_t es = 8;
    int64_t ncell = 0;
    void *U = nullptr;
    double *src = nullptr;
    double *stray = nullptr;
    bool U_set = false;
    bool stray_dirty = false;
    // st_field_zero_source was called and nothing has written `src` since: the memory
    // still holds the old sums.  The next sorted owner-flush pass starts every brick's
    // sums from zero and stores every cell (materialize_zero writes the zeros for
    // every other path first).
    bool src_zero = false;
    int pins = 0;  // live graphs, batches and adjoints using this field
};

struct st_batch;

struct st_graph {
    st_ctx *ctx = nullptr;
    st_field *field = nullptr;
    st_cloud *cloud = nullptr;
    std::vector<st_field *> more_fields;  // batch graphs: every field the graph deposits into
    std::vector<st_cloud *> more_clouds;  // batch graphs: every cloud the graph steps
    st_batch *batch = nullptr;            // batch graphs: the batch whose device tables they use
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool owner_flush = false;
};


namespace st {
namespace host {

const char *last_error_text();
int check_ctx(st_ctx *ctx);
void retain_ctx(st_ctx *ctx);
// Drops a reference; frees a closed context with no references left.
void release_ctx(st_ctx *ctx);
// Makes the context's device current and waits for its stream (destroy paths).
void quiesce_ctx(st_ctx *ctx);
int alloc_set(st_cloud *c, int which);
// Orders parcels [n0, n0 + k) of a sorted cloud by cell of its sorted mesh (parcels just
// injected: they form its unsorted tail); queued on the context stream (st_inject.hip).
int sort_injected(st_cloud *c, int64_t n0, int64_t k);
// Queues the copy of k injected parcels from pinned, device-mapped host memory (nf
// arrays of k values: d x y z u v w [rho]) into the cloud at [n0, n0 + k), with their
// upload-order indices n0 .. n0 + k - 1 (st_inject.hip).
int copy_injected(st_cloud *c, const void *stage, int64_t n0, int64_t k);
// The kernels' geometry of a uniform mesh (origin, 1/delta, counts).
Geo make_geo(const st_mesh &m);
int validate_params(const st_params &p);
bool sort_matches(const st_cloud *c, const st_field *f, const st_params &p);
int sort_any(st_cloud *c, st_field *f, const st_params &p);
int prepare(st_ctx *ctx, st_cloud *c, st_field *f, double dt_fluid, int nsub, const st_params *pin, st_params &p);
// Where a step's statistics go: per-workgroup slots (cap of them), the reduced bounding
// box and stray count, and the gate of the device-side re-sort decision.
struct StatsBuf {
    double *part = nullptr;
    int32_t cap = 0;
    double *bbox = nullptr;
    unsigned int *gate = nullptr;    // gate[0]: re-sort gate, gate[1]: stray count
    bool may_alloc = true;  // the device re-sort may size the cloud's sort buffers first
};
// The cloud's own statistics buffers (slots grown to nslots on demand).
int cloud_stats(st_cloud *c, int32_t nslots, StatsBuf &sb);
// Statistics slots one step of the cloud uses with these params (sorted: one per brick,
// then one per workgroup of the unsorted tail; plain: one per workgroup).
int32_t stats_slots(const st_cloud *c, const st_params &p);
// Workgroups over the unsorted tail in a sorted step's launches (0: no tail).
int32_t tail_blocks(const st_cloud *c, const st_params &p);
// Whether the sort buffers hold the cloud's current parcel count and brick geometry.
bool sort_buffers_fit(const st_cloud *c);
// Queues nsub substeps of length dt_sub (exact in T) on the context stream, split in
// passes of p.substeps_per_pass.  With ck (sorted method, FP64) the first pass also
// writes the state each parcel starts with into ck (6 arrays of c->n, device order).
// With device_resort (sorted method) a re-sort of the cloud is queued behind the step,
// gated on the share of parcels found outside their brick exceeding p.resort_fraction.
template <class T>
int run_passes(st_ctx *ctx, st_cloud *c, st_field *f, double dt_sub, int nsub, const st_params &p, bool capturing,
               const StatsBuf &sb, bool device_resort, T *ck = nullptr, const st::Soa<T> *in = nullptr,
               const st::Soa<T> *out = nullptr);
// Looks up the device's copy of the no-allocation entry point's statistics scratch
// (st_noalloc.hip); called by st_ctx_create with the device current.
int noalloc_setup(int device);
// Pins and unpins a batch (graphs of it hold its device tables; st_batch_api.cpp).
void batch_pin(st_batch *b);
void batch_unpin(st_batch *b);
// Pins and unpins a cloud / field (see st_cloud::pins).
void pin(st_cloud *c);
void unpin(st_cloud *c);
void pin(st_field *f);
void unpin(st_field *f);
// The substep length the reference uses: Time stores dt as `scalar` (:272-275) and the
// kernel divides by nSteps (:1261).
inline double substep_dt(int prec, double dt_fluid, int nsub) {
    return prec == ST_FP64 ? dt_fluid / (double)nsub : (double)((float)dt_fluid / (float)nsub);
}
template <class T>
Phys<T> make_phys(const st_cloud *c, const st_field *f, double dt_sub, double cd);  // st_api.cpp
int merge_source(st_field *f);
// Writes a pending zero of the source (st_field::src_zero) to memory.
int materialize_zero(st_field *f);
// Timed region of one call on the context stream: start() records the start event of the
// next ring slot, stop() its end event and launch count and makes it the newest slot.
hipError_t timer_start(st_ctx *ctx, hipStream_t s);
hipError_t timer_stop(st_ctx *ctx, hipStream_t s, int launches);
// ST_E_INVALID unless the brick's halo tile fits in one workgroup's LDS.
int check_tile(const st_ctx *ctx, int prec, const int32_t brick[3]);
// Statistics slots for nslots workgroups (grown on demand).
int ensure_part(st_cloud *c, int32_t nslots);

template <class T>
inline Soa<T> soa(st_cloud *c, int which) {
    Soa<T> s;
    void **b = c->buf[which];
    s.d = (T *)b[0];
    s.x = (T *)b[1];
    s.y = (T *)b[2];
    s.z = (T *)b[3];
    s.u = (T *)b[4];
    s.v = (T *)b[5];
    s.w = (T *)b[6];
    s.rho = c->has_rho ? (const T *)b[7] : nullptr;
    return s;
}

}  // namespace host
}  // namespace st
