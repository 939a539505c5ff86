/*
 * scaletrack.h -- C-ABI of the MI355X (gfx950) Lagrangian parcel step.
 *
 * The library replaces the device side of SCALE-TRACK's tracker: the per-chunk
 * kernel launch `kernel(chunk, eulerian, mesh, executor; threads, blocks)`
 * (run/icoJuliaParcelFoam/cavity3D/asyncParallelTracking.jl:1098-1125, compiled at
 * :963-972), the kernel itself `evolve_on_device!`/`evolve_particle!` (:1257-1365),
 * the bounding-box epilogue (:749-797, :1245-1255), the UTrans `fill!` (:1064-1067)
 * and the U/UTrans `copyto!` calls (:1076-1083, :1148-1152), and the synchronous
 * driver of syncSerialTracking.jl:325-345.
 *
 * Conventions: every function returns ST_OK (0) or a negative st_status; nothing
 * aborts or throws across the ABI; st_last_error() gives the calling thread's last
 * message.  The caller owns host arrays; the library owns device buffers behind
 * opaque handles.  Work is queued on the context's HIP stream; functions that return
 * host data synchronise that stream.  A handle must not be used from two threads at
 * once.  Cell fields are AoS xyz per cell, cells numbered x fastest (the reference's
 * VectorField, asyncParallelTracking.jl:186-205, and OpenFOAM's blockMesh order).
 *
 * Lifetimes: a graph, batch or adjoint keeps raw device pointers of the clouds and
 * fields it was made from.  While one is alive those clouds and fields refuse (with
 * ST_E_STATE, before any device work) to be destroyed, to be sorted by another brick,
 * mesh or boundary form, or to grow a buffer the graph or batch captured; destroy the
 * graph, batch or adjoint first.  An adjoint's own forward may still sort the cloud for
 * the adjoint's parameters when no other graph, batch or adjoint holds it.  A batch
 * refuses to be destroyed while a graph of it is alive.
 */
#ifndef SCALETRACK_H
#define SCALETRACK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ST_ABI_VERSION 2

typedef enum st_status {
    ST_OK = 0,
    ST_E_INVALID = -1,     /* bad argument: null pointer, size, enum, non-finite value */
    ST_E_DEVICE = -2,      /* HIP runtime error (no device, launch failure, ...) */
    ST_E_NOMEM = -3,       /* device or host allocation failed */
    ST_E_STATE = -4,       /* call out of order (e.g. evolve before upload) */
    ST_E_UNSUPPORTED = -5, /* valid request this build does not provide */
    ST_E_RANGE = -6        /* parcel outside the mesh, or size above a limit */
} st_status;

/* Precision forms.  ST_FP64: parcel state, fluid velocity, drag, face test and source
 * all in double (the reference with its `scalar` alias, :172, set to Float64).
 * ST_FP32: parcel state, fluid velocity and drag in float; the face-crossing test
 * (x - origin) * rDelta and the momentum source in double. */
typedef enum st_precision { ST_FP64 = 0, ST_FP32 = 1 } st_precision;

/* ST_WALL: the reference's specular bounce (:1306-1335) on every face.  ST_PERIODIC:
 * extension, every axis periodic.  ST_CHANNEL: extension, walls on y and z while the x
 * planes at origin and ending are open -- a substep that ends outside them (tested
 * every substep, before the bounce test) ends the parcel: its residence deposit
 * m (s0 - v') goes into the cell it was in, its state is stored as advanced, and its
 * diameter is stored negated, which marks it removed: later steps leave it untouched,
 * downloads return it with d < 0, st_cloud_compact drops it.  The reference has no
 * open boundary; its chunks keep every parcel (asyncParallelTracking.jl:329-364). */
typedef enum st_boundary { ST_WALL = 0, ST_PERIODIC = 1, ST_CHANNEL = 2 } st_boundary;

/* ST_METHOD_PLAIN: reference-shaped -- one thread per parcel in the caller's order,
 * cell velocity gathered from HBM, three FP64 atomics per deposit.
 * ST_METHOD_SORTED: parcels kept sorted by mesh brick; one workgroup per brick stages
 * the brick's cell velocities and sources in LDS and writes the source once. */
typedef enum st_method { ST_METHOD_PLAIN = 0, ST_METHOD_SORTED = 1 } st_method;

/* ST_FLUSH_OWNER: a brick's source is added with plain read-modify-write by the one
 * workgroup owning it; deposits from parcels that drifted out of their brick go to a
 * second FP64 array with atomics, merged when the source is read.
 * ST_FLUSH_ATOMIC: every brick adds its LDS sums with FP64 atomics into one array. */
typedef enum st_flush { ST_FLUSH_OWNER = 0, ST_FLUSH_ATOMIC = 1 } st_flush;

typedef struct st_ctx st_ctx;
typedef struct st_cloud st_cloud;
typedef struct st_field st_field;
typedef struct st_graph st_graph;

/* Uniform hexahedral box (the reference's Mesh, :222-270, decomposition (1,1,1)). */
typedef struct st_mesh {
    int32_t n[3];
    double origin[3];
    double ending[3];
} st_mesh;

/* Per-call options.  st_params_default() fills the defaults noted here. */
typedef struct st_params {
    int32_t boundary;          /* st_boundary, default ST_WALL */
    int32_t method;            /* st_method, default ST_METHOD_SORTED */
    int32_t flush;             /* st_flush, default ST_FLUSH_OWNER */
    int32_t substeps_per_pass; /* substeps fused in registers per launch; 0 = all nsub */
    double cd;                 /* drag-coefficient multiplier, default 1 (reference drag) */
    int32_t brick[3];          /* cells per brick (sorted method); default 32,8,2 */
    int32_t block;             /* threads per workgroup; default 512 */
    double resort_fraction;    /* sorted method: after each step the device re-sorts the
                                  cloud when the share of parcels outside their brick in the
                                  step's last pass exceeds this (decided on the device, also
                                  inside captured graphs); default 0.08 */
    int32_t passes_per_launch; /* sorted method: passes run by one persistent launch, each
                                  workgroup owning the same bricks in every pass; 1 = one
                                  launch per pass (default) */
} st_params;

/* ---- library ---------------------------------------------------------------- */
int st_version(void);                     /* ST_ABI_VERSION */
const char *st_last_error(void);          /* thread-local; "" when none */
void st_params_default(st_params *p);
int st_device_count(int32_t *count);

/* ---- context: one device, one stream --------------------------------------- */
int st_ctx_create(int32_t device, st_ctx **out);
int st_ctx_destroy(st_ctx *ctx);
int st_ctx_synchronize(st_ctx *ctx);

/* ---- parcel cloud: the reference's Chunk (:329-364) ------------------------- */
/* n parcels; mu_c, rho_d and rho_d_by_rho_c are the chunk scalars of
 * allocate_chunk(executor, n, mu_c, rho_d, rho_d/rho_c).  per_parcel_rho != 0 adds a
 * per-parcel density array (extension) with rho_c used for the density ratio. */
int st_cloud_create(st_ctx *ctx, int32_t precision, int64_t n, double mu_c, double rho_d,
                    double rho_d_by_rho_c, double rho_c, int32_t per_parcel_rho,
                    st_cloud **out);
int st_cloud_destroy(st_cloud *cloud);
/* Host SoA in the cloud's precision (double or float).  rho may be NULL unless the
 * cloud was created with per_parcel_rho. */
int st_cloud_upload(st_cloud *cloud, const void *d, const void *x, const void *y,
                    const void *z, const void *u, const void *v, const void *w,
                    const void *rho);
/* Returns parcels in the order they were uploaded, whatever the device order. */
int st_cloud_download(st_cloud *cloud, void *d, void *x, void *y, void *z, void *u,
                      void *v, void *w);
int st_cloud_count(const st_cloud *cloud, int64_t *n);
/* Extension (injection and removal; the reference's chunks keep a fixed N,
 * asyncParallelTracking.jl:329-364).  st_cloud_reserve grows the device arrays to hold
 * `capacity` parcels.  st_cloud_inject appends k parcels (host SoA in the cloud's
 * precision; rho as in st_cloud_upload) after the current ones: they take the next
 * upload-order indices, so st_cloud_count grows by k and downloads return them last.
 * A sorted cloud stays sorted: the injected parcels are checked against its mesh at
 * once (ST_E_RANGE, and nothing is appended, when a cell index
 * floor((x - origin) / delta) lies outside [0, N); an unsorted cloud checks them at its
 * next step), ordered by cell on the device (queued on the context stream), and sorted
 * steps advance them as an unsorted tail (workgroups after the bricks in each pass's
 * launch, one thread per parcel) until a re-sort the device decides on
 * (params->resort_fraction) takes them into the bricks.
 * st_cloud_compact drops the parcels an open boundary removed (ST_CHANNEL; d < 0),
 * keeping the others in their upload order with their indices renumbered from 0, and
 * returns their number (may be NULL); the first sorted step after it sorts the cloud.
 * All three return once the host arrays are no longer needed, and refuse (ST_E_STATE)
 * while a graph, batch or adjoint holds the cloud. */
int st_cloud_reserve(st_cloud *cloud, int64_t capacity);
int st_cloud_inject(st_cloud *cloud, int64_t k, const void *d, const void *x, const void *y,
                    const void *z, const void *u, const void *v, const void *w, const void *rho);
int st_cloud_compact(st_cloud *cloud, int64_t *n_live);

/* ---- Eulerian fields: the reference's TwoWayEulerian (:207-220) ------------- */
int st_field_create(st_ctx *ctx, const st_mesh *mesh, int32_t precision, st_field **out);
int st_field_destroy(st_field *field);
/* U: host AoS xyz, ncell*3 elements of elem_bytes (4 or 8) each; converted to the
 * field's precision. */
int st_field_upload_U(st_field *field, const void *U, int32_t elem_bytes);
/* UTrans = 0 (asyncParallelTracking.jl:1064-1067, syncSerialTracking.jl:553).  The
 * zeros are stored by the next step's first sorted owner-flush pass, which writes every
 * cell; every other call that uses the source stores them first. */
int st_field_zero_source(st_field *field);
/* Completes UTrans on the device: stores a pending zero and adds the sums of deposits
 * made outside their brick (ST_FLUSH_OWNER) into it, queued on the context stream,
 * without a host copy.  For a caller that reads the source from device memory;
 * st_field_download_source does it itself. */
int st_field_merge_source(st_field *field);
/* UTrans: host AoS xyz, ncell*3 elements of elem_bytes (8 = the FP64 sums, 4 =
 * rounded to float for a single-precision caller). */
int st_field_download_source(st_field *field, void *out, int32_t elem_bytes);

/* ---- the parcel step -------------------------------------------------------- */
/* Advances every parcel of `cloud` by nsub substeps of dt_fluid / nsub (the
 * reference's nSteps = 10 and time[1].Delta_t / nSteps, :1260-1261) through the
 * frozen cell velocity of `field`, adding the momentum source
 * sum (m_d / rho_c) (v_in - v_out) per cell into the field's UTrans.  Queued on the
 * context stream.  bbox_out (may be NULL) receives min xyz, max xyz of the new
 * positions, and then the call synchronises.  params may be NULL (defaults). */
int st_evolve(st_ctx *ctx, st_cloud *cloud, st_field *field, double dt_fluid, int32_t nsub,
              const st_params *params, double bbox_out[6]);

/* No counterpart in the reference.  st_evolve for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same step, with its statistics slots, bounding box, stray count and
 * re-sort gate in one fixed library-owned device array of 4 MiB + 256 B per GPU instead
 * of buffers grown on demand.  Limits: at most 65536 statistics slots (one per brick
 * with the sorted method, one per workgroup with the plain one; ST_E_RANGE above), and
 * with the sorted method the cloud must already be sorted for this mesh and brick
 * (st_cloud_sort or an earlier st_evolve; ST_E_STATE otherwise).  Every call on one GPU
 * shares that array and it keeps the last call's contents: calls that overlap on one
 * GPU (two contexts or threads) race. */
int st_evolve_noalloc(st_ctx *ctx, st_cloud *cloud, st_field *field, double dt_fluid, int32_t nsub,
                      const st_params *params, double bbox_out[6]);

/* Share of parcels that were outside their sorted brick during the last sorted pass
 * (synchronises); -1 if unknown. */
int st_cloud_stray_fraction(st_cloud *cloud, double *fraction);
/* Re-sorts the device has decided on for this cloud so far (params->resort_fraction);
 * synchronises.  No counterpart in the reference, whose Hilbert-ordered chunks keep
 * their initial order (asyncParallelTracking.jl:559-638). */
int st_cloud_resort_count(st_cloud *cloud, int64_t *count);
/* Forces a sort of the cloud by brick of `field`'s mesh before the next step. */
int st_cloud_sort(st_ctx *ctx, st_cloud *cloud, st_field *field, const st_params *params);

/* ---- HIP graph of one fluid step -------------------------------------------- */
/* Captures zeroing UTrans plus the nsub substeps as one hipGraph (sorted method: the
 * cloud is sorted first, outside the graph; the device-decided re-sort behind the step
 * is part of the graph). */
int st_graph_capture(st_ctx *ctx, st_cloud *cloud, st_field *field, double dt_fluid,
                     int32_t nsub, const st_params *params, int32_t zero_source,
                     st_graph **out);
int st_graph_launch(st_graph *graph);
int st_graph_destroy(st_graph *graph);

/* ---- timing helper ---------------------------------------------------------- */
/* Average device time (ms) of the last st_evolve's dominant kernel launches, from HIP
 * events recorded on the context stream around them; -1 if not recorded. */
int st_last_kernel_ms(st_ctx *ctx, double *ms, int32_t *launches);

/* ---- batches of clouds ---------------------------------------------------------- */
/* Extension: n independent clouds, each with its own field and mesh, advanced together
 * by one persistent launch per pass.  The reference runs one kernel launch per chunk
 * (asyncParallelTracking.jl:1098-1125); a batch replaces those launches.  All clouds
 * share the precision and params (boundary, flush, block, substeps_per_pass, cd).
 * target_parcels_per_brick > 0 picks each cloud's brick so that a brick holds about
 * that many parcels (0: params->brick for all).  blocks_per_cu > 0 caps the persistent
 * grid at that many workgroups per CU and, with target_parcels_per_brick, also bounds
 * every cloud's brick so that its tile fits that share of a CU's LDS (0: as many as
 * fit, bricks bounded by one workgroup's LDS only).  Clouds are sorted at creation;
 * st_batch_sort re-sorts. */
typedef struct st_batch st_batch;
int st_batch_create(st_ctx *ctx, int32_t n, st_cloud *const *clouds, st_field *const *fields,
                    const st_params *params, int32_t target_parcels_per_brick,
                    int32_t blocks_per_cu, st_batch **out);
int st_batch_destroy(st_batch *batch);
int st_batch_sort(st_batch *batch);
/* nsub substeps of dt_fluid / nsub for every cloud; sources accumulate per field. */
int st_batch_evolve(st_batch *batch, double dt_fluid, int32_t nsub);
/* One fluid step of the whole batch as a hipGraph (optionally zeroing every UTrans). */
int st_batch_graph_capture(st_batch *batch, double dt_fluid, int32_t nsub, int32_t zero_source,
                           st_graph **out);
/* Work items, persistent workgroups per pass (summed over the launches of clouds with
 * and without a per-parcel density), and the brick of each cloud (3*n, caller order);
 * any output may be NULL. */
int st_batch_info(const st_batch *batch, int32_t *total_items, int32_t *grid,
                  int32_t *brick_of_cloud);

/* Extension: autotuner of a batch.  Times every candidate (parcels per brick in
 * {1024 .. 16384}, block in {256, 512, 1024}, persistent workgroups per CU capped at
 * {all, 4, 3, 2}) as production runs it: one fluid step of 10 substeps (passes of
 * base->substeps_per_pass) captured as a hipGraph, one warm-up replay, then
 * max(trials, 5) timed replays; the candidate's figure is the median replay per pass.
 * The clouds and sources are restored afterwards.  The result is cached by device and
 * workload in cache_path (JSON lines; NULL or "" disables the cache). */
typedef struct st_tune_result {
    int32_t target_parcels_per_brick;
    int32_t block;
    int32_t blocks_per_cu;
    int32_t from_cache;  /* 1 if read from cache_path */
    double ms_per_pass;            /* the winner's median replay per pass */
    double runner_up_ms_per_pass;  /* the second-best candidate's */
} st_tune_result;
int st_autotune_batch(st_ctx *ctx, int32_t n, st_cloud *const *clouds, st_field *const *fields,
                      const st_params *base, double dt_fluid, int32_t trials, const char *cache_path,
                      st_tune_result *out);

/* ---- adjoint (FP64) -------------------------------------------------------------- */
/* Extension: reverse-mode gradients of a loss of the deposited source (and optionally of
 * the final parcel state) with respect to the per-parcel diameter, the per-parcel density
 * and the drag-coefficient multiplier, through nsub substeps run as passes of
 * checkpoint_every substeps with (x, v) saved before each pass and recomputed in the
 * reverse sweep.  The derivative is the almost-everywhere one: cells, bounces and wraps
 * are held at the forward run's.  Sequence: create, forward, seed (source misfit or
 * cotangents), backward, gradients; after one forward, seeding, backward and gradients
 * may repeat (each backward starts from the seeds last given, which stay until the next
 * forward).  Host arrays are in upload order. */
typedef struct st_adjoint st_adjoint;
int st_adjoint_create(st_ctx *ctx, st_cloud *cloud, st_field *field, int32_t nsub,
                      int32_t checkpoint_every, const st_params *params, st_adjoint **out);
int st_adjoint_destroy(st_adjoint *adj);
/* Zeroes UTrans and runs the forward step (substep = dt_fluid / nsub). */
int st_adjoint_forward(st_adjoint *adj, double dt_fluid);
/* loss = sum over cells and components of (UTrans - target)^2; seeds dL/dUTrans. */
int st_adjoint_source_misfit(st_adjoint *adj, const double *target, double *loss);
/* Seeds dL/dUTrans directly (AoS xyz, ncell*3). */
int st_adjoint_set_source_cotangent(st_adjoint *adj, const double *dl_dsource);
/* Seeds dL/d(final x, y, z) and dL/d(final u, v, w), SoA 3*n each; either may be NULL. */
int st_adjoint_set_state_cotangent(st_adjoint *adj, const double *xbar, const double *vbar);
int st_adjoint_backward(st_adjoint *adj);
/* Any output may be NULL: grad_d, grad_rho (n each; rho needs per-parcel density),
 * grad_cd (scalar), grad_x0 and grad_v0 (SoA 3*n, gradients of the initial state). */
int st_adjoint_gradients(st_adjoint *adj, double *grad_d, double *grad_rho, double *grad_cd,
                         double *grad_x0, double *grad_v0);

/* ---- host memory -------------------------------------------------------------- */
/* No counterpart in the reference, whose host arrays are the host language's own.
 * Page-locked host memory, which every entry point taking host arrays copies to and from
 * by DMA at the bus's rate (pageable arrays are staged through the runtime's bounce
 * buffers at a fraction of it).  Allocating pins the pages and costs more than one
 * transfer of the same size: allocate once and reuse.  st_host_alloc returns
 * ST_E_INVALID for bytes < 1 or a null out, ST_E_NOMEM when the memory cannot be
 * allocated and pinned, ST_E_DEVICE without a usable device; st_host_free(NULL) does
 * nothing. */
int st_host_alloc(int64_t bytes, void **out);
int st_host_free(void *ptr);

/* ---- diagnostics -------------------------------------------------------------- */
/* No counterpart in the reference.  Streams `bytes` from one device buffer to another
 * `reps` times (16-byte loads and stores) and returns read + write GB/s: the measured
 * copy bandwidth the parcel step's HBM figures are compared with, and the known-size
 * kernel that hardware traffic counters are calibrated on. */
int st_diag_copy_bandwidth(st_ctx *ctx, int64_t bytes, int32_t reps, double *gbs);
/* Per-launch device time (ms) of the dominant kernels of each of the last timed calls
 * (st_evolve, st_batch_evolve, st_adjoint_backward, st_diag_copy_bandwidth), oldest
 * first: up to min(max, 64) calls, their count in *count, launches per call in
 * launches[].  Lets a caller time consecutive calls with one synchronisation. */
int st_diag_kernel_times(st_ctx *ctx, int32_t max, double *ms, int32_t *launches, int32_t *count);

#ifdef __cplusplus
}
#endif
#endif
