// st_internal.h -- device-side data structures of the gfx950 parcel step.
//
// Layout in HBM (all arrays are separate allocations, 256-byte aligned by hipMalloc):
//   parcel state  SoA: d, x, y, z, u, v, w [, rho], element T (float or double)
//   pid           int32 original index of each device slot (un-permutes downloads)
//   U             AoS xyz per cell, element T, cells x fastest
//   src, stray    AoS xyz per cell, double
//   brick_start   int32[nbricks + 1] CSR offsets of the brick-sorted parcel array
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace st {

constexpr double kPi = 3.14159265358979323846;

// Mesh as the kernels see it: the face test (x - o) * rd runs in double in both
// precision forms.  rd = 1 / (L / N) is computed once on the host in double, as the
// reference's construct_mesh does in its scalar type (asyncParallelTracking.jl:239-241).
struct Geo {
    int32_t n[3];
    uint32_t nxny;
    double o[3];
    double rd[3];
};

// Parcel-phase constants in the state precision T.
template <class T>
struct Phys {
    T dt;      // substep = (T)dt_fluid / (T)nsub
    T mu;      // carrier viscosity
    T rho;     // chunk parcel density
    T cd;      // drag-coefficient multiplier
    T L[3];    // periods (periodic boundary)
    double rdr;    // chunk rho_d / rho_c
    double rho_c;  // carrier density for per-parcel rho
};

template <class T>
struct Soa {
    T *d, *x, *y, *z, *u, *v, *w;
    const T *rho;  // nullptr unless per-parcel density
};

// Per-workgroup statistics of a pass, reduced by a separate small kernel (no
// same-address atomics): slot k holds min xyz, max xyz of the new positions and the
// number of parcels found outside their brick.
constexpr int kStatSlot = 8;  // doubles per slot

// Where one cloud's per-slot statistics are reduced to.  With `resort` set, the
// reduction also writes resort[0] = 1 when more than `threshold` parcels were outside
// their brick (else 0): the gate of the re-sort queued behind the pass, so the
// decision is taken on the device without a host round trip (a threshold of
// 0xffffffff writes 0: no re-sort).
struct StatsTarget {
    const double *part;  // nslots * kStatSlot
    int32_t nslots;
    double *bbox;        // 6: min xyz, max xyz
    unsigned int *strays;
    unsigned int *resort;  // or nullptr
    uint32_t threshold;
    unsigned int *host;    // pinned host words [0] resort[0], [1] strays[0], or nullptr
};

template <class T>
struct PassArgs {
    Soa<T> s;
    const T *U;
    double *src;    // owner sums (or everything, ST_FLUSH_ATOMIC)
    double *stray;  // deposits outside the depositing parcel's brick (ST_FLUSH_OWNER)
    const int32_t *brick_start;
    Geo g;
    Phys<T> ph;
    double *part;    // per-workgroup statistics slots, or nullptr (no statistics this pass)
    Soa<T> out;      // where the checkpointing pass stores the new state (x..w; may be s itself);
                     // every other pass stores into s
    T *ck;           // checkpoint of the state each parcel starts the pass with (6 arrays of np,
                     // device order), or nullptr; sorted method, FP64 only
    int32_t nbricks;
    int32_t nb[3];   // bricks per direction
    int32_t bd[3];   // cells per brick per direction
    int32_t tile_cells;  // (bd[0]+2)*(bd[1]+2)*(bd[2]+2): a brick with a halo of one cell
    int32_t nsteps;
    int32_t flush;   // 0 owner, 1 atomic
    int32_t zero_src;  // owner flush: `src` counts as zero (not read); every own cell is stored
    uint32_t np;
    // steps over a sorted cloud's unsorted tail: the first parcel is read from
    // here (the brick ranges' end, brick_start[nbricks]); nullptr otherwise
    const int32_t *tail0;
};

// Order-preserving map of a double onto uint64 (for atomicMin/atomicMax).
__host__ __device__ inline unsigned long long order_key(double x) {
    union { double d; unsigned long long u; } c;
    c.d = x;
    return (c.u >> 63) ? ~c.u : (c.u | 0x8000000000000000ull);
}
__host__ __device__ inline double order_key_decode(unsigned long long k) {
    union { double d; unsigned long long u; } c;
    c.u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return c.d;
}

// Static LDS the pass kernels declare besides the tile, rounded up: the statistics
// reduction's 7 x 16 doubles (896 B) and the batch kernel's item slot.
constexpr size_t kPassStaticLds = 1024;

// LDS bytes of the halo tile of a brick (velocities in T, sums in double).
template <class T>
inline size_t tile_lds_bytes(const int32_t bd[3]) {
    return (size_t)(bd[0] + 2) * (bd[1] + 2) * (bd[2] + 2) * 3 * (sizeof(double) + sizeof(T));
}

}  // namespace st

// Host-side launchers (st_step.hip).
namespace st {
// bnd: the st_boundary form (wall 0, periodic 1, channel 2).
template <class T>
hipError_t launch_pass_sorted(const PassArgs<T> &a, int bnd, int block, hipStream_t s);
template <class T>
hipError_t launch_pass_plain(const PassArgs<T> &a, int bnd, int block, hipStream_t s, uint32_t grid = 0);
// A sorted pass and the pass over the cloud's unsorted tail (a.tail0 set, tail_blocks
// workgroups of `block` parcels) in one launch; a.part holds nbricks + tail_blocks slots.
template <class T>
hipError_t launch_pass_sorted_tail(const PassArgs<T> &a, uint32_t tail_blocks, int bnd, int block, hipStream_t s);
// npass passes of a.nsteps substeps each in one persistent launch of `grid` workgroups
// (each owning the same bricks in every pass); a.part is filled in the last pass.
template <class T>
hipError_t launch_pass_multi(const PassArgs<T> &a, int32_t npass, int bnd, int block, int grid, hipStream_t s);
template <class T>
int pass_multi_blocks_per_cu(int bnd, int rho, int block, size_t lds);
// Persistent pass over clouds [0, ncloud) of descs: work items item0[0] .. item0[ncloud]-1
// (counter must hold item0[0] at launch); rho: every one of these clouds carries a
// per-parcel density array (else none does).
template <class T>
// zero_src: every cloud's source counts as zero in this pass (PassArgs::zero_src).
hipError_t launch_pass_batch(const PassArgs<T> *descs, const int32_t *item0, int32_t ncloud, int32_t nsteps,
                             int stats, unsigned int *counter, int periodic, int rho, int block, int grid,
                             size_t lds, int zero_src, hipStream_t s);
template <class T>
int batch_max_blocks_per_cu(int periodic, int rho, int block, size_t lds);
hipError_t launch_reduce_stats(const StatsTarget &t, hipStream_t s);
hipError_t launch_reduce_stats_many(const StatsTarget *targets, int32_t n, hipStream_t s);
// Every kernel of the sort takes a gate: nullptr runs it; otherwise it does nothing
// unless gate[0] != 0 (the device-side re-sort decision of StatsTarget::resort).
template <class T>
hipError_t launch_brick_keys(const Soa<T> &s, uint32_t np, Geo g, const int32_t nb[3],
                             const int32_t bd[3], int periodic, uint32_t *keys,
                             uint32_t *counts, const unsigned int *gate, hipStream_t st);
// Counting sort by brick in per-workgroup chunks (nbins * 4 B of LDS per workgroup):
// keys, per-chunk histograms (nrows x nbins in hist), brick starts and the scatter.
template <class T>
hipError_t launch_sort_chunked(const Soa<T> &in, const int32_t *pid_in, Soa<T> out, int32_t *pid_out, uint32_t np,
                               Geo g, const int32_t nb[3], const int32_t bd[3], int periodic, int32_t nrows,
                               uint32_t *keys, uint32_t *hist, uint32_t *counts, int32_t *start,
                               uint32_t *cursor, const unsigned int *gate, hipStream_t s);
hipError_t launch_scan(const uint32_t *counts, int32_t nbins, int32_t *start, uint32_t *cursor,
                       const unsigned int *gate, hipStream_t s);
template <class T>
hipError_t launch_scatter(const Soa<T> &in, const int32_t *pid_in, Soa<T> out, int32_t *pid_out,
                          const uint32_t *keys, uint32_t *cursor, uint32_t np, const unsigned int *gate,
                          hipStream_t s);
// Copies the sorted set back over the cloud's own arrays (so device pointers never
// change), gated like the sort; a gated copy that runs adds one to *count.
template <class T>
hipError_t launch_copy_back(const Soa<T> &from, const int32_t *pid_from, Soa<T> to, int32_t *pid_to, uint32_t np,
                            const unsigned int *gate, unsigned int *count, hipStream_t s);
template <class T>
hipError_t launch_unpermute(const Soa<T> &in, const int32_t *pid, Soa<T> out, uint32_t np,
                            hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s, const unsigned int *gate = nullptr);
hipError_t launch_fill_f64(double *p, double value, size_t n, hipStream_t s);
// Zeroes count arrays of doubles in one launch (graph-capturable).  chunk0[i] is the
// first chunk of item i in chunks of fill_many_chunks(1) elements; chunk0[count] = chunks.
struct FillItem {
    double *p;
    size_t n;
};
uint32_t fill_many_chunks(size_t n);
hipError_t launch_fill_many(const FillItem *items, const uint32_t *chunk0, int32_t count, uint32_t chunks,
                            hipStream_t s);
hipError_t launch_iota(int32_t *p, uint32_t n, hipStream_t s);
hipError_t launch_iota_from(int32_t *p, uint32_t n, int32_t base, hipStream_t s);
// Tiles of the exclusive scan over n elements (entries of its tile-sum array).
uint32_t scan_tiles(uint32_t n);
// Drops removed parcels (d <= 0) of `in` into `out`, keeping the device order of the
// rest and renumbering their upload-order indices among themselves.  offsets_dev
// (scan_tiles(np) + 1 entries) ends with the number of parcels kept.  Scratch:
// live_dev, live_up, pos_dev, new_up (np each), tile_sums (scan_tiles(np)),
// offsets_up (scan_tiles(np) + 1), cursor (scan_tiles(np)).
template <class T>
hipError_t launch_compact(const Soa<T> &in, const int32_t *pid_in, Soa<T> out, int32_t *pid_out, uint32_t np,
                          uint32_t *live_dev, uint32_t *live_up, uint32_t *pos_dev, uint32_t *new_up,
                          uint32_t *tile_sums, int32_t *offsets_dev, int32_t *offsets_up, uint32_t *cursor,
                          hipStream_t s);
hipError_t launch_merge(double *src, double *stray, size_t n, hipStream_t s);
hipError_t launch_copy(const void *a, void *b, size_t bytes, hipStream_t s);

// Adjoint (st_adjoint.hip), FP64.
struct AdjArgsHost {
    const double *d, *rho, *ck, *U, *G, *UG;
    const double *xb;     // the final-position cotangent, or nullptr for zero
    const double *vb_in;  // the velocity cotangent the segment starts from (vseed or vb), or
                          // nullptr for zero
    double *vb, *fbar, *mbar;
    bool first;           // fbar and mbar start at zero (not read)
    Geo g;
    Phys<double> ph;
    int32_t K;
    uint32_t np;
    int periodic;
};
hipError_t launch_adjoint_segment(const AdjArgsHost &h, hipStream_t s);
// dst[6c .. 6c+5] = a[3c .. 3c+2], b[3c .. 3c+2] for c < n (b == nullptr: zeros).
hipError_t launch_interleave3(const double *a, const double *b, double *dst, size_t n, hipStream_t s);
// The two sums below (dL/dcd, the loss) are taken without atomics: at most kSumBlocks
// workgroups each write a partial to part[], and one workgroup adds them in a fixed
// order into the output (so they repeat bit for bit from run to run).
constexpr int kSumBlocks = 1024;
hipError_t launch_adjoint_grads(const double *d, const double *rho, const double *fbar, const double *mbar,
                                double *gd, double *grho, double *gcd, double *part, const Phys<double> &ph,
                                uint32_t np, hipStream_t s);
hipError_t launch_misfit(const double *S, const double *T, double *G, double *loss, double *part, size_t n,
                         hipStream_t s);
hipError_t launch_save6(const double *const f[6], double *dst, uint32_t n, hipStream_t s);
hipError_t launch_gather_pid(const double *src, const int32_t *pid, double *dst, uint32_t n, int k,
                             hipStream_t s);
hipError_t launch_scatter_pid(const double *src, const int32_t *pid, double *dst, uint32_t n, hipStream_t s);
template <class T>
hipError_t launch_range_check(const Soa<T> &s, uint32_t p0, uint32_t np, Geo g, unsigned int *bad,
                              hipStream_t st);
}  // namespace st
