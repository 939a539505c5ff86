// st_step.hip -- gfx950 kernels of the Lagrangian parcel step.
//
// Physics and event order follow evolve_particle! (run/icoJuliaParcelFoam/cavity3D/
// asyncParallelTracking.jl:1271-1365): cell-constant carrier velocity, implicit-Euler
// Stokes drag, relocation, specular wall bounce gated on the folded cell index
// changing, and a momentum deposit into the cell each residence started in.  The
// whole library is compiled with -ffp-contract=off, so each expression below rounds
// exactly as the reference's left-to-right Julia arithmetic does.
//
// Two launch shapes:
//  * k_pass_plain  -- one thread per parcel in device order, velocity gathered from
//    HBM, three FP64 atomics per deposit (the reference's shape, :1257-1269, :710-728).
//  * k_pass_sorted -- parcels sorted by mesh brick; one workgroup per brick stages the
//    brick's cell velocities in LDS, sums its deposits in LDS (FP64 ds_add), and adds
//    them to HBM once per cell.  Parcels that drifted out of their brick fall back to
//    HBM gathers and FP64 atomics into a separate array.
#include "st_device.h"

#include <type_traits>

namespace st {

// Calls f(std::integral_constant<int, BND>, std::bool_constant<RHO>) for the runtime
// boundary form (kWall, kPeriodic, kChannel) and density form: the launchers pick the
// kernel instance with it.
template <class F>
inline void with_forms(int bnd, bool rho, F &&f) {
    using W = std::integral_constant<int, kWall>;
    using P = std::integral_constant<int, kPeriodic>;
    using C = std::integral_constant<int, kChannel>;
    if (bnd == kPeriodic) {
        if (rho) f(P{}, std::true_type{}); else f(P{}, std::false_type{});
    } else if (bnd == kChannel) {
        if (rho) f(C{}, std::true_type{}); else f(C{}, std::false_type{});
    } else {
        if (rho) f(W{}, std::true_type{}); else f(W{}, std::false_type{});
    }
}

// Everything in HBM: the plain method.
//
// An environment hands out a Slot per residence (where the cell's velocity and sums
// live), so a parcel's deposit into the cell it leaves reuses what was worked out when
// it entered that cell.
template <class T>
struct GlobalEnv {
    const T *__restrict__ U;
    double *__restrict__ src;
    struct Slot {};
    __device__ __forceinline__ void enter(const Cell &) {}
    __device__ __forceinline__ Slot slot(const Cell &) const { return Slot{}; }
    __device__ __forceinline__ void load_u(const Cell &cl, const Slot &, T &ux, T &uy, T &uz) const {
        const T *q = U + 3u * cl.c;
        ux = q[0];
        uy = q[1];
        uz = q[2];
    }
    __device__ __forceinline__ void deposit(const Cell &cl, const Slot &, double a, double b, double c) {
        double *q = src + 3u * cl.c;
        atomic_add_global(q, a);
        atomic_add_global(q + 1, b);
        atomic_add_global(q + 2, c);
    }
};

// LDS-qualified pointers: reads through them are ds_read, never flat loads (a flat
// load of the tile would share the vector-memory counter with the HBM loads in flight).
#define ST_LDS __attribute__((address_space(3)))

// A brick staged in LDS with a halo of one cell on every side, so parcels that drifted
// into a neighbouring brick still read their cell velocity and sum their deposits in
// LDS.  Only parcels further away use HBM; that path waits for its own gather at once,
// because vector-memory operations complete in order and a wait placed later would
// also wait for the next parcel's loads and this parcel's stores.
template <class T>
struct TileEnv {
    const T *__restrict__ U;
    double *__restrict__ out;  // stray array (owner flush) or the source (atomic flush)
    const ST_LDS T *Ut;        // LDS velocity tile, AoS
    ST_LDS double *St;         // LDS source tile, AoS
    int32_t h0, h1, h2;        // tile origin (cells) = brick origin - 1
    uint32_t w0, w1, w2;       // tile extent = brick extent + 2
    uint32_t e0, e1, e2;       // brick extent
    uint32_t strays;
    uint32_t nxny, nx;         // mesh strides (far path only)
    // Tile index of a cell.  The tile extents are far below 2^24, so the index is formed
    // with full-rate 24-bit multiplies; it is garbage (and unused) outside the tile.
    __device__ __forceinline__ bool local(const Cell &cl, uint32_t &l) const {
        const uint32_t a = (uint32_t)(cl.f[0] - h0);
        const uint32_t b = (uint32_t)(cl.f[1] - h1);
        const uint32_t c = (uint32_t)(cl.f[2] - h2);
        l = a + __umul24(w0, b + __umul24(w1, c));
        return a < w0 && b < w1 && c < w2;
    }
    __device__ __forceinline__ uint32_t global(const Cell &cl) const {
        return (uint32_t)cl.f[2] * nxny + (uint32_t)cl.f[1] * nx + (uint32_t)cl.f[0];
    }
    __device__ __forceinline__ void enter(const Cell &cl) {
        const uint32_t a = (uint32_t)(cl.f[0] - h0 - 1);
        const uint32_t b = (uint32_t)(cl.f[1] - h1 - 1);
        const uint32_t c = (uint32_t)(cl.f[2] - h2 - 1);
        if (!(a < e0 && b < e1 && c < e2)) ++strays;
    }
    struct Slot {
        uint32_t l;  // tile index (meaningful when in)
        bool in;     // the cell lies in the tile
    };
    __device__ __forceinline__ Slot slot(const Cell &cl) const {
        Slot sl;
        sl.in = local(cl, sl.l);
        return sl;
    }
    __device__ __forceinline__ void load_u(const Cell &cl, const Slot &sl, T &ux, T &uy, T &uz) const {
        if (sl.in) {
            const uint32_t l = sl.l;
            ux = Ut[3u * l];
            uy = Ut[3u * l + 1];
            uz = Ut[3u * l + 2];
        } else {
            const T *q = U + 3u * global(cl);
            ux = q[0];
            uy = q[1];
            uz = q[2];
            __builtin_amdgcn_s_waitcnt(0);
        }
    }
    __device__ __forceinline__ void deposit(const Cell &cl, const Slot &sl, double a, double b, double c) {
        if (sl.in) {
            ST_LDS double *q = St + 3u * sl.l;
            atomic_add_lds(q, a);
            atomic_add_lds(q + 1, b);
            atomic_add_lds(q + 2, c);
        } else {
            double *q = out + 3u * global(cl);
            atomic_add_global(q, a);
            atomic_add_global(q + 1, b);
            atomic_add_global(q + 2, c);
        }
    }
};

template <class T>
struct Box {
    T lo[3], hi[3];
    __device__ __forceinline__ void init() {
        for (int c = 0; c < 3; ++c) {
            lo[c] = (T)INFINITY;
            hi[c] = -(T)INFINITY;
        }
    }
    // Compare-and-select: a NaN coordinate is skipped, as fmin/fmax skip it, without
    // the operand canonicalisation fmin/fmax cost.
    __device__ __forceinline__ void add(T x, T y, T z) {
        lo[0] = x < lo[0] ? x : lo[0];
        lo[1] = y < lo[1] ? y : lo[1];
        lo[2] = z < lo[2] ? z : lo[2];
        hi[0] = x > hi[0] ? x : hi[0];
        hi[1] = y > hi[1] ? y : hi[1];
        hi[2] = z > hi[2] ? z : hi[2];
    }
};

// Marks a parcel removed (channel form): its diameter stored negated.  The diameter is
// read again here rather than kept in a register through the substeps, on this rare
// path only (the wait for it also waits for the loads in flight).
template <class T>
__device__ __forceinline__ void mark_removed(const Soa<T> &s, uint32_t off) {
    st_off(s.d, off, -ld_off(s.d, off));
}

struct NoPrefetch {
    __device__ __forceinline__ void operator()() const {}
};

// One parcel, nsteps substeps (asyncParallelTracking.jl:1271-1365).  `prefetch` runs
// once the parcel's starting cell velocity has been read, so loads it issues (the next
// parcel's state) stay in flight across the substeps: vector-memory loads complete in
// order, and a later HBM gather would otherwise wait for them.
template <class T, int BND, bool RHO, class Env, class Prefetch>
__device__ __forceinline__ void step_parcel(const Geo &g, const Phys<T> &ph, const Soa<T> &s,
                                            const ParcelIn<T> &in, uint32_t p, int32_t nsteps,
                                            Env &env, Box<T> &box, const Prefetch &prefetch) {
    // removed earlier (channel form): no substep, nothing stored, deposited or counted,
    // and the next parcel's loads are issued from the same place as for the others
    const bool live = BND != kChannel || in.d > (T)0;
    double m;
    T f, den;
    parcel_consts(ph, RHO, in.d, in.rho, m, f, den);
    const Quot<T> q(den);

    T px = in.x, py = in.y, pz = in.z;
    T vx = in.u, vy = in.v, vz = in.w;

    int32_t I[3];
    Cell cell;
    locate<T, BND == kPeriodic>(g, px, py, pz, I, cell);  // :1286
    const int32_t part0 = cell.part;
    if (live) env.enter(cell);
    typename Env::Slot sl = env.slot(cell);
    T ux = (T)0, uy = (T)0, uz = (T)0;
    if (live) env.load_u(cell, sl, ux, uy, uz);  // :1287-1288
    double s0x = (double)vx, s0y = (double)vy, s0z = (double)vz;  // :1289
    prefetch();

    bool exited = false;
    const int32_t steps = live ? nsteps : 0;
    for (int32_t t = 1; t <= steps; ++t) {
        T nx, ny, nz, nvx, nvy, nvz;
        Cell nc;
        const unsigned b = advance<T, BND>(g, ph, f, q, ux, uy, uz, px, py, pz, vx, vy, vz, cell,
                                           part0, nx, ny, nz, nvx, nvy, nvz, nc);
        if (BND == kChannel && (b & kExit)) {  // the residence, and the parcel, end here
            vx = nvx;
            vy = nvy;
            vz = nvz;
            px = nx;
            py = ny;
            pz = nz;
            exited = true;
            break;
        }
        if (b & 1u) s0x = s0x - 2.0 * (double)vx;  // update_at_boundary!, :730-736
        if (b & 2u) s0y = s0y - 2.0 * (double)vy;
        if (b & 4u) s0z = s0z - 2.0 * (double)vz;
        if (t != nsteps && !same_cell(nc, cell)) {  // :1342-1348
            env.deposit(cell, sl, m * (s0x - (double)nvx), m * (s0y - (double)nvy), m * (s0z - (double)nvz));
            s0x = (double)nvx;
            s0y = (double)nvy;
            s0z = (double)nvz;
            cell = nc;
            sl = env.slot(cell);
            env.load_u(cell, sl, ux, uy, uz);
        }
        vx = nvx;  // :1351-1352
        vy = nvy;
        vz = nvz;
        px = nx;
        py = ny;
        pz = nz;
    }
    // :1355-1361
    if (!live) return;
    const uint32_t off = p * (uint32_t)sizeof(T);
    st_off(s.x, off, px);
    st_off(s.y, off, py);
    st_off(s.z, off, pz);
    st_off(s.u, off, vx);
    st_off(s.v, off, vy);
    st_off(s.w, off, vz);
    env.deposit(cell, sl, m * (s0x - (double)vx), m * (s0y - (double)vy), m * (s0z - (double)vz));
    if (BND == kChannel && exited)
        mark_removed(s, off);
    else
        box.add(px, py, pz);  // :1362
}

// The same step for a pass of exactly one substep (the schedule of one launch per
// substep).  With t == nSteps the reference never deposits or refreshes the velocity
// inside the loop (:1342) and never uses the new cell, so the relocation after a bounce
// or wrap (:1333-1335) has no effect on any output and is left out; every value that
// is written is computed by the same operations as in step_parcel.
template <class T, int BND, bool RHO, class Env, class Prefetch>
__device__ __forceinline__ void step_parcel_one(const Geo &g, const Phys<T> &ph, const Soa<T> &s,
                                                const ParcelIn<T> &in, uint32_t p, Env &env, Box<T> &box,
                                                bool stats, const Prefetch &prefetch) {
    // removed earlier (channel form): stepped like the others but nothing is stored,
    // deposited or counted, so the next parcel's loads are issued from one place
    const bool live = BND != kChannel || in.d > (T)0;
    double m;
    T f, den;
    parcel_consts(ph, RHO, in.d, in.rho, m, f, den);
    const T px = in.x, py = in.y, pz = in.z;
    const T vx = in.u, vy = in.v, vz = in.w;

    int32_t I[3];
    Cell cell;
    locate<T, BND == kPeriodic>(g, px, py, pz, I, cell);  // :1286
    if (live) env.enter(cell);
    const typename Env::Slot sl = env.slot(cell);
    T ux = (T)0, uy = (T)0, uz = (T)0;
    if (live) env.load_u(cell, sl, ux, uy, uz);  // :1287-1288
    prefetch();

    // :1295-1296
    const Quot<T> q(den);
    T nvx, nvy, nvz;
    q.three(vx + f * ux, vy + f * uy, vz + f * uz, nvx, nvy, nvz);
    T nx = px + nvx * ph.dt;
    T ny = py + nvy * ph.dt;
    T nz = pz + nvz * ph.dt;
    double s0x = (double)vx, s0y = (double)vy, s0z = (double)vz;  // :1289
    bool exited = false;
    // Walls and open planes: only a new position outside [0, N) on some axis can bounce
    // or leave (a bounce needs an index out of range), so the per-axis side tests skip
    // the cell lookup, the fold and the same-cell test for a position inside; outside,
    // the reference's steps run unchanged.
    if (BND != kPeriodic && (axis_side((double)nx, g.o[0], g.rd[0], g.n[0]) |
                             axis_side((double)ny, g.o[1], g.rd[1], g.n[1]) |
                             axis_side((double)nz, g.o[2], g.rd[2], g.n[2])) != 0) {
        Cell nc;
        locate<T, false>(g, nx, ny, nz, I, nc);  // :1303
        if (BND == kChannel && (uint32_t)I[0] >= (uint32_t)g.n[0]) {
            exited = true;  // out through an open plane: no bounce, the state as advanced
        } else if (!same_cell(nc, cell) || nc.part != cell.part) {  // :1306
            if (I[0] < 0 || I[0] >= g.n[0]) {  // :1309-1329, :730-736
                nx = px;
                nvx = -vx;
                s0x = s0x - 2.0 * (double)vx;
            }
            if (I[1] < 0 || I[1] >= g.n[1]) {
                ny = py;
                nvy = -vy;
                s0y = s0y - 2.0 * (double)vy;
            }
            if (I[2] < 0 || I[2] >= g.n[2]) {
                nz = pz;
                nvz = -vz;
                s0z = s0z - 2.0 * (double)vz;
            }
        }
    } else if (BND == kPeriodic) {
        // periodic extension: shift each axis that left [0, N) by one period
        const int s0 = axis_side((double)nx, g.o[0], g.rd[0], g.n[0]);
        const int s1 = axis_side((double)ny, g.o[1], g.rd[1], g.n[1]);
        const int s2 = axis_side((double)nz, g.o[2], g.rd[2], g.n[2]);
        if (s0 < 0) nx = nx + ph.L[0];
        else if (s0 > 0) nx = nx - ph.L[0];
        if (s1 < 0) ny = ny + ph.L[1];
        else if (s1 > 0) ny = ny - ph.L[1];
        if (s2 < 0) nz = nz + ph.L[2];
        else if (s2 > 0) nz = nz - ph.L[2];
    }
    // :1355-1361
    if (!live) return;
    const uint32_t off = p * (uint32_t)sizeof(T);
    st_off(s.x, off, nx);
    st_off(s.y, off, ny);
    st_off(s.z, off, nz);
    st_off(s.u, off, nvx);
    st_off(s.v, off, nvy);
    st_off(s.w, off, nvz);
    env.deposit(cell, sl, m * (s0x - (double)nvx), m * (s0y - (double)nvy), m * (s0z - (double)nvz));
    if (BND == kChannel && exited)
        mark_removed(s, off);
    else if (stats)
        box.add(nx, ny, nz);  // :1362
}

// Workgroup reduction of the box and the stray count into this workgroup's statistics
// slot (plain stores).  Every thread of the workgroup must call it (it has barriers).
template <class T>
__device__ __forceinline__ void flush_stats(Box<T> &box, uint32_t strays, double *slot) {
    __shared__ double red[7][1024 / kWave];
    static_assert(sizeof(red) + sizeof(int32_t) <= kPassStaticLds, "static LDS above the budget check_tile uses");
    double v[7];
    for (int c = 0; c < 3; ++c) {
        v[c] = (double)box.lo[c];
        v[3 + c] = (double)box.hi[c];
    }
    v[6] = (double)strays;
    for (int off = kWave / 2; off > 0; off >>= 1) {
        for (int c = 0; c < 3; ++c) {
            v[c] = fmin(v[c], __shfl_xor(v[c], off, kWave));
            v[3 + c] = fmax(v[3 + c], __shfl_xor(v[3 + c], off, kWave));
        }
        v[6] += __shfl_xor(v[6], off, kWave);
    }
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
    if (lane == 0)
        for (int c = 0; c < 7; ++c) red[c][wid] = v[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nw; ++w) {
            for (int c = 0; c < 3; ++c) {
                v[c] = fmin(v[c], red[c][w]);
                v[3 + c] = fmax(v[3 + c], red[3 + c][w]);
            }
            v[6] += red[6][w];
        }
        for (int c = 0; c < 7; ++c) slot[c] = v[c];
    }
}

// The plain method's environment for a pass of one substep, where a parcel deposits
// exactly once (at the end of the pass): the deposit is kept, for a wave-wide
// reduction after the step, instead of being added at once.
template <class T>
struct DeferEnv : GlobalEnv<T> {
    uint32_t cell = 0xffffffffu;  // global cell index of the deposit, or none
    double s[3] = {0.0, 0.0, 0.0};
    __device__ __forceinline__ void deposit(const Cell &cl, const typename GlobalEnv<T>::Slot &, double a, double b,
                                            double c) {
        cell = cl.c;
        s[0] = a;
        s[1] = b;
        s[2] = c;
    }
};

// Adds one deposit per lane (cell 0xffffffff: none) into `out` with FP64 atomics after a
// segmented reduction across the wave: lanes holding the same cell as their neighbours
// (parcels ordered by cell, as injected parcels are) are summed first, and only the
// first lane of each run of equal cells adds, so a run of r lanes costs 3 atomics
// instead of 3 r.  The runs are found with ballots and summed with a suffix sum by
// doubling, each lane adding the lane d above it while that lane is in its run.  When
// no two neighbouring lanes share a cell, every lane adds its own deposit.  Every lane
// of the wave must call it.
__device__ __forceinline__ void wave_deposit(double *out, uint32_t cell, double a, double b, double c) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const uint32_t up = __shfl_up(cell, 1, kWave);
    const bool head = lane == 0 || up != cell;
    const unsigned long long runs = __ballot(!head && cell != 0xffffffffu);
    if (runs != 0ull) {
        const uint32_t down = __shfl_down(cell, 1, kWave);
        const unsigned long long tails = __ballot(lane == kWave - 1 || down != cell);
        const int end = lane + __builtin_ctzll(tails >> lane);  // last lane of this lane's run
        for (int d = 1; d < kWave; d <<= 1) {
            const double ta = __shfl_down(a, d, kWave);
            const double tb = __shfl_down(b, d, kWave);
            const double tc = __shfl_down(c, d, kWave);
            if (lane + d <= end) {
                a += ta;
                b += tb;
                c += tc;
            }
        }
    }
    if (head && cell != 0xffffffffu) {
        double *q = out + 3u * cell;
        atomic_add_global(q, a);
        atomic_add_global(q + 1, b);
        atomic_add_global(q + 2, c);
    }
}

// One workgroup of the plain method: one thread per parcel, parcels base + blk *
// blockDim.x + threadIdx.x, deposits with FP64 atomics into `out`, statistics into
// `slot` (when statistics are taken).  Over a sorted cloud's unsorted tail (a.tail0
// set) each live parcel counts as outside its brick.  With WAVE_SUMS and one substep
// per pass, each wave reduces its deposits before the atomics (wave_deposit).
template <class T, int BND, bool RHO, bool WAVE_SUMS = false>
__device__ __forceinline__ void plain_block(const PassArgs<T> &a, uint32_t base, uint32_t blk, double *out,
                                            double *slot) {
    const uint32_t p = base + blk * blockDim.x + threadIdx.x;
    Box<T> box;
    box.init();
    uint32_t tail = 0;
    if (WAVE_SUMS && a.nsteps == 1) {
        DeferEnv<T> env;
        env.U = a.U;
        env.src = out;
        if (p < a.np) {
            ParcelIn<T> in;
            load_parcel<T, RHO>(a.s, p, a.ph.rho, in);
            if (a.tail0 && in.d > (T)0) tail = 1;
            step_parcel_one<T, BND, RHO>(a.g, a.ph, a.s, in, p, env, box, true, NoPrefetch{});
        }
        wave_deposit(out, env.cell, env.s[0], env.s[1], env.s[2]);
    } else {
        GlobalEnv<T> env{a.U, out};
        if (p < a.np) {
            ParcelIn<T> in;
            load_parcel<T, RHO>(a.s, p, a.ph.rho, in);
            if (a.tail0 && in.d > (T)0) tail = 1;
            if (a.nsteps == 1)
                step_parcel_one<T, BND, RHO>(a.g, a.ph, a.s, in, p, env, box, true, NoPrefetch{});
            else
                step_parcel<T, BND, RHO>(a.g, a.ph, a.s, in, p, a.nsteps, env, box, NoPrefetch{});
        }
    }
    if (slot) flush_stats(box, tail, slot);
}

// One thread per parcel, from parcel 0 or, over a sorted cloud's unsorted tail, from
// tail0[0] on.
template <class T, int BND, bool RHO>
__global__ void __launch_bounds__(1024) k_pass_plain(PassArgs<T> a) {
    const uint32_t base = a.tail0 ? (uint32_t)a.tail0[0] : 0u;
    plain_block<T, BND, RHO>(a, base, blockIdx.x, a.src, a.part ? a.part + (size_t)blockIdx.x * kStatSlot : nullptr);
}

// The adjoint's checkpoint of a parcel (the state it starts a segment with), written
// from the registers the pass loaded it into instead of by a separate copy kernel.
template <class T>
__device__ __forceinline__ void save_checkpoint(T *ck, uint32_t n, uint32_t p, const ParcelIn<T> &in) {
    ck[p] = in.x;
    ck[n + p] = in.y;
    ck[2 * n + p] = in.z;
    ck[3 * n + p] = in.u;
    ck[4 * n + p] = in.v;
    ck[5 * n + p] = in.w;
}

// Tile cells handled per thread per round of the staging and flush loops.
constexpr int kStage = 4;
// Parcels per own cell from which a brick's source sums start from the old values.
constexpr int32_t kSeedDensity = 6;

// Index -> (i, j, k) of a tile of extent wx * wy * (any) cells, x fastest.  The
// quotients come from a float reciprocal: for t below 2^22 the rounding error of
// (t + 1/2) * (1/w) stays far below the 1/(2w) distance of t + 1/2 from a multiple
// of w, so truncation gives the exact integer quotient.
struct TileShape {
    int32_t wx, wxy;
    float rwx, rwxy;
    __device__ TileShape(int32_t x, int32_t y) : wx(x), wxy(x * y), rwx(1.0f / (float)x), rwxy(1.0f / (float)(x * y)) {}
    __device__ __forceinline__ void split(int32_t t, int32_t &i, int32_t &j, int32_t &k) const {
        k = (int32_t)(((float)t + 0.5f) * rwxy);
        const int32_t r = t - k * wxy;
        j = (int32_t)(((float)r + 0.5f) * rwx);
        i = r - j * wx;
    }
};

// One brick of a sorted cloud: stage its velocities in LDS, step its parcels, add its
// source sums to HBM once per cell.  Every thread of the workgroup must call it.
template <class T, int BND, bool RHO, bool CKPT = false>
__device__ __forceinline__ void process_brick(const PassArgs<T> &a, int32_t brick, int32_t nsteps,
                                              double *lds, double *part) {
    const int32_t bi = brick % a.nb[0];
    const int32_t bj = (brick / a.nb[0]) % a.nb[1];
    const int32_t bk = brick / (a.nb[0] * a.nb[1]);

    TileEnv<T> env;
    env.U = a.U;
    env.out = a.flush == 0 ? a.stray : a.src;
    const int32_t o0 = bi * a.bd[0], o1 = bj * a.bd[1], o2 = bk * a.bd[2];
    const int32_t ex = min(a.bd[0], a.g.n[0] - o0);
    const int32_t ey = min(a.bd[1], a.g.n[1] - o1);
    const int32_t ez = min(a.bd[2], a.g.n[2] - o2);
    env.h0 = o0 - 1;
    env.h1 = o1 - 1;
    env.h2 = o2 - 1;
    const int32_t wx = ex + 2, wy = ey + 2, wz = ez + 2;
    env.w0 = (uint32_t)wx;
    env.w1 = (uint32_t)wy;
    env.w2 = (uint32_t)wz;
    env.e0 = (uint32_t)ex;
    env.e1 = (uint32_t)ey;
    env.e2 = (uint32_t)ez;
    env.strays = 0;
    env.nxny = a.g.nxny;
    env.nx = (uint32_t)a.g.n[0];
    const int32_t ncell = wx * wy * wz;
    ST_LDS double *St = (ST_LDS double *)lds;
    ST_LDS T *Ut = (ST_LDS T *)(St + 3 * a.tile_cells);
    env.St = St;
    env.Ut = Ut;

    // the brick ranges do not change during a launch: read them as constants (scalar
    // loads), so no vector-memory wait is spent on them in a persistent launch either
    typedef __attribute__((address_space(4))) const int32_t ConstI32;
    const ConstI32 *bs_k = (const ConstI32 *)a.brick_start;
    const int32_t p0 = bs_k[brick], p1 = bs_k[brick + 1];
    const int32_t bs = (int32_t)blockDim.x;
    const uint32_t last = (uint32_t)(p1 - 1);
    int32_t p = p0 + (int32_t)threadIdx.x;
    // The first parcel's state is requested before the tile is staged, so its latency
    // overlaps the staging (the brick is non-empty: a clamped index is valid).
    ParcelIn<T> A, B;
    if (p1 > p0) load_parcel<T, RHO>(a.s, min((uint32_t)p, last), a.ph.rho, A);

    // Owner flush of a dense brick (at least kSeedDensity parcels per own cell, so that
    // nearly every own cell takes a deposit: with parcels spread at random a fraction
    // of about exp(-kSeedDensity) does not): the own cells' sums start from the values in
    // `src`, read with the velocities, and the flush stores them without a second
    // memory round trip.  Only this workgroup touches its own cells of `src` during
    // the launch, so the value read here is the one the flush overwrites.  A sparse
    // brick keeps zero-started sums and updates only the cells that took a deposit.
    // With zero_src the source counts as zero: every brick starts its own cells' sums
    // from zero without reading `src` and stores all of them, deposits or not.
    const bool seeded = a.flush == 0 && (a.zero_src != 0 || p1 - p0 >= kSeedDensity * ex * ey * ez);

    // Stage the tile's velocities and clear its sums, kStage cells per thread at a
    // time with every load issued before the first LDS write (one memory round trip per
    // kStage*blockDim cells).  Halo cells outside the mesh are never looked up (every
    // cell index is folded or clamped into the mesh): they read a clamped cell and
    // store zeros, so the velocity loads need no condition.  Two lanes may write the
    // same cell (the tile's last) with the same values: no barrier separates LDS writes
    // of equal data, and the barrier below orders them before any read.
    const TileShape ts(wx, wy);
    for (int32_t t0 = 0; t0 < ncell; t0 += kStage * bs) {
        T u[kStage][3];
        double o[kStage][3];
        bool in[kStage], own[kStage];
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int32_t t = min(t0 + k * bs + (int32_t)threadIdx.x, ncell - 1);
            int32_t li, lj, lk;
            ts.split(t, li, lj, lk);
            const int32_t gi = env.h0 + li, gj = env.h1 + lj, gk = env.h2 + lk;
            in[k] = (uint32_t)gi < (uint32_t)a.g.n[0] && (uint32_t)gj < (uint32_t)a.g.n[1] &&
                    (uint32_t)gk < (uint32_t)a.g.n[2];
            own[k] = seeded && (uint32_t)(li - 1) < (uint32_t)ex && (uint32_t)(lj - 1) < (uint32_t)ey &&
                     (uint32_t)(lk - 1) < (uint32_t)ez;
            const uint32_t c = (uint32_t)clampi(gk, a.g.n[2]) * a.g.nxny +
                               (uint32_t)clampi(gj, a.g.n[1]) * (uint32_t)a.g.n[0] + (uint32_t)clampi(gi, a.g.n[0]);
            const T *q = a.U + 3u * c;
            u[k][0] = q[0];
            u[k][1] = q[1];
            u[k][2] = q[2];
            o[k][0] = o[k][1] = o[k][2] = 0.0;
            if (own[k] && !a.zero_src) {
                const double *r = a.src + 3u * c;
                o[k][0] = r[0];
                o[k][1] = r[1];
                o[k][2] = r[2];
            }
        }
        // Lanes past the tile rewrite its last cell with the same values, so every load
        // is consumed on every path and none is left in flight around the loop.
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int32_t t = min(t0 + k * bs + (int32_t)threadIdx.x, ncell - 1);
            Ut[3 * t] = in[k] ? u[k][0] : (T)0;
            Ut[3 * t + 1] = in[k] ? u[k][1] : (T)0;
            Ut[3 * t + 2] = in[k] ? u[k][2] : (T)0;
            St[3 * t] = o[k][0];
            St[3 * t + 1] = o[k][1];
            St[3 * t + 2] = o[k][2];
        }
    }
    __syncthreads();

    Box<T> box;
    box.init();
    // Software pipeline: the next parcel's state is in flight while one steps.
    // Vector-memory counters complete in order, and the compiler places each wait for
    // the worst path into it, so the loop is shaped to leave it no pessimistic path:
    //  * the loads are unconditional (indices clamped to the brick's last parcel, a
    //    valid address once the brick is non-empty) -- a path that skipped them would
    //    make every iteration wait for all loads in flight, the prefetch included;
    //  * two register sets alternate (unrolled by two), so no copy of a register with a
    //    load in flight forces a wait at the end of an iteration;
    //  * the first loads are waited for before the loop, so the loop's entry path has
    //    nothing in flight and the waits inside count past earlier parcels' stores
    //    instead of waiting for them.
    // (A second parcel in flight per thread measured no faster and costs the FP64 and
    // batch kernels occupancy.)
    // the checkpointing pass may store the new state into other arrays than it loads
    const Soa<T> so = CKPT && a.out.x != nullptr ? a.out : a.s;
    if (p1 > p0) {
        __builtin_amdgcn_s_waitcnt(0);
        auto ahead = [&](int32_t q) { return min((uint32_t)(q + bs), last); };
        if (nsteps == 1) {
            const bool stats = part != nullptr;
            auto one = [&](const ParcelIn<T> &cur, ParcelIn<T> &fill) {
                const uint32_t q = ahead(p);
                if (CKPT && a.ck) save_checkpoint(a.ck, a.np, (uint32_t)p, cur);
                step_parcel_one<T, BND, RHO>(a.g, a.ph, so, cur, (uint32_t)p, env, box, stats,
                                                  [&] { load_parcel<T, RHO>(a.s, q, a.ph.rho, fill); });
                p += bs;
            };
            for (;;) {
                if (p >= p1) break;
                one(A, B);
                if (p >= p1) break;
                one(B, A);
            }
        } else {
            auto many = [&](const ParcelIn<T> &cur, ParcelIn<T> &fill) {
                const uint32_t q = ahead(p);
                if (CKPT && a.ck) save_checkpoint(a.ck, a.np, (uint32_t)p, cur);
                step_parcel<T, BND, RHO>(a.g, a.ph, so, cur, (uint32_t)p, nsteps, env, box,
                                              [&] { load_parcel<T, RHO>(a.s, q, a.ph.rho, fill); });
                p += bs;
            };
            for (;;) {
                if (p >= p1) break;
                many(A, B);
                if (p >= p1) break;
                many(B, A);
            }
        }
    }
    __syncthreads();

    // Add the tile's sums to HBM, kStage cells per thread at a time with every load of
    // the read-modify-writes issued before the first add.  With the owner flush only
    // this workgroup touches the brick's own cells of `src` during the launch
    // (deposits from other workgroups' halos and strays go to `stray`), so a plain
    // read-modify-write is race-free there; the halo's sums go to `stray` with atomics.
    // A cell with a nonzero sum lies inside the mesh (deposits go to folded or clamped
    // cells only).  A seeded brick's own sums already hold the old values: they are
    // stored without a read.
    for (int32_t t0 = 0; t0 < ncell; t0 += kStage * bs) {
        double v[kStage][3], old[kStage][3];
        uint32_t c[kStage];
        bool live[kStage], nz[kStage], own[kStage];
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int32_t t = t0 + k * bs + (int32_t)threadIdx.x;
            const int32_t tc = min(t, ncell - 1);
            v[k][0] = St[3 * tc];
            v[k][1] = St[3 * tc + 1];
            v[k][2] = St[3 * tc + 2];
            live[k] = t < ncell;
            nz[k] = live[k] && (v[k][0] != 0.0 || v[k][1] != 0.0 || v[k][2] != 0.0);
            int32_t li, lj, lk;
            ts.split(tc, li, lj, lk);
            c[k] = (uint32_t)(env.h2 + lk) * a.g.nxny + (uint32_t)(env.h1 + lj) * (uint32_t)a.g.n[0] +
                   (uint32_t)(env.h0 + li);
            own[k] = a.flush == 0 && (uint32_t)(li - 1) < (uint32_t)ex && (uint32_t)(lj - 1) < (uint32_t)ey &&
                     (uint32_t)(lk - 1) < (uint32_t)ez;
            old[k][0] = old[k][1] = old[k][2] = 0.0;
            if (!seeded && nz[k] && own[k]) {
                const double *q = a.src + 3u * c[k];
                old[k][0] = q[0];
                old[k][1] = q[1];
                old[k][2] = q[2];
            }
        }
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            if (!(seeded && own[k] ? live[k] : nz[k])) continue;
            if (own[k]) {
                double *q = a.src + 3u * c[k];
                q[0] = seeded ? v[k][0] : old[k][0] + v[k][0];
                q[1] = seeded ? v[k][1] : old[k][1] + v[k][1];
                q[2] = seeded ? v[k][2] : old[k][2] + v[k][2];
            } else {
                double *q = env.out + 3u * c[k];
                atomic_add_global(q, v[k][0]);
                atomic_add_global(q + 1, v[k][1]);
                atomic_add_global(q + 2, v[k][2]);
            }
        }
    }
    if (part) flush_stats(box, env.strays, part + (size_t)brick * kStatSlot);
}

template <class T, int BND, bool RHO, bool CKPT = false>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 4))) k_pass_sorted(PassArgs<T> a) {
    extern __shared__ double lds[];
    // Bricks in dispatch order: workgroups go round-robin over the XCDs, so the bricks
    // in flight at any time form one contiguous window of the parcel arrays (measured
    // faster on MI355X than giving each XCD its own contiguous range of bricks, which
    // spreads the concurrent streams over eight windows).
    process_brick<T, BND, RHO, CKPT>(a, blockIdx.x, a.nsteps, lds, a.part);
}

// A pass over a sorted cloud with an unsorted tail (parcels injected since its last
// sort) in one launch: workgroups 0 .. nbricks - 1 step the bricks, the ones after them
// the tail, one thread per parcel from a.tail0[0] on.  With G workgroups of which
// n = G - nbricks are tail workgroups, workgroup w runs tail workgroup floor(w n / G)
// when floor((w + 1) n / G) exceeds it, else brick w - floor(w n / G): the tail's
// latency-bound workgroups are spread evenly through dispatch order and overlap the
// bricks' bandwidth-bound sweep (tail workgroups last, or first, are slower: the drain
// lengthens, or the sweep starts late).
// Tail deposits go where deposits from outside a brick go -- `stray` with atomics under
// the owner flush (bricks update their own cells of `src` with plain read-modify-writes
// in the same launch), `src` with atomics under the atomic flush.  Statistics slots:
// one per brick, then one per tail workgroup.
template <class T, int BND, bool RHO>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 4)))
k_pass_sorted_tail(PassArgs<T> a) {
    extern __shared__ double lds[];
    const uint64_t G = gridDim.x, n = G - (uint64_t)a.nbricks, w = blockIdx.x;
    const uint32_t k0 = (uint32_t)(w * n / G), k1 = (uint32_t)((w + 1) * n / G);
    if (k1 > k0) {
        plain_block<T, BND, RHO, true>(a, (uint32_t)a.tail0[0], k0, a.flush == 0 ? a.stray : a.src,
                                       a.part ? a.part + (size_t)(a.nbricks + k0) * kStatSlot : nullptr);
    } else {
        process_brick<T, BND, RHO, false>(a, (int32_t)(w - k0), a.nsteps, lds, a.part);
    }
}

// Several passes of one sorted cloud in one persistent launch.  Workgroup w owns bricks
// w, w + gridDim.x, ... in every pass and runs its bricks pass by pass, so a brick's
// pass k + 1 follows its pass k in the same workgroup: the parcels of a brick, and its
// own cells of `src` (owner flush), are read and written by that workgroup alone, with
// its barriers ordering them, and everything shared (halo and stray sums) goes to
// `stray` with atomics.  Pass-major order keeps each pass a full sweep over the
// parcel arrays in HBM.  `part` (statistics) is filled in the last pass only.
template <class T, int BND, bool RHO>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 4)))
k_pass_multi(PassArgs<T> a, int32_t npass) {
    extern __shared__ double lds[];
    for (int32_t k = 0; k < npass; ++k) {
        double *part = k == npass - 1 ? a.part : nullptr;
        for (int32_t brick = blockIdx.x; brick < a.nbricks; brick += gridDim.x) {
            process_brick<T, BND, RHO>(a, brick, a.nsteps, lds, part);
            __syncthreads();  // the flush's tile reads are done before the next staging
        }
    }
}

template <class T>
hipError_t launch_pass_multi(const PassArgs<T> &a, int32_t npass, int bnd, int block, int grid, hipStream_t s) {
    const size_t lds = (size_t)a.tile_cells * 3 * (sizeof(double) + sizeof(T));
    with_forms(bnd, a.s.rho != nullptr, [&](auto B, auto R) {
        k_pass_multi<T, decltype(B)::value, decltype(R)::value><<<grid, block, lds, s>>>(a, npass);
    });
    return hipGetLastError();
}
template hipError_t launch_pass_multi<float>(const PassArgs<float> &, int32_t, int, int, int, hipStream_t);
template hipError_t launch_pass_multi<double>(const PassArgs<double> &, int32_t, int, int, int, hipStream_t);

template <class T>
int pass_multi_blocks_per_cu(int bnd, int rho, int block, size_t lds) {
    int n = 0;
    hipError_t e = hipSuccess;
    with_forms(bnd, rho != 0, [&](auto B, auto R) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pass_multi<T, decltype(B)::value, decltype(R)::value>,
                                                         block, lds);
    });
    return e == hipSuccess ? n : 0;
}
template int pass_multi_blocks_per_cu<float>(int, int, int, size_t);
template int pass_multi_blocks_per_cu<double>(int, int, int, size_t);

// Many clouds in one persistent launch: each workgroup takes (cloud, brick) work items
// from a global counter until none are left.  item0[c] is the first item of cloud c
// (clouds ordered densest first, so the longest items are handed out first); every
// workgroup leaves once the counter passes item0[ncloud].
//
// A workgroup claims its next item as it starts the current one: the counter's
// atomic (a round trip to the memory side, since the counter is shared by all XCDs)
// then completes under the brick's tile staging instead of stalling the workgroup
// between items.  The FP32 instances are held to 80 VGPRs (6 waves per SIMD), the
// budget of the single-cloud kernel: the loop over items otherwise keeps enough
// per-thread values live across items to drop the occupancy by a wave.
template <class T, int BND, bool RHO>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 4)))
k_pass_batch(const PassArgs<T> *__restrict__ descs, const int32_t *__restrict__ item0, int32_t ncloud,
             int32_t nsteps, int stats, unsigned int *counter, int zero_src) {
    extern __shared__ double lds[];
    __shared__ int32_t s_item;
    const int32_t total = item0[ncloud];
    if (threadIdx.x == 0) s_item = (int32_t)atomicAdd(counter, 1u);
    __syncthreads();
    int32_t item = __builtin_amdgcn_readfirstlane(s_item);
    while (item < total) {
        int32_t next = 0;
        if (threadIdx.x == 0) next = (int32_t)atomicAdd(counter, 1u);
        int32_t lo = 0, hi = ncloud - 1;  // last cloud with item0 <= item
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (item0[mid] <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int32_t c = __builtin_amdgcn_readfirstlane(lo);
        PassArgs<T> a = descs[c];
        a.zero_src = zero_src;
        process_brick<T, BND, RHO>(a, item - item0[c], nsteps, lds, stats ? a.part : nullptr);
        __syncthreads();  // the flush's tile reads are done before the next staging
        if (threadIdx.x == 0) s_item = next;
        __syncthreads();
        item = __builtin_amdgcn_readfirstlane(s_item);
    }
}

template <class T>
hipError_t launch_pass_batch(const PassArgs<T> *descs, const int32_t *item0, int32_t ncloud, int32_t nsteps,
                             int stats, unsigned int *counter, int periodic, int rho, int block, int grid,
                             size_t lds, int zero_src, hipStream_t s) {
    if (periodic && rho)
        k_pass_batch<T, true, true><<<grid, block, lds, s>>>(descs, item0, ncloud, nsteps, stats, counter, zero_src);
    else if (periodic)
        k_pass_batch<T, true, false><<<grid, block, lds, s>>>(descs, item0, ncloud, nsteps, stats, counter, zero_src);
    else if (rho)
        k_pass_batch<T, false, true><<<grid, block, lds, s>>>(descs, item0, ncloud, nsteps, stats, counter, zero_src);
    else
        k_pass_batch<T, false, false><<<grid, block, lds, s>>>(descs, item0, ncloud, nsteps, stats, counter,
                                                               zero_src);
    return hipGetLastError();
}
template hipError_t launch_pass_batch<float>(const PassArgs<float> *, const int32_t *, int32_t, int32_t, int,
                                             unsigned int *, int, int, int, int, size_t, int, hipStream_t);
template hipError_t launch_pass_batch<double>(const PassArgs<double> *, const int32_t *, int32_t, int32_t, int,
                                              unsigned int *, int, int, int, int, size_t, int, hipStream_t);

template <class T>
int batch_max_blocks_per_cu(int periodic, int rho, int block, size_t lds) {
    int n = 0;
    hipError_t e;
    if (periodic && rho)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pass_batch<T, true, true>, block, lds);
    else if (periodic)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pass_batch<T, true, false>, block, lds);
    else if (rho)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pass_batch<T, false, true>, block, lds);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pass_batch<T, false, false>, block, lds);
    return e == hipSuccess ? n : 0;
}
template int batch_max_blocks_per_cu<float>(int, int, int, size_t);
template int batch_max_blocks_per_cu<double>(int, int, int, size_t);

template <class T>
hipError_t launch_pass_plain(const PassArgs<T> &a, int bnd, int block, hipStream_t s, uint32_t grid) {
    if (a.np == 0) return hipSuccess;
    if (grid == 0) grid = (a.np + block - 1) / block;
    with_forms(bnd, a.s.rho != nullptr, [&](auto B, auto R) {
        k_pass_plain<T, decltype(B)::value, decltype(R)::value><<<grid, block, 0, s>>>(a);
    });
    return hipGetLastError();
}

template <class T>
hipError_t launch_pass_sorted(const PassArgs<T> &a, int bnd, int block, hipStream_t s) {
    const size_t lds = (size_t)a.tile_cells * 3 * (sizeof(double) + sizeof(T));
    if constexpr (sizeof(T) == 8) {
        if (a.ck || (a.out.x && a.out.x != a.s.x)) {  // the adjoint's forward: wall and periodic forms only
            with_forms(bnd == kPeriodic ? kPeriodic : kWall, a.s.rho != nullptr, [&](auto B, auto R) {
                k_pass_sorted<T, decltype(B)::value, decltype(R)::value, true><<<a.nbricks, block, lds, s>>>(a);
            });
            return hipGetLastError();
        }
    }
    with_forms(bnd, a.s.rho != nullptr, [&](auto B, auto R) {
        k_pass_sorted<T, decltype(B)::value, decltype(R)::value><<<a.nbricks, block, lds, s>>>(a);
    });
    return hipGetLastError();
}

template <class T>
hipError_t launch_pass_sorted_tail(const PassArgs<T> &a, uint32_t tail_blocks, int bnd, int block, hipStream_t s) {
    const size_t lds = (size_t)a.tile_cells * 3 * (sizeof(double) + sizeof(T));
    const uint32_t grid = (uint32_t)a.nbricks + tail_blocks;
    with_forms(bnd, a.s.rho != nullptr, [&](auto B, auto R) {
        k_pass_sorted_tail<T, decltype(B)::value, decltype(R)::value><<<grid, block, lds, s>>>(a);
    });
    return hipGetLastError();
}
template hipError_t launch_pass_sorted_tail<float>(const PassArgs<float> &, uint32_t, int, int, hipStream_t);
template hipError_t launch_pass_sorted_tail<double>(const PassArgs<double> &, uint32_t, int, int, hipStream_t);

template hipError_t launch_pass_plain<float>(const PassArgs<float> &, int, int, hipStream_t, uint32_t);
template hipError_t launch_pass_plain<double>(const PassArgs<double> &, int, int, hipStream_t, uint32_t);
template hipError_t launch_pass_sorted<float>(const PassArgs<float> &, int, int, hipStream_t);
template hipError_t launch_pass_sorted<double>(const PassArgs<double> &, int, int, hipStream_t);

// ---------------------------------------------------------------------------------
// Small utility kernels (graph-capturable: no hipMemsetAsync anywhere).

// Reduces one cloud's statistics slots: one workgroup of kReduceBlock threads per
// target.  Each thread requests kReducePer slots before it combines any of them, so
// every kReduceBlock * kReducePer slots (4096) cost one memory round trip instead of one
// per slot a thread visits.  min, max and the sum of integer counts are exact in any
// order.
constexpr int kReduceBlock = 1024;
constexpr int kReducePer = 4;

__device__ __forceinline__ void reduce_stats(const StatsTarget &t) {
    __shared__ double red[7][kReduceBlock / kWave];
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int32_t k0 = threadIdx.x; k0 < t.nslots; k0 += kReducePer * kReduceBlock) {
        double q[kReducePer][7];
        bool ok[kReducePer];
#pragma unroll
        for (int r = 0; r < kReducePer; ++r) {
            const int32_t k = k0 + r * kReduceBlock;
            ok[r] = k < t.nslots;
            const double *slot = t.part + (size_t)(ok[r] ? k : k0) * kStatSlot;  // k0 < nslots
#pragma unroll
            for (int c = 0; c < 7; ++c) q[r][c] = slot[c];
        }
#pragma unroll
        for (int r = 0; r < kReducePer; ++r) {
            if (!ok[r]) continue;
            for (int c = 0; c < 3; ++c) {
                v[c] = fmin(v[c], q[r][c]);
                v[3 + c] = fmax(v[3 + c], q[r][3 + c]);
            }
            v[6] += q[r][6];
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        for (int c = 0; c < 3; ++c) {
            v[c] = fmin(v[c], __shfl_xor(v[c], off, kWave));
            v[3 + c] = fmax(v[3 + c], __shfl_xor(v[3 + c], off, kWave));
        }
        v[6] += __shfl_xor(v[6], off, kWave);
    }
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    if (lane == 0)
        for (int c = 0; c < 7; ++c) red[c][wid] = v[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kReduceBlock / kWave; ++w) {
            for (int c = 0; c < 3; ++c) {
                v[c] = fmin(v[c], red[c][w]);
                v[3 + c] = fmax(v[3 + c], red[3 + c][w]);
            }
            v[6] += red[6][w];
        }
        for (int c = 0; c < 6; ++c) t.bbox[c] = v[c];
        const unsigned int strays = (unsigned int)v[6];
        const unsigned int gate = t.resort && strays > t.threshold ? 1u : 0u;
        t.strays[0] = strays;
        if (t.resort) t.resort[0] = gate;
        if (t.host) {  // straight to the host: no copy launch behind the step
            t.host[0] = gate;
            t.host[1] = strays;
        }
    }
}

__global__ void __launch_bounds__(kReduceBlock) k_reduce_stats(StatsTarget t) { reduce_stats(t); }
__global__ void __launch_bounds__(kReduceBlock) k_reduce_stats_many(const StatsTarget *ts) {
    reduce_stats(ts[blockIdx.x]);
}

hipError_t launch_reduce_stats(const StatsTarget &t, hipStream_t s) {
    k_reduce_stats<<<1, kReduceBlock, 0, s>>>(t);
    return hipGetLastError();
}

hipError_t launch_reduce_stats_many(const StatsTarget *targets, int32_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_reduce_stats_many<<<n, kReduceBlock, 0, s>>>(targets);
    return hipGetLastError();
}

// A gated kernel does nothing unless its gate is set (the gate is uniform: every
// thread returns or none does).
__device__ __forceinline__ bool gated_off(const unsigned int *gate) { return gate && gate[0] == 0u; }

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n, const unsigned int *gate) {
    if (gated_off(gate)) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_f64(double *p, double v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_iota(int32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}
__global__ void k_merge(double *src, double *stray, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double s = stray[i];
        if (s != 0.0) {
            src[i] += s;
            stray[i] = 0.0;
        }
    }
}

// Zeroes many arrays in one launch: workgroup w clears chunk w of the concatenation
// (items ordered, chunk0[i] = first chunk of item i, chunk0[count] = all chunks).
constexpr size_t kFillChunk = 16384;
__global__ void __launch_bounds__(256) k_fill_many(const FillItem *items, const uint32_t *chunk0, int32_t count) {
    const uint32_t w = blockIdx.x;
    int32_t lo = 0, hi = count - 1;  // last item with chunk0 <= w
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (chunk0[mid] <= w)
            lo = mid;
        else
            hi = mid - 1;
    }
    const FillItem it = items[lo];
    const size_t b0 = (size_t)(w - chunk0[lo]) * kFillChunk;
    const size_t b1 = min(b0 + kFillChunk, it.n);
    for (size_t i = b0 + threadIdx.x; i < b1; i += 256) it.p[i] = 0.0;
}

uint32_t fill_many_chunks(size_t n) { return (uint32_t)((n + kFillChunk - 1) / kFillChunk); }

hipError_t launch_fill_many(const FillItem *items, const uint32_t *chunk0, int32_t count, uint32_t chunks,
                            hipStream_t s) {
    if (count <= 0 || chunks == 0) return hipSuccess;
    k_fill_many<<<chunks, 256, 0, s>>>(items, chunk0, count);
    return hipGetLastError();
}

static unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g == 0) g = 1;
    return (unsigned)g;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s, const unsigned int *gate) {
    if (!n) return hipSuccess;
    k_fill_u32<<<grid_for(n, 256), 256, 0, s>>>(p, v, n, gate);
    return hipGetLastError();
}
hipError_t launch_fill_f64(double *p, double v, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    k_fill_f64<<<grid_for(n, 256), 256, 0, s>>>(p, v, n);
    return hipGetLastError();
}
hipError_t launch_iota(int32_t *p, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    k_iota<<<(n + 255) / 256, 256, 0, s>>>(p, n);
    return hipGetLastError();
}
hipError_t launch_merge(double *src, double *stray, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    k_merge<<<grid_for(n, 256), 256, 0, s>>>(src, stray, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Counting sort of parcels by brick.

// Wave-aggregated atomicAdd: lanes with equal keys share one atomic; each active lane
// gets base + its rank among them (lanes of a group in lane order).  The groups are
// found first with ballots alone; then every group's first lane adds its group's size
// in one atomic instruction for the whole wave, so a wave whose 64 keys all differ
// (an unsorted cloud) waits for one memory round trip, not 64 in a row.  All lanes of
// the wave must call it.
__device__ __forceinline__ uint32_t wave_claim(uint32_t *ctr, uint32_t key, bool active) {
    const int lane = threadIdx.x % kWave;
    unsigned long long pending = __ballot(active);
    int my_leader = lane;
    uint32_t rank = 0, group = 0;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t lkey = __shfl(key, leader, kWave);
        const bool mine = ((pending >> lane) & 1ull) && key == lkey;
        const unsigned long long grp = __ballot(mine);
        if (mine) {
            my_leader = leader;
            rank = (uint32_t)__popcll(grp & ((1ull << lane) - 1ull));
        }
        if (lane == leader) group = (uint32_t)__popcll(grp);
        pending &= ~grp;
    }
    uint32_t base = 0;
    if (group) base = atomicAdd(ctr + key, group);  // a leader's own key is its group's key
    base = __shfl(base, my_leader, kWave);
    return base + rank;
}

template <class T, bool PERIODIC>
__global__ void k_brick_keys(Soa<T> s, uint32_t np, Geo g, int32_t nb0, int32_t nb1, int32_t bd0,
                             int32_t bd1, int32_t bd2, uint32_t *keys, uint32_t *counts,
                             const unsigned int *gate) {
    if (gated_off(gate)) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = p < np;
    uint32_t key = 0;
    if (active) {
        int32_t I[3];
        Cell cl;
        locate<T, PERIODIC>(g, s.x[p], s.y[p], s.z[p], I, cl);
        key = (uint32_t)(cl.f[0] / bd0) +
              (uint32_t)nb0 * ((uint32_t)(cl.f[1] / bd1) + (uint32_t)nb1 * (uint32_t)(cl.f[2] / bd2));
        keys[p] = key;
    }
    wave_claim(counts, key, active);
}

template <class T>
hipError_t launch_brick_keys(const Soa<T> &s, uint32_t np, Geo g, const int32_t nb[3],
                             const int32_t bd[3], int periodic, uint32_t *keys, uint32_t *counts,
                             const unsigned int *gate, hipStream_t st) {
    if (!np) return hipSuccess;
    const uint32_t grid = (np + 255) / 256;
    if (periodic)
        k_brick_keys<T, true><<<grid, 256, 0, st>>>(s, np, g, nb[0], nb[1], bd[0], bd[1], bd[2], keys,
                                                     counts, gate);
    else
        k_brick_keys<T, false><<<grid, 256, 0, st>>>(s, np, g, nb[0], nb[1], bd[0], bd[1], bd[2], keys,
                                                      counts, gate);
    return hipGetLastError();
}
template hipError_t launch_brick_keys<float>(const Soa<float> &, uint32_t, Geo, const int32_t *,
                                             const int32_t *, int, uint32_t *, uint32_t *,
                                             const unsigned int *, hipStream_t);
template hipError_t launch_brick_keys<double>(const Soa<double> &, uint32_t, Geo, const int32_t *,
                                              const int32_t *, int, uint32_t *, uint32_t *,
                                              const unsigned int *, hipStream_t);

// Exclusive scan of the bin counts in one workgroup of 1024 threads: each thread
// sums a contiguous segment, the 1024 partial sums are scanned in LDS, then each
// thread writes its segment.  start has nbins + 1 entries; cursor = start[0..nbins).
__global__ void __launch_bounds__(1024) k_scan(const uint32_t *counts, int32_t nbins, int32_t *start,
                                               uint32_t *cursor, const unsigned int *gate) {
    if (gated_off(gate)) return;
    __shared__ uint32_t part[1024];
    const int32_t t = threadIdx.x;
    const int32_t seg = (nbins + 1023) / 1024;
    const int32_t b0 = min(t * seg, nbins), b1 = min(b0 + seg, nbins);
    uint32_t sum = 0;
    for (int32_t b = b0; b < b1; ++b) sum += counts[b];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t add = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;  // exclusive prefix of this segment
    for (int32_t b = b0; b < b1; ++b) {
        start[b] = (int32_t)run;
        cursor[b] = run;
        run += counts[b];
    }
    if (t == 1023) start[nbins] = (int32_t)part[1023];
}

hipError_t launch_scan(const uint32_t *counts, int32_t nbins, int32_t *start, uint32_t *cursor,
                       const unsigned int *gate, hipStream_t s) {
    k_scan<<<1, 1024, 0, s>>>(counts, nbins, start, cursor, gate);
    return hipGetLastError();
}

// Chunked counting sort, for brick counts whose histogram fits one workgroup's LDS.
// Workgroup g owns parcels [g C, (g+1) C): it counts its keys in LDS and writes its
// row of hist (g-major, nbins per row); k_col_prefix turns every column (brick) into
// the exclusive prefix over workgroups and the brick totals; the scatter then claims
// slots with LDS atomics on start[b] + hist[g][b].  No global atomics at all, where
// one counter per brick would take every parcel's atomic.
constexpr int kSortBlock = 1024;

template <class T, bool PERIODIC>
__global__ void __launch_bounds__(kSortBlock) k_sort_count(Soa<T> s, uint32_t np, uint32_t chunk, Geo g,
                                                           int32_t nb0, int32_t nb1, int32_t bd0, int32_t bd1,
                                                           int32_t bd2, int32_t nbins, uint32_t *keys,
                                                           uint32_t *hist, const unsigned int *gate) {
    if (gated_off(gate)) return;
    extern __shared__ uint32_t h[];
    for (int32_t b = threadIdx.x; b < nbins; b += kSortBlock) h[b] = 0u;
    __syncthreads();
    const uint32_t p0 = blockIdx.x * chunk, p1 = min(p0 + chunk, np);
    for (uint32_t p = p0 + threadIdx.x; p < p1; p += kSortBlock) {
        int32_t I[3];
        Cell cl;
        locate<T, PERIODIC>(g, s.x[p], s.y[p], s.z[p], I, cl);
        const uint32_t key = (uint32_t)(cl.f[0] / bd0) +
                             (uint32_t)nb0 * ((uint32_t)(cl.f[1] / bd1) + (uint32_t)nb1 * (uint32_t)(cl.f[2] / bd2));
        keys[p] = key;
        atomicAdd(&h[key], 1u);
    }
    __syncthreads();
    uint32_t *row = hist + (size_t)blockIdx.x * nbins;
    for (int32_t b = threadIdx.x; b < nbins; b += kSortBlock) row[b] = h[b];
}

// Per brick: the exclusive prefix of its counts over the workgroups (in place) and the
// total (into counts, for k_scan).
__global__ void k_col_prefix(uint32_t *__restrict__ hist, int32_t nbins, int32_t nrows, uint32_t *counts,
                             const unsigned int *gate) {
    if (gated_off(gate)) return;
    const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    constexpr int kBatch = 16;  // loads in flight per thread
    uint32_t run = 0;
    for (int32_t r0 = 0; r0 < nrows; r0 += kBatch) {
        uint32_t v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) v[k] = r0 + k < nrows ? hist[(size_t)(r0 + k) * nbins + b] : 0u;
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            if (r0 + k < nrows) hist[(size_t)(r0 + k) * nbins + b] = run;
            run += v[k];
        }
    }
    counts[b] = run;
}

template <class T>
__global__ void __launch_bounds__(kSortBlock) k_sort_scatter(Soa<T> in, const int32_t *pid_in, Soa<T> out,
                                                             int32_t *pid_out, const uint32_t *keys,
                                                             const uint32_t *hist, const int32_t *start,
                                                             int32_t nbins, uint32_t np, uint32_t chunk,
                                                             const unsigned int *gate) {
    if (gated_off(gate)) return;
    extern __shared__ uint32_t cur[];
    const uint32_t *row = hist + (size_t)blockIdx.x * nbins;
    for (int32_t b = threadIdx.x; b < nbins; b += kSortBlock) cur[b] = (uint32_t)start[b] + row[b];
    __syncthreads();
    const uint32_t p0 = blockIdx.x * chunk, p1 = min(p0 + chunk, np);
    // kUnroll parcels per thread per round, every load issued before the first store
    // (two for doubles, which keeps two 1024-thread workgroups per CU)
    constexpr int kUnroll = sizeof(T) == 4 ? 4 : 2;
    for (uint32_t pb = p0 + threadIdx.x; pb < p1; pb += kUnroll * kSortBlock) {
        uint32_t q[kUnroll];
        T d[kUnroll], x[kUnroll], y[kUnroll], z[kUnroll], u[kUnroll], v[kUnroll], w[kUnroll], r[kUnroll];
        int32_t id[kUnroll];
        bool on[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) {
            const uint32_t p = pb + k * kSortBlock;
            on[k] = p < p1;
            const uint32_t pc = on[k] ? p : p0;
            q[k] = keys[pc];
            d[k] = in.d[pc];
            x[k] = in.x[pc];
            y[k] = in.y[pc];
            z[k] = in.z[pc];
            u[k] = in.u[pc];
            v[k] = in.v[pc];
            w[k] = in.w[pc];
            r[k] = in.rho ? in.rho[pc] : (T)0;
            id[k] = pid_in[pc];
        }
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) {
            if (!on[k]) continue;
            const uint32_t o = atomicAdd(&cur[q[k]], 1u);
            out.d[o] = d[k];
            out.x[o] = x[k];
            out.y[o] = y[k];
            out.z[o] = z[k];
            out.u[o] = u[k];
            out.v[o] = v[k];
            out.w[o] = w[k];
            if (in.rho) ((T *)out.rho)[o] = r[k];
            pid_out[o] = id[k];
        }
    }
}

template <class T>
hipError_t launch_sort_chunked(const Soa<T> &in, const int32_t *pid_in, Soa<T> out, int32_t *pid_out, uint32_t np,
                               Geo g, const int32_t nb[3], const int32_t bd[3], int periodic, int32_t nrows,
                               uint32_t *keys, uint32_t *hist, uint32_t *counts, int32_t *start,
                               uint32_t *cursor, const unsigned int *gate, hipStream_t s) {
    const int32_t nbins = nb[0] * nb[1] * nb[2];
    const uint32_t chunk = (np + (uint32_t)nrows - 1) / (uint32_t)nrows;
    const size_t lds = (size_t)nbins * 4;
    if (periodic)
        k_sort_count<T, true><<<nrows, kSortBlock, lds, s>>>(in, np, chunk, g, nb[0], nb[1], bd[0], bd[1], bd[2],
                                                             nbins, keys, hist, gate);
    else
        k_sort_count<T, false><<<nrows, kSortBlock, lds, s>>>(in, np, chunk, g, nb[0], nb[1], bd[0], bd[1], bd[2],
                                                              nbins, keys, hist, gate);
    k_col_prefix<<<grid_for((size_t)nbins, 256), 256, 0, s>>>(hist, nbins, nrows, counts, gate);
    k_scan<<<1, 1024, 0, s>>>(counts, nbins, start, cursor, gate);
    if (np) k_sort_scatter<T><<<nrows, kSortBlock, lds, s>>>(in, pid_in, out, pid_out, keys, hist, start, nbins, np,
                                                             chunk, gate);
    return hipGetLastError();
}
template hipError_t launch_sort_chunked<float>(const Soa<float> &, const int32_t *, Soa<float>, int32_t *, uint32_t,
                                               Geo, const int32_t[3], const int32_t[3], int, int32_t, uint32_t *,
                                               uint32_t *, uint32_t *, int32_t *, uint32_t *,
                                               const unsigned int *, hipStream_t);
template hipError_t launch_sort_chunked<double>(const Soa<double> &, const int32_t *, Soa<double>, int32_t *,
                                                uint32_t, Geo, const int32_t[3], const int32_t[3], int, int32_t,
                                                uint32_t *, uint32_t *, uint32_t *, int32_t *, uint32_t *,
                                                const unsigned int *, hipStream_t);


template <class T>
__global__ void k_scatter(Soa<T> in, const int32_t *pid_in, Soa<T> out, int32_t *pid_out,
                          const uint32_t *keys, uint32_t *cursor, uint32_t np, const unsigned int *gate) {
    if (gated_off(gate)) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = p < np;
    const uint32_t key = active ? keys[p] : 0u;
    const uint32_t q = wave_claim(cursor, key, active);
    if (!active) return;
    out.d[q] = in.d[p];
    out.x[q] = in.x[p];
    out.y[q] = in.y[p];
    out.z[q] = in.z[p];
    out.u[q] = in.u[p];
    out.v[q] = in.v[p];
    out.w[q] = in.w[p];
    if (in.rho) ((T *)out.rho)[q] = in.rho[p];
    pid_out[q] = pid_in[p];
}

template <class T>
hipError_t launch_scatter(const Soa<T> &in, const int32_t *pid_in, Soa<T> out, int32_t *pid_out,
                          const uint32_t *keys, uint32_t *cursor, uint32_t np, const unsigned int *gate,
                          hipStream_t s) {
    if (!np) return hipSuccess;
    k_scatter<T><<<(np + 255) / 256, 256, 0, s>>>(in, pid_in, out, pid_out, keys, cursor, np, gate);
    return hipGetLastError();
}
template hipError_t launch_scatter<float>(const Soa<float> &, const int32_t *, Soa<float>, int32_t *,
                                          const uint32_t *, uint32_t *, uint32_t, const unsigned int *,
                                          hipStream_t);
template hipError_t launch_scatter<double>(const Soa<double> &, const int32_t *, Soa<double>,
                                           int32_t *, const uint32_t *, uint32_t *, uint32_t,
                                           const unsigned int *, hipStream_t);

// The sorted set copied back over the cloud's arrays, four elements per thread per
// round with every load issued before the first store.
template <class T>
__global__ void __launch_bounds__(256) k_copy_back(Soa<T> from, const int32_t *pid_from, Soa<T> to, int32_t *pid_to,
                                                   uint32_t np, const unsigned int *gate, unsigned int *count) {
    if (gated_off(gate)) return;
    if (gate && count && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(count, 1u);
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool rho = from.rho != nullptr;
    constexpr int kU = 4;
    for (uint32_t p0 = blockIdx.x * blockDim.x + threadIdx.x; p0 < np; p0 += kU * stride) {
        T d[kU], x[kU], y[kU], z[kU], u[kU], v[kU], w[kU], r[kU];
        int32_t id[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const uint32_t p = min(p0 + k * stride, np - 1);
            d[k] = from.d[p];
            x[k] = from.x[p];
            y[k] = from.y[p];
            z[k] = from.z[p];
            u[k] = from.u[p];
            v[k] = from.v[p];
            w[k] = from.w[p];
            r[k] = rho ? from.rho[p] : (T)0;
            id[k] = pid_from[p];
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const uint32_t p = p0 + k * stride;
            if (p >= np) continue;
            to.d[p] = d[k];
            to.x[p] = x[k];
            to.y[p] = y[k];
            to.z[p] = z[k];
            to.u[p] = u[k];
            to.v[p] = v[k];
            to.w[p] = w[k];
            if (rho) ((T *)to.rho)[p] = r[k];
            pid_to[p] = id[k];
        }
    }
}

template <class T>
hipError_t launch_copy_back(const Soa<T> &from, const int32_t *pid_from, Soa<T> to, int32_t *pid_to, uint32_t np,
                            const unsigned int *gate, unsigned int *count, hipStream_t s) {
    if (!np) return hipSuccess;
    const unsigned grid = grid_for(((size_t)np + 3) / 4, 256);
    k_copy_back<T><<<grid, 256, 0, s>>>(from, pid_from, to, pid_to, np, gate, count);
    return hipGetLastError();
}
template hipError_t launch_copy_back<float>(const Soa<float> &, const int32_t *, Soa<float>, int32_t *, uint32_t,
                                            const unsigned int *, unsigned int *, hipStream_t);
template hipError_t launch_copy_back<double>(const Soa<double> &, const int32_t *, Soa<double>, int32_t *,
                                             uint32_t, const unsigned int *, unsigned int *, hipStream_t);

template <class T>
__global__ void k_unpermute(Soa<T> in, const int32_t *pid, Soa<T> out, uint32_t np) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint32_t q = (uint32_t)pid[p];
    out.d[q] = in.d[p];
    out.x[q] = in.x[p];
    out.y[q] = in.y[p];
    out.z[q] = in.z[p];
    out.u[q] = in.u[p];
    out.v[q] = in.v[p];
    out.w[q] = in.w[p];
}

template <class T>
hipError_t launch_unpermute(const Soa<T> &in, const int32_t *pid, Soa<T> out, uint32_t np,
                            hipStream_t s) {
    if (!np) return hipSuccess;
    k_unpermute<T><<<(np + 255) / 256, 256, 0, s>>>(in, pid, out, np);
    return hipGetLastError();
}
template hipError_t launch_unpermute<float>(const Soa<float> &, const int32_t *, Soa<float>, uint32_t,
                                            hipStream_t);
template hipError_t launch_unpermute<double>(const Soa<double> &, const int32_t *, Soa<double>,
                                             uint32_t, hipStream_t);

// ---------------------------------------------------------------------------------
// Injection and compaction of a cloud.

__global__ void k_iota_from(int32_t *p, uint32_t n, int32_t base) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = base + (int32_t)i;
}

hipError_t launch_iota_from(int32_t *p, uint32_t n, int32_t base, hipStream_t s) {
    if (!n) return hipSuccess;
    k_iota_from<<<(n + 255) / 256, 256, 0, s>>>(p, n, base);
    return hipGetLastError();
}

// live_dev[p] = 1 for a parcel not removed (d > 0) in device order; live_up[pid[p]] the
// same in upload order.
template <class T>
__global__ void k_live_flags(const T *d, const int32_t *pid, uint32_t np, uint32_t *live_dev, uint32_t *live_up) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint32_t on = d[p] > (T)0 ? 1u : 0u;
    live_dev[p] = on;
    live_up[(uint32_t)pid[p]] = on;
}

// Exclusive scan of n uint32 in tiles of kScanTile: tile sums, one scan of the sums
// (k_scan), then each tile's own scan from its offset.
constexpr int kScanPer = 16;
constexpr int kScanTile = 256 * kScanPer;

__global__ void __launch_bounds__(256) k_tile_sums(const uint32_t *in, uint32_t n, uint32_t *sums) {
    __shared__ uint32_t red[256 / kWave];
    const uint32_t t0 = blockIdx.x * (uint32_t)kScanTile;
    uint32_t v = 0;
    for (int k = 0; k < kScanPer; ++k) {
        const uint32_t i = t0 + k * 256u + threadIdx.x;
        if (i < n) v += in[i];
    }
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(256) k_tile_scan(const uint32_t *in, uint32_t n, const int32_t *offsets,
                                                   uint32_t *out) {
    __shared__ uint32_t part[256];
    const uint32_t base = blockIdx.x * (uint32_t)kScanTile + threadIdx.x * kScanPer;  // a contiguous run per thread
    uint32_t v[kScanPer], sum = 0;
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        sum += v[k];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t add = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = (uint32_t)offsets[blockIdx.x] + part[threadIdx.x] - sum;
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// out[i] = in[0] + ... + in[i-1]; offsets[t] receives each tile's start and
// offsets[ntiles] the total (tile_sums holds ntiles = ceil(n / kScanTile) entries).
hipError_t launch_exclusive_scan(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *tile_sums,
                                 int32_t *offsets, uint32_t *cursor, hipStream_t s) {
    const uint32_t tiles = n ? (n + kScanTile - 1) / kScanTile : 1;
    if (n) k_tile_sums<<<tiles, 256, 0, s>>>(in, n, tile_sums);
    else k_fill_u32<<<1, 256, 0, s>>>(tile_sums, 0u, 1, nullptr);
    k_scan<<<1, 1024, 0, s>>>(tile_sums, (int32_t)tiles, offsets, cursor, nullptr);
    if (n) k_tile_scan<<<tiles, 256, 0, s>>>(in, n, offsets, out);
    return hipGetLastError();
}

uint32_t scan_tiles(uint32_t n) { return n ? (n + kScanTile - 1) / kScanTile : 1; }

// Live parcels moved to their place in a compacted order (device order kept), with
// their upload-order index renumbered among the live ones.
template <class T>
__global__ void k_compact_scatter(Soa<T> in, const int32_t *pid_in, Soa<T> out, int32_t *pid_out,
                                  const uint32_t *pos_dev, const uint32_t *new_up, uint32_t np) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || !(in.d[p] > (T)0)) return;
    const uint32_t q = pos_dev[p];
    out.d[q] = in.d[p];
    out.x[q] = in.x[p];
    out.y[q] = in.y[p];
    out.z[q] = in.z[p];
    out.u[q] = in.u[p];
    out.v[q] = in.v[p];
    out.w[q] = in.w[p];
    if (in.rho) ((T *)out.rho)[q] = in.rho[p];
    pid_out[q] = (int32_t)new_up[(uint32_t)pid_in[p]];
}

template <class T>
hipError_t launch_compact(const Soa<T> &in, const int32_t *pid_in, Soa<T> out, int32_t *pid_out, uint32_t np,
                          uint32_t *live_dev, uint32_t *live_up, uint32_t *pos_dev, uint32_t *new_up,
                          uint32_t *tile_sums, int32_t *offsets_dev, int32_t *offsets_up, uint32_t *cursor,
                          hipStream_t s) {
    if (!np) return hipSuccess;
    const unsigned g = (np + 255) / 256;
    k_live_flags<T><<<g, 256, 0, s>>>(in.d, pid_in, np, live_dev, live_up);
    hipError_t e = launch_exclusive_scan(live_dev, np, pos_dev, tile_sums, offsets_dev, cursor, s);
    if (e == hipSuccess) e = launch_exclusive_scan(live_up, np, new_up, tile_sums, offsets_up, cursor, s);
    if (e != hipSuccess) return e;
    k_compact_scatter<T><<<g, 256, 0, s>>>(in, pid_in, out, pid_out, pos_dev, new_up, np);
    return hipGetLastError();
}
template hipError_t launch_compact<float>(const Soa<float> &, const int32_t *, Soa<float>, int32_t *, uint32_t,
                                          uint32_t *, uint32_t *, uint32_t *, uint32_t *, uint32_t *, int32_t *,
                                          int32_t *, uint32_t *, hipStream_t);
template hipError_t launch_compact<double>(const Soa<double> &, const int32_t *, Soa<double>, int32_t *, uint32_t,
                                           uint32_t *, uint32_t *, uint32_t *, uint32_t *, uint32_t *, int32_t *,
                                           int32_t *, uint32_t *, hipStream_t);

// Count parcels whose position is outside [origin, ending) on some axis, from parcel
// p0 on (removed parcels, d <= 0, are not checked).
template <class T>
__global__ void k_range_check(Soa<T> s, uint32_t p0, uint32_t np, Geo g, unsigned int *bad) {
    const uint32_t p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || !(s.d[p] > (T)0)) return;
    const int32_t i = axis_floor((double)s.x[p], g.o[0], g.rd[0]);
    const int32_t j = axis_floor((double)s.y[p], g.o[1], g.rd[1]);
    const int32_t k = axis_floor((double)s.z[p], g.o[2], g.rd[2]);
    if ((uint32_t)i >= (uint32_t)g.n[0] || (uint32_t)j >= (uint32_t)g.n[1] ||
        (uint32_t)k >= (uint32_t)g.n[2])
        atomicAdd(bad, 1u);
}

template <class T>
hipError_t launch_range_check(const Soa<T> &s, uint32_t p0, uint32_t np, Geo g, unsigned int *bad,
                              hipStream_t st) {
    if (np <= p0) return hipSuccess;
    k_range_check<T><<<(np - p0 + 255) / 256, 256, 0, st>>>(s, p0, np, g, bad);
    return hipGetLastError();
}
template hipError_t launch_range_check<float>(const Soa<float> &, uint32_t, uint32_t, Geo, unsigned int *,
                                              hipStream_t);
template hipError_t launch_range_check<double>(const Soa<double> &, uint32_t, uint32_t, Geo, unsigned int *,
                                               hipStream_t);

// Streaming copy (bandwidth calibration): each thread moves 4 independent 16-byte
// vectors per iteration, so four loads are in flight before the first store.
__global__ void __launch_bounds__(256) k_copy(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const uint4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        b[i] = v0;
        b[i + stride] = v1;
        b[i + 2 * stride] = v2;
        b[i + 3 * stride] = v3;
    }
    for (; i < n; i += stride) b[i] = a[i];
}

hipError_t launch_copy(const void *a, void *b, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    if (!n) return hipSuccess;
    size_t g = (n + 1023) / 1024;
    if (g > 256 * 16) g = 256 * 16;
    k_copy<<<(unsigned)g, 256, 0, s>>>((const uint4 *)a, (uint4 *)b, n);
    return hipGetLastError();
}

}  // namespace st
