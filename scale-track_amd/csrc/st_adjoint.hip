// st_adjoint.hip -- reverse (adjoint) kernels of the parcel step, FP64.
//
// The loss L depends on the parcels through the deposited source S (and, for an
// endpoint loss, the final positions and velocities).  Cells, bounces and wraps are
// piecewise constant in the parameters, so the derivative is the almost-everywhere one
// with the event sequence the forward run took.  Per substep and axis a, with
// v' = (v + f u) / (1 + f), deposit D_a = m (v_a - v'_a) into the starting cell c,
// g = dL/dS and adjoints vb (of v'), xb (of x, constant through the substeps):
//   not bounced:  w = vb_a + xb_a dt - m g_a;   vb_a <- w / (1 + f) + m g_a
//                 fbar += w (u_a - v'_a) / (1 + f);   mbar += g_a (v_a - v'_a)
//   bounced:      v'_a = -v_a and D_a = 0, so vb_a <- -vb_a.
// The deposit the reference makes per residence (asyncParallelTracking.jl:1342-1361,
// with the bounce correction of :730-736) is the sum of these per-substep deposits,
// so this is the derivative of exactly the forward's loss.
//
// Checkpoints: the forward saves (x, v) at the start of every segment of K substeps.
// One launch per segment, last segment first.  Inside a segment each thread
// recomputes its parcel forward from the checkpoint and keeps B substeps at a time in
// registers (nested recompute: segment blocks of B are replayed last block first).
#include "st_device.h"

#include <algorithm>

namespace st {

struct AdjArgs {
    const double *d, *rho;  // per-parcel diameter, density (rho may be null)
    const double *ck;       // checkpoint of this segment: x, y, z, u, v, w (SoA, n each)
    const double *U;        // cell velocity, AoS
    const double *G;        // dL/dS, AoS
    const double *UG;       // per cell U xyz then G xyz (48 B, 16-byte aligned)
    const double *xb;       // adjoint of x, y, z (SoA, 3n; constant through the substeps), or
                            // nullptr for zero
    const double *vb_in;    // adjoint of u, v, w the segment starts from (SoA, 3n), or nullptr
                            // for zero
    double *vb;             // adjoint of u, v, w the segment ends with (may equal vb_in)
    double *fbar, *mbar;    // accumulated adjoints of f and m per parcel
    int first;              // the sweep's first segment: fbar and mbar start at zero
    Geo g;
    Phys<double> ph;
    int32_t K;              // substeps in this segment
    uint32_t np;
};

__device__ __forceinline__ void load_ug(const double *ug, uint32_t c, double2 &r0, double2 &r1, double2 &r2) {
    const double2 *q = (const double2 *)(ug + 6u * c);
    r0 = q[0];
    r1 = q[1];
    r2 = q[2];
}

// Both instances run at 3 waves per SIMD; the kernel is bound by latency, not issue.
// The periodic one fits 166 VGPRs without spills when the compiler is allowed more (a
// cap of 168 makes it spill 36 B); the wall one is held to 168 VGPRs and spills 84 B,
// which measured faster than 2 waves per SIMD without spills.
//
// Workgroups are dispatched to the eight XCDs in turn, and each XCD has its own L2.
// Workgroup w therefore takes parcel block (w mod 8) * (grid / 8) + w / 8, so each XCD
// steps one contiguous eighth of the brick-sorted parcels: its gathers of U and G stay
// within one region of the mesh, and its L2 holds that region.
template <bool PERIODIC, int B>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PERIODIC ? 2 : 3)))
k_adjoint_segment(AdjArgs a) {
    uint32_t w = blockIdx.x;
    const uint32_t full = gridDim.x & ~7u;
    if (w < full) w = (w & 7u) * (full >> 3) + (w >> 3);
    const uint32_t p = w * blockDim.x + threadIdx.x;
    if (p >= a.np) return;
    const uint32_t n = a.np;
    const double dia = a.d[p];
    const double rho = a.rho ? a.rho[p] : a.ph.rho;
    double m, f, den;
    parcel_consts<double>(a.ph, a.rho != nullptr, dia, rho, m, f, den);
    const double dt = a.ph.dt;
    // The reverse sweep multiplies by 1 / (1 + f) instead of dividing (its results are
    // compared with the derivative to rounding, not bit for bit); the replayed forward
    // keeps the forward's exact divisions, so it takes the forward's decisions.
    const double rden = 1.0 / den;
    const Quot<double> qden(den);

    double xb0 = 0.0, xb1 = 0.0, xb2 = 0.0, vb0 = 0.0, vb1 = 0.0, vb2 = 0.0;
    if (a.xb) {
        xb0 = a.xb[p];
        xb1 = a.xb[n + p];
        xb2 = a.xb[2 * n + p];
    }
    if (a.vb_in) {
        vb0 = a.vb_in[p];
        vb1 = a.vb_in[n + p];
        vb2 = a.vb_in[2 * n + p];
    }
    double fb = 0.0, mb = 0.0;  // fb without its factor 1 / (1 + f), applied at the end

    const double cx = a.ck[p], cy = a.ck[n + p], cz = a.ck[2 * n + p];
    const double cu = a.ck[3 * n + p], cv = a.ck[4 * n + p], cw = a.ck[5 * n + p];

    const int32_t nblk = (a.K + B - 1) / B;
    for (int32_t blk = nblk - 1; blk >= 0; --blk) {
        // replay from the checkpoint to the start of this block (as the forward pass of
        // K substeps did: cell and velocity refreshed on each change before substep K)
        double px = cx, py = cy, pz = cz, vx = cu, vy = cv, vz = cw;
        int32_t I[3];
        Cell cell;
        locate<double, PERIODIC>(a.g, px, py, pz, I, cell);
        const int32_t part0 = cell.part;
        const double *q = a.U + 3u * cell.c;
        double ux = q[0], uy = q[1], uz = q[2];
        // Before each run of substeps (periodic form), the velocity of the cell the
        // parcel would reach by drifting at its cell's velocity for the whole run is
        // requested when that cell differs from the current one: a change into that cell
        // then reads these registers instead of waiting for a load.  (In the wall form
        // the extra registers cost more spills than the waits they save.)
        uint32_t pc = 0xffffffffu;
        double pu0 = 0.0, pu1 = 0.0, pu2 = 0.0;
        auto prefetch_ahead = [&](int32_t steps) {
            const double h = (double)steps * dt;
            int32_t J[3];
            J[0] = axis_floor(px + h * ux, a.g.o[0], a.g.rd[0]);
            J[1] = axis_floor(py + h * uy, a.g.o[1], a.g.rd[1]);
            J[2] = axis_floor(pz + h * uz, a.g.o[2], a.g.rd[2]);
            Cell pcl;
            cell_clamped(a.g, J, pcl);
            if (pcl.c != cell.c) {
                pc = pcl.c;
                const double *qp = a.U + 3u * pc;
                pu0 = qp[0];
                pu1 = qp[1];
                pu2 = qp[2];
            }
        };
        auto enter = [&](const Cell &nc) {
            cell = nc;
            if (nc.c == pc) {
                ux = pu0;
                uy = pu1;
                uz = pu2;
            } else {
                q = a.U + 3u * cell.c;
                ux = q[0];
                uy = q[1];
                uz = q[2];
            }
        };
        int32_t t = 1;
        const int32_t t0 = blk * B;
        if (PERIODIC && t0 > 0) prefetch_ahead(t0);
        for (; t <= t0; ++t) {
            double nx, ny, nz, nvx, nvy, nvz;
            Cell nc;
            advance<double, PERIODIC>(a.g, a.ph, f, qden, ux, uy, uz, px, py, pz, vx, vy, vz, cell, part0, nx,
                                      ny, nz, nvx, nvy, nvz, nc);
            if (t != a.K && nc.c != cell.c) enter(nc);
            px = nx; py = ny; pz = nz;
            vx = nvx; vy = nvy; vz = nvz;
        }
        // record this block
        const int32_t nb = min(B, a.K - t0);
        double rv[B][3];
        uint32_t rc[B];
        unsigned rb = 0;
        if (PERIODIC) prefetch_ahead(nb);
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (j < nb) {
                rv[j][0] = vx;
                rv[j][1] = vy;
                rv[j][2] = vz;
                rc[j] = cell.c;
                double nx, ny, nz, nvx, nvy, nvz;
                Cell nc;
                const unsigned bnc = advance<double, PERIODIC>(a.g, a.ph, f, qden, ux, uy, uz, px, py, pz, vx, vy,
                                                               vz, cell, part0, nx, ny, nz, nvx, nvy, nvz, nc);
                rb |= bnc << (3 * j);
                if (t != a.K && nc.c != cell.c) enter(nc);
                px = nx; py = ny; pz = nz;
                vx = nvx; vy = nvy; vz = nvz;
                ++t;
            }
        }
        // reverse this block.  A parcel keeps its cell for many substeps, so each
        // substep's U and G (one 48-byte record per cell) are read only when the cell
        // differs from the later substep's.  The records of the block's last and first
        // cells are requested together before the sweep; only a parcel that visited a
        // third cell inside the block waits for a load in the sweep.
        double nvx = vx, nvy = vy, nvz = vz;  // velocity after the substep being reversed
        const uint32_t cl = rc[nb - 1], cf = rc[0];  // nb >= 1: every block holds a substep
        double2 r0, r1, r2, f0, f1, f2;
        load_ug(a.UG, cl, r0, r1, r2);
        load_ug(a.UG, cf, f0, f1, f2);
        uint32_t ccur = cl;
#pragma unroll
        for (int j = B - 1; j >= 0; --j) {
            if (j < nb) {
                if (rc[j] != ccur) {
                    ccur = rc[j];
                    if (ccur == cf) {
                        r0 = f0;
                        r1 = f1;
                        r2 = f2;
                    } else {
                        load_ug(a.UG, ccur, r0, r1, r2);
                    }
                }
                const double uu[3] = {r0.x, r0.y, r1.x};
                const double gg[3] = {r1.y, r2.x, r2.y};
                const unsigned bj = (rb >> (3 * j)) & 7u;
                const double pv[3] = {rv[j][0], rv[j][1], rv[j][2]};
                const double nv[3] = {nvx, nvy, nvz};
                double *vbp[3] = {&vb0, &vb1, &vb2};
                const double xbv[3] = {xb0, xb1, xb2};
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    if (bj & (1u << ax)) {
                        *vbp[ax] = -*vbp[ax];
                    } else {
                        const double ga = gg[ax];
                        const double w = *vbp[ax] + xbv[ax] * dt - m * ga;
                        *vbp[ax] = w * rden + m * ga;
                        fb += w * (uu[ax] - nv[ax]);
                        mb += ga * (pv[ax] - nv[ax]);
                    }
                }
                nvx = pv[0];
                nvy = pv[1];
                nvz = pv[2];
            }
        }
    }
    a.vb[p] = vb0;
    a.vb[n + p] = vb1;
    a.vb[2 * n + p] = vb2;
    a.fbar[p] = (a.first ? 0.0 : a.fbar[p]) + fb * rden;
    a.mbar[p] = (a.first ? 0.0 : a.mbar[p]) + mb;
}

struct GradArgs {
    const double *d, *rho;
    const double *fbar, *mbar;
    double *gd, *grho;  // per parcel (grho null with a chunk density)
    double *part;       // one partial of dL/dcd per workgroup
    Phys<double> ph;
    uint32_t np;
};

// d f/d d = -2 f / d, d f/d rho = -f / rho, d f/d cd = f / cd (= the factor without cd),
// d m/d d = 3 m / d, d m/d rho = m / rho (per-parcel density only).
__device__ __forceinline__ void block_partial(double v, double *part) {
    __shared__ double red[256 / kWave];
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// out = sum of part[0 .. n), n <= kSumBlocks, in a fixed order (one workgroup).
__global__ void __launch_bounds__(kSumBlocks) k_sum_partials(const double *part, int n, double *out) {
    __shared__ double red[kSumBlocks / kWave];
    double v = (int)threadIdx.x < n ? part[threadIdx.x] : 0.0;
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kSumBlocks / kWave; ++w) s += red[w];
        out[0] = s;
    }
}

__global__ void __launch_bounds__(256) k_adjoint_grads(GradArgs a) {
    double gc = 0.0;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < a.np; p += gridDim.x * blockDim.x) {
        const double dia = a.d[p];
        const double rho = a.rho ? a.rho[p] : a.ph.rho;
        double m, f, den;
        parcel_consts<double>(a.ph, a.rho != nullptr, dia, rho, m, f, den);
        const double fref = ((18.0 * a.ph.mu) * a.ph.dt) / (rho * (dia * dia));
        const double fb = a.fbar[p], mb = a.mbar[p];
        a.gd[p] = fb * (-2.0 * f / dia) + mb * (3.0 * m / dia);
        if (a.grho) a.grho[p] = fb * (-f / rho) + mb * (m / rho);
        gc += fb * fref;
    }
    block_partial(gc, a.part);
}

// G = 2 (S - T) and L += sum (S - T)^2 (the source-misfit loss).
__global__ void __launch_bounds__(256) k_misfit(const double *S, const double *T, double *G, double *part,
                                                size_t n) {
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double r = S[i] - T[i];
        G[i] = 2.0 * r;
        acc += r * r;
    }
    block_partial(acc, part);
}

// dst[k*n + p] = src[k][p] for the six state arrays (a checkpoint).
__global__ void k_save6(const double *x, const double *y, const double *z, const double *u, const double *v,
                        const double *w, double *dst, uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    dst[p] = x[p];
    dst[n + p] = y[p];
    dst[2 * n + p] = z[p];
    dst[3 * n + p] = u[p];
    dst[4 * n + p] = v[p];
    dst[5 * n + p] = w[p];
}

// Device order <- upload order, k arrays of n: dst[j*n + p] = src[j*n + pid[p]].
__global__ void k_gather_pid(const double *src, const int32_t *pid, double *dst, uint32_t n, int k) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t q = (uint32_t)pid[p];
    for (int j = 0; j < k; ++j) dst[(size_t)j * n + p] = src[(size_t)j * n + q];
}

// Upload order <- device order: dst[pid[p]] = src[p].
__global__ void k_scatter_pid(const double *src, const int32_t *pid, double *dst, uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    dst[pid[p]] = src[p];
}

__global__ void k_interleave3(const double *a, const double *b, double *dst, size_t n) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        double *q = dst + 6 * c;
        q[0] = a[3 * c];
        q[1] = a[3 * c + 1];
        q[2] = a[3 * c + 2];
        q[3] = b ? b[3 * c] : 0.0;  // b == nullptr: zero
        q[4] = b ? b[3 * c + 1] : 0.0;
        q[5] = b ? b[3 * c + 2] : 0.0;
    }
}

hipError_t launch_interleave3(const double *a, const double *b, double *dst, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    k_interleave3<<<(unsigned)blocks, 256, 0, s>>>(a, b, dst, n);
    return hipGetLastError();
}

hipError_t launch_adjoint_segment(const AdjArgsHost &h, hipStream_t s) {
    AdjArgs a;
    a.d = h.d;
    a.rho = h.rho;
    a.ck = h.ck;
    a.U = h.U;
    a.G = h.G;
    a.UG = h.UG;
    a.xb = h.xb;
    a.vb_in = h.vb_in;
    a.vb = h.vb;
    a.first = h.first ? 1 : 0;
    a.fbar = h.fbar;
    a.mbar = h.mbar;
    a.g = h.g;
    a.ph = h.ph;
    a.K = h.K;
    a.np = h.np;
    if (!a.np) return hipSuccess;
    const unsigned grid = (a.np + 255) / 256;
    if (h.periodic)
        k_adjoint_segment<true, 10><<<grid, 256, 0, s>>>(a);
    else
        k_adjoint_segment<false, 10><<<grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_adjoint_grads(const double *d, const double *rho, const double *fbar, const double *mbar,
                                double *gd, double *grho, double *gcd, double *part, const Phys<double> &ph,
                                uint32_t np, hipStream_t s) {
    const int grid = (int)std::min<uint32_t>((np + 255) / 256, (uint32_t)kSumBlocks);
    if (grid > 0) {
        GradArgs a{d, rho, fbar, mbar, gd, grho, part, ph, np};
        k_adjoint_grads<<<grid, 256, 0, s>>>(a);
    }
    k_sum_partials<<<1, kSumBlocks, 0, s>>>(part, grid, gcd);
    return hipGetLastError();
}

hipError_t launch_misfit(const double *S, const double *T, double *G, double *loss, double *part, size_t n,
                         hipStream_t s) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, (size_t)kSumBlocks);
    if (grid > 0) k_misfit<<<grid, 256, 0, s>>>(S, T, G, part, n);
    k_sum_partials<<<1, kSumBlocks, 0, s>>>(part, grid, loss);
    return hipGetLastError();
}

hipError_t launch_save6(const double *const f[6], double *dst, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    k_save6<<<(n + 255) / 256, 256, 0, s>>>(f[0], f[1], f[2], f[3], f[4], f[5], dst, n);
    return hipGetLastError();
}

hipError_t launch_gather_pid(const double *src, const int32_t *pid, double *dst, uint32_t n, int k,
                             hipStream_t s) {
    if (!n) return hipSuccess;
    k_gather_pid<<<(n + 255) / 256, 256, 0, s>>>(src, pid, dst, n, k);
    return hipGetLastError();
}

hipError_t launch_scatter_pid(const double *src, const int32_t *pid, double *dst, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    k_scatter_pid<<<(n + 255) / 256, 256, 0, s>>>(src, pid, dst, n);
    return hipGetLastError();
}

}  // namespace st
