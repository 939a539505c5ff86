// st_device.h -- device functions shared by the forward and adjoint kernels: cell
// location, wall fold, substep kinematics and per-parcel constants.  Both kernels call
// the same code, so the adjoint's forward recompute takes exactly the forward's
// decisions (cells, bounces, wraps).
#pragma once

#include "st_internal.h"

namespace st {

constexpr int kWave = 64;

// Boundary forms of the kernels (the st_boundary values): the reference's walls; the
// periodic extension; and the channel extension -- walls on y and z, while the x planes
// at origin and ending are open: a parcel whose new position lies outside them leaves
// the domain (kExit below) and is marked removed by a negative diameter.
constexpr int kWall = 0, kPeriodic = 1, kChannel = 2;
// Bit of advance()'s result: the substep carried the parcel out through an open plane.
constexpr unsigned kExit = 8u;

__device__ __forceinline__ int32_t axis_floor(double x, double o, double rd) {
    return (int32_t)floor((x - o) * rd);
}

// Which side of the mesh an axis coordinate lies on: -1 below, +1 at or above the end,
// 0 inside.  floor(t) < 0 exactly when t < 0, and floor(t) >= n exactly when t >= n
// (n an integer), so this takes the decisions axis_floor's index would take, with
// the floor and the integer conversion left out.
__device__ __forceinline__ int axis_side(double x, double o, double rd, int32_t n) {
    const double t = (x - o) * rd;
    return t < 0.0 ? -1 : (t >= (double)n ? 1 : 0);
}

// divrem by the partition size (= N for one partition) with truncation toward zero,
// then abs() of both parts (asyncParallelTracking.jl:671-696).
__device__ __forceinline__ void fold(int32_t I, int32_t N, int32_t &f, int32_t &p) {
    if ((uint32_t)I < (uint32_t)N) {
        f = I;
        p = 0;
    } else {
        const int32_t q = I / N;
        const int32_t r = I - q * N;
        f = r < 0 ? -r : r;
        p = q < 0 ? -q : q;
    }
}

// i clamped to [0, n) (n >= 1): the median of i, 0 and n - 1, one v_med3_i32.
__device__ __forceinline__ int32_t clampi(int32_t i, int32_t n) {
    return min(max(i, 0), n - 1);
}

struct Cell {
    int32_t f[3];  // folded per-axis indices, always in [0, N)
    uint32_t c;    // linear cell, x fastest
    int32_t part;  // folded partition index (wall mode)
};

// Linear cell of folded indices; kernels that need it only off their fast path call
// this there instead of reading Cell::c (two quarter-rate integer multiplies).
__device__ __forceinline__ uint32_t lin(const Geo &g, const Cell &cl) {
    return (uint32_t)cl.f[2] * g.nxny + (uint32_t)cl.f[1] * (uint32_t)g.n[0] + (uint32_t)cl.f[0];
}

// Same folded cell: the linear index is one-to-one on folded indices, so comparing the
// three axes decides what comparing Cell::c does.
__device__ __forceinline__ bool same_cell(const Cell &a, const Cell &b) {
    return (a.f[0] == b.f[0]) & (a.f[1] == b.f[1]) & (a.f[2] == b.f[2]);  // one mask, no branches
}

__device__ __forceinline__ void cell_wall(const Geo &g, const int32_t I[3], Cell &cl) {
    if (((uint32_t)I[0] < (uint32_t)g.n[0]) & ((uint32_t)I[1] < (uint32_t)g.n[1]) &
        ((uint32_t)I[2] < (uint32_t)g.n[2])) {  // inside the mesh: one branch for the three axes
        cl.f[0] = I[0];
        cl.f[1] = I[1];
        cl.f[2] = I[2];
        cl.part = 0;
    } else {
        int32_t p0, p1, p2;
        fold(I[0], g.n[0], cl.f[0], p0);
        fold(I[1], g.n[1], cl.f[1], p1);
        fold(I[2], g.n[2], cl.f[2], p2);
        cl.part = p0 + p1 + p2;
    }
    cl.c = lin(g, cl);
}

__device__ __forceinline__ void cell_clamped(const Geo &g, const int32_t I[3], Cell &cl) {
    cl.f[0] = clampi(I[0], g.n[0]);
    cl.f[1] = clampi(I[1], g.n[1]);
    cl.f[2] = clampi(I[2], g.n[2]);
    cl.part = 0;
    cl.c = lin(g, cl);
}

template <class T, bool PERIODIC>
__device__ __forceinline__ void locate(const Geo &g, T x, T y, T z, int32_t I[3], Cell &cl) {
    I[0] = axis_floor((double)x, g.o[0], g.rd[0]);
    I[1] = axis_floor((double)y, g.o[1], g.rd[1]);
    I[2] = axis_floor((double)z, g.o[2], g.rd[2]);
    if (PERIODIC)
        cell_clamped(g, I, cl);
    else
        cell_wall(g, I, cl);
}

__device__ __forceinline__ void atomic_add_global(double *p, double v) {
    unsafeAtomicAdd(p, v);  // global_atomic_add_f64 (coarse-grained hipMalloc memory)
}
__device__ __forceinline__ void atomic_add_lds(__attribute__((address_space(3))) double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The persistent state of one parcel as read at the start of a pass.
template <class T>
struct ParcelIn {
    T d, x, y, z, u, v, w, rho;
};

// Element access through a 32-bit byte offset from a workgroup-uniform base: the loads
// and stores then use the scalar-base addressing form and need no 64-bit address
// arithmetic per array.  Every parcel array is below 4 GiB (checked at cloud creation).
template <class T>
__device__ __forceinline__ T ld_off(const T *base, uint32_t off) {
    return *(const T *)((const char *)base + off);
}
template <class T>
__device__ __forceinline__ void st_off(T *base, uint32_t off, T v) {
    *(T *)((char *)base + off) = v;
}

// RHO: the cloud carries a per-parcel density array.  It is a template parameter, not
// a test of s.rho, so that no load in the hot loops is conditional: the compiler's
// wait-count placement treats a register that a load may or may not have written as
// pending and drains every load in flight before its first use.
template <class T, bool RHO>
__device__ __forceinline__ void load_parcel(const Soa<T> &s, uint32_t p, T rho_chunk, ParcelIn<T> &in) {
    const uint32_t off = p * (uint32_t)sizeof(T);
    in.d = ld_off(s.d, off);
    in.x = ld_off(s.x, off);
    in.y = ld_off(s.y, off);
    in.z = ld_off(s.z, off);
    in.u = ld_off(s.u, off);
    in.v = ld_off(s.v, off);
    in.w = ld_off(s.w, off);
    in.rho = RHO ? ld_off(s.rho, off) : rho_chunk;
}

// Division by a parcel's implicit-Euler denominator 1 + f (:1295), correctly rounded in
// both precision forms.  Three velocity components per substep share the denominator.
template <class T>
struct Quot;
// For doubles, with r = RN(1 / den) and q0 = RN(a r) (within one ulp of a / den), the
// residual t = q0 den - a is exact as an FMA and RN(q0 - t r) is RN(a / den)
// (Markstein's theorem: r within half an ulp of 1 / den, q0 within one ulp, no under-
// or overflow).  Written this way a zero numerator keeps its sign (t = +0, and
// -0 + q0 = q0 for either zero).  A nonzero quotient below 2^-969 (where the residual
// could underflow), or one above 2^1000 or NaN, sends all three components to the IEEE
// division.  Per substep: 3 multiplies and 6 FMAs in place of three division sequences.
template <>
struct Quot<double> {
    double den, r;
    __device__ __forceinline__ explicit Quot(double d) : den(d), r(1.0 / d) {}
    __device__ __forceinline__ static bool tiny(double q) { return (fabs(q) < 0x1p-969) & (q != 0.0); }
    __device__ __forceinline__ void three(double ax, double ay, double az, double &rx, double &ry,
                                          double &rz) const {
        const double qx = ax * r, qy = ay * r, qz = az * r;
        rx = fma(-fma(qx, den, -ax), r, qx);
        ry = fma(-fma(qy, den, -ay), r, qy);
        rz = fma(-fma(qz, den, -az), r, qz);
        const bool huge = !(fmax(fmax(fabs(qx), fabs(qy)), fabs(qz)) <= 0x1p1000);
        if ((int)tiny(qx) | (int)tiny(qy) | (int)tiny(qz) | (int)huge) {  // one mask, no branches
            rx = ax / den;
            ry = ay / den;
            rz = az / den;
        }
    }
};
// For floats, a / den == (float)((double)a * RN64(1 / den)) whenever the quotient is
// zero or outside the FP32 subnormal range.  The double product is within 2^-52
// (relative) of a / den.  With a = A 2^p and den = D 2^q (A, D odd, below 2^24) and
// a float rounding midpoint m = M 2^r (M odd, 25 bits), A 2^p - M D 2^(q+r) is a
// nonzero multiple of 2^min(p, q+r) unless a / den == m, which would need the odd A to
// equal M D >= M > 2^24; so a / den is never a midpoint and never within 2^-49
// (relative) of one, and the product rounds to the same float.  Midpoints of the
// subnormal range (M may be small there) take the IEEE float division.  One double
// reciprocal per parcel then replaces three float divisions per substep.
template <>
struct Quot<float> {
    float den;
    double r;
    __device__ __forceinline__ explicit Quot(float d) : den(d), r(1.0 / (double)d) {}
    // The three components of a substep with one test: when the smallest rounded
    // product is at least 2^-126 every product is correct (a product that rounds up to
    // 2^-126 lies in [2^-127, 2^-126), where a quotient can sit on a midpoint only if
    // den is a power of two, and then the product is exact).  Otherwise each component
    // whose rounded product is below 2^-126 and whose numerator is nonzero takes the
    // float division (a zero numerator gives the same signed zero either way).
    __device__ __forceinline__ void three(float ax, float ay, float az, float &rx, float &ry, float &rz) const {
        rx = (float)((double)ax * r);
        ry = (float)((double)ay * r);
        rz = (float)((double)az * r);
        if (fminf(fminf(fabsf(rx), fabsf(ry)), fabsf(rz)) < 0x1p-126f) {
            if (fabsf(rx) < 0x1p-126f && ax != 0.0f) rx = ax / den;
            if (fabsf(ry) < 0x1p-126f && ay != 0.0f) ry = ay / den;
            if (fabsf(rz) < 0x1p-126f && az != 0.0f) rz = az / den;
        }
    }
};

// Kinematics of one substep (asyncParallelTracking.jl:1294-1335): implicit-Euler
// drag update and advance, relocation, then the specular wall bounce -- only when the
// folded cell or partition changed -- or the periodic wrap.  On return (n*) is the new
// state, nc the cell after relocation, and the result has bit a set for each axis a
// that bounced (the caller applies the source correction s0_a -= 2 v_a of :734).  In
// the channel form a new position outside the open x planes (tested on every substep,
// before the reference's bounce gate) returns kExit alone: no bounce, the state as
// advanced.
template <class T, int BND>
__device__ __forceinline__ unsigned advance(const Geo &g, const Phys<T> &ph, T f, const Quot<T> &den, T ux, T uy,
                                            T uz, T px, T py, T pz, T vx, T vy, T vz, const Cell &cell,
                                            int32_t part0, T &nx, T &ny, T &nz, T &nvx, T &nvy, T &nvz,
                                            Cell &nc) {
    // :1295-1296
    den.three(vx + f * ux, vy + f * uy, vz + f * uz, nvx, nvy, nvz);
    nx = px + nvx * ph.dt;
    ny = py + nvy * ph.dt;
    nz = pz + nvz * ph.dt;
    int32_t I[3];
    unsigned bounced = 0;
    if (BND != kPeriodic) {
        locate<T, false>(g, nx, ny, nz, I, nc);  // :1303
        if (BND == kChannel && (uint32_t)I[0] >= (uint32_t)g.n[0]) return kExit;
        if (!same_cell(nc, cell) || nc.part != part0) {  // :1306
            if (I[0] < 0 || I[0] >= g.n[0]) {  // :1309-1329
                nx = px;
                nvx = -vx;
                bounced |= 1u;
            }
            if (I[1] < 0 || I[1] >= g.n[1]) {
                ny = py;
                nvy = -vy;
                bounced |= 2u;
            }
            if (I[2] < 0 || I[2] >= g.n[2]) {
                nz = pz;
                nvz = -vz;
                bounced |= 4u;
            }
            if (bounced) locate<T, false>(g, nx, ny, nz, I, nc);  // :1333-1335
        }
    } else {
        // periodic extension: wrap each axis that left [0, N) by one period (one
        // unsigned test per axis and one branch for the three on the common path)
        I[0] = axis_floor((double)nx, g.o[0], g.rd[0]);
        I[1] = axis_floor((double)ny, g.o[1], g.rd[1]);
        I[2] = axis_floor((double)nz, g.o[2], g.rd[2]);
        if (((uint32_t)I[0] >= (uint32_t)g.n[0]) | ((uint32_t)I[1] >= (uint32_t)g.n[1]) |
            ((uint32_t)I[2] >= (uint32_t)g.n[2])) {
            if (I[0] < 0) {
                nx = nx + ph.L[0];
                I[0] = axis_floor((double)nx, g.o[0], g.rd[0]);
            } else if (I[0] >= g.n[0]) {
                nx = nx - ph.L[0];
                I[0] = axis_floor((double)nx, g.o[0], g.rd[0]);
            }
            if (I[1] < 0) {
                ny = ny + ph.L[1];
                I[1] = axis_floor((double)ny, g.o[1], g.rd[1]);
            } else if (I[1] >= g.n[1]) {
                ny = ny - ph.L[1];
                I[1] = axis_floor((double)ny, g.o[1], g.rd[1]);
            }
            if (I[2] < 0) {
                nz = nz + ph.L[2];
                I[2] = axis_floor((double)nz, g.o[2], g.rd[2]);
            } else if (I[2] >= g.n[2]) {
                nz = nz - ph.L[2];
                I[2] = axis_floor((double)nz, g.o[2], g.rd[2]);
            }
            // a wrapped coordinate that rounds onto the far face is clamped, as a
            // fresh location would be
            I[0] = clampi(I[0], g.n[0]);
            I[1] = clampi(I[1], g.n[1]);
            I[2] = clampi(I[2], g.n[2]);
        }
        nc.f[0] = I[0];  // in [0, N) on both paths: the clamp of cell_clamped is a no-op
        nc.f[1] = I[1];
        nc.f[2] = I[2];
        nc.part = 0;
        nc.c = lin(g, nc);
    }
    return bounced;
}

// Per-parcel constants (:1280 and :1294): mass over carrier density m (double in both
// precision forms), drag factor f = cd * ((18 mu) dt) / (rho d^2) and 1 + f.  f does
// not depend on position or velocity, so one value serves every substep.
template <class T>
__device__ __forceinline__ void parcel_consts(const Phys<T> &ph, bool per_parcel_rho, T dia, T rho,
                                              double &m, T &f, T &den) {
    const double rdr = per_parcel_rho ? (double)rho / ph.rho_c : ph.rdr;
    const double dd = (double)dia;
    m = ((rdr * kPi) * ((dd * dd) * dd)) / 6.0;
    f = (((T)18 * ph.mu) * ph.dt) / (rho * (dia * dia));
    f = ph.cd * f;
    den = (T)1 + f;
}

}  // namespace st
