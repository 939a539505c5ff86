/*
 * st_oracle.cpp -- CPU restatement of SCALE-TRACK's Lagrangian parcel step.
 *
 * TEST INFRASTRUCTURE ONLY (see st_oracle.h): the checker that the HIP library is
 * compared against, and the CPU baseline that bench.py times.  Never linked into
 * the product.
 *
 * Written as C++17 (built with g++, -ffp-contract=off) only so that one templated
 * step can be instantiated for double, for float state with double face test and
 * source, and for forward-mode dual numbers (the gradient check).  Operation order
 * follows the reference's Julia expressions literally, because Julia evaluates
 * left to right and never contracts a*b+c into an FMA without @fastmath/@muladd:
 *   m    = ((rho_d/rho_c * pi) * d^3) / 6          asyncParallelTracking.jl:1280
 *   f    = ((18 * mu) * dt) / (rho * (d*d))        asyncParallelTracking.jl:1294
 *   v'   = (v + f*u) / (1 + f)                     asyncParallelTracking.jl:1295
 *   x'   = x + v' * dt                             asyncParallelTracking.jl:1296
 * d^3 and d^2 are Julia literal_pow, i.e. (d*d)*d and d*d.
 */
#include "st_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPi = 3.14159265358979323846; /* Float64(pi), correctly rounded */

/* ---- forward-mode dual number (value, one tangent) ------------------------- */
struct Dual {
    double v, g;
    Dual() : v(0), g(0) {}
    Dual(double x) : v(x), g(0) {}  // NOLINT: implicit from constants
    Dual(double x, double t) : v(x), g(t) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.g + b.g}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.g - b.g}; }
inline Dual operator-(Dual a) { return {-a.v, -a.g}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.g * b.v + a.v * b.g}; }
inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return {q, (a.g - q * b.g) / b.v};
}

inline double val(double x) { return x; }
inline double val(float x) { return (double)x; }
inline double val(const Dual &x) { return x.v; }

/* Source / face-test type of each state type: double for float and double, Dual for Dual. */
template <class S> struct Wide { using type = double; };
template <> struct Wide<Dual> { using type = Dual; };
/* Storage type of the cell velocity U for each state type. */
template <class S> struct Store { using type = S; };
template <> struct Store<Dual> { using type = double; };

template <class S> inline typename Wide<S>::type widen(S x) { return (typename Wide<S>::type)x; }

/* Uniform box mesh, asyncParallelTracking.jl:222-270 (construct_mesh):
 * L = ending - origin, Delta = L / N, rDelta = 1 / Delta, all in the face-test
 * precision (double here).  Decomposition (1,1,1): partitionN = N. */
struct MeshD {
    int32_t n[3];
    int64_t nxny;
    double o[3], L[3], delta[3], rdelta[3];
    explicit MeshD(const int32_t nn[3], const double org[3], const double end[3]) {
        for (int a = 0; a < 3; ++a) {
            n[a] = nn[a];
            o[a] = org[a];
            L[a] = end[a] - org[a];
            delta[a] = L[a] / (double)nn[a];
            rdelta[a] = 1.0 / delta[a];
        }
        nxny = (int64_t)n[0] * n[1];
    }
};

/* locate_ijk, asyncParallelTracking.jl:663-677: floor((x - o) * rDelta) per axis
 * (IEEE floor, never truncation -- the reference's comment at :664-666 explains
 * why unsafe_trunc is wrong for -1 < n < 0). */
inline int32_t axis_index(double xa, double o, double rd) { return (int32_t)std::floor((xa - o) * rd); }

struct Loc {
    int32_t ijk[3];
    int64_t cell; /* 0-based folded linear index (the reference's posI - 1) */
    int64_t part; /* 0-based folded partition index (partitionI - 1) */
};

/* locate, asyncParallelTracking.jl:679-697, for decomposition (1,1,1):
 * divrem(I, N) truncates toward zero (as C's / and %); the linear indices fold the
 * local and partition indices with abs(), x fastest. */
inline Loc locate_wall(const MeshD &m, double x, double y, double z) {
    Loc l;
    const double p[3] = {x, y, z};
    int64_t loc[3], par[3];
    for (int a = 0; a < 3; ++a) {
        l.ijk[a] = axis_index(p[a], m.o[a], m.rdelta[a]);
        par[a] = l.ijk[a] / m.n[a];
        loc[a] = l.ijk[a] % m.n[a];
    }
    l.cell = std::llabs(loc[2]) * m.nxny + std::llabs(loc[1]) * m.n[0] + std::llabs(loc[0]);
    /* decompositionXTimesY = decomposition.x = 1 */
    l.part = std::llabs(par[2]) + std::llabs(par[1]) + std::llabs(par[0]);
    return l;
}

/* Periodic extension: index clamped into [0, N) (a wrapped coordinate can round onto
 * the far face `ending`, which belongs to the last cell). */
inline int32_t clamp_index(int32_t i, int32_t n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

inline int64_t linear(const MeshD &m, const int32_t ijk[3]) {
    return (int64_t)ijk[2] * m.nxny + (int64_t)ijk[1] * m.n[0] + ijk[0];
}

/* Event word of one substep, for recording and replaying the piecewise-constant
 * decisions (frozen-event mode of the gradient checks):
 * bits 0-31 cell after the substep, 32-34 bounce per axis, 35-40 wrap per axis
 * (2 bits: 1 = +L, 2 = -L). */
inline int64_t ev_pack(int64_t cell, const int bounce[3], const int wrap[3]) {
    int64_t e = cell & 0xffffffffLL;
    for (int c = 0; c < 3; ++c) {
        if (bounce[c]) e |= (int64_t)1 << (32 + c);
        e |= (int64_t)wrap[c] << (35 + 2 * c);
    }
    return e;
}

template <class S>
struct Vec3 {
    S c[3];
};

template <class S>
struct Consts {
    S mu, rho_chunk, dt, cd, one, eighteen;
    S L[3];
};

template <class S>
using DW = typename Wide<S>::type;

/* The step of one parcel for nsteps substeps: asyncParallelTracking.jl:1271-1365.
 * Returns true when the parcel left through an open x plane (channel extension). */
template <class S>
bool evolve_particle(const st_oracle_args &a, const MeshD &m, const Consts<S> &k, S dia, S rho,
                     DW<S> rdr, Vec3<S> &pos, Vec3<S> &vel, const typename Store<S>::type *Uc,
                     int nsteps, DW<S> *src, int64_t *ev) {
    using D = DW<S>;
    const bool periodic = a.boundary == ST_O_PERIODIC;
    const int evmode = ev ? a.events_mode : 0;
    /* :1280, in the source precision */
    const D dd = widen(dia);
    const D mByRho = ((rdr * D(kPi)) * ((dd * dd) * dd)) / D(6.0);

    Vec3<S> posNew = pos, velNew = vel;

    /* :1286-1289 */
    int64_t posI, partI;
    if (!periodic) {
        Loc l = locate_wall(m, val(pos.c[0]), val(pos.c[1]), val(pos.c[2]));
        posI = l.cell;
        partI = l.part;
    } else {
        int32_t ijk[3];
        for (int ax = 0; ax < 3; ++ax)
            ijk[ax] = clamp_index(axis_index(val(pos.c[ax]), m.o[ax], m.rdelta[ax]), m.n[ax]);
        posI = linear(m, ijk);
        partI = 0;
    }
    Vec3<S> uc{{S(Uc[3 * posI]), S(Uc[3 * posI + 1]), S(Uc[3 * posI + 2])}};
    D s0[3] = {widen(vel.c[0]), widen(vel.c[1]), widen(vel.c[2])};
    bool exited = false;

    for (int t = 1; t <= nsteps; ++t) {
        /* :1294-1296 implicit Euler with Stokes drag */
        S f = ((k.eighteen * k.mu) * k.dt) / (rho * (dia * dia));
        f = k.cd * f; /* drag-coefficient multiplier (extension); exact for cd = 1 */
        const S denom = k.one + f;
        for (int c = 0; c < 3; ++c) velNew.c[c] = (vel.c[c] + f * uc.c[c]) / denom;
        for (int c = 0; c < 3; ++c) posNew.c[c] = pos.c[c] + velNew.c[c] * k.dt;

        int bounce[3] = {0, 0, 0}, wrap[3] = {0, 0, 0};
        int64_t cellNew = posI;
        if (evmode == 2) {
            /* replay the recorded decisions */
            const int64_t e = ev[t - 1];
            cellNew = e & 0xffffffffLL;
            for (int c = 0; c < 3; ++c) {
                bounce[c] = (int)((e >> (32 + c)) & 1);
                wrap[c] = (int)((e >> (35 + 2 * c)) & 3);
                if (bounce[c]) {
                    posNew.c[c] = pos.c[c];
                    velNew.c[c] = -vel.c[c];
                    s0[c] = s0[c] - D(2.0) * widen(vel.c[c]);
                }
                if (wrap[c] == 1) posNew.c[c] = posNew.c[c] + k.L[c];
                if (wrap[c] == 2) posNew.c[c] = posNew.c[c] - k.L[c];
            }
        } else if (!periodic) {
            /* :1303 */
            Loc l = locate_wall(m, val(posNew.c[0]), val(posNew.c[1]), val(posNew.c[2]));
            if (a.boundary == ST_O_CHANNEL && (l.ijk[0] < 0 || l.ijk[0] >= m.n[0])) {
                /* channel extension: the new position lies outside an open x plane,
                 * tested every substep before the bounce gate.  The parcel leaves with
                 * its state as advanced (no bounce); the deposit after the loop ends
                 * its residence. */
                vel = velNew;
                pos = posNew;
                exited = true;
                break;
            }
            /* :1306 */
            if (l.cell != posI || l.part != partI) {
                bool hit = false;
                for (int c = 0; c < 3; ++c) {
                    /* :1309-1311 */
                    if ((l.ijk[c] < 0) || (l.ijk[c] >= m.n[c])) {
                        /* :1313-1329 and update_at_boundary!, :730-736 */
                        posNew.c[c] = pos.c[c];
                        velNew.c[c] = -vel.c[c];
                        s0[c] = s0[c] - D(2.0) * widen(vel.c[c]);
                        bounce[c] = 1;
                        hit = true;
                    }
                }
                /* :1333-1335 */
                if (hit) l = locate_wall(m, val(posNew.c[0]), val(posNew.c[1]), val(posNew.c[2]));
                /* :1337-1340 re-reads the old partition; a no-op for one partition,
                 * and partitionI is never updated inside the loop. */
            }
            cellNew = l.cell;
        } else {
            /* periodic extension: wrap each axis that left [0, N) by one period */
            int32_t ijk[3];
            for (int c = 0; c < 3; ++c) {
                int32_t r = axis_index(val(posNew.c[c]), m.o[c], m.rdelta[c]);
                if (r < 0) {
                    posNew.c[c] = posNew.c[c] + k.L[c];
                    r = axis_index(val(posNew.c[c]), m.o[c], m.rdelta[c]);
                    wrap[c] = 1;
                } else if (r >= m.n[c]) {
                    posNew.c[c] = posNew.c[c] - k.L[c];
                    r = axis_index(val(posNew.c[c]), m.o[c], m.rdelta[c]);
                    wrap[c] = 2;
                }
                ijk[c] = clamp_index(r, m.n[c]);
            }
            cellNew = linear(m, ijk);
        }
        if (evmode == 1) ev[t - 1] = ev_pack(cellNew, bounce, wrap);
        /* :1342-1348 (the wall form only reaches here with a changed folded index) */
        if (t != nsteps && cellNew != posI) {
            for (int c = 0; c < 3; ++c) src[3 * posI + c] = src[3 * posI + c] + mByRho * (s0[c] - widen(velNew.c[c]));
            for (int c = 0; c < 3; ++c) s0[c] = widen(velNew.c[c]);
            for (int c = 0; c < 3; ++c) uc.c[c] = S(Uc[3 * cellNew + c]);
            posI = cellNew;
        }
        /* :1351-1352 */
        vel = velNew;
        pos = posNew;
    }
    /* :1361 */
    for (int c = 0; c < 3; ++c) src[3 * posI + c] = src[3 * posI + c] + mByRho * (s0[c] - widen(velNew.c[c]));
    return exited;
}

/* :1362 bounding box of the positions a call ends with (after all its passes) */
template <class S>
void bbox_add(double bb[6], const Vec3<S> &pos) {
    for (int c = 0; c < 3; ++c) {
        const double p = val(pos.c[c]);
        bb[c] = std::min(bb[c], p);
        bb[3 + c] = std::max(bb[3 + c], p);
    }
}

template <class S>
Consts<S> make_consts(const st_oracle_args *a) {
    Consts<S> k;
    k.mu = (S)a->mu_c;
    k.rho_chunk = (S)a->rho_d;
    /* Time stores dt as `scalar` (:272-275); the substep is dt / nSteps (:1261). */
    k.dt = (S)a->dt_fluid / (S)a->nsub;
    k.cd = (S)a->cd;
    k.one = (S)1;
    k.eighteen = (S)18;
    for (int c = 0; c < 3; ++c) k.L[c] = (S)(a->ending[c] - a->origin[c]);
    return k;
}

void bbox_init(double bb[6]) {
    for (int c = 0; c < 3; ++c) {
        bb[c] = std::numeric_limits<double>::infinity();
        bb[3 + c] = -std::numeric_limits<double>::infinity();
    }
}

/* One parcel over all passes of the call, float or double state. */
template <class S>
void run_parcel(st_oracle_args *a, const MeshD &m, const Consts<S> &k, int64_t i, double *src, double bb[6]) {
    S *d = (S *)a->d;
    S *X = (S *)a->x, *Y = (S *)a->y, *Z = (S *)a->z;
    S *U = (S *)a->u, *V = (S *)a->v, *W = (S *)a->w;
    const S dia = d[i];
    if (!(dia > (S)0)) return; /* removed by an open boundary earlier: never touched again */
    S rho = k.rho_chunk;
    double rdr = a->rho_d_by_rho_c;
    if (a->rho) {
        rho = ((const S *)a->rho)[i];
        rdr = (double)rho / a->rho_c; /* per-parcel density: extension */
    }
    Vec3<S> pos{{X[i], Y[i], Z[i]}}, vel{{U[i], V[i], W[i]}};
    const int per = a->substeps_per_pass > 0 ? a->substeps_per_pass : a->nsub;
    bool exited = false;
    for (int done = 0; done < a->nsub && !exited; done += per) {
        const int ns = std::min(per, a->nsub - done);
        int64_t *ev = a->events ? a->events + i * (int64_t)a->nsub + done : nullptr;
        exited = evolve_particle<S>(*a, m, k, dia, rho, rdr, pos, vel, (const S *)a->U, ns, src, ev);
    }
    if (exited)
        d[i] = -dia; /* marks the parcel removed; the bounding box holds parcels still in */
    else
        bbox_add(bb, pos);
    /* :1355-1360 */
    X[i] = pos.c[0];
    Y[i] = pos.c[1];
    Z[i] = pos.c[2];
    U[i] = vel.c[0];
    V[i] = vel.c[1];
    W[i] = vel.c[2];
}

template <class S>
int evolve_all(st_oracle_args *a) {
    MeshD m(a->n, a->origin, a->ending);
    const int64_t ncell = (int64_t)a->n[0] * a->n[1] * a->n[2];
    const Consts<S> k = make_consts<S>(a);
    double bb[6];
    bbox_init(bb);
    int nthreads = a->nthreads > 0 ? a->nthreads : 1;
#ifndef _OPENMP
    nthreads = 1;
#endif
    if (nthreads == 1) {
        /* syncSerialTracking.jl:318-320 loop shape, with the "+=" deposit of
         * asyncParallelTracking.jl:706.  Passes are an outer loop per parcel: each
         * parcel's trajectory is independent, and the source is a sum. */
        for (int64_t i = 0; i < a->np; ++i) run_parcel<S>(a, m, k, i, a->source, bb);
    } else {
#ifdef _OPENMP
        std::vector<std::vector<double>> priv(nthreads);
        std::vector<std::array<double, 6>> pbb(nthreads);
#pragma omp parallel num_threads(nthreads)
        {
            const int tid = omp_get_thread_num();
            std::vector<double> &s = priv[tid];
            s.assign(3 * ncell, 0.0);
            double b[6];
            bbox_init(b);
#pragma omp for schedule(static)
            for (int64_t i = 0; i < a->np; ++i) run_parcel<S>(a, m, k, i, s.data(), b);
            std::memcpy(pbb[tid].data(), b, sizeof b);
        }
        for (int t = 0; t < nthreads; ++t) {
            const std::vector<double> &s = priv[t];
            if (s.empty()) continue;
            for (int64_t j = 0; j < 3 * ncell; ++j) a->source[j] += s[j];
            for (int c = 0; c < 3; ++c) {
                bb[c] = std::min(bb[c], pbb[t][c]);
                bb[3 + c] = std::max(bb[3 + c], pbb[t][3 + c]);
            }
        }
#endif
    }
    std::memcpy(a->bbox, bb, sizeof bb);
    return 0;
}

/* Dual-number run (FP64 values): tangents seeded by tan_d, tan_rho, tan_cd and the
 * incoming state tangents tan_state; returns the source tangent and the state
 * tangents. */
int evolve_dual(st_oracle_args *a) {
    MeshD m(a->n, a->origin, a->ending);
    const int64_t ncell = (int64_t)a->n[0] * a->n[1] * a->n[2];
    Consts<Dual> k = make_consts<Dual>(a);
    k.cd = Dual(a->cd, a->tan_cd);
    std::vector<Dual> src(3 * ncell);
    for (int64_t j = 0; j < 3 * ncell; ++j) src[j] = Dual(a->source[j], 0.0);
    double bb[6];
    bbox_init(bb);
    double *d = (double *)a->d;
    double *X = (double *)a->x, *Y = (double *)a->y, *Z = (double *)a->z;
    double *U = (double *)a->u, *V = (double *)a->v, *W = (double *)a->w;
    double *ts = a->tan_state;
    const int64_t n = a->np;
    const int per = a->substeps_per_pass > 0 ? a->substeps_per_pass : a->nsub;
    for (int64_t i = 0; i < n; ++i) {
        const Dual dia(d[i], a->tan_d ? a->tan_d[i] : 0.0);
        Dual rho = k.rho_chunk;
        Dual rdr(a->rho_d_by_rho_c);
        if (a->rho) {
            rho = Dual(((const double *)a->rho)[i], a->tan_rho ? a->tan_rho[i] : 0.0);
            rdr = rho / Dual(a->rho_c);
        }
        Vec3<Dual> pos{{Dual(X[i], ts ? ts[i] : 0), Dual(Y[i], ts ? ts[n + i] : 0), Dual(Z[i], ts ? ts[2 * n + i] : 0)}};
        Vec3<Dual> vel{{Dual(U[i], ts ? ts[3 * n + i] : 0), Dual(V[i], ts ? ts[4 * n + i] : 0),
                        Dual(W[i], ts ? ts[5 * n + i] : 0)}};
        for (int done = 0; done < a->nsub; done += per) {
            const int ns = std::min(per, a->nsub - done);
            int64_t *ev = a->events ? a->events + i * (int64_t)a->nsub + done : nullptr;
            evolve_particle<Dual>(*a, m, k, dia, rho, rdr, pos, vel, (const double *)a->U, ns, src.data(), ev);
        }
        bbox_add(bb, pos);
        X[i] = pos.c[0].v;
        Y[i] = pos.c[1].v;
        Z[i] = pos.c[2].v;
        U[i] = vel.c[0].v;
        V[i] = vel.c[1].v;
        W[i] = vel.c[2].v;
        if (ts) {
            for (int c = 0; c < 3; ++c) {
                ts[c * n + i] = pos.c[c].g;
                ts[(3 + c) * n + i] = vel.c[c].g;
            }
        }
    }
    for (int64_t j = 0; j < 3 * ncell; ++j) {
        a->source[j] = src[j].v;
        a->tan_source[j] += src[j].g;
    }
    std::memcpy(a->bbox, bb, sizeof bb);
    return 0;
}

}  // namespace

extern "C" int st_oracle_evolve(st_oracle_args *a) {
    if (!a || a->nsub <= 0 || a->np < 0) return -1;
    for (int c = 0; c < 3; ++c)
        if (a->n[c] <= 0 || !(a->ending[c] > a->origin[c])) return -1;
    if (a->events_mode != 0 && !a->events) return -1;
    if (a->boundary != ST_O_WALL && a->boundary != ST_O_PERIODIC && a->boundary != ST_O_CHANNEL) return -1;
    if (a->boundary == ST_O_CHANNEL && a->events_mode != 0) return -1;
    if (a->tan_source) {
        if (a->precision != ST_O_FP64 || a->boundary == ST_O_CHANNEL) return -1;
        return evolve_dual(a);
    }
    if (a->events && a->nthreads > 1) return -1;
    if (a->precision == ST_O_FP64) return evolve_all<double>(a);
    if (a->precision == ST_O_FP32) return evolve_all<float>(a);
    return -1;
}

extern "C" int st_oracle_locate(const int32_t n[3], const double origin[3], const double ending[3], double x,
                                double y, double z, int32_t ijk[3], int64_t *cell, int64_t *partition) {
    MeshD m(n, origin, ending);
    Loc l = locate_wall(m, x, y, z);
    for (int c = 0; c < 3; ++c) ijk[c] = l.ijk[c];
    *cell = l.cell;
    *partition = l.part;
    return 0;
}

/* estimate_source!, asyncParallelTracking.jl:738-747 */
extern "C" void st_oracle_extrapolate(double *cur, double *prev, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        cur[i] = 2.0 * cur[i] - prev[i];
        prev[i] = cur[i];
    }
}
