/*
 * st_oracle.h -- CPU restatement of SCALE-TRACK's Lagrangian parcel step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library and its
 * host layer) links, loads or calls this code.  Only tests/, the smoke check of
 * __graft_entry__.py and the cpu_baseline leg of bench.py use it, and only as
 * the checker or the timed CPU baseline.
 *
 * Algorithm source (reference, Julia):
 *   run/icoJuliaParcelFoam/cavity3D/asyncParallelTracking.jl
 *     :172-173   scalar/label aliases (restated with Float64 for the oracle)
 *     :222-270   Mesh / construct_mesh
 *     :663-697   locate_ijk / locate
 *     :699-736   set_parcel_state, update! (CPU "+=" form, :703-708), update_at_boundary!
 *     :1257-1365 evolve_on_device! / evolve_particle!
 *     :738-747   estimate_source! (ConstExtrapolator)
 *   run/icoJuliaParcelFoam/cavity3D/syncSerialTracking.jl:312-323 (serial loop shape)
 *
 * Pinning status of each exported function (how a plausible mistake would be caught):
 *   st_oracle_locate        pinned: brute force over every cell of tiny meshes, exact
 *                           face/overflow/underflow cases (tests/test_oracle_locate.py)
 *   st_oracle_evolve        pinned: closed-form drag/advance series, f->0 and f->inf
 *                           limits, hand-derived wall/edge/corner bounces, momentum
 *                           invariant, linearity over parcels, pass-splitting telescoping,
 *                           periodic wrap cases (tests/test_oracle_*.py)
 *   channel boundary        pinned: hand cases of exits through the outlet and the inlet,
 *                           at a wall, on the first and last substeps, and of removed
 *                           parcels left alone; equals the wall form for parcels that
 *                           never reach the open planes; momentum balance with exits
 *                           (tests/test_oracle_channel.py)
 *   st_oracle_extrapolate   pinned: closed form est_n = 2 S_{n-1} - est_{n-1}
 *   st_oracle_adjoint_*     pinned: central finite differences with frozen events and
 *                           forward-mode dual numbers of the same templated step
 */
#ifndef ST_ORACLE_H
#define ST_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Precision forms.  ST_O_FP64: every quantity in double (the reference with its
 * `scalar` alias set to Float64).  ST_O_FP32: parcel state, fluid velocity and drag
 * in float; face test (x-o)*rDelta and the momentum source in double. */
enum { ST_O_FP64 = 0, ST_O_FP32 = 1 };
/* Boundary forms: wall = the reference's specular bounce; periodic = extension;
 * channel = extension: walls on y and z, the x planes at origin and ending open.  A
 * substep whose new position lies outside them (tested every substep, before the bounce
 * gate) ends the parcel: its state is kept as advanced, its residence deposit
 * m (s0 - v') goes into the cell it was in, its diameter is stored negated, and later
 * calls skip it (any parcel with d <= 0). */
enum { ST_O_WALL = 0, ST_O_PERIODIC = 1, ST_O_CHANNEL = 2 };

typedef struct st_oracle_args {
    int32_t precision;          /* ST_O_FP64 / ST_O_FP32 */
    int32_t boundary;           /* ST_O_WALL / ST_O_PERIODIC */
    int32_t n[3];               /* cells per direction */
    double origin[3];
    double ending[3];
    int64_t np;                 /* parcels */
    /* parcel SoA, element type double (FP64) or float (FP32); updated in place */
    void *d, *x, *y, *z, *u, *v, *w;
    const void *rho;            /* optional per-parcel density (same element type), or NULL */
    double mu_c;                /* carrier dynamic viscosity */
    double rho_d;               /* chunk parcel density (used when rho == NULL) */
    double rho_d_by_rho_c;      /* chunk density ratio (used when rho == NULL) */
    double rho_c;               /* carrier density (used when rho != NULL) */
    double cd;                  /* drag-coefficient multiplier, 1 = reference drag */
    const void *U;              /* cell velocity, AoS xyz, element type as the state */
    double *source;             /* cell momentum source, AoS xyz, accumulated (+=) */
    double dt_fluid;            /* fluid time step; parcel substep = dt_fluid / nsub */
    int32_t nsub;               /* substeps per call (the reference's nSteps = 10) */
    int32_t substeps_per_pass;  /* substeps fused per pass (nsub = reference schedule) */
    int32_t nthreads;           /* 1 = the reference's serial loop; >1 OpenMP */
    double bbox[6];             /* out: min xyz, max xyz of the new positions */
    /* Frozen events (gradient checks): np * nsub words, parcel-major.  events_mode
     * 1 records each substep's cell, bounce and wrap decisions; 2 replays them instead
     * of locating, so a perturbed run follows the recorded event sequence. */
    int64_t *events;
    int32_t events_mode;
    /* Forward-mode dual numbers (FP64 only; active when tan_source != NULL): tangent
     * seeds of the per-parcel diameter and density and of cd; tan_state (6*np, SoA
     * x,y,z,u,v,w, may be NULL) carries state tangents in and out; the source tangent
     * is added into tan_source (ncell*3). */
    const double *tan_d;
    const double *tan_rho;
    double tan_cd;
    double *tan_source;
    double *tan_state;
} st_oracle_args;

int st_oracle_evolve(st_oracle_args *a);

/* Cell location of one point, FP64: global indices, 0-based folded linear cell index
 * and 0-based partition index for decomposition (1,1,1). */
int st_oracle_locate(const int32_t n[3], const double origin[3], const double ending[3],
                     double x, double y, double z, int32_t ijk[3], int64_t *cell,
                     int64_t *partition);

/* Constant extrapolator, in place: cur <- 2 cur - prev; prev <- cur. */
void st_oracle_extrapolate(double *cur, double *prev, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
