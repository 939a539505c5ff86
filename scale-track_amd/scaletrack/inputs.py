"""Seeded synthetic inputs of the parcel step, generated once on the CPU in FP64.

The same arrays are handed to the device and to any CPU check, so no transcendental
function is ever evaluated twice on different hardware.  The generator is SplitMix64
(Steele, Lea, Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014)
in its counter form: output i of a stream with state s is mix(s + (i + 1) * GAMMA).
Julia's default RNG stream for the reference's seed 19891
(run/icoJuliaParcelFoam/cavity3D/asyncParallelTracking.jl:504-531) cannot be
reproduced without Julia, so the seed is kept and the generator is this one.
"""
import math

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
SEED = 19891  # the reference's seed (asyncParallelTracking.jl:514)


def splitmix64(state, n, offset=0):
    """n outputs of SplitMix64 started at `state`, skipping the first `offset`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(state) + i * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def stream_state(seed, stream):
    """Initial state of an independent stream: mix of (seed, stream)."""
    return int(splitmix64(np.uint64(seed) ^ np.uint64(stream * 0x632BE59BD9B4E019 % 2**64), 1)[0])


def uniform(seed, stream, n):
    """Uniform doubles in [0, 1) with 53 random bits."""
    z = splitmix64(stream_state(seed, stream), n)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, stream, n):
    """Standard normals by Box-Muller from two uniform streams."""
    u1 = 1.0 - uniform(seed, stream, n)  # (0, 1]
    u2 = uniform(seed, stream + 1, n)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def positions(seed, n, origin, ending, dtype=np.float64, stream=10):
    """Uniform positions in [origin, ending) per axis; after rounding to dtype every
    coordinate is still strictly below `ending`."""
    out = []
    for a in range(3):
        x = origin[a] + uniform(seed, stream + a, n) * (ending[a] - origin[a])
        x = x.astype(dtype)
        hi = np.asarray(ending[a], dtype=dtype)
        bad = x >= hi
        if bad.any():
            x[bad] = np.nextafter(hi, np.asarray(-np.inf, dtype=dtype))
        lo = np.asarray(origin[a], dtype=dtype)
        if float(lo) < origin[a]:  # origin rounded down to dtype
            lo = np.nextafter(lo, np.asarray(np.inf, dtype=dtype))
        x = np.maximum(x, lo)
        out.append(x)
    return out


def _ordered(x, it):
    """IEEE values of one width mapped to integers in the same order."""
    i = x.view(it)
    return np.where(i >= 0, i, np.iinfo(it).min - i)


def _from_ordered(k, it, dtype):
    return np.where(k >= 0, k, np.iinfo(it).min - k).astype(it).view(dtype)


def into_cells(x, origin, ending, ncell):
    """`x` with every coordinate outside the cells [0, ncell) of the library's locate,
    floor(((double)x - origin) * rdelta) with rdelta = 1 / ((ending - origin) / ncell)
    in double precision, replaced by the nearest value of its type inside them.  A
    coordinate within rounding of `ending` (or, after rounding to float, just below
    `origin`) can otherwise compute to cell ncell (or -1) although it lies inside
    [origin, ending); the library refuses such a parcel on upload.  The nearest value
    is found by bisection over the values of the type (the cell is monotone in x)."""
    x = np.array(x, copy=True)
    dt = x.dtype
    it = np.int32 if dt == np.float32 else np.int64
    rd = 1.0 / ((ending - origin) / ncell)
    delta = (ending - origin) / ncell

    def cell(v):
        return np.floor((v.astype(np.float64) - origin) * rd)

    for side in (1, -1):
        bad = np.flatnonzero(cell(x) >= ncell) if side > 0 else np.flatnonzero(cell(x) < 0)
        if bad.size == 0:
            continue
        centre = np.asarray(origin + (ncell - 0.5 if side > 0 else 0.5) * delta, dtype=dt)
        good = np.full(bad.size, _ordered(centre.reshape(1), it)[0])
        edge = _ordered(x[bad], it)
        lo, hi = (good, edge) if side > 0 else (edge, good)  # lo < hi in value
        while np.any(hi - lo > 1):
            mid = lo + (hi - lo) // 2
            out = cell(_from_ordered(mid, it, dt))
            past = out >= ncell if side > 0 else out < 0
            # keep the in-cell end on the good side, the outside end on the bad side
            if side > 0:
                hi = np.where(past, mid, hi)
                lo = np.where(past, lo, mid)
            else:
                lo = np.where(past, mid, lo)
                hi = np.where(past, hi, mid)
        x[bad] = _from_ordered(lo if side > 0 else hi, it, dt)
    c = cell(x)
    if np.any((c < 0) | (c >= ncell)):
        raise ValueError("coordinates could not be moved into the mesh's cells")
    return x


def lognormal_diameters(seed, n, median=30e-6, sigma=0.5, dtype=np.float64, stream=20):
    return (median * np.exp(sigma * normal(seed, stream, n))).astype(dtype)


def cell_centres(n, origin, ending):
    """Cell-centre coordinates, each of shape (ncell,), cells numbered x fastest."""
    axes = []
    for a in range(3):
        d = (ending[a] - origin[a]) / n[a]
        axes.append(origin[a] + (np.arange(n[a]) + 0.5) * d)
    zc, yc, xc = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return xc.ravel(), yc.ravel(), zc.ravel()


def taylor_green(n, origin, ending, u0=1.0, phase=(0.0, 0.0, 0.0), dtype=np.float64):
    """Frozen Taylor-Green velocity at cell centres, AoS (ncell, 3):
    u = U0 sin(kx) cos(ky) cos(kz), v = -U0 cos(kx) sin(ky) cos(kz), w = 0,
    k = 2 pi / L per axis."""
    x, y, z = cell_centres(n, origin, ending)
    kx = 2 * np.pi / (ending[0] - origin[0])
    ky = 2 * np.pi / (ending[1] - origin[1])
    kz = 2 * np.pi / (ending[2] - origin[2])
    ax, ay, az = kx * x + phase[0], ky * y + phase[1], kz * z + phase[2]
    U = np.empty((x.size, 3), dtype=np.float64)
    U[:, 0] = u0 * np.sin(ax) * np.cos(ay) * np.cos(az)
    U[:, 1] = -u0 * np.cos(ax) * np.sin(ay) * np.cos(az)
    U[:, 2] = 0.0
    return U.astype(dtype)


def fourier_field(n, origin, ending, modes=8, seed=SEED, amplitude=1.0, kmax=4, dtype=np.float64,
                  stream=40):
    """Divergence-free random field from `modes` Fourier modes Re(A_m exp(i k_m . x))
    with A_m perpendicular to k_m, evaluated at cell centres; AoS (ncell, 3)."""
    x, y, z = cell_centres(n, origin, ending)
    L = [ending[a] - origin[a] for a in range(3)]
    U = np.zeros((x.size, 3))
    r = uniform(seed, stream, modes * 9)
    for m in range(modes):
        q = r[9 * m: 9 * m + 9]
        kint = np.floor(q[0:3] * (2 * kmax + 1)) - kmax
        if not kint.any():
            kint[0] = 1.0
        k = 2 * np.pi * kint / np.asarray(L)
        a = q[3:6] - 0.5
        a = a - k * (a @ k) / (k @ k)  # project out the k component
        na = np.linalg.norm(a)
        if na == 0:
            continue
        a = a / na * amplitude / math.sqrt(modes)
        ph = 2 * np.pi * q[6]
        arg = k[0] * x + k[1] * y + k[2] * z + ph
        U += np.cos(arg)[:, None] * a[None, :]
    return U.astype(dtype)


def cavity_field(n, origin, ending, lid=1.0, dtype=np.float64):
    """A frozen lid-driven-cavity-like velocity for the 2D cavity case: the
    divergence-free field of the stream function psi = lid * Ly * f(sx) * g(sy), with
    sx, sy the coordinates scaled to [0, 1], f = sin^2(pi sx), g = sy^2 (sy - 1);
    u = dpsi/dy, v = -dpsi/dx, w = 0.  It vanishes on the walls and u = lid * f(sx)
    under the lid."""
    x, y, _ = cell_centres(n, origin, ending)
    Lx = ending[0] - origin[0]
    Ly = ending[1] - origin[1]
    sx = (x - origin[0]) / Lx
    sy = (y - origin[1]) / Ly
    f = np.sin(np.pi * sx) ** 2
    df = 2 * np.pi * np.sin(np.pi * sx) * np.cos(np.pi * sx)
    g = sy ** 2 * (sy - 1.0)
    dg = 3 * sy ** 2 - 2 * sy
    U = np.zeros((x.size, 3))
    U[:, 0] = lid * f * dg
    U[:, 1] = -lid * df * g * (Ly / Lx)
    return U.astype(dtype)


class Case:
    """One synthetic workload: mesh, frozen carrier velocity and a parcel cloud."""

    def __init__(self, n, origin, ending, U, d, x, y, z, u, v, w, rho=None, mu_c=1e-3,
                 rho_d=1000.0, rho_c=1000.0, dt_fluid=1e-3, nsub=10, boundary=0, name=""):
        self.n, self.origin, self.ending = tuple(n), tuple(origin), tuple(ending)
        self.U = U
        x, y, z = [into_cells(c, self.origin[a], self.ending[a], self.n[a]) for a, c in enumerate((x, y, z))]
        self.d, self.x, self.y, self.z, self.u, self.v, self.w = d, x, y, z, u, v, w
        self.rho = rho
        self.mu_c, self.rho_d, self.rho_c = mu_c, rho_d, rho_c
        self.rho_d_by_rho_c = rho_d / rho_c
        self.dt_fluid, self.nsub, self.boundary = dt_fluid, nsub, boundary
        self.name = name

    @property
    def np(self):
        return self.d.size

    @property
    def ncell(self):
        return self.n[0] * self.n[1] * self.n[2]

    def state(self):
        return [a.copy() for a in (self.d, self.x, self.y, self.z, self.u, self.v, self.w)]


def taylor_green_box(np_parcels=1 << 24, ncell=128, dtype=np.float64, seed=SEED, boundary=1,
                     dt_fluid=1e-3, nsub=10, name="taylor-green box"):
    """Periodic box [0,1]^3 of ncell^3 cells with frozen Taylor-Green flow (U0 = 1 m/s)
    and parcels at uniform positions, log-normal diameters (median 30 um, sigma 0.5),
    rho_p = 1000 kg/m^3, zero initial velocity; carrier mu = 1e-3, rho = 1e3 (the
    reference defaults of setup.jl:12-13)."""
    n = (ncell, ncell, ncell)
    o, e = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    U = taylor_green(n, o, e, dtype=dtype)
    x, y, z = positions(seed, np_parcels, o, e, dtype=dtype)
    d = lognormal_diameters(seed, np_parcels, dtype=dtype)
    zero = np.zeros(np_parcels, dtype=dtype)
    return Case(n, o, e, U, d, x, y, z, zero, zero.copy(), zero.copy(), mu_c=1e-3, rho_d=1000.0,
                rho_c=1000.0, dt_fluid=dt_fluid, nsub=nsub, boundary=boundary, name=name)


def cavity_case(np_parcels=10000, dtype=np.float64, seed=SEED, name="2D lid-driven cavity"):
    """20x20x1 cells on a 0.1 m square (0.01 m deep), lid 1 m/s, nu = 0.01 m^2/s;
    parcels d = 50 um, rho = 2500 kg/m^3 at uniform positions, zero velocity; walls;
    dt = 0.005 s.  The carrier field is frozen (no fluid solver here)."""
    n = (20, 20, 1)
    o, e = (0.0, 0.0, 0.0), (0.1, 0.1, 0.01)
    U = cavity_field(n, o, e, dtype=dtype)
    x, y, z = positions(seed, np_parcels, o, e, dtype=dtype)
    d = np.full(np_parcels, 50e-6, dtype=dtype)
    zero = np.zeros(np_parcels, dtype=dtype)
    rho_c = 1.0
    mu_c = 0.01 * rho_c  # nu = 0.01 m^2/s
    return Case(n, o, e, U, d, x, y, z, zero, zero.copy(), zero.copy(), mu_c=mu_c, rho_d=2500.0,
                rho_c=rho_c, dt_fluid=0.005, nsub=10, boundary=0, name=name)


def batch_cases(nclouds=64, seed=SEED, dtype=np.float32):
    """Many independent periodic clouds: parcel counts log-uniform in [1e3, 2e6], meshes
    n^3 with n uniform in {32..96}, Taylor-Green flow with seeded random phases per cloud
    (the batched workload)."""
    u = uniform(seed, 100, 5 * nclouds).reshape(nclouds, 5)
    counts = np.rint(np.exp(np.log(1e3) + u[:, 0] * (np.log(2e6) - np.log(1e3)))).astype(np.int64)
    cells = (32 + np.floor(u[:, 1] * 65)).astype(np.int64)
    cases = []
    for i in range(nclouds):
        c = taylor_green_box(np_parcels=int(counts[i]), ncell=int(cells[i]), dtype=dtype, boundary=1,
                             seed=seed + 1000 + i)
        c.U = taylor_green(c.n, c.origin, c.ending, phase=tuple(2 * np.pi * u[i, 2:5]), dtype=dtype)
        cases.append(c)
    return cases


def adjoint_case(np_parcels=1 << 20, ncell=64, nsub=200, seed=SEED, boundary=1):
    """The adjoint workload: ncell^3 cells on [0,1]^3 with a frozen divergence-free field
    of 8 seeded Fourier modes; parcels at uniform positions, log-normal diameters, per-
    parcel density 1000 (1 + U[0, 0.5)) kg/m^3, starting at rest; nsub substeps of
    1e-4 s in one call.  The loss target is target_source(case, seed)."""
    n = (ncell,) * 3
    o, e = (0.0,) * 3, (1.0,) * 3
    U = fourier_field(n, o, e, modes=8, seed=seed)
    x, y, z = positions(seed, np_parcels, o, e)
    d = lognormal_diameters(seed, np_parcels)
    rho = 1000.0 * (1.0 + 0.5 * uniform(seed, 50, np_parcels))
    zero = np.zeros(np_parcels)
    return Case(n, o, e, U, d, x, y, z, zero, zero.copy(), zero.copy(), rho=rho, mu_c=1e-3, rho_d=1000.0,
                rho_c=1000.0, dt_fluid=nsub * 1e-4, nsub=nsub, boundary=boundary, name="adjoint")


def channel_profile(n, origin, ending, u0=1.0):
    """Laminar duct inflow with walls on y and z: u = (9/4) u0 16 eta (1 - eta) zeta (1 - zeta)
    at cell centres (eta, zeta the y, z coordinates scaled to [0, 1]), whose mean over the
    cross-section is u0; v = w = 0.  AoS (ncell, 3)."""
    x, y, z = cell_centres(n, origin, ending)
    eta = (y - origin[1]) / (ending[1] - origin[1])
    zeta = (z - origin[2]) / (ending[2] - origin[2])
    U = np.zeros((x.size, 3))
    U[:, 0] = 2.25 * u0 * 16.0 * eta * (1.0 - eta) * zeta * (1.0 - zeta)
    return U


def channel_field(n, origin, ending, u0=1.0, perturbation=0.05, seed=SEED, dtype=np.float64):
    """The channel's frozen carrier: the laminar profile plus a divergence-free random
    perturbation (8 Fourier modes) of RMS perturbation * u0."""
    U = channel_profile(n, origin, ending, u0)
    P = fourier_field(n, origin, ending, modes=8, seed=seed, stream=70)
    rms = float(np.sqrt((P ** 2).sum(1).mean()))
    if rms > 0:
        U += P * (perturbation * u0 / rms)
    return U.astype(dtype)


def cells_of(case, x, y, z):
    """Linear cell index (x fastest) of positions inside the mesh."""
    idx = []
    for a, c in enumerate((x, y, z)):
        d = (case.ending[a] - case.origin[a]) / case.n[a]
        idx.append(np.clip(np.floor((c.astype(np.float64) - case.origin[a]) * (1.0 / d)), 0, case.n[a] - 1)
                   .astype(np.int64))
    return idx[2] * case.n[0] * case.n[1] + idx[1] * case.n[0] + idx[0]


def channel_case(np_parcels=1 << 25, n=(256, 128, 128), length=(2.0, 1.0, 1.0), u0=1.0, perturbation=0.05,
                 dtype=np.float32, seed=SEED, dt_fluid=1e-3, nsub=10, name="channel"):
    """The channel workload at one GPU's shape: n cells (256 x 128 x 128, uniform cubes of
    1/128 m) on a 2 x 1 x 1 m duct, walls on y and z, open inlet (x = 0) and outlet
    (x = 2 m): boundary form 2 (ST_CHANNEL).  Frozen carrier: laminar profile of bulk
    velocity u0 with seeded 5% perturbations.  np_parcels parcels fill the duct at
    seeded uniform positions -- the state the seeded inflow builds up -- moving with
    the carrier velocity of their cell, log-normal diameters (median 30 um, sigma 0.5),
    rho_p = 1000 kg/m^3; carrier mu = 1e-3, rho = 1e3."""
    o = (0.0, 0.0, 0.0)
    e = tuple(float(v) for v in length)
    U = channel_field(n, o, e, u0, perturbation, seed, dtype=dtype)
    x, y, z = positions(seed, np_parcels, o, e, dtype=dtype)
    d = lognormal_diameters(seed, np_parcels, dtype=dtype)
    c = Case(n, o, e, U, d, x, y, z, np.zeros(np_parcels, dtype), np.zeros(np_parcels, dtype),
             np.zeros(np_parcels, dtype), mu_c=1e-3, rho_d=1000.0, rho_c=1000.0, dt_fluid=dt_fluid, nsub=nsub,
             boundary=2, name=name)
    cell = cells_of(c, c.x, c.y, c.z)
    c.u, c.v, c.w = [np.ascontiguousarray(U[cell, k]) for k in range(3)]
    return c


def channel_injection(case, step, count, seed=SEED):
    """The parcels entering the channel during fluid step `step`: count parcels spread
    uniformly over the inlet plane, each placed at x = origin + xi u dt (the distance the
    carrier at its inlet cell moves it during the step, xi uniform in [0, 1)) so the
    inflow rate is uniform in time, moving with that carrier velocity, with log-normal
    diameters; seeded by (seed, step).  Returns d, x, y, z, u, v, w in the case's dtype."""
    dt = case.d.dtype
    s = seed * 1000003 + step
    yz = [case.origin[a] + uniform(s, 80 + a, count) * (case.ending[a] - case.origin[a]) for a in (1, 2)]
    d1 = (case.ending[0] - case.origin[0]) / case.n[0]
    x0 = np.full(count, case.origin[0] + 0.5 * d1)
    cell = cells_of(case, x0, yz[0], yz[1])
    uin = case.U[cell].astype(np.float64)
    xi = uniform(s, 83, count)
    x = case.origin[0] + xi * np.maximum(uin[:, 0], 0.0) * case.dt_fluid
    x = np.minimum(x, case.origin[0] + 0.999 * d1)
    y = into_cells(yz[0].astype(dt), case.origin[1], case.ending[1], case.n[1])
    z = into_cells(yz[1].astype(dt), case.origin[2], case.ending[2], case.n[2])
    x = into_cells(x.astype(dt), case.origin[0], case.ending[0], case.n[0])
    d = lognormal_diameters(s, count, dtype=dt, stream=84)
    return [d, x, y, z] + [np.ascontiguousarray(uin[:, k].astype(dt)) for k in range(3)]


def target_source(case, seed=SEED, scale=1e-12):
    """Seeded target field of the source-misfit loss L = |S - T|^2."""
    return scale * normal(seed, 60, case.ncell * 3).reshape(case.ncell, 3)
