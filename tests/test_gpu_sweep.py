"""Randomised configurations of the GPU paths against the oracle.

Each seed draws a whole configuration:
- mesh dims 1..40 per axis, a non-cubic box at an arbitrary origin;
- a divergence-free carrier of random amplitude;
- 1..30000 parcels with random diameters and velocities (per-parcel density or not);
- random carrier properties and c_D multiplier;
- 1..12 substeps in passes of random length;
- boundary, precision form, method, brick shape, block size and flush mode.

The batch sweep draws 1..6 clouds per seed (empty, one-parcel and large clouds, with and
without a per-parcel density, each on its own random mesh) with a random brick target,
per-CU cap, block, flush mode and substeps per pass, and checks every cloud.

Two calls are made per configuration, the source zeroed in between.  The bars are
those of the parity tests: state and bounding box bit-exact; the source within 1e-12
of its largest magnitude.  The seeds are fixed, so a failure names its configuration.
"""
import numpy as np
import pytest

import scaletrack as stk
from scaletrack import inputs
from oracle import st_oracle

pytestmark = pytest.mark.gpu

LDS_LIMIT = 65536


def draw(seed):
    rng = np.random.default_rng(1000 + seed)
    precision = stk.FP32 if rng.random() < 0.5 else stk.FP64
    dt = np.float32 if precision == stk.FP32 else np.float64
    boundary = stk.PERIODIC if rng.random() < 0.5 else stk.WALL
    n = tuple(int(v) for v in rng.integers(1, 41, size=3))
    if rng.random() < 0.25:
        n = (n[0], n[1], 1)  # a 2D mesh
    origin = tuple(float(v) for v in rng.uniform(-2.0, 2.0, size=3))
    ending = tuple(o + float(L) for o, L in zip(origin, rng.uniform(0.05, 3.0, size=3)))
    U = inputs.fourier_field(n, origin, ending, modes=6, seed=seed, amplitude=float(10 ** rng.uniform(-1, 1)),
                             dtype=dt)
    big = int(rng.integers(100, 30001))
    npc = int(rng.choice([1, 2, 63, 64, 65, big, big, big, big, big]))
    x, y, z = inputs.positions(seed, npc, origin, ending, dtype=dt)
    d = inputs.lognormal_diameters(seed, npc, median=float(10 ** rng.uniform(-5.3, -3.7)), dtype=dt)
    vs = float(10 ** rng.uniform(-2, 0.5))
    u, v, w = [(rng.standard_normal(npc) * vs).astype(dt) for _ in range(3)]
    rho_d = float(rng.uniform(500.0, 5000.0))
    rho = (rho_d * rng.uniform(0.5, 2.0, size=npc)).astype(dt) if rng.random() < 0.4 else None
    nsub = int(rng.integers(1, 13))
    case = inputs.Case(n, origin, ending, U, d, x, y, z, u, v, w, rho=rho, mu_c=float(10 ** rng.uniform(-5, -2)),
                       rho_d=rho_d, rho_c=float(rng.uniform(1.0, 1000.0)),
                       dt_fluid=float(10 ** rng.uniform(-4, -2)), nsub=nsub, boundary=boundary,
                       name=f"sweep {seed}")
    method = stk.SORTED if rng.random() < 0.7 else stk.PLAIN
    per = int(rng.choice([0, 1, 2, 3, nsub]))
    elem = 8 if precision == stk.FP64 else 4
    while True:
        brick = tuple(int(v) for v in rng.integers(1, 17, size=3))
        tile = (brick[0] + 2) * (brick[1] + 2) * (brick[2] + 2)
        if brick[0] * brick[1] * brick[2] <= 2048 and tile * 3 * (8 + elem) <= LDS_LIMIT:
            break
    p = stk.params(boundary=boundary, method=method, substeps_per_pass=per, brick=brick,
                   block=int(rng.choice([64, 128, 256, 512, 1024])),
                   flush=stk.FLUSH_OWNER if rng.random() < 0.7 else stk.FLUSH_ATOMIC,
                   cd=float(rng.uniform(0.5, 2.0)))
    return case, precision, p


@pytest.mark.parametrize("seed", range(96))
def test_random_configuration(ctx, seed):
    case, precision, p = draw(seed)
    ident = (f"seed {seed}: n={case.n} np={case.np} {'fp32' if precision == stk.FP32 else 'fp64'} "
             f"{'periodic' if case.boundary else 'wall'} method={p.method} brick={tuple(p.brick)} "
             f"block={p.block} flush={p.flush} nsub={case.nsub} per={p.substeps_per_pass} "
             f"rho={'parcel' if case.rho is not None else 'chunk'}")
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), precision)
    field.upload_U(case.U)
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision, rho_p=case.rho,
                      rho_c=case.rho_c)
    chunk.set_state(*case.state())
    chunk.upload()
    st = case.state()
    oprec = st_oracle.FP32 if precision == stk.FP32 else st_oracle.FP64
    for call in range(2):
        field.zero_source()
        bb = stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
        src = field.download_source()
        ref_src, ref_bb, st = st_oracle.evolve_case(case, state=st, precision=oprec,
                                                   substeps_per_pass=p.substeps_per_pass, cd=p.cd)
        scale = max(np.abs(ref_src).max(), 1e-300)
        assert np.abs(src - ref_src).max() <= 1e-12 * scale, f"{ident}, call {call}: source"
        np.testing.assert_array_equal(np.concatenate(bb), ref_bb, err_msg=f"{ident}, call {call}: bbox")
    for name, a, b in zip("dxyzuvw", chunk.download(), st):
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{ident}: {name} differs for {bad.size} parcels, first {bad[:3]}"
    chunk.close()
    field.close()


def draw_batch(seed):
    """1..6 clouds sharing precision and params, each with its own random mesh, cloud and
    density form; random brick target, block, per-CU cap and substeps per pass."""
    rng = np.random.default_rng(5000 + seed)
    precision = stk.FP32 if rng.random() < 0.5 else stk.FP64
    dt = np.float32 if precision == stk.FP32 else np.float64
    boundary = stk.PERIODIC if rng.random() < 0.5 else stk.WALL
    cases = []
    for i in range(int(rng.integers(1, 7))):
        n = tuple(int(v) for v in rng.integers(1, 33, size=3))
        origin = tuple(float(v) for v in rng.uniform(-1.0, 1.0, size=3))
        ending = tuple(o + float(L) for o, L in zip(origin, rng.uniform(0.1, 2.0, size=3)))
        U = inputs.fourier_field(n, origin, ending, modes=5, seed=100 * seed + i,
                                 amplitude=float(10 ** rng.uniform(-1, 0.7)), dtype=dt)
        npc = int(rng.choice([0, 1, 65, int(rng.integers(100, 20001)), int(rng.integers(100, 20001))]))
        x, y, z = inputs.positions(100 * seed + i, npc, origin, ending, dtype=dt)
        d = inputs.lognormal_diameters(100 * seed + i, npc, median=float(10 ** rng.uniform(-5, -3.8)), dtype=dt)
        u, v, w = [(rng.standard_normal(npc) * 0.3).astype(dt) for _ in range(3)]
        rho = (1500.0 * rng.uniform(0.5, 2.0, size=npc)).astype(dt) if rng.random() < 0.3 else None
        cases.append(inputs.Case(n, origin, ending, U, d, x, y, z, u, v, w, rho=rho, mu_c=1.8e-5, rho_d=1500.0,
                                 rho_c=1.2, dt_fluid=2e-3, nsub=5, boundary=boundary))
    p = stk.params(boundary=boundary, substeps_per_pass=int(rng.choice([0, 1, 2])),
                   block=int(rng.choice([128, 256, 512, 1024])),
                   flush=stk.FLUSH_OWNER if rng.random() < 0.7 else stk.FLUSH_ATOMIC)
    target = int(rng.choice([0, 64, 512, 4096]))
    per_cu = int(rng.choice([0, 1, 2, 3, 4]))
    return cases, precision, p, target, per_cu


@pytest.mark.parametrize("seed", range(24))
def test_random_batch(ctx, seed):
    cases, precision, p, target, per_cu = draw_batch(seed)
    ident = (f"batch seed {seed}: {len(cases)} clouds, nps={[c.np for c in cases]}, "
             f"{'fp32' if precision == stk.FP32 else 'fp64'} target={target} per_cu={per_cu} "
             f"block={p.block} flush={p.flush} per={p.substeps_per_pass}")
    objs = []
    for c in cases:
        field = stk.TwoWayEulerian(ctx, stk.Mesh(c.n, c.origin, c.ending), precision)
        field.upload_U(c.U)
        chunk = stk.Chunk(ctx, c.np, c.mu_c, c.rho_d, c.rho_d_by_rho_c, precision, rho_p=c.rho, rho_c=c.rho_c)
        chunk.set_state(*c.state())
        chunk.upload()
        objs.append((chunk, field))
    batch = stk.Batch(ctx, [o[0] for o in objs], [o[1] for o in objs], p, target_parcels_per_brick=target,
                      blocks_per_cu=per_cu)
    states = [c.state() for c in cases]
    oprec = st_oracle.FP32 if precision == stk.FP32 else st_oracle.FP64
    for call in range(2):
        for _, f in objs:
            f.zero_source()
        batch.evolve(cases[0].dt_fluid, cases[0].nsub)
        for i, (c, (ch, f)) in enumerate(zip(cases, objs)):
            ref_src, _, states[i] = st_oracle.evolve_case(c, state=states[i], precision=oprec,
                                                          substeps_per_pass=p.substeps_per_pass)
            src = f.download_source()
            scale = max(np.abs(ref_src).max(), 1e-300)
            assert np.abs(src - ref_src).max() <= 1e-12 * scale, f"{ident}: cloud {i} call {call} source"
    for i, (ch, _) in enumerate(objs):
        for name, a, b in zip("dxyzuvw", ch.download(), states[i]):
            assert np.array_equal(a, b), f"{ident}: cloud {i} {name} differs"
    batch.close()
    for ch, f in objs:
        ch.close()
        f.close()


def draw_channel(seed):
    """A random channel-form configuration (walls on y and z, open x planes): mesh,
    carrier with a mean flow along x that carries parcels out through the outlet, cloud
    (optionally with a per-parcel density), substeps, method, brick, block, flush and
    re-sort share; and the parcels injected before each call (0 to 3000, spread over
    the first fifth of the duct) and the calls after which removed parcels are dropped."""
    rng = np.random.default_rng(9000 + seed)
    precision = stk.FP32 if rng.random() < 0.5 else stk.FP64
    dt = np.float32 if precision == stk.FP32 else np.float64
    n = (int(rng.integers(2, 41)), int(rng.integers(1, 21)), int(rng.integers(1, 21)))
    origin = tuple(float(v) for v in rng.uniform(-1.0, 1.0, size=3))
    ending = tuple(o + float(L) for o, L in zip(origin, rng.uniform(0.05, 2.0, size=3)))
    amp = float(10 ** rng.uniform(-1, 0.5))
    U = inputs.fourier_field(n, origin, ending, modes=5, seed=seed, amplitude=0.3 * amp)
    U[:, 0] += amp
    U = U.astype(dt)
    npc = int(rng.choice([0, 1, 64, int(rng.integers(100, 15001)), int(rng.integers(100, 15001))]))
    x, y, z = inputs.positions(seed, npc, origin, ending, dtype=dt)
    d = inputs.lognormal_diameters(seed, npc, median=float(10 ** rng.uniform(-5, -3.8)), dtype=dt)
    u, v, w = [(rng.standard_normal(npc) * 0.3 * amp).astype(dt) for _ in range(3)]
    rho = (1500.0 * rng.uniform(0.5, 2.0, size=npc)).astype(dt) if rng.random() < 0.3 else None
    L0 = ending[0] - origin[0]
    nsub = int(rng.integers(1, 13))
    # a fluid step moves the mean flow over about 0.3 to 3 duct lengths' worth of cells
    dt_fluid = float(L0 / amp * 10 ** rng.uniform(-1.5, -0.5))
    case = inputs.Case(n, origin, ending, U, d, x, y, z, u, v, w, rho=rho, mu_c=1.8e-5, rho_d=1500.0, rho_c=1.2,
                       dt_fluid=dt_fluid, nsub=nsub, boundary=2, name=f"channel sweep {seed}")
    elem = 8 if precision == stk.FP64 else 4
    while True:
        brick = tuple(int(v) for v in rng.integers(1, 17, size=3))
        tile = (brick[0] + 2) * (brick[1] + 2) * (brick[2] + 2)
        if brick[0] * brick[1] * brick[2] <= 2048 and tile * 3 * (8 + elem) <= LDS_LIMIT:
            break
    p = stk.params(boundary=stk.CHANNEL, method=stk.SORTED if rng.random() < 0.7 else stk.PLAIN,
                   substeps_per_pass=int(rng.choice([0, 1, 2, nsub])), brick=brick,
                   block=int(rng.choice([64, 128, 256, 512, 1024])),
                   flush=stk.FLUSH_OWNER if rng.random() < 0.7 else stk.FLUSH_ATOMIC,
                   resort_fraction=float(rng.choice([0.0, 0.02, 0.1, 1.0])))
    injections = []
    for call in range(4):
        k = int(rng.choice([0, 1, 50, int(rng.integers(100, 3001))]))
        xi, yi, zi = inputs.positions(100 * seed + call, k, origin, ending, dtype=dt)
        xi = (origin[0] + (xi.astype(np.float64) - origin[0]) * 0.2).astype(dt)
        lo = dt(origin[0])
        if float(lo) < origin[0]:  # the nearest value of the state type at or above the inlet
            lo = np.nextafter(lo, dt(np.inf))
        xi = np.maximum(xi, lo)
        di = inputs.lognormal_diameters(100 * seed + call, k, median=float(10 ** rng.uniform(-5, -3.8)), dtype=dt)
        ui = np.full(k, amp, dtype=dt)
        vi, wi = np.zeros(k, dtype=dt), np.zeros(k, dtype=dt)
        rhoi = (1500.0 * rng.uniform(0.5, 2.0, size=k)).astype(dt) if rho is not None else None
        injections.append(((di, xi, yi, zi, ui, vi, wi), rhoi, bool(rng.random() < 0.4)))
    return case, precision, p, injections


@pytest.mark.parametrize("seed", range(24))
def test_random_channel(ctx, seed):
    """Random channel runs: four calls, parcels injected before each and the removed ones
    dropped after some; the cloud in upload order, removal marks included, equals the
    oracle's list kept the same way (appended, then filtered in order)."""
    case, precision, p, injections = draw_channel(seed)
    ident = (f"channel seed {seed}: n={case.n} np={case.np} {'fp32' if precision == stk.FP32 else 'fp64'} "
             f"method={p.method} brick={tuple(p.brick)} block={p.block} flush={p.flush} nsub={case.nsub} "
             f"per={p.substeps_per_pass} resort={p.resort_fraction} "
             f"rho={'parcel' if case.rho is not None else 'chunk'}")
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), precision)
    field.upload_U(case.U)
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision, rho_p=case.rho,
                      rho_c=case.rho_c)
    chunk.set_state(*case.state())
    chunk.upload()
    st = case.state()
    rho = case.rho
    oprec = st_oracle.FP32 if precision == stk.FP32 else st_oracle.FP64
    for call, (new, rho_new, drop) in enumerate(injections):
        if new[0].size:
            chunk.inject(*new, rho=rho_new)
            st = [np.concatenate([a, b]) for a, b in zip(st, new)]
            if rho is not None:
                rho = np.concatenate([rho, rho_new])
        field.zero_source()
        bb = stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
        src = field.download_source()
        ref_src, ref_bb = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid,
                                           nsub=case.nsub, precision=oprec, boundary=st_oracle.CHANNEL,
                                           mu_c=case.mu_c, rho_d=case.rho_d, rho_d_by_rho_c=case.rho_d_by_rho_c,
                                           rho_c=case.rho_c, rho=rho, substeps_per_pass=p.substeps_per_pass)
        scale = max(np.abs(ref_src).max(), 1e-300)
        assert np.abs(src - ref_src).max() <= 1e-12 * scale, f"{ident}, call {call}: source"
        np.testing.assert_array_equal(np.concatenate(bb), ref_bb, err_msg=f"{ident}, call {call}: bbox")
        for name, a, b in zip("dxyzuvw", chunk.download(), st):
            bad = np.flatnonzero(a != b)
            assert bad.size == 0, f"{ident}, call {call}: {name} differs for {bad.size} parcels, first {bad[:3]}"
        if drop:
            live = st[0] > 0
            st = [a[live] for a in st]
            if rho is not None:
                rho = rho[live]
            assert chunk.compact() == st[0].size, f"{ident}, call {call}: live count"
    chunk.close()
    field.close()
