"""Batched persistent launches and hipGraph replay against the oracle.

A batch advances each of its clouds exactly as a single-cloud step would: parcel
states bit-exact against the oracle, sources within 1e-12 of the largest |source|.
Graph replays are checked one by one against eager steps, with eager work and
synchronisations between the replays.
"""
import numpy as np
import pytest

import scaletrack as stk
from oracle import st_oracle
from scaletrack import inputs

pytestmark = pytest.mark.gpu


def uneven_cases(dtype, count=6, seed=5):
    cases = []
    for i in range(count):
        ncell = [8, 12, 16, 20, 10, 14][i % 6]
        npar = [200, 20000, 3000, 50000, 1, 7000][i % 6]
        c = inputs.taylor_green_box(np_parcels=npar, ncell=ncell, dtype=dtype, boundary=1, seed=seed + i)
        c.U = inputs.taylor_green(c.n, c.origin, c.ending, phase=(0.3 * i, 0.7 * i, 1.1 * i), dtype=dtype)
        cases.append(c)
    return cases


def gpu_cloud(ctx, case, precision):
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    field = stk.TwoWayEulerian(ctx, mesh, precision)
    field.upload_U(case.U)
    field.zero_source()
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision)
    chunk.set_state(*case.state())
    chunk.upload()
    return chunk, field


def check_against_oracle(case, precision, chunk, field, states, srcs):
    got = chunk.download()
    src = field.download_source()
    ref_src, _, ref_state = st_oracle.evolve_case(case, state=states, precision=precision)
    for a, b in zip(got, ref_state):
        np.testing.assert_array_equal(a, b)
    scale = max(np.abs(ref_src).max(), 1e-300)
    assert np.abs(src - ref_src).max() <= 1e-12 * scale
    return ref_state


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
@pytest.mark.parametrize("target,per", [(512, 1), (4096, 1), (2048, 0)])
def test_batch_matches_oracle(ctx, precision, target, per):
    dt = np.float64 if precision == stk.FP64 else np.float32
    cases = uneven_cases(dt)
    objs = [gpu_cloud(ctx, c, precision) for c in cases]
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=per, block=256)
    batch = stk.Batch(ctx, [o[0] for o in objs], [o[1] for o in objs], p, target_parcels_per_brick=target)
    info = batch.info()
    assert info["items"] > 0 and info["grid"] > 0
    states = [c.state() for c in cases]
    for step in range(2):
        for _, f in objs:
            f.zero_source()
        batch.evolve(cases[0].dt_fluid, cases[0].nsub)
        for i, (c, (ch, f)) in enumerate(zip(cases, objs)):
            ref_src, _, states[i] = st_oracle.evolve_case(c, state=states[i], precision=precision,
                                                          substeps_per_pass=per)
            got = ch.download()
            for a, b in zip(got, states[i]):
                np.testing.assert_array_equal(a, b)
            src = f.download_source()
            scale = max(np.abs(ref_src).max(), 1e-300)
            assert np.abs(src - ref_src).max() <= 1e-12 * scale, (step, i)
    batch.close()


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
def test_batch_mixed_density_forms(ctx, precision):
    """Clouds with and without a per-parcel density in one batch: each group runs its
    own kernel instance (one launch per group and pass), and every cloud still matches
    the oracle; the work-item counter restarts at each group's first item."""
    dt = np.float64 if precision == stk.FP64 else np.float32
    cases = uneven_cases(dt)
    rng = np.random.default_rng(11)
    objs = []
    for i, c in enumerate(cases):
        if i % 2 == 0:
            c.rho = (1000.0 + 300.0 * rng.standard_normal(c.np)).clip(200.0, 3000.0).astype(dt)
        mesh = stk.Mesh(c.n, c.origin, c.ending)
        field = stk.TwoWayEulerian(ctx, mesh, precision)
        field.upload_U(c.U)
        field.zero_source()
        chunk = stk.Chunk(ctx, c.np, c.mu_c, c.rho_d, c.rho_d_by_rho_c, precision, rho_p=c.rho, rho_c=c.rho_c)
        chunk.set_state(*c.state())
        chunk.upload()
        objs.append((chunk, field))
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1, block=256)
    batch = stk.Batch(ctx, [o[0] for o in objs], [o[1] for o in objs], p, target_parcels_per_brick=2048)
    states = [c.state() for c in cases]
    for step in range(2):
        for _, f in objs:
            f.zero_source()
        batch.evolve(cases[0].dt_fluid, cases[0].nsub)
        for i, (c, (ch, f)) in enumerate(zip(cases, objs)):
            ref_src, _, states[i] = st_oracle.evolve_case(c, state=states[i], precision=precision,
                                                          substeps_per_pass=1)
            for a, b in zip(ch.download(), states[i]):
                np.testing.assert_array_equal(a, b)
            src = f.download_source()
            scale = max(np.abs(ref_src).max(), 1e-300)
            assert np.abs(src - ref_src).max() <= 1e-12 * scale, (step, i)
    batch.close()
    for ch, f in objs:
        ch.close()
        f.close()


def test_batch_graph_replays(ctx):
    cases = uneven_cases(np.float32, count=4)
    objs = [gpu_cloud(ctx, c, stk.FP32) for c in cases]
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1)
    batch = stk.Batch(ctx, [o[0] for o in objs], [o[1] for o in objs], p, target_parcels_per_brick=1024)
    g = batch.graph(cases[0].dt_fluid, cases[0].nsub, zero_source=True)
    states = [c.state() for c in cases]
    for replay in range(3):
        g.launch()
        ctx.synchronize()
        for i, (c, (ch, f)) in enumerate(zip(cases, objs)):
            ref_src, _, states[i] = st_oracle.evolve_case(c, state=states[i], precision=stk.FP32,
                                                          substeps_per_pass=1)
            for a, b in zip(ch.download(), states[i]):
                np.testing.assert_array_equal(a, b)
            src = f.download_source()
            assert np.abs(src - ref_src).max() <= 1e-12 * max(np.abs(ref_src).max(), 1e-300), (replay, i)
        # eager work between replays: one more eager step of the whole batch
        for _, f in objs:
            f.zero_source()
        batch.evolve(cases[0].dt_fluid, cases[0].nsub)
        for i, c in enumerate(cases):
            _, _, states[i] = st_oracle.evolve_case(c, state=states[i], precision=stk.FP32, substeps_per_pass=1)
    g.close()
    batch.close()


def test_single_cloud_graph_replays(ctx):
    case = inputs.taylor_green_box(np_parcels=20000, ncell=16, boundary=1)
    chunk, field = gpu_cloud(ctx, case, stk.FP64)
    p = stk.params(boundary=1, substeps_per_pass=1)
    g = stk.Graph(ctx, chunk, field, case.dt_fluid, case.nsub, p, zero_source=True)
    state = case.state()
    for replay in range(3):
        g.launch()
        ref_src, _, state = st_oracle.evolve_case(case, state=state, substeps_per_pass=1)
        for a, b in zip(chunk.download(), state):
            np.testing.assert_array_equal(a, b)
        src = field.download_source()
        assert np.abs(src - ref_src).max() <= 1e-12 * np.abs(ref_src).max(), replay
        field.zero_source()
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
        _, _, state = st_oracle.evolve_case(case, state=state, substeps_per_pass=1)
    g.close()


def test_autotune_restores_state_and_caches(ctx, tmp_path):
    cases = uneven_cases(np.float32, count=3)
    objs = [gpu_cloud(ctx, c, stk.FP32) for c in cases]
    chunks, fields = [o[0] for o in objs], [o[1] for o in objs]
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1)
    cache = tmp_path / "tune.jsonl"
    best = stk.autotune_batch(ctx, chunks, fields, p, cases[0].dt_fluid, trials=1, cache_path=cache)
    assert not best["from_cache"] and best["ms_per_pass"] > 0
    again = stk.autotune_batch(ctx, chunks, fields, p, cases[0].dt_fluid, trials=1, cache_path=cache)
    assert again["from_cache"] and again["block"] == best["block"]
    # tuning left the clouds untouched: a tuned batch step still matches the oracle
    batch = stk.Batch(ctx, chunks, fields, stk.params(boundary=stk.PERIODIC, substeps_per_pass=1,
                                                      block=best["block"]),
                      target_parcels_per_brick=best["target_parcels_per_brick"],
                      blocks_per_cu=best["blocks_per_cu"])
    for _, f in objs:
        f.zero_source()
    batch.evolve(cases[0].dt_fluid, cases[0].nsub)
    for c, (ch, f) in zip(cases, objs):
        ref_src, _, ref_state = st_oracle.evolve_case(c, precision=stk.FP32, substeps_per_pass=1)
        for a, b in zip(ch.download(), ref_state):
            np.testing.assert_array_equal(a, b)
        src = f.download_source()
        assert np.abs(src - ref_src).max() <= 1e-12 * max(np.abs(ref_src).max(), 1e-300)
    batch.close()


@pytest.mark.parametrize("seed", range(12))
def test_random_single_cloud_graph_replays(ctx, seed):
    """A random configuration captured in a graph and replayed three times, an eager
    step between replays.  With resort_fraction 0 every eager step re-sorts the cloud,
    in place, so the replays run on a brick order the capture never saw."""
    rng = np.random.default_rng(11000 + seed)
    precision = stk.FP32 if rng.random() < 0.5 else stk.FP64
    dt = np.float32 if precision == stk.FP32 else np.float64
    boundary = int(rng.integers(0, 2))
    n = tuple(int(v) for v in rng.integers(2, 25, size=3))
    o = tuple(float(v) for v in rng.uniform(-1.0, 1.0, size=3))
    e = tuple(a + float(L) for a, L in zip(o, rng.uniform(0.2, 2.0, size=3)))
    npc = int(rng.choice([1, 64, int(rng.integers(100, 20001))]))
    x, y, z = inputs.positions(700 + seed, npc, o, e, dtype=dt)
    d = inputs.lognormal_diameters(700 + seed, npc, median=float(10 ** rng.uniform(-5, -4)), dtype=dt)
    u, v, w = [(rng.standard_normal(npc) * 0.3).astype(dt) for _ in range(3)]
    U = inputs.fourier_field(n, o, e, modes=5, seed=800 + seed, amplitude=float(10 ** rng.uniform(-1, 0.5)), dtype=dt)
    nsub = int(rng.integers(1, 9))
    case = inputs.Case(n, o, e, U, d, x, y, z, u, v, w, dt_fluid=float(10 ** rng.uniform(-4, -2.5)), nsub=nsub,
                       boundary=boundary)
    per = int(rng.choice([0, 1, 2]))
    p = stk.params(boundary=boundary, method=stk.SORTED if rng.random() < 0.7 else stk.PLAIN,
                   substeps_per_pass=per, block=int(rng.choice([128, 256, 512])),
                   resort_fraction=float(rng.choice([0.0, 0.08, 1e9])))
    chunk, field = gpu_cloud(ctx, case, precision)
    oprec = st_oracle.FP32 if precision == stk.FP32 else st_oracle.FP64
    g = stk.Graph(ctx, chunk, field, case.dt_fluid, case.nsub, p, zero_source=True)
    state = case.state()
    for replay in range(3):
        g.launch()
        ref_src, _, state = st_oracle.evolve_case(case, state=state, precision=oprec, substeps_per_pass=per)
        for name, a, b in zip("dxyzuvw", chunk.download(), state):
            assert np.array_equal(a, b), (seed, replay, name)
        src = field.download_source()
        assert np.abs(src - ref_src).max() <= 1e-12 * max(np.abs(ref_src).max(), 1e-300), (seed, replay)
        field.zero_source()
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
        _, _, state = st_oracle.evolve_case(case, state=state, precision=oprec, substeps_per_pass=per)
    g.close()
    chunk.close()
    field.close()
