"""The GPU paths against the oracle at the full sizes of the workloads the benchmark is
quoted on (the oracle runs on the host's cores with OpenMP, or serially where it
records decisions):

* 2D lid-driven cavity, 20x20x1 cells, 10,000 parcels, 100 fluid steps of 10 substeps,
  walls, FP64, both methods: every parcel bit-exact after every step, each step's
  source within 1e-12;
* periodic Taylor-Green box, 2^24 parcels on 128^3 cells, one fluid step of 10
  one-substep passes, FP64 and the FP32 form: state bit-exact, source within 1e-12,
  and the momentum balance sum(S) = sum(m (v_start - v_end)) over the whole cloud;
* 64 uneven clouds (1e3 .. 2e6 parcels on 32^3 .. 96^3 cells) in one persistent batch,
  FP32 form, one fluid step: every cloud bit-exact and within 1e-12; the same clouds
  autotuned, captured as a hipGraph and replayed twice;
* adjoint of the source misfit at 2^20 parcels on 64^3 cells over 200 substeps with a
  checkpoint every 20: directional derivative against the oracle's dual numbers
  (1e-9) and against central differences with eps = 1e-6 and frozen decisions (1e-4).
"""
import os

import numpy as np
import pytest

import scaletrack as stk
from oracle import st_oracle
from scaletrack import inputs
from scaletrack.adjoint import source_misfit_rrule

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
SRC_RTOL = 1e-12


def device_cloud(ctx, case, precision):
    dt = np.float64 if precision == stk.FP64 else np.float32
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    field = stk.TwoWayEulerian(ctx, mesh, precision)
    field.upload_U(case.U.astype(dt))
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision, rho_p=case.rho,
                      rho_c=case.rho_c)
    chunk.set_state(*[a.astype(dt) for a in case.state()])
    chunk.upload()
    return chunk, field


def assert_source(got, ref, what):
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max() / scale
    assert err <= SRC_RTOL, f"{what}: source max rel err {err:.3e}"


@pytest.mark.parametrize("method", [stk.PLAIN, stk.SORTED])
def test_cavity_100_fluid_steps(ctx, method):
    case = inputs.cavity_case(np_parcels=10000)
    chunk, field = device_cloud(ctx, case, stk.FP64)
    p = stk.params(boundary=stk.WALL, method=method, substeps_per_pass=1)
    state = case.state()
    for step in range(100):
        field.zero_source()
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
        ref_src, _, state = st_oracle.evolve_case(case, state=state, substeps_per_pass=1)
        assert_source(field.download_source(), ref_src, f"step {step}")
    for name, a, b in zip("dxyzuvw", chunk.download(), state):
        assert np.array_equal(a, b), f"{name} differs after 100 steps"
    chunk.close()
    field.close()


@pytest.mark.parametrize("precision,passes_per_launch", [(stk.FP64, 1), (stk.FP32, 1), (stk.FP64, 10), (stk.FP32, 10)])
def test_taylor_green_16M(ctx, precision, passes_per_launch):
    """The headline workload, one fluid step of ten one-substep passes, as ten launches
    or as one launch of the per-XCD work queue."""
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = inputs.taylor_green_box(np_parcels=1 << 24, ncell=128, dtype=dt, boundary=stk.PERIODIC)
    chunk, field = device_cloud(ctx, case, precision)
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1, passes_per_launch=passes_per_launch)
    field.zero_source()
    stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
    src = field.download_source()
    got = chunk.download()
    chunk.close()
    field.close()
    ref_src, _, ref = st_oracle.evolve_case(case, precision=st_oracle.FP64 if precision == stk.FP64 else st_oracle.FP32,
                                            substeps_per_pass=1, nthreads=THREADS)
    for name, a, b in zip("dxyzuvw", got, ref):
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{name}: {bad.size} parcels differ"
    assert_source(src, ref_src, "16M")
    # momentum balance (no walls): the deposits of a parcel telescope to m (v_start - v_end)
    m = (case.rho_d_by_rho_c * np.pi) * case.d.astype(np.float64) ** 3 / 6.0
    v0 = np.stack([case.u, case.v, case.w], 1).astype(np.float64)
    v1 = np.stack(got[4:7], 1).astype(np.float64)
    want = (m[:, None] * (v0 - v1)).sum(0)
    np.testing.assert_allclose(src.sum(0), want, rtol=1e-10, atol=1e-12 * np.abs(src).sum())


def test_taylor_green_16M_graph(ctx):
    """The headline workload as a captured hipGraph of one fluid step (source zeroed,
    ten one-substep passes), replayed three times with eager work and synchronisations
    between the replays: after every replay the state equals, bit for bit, that of the
    same steps run eagerly on a second copy of the cloud, and the source agrees to the
    order of the FP64 additions."""
    case = inputs.taylor_green_box(np_parcels=1 << 24, ncell=128, dtype=np.float32, boundary=stk.PERIODIC)
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1)
    gchunk, gfield = device_cloud(ctx, case, stk.FP32)
    echunk, efield = device_cloud(ctx, case, stk.FP32)
    graph = stk.Graph(ctx, gchunk, gfield, case.dt_fluid, case.nsub, p, zero_source=True)
    for replay in range(3):
        graph.launch()
        ctx.synchronize()
        efield.zero_source()
        stk.evolve_device(ctx, echunk, efield, case.dt_fluid, case.nsub, p, want_bbox=False)
        for name, a, b in zip("dxyzuvw", gchunk.download(), echunk.download()):
            assert np.array_equal(a, b), f"replay {replay}: {name} differs"
        assert_source(gfield.download_source(), efield.download_source(), f"replay {replay}")
    graph.close()
    for o in (gchunk, gfield, echunk, efield):
        o.close()


def test_batch_64_clouds(ctx):
    cases = inputs.batch_cases(64, dtype=np.float32)
    objs = [device_cloud(ctx, c, stk.FP32) for c in cases]
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1)
    batch = stk.Batch(ctx, [o[0] for o in objs], [o[1] for o in objs], p, target_parcels_per_brick=4096)
    for _, f in objs:
        f.zero_source()
    batch.evolve(cases[0].dt_fluid, cases[0].nsub)
    for i, (c, (ch, f)) in enumerate(zip(cases, objs)):
        ref_src, _, ref = st_oracle.evolve_case(c, precision=st_oracle.FP32, substeps_per_pass=1, nthreads=THREADS)
        for name, a, b in zip("dxyzuvw", ch.download(), ref):
            assert np.array_equal(a, b), f"cloud {i}: {name} differs"
        assert_source(f.download_source(), ref_src, f"cloud {i}")
    batch.close()
    for ch, f in objs:
        ch.close()
        f.close()


def test_batch_64_clouds_tuned_graph(ctx, tmp_path):
    """The batch as production runs it: the autotuner picks the brick target, block and
    workgroups per CU for the 64 clouds (restoring them afterwards), the batch is built
    with its choice, one fluid step (every UTrans zeroed, 10 one-substep passes) is
    captured as a hipGraph and replayed twice, with the sources and parcels downloaded
    between the replays.  After each replay every cloud's source is the oracle's for
    that fluid step (1e-12), and after the second every parcel is bit-exact."""
    cases = inputs.batch_cases(64, dtype=np.float32)
    objs = [device_cloud(ctx, c, stk.FP32) for c in cases]
    chunks, fields = [o[0] for o in objs], [o[1] for o in objs]
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1)
    tuned = stk.autotune_batch(ctx, chunks, fields, p, dt_fluid=cases[0].dt_fluid, trials=5,
                               cache_path=tmp_path / "tune.jsonl")
    assert tuned["ms_per_pass"] > 0 and tuned["runner_up_ms_per_pass"] >= tuned["ms_per_pass"]
    batch = stk.Batch(ctx, chunks, fields, stk.params(boundary=stk.PERIODIC, substeps_per_pass=1, block=tuned["block"]),
                      target_parcels_per_brick=tuned["target_parcels_per_brick"], blocks_per_cu=tuned["blocks_per_cu"])
    g = batch.graph(cases[0].dt_fluid, cases[0].nsub, zero_source=True)
    states = [c.state() for c in cases]
    for replay in range(2):
        g.launch()
        for i, (c, (ch, f)) in enumerate(zip(cases, objs)):
            ref_src, _ = st_oracle.evolve(states[i], c.U, c.n, c.origin, c.ending, c.dt_fluid, nsub=c.nsub,
                                          precision=st_oracle.FP32, boundary=c.boundary, mu_c=c.mu_c, rho_d=c.rho_d,
                                          rho_d_by_rho_c=c.rho_d_by_rho_c, rho_c=c.rho_c, rho=c.rho,
                                          substeps_per_pass=1, nthreads=THREADS)
            assert_source(f.download_source(), ref_src, f"replay {replay}, cloud {i}")
            got = ch.download()
            if replay == 1:
                for name, a, b in zip("dxyzuvw", got, states[i]):
                    assert np.array_equal(a, b), f"cloud {i}: {name} differs after two replays"
    g.close()
    batch.close()
    for ch, f in objs:
        ch.close()
        f.close()


def test_adjoint_config_size(ctx):
    case = inputs.adjoint_case(np_parcels=1 << 20, ncell=64, nsub=200)
    T = inputs.target_source(case)
    chunk, field = device_cloud(ctx, case, stk.FP64)
    loss, pullback = source_misfit_rrule(ctx, chunk, field, T, case.dt_fluid, case.nsub, 20,
                                         stk.params(boundary=stk.PERIODIC))
    _, gd, grho, gcd = pullback(1.0)
    chunk.close()
    field.close()
    rng = np.random.default_rng(17)
    dd = rng.normal(size=case.np) * case.d
    dr = rng.normal(size=case.np) * case.rho
    dc = 0.3
    got = np.dot(gd, dd) + np.dot(grho, dr) + gcd * dc

    # forward-mode dual numbers through the oracle
    st = case.state()
    tsrc = np.zeros((case.ncell, 3))
    src, _ = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                              boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                              rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c,
                              rho=case.rho, tan_d=dd, tan_rho=dr, tan_cd=dc, tan_source=tsrc)
    assert abs(loss - np.sum((src - T) ** 2)) <= 1e-12 * loss
    dual = np.sum(2.0 * (src - T) * tsrc)
    assert abs(got - dual) <= 1e-9 * abs(dual), (got, dual)

    # central differences with the forward run's decisions frozen
    ev = np.zeros((case.np, case.nsub), dtype=np.int64)
    st_oracle.evolve(case.state(), case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                     boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                     rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c, rho=case.rho,
                     events=ev, events_mode=1)
    h = 1e-6

    def L(sign):
        st = case.state()
        st[0] = case.d + sign * h * dd
        s, _ = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                                boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                                rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c,
                                cd=1.0 + sign * h * dc, rho=case.rho + sign * h * dr, events=ev, events_mode=2)
        return np.sum((s - T) ** 2)

    fd = (L(1.0) - L(-1.0)) / (2.0 * h)
    assert abs(got - fd) <= 1e-4 * abs(fd), (got, fd)
