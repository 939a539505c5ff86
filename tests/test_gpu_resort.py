"""Re-sorts decided on the device, steps queued back to back and replayed from graphs,
the multi-pass persistent launch, and the lifetime rules of graphs, batches and
adjoints.

Parity bars as everywhere: parcel state bit-exact against the oracle, sources within
1e-12 of the largest |source| (only the order of FP64 additions differs).
"""
import ctypes as C

import numpy as np
import pytest

import scaletrack as stk
from oracle import st_oracle
from scaletrack import _lib, inputs
from scaletrack.adjoint import Adjoint

pytestmark = pytest.mark.gpu

SRC_RTOL = 1e-12


def device_cloud(ctx, case, precision):
    dt = np.float64 if precision == stk.FP64 else np.float32
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), precision)
    field.upload_U(case.U.astype(dt))
    field.zero_source()
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision, rho_p=case.rho,
                      rho_c=case.rho_c)
    chunk.set_state(*[a.astype(dt) for a in case.state()])
    chunk.upload()
    return chunk, field


def oracle_calls(case, precision, calls):
    """`calls` fluid steps of the oracle, one pass per substep, the source summed over
    them (never zeroed)."""
    st = case.state()
    src = np.zeros((case.ncell, 3))
    for _ in range(calls):
        _, _, st = st_oracle.evolve_case(case, state=st, source=src, precision=precision, substeps_per_pass=1)
    return st, src


def assert_match(got_state, got_src, ref_state, ref_src, what):
    for name, a, b in zip("dxyzuvw", got_state, ref_state):
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{what}: {name} differs for {bad.size} parcels (first {bad[:4]})"
    err = np.abs(got_src - ref_src).max() / max(np.abs(ref_src).max(), 1e-300)
    assert err <= SRC_RTOL, f"{what}: source max rel err {err:.3e}"


def drifting_case(precision, boundary=stk.PERIODIC, n=1 << 16):
    """32^3 Taylor-Green box with fluid steps of 1e-2 s: parcels cross the 4^3-cell bricks
    of these tests within a few steps."""
    dt = np.float64 if precision == stk.FP64 else np.float32
    return inputs.taylor_green_box(np_parcels=n, ncell=32, dtype=dt, boundary=boundary, dt_fluid=1e-2, nsub=10)


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
def test_device_resort_back_to_back(ctx, precision):
    """30 fluid steps queued without any synchronisation: the device re-sorts the cloud
    whenever more than 2% of its parcels were outside their brick, and the results
    still equal the oracle's."""
    case = drifting_case(precision)
    chunk, field = device_cloud(ctx, case, precision)
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1, brick=(4, 4, 4), resort_fraction=0.02)
    calls = 30
    for _ in range(calls):
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
    resorts = chunk.resort_count()
    frac = chunk.stray_fraction()
    state = [a.copy() for a in chunk.download()]
    src = field.download_source()
    chunk.close()
    field.close()
    assert resorts >= 3, f"only {resorts} re-sorts in {calls} steps"
    # the share measured at the start of the last pass: at most the threshold plus one
    # fluid step's drift (re-sorts are taken after every step that exceeded it)
    assert 0 <= frac < 0.02 + 0.2, frac
    ref_state, ref_src = oracle_calls(case, st_oracle.FP64 if precision == stk.FP64 else st_oracle.FP32, calls)
    assert_match(state, src, ref_state, ref_src, f"{calls} back-to-back steps")


def test_no_resort_above_threshold(ctx):
    case = drifting_case(stk.FP32, n=1 << 14)
    chunk, field = device_cloud(ctx, case, stk.FP32)
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1, brick=(4, 4, 4), resort_fraction=1.0)
    for _ in range(5):
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
    assert chunk.resort_count() == 0
    assert chunk.stray_fraction() > 0.02
    chunk.close()
    field.close()


@pytest.mark.parametrize("boundary", [stk.WALL, stk.PERIODIC])
def test_graph_replays_resort_on_device(ctx, boundary):
    """A captured fluid step carries the gated re-sort: replays re-sort the cloud
    without the host, with eager steps and synchronisations between replays."""
    case = drifting_case(stk.FP64, boundary=boundary)
    chunk, field = device_cloud(ctx, case, stk.FP64)
    p = stk.params(boundary=boundary, substeps_per_pass=1, brick=(4, 4, 4), resort_fraction=0.02)
    g = stk.Graph(ctx, chunk, field, case.dt_fluid, case.nsub, p, zero_source=False)
    calls = 0
    for k in range(12):
        g.launch()
        calls += 1
        if k % 4 == 3:
            stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
            calls += 1
            ctx.synchronize()
    resorts = chunk.resort_count()
    state = [a.copy() for a in chunk.download()]
    src = field.download_source()
    g.close()
    chunk.close()
    field.close()
    assert resorts >= 2, resorts
    ref_state, ref_src = oracle_calls(case, st_oracle.FP64, calls)
    assert_match(state, src, ref_state, ref_src, f"{calls} graph and eager steps")


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
@pytest.mark.parametrize("boundary", [stk.WALL, stk.PERIODIC])
@pytest.mark.parametrize("group,brick", [(10, (32, 8, 2)), (3, (4, 4, 4)), (10, (8, 4, 2))])
def test_multi_pass_launch(ctx, precision, boundary, group, brick):
    """Several one-substep passes in one persistent launch (groups of `group`, the last
    group shorter) give the oracle's results exactly."""
    case = drifting_case(precision, boundary=boundary, n=50000)
    chunk, field = device_cloud(ctx, case, precision)
    p = stk.params(boundary=boundary, substeps_per_pass=1, brick=brick, passes_per_launch=group,
                   resort_fraction=0.05)
    calls = 3
    for _ in range(calls):
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
    state = [a.copy() for a in chunk.download()]
    src = field.download_source()
    chunk.close()
    field.close()
    ref_state, ref_src = oracle_calls(case, st_oracle.FP64 if precision == stk.FP64 else st_oracle.FP32, calls)
    assert_match(state, src, ref_state, ref_src, f"passes_per_launch {group}")


def _status(fn, *args):
    r = fn(*args)
    return r, _lib.lib.st_last_error().decode()


def test_graph_pins_cloud_and_field(ctx):
    case = drifting_case(stk.FP32, n=1 << 16)
    chunk, field = device_cloud(ctx, case, stk.FP32)
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1, brick=(4, 4, 4))
    g = stk.Graph(ctx, chunk, field, case.dt_fluid, case.nsub, p)
    lib = _lib.lib
    # another brick would re-sort under the graph: refused before any device work
    q = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1, brick=(8, 4, 2))
    r, msg = _status(lib.st_evolve, ctx.handle, chunk.handle, field.handle, case.dt_fluid, case.nsub, C.byref(q),
                     None)
    assert r == _lib.ST_E_STATE and "destroy them first" in msg
    r, _ = _status(lib.st_cloud_sort, ctx.handle, chunk.handle, field.handle, C.byref(q))
    assert r == _lib.ST_E_STATE
    # the plain method with a small block would grow the captured statistics slots
    q = stk.params(boundary=stk.PERIODIC, method=stk.PLAIN, substeps_per_pass=1, block=64)
    r, msg = _status(lib.st_evolve, ctx.handle, chunk.handle, field.handle, case.dt_fluid, case.nsub, C.byref(q),
                     None)
    assert r == _lib.ST_E_STATE and "statistics slots" in msg
    assert _status(lib.st_cloud_destroy, chunk.handle)[0] == _lib.ST_E_STATE
    assert _status(lib.st_field_destroy, field.handle)[0] == _lib.ST_E_STATE
    # the same params still step, and the graph still runs
    stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
    g.launch()
    ctx.synchronize()
    g.close()
    q = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1, brick=(8, 4, 2))
    stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, q)
    chunk.close()
    field.close()


def test_batch_and_its_graph_pin(ctx):
    cases = [drifting_case(stk.FP32, n=n) for n in (5000, 12000)]
    pairs = [device_cloud(ctx, c, stk.FP32) for c in cases]
    chunks = [c for c, _ in pairs]
    fields = [f for _, f in pairs]
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1)
    batch = stk.Batch(ctx, chunks, fields, p, target_parcels_per_brick=1024)
    g = batch.graph(1e-2, 10)
    lib = _lib.lib
    assert _status(lib.st_batch_destroy, batch.handle)[0] == _lib.ST_E_STATE
    assert _status(lib.st_cloud_destroy, chunks[0].handle)[0] == _lib.ST_E_STATE
    g.launch()
    g.close()
    assert _status(lib.st_cloud_destroy, chunks[1].handle)[0] == _lib.ST_E_STATE  # still in the batch
    batch.close()
    for c in chunks:
        c.close()
    for f in fields:
        f.close()


def test_adjoint_refuses_a_changed_cloud(ctx):
    case = inputs.adjoint_case(np_parcels=4096, ncell=16, nsub=20)
    chunk, field = device_cloud(ctx, case, stk.FP64)
    p = stk.params(boundary=stk.PERIODIC)
    adj = Adjoint(ctx, chunk, field, 20, 5, p)
    adj.forward(case.dt_fluid)
    adj.source_misfit(inputs.target_source(case))
    stk.evolve_device(ctx, chunk, field, case.dt_fluid, 10, p)
    with pytest.raises(_lib.ScaleTrackError) as ei:
        adj.backward()
    assert ei.value.status == _lib.ST_E_STATE
    adj.forward(case.dt_fluid)
    adj.source_misfit(inputs.target_source(case))
    adj.backward()
    adj.gradients()
    adj.close()
    chunk.close()
    field.close()
