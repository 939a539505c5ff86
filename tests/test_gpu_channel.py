"""The channel workload (walls on y and z, open inlet and outlet) on the GPU against the
oracle's channel form: parcels that leave are removed (negative diameter) and never
touched again; parcels enter through st_cloud_inject and removed ones are dropped by
st_cloud_compact.  The bars are the usual ones: state (removed marks included) and
bounding box bit-exact, sources within 1e-12 of the largest |source|.

The last two tests run the channel at one GPU's shape: 2^25 parcels on 256 x 128 x 128
cells, one fluid step of ten one-substep passes, in the FP32 form with open planes and
in FP64 with walls on every face, the oracle on the host's cores.
"""
import os

import numpy as np
import pytest

import scaletrack as stk
from oracle import st_oracle
from scaletrack import coupling, inputs

pytestmark = pytest.mark.gpu

SRC_RTOL = 1e-12
THREADS = max(1, min(16, os.cpu_count() or 1))


def oprec(precision):
    return st_oracle.FP64 if precision == stk.FP64 else st_oracle.FP32


def device(ctx, case, precision):
    dt = np.float64 if precision == stk.FP64 else np.float32
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), precision)
    field.upload_U(case.U.astype(dt))
    field.zero_source()
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision)
    chunk.set_state(*[a.astype(dt) for a in case.state()])
    chunk.upload()
    return chunk, field


def oracle_step(case, st, precision, substeps_per_pass=1, boundary=st_oracle.CHANNEL, nthreads=1):
    src = np.zeros((case.ncell, 3))
    _, bb = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                             precision=oprec(precision), boundary=boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                             rho_d_by_rho_c=case.rho_d_by_rho_c, source=src, substeps_per_pass=substeps_per_pass,
                             nthreads=nthreads)
    return src, bb


def assert_step(got, gsrc, gbb, ref, rsrc, rbb, what):
    for name, a, b in zip("dxyzuvw", got, ref):
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{what}: {name} differs for {bad.size} parcels (first {bad[:4]})"
    err = np.abs(gsrc - rsrc).max() / max(np.abs(rsrc).max(), 1e-300)
    assert err <= SRC_RTOL, f"{what}: source rel err {err:.3e}"
    if gbb is not None:
        np.testing.assert_array_equal(np.concatenate(gbb), rbb)


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
@pytest.mark.parametrize("method", [stk.PLAIN, stk.SORTED])
@pytest.mark.parametrize("per", [1, 0])
def test_channel_parity(ctx, precision, method, per):
    """Fluid steps of 0.02 s on a 32 x 16 x 16 channel: parcels near the outlet leave
    during each call; removed parcels stay untouched in later calls."""
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = inputs.channel_case(np_parcels=1 << 16, n=(32, 16, 16), dtype=dt, dt_fluid=2e-2)
    chunk, field = device(ctx, case, precision)
    p = stk.params(boundary=stk.CHANNEL, method=method, substeps_per_pass=per, brick=(8, 4, 4),
                   resort_fraction=0.05)
    st = case.state()
    for k in range(4):
        field.zero_source()
        bb = stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
        got = chunk.download()
        rsrc, rbb = oracle_step(case, st, precision, substeps_per_pass=per)
        assert_step(got, field.download_source(), bb, st, rsrc, rbb, f"call {k}")
    removed = int((st[0] < 0).sum())
    assert removed > 200, removed
    chunk.close()
    field.close()


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
def test_inject_and_compact(ctx, precision):
    """A channel filled by its inflow: each fluid step injects parcels at the inlet, every
    third step drops the removed ones.  The cloud, in upload order, equals the oracle's
    list kept the same way (appended, then filtered with its order kept)."""
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = inputs.channel_case(np_parcels=3000, n=(16, 8, 8), length=(0.25, 1.0, 1.0), dtype=dt, dt_fluid=2e-2)
    chunk, field = device(ctx, case, precision)
    p = stk.params(boundary=stk.CHANNEL, substeps_per_pass=1, brick=(4, 4, 4), resort_fraction=0.05)
    st = case.state()
    drops = 0
    for k in range(1, 13):
        new = inputs.channel_injection(case, k, 800)
        chunk.inject(*new)
        st = [np.concatenate([a, b]) for a, b in zip(st, new)]
        field.zero_source()
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
        rsrc, _ = oracle_step(case, st, precision)
        got = chunk.download()
        assert_step(got, field.download_source(), None, st, rsrc, None, f"step {k}")
        if k % 3 == 0:
            live = st[0] > 0
            drops += int((~live).sum())
            st = [a[live] for a in st]
            assert chunk.compact() == st[0].size
            got = chunk.download()
            for a, b in zip(got, st):
                np.testing.assert_array_equal(a, b)
    assert drops > 500, drops
    chunk.close()
    field.close()


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
@pytest.mark.parametrize("resort_fraction,flush", [(0.02, stk.FLUSH_OWNER), (1.0, stk.FLUSH_OWNER),
                                                   (1.0, stk.FLUSH_ATOMIC)])
def test_injected_tail(ctx, precision, resort_fraction, flush):
    """Parcels injected into a sorted cloud are stepped as its unsorted tail (one thread
    per parcel, in workgroups after the bricks in each pass) until a re-sort decided on
    the device takes them in (fraction 0.02: within a few steps; 1.0: never, the tail
    grows for the whole run), under the owner flush (tail deposits into the stray array)
    and the atomic flush (tail deposits into the source).
    Steps are queued back to back; every step's parcels and the summed source equal the
    oracle's.  A graph captured with a tail and the no-allocation step take part too."""
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = inputs.channel_case(np_parcels=20000, n=(32, 16, 16), length=(1.0, 1.0, 1.0), dtype=dt, dt_fluid=1e-2)
    chunk, field = device(ctx, case, precision)
    p = stk.params(boundary=stk.CHANNEL, substeps_per_pass=1, brick=(8, 4, 4), resort_fraction=resort_fraction,
                   flush=flush)
    st = case.state()
    src = np.zeros((case.ncell, 3))
    from scaletrack import _lib
    import ctypes as C
    for k in range(1, 9):
        new = inputs.channel_injection(case, k, 1500)
        chunk.inject(*new)
        st = [np.concatenate([a, b]) for a, b in zip(st, new)]
        for _ in range(2):
            if k == 4:  # the no-allocation step on a cloud with a tail
                _lib.check(_lib.lib.st_evolve_noalloc(ctx.handle, chunk.handle, field.handle, case.dt_fluid,
                                                      case.nsub, C.byref(p), None))
            else:
                stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
            s, _ = oracle_step(case, st, precision)
            src += s
        if k == 6:  # a captured step with the tail, replayed twice
            g = stk.Graph(ctx, chunk, field, case.dt_fluid, case.nsub, p, zero_source=False)
            for _ in range(2):
                g.launch()
                s, _ = oracle_step(case, st, precision)
                src += s
            g.close()
    got = chunk.download()
    assert_step(got, field.download_source(), None, st, src, None, f"fraction {resort_fraction}")
    resorts = chunk.resort_count()
    if resort_fraction < 1:
        assert resorts >= 2, resorts
    else:
        assert resorts == 0
    chunk.close()
    field.close()


def test_async_tracker_channel_100_steps(ctx):
    """The async tracker drives the channel for 100 fluid steps with an inflow of 400
    parcels per step and compaction every 5 steps: after every evolve_cloud the solver
    sees the extrapolated source of the previous parcel step, and at the end the cloud
    equals the oracle's."""
    case = inputs.channel_case(np_parcels=2000, n=(16, 8, 8), length=(0.5, 1.0, 1.0), dtype=np.float64,
                               dt_fluid=1e-2)
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    st = case.state()
    host = {"N": case.np, "mu_c": case.mu_c, "rho": case.rho_d, "rho_d_by_rho_c": case.rho_d_by_rho_c}
    for k, a in zip(("d", "X", "Y", "Z", "U", "V", "W"), st):
        host[k] = a.copy()
    tr = coupling.AsyncTracker(mesh, [host], precision=stk.FP64, nsub=case.nsub,
                               params=stk.params(boundary=stk.CHANNEL, substeps_per_pass=1, brick=(4, 4, 4)),
                               injector=lambda step: [inputs.channel_injection(case, step, 400)], compact_every=5)
    try:
        est_prev = np.zeros((case.ncell, 3))
        sources = []
        steps = 100
        for k in range(steps + 1):
            tr.U[...] = case.U
            tr.evolve_cloud(case.dt_fluid)
            if k >= 1:
                est = 2.0 * sources[k - 1] - est_prev
                scale = max(float(np.abs(est).max()), 1e-300)
                assert np.abs(tr.UTrans - est).max() <= 1e-12 * scale, f"call {k + 1}"
                est_prev = est
            if k < steps:
                new = inputs.channel_injection(case, k + 1, 400)
                st = [np.concatenate([a, b]) for a, b in zip(st, new)]
                s, _ = oracle_step(case, st, stk.FP64)
                sources.append(s)
                if (k + 1) % 5 == 0:
                    live = st[0] > 0
                    st = [a[live] for a in st]
    finally:
        tr.finalize_evolve()
    final = tr.executor.final[0]
    # the tracker ran one more parcel step than the solver consumed (the last call's)
    new = inputs.channel_injection(case, steps + 1, 400)
    st = [np.concatenate([a, b]) for a, b in zip(st, new)]
    oracle_step(case, st, stk.FP64)
    for name, a, b in zip("dxyzuvw", final, st):
        np.testing.assert_array_equal(a, b, err_msg=name)


def _full(ctx, precision, boundary):
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = inputs.channel_case(np_parcels=1 << 25, dtype=dt)
    case.boundary = boundary
    chunk, field = device(ctx, case, precision)
    p = stk.params(boundary=boundary, substeps_per_pass=1)
    field.zero_source()
    bb = stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
    got = [a.copy() for a in chunk.download()]
    gsrc = field.download_source()
    chunk.close()
    field.close()
    st = case.state()
    rsrc, rbb = oracle_step(case, st, precision, boundary=boundary, nthreads=THREADS)
    # the oracle's OpenMP sum runs in another order than its serial one: 1e-12 still holds
    assert_step(got, gsrc, bb, st, rsrc, rbb, "full channel step")
    return st


def test_channel_full_shape_fp32(ctx):
    st = _full(ctx, stk.FP32, stk.CHANNEL)
    assert int((st[0] < 0).sum()) > 1000  # the outlet takes parcels in one fluid step


def test_channel_full_shape_fp64_walls(ctx):
    _full(ctx, stk.FP64, stk.WALL)


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
def test_injected_parcels_outside_the_mesh(ctx, precision):
    """A parcel injected outside the mesh (here exactly on the outlet plane, whose cell
    index is N) is refused: by st_cloud_inject itself when the cloud is sorted (it
    knows its mesh; nothing is appended), else by the next step's check on the device,
    as for an upload."""
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = inputs.channel_case(np_parcels=2000, n=(16, 8, 8), length=(0.5, 1.0, 1.0), dtype=dt, dt_fluid=1e-2)
    p = stk.params(boundary=stk.CHANNEL, substeps_per_pass=1, brick=(4, 4, 4))
    for sort_first in (True, False):
        chunk, field = device(ctx, case, precision)
        if sort_first:
            stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
        new = inputs.channel_injection(case, 1, 100)
        new[1][17] = dt(case.ending[0])
        n0 = chunk.count()
        if sort_first:
            with pytest.raises(_lib_error(), match="outside the mesh"):
                chunk.inject(*new)
            assert chunk.count() == n0
        else:
            chunk.inject(*new)
            with pytest.raises(_lib_error(), match="outside the mesh"):
                stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p, want_bbox=False)
        chunk.close()
        field.close()


def _lib_error():
    from scaletrack import _lib
    return _lib.ScaleTrackError
