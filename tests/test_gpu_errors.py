"""Failure paths of the C-ABI on the GPU: every bad argument, out-of-order call and
non-finite input is refused with its status code before any device work, the message
names the cause, and the same context, cloud and field then take a valid step that
matches the oracle bit for bit.  This is the list of the reference's run-time
self-checks (allocate_array_j's size and type checks, asyncParallelTracking.jl:1444-1465;
the solver's refusal of a missing entry point, createFields.H:28) widened to the whole
boundary: wrong element size or field length, null pointers, NaN/Inf or out-of-domain
parcels, nsub <= 0, a bad time step, mismatched handles and use after destroy."""
import ctypes as C

import numpy as np
import pytest

import scaletrack as stk
from scaletrack import _lib, inputs
from oracle import st_oracle

pytestmark = pytest.mark.gpu

lib = _lib.lib


def status(fn, *args):
    r = fn(*args)
    return r, lib.st_last_error().decode()


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def small_case(precision, boundary=stk.WALL, n=2000):
    dt = np.float64 if precision == stk.FP64 else np.float32
    return inputs.taylor_green_box(np_parcels=n, ncell=8, dtype=dt, boundary=boundary)


def make(ctx, case, precision):
    dt = np.float64 if precision == stk.FP64 else np.float32
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), precision)
    field.upload_U(case.U.astype(dt))
    field.zero_source()
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision)
    chunk.set_state(*[a.astype(dt) for a in case.state()])
    chunk.upload()
    return chunk, field


def assert_valid_step(ctx, chunk, field, case, precision, p):
    """One fluid step after the refused calls: state bit-exact, source to 1e-12."""
    field.zero_source()
    stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
    got = chunk.download()
    src = field.download_source()
    ref_src, _, ref_state = st_oracle.evolve_case(
        case, precision=st_oracle.FP64 if precision == stk.FP64 else st_oracle.FP32,
        boundary=p.boundary)
    for g, r in zip(got, ref_state):
        assert np.array_equal(g, r)
    assert np.abs(src - ref_src).max() <= 1e-12 * np.abs(ref_src).max()


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
def test_step_arguments_refused_then_valid_step(ctx, precision):
    case = small_case(precision)
    chunk, field = make(ctx, case, precision)
    p = stk.params(boundary=stk.WALL)
    bad = [
        ((ctx.handle, chunk.handle, field.handle, case.dt_fluid, 0, C.byref(p), None), _lib.ST_E_INVALID, "nsub"),
        ((ctx.handle, chunk.handle, field.handle, case.dt_fluid, -3, C.byref(p), None), _lib.ST_E_INVALID, "nsub"),
        ((ctx.handle, chunk.handle, field.handle, 0.0, 10, C.byref(p), None), _lib.ST_E_INVALID, "dt_fluid"),
        ((ctx.handle, chunk.handle, field.handle, -1e-3, 10, C.byref(p), None), _lib.ST_E_INVALID, "dt_fluid"),
        ((ctx.handle, chunk.handle, field.handle, float("nan"), 10, C.byref(p), None), _lib.ST_E_INVALID,
         "dt_fluid"),
        ((ctx.handle, chunk.handle, field.handle, float("inf"), 10, C.byref(p), None), _lib.ST_E_INVALID,
         "dt_fluid"),
        ((ctx.handle, None, field.handle, case.dt_fluid, 10, C.byref(p), None), _lib.ST_E_INVALID, "null"),
        ((ctx.handle, chunk.handle, None, case.dt_fluid, 10, C.byref(p), None), _lib.ST_E_INVALID, "null"),
        ((None, chunk.handle, field.handle, case.dt_fluid, 10, C.byref(p), None), _lib.ST_E_INVALID, ""),
    ]
    for args, code, word in bad:
        for fn in (lib.st_evolve, lib.st_evolve_noalloc):
            r, msg = status(fn, *args)
            assert r == code, (args, r, msg)
            assert word in msg
    # parameters out of their ranges
    for k, v, word in (("boundary", 7, "boundary"), ("method", 5, "method"), ("flush", 9, "flush"),
                       ("substeps_per_pass", -1, "substeps_per_pass"), ("cd", float("nan"), "cd"),
                       ("cd", -1.0, "cd"), ("block", 96, "block"), ("block", 2048, "block"),
                       ("brick", (0, 8, 2), "brick"), ("brick", (65, 8, 2), "brick"),
                       ("resort_fraction", float("nan"), "resort_fraction"),
                       ("passes_per_launch", 0, "passes_per_launch")):
        q = stk.params(**{"boundary": stk.WALL, k: v})
        r, msg = status(lib.st_evolve, ctx.handle, chunk.handle, field.handle, case.dt_fluid, 10, C.byref(q), None)
        assert r == _lib.ST_E_INVALID and word in msg, (k, v, r, msg)
    # nothing ran: the source is still zero and the state is the uploaded one
    assert np.all(field.download_source() == 0.0)
    for g, r in zip(chunk.download(), case.state()):
        assert np.array_equal(g, r.astype(g.dtype))
    assert_valid_step(ctx, chunk, field, case, precision, p)
    chunk.close()
    field.close()


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
def test_uploads_refused_then_valid_step(ctx, precision):
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = small_case(precision)
    chunk, field = make(ctx, case, precision)
    state = [a.astype(dt) for a in case.state()]
    # field: element size, null array, non-finite value
    U = case.U.astype(dt)
    for eb in (0, 2, 16):
        r, msg = status(lib.st_field_upload_U, field.handle, ptr(U), eb)
        assert r == _lib.ST_E_INVALID, (eb, msg)
    r, msg = status(lib.st_field_upload_U, field.handle, None, U.itemsize)
    assert r == _lib.ST_E_INVALID
    for v in (np.nan, np.inf, -np.inf):
        Ub = U.copy()
        Ub.reshape(-1)[17] = v
        r, msg = status(lib.st_field_upload_U, field.handle, ptr(Ub), U.itemsize)
        assert r == _lib.ST_E_INVALID and "non-finite" in msg
    # cloud: null arrays, non-finite values, diameters that are not positive
    for k in range(7):
        arrs = [ptr(a) for a in state]
        arrs[k] = None
        r, msg = status(lib.st_cloud_upload, chunk.handle, *arrs, None)
        assert r == _lib.ST_E_INVALID, (k, msg)
    for k in range(1, 7):
        for v in (np.nan, np.inf):
            bad = [a.copy() for a in state]
            bad[k][123] = v
            r, msg = status(lib.st_cloud_upload, chunk.handle, *[ptr(a) for a in bad], None)
            assert r == _lib.ST_E_INVALID and "non-finite" in msg, (k, v, msg)
    for v in (0.0, -1e-5, np.nan):
        bad = [a.copy() for a in state]
        bad[0][5] = v
        r, msg = status(lib.st_cloud_upload, chunk.handle, *[ptr(a) for a in bad], None)
        assert r == _lib.ST_E_INVALID, (v, msg)
    # source download: element sizes and a null output
    out = np.empty(case.ncell * 3, np.float64)
    for eb in (0, 2, 16):
        r, msg = status(lib.st_field_download_source, field.handle, ptr(out), eb)
        assert r == _lib.ST_E_INVALID
    r, msg = status(lib.st_field_download_source, field.handle, None, 8)
    assert r == _lib.ST_E_INVALID
    # the refused uploads left the earlier ones in place
    assert_valid_step(ctx, chunk, field, case, precision, stk.params(boundary=stk.WALL))
    chunk.close()
    field.close()


def test_out_of_domain_parcels_refused(ctx):
    """Wall and channel forms take the reference's cell index floor((x - o) / delta):
    a parcel outside [origin, ending) on any axis is refused at the step (ST_E_RANGE)
    and nothing is stepped; the same upload moved inside then steps normally."""
    precision = stk.FP64
    case = small_case(precision)
    chunk, field = make(ctx, case, precision)
    state = [a.copy() for a in case.state()]
    for axis, v in ((1, case.ending[0]), (2, case.origin[1] - 1e-9), (3, case.ending[2] + 0.5)):
        bad = [a.copy() for a in state]
        bad[axis][7] = v
        chunk.set_state(*bad)
        chunk.upload()
        for method in (stk.PLAIN, stk.SORTED):
            p = stk.params(boundary=stk.WALL, method=method)
            r, msg = status(lib.st_evolve, ctx.handle, chunk.handle, field.handle, case.dt_fluid, 10, C.byref(p),
                            None)
            assert r == _lib.ST_E_RANGE and "outside" in msg, (axis, method, msg)
        assert np.all(field.download_source() == 0.0)
    chunk.set_state(*state)
    chunk.upload()
    assert_valid_step(ctx, chunk, field, case, precision, stk.params(boundary=stk.WALL))
    chunk.close()
    field.close()


def test_out_of_order_and_mismatched_handles(ctx):
    case = small_case(stk.FP64)
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    p = stk.params(boundary=stk.WALL)
    field = stk.TwoWayEulerian(ctx, mesh, stk.FP64)
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, stk.FP64)
    # step before any upload, then before U
    r, msg = status(lib.st_evolve, ctx.handle, chunk.handle, field.handle, case.dt_fluid, 10, C.byref(p), None)
    assert r == _lib.ST_E_STATE and "uploaded" in msg
    chunk.set_state(*case.state())
    chunk.upload()
    r, msg = status(lib.st_evolve, ctx.handle, chunk.handle, field.handle, case.dt_fluid, 10, C.byref(p), None)
    assert r == _lib.ST_E_STATE and "U" in msg
    field.upload_U(case.U)
    field.zero_source()
    # a field of the other precision
    f32 = stk.TwoWayEulerian(ctx, mesh, stk.FP32)
    f32.upload_U(case.U.astype(np.float32))
    r, msg = status(lib.st_evolve, ctx.handle, chunk.handle, f32.handle, case.dt_fluid, 10, C.byref(p), None)
    assert r == _lib.ST_E_INVALID and "precision" in msg
    f32.close()
    # a cloud of another context
    other = stk.Context(0)
    try:
        r, msg = status(lib.st_evolve, other.handle, chunk.handle, field.handle, case.dt_fluid, 10, C.byref(p),
                        None)
        assert r == _lib.ST_E_INVALID and "context" in msg
    finally:
        other.close()
    assert_valid_step(ctx, chunk, field, case, stk.FP64, p)
    chunk.close()
    field.close()


def test_context_used_after_destroy(ctx):
    """A destroyed context refuses every call (ST_E_STATE) and stays allocated while
    clouds or fields made on it live; destroying them afterwards succeeds."""
    case = small_case(stk.FP64, n=500)
    c2 = stk.Context(0)
    chunk, field = make(c2, case, stk.FP64)
    handle = c2.handle
    assert lib.st_ctx_destroy(handle) == _lib.ST_OK
    assert lib.st_ctx_destroy(handle) == _lib.ST_OK  # a second destroy is a no-op
    p = stk.params(boundary=stk.WALL)
    r, msg = status(lib.st_evolve, handle, chunk.handle, field.handle, case.dt_fluid, 10, C.byref(p), None)
    assert r == _lib.ST_E_STATE and "destroyed" in msg
    r, msg = status(lib.st_ctx_synchronize, handle)
    assert r == _lib.ST_E_STATE
    h = C.c_void_p()
    r, msg = status(lib.st_cloud_create, handle, stk.FP64, 10, 1e-3, 1000.0, 1.0, 0.0, 0, C.byref(h))
    assert r == _lib.ST_E_STATE and not h.value
    assert lib.st_cloud_destroy(chunk.handle) == _lib.ST_OK
    assert lib.st_field_destroy(field.handle) == _lib.ST_OK
    chunk.handle = None
    field.handle = None
    c2.handle = None
