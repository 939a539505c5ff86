"""st_evolve_noalloc (the step for callers that cannot allocate) against st_evolve,
both called through the C-ABI on identical clouds and fields.

Two calls in a row per form, the second on a smaller cloud and mesh (fewer statistics
slots than the first left in the shared scratch array).  Bars: parcel state and bounding
box bit-exact, sources within 1e-12 of the largest |source| (only the order of FP64
atomic additions may differ between two runs).
"""
import ctypes as C

import numpy as np
import pytest

import scaletrack as stk
from scaletrack import _lib, inputs

pytestmark = pytest.mark.gpu

SRC_RTOL = 1e-12
lib = _lib.lib


def make(ctx, case, precision):
    dt = np.float64 if precision == stk.FP64 else np.float32
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), precision)
    field.upload_U(case.U.astype(dt))
    field.zero_source()
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision)
    chunk.set_state(*[a.astype(dt) for a in case.state()])
    chunk.upload()
    return chunk, field


def step(fn, ctx, chunk, field, case, p):
    bb = (C.c_double * 6)()
    _lib.check(fn(ctx.handle, chunk.handle, field.handle, case.dt_fluid, case.nsub, C.byref(p), bb))
    return np.array(bb[:])


@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
@pytest.mark.parametrize("method", [stk.SORTED, stk.PLAIN])
def test_noalloc_matches_evolve(ctx, precision, method):
    dt = np.float64 if precision == stk.FP64 else np.float32
    cases = [inputs.taylor_green_box(np_parcels=1 << 17, ncell=32, dtype=dt, boundary=stk.PERIODIC, dt_fluid=1e-2),
             inputs.taylor_green_box(np_parcels=3000, ncell=12, dtype=dt, boundary=stk.WALL, dt_fluid=1e-2)]
    for k, case in enumerate(cases):
        p = stk.params(boundary=case.boundary, method=method, substeps_per_pass=1, brick=(8, 4, 2),
                       block=256, resort_fraction=0.02)
        out = []
        for fn in (lib.st_evolve, lib.st_evolve_noalloc):
            chunk, field = make(ctx, case, precision)
            if method == stk.SORTED:
                chunk.sort(ctx, field, p)  # the no-allocation step takes a sorted cloud
            bbs = [step(fn, ctx, chunk, field, case, p) for _ in range(3)]
            out.append(([a.copy() for a in chunk.download()], field.download_source(), bbs,
                        chunk.resort_count() if method == stk.SORTED else 0))
            chunk.close()
            field.close()
        (s1, src1, bb1, r1), (s2, src2, bb2, r2) = out
        for name, a, b in zip("dxyzuvw", s1, s2):
            assert np.array_equal(a, b), f"call {k}: {name} differs"
        err = np.abs(src1 - src2).max() / max(np.abs(src1).max(), 1e-300)
        assert err <= SRC_RTOL, f"call {k}: source rel err {err:.3e}"
        for a, b in zip(bb1, bb2):
            np.testing.assert_array_equal(a, b)
        assert r1 == r2


def test_noalloc_refusals(ctx):
    case = inputs.taylor_green_box(np_parcels=1 << 12, ncell=16, dtype=np.float32, boundary=stk.PERIODIC)
    chunk, field = make(ctx, case, stk.FP32)
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1, brick=(4, 4, 4))
    bb = (C.c_double * 6)()
    # not yet sorted for this brick: the sort would allocate
    assert lib.st_evolve_noalloc(ctx.handle, chunk.handle, field.handle, 1e-3, 10, C.byref(p), bb) == _lib.ST_E_STATE
    assert b"st_cloud_sort" in lib.st_last_error()
    chunk.sort(ctx, field, p)
    _lib.check(lib.st_evolve_noalloc(ctx.handle, chunk.handle, field.handle, 1e-3, 10, C.byref(p), bb))
    chunk.close()
    field.close()
    # more statistics slots than the scratch holds: 65536 * 64 + 1 parcels, 64 per workgroup
    n = 65536 * 64 + 1
    big = stk.Chunk(ctx, n, 1e-3, 1000.0, 1.0, stk.FP32)
    big.set_state(np.full(n, 3e-5, np.float32), *[np.full(n, 0.5, np.float32)] * 3,
                  *[np.zeros(n, np.float32)] * 3)
    big.upload()
    fld = stk.TwoWayEulerian(ctx, stk.Mesh(16, 0.0, 1.0), stk.FP32)
    fld.upload_U(np.zeros((16 ** 3, 3), np.float32))
    q = stk.params(boundary=stk.PERIODIC, method=stk.PLAIN, block=64)
    assert lib.st_evolve_noalloc(ctx.handle, big.handle, fld.handle, 1e-3, 10, C.byref(q), bb) == _lib.ST_E_RANGE
    assert b"65536" in lib.st_last_error()
    big.close()
    fld.close()
