"""GPU (libscaletrack.so) against the CPU oracle on the same seeded inputs.

FP64: positions, velocities and cells bit-exact (both sides compile without FMA
contraction and use IEEE division and floor); sources within 1e-12 relative to the
largest |source| (only the order of the FP64 additions differs).
FP32 form: bit-exact against the oracle's FP32 form as well (same operations).
"""
import numpy as np
import pytest

import scaletrack as stk
from scaletrack import inputs
from oracle import st_oracle

pytestmark = pytest.mark.gpu

SRC_RTOL = 1e-12


def run_gpu(ctx, case, precision, p, calls=1, zero_each=True):
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    field = stk.TwoWayEulerian(ctx, mesh, precision)
    field.upload_U(case.U)
    field.zero_source()
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision,
                      rho_p=case.rho, rho_c=case.rho_c)
    chunk.set_state(*case.state())
    chunk.upload()
    srcs, bbs = [], []
    for _ in range(calls):
        if zero_each:
            field.zero_source()
        bb = stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
        srcs.append(field.download_source().copy())
        bbs.append(bb)
    state = [a.copy() for a in chunk.download()]
    field.close()
    chunk.close()
    return state, srcs, bbs


def run_oracle(case, precision, calls=1, substeps_per_pass=0, boundary=None, cd=1.0):
    st = case.state()
    srcs, bbs = [], []
    for _ in range(calls):
        src, bb, st = st_oracle.evolve_case(case, state=st, precision=precision,
                                            substeps_per_pass=substeps_per_pass,
                                            boundary=boundary, cd=cd)
        srcs.append(src)
        bbs.append(bb)
    return st, srcs, bbs


def assert_parity(g, r, exact=True):
    gs, gsrc, gbb = g
    rs, rsrc, rbb = r
    names = ["d", "x", "y", "z", "u", "v", "w"]
    for name, a, b in zip(names, gs, rs):
        if exact:
            bad = np.flatnonzero(a != b)
            assert bad.size == 0, f"{name}: {bad.size} parcels differ, first {bad[:5]} gpu={a[bad[:3]]} ref={b[bad[:3]]}"
    for k, (a, b) in enumerate(zip(gsrc, rsrc)):
        scale = max(np.abs(b).max(), 1e-300)
        err = np.abs(a - b).max() / scale
        assert err <= SRC_RTOL, f"call {k}: source max rel err {err:.3e}"
    for a, b in zip(gbb, rbb):
        np.testing.assert_array_equal(np.concatenate(a), b)


CASES = {
    "tg16-wall": lambda dt: inputs.taylor_green_box(np_parcels=20000, ncell=16, dtype=dt, boundary=0),
    "tg16-periodic": lambda dt: inputs.taylor_green_box(np_parcels=20000, ncell=16, dtype=dt, boundary=1),
    "cavity": lambda dt: inputs.cavity_case(np_parcels=10000, dtype=dt),
}


@pytest.mark.parametrize("method", [stk.PLAIN, stk.SORTED])
@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
@pytest.mark.parametrize("name", list(CASES))
def test_parity_basic(ctx, name, precision, method):
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = CASES[name](dt)
    p = stk.params(boundary=case.boundary, method=method)
    g = run_gpu(ctx, case, precision, p, calls=3)
    r = run_oracle(case, precision, calls=3)
    assert_parity(g, r)


@pytest.mark.parametrize("per", [1, 3])
def test_parity_pass_split(ctx, per):
    """Substeps split over several launches: the oracle runs the same split."""
    case = inputs.taylor_green_box(np_parcels=20000, ncell=16, boundary=1)
    p = stk.params(boundary=1, substeps_per_pass=per)
    g = run_gpu(ctx, case, stk.FP64, p, calls=2)
    r = run_oracle(case, stk.FP64, calls=2, substeps_per_pass=per)
    assert_parity(g, r)


@pytest.mark.parametrize("flush", [stk.FLUSH_OWNER, stk.FLUSH_ATOMIC])
@pytest.mark.parametrize("brick", [(4, 4, 4), (8, 8, 8), (5, 3, 7), (16, 2, 1)])
def test_parity_bricks(ctx, flush, brick):
    case = inputs.taylor_green_box(np_parcels=30000, ncell=20, boundary=0)
    p = stk.params(boundary=0, flush=flush, brick=brick, block=128)
    g = run_gpu(ctx, case, stk.FP64, p, calls=2)
    r = run_oracle(case, stk.FP64, calls=2)
    assert_parity(g, r)


HAND = __import__("json").load(open(__import__("os").path.join(
    __import__("os").path.dirname(__file__), "golden", "hand_cases.json")))["cases"]


@pytest.mark.parametrize("method", [stk.PLAIN, stk.SORTED])
@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
@pytest.mark.parametrize("case", HAND, ids=[c["name"][:40] for c in HAND])
def test_hand_cases_gpu(ctx, case, precision, method):
    """The hand-derived single-parcel cases (tests/golden/hand_cases.json) on the GPU."""
    import math

    dt = np.float64 if precision == stk.FP64 else np.float32
    m = case["mesh"]
    ph = case["physics"]
    par = case["parcel"]
    mesh = stk.Mesh(m["n"], m["origin"], m["ending"])
    field = stk.TwoWayEulerian(ctx, mesh, precision)
    field.upload_U(np.tile(np.asarray(case["U_uniform"], dtype=dt), (mesh.ncell, 1)))
    field.zero_source()
    chunk = stk.Chunk(ctx, 1, ph["mu_c"], ph["rho_d"], ph["rho_d_by_rho_c"], precision)
    chunk.set_state(*[np.array([v], dtype=dt) for v in [par["d"]] + par["x"] + par["v"]])
    chunk.upload()
    bnd = stk.PERIODIC if case.get("boundary") == "periodic" else stk.WALL
    stk.evolve_device(ctx, chunk, field, ph["dt_fluid"], ph["nsub"],
                      stk.params(boundary=bnd, method=method, brick=(2, 2, 2), block=64))
    st = chunk.download()
    src = field.download_source()
    ex = case["expect"]
    assert [float(st[1][0]), float(st[2][0]), float(st[3][0])] == ex["x"]
    assert [float(st[4][0]), float(st[5][0]), float(st[6][0])] == ex["v"]
    mass = ph["rho_d_by_rho_c"] * math.pi * par["d"] ** 3 / 6
    exp = np.zeros((mesh.ncell, 3))
    for c, vals in ex["src_over_m"].items():
        exp[int(c)] = np.asarray(vals) * mass
    np.testing.assert_allclose(src, exp, rtol=1e-6 if precision == stk.FP32 else 1e-15, atol=0)


def test_kernel_times_ring(ctx):
    """st_diag_kernel_times: one entry per timed call, oldest first, the newest equal to
    st_last_kernel_ms; the ring keeps the last 64 calls."""
    case = inputs.taylor_green_box(np_parcels=20000, ncell=16, boundary=1)
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    field = stk.TwoWayEulerian(ctx, mesh, stk.FP64)
    field.upload_U(case.U)
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, stk.FP64)
    chunk.set_state(*case.state())
    chunk.upload()
    p = stk.params(boundary=stk.PERIODIC, substeps_per_pass=1)
    for k in range(70):
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, 1 + k % 3, p, want_bbox=False)
    times = ctx.kernel_times(100)
    assert len(times) == 64
    assert [n for _, n in times] == [1 + k % 3 for k in range(6, 70)]
    assert all(ms > 0 for ms, _ in times)
    last = ctx.last_kernel_ms()
    assert last == times[-1]
    assert len(ctx.kernel_times(5)) == 5 and ctx.kernel_times(5)[-1] == times[-1]
    chunk.close()
    field.close()
