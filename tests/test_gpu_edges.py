"""Edge cases of the GPU paths: empty and one-parcel clouds, a non-cubic mesh that no
brick divides, and the size and range limits the library turns into errors."""
import ctypes as C

import numpy as np
import pytest

import scaletrack as stk
from scaletrack import _lib, inputs
from oracle import st_oracle

pytestmark = pytest.mark.gpu


def cloud(ctx, case, precision):
    dt = np.float64 if precision == stk.FP64 else np.float32
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), precision)
    field.upload_U(case.U.astype(dt))
    field.zero_source()
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, precision)
    chunk.set_state(*[a.astype(dt) for a in case.state()])
    chunk.upload()
    return chunk, field


def odd_case(np_parcels, dtype, boundary, seed=31):
    n = (37, 23, 5)
    o, e = (0.0, -0.1, 0.0), (1.0, 0.6, 0.2)
    U = inputs.taylor_green(n, o, e, phase=(0.4, 1.3, 2.0), dtype=dtype)
    x, y, z = inputs.positions(seed, np_parcels, o, e, dtype=dtype)
    d = inputs.lognormal_diameters(seed, np_parcels, dtype=dtype)
    zero = np.zeros(np_parcels, dtype=dtype)
    return inputs.Case(n, o, e, U, d, x, y, z, zero, zero.copy(), zero.copy(), boundary=boundary)


@pytest.mark.parametrize("method", [stk.PLAIN, stk.SORTED])
@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
def test_empty_cloud(ctx, method, precision):
    case = odd_case(0, np.float64 if precision == stk.FP64 else np.float32, 0)
    chunk, field = cloud(ctx, case, precision)
    bb = stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, stk.params(method=method))
    assert np.all(field.download_source() == 0.0)
    assert np.all(bb[0] == np.inf) and np.all(bb[1] == -np.inf)
    chunk.close()
    field.close()


@pytest.mark.parametrize("method", [stk.PLAIN, stk.SORTED])
@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
@pytest.mark.parametrize("boundary", [stk.WALL, stk.PERIODIC])
@pytest.mark.parametrize("np_parcels", [1, 5000])
def test_non_cubic_mesh_partial_bricks(ctx, method, precision, boundary, np_parcels):
    """37x23x5 cells with the default 32x8x2 brick: partial bricks on every axis, and
    with one parcel all bricks but one are empty."""
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = odd_case(np_parcels, dt, boundary)
    chunk, field = cloud(ctx, case, precision)
    p = stk.params(boundary=boundary, method=method, substeps_per_pass=1)
    state = case.state()
    for step in range(3):
        field.zero_source()
        bb = stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
        ref_src, ref_bb, state = st_oracle.evolve_case(
            case, state=state, precision=st_oracle.FP64 if precision == stk.FP64 else st_oracle.FP32,
            substeps_per_pass=1)
        src = field.download_source()
        scale = max(np.abs(ref_src).max(), 1e-300)
        assert np.abs(src - ref_src).max() <= 1e-12 * scale, step
        np.testing.assert_array_equal(np.concatenate(bb), ref_bb)
    for name, a, b in zip("dxyzuvw", chunk.download(), state):
        assert np.array_equal(a, b), name
    chunk.close()
    field.close()


def test_cloud_size_limit(ctx):
    """Parcel arrays are addressed with 32-bit byte offsets: a cloud whose arrays reach
    4 GiB is refused before any allocation."""
    h = C.c_void_p()
    for prec, n in ((stk.FP64, 1 << 29), (stk.FP32, 1 << 30)):
        r = _lib.lib.st_cloud_create(ctx.handle, prec, n, 1e-3, 1000.0, 1.0, 1000.0, 0, C.byref(h))
        assert r == _lib.ST_E_RANGE
        assert b"exceed 4 GiB" in _lib.lib.st_last_error()
    r = _lib.lib.st_cloud_create(ctx.handle, stk.FP64, -1, 1e-3, 1000.0, 1.0, 1000.0, 0, C.byref(h))
    assert r == _lib.ST_E_RANGE


def test_brick_limits(ctx):
    """A brick above 2048 cells, or whose tile with its halo needs more LDS than one
    workgroup may use, is refused with the reason."""
    case = odd_case(100, np.float64, 0)
    chunk, field = cloud(ctx, case, stk.FP64)
    with pytest.raises(_lib.ScaleTrackError, match="at most 2048"):
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, stk.params(brick=(64, 64, 64)))
    with pytest.raises(_lib.ScaleTrackError, match="LDS"):
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, stk.params(brick=(64, 32, 1)))
    chunk.close()
    field.close()


def test_wall_parcel_outside_mesh_is_refused_by_the_sort(ctx):
    """Wall mode keeps parcels inside the box; a parcel uploaded outside it has no brick,
    and the sorted method reports it instead of misplacing it."""
    case = odd_case(100, np.float64, 0)
    case.x[7] = 1.5
    chunk, field = cloud(ctx, case, stk.FP64)
    with pytest.raises(_lib.ScaleTrackError, match="outside the mesh"):
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, stk.params(method=stk.SORTED))
    chunk.close()
    field.close()


@pytest.mark.parametrize("method", [stk.PLAIN, stk.SORTED])
@pytest.mark.parametrize("boundary", [stk.WALL, stk.PERIODIC])
def test_fp32_quotients_on_subnormal_midpoints(ctx, method, boundary):
    """FP32 form with 1 + f = 98 exactly (mu 1, rho 18, d 1, one substep of 97 s) and a
    carrier at rest: a parcel velocity (2j+1) 49 2^-149 makes v / (1 + f) an exact
    midpoint between two float subnormals, where the double product alone rounds to the
    wrong neighbour for many j.  Mixed with zeros and normal velocities, so the shared
    test of the three components is taken for some parcels and not for others.  The
    kernels must reproduce the oracle's IEEE float division bit for bit."""
    rng = np.random.default_rng(41)
    npc = 4096
    n, o, e = (4, 4, 4), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    x, y, z = inputs.positions(41, npc, o, e, dtype=np.float32)

    def comp():
        j = rng.integers(0, 1 << 14, size=npc).astype(np.float64)
        mid = (2 * j + 1) * 49 * 2.0 ** -149 * rng.choice([-1.0, 1.0], size=npc)
        kind = rng.integers(0, 4, size=npc)
        vals = np.where(kind == 0, 0.0, np.where(kind == 1, rng.normal(size=npc), mid))
        return vals.astype(np.float32)

    u, v, w = comp(), comp(), comp()
    d = np.ones(npc, dtype=np.float32)
    U = np.zeros((64, 3), dtype=np.float32)
    case = inputs.Case(n, o, e, U, d, x, y, z, u, v, w, mu_c=1.0, rho_d=18.0, rho_c=18.0, dt_fluid=97.0, nsub=1,
                       boundary=boundary)
    assert np.float32(np.float32(18.0 * 1.0) * np.float32(97.0)) / np.float32(18.0) == np.float32(97.0)
    chunk, field = cloud(ctx, case, stk.FP32)
    stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, stk.params(boundary=boundary, method=method))
    ref_src, _, ref = st_oracle.evolve_case(case, precision=st_oracle.FP32)
    got = chunk.download()
    for name, a, b in zip("dxyzuvw", got, ref):
        bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, (name, bad[:5], a[bad[:5]], b[bad[:5]])
    src = field.download_source()
    scale = max(np.abs(ref_src).max(), 1e-300)
    assert np.abs(src - ref_src).max() <= 1e-12 * scale
    assert np.any(got[4] != 0) and np.any(np.abs(got[4][got[4] != 0]) < 2.0 ** -126)  # subnormal results present
    chunk.close()
    field.close()


@pytest.mark.parametrize("method", [stk.PLAIN, stk.SORTED])
@pytest.mark.parametrize("fused", [1, 4])
def test_fp64_quotients_tiny_zero_and_near_midpoints(ctx, method, fused):
    """FP64 with 1 + f = 98 exactly (mu 97, rho 18, d 1, substeps of 1 s) and a carrier
    at rest, so the drag numerator is the parcel velocity: velocities below 2^-969 (the
    fallback), zero, ordinary, or placed so v / 98 lies next to a rounding midpoint.  The
    reciprocal-and-FMA quotient and its fallback must reproduce the oracle's divisions
    bit for bit over four substeps.  Quotients stay above the double subnormal range
    (2.2e-308): there the GPU's FP64 division itself can differ from IEEE in the last
    place."""
    from fractions import Fraction
    import math

    rng = np.random.default_rng(43)
    npc = 4096
    n, o, e = (4, 4, 4), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    x, y, z = inputs.positions(43, npc, o, e, dtype=np.float64)

    def comp():
        out = np.empty(npc)
        for i in range(npc):
            k = rng.integers(0, 5)
            if k == 0:
                out[i] = rng.uniform(-1, 1) * 10.0 ** rng.uniform(-295, -290)  # below 2^-969
            elif k == 1:
                out[i] = 0.0
            elif k == 2:
                out[i] = rng.normal()
            else:
                q = float(rng.uniform(1, 2) * 2.0 ** int(rng.integers(-40, -5)))
                out[i] = float(Fraction(q + math.ulp(q) / 2) * 98)
        return out

    u, v, w = comp(), comp(), comp()
    d = np.ones(npc)
    U = np.zeros((64, 3))
    case = inputs.Case(n, o, e, U, d, x, y, z, u, v, w, mu_c=97.0, rho_d=18.0, rho_c=18.0, dt_fluid=4.0,
                       nsub=4, boundary=stk.PERIODIC)
    chunk, field = cloud(ctx, case, stk.FP64)
    stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub,
                      stk.params(boundary=stk.PERIODIC, method=method, substeps_per_pass=fused))
    ref_src, _, ref = st_oracle.evolve_case(case, precision=st_oracle.FP64)
    for name, a, b in zip("dxyzuvw", chunk.download(), ref):
        bad = np.flatnonzero(a.view(np.uint64) != b.view(np.uint64))
        assert bad.size == 0, (name, bad[:5], a[bad[:5]], b[bad[:5]])
    src = field.download_source()
    assert np.abs(src - ref_src).max() <= 1e-12 * max(np.abs(ref_src).max(), 1e-300)
    chunk.close()
    field.close()


def layered_case(dtype, boundary, seed=47):
    """24x16x8 cells in 8x8x4 bricks (three brick columns along x): ~10 parcels per
    cell in the first column, ~4 in the second, ~0.2 in the third."""
    n = (24, 16, 8)
    o, e = (0.0, 0.0, 0.0), (1.0, 0.75, 0.4)
    third = (e[0] - o[0]) / 3.0
    parts = []
    for k, count in enumerate((10240, 4096, 400)):
        lo = (o[0] + k * third, o[1], o[2])
        hi = (o[0] + (k + 1) * third, e[1], e[2])
        parts.append(inputs.positions(seed + k, count, lo, hi, dtype=dtype))
    x, y, z = [np.concatenate([p[a] for p in parts]) for a in range(3)]
    npc = x.size
    d = inputs.lognormal_diameters(seed, npc, dtype=dtype)
    zero = np.zeros(npc, dtype=dtype)
    U = inputs.taylor_green(n, o, e, phase=(0.3, 0.9, 0.0), dtype=dtype)
    return inputs.Case(n, o, e, U, d, x, y, z, zero, zero.copy(), zero.copy(), boundary=boundary,
                       dt_fluid=1e-2, nsub=5)


@pytest.mark.parametrize("per", [1, 0])
@pytest.mark.parametrize("precision", [stk.FP64, stk.FP32])
@pytest.mark.parametrize("boundary", [stk.WALL, stk.PERIODIC])
def test_owner_flush_dense_and_sparse_bricks_accumulate(ctx, per, precision, boundary):
    """The owner flush of a brick with at least 6 parcels per own cell starts its sums
    from the source already in HBM; sparser bricks add their sums to it.  Bricks of
    both kinds share the launch, and the source is never zeroed between calls, so
    every call accumulates on what the earlier ones left: the result must be the sum
    of the oracle's per-call sources."""
    dt = np.float64 if precision == stk.FP64 else np.float32
    case = layered_case(dt, boundary)
    cells_per_brick = 8 * 8 * 4
    counts = np.bincount(np.minimum((case.x.astype(np.float64) * 3.0).astype(int), 2), minlength=3) / 4
    assert counts[0] >= 6 * cells_per_brick and counts[1] < 6 * cells_per_brick  # both kinds present
    chunk, field = cloud(ctx, case, precision)
    p = stk.params(boundary=boundary, method=stk.SORTED, substeps_per_pass=per, brick=(8, 8, 4))
    state = case.state()
    total = None
    oprec = st_oracle.FP64 if precision == stk.FP64 else st_oracle.FP32
    for call in range(3):
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, p)
        ref_src, _, state = st_oracle.evolve_case(case, state=state, precision=oprec, substeps_per_pass=per)
        total = ref_src if total is None else total + ref_src
        src = field.download_source()
        scale = max(np.abs(total).max(), 1e-300)
        assert np.abs(src - total).max() <= 1e-12 * scale, f"call {call}"
    for name, a, b in zip("dxyzuvw", chunk.download(), state):
        assert np.array_equal(a, b), name
    chunk.close()
    field.close()


def test_tile_lds_check_counts_static_lds(ctx):
    """An FP64 32 x 8 x 8 brick: its halo tile (34 * 10 * 10 cells * 48 B = 163,200 B) fits
    the 160 KiB a workgroup may use, but not with the kernel's own static LDS; the step
    refuses it with ST_E_INVALID before launching, instead of failing on the device."""
    from scaletrack import _lib

    case = inputs.taylor_green_box(np_parcels=4096, ncell=32, boundary=1)
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    field = stk.TwoWayEulerian(ctx, mesh, stk.FP64)
    field.upload_U(case.U)
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, stk.FP64)
    chunk.set_state(*case.state())
    chunk.upload()
    with pytest.raises(_lib.ScaleTrackError) as ei:
        stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, stk.params(boundary=1, brick=(32, 8, 8)))
    assert ei.value.status == _lib.ST_E_INVALID and "LDS" in str(ei.value)
    stk.evolve_device(ctx, chunk, field, case.dt_fluid, case.nsub, stk.params(boundary=1, brick=(32, 8, 4)))
    chunk.close()
    field.close()
