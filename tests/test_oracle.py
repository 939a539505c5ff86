"""Pins the CPU oracle (oracle/st_oracle.cpp) without running the reference.

The reference (Julia) cannot run here and has no tests or golden vectors, so each
part of the oracle is pinned by what the mathematics or a hand derivation fixes:
  * locate: brute force over every cell of tiny meshes with exact rational arithmetic,
    plus the exact face / underflow / overflow rules of asyncParallelTracking.jl:663-697;
  * drag and advance: the closed-form geometric series of the implicit Euler step in a
    constant velocity, and its f -> 0 (ballistic) and f -> inf (tracer) limits;
  * mass: the sphere volume (4/3) pi r^3, an expression different from the oracle's;
  * bounce, last-substep rule, folded-index quirk, periodic wrap: hand-derived cases
    in tests/golden/hand_cases.json (exact binary numbers);
  * deposit: momentum balance, linearity over parcels, pass-splitting (telescoping);
  * extrapolator: the closed form est_n = 2 S_{n-1} - est_{n-1}.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import st_oracle
from scaletrack import inputs

HERE = os.path.dirname(os.path.abspath(__file__))
HAND = json.load(open(os.path.join(HERE, "golden", "hand_cases.json")))["cases"]


# ---------------------------------------------------------------- locate

MESHES = [
    ((2, 2, 2), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
    ((3, 4, 5), (-0.3, 0.1, 2.0), (0.7, 0.5, 2.6)),
    ((20, 20, 1), (0.0, 0.0, 0.0), (0.1, 0.1, 0.01)),
]


@pytest.mark.parametrize("n,o,e", MESHES)
def test_locate_every_cell_centre(n, o, e):
    """Every cell centre maps to its own (i, j, k) and to k*Nx*Ny + j*Nx + i (x fastest,
    OpenFOAM blockMesh order), in partition 0."""
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                p = [o[a] + (idx + 0.5) * (e[a] - o[a]) / n[a] for a, idx in enumerate((i, j, k))]
                ijk, cell, part = st_oracle.locate(n, o, e, *p)
                assert ijk == (i, j, k)
                assert cell == k * n[0] * n[1] + j * n[0] + i
                assert part == 0


@pytest.mark.parametrize("n,o,e", MESHES)
def test_locate_random_points_exact(n, o, e):
    """Random points (away from faces by more than 1e-9 cells) against the exact
    rational floor((x - o) / delta)."""
    rng = np.random.default_rng(3)
    for _ in range(2000):
        p = [o[a] + rng.uniform(-0.2, 1.2) * (e[a] - o[a]) for a in range(3)]
        exact = []
        skip = False
        for a in range(3):
            t = (Fraction(p[a]) - Fraction(o[a])) * n[a] / (Fraction(e[a]) - Fraction(o[a]))
            fl = math.floor(t)
            if min(t - fl, fl + 1 - t) < Fraction(1, 10**9):
                skip = True
            exact.append(fl)
        if skip:
            continue
        ijk, cell, part = st_oracle.locate(n, o, e, *p)
        assert list(ijk) == exact
        # fold: divrem truncates toward zero, then abs of both parts
        loc = [abs(int(math.fmod(I, N))) for I, N in zip(exact, n)]
        par = [abs(int(I / N)) for I, N in zip(exact, n)]
        assert cell == loc[2] * n[0] * n[1] + loc[1] * n[0] + loc[0]
        assert part == sum(par)


def test_locate_faces_and_outside():
    n, o, e = (4, 4, 4), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    # exactly on an interior face: the upper cell (IEEE floor)
    assert st_oracle.locate(n, o, e, 0.5, 0.25, 0.75)[0] == (2, 1, 3)
    # one ulp below the face: the lower cell
    assert st_oracle.locate(n, o, e, np.nextafter(0.5, 0), 0.1, 0.1)[0][0] == 1
    # x in (-delta, 0): I = -1 (floor, not truncation), folded to local 1, partition 0
    ijk, cell, part = st_oracle.locate(n, o, e, -0.1, 0.1, 0.1)
    assert ijk == (-1, 0, 0) and cell == 1 and part == 0
    # x in [ending, ending + delta): I = N, local 0, partition 1
    ijk, cell, part = st_oracle.locate(n, o, e, 1.0, 0.1, 0.1)
    assert ijk == (4, 0, 0) and cell == 0 and part == 1
    ijk, cell, part = st_oracle.locate(n, o, e, 0.1, 0.1, 1.1)
    assert ijk == (0, 0, 4) and cell == 0 and part == 1


# ---------------------------------------------------------------- drag & advance

def one_parcel(d, x, v, U, n=(1, 1, 1), o=(0.0, 0.0, 0.0), e=(1.0, 1.0, 1.0), **kw):
    st = [np.array([d])] + [np.array([c], dtype=np.float64) for c in list(x) + list(v)]
    Uf = np.tile(np.asarray(U, dtype=np.float64), (n[0] * n[1] * n[2], 1))
    src, bb = st_oracle.evolve(st, Uf, n, o, e, **kw)
    return st, src, bb


def test_closed_form_series_in_one_cell():
    """Constant u, r = 1/(1+f): v_n = u + (v0 - u) r^n and
    x_n = x0 + dt [n u + (v0 - u) r (1 - r^n) / (1 - r)]."""
    mu, rho, d = 1.8e-5, 1200.0, 40e-6
    dtf, nsub = 2e-3, 10
    dt = dtf / nsub
    f = 18 * mu * dt / (rho * d * d)
    r = 1 / (1 + f)
    u = np.array([0.3, -0.2, 0.1])
    v0 = np.array([-0.1, 0.05, 0.2])
    x0 = np.array([0.4, 0.5, 0.6])
    st, src, _ = one_parcel(d, x0, v0, u, dt_fluid=dtf, nsub=nsub, mu_c=mu, rho_d=rho,
                            rho_d_by_rho_c=rho / 1.2)
    vn = u + (v0 - u) * r**nsub
    xn = x0 + dt * (nsub * u + (v0 - u) * r * (1 - r**nsub) / (1 - r))
    got_v = np.array([st[4][0], st[5][0], st[6][0]])
    got_x = np.array([st[1][0], st[2][0], st[3][0]])
    np.testing.assert_allclose(got_v, vn, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(got_x, xn, rtol=1e-14)
    # deposit over the single residence: (mass / rho_c) (v0 - v_n), mass of a sphere
    mass_by_rhoc = (rho / 1.2) * (4.0 / 3.0) * math.pi * (d / 2) ** 3
    np.testing.assert_allclose(src[0], mass_by_rhoc * (v0 - vn), rtol=1e-12)


def test_ballistic_limit():
    """mu = 0: f = 0, velocity unchanged, straight-line motion, no deposit."""
    v0 = np.array([0.25, -0.5, 0.125])
    st, src, _ = one_parcel(1e-4, (0.5, 0.5, 0.5), v0, (1.0, 1.0, 1.0), dt_fluid=0.08, nsub=8, mu_c=0.0)
    np.testing.assert_array_equal([st[4][0], st[5][0], st[6][0]], v0)
    np.testing.assert_allclose([st[1][0], st[2][0], st[3][0]], 0.5 + v0 * 0.08, rtol=1e-15)
    assert not src.any()


def test_tracer_limit():
    """f -> infinity (tiny parcel): the parcel moves with the carrier."""
    u = np.array([0.25, -0.125, 0.5])
    st, src, _ = one_parcel(1e-9, (0.25, 0.5, 0.25), (0, 0, 0), u, dt_fluid=0.1, nsub=4, mu_c=1.0,
                            rho_d=1.0)
    np.testing.assert_allclose([st[4][0], st[5][0], st[6][0]], u, rtol=1e-12)
    np.testing.assert_allclose([st[1][0], st[2][0], st[3][0]], np.array([0.25, 0.5, 0.25]) + u * 0.1,
                               rtol=1e-12)


def test_cd_multiplier_scales_f():
    """cd multiplies the Stokes factor: cd = 2 at dt equals cd = 1 for a factor 2f,
    which the closed form gives as r = 1/(1+2f)."""
    mu, rho, d, dtf = 1e-3, 1000.0, 30e-6, 1e-4
    u = np.array([1.0, 0.0, 0.0])
    st, _, _ = one_parcel(d, (0.5, 0.5, 0.5), (0, 0, 0), u, dt_fluid=dtf, nsub=1, mu_c=mu, rho_d=rho,
                          cd=2.0)
    f = 2 * 18 * mu * dtf / (rho * d * d)
    np.testing.assert_allclose(st[4][0], f / (1 + f), rtol=1e-14)


# ---------------------------------------------------------------- hand cases

def run_hand(case, precision=st_oracle.FP64):
    dt = np.float64 if precision == st_oracle.FP64 else np.float32
    m = case["mesh"]
    n = tuple(m["n"])
    ph = case["physics"]
    par = case["parcel"]
    st = [np.array([par["d"]], dtype=dt)] + [np.array([c], dtype=dt) for c in par["x"] + par["v"]]
    U = np.tile(np.asarray(case["U_uniform"], dtype=dt), (n[0] * n[1] * n[2], 1))
    bnd = st_oracle.PERIODIC if case.get("boundary") == "periodic" else st_oracle.WALL
    src, _ = st_oracle.evolve(st, U, n, m["origin"], m["ending"], ph["dt_fluid"], nsub=ph["nsub"],
                              precision=precision, boundary=bnd, mu_c=ph["mu_c"], rho_d=ph["rho_d"],
                              rho_d_by_rho_c=ph["rho_d_by_rho_c"])
    return st, src


def expected_hand(case):
    n = case["mesh"]["n"]
    d = case["parcel"]["d"]
    mass = case["physics"]["rho_d_by_rho_c"] * math.pi * d**3 / 6
    exp = np.zeros((n[0] * n[1] * n[2], 3))
    for c, vals in case["expect"]["src_over_m"].items():
        exp[int(c)] = np.asarray(vals) * mass
    return exp


@pytest.mark.parametrize("case", HAND, ids=[c["name"][:40] for c in HAND])
def test_hand_cases(case):
    st, src = run_hand(case)
    ex = case["expect"]
    assert [st[1][0], st[2][0], st[3][0]] == ex["x"]
    assert [st[4][0], st[5][0], st[6][0]] == ex["v"]
    np.testing.assert_allclose(src, expected_hand(case), rtol=1e-15, atol=0)


@pytest.mark.parametrize("case", HAND, ids=[c["name"][:40] for c in HAND])
def test_hand_cases_fp32_form(case):
    """The FP32 form follows the same rules (all numbers are exact in float too)."""
    st, src = run_hand(case, st_oracle.FP32)
    ex = case["expect"]
    assert [float(st[1][0]), float(st[2][0]), float(st[3][0])] == ex["x"]
    assert [float(st[4][0]), float(st[5][0]), float(st[6][0])] == ex["v"]
    np.testing.assert_allclose(src, expected_hand(case), rtol=1e-6, atol=1e-12)


# ---------------------------------------------------------------- invariants

def mass_by_rhoc(case, d):
    return case.rho_d_by_rho_c * (4.0 / 3.0) * math.pi * (d / 2) ** 3


def test_momentum_balance_periodic():
    """No walls: sum over cells of UTrans = sum over parcels of (m/rho_c)(v_start - v_end)."""
    case = inputs.taylor_green_box(np_parcels=20000, ncell=16, boundary=1)
    st0 = case.state()
    src, _, st = st_oracle.evolve_case(case)
    m = mass_by_rhoc(case, st0[0])
    dv = np.stack([st0[4] - st[4], st0[5] - st[5], st0[6] - st[6]], 1)
    expect = (m[:, None] * dv).sum(0)
    scale = np.abs(m[:, None] * dv).sum(0)
    assert np.all(np.abs(src.sum(0) - expect) <= 1e-12 * scale)


def test_linearity_over_parcels():
    """The source of N parcels is the sum of the N single-parcel sources, and every
    parcel's trajectory is independent of the others."""
    case = inputs.taylor_green_box(np_parcels=64, ncell=8, boundary=0)
    src_all, _, st_all = st_oracle.evolve_case(case)
    acc = np.zeros_like(src_all)
    for i in range(case.np):
        st = [a[i:i + 1].copy() for a in case.state()]
        src_i, _ = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid,
                                    nsub=case.nsub, boundary=case.boundary, mu_c=case.mu_c,
                                    rho_d=case.rho_d, rho_d_by_rho_c=case.rho_d_by_rho_c)
        acc += src_i
        for a, b in zip(st, st_all):
            assert a[0] == b[i]
    np.testing.assert_allclose(acc, src_all, rtol=1e-12, atol=1e-30)


@pytest.mark.parametrize("boundary", [0, 1])
def test_pass_splitting_telescopes(boundary):
    """One launch of 10 substeps equals 10 launches of one substep: identical parcel
    state, and the same source up to rounding (sum of per-substep deposits into the
    starting cell = the residence deposits)."""
    case = inputs.taylor_green_box(np_parcels=5000, ncell=16, boundary=boundary)
    src_f, _, st_f = st_oracle.evolve_case(case)
    src_s, _, st_s = st_oracle.evolve_case(case, substeps_per_pass=1)
    for a, b in zip(st_f, st_s):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(src_s, src_f, rtol=0, atol=1e-12 * np.abs(src_f).max())


def test_openmp_equals_serial():
    case = inputs.taylor_green_box(np_parcels=20000, ncell=16, boundary=0)
    src1, bb1, st1 = st_oracle.evolve_case(case, nthreads=1)
    src4, bb4, st4 = st_oracle.evolve_case(case, nthreads=4)
    for a, b in zip(st1, st4):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(bb1, bb4)
    np.testing.assert_allclose(src4, src1, rtol=0, atol=1e-13 * np.abs(src1).max())


def test_fp32_form_close_to_fp64():
    case = inputs.taylor_green_box(np_parcels=20000, ncell=16, boundary=1)
    src64, _, st64 = st_oracle.evolve_case(case)
    c32 = inputs.taylor_green_box(np_parcels=20000, ncell=16, boundary=1, dtype=np.float32)
    src32, _, st32 = st_oracle.evolve_case(c32, precision=st_oracle.FP32)
    np.testing.assert_allclose(st32[1], st64[1], atol=1e-6)
    rel = np.linalg.norm(src32 - src64) / np.linalg.norm(src64)
    assert rel < 1e-2


def _as_f64_case(c32):
    """The FP32 case's (FP32-representable) inputs as a float64 case."""
    st = [a.astype(np.float64) for a in c32.state()]
    return inputs.Case(c32.n, c32.origin, c32.ending, c32.U.astype(np.float64), *st, mu_c=c32.mu_c,
                       rho_d=c32.rho_d, rho_c=c32.rho_c, dt_fluid=c32.dt_fluid, nsub=c32.nsub,
                       boundary=c32.boundary)


@pytest.mark.parametrize("boundary", [0, 1])
def test_fp32_form_arithmetic_error_under_frozen_decisions(boundary):
    """Accuracy of the FP32 form's arithmetic, separated from its face decisions.

    With a cell-constant carrier velocity the per-cell source is a discontinuous
    function of the parcel positions: a parcel whose FP32 position lands on the other
    side of a face than the FP64 one deposits into another cell, which alone moves the
    per-cell relative L2 error by ~1/(n sqrt(Ncells)) at n parcels per cell.  Here the
    FP32 run records its decisions (cell, bounce, wrap per substep) and the FP64 run
    replays them from the same FP32-representable inputs, over one call of 50 substeps
    of 1e-4 s (a call starts by locating each parcel from its position, so one call
    keeps every decision frozen): the remaining difference is the FP32 arithmetic
    (state rounding, FP32 drag and advance) and must stay below 1e-5 relative per cell.
    The free-running pair differs by far more: that error is the decisions'."""
    c32 = inputs.taylor_green_box(np_parcels=100000, ncell=24, boundary=boundary, dtype=np.float32,
                                  dt_fluid=5e-3, nsub=50)
    c64 = _as_f64_case(c32)
    ev = np.zeros((c32.np, c32.nsub), dtype=np.int64)
    s32, _, _ = st_oracle.evolve_case(c32, precision=st_oracle.FP32, events=ev, events_mode=1)
    s64, _, _ = st_oracle.evolve_case(c64, events=ev, events_mode=2)
    sfree, _, _ = st_oracle.evolve_case(c64)
    frozen = np.linalg.norm(s32 - s64) / np.linalg.norm(s64)
    free = np.linalg.norm(s32 - sfree) / np.linalg.norm(sfree)
    assert frozen < 1e-5, frozen
    assert np.abs(s32 - s64).max() < 1e-5 * np.abs(s64).max()
    assert free > 10 * frozen


def test_extrapolator_closed_form():
    """estimate_source! (asyncParallelTracking.jl:738-747): est_n = 2 S_n - est_{n-1}."""
    rng = np.random.default_rng(0)
    prev = np.zeros(12)
    est_prev = np.zeros(12)
    for _ in range(5):
        s = rng.normal(size=12)
        cur = s.copy()
        st_oracle.extrapolate(cur, prev)
        np.testing.assert_allclose(cur, 2 * s - est_prev, rtol=1e-15)
        np.testing.assert_array_equal(prev, cur)
        est_prev = cur.copy()
