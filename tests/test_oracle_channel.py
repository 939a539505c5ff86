"""The oracle's channel form (walls on y and z, open x planes; an extension -- the
reference keeps every parcel of a chunk, asyncParallelTracking.jl:329-364) pinned by
hand-derived exits, by equality with the wall form wherever the open planes are never
reached, and by the momentum balance with exits.

Hand cases use a drag-free carrier (mu = 0: f = 0 and v' = v exactly) on a mesh of
1/8-m cells with substeps of 1/1024 s, so every position is an exact binary number.
"""
import numpy as np
import pytest

from oracle import st_oracle
from scaletrack import inputs

N = (4, 2, 2)
O = (0.0, 0.0, 0.0)
E = (0.5, 0.25, 0.25)
DT_FLUID = 10.0 / 1024.0  # 10 substeps of 1/1024 s


def one(x, v, d=1e-4, nsub=10, U=None, mu=0.0, boundary=st_oracle.CHANNEL, precision=st_oracle.FP64):
    dt = np.float64 if precision == st_oracle.FP64 else np.float32
    st = [np.array([d], dt)] + [np.array([a], dt) for a in x] + [np.array([a], dt) for a in v]
    U = np.zeros((16, 3), dt) if U is None else U
    src, bb = st_oracle.evolve(st, U, N, O, E, DT_FLUID, nsub=nsub, precision=precision, boundary=boundary,
                               mu_c=mu, rho_d=1000.0, rho_d_by_rho_c=1.0)
    return st, src, bb


def cell(i, j, k):
    return k * N[0] * N[1] + j * N[0] + i


@pytest.mark.parametrize("precision", [st_oracle.FP64, st_oracle.FP32])
def test_exit_through_the_outlet(precision):
    # x0 = 0.5 - 1/32 at 8 m/s: 1/128 per substep, x = 0.5 (cell 4 = N) after substep 4
    st, src, bb = one((0.46875, 0.1, 0.1), (8.0, 0.0, 0.0), precision=precision)
    assert st[0][0] == -1e-4 if precision == st_oracle.FP64 else st[0][0] == -np.float32(1e-4)
    assert [st[1][0], st[2][0], st[3][0]] == [0.5, 0.1 if precision == st_oracle.FP64 else np.float32(0.1),
                                             0.1 if precision == st_oracle.FP64 else np.float32(0.1)]
    assert [st[4][0], st[5][0], st[6][0]] == [8.0, 0.0, 0.0]
    assert not src.any()  # no drag: nothing deposited
    assert np.all(np.isinf(bb))  # the removed parcel is not in the bounding box


def test_exit_through_the_inlet_and_deposit_cell():
    # x0 = 1/32 at -8 m/s: x = 0 after substep 4 (still cell 0), -1/128 after substep 5
    # (cell -1): it leaves at substep 5.  A carrier of -8 m/s in its cell makes f > 0
    # irrelevant to the path (v' = v) and the deposit m (s0 - v') zero in x; a carrier
    # in y gives a deposit into cell (0, 0, 0) alone.
    U = np.zeros((16, 3))
    U[cell(0, 0, 0)] = (-8.0, 1.0, 0.0)
    st, src, _ = one((0.03125, 0.1, 0.1), (-8.0, 0.0, 0.0), U=U, mu=1e-3, nsub=10)
    assert st[0][0] < 0
    assert st[1][0] == -0.0078125
    nz = np.flatnonzero(np.abs(src).sum(1))
    assert list(nz) == [cell(0, 0, 0)]
    # the y velocity taken from the carrier over 5 substeps is deposited with the
    # opposite sign: m (0 - v_y) with m = (pi d^3 / 6) (rho_d / rho_c)
    m = (1.0 * np.pi) * ((1e-4 * 1e-4) * 1e-4) / 6.0
    assert src[cell(0, 0, 0), 1] == m * (0.0 - st[5][0])
    assert st[5][0] > 0


def test_exit_takes_precedence_over_a_wall():
    # leaves through the outlet on the substep it also crosses y = 0.25: no bounce
    st, src, _ = one((0.49609375, 0.248046875, 0.1), (1.0, 2.0, 0.0), nsub=10)
    # substep 1: x = 0.49609375 + 1/1024 (inside), y = 0.248046875 + 2/1024 = 0.25 -> wall
    # bounce (the parcel stays inside); it exits later through x = 0.5
    assert st[0][0] < 0
    assert st[1][0] >= 0.5
    st2, _, _ = one((0.498046875, 0.248046875, 0.1), (2.0, 2.0, 0.0), nsub=10)
    # substep 1 ends at x = 0.5 and y = 0.25: out in x -> removed without a bounce in y
    assert st2[0][0] < 0
    assert [st2[1][0], st2[2][0]] == [0.5, 0.25]
    assert [st2[4][0], st2[5][0]] == [2.0, 2.0]


def test_exit_on_the_last_substep():
    # x reaches 0.5 exactly on substep 10
    x0 = 0.5 - 10 / 128.0
    st, _, _ = one((x0, 0.1, 0.1), (8.0, 0.0, 0.0), nsub=10)
    assert st[0][0] < 0 and st[1][0] == 0.5
    st, _, _ = one((x0 - 1 / 1024.0, 0.1, 0.1), (8.0, 0.0, 0.0), nsub=10)
    assert st[0][0] > 0 and st[1][0] == 0.5 - 1 / 1024.0


def test_removed_parcels_are_left_alone():
    st, _, _ = one((0.46875, 0.1, 0.1), (8.0, 0.0, 0.0))
    before = [a.copy() for a in st]
    U = np.ones((16, 3))
    src = np.zeros((16, 3))
    st_oracle.evolve(st, U, N, O, E, DT_FLUID, nsub=10, boundary=st_oracle.CHANNEL, mu_c=1e-3, rho_d=1000.0,
                     rho_d_by_rho_c=1.0, source=src)
    for a, b in zip(st, before):
        assert a[0] == b[0]
    assert not src.any()


@pytest.mark.parametrize("precision", [st_oracle.FP64, st_oracle.FP32])
def test_channel_equals_walls_away_from_the_open_planes(precision):
    """Parcels that never reach x = 0 or x = L follow the wall form bit for bit."""
    dt = np.float64 if precision == st_oracle.FP64 else np.float32
    case = inputs.channel_case(np_parcels=20000, n=(32, 16, 16), dtype=dt)
    keep = (case.x > 0.2) & (case.x < 1.8)  # far more than one call's travel from the planes
    st0 = [a[keep].copy() for a in case.state()]
    res = {}
    for bnd in (st_oracle.WALL, st_oracle.CHANNEL):
        st = [a.copy() for a in st0]
        src, bb = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                                   precision=precision, boundary=bnd, mu_c=case.mu_c, rho_d=case.rho_d,
                                   rho_d_by_rho_c=case.rho_d_by_rho_c)
        res[bnd] = (st, src, bb)
    for a, b in zip(res[0][0], res[2][0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(res[0][1], res[2][1])
    np.testing.assert_array_equal(res[0][2], res[2][2])
    assert np.all(res[2][0][0] > 0)


def test_momentum_balance_with_exits():
    """Away from the walls (no bounces): the summed source equals the summed momentum
    change m (v_start - v_end) of every parcel, the exited ones with their exit velocity."""
    case = inputs.channel_case(np_parcels=20000, n=(16, 8, 8), dtype=np.float64, dt_fluid=0.05, nsub=10)
    inner = (case.y > 0.3) & (case.y < 0.7) & (case.z > 0.3) & (case.z < 0.7)
    st = [a[inner].copy() for a in case.state()]
    v0 = np.stack(st[4:7], 1).copy()
    d0 = st[0].copy()
    src, _ = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                              boundary=st_oracle.CHANNEL, mu_c=case.mu_c, rho_d=case.rho_d,
                              rho_d_by_rho_c=case.rho_d_by_rho_c)
    gone = st[0] < 0
    assert gone.sum() > 100  # 0.5 s at ~1 m/s: parcels near the outlet leave
    np.testing.assert_array_equal(np.abs(st[0]), d0)
    m = case.rho_d_by_rho_c * np.pi * d0 ** 3 / 6.0
    dv = v0 - np.stack(st[4:7], 1)
    expect = (m[:, None] * dv).sum(0)
    scale = np.abs(m[:, None] * dv).sum(0)
    assert np.all(np.abs(src.sum(0) - expect) <= 1e-12 * scale)
    assert np.all((st[1][gone] < 0) | (st[1][gone] >= case.ending[0]))
    assert np.all((st[1][~gone] >= 0) & (st[1][~gone] < case.ending[0]))
