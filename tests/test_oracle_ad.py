"""Pins the oracle's forward-mode dual numbers and its frozen-event replay, which the
GPU adjoint is checked against.

* Replaying recorded events without a perturbation reproduces the run bit for bit.
* The dual-number directional derivative of the source-misfit loss equals central
  finite differences (eps = 1e-6, FP64) taken with the event sequence frozen, for
  per-parcel diameter, per-parcel density and the global drag coefficient.
"""
import numpy as np
import pytest

from oracle import st_oracle
from scaletrack import inputs


def adjoint_case(np_parcels=400, ncell=8, boundary=1, nsub=40, seed=7):
    n = (ncell, ncell, ncell)
    o, e = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    U = inputs.fourier_field(n, o, e, modes=8, seed=seed)
    x, y, z = inputs.positions(seed, np_parcels, o, e)
    d = inputs.lognormal_diameters(seed, np_parcels, median=60e-6)
    rho = 1000.0 * (1.0 + 0.5 * inputs.uniform(seed, 50, np_parcels))
    zero = np.zeros(np_parcels)
    c = inputs.Case(n, o, e, U, d, x, y, z, zero, zero.copy(), zero.copy(), rho=rho, mu_c=1e-3,
                    rho_d=1000.0, rho_c=1000.0, dt_fluid=nsub * 1e-4, nsub=nsub, boundary=boundary)
    return c


def run(case, d=None, rho=None, cd=1.0, events=None, mode=0, tan=None):
    st = case.state()
    if d is not None:
        st[0] = d.copy()
    kw = {}
    if tan is not None:
        kw = dict(tan_d=tan[0], tan_rho=tan[1], tan_cd=tan[2],
                  tan_source=np.zeros((case.ncell, 3)))
    src, _ = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid,
                              nsub=case.nsub, boundary=case.boundary, mu_c=case.mu_c,
                              rho_d=case.rho_d, rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c, cd=cd,
                              rho=case.rho if rho is None else rho, events=events,
                              events_mode=mode, **kw)
    return src, st, kw.get("tan_source")


def target(case):
    return 1e-12 * inputs.normal(11, 60, case.ncell * 3).reshape(case.ncell, 3)


@pytest.mark.parametrize("boundary", [0, 1])
def test_replay_is_exact(boundary):
    case = adjoint_case(boundary=boundary)
    ev = np.zeros((case.np, case.nsub), dtype=np.int64)
    src1, st1, _ = run(case, events=ev, mode=1)
    src2, st2, _ = run(case, events=ev, mode=2)
    np.testing.assert_array_equal(src1, src2)
    for a, b in zip(st1, st2):
        np.testing.assert_array_equal(a, b)
    src0, st0, _ = run(case)
    np.testing.assert_array_equal(src0, src1)


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("which", ["d", "rho", "cd", "all"])
def test_dual_matches_frozen_event_fd(boundary, which):
    case = adjoint_case(boundary=boundary)
    T = target(case)
    rng = np.random.default_rng(5)
    dd = rng.normal(size=case.np) * case.d if which in ("d", "all") else np.zeros(case.np)
    dr = rng.normal(size=case.np) * case.rho if which in ("rho", "all") else np.zeros(case.np)
    dc = 1.0 if which in ("cd", "all") else 0.0
    ev = np.zeros((case.np, case.nsub), dtype=np.int64)
    src, _, tsrc = run(case, events=ev, mode=1, tan=(dd, dr, dc))
    dual = np.sum(2.0 * (src - T) * tsrc)
    h = 1e-6

    def loss(sign):
        s, _, _ = run(case, d=case.d + sign * h * dd, rho=case.rho + sign * h * dr, cd=1.0 + sign * h * dc,
                      events=ev, mode=2)
        return np.sum((s - T) ** 2)

    fd = (loss(+1) - loss(-1)) / (2 * h)
    assert abs(fd - dual) <= 1e-6 * abs(dual) + 1e-30, (fd, dual)


def test_dual_state_tangent_of_velocity():
    """The state tangent dv/dcd of one parcel in a constant field matches the closed
    form v_n = u + (v0 - u) r^n, r = 1/(1+f), differentiated in f."""
    n = (1, 1, 1)
    u = np.array([[0.5, -0.25, 0.125]])
    st = [np.array([40e-6]), np.array([0.5]), np.array([0.5]), np.array([0.5]), np.zeros(1), np.zeros(1),
          np.zeros(1)]
    ts = np.zeros((6, 1))
    tsrc = np.zeros((1, 3))
    nsub, dtf, mu, rho = 5, 5e-5, 1e-3, 1000.0
    st_oracle.evolve(st, u, n, (0, 0, 0), (1, 1, 1), dtf, nsub=nsub, mu_c=mu, rho_d=rho, cd=1.0,
                     tan_cd=1.0, tan_source=tsrc, tan_state=ts)
    dt = dtf / nsub
    f = 18 * mu * dt / (rho * 40e-6 ** 2)
    # v_n = u (1 - r^n), dv_n/dcd = -u n r^(n-1) dr/df * df/dcd, dr/df = -r^2, df/dcd = f
    r = 1 / (1 + f)
    dv = u[0] * nsub * r ** (nsub + 1) * f
    np.testing.assert_allclose(ts[3:6, 0], dv, rtol=1e-10)
