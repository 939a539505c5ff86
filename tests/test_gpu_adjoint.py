"""GPU adjoint (st_adjoint_*) against the oracle's forward-mode dual numbers and
against central finite differences (eps = 1e-6, FP64) with the event sequence frozen.

Bars: adjoint vs dual numbers 1e-9 relative (same derivative, different summation
order); adjoint vs frozen-event finite differences 1e-4 relative.
"""
import numpy as np
import pytest

import scaletrack as stk
from scaletrack.adjoint import Adjoint, source_misfit_rrule
from oracle import st_oracle
from scaletrack import _lib, inputs

pytestmark = pytest.mark.gpu


def make_case(np_parcels=3000, ncell=12, boundary=1, nsub=40, seed=7):
    n = (ncell, ncell, ncell)
    o, e = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    U = inputs.fourier_field(n, o, e, modes=8, seed=seed)
    x, y, z = inputs.positions(seed, np_parcels, o, e)
    d = inputs.lognormal_diameters(seed, np_parcels, median=60e-6)
    rho = 1000.0 * (1.0 + 0.5 * inputs.uniform(seed, 50, np_parcels))
    zero = np.zeros(np_parcels)
    return inputs.Case(n, o, e, U, d, x, y, z, zero, zero.copy(), zero.copy(), rho=rho, mu_c=1e-3,
                       rho_d=1000.0, rho_c=1000.0, dt_fluid=nsub * 1e-4, nsub=nsub, boundary=boundary)


def target(case):
    return 1e-11 * inputs.normal(11, 60, case.ncell * 3).reshape(case.ncell, 3)


def gpu_objects(ctx, case, cd=1.0):
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    field = stk.TwoWayEulerian(ctx, mesh, stk.FP64)
    field.upload_U(case.U)
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, stk.FP64, rho_p=case.rho,
                      rho_c=case.rho_c)
    chunk.set_state(*case.state())
    chunk.upload()
    return field, chunk


def oracle_dual(case, direction, cd=1.0, xw=None, vw=None, T=None):
    st = case.state()
    tsrc = np.zeros((case.ncell, 3))
    ts = np.zeros((6, case.np))
    src, _ = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                              boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                              rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c,
                              cd=cd, rho=case.rho, tan_d=direction[0], tan_rho=direction[1],
                              tan_cd=direction[2], tan_source=tsrc, tan_state=ts)
    val = 0.0
    if T is not None:
        val += np.sum(2.0 * (src - T) * tsrc)
    if xw is not None:
        val += np.sum(xw * ts[0:3])
    if vw is not None:
        val += np.sum(vw * ts[3:6])
    return val


@pytest.mark.parametrize("method", [stk.PLAIN, stk.SORTED])
@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("K", [10, 7, 40])
def test_adjoint_vs_dual_source_misfit(ctx, boundary, method, K):
    case = make_case(boundary=boundary)
    T = target(case)
    field, chunk = gpu_objects(ctx, case)
    p = stk.params(boundary=boundary, method=method, brick=(4, 4, 4))
    adj = Adjoint(ctx, chunk, field, case.nsub, K, p)
    adj.forward(case.dt_fluid)
    adj.source_misfit(T)
    adj.backward()
    g = adj.gradients()
    rng = np.random.default_rng(3)
    dd = rng.normal(size=case.np) * case.d
    dr = rng.normal(size=case.np) * case.rho
    dc = 0.7
    ref = oracle_dual(case, (dd, dr, dc), T=T)
    got = np.dot(g["d"], dd) + np.dot(g["rho"], dr) + g["cd"] * dc
    assert abs(got - ref) <= 1e-9 * abs(ref), (got, ref)
    # per-component: cd alone and one parcel's d alone
    ref_cd = oracle_dual(case, (None, None, 1.0), T=T)
    assert abs(g["cd"] - ref_cd) <= 1e-9 * abs(ref_cd)
    e = np.zeros(case.np)
    e[17] = 1.0
    ref_d17 = oracle_dual(case, (e, None, 0.0), T=T)
    assert abs(g["d"][17] - ref_d17) <= 1e-8 * abs(ref_d17) + 1e-30


@pytest.mark.parametrize("boundary", [0, 1])
def test_adjoint_endpoint_loss(ctx, boundary):
    """L = sum(wx * x_T) + sum(wv * v_T): the state cotangents propagate through."""
    case = make_case(boundary=boundary, np_parcels=1500)
    field, chunk = gpu_objects(ctx, case)
    rng = np.random.default_rng(9)
    wx = rng.normal(size=(3, case.np))
    wv = rng.normal(size=(3, case.np))
    adj = Adjoint(ctx, chunk, field, case.nsub, 10, stk.params(boundary=boundary))
    adj.forward(case.dt_fluid)
    adj.set_state_cotangent(wx, wv)
    adj.backward()
    g = adj.gradients(initial_state=True)
    dd = rng.normal(size=case.np) * case.d
    ref = oracle_dual(case, (dd, None, 1.0), xw=wx, vw=wv)
    got = np.dot(g["d"], dd) + g["cd"]
    assert abs(got - ref) <= 1e-9 * abs(ref), (got, ref)
    # x passes through unchanged; backward is repeatable
    np.testing.assert_allclose(g["x0"], wx, rtol=0, atol=0)
    adj.backward()
    g2 = adj.gradients()
    np.testing.assert_array_equal(g2["d"], g["d"])


@pytest.mark.parametrize("boundary", [0, 1])
def test_adjoint_vs_frozen_event_fd(ctx, boundary):
    """The accuracy bar of the adjoint: central differences (eps = 1e-6) agree within 1e-4."""
    case = make_case(boundary=boundary, np_parcels=2000)
    T = target(case)
    field, chunk = gpu_objects(ctx, case)
    loss, pullback = source_misfit_rrule(ctx, chunk, field, T, case.dt_fluid, case.nsub, 20,
                                         stk.params(boundary=boundary))
    _, gd, grho, gcd = pullback(1.0)
    ev = np.zeros((case.np, case.nsub), dtype=np.int64)
    st = case.state()
    src0, _ = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                               boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                               rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c,
                               rho=case.rho, events=ev, events_mode=1)
    assert abs(loss - np.sum((src0 - T) ** 2)) <= 1e-12 * loss
    h = 1e-6
    rng = np.random.default_rng(4)
    dd = rng.normal(size=case.np) * case.d
    dr = rng.normal(size=case.np) * case.rho

    def L(sign, which):
        st = case.state()
        d = case.d + sign * h * dd if which == "d" else case.d
        st[0] = d.copy()
        rho = case.rho + sign * h * dr if which == "rho" else case.rho
        cd = 1.0 + sign * h if which == "cd" else 1.0
        s, _ = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                                boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                                rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c,
                                cd=cd, rho=rho, events=ev, events_mode=2)
        return np.sum((s - T) ** 2)

    for which, adj_val in (("d", np.dot(gd, dd)), ("rho", np.dot(grho, dr)), ("cd", gcd)):
        fd = (L(+1, which) - L(-1, which)) / (2 * h)
        assert abs(fd - adj_val) <= 1e-4 * abs(fd), (which, fd, adj_val)


def test_adjoint_forward_resorts_a_cloud_only_it_holds(ctx):
    """A cloud stepped (and so sorted) for the wall form and another brick, then given to
    an adjoint of the periodic form: the adjoint's forward re-sorts it for its own
    parameters (its own hold on the cloud does not refuse that), and the loss and
    gradients equal those of the same adjoint on a fresh upload of the same state.  A
    graph holding the cloud still refuses the re-sort."""
    case = make_case(np_parcels=2000, boundary=1, nsub=20)
    T = target(case)
    p = stk.params(boundary=1, brick=(4, 4, 4))
    runs = []
    for presorted in (False, True):
        field, chunk = gpu_objects(ctx, case)
        if presorted:
            chunk.sort(ctx, field, stk.params(boundary=0, brick=(2, 3, 4)))
        adj = Adjoint(ctx, chunk, field, case.nsub, 5, p)
        adj.forward(case.dt_fluid)
        loss = adj.source_misfit(T)
        adj.backward()
        runs.append((loss, adj.gradients()))
        adj.close()
        if presorted:
            chunk.sort(ctx, field, stk.params(boundary=0, brick=(2, 3, 4)))
            graph = stk.Graph(ctx, chunk, field, case.dt_fluid, case.nsub, stk.params(boundary=0, brick=(2, 3, 4)))
            held = Adjoint(ctx, chunk, field, case.nsub, 5, p)
            with pytest.raises(stk.ScaleTrackError) as ei:
                held.forward(case.dt_fluid)
            assert ei.value.status == _lib.ST_E_STATE and "boundary form" in str(ei.value)
            held.close()
            graph.close()
        chunk.close()
        field.close()
    (l0, g0), (l1, g1) = runs
    assert abs(l1 - l0) <= 1e-12 * abs(l0)
    for k in ("d", "rho", "cd"):
        a, b = np.asarray(g0[k], float), np.asarray(g1[k], float)
        assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max(), k


def test_gradients_into_page_locked_arrays(ctx):
    """The target and the gradients on page-locked host arrays (st_host_alloc through
    pinned_empty / pinned_copy), reused over two runs, the first with two losses (two
    backward sweeps) after one forward.  After the same forward every per-parcel gradient
    equals, bit for bit, that of pageable numpy arrays after a single sweep (the loss and
    dL/dcd, sums taken with atomics, to 1e-12); after a second forward, whose source sums
    add in another order, to 1e-9.  A mis-shaped output is refused before anything is
    copied."""
    case = make_case(np_parcels=3000, boundary=1, nsub=20)
    T = target(case)
    Tp = stk.pinned_copy(T)
    assert np.array_equal(Tp, T) and Tp.flags.c_contiguous
    p = stk.params(boundary=1, brick=(4, 4, 4))
    field, chunk = gpu_objects(ctx, case)
    adj = Adjoint(ctx, chunk, field, case.nsub, 5, p)
    n = case.np
    out = {"d": stk.pinned_empty(n), "rho": stk.pinned_empty(n), "x0": stk.pinned_empty((3, n)),
           "v0": stk.pinned_empty((3, n))}
    ref = None
    for run in range(2):
        chunk.set_state(*case.state())
        chunk.upload()
        adj.forward(case.dt_fluid)
        if ref is None:
            loss = adj.source_misfit(T)
            adj.backward()
            ref = (loss, adj.gradients(initial_state=True))
        loss_p = adj.source_misfit(Tp)
        adj.backward()
        g = adj.gradients(initial_state=True, out=out)
        tol = 1e-12 if run == 0 else 1e-9
        assert abs(loss_p - ref[0]) <= tol * abs(ref[0])
        for k in ("d", "rho", "x0", "v0"):
            assert g[k] is out[k]
            if run == 0:
                assert np.array_equal(g[k], ref[1][k]), k
            else:
                assert np.abs(g[k] - ref[1][k]).max() <= tol * np.abs(ref[1][k]).max(), k
        assert abs(g["cd"] - ref[1]["cd"]) <= tol * abs(ref[1]["cd"])
    with pytest.raises(ValueError):
        adj.gradients(out={"d": stk.pinned_empty(n - 1)})
    adj.close()
    chunk.close()
    field.close()
