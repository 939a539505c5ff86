"""Randomised configurations of the GPU adjoint against the oracle's forward-mode dual
numbers.

Each seed draws a mesh (2..20 cells per axis, a non-cubic box at an arbitrary origin), a
divergence-free carrier of random amplitude, a cloud of 1..4000 parcels with random
diameters, velocities and (for some seeds) per-parcel densities, a drag multiplier,
1..60 substeps, a checkpoint interval K in [1, nsub], the boundary and the method.  The
loss is the source misfit against a random target, the endpoint loss of random state
cotangents, or both.

Checked per seed, as directional derivatives along random directions of d, of rho and
of c_D alone: the adjoint's gradient dotted with the direction equals the oracle's
dual-number tangent of the loss.  The bar is 1e-8 of the sum of the magnitudes of the
dot product's terms (the scale at which summation order shows).  For the scalar c_D it
is 1e-8 of |dL/dln c_D| plus the loss's sensitivity to relative changes of d.
"""
import numpy as np
import pytest

import scaletrack as stk
from scaletrack import inputs
from scaletrack.adjoint import Adjoint
from oracle import st_oracle

pytestmark = pytest.mark.gpu


def draw(seed):
    rng = np.random.default_rng(7000 + seed)
    n = tuple(int(v) for v in rng.integers(2, 21, size=3))
    origin = tuple(float(v) for v in rng.uniform(-1.0, 1.0, size=3))
    ending = tuple(o + float(L) for o, L in zip(origin, rng.uniform(0.2, 2.0, size=3)))
    boundary = stk.PERIODIC if rng.random() < 0.5 else stk.WALL
    U = inputs.fourier_field(n, origin, ending, modes=6, seed=300 + seed,
                             amplitude=float(10 ** rng.uniform(-1, 0.7)))
    npc = int(rng.choice([1, 2, 65, int(rng.integers(100, 4001))]))
    x, y, z = inputs.positions(300 + seed, npc, origin, ending)
    d = inputs.lognormal_diameters(300 + seed, npc, median=float(10 ** rng.uniform(-4.7, -3.8)))
    u, v, w = [rng.standard_normal(npc) * float(10 ** rng.uniform(-2, 0)) for _ in range(3)]
    rho_d = float(rng.uniform(500.0, 3000.0))
    rho = rho_d * rng.uniform(0.5, 2.0, size=npc) if rng.random() < 0.5 else None
    nsub = int(rng.integers(1, 61))
    case = inputs.Case(n, origin, ending, U, d, x, y, z, u, v, w, rho=rho, mu_c=float(10 ** rng.uniform(-4, -2.5)),
                       rho_d=rho_d, rho_c=float(rng.uniform(1.0, 1000.0)),
                       dt_fluid=float(nsub * 10 ** rng.uniform(-5, -3)), nsub=nsub, boundary=boundary)
    K = int(rng.integers(1, nsub + 1))
    method = stk.SORTED if rng.random() < 0.5 else stk.PLAIN
    cd = float(rng.uniform(0.5, 2.0))
    loss = ["source", "endpoint", "both"][int(rng.integers(0, 3))]
    return case, K, method, cd, loss, rng


def dual(case, cd, direction, T=None, xw=None, vw=None):
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
        val += np.sum(xw * ts[0:3]) + np.sum(vw * ts[3:6])
    return val, src


@pytest.mark.parametrize("seed", range(40))
def test_random_adjoint(ctx, seed):
    case, K, method, cd, loss, rng = draw(seed)
    ident = (f"adjoint seed {seed}: n={case.n} np={case.np} nsub={case.nsub} K={K} "
             f"{'periodic' if case.boundary else 'wall'} method={method} loss={loss} "
             f"rho={'parcel' if case.rho is not None else 'chunk'}")
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), stk.FP64)
    field.upload_U(case.U)
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, stk.FP64, rho_p=case.rho,
                      rho_c=case.rho_c)
    chunk.set_state(*case.state())
    chunk.upload()
    p = stk.params(boundary=case.boundary, method=method, brick=(4, 4, 2), cd=cd)
    adj = Adjoint(ctx, chunk, field, case.nsub, K, p)
    adj.forward(case.dt_fluid)

    _, src = dual(case, cd, (None, None, 0.0))
    T = xw = vw = None
    if loss in ("source", "both"):
        # a target near the run's own source, so the misfit weights every cell
        T = (src * (1.0 + 0.3 * rng.standard_normal(src.shape))
             + 1e-3 * np.abs(src).max() * rng.standard_normal(src.shape))
        adj.source_misfit(T)
    if loss in ("endpoint", "both"):
        xw = rng.standard_normal((3, case.np))
        vw = rng.standard_normal((3, case.np)) * 1e-2
        adj.set_state_cotangent(xw, vw)
    adj.backward()
    g = adj.gradients()

    dd = rng.standard_normal(case.np) * case.d
    ref, _ = dual(case, cd, (dd, None, 0.0), T, xw, vw)
    got = np.dot(g["d"], dd)
    scale = np.sum(np.abs(g["d"] * dd))
    assert abs(got - ref) <= 1e-8 * scale + 1e-300, f"{ident}: d ({got} vs {ref}, scale {scale})"

    if case.rho is not None:
        dr = rng.standard_normal(case.np) * case.rho
        ref, _ = dual(case, cd, (None, dr, 0.0), T, xw, vw)
        got = np.dot(g["rho"], dr)
        scale = np.sum(np.abs(g["rho"] * dr))
        assert abs(got - ref) <= 1e-8 * scale + 1e-300, f"{ident}: rho ({got} vs {ref}, scale {scale})"

    # c_D enters only through f = c_D * (18 mu dt) / (rho d^2), so dL/dln(c_D) is of the
    # order of the loss's sensitivity to relative changes of d, sum |dL/dd * d|; that is
    # the scale of its rounding too (a parcel relaxed onto the carrier leaves dL/dc_D at
    # the rounding level of L, where oracle and adjoint agree only to that scale)
    ref, _ = dual(case, cd, (None, None, 1.0), T, xw, vw)
    scale = abs(ref) * cd + np.sum(np.abs(g["d"] * case.d))
    assert abs(g["cd"] - ref) * cd <= 1e-8 * scale + 1e-300, f"{ident}: cd ({g['cd']} vs {ref}, scale {scale})"
    adj.close()
    chunk.close()
    field.close()
