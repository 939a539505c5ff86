"""The asynchronous tracker (evolve_cloud, lag-one exchange, extrapolator) driving the
real device step: the host UTrans seen by the solver after each evolve_cloud equals the
oracle's source of the previous parcel step, extrapolated."""
import numpy as np
import pytest

import scaletrack as stk
from scaletrack import coupling, inputs
from oracle import st_oracle

pytestmark = pytest.mark.gpu


def test_async_tracker_against_oracle(ctx):
    case = inputs.taylor_green_box(np_parcels=5000, ncell=8, boundary=0)
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    st = case.state()
    host = {"N": case.np, "mu_c": case.mu_c, "rho": case.rho_d, "rho_d_by_rho_c": case.rho_d_by_rho_c}
    for k, a in zip(("d", "X", "Y", "Z", "U", "V", "W"), st):
        host[k] = a.copy()
    p = stk.params(boundary=stk.WALL)
    tr = coupling.AsyncTracker(mesh, [host], precision=stk.FP64, nsub=case.nsub, params=p)
    tr.U[...] = case.U
    est_prev = np.zeros((case.ncell, 3))
    ref_state = case.state()
    srcs = []
    for n in range(1, 5):
        tr.evolve_cloud(case.dt_fluid)
        if n >= 2:
            s_true = srcs[n - 2]
            est = 2.0 * s_true - est_prev
            scale = np.abs(est).max()
            assert np.abs(tr.UTrans - est).max() <= 1e-12 * scale, n
            est_prev = est
        else:
            assert not tr.UTrans.any()
        # the oracle's parcel step n, run with the field the tracker uploaded (frozen)
        s, _, ref_state = st_oracle.evolve_case(case, state=ref_state)
        srcs.append(s)
    tr.finalize_evolve()


def draw_coupled(seed):
    """A random coupled run: mesh, precision, boundary, method and substeps; 1..3
    chunks with their own density, viscosity and density ratio sharing one field; and
    3..6 fluid steps, each with its own carrier field and time step."""
    rng = np.random.default_rng(9000 + seed)
    precision = stk.FP32 if rng.random() < 0.5 else stk.FP64
    dt = np.float32 if precision == stk.FP32 else np.float64
    boundary = stk.PERIODIC if rng.random() < 0.5 else stk.WALL
    n = tuple(int(v) for v in rng.integers(2, 17, size=3))
    origin = tuple(float(v) for v in rng.uniform(-1.0, 1.0, size=3))
    ending = tuple(o + float(L) for o, L in zip(origin, rng.uniform(0.2, 2.0, size=3)))
    nsub = int(rng.integers(1, 13))
    steps = int(rng.integers(3, 7))
    Us = [inputs.fourier_field(n, origin, ending, modes=4, seed=500 + 10 * seed + k,
                               amplitude=float(10 ** rng.uniform(-1, 0.5)), dtype=dt) for k in range(steps)]
    dts = [float(10 ** rng.uniform(-4, -2.5)) for _ in range(steps)]
    chunks = []
    for c in range(int(rng.integers(1, 4))):
        npc = int(rng.choice([1, 64, int(rng.integers(100, 3001))]))
        x, y, z = inputs.positions(600 + 10 * seed + c, npc, origin, ending, dtype=dt)
        d = inputs.lognormal_diameters(600 + 10 * seed + c, npc, median=float(10 ** rng.uniform(-5, -4)), dtype=dt)
        u, v, w = [(rng.standard_normal(npc) * 0.1).astype(dt) for _ in range(3)]
        rho_d = float(rng.uniform(500.0, 3000.0))
        chunks.append(inputs.Case(n, origin, ending, Us[0], d, x, y, z, u, v, w,
                                  mu_c=float(10 ** rng.uniform(-4, -2.5)), rho_d=rho_d,
                                  rho_c=rho_d / float(rng.uniform(0.5, 4.0)), nsub=nsub, boundary=boundary))
    method = stk.SORTED if rng.random() < 0.6 else stk.PLAIN
    return precision, dt, boundary, n, origin, ending, nsub, Us, dts, chunks, method


@pytest.mark.parametrize("seed", range(16))
def test_async_tracker_random(ctx, seed):
    """Parcel step k runs with the carrier field and time step the solver left before
    its k-th evolve_cloud; after call k + 1 the solver sees the extrapolated source of
    step k, 2 S_k - (the previous estimate), in the field's precision.  Every chunk
    deposits into the one field.  The bar is 1e-12 of the largest estimate in FP64 and
    1e-6 (a few float ulps: the source is rounded to float on download) in FP32."""
    precision, dt, boundary, n, origin, ending, nsub, Us, dts, chunks, method = draw_coupled(seed)
    ident = f"coupled seed {seed}: n={n} chunks={[c.np for c in chunks]} steps={len(Us)} nsub={nsub} " \
            f"{'fp32' if precision == stk.FP32 else 'fp64'} {'periodic' if boundary else 'wall'} method={method}"
    mesh = stk.Mesh(n, origin, ending)
    hosts, states = [], []
    for c in chunks:
        st = c.state()
        h = {"N": c.np, "mu_c": c.mu_c, "rho": c.rho_d, "rho_d_by_rho_c": c.rho_d_by_rho_c}
        for k, a in zip(("d", "X", "Y", "Z", "U", "V", "W"), st):
            h[k] = a.copy()
        hosts.append(h)
        states.append(st)
    oprec = st_oracle.FP32 if precision == stk.FP32 else st_oracle.FP64
    tr = coupling.AsyncTracker(mesh, hosts, precision=precision, nsub=nsub,
                               params=stk.params(boundary=boundary, method=method))
    try:
        est_prev = np.zeros((mesh.ncell, 3), dtype=dt)
        sources = []
        for k in range(len(Us) + 1):
            if k < len(Us):
                tr.U[...] = Us[k]
                tr.evolve_cloud(dts[k])
            else:
                tr.evolve_cloud(dts[-1])  # the last call only collects step k - 1's source
            if k >= 1:
                est = 2.0 * sources[k - 1].astype(dt) - est_prev
                scale = max(float(np.abs(est).max()), 1e-300)
                tol = (1e-12 if precision == stk.FP64 else 1e-6) * scale
                assert np.abs(tr.UTrans - est).max() <= tol, f"{ident}: call {k + 1}"
                est_prev = est
            else:
                assert not tr.UTrans.any()
            if k < len(Us):
                total = np.zeros((mesh.ncell, 3))
                for i, c in enumerate(chunks):
                    c.U, c.dt_fluid = Us[k], dts[k]
                    s, _, states[i] = st_oracle.evolve_case(c, state=states[i], precision=oprec)
                    total += s
                sources.append(total)
    finally:
        tr.finalize_evolve()
