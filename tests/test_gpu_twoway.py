"""Config 1 run two-way: the icoFoam-style fluid step (scaletrack.piso) drives the
async tracker through the reference's loop (icoJuliaParcelFoam.C:118-186) on the 2D
lid-driven cavity (0.1 m, 20 x 20 x 1 cells, lid 1 m/s, nu = 0.01 m^2/s; 10,000
parcels d = 50 um, rho = 2500 kg/m^3; dt = 0.005 s; 100 fluid steps), the solver
reading UTrans from and writing U into the tracker's own host fields.

Every parcel step is checked against the oracle run with the velocity that step was
given (recorded when the solver hands it over): after call k + 1 the solver sees the
extrapolated source of parcel step k to 1e-12, and the final cloud equals the oracle's
bit for bit.  The parcels' source feeds back into the flow: the run's velocity differs
from a one-way run's.
"""
import numpy as np
import pytest

import scaletrack as stk
from oracle import st_oracle
from scaletrack import coupling, inputs
from scaletrack.piso import IcoFoam, Patch

pytestmark = pytest.mark.gpu

PATCHES = {(0, 0): Patch("wall"), (0, 1): Patch("wall"), (1, 0): Patch("wall"),
           (1, 1): Patch("wall", (1.0, 0.0, 0.0)), (2, 0): Patch("empty"), (2, 1): Patch("empty")}


@pytest.mark.parametrize("method", [stk.SORTED, stk.PLAIN])
@pytest.mark.parametrize("per_volume", [False, True])
def test_cavity_two_way_100_steps(ctx, method, per_volume):
    case = inputs.cavity_case(np_parcels=10000)
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    host = {"N": case.np, "mu_c": case.mu_c, "rho": case.rho_d, "rho_d_by_rho_c": case.rho_d_by_rho_c}
    for k, a in zip(("d", "X", "Y", "Z", "U", "V", "W"), case.state()):
        host[k] = a.copy()
    tr = coupling.AsyncTracker(mesh, [host], precision=stk.FP64, nsub=case.nsub,
                               params=stk.params(boundary=stk.WALL, method=method, substeps_per_pass=1))
    U, _ = coupling.borrow(tr, "U")
    fluid = IcoFoam(case.n, case.origin, case.ending, 0.01, PATCHES, U=U)
    seen = []
    uts = []

    def record(k, Uk):
        seen.append(Uk.copy())
        if k >= 1:
            uts.append(tr.UTrans.copy())  # what fluid step k - 1 was given

    steps = 100
    try:
        coupling.run_coupled(fluid, tr, case.dt_fluid, steps, per_volume=per_volume, on_step=record)
        uts.append(tr.UTrans.copy())
    finally:
        tr.finalize_evolve()
    st = case.state()
    est_prev = np.zeros((case.ncell, 3))
    for k in range(steps):
        c = inputs.Case(case.n, case.origin, case.ending, seen[k], *st, mu_c=case.mu_c, rho_d=case.rho_d,
                        rho_c=case.rho_c, dt_fluid=case.dt_fluid, nsub=case.nsub, boundary=0)
        s, _, st = st_oracle.evolve_case(c, state=st, substeps_per_pass=1)
        if k + 1 < steps:
            est = 2.0 * s - est_prev
            got = uts[k + 1]
            assert np.abs(got - est).max() <= 1e-12 * max(np.abs(est).max(), 1e-300), k
            est_prev = est
    for name, a, b in zip("dxyzuvw", tr.executor.final[0], st):
        np.testing.assert_array_equal(a, b, err_msg=name)
    # the feedback: the same flow without the source differs
    one_way = IcoFoam(case.n, case.origin, case.ending, 0.01, PATCHES)
    for _ in range(steps):
        one_way.step(case.dt_fluid)
    diff = np.abs(one_way.U - U).max()
    assert diff > 0.0
    if per_volume:
        assert diff > 1e-4, diff
    assert np.abs(fluid.divergence()).max() < 1e-12


def test_channel_two_way_100_steps(ctx):
    """The channel workload's coupled loop at a reduced mesh (32 x 16 x 16 cells on the
    2 x 1 x 1 m duct): laminar inflow through the inlet patch, zero-gradient outlet,
    walls on y and z, FP32 tracking with the channel boundary form, 500 parcels
    injected per fluid step and removed ones compacted every 10 steps; 100 fluid steps
    of 0.01 s.  Each parcel step is checked against the oracle's FP32 form with the
    velocity it was given."""
    n, L = (32, 16, 16), (2.0, 1.0, 1.0)
    case = inputs.channel_case(np_parcels=4000, n=n, length=L, dtype=np.float32, dt_fluid=1e-2)
    mesh = stk.Mesh(case.n, case.origin, case.ending)
    host = {"N": case.np, "mu_c": case.mu_c, "rho": case.rho_d, "rho_d_by_rho_c": case.rho_d_by_rho_c}
    for k, a in zip(("d", "X", "Y", "Z", "U", "V", "W"), case.state()):
        host[k] = a.copy()
    inject = 500
    tr = coupling.AsyncTracker(mesh, [host], precision=stk.FP32, nsub=case.nsub,
                               params=stk.params(boundary=stk.CHANNEL, substeps_per_pass=1, brick=(8, 8, 4)),
                               injector=lambda step: [inputs.channel_injection(case, step, inject)],
                               compact_every=10)
    U, _ = coupling.borrow(tr, "U")
    U[...] = case.U
    # inlet profile at the inlet face centres, (nz, ny, 1, 3)
    prof = inputs.channel_profile((1,) + n[1:], (0.0,) + case.origin[1:], (L[0] / n[0],) + case.ending[1:])
    patches = {(0, 0): Patch("inlet", prof.reshape(n[2], n[1], 1, 3)), (0, 1): Patch("outlet"),
               (1, 0): Patch("wall"), (1, 1): Patch("wall"), (2, 0): Patch("wall"), (2, 1): Patch("wall")}
    fluid = IcoFoam(n, case.origin, case.ending, 0.01, patches, U=U)
    seen = []
    steps = 100
    try:
        coupling.run_coupled(fluid, tr, case.dt_fluid, steps, per_volume=True,
                             on_step=lambda k, Uk: seen.append(Uk.copy()))
    finally:
        tr.finalize_evolve()
    st = case.state()
    for k in range(steps):
        new = inputs.channel_injection(case, k + 1, inject)
        st = [np.concatenate([a, b]) for a, b in zip(st, new)]
        c = inputs.Case(case.n, case.origin, case.ending, seen[k], *st, mu_c=case.mu_c, rho_d=case.rho_d,
                        rho_c=case.rho_c, dt_fluid=case.dt_fluid, nsub=case.nsub, boundary=2)
        _, _, st = st_oracle.evolve_case(c, state=st, precision=st_oracle.FP32, substeps_per_pass=1)
        if (k + 1) % 10 == 0:
            live = st[0] > 0
            st = [a[live] for a in st]
    for name, a, b in zip("dxyzuvw", tr.executor.final[0], st):
        np.testing.assert_array_equal(a, b, err_msg=name)
    assert st[0].size > 4000 + 100 * inject * 0.5
    inflow, outflow = fluid.phi[0][1].sum(), fluid.phi[0][2].sum()
    assert abs(outflow - inflow) < 1e-8 * abs(inflow)
