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
