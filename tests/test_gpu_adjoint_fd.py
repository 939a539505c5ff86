"""The adjoint against free-running central differences (eps = 1e-6, FP64, nothing
frozen) on the adjoint workload: 2^20 parcels on 64^3 cells, 200 substeps, a checkpoint
every 20, loss = |S - T|^2.

A perturbation of a global direction (every parcel's d, rho, or c_D) moves a few of
the 2^20 parcels across a face one substep earlier or later (2 to 5 parcels at
eps = 1e-6), and each such event flip changes the loss by a jump larger than the whole
smooth change; the global free-running quotient is then off by O(1), which the
frozen-decision check of test_gpu_configs.py leaves out by construction.  Perturbing
one parcel at a time avoids that: a parcel's own flip is a rare event, and the check is
exact free-running arithmetic -- only that parcel's trajectory changes, so
L(+-eps) = |S - S_i + S_i(+-eps) - T|^2 needs the full source once and one single-parcel
run per side.  Each sampled parcel's adjoint dL/dd_i and dL/drho_i must match its
quotient to 1e-4 relative, and the parcels whose events did flip are counted and left
out (none are expected among the sample).
"""
import os

import numpy as np
import pytest

import scaletrack as stk
from oracle import st_oracle
from scaletrack import inputs
from scaletrack.adjoint import source_misfit_rrule

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
EPS = 1e-6
SAMPLE = 160


def _one(case, i, d, rho):
    """Source and events of parcel i alone with diameter d and density rho."""
    st = [np.array([a[i]]) for a in case.state()]
    st[0][0] = d
    ev = np.zeros((1, case.nsub), dtype=np.int64)
    src, _ = st_oracle.evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                              boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                              rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c, rho=np.array([rho]),
                              events=ev, events_mode=1)
    return src, ev


def test_adjoint_per_parcel_free_running_fd(ctx):
    case = inputs.adjoint_case(np_parcels=1 << 20, ncell=64, nsub=200)
    T = inputs.target_source(case)
    field = stk.TwoWayEulerian(ctx, stk.Mesh(case.n, case.origin, case.ending), stk.FP64)
    field.upload_U(case.U)
    chunk = stk.Chunk(ctx, case.np, case.mu_c, case.rho_d, case.rho_d_by_rho_c, stk.FP64, rho_p=case.rho,
                      rho_c=case.rho_c)
    chunk.set_state(*case.state())
    chunk.upload()
    _, pullback = source_misfit_rrule(ctx, chunk, field, T, case.dt_fluid, case.nsub, 20,
                                      stk.params(boundary=stk.PERIODIC))
    _, gd, grho, _ = pullback(1.0)
    chunk.close()
    field.close()

    S, _ = st_oracle.evolve(case.state(), case.U, case.n, case.origin, case.ending, case.dt_fluid, nsub=case.nsub,
                            boundary=case.boundary, mu_c=case.mu_c, rho_d=case.rho_d,
                            rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c, rho=case.rho, nthreads=THREADS)
    rng = np.random.default_rng(23)
    flips, checked, worst = 0, 0, 0.0
    for i in rng.choice(case.np, SAMPLE, replace=False):
        d0, r0 = case.d[i], case.rho[i]
        s0, ev0 = _one(case, i, d0, r0)
        R = S - s0 - T  # everything but parcel i
        for which, grad, base in (("d", gd[i], d0), ("rho", grho[i], r0)):
            dl, evs = [], []
            for sign in (1.0, -1.0):
                val = base * (1.0 + sign * EPS)
                si, ev = _one(case, i, val if which == "d" else d0, val if which == "rho" else r0)
                touched = np.flatnonzero(np.abs(si).sum(1) + np.abs(s0).sum(1))
                dl.append(np.sum((R[touched] + si[touched]) ** 2) - np.sum((R[touched] + s0[touched]) ** 2))
                evs.append(ev)
            if not np.array_equal(evs[0], evs[1]):
                flips += 1
                continue
            fd = (dl[0] - dl[1]) / (2.0 * EPS * base)
            err = abs(grad - fd) / max(abs(fd), 1e-300)
            worst = max(worst, err)
            checked += 1
            assert err <= 1e-4, (int(i), which, grad, fd, err)
    print(f"per-parcel free-running FD: {checked} checks, worst rel err {worst:.2e}, {flips} flipped and left out")
    assert checked >= 2 * SAMPLE - 4
