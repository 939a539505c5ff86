"""Host-side logic that needs no GPU: the lag-one exchange of evolve_cloud with the
fluid solver, the constant extrapolator, allocate_array_j's checks, the timing file and
the VTK particle output.  The exchange is driven with a test-only executor whose
"step" is a known function of the uploaded U, so the timing of every copy is visible.
"""
import ctypes as C
import os

import numpy as np
import pytest

from scaletrack import coupling, vtk
from scaletrack.tracking import FP64, Mesh


class RecordingExecutor:
    """Stands in for the device in these CPU tests: source_k = 2 U_k + k."""

    def __init__(self, ncell):
        self.ncell = ncell
        self.k = 0
        self.U = None
        self.src = np.zeros((ncell, 3))
        self.uploads = []

    def start(self):
        pass

    def zero_source(self):
        self.src[...] = 0.0

    def upload_U(self, U):
        self.U = U.copy()
        self.uploads.append(self.U.copy())

    def evolve(self, dt):
        self.k += 1
        self.src += 2.0 * self.U + self.k
        return []

    def download_source(self, out):
        out[...] = self.src

    def stop(self):
        pass


def test_async_exchange_is_lag_one_with_extrapolation():
    mesh = Mesh((2, 2, 1), 0.0, 1.0)
    ex = RecordingExecutor(mesh.ncell)
    tr = coupling.AsyncTracker(mesh, [], precision=FP64, executor=ex)
    U = np.frombuffer((C.c_double * (mesh.ncell * 3)).from_address(tr.allocate_array_j("U", 4, 3, 8)),
                      dtype=np.float64).reshape(-1, 3)
    S = np.frombuffer((C.c_double * (mesh.ncell * 3)).from_address(tr.allocate_array_j("UTrans", 4, 3, 8)),
                      dtype=np.float64).reshape(-1, 3)
    U[...] = 1.0  # initial field
    est_prev = np.zeros((mesh.ncell, 3))
    seen = []
    for n in range(1, 6):
        tr.evolve_cloud(1e-3)
        seen.append(S.copy())
        # the fluid step n writes the field of the end of step n
        U[...] = 10.0 * n
    tr.finalize_evolve()
    # parcel step k used the field of the end of fluid step k-1
    assert [u[0, 0] for u in ex.uploads[:5]] == [1.0, 10.0, 20.0, 30.0, 40.0]
    # after evolve_cloud(n) the solver sees the extrapolated source of parcel step n-1
    np.testing.assert_array_equal(seen[0], 0.0)
    for n in range(2, 6):
        k = n - 1
        s_true = 2.0 * ex.uploads[k - 1] + k
        est = 2.0 * s_true - est_prev
        np.testing.assert_array_equal(seen[n - 1], est)
        est_prev = est


def test_allocate_array_checks():
    mesh = Mesh((2, 2, 1), 0.0, 1.0)
    tr = coupling.AsyncTracker(mesh, [], precision=FP64, executor=RecordingExecutor(mesh.ncell))
    with pytest.raises(ValueError, match="different from the extern type"):
        tr.allocate_array_j("U", 4, 3, 4)
    with pytest.raises(ValueError, match="Size of Julia mesh"):
        tr.allocate_array_j("U", 5, 3, 8)
    assert tr.allocate_array_j("UTrans", 4, 3, 8) == tr.UTrans.ctypes.data
    tr.finalize_evolve()


def test_extrapolator_matches_reference_rule():
    ex = coupling.ConstExtrapolator(3)
    s1 = np.arange(9.0).reshape(3, 3)
    cur = s1.copy()
    ex.estimate_source(cur)
    np.testing.assert_array_equal(cur, 2 * s1)
    s2 = np.ones((3, 3))
    cur2 = s2.copy()
    ex.estimate_source(cur2)
    np.testing.assert_array_equal(cur2, 2 * s2 - 2 * s1)


def test_timings_file(tmp_path):
    mesh = Mesh((1, 1, 1), 0.0, 1.0)
    tr = coupling.AsyncTracker(mesh, [], precision=FP64, executor=RecordingExecutor(1),
                               nSkipInitTimeStepsProfiling=0)
    for _ in range(4):
        tr.evolve_cloud(1e-3)
    tr.finalize_evolve(stats_dir=str(tmp_path))
    lines = open(os.path.join(tmp_path, "stats_np0001")).read().splitlines()
    assert lines[0].split() == ["nTimeSteps", "tAllTotal", "tEulerTotal", "tEvolveTotal",
                                "tDeviceComputeTotal", "tAllMean", "tEulerMean", "tEvolveMean",
                                "tDeviceComputeMean", "tAllStd", "tEulerStd", "tEvolveStd",
                                "tDeviceComputeStd"]
    vals = lines[1].split()
    assert len(vals) == 13 and int(vals[0]) >= 3


def test_tracker_thread_error_surfaces():
    class Broken(RecordingExecutor):
        def evolve(self, dt):
            raise RuntimeError("device lost")

    mesh = Mesh((1, 1, 1), 0.0, 1.0)
    tr = coupling.AsyncTracker(mesh, [], precision=FP64, executor=Broken(1))
    tr.evolve_cloud(1e-3)
    with pytest.raises(RuntimeError, match="tracker thread failed"):
        tr.evolve_cloud(1e-3)


def test_vtu_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    n = 37

    class Ch:
        pass

    ch = Ch()
    for f in ("X", "Y", "Z", "U", "V", "W", "d"):
        setattr(ch, f, rng.random(n))
    coll = vtk.ParaviewCollection(str(tmp_path / "particleTimeSeries.pvd"))
    p = vtk.write_chunk(ch, 0.0123456, coll, directory=str(tmp_path / "dataVTK"))
    assert p.endswith("particleFields_0.01235.vtu")
    coll.save()
    a = vtk.read_vtu(p)
    np.testing.assert_array_equal(a["Points"], np.stack([ch.X, ch.Y, ch.Z], 1))
    np.testing.assert_array_equal(a["U"], np.stack([ch.U, ch.V, ch.W], 1))
    np.testing.assert_array_equal(a["d"], ch.d)
    np.testing.assert_array_equal(a["types"], np.full(n, vtk.VTK_VERTEX))
    text = open(coll.path).read()
    assert 'timestep="0.01235"' in text and "dataVTK/particleFields_0.01235.vtu" in text


def _cell(x, o, e, n):
    return np.floor((np.asarray(x, dtype=np.float64) - o) * (1.0 / ((e - o) / n)))


def test_into_cells_moves_only_coordinates_outside_the_cells():
    """A coordinate inside [origin, ending) can compute to cell -1 or N through the
    rounding of (x - origin) * rdelta, or of origin itself to float.  into_cells moves
    exactly those, to the nearest value inside the cells, and leaves every other
    coordinate alone."""
    from scaletrack import inputs

    # upper face, double: the largest double below ending computes to cell N
    o, e, n = -0.9535514630027344, -0.023002589931070516, 17
    top = np.nextafter(e, -np.inf)
    assert top < e and _cell(top, o, e, n) == n
    x = np.array([top, o, 0.5 * (o + e)])
    y = inputs.into_cells(x, o, e, n)
    assert np.all((_cell(y, o, e, n) >= 0) & (_cell(y, o, e, n) < n))
    assert y[0] < top and _cell(np.nextafter(y[0], np.inf), o, e, n) == n  # the nearest value inside
    np.testing.assert_array_equal(y[1:], x[1:])

    # lower face, float: origin rounds below itself in float32 and computes to cell -1
    o, e, n = -1.1443731422599845, 1.407537619706154, 26
    low = np.float32(o)
    assert float(low) < o and _cell(low, o, e, n) == -1
    y = inputs.into_cells(np.array([low, np.float32(0.0)], dtype=np.float32), o, e, n)
    assert y.dtype == np.float32 and _cell(y[0], o, e, n) == 0 and y[0] == np.nextafter(low, np.float32(1))
    assert y[1] == 0.0


def test_positions_and_cases_lie_in_the_cells():
    """positions() keeps float coordinates at or above origin; a Case moves the rest of
    its coordinates into the cells, so every synthetic cloud passes the upload check."""
    from scaletrack import inputs

    o, e, n = (-1.1443731422599845, -0.9535514630027344, 0.0), (1.407537619706154, -0.023002589931070516, 1.0), (26, 17, 4)
    for dt in (np.float32, np.float64):
        xs = inputs.positions(5, 20000, o, e, dtype=dt)
        for a in range(3):
            assert np.all(xs[a].astype(np.float64) >= o[a]) and np.all(xs[a].astype(np.float64) < e[a])
        xs[1][0] = np.nextafter(e[1], -np.inf)  # computes to cell N in double
        z = np.zeros(20000, dtype=dt)
        case = inputs.Case(n, o, e, np.zeros((26 * 17 * 4, 3), dtype=dt), z + 1e-5, *xs, z, z, z)
        for a, arr in enumerate((case.x, case.y, case.z)):
            c = _cell(arr, o[a], e[a], n[a])
            assert np.all((c >= 0) & (c < n[a])), a
