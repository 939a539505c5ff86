"""ctypes wrapper of the CPU oracle (liboracle built from st_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's CPU
baseline leg.  The product (scale-track_amd/) never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libst_oracle.so")

FP64, FP32 = 0, 1
WALL, PERIODIC, CHANNEL = 0, 1, 2


class Args(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("boundary", C.c_int32),
        ("n", C.c_int32 * 3),
        ("origin", C.c_double * 3),
        ("ending", C.c_double * 3),
        ("np", C.c_int64),
        ("d", C.c_void_p), ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
        ("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
        ("rho", C.c_void_p),
        ("mu_c", C.c_double),
        ("rho_d", C.c_double),
        ("rho_d_by_rho_c", C.c_double),
        ("rho_c", C.c_double),
        ("cd", C.c_double),
        ("U", C.c_void_p),
        ("source", C.c_void_p),
        ("dt_fluid", C.c_double),
        ("nsub", C.c_int32),
        ("substeps_per_pass", C.c_int32),
        ("nthreads", C.c_int32),
        ("bbox", C.c_double * 6),
        ("events", C.c_void_p),
        ("events_mode", C.c_int32),
        ("tan_d", C.c_void_p),
        ("tan_rho", C.c_void_p),
        ("tan_cd", C.c_double),
        ("tan_source", C.c_void_p),
        ("tan_state", C.c_void_p),
    ]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    lib.st_oracle_evolve.restype = C.c_int
    lib.st_oracle_evolve.argtypes = [C.POINTER(Args)]
    lib.st_oracle_locate.restype = C.c_int
    lib.st_oracle_locate.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.st_oracle_extrapolate.restype = None
    lib.st_oracle_extrapolate.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    return lib


lib = _load()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def evolve(state, U, n, origin, ending, dt_fluid, nsub=10, precision=FP64, boundary=WALL,
           mu_c=1e-3, rho_d=1000.0, rho_d_by_rho_c=1.0, rho_c=1000.0, cd=1.0, rho=None,
           source=None, substeps_per_pass=0, nthreads=1, events=None, events_mode=0,
           tan_d=None, tan_rho=None, tan_cd=0.0, tan_source=None, tan_state=None):
    """Advance `state` = [d, x, y, z, u, v, w] (numpy arrays, modified in place) and add
    the momentum source into `source` (float64, (ncell, 3)).  Returns (source, bbox).

    events: int64 array (np, nsub) recorded (events_mode=1) or replayed (2).
    tan_source: float64 (ncell, 3); when given, the run is forward-mode dual numbers
    seeded by tan_d, tan_rho (per parcel) and tan_cd, with state tangents in/out in
    tan_state (6, np) if given."""
    dt = np.float64 if precision == FP64 else np.float32
    for a in state:
        assert a.dtype == dt and a.flags.c_contiguous
    U = np.ascontiguousarray(U, dtype=dt)
    ncell = int(n[0]) * int(n[1]) * int(n[2])
    if source is None:
        source = np.zeros((ncell, 3), dtype=np.float64)
    assert source.dtype == np.float64 and source.flags.c_contiguous and source.size == 3 * ncell
    if rho is not None:
        rho = np.ascontiguousarray(rho, dtype=dt)
    a = Args()
    a.precision, a.boundary = precision, boundary
    for i in range(3):
        a.n[i], a.origin[i], a.ending[i] = int(n[i]), float(origin[i]), float(ending[i])
    a.np = state[0].size
    a.d, a.x, a.y, a.z, a.u, a.v, a.w = [_p(s) for s in state]
    a.rho = _p(rho)
    a.mu_c, a.rho_d, a.rho_d_by_rho_c, a.rho_c, a.cd = mu_c, rho_d, rho_d_by_rho_c, rho_c, cd
    a.U, a.source = _p(U), _p(source)
    a.dt_fluid, a.nsub, a.substeps_per_pass, a.nthreads = dt_fluid, nsub, substeps_per_pass, nthreads
    keep = []
    if events is not None:
        assert events.dtype == np.int64 and events.size == a.np * nsub and events.flags.c_contiguous
        a.events, a.events_mode = _p(events), events_mode
    if tan_source is not None:
        assert tan_source.dtype == np.float64 and tan_source.size == 3 * ncell
        for name, arr in (("tan_d", tan_d), ("tan_rho", tan_rho)):
            if arr is not None:
                arr = np.ascontiguousarray(arr, dtype=np.float64)
                keep.append(arr)
                setattr(a, name, _p(arr))
        if tan_state is not None:
            assert tan_state.dtype == np.float64 and tan_state.shape == (6, a.np) and tan_state.flags.c_contiguous
            a.tan_state = _p(tan_state)
        a.tan_cd = tan_cd
        a.tan_source = _p(tan_source)
    r = lib.st_oracle_evolve(C.byref(a))
    if r != 0:
        raise ValueError("st_oracle_evolve rejected its arguments")
    return source, np.array(a.bbox[:])


def locate(n, origin, ending, x, y, z):
    """(I, J, K), 0-based folded cell index, 0-based partition index."""
    nn = (C.c_int32 * 3)(*n)
    oo = (C.c_double * 3)(*origin)
    ee = (C.c_double * 3)(*ending)
    ijk = (C.c_int32 * 3)()
    cell, part = C.c_int64(), C.c_int64()
    lib.st_oracle_locate(nn, oo, ee, x, y, z, ijk, C.byref(cell), C.byref(part))
    return tuple(ijk), cell.value, part.value


def extrapolate(cur, prev):
    lib.st_oracle_extrapolate(_p(cur), _p(prev), cur.size)


def evolve_case(case, state=None, source=None, precision=FP64, boundary=None, **kw):
    """Run the oracle on an inputs.Case (or a given state copy)."""
    st = case.state() if state is None else state
    return evolve(st, case.U, case.n, case.origin, case.ending, case.dt_fluid,
                  nsub=kw.pop("nsub", case.nsub), precision=precision,
                  boundary=case.boundary if boundary is None else boundary, mu_c=case.mu_c,
                  rho_d=case.rho_d, rho_d_by_rho_c=case.rho_d_by_rho_c, rho_c=case.rho_c,
                  rho=case.rho, source=source, **kw) + (st,)
