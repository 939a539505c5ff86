"""The reference's coupling API and its asynchronous exchange with the fluid solver,
driving the device step through the C-ABI (one rank, one GPU).

Reference (run/icoJuliaParcelFoam/cavity3D/asyncParallelTracking.jl):
  allocate_array_j         :1436-1483   hands the host U / UTrans storage to the solver
  evolve_cloud             :1485-1528   one fluid step's call into the tracker
  evolve!(control, ::GPU)  :1206-1223   the solver-side half of the handshake
  Master loop              :933-1190    the tracker thread: copy U, run the step, wait,
                                        copy UTrans back, estimate_source!
  estimate_source!         :738-747     constant extrapolator
  finalize_evolve          :1225-1243
  save_timings             :73-109      stats_np#### file

Timing of the exchange (lag one): the parcel step run during fluid step n uses the U
of the end of step n-1, and fluid step n reads the extrapolated source of the parcel
step run during step n-1.  Events reset automatically when a waiter takes them, like
Julia's Event(true).
"""
import ctypes as C
import math
import os
import threading
import time

import numpy as np

from . import tracking as trk


class AutoResetEvent:
    """notify() wakes one waiter (or the next one to wait); the event then resets."""

    def __init__(self):
        self._cv = threading.Condition()
        self._set = False

    def notify(self):
        with self._cv:
            self._set = True
            self._cv.notify()

    def wait(self, timeout=None):
        with self._cv:
            if not self._cv.wait_for(lambda: self._set, timeout):
                raise TimeoutError("event not notified in time")
            self._set = False


class ConstExtrapolator:
    """estimate_source! (:738-747): S <- 2 S - prev; prev <- S."""

    def __init__(self, n, dtype=np.float64):
        self.prevUTrans = np.zeros((n, 3), dtype=dtype)

    def estimate_source(self, cur):
        cur[...] = 2.0 * cur - self.prevUTrans
        self.prevUTrans[...] = cur


class Events:
    def __init__(self):
        self.U_copied = AutoResetEvent()
        self.S_copied = AutoResetEvent()
        self.Eulerian_computed = AutoResetEvent()


class GpuExecutor:
    """The device side of the tracker: one st_ctx owned by the tracker thread, one
    cloud per chunk and one field.

    Extension (an inflow; the reference's chunks keep a fixed N): `injector(step)`, if
    given, returns for parcel step `step` (1, 2, ...) one entry per chunk, None or the
    arrays d, x, y, z, u, v, w of parcels entering before the step (st_cloud_inject);
    with `compact_every` = k > 0 the parcels an open boundary removed are dropped after
    every k-th step (st_cloud_compact).  `final` holds each chunk's parcels (upload
    order) when the executor stops."""

    def __init__(self, mesh, chunks_host, precision, nsub=10, params=None, device=0, injector=None,
                 compact_every=0):
        self.mesh, self.chunks_host, self.precision = mesh, chunks_host, precision
        self.nsub, self.params, self.device = nsub, params, device
        self.injector, self.compact_every = injector, int(compact_every)
        self.step = 0
        self.final = None

    def start(self):
        self.ctx = trk.Context(self.device)
        self.field = trk.TwoWayEulerian(self.ctx, self.mesh, self.precision)
        self.chunks = []
        for h in self.chunks_host:
            c = trk.Chunk(self.ctx, h["N"], h["mu_c"], h["rho"], h["rho_d_by_rho_c"], self.precision)
            c.set_state(*[h[k] for k in ("d", "X", "Y", "Z", "U", "V", "W")])
            c.upload()
            self.chunks.append(c)

    def zero_source(self):
        self.field.zero_source()

    def upload_U(self, U):
        self.field.upload_U(U)

    def evolve(self, dt):
        self.step += 1
        if self.injector is not None:
            for c, new in zip(self.chunks, self.injector(self.step)):
                if new is not None and len(new[0]):
                    c.inject(*new)
        boxes = []
        for c in self.chunks:
            boxes.append(trk.evolve_device(self.ctx, c, self.field, dt, self.nsub, self.params, want_bbox=True))
        if self.compact_every > 0 and self.step % self.compact_every == 0:
            for c in self.chunks:
                c.compact()
        return boxes

    def download_source(self, out):
        out[...] = self.field.download_source(dtype=out.dtype)

    def download_chunks(self):
        return [c.download() for c in self.chunks]

    def stop(self):
        try:
            self.final = [[a.copy() for a in c.download()] for c in self.chunks]
        finally:
            self._close()

    def _close(self):
        for c in self.chunks:
            c.close()
        self.field.close()
        self.ctx.close()


class AsyncTracker:
    """Single-rank tracker with the reference's C-callable API.

    `chunks_host` is a list of dicts with the chunk scalars N, mu_c, rho,
    rho_d_by_rho_c and the arrays d, X, Y, Z, U, V, W.  The tracker thread starts in
    the constructor, as init_async_evolve! is spawned at load time (:1625-1629)."""

    def __init__(self, mesh, chunks_host, precision=trk.FP32, nsub=10, params=None, device=0,
                 executor=None, nSkipInitTimeStepsProfiling=2, injector=None, compact_every=0):
        self.mesh = mesh
        self.precision = precision
        dtype = np.float64 if precision == trk.FP64 else np.float32
        self.scalar = dtype
        # host Eulerian fields: the storage handed to the solver by allocate_array_j
        self.U = np.zeros((mesh.ncell, 3), dtype=dtype)
        self.UTrans = np.zeros((mesh.ncell, 3), dtype=dtype)
        self.events = Events()
        self.eulerian_lock = threading.Lock()
        self.chunk_transfers = threading.Lock()
        self.extrapolator = ConstExtrapolator(mesh.ncell, dtype)
        self.executor = executor or GpuExecutor(mesh, chunks_host, precision, nsub, params, device, injector,
                                                compact_every)
        self.iTimeStep = 0
        self.dt = None
        self.boxes = []
        self.reg = {"dtVecAll": [], "dtVecEuler": [], "dtVecEvolve": [], "dtVecDeviceCompute": [],
                    "nSkipInitTimeStepsProfiling": nSkipInitTimeStepsProfiling}
        self._t_all = time.time()
        self._t_euler = time.time()
        self._stop = False
        self._error = None
        self._ready = threading.Event()
        self.eulerian_lock.acquire()  # the solver thread owns the Eulerian between calls (:1592)
        self._thread = threading.Thread(target=self._master, name="tracker", daemon=True)
        self._thread.start()
        self._ready.wait()
        if self._error:
            raise self._error

    # ---- solver-side API ------------------------------------------------------
    def allocate_array_j(self, name, size, nComponents, typeByteSize):
        """Pointer to the host field `name` ("U" or "UTrans") after the reference's checks
        (:1444-1465): element size and cell count must match."""
        if int(typeByteSize) != np.dtype(self.scalar).itemsize:
            raise ValueError(f"Size of Julia type (in Bytes): {np.dtype(self.scalar).itemsize} is different "
                             f"from the extern type: {typeByteSize}")
        if int(size) != self.mesh.ncell:
            raise ValueError(f"Size of Julia mesh: {self.mesh.ncell} is different from field's {name} size: "
                             f"{int(size)}")
        if name not in ("U", "UTrans") or int(nComponents) != 3:
            raise ValueError(f"no field {name} with {nComponents} components")
        return getattr(self, name).ctypes.data

    def evolve_cloud(self, dt):
        """One fluid step's call (:1485-1528): timings, time update, then the handshake of
        evolve!(control, ::GPU) (:1206-1223)."""
        self._check()
        self.iTimeStep += 1
        now = time.time()
        dt_all, dt_euler = now - self._t_all, now - self._t_euler
        self._t_all = now
        if self.iTimeStep > self.reg["nSkipInitTimeStepsProfiling"]:
            self.reg["dtVecAll"].append(dt_all)
            self.reg["dtVecEuler"].append(dt_euler)
        self.dt = float(dt)
        self.eulerian_lock.release()
        if self.iTimeStep > 1:
            self.events.Eulerian_computed.notify()
            self._wait(self.events.S_copied)
        self._wait(self.events.U_copied)
        self.eulerian_lock.acquire()
        self._t_euler = time.time()

    def finalize_evolve(self, stats_dir=None):
        """Drain the tracker thread (:1225-1243) and write the timing statistics."""
        if stats_dir is not None:
            self.save_timings(stats_dir)
        self._stop = True
        self.eulerian_lock.release()
        self.events.Eulerian_computed.notify()
        self._thread.join(timeout=600)
        self._check()

    # ---- tracker thread --------------------------------------------------------
    def _wait(self, ev):
        while True:
            try:
                ev.wait(timeout=1.0)
                return
            except TimeoutError:
                self._check()

    def _check(self):
        if self._error is not None:
            raise RuntimeError("tracker thread failed") from self._error

    def _master(self):
        try:
            self.executor.start()
        except Exception as e:  # surfaced by the constructor
            self._error = e
            self._ready.set()
            return
        self._ready.set()
        step = 0
        try:
            while True:
                self.executor.zero_source()
                with self.eulerian_lock:
                    if self._stop:
                        break
                    self.executor.upload_U(self.U)
                self.events.U_copied.notify()
                t0 = time.time()
                with self.chunk_transfers:
                    self.boxes = self.executor.evolve(self.dt)
                dt_dev = time.time() - t0
                step += 1
                if step > self.reg["nSkipInitTimeStepsProfiling"]:
                    self.reg["dtVecDeviceCompute"].append(dt_dev)
                    self.reg["dtVecEvolve"].append(dt_dev)
                self.events.Eulerian_computed.wait()
                if self._stop:
                    self.events.S_copied.notify()
                    self.events.U_copied.notify()
                    break
                with self.eulerian_lock:
                    self.executor.download_source(self.UTrans)
                    self.extrapolator.estimate_source(self.UTrans)
                self.events.S_copied.notify()
        except Exception as e:
            self._error = e
        finally:
            try:
                self.executor.stop()
            except Exception:
                pass

    # ---- output ----------------------------------------------------------------
    def save_timings(self, directory, nranks=1):
        """stats_np#### of :73-109: counts, totals, means and standard deviations."""
        r = self.reg
        n = min(len(r["dtVecEuler"]), len(r["dtVecEvolve"]))
        cols = [r["dtVecAll"][:n], r["dtVecEuler"][:n], r["dtVecEvolve"][:n], r["dtVecDeviceCompute"][:n]]
        vals = [sum(c) for c in cols] + [float(np.mean(c)) if c else 0.0 for c in cols] + \
               [float(np.std(c, ddof=1)) if len(c) > 1 else 0.0 for c in cols]
        path = os.path.join(directory, f"stats_np{nranks:04d}")
        with open(path, "w") as io:
            io.write("nTimeSteps tAllTotal tEulerTotal tEvolveTotal tDeviceComputeTotal "
                     "tAllMean tEulerMean tEvolveMean tDeviceComputeMean "
                     "tAllStd tEulerStd tEvolveStd tDeviceComputeStd\n")
            io.write(str(n) + "".join(" " + _sig4(v) for v in vals) + "\n")
        return path


def borrow(tracker, name):
    """The solver-side view of a tracker field: the pointer allocate_array_j hands out,
    wrapped without a copy (juliaWrapper::reallocate swaps OpenFOAM's storage to it,
    juliaWrapper.H:119-172).  Returns the (ncell, 3) array and the pointer, which
    check_pointers compares against later."""
    dt = np.dtype(tracker.scalar)
    ptr = tracker.allocate_array_j(name, tracker.mesh.ncell, 3, dt.itemsize)
    buf = (C.c_char * (tracker.mesh.ncell * 3 * dt.itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dt).reshape(tracker.mesh.ncell, 3), ptr


def check_pointers(tracker, expected):
    """juliaWrapper::checkPointers (juliaWrapper.H:78-97): the storage the solver borrowed
    must still be the tracker's."""
    for name, ptr in expected.items():
        if getattr(tracker, name).ctypes.data != ptr:
            raise RuntimeError(f"Pointer of field {name} was changed by OpenFOAM")


def run_coupled(solver, tracker, dt, steps, per_volume=False, on_step=None):
    """The solver's time loop of icoJuliaParcelFoam.C:118-186 with the tracker: each step
    checks the borrowed pointers, calls evolve_cloud(dt) (the lag-one exchange), then
    solves the fluid step with the momentum source UTrans / dt, writing U into the
    tracker's storage.  `solver.U` must be the borrowed U (see borrow).  on_step(k, U)
    is called with the velocity the parcel step k is about to be given."""
    ptrs = {"U": tracker.U.ctypes.data, "UTrans": tracker.UTrans.ctypes.data}
    for k in range(steps):
        check_pointers(tracker, ptrs)
        if on_step is not None:
            on_step(k, tracker.U)
        tracker.evolve_cloud(dt)
        solver.step(dt, tracker.UTrans, per_volume=per_volume)


def _sig4(x):
    """round(x, sigdigits=4) printed as Julia prints a Float32."""
    if x == 0 or not math.isfinite(x):
        return repr(float(x))
    d = 4 - int(math.floor(math.log10(abs(x)))) - 1
    return repr(float(np.float32(round(x, d))))
