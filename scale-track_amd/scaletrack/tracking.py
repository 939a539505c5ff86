"""Host layer over the C-ABI, named after the reference's tracking API.

Reference types and calls mirrored here (run/icoJuliaParcelFoam/cavity3D/):
  Mesh / construct_mesh           asyncParallelTracking.jl:222-270
  Chunk (parcel SoA)              asyncParallelTracking.jl:329-364
  TwoWayEulerian (U, UTrans)      asyncParallelTracking.jl:207-220
  Time                            asyncParallelTracking.jl:272-275
  evolve!(chunk, eulerian, mesh, dt, ::GPU)
                                  syncSerialTracking.jl:325-345
Every device computation goes through libscaletrack.so; there is no other path.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

FP64, FP32 = _lib.ST_FP64, _lib.ST_FP32
WALL, PERIODIC, CHANNEL = _lib.ST_WALL, _lib.ST_PERIODIC, _lib.ST_CHANNEL
PLAIN, SORTED = _lib.ST_METHOD_PLAIN, _lib.ST_METHOD_SORTED
FLUSH_OWNER, FLUSH_ATOMIC = _lib.ST_FLUSH_OWNER, _lib.ST_FLUSH_ATOMIC


def _dtype(precision):
    return np.float64 if precision == FP64 else np.float32


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class _PinnedBlock:
    """One st_host_alloc block exposed through the array interface; freed when the last
    array viewing it goes."""

    def __init__(self, shape, dtype):
        dt = np.dtype(dtype)
        shape = tuple(int(k) for k in (shape if np.iterable(shape) else (shape,)))
        nbytes = int(np.prod(shape, dtype=np.int64)) * dt.itemsize
        p = C.c_void_p()
        check(lib.st_host_alloc(max(nbytes, 1), C.byref(p)))
        self.ptr = p
        self.__array_interface__ = {"shape": shape, "typestr": dt.str, "data": (p.value, False), "version": 3}

    def __del__(self):
        ptr = getattr(self, "ptr", None)
        if ptr is not None and ptr.value and lib is not None:  # lib is gone at interpreter exit
            lib.st_host_free(ptr)
        self.ptr = None


def pinned_empty(shape, dtype=np.float64):
    """An uninitialised C-contiguous numpy array on page-locked host memory
    (st_host_alloc).  Copies between it and the device run by DMA at the bus's rate
    instead of through the runtime's bounce buffers.  Pinning costs more than one
    transfer of the same size: allocate once and reuse."""
    return np.asarray(_PinnedBlock(shape, dtype))


def pinned_copy(a):
    """`a` copied into a new page-locked array of the same shape and dtype."""
    a = np.asarray(a)
    out = pinned_empty(a.shape, a.dtype)
    out[...] = a
    return out


def device_count():
    n = C.c_int32(0)
    check(lib.st_device_count(C.byref(n)))
    return n.value


class Context:
    """One device and one HIP stream (st_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib.st_ctx_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        check(lib.st_ctx_synchronize(self.handle))

    def last_kernel_ms(self):
        ms, n = C.c_double(), C.c_int32()
        check(lib.st_last_kernel_ms(self.handle, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernel_times(self, max_calls=64):
        """Per-launch kernel ms and launch counts of the last timed calls, oldest first
        (one synchronisation for all of them)."""
        ms = (C.c_double * max_calls)()
        nl = (C.c_int32 * max_calls)()
        cnt = C.c_int32()
        check(lib.st_diag_kernel_times(self.handle, int(max_calls), ms, nl, C.byref(cnt)))
        return [(ms[i], nl[i]) for i in range(cnt.value)]

    def copy_bandwidth(self, nbytes=1 << 30, reps=10):
        """Measured device-to-device streaming copy, read + write GB/s."""
        g = C.c_double()
        check(lib.st_diag_copy_bandwidth(self.handle, int(nbytes), int(reps), C.byref(g)))
        return g.value

    def close(self):
        if self.handle:
            lib.st_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """Uniform box of N[0] x N[1] x N[2] hexahedral cells between origin and ending
    (construct_mesh, asyncParallelTracking.jl:237-262).  L = ending - origin,
    delta = L / N, rdelta = 1 / delta.  The device path handles decomposition (1,1,1),
    which is what one GPU runs (:1549-1552)."""

    def __init__(self, N, origin, ending, decomposition=(1, 1, 1)):
        N = [int(v) for v in (N if np.ndim(N) else (N, N, N))]
        origin = [float(v) for v in (origin if np.ndim(origin) else (origin,) * 3)]
        ending = [float(v) for v in (ending if np.ndim(ending) else (ending,) * 3)]
        decomposition = [int(v) for v in decomposition]
        if sum(n % d for n, d in zip(N, decomposition)) != 0:
            raise ValueError(
                "Each partition must have the same number of cells per direction.  The "
                f"current setup does not comply:\n N = {N} decomposition = {decomposition}")
        if decomposition != [1, 1, 1]:
            raise NotImplementedError("the device path runs one partition: decomposition (1,1,1)")
        self.N = tuple(N)
        self.origin = tuple(origin)
        self.ending = tuple(ending)
        self.L = tuple(e - o for o, e in zip(origin, ending))
        self.delta = tuple(l / n for l, n in zip(self.L, N))
        self.rdelta = tuple(1.0 / d for d in self.delta)
        self.decomposition = tuple(decomposition)
        self.partitionN = tuple(n // d for n, d in zip(N, decomposition))

    @property
    def ncell(self):
        return self.N[0] * self.N[1] * self.N[2]

    def st_mesh(self):
        m = _lib.st_mesh()
        for a in range(3):
            m.n[a] = self.N[a]
            m.origin[a] = self.origin[a]
            m.ending[a] = self.ending[a]
        return m


def construct_mesh(N, origin, ending, decomposition=(1, 1, 1)):
    return Mesh(N, origin, ending, decomposition)


class Time:
    """Simulation time and fluid step of a chunk (asyncParallelTracking.jl:272-275)."""

    def __init__(self, t=0.0, dt=0.0):
        self.t, self.dt = t, dt


class TwoWayEulerian:
    """Cell velocity U (input) and momentum source UTrans (output), AoS xyz per cell
    (asyncParallelTracking.jl:207-220), with their device copies."""

    def __init__(self, ctx, mesh, precision=FP64):
        self.ctx, self.mesh, self.precision = ctx, mesh, precision
        self.N = mesh.ncell
        h = C.c_void_p()
        m = mesh.st_mesh()
        check(lib.st_field_create(ctx.handle, C.byref(m), int(precision), C.byref(h)))
        self.handle = h
        self.U = np.zeros((self.N, 3), dtype=_dtype(precision))
        self.UTrans = np.zeros((self.N, 3), dtype=np.float64)

    def upload_U(self, U=None):
        if U is not None:
            self.U[...] = U
        a = np.ascontiguousarray(self.U)
        check(lib.st_field_upload_U(self.handle, _ptr(a), a.dtype.itemsize))

    def zero_source(self):
        check(lib.st_field_zero_source(self.handle))

    def merge_source(self):
        """Complete UTrans on the device (st_field_merge_source), without a host copy."""
        check(lib.st_field_merge_source(self.handle))

    def download_source(self, dtype=np.float64):
        out = np.empty((self.N, 3), dtype=dtype)
        check(lib.st_field_download_source(self.handle, _ptr(out), out.dtype.itemsize))
        if dtype == np.float64:
            self.UTrans[...] = out
        return out

    def close(self):
        if self.handle:
            check(lib.st_field_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Chunk:
    """A parcel cloud (asyncParallelTracking.jl:329-364): chunk scalars N, mu_c, rho,
    rho_d_by_rho_c, a bounding box and a Time, and per-parcel SoA arrays d, X, Y, Z,
    U, V, W on the host mirrored on the device.  `rho_p` (per-parcel density) and
    `rho_c` are extensions."""

    FIELDS = ("d", "X", "Y", "Z", "U", "V", "W")

    def __init__(self, ctx, N, mu_c, rho, rho_d_by_rho_c, precision=FP64, rho_p=None, rho_c=None):
        self.ctx, self.N, self.precision = ctx, int(N), precision
        self.mu_c, self.rho, self.rho_d_by_rho_c = float(mu_c), float(rho), float(rho_d_by_rho_c)
        self.rho_c = float(rho_c) if rho_c is not None else (
            self.rho / self.rho_d_by_rho_c if self.rho_d_by_rho_c else 0.0)
        dt = _dtype(precision)
        for f in self.FIELDS:
            setattr(self, f, np.zeros(self.N, dtype=dt))
        self.rho_p = None if rho_p is None else np.ascontiguousarray(rho_p, dtype=dt)
        self.boundingBox = (np.full(3, np.inf), np.full(3, -np.inf))
        self.time = Time()
        h = C.c_void_p()
        check(lib.st_cloud_create(ctx.handle, int(precision), self.N, self.mu_c, self.rho,
                                  self.rho_d_by_rho_c, self.rho_c, int(rho_p is not None), C.byref(h)))
        self.handle = h

    def set_state(self, d, x, y, z, u, v, w):
        for f, a in zip(self.FIELDS, (d, x, y, z, u, v, w)):
            getattr(self, f)[...] = a

    def upload(self):
        arrs = [np.ascontiguousarray(getattr(self, f)) for f in self.FIELDS]
        check(lib.st_cloud_upload(self.handle, *[_ptr(a) for a in arrs], _ptr(self.rho_p)))

    def download(self):
        """Device parcels into the host arrays, in upload order (parcels an open boundary
        removed come back with d < 0)."""
        n = self.count()
        if n != self.N:
            self._resize(n)
        arrs = [getattr(self, f) for f in self.FIELDS]
        check(lib.st_cloud_download(self.handle, *[_ptr(a) for a in arrs]))
        return arrs

    def count(self):
        n = C.c_int64()
        check(lib.st_cloud_count(self.handle, C.byref(n)))
        return n.value

    def _resize(self, n):
        dt = _dtype(self.precision)
        for f in self.FIELDS:
            setattr(self, f, np.zeros(n, dtype=dt))
        self.N = n

    def reserve(self, capacity):
        check(lib.st_cloud_reserve(self.handle, int(capacity)))

    def inject(self, d, x, y, z, u, v, w, rho=None):
        """Append parcels after the current ones (st_cloud_inject); they take the next
        upload-order indices."""
        dt = _dtype(self.precision)
        arrs = [np.ascontiguousarray(a, dtype=dt) for a in (d, x, y, z, u, v, w)]
        r = None if rho is None else np.ascontiguousarray(rho, dtype=dt)
        k = arrs[0].size
        check(lib.st_cloud_inject(self.handle, k, *[_ptr(a) for a in arrs], _ptr(r)))
        # the host arrays are resized (not refilled) by the next download
        if r is not None and self.rho_p is not None:
            self.rho_p = np.concatenate([self.rho_p, r])

    def compact(self):
        """Drop the parcels an open boundary removed (st_cloud_compact); returns how many
        are left.  The host arrays are resized (their contents are refreshed by the next
        download)."""
        n = C.c_int64()
        check(lib.st_cloud_compact(self.handle, C.byref(n)))
        if n.value != self.N:
            self._resize(n.value)
        return n.value

    def sort(self, ctx, eulerian, p=None):
        """Sort the device parcels by brick of `eulerian`'s mesh now (st_cloud_sort);
        otherwise the first sorted step after an upload sorts."""
        check(lib.st_cloud_sort(ctx.handle, self.handle, eulerian.handle, C.byref(p) if p is not None else None))

    def stray_fraction(self):
        f = C.c_double()
        check(lib.st_cloud_stray_fraction(self.handle, C.byref(f)))
        return f.value

    def resort_count(self):
        """Re-sorts the device decided on so far (st_cloud_resort_count)."""
        n = C.c_int64()
        check(lib.st_cloud_resort_count(self.handle, C.byref(n)))
        return n.value

    def close(self):
        if self.handle:
            check(lib.st_cloud_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def params(**kw):
    """st_params with the library defaults, overridden by keyword."""
    return _lib.default_params(**kw)


def increment_time(chunk, dt):
    """increment_time! (asyncParallelTracking.jl:497-502)."""
    chunk.time = Time(chunk.time.t + dt, dt)


def evolve_device(ctx, chunk, eulerian, dt_fluid, nsub=10, p=None, want_bbox=True):
    """One st_evolve call: advance the chunk by nsub substeps of dt_fluid / nsub and add
    the momentum source into the field's UTrans on the device."""
    bb = (C.c_double * 6)() if want_bbox else None
    check(lib.st_evolve(ctx.handle, chunk.handle, eulerian.handle, float(dt_fluid), int(nsub),
                        C.byref(p) if p is not None else None, bb))
    if want_bbox:
        chunk.boundingBox = (np.array(bb[0:3]), np.array(bb[3:6]))
    return chunk.boundingBox if want_bbox else None


def evolve(chunk, eulerian, mesh, dt, nsub=10, p=None):
    """evolve!(chunk, eulerian, mesh, dt, ::GPU) of syncSerialTracking.jl:325-345:
    copy U to the device, advance the time, run the step, wait, copy UTrans back.
    As in syncSerialTracking.jl:553 the caller zeroes UTrans per fluid step (see
    evolve_cloud); this call accumulates."""
    eulerian.upload_U()
    increment_time(chunk, dt)
    evolve_device(eulerian.ctx, chunk, eulerian, dt, nsub, p, want_bbox=True)
    eulerian.download_source()


class Graph:
    """A captured fluid step (st_graph): optional UTrans zeroing plus nsub substeps."""

    def __init__(self, ctx, chunk, eulerian, dt_fluid, nsub=10, p=None, zero_source=True):
        h = C.c_void_p()
        check(lib.st_graph_capture(ctx.handle, chunk.handle, eulerian.handle, float(dt_fluid),
                                   int(nsub), C.byref(p) if p is not None else None,
                                   int(bool(zero_source)), C.byref(h)))
        self.handle = h
        self._keep = (chunk, eulerian)  # the library refuses to free them while the graph lives

    def launch(self):
        check(lib.st_graph_launch(self.handle))

    def close(self):
        if self.handle:
            lib.st_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent clouds advanced together by one persistent launch per pass (st_batch).
    `target_parcels_per_brick` > 0 sizes each cloud's brick to hold about that many
    parcels; `blocks_per_cu` > 0 caps the persistent grid at that many workgroups per CU
    and bounds each cloud's brick so its tile fits that share of a CU's LDS (0: as many
    as fit)."""

    def __init__(self, ctx, chunks, fields, p=None, target_parcels_per_brick=4096, blocks_per_cu=0):
        n = len(chunks)
        assert n == len(fields) and n > 0
        self.ctx, self.chunks, self.fields = ctx, list(chunks), list(fields)
        arr_c = (C.c_void_p * n)(*[c.handle.value for c in chunks])
        arr_f = (C.c_void_p * n)(*[f.handle.value for f in fields])
        h = C.c_void_p()
        check(lib.st_batch_create(ctx.handle, n, C.cast(arr_c, C.POINTER(C.c_void_p)),
                                  C.cast(arr_f, C.POINTER(C.c_void_p)),
                                  C.byref(p) if p is not None else None, int(target_parcels_per_brick),
                                  int(blocks_per_cu), C.byref(h)))
        self.handle = h

    def evolve(self, dt_fluid, nsub=10):
        check(lib.st_batch_evolve(self.handle, float(dt_fluid), int(nsub)))

    def sort(self):
        check(lib.st_batch_sort(self.handle))

    def info(self):
        n = len(self.chunks)
        items, grid = C.c_int32(), C.c_int32()
        bricks = (C.c_int32 * (3 * n))()
        check(lib.st_batch_info(self.handle, C.byref(items), C.byref(grid), bricks))
        return {"items": items.value, "grid": grid.value,
                "bricks": [tuple(bricks[3 * i:3 * i + 3]) for i in range(n)]}

    def graph(self, dt_fluid, nsub=10, zero_source=True):
        g = Graph.__new__(Graph)
        h = C.c_void_p()
        check(lib.st_batch_graph_capture(self.handle, float(dt_fluid), int(nsub), int(bool(zero_source)),
                                         C.byref(h)))
        g.handle = h
        g._keep = (self,)
        return g

    def close(self):
        if self.handle:
            check(lib.st_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_batch(ctx, chunks, fields, p=None, dt_fluid=1e-3, trials=5, cache_path=None):
    """st_autotune_batch: the fastest (target parcels per brick, block, workgroups per CU)
    for these clouds, timed on a snapshot that is restored afterwards; cached per device
    and workload in `cache_path` (JSON lines) when given."""
    n = len(chunks)
    arr_c = (C.c_void_p * n)(*[c.handle.value for c in chunks])
    arr_f = (C.c_void_p * n)(*[f.handle.value for f in fields])
    r = _lib.st_tune_result()
    check(lib.st_autotune_batch(ctx.handle, n, C.cast(arr_c, C.POINTER(C.c_void_p)),
                                C.cast(arr_f, C.POINTER(C.c_void_p)), C.byref(p) if p is not None else None,
                                float(dt_fluid), int(trials),
                                None if cache_path is None else str(cache_path).encode(), C.byref(r)))
    return {"target_parcels_per_brick": r.target_parcels_per_brick, "block": r.block,
            "blocks_per_cu": r.blocks_per_cu, "ms_per_pass": r.ms_per_pass,
            "runner_up_ms_per_pass": r.runner_up_ms_per_pass, "from_cache": bool(r.from_cache)}
