"""Reverse-mode gradients of the parcel step (st_adjoint_* of the C-ABI), FP64.

`source_misfit_rrule` follows the ChainRulesCore `rrule` contract: it returns the
primal value and a pullback that maps the loss cotangent to the parameter cotangents.
The derivative is the almost-everywhere one: the forward run's cells, bounces and
wraps are held fixed.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Adjoint:
    """Checkpointed adjoint of `nsub` substeps of a chunk through a field; the state is
    saved every `checkpoint_every` substeps and recomputed in the reverse sweep."""

    def __init__(self, ctx, chunk, field, nsub, checkpoint_every, p=None):
        self.ctx, self.chunk, self.field = ctx, chunk, field
        self.nsub, self.K = int(nsub), int(checkpoint_every)
        h = C.c_void_p()
        check(lib.st_adjoint_create(ctx.handle, chunk.handle, field.handle, self.nsub, self.K,
                                    C.byref(p) if p is not None else None, C.byref(h)))
        self.handle = h

    def forward(self, dt_fluid):
        """Zero UTrans and advance the chunk by nsub substeps of dt_fluid / nsub."""
        check(lib.st_adjoint_forward(self.handle, float(dt_fluid)))

    def source_misfit(self, target):
        """loss = sum (UTrans - target)^2, and dL/dUTrans as the backward seed."""
        t = np.ascontiguousarray(target, dtype=np.float64)
        assert t.size == self.field.N * 3
        loss = C.c_double()
        check(lib.st_adjoint_source_misfit(self.handle, _ptr(t), C.byref(loss)))
        return loss.value

    def set_source_cotangent(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        assert g.size == self.field.N * 3
        check(lib.st_adjoint_set_source_cotangent(self.handle, _ptr(g)))

    def set_state_cotangent(self, xbar=None, vbar=None):
        """Cotangents of the final positions and velocities, arrays of shape (3, n)."""
        xb = None if xbar is None else np.ascontiguousarray(xbar, dtype=np.float64)
        vb = None if vbar is None else np.ascontiguousarray(vbar, dtype=np.float64)
        check(lib.st_adjoint_set_state_cotangent(self.handle, _ptr(xb), _ptr(vb)))

    def backward(self):
        check(lib.st_adjoint_backward(self.handle))

    def gradients(self, initial_state=False, out=None):
        """Per-parcel gradients in upload order.  `out` may hold preallocated arrays
        ("d", "rho", "x0", "v0"; float64, C-contiguous, of shapes (n,) and (3, n)), for
        instance page-locked ones from `pinned_empty`, which the download fills."""
        n = self.chunk.N
        out = {} if out is None else out

        def dst(key, shape):
            a = out.get(key)
            if a is None:
                return np.empty(shape)
            if a.shape != shape or a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError(f"out[{key!r}] must be a C-contiguous float64 array of shape {shape}")
            return a

        gd = dst("d", (n,))
        grho = dst("rho", (n,)) if self.chunk.rho_p is not None else None
        gcd = C.c_double()
        gx = dst("x0", (3, n)) if initial_state else None
        gv = dst("v0", (3, n)) if initial_state else None
        check(lib.st_adjoint_gradients(self.handle, _ptr(gd), _ptr(grho), C.byref(gcd), _ptr(gx), _ptr(gv)))
        out = {"d": gd, "rho": grho, "cd": gcd.value}
        if initial_state:
            out["x0"], out["v0"] = gx, gv
        return out

    def last_ms(self):
        return self.ctx.last_kernel_ms()

    def close(self):
        if self.handle:
            check(lib.st_adjoint_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def source_misfit_rrule(ctx, chunk, field, target, dt_fluid, nsub, checkpoint_every, p=None):
    """rrule of L(d, rho, cd) = sum over cells (UTrans - target)^2 after nsub substeps.

    Returns (L, pullback) with pullback(Lbar) -> (NoTangent, d_bar, rho_bar, cd_bar);
    the chunk is advanced by the forward run.  The first pullback runs the reverse sweep
    and releases the adjoint's device buffers (and its hold on the chunk and field);
    the gradients are linear in Lbar, so later calls scale the kept ones."""
    adj = Adjoint(ctx, chunk, field, nsub, checkpoint_every, p)
    try:
        adj.forward(dt_fluid)
        loss = adj.source_misfit(target)
    except Exception:
        adj.close()
        raise
    kept = {}

    def pullback(lbar=1.0):
        if not kept:
            try:
                adj.backward()
                kept.update(adj.gradients())
            finally:
                adj.close()
        rho = None if kept["rho"] is None else lbar * kept["rho"]
        return None, lbar * kept["d"], rho, lbar * kept["cd"]

    return loss, pullback
