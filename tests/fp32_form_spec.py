"""A single-parcel statement of the FP32 precision form, for tiny hand-built inputs.

It follows evolve_particle! (run/icoJuliaParcelFoam/cavity3D/asyncParallelTracking.jl:
1271-1365, wall boundary, one partition) operation by operation, in numpy float32 for
the parcel state and drag and Python floats (IEEE double) for the rest, and it can take
each of the form's three double-precision rules the other way:

* face    -- the cell test floor((x - origin) * rDelta): "double" (the form) evaluates it
             on the float coordinate widened to double; "float" in float arithmetic;
* s0      -- the residence's starting velocity and its bounce corrections s0 -= 2 v:
             "double" (the form) or "float";
* source  -- each deposit m (s0 - v') and the per-cell sums: "double" (the form) or
             "float".

The tests use it to show that an input separates the form from each alternative (the
alternatives give other results) and that the oracle's FP32 form gives the form's.
Products and quotients of two float32 values are computed by numpy as correctly rounded
float32 operations (a double evaluation rounded to float32 is exact for +, -, *, /).
"""
import math

import numpy as np

F = np.float32
PI = 3.14159265358979323846


def _floor_index(p, o, rd, face):
    if face == "double":
        return math.floor((float(p) - o) * rd)
    return math.floor(float(F(F(p) - F(o)) * F(rd)))


def _trunc_divrem(i, n):
    q = int(i / n)  # truncation toward zero, as Julia's divrem and C's / and %
    return q, i - q * n


class Mesh:
    def __init__(self, n, origin, ending):
        self.n = tuple(int(v) for v in n)
        self.o = tuple(float(v) for v in origin)
        e = tuple(float(v) for v in ending)
        self.rd = tuple(1.0 / ((e[a] - self.o[a]) / self.n[a]) for a in range(3))
        self.nxny = self.n[0] * self.n[1]

    def locate(self, pos, face="double"):
        """(I per axis, folded cell, folded partition) as locate (:663-697)."""
        I = [_floor_index(pos[a], self.o[a], self.rd[a], face) for a in range(3)]
        f, part = [], 0
        for a in range(3):
            q, r = _trunc_divrem(I[a], self.n[a])
            f.append(abs(r))
            part += abs(q)
        return I, f[2] * self.nxny + f[1] * self.n[0] + f[0], part


def step(mesh, U, d, x, v, dt_fluid, nsub, mu, rho, rdr, face="double", s0="double", source="double", into=None):
    """One parcel, one call of nsub substeps.  U: (ncell, 3) float32.  Returns the final
    position and velocity (float32 triples) and {cell: [3 sums]}: the parcel's own
    deposits, or `into` (a dict of earlier parcels' sums) with them added one by one."""
    d = F(d)
    x = [F(a) for a in x]
    v = [F(a) for a in v]
    dt = F(F(dt_fluid) / F(nsub))
    dd = float(d)
    m = ((rdr * PI) * ((dd * dd) * dd)) / 6.0
    f = F(F(F(18) * F(mu)) * dt) / F(F(rho) * F(d * d))
    f = F(F(1) * f)
    den = F(F(1) + f)
    wide = float if s0 == "double" else F
    I, c, part0 = mesh.locate(x, face)
    u = [F(U[c][a]) for a in range(3)]
    s = [wide(a) for a in v]
    src = {} if into is None else into

    def deposit(cell, vnew):
        for a in range(3):
            if source == "double":
                dep = m * (float(s[a]) - float(vnew[a]))
                src.setdefault(cell, [0.0, 0.0, 0.0])[a] += dep
            else:
                dep = F(F(m) * F(F(s[a]) - F(vnew[a])))
                cur = src.setdefault(cell, [F(0), F(0), F(0)])
                cur[a] = F(F(cur[a]) + dep)

    for t in range(1, nsub + 1):
        vn = [F(F(v[a] + F(f * u[a])) / den) for a in range(3)]
        xn = [F(x[a] + F(vn[a] * dt)) for a in range(3)]
        In, cn, pn = mesh.locate(xn, face)
        if cn != c or pn != part0:
            hit = False
            for a in range(3):
                if In[a] < 0 or In[a] >= mesh.n[a]:
                    xn[a] = x[a]
                    vn[a] = F(-v[a])
                    s[a] = wide(s[a] - wide(2.0) * wide(v[a]))
                    hit = True
            if hit:
                In, cn, pn = mesh.locate(xn, face)
        if t != nsub and cn != c:
            deposit(c, vn)
            s = [wide(a) for a in vn]
            u = [F(U[cn][a]) for a in range(3)]
            c = cn
        v, x = vn, xn
    deposit(c, v)
    return x, v, src
