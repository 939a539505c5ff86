"""A minimal icoFoam-style fluid step on uniform hexahedral meshes: the caller of the
parcel step, standing in for OpenFOAM's icoJuliaParcelFoam (OpenFOAM is not available
here).  It follows sol/icoJuliaParcelFoam/icoJuliaParcelFoam.C:118-186 with the case's
settings (system/fvSchemes, system/fvSolution of run/icoJuliaParcelFoam/cavity3D):

  UEqn:  ddt(U) + div(phi, U) - laplacian(nu, U) == UTrans / dt       (:129-136)
         Euler in time, Gauss linear convection and diffusion (orthogonal mesh);
  no momentum predictor (momentumPredictor no), then nCorrectors = 2 PISO correctors:
         rAU = 1 / A,  HbyA = rAU H(U)  (fixed-value boundaries take their value),
         phiHbyA = flux(HbyA),
         laplacian(rAU_f, p) == div(phiHbyA)  (pRefCell 0 = 0 when no boundary fixes p),
         phi = phiHbyA - pEqn.flux(),  U = HbyA - rAU grad(p)   (Gauss linear).
  The time-derivative flux correction fvc::ddtCorr of :151 is left out.

The source term is taken as the reference writes it: UTrans / dt is added to the
momentum equation per unit volume (it enters the cell's equation multiplied by the cell
volume), so one step's deposit changes the velocity by about UTrans itself.  UTrans as
the tracker deposits it is sum (m / rho_c)(v_in - v_out) per cell, a volume times a
velocity (asyncParallelTracking.jl:710-728): `per_volume=True` divides it by the cell
volume first, which makes the coupling dimensionally consistent.

Boundaries, per (axis, side): "wall" (fixed velocity `value`, zero pressure gradient),
"inlet" (fixed velocity, per-face array or vector, zero pressure gradient), "outlet"
(zero velocity gradient, fixed pressure 0), "empty" (2D: no flux, no diffusion), or
"periodic" (both sides of the axis).

U is AoS (ncell, 3), cells x fastest, as the tracker's fields.  `U` may be an array the
solver does not own (the host field the tracker hands out through allocate_array_j):
the step writes the new velocity into it in place, as OpenFOAM writes into the memory
Julia lent it (juliaWrapper.H:119-172, icoJuliaParcelFoam.C:182-184).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

AXES = (0, 1, 2)


class Patch:
    def __init__(self, kind, value=None):
        assert kind in ("wall", "inlet", "outlet", "empty", "periodic")
        self.kind = kind
        self.value = value


class IcoFoam:
    """Incompressible laminar fluid step (icoFoam's PISO) on n[0] x n[1] x n[2] cells of
    the box [origin, ending]."""

    def __init__(self, n, origin, ending, nu, patches, U=None, n_correctors=2, p_tol=1e-10):
        self.n = tuple(int(v) for v in n)
        self.o = np.asarray(origin, float)
        self.e = np.asarray(ending, float)
        self.d = (self.e - self.o) / np.asarray(self.n)
        self.V = float(np.prod(self.d))
        self.A = (self.d[1] * self.d[2], self.d[0] * self.d[2], self.d[0] * self.d[1])
        self.nu = float(nu)
        self.ncell = self.n[0] * self.n[1] * self.n[2]
        self.patches = {}
        for a in AXES:
            for s in (0, 1):
                pt = patches.get((a, s), Patch("wall"))
                self.patches[(a, s)] = pt if isinstance(pt, Patch) else Patch(*pt)
            if (self.patches[(a, 0)].kind == "periodic") != (self.patches[(a, 1)].kind == "periodic"):
                raise ValueError("a periodic axis needs periodic patches on both sides")
        self.U = np.zeros((self.ncell, 3)) if U is None else U
        self.p = np.zeros(self.ncell)
        self.n_correctors = n_correctors
        self.p_tol = p_tol
        self.fixes_p = any(pt.kind == "outlet" for pt in self.patches.values())
        self.phi = self._flux(self._grid(self.U), boundary_values=True)
        self.continuity = []

    # ---- index helpers -----------------------------------------------------
    def _grid(self, U):
        nx, ny, nz = self.n
        return U.reshape(nz, ny, nx, -1) if U.ndim == 2 else U.reshape(nz, ny, nx)

    @staticmethod
    def _ax(a):
        return 2 - a  # array axis of mesh axis a in (z, y, x) order

    def _bvalue(self, a, s, comp_shape):
        """Fixed boundary velocity of patch (a, s) on its faces, shape of a face slab x 3."""
        v = self.patches[(a, s)].value
        if v is None:
            return np.zeros(comp_shape)
        v = np.asarray(v, float)
        return np.broadcast_to(v, comp_shape).copy()

    def _face_slab_shape(self, a, ncomp=3):
        shp = [self.n[2], self.n[1], self.n[0]]
        shp[self._ax(a)] = 1
        return tuple(shp) + ((ncomp,) if ncomp else ())

    def _take(self, X, a, idx):
        ax = self._ax(a)
        sl = [slice(None)] * X.ndim
        sl[ax] = slice(idx, idx + 1) if idx >= 0 else slice(self.n[a] - 1, self.n[a])
        return X[tuple(sl)]

    def _neighbours(self, X, a):
        """(lower, upper) cell values across every interior face of axis a, and for a
        periodic axis the wrap face too (as its last face)."""
        ax = self._ax(a)
        n = self.n[a]
        lo = np.take(X, np.arange(0, n - 1), axis=ax)
        hi = np.take(X, np.arange(1, n), axis=ax)
        if self.patches[(a, 0)].kind == "periodic":
            lo = np.concatenate([lo, np.take(X, [n - 1], axis=ax)], axis=ax)
            hi = np.concatenate([hi, np.take(X, [0], axis=ax)], axis=ax)
        return lo, hi

    # ---- fluxes --------------------------------------------------------------
    def _flux(self, Ug, boundary_values):
        """Face fluxes of a cell velocity field, Gauss linear interpolation.  Returns per
        axis (interior faces [with the wrap face last], low boundary, high boundary);
        boundary faces carry the patch value (fixed-value patches) or the cell value."""
        out = {}
        for a in AXES:
            lo, hi = self._neighbours(Ug[..., a], a)
            interior = 0.5 * (lo + hi) * self.A[a]
            bnd = []
            for s in (0, 1):
                pt = self.patches[(a, s)]
                cell = self._take(Ug[..., a], a, 0 if s == 0 else -1)
                if pt.kind in ("wall", "inlet") and boundary_values:
                    val = self._bvalue(a, s, self._face_slab_shape(a))[..., a]
                    bnd.append(val * self.A[a])
                elif pt.kind == "outlet":
                    bnd.append(cell * self.A[a])
                elif pt.kind in ("wall", "inlet"):
                    bnd.append(cell * self.A[a])
                else:
                    bnd.append(np.zeros_like(cell))
            out[a] = (interior, bnd[0], bnd[1])
        return out

    def divergence(self, phi=None):
        """Net outward flux of every cell (the discrete continuity error)."""
        phi = self.phi if phi is None else phi
        div = np.zeros((self.n[2], self.n[1], self.n[0]))
        for a in AXES:
            ax = self._ax(a)
            interior, b0, b1 = phi[a]
            n = self.n[a]
            per = self.patches[(a, 0)].kind == "periodic"
            inner = interior[tuple(slice(0, n - 1) if k == ax else slice(None) for k in range(3))]
            up = np.concatenate([inner, b1 if not per else np.take(interior, [n - 1], axis=ax)], axis=ax)
            down = np.concatenate([b0 if not per else np.take(interior, [n - 1], axis=ax), inner], axis=ax)
            # outward: + on the upper face, - on the lower face (b0 is an inflow-positive
            # flux in +x direction, so it enters with -)
            div += up - down
        return div.reshape(-1)

    # ---- momentum matrix -------------------------------------------------------
    def _idx(self):
        return np.arange(self.ncell).reshape(self.n[2], self.n[1], self.n[0])

    def _momentum(self, U_old, S):
        """Diagonal, off-diagonal (sparse) and source of the UEqn (per component source)."""
        idx = self._idx()
        diag = np.full(self.ncell, self.V / self._dt)
        b = (self.V / self._dt) * U_old + self.V * S
        rows, cols, vals = [], [], []
        for a in AXES:
            ax = self._ax(a)
            interior, b0, b1 = self.phi[a]
            P, N = self._neighbours(idx, a)
            P, N, phif = P.reshape(-1), N.reshape(-1), interior.reshape(-1)
            g = self.nu * self.A[a] / self.d[a]
            np.add.at(diag, P, 0.5 * phif + g)
            np.add.at(diag, N, -0.5 * phif + g)
            rows += [P, N]
            cols += [N, P]
            vals += [0.5 * phif - g, -0.5 * phif - g]
            for s, bphi in ((0, b0), (1, b1)):
                pt = self.patches[(a, s)]
                if pt.kind in ("empty", "periodic"):
                    continue
                cells = self._take(idx, a, 0 if s == 0 else -1).reshape(-1)
                sign = -1.0 if s == 0 else 1.0
                out = sign * bphi.reshape(-1)  # outward flux
                if pt.kind in ("wall", "inlet"):
                    Ub = self._bvalue(a, s, self._face_slab_shape(a)).reshape(-1, 3)
                    g2 = 2.0 * self.nu * self.A[a] / self.d[a]
                    np.add.at(diag, cells, g2)
                    np.add.at(b, cells, (g2 - out)[:, None] * Ub)
                else:  # outlet: zero gradient
                    np.add.at(diag, cells, out)
        off = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                            shape=(self.ncell, self.ncell))
        return diag, off, b

    # ---- pressure ------------------------------------------------------------
    def _pressure(self, rAU, phiH):
        """Solve laplacian(rAU_f, p) == div(phiHbyA); returns p and the face fluxes of
        rAU_f grad(p) (the pEqn flux) in the layout of _flux."""
        idx = self._idx()
        rg = self._grid(rAU)
        rows, cols, vals = [], [], []
        diag = np.zeros(self.ncell)
        coef = {}
        for a in AXES:
            P, N = self._neighbours(idx, a)
            rP, rN = self._neighbours(rg, a)
            c = (0.5 * (rP + rN) * self.A[a] / self.d[a]).reshape(-1)
            P, N = P.reshape(-1), N.reshape(-1)
            np.add.at(diag, P, -c)
            np.add.at(diag, N, -c)
            rows += [P, N]
            cols += [N, P]
            vals += [c, c]
            coef[a] = [c, None, None]
            for s in (0, 1):
                if self.patches[(a, s)].kind == "outlet":
                    cells = self._take(idx, a, 0 if s == 0 else -1).reshape(-1)
                    cb = 2.0 * self._take(rg, a, 0 if s == 0 else -1).reshape(-1) * self.A[a] / self.d[a]
                    np.add.at(diag, cells, -cb)
                    coef[a][1 + s] = cb
        rhs = self.divergence(phiH)
        M = sp.csr_matrix((np.concatenate(vals + [diag]),
                           (np.concatenate(rows + [np.arange(self.ncell)]),
                            np.concatenate(cols + [np.arange(self.ncell)]))), shape=(self.ncell, self.ncell))
        if not self.fixes_p:  # pRefCell 0, pRefValue 0
            M = M.tolil()
            M[0, :] = 0.0
            M[0, 0] = 1.0
            M = M.tocsr()
            rhs = rhs.copy()
            rhs[0] = 0.0
        p = spla.spsolve(M.tocsc(), rhs) if self.ncell <= 200000 else \
            spla.cg(-M, -rhs, rtol=self.p_tol, maxiter=20000)[0]
        pg = self._grid(p)
        flux = {}
        for a in AXES:
            pP, pN = self._neighbours(pg, a)
            shp = pP.shape
            interior = coef[a][0].reshape(shp) * (pN - pP)
            bnd = []
            for s in (0, 1):
                cell = self._take(pg, a, 0 if s == 0 else -1)
                if coef[a][1 + s] is not None:
                    # outward gradient (0 - p_P) / (d/2), as a flux along +axis
                    sign = -1.0 if s == 0 else 1.0
                    bnd.append(sign * coef[a][1 + s].reshape(cell.shape) * (0.0 - cell))
                else:
                    bnd.append(np.zeros_like(cell))
            flux[a] = (interior, bnd[0], bnd[1])
        return p, flux

    def _grad(self, p):
        """Gauss linear gradient of p (zero-gradient boundaries, outlets at 0)."""
        pg = self._grid(p)
        gr = np.zeros((self.n[2], self.n[1], self.n[0], 3))
        for a in AXES:
            ax = self._ax(a)
            n = self.n[a]
            per = self.patches[(a, 0)].kind == "periodic"
            pP, pN = self._neighbours(pg, a)
            pf = 0.5 * (pP + pN)
            inner = np.take(pf, np.arange(0, n - 1), axis=ax)
            faces = []
            for s in (0, 1):
                pt = self.patches[(a, s)]
                cell = self._take(pg, a, 0 if s == 0 else -1)
                if per:
                    faces.append(np.take(pf, [n - 1], axis=ax))
                elif pt.kind == "outlet":
                    faces.append(np.zeros_like(cell))
                elif pt.kind == "empty":
                    faces.append(None)
                else:
                    faces.append(cell)
            if faces[0] is None:
                continue
            up = np.concatenate([inner, faces[1]], axis=ax)
            down = np.concatenate([faces[0], inner], axis=ax)
            gr[..., a] = (up - down) * self.A[a] / self.V
        return gr.reshape(-1, 3)

    # ---- the step --------------------------------------------------------------
    def step(self, dt, UTrans=None, per_volume=False):
        """One fluid step of dt with the momentum source UTrans / dt (icoJuliaParcelFoam.C
        :118-186).  Updates self.U in place, self.p and self.phi."""
        self._dt = float(dt)
        U_old = np.array(self.U, dtype=float)
        S = np.zeros((self.ncell, 3)) if UTrans is None else np.asarray(UTrans, float) / self._dt
        if per_volume:
            S = S / self.V
        diag, off, b = self._momentum(U_old, S)
        rAU = self.V / diag  # 1 / A with A = diag / V
        U = U_old
        for _ in range(self.n_correctors):
            H = (b - off @ U) / self.V
            HbyA = rAU[:, None] * H
            phiH = self._flux(self._grid(HbyA), boundary_values=True)
            p, pflux = self._pressure(rAU, phiH)
            self.phi = {a: tuple(x - y for x, y in zip(phiH[a], pflux[a])) for a in AXES}
            U = HbyA - rAU[:, None] * self._grad(p)
        self.p = p
        self.U[...] = U
        self.continuity.append(float(np.abs(self.divergence()).sum() * self._dt / self.V / self.ncell))
        return self.U
