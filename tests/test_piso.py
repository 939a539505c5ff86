"""The icoFoam-style fluid step (the caller of the parcel step) pinned by closed forms:
Couette flow (linear profile, exact for the discrete operators), plane Poiseuille flow
driven by the momentum source (the discrete steady state solved independently, and the
parabola to second order), a divergence-free cavity, and mass conservation through an
inlet and an outlet."""
import numpy as np
import pytest

from scaletrack.piso import IcoFoam, Patch


def channel_patches(top=(0.0, 0.0, 0.0)):
    return {(0, 0): Patch("periodic"), (0, 1): Patch("periodic"),
            (1, 0): Patch("wall"), (1, 1): Patch("wall", top),
            (2, 0): Patch("empty"), (2, 1): Patch("empty")}


def profile(f, ny):
    return f.U.reshape(1, ny, -1, 3)[0, :, 0, 0]


def test_couette_linear_profile():
    ny, H, nu = 16, 1.0, 0.1
    f = IcoFoam((4, ny, 1), (0, 0, 0), (0.5, H, 0.1), nu, channel_patches(top=(1.0, 0.0, 0.0)))
    for _ in range(1200):
        f.step(0.5)
    y = (np.arange(ny) + 0.5) / ny * H
    np.testing.assert_allclose(profile(f, ny), y / H, rtol=0, atol=1e-12)
    assert np.abs(f.U[:, 1:]).max() < 1e-12


def discrete_poiseuille(ny, H, nu, g):
    """The steady discrete solution: nu (u[j+1] - 2u[j] + u[j-1]) / d^2 + g = 0 inside,
    with the wall face value 0 at distance d / 2 (the Gauss linear wall stencil)."""
    d = H / ny
    M = np.zeros((ny, ny))
    for j in range(ny):
        M[j, j] = -2.0
        if j > 0:
            M[j, j - 1] = 1.0
        else:
            M[j, j] -= 1.0  # (u1 - u0)/d - (u0 - 0)/(d/2) = u1 - 3 u0 over d
        if j < ny - 1:
            M[j, j + 1] = 1.0
        else:
            M[j, j] -= 1.0
    return np.linalg.solve(nu * M / d ** 2, -g * np.ones(ny))


@pytest.mark.parametrize("ny", [8, 16])
def test_poiseuille_driven_by_the_source(ny):
    """A body force g enters through the momentum source as UTrans = g dt per step
    (icoJuliaParcelFoam.C:135 adds UTrans / dt): the steady state is the discrete
    Poiseuille profile, within second order of u = g y (H - y) / (2 nu)."""
    H, nu, g, dt = 1.0, 0.1, 0.2, 0.5
    f = IcoFoam((4, ny, 1), (0, 0, 0), (0.5, H, 0.1), nu, channel_patches())
    S = np.zeros((f.ncell, 3))
    S[:, 0] = g * dt
    for _ in range(1500):
        f.step(dt, S)
    u = profile(f, ny)
    np.testing.assert_allclose(u, discrete_poiseuille(ny, H, nu, g), rtol=1e-9)
    y = (np.arange(ny) + 0.5) / ny * H
    exact = g * y * (H - y) / (2 * nu)
    err = np.abs(u - exact).max() / exact.max()
    assert err < 3.0 / ny ** 2, err


def test_poiseuille_second_order():
    H, nu, g = 1.0, 0.1, 0.2
    errs = []
    for ny in (8, 16, 32):
        u = discrete_poiseuille(ny, H, nu, g)
        y = (np.arange(ny) + 0.5) / ny * H
        errs.append(np.abs(u - g * y * (H - y) / (2 * nu)).max())
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5


def test_cavity_stays_divergence_free():
    """The config-1 cavity (0.1 m, 20 x 20 x 1, lid 1 m/s, nu = 0.01, dt = 0.005): the
    face fluxes stay discretely divergence-free and the flow turns in the lid's sense."""
    patches = {(0, 0): Patch("wall"), (0, 1): Patch("wall"), (1, 0): Patch("wall"),
               (1, 1): Patch("wall", (1.0, 0.0, 0.0)), (2, 0): Patch("empty"), (2, 1): Patch("empty")}
    U = np.zeros((400, 3))
    f = IcoFoam((20, 20, 1), (0, 0, 0), (0.1, 0.1, 0.01), 0.01, patches, U=U)
    for _ in range(100):
        f.step(0.005)
    assert f.U is U  # written in place: the caller's storage
    div = np.abs(f.divergence()).max()
    assert div < 1e-12 * 1.0 * f.A[0], div
    assert max(f.continuity) < 1e-12
    g = U.reshape(20, 20, 3)
    assert g[19, 10, 0] > 0.2 and g[2, 10, 0] < 0  # along the lid at the top, back at the bottom
    assert np.abs(U[:, 2]).max() == 0.0


def test_inlet_outlet_mass_conservation():
    n = (16, 8, 1)
    patches = {(0, 0): Patch("inlet", (1.0, 0.0, 0.0)), (0, 1): Patch("outlet"),
               (1, 0): Patch("wall"), (1, 1): Patch("wall"), (2, 0): Patch("empty"), (2, 1): Patch("empty")}
    f = IcoFoam(n, (0, 0, 0), (2.0, 1.0, 0.1), 0.05, patches)
    for _ in range(200):
        f.step(0.05)
    inflow = f.phi[0][1].sum()
    outflow = f.phi[0][2].sum()
    assert abs(outflow - inflow) < 1e-10 * abs(inflow)
    assert abs(inflow - 1.0 * 1.0 * 0.1) < 1e-12
    assert np.abs(f.divergence()).max() < 1e-10
    p = f.p.reshape(1, 8, 16)[0]
    assert np.all(np.diff(p.mean(0)) < 0)  # pressure falls along the channel
