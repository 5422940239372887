"""Physics-level checks of the GPU step against closed forms, independent of
the CPU oracle: the reference's own validation cases (Couette velocity and
stress, lbm.cpp:3195-3251; Poiseuille, lbm.cpp:3355-3414) with the
geometry of half-way bounce-back (walls half a cell outside the first and
last fluid rows, H = Ny), and the viscous decay of the Taylor-Green vortex.
Relative L2 is the reference's error form (lbm.cpp:3219-3231).
"""
import math

import numpy as np
import pytest

import cases
import lbm_amd

pytestmark = pytest.mark.gpu


def _run(case, steps, storage):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.init()
    lat.step(steps)
    assert lat.sync() == 0
    return lat


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_couette_transient_series(storage):
    """Impulsively started Couette flow at t = 500 (about a third of the
    way to steady): u(y, t) = U y/H - sum_n 2U/(n pi) (-1)^(n+1)
    sin(n pi y/H) exp(-nu (n pi/H)^2 t), y = j - 1/2, to 1e-3 relative L2
    (the lattice's own discretisation error is 5e-4 here)."""
    U, tau, Ny, t = 0.01, 0.8, 20, 500
    lat = _run(cases.couette(Nx=2, Ny=Ny, Nz=2, tau=tau, u_wall=U), t, storage)
    ux = lat.macro()[1][1, 1:-1, 1]
    lat.close()
    nu = (tau - 0.5) / 3
    y = np.arange(1, Ny + 1) - 0.5
    s = sum(2 * U / (n * math.pi) * (-1) ** (n + 1) * np.sin(n * math.pi * y / Ny)
            * math.exp(-nu * (n * math.pi / Ny) ** 2 * t) for n in range(1, 400))
    ua = U * y / Ny - s
    assert math.sqrt(((ux - ua) ** 2).sum() / (ua ** 2).sum()) < 1e-3


def test_couette_steady_shear_stress():
    """Steady Couette flow: the viscous stress of the MORE_DETAILS fields is
    sigma_xy = rho nu U / H in every row, to 1e-8."""
    U, tau, Ny = 0.001, 0.8, 10
    lat = _run(cases.couette(Nx=2, Ny=Ny, Nz=2, tau=tau, u_wall=U), 8000, "f64")
    sig, _ = lat.stress()
    lat.close()
    want = (tau - 0.5) / 3 * U / Ny
    assert np.abs(sig[1, 1:-1, 1, 0, 1] / want - 1).max() < 1e-8


@pytest.mark.parametrize("storage,tol", [("f64", 1e-9), ("f32", 2e-4)])
def test_poiseuille_exact_curvature(storage, tol):
    """Body-force-driven Poiseuille flow at steady state: the discrete
    profile is exactly parabolic, u_(j+1) - 2 u_j + u_(j-1) = -g/nu in every
    interior row, and within 5e-3 of g/(2 nu) y (H - y) (the remainder is
    the bounce-back wall slip at this tau)."""
    g, Ny, tau = 1e-6, 14, 0.8
    lat = _run(cases.poiseuille(2, Ny, 2, tau=tau, g=g), 20000, storage)
    ux = lat.macro()[1][1, 1:-1, 1]
    lat.close()
    nu = (tau - 0.5) / 3
    curv = ux[2:] - 2 * ux[1:-1] + ux[:-2]
    assert np.abs(curv / (-g / nu) - 1).max() < tol
    y = np.arange(1, Ny + 1) - 0.5
    ua = g / (2 * nu) * y * (Ny - y)
    assert np.abs(ux / ua - 1).max() < 5e-3


def test_taylor_green_decay_converges_second_order():
    """Low-Re 3D Taylor-Green vortex (u0 = 0.02, Re = 5): the kinetic energy
    decays as exp(-6 nu k^2 t) (each velocity component is one mode with
    |k|^2 = 3 k^2).  At t = 1/(12 nu k^2) the lattice's deviation from that
    law falls by about 4 from 32^3 to 64^3 (second order) and is below 1e-2
    at 64^3."""
    err = {}
    for N in (32, 64):
        case = cases.taylor_green(N, u0=0.02, Re=5.0)
        nu = (case.tau - 0.5) / 3
        k = 2 * math.pi / N
        T = int(round(0.5 / (6 * nu * k * k)))
        lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f32")
        lat.upload(**case.fields())
        lat.init()
        e0 = lat.reduce(lbm_amd.REDUCE_KINETIC)
        lat.step(T)
        assert lat.sync() == 0
        e1 = lat.reduce(lbm_amd.REDUCE_KINETIC)
        lat.close()
        err[N] = e1 / e0 / math.exp(-6 * nu * k * k * T) - 1
    assert abs(err[64]) < 1e-2
    assert 2.5 < err[32] / err[64] < 5.5, err
