"""GPU D3Q7 temperature (LBM_THERMAL_D3Q7) against the CPU oracle's
restatement of the same scheme, through the C-ABI.

The reference has no D3Q7; the oracle's version (oracle/lbm_oracle.h) is
pinned by closed forms in tests/test_oracle_pins.py.  Bars: FP64 storage --
u within relative L2 1e-11, rho, T and g within 1e-12 absolute (T is O(1));
FP32 and 16-bit storage -- the bounds stated in each test.
"""
import math

import numpy as np
import pytest

import cases
import lbm_amd
from conftest import rel_l2_u

pytestmark = pytest.mark.gpu


def _gpu(case, th, storage="f64", T0=float("nan")):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, boussinesq=th.boussinesq, g_beta=th.g_beta,
                       T_ref=th.T_ref, T0=T0, j_dot_n=th.j_dot_n, model="d3q7")
    lat.init()
    return lat


def _oracle(oracle_mod, case, th):
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.q7_enable(th.tau_g, th.boussinesq, th.g_beta, th.T_ref, j_dot_n=th.j_dot_n)
    o.set_cells(**case.fields())
    o.set_q7(th.T, th.type_energy)
    o.init()
    return o


def _cmp(lat, o, case, tol_u, tol_rho, tol_T):
    rho, ux, uy, uz = lat.macro()
    r_rho, r_ux, r_uy, r_uz = o.macro()
    T, qx, qy, qz = lat.thermal()
    r_T, r_qx, r_qy, r_qz = o.q7_thermal()
    fl = case.type == cases.TYPE_F
    eu = rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), fl)
    er = float(np.abs(rho - r_rho)[fl].max())
    eT = float(np.abs(T - r_T)[fl].max())
    eq = max(float(np.abs(a - b)[fl].max()) for a, b in ((qx, r_qx), (qy, r_qy), (qz, r_qz)))
    assert eu <= tol_u, f"rel L2(u) = {eu:.3e} > {tol_u}"
    assert er <= tol_rho, f"max|drho| = {er:.3e} > {tol_rho}"
    assert eT <= tol_T, f"max|dT| = {eT:.3e} > {tol_T}"
    assert eq <= tol_T, f"max|dq| = {eq:.3e} > {tol_T}"
    return eu, er, eT


CASES = {
    "conduction": lambda: cases.q7_conduction(Nx=3, Ny=8, Nz=2, tau_g=0.9),
    "conduction_adiabatic_top": lambda: cases.q7_conduction(Nx=3, Ny=8, Nz=2, tau_g=1.2, top=cases.TYPE_Q),
    "sine_advection": lambda: cases.q7_sine(Np=16, U=0.04, tau_g=0.8),
    "random": lambda: cases.q7_random_case(7, 6, 5, seed=9),
    "random_no_buoyancy": lambda: cases.q7_random_case(6, 5, 7, seed=13, boussinesq=False),
    "conduction_flux_bottom": lambda: cases.q7_conduction(Nx=3, Ny=8, Nz=2, tau_g=1.1, bottom=cases.TYPE_Q,
                                                          j_dot_n=2e-3),
    "random_flux_walls": lambda: cases.q7_random_case(7, 6, 5, seed=17, j_dot_n=-5e-4),
    "rayleigh_benard": lambda: cases.q7_rayleigh_benard(Nx=16, Ny=8, Nz=1, Ra=5e3),
    # relative temperatures about T_ref = 0: negative T is no range error
    "conduction_relative": lambda: cases.q7_conduction(Nx=3, Ny=8, Nz=2, tau_g=0.9, Tb=0.5, Tt=-0.5),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_q7_fused_fp64(oracle_mod, name):
    case, th = CASES[name]()
    steps = 40
    o = _oracle(oracle_mod, case, th)
    assert o.run(steps) == 0
    lat = _gpu(case, th)
    lat.step(steps)
    assert lat.sync() == 0
    _cmp(lat, o, case, 1e-11, 1e-12, 1e-12)
    fl = case.type == cases.TYPE_F
    g = lat.g_populations()
    assert g.shape == case.shape + (7,)
    assert np.abs(g - o.q7_g())[fl].max() < 1e-12
    lat.close()


@pytest.mark.parametrize("name", ["conduction", "random", "rayleigh_benard", "random_flux_walls"])
def test_q7_reference_calls_fp64(oracle_mod, name):
    """Collision -> Streaming -> MacroProp as separate calls, against the
    oracle's same calls."""
    case, th = CASES[name]()
    o = _oracle(oracle_mod, case, th)
    lat = _gpu(case, th)
    for _ in range(12):
        assert o.collision() == 0
        o.streaming()
        assert o.macroprop() == 0
        lat.collision()
        lat.streaming()
        lat.macroprop()
    assert lat.sync() == 0
    _cmp(lat, o, case, 1e-11, 1e-12, 1e-12)
    lat.close()


def test_q7_sine_mode_exact_factor_on_device():
    """The closed form of tests/test_oracle_pins.py on the device: tau_g = 1,
    uniform flow U, lambda = 3/4 + cos(k)/4 - i U sin(k) per step."""
    U = 0.04
    case, th = cases.q7_sine(Np=32, U=U)
    lat = _gpu(case, th)
    n = 30
    lat.step(n)
    assert lat.sync() == 0
    T = lat.thermal()[0]
    p = case.params
    lam = 0.75 + 0.25 * math.cos(p["k"]) - 1j * U * math.sin(p["k"])
    x = np.arange(case.Nx + 2)
    want = p["T0"] + p["A"] * np.imag(lam ** n * np.exp(1j * p["k"] * x))
    assert np.abs(T[:, 1, 1] - want).max() < 1e-13
    lat.close()


def test_q7_conduction_fp32_steady_profile():
    """FP32 storage (g - w T0): the steady Dirichlet profile to 1e-5 of the
    wall temperature difference."""
    case, th = cases.q7_conduction(Nx=3, Ny=12, Nz=2, tau_g=1.0, Tb=1.0, Tt=0.25)
    lat = _gpu(case, th, storage="f32")
    lat.step(8000)
    assert lat.sync() == 0
    T = lat.thermal()[0]
    j = np.arange(1, case.Ny + 1)
    want = 1.0 + (0.25 - 1.0) * (j - 0.5) / case.Ny
    assert np.abs(T[:, 1:-1, :] - want[None, :, None]).max() < 1e-5 * 0.75
    lat.close()


def test_q7_random_fp32(oracle_mod):
    """FP32 storage and arithmetic against the FP64 oracle after 200 steps."""
    case, th = CASES["random"]()
    o = _oracle(oracle_mod, case, th)
    assert o.run(200) == 0
    lat = _gpu(case, th, storage="f32")
    lat.step(200)
    assert lat.sync() == 0
    fl = case.type == cases.TYPE_F
    _, ux, uy, uz = lat.macro()
    _, r_ux, r_uy, r_uz = o.macro()
    assert rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), fl) < 1e-5
    T, r_T = lat.thermal()[0], o.q7_thermal()[0]
    assert np.abs(T - r_T)[fl].max() < 1e-5
    lat.close()


def test_q7_rayleigh_benard_convects():
    """2D Rayleigh-Benard at Ra = 1e4, Pr = 0.71, H = 24, rolls of wavelength
    2H (x periodic over 48 cells): steady convection with a Nusselt number
    of 2.63 at this resolution (conduction alone gives 1; Ra = 2e4 gives
    3.2 on the oracle).  Nu = 1 + <u_y T>/(alpha dT/H) over the fluid."""
    case, th = cases.q7_rayleigh_benard(Nx=48, Ny=24, Nz=1, Ra=1e4, Pr=0.71, tau=0.6)
    lat = _gpu(case, th)
    lat.step(30000)
    assert lat.sync() == 0
    T = lat.thermal()[0]
    _, ux, uy, uz = lat.macro()
    fl = case.type == cases.TYPE_F
    p = th.params
    Nu = 1 + (uy * T)[fl].mean() / (p["alpha"] / p["H"])
    assert 2.4 < Nu < 2.8, Nu
    lat.close()


def test_q7_temperature_range_error():
    """The D3Q7 guard flags |T| > 100 (blow-up), of either sign."""
    for T in (150.0, -150.0):
        case, th = cases.q7_conduction(Tb=T, Tt=T, T_fluid=T)
        lat = _gpu(case, th)
        lat.step(2)
        assert lat.sync() == lbm_amd.LBM_ERR_TEMPERATURE_RANGE
        lat.close()


@pytest.mark.parametrize("storage,tol_u,tol_T", [("f16s", 1e-3, 2e-4), ("f16c", 5e-4, 1e-4)])
def test_q7_random_fp16(oracle_mod, storage, tol_u, tol_T):
    """16-bit storage (g - w T0 in the units of f) against the FP64 oracle
    after 200 steps: the storages round each population to 11 (F16S) or 12
    (F16C) significant bits."""
    case, th = CASES["random"]()
    o = _oracle(oracle_mod, case, th)
    assert o.run(200) == 0
    lat = _gpu(case, th, storage=storage)
    lat.step(200)
    assert lat.sync() == 0
    fl = case.type == cases.TYPE_F
    _, ux, uy, uz = lat.macro()
    _, r_ux, r_uy, r_uz = o.macro()
    eu = rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), fl)
    T, r_T = lat.thermal()[0], o.q7_thermal()[0]
    eT = float(np.abs(T - r_T)[fl].max())
    print(f"d3q7 random {storage}: rel L2(u) {eu:.3e}, max|dT| {eT:.3e}")
    assert eu < tol_u
    assert eT < tol_T
    lat.close()


def test_q7_f16c_flags_temperature_beyond_its_range():
    """F16C holds |g - w T0| < 0.5, |T - T0| < 2 at the rest weight 1/4: a
    cell past 1.9 flags LBM_ERR_TEMPERATURE_RANGE rather than let its g
    saturate.  The fluid at 0 (T0 = 0) heated by a wall at 5 passes it
    within a few steps; F16S (|g - w T0| < 16) runs the case."""
    case, th = cases.q7_conduction(Nx=3, Ny=8, Nz=2, tau_g=0.9, Tb=5.0, Tt=0.0, T_fluid=0.0)
    for st, want in (("f16c", lbm_amd.LBM_ERR_TEMPERATURE_RANGE), ("f16s", 0), ("f32", 0)):
        lat = _gpu(case, th, storage=st)
        lat.step(50)
        assert lat.sync() == want, st
        lat.close()


@pytest.mark.parametrize("storage,tol", [("f16s", 1e-2), ("f16c", 5e-3)])
def test_q7_conduction_fp16_steady_profile(storage, tol):
    """16-bit storage: the steady Dirichlet profile between walls at 1 and
    0.25 to `tol` of the wall temperature difference.  At a steady state
    every step rounds the same populations the same way, so the rounding
    acts as a fixed source over the diffusion time Ny^2/alpha (about 1200
    steps here): the error is that many roundings of the stored g, not
    one."""
    case, th = cases.q7_conduction(Nx=3, Ny=12, Nz=2, tau_g=1.0, Tb=1.0, Tt=0.25)
    lat = _gpu(case, th, storage=storage)
    lat.step(8000)
    assert lat.sync() == 0
    T = lat.thermal()[0]
    j = np.arange(1, case.Ny + 1)
    want = 1.0 + (0.25 - 1.0) * (j - 0.5) / case.Ny
    err = float(np.abs(T[:, 1:-1, :] - want[None, :, None]).max())
    print(f"d3q7 conduction {storage}: max|dT| {err:.3e}")
    assert err < tol * 0.75
    lat.close()


def test_q7_flux_wall_fp32_steady_profile():
    """FP32 storage: a TYPE_Q wall letting q into the fluid against a
    Dirichlet wall at Tt gives T_j = Tt + (q/alpha)(Ny + 1/2 - j), alpha =
    (tau_g - 1/2)/4, to 1e-5 of the profile's span."""
    q, tau_g = 2e-3, 1.0
    case, th = cases.q7_conduction(Nx=3, Ny=12, Nz=2, tau_g=tau_g, Tb=0.5, Tt=0.25, bottom=cases.TYPE_Q,
                                   j_dot_n=q)
    lat = _gpu(case, th, storage="f32")
    lat.step(12000)
    assert lat.sync() == 0
    T = lat.thermal()[0]
    alpha = (tau_g - 0.5) / 4.0
    j = np.arange(1, case.Ny + 1)
    want = 0.25 + q / alpha * (case.Ny + 0.5 - j)
    span = q / alpha * case.Ny
    assert np.abs(T[:, 1:-1, :] - want[None, :, None]).max() < 1e-5 * span
    lat.close()


@pytest.mark.timeout(300)
@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 1e-9), ("f16s", 5e-6), ("f16c", 1e-6)])
def test_q7_tgv256_heat_and_mass_conserved(storage, tol):
    """At the benchmark's full size (Taylor-Green 256^3 carrying T = 0.5 +
    0.1 sin sin sin, Boussinesq on, periodic): streaming moves every g and
    f to another cell and the collisions keep each cell's T and rho, so
    sum T and sum rho over the lattice are invariants of the exact scheme;
    after 200 steps they hold to `tol` relative (the storage's rounding),
    and the T field stays within its initial range."""
    n = 256
    case = cases.taylor_green(n, u0=0.04, Re=1600.0)
    k = 2 * np.pi / n
    x = np.arange(n, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    T = 0.5 + 0.1 * np.sin(k * X) * np.sin(k * Y) * np.sin(k * Z)
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.upload_thermal(T, None, tau_g=0.6, boussinesq=True, g_beta=1e-5, T_ref=0.5, model="d3q7")
    lat.init()
    lat.step(1)
    assert lat.sync() == 0
    rho0 = float(lat.macro()[0].sum(dtype=np.float64))
    T0 = float(lat.thermal()[0].sum(dtype=np.float64))
    lat.step(200)
    assert lat.sync() == 0
    rho1 = float(lat.macro()[0].sum(dtype=np.float64))
    Tn = lat.thermal()[0]
    T1 = float(Tn.sum(dtype=np.float64))
    print(f"d3q7 256^3 {storage}: sum T drift {abs(T1 / T0 - 1):.3e}, sum rho drift {abs(rho1 / rho0 - 1):.3e}")
    assert abs(T1 / T0 - 1) < tol
    assert abs(rho1 / rho0 - 1) < tol
    assert Tn.min() > 0.4 - 1e-3 and Tn.max() < 0.6 + 1e-3
    lat.close()

