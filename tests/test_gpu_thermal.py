"""GPU parity of the thermal branch (ENERGY: lbm.cpp:121-178, 263-289,
636-653, 777-806, 964-1028) against the CPU oracle, through the C-ABI.

Bars: FP64 storage -- u within relative L2 1e-11, rho and T within 1e-12
absolute (T is O(1) in these cases), g within 1e-12; FP32 storage -- the
bounds stated in the test.  The device computes the strain rate and the
equilibria with another operation order than the reference's, hence
roundoff-level rather than bit-exact bars.
"""
import numpy as np
import pytest

import cases
import lbm_amd
from conftest import rel_l2_u

pytestmark = pytest.mark.gpu


def _gpu(case, th, storage="f64"):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, nu=th.nu, cp=th.cp, j_dot_n=th.j_dot_n,
                       boussinesq=th.boussinesq, g_beta=th.g_beta, T_ref=th.T_ref)
    lat.init()
    return lat


def _oracle(oracle_mod, case, th):
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.thermal_enable(th.tau_g, th.nu, th.cp, th.j_dot_n, th.boussinesq, th.g_beta, th.T_ref)
    o.set_cells(**case.fields())
    o.set_thermal(th.T, th.type_energy)
    o.init()
    return o


def _cmp(lat, o, case, tol_u, tol_rho, tol_T):
    rho, ux, uy, uz = lat.macro()
    r_rho, r_ux, r_uy, r_uz = o.macro()
    T = lat.thermal()[0]
    r_T = o.thermal()[0]
    fl = case.type == cases.TYPE_F
    eu = rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), fl)
    er = float(np.abs(rho - r_rho)[fl].max())
    eT = float(np.abs(T - r_T)[fl].max())
    assert eu <= tol_u, f"rel L2(u) = {eu:.3e} > {tol_u}"
    assert er <= tol_rho, f"max|drho| = {er:.3e} > {tol_rho}"
    assert eT <= tol_T, f"max|dT| = {eT:.3e} > {tol_T}"
    return eu, er, eT


CASES = {
    "couette": lambda: cases.couette_thermal(Ny=12, Re=10.0),
    "neumann": lambda: cases.conduction_channel(bottom=cases.TYPE_Q, top=cases.TYPE_T),
    "rayleigh_benard": lambda: cases.rayleigh_benard(),
    "random": lambda: cases.random_thermal_case(6, 5, 4, seed=5),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_thermal_fused_fp64(oracle_mod, name):
    case, th = CASES[name]()
    steps = 40
    o = _oracle(oracle_mod, case, th)
    assert o.run(steps) == 0
    lat = _gpu(case, th)
    lat.step(steps)
    assert lat.sync() == 0
    _cmp(lat, o, case, 1e-11, 1e-12, 1e-12)
    fl = case.type == cases.TYPE_F
    assert np.abs(lat.g_populations() - o.g())[fl].max() < 1e-12
    lat.close()


@pytest.mark.parametrize("name", ["couette", "rayleigh_benard", "random"])
def test_thermal_reference_calls_fp64(oracle_mod, name):
    """Collision -> Streaming -> MacroProp as separate calls (the drop-in
    path), against the oracle's same calls."""
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


def test_thermal_couette_fp32(oracle_mod):
    """FP32 storage (g - w T0, FP32 arithmetic) against the FP64 oracle: the
    temperature rise of the heated Couette flow (about 1e-4 here) to 1e-3 of
    itself."""
    case, th = CASES["couette"]()
    steps = 3000
    o = _oracle(oracle_mod, case, th)
    assert o.run(steps) == 0
    lat = _gpu(case, th, storage="f32")
    lat.step(steps)
    assert lat.sync() == 0
    fl = case.type == cases.TYPE_F
    T, r_T = lat.thermal()[0], o.thermal()[0]
    rise = float(np.abs(r_T[fl] - th.params["Tb"]).max())
    err = float(np.abs(T - r_T)[fl].max())
    assert err <= 1e-3 * rise, f"{err:.3e} vs rise {rise:.3e}"
    _, ux, uy, uz = lat.macro()
    _, r_ux, r_uy, r_uz = o.macro()
    assert rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), fl) < 1e-5
    lat.close()


def test_thermal_temperature_range_error():
    """T > 100 after MacroProp: LBM_ERR_TEMPERATURE_RANGE (lbm.cpp:794-797)."""
    case, th = cases.conduction_channel(Tb=150.0, Tt=150.0, T_fluid=150.0)
    lat = _gpu(case, th)
    lat.step(2)
    assert lat.sync() == lbm_amd.LBM_ERR_TEMPERATURE_RANGE
    lat.close()


def test_thermal_rejects_fp16_and_unset_walls():
    case, th = CASES["couette"]()
    for st in ("f16s", "f16c"):
        lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=st)
        lat.upload(**case.fields())
        with pytest.raises(lbm_amd.LBMError) as e:
            lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, nu=th.nu, cp=th.cp)
        assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
        lat.close()
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    lat.upload(**case.fields())
    with pytest.raises(lbm_amd.LBMError) as e:  # walls without type_Energy
        lat.upload_thermal(th.T, None, tau_g=th.tau_g, nu=th.nu, cp=th.cp)
    assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
    lat.close()


def test_thermal_rejects_inflow_write_into_fluid():
    """The k = 0 inlet's back-face write sets g of cell (i,j,Nz+1) too
    (lbm.cpp:611-617); the ENERGY kernels do not apply it, so a set-up where
    that cell is fluid or solid is rejected at lbm_init."""
    case, th = cases.couette_thermal()
    case.type[1, 3, 0] = cases.TYPE_IN  # fluid above it at k = Nz+1 (periodic z)
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, nu=th.nu, cp=th.cp)
    with pytest.raises(lbm_amd.LBMError) as e:
        lat.init()
    assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
    lat.close()
