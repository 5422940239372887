"""A flexible filament coupled to the lattice through the IBM (lbm_fsi_*)
against the oracle's same coupled loop, through the C-ABI.

Per step, on both sides: the filament's nodes are the markers (weights ds,
half at the ends), with the desired velocity of their last displacement and
the density of the cell they sit in; the IBM step, Collision, Streaming;
then the filament step under -sum F_b per unit length.  FP64: filament and
fields within roundoff bars (the device sums the IBM terms in the
reference's order, the fused step differs from the oracle's three sweeps by
rounding only).
"""
import math

import numpy as np
import pytest

import cases
import lbm_amd
from conftest import rel_l2_u

pytestmark = pytest.mark.gpu

FIL = dict(n=17, length=16.0, line_density=24.0, bending=1.6e-3)


def _case(U=0.02):
    t, rho, ux, uy, uz = cases._blank(46, 22, 1)
    ux[:] = U
    return cases.Case("fsi_periodic", 46, 22, 1, 0.6, t, rho, ux, uy, uz)


def _oracle_run(oracle_mod, case, fil_args, box, m_max, nsteps, th=None):
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    if th is not None:
        o.q7_enable(th.tau_g, th.boussinesq, th.g_beta, th.T_ref)
    o.set_cells(**case.fields())
    if th is not None:
        o.set_q7(th.T, th.type_energy)
    o.init()
    f = oracle_mod.Filament(**fil_args)
    n = fil_args["n"]
    ds = np.full(n, fil_args["length"] / (n - 1))
    ds[0] = ds[-1] = ds[1] / 2
    out = []
    for _ in range(nsteps):
        x, xp, _ = f.get()
        _, _, fsum = o.fsi_ibm_step(x, x - xp, ds, box, m_max)
        assert o.collision() == 0
        o.streaming()
        f.step(fsum, nsteps=1)
        out.append(np.concatenate([(fsum * ds[:, None]).sum(0), f.get()[0][-1]]))
    return o, f, np.array(out)


@pytest.mark.parametrize("angle", [0.0, 30.0])
def test_fsi_filament_in_uniform_flow_fp64(oracle_mod, angle):
    a = math.radians(angle)
    fil_args = dict(FIL, clamp=(10.0, 8.0, 1.0), tangent=(math.cos(a), math.sin(a), 0.0))
    box, m_max, nsteps = (6, 30, 3, 21, 1, 1), 5, 60
    case = _case()
    o, of, oout = _oracle_run(oracle_mod, case, fil_args, box, m_max, nsteps)
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64")
    lat.upload(**case.fields())
    lat.init()
    f = lbm_amd.Filament(**fil_args)
    lat.fsi_setup(f, box, m_max=m_max)
    out = lat.fsi_run(nsteps)
    assert lat.sync() == 0
    x, xp, T = f.get()
    rx, rxp, rT = of.get()
    assert np.abs(x - rx).max() < 1e-11 and np.abs(xp - rxp).max() < 1e-11
    assert np.abs(T - rT).max() < 1e-10 * max(1.0, np.abs(rT).max())
    scale = np.abs(oout[:, :3]).max()
    assert np.abs(out[:, :3] - oout[:, :3]).max() < 1e-11 * scale
    assert np.abs(out[:, 3:] - oout[:, 3:]).max() < 1e-11
    rho, ux, uy, uz = lat.macro()
    r_rho, r_ux, r_uy, r_uz = o.macro()
    fl = case.type == cases.TYPE_F
    assert rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), fl) < 1e-11
    assert np.abs(rho - r_rho)[fl].max() < 1e-12
    # the flow drags the filament downstream; an inclined one turns toward it
    assert out[-1, 0] > 0
    if angle:
        assert out[-1, 4] < x[0, 1] + 16.0 * math.sin(a)
    lat.close()
    f.close()


def _heated_channel(U=0.02, seed=5):
    """The periodic-x channel between a heated Dirichlet wall (j = 0, T = 1)
    and an adiabatic one (j = Ny+1), fluid at seeded T in [0.2, 0.8), the
    Boussinesq force about T_ref = 0.5."""
    case = _case(U)
    case.type[:, 0, :] = cases.TYPE_S
    case.type[:, -1, :] = cases.TYPE_S
    case.ux[:, 0, :] = case.ux[:, -1, :] = 0.0
    rng = np.random.default_rng(seed)
    T = 0.2 + 0.6 * rng.random(case.shape)
    T[:, 0, :] = 1.0
    te = np.zeros(case.shape, np.int16)
    te[:, 0, :] = cases.TYPE_T
    te[:, -1, :] = cases.TYPE_Q
    return case, cases.q7_thermal(T, te, cases.q7_tau(0.02 / 0.71), boussinesq=True, g_beta=4e-4, T_ref=0.5)


def test_fsi_with_d3q7_buoyancy_fp64(oracle_mod):
    """The filament coupling together with D3Q7 temperature and Boussinesq
    buoyancy (a heated wall below the filament): the IBM step's force plus
    g_beta (T - T_ref) drives the fused D3Q19+D3Q7 step, against the
    oracle's IBM step, Collision and Streaming with D3Q7 on."""
    case, th = _heated_channel()
    fil_args = dict(FIL, clamp=(10.0, 8.0, 1.0), tangent=(0.8, 0.6, 0.0))
    box, m_max, nsteps = (6, 30, 3, 21, 1, 1), 5, 50
    o, of, oout = _oracle_run(oracle_mod, case, fil_args, box, m_max, nsteps, th)
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64")
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, boussinesq=True, g_beta=th.g_beta,
                       T_ref=th.T_ref, model="d3q7")
    lat.init()
    f = lbm_amd.Filament(**fil_args)
    lat.fsi_setup(f, box, m_max=m_max)
    out = lat.fsi_run(nsteps)
    assert lat.sync() == 0
    x, xp, T = f.get()
    rx, rxp, rT = of.get()
    assert np.abs(x - rx).max() < 1e-11 and np.abs(xp - rxp).max() < 1e-11
    scale = np.abs(oout[:, :3]).max()
    assert np.abs(out[:, :3] - oout[:, :3]).max() < 1e-11 * scale
    fl = case.type == cases.TYPE_F
    rho, ux, uy, uz = lat.macro()
    r_rho, r_ux, r_uy, r_uz = o.macro()
    assert rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), fl) < 1e-11
    assert np.abs(rho - r_rho)[fl].max() < 1e-12
    Tf, r_Tf = lat.thermal()[0], o.q7_thermal()[0]
    assert np.abs(Tf - r_Tf)[fl].max() < 1e-12
    assert np.abs(lat.g_populations() - o.q7_g())[fl].max() < 1e-12
    lat.close()
    f.close()


def _setup_c3(case, th, fil, box, storage):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, boussinesq=th.boussinesq, g_beta=th.g_beta,
                       T_ref=th.T_ref, T0=0.0, model="d3q7")
    lat.init()
    f = lbm_amd.Filament(**fil)
    lat.fsi_setup(f, box, m_max=5)
    return lat, f


def test_fsi_thermal_channel_3d_fp64(oracle_mod):
    """The heated channel with inflow, outflow, Dirichlet walls and a
    clamped filament (cases.fsi_thermal_channel) at a small 3D size: the
    coupled D3Q19 + D3Q7 + filament step against the oracle at FP64."""
    case, th, fil, box = cases.fsi_thermal_channel(40, 20, 20, Re=20.0, Ra=1e5, n_markers=32)
    nsteps = 30
    o, of, oout = _oracle_run(oracle_mod, case, dict(fil), box, 5, nsteps, th)
    lat, f = _setup_c3(case, th, fil, box, "f64")
    out = lat.fsi_run(nsteps)
    assert lat.sync() == 0
    x, xp, _ = f.get()
    rx, rxp, _ = of.get()
    assert np.abs(x - rx).max() < 1e-11 and np.abs(xp - rxp).max() < 1e-11
    assert np.abs(out[:, :3] - oout[:, :3]).max() < 1e-11 * np.abs(oout[:, :3]).max()
    fl = case.type == cases.TYPE_F
    _, ux, uy, uz = lat.macro()
    _, r_ux, r_uy, r_uz = o.macro()
    assert rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), fl) < 1e-11
    assert np.abs(lat.thermal()[0] - o.q7_thermal()[0])[fl].max() < 1e-12
    lat.close()
    f.close()


def test_fsi_thermal_channel_3d_fp32_vs_fp64():
    """The same coupled run over 400 steps with FP32 populations against
    FP64 populations on the device: the filament within 1e-4 cells, u and
    T within 1e-4 relative."""
    case, th, fil, box = cases.fsi_thermal_channel(40, 20, 20, Re=20.0, Ra=1e5, n_markers=32)
    res = {}
    for st in ("f64", "f32"):
        lat, f = _setup_c3(case, th, fil, box, st)
        out = lat.fsi_run(400)
        assert lat.sync() == 0
        res[st] = (f.get()[0], lat.macro(), lat.thermal()[0], out)
        lat.close()
        f.close()
    (x64, m64, T64, o64), (x32, m32, T32, o32) = res["f64"], res["f32"]
    fl = case.type == cases.TYPE_F
    assert np.isfinite(o32).all()
    assert np.abs(x32 - x64).max() < 1e-4
    assert rel_l2_u(m32[1:], m64[1:], fl) < 1e-4
    assert np.abs(T32 - T64)[fl].max() < 1e-4
    # the filament bent downstream and the heated wall warmed the fluid
    assert np.abs(x64 - (np.array(fil["clamp"]) + np.outer(np.linspace(0, fil["length"], fil["n"]),
                                                            fil["tangent"]))).max() > 1e-2
    assert T64[fl].mean() > th.T[fl].mean() + 1e-3
    # the smooth start keeps T inside the walls' range
    assert T64[fl].min() >= 0.0 and T32[fl].min() >= 0.0 and T64[fl].max() <= 1.0


@pytest.mark.parametrize("storage,tol_x,tol_u,tol_T", [("f16s", 5e-3, 3e-3, 3e-3), ("f16c", 3e-3, 1e-3, 1.5e-3)])
def test_fsi_thermal_channel_3d_fp16_vs_fp64(storage, tol_x, tol_u, tol_T):
    """The coupled run over 400 steps with 16-bit f and g (the IBM, the
    buoyancy and the filament in FP32/FP64 arithmetic) against FP64
    populations: the filament, u and T within the stated bars."""
    case, th, fil, box = cases.fsi_thermal_channel(40, 20, 20, Re=20.0, Ra=1e5, n_markers=32)
    res = {}
    for st in ("f64", storage):
        lat, f = _setup_c3(case, th, fil, box, st)
        out = lat.fsi_run(400)
        assert lat.sync() == 0
        res[st] = (f.get()[0], lat.macro(), lat.thermal()[0], out)
        lat.close()
        f.close()
    (x64, m64, T64, _), (x16, m16, T16, o16) = res["f64"], res[storage]
    fl = case.type == cases.TYPE_F
    assert np.isfinite(o16).all()
    ex = float(np.abs(x16 - x64).max())
    eu = rel_l2_u(m16[1:], m64[1:], fl)
    eT = float(np.abs(T16 - T64)[fl].max())
    print(f"fsi heated channel {storage}: max|dx| {ex:.3e}, rel L2(u) {eu:.3e}, max|dT| {eT:.3e}")
    assert ex < tol_x
    assert eu < tol_u
    assert eT < tol_T


def test_fsi_long_run_stays_inextensible():
    """Many coupled steps at FP32 flow storage: the filament keeps its
    segment lengths (FP64 structure) and settles downstream of its clamp."""
    case = _case()
    a = math.radians(45.0)
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f32")
    lat.upload(**case.fields())
    lat.init()
    f = lbm_amd.Filament(**FIL, clamp=(10.0, 8.0, 1.0), tangent=(math.cos(a), math.sin(a), 0.0))
    lat.fsi_setup(f, (6, 30, 3, 21, 1, 1), m_max=5)
    out = lat.fsi_run(1500)
    assert lat.sync() == 0
    x = f.get()[0]
    assert np.abs(np.linalg.norm(np.diff(x, axis=0), axis=1) - 1.0).max() < 1e-10
    assert np.isfinite(out).all()
    # the tip has turned from 45 degrees toward the flow direction
    tip = x[-1] - x[0]
    assert math.degrees(math.atan2(tip[1], tip[0])) < 40.0
    lat.close()
    f.close()


def test_fsi_requires_setup_and_device_state():
    case = _case()
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    lat.upload(**case.fields())
    f = lbm_amd.Filament(**FIL, clamp=(10.0, 8.0, 1.0))
    with pytest.raises(lbm_amd.LBMError) as e:  # before lbm_init
        lat.fsi_setup(f, (6, 30, 3, 21, 1, 1))
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    lat.init()
    with pytest.raises(lbm_amd.LBMError) as e:  # no coupling yet
        lat.fsi_run(1)
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    lat.close()
    f.close()
