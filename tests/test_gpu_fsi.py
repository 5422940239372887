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


def _oracle_run(oracle_mod, case, fil_args, box, m_max, nsteps):
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.set_cells(**case.fields())
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
