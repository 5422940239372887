"""The device-packed .vtr output (lbm_output_vtk) against the restated
OutputVTK (oracle/vtk_oracle.py) on the oracle's state after the same steps.

Bar: the XML header and every integer-valued block (CellType,
EnergyCellType, coordinates) byte-identical; float blocks within 1 ulp of
float32 (the FP64 states agree to roundoff, which can move a value across a
float32 rounding boundary)."""
import numpy as np
import pytest

import cases
import lbm_amd
import vtk_oracle

pytestmark = pytest.mark.gpu


def _ulp_close(a, b, ulps=1):
    """Within `ulps` float32 ulps of each value, or within float32 epsilon of
    the block's largest magnitude (components that are zero up to FP64
    roundoff carry no digits to compare)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    # map the sign-magnitude float order onto integers
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    floor = np.finfo(np.float32).eps * float(np.abs(b.astype(np.float64)).max())
    ok = (np.abs(ia - ib) <= ulps) | (np.abs(a.astype(np.float64) - b.astype(np.float64)) <= floor)
    return bool(ok.all())


def _compare(path_gpu, path_ref):
    hg, eg, ag = vtk_oracle.read_vtr(path_gpu)
    hr, er, ar = vtk_oracle.read_vtr(path_ref)
    assert hg == hr
    assert list(ag) == list(ar)
    for name in ag:
        if name in ("CellType", "EnergyCellType") or name.startswith("Coordinate"):
            assert np.array_equal(ag[name][1], ar[name][1]), name
        else:
            assert _ulp_close(ag[name][1], ar[name][1]), name


def test_vtk_plain_cavity(oracle_mod, tmp_path):
    case = cases.cavity(10, u_lid=0.05, Re=100.0)
    ref = oracle_mod.run_case(case, 30)
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    lat.upload(**case.fields())
    lat.init()
    lat.step(30)
    pg, pr = tmp_path / "g.vtr", tmp_path / "r.vtr"
    lat.output_vtk(str(pg), u_scale=0.05)
    rho, ux, uy, uz = ref.macro()
    vtk_oracle.write_vtr(pr, case.Nx, case.Ny, case.Nz, rho, ux, uy, uz, case.type, u_scale=0.05)
    _compare(pg, pr)
    lat.close()


def test_vtk_cylinder_with_inlet_and_outlets(oracle_mod, tmp_path):
    """The bounce-back cylinder (inlet at i = 0, outlets on the other x and
    y faces): every block, the inlet and outlet cells' density and velocity
    included, from the face loops on the device."""
    case = cases.cylinder_bb(40, 24, 1, D=8.0, u_max=0.05, Re=30.0)
    ref = oracle_mod.run_case(case, 40)
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    lat.upload(**case.fields())
    lat.init()
    lat.step(40)
    pg, pr = tmp_path / "g.vtr", tmp_path / "r.vtr"
    lat.output_vtk(str(pg), u_scale=0.05)
    rho, ux, uy, uz = ref.macro()
    vtk_oracle.write_vtr(pr, case.Nx, case.Ny, case.Nz, rho, ux, uy, uz, case.type, u_scale=0.05)
    _compare(pg, pr)
    lat.close()


def test_vtk_ibm_force_density(oracle_mod, tmp_path):
    case, mk, box = cases.ibm_cylinder(40, 30, 1, D=8.0, Re=20.0, perturb=0.01)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.set_cells(**case.fields())
    o.init()
    omk = {k: v.copy() for k, v in mk.items()}
    for _ in range(5):
        e, _ = o.ibm_step(omk, (0, 0, 0), (0, 0, 0), box, m_max=20)
        assert e == 0 and o.collision() == 0
        o.streaming()
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    lat.upload(**case.fields())
    lat.init()
    lat.ibm_setup(mk, box, m_max=20)
    for _ in range(4):
        lat.ibm_apply()
        lat.collision()
        lat.streaming()
    # the 5th step: IBLBM, then the output sees the IBM force (the reference
    # writes after its Streaming, with fluid.F as IBLBM left it)
    lat.ibm_apply()
    lat.collision()
    lat.streaming()
    pg, pr = tmp_path / "g.vtr", tmp_path / "r.vtr"
    lat.output_vtk(str(pg), u_scale=0.04, force=True)
    rho, ux, uy, uz = o.macro()
    vtk_oracle.write_vtr(pr, case.Nx, case.Ny, case.Nz, rho, ux, uy, uz, case.type, u_scale=0.04,
                         force=o.force())
    hg, _, ag = vtk_oracle.read_vtr(pg)
    hr, _, ar = vtk_oracle.read_vtr(pr)
    assert hg == hr
    fl = np.transpose(case.type[:, :, 1:2], (2, 1, 0)).ravel() == cases.TYPE_F
    for name in ("Density", "Velocity"):
        nc = ag[name][0]
        assert _ulp_close(ag[name][1].reshape(-1, nc)[fl], ar[name][1].reshape(-1, nc)[fl]), name
    Fg = ag["Force density"][1].reshape(-1, 3)[fl]
    Fr = ar["Force density"][1].reshape(-1, 3)[fl]
    assert np.abs(Fg - Fr).max() <= 1e-6 * max(1e-30, np.abs(Fr).max())
    assert np.abs(Fr).max() > 0
    lat.close()


def test_vtk_thermal_ratio_blocks(oracle_mod, tmp_path):
    case, th = cases.couette_thermal(Ny=12, Re=10.0)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.thermal_enable(th.tau_g, th.nu, th.cp)
    o.set_cells(**case.fields())
    o.set_thermal(th.T, th.type_energy)
    o.init()
    assert o.run(50) == 0
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, nu=th.nu, cp=th.cp)
    lat.init()
    lat.step(50)
    p = th.params
    kw = dict(u_scale=p["u_wall"], velocity_name=1, T_mode=1, T_a=p["Tb"], T_b=p["dT"])
    pg, pr = tmp_path / "g.vtr", tmp_path / "r.vtr"
    lat.output_vtk(str(pg), thermal=True, **kw)
    rho, ux, uy, uz = o.macro()
    T, qx, qy, qz = o.thermal()
    vtk_oracle.write_vtr(pr, case.Nx, case.Ny, case.Nz, rho, ux, uy, uz, case.type, T=T, q=(qx, qy, qz),
                         type_energy=th.type_energy, **kw)
    hg, _, ag = vtk_oracle.read_vtr(pg)
    hr, _, ar = vtk_oracle.read_vtr(pr)
    assert hg == hr
    fl = np.transpose(case.type[:, :, 1:-1], (2, 1, 0)).ravel() == cases.TYPE_F
    for name in ("Density", "Velocity Ratio", "Temperature Ratio", "Heat flux"):
        nc = ag[name][0]
        # the temperature ratio divides a 1e-4 difference: allow for the FP64
        # roundoff of T being magnified there
        tol = 4 if name == "Temperature Ratio" else 1
        assert _ulp_close(ag[name][1].reshape(-1, nc)[fl], ar[name][1].reshape(-1, nc)[fl], tol), name
    assert np.array_equal(ag["EnergyCellType"][1], ar["EnergyCellType"][1])
    lat.close()


def test_vtk_async_copy_keeps_the_state_of_begin(tmp_path):
    """lbm_output_vtk_begin packs the state at the call; steps taken before
    lbm_output_vtk_end do not change the file."""
    case = cases.cavity(12, u_lid=0.05, Re=100.0)
    a = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    b = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    for lat in (a, b):
        lat.upload(**case.fields())
        lat.init()
        lat.step(20)
    a.output_vtk_begin(u_scale=0.05)
    a.step(15)
    a.output_vtk_end(str(tmp_path / "a.vtr"))
    b.output_vtk(str(tmp_path / "b.vtr"), u_scale=0.05)
    assert (tmp_path / "a.vtr").read_bytes() == (tmp_path / "b.vtr").read_bytes()
    a.close()
    b.close()
