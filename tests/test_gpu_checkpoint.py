"""Checkpoint and resume (lbm_checkpoint_size/save/load; the reference has
none, its only resume being the interactive prompt that extends the step
count, main.cpp:397-410).

Each test steps a lattice, takes a checkpoint, steps on; a second lattice
built by the same set-up restores the checkpoint and takes the same steps.
Bar: bit for bit -- populations (cells of every type), rho and u, and the
thermal fields, running statistics and status where the lattice has them.
"""
import numpy as np
import pytest

import cases
import lbm_amd

pytestmark = pytest.mark.gpu


def _plain(case, storage):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.init()
    return lat


def _energy(case, th, storage):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, nu=th.nu, cp=th.cp, j_dot_n=th.j_dot_n,
                       boussinesq=th.boussinesq, g_beta=th.g_beta, T_ref=th.T_ref)
    lat.init()
    return lat


def _q7(case, th, storage):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, boussinesq=th.boussinesq, g_beta=th.g_beta,
                       T_ref=th.T_ref, model="d3q7")
    lat.init()
    return lat


def _state(lat, thermal):
    out = [lat.populations(), *lat.macro()]
    if thermal:
        out += list(lat.thermal()) + [lat.g_populations()]
    return out


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(np.asarray(x), np.asarray(y))


def _resume(make, n0, n1, thermal=False, stats=False):
    """Steps n0, checkpoint, n1 more; a fresh lattice restores and steps n1."""
    a = make()
    if stats:
        a.step(n0)
        a.stats_update(n0, n0)
    else:
        a.step(n0)
    assert a.sync() == 0
    ck = a.checkpoint()
    a.step(n1)
    if stats:
        a.stats_update(n0 + n1, n1)
    assert a.sync() == 0
    want = _state(a, thermal)
    want_stats = a.stats() if stats else None
    a.close()
    b = make()
    b.restore(ck)
    b.step(n1)
    if stats:
        b.stats_update(n0 + n1, n1)
    assert b.sync() == 0
    _same(_state(b, thermal), want)
    if stats:
        _same(b.stats(), want_stats)
    b.close()
    return ck


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s", "f16c"])
def test_resume_cavity(storage):
    case = cases.cavity(20, u_lid=0.05, Re=100.0)
    _resume(lambda: _plain(case, storage), 17, 13)


def test_resume_channel_with_faces_and_statistics():
    """Inlet/outlet slots (the face loops' FP64 populations) and the
    running statistics carried across the checkpoint."""
    case = cases.channel(24, 12, 6, u_in=0.04, Re=20.0, perturb=0.01, seed=3)
    _resume(lambda: _plain(case, "f32"), 10, 10, stats=True)


def test_resume_forced_periodic():
    case = cases.poiseuille(Nx=2, Ny=16, Nz=2, tau=0.9, g=1e-5)
    _resume(lambda: _plain(case, "f64"), 25, 25)


@pytest.mark.parametrize("storage", ["f64", "f16c"])
def test_resume_energy_thermal(storage):
    """The ENERGY branch: g, T, q and the previous collision's Phi and u
    (its dF/dt history) carried across."""
    case, th = cases.couette_thermal(Ny=12, Re=10.0)
    _resume(lambda: _energy(case, th, storage), 15, 15, thermal=True)


@pytest.mark.parametrize("storage", ["f32", "f16c"])
def test_resume_d3q7(storage):
    case, th = cases.q7_random_case(7, 6, 5, seed=9)
    _resume(lambda: _q7(case, th, storage), 12, 12, thermal=True)


def test_checkpoint_rejects_mismatch_and_ibm():
    case = cases.cavity(12, u_lid=0.05, Re=100.0)
    a = _plain(case, "f32")
    a.step(3)
    ck = a.checkpoint()
    # another storage, another size
    b = _plain(case, "f64")
    with pytest.raises(lbm_amd.LBMError):
        b.restore(ck)
    b.close()
    other = cases.cavity(14, u_lid=0.05, Re=100.0)
    b = _plain(other, "f32")
    with pytest.raises(lbm_amd.LBMError):
        b.restore(ck)
    # a truncated buffer
    with pytest.raises(lbm_amd.LBMError):
        a.restore(ck[:-8])
    b.close()
    a.close()
    # before lbm_init
    c = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f32")
    c.upload(**case.fields())
    with pytest.raises(lbm_amd.LBMError):
        c.checkpoint()
    c.close()
    # an immersed body that moved: its marker positions are state of their own
    icase, mk, box = cases.ibm_cylinder(36, 24, 2, D=8.0, Re=20.0, center=(14.0, 12.0))
    d = _plain(icase, "f32")
    d.ibm_setup(mk, box, m_max=4)
    d.ibm_run(2, vcm=np.tile([0.001, 0.0, 0.0], (3, 1)))
    with pytest.raises(lbm_amd.LBMError):
        d.checkpoint()
    d.close()


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_resume_ibm_cylinder_at_rest(storage):
    """A rigid body at rest: every IBM step recomputes the markers' u_b and
    F_b from the populations, so the lattice state resumes the IBM run; the
    restoring lattice repeats the same lbm_ibm_setup.  Fields and the drag,
    lift and interior density series bit for bit."""
    icase, mk, box = cases.ibm_cylinder(36, 24, 2, D=8.0, Re=20.0, center=(14.0, 12.0))

    def make():
        lat = _plain(icase, storage)
        lat.ibm_setup(mk, box, m_max=4)
        return lat

    a = make()
    a.ibm_run(15)
    assert a.sync() == 0
    ck = a.checkpoint()
    want_out = a.ibm_run(10)
    assert a.sync() == 0
    want = _state(a, False)
    a.close()
    b = make()
    b.restore(ck)
    got_out = b.ibm_run(10)
    assert b.sync() == 0
    np.testing.assert_array_equal(got_out, want_out)
    _same(_state(b, False), want)
    b.close()
