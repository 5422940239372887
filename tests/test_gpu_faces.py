"""GPU face loops of Streaming (lbm.cpp:299-625) and MacroProp of the
inlet/outlet (and other non-fluid, non-solid) cells against the CPU oracle,
through the C-ABI: rho, u of every cell and the populations of the fluid
and the other-type cells.

The oracle's face loops are pinned by the copy, order and inflow closed
forms of tests/test_oracle_pins.py.  Bars: FP64 storage -- rho within
1e-13, u within 1e-12 of the field's largest magnitude, populations within
1e-13 (the device's fluid sums are pairwise where the reference's are
serial); FP32 storage -- the bounds stated in the test.
"""
import numpy as np
import pytest

import cases
import lbm_amd

pytestmark = pytest.mark.gpu


def _gpu(case, storage="f64"):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.init()
    return lat


def _cmp_all(lat, o, case, tol_rho=1e-13, tol_u=1e-12, tol_f=1e-13, check_f=True):
    """rho, u of every cell but the solids; populations of fluid and
    other-type cells (a solid's populations carry bounce-back values)."""
    got, ref = lat.macro(), o.macro()
    m = case.type != cases.TYPE_S
    assert np.abs(got[0] - ref[0])[m].max() <= tol_rho, "rho"
    scale = max(float(np.abs(np.stack(ref[1:]))[..., m].max()), 1e-30)
    for q in (1, 2, 3):
        err = float(np.abs(got[q] - ref[q])[m].max()) / scale
        assert err <= tol_u, f"u{q}: {err:.3e}"
    if check_f:
        f, rf = lat.populations(), o.f()
        assert np.abs(f - rf)[m].max() <= tol_f, "f"


FACE_CASES = {
    "random_shell": lambda: cases.face_case(6, 5, 4, seed=3),
    "outlets_only": lambda: cases.face_case(5, 4, 4, seed=8, shell=(cases.TYPE_OUT,)),
    "inlets_outlets": lambda: cases.face_case(5, 6, 3, seed=11, shell=(cases.TYPE_OUT, cases.TYPE_IN)),
    "slab_nz1": lambda: cases.face_case(7, 5, 1, seed=4),
    "random_mask": lambda: cases.random_mask_case(6, 5, 4, seed=21),
    "cylinder_2d": lambda: cases.cylinder_bb(40, 24, 1, D=8.0, u_max=0.05, Re=30.0),
    "sphere_3d": lambda: cases.sphere_bb(30, 18, 18, D=6.0, u_max=0.05, Re=30.0),
}


@pytest.mark.parametrize("name", sorted(FACE_CASES))
def test_faces_fused_steps_fp64(oracle_mod, name):
    """lbm_step (fused launches, one face pass per step) against the
    reference's loop of Collision, Streaming, MacroProp."""
    case = FACE_CASES[name]()
    steps = 25
    ref = oracle_mod.run_case(case, steps)
    lat = _gpu(case)
    lat.step(steps)
    assert lat.sync() == 0
    _cmp_all(lat, ref, case)
    lat.close()


@pytest.mark.parametrize("seed", [21, 22, 23, 24])
def test_faces_random_shells_all_types_fp64(oracle_mod, seed):
    """Shells drawn from every cell type the face loops meet (outlet, inlet,
    solid, fluid, TYPE_P, TYPE_L), interior outlet/P cells, 30 fused steps:
    rho, u and f of every non-solid cell."""
    case = cases.face_case(5, 4, 3, seed=seed, shell=(cases.TYPE_OUT, cases.TYPE_IN, cases.TYPE_S, cases.TYPE_F, 4, 5))
    ref = oracle_mod.run_case(case, 30)
    lat = _gpu(case)
    lat.step(30)
    assert lat.sync() == 0
    _cmp_all(lat, ref, case)
    lat.close()


@pytest.mark.parametrize("name", ["random_shell", "inlets_outlets"])
def test_faces_reference_calls_fp64(oracle_mod, name):
    """Collision, Streaming, MacroProp as separate calls, compared after
    every Streaming (the face loops ran, rho, u still the last MacroProp's)
    and every MacroProp.  MacroProp right after Init takes the moments of
    every cell, so the wall velocities of the first Streaming are the
    solids' moments; the solids that the k = 0 inflow loop rewrites at that
    Streaming change their moments once more, which the second Streaming's
    wall velocities carry (a regression case: these tables were refreshed
    only once)."""
    case = FACE_CASES[name]()
    ref = oracle_mod.run_case(case, 0)
    lat = _gpu(case)
    assert ref.macroprop() == 0
    lat.macroprop()
    _cmp_all(lat, ref, case)
    for _ in range(6):
        assert ref.collision() == 0
        ref.streaming()
        lat.collision()
        lat.streaming()
        _cmp_all(lat, ref, case)
        assert ref.macroprop() == 0
        lat.macroprop()
        _cmp_all(lat, ref, case)
    lat.close()


@pytest.mark.parametrize("mode", ["field", "uniform"])
def test_faces_with_force_fp64(oracle_mod, mode):
    """The inlet's u gains F/(2 rho) per step (MacroProp's half force on its
    own equilibrium populations); per-cell and uniform forces."""
    case = cases.face_case(6, 4, 4, seed=6, shell=(cases.TYPE_OUT, cases.TYPE_IN, cases.TYPE_F))
    rng = np.random.default_rng(2)
    if mode == "field":
        F = 2e-5 * rng.standard_normal((3,) + case.shape)
    else:
        F = np.broadcast_to(np.array([2e-5, -1e-5, 5e-6])[:, None, None, None], (3,) + case.shape).copy()
    case.Fx, case.Fy, case.Fz = F[0], F[1], F[2]
    steps = 20
    ref = oracle_mod.run_case(case, steps)
    lat = _gpu(case)
    lat.step(steps)
    assert lat.sync() == 0
    _cmp_all(lat, ref, case)
    lat.close()


def _inlet_over_solid_case(mode):
    """k = 0 inlets under k = Nz+1 solids over most of the (i, j) plane (the
    back-face inflow loop writes feq2 of each inlet into its solid at every
    Streaming, lbm.cpp:599-611), a forced lattice: the inlets' u grows by
    F/(2 rho) per MacroProp, so the solids' moments and the wall velocities
    of the fluid cells below them change every step."""
    case = cases.face_case(6, 5, 4, seed=9, shell=(cases.TYPE_OUT, cases.TYPE_IN, cases.TYPE_S, cases.TYPE_F))
    case.type[1:-1, 1:-1, 0] = cases.TYPE_IN
    case.type[1:-1, 1:-1, -1] = cases.TYPE_S
    case.type[2, 3, -1] = cases.TYPE_F  # a fluid cell in the solid layer: the fluid write too
    rng = np.random.default_rng(4)
    if mode == "field":
        F = 3e-5 * rng.standard_normal((3,) + case.shape)
    else:
        F = np.broadcast_to(np.array([3e-5, -2e-5, 1e-5])[:, None, None, None], (3,) + case.shape).copy()
    case.Fx, case.Fy, case.Fz = F[0], F[1], F[2]
    return case


@pytest.mark.parametrize("mode", ["field", "uniform"])
def test_faces_forced_inlets_over_solids_fp64(oracle_mod, mode):
    """Fused steps and the reference's separate calls against the oracle with
    forced k = 0 inlets over k = Nz+1 solids: every Streaming rewrites the
    solids from the inlets' previous MacroProp (a regression case: the solids
    were written once, from the inlets' values one MacroProp late)."""
    case = _inlet_over_solid_case(mode)
    steps = 20
    ref = oracle_mod.run_case(case, steps)
    lat = _gpu(case)
    lat.step(steps)
    assert lat.sync() == 0
    _cmp_all(lat, ref, case)
    lat.close()
    ref = oracle_mod.run_case(case, 0)
    lat = _gpu(case)
    for _ in range(5):
        assert ref.collision() == 0
        ref.streaming()
        lat.collision()
        lat.streaming()
        assert ref.macroprop() == 0
        lat.macroprop()
        _cmp_all(lat, ref, case)
    lat.close()


def test_faces_on_d3q7_context(oracle_mod):
    """The flow populations of a D3Q19 + D3Q7 context (26-plane buffers)."""
    case, th = cases.q7_random_case(7, 6, 5, seed=9)
    t = cases.face_case(7, 6, 5, seed=5).type
    case.type[:] = np.where(case.type == cases.TYPE_S, cases.TYPE_S, t)
    th.type_energy = np.where(case.type == cases.TYPE_S, th.type_energy, 0).astype(np.int16)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.q7_enable(th.tau_g, th.boussinesq, th.g_beta, th.T_ref)
    o.set_cells(**case.fields())
    o.set_q7(th.T, th.type_energy)
    o.init()
    assert o.run(20) == 0
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64")
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, boussinesq=th.boussinesq, g_beta=th.g_beta,
                       T_ref=th.T_ref, model="d3q7")
    lat.init()
    lat.step(20)
    assert lat.sync() == 0
    _cmp_all(lat, o, case)
    lat.close()


def test_faces_ibm_run_fp64(oracle_mod):
    """An IBM run on the reference's cylinder channel (inlet at i = 0,
    outlets on the other x and y faces): the IBM step's unforced moments of
    the inlet/outlet cells (lbm.cpp:1262-1285), no MacroProp."""
    case, mk, box = cases.ibm_cylinder(40, 30, 1, D=8.0, Re=20.0, perturb=0.01)
    nsteps = 10
    lat = _gpu(case)
    lat.ibm_setup(mk, box, m_max=20)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.set_cells(**case.fields())
    o.init()
    omk = {k: v.copy() for k, v in mk.items()}
    z = np.zeros(3)
    for _ in range(nsteps):
        e, _ = o.ibm_step(omk, z, z, box, m_max=20)
        assert e == 0
        assert o.collision() == 0
        o.streaming()
    lat.ibm_run(nsteps)
    assert lat.sync() == 0
    f, rf = lat.populations(), o.f()
    m = case.type != cases.TYPE_S
    assert np.abs(f - rf)[m].max() <= 1e-13
    # after the run the inlet/outlet cells hold the last IBM step's moments
    io = (case.type == cases.TYPE_IN) | (case.type == cases.TYPE_OUT)
    got, ref = lat.macro(), o.macro()
    assert np.abs(got[0] - ref[0])[io].max() <= 1e-13
    for q in (1, 2, 3):
        assert np.abs(got[q] - ref[q])[io].max() <= 1e-13
    # and the next IBM step takes the moments of the current populations
    e, _ = o.ibm_step(omk, z, z, box, m_max=20)
    assert e == 0
    lat.ibm_apply()
    got, ref = lat.macro(), o.macro()
    assert np.abs(got[0] - ref[0])[io].max() <= 1e-13
    for q in (1, 2, 3):
        assert np.abs(got[q] - ref[q])[io].max() <= 1e-13
    lat.close()


@pytest.mark.parametrize("storage,tol", [("f32", 1e-6), ("f16s", 1e-4), ("f16c", 5e-5)])
def test_faces_reduced_storage_against_fp64(storage, tol):
    """FP32 and 16-bit storages keep the inlet/outlet cells in FP64: their
    rho, u follow FP64 storage's to the fluid's own storage error (absolute
    bound stated per storage, over every non-solid cell after 40 steps)."""
    case = FACE_CASES["inlets_outlets"]()
    out = {}
    for st in ("f64", storage):
        lat = _gpu(case, st)
        lat.step(40)
        assert lat.sync() == 0
        out[st] = lat.macro()
        lat.close()
    m = case.type != cases.TYPE_S
    errs = [float(np.abs(out[storage][q] - out["f64"][q])[m].max()) for q in range(4)]
    print(storage, "max |d rho|, |d u|:", errs, flush=True)
    assert max(errs) < tol


def test_faces_density_range_flag():
    """MacroProp's density check (lbm.cpp:772-775) covers the other-type
    cells: an outlet whose copied populations put its rho above 100."""
    case = cases.face_case(4, 4, 4, seed=1, shell=(cases.TYPE_OUT,), p_solid=0.0)
    case.rho[:] = 1.0
    case.rho[0, 2, 2] = 300.0  # the outlet keeps 16 of its set-up populations
    lat = _gpu(case)
    lat.step(1)
    code, first_bad = lat.status()
    assert code == lbm_amd.LBM_ERR_DENSITY_RANGE
    assert first_bad >= 0
    lat.close()


def test_faces_stress_fp64(oracle_mod):
    """MORE_DETAILS stress of the other-type cells from their own
    populations (lbm.cpp:681-705)."""
    case = FACE_CASES["random_shell"]()
    ref = oracle_mod.run_case(case, 12)
    lat = _gpu(case)
    lat.step(12)
    assert lat.sync() == 0
    sig, _ = lat.stress()
    r_sig, _ = ref.stress()
    m = case.type != cases.TYPE_S
    scale = float(np.abs(r_sig[m]).max())
    assert float(np.abs(sig - r_sig)[m].max()) / scale < 1e-12
    lat.close()


def test_batch_rejects_lattices_with_inlets():
    case = FACE_CASES["inlets_outlets"]()
    lat = _gpu(case)
    with pytest.raises(lbm_amd.LBMError) as e:
        lbm_amd.Batch([lat])
    assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
    lat.close()
