"""GPU parity of the implicit immersed-boundary step (IBLBM, lbm.cpp:1135-1417)
against the CPU oracle, through the C-ABI.

Each test runs the reference's IBM time loop (main.cpp:199-205: IBLBM,
Collision, Streaming; no MacroProp) on both sides from the same seeded
fields and markers:
  - the step-by-step path (lbm_ibm_apply, lbm_collide, lbm_stream);
  - the fused path (lbm_ibm_run: box moments pulled from f*, one fused
    stream+collide launch per step).
Bars: FP64 storage -- per-step Drag, Lift and interior density within
1e-11 of the oracle's (relative to the series' magnitude), fluid u within
relative L2 1e-11, rho within 1e-12, marker u_b and F_b within 1e-10 of
their magnitude; FP32/FP16 storage -- the bounds stated in the test.  The
device sums drag and lift as trees, the reference serially, hence the
roundoff-level (not bit-exact) bars.
"""
import math

import numpy as np
import pytest

import cases
import lbm_amd
from conftest import rel_l2_u

pytestmark = pytest.mark.gpu


def _setup(case, mk, box, storage="f64", m_max=20, vsolid=0.0, moving=False):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.init()
    lat.ibm_setup(mk, box, m_max=m_max, vsolid=vsolid, moving=moving)
    return lat


def _oracle(oracle_mod, case):
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.set_cells(**case.fields())
    o.init()
    return o


def _oracle_run(o, mk, box, vcm, m_max=20, vsolid=0.0, scheme="implicit"):
    """vcm: (nsteps+1, 3); returns (nsteps, 3) Drag, Lift, interior rho."""
    outs = []
    for s in range(vcm.shape[0] - 1):
        e, out = o.ibm_step(mk, vcm[s + 1], vcm[s], box, m_max=m_max, vsolid=vsolid, scheme=scheme)
        assert e == 0
        assert o.collision() == 0
        o.streaming()
        outs.append(out.copy())
    return np.array(outs)


def _cmp_fields(lat, o, case, tol_u, tol_rho):
    rho, ux, uy, uz = lat.macro()
    r_rho, r_ux, r_uy, r_uz = o.macro()
    m = case.type == cases.TYPE_F
    eu = rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), m)
    er = float(np.abs(rho - r_rho)[m].max())
    assert eu <= tol_u, f"rel L2(u) = {eu:.3e} > {tol_u}"
    assert er <= tol_rho, f"max|drho| = {er:.3e} > {tol_rho}"
    return eu, er


def _cmp_series(got, ref, tol, what):
    for c, name in enumerate(("drag", "lift", "rho_interior")):
        scale = max(1.0, float(np.abs(ref[:, c]).max()))
        err = float(np.abs(got[:, c] - ref[:, c]).max()) / scale
        assert err <= tol, f"{what} {name}: max err {err:.3e} > {tol}"


def _cmp_markers(lat, mk, tol):
    g = lat.ibm_markers()
    for k in ("x", "y", "z"):
        assert np.abs(g[k] - mk[k]).max() <= 1e-13, k
    for k in ("ub", "Fb"):
        scale = max(1e-30, float(np.abs(mk[k]).max()))
        err = float(np.abs(g[k] - mk[k]).max()) / scale
        assert err <= tol, f"marker {k}: {err:.3e} > {tol}"


IBM_CASES = {
    # the reference's configuration: a 2D slab (Nz = 1, box k = 1..1)
    "slab": lambda: cases.ibm_cylinder(40, 30, 1, D=8.0, Re=20.0, perturb=0.01),
    # the 3D generalisation: one marker ring per layer, box k = 1..Nz
    "span": lambda: cases.ibm_cylinder(36, 24, 6, D=8.0, Re=20.0, perturb=0.01, center=(14.0, 12.0)),
}


def _transverse_vcm(nsteps, amp=0.1, freq=0.05):
    """V_CM of Cylinder_Transverse_Oscillation (lbm.cpp:1145-1147) at
    t = step*dt: (0, -2 pi f A sin(2 pi f t), 0); row 0 is the initial 0."""
    v = np.zeros((nsteps + 1, 3))
    t = np.arange(1, nsteps + 1, dtype=np.float64)
    v[1:, 1] = -2 * cases.PI_REF * freq * amp * np.sin(2 * cases.PI_REF * freq * t)
    return v


@pytest.mark.parametrize("name", sorted(IBM_CASES))
def test_ibm_apply_path_fp64(oracle_mod, name):
    case, mk, box = IBM_CASES[name]()
    nsteps = 12
    lat = _setup(case, mk, box)
    o = _oracle(oracle_mod, case)
    omk = {k: v.copy() for k, v in mk.items()}
    vcm = np.zeros((nsteps + 1, 3))
    ref = _oracle_run(o, omk, box, vcm)
    got = []
    for s in range(nsteps):
        got.append(lat.ibm_apply(vcm[s + 1], vcm[s]))
        lat.collision()
        lat.streaming()
    got = np.array(got)
    assert lat.sync() == 0
    _cmp_series(got, ref, 1e-11, name)
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    _cmp_markers(lat, omk, 1e-10)
    lat.close()


@pytest.mark.parametrize("name", sorted(IBM_CASES))
def test_ibm_fused_run_fp64(oracle_mod, name):
    case, mk, box = IBM_CASES[name]()
    nsteps = 16
    lat = _setup(case, mk, box)
    o = _oracle(oracle_mod, case)
    omk = {k: v.copy() for k, v in mk.items()}
    ref = _oracle_run(o, omk, box, np.zeros((nsteps + 1, 3)))
    # split into two runs: the second starts from f (ST_F) again
    got = np.concatenate([lat.ibm_run(7), lat.ibm_run(nsteps - 7)])
    assert lat.sync() == 0
    _cmp_series(got, ref, 1e-11, name)
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    _cmp_markers(lat, omk, 1e-10)
    lat.close()


@pytest.mark.parametrize("name,m_max", [("span", 1), ("span", 2), ("span", 3), ("span", 4), ("span", 7),
                                         ("dense", 2), ("dense", 3), ("dense", 20)])
def test_ibm_iteration_counts_fp64(oracle_mod, name, m_max):
    """The m_max passes run as a single pass (m_max odd) plus products by
    (I - H)^2 and a closing launch: every branch of that schedule (no
    product, one, several; with and without the single pass) against the
    oracle's m_max spread/interpolate passes, fused and step-by-step.
    "dense": 50 markers per ring of D = 8 (spacing 0.5), whose marker rows
    are longer than one wavefront (65..128 entries)."""
    if name == "dense":
        case, mk, box = cases.ibm_cylinder(36, 24, 6, D=8.0, Re=20.0, perturb=0.01, center=(14.0, 12.0),
                                           n_theta=50)
    else:
        case, mk, box = IBM_CASES[name]()
    nsteps = 6
    o = _oracle(oracle_mod, case)
    omk = {k: v.copy() for k, v in mk.items()}
    ref = _oracle_run(o, omk, box, np.zeros((nsteps + 1, 3)), m_max=m_max)
    lat = _setup(case, mk, box, m_max=m_max)
    got = lat.ibm_run(nsteps)
    assert lat.sync() == 0
    _cmp_series(got, ref, 1e-11, f"m_max={m_max}")
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    _cmp_markers(lat, omk, 1e-10)
    lat.close()
    lat = _setup(case, mk, box, m_max=m_max)
    got = []
    for s in range(nsteps):
        got.append(lat.ibm_apply(np.zeros(3), np.zeros(3)))
        lat.collision()
        lat.streaming()
    assert lat.sync() == 0
    _cmp_series(np.array(got), ref, 1e-11, f"m_max={m_max} apply")
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    lat.close()


def test_ibm_moving_body_added_mass_fp64(oracle_mod):
    """Transverse oscillation: the markers move every step (nearest points,
    interior cells and tiles recomputed) and the added-mass term is live."""
    case, mk, box = cases.ibm_cylinder(48, 36, 1, D=8.0, Re=20.0, perturb=0.0,
                                       center=(14.0, 18.0))
    vs = case.params["vsolid"]
    nsteps = 14
    vcm = _transverse_vcm(nsteps)
    o = _oracle(oracle_mod, case)
    omk = {k: v.copy() for k, v in mk.items()}
    ref = _oracle_run(o, omk, box, vcm, vsolid=vs)
    assert np.all(np.isfinite(ref))
    lat = _setup(case, mk, box, vsolid=vs)
    got = lat.ibm_run(nsteps, vcm)
    assert lat.sync() == 0
    _cmp_series(got, ref, 1e-11, "moving")
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    _cmp_markers(lat, omk, 1e-10)
    lat.close()
    # the same through the step-by-step path
    lat = _setup(case, mk, box, vsolid=vs)
    got2 = np.array([_apply_step(lat, vcm[s + 1], vcm[s]) for s in range(nsteps)])
    _cmp_series(got2, ref, 1e-11, "moving/apply")
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    lat.close()


def test_ibm_box_follows_moving_body_fp64(oracle_mod):
    """Cylinder_Inline_Oscillating motion, V_CM = (-u cos(2 pi f t), 0, 0)
    (lbm.cpp:1143-1144), with the reference's box that follows the centre,
    [floor(c - D), ceil(c + D)] (lbm.cpp:1162-1166): the box shifts and
    grows during the run."""
    D = 8.0
    case, mk, box0 = cases.ibm_cylinder(60, 40, 1, D=D, Re=20.0, center=(30.0, 20.0))
    nsteps = 16
    u, fr = 0.1, 0.01
    vcm = np.zeros((nsteps + 1, 3))
    t = np.arange(1, nsteps + 1, dtype=np.float64)
    vcm[1:, 0] = -u * np.cos(2 * cases.PI_REF * fr * t)
    vs = case.params["vsolid"]
    o = _oracle(oracle_mod, case)
    omk = {k: v.copy() for k, v in mk.items()}
    cx, cy = 30.0, 20.0
    ref, boxes = [], []
    for s in range(nsteps):
        cx += (vcm[s + 1, 0] + vcm[s, 0]) * 1.0 / 2
        cy += (vcm[s + 1, 1] + vcm[s, 1]) * 1.0 / 2
        box = (math.floor(cx - D), math.ceil(cx + D), math.floor(cy - D), math.ceil(cy + D), 1, 1)
        boxes.append(box)
        e, out = o.ibm_step(omk, vcm[s + 1], vcm[s], box, m_max=20, vsolid=vs)
        assert e == 0 and o.collision() == 0
        o.streaming()
        ref.append(out.copy())
    ref = np.array(ref)
    assert len(set(boxes)) > 1 and boxes[-1] != box0  # the box did move
    for fused in (True, False):
        lat = _setup(case, mk, box0, vsolid=vs)
        lat.ibm_setup(mk, box0, m_max=20, vsolid=vs, follow=(30.0, 20.0, D))
        if fused:
            got = lat.ibm_run(nsteps, vcm)
        else:
            got = np.array([_apply_step(lat, vcm[s + 1], vcm[s]) for s in range(nsteps)])
        assert lat.sync() == 0
        b, cen = lat.ibm_box()
        assert b == boxes[-1]
        assert np.allclose(cen, (cx, cy), rtol=0, atol=1e-13)
        _cmp_series(got, ref, 1e-11, f"follow fused={fused}")
        _cmp_fields(lat, o, case, 1e-11, 1e-12)
        _cmp_markers(lat, omk, 1e-10)
        lat.close()


def _apply_step(lat, v, vb):
    out = lat.ibm_apply(v, vb)
    lat.collision()
    lat.streaming()
    return out


@pytest.mark.parametrize("storage,tol_u,tol_series", [("f32", 1e-5, 1e-4), ("f16s", 5e-3, 2e-2),
                                                      ("f16c", 5e-3, 2e-2)])
def test_ibm_reduced_storage(oracle_mod, storage, tol_u, tol_series):
    case, mk, box = IBM_CASES["slab"]()
    nsteps = 40
    o = _oracle(oracle_mod, case)
    omk = {k: v.copy() for k, v in mk.items()}
    ref = _oracle_run(o, omk, box, np.zeros((nsteps + 1, 3)))
    lat = _setup(case, mk, box, storage=storage)
    got = lat.ibm_run(nsteps)
    assert lat.sync() == 0
    _cmp_series(got, ref, tol_series, storage)
    _cmp_fields(lat, o, case, tol_u, 50 * tol_u)
    lat.close()


def test_ibm_m_max_zero_is_unforced(oracle_mod):
    """m_max = 0: no force passes; the step is the plain BGK step with the
    unforced moments (lbm.cpp:1262-1285) and F = 0."""
    case, mk, box = IBM_CASES["slab"]()
    nsteps = 6
    o = _oracle(oracle_mod, case)
    omk = {k: v.copy() for k, v in mk.items()}
    ref = _oracle_run(o, omk, box, np.zeros((nsteps + 1, 3)), m_max=0)
    lat = _setup(case, mk, box, m_max=0)
    got = lat.ibm_run(nsteps)
    assert np.all(got[:, :2] == 0.0)
    _cmp_series(got, ref, 1e-12, "m_max=0")
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    lat.close()


def test_ibm_call_order_and_box_checks():
    case, mk, box = IBM_CASES["slab"]()
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau)
    lat.upload(**case.fields())
    with pytest.raises(lbm_amd.LBMError) as e:  # before lbm_init
        lat.ibm_setup(mk, box)
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    lat.init()
    with pytest.raises(lbm_amd.LBMError) as e:  # the box reaches the inlet face
        lat.ibm_setup(mk, (0,) + tuple(box[1:]))
    assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
    with pytest.raises(lbm_amd.LBMError) as e:  # outside the lattice
        lat.ibm_setup(mk, box[:1] + (case.Nx + 2,) + tuple(box[2:]))
    assert e.value.code == lbm_amd.LBM_ERR_ARG
    lat.ibm_setup(mk, box)
    with pytest.raises(lbm_amd.LBMError) as e:  # collision needs IBLBM first
        lat.collision()
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    with pytest.raises(lbm_amd.LBMError) as e:  # the plain step has no IBM
        lat.step(1)
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    lat.ibm_apply()
    lat.collision()
    with pytest.raises(lbm_amd.LBMError) as e:  # IBLBM reads f, not f*
        lat.ibm_apply()
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    lat.streaming()
    # re-initialising keeps the body; re-uploading drops it
    lat.init()
    out = lat.ibm_apply()
    assert math.isfinite(out[0])
    lat.upload(**case.fields())
    lat.init()
    lat.step(2)
    assert lat.sync() == 0
    lat.close()


@pytest.mark.parametrize("path", ["fused", "apply"])
def test_ibm_with_d3q7_buoyancy_fp64(oracle_mod, path):
    """The rigid IBM on a D3Q19 + D3Q7 context: each IBM step ends with the
    D3Q7 temperature of every fluid cell and its Boussinesq force added to
    the IBM's force and velocity; the oracle's IBM step does the same."""
    case, mk, box = cases.ibm_cylinder(48, 36, 1, D=8.0, Re=20.0, perturb=0.0, center=(14.0, 18.0))
    rng = np.random.default_rng(31)
    T = 0.2 + 0.6 * rng.random(case.shape)
    th = cases.q7_thermal(T, np.zeros(case.shape, np.int16), 0.9, boussinesq=True, g_beta=1e-3, T_ref=0.5)
    nsteps = 12
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.q7_enable(th.tau_g, True, th.g_beta, th.T_ref)
    o.set_cells(**case.fields())
    o.set_q7(th.T, th.type_energy)
    o.init()
    omk = {k: v.copy() for k, v in mk.items()}
    ref = _oracle_run(o, omk, box, np.zeros((nsteps + 1, 3)))
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64")
    lat.upload(**case.fields())
    lat.upload_thermal(th.T, th.type_energy, tau_g=th.tau_g, boussinesq=True, g_beta=th.g_beta,
                       T_ref=th.T_ref, model="d3q7")
    lat.init()
    lat.ibm_setup(mk, box, m_max=20)
    if path == "fused":
        got = lat.ibm_run(nsteps)
    else:
        got = []
        for _ in range(nsteps):
            got.append(lat.ibm_apply())
            lat.collision()
            lat.streaming()
        got = np.array(got)
    assert lat.sync() == 0
    _cmp_series(got, ref, 1e-11, "d3q7")
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    fl = case.type == cases.TYPE_F
    assert np.abs(lat.thermal()[0] - o.q7_thermal()[0])[fl].max() < 1e-12
    lat.close()


@pytest.mark.parametrize("name", sorted(IBM_CASES))
@pytest.mark.parametrize("path", ["fused", "apply"])
def test_ibm_explicit_scheme_fp64(oracle_mod, name, path):
    """The explicit branch (one pass, F = +spread, density at the marker's
    local nearest point, cell-summed drag and lift, lbm.cpp:1171-1258)
    against the oracle's restatement, fused and step by step."""
    case, mk, box = IBM_CASES[name]()
    nsteps = 12
    o = _oracle(oracle_mod, case)
    omk = {k: v.copy() for k, v in mk.items()}
    ref = _oracle_run(o, omk, box, np.zeros((nsteps + 1, 3)), scheme="explicit")
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64")
    lat.upload(**case.fields())
    lat.init()
    lat.ibm_setup(mk, box, scheme="explicit")
    if path == "fused":
        got = lat.ibm_run(nsteps)
    else:
        got = np.array([_apply_step(lat, (0, 0, 0), (0, 0, 0)) for _ in range(nsteps)])
    assert lat.sync() == 0
    _cmp_series(got, ref, 1e-11, "explicit " + name)
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    _cmp_markers(lat, omk, 1e-10)
    lat.close()


def test_ibm_explicit_moving_markers_fixed_box_fp64(oracle_mod):
    """Transverse oscillation of the markers inside a box that stays in
    place, with added mass: the explicit branch against the oracle."""
    case, mk, box = cases.ibm_cylinder(48, 36, 1, D=8.0, Re=20.0, perturb=0.0, center=(14.0, 18.0))
    vs = case.params["vsolid"]
    nsteps = 14
    vcm = _transverse_vcm(nsteps)
    o = _oracle(oracle_mod, case)
    omk = {k: v.copy() for k, v in mk.items()}
    ref = _oracle_run(o, omk, box, vcm, vsolid=vs, scheme="explicit")
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64")
    lat.upload(**case.fields())
    lat.init()
    lat.ibm_setup(mk, box, vsolid=vs, moving=True, scheme="explicit")
    got = lat.ibm_run(nsteps, vcm)
    assert lat.sync() == 0
    _cmp_series(got, ref, 1e-11, "explicit moving")
    _cmp_fields(lat, o, case, 1e-11, 1e-12)
    _cmp_markers(lat, omk, 1e-10)
    lat.close()


def test_ibm_explicit_rejects_follow_and_forced_setups():
    case, mk, box = IBM_CASES["slab"]()
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64")
    lat.upload(**case.fields())
    lat.init()
    with pytest.raises(lbm_amd.LBMError) as e:
        lat.ibm_setup(mk, box, scheme="explicit", follow=(case.Nx / 5.0, case.Ny / 2.0, 8.0))
    assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
    lat.close()
    f = dict(case.fields())
    f["Fx"] = np.full(case.shape, 1e-6)
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64")
    lat.upload(**f)
    lat.init()
    with pytest.raises(lbm_amd.LBMError) as e:
        lat.ibm_setup(mk, box, scheme="explicit")
    assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
    lat.close()


def test_ibm_explicit_rejects_density_points_outside_the_box(oracle_mod):
    """A marker on a lattice point has no neighbour ahead along its normal
    (floor = ceil on every axis), so the explicit branch reads the density
    of cell (0, 0, 1), the start value of its search (lbm.cpp:1216-1225):
    the oracle does; the device, which computes the IBM step's density only
    inside the box, refuses the set-up instead of returning NaN forces."""
    case, mk, box = IBM_CASES["slab"]()
    mk = {k: v.copy() for k, v in mk.items()}
    mk["x"][0] = float(round(mk["x"][0]))
    mk["y"][0] = float(round(mk["y"][0]))
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.set_cells(**case.fields())
    o.init()
    omk = {k: v.copy() for k, v in mk.items()}
    err, _ = o.ibm_step(omk, (0, 0, 0), (0, 0, 0), box, scheme="explicit")
    assert err == 0 and np.isfinite(omk["Fb"]).all()
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64")
    lat.upload(**case.fields())
    lat.init()
    with pytest.raises(lbm_amd.LBMError) as e:
        lat.ibm_setup(mk, box, scheme="explicit")
    assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
    lat.close()
