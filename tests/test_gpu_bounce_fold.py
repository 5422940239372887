"""The two layouts of Streaming's half-way bounce-back (lbm.cpp:268-275)
give the same step, through the C-ABI.

LBM_BB_AUTO resolves the wall links into all-solid faces of the padded
lattice in the pull address: the fluid cell reads its own f*(x, opp m) and
adds the moving-wall term 6 w_m (u_w . c_m).  LBM_BB_SLOTS sends every wall
link through the solid's slot, written by the wall cell after its collision
with the same expression.  Bars:
  - FP64 and FP32 storage: bit for bit (same operations on the same values);
  - F16S/F16C: resting walls bit for bit; at a moving wall the slot layout
    rounds f* + 6 w (u_w . c) to 16 bits, the folded one rounds f* and adds
    the term in FP32, so each is held to its error against FP64 storage;
  - the folded cases also against the CPU oracle at FP64 (<= 1e-11 rel L2).
"""
import numpy as np
import pytest

import cases
import lbm_amd
from conftest import rel_l2_u

pytestmark = pytest.mark.gpu

ALL_FACES = 63
YZ_FACES = 4 | 8 | 16 | 32
X_KEEP = (1 | 2) << 8  # the inlet (i = 0) and outlet (i = Nx+1) faces


def _run(case, storage, bb, steps, ibm=None):
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage, bounce_back=bb)
    lat.upload(**case.fields())
    lat.init()
    faces = lat.bounce_back_faces()
    if ibm is not None:
        ibm(lat)
        out = lat.ibm_run(steps)
    else:
        lat.step(steps)
        out = None
    code, cell = lat.status()
    assert code == 0, f"status {code} at {cell}"
    m = lat.macro()
    lat.close()
    return faces, m, out


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y), f"max diff {np.abs(x - y).max():.3e}"


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_cavity_fold_equals_slots_bitwise(storage):
    case = cases.cavity(20, u_lid=0.05, Re=100.0)
    fa, a, _ = _run(case, storage, lbm_amd.BB_AUTO, 300)
    fs, b, _ = _run(case, storage, lbm_amd.BB_SLOTS, 300)
    assert fa == (ALL_FACES, 0)  # every face solid, no slot left
    assert fs == (0, 1)
    _same(a, b)


# measured on the MI355X, 20^3 cavity after 300 steps, rel L2 of u against
# FP64 storage: F16S fold 2.6e-3, slots 6.2e-3; F16C fold 1.13e-3, slots
# 1.31e-3 (the slot layout rounds the lid term into the 16-bit value)
@pytest.mark.parametrize("storage,tol_fold,tol_slot", [("f16s", 3.5e-3, 8e-3), ("f16c", 1.5e-3, 2e-3)])
def test_cavity_fold_vs_slots_16bit(storage, tol_fold, tol_slot):
    # at the moving lid the two layouts round in different places, so each
    # is measured against FP64 storage (which equals the oracle to 1e-11),
    # and the fold must be no less accurate than the slots
    case = cases.cavity(20, u_lid=0.05, Re=100.0)
    _, ref, _ = _run(case, "f64", lbm_amd.BB_AUTO, 300)
    _, a, _ = _run(case, storage, lbm_amd.BB_AUTO, 300)
    _, b, _ = _run(case, storage, lbm_amd.BB_SLOTS, 300)
    fl = case.type == cases.TYPE_F
    e_fold = rel_l2_u(a[1:], ref[1:], fl)
    e_slot = rel_l2_u(b[1:], ref[1:], fl)
    assert e_fold <= tol_fold and e_slot <= tol_slot, f"fold {e_fold:.3e}, slots {e_slot:.3e}"
    assert e_fold <= 1.1 * e_slot, f"fold {e_fold:.3e} less accurate than slots {e_slot:.3e}"


def test_cavity_resting_walls_16bit_bitwise():
    # no moving wall: the folded pull reads exactly what the slot held
    case = cases.cavity(16, u_lid=0.05, Re=100.0)
    case.ux[:, -1, :] = 0.0
    rng = np.random.default_rng(3)
    fl = case.type == cases.TYPE_F
    case.ux[fl] = 0.02 * rng.standard_normal(int(fl.sum()))
    case.uy[fl] = 0.02 * rng.standard_normal(int(fl.sum()))
    for st in ("f16s", "f16c"):
        _, a, _ = _run(case, st, lbm_amd.BB_AUTO, 200)
        _, b, _ = _run(case, st, lbm_amd.BB_SLOTS, 200)
        _same(a, b)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_channel_faces_fold_equals_slots_bitwise(storage):
    # inlet/outlet faces at i = 0, Nx+1 (face loops, frozen links), walls on
    # the j and k faces: the face loops re-read the folded pulled values
    case = cases.channel(30, 14, 10, perturb=0.01)
    fa, a, _ = _run(case, storage, lbm_amd.BB_AUTO, 200)
    _, b, _ = _run(case, storage, lbm_amd.BB_SLOTS, 200)
    assert fa == (YZ_FACES | X_KEEP, 0)
    _same(a, b)


def test_mixed_fold_and_slots_bitwise():
    # walls on every face plus solids inside: folded faces and slot links in
    # one lattice, with the moving lid
    case = cases.cavity(18, u_lid=0.05, Re=100.0)
    rng = np.random.default_rng(7)
    inner = np.zeros(case.type.shape, bool)
    inner[3:-3, 3:-3, 3:-3] = rng.random(inner[3:-3, 3:-3, 3:-3].shape) < 0.08
    case.type[inner] = cases.TYPE_S
    fa, a, _ = _run(case, "f32", lbm_amd.BB_AUTO, 200)
    _, b, _ = _run(case, "f32", lbm_amd.BB_SLOTS, 200)
    assert fa == (ALL_FACES, 1)
    _same(a, b)


def test_ibm_channel_fold_equals_slots_bitwise():
    # force mode 3: the IBM box moments pull the box cells next to the k
    # walls through the fold
    case = cases.channel(40, 24, 6, u_in=0.04, Re=100.0, D=8.0, perturb=0.01)
    mk = cases.cylinder_markers(12.0, 12.0, 8.0, 24, np.arange(1, case.Nz + 1, dtype=np.float64))

    def ibm(lat):
        lat.ibm_setup(mk, box=(4, 20, 4, 20, 1, case.Nz), m_max=5)

    fa, a, ra = _run(case, "f64", lbm_amd.BB_AUTO, 30, ibm=ibm)
    _, b, rb = _run(case, "f64", lbm_amd.BB_SLOTS, 30, ibm=ibm)
    assert fa[0] == YZ_FACES | X_KEEP
    _same(a, b)
    assert np.array_equal(np.asarray(ra), np.asarray(rb))


def test_momentum_exchange_fold_equals_slots():
    # MEA over a box that touches the moving lid and two walls
    case = cases.cavity(16, u_lid=0.05, Re=100.0)
    out = []
    for bb in (lbm_amd.BB_AUTO, lbm_amd.BB_SLOTS):
        lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage="f64", bounce_back=bb)
        lat.upload(**case.fields())
        lat.init()
        lat.step(50)
        out.append(np.asarray(lat.momentum_exchange((1, 6, 10, 16, 1, 16))))
        lat.close()
    assert np.array_equal(out[0], out[1])
    assert np.abs(out[0]).max() > 0


def test_folded_cavity_against_oracle_fp64(oracle_mod):
    case = cases.cavity(12, u_lid=0.05, Re=100.0)
    _, (rho, ux, uy, uz), _ = _run(case, "f64", lbm_amd.BB_AUTO, 120)
    ref = oracle_mod.run_case(case, 120)
    r_rho, r_ux, r_uy, r_uz = ref.macro()
    fl = case.type == cases.TYPE_F
    assert rel_l2_u((ux, uy, uz), (r_ux, r_uy, r_uz), fl) <= 1e-11
    assert np.abs(rho - r_rho)[fl].max() <= 1e-12


def test_bounce_back_mode_argument():
    with pytest.raises(lbm_amd.LBMError):
        lbm_amd.Lattice(4, 4, 4, 0.8, bounce_back=2)
    lat = lbm_amd.Lattice(4, 4, 4, 0.8)
    with pytest.raises(lbm_amd.LBMError):
        lat.bounce_back_faces()  # before lbm_init
    lat.close()
