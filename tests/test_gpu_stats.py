"""GPU MORE_DETAILS fields (lbm_download_stress, lbm_stats_update,
lbm_download_stats: lbm.cpp:681-705, 865-939) against the CPU oracle's
restatement, through the C-ABI.

The oracle's versions are pinned by the isotropic stress of an equilibrium,
the Couette shear stress, and the steady-field and linear-signal closed
forms of the statistics (tests/test_oracle_pins.py).  Bars: FP64 storage --
stress within 1e-12 of its largest magnitude over TYPE_F cells, statistics
within 1e-11 (relative to each field's largest magnitude); FP32 storage --
the bound stated in the test.
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


def _rel(a, b, m):
    scale = max(float(np.abs(b[m]).max()), 1e-300)
    return float(np.abs(a - b)[m].max()) / scale


CASES = {
    "cavity": lambda: cases.cavity(10, u_lid=0.08, Re=50.0),
    "random": lambda: cases.random_mask_case(8, 7, 6, seed=4),
    "tgv": lambda: cases.taylor_green(16, u0=0.05, Re=100.0),
    "cylinder": lambda: cases.cylinder_bb(40, 24, 1, D=8.0, u_max=0.05, Re=30.0),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_stress_fp64(oracle_mod, name):
    case = CASES[name]()
    steps = 25
    ref = oracle_mod.run_case(case, steps)
    lat = _gpu(case)
    lat.step(steps)
    assert lat.sync() == 0
    sig, p = lat.stress()
    r_sig, r_p = ref.stress()
    fl = case.type == cases.TYPE_F
    assert _rel(sig, r_sig, fl) < 1e-12
    assert np.abs(p - r_p)[fl].max() < 1e-15
    assert np.array_equal(sig, np.swapaxes(sig, -1, -2))
    lat.close()


@pytest.mark.parametrize("name", ["cavity", "cylinder"])
def test_stats_every_nt_steps_fp64(oracle_mod, name):
    """The reference's loop with MORE_DETAILS: Collision, Streaming,
    MacroProp, and Sup_MacroProp(step) every Nt steps; every non-solid cell
    (the cylinder's inlet and outlet cells included)."""
    case = CASES[name]()
    nt, steps = 5, 40
    ref = oracle_mod.run_case(case, 0)
    lat = _gpu(case)
    for step in range(nt, steps + 1, nt):
        assert ref.run(nt) == 0
        ref.stats_update(step, nt)
        lat.step(nt)
        lat.stats_update(step, nt)
    assert lat.sync() == 0
    fl = case.type != cases.TYPE_S
    for got, want, name in zip(lat.stats(), ref.stats(), ("avg_u", "rms_u", "avg_p", "rms_p", "avg_rs", "rms_rs")):
        assert _rel(got, want, fl) < 1e-11, name
    lat.close()


def test_stats_through_reference_calls_fp64(oracle_mod):
    """Collision -> Streaming -> MacroProp as separate calls, Nt = 1."""
    case = CASES["random"]()
    ref = oracle_mod.run_case(case, 0)
    lat = _gpu(case)
    for step in range(1, 9):
        assert ref.collision() == 0
        ref.streaming()
        assert ref.macroprop() == 0
        ref.stats_update(step, 1)
        lat.collision()
        lat.streaming()
        lat.macroprop()
        lat.stats_update(step, 1)
    fl = case.type == cases.TYPE_F
    for got, want in zip(lat.stats(), ref.stats()):
        assert _rel(got, want, fl) < 1e-11
    sig, _ = lat.stress()
    assert _rel(sig, ref.stress()[0], fl) < 1e-12
    lat.close()


def test_stress_fp32_against_fp64():
    case = CASES["tgv"]()
    out = {}
    for st in ("f64", "f32"):
        lat = _gpu(case, st)
        lat.step(100)
        assert lat.sync() == 0
        out[st] = lat.stress()[0]
        lat.close()
    fl = case.type == cases.TYPE_F
    # the deviatoric part: sigma minus its isotropic equilibrium part
    iso = lambda s: s - np.trace(s, axis1=-2, axis2=-1)[..., None, None] / 3 * np.eye(3)  # noqa: E731
    assert _rel(iso(out["f32"]), iso(out["f64"]), fl) < 2e-3
    assert _rel(out["f32"], out["f64"], fl) < 1e-6


def test_stats_state_checks():
    case = CASES["cavity"]()
    lat = _gpu(case)
    with pytest.raises(lbm_amd.LBMError) as e:
        lat.stats()
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    with pytest.raises(lbm_amd.LBMError) as e:
        lat.stats_update(3, 5)  # n < Nt
    assert e.value.code == lbm_amd.LBM_ERR_ARG
    lat.collision()
    with pytest.raises(lbm_amd.LBMError) as e:
        lat.stress()
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    lat.close()
