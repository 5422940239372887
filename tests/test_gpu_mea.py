"""GPU momentum exchange (lbm_momentum_exchange: MEA_CD / MEA_CL,
lbm.cpp:3085-3163) against the CPU oracle's restatement, through the C-ABI.

The oracle's version is pinned by the rest-fluid wall pressure and the
moving-lid closed form in tests/test_oracle_pins.py.  Bars: FP64 storage --
each component within 1e-12 of the largest component's magnitude (the
device sums per block in a tree, the oracle serially); FP32 storage -- the
bound stated in the test.
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


def _close(P, R, rtol=1e-12):
    scale = max(np.abs(R).max(), 1e-300)
    assert np.abs(P - R).max() <= rtol * scale, f"{P} vs {R}"


CASES = {
    "cylinder_2d": lambda: cases.cylinder_bb(60, 40, 1, D=10.0, u_max=0.05, Re=40.0),
    "cylinder_3d_periodic_z": lambda: cases.cylinder_bb(40, 30, 6, D=8.0, u_max=0.05, Re=30.0),
    "sphere": lambda: cases.sphere_bb(40, 24, 24, D=8.0, u_max=0.05, Re=30.0),
    "cavity": lambda: cases.cavity(10, u_lid=0.05, Re=50.0),
    "random": lambda: cases.random_mask_case(8, 7, 6, seed=21),
}


def _full(case):
    return (0, case.Nx + 1, 0, case.Ny + 1, 0, case.Nz + 1)


@pytest.mark.parametrize("name", sorted(CASES))
def test_mea_fused_steps_fp64(oracle_mod, name):
    """After lbm_step (fused launches) the device sums equal the oracle's
    after the same number of Collision/Streaming/MacroProp steps, over the
    reference's box and over the whole lattice."""
    case = CASES[name]()
    steps = 30
    ref = oracle_mod.run_case(case, steps)
    lat = _gpu(case)
    lat.step(steps)
    assert lat.sync() == 0
    boxes = [_full(case)]
    if "mea_box" in case.params:
        boxes.append(case.params["mea_box"])
    for box in boxes:
        _close(lat.momentum_exchange(box), ref.mea(box))
    lat.close()


def test_mea_reference_calls_and_moving_lid_fp64(oracle_mod):
    """Collision -> Streaming -> MacroProp as separate calls on a cavity
    with a moving lid (wall velocity on the links): equal after each step."""
    case = cases.cavity(8, u_lid=0.08, Re=40.0)
    ref = oracle_mod.run_case(case, 0)
    lat = _gpu(case)
    box = _full(case)
    for _ in range(6):
        assert ref.collision() == 0
        ref.streaming()
        assert ref.macroprop() == 0
        lat.collision()
        lat.streaming()
        lat.macroprop()
        _close(lat.momentum_exchange(box), ref.mea(box))
    lat.close()


def test_mea_on_d3q7_context(oracle_mod):
    """The flow populations of a D3Q19 + D3Q7 context (26-plane buffers)."""
    case, th = cases.q7_random_case(7, 6, 5, seed=9)
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
    _close(lat.momentum_exchange(_full(case)), o.mea(_full(case)))
    lat.close()


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f16c", 2e-3)])
def test_mea_reduced_storage(storage, tol):
    """FP32 and F16C storage against FP64 storage on the device after 200
    steps of the 2D cylinder (drag within the stated relative bound)."""
    case = CASES["cylinder_2d"]()
    box = case.params["mea_box"]
    res = {}
    for st in ("f64", storage):
        lat = _gpu(case, st)
        lat.step(200)
        assert lat.sync() == 0
        res[st] = lat.momentum_exchange(box)
        lat.close()
    assert abs(res[storage][0] - res["f64"][0]) <= tol * abs(res["f64"][0])
    assert abs(res[storage][1] - res["f64"][1]) <= tol * abs(res["f64"][0])


def test_mea_state_and_box_checks():
    case = CASES["cavity"]()
    lat = _gpu(case)
    box = _full(case)
    with pytest.raises(lbm_amd.LBMError) as e:  # after lbm_init: no streaming yet
        lat.momentum_exchange(box)
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    lat.collision()
    with pytest.raises(lbm_amd.LBMError) as e:  # f* in the buffer
        lat.momentum_exchange(box)
    assert e.value.code == lbm_amd.LBM_ERR_STATE
    lat.streaming()
    lat.momentum_exchange(box)
    for bad in [(0, case.Nx + 2, 0, 1, 0, 1), (3, 2, 0, 1, 0, 1), (-1, 1, 0, 1, 0, 1)]:
        with pytest.raises(lbm_amd.LBMError) as e:
            lat.momentum_exchange(bad)
        assert e.value.code == lbm_amd.LBM_ERR_ARG
    lat.close()
