"""GPU flexible-filament structure solver (lbm_filament_*) against the CPU
oracle's restatement, through the C-ABI.

The device follows the oracle's operations in the same order with FP64
contraction off: positions and tensions must agree bitwise.  The closed
forms that pin the oracle (tests/test_oracle_pins.py) are re-checked on the
device.
"""
import math

import numpy as np
import pytest

import lbm_amd

pytestmark = pytest.mark.gpu

ARGS = dict(n=41, length=40.0, line_density=1.5, bending=3.0, damping=0.01, body=(0.0, -2e-5, 1e-5),
            clamp=(10.0, 20.0, 5.0), tangent=(0.6, 0.8, 0.0))


def test_filament_bitwise_vs_oracle(oracle_mod):
    rng = np.random.default_rng(17)
    dev = lbm_amd.Filament(**ARGS)
    ref = oracle_mod.Filament(**ARGS)
    assert dev.substeps == ref.substeps
    for call in range(60):
        F = 2e-3 * rng.standard_normal((ARGS["n"], 3)) * math.sin(0.1 * call)
        dev.step(F, nsteps=1)
        ref.step(F, nsteps=1)
    dev.step(None, nsteps=20)
    ref.step(None, nsteps=20)
    xd, xpd, td = dev.get()
    xr, xpr, tr = ref.get()
    for a, b, what in ((xd, xr, "x"), (xpd, xpr, "x_prev"), (td, tr, "T")):
        assert np.array_equal(a, b), f"{what}: max diff {np.abs(a - b).max():.3e}"
    # the field moved, and the segments kept their length
    assert np.abs(xd - (np.array(ARGS["clamp"]) + np.outer(np.arange(41.0), ARGS["tangent"]))).max() > 1e-3
    assert np.abs(np.linalg.norm(np.diff(xd, axis=0), axis=1) - 1.0).max() < 1e-12
    dev.close()


def test_filament_set_and_multistep_call(oracle_mod):
    """A set curved state and one call of many steps equal the oracle's."""
    n, L = 33, 24.0
    args = dict(n=n, length=L, line_density=0.8, bending=0.5, damping=0.0, clamp=(0.0, 0.0, 0.0),
                tangent=(1.0, 0.0, 0.0))
    ds = L / (n - 1)
    s = np.arange(n) * ds
    R = 30.0
    x = np.stack([R * np.sin(s / R), R * (1 - np.cos(s / R)), 0.1 * np.sin(s / 7)], axis=1)
    dev, ref = lbm_amd.Filament(**args), oracle_mod.Filament(**args)
    dev.set(x)
    ref.set(x)
    dev.step(None, nsteps=400)
    ref.step(None, nsteps=400)
    assert np.array_equal(dev.get()[0], ref.get()[0])
    dev.close()


def test_filament_cantilever_on_device():
    """Static deflection qL^4/(8EI) and the first mode 1.8751^2 sqrt(EI/(m_l L^4))."""
    L, EI, ml = 40.0, 3.0, 1.5
    w1 = 1.8751040687 ** 2 * math.sqrt(EI / (ml * L ** 4))
    q = 0.08 * EI / L ** 3
    f = lbm_amd.Filament(41, L, ml, EI, damping=2 * ml * w1, clamp=(10.0, 20.0, 5.0))
    F = np.zeros((41, 3))
    F[:, 1] = -q
    f.step(F, nsteps=24000)
    x = f.get()[0]
    assert abs((20.0 - x[-1, 1]) / (q * L ** 4 / (8 * EI)) - 1) < 2e-3
    g = lbm_amd.Filament(41, L, ml, EI, clamp=(10.0, 20.0, 5.0))
    g.set(x)
    period = 2 * math.pi / w1
    tip = []
    for _ in range(int(3 * period) // 10):
        g.step(None, nsteps=10)
        tip.append(g.get()[0][-1, 1] - 20.0)
    tip = np.array(tip)
    z = np.flatnonzero((tip[:-1] < 0) & (tip[1:] >= 0))
    t = 10.0 * (z + (-tip[z]) / (tip[z + 1] - tip[z]))
    assert abs(2 * math.pi / np.mean(np.diff(t)) / w1 - 1) < 1e-2
    f.close()
    g.close()


def test_filament_argument_checks():
    with pytest.raises(lbm_amd.LBMError) as e:
        lbm_amd.Filament(3, 10.0, 1.0, 1.0)
    assert e.value.code == lbm_amd.LBM_ERR_ARG
    with pytest.raises(lbm_amd.LBMError) as e:
        lbm_amd.Filament(300, 10.0, 1.0, 1.0)
    assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
    with pytest.raises(lbm_amd.LBMError) as e:
        lbm_amd.Filament(10, 10.0, 1.0, 1.0, tangent=(1.0, 1.0, 0.0))
    assert e.value.code == lbm_amd.LBM_ERR_ARG
