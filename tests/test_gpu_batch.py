"""Ensembles (lbm_batch_*): one launch per step over many lattices must give
every lattice exactly the state that stepping it alone gives (bit-exact:
the same kernel body runs on the same data), for every launch shape, over
repeated graph replays with eager work and synchronisation between them."""
import numpy as np
import pytest

import cases
import lbm_amd

pytestmark = pytest.mark.gpu

SPECS = [(8, 0.05, 100.0), (12, 0.03, 400.0), (16, 0.07, 60.0), (10, 0.02, 900.0), (20, 0.05, 200.0)]


def _lattices(storage):
    out = []
    for N, u, Re in SPECS:
        case = cases.cavity(N, u_lid=u, Re=Re)
        lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
        lat.upload(**case.fields())
        lat.init()
        out.append(lat)
    return out


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s", "f16c"])
@pytest.mark.parametrize("block,per_cu", [(256, 0), (64, 2), (128, 8)])
def test_batch_matches_lattices_stepped_alone(storage, block, per_cu):
    alone = _lattices(storage)
    ens = _lattices(storage)
    batch = lbm_amd.Batch(ens)
    batch.set_launch(block, per_cu)
    for n in (3, 5, 1, 8, 2):  # graph pairs, odd remainders, single steps
        batch.step(n)
        for lat in alone:
            lat.step(n)
        for a, b in zip(alone, ens):
            assert a.sync() == 0 and b.sync() == 0
            for x, y in zip(a.macro(), b.macro()):
                assert np.array_equal(x, y)
    for a, b in zip(alone, ens):
        assert np.array_equal(a.populations(), b.populations())
    batch.close()
    for lat in alone + ens:
        lat.close()


def test_batch_then_single_lattice_calls():
    """After batched steps each context carries on alone (buffers and state
    handed back), and the reference-style calls still work."""
    alone = _lattices("f32")
    ens = _lattices("f32")
    batch = lbm_amd.Batch(ens)
    batch.step(7)
    for lat in alone:
        lat.step(7)
    for a, b in zip(alone, ens):
        a.collision(); a.streaming(); a.macroprop()
        b.collision(); b.streaming(); b.macroprop()
        for x, y in zip(a.macro(), b.macro()):
            assert np.array_equal(x, y)
    batch.close()


def test_batch_autotune_and_rejections():
    ens = _lattices("f32")
    twins = _lattices("f32")
    for lat in ens + twins:
        lat.step(3)
    f0 = [lat.populations() for lat in ens]
    batch = lbm_amd.Batch(ens)
    b, p, ms = batch.autotune(4)
    assert b in (64, 128, 256) and p in (0, 1, 2, 4, 8) and ms > 0
    assert batch.get_launch() == (b, p)
    # dry runs: the lattices are as before, and step on as their twins do
    for lat, f in zip(ens, f0):
        assert np.array_equal(lat.populations(), f)
    batch.step(5)
    for lat, twin in zip(ens, twins):
        twin.step(5)
        assert np.array_equal(lat.populations(), twin.populations())
    batch.close()
    mixed = _lattices("f32")[:2] + _lattices("f64")[:1]
    with pytest.raises(lbm_amd.LBMError) as e:
        lbm_amd.Batch(mixed)
    assert e.value.code == lbm_amd.LBM_ERR_UNSUPPORTED
    lat = ens[0]
    assert lat.autotune(4)[0] in (64, 128, 256)
