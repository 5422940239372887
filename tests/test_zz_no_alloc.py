"""lbm_reduce_fp64_noalloc (no device allocation: partial sums in the
library's fixed device array) against lbm_reduce_fp64 through the C-ABI, bit
for bit: two calls in a row, the second on a smaller lattice (fewer partial
blocks: nothing of the first call's partials may leak into it), for each
quantity and storage."""
import ctypes as C

import numpy as np
import pytest

import cases
import lbm_amd

pytestmark = pytest.mark.gpu


def _lat(N, storage, steps):
    case = cases.taylor_green(N, u0=0.04, Re=200.0)
    lat = lbm_amd.Lattice(case.Nx, case.Ny, case.Nz, case.tau, storage=storage)
    lat.upload(**case.fields())
    lat.init()
    lat.step(steps)
    assert lat.sync() == 0
    return lat


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_noalloc_reduction_matches_the_allocating_one(storage):
    big = _lat(96, storage, 7)     # 1,568 partial blocks
    small = _lat(20, storage, 3)   # 4 partial blocks
    for what in (lbm_amd.REDUCE_MASS, lbm_amd.REDUCE_MOMENTUM, lbm_amd.REDUCE_KINETIC):
        for lat in (big, small, big, small):
            a = lat.reduce(what)
            b = lat.reduce_noalloc(what)
            assert np.array_equal(np.atleast_1d(a), np.atleast_1d(b)), (storage, what)
    big.close()
    small.close()


def test_noalloc_reduction_argument_errors():
    lib = lbm_amd.load_library()
    out = np.zeros(3)
    p = out.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.lbm_reduce_fp64_noalloc(None, 0, p) == lbm_amd.LBM_ERR_ARG
    lat = _lat(12, "f32", 1)
    assert lib.lbm_reduce_fp64_noalloc(lat._h, 3, p) == lbm_amd.LBM_ERR_ARG
    assert lib.lbm_reduce_fp64_noalloc(lat._h, 0, None) == lbm_amd.LBM_ERR_ARG
    lat.close()
