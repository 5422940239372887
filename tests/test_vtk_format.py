"""The .vtr format of the reference's OutputVTK (lbm.cpp:2290-2468), checked
on the CPU: the restated writer (oracle/vtk_oracle.py) produces a file whose
DataArray offsets chain exactly (each block is a UInt32 byte count plus its
data and ends where the next offset points), whose extents cover cells
i = 0..Nx+1, j = 0..Ny+1, k = 1..Nz, and whose blocks decode back to the
fields in that cell order (i fastest)."""
import numpy as np

import vtk_oracle


def _fields(Nx, Ny, Nz, seed=3):
    rng = np.random.default_rng(seed)
    sh = (Nx + 2, Ny + 2, Nz + 2)
    rho = 1 + 0.01 * rng.standard_normal(sh)
    u = [0.05 * rng.standard_normal(sh) for _ in range(3)]
    t = rng.integers(0, 4, sh).astype(np.int16)
    return rho, u, t


def test_vtr_offsets_extents_and_cell_order(tmp_path):
    Nx, Ny, Nz = 5, 4, 3
    rho, u, t = _fields(Nx, Ny, Nz)
    F = [np.full(rho.shape, v) for v in (1.0, -2.0, 3.0)]
    p = tmp_path / "a.vtr"
    vtk_oracle.write_vtr(p, Nx, Ny, Nz, rho, *u, t, u_scale=0.05, force=F)
    header, ext, arrays = vtk_oracle.read_vtr(p)
    assert ext == (0, Nx + 2, 0, Ny + 2, 0, Nz)
    assert list(arrays) == ["Density", "Velocity", "Force density", "CellType", "CoordinateX",
                            "CoordinateY", "CoordinateZ"]
    n = (Nx + 2) * (Ny + 2) * Nz
    nc, d = arrays["Density"]
    assert nc == 1 and d.size == n
    # cell (i, j, k) is entry ((k-1)(Ny+2) + j)(Nx+2) + i
    for (i, j, k) in [(0, 0, 1), (Nx + 1, 2, 1), (3, Ny + 1, Nz)]:
        e = ((k - 1) * (Ny + 2) + j) * (Nx + 2) + i
        assert d[e] == np.float32(rho[i, j, k])
        assert arrays["CellType"][1][e] == float(t[i, j, k])
        v = arrays["Velocity"][1].reshape(n, 3)[e]
        assert v[1] == np.float32(np.float64(np.float32(u[1][i, j, k])) / 0.05)
    assert np.all(arrays["Force density"][1].reshape(n, 3) == np.float32([1, -2, 3]))
    assert np.array_equal(arrays["CoordinateX"][1], np.arange(Nx + 3, dtype=np.float32))
    assert np.array_equal(arrays["CoordinateZ"][1], np.arange(Nz + 1, dtype=np.float32))


def test_vtr_thermal_blocks_and_poiseuille_coordinates(tmp_path):
    Nx, Ny, Nz = 3, 6, 2
    rho, u, t = _fields(Nx, Ny, Nz, seed=5)
    T = 0.5 + 0.1 * np.random.default_rng(1).random(rho.shape)
    q = [np.zeros(rho.shape), np.ones(rho.shape), -np.ones(rho.shape)]
    te = np.where(t == 1, 6, 0).astype(np.int16)
    p = tmp_path / "b.vtr"
    vtk_oracle.write_vtr(p, Nx, Ny, Nz, rho, *u, t, T=T, q=q, type_energy=te, T_mode=1, T_a=0.5,
                         T_b=0.1, poiseuille_coords=True, velocity_name=1)
    header, ext, arrays = vtk_oracle.read_vtr(p)
    assert list(arrays)[:6] == ["Density", "Velocity Ratio", "Temperature Ratio", "Heat flux",
                                "CellType", "EnergyCellType"]
    n = (Nx + 2) * (Ny + 2) * Nz
    e = ((1 - 1) * (Ny + 2) + 2) * (Nx + 2) + 1
    assert arrays["Temperature Ratio"][1][e] == np.float32(np.float64(np.float32(T[1, 2, 1] - 0.5)) / 0.1)
    assert np.all(arrays["Heat flux"][1].reshape(n, 3)[:, 1] == 1.0)
    cy = arrays["CoordinateY"][1]
    assert cy[0] == -1.0 and cy[Ny + 2] == np.float32(1.0) and cy.size == Ny + 3
