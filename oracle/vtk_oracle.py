"""Restatement of the reference's OutputVTK (lbm.cpp:2290-2468) and a reader
of the file format.

TEST INFRASTRUCTURE ONLY: writes the .vtr bytes the reference would write
for a given lattice state (the oracle's fields), to check the device writer
against.  Cell data blocks cover i = 0..Nx+1, j = 0..Ny+1, k = 1..Nz with i
fastest; every block is a UInt32 byte count followed by Float32 values.
"""
from __future__ import annotations

import struct

import numpy as np


def _f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32)


def write_vtr(path, Nx, Ny, Nz, rho, ux, uy, uz, type_, *, u_scale=1.0, velocity_name=0,
              force=None, T=None, q=None, type_energy=None, T_mode=0, T_a=0.0, T_b=1.0,
              poiseuille_coords=False):
    """Arrays are indexed [i, j, k] over the padded lattice (Nx+2, Ny+2,
    Nz+2).  force = (Fx, Fy, Fz) or None (the IBM build writes it); T, q,
    type_energy for the ENERGY build."""
    zcomp = Nz
    ncell = (Nx + 2) * (Ny + 2) * zcomp

    def cells(a):  # [i, j, k=1..Nz] -> file order (k, j, i), i fastest
        return np.ascontiguousarray(np.transpose(np.asarray(a)[:, :, 1:Nz + 1], (2, 1, 0))).ravel()

    blocks = [("Density", 1, _f32(cells(rho)))]
    vname = {0: "Velocity", 1: "Velocity Ratio", 2: "Velocity (Dimensional)"}[velocity_name]
    # (float)u / u_scale: the float converts back to double (lbm.cpp:2451)
    vel = np.stack([(_f32(cells(u)).astype(np.float64) / u_scale).astype(np.float32)
                    for u in (ux, uy, uz)], axis=1).ravel()
    blocks.append((vname, 3, vel))
    if force is not None:
        blocks.append(("Force density", 3, np.stack([_f32(cells(f)) for f in force], axis=1).ravel()))
    if T is not None:
        Tc = cells(T)
        if T_mode == 1:
            tv = (_f32(Tc - T_a).astype(np.float64) / T_b).astype(np.float32)
        elif T_mode == 2:
            tv = (_f32(Tc).astype(np.float64) / T_a - 1).astype(np.float32)
        else:
            tv = _f32(Tc)
        blocks.append(("Temperature Ratio" if T_mode else "Temperature", 1, tv))
        blocks.append(("Heat flux", 3, np.stack([_f32(cells(x)) for x in q], axis=1).ravel()))
    blocks.append(("CellType", 1, cells(type_).astype(np.int32).astype(np.float32)))
    if T is not None:
        blocks.append(("EnergyCellType", 1, cells(type_energy).astype(np.int32).astype(np.float32)))

    H = Ny + 2
    cx = np.arange(Nx + 3, dtype=np.float64).astype(np.float32)
    if poiseuille_coords:
        cy = (np.arange(Ny + 3, dtype=np.float64).astype(np.float32).astype(np.float64) / H * 2 - 1).astype(np.float32)
        cz = (np.arange(zcomp + 1, dtype=np.float64).astype(np.float32).astype(np.float64) / H * 2 - 1).astype(np.float32)
    else:
        cy = np.arange(Ny + 3, dtype=np.float64).astype(np.float32)
        cz = np.arange(zcomp + 1, dtype=np.float64).astype(np.float32)

    out = []
    off = 0
    out.append('<?xml version="1.0"?>\n')
    out.append('<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian" header_type="UInt32">\n')
    out.append(f'  <RectilinearGrid WholeExtent="0 {Nx + 2} 0 {Ny + 2} 0 {zcomp}">\n')
    out.append(f'  <Piece Extent="0 {Nx + 2} 0 {Ny + 2} 0 {zcomp}">\n')
    out.append("    <PointData>\n")
    out.append("    </PointData>\n")
    out.append("    <CellData>\n")
    for name, nc, _ in blocks:
        comp = ' NumberOfComponents="3"' if nc == 3 else ""
        out.append(f'      <DataArray type="Float32" Name="{name}"{comp} format="appended" offset="{off}" />\n')
        off += 4 + nc * 4 * ncell
    out.append("    </CellData>\n")
    out.append("    <Coordinates>\n")
    for name, arr in (("CoordinateX", cx), ("CoordinateY", cy), ("CoordinateZ", cz)):
        out.append(f'      <DataArray type="Float32" Name="{name}" format="appended" offset="{off}" />\n')
        off += 4 + 4 * arr.size
    out.append("    </Coordinates>\n")
    out.append("  </Piece>\n")
    out.append("  </RectilinearGrid>\n")
    out.append('  <AppendedData encoding="raw">')
    out.append("_")
    data = bytearray("".join(out).encode())
    for _, nc, arr in blocks:
        data += struct.pack("<I", nc * 4 * ncell)
        data += np.asarray(arr, dtype="<f4").tobytes()
    for arr in (cx, cy, cz):
        data += struct.pack("<I", 4 * arr.size)
        data += np.asarray(arr, dtype="<f4").tobytes()
    data += b"  </AppendedData>\n</VTKFile>\n"
    with open(path, "wb") as fh:
        fh.write(bytes(data))


def read_vtr(path):
    """Parses a raw-appended VTK RectilinearGrid file: returns (header text,
    extents, {name: (ncomp, float32 array)}) after checking that every
    DataArray offset points at a UInt32 byte count whose block ends where
    the next one starts."""
    import re

    raw = open(path, "rb").read()
    mark = raw.index(b'<AppendedData encoding="raw">') + len(b'<AppendedData encoding="raw">')
    assert raw[mark:mark + 1] == b"_"
    header = raw[:mark].decode()
    base = mark + 1
    ext = tuple(int(v) for v in re.search(r'WholeExtent="([^"]+)"', header).group(1).split())
    arrays = {}
    entries = re.findall(r'<DataArray type="Float32" Name="([^"]+)"( NumberOfComponents="(\d)")? '
                         r'format="appended" offset="(\d+)" />', header)
    pos = base
    for name, _, nc, off in entries:
        off = int(off)
        assert base + off == pos, f"{name}: offset {off} is not where the previous block ended"
        (nbytes,) = struct.unpack("<I", raw[pos:pos + 4])
        arr = np.frombuffer(raw[pos + 4:pos + 4 + nbytes], dtype="<f4").copy()
        arrays[name] = (int(nc) if nc else 1, arr)
        pos += 4 + nbytes
    assert raw[pos:] == b"  </AppendedData>\n</VTKFile>\n"
    return header, ext, arrays
