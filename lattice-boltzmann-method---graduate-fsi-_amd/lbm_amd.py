"""ctypes binding of liblbm_hip.so (the C-ABI declared in include/lbm_hip.h).

This is plumbing for Python callers (tests, the benchmark): every call goes
straight to the gfx950 library.  There is no fallback: if the library is not
built or cannot be loaded, importing this module's Lattice raises.

Array convention: numpy arrays of shape (Nx+2, Ny+2, Nz+2) in C order, i.e.
the reference's LBM::fluid[i][j][k] indexing with k fastest (lbm.cpp:17-28).
Populations are (Nx+2, Ny+2, Nz+2, 19) in the reference's D3Q19 order.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_hip.so")

LBM_OK = 0
LBM_ERR_FEQ_NEGATIVE = 1
LBM_ERR_DENSITY_RANGE = 2
LBM_ERR_TEMPERATURE_RANGE = 3
LBM_ERR_ARG = -1
LBM_ERR_STATE = -2
LBM_ERR_HIP = -3
LBM_ERR_NOMEM = -4
LBM_ERR_UNSUPPORTED = -5

TYPE_F, TYPE_S, TYPE_OUT, TYPE_IN, TYPE_P, TYPE_L, TYPE_T, TYPE_Q = range(8)

STORE_F64, STORE_F32, STORE_F16S, STORE_F16C = 0, 1, 2, 3
STORAGE_NAMES = {"f64": STORE_F64, "f32": STORE_F32, "f16s": STORE_F16S, "f16c": STORE_F16C}

REDUCE_MASS, REDUCE_MOMENTUM, REDUCE_KINETIC = 0, 1, 2
# lbm_desc.bounce_back: solid shell faces folded into the pull address, or
# every wall link through the solid's slot
BB_AUTO, BB_SLOTS = 0, 1


class LBMError(RuntimeError):
    def __init__(self, code: int, what: str, msg: str):
        super().__init__(f"{what}: {msg} (code {code})")
        self.code = code


class _Desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("lattice", C.c_int32),
        ("storage", C.c_int32),
        ("device", C.c_int32),
        ("tau", C.c_double),
        ("bounce_back", C.c_int32),
        ("reserved", C.c_int32 * 7),
    ]


class _IbmDesc(C.Structure):
    _fields_ = [
        ("nmarkers", C.c_int32),
        ("m_max", C.c_int32),
        ("box", C.c_int32 * 6),
        ("moving", C.c_int32),
        ("follow", C.c_int32),
        ("vsolid", C.c_double),
        ("center", C.c_double * 2),
        ("half", C.c_double),
        ("scheme", C.c_int32),
    ]


class _ThermalDesc(C.Structure):
    _fields_ = [
        ("tau_g", C.c_double),
        ("nu", C.c_double),
        ("cp", C.c_double),
        ("j_dot_n", C.c_double),
        ("g_beta", C.c_double),
        ("T_ref", C.c_double),
        ("T0", C.c_double),
        ("boussinesq", C.c_int32),
        ("model", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


THERMAL_ENERGY, THERMAL_D3Q7 = 0, 1


class _FilamentDesc(C.Structure):
    _fields_ = [
        ("nnodes", C.c_int32),
        ("substeps", C.c_int32),
        ("length", C.c_double),
        ("line_density", C.c_double),
        ("bending", C.c_double),
        ("damping", C.c_double),
        ("body_force", C.c_double * 3),
        ("clamp", C.c_double * 3),
        ("tangent", C.c_double * 3),
    ]


class _VtkDesc(C.Structure):
    _fields_ = [
        ("u_scale", C.c_double),
        ("T_a", C.c_double),
        ("T_b", C.c_double),
        ("velocity_name", C.c_int32),
        ("T_mode", C.c_int32),
        ("force", C.c_int32),
        ("thermal", C.c_int32),
        ("poiseuille_coords", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


def vtk_desc(u_scale=1.0, T_a=0.0, T_b=1.0, velocity_name=0, T_mode=0, force=False, thermal=False,
             poiseuille_coords=False):
    return _VtkDesc(float(u_scale), float(T_a), float(T_b), int(velocity_name), int(T_mode),
                    int(bool(force)), int(bool(thermal)), int(bool(poiseuille_coords)),
                    (C.c_int32 * 3)())


# name -> (restype, argtypes) for every entry point of include/lbm_hip.h
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_SIGNATURES = {
    "lbm_create": (C.c_int, [C.POINTER(_Desc), C.POINTER(_P)]),
    "lbm_destroy": (C.c_int, [_P]),
    "lbm_upload_cells": (C.c_int, [_P, C.POINTER(C.c_int16), _D, _D, _D, _D, _D, _D, _D]),
    "lbm_init": (C.c_int, [_P]),
    "lbm_collide": (C.c_int, [_P]),
    "lbm_stream": (C.c_int, [_P]),
    "lbm_macro": (C.c_int, [_P]),
    "lbm_step": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "lbm_sync": (C.c_int, [_P]),
    "lbm_status": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "lbm_download_macro": (C.c_int, [_P, _D, _D, _D, _D]),
    "lbm_download_f": (C.c_int, [_P, _D]),
    "lbm_download_force": (C.c_int, [_P, _D, _D, _D]),
    "lbm_reduce_fp64": (C.c_int, [_P, C.c_int32, _D]),
    "lbm_export_links": (
        C.c_int,
        [_P, C.POINTER(C.c_int16), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)],
    ),
    "lbm_ncells": (C.c_int64, [_P]),
    "lbm_nfluid": (C.c_int64, [_P]),
    "lbm_storage": (C.c_int32, [_P]),
    "lbm_time_steps": (C.c_int, [_P, C.c_int32, _D]),
    "lbm_bytes_per_cell": (C.c_double, [_P]),
    "lbm_set_launch": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "lbm_get_launch": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "lbm_bounce_back_faces": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "lbm_error_string": (C.c_char_p, [C.c_int]),
    "lbm_copy_probe": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, _D]),
    "lbm_storage_roundtrip": (C.c_int, [C.c_int32, C.c_int32, _D, _D, C.c_int64]),
    "lbm_ibm_setup": (C.c_int, [_P, C.POINTER(_IbmDesc), _D, _D, _D, _D, _D, _D, _D]),
    "lbm_ibm_apply": (C.c_int, [_P, _D, _D, _D]),
    "lbm_ibm_run": (C.c_int, [_P, C.c_int32, _D, _D]),
    "lbm_ibm_markers": (C.c_int, [_P, _D, _D, _D, _D, _D]),
    "lbm_ibm_box": (C.c_int, [_P, C.POINTER(C.c_int32), _D]),
    "lbm_upload_thermal": (C.c_int, [_P, C.POINTER(_ThermalDesc), _D, C.POINTER(C.c_int16)]),
    "lbm_download_thermal": (C.c_int, [_P, _D, _D, _D, _D]),
    "lbm_download_g": (C.c_int, [_P, _D]),
    "lbm_output_vtk_begin": (C.c_int, [_P, C.POINTER(_VtkDesc)]),
    "lbm_output_vtk_end": (C.c_int, [_P, C.c_char_p]),
    "lbm_output_vtk": (C.c_int, [_P, C.POINTER(_VtkDesc), C.c_char_p]),
    "lbm_batch_create": (C.c_int, [C.POINTER(_P), C.c_int32, C.POINTER(_P)]),
    "lbm_batch_destroy": (C.c_int, [_P]),
    "lbm_batch_step": (C.c_int, [_P, C.c_int32]),
    "lbm_batch_time_steps": (C.c_int, [_P, C.c_int32, _D]),
    "lbm_batch_set_launch": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "lbm_batch_get_launch": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "lbm_batch_nfluid": (C.c_int64, [_P]),
    "lbm_batch_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _D]),
    "lbm_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "lbm_reduce_fp64_noalloc": (C.c_int, [_P, C.c_int32, _D]),
    "lbm_checkpoint_size": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "lbm_checkpoint_save": (C.c_int, [_P, C.c_void_p, C.c_uint64]),
    "lbm_checkpoint_load": (C.c_int, [_P, C.c_void_p, C.c_uint64]),
    "lbm_fsi_setup": (C.c_int, [_P, _P, C.POINTER(_IbmDesc)]),
    "lbm_fsi_run": (C.c_int, [_P, C.c_int32, _D]),
    "lbm_momentum_exchange": (C.c_int, [_P, C.POINTER(C.c_int32), _D]),
    "lbm_download_stress": (C.c_int, [_P, _D, _D]),
    "lbm_stats_update": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "lbm_download_stats": (C.c_int, [_P, _D, _D, _D, _D, _D, _D]),
    "lbm_filament_create": (C.c_int, [C.c_int32, C.POINTER(_FilamentDesc), C.POINTER(_P)]),
    "lbm_filament_destroy": (C.c_int, [_P]),
    "lbm_filament_substeps": (C.c_int, [_P]),
    "lbm_filament_set": (C.c_int, [_P, _D, _D]),
    "lbm_filament_get": (C.c_int, [_P, _D, _D, _D]),
    "lbm_filament_step": (C.c_int, [_P, _D, C.c_double, C.c_int32]),
}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load liblbm_hip.so and bind every entry point.  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run __graft_entry__.build() (hipcc, gfx950) first"
        )
    lib = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _check(code: int, what: str) -> int:
    if code != LBM_OK:
        msg = load_library().lbm_error_string(code).decode()
        raise LBMError(code, what, msg)
    return code


def copy_probe(nbytes: int = 1 << 30, reps: int = 20, device: int = 0) -> float:
    """GB/s of a plain 16-byte-per-lane device copy of nbytes (read + write)."""
    g = C.c_double(0.0)
    _check(load_library().lbm_copy_probe(device, int(nbytes), int(reps), C.byref(g)), "lbm_copy_probe")
    return g.value


def storage_roundtrip(values, storage, device: int = 0) -> np.ndarray:
    """What population storage `storage` keeps of each value: encoded on the
    device and decoded back (values are f - w_i for the shifted formats)."""
    x = np.ascontiguousarray(values, dtype=np.float64).ravel()
    out = np.empty_like(x)
    st = STORAGE_NAMES[storage] if isinstance(storage, str) else int(storage)
    _check(load_library().lbm_storage_roundtrip(device, st, x.ctypes.data_as(_D), out.ctypes.data_as(_D),
                                                x.size), "lbm_storage_roundtrip")
    return out.reshape(np.shape(values))


class Filament:
    """A flexible filament on the device (lbm_filament_*)."""

    def __init__(self, n, length, line_density, bending, damping=0.0, body=(0.0, 0.0, 0.0),
                 clamp=(0.0, 0.0, 0.0), tangent=(1.0, 0.0, 0.0), substeps=0, device=0):
        self._lib = load_library()
        d = _FilamentDesc(int(n), int(substeps), float(length), float(line_density), float(bending),
                          float(damping), (C.c_double * 3)(*body), (C.c_double * 3)(*clamp),
                          (C.c_double * 3)(*tangent))
        h = _P()
        _check(self._lib.lbm_filament_create(int(device), C.byref(d), C.byref(h)), "lbm_filament_create")
        self._h = h
        self.n = int(n)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lbm_filament_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def substeps(self):
        return self._lib.lbm_filament_substeps(self._h)

    def set(self, x_now, x_prev=None):
        a = np.ascontiguousarray(x_now, dtype=np.float64).reshape(self.n, 3)
        b = None if x_prev is None else np.ascontiguousarray(x_prev, dtype=np.float64).reshape(self.n, 3)
        _check(self._lib.lbm_filament_set(self._h, _dp(a), _dp(b)), "lbm_filament_set")

    def get(self):
        x, xp, t = np.empty((self.n, 3)), np.empty((self.n, 3)), np.empty(self.n - 1)
        _check(self._lib.lbm_filament_get(self._h, _dp(x), _dp(xp), _dp(t)), "lbm_filament_get")
        return x, xp, t

    def step(self, force=None, dt=1.0, nsteps=1):
        F = None if force is None else np.ascontiguousarray(force, dtype=np.float64).reshape(self.n, 3)
        _check(self._lib.lbm_filament_step(self._h, _dp(F), float(dt), int(nsteps)), "lbm_filament_step")


class Lattice:
    """One device lattice, the counterpart of the reference's LBM object.

    Method names follow LBM (LBM.h:620-664): init / collision / streaming /
    macroprop, plus step(n) (the fused fast path) and downloads.
    """

    def __init__(self, Nx: int, Ny: int, Nz: int, tau: float, storage="f64", device=0,
                 bounce_back=BB_AUTO):
        lib = load_library()
        if isinstance(storage, str):
            storage = STORAGE_NAMES[storage]
        d = _Desc(Nx, Ny, Nz, 19, storage, device, float(tau), int(bounce_back))
        h = _P()
        _check(lib.lbm_create(C.byref(d), C.byref(h)), "lbm_create")
        self._lib = lib
        self._h = h
        self.Nx, self.Ny, self.Nz, self.tau = Nx, Ny, Nz, float(tau)
        self.storage = storage
        self.shape = (Nx + 2, Ny + 2, Nz + 2)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- set-up ---------------------------------------------------------------
    def _f64(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), self.shape))
        return a

    def upload(self, type=None, rho=None, ux=None, uy=None, uz=None, Fx=None, Fy=None, Fz=None):
        t = None
        if type is not None:
            t = np.ascontiguousarray(np.broadcast_to(np.asarray(type, dtype=np.int16), self.shape))
        arrs = [self._f64(a) for a in (rho, ux, uy, uz, Fx, Fy, Fz)]
        tp = None if t is None else t.ctypes.data_as(C.POINTER(C.c_int16))
        _check(self._lib.lbm_upload_cells(self._h, tp, *[_dp(a) for a in arrs]), "lbm_upload_cells")

    # -- the reference's methods ------------------------------------------------
    def init(self):
        _check(self._lib.lbm_init(self._h), "lbm_init")

    def collision(self):
        _check(self._lib.lbm_collide(self._h), "lbm_collide")

    def streaming(self):
        _check(self._lib.lbm_stream(self._h), "lbm_stream")

    def macroprop(self):
        _check(self._lib.lbm_macro(self._h), "lbm_macro")

    def step(self, n: int = 1):
        _check(self._lib.lbm_step(self._h, int(n), 1), "lbm_step")

    def sync(self) -> int:
        return self._lib.lbm_sync(self._h)

    def status(self):
        fb = C.c_int64(-1)
        code = self._lib.lbm_status(self._h, C.byref(fb))
        return code, fb.value

    def time_steps(self, n: int) -> float:
        ms = C.c_double(0.0)
        _check(self._lib.lbm_time_steps(self._h, int(n), C.byref(ms)), "lbm_time_steps")
        return ms.value

    # -- downloads -------------------------------------------------------------------
    def macro(self):
        out = [np.empty(self.shape, np.float64) for _ in range(4)]
        _check(self._lib.lbm_download_macro(self._h, *[_dp(a) for a in out]), "lbm_download_macro")
        return tuple(out)

    def force(self):
        out = [np.empty(self.shape, np.float64) for _ in range(3)]
        _check(self._lib.lbm_download_force(self._h, *[_dp(a) for a in out]), "lbm_download_force")
        return tuple(out)

    def populations(self):
        f = np.empty(self.shape + (19,), np.float64)
        _check(self._lib.lbm_download_f(self._h, _dp(f)), "lbm_download_f")
        return f

    def links(self):
        t = np.empty(self.shape, np.int16)
        src = np.empty(self.shape + (19,), np.int32)
        kind = np.empty(self.shape + (19,), np.uint8)
        _check(
            self._lib.lbm_export_links(
                self._h,
                t.ctypes.data_as(C.POINTER(C.c_int16)),
                src.ctypes.data_as(C.POINTER(C.c_int32)),
                kind.ctypes.data_as(C.POINTER(C.c_uint8)),
            ),
            "lbm_export_links",
        )
        return t, src, kind

    def reduce(self, what: int):
        out = np.zeros(3, np.float64)
        _check(self._lib.lbm_reduce_fp64(self._h, int(what), _dp(out)), "lbm_reduce_fp64")
        return out[:3] if what == REDUCE_MOMENTUM else float(out[0])

    def checkpoint(self) -> np.ndarray:
        """The evolving state (lbm_checkpoint_save) as a uint8 array."""
        n = C.c_uint64(0)
        _check(self._lib.lbm_checkpoint_size(self._h, C.byref(n)), "lbm_checkpoint_size")
        buf = np.empty(int(n.value), dtype=np.uint8)
        _check(self._lib.lbm_checkpoint_save(self._h, buf.ctypes.data, n.value), "lbm_checkpoint_save")
        return buf

    def restore(self, buf: np.ndarray):
        """lbm_checkpoint_load of a checkpoint() of a lattice built by the same set-up."""
        b = np.ascontiguousarray(buf, dtype=np.uint8)
        _check(self._lib.lbm_checkpoint_load(self._h, b.ctypes.data, b.size), "lbm_checkpoint_load")

    def reduce_noalloc(self, what: int):
        """reduce() through lbm_reduce_fp64_noalloc (no device allocation)."""
        out = np.zeros(3, np.float64)
        _check(self._lib.lbm_reduce_fp64_noalloc(self._h, int(what), _dp(out)), "lbm_reduce_fp64_noalloc")
        return out[:3] if what == REDUCE_MOMENTUM else float(out[0])

    # -- thermal branch (ENERGY, lbm.cpp:121-178, 263-289, 636-653, 777-806) ----
    def upload_thermal(self, T=None, type_energy=None, *, tau_g, nu=0.0, cp=1.0, j_dot_n=0.0,
                       boussinesq=False, g_beta=0.0, T_ref=0.0, T0=float("nan"), model="energy"):
        """After upload(), before init(): lb.fluid[i][j][k].T and .type_Energy
        and the case's thermal constants.  model "energy": the reference's
        same-lattice D3Q19 g; "d3q7": D3Q7 advection-diffusion (nu, cp
        unused)."""
        m = {"energy": THERMAL_ENERGY, "d3q7": THERMAL_D3Q7}[model] if isinstance(model, str) else int(model)
        self._ng = 7 if m == THERMAL_D3Q7 else 19
        d = _ThermalDesc(float(tau_g), float(nu), float(cp), float(j_dot_n), float(g_beta),
                         float(T_ref), float(T0), int(bool(boussinesq)), m, (C.c_int32 * 2)())
        Ta = None if T is None else self._f64(T)
        te = None
        if type_energy is not None:
            te = np.ascontiguousarray(np.broadcast_to(np.asarray(type_energy, np.int16), self.shape))
        _check(self._lib.lbm_upload_thermal(self._h, C.byref(d), _dp(Ta),
                                            None if te is None else te.ctypes.data_as(C.POINTER(C.c_int16))),
               "lbm_upload_thermal")

    def thermal(self):
        """(T, qx, qy, qz) as of the last MacroProp."""
        out = [np.empty(self.shape) for _ in range(4)]
        _check(self._lib.lbm_download_thermal(self._h, *[_dp(a) for a in out]), "lbm_download_thermal")
        return tuple(out)

    def g_populations(self):
        a = np.empty(self.shape + (getattr(self, "_ng", 19),))
        _check(self._lib.lbm_download_g(self._h, _dp(a)), "lbm_download_g")
        return a

    # -- macroscopic output (OutputVTK, lbm.cpp:2290-2468) ----------------------
    def output_vtk(self, path, **kw):
        d = vtk_desc(**kw)
        _check(self._lib.lbm_output_vtk(self._h, C.byref(d), os.fsencode(path)), "lbm_output_vtk")

    def output_vtk_begin(self, **kw):
        d = vtk_desc(**kw)
        _check(self._lib.lbm_output_vtk_begin(self._h, C.byref(d)), "lbm_output_vtk_begin")

    def output_vtk_end(self, path):
        _check(self._lib.lbm_output_vtk_end(self._h, os.fsencode(path)), "lbm_output_vtk_end")

    # -- a flexible filament coupled through the IBM ----------------------------
    def fsi_setup(self, filament: "Filament", box, m_max: int = 20):
        d = _IbmDesc(0, int(m_max), (C.c_int32 * 6)(*[int(b) for b in box]), 1, 0, 0.0,
                     (C.c_double * 2)(0.0, 0.0), 0.0)
        _check(self._lib.lbm_fsi_setup(self._h, filament._h, C.byref(d)), "lbm_fsi_setup")
        self._filament = filament  # the coupling reads it every step

    def fsi_run(self, nsteps: int):
        """(nsteps, 6): total force on the filament and its free end, per step."""
        out = np.zeros((int(nsteps), 6))
        _check(self._lib.lbm_fsi_run(self._h, int(nsteps), _dp(out)), "lbm_fsi_run")
        return out

    # -- immersed boundary (IBLBM, lbm.cpp:1135-1417) ---------------------------
    def ibm_setup(self, markers: dict, box, m_max: int = 20, vsolid: float = 0.0, moving=False,
                  follow=None, scheme: str = "implicit"):
        """markers: dict of float64 arrays x, y, z, nx, ny, nz, ds; box:
        (xmin, xmax, ymin, ymax, zmin, zmax) inclusive lattice indices.
        follow: None, or (center_x, center_y, half) for a box that follows
        the body as the reference's does (then only zmin, zmax of box are
        used).  scheme: "implicit" (the reference's default) or "explicit"."""
        n = int(markers["x"].size)
        fc = (0.0, 0.0, 0.0) if follow is None else tuple(float(v) for v in follow)
        d = _IbmDesc(n, int(m_max), (C.c_int32 * 6)(*[int(b) for b in box]), int(bool(moving)),
                     int(follow is not None), float(vsolid), (C.c_double * 2)(fc[0], fc[1]), fc[2],
                     {"implicit": 0, "explicit": 1}[scheme])
        arrs = [np.ascontiguousarray(markers[k], dtype=np.float64) for k in
                ("x", "y", "z", "nx", "ny", "nz", "ds")]
        self._ibm_n = n
        _check(self._lib.lbm_ibm_setup(self._h, C.byref(d), *[_dp(a) for a in arrs]), "lbm_ibm_setup")

    def stress(self):
        """MORE_DETAILS: (sigma (..., 3, 3), pressure) of every cell from the
        last MacroProp's state (lbm.cpp:681-705)."""
        sig = np.empty(self.shape + (3, 3))
        p = np.empty(self.shape)
        _check(self._lib.lbm_download_stress(self._h, _dp(sig), _dp(p)), "lbm_download_stress")
        return sig, p

    def stats_update(self, n, nt):
        """Sup_MacroProp(n) with output interval nt (lbm.cpp:865-939)."""
        _check(self._lib.lbm_stats_update(self._h, int(n), int(nt)), "lbm_stats_update")

    def stats(self):
        """(avg_u, rms_u, avg_p, rms_p, avg_rs, rms_rs), shapes (...,3),
        (...,3), (...), (...), (...,3,3), (...,3,3)."""
        out = [np.empty(self.shape + (3,)), np.empty(self.shape + (3,)), np.empty(self.shape),
               np.empty(self.shape), np.empty(self.shape + (3, 3)), np.empty(self.shape + (3, 3))]
        _check(self._lib.lbm_download_stats(self._h, *[_dp(a) for a in out]), "lbm_download_stats")
        return tuple(out)

    def momentum_exchange(self, box):
        """MEA_CD / MEA_CL's raw momentum sums (Px, Py, Pz) over the
        inclusive box (x0, x1, y0, y1, z0, z1), lbm.cpp:3085-3163."""
        b = (C.c_int32 * 6)(*[int(v) for v in box])
        out = np.zeros(3)
        _check(self._lib.lbm_momentum_exchange(self._h, b, _dp(out)), "lbm_momentum_exchange")
        return out

    def ibm_box(self):
        b = (C.c_int32 * 6)()
        cen = np.zeros(2)
        _check(self._lib.lbm_ibm_box(self._h, b, _dp(cen)), "lbm_ibm_box")
        return tuple(b[:]), cen

    def ibm_apply(self, vcm=(0.0, 0.0, 0.0), vcm_before=(0.0, 0.0, 0.0)):
        out = np.zeros(3)
        v = np.asarray(vcm, np.float64)
        b = np.asarray(vcm_before, np.float64)
        _check(self._lib.lbm_ibm_apply(self._h, _dp(v), _dp(b), _dp(out)), "lbm_ibm_apply")
        return out

    def ibm_run(self, nsteps: int, vcm=None):
        """vcm: None (body at rest) or array (nsteps+1, 3).  Returns (nsteps, 3):
        Drag, Lift, interior density per step."""
        out = np.zeros((int(nsteps), 3))
        v = None if vcm is None else np.ascontiguousarray(vcm, np.float64)
        _check(self._lib.lbm_ibm_run(self._h, int(nsteps), _dp(v), _dp(out)), "lbm_ibm_run")
        return out

    def ibm_markers(self):
        n = self._ibm_n
        x, y, z = np.empty(n), np.empty(n), np.empty(n)
        ub, Fb = np.empty(3 * n), np.empty(3 * n)
        _check(self._lib.lbm_ibm_markers(self._h, _dp(x), _dp(y), _dp(z), _dp(ub), _dp(Fb)),
               "lbm_ibm_markers")
        return dict(x=x, y=y, z=z, ub=ub, Fb=Fb)

    def autotune(self, nsteps: int = 20):
        """Dry-run tuning of the launch shape (the lattice's state is not
        changed); returns (threads per block, cells per thread), now set."""
        b, v = C.c_int32(0), C.c_int32(0)
        _check(self._lib.lbm_autotune(self._h, int(nsteps), C.byref(b), C.byref(v)), "lbm_autotune")
        return b.value, v.value

    def set_launch(self, block: int, cells_per_thread: int = 1):
        _check(self._lib.lbm_set_launch(self._h, int(block), int(cells_per_thread)), "lbm_set_launch")

    def get_launch(self):
        b, v = C.c_int32(0), C.c_int32(0)
        _check(self._lib.lbm_get_launch(self._h, C.byref(b), C.byref(v)), "lbm_get_launch")
        return b.value, v.value

    def bounce_back_faces(self):
        """(faces, slot_links) after init: the folded-face bits (i = 0, Nx+1,
        j = 0, Ny+1, k = 0, Nz+1) and whether some wall link uses a slot."""
        f, s = C.c_int32(), C.c_int32()
        _check(self._lib.lbm_bounce_back_faces(self._h, C.byref(f), C.byref(s)), "lbm_bounce_back_faces")
        return f.value, s.value

    @property
    def ncells(self) -> int:
        return int(self._lib.lbm_ncells(self._h))

    @property
    def nfluid(self) -> int:
        return int(self._lib.lbm_nfluid(self._h))

    @property
    def bytes_per_cell(self) -> float:
        return float(self._lib.lbm_bytes_per_cell(self._h))


class Batch:
    """An ensemble of Lattice objects stepped by one launch per step
    (lbm_batch_*).  The lattices must stay alive while the batch is used."""

    def __init__(self, lattices):
        self._lib = load_library()
        self.lattices = list(lattices)
        arr = (_P * len(self.lattices))(*[lat._h for lat in self.lattices])
        h = _P()
        _check(self._lib.lbm_batch_create(arr, len(self.lattices), C.byref(h)), "lbm_batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int):
        _check(self._lib.lbm_batch_step(self._h, int(n)), "lbm_batch_step")

    def time_steps(self, n: int) -> float:
        ms = C.c_double(0)
        _check(self._lib.lbm_batch_time_steps(self._h, int(n), C.byref(ms)), "lbm_batch_time_steps")
        return ms.value

    def set_launch(self, block: int, per_cu: int):
        _check(self._lib.lbm_batch_set_launch(self._h, int(block), int(per_cu)), "lbm_batch_set_launch")

    def get_launch(self):
        b, p = C.c_int32(0), C.c_int32(0)
        _check(self._lib.lbm_batch_get_launch(self._h, C.byref(b), C.byref(p)), "lbm_batch_get_launch")
        return b.value, p.value

    def autotune(self, nsteps: int = 20):
        b, p, ms = C.c_int32(0), C.c_int32(0), C.c_double(0)
        _check(self._lib.lbm_batch_autotune(self._h, int(nsteps), C.byref(b), C.byref(p), C.byref(ms)),
               "lbm_batch_autotune")
        return b.value, p.value, ms.value

    @property
    def nfluid(self) -> int:
        return int(self._lib.lbm_batch_nfluid(self._h))
