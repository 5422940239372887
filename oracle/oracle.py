"""ctypes binding of the CPU oracle (liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline.  The product
library never loads it.

Arrays use the reference's LBM::fluid[i][j][k] order: numpy shape
(Nx+2, Ny+2, Nz+2) in C order; populations (Nx+2, Ny+2, Nz+2, q).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_D = C.POINTER(C.c_double)
_P = C.c_void_p
_SIG = {
    "orc_create": (_P, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]),
    "orc_destroy": (None, [_P]),
    "orc_ncells": (C.c_int64, [_P]),
    "orc_sizeof_cell": (C.c_int, []),
    "orc_q": (C.c_int, [_P]),
    "orc_tables": (None, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                          C.POINTER(C.c_int), _D]),
    "orc_set_cells": (None, [_P, C.POINTER(C.c_int16), _D, _D, _D, _D, _D, _D, _D]),
    "orc_get_macro": (None, [_P, _D, _D, _D, _D]),
    "orc_get_force": (None, [_P, _D, _D, _D]),
    "orc_get_fm": (None, [_P, _D, _D, _D]),
    "orc_get_f": (None, [_P, _D]),
    "orc_get_fstar": (None, [_P, _D]),
    "orc_set_f": (None, [_P, _D]),
    "orc_set_fstar": (None, [_P, _D]),
    "orc_feq": (C.c_double, [_P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "orc_init": (None, [_P]),
    "orc_collision": (C.c_int, [_P]),
    "orc_streaming": (None, [_P]),
    "orc_macroprop": (C.c_int, [_P]),
    "orc_run_plain": (C.c_int, [_P, C.c_int]),
    "orc_wall_velocity": (None, [_P, C.c_int, C.c_int, C.c_int, _D]),
    "orc_export_links": (None, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_uint8)]),
    "orc_set_threads": (None, [C.c_int]),
    "orc_get_threads": (C.c_int, []),
    "orc_ibm_kernel": (C.c_double, [C.c_double]),
    "orc_ibm_nearest": (None, [C.c_int] * 6 + [C.c_double] * 6 + [C.POINTER(C.c_int)] * 3),
    "orc_ibm_step": (C.c_int, [_P, C.c_int, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]
                     + [C.c_int] * 7 + [C.c_double, _D]),
    "orc_thermal_enable": (C.c_int, [_P] + [C.c_double] * 4 + [C.c_int] + [C.c_double] * 2),
    "orc_set_thermal": (None, [_P, _D, C.POINTER(C.c_int16)]),
    "orc_get_thermal": (None, [_P, _D, _D, _D, _D]),
    "orc_get_g": (None, [_P, _D]),
    "orc_q7_enable": (C.c_int, [_P, C.c_double, C.c_int, C.c_double, C.c_double]),
    "orc_q7_set_flux": (None, [_P, C.c_double]),
    "orc_q7_set": (None, [_P, _D, C.POINTER(C.c_int16)]),
    "orc_q7_get": (None, [_P, _D, _D, _D, _D]),
    "orc_q7_get_g": (None, [_P, _D]),
    "orc_fsi_ibm_step": (C.c_int, [_P, C.c_int, _D, _D, _D, _D, _D, _D, _D, _D] + [C.c_int] * 7),
    "orc_mea": (None, [_P, C.POINTER(C.c_int), _D]),
    "orc_ibm_step_explicit": (C.c_int, [_P, C.c_int, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]
                              + [C.c_int] * 6 + [C.c_double, _D]),
    "orc_stress": (None, [_P, _D, _D]),
    "orc_stats_update": (None, [_P, C.c_int, C.c_int]),
    "orc_get_stats": (None, [_P, _D, _D, _D, _D, _D, _D]),
    "orc_fil_create": (_P, [C.c_int] + [C.c_double] * 4 + [_D, _D, _D, C.c_int]),
    "orc_fil_destroy": (None, [_P]),
    "orc_fil_substeps": (C.c_int, [_P]),
    "orc_fil_set": (None, [_P, _D, _D]),
    "orc_fil_get": (None, [_P, _D, _D, _D]),
    "orc_fil_step": (None, [_P, _D, C.c_double, C.c_int]),
    "orc_set_g": (None, [_P, _D]),
    "orc_wall_temperature": (C.c_double, [_P, C.c_int, C.c_int, C.c_int]),
    "orc_geq": (C.c_double, [_P] + [C.c_int] * 5 + [C.c_double] * 3),
}

_libs: dict[str, C.CDLL] = {}


def load(opt: str = "O3") -> C.CDLL:
    name = "liblbm_oracle.so" if opt == "O3" else "liblbm_oracle_O0.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = C.CDLL(path)
    for k, (res, args) in _SIG.items():
        fn = getattr(lib, k)
        fn.restype = res
        fn.argtypes = args
    _libs[name] = lib
    return lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a, shape):
    if a is None:
        return None
    return np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shape))


class Oracle:
    """The reference's LBM object restated on the CPU (FP64, OpenMP)."""

    def __init__(self, Nx, Ny, Nz, tau, q=19, opt="O3"):
        self.lib = load(opt)
        self.h = self.lib.orc_create(Nx, Ny, Nz, float(tau), q)
        if not self.h:
            raise ValueError("orc_create failed")
        self.Nx, self.Ny, self.Nz, self.tau, self.q = Nx, Ny, Nz, float(tau), q
        self.shape = (Nx + 2, Ny + 2, Nz + 2)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_cells(self, type=None, rho=None, ux=None, uy=None, uz=None, Fx=None, Fy=None, Fz=None):
        t = None
        if type is not None:
            t = np.ascontiguousarray(np.broadcast_to(np.asarray(type, np.int16), self.shape))
        arrs = [_f64(a, self.shape) for a in (rho, ux, uy, uz, Fx, Fy, Fz)]
        self.lib.orc_set_cells(self.h, None if t is None else t.ctypes.data_as(C.POINTER(C.c_int16)),
                               *[_dp(a) for a in arrs])

    def init(self):
        self.lib.orc_init(self.h)

    def collision(self):
        return self.lib.orc_collision(self.h)

    def streaming(self):
        self.lib.orc_streaming(self.h)

    def macroprop(self):
        return self.lib.orc_macroprop(self.h)

    def run(self, n):
        return self.lib.orc_run_plain(self.h, int(n))

    def macro(self):
        out = [np.empty(self.shape) for _ in range(4)]
        self.lib.orc_get_macro(self.h, *[_dp(a) for a in out])
        return tuple(out)

    def force(self):
        out = [np.empty(self.shape) for _ in range(3)]
        self.lib.orc_get_force(self.h, *[_dp(a) for a in out])
        return tuple(out)

    def fm(self):
        out = [np.empty(self.shape) for _ in range(3)]
        self.lib.orc_get_fm(self.h, *[_dp(a) for a in out])
        return tuple(out)

    def f(self):
        a = np.empty(self.shape + (self.q,))
        self.lib.orc_get_f(self.h, _dp(a))
        return a

    def fstar(self):
        a = np.empty(self.shape + (self.q,))
        self.lib.orc_get_fstar(self.h, _dp(a))
        return a

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == self.shape + (self.q,)
        self.lib.orc_set_f(self.h, _dp(f))

    def set_fstar(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == self.shape + (self.q,)
        self.lib.orc_set_fstar(self.h, _dp(f))

    # -- thermal branch (ENERGY) -------------------------------------------------
    def thermal_enable(self, tau_g, nu, cp, j_dot_n=0.0, boussinesq=False, g_beta=0.0, T_ref=0.0):
        self.lib.orc_thermal_enable(self.h, float(tau_g), float(nu), float(cp), float(j_dot_n),
                                    int(bool(boussinesq)), float(g_beta), float(T_ref))

    def set_thermal(self, T=None, type_energy=None):
        Ta = None if T is None else _f64(T, self.shape)
        te = None
        if type_energy is not None:
            te = np.ascontiguousarray(np.broadcast_to(np.asarray(type_energy, np.int16), self.shape))
        self.lib.orc_set_thermal(self.h, _dp(Ta), None if te is None else te.ctypes.data_as(C.POINTER(C.c_int16)))

    def thermal(self):
        out = [np.empty(self.shape) for _ in range(4)]
        self.lib.orc_get_thermal(self.h, *[_dp(a) for a in out])
        return tuple(out)

    def g(self):
        a = np.empty(self.shape + (self.q,))
        self.lib.orc_get_g(self.h, _dp(a))
        return a

    def set_g(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        assert g.shape == self.shape + (self.q,)
        self.lib.orc_set_g(self.h, _dp(g))

    # -- D3Q7 temperature (no reference counterpart) ----------------------------
    def q7_enable(self, tau_g, boussinesq=False, g_beta=0.0, T_ref=0.0, j_dot_n=0.0):
        assert self.lib.orc_q7_enable(self.h, float(tau_g), int(bool(boussinesq)), float(g_beta), float(T_ref))
        self.lib.orc_q7_set_flux(self.h, float(j_dot_n))

    def set_q7(self, T=None, type_energy=None):
        """After set_cells (the buoyancy adds to the set-up F_y read here)."""
        Ta = None if T is None else _f64(T, self.shape)
        te = None
        if type_energy is not None:
            te = np.ascontiguousarray(np.broadcast_to(np.asarray(type_energy, np.int16), self.shape))
        self.lib.orc_q7_set(self.h, _dp(Ta), None if te is None else te.ctypes.data_as(C.POINTER(C.c_int16)))

    def q7_thermal(self):
        out = [np.empty(self.shape) for _ in range(4)]
        self.lib.orc_q7_get(self.h, *[_dp(a) for a in out])
        return tuple(out)

    def q7_g(self):
        a = np.empty(self.shape + (7,))
        self.lib.orc_q7_get_g(self.h, _dp(a))
        return a

    def fsi_ibm_step(self, x, udes, ds, box, m_max):
        """IBM step of a coupled filament: x (n,3) marker positions, udes
        (n,3), ds (n,); returns (ub, Fb, fsum), each (n,3)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = x.shape[0]
        mx, my, mz = (np.ascontiguousarray(x[:, d]) for d in range(3))
        ud = np.ascontiguousarray(udes, dtype=np.float64).reshape(n, 3)
        dsa = np.ascontiguousarray(ds, dtype=np.float64)
        ub, Fb, fsum = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        err = self.lib.orc_fsi_ibm_step(self.h, n, _dp(mx), _dp(my), _dp(mz), _dp(ud), _dp(dsa), _dp(ub), _dp(Fb),
                                        _dp(fsum), *[int(b) for b in box], int(m_max))
        assert err == 0, "a marker left the box"
        return ub, Fb, fsum

    def stress(self):
        """MORE_DETAILS: (sigma (..., 3, 3), pressure) of every cell."""
        sig = np.empty(self.shape + (3, 3))
        p = np.empty(self.shape)
        self.lib.orc_stress(self.h, _dp(sig), _dp(p))
        return sig, p

    def stats_update(self, n, nt):
        """Sup_MacroProp(n) with output interval nt."""
        self.lib.orc_stats_update(self.h, int(n), int(nt))

    def stats(self):
        """(avg_u, rms_u, avg_p, rms_p, avg_rs, rms_rs), shapes (...,3),
        (...,3), (...), (...), (...,3,3), (...,3,3)."""
        out = [np.empty(self.shape + (3,)), np.empty(self.shape + (3,)), np.empty(self.shape),
               np.empty(self.shape), np.empty(self.shape + (3, 3)), np.empty(self.shape + (3, 3))]
        self.lib.orc_get_stats(self.h, *[_dp(a) for a in out])
        return tuple(out)

    def mea(self, box):
        """MEA_CD / MEA_CL's raw momentum sums (Px, Py, Pz) over the box."""
        b = (C.c_int * 6)(*[int(v) for v in box])
        out = np.zeros(3)
        self.lib.orc_mea(self.h, b, _dp(out))
        return out

    def wall_temperature(self, i, j, k):
        return self.lib.orc_wall_temperature(self.h, i, j, k)

    def geq(self, i, j, k, l, as_solid=False, uw=(0.0, 0.0, 0.0)):
        return self.lib.orc_geq(self.h, i, j, k, l, int(bool(as_solid)), *[float(v) for v in uw])

    def feq(self, rho, ux, uy, uz, l):
        return self.lib.orc_feq(self.h, rho, ux, uy, uz, l)

    def wall_velocity(self, i, j, k):
        out = np.zeros(3)
        self.lib.orc_wall_velocity(self.h, i, j, k, _dp(out))
        return out

    def links(self):
        src = np.empty(self.shape + (self.q,), np.int32)
        kind = np.empty(self.shape + (self.q,), np.uint8)
        self.lib.orc_export_links(self.h, src.ctypes.data_as(C.POINTER(C.c_int32)),
                                  kind.ctypes.data_as(C.POINTER(C.c_uint8)))
        return src, kind

    def tables(self):
        q = self.q
        cx, cy, cz, opp = [(C.c_int * q)() for _ in range(4)]
        w = (C.c_double * q)()
        self.lib.orc_tables(self.h, cx, cy, cz, opp, w)
        return (np.array(cx[:]), np.array(cy[:]), np.array(cz[:]), np.array(opp[:]),
                np.array(w[:]))

    def ibm_step(self, markers: dict, vcm, vcm_before, box, m_max=20, vsolid=0.0, scheme="implicit"):
        """One IBLBM call (implicit, or the explicit branch); markers is a
        dict of float64 arrays x, y, z, nx, ny, nz, ds (updated in place:
        x, y, z, ub, Fb)."""
        n = markers["x"].size
        for k in ("ub", "Fb"):
            if k not in markers:
                markers[k] = np.zeros(3 * n)
        out = np.zeros(3)
        vcm = np.asarray(vcm, np.float64)
        vcb = np.asarray(vcm_before, np.float64)
        if scheme == "explicit":
            err = self.lib.orc_ibm_step_explicit(
                self.h, n, _dp(markers["x"]), _dp(markers["y"]), _dp(markers["z"]),
                _dp(markers["nx"]), _dp(markers["ny"]), _dp(markers["nz"]),
                _dp(markers["ub"]), _dp(markers["Fb"]), _dp(markers["ds"]),
                _dp(vcm), _dp(vcb), *[int(b) for b in box], float(vsolid), _dp(out))
            return err, out
        err = self.lib.orc_ibm_step(
            self.h, n, _dp(markers["x"]), _dp(markers["y"]), _dp(markers["z"]),
            _dp(markers["nx"]), _dp(markers["ny"]), _dp(markers["nz"]),
            _dp(markers["ub"]), _dp(markers["Fb"]), _dp(markers["ds"]),
            _dp(vcm), _dp(vcb), *[int(b) for b in box], int(m_max), float(vsolid), _dp(out))
        return err, out


class Filament:
    """The oracle's flexible filament (specification in lbm_oracle.h)."""

    def __init__(self, n, length, line_density, bending, damping=0.0, body=(0.0, 0.0, 0.0),
                 clamp=(0.0, 0.0, 0.0), tangent=(1.0, 0.0, 0.0), substeps=0):
        self.lib = load()
        arr = lambda v: np.ascontiguousarray(v, dtype=np.float64)  # noqa: E731
        self._keep = [arr(body), arr(clamp), arr(tangent)]
        self.h = self.lib.orc_fil_create(int(n), float(length), float(line_density), float(bending),
                                         float(damping), *[_dp(a) for a in self._keep], int(substeps))
        assert self.h
        self.n = int(n)

    def close(self):
        if self.h:
            self.lib.orc_fil_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def substeps(self):
        return self.lib.orc_fil_substeps(self.h)

    def set(self, x_now, x_prev=None):
        a = np.ascontiguousarray(x_now, dtype=np.float64).reshape(self.n, 3)
        b = None if x_prev is None else np.ascontiguousarray(x_prev, dtype=np.float64).reshape(self.n, 3)
        self.lib.orc_fil_set(self.h, _dp(a), _dp(b))

    def get(self):
        x, xp, t = np.empty((self.n, 3)), np.empty((self.n, 3)), np.empty(self.n - 1)
        self.lib.orc_fil_get(self.h, _dp(x), _dp(xp), _dp(t))
        return x, xp, t

    def step(self, force=None, dt=1.0, nsteps=1):
        F = None if force is None else np.ascontiguousarray(force, dtype=np.float64).reshape(self.n, 3)
        self.lib.orc_fil_step(self.h, _dp(F), float(dt), int(nsteps))


def kernel(r: float) -> float:
    return load().orc_ibm_kernel(float(r))


def nearest(box, m, nvec):
    i, j, k = C.c_int(-1), C.c_int(-1), C.c_int(1)
    load().orc_ibm_nearest(*[int(b) for b in box], *[float(v) for v in m],
                           *[float(v) for v in nvec], C.byref(i), C.byref(j), C.byref(k))
    return i.value, j.value, k.value


def sizeof_cell() -> int:
    return load().orc_sizeof_cell()


def set_threads(n: int):
    load().orc_set_threads(int(n))
    load("O0").orc_set_threads(int(n))


def run_case(case, nsteps: int, opt="O3") -> "Oracle":
    o = Oracle(case.Nx, case.Ny, case.Nz, case.tau, opt=opt)
    o.set_cells(**case.fields())
    o.init()
    if nsteps:
        e = o.run(nsteps)
        if e:
            raise RuntimeError(f"oracle aborted with code {e}")
    return o
