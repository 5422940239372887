"""Runtime case set-ups: the role of the reference's main_setup_* functions.

The reference hard-codes its cases behind compile-time #defines
(LBM.h:8-72, lbm.cpp:1429-2031).  Here each case is a function returning the
fields a setup writes into lb.fluid[i][j][k] -- type, density, ux, uy, uz and
F -- as numpy arrays of shape (Nx+2, Ny+2, Nz+2), plus the lattice size and
tau.  The same arrays feed the GPU library and the CPU oracle, so both start
from identical (seeded) inputs.

Cell-type precedence follows the reference's setup loops: the later `if`
wins (e.g. lbm.cpp:1838-1841).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

TYPE_F, TYPE_S, TYPE_OUT, TYPE_IN = 0, 1, 2, 3
TYPE_T, TYPE_Q = 6, 7
# LBM.h:74 and lbm.cpp:1859 use this literal for pi
PI_REF = 3.14159265

_MASK64 = (1 << 64) - 1
_GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(x):
    """splitmix64 finaliser (Steele, Lea & Flood 2014) on uint64 numpy arrays."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(_GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def counter_uniform(seed: int, counter) -> np.ndarray:
    """Counter-based uniform draws in [-1, 1): u = splitmix64(seed*G + counter)."""
    c = np.asarray(counter, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = np.uint64((seed * _GOLDEN) & _MASK64)
        z = splitmix64(base + c)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0


@dataclass
class Case:
    name: str
    Nx: int
    Ny: int
    Nz: int
    tau: float
    type: np.ndarray
    rho: np.ndarray
    ux: np.ndarray
    uy: np.ndarray
    uz: np.ndarray
    Fx: np.ndarray | None = None
    Fy: np.ndarray | None = None
    Fz: np.ndarray | None = None
    params: dict = field(default_factory=dict)

    @property
    def shape(self):
        return (self.Nx + 2, self.Ny + 2, self.Nz + 2)

    def fields(self):
        return dict(type=self.type, rho=self.rho, ux=self.ux, uy=self.uy, uz=self.uz,
                    Fx=self.Fx, Fy=self.Fy, Fz=self.Fz)

    @property
    def nfluid(self) -> int:
        return int((self.type == TYPE_F).sum())


def _blank(Nx, Ny, Nz):
    shape = (Nx + 2, Ny + 2, Nz + 2)
    return (np.zeros(shape, np.int16), np.ones(shape), np.zeros(shape), np.zeros(shape),
            np.zeros(shape))


def cavity(N: int = 64, u_lid: float = 0.05, Re: float = 100.0, Ny: int | None = None,
           Nz: int | None = None) -> Case:
    """Lid-driven cavity LBM(N,N,N): all six ghost faces TYPE_S, the whole
    j = Ny+1 face moves with ux = u_lid (the moving-wall pattern of the
    Couette setup, lbm.cpp:1536-1584), rho = 1 and u = 0 elsewhere.
    nu = u_lid*N/Re, tau = 0.5 + nu/cs2 (LBM.h:100-101)."""
    Nx = N
    Ny = N if Ny is None else Ny
    Nz = N if Nz is None else Nz
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[0, :, :] = TYPE_S
    t[-1, :, :] = TYPE_S
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    t[:, :, 0] = TYPE_S
    t[:, :, -1] = TYPE_S
    ux[:, -1, :] = u_lid
    nu = u_lid * N / Re
    tau = 0.5 + 3.0 * nu
    return Case("cavity", Nx, Ny, Nz, tau, t, rho, ux, uy, uz,
                params=dict(u_lid=u_lid, Re=Re, nu=nu))


def taylor_green(Np: int = 256, u0: float = 0.04, Re: float = 1600.0) -> Case:
    """Periodic Taylor-Green vortex on a Np^3 periodic box.  The reference
    wraps modulo N+2 over its padded array (lbm.cpp:197-199), so the box is
    LBM(Np-2, Np-2, Np-2) with every cell TYPE_F.  k = 2 pi / Np,
    u = u0 sin kx cos ky cos kz, v = -u0 cos kx sin ky cos kz, w = 0,
    p = rho0 u0^2/16 (cos 2kx + cos 2ky)(cos 2kz + 2), rho = 1 + p/cs2,
    nu = u0 Np / (2 pi) / Re."""
    n = Np - 2
    k = 2.0 * math.pi / Np
    x = np.arange(Np, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    ux = u0 * np.sin(k * X) * np.cos(k * Y) * np.cos(k * Z)
    uy = -u0 * np.cos(k * X) * np.sin(k * Y) * np.cos(k * Z)
    uz = np.zeros_like(ux)
    p = u0 * u0 / 16.0 * (np.cos(2 * k * X) + np.cos(2 * k * Y)) * (np.cos(2 * k * Z) + 2.0)
    rho = 1.0 + 3.0 * p
    t = np.zeros(ux.shape, np.int16)
    nu = u0 * Np / (2.0 * math.pi) / Re
    tau = 0.5 + 3.0 * nu
    return Case("taylor_green", n, n, n, tau, t, rho, ux, uy, uz,
                params=dict(u0=u0, Re=Re, nu=nu, k=k))


def channel(Nx: int = 256, Ny: int = 128, Nz: int = 64, u_in: float = 0.04, Re: float = 100.0,
            D: float = 20.0, seed: int = 1, perturb: float = 0.01) -> Case:
    """Channel in the pattern of main_setup_Cylinder_IBM (lbm.cpp:1828-1855):
    i = 0 TYPE_IN, i = Nx+1 TYPE_OUT, then the j and k ghost faces TYPE_S
    (walls override inlet/outlet on the edges).  F/IN/OUT cells start at
    rho = 1, u = u_in; TYPE_F cells carry a seeded +-perturb relative
    perturbation: u = u_in(1 + p xi_x), v = p u_in xi_y, w = p u_in xi_z with
    xi = counter_uniform(seed, 3*cell + component)."""
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[0, :, :] = TYPE_IN
    t[-1, :, :] = TYPE_OUT
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    t[:, :, 0] = TYPE_S
    t[:, :, -1] = TYPE_S
    live = t != TYPE_S
    ux[live] = u_in
    if perturb:
        n = t.size
        cell = np.arange(n, dtype=np.uint64).reshape(t.shape)
        xi = [counter_uniform(seed, 3 * cell + np.uint64(c)) for c in range(3)]
        f = t == TYPE_F
        ux[f] = u_in * (1.0 + perturb * xi[0][f])
        uy[f] = perturb * u_in * xi[1][f]
        uz[f] = perturb * u_in * xi[2][f]
    nu = u_in * D / Re
    tau = 0.5 + 3.0 * nu
    return Case("channel", Nx, Ny, Nz, tau, t, rho, ux, uy, uz,
                params=dict(u_in=u_in, Re=Re, D=D, nu=nu, seed=seed))


def couette(Nx: int = 2, Ny: int = 30, Nz: int = 2, tau: float = 0.9, u_wall: float = 0.01) -> Case:
    """main_setup_Couette (lbm.cpp:1536-1584): x and z periodic through TYPE_F
    ghosts, j = 0 and j = Ny+1 TYPE_S, top wall ux = u_wall."""
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    ux[:, -1, :] = u_wall
    return Case("couette", Nx, Ny, Nz, tau, t, rho, ux, uy, uz, params=dict(u_wall=u_wall))


@dataclass
class Thermal:
    """The thermal fields a setup writes (lb.fluid[i][j][k].T, .type_Energy)
    and the ENERGY case constants (LBM.h:150-161, 206-220)."""
    T: np.ndarray
    type_energy: np.ndarray
    tau_g: float
    nu: float
    cp: float
    j_dot_n: float = 0.0
    boussinesq: bool = False
    g_beta: float = 0.0
    T_ref: float = 0.0
    params: dict = field(default_factory=dict)


def thermal_constants(nu: float, Pr: float, rho0: float = 1.0):
    """alpha = nu/Pr, tau_g = 0.5 + alpha/cs2, k = cs2 tau_g, cp = k/(rho0
    alpha) (LBM.h:153-156)."""
    cs2 = 1.0 / 3.0
    alpha = nu / Pr
    tau_g = 0.5 + alpha / (cs2 * 1.0)
    k = cs2 * tau_g * 1.0
    cp = k / (rho0 * alpha)
    return alpha, tau_g, k, cp


def couette_thermal(Nx: int = 2, Ny: int = 20, Nz: int = 2, tau: float = 0.9, Re: float = 10.0,
                    Pr: float = 0.7, Br: float = 10.0, Tb: float = 0.5):
    """main_setup_Couette with ENERGY (lbm.cpp:1536-1584, LBM.h:165-220):
    walls at j = 0, Ny+1 (TYPE_S, type_Energy TYPE_T), the top one moving
    with u_wall = Re nu/(Ny+1); T = Tb on the fluid and the bottom wall,
    Tt = Tb + u_wall^2/(cp Ec) on the top wall, Ec = Br/Pr.  Returns
    (case, thermal)."""
    cs2 = 1.0 / 3.0
    nu = cs2 * (tau - 0.5)
    u_wall = Re * nu / (Ny + 1)
    alpha, tau_g, k, cp = thermal_constants(nu, Pr)
    Ec = Br / Pr
    dT = u_wall ** 2 / (cp * Ec)
    Tt = Tb + dT
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    ux[:, -1, :] = u_wall
    T = np.full(t.shape, Tb)
    T[:, -1, :] = Tt
    te = np.zeros(t.shape, np.int16)
    te[:, 0, :] = TYPE_T
    te[:, -1, :] = TYPE_T
    case = Case("couette_thermal", Nx, Ny, Nz, tau, t, rho, ux, uy, uz, params=dict(u_wall=u_wall))
    th = Thermal(T, te, tau_g, nu, cp, params=dict(Tb=Tb, Tt=Tt, dT=dT, Pr=Pr, Br=Br, alpha=alpha,
                                                   u_wall=u_wall))
    return case, th


def conduction_channel(Nx: int = 2, Ny: int = 12, Nz: int = 2, tau: float = 0.8, Pr: float = 0.7,
                       bottom: int = TYPE_T, top: int = TYPE_T, Tb: float = 1.0, Tt: float = 0.5,
                       T_fluid=None, seed: int = 3):
    """Fluid at rest between walls at j = 0 and j = Ny+1 (TYPE_S) with
    type_Energy `bottom` / `top` (TYPE_T Dirichlet or TYPE_Q Neumann), set-up
    temperatures Tb, Tt; x and z periodic.  The fluid starts at T_fluid
    (default: a seeded profile between Tt and Tb)."""
    cs2 = 1.0 / 3.0
    nu = cs2 * (tau - 0.5)
    alpha, tau_g, k, cp = thermal_constants(nu, Pr)
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    if T_fluid is None:
        xi = counter_uniform(seed, np.arange(t.size, dtype=np.uint64)).reshape(t.shape)
        T = 0.5 * (Tb + Tt) + 0.25 * (Tb - Tt) * xi
    else:
        T = np.full(t.shape, float(T_fluid))
    T[:, 0, :] = Tb
    T[:, -1, :] = Tt
    te = np.zeros(t.shape, np.int16)
    te[:, 0, :] = bottom
    te[:, -1, :] = top
    case = Case("conduction", Nx, Ny, Nz, tau, t, rho, ux, uy, uz)
    return case, Thermal(T, te, tau_g, nu, cp, params=dict(Tb=Tb, Tt=Tt, alpha=alpha))


def rayleigh_benard(Nx: int = 12, Ny: int = 8, Nz: int = 4, tau: float = 0.8, Pr: float = 0.71,
                    Tb: float = 2.0, Tt: float = 1.0, g_beta: float = 1e-3):
    """The pattern of main_setup_RayleighBenard (lbm.cpp:1586-1640): all
    ghost faces TYPE_S, the i and k faces Neumann (TYPE_Q), the j faces
    Dirichlet (TYPE_T, set last so they win on the edges) with the bottom at
    Tb and the top at Tt; fluid at Tt with F_y = g_beta (T - T*), T* =
    (Tb + Tt)/2, updated by every MacroProp (Boussinesq)."""
    cs2 = 1.0 / 3.0
    nu = cs2 * (tau - 0.5)
    alpha, tau_g, k, cp = thermal_constants(nu, Pr)
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    te = np.zeros(t.shape, np.int16)
    for sl in ((slice(None), slice(None), 0), (slice(None), slice(None), -1),
               (0, slice(None), slice(None)), (-1, slice(None), slice(None))):
        t[sl] = TYPE_S
        te[sl] = TYPE_Q
    for sl in ((slice(None), 0, slice(None)), (slice(None), -1, slice(None))):
        t[sl] = TYPE_S
        te[sl] = TYPE_T
    T = np.full(t.shape, Tt)
    T[:, 0, :] = Tb
    T[:, -1, :] = Tt
    T_ref = 0.5 * (Tb + Tt)
    Fy = np.zeros(t.shape)
    Fy[t == TYPE_F] = g_beta * (T[t == TYPE_F] - T_ref)
    case = Case("rayleigh_benard", Nx, Ny, Nz, tau, t, rho, ux, uy, uz, Fx=np.zeros(t.shape), Fy=Fy,
                Fz=np.zeros(t.shape))
    th = Thermal(T, te, tau_g, nu, cp, boussinesq=True, g_beta=g_beta, T_ref=T_ref,
                 params=dict(Tb=Tb, Tt=Tt))
    return case, th


def random_thermal_case(Nx: int, Ny: int, Nz: int, seed: int = 5, tau: float = 0.8,
                        Pr: float = 0.7):
    """random_mask_case with temperatures: every cell T in [0.4, 0.9), every
    solid a Dirichlet (TYPE_T) wall.  Exercises the thermal pull (wrap,
    frozen links next to inlet cells, moving Dirichlet walls, the averaged
    wall temperature).  No inlet on the k = 0 layer (thermal_face_case has
    them)."""
    case = random_mask_case(Nx, Ny, Nz, seed=seed, tau=tau)
    case.type[:, :, 0][case.type[:, :, 0] == TYPE_IN] = TYPE_F
    return case, thermal_fields(case, seed=seed + 1000, Pr=Pr)


def thermal_fields(case: Case, seed: int, Pr: float = 0.7, boussinesq: bool = False,
                   g_beta: float = 0.0, T_ref: float = 0.65) -> Thermal:
    """Temperatures for any case: every cell T in [0.4, 0.9), every solid a
    Dirichlet (TYPE_T) wall, the ENERGY constants of case.tau and Pr.  With
    the Boussinesq force the case gets F = (0, g_beta (T - T_ref), 0) in
    every cell (force mode 1), which MacroProp then rewrites."""
    rng = np.random.default_rng(seed)
    T = 0.4 + 0.5 * rng.random(case.shape)
    te = np.where(case.type == TYPE_S, TYPE_T, TYPE_F).astype(np.int16)
    nu = (case.tau - 0.5) / 3.0
    alpha, tau_g, k, cp = thermal_constants(nu, Pr)
    if boussinesq:
        case.Fx = np.zeros(case.shape)
        case.Fy = g_beta * (T - T_ref)
        case.Fz = np.zeros(case.shape)
    return Thermal(T, te, tau_g, nu, cp, boussinesq=boussinesq, g_beta=g_beta, T_ref=T_ref)


def thermal_face_case(Nx: int, Ny: int, Nz: int, seed: int = 3,
                      shell=(TYPE_OUT, TYPE_IN, TYPE_S, TYPE_F, 4), back_inlets=(), boussinesq=False,
                      g_beta: float = 1e-3):
    """face_case with temperatures (thermal_fields): the face loops of
    Streaming on g as well as f (lbm.cpp:299-625, ENERGY), MacroProp's T of
    the inlet/outlet cells.  back_inlets: (i, j, type) triples that put a
    TYPE_IN cell at (i, j, 0) and a cell of `type` at (i, j, Nz+1), whose f
    and g the back-face inflow write then replaces every step."""
    case = face_case(Nx, Ny, Nz, seed=seed, shell=shell)
    for i, j, top in back_inlets:
        case.type[i, j, 0] = TYPE_IN
        case.type[i, j, -1] = top
    return case, thermal_fields(case, seed=seed + 2000, boussinesq=boussinesq, g_beta=g_beta)


# ---- D3Q7 temperature (LBM_THERMAL_D3Q7; no reference counterpart) -------

def q7_tau(alpha: float) -> float:
    """tau_g of D3Q7 for diffusivity alpha: alpha = cs2 (tau_g - 1/2), cs2 = 1/4."""
    return 0.5 + alpha / 0.25


def q7_thermal(T, type_energy, tau_g, boussinesq=False, g_beta=0.0, T_ref=0.0, j_dot_n=0.0, **params) -> Thermal:
    return Thermal(T, type_energy, tau_g, nu=0.0, cp=1.0, j_dot_n=j_dot_n, boussinesq=boussinesq, g_beta=g_beta,
                   T_ref=T_ref, params=dict(model="d3q7", **params))


def q7_conduction(Nx: int = 2, Ny: int = 8, Nz: int = 2, tau: float = 0.8, tau_g: float = 1.0,
                  Tb: float = 1.0, Tt: float = 0.0, bottom: int = TYPE_T, top: int = TYPE_T,
                  T_fluid=None, seed: int = 4, j_dot_n: float = 0.0):
    """Fluid at rest between walls at j = 0 (T = Tb) and j = Ny+1 (T = Tt),
    type_Energy `bottom` / `top` (TYPE_T Dirichlet, TYPE_Q the heat flux
    j_dot_n into the fluid, anything else adiabatic), x and z periodic; the
    fluid starts at T_fluid (default: seeded values between Tt and Tb).  With
    two Dirichlet walls the steady profile is T_j = Tb + (Tt - Tb)(j - 1/2)/Ny
    (walls half-way between cells)."""
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    if T_fluid is None:
        xi = counter_uniform(seed, np.arange(t.size, dtype=np.uint64)).reshape(t.shape)
        T = 0.5 * (Tb + Tt) + 0.5 * (Tb - Tt) * xi
    else:
        T = np.full(t.shape, float(T_fluid))
    T[:, 0, :] = Tb
    T[:, -1, :] = Tt
    te = np.zeros(t.shape, np.int16)
    te[:, 0, :] = bottom
    te[:, -1, :] = top
    case = Case("q7_conduction", Nx, Ny, Nz, tau, t, rho, ux, uy, uz)
    return case, q7_thermal(T, te, tau_g, j_dot_n=j_dot_n, Tb=Tb, Tt=Tt)


def q7_sine(Np: int = 16, U: float = 0.0, A: float = 0.1, T0: float = 0.5, tau_g: float = 1.0,
            tau: float = 0.8, Nyz: int = 2):
    """All-fluid periodic box of Np x (Nyz+2)^2 cells, uniform flow ux = U,
    T = T0 + A sin(k x), k = 2 pi/Np.  At tau_g = 1 one step maps the mode
    by exactly lambda = 3/4 + cos(k)/4 - i U sin(k): T_n = T0 + A Im(lambda^n
    e^{ikx})."""
    t, rho, ux, uy, uz = _blank(Np - 2, Nyz, Nyz)
    ux[:] = U
    k = 2.0 * math.pi / Np
    x = np.arange(Np, dtype=np.float64)[:, None, None]
    T = T0 + A * np.sin(k * x) * np.ones(t.shape)
    case = Case("q7_sine", Np - 2, Nyz, Nyz, tau, t, rho, ux, uy, uz, params=dict(k=k, U=U, A=A, T0=T0))
    return case, q7_thermal(T, np.zeros(t.shape, np.int16), tau_g)


def q7_random_case(Nx: int, Ny: int, Nz: int, seed: int = 9, tau: float = 0.8, tau_g: float = 0.9,
                   boussinesq: bool = True, g_beta: float = 2e-3, j_dot_n: float = 0.0):
    """random_mask_case with temperatures in [0.2, 0.8): half of the solids
    Dirichlet (TYPE_T), half TYPE_Q (the heat flux j_dot_n, adiabatic at 0),
    inlet cells freezing links, and the Boussinesq force about T_ref = 0.5."""
    case = random_mask_case(Nx, Ny, Nz, seed=seed, tau=tau)
    rng = np.random.default_rng(seed + 2000)
    T = 0.2 + 0.6 * rng.random(case.shape)
    te = np.zeros(case.shape, np.int16)
    solid = case.type == TYPE_S
    te[solid] = np.where(rng.random(case.shape) < 0.5, TYPE_T, TYPE_Q)[solid]
    return case, q7_thermal(T, te, tau_g, boussinesq=boussinesq, g_beta=g_beta, T_ref=0.5, j_dot_n=j_dot_n)


def q7_rayleigh_benard(Nx: int = 48, Ny: int = 24, Nz: int = 1, Ra: float = 1e4, Pr: float = 0.71,
                       tau: float = 0.6, Tb: float = 1.0, Tt: float = 0.0, seed: int = 6,
                       perturb: float = 1e-3):
    """Rayleigh-Benard in a slab heated from below: Dirichlet walls at j = 0
    (Tb) and j = Ny+1 (Tt), x and z periodic.  H = Ny, nu = cs2 (tau - 1/2),
    alpha = nu/Pr, g_beta = Ra nu alpha / ((Tb - Tt) H^3), F_y = g_beta (T -
    T_ref), T_ref = (Tb + Tt)/2; the fluid starts on the conduction profile
    with a seeded perturbation."""
    nu = (tau - 0.5) / 3.0
    alpha = nu / Pr
    H = float(Ny)
    g_beta = Ra * nu * alpha / ((Tb - Tt) * H ** 3)
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    j = np.arange(Ny + 2, dtype=np.float64)[None, :, None]
    T = (Tb + (Tt - Tb) * (j - 0.5) / Ny) * np.ones(t.shape)
    xi = counter_uniform(seed, np.arange(t.size, dtype=np.uint64)).reshape(t.shape)
    T = T + perturb * xi
    T[:, 0, :] = Tb
    T[:, -1, :] = Tt
    te = np.zeros(t.shape, np.int16)
    te[:, 0, :] = TYPE_T
    te[:, -1, :] = TYPE_T
    case = Case("q7_rayleigh_benard", Nx, Ny, Nz, tau, t, rho, ux, uy, uz)
    return case, q7_thermal(T, te, q7_tau(alpha), boussinesq=True, g_beta=g_beta,
                            T_ref=0.5 * (Tb + Tt), Ra=Ra, Pr=Pr, nu=nu, alpha=alpha, H=H)


def poiseuille(Nx: int = 2, Ny: int = 30, Nz: int = 2, tau: float = 0.9, g: float = 1e-6) -> Case:
    """Channel between TYPE_S walls at j = 0, Ny+1 (main_setup_Poiseuille,
    lbm.cpp:1495-1530) driven by a uniform body force Fx = g in every cell."""
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    Fx = np.full(t.shape, g)
    return Case("poiseuille", Nx, Ny, Nz, tau, t, rho, ux, uy, uz, Fx=Fx,
                Fy=np.zeros(t.shape), Fz=np.zeros(t.shape), params=dict(g=g))


def random_mask_case(Nx: int, Ny: int, Nz: int, seed: int = 7, p_solid: float = 0.15,
                     p_other: float = 0.05, u_amp: float = 0.03, tau: float = 0.8) -> Case:
    """A lattice with random solid / inlet-outlet / fluid cells, random moving
    solid velocities and random fluid velocities.  Exercises every streaming
    branch (stream, bounce-back with averaged wall velocity, frozen link) and
    the periodic wrap of the ghost shell."""
    rng = np.random.default_rng(seed)
    shape = (Nx + 2, Ny + 2, Nz + 2)
    r = rng.random(shape)
    t = np.zeros(shape, np.int16)
    t[r < p_solid] = TYPE_S
    t[(r >= p_solid) & (r < p_solid + p_other)] = np.int16(TYPE_IN)
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    ux = u_amp * rng.standard_normal(shape)
    uy = u_amp * rng.standard_normal(shape)
    uz = u_amp * rng.standard_normal(shape)
    # half of the solids are stationary
    still = (t == TYPE_S) & (rng.random(shape) < 0.5)
    ux[still] = uy[still] = uz[still] = 0.0
    return Case("random_mask", Nx, Ny, Nz, tau, t, rho, ux, uy, uz, params=dict(seed=seed))


def face_case(Nx: int, Ny: int, Nz: int, seed: int = 3, p_solid: float = 0.1, u_amp: float = 0.03,
              tau: float = 0.8, shell=(TYPE_OUT, TYPE_IN, TYPE_S, TYPE_F, 4)) -> Case:
    """Every shell cell drawn from `shell` (outlets, inlets, solids, fluid,
    TYPE_P), the interior fluid with random solids and a few interior
    outlet / TYPE_P cells: the face loops of Streaming (lbm.cpp:299-625) on
    every face, edge and corner, in both orders of neighbouring types, with
    fluid, solid and other-type inner neighbours."""
    rng = np.random.default_rng(seed)
    shape = (Nx + 2, Ny + 2, Nz + 2)
    t = np.zeros(shape, np.int16)
    inner = (slice(1, -1),) * 3
    r = rng.random(shape)
    t[r < p_solid] = TYPE_S
    t[(r >= p_solid) & (r < p_solid + 0.02)] = np.int16(TYPE_OUT)
    t[(r >= p_solid + 0.02) & (r < p_solid + 0.04)] = np.int16(4)
    shell_t = np.asarray(shell, np.int16)[rng.integers(0, len(shell), shape)]
    keep = t[inner].copy()
    t[...] = shell_t
    t[inner] = keep
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    ux = u_amp * rng.standard_normal(shape)
    uy = u_amp * rng.standard_normal(shape)
    uz = u_amp * rng.standard_normal(shape)
    still = (t == TYPE_S) & (rng.random(shape) < 0.5)
    ux[still] = uy[still] = uz[still] = 0.0
    return Case("faces", Nx, Ny, Nz, tau, t, rho, ux, uy, uz, params=dict(seed=seed))


def cylinder_bb(Nx: int = 100, Ny: int = 100, Nz: int = 1, D: float = 20.0, u_max: float = 0.1,
                Re: float = 200.0, center=None) -> Case:
    """main_setup_Cylinder with Generate_Cylinder (lbm.cpp:1429-1487): cells
    i = 1..Nx, j = 1..Ny, k = 0..Nz whose centre (i - 1/2, j - 1/2) lies
    within D/2 of the axis (Nx/5, Ny/2) become TYPE_S; then i = 0 TYPE_IN,
    i = Nx+1 and j = 0, Ny+1 TYPE_OUT, the k ghost layers TYPE_F when
    Nz != 1; F/IN/OUT cells start at rho = 1, u = (u_max, 0, 0), solids at
    the LATTICE defaults rho = 1, u = 0 (LBM.h:559-563).  tau = 1/2 + nu/cs2
    with nu = u_max D / Re (LBM.h:102-103).  params['mea_box'] is the
    reference's MEA box (the int conversion of centre -+ D, k = 1..Nz as in
    main.cpp:246 for Nz = 1)."""
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    cx, cy = center if center is not None else (Nx / 5.0, Ny / 2.0)
    i = np.arange(Nx + 2)[:, None, None]
    j = np.arange(Ny + 2)[None, :, None]
    k = np.arange(Nz + 2)[None, None, :]
    rx, ry = i * 1.0 - 0.5, j * 1.0 - 0.5
    inside = np.sqrt((cx - rx) ** 2 + (cy - ry) ** 2) <= D / 2.0
    region = (i >= 1) & (i <= Nx) & (j >= 1) & (j <= Ny) & (k >= 0) & (k <= Nz)
    t[np.broadcast_to(inside & region, t.shape)] = TYPE_S
    t[0, :, :] = TYPE_IN
    t[-1, :, :] = TYPE_OUT
    t[:, 0, :] = TYPE_OUT
    t[:, -1, :] = TYPE_OUT
    if Nz != 1:
        t[:, :, 0] = TYPE_F
        t[:, :, -1] = TYPE_F
    live = t != TYPE_S
    ux[live] = u_max
    nu = u_max * D / Re
    tau = 0.5 + nu / (1.0 / 3.0)
    box = (max(0, int(cx - D)), int(cx + D), max(0, int(cy - D)), int(cy + D), 1, Nz)
    return Case("cylinder_bb", Nx, Ny, Nz, tau, t, rho, ux, uy, uz,
                params=dict(u_max=u_max, Re=Re, D=D, nu=nu, center=(cx, cy), mea_box=box,
                            mea_norm=0.5 * 1.0 * u_max ** 2 * D))


def sphere_bb(Nx: int = 800, Ny: int = 400, Nz: int = 400, D: float = 20.0, u_max: float = 0.1,
              Re: float = 200.0, center=None) -> Case:
    """main_setup_Sphere with Generate_Sphere (lbm.cpp:1759-1820): cells
    i = 1..Nx, j = 1..Ny, k = 0..Nz whose centre (i - 1/2, j - 1/2, k - 1/2)
    lies within D/2 of (Nx/5, Ny/2, Nz/2) become TYPE_S (rho = 1, u = 0);
    then i = 0 TYPE_IN and the other five faces TYPE_OUT; F/IN/OUT cells
    start at rho = 1, u = (u_max, 0, 0).  params['mea_box'] is the
    reference's MEA box (main.cpp:369) and params['mea_norm'] the
    C_D/C_L normalisation 1/2 rho0 u^2 D^2 PI/4 (lbm.cpp:3122)."""
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    cx, cy, cz = center if center is not None else (Nx / 5.0, Ny / 2.0, Nz / 2.0)
    i = np.arange(Nx + 2)[:, None, None]
    j = np.arange(Ny + 2)[None, :, None]
    k = np.arange(Nz + 2)[None, None, :]
    d = np.sqrt((cx - (i - 0.5)) ** 2 + (cy - (j - 0.5)) ** 2 + (cz - (k - 0.5)) ** 2)
    region = (i >= 1) & (i <= Nx) & (j >= 1) & (j <= Ny) & (k >= 0) & (k <= Nz)
    t[(d <= D / 2.0) & region] = TYPE_S
    t[0, :, :] = TYPE_IN
    t[-1, :, :] = TYPE_OUT
    t[:, 0, :] = TYPE_OUT
    t[:, -1, :] = TYPE_OUT
    t[:, :, 0] = TYPE_OUT
    t[:, :, -1] = TYPE_OUT
    live = t != TYPE_S
    ux[live] = u_max
    nu = u_max * D / Re
    tau = 0.5 + nu / (1.0 / 3.0)
    box = (int(cx - D), int(cx + D), int(cy - D), int(cy + D), int(cz - D), int(cz + D))
    return Case("sphere_bb", Nx, Ny, Nz, tau, t, rho, ux, uy, uz,
                params=dict(u_max=u_max, Re=Re, D=D, nu=nu, center=(cx, cy, cz), mea_box=box,
                            mea_norm=0.5 * 1.0 * u_max ** 2 * D * D * PI_REF / 4))


def cylinder_markers(center_x: float, center_y: float, D: float, n_theta: int,
                     z_levels) -> dict:
    """Markers on a cylinder of diameter D along z, in the pattern of
    Generate_Cylinder_Marker (lbm.cpp:1857-1877): x = r cos(2 pi n/N) +
    center_x + 0.5, y = r sin(2 pi n/N) + center_y + 0.5 (PI = 3.14159265),
    counter-clockwise, one ring per z level.  Normals follow
    MARKER::Normal_Marker (lbm.cpp:1048-1069): t1 = unit chord to the next
    marker of the ring, n = t1 x z_hat.  ds = chord length to the next
    marker (main.cpp:123-131) times the axial spacing (1 for a single 2D
    ring)."""
    z_levels = np.atleast_1d(np.asarray(z_levels, dtype=np.float64))
    r = D / 2.0
    a = 2 * PI_REF / n_theta * np.arange(n_theta)
    x = r * np.cos(a) + center_x + 0.5
    y = r * np.sin(a) + center_y + 0.5
    x2, y2 = np.roll(x, -1), np.roll(y, -1)
    dx_, dy_ = x2 - x, y2 - y
    norm = np.sqrt(dx_ ** 2 + dy_ ** 2)
    t1x, t1y = dx_ / norm, dy_ / norm
    # (t1x, t1y, 0) x (0, 0, 1) = (t1y, -t1x, 0)
    nxr, nyr = t1y, -t1x
    chord = np.sqrt((x - x2) ** 2 + (y - y2) ** 2)
    dz = 1.0 if z_levels.size < 2 else float(z_levels[1] - z_levels[0])
    nz_ = z_levels.size
    return dict(
        x=np.tile(x, nz_), y=np.tile(y, nz_), z=np.repeat(z_levels, n_theta),
        nx=np.tile(nxr, nz_), ny=np.tile(nyr, nz_), nz=np.zeros(n_theta * nz_),
        ds=np.tile(chord * dz, nz_),
    )


def ibm_cylinder(Nx: int = 60, Ny: int = 40, Nz: int = 1, D: float = 10.0, u_max: float = 0.04,
                 Re: float = 100.0, center=None, n_theta: int | None = None, seed: int = 1,
                 perturb: float = 0.0):
    """main_setup_Cylinder_IBM (lbm.cpp:1828-1855) with its markers: i = 0
    TYPE_IN, i = Nx+1 and j = 0, Ny+1 TYPE_OUT (slip faces), the k ghost
    layers TYPE_F (periodic); every cell starts at rho = 1, u = (u_max, 0, 0),
    TYPE_F cells optionally with the seeded perturbation of channel().  A
    cylinder of diameter D along z (one marker ring per lattice layer
    k = 1..Nz) and the reference's box [floor(c - D), ceil(c + D)] in x and
    y (lbm.cpp:1165-1166), k = 1..Nz.  Returns (case, markers, box)."""
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[0, :, :] = TYPE_IN
    t[-1, :, :] = TYPE_OUT
    t[:, 0, :] = TYPE_OUT
    t[:, -1, :] = TYPE_OUT
    ux[:] = u_max
    if perturb:
        cell = np.arange(t.size, dtype=np.uint64).reshape(t.shape)
        xi = [counter_uniform(seed, 3 * cell + np.uint64(c)) for c in range(3)]
        f = t == TYPE_F
        ux[f] = u_max * (1.0 + perturb * xi[0][f])
        uy[f] = perturb * u_max * xi[1][f]
        uz[f] = perturb * u_max * xi[2][f]
    cx, cy = center if center is not None else (Nx / 4.0, Ny / 2.0)
    if n_theta is None:
        n_theta = int(round(PI_REF * D))
    mk = cylinder_markers(cx, cy, D, n_theta, np.arange(1, Nz + 1, dtype=np.float64))
    box = (int(math.floor(cx - D)), int(math.ceil(cx + D)), int(math.floor(cy - D)),
           int(math.ceil(cy + D)), 1, Nz)
    nu = u_max * D / Re
    tau = 0.5 + 3.0 * nu
    case = Case("ibm_cylinder", Nx, Ny, Nz, tau, t, rho, ux, uy, uz,
                params=dict(u_max=u_max, Re=Re, D=D, nu=nu, center=(cx, cy), n_theta=n_theta,
                            vsolid=PI_REF * D * D / 4.0 * Nz))
    return case, mk, box


def fsi_thermal_channel(Nx: int = 192, Ny: int = 96, Nz: int = 96, u_in: float = 0.03, Re: float = 100.0,
                        Pr: float = 0.71, Ra: float = 1e4, n_markers: int = 128, length: float | None = None,
                        density_ratio: float = 1.5, K_b: float = 1e-3, incline_deg: float = 5.0,
                        seed: int = 1, perturb: float = 0.01, T_cold: float = 0.0, T_hot: float = 1.0,
                        layer: float = 2.0):
    """A flexible filament clamped at its leading end in a heated channel:
    i = 0 TYPE_IN, i = Nx+1 TYPE_OUT, y and z walls TYPE_S (channel()'s
    pattern); every wall Dirichlet, j = 0 at T_hot, the others at T_cold;
    Boussinesq F_y = g_beta (T - T_cold) with g_beta = Ra nu alpha / (dT H^3)
    over the height H = Ny.  The fluid and the inflow start at T_cold plus
    a thermal layer over the heated wall, (T_hot - T_cold) exp(-y/layer)
    at the distance y from it: the over-relaxed D3Q7 BGK (tau_g near 0.58
    at Pr 0.71) undershoots a temperature step by several per cent, a
    smooth start does not.  With T_cold = 0 most of the channel sits at the
    bottom of the range check (T < 0 flags LBM_ERR_TEMPERATURE_RANGE):
    upload it with the storage shift T0 = 0, so that zero stays exactly
    zero in FP32 and 16-bit storage.

    The filament (length L, default Ny/2) is clamped at (Nx/4, Ny/2, Nz/2),
    inclined by incline_deg in x-y, and follows the usual flapping-filament
    scales: m_l = density_ratio rho0 L, EI = K_b rho0 u_in^2 L^3,
    Re = u_in L / nu; Pr = nu / alpha (D3Q7: alpha = (tau_g - 1/2)/4).  The
    IBM box spans the filament's reach (L/2 + 4 across the flow, 4 cells
    about the clamp plane in z).  Returns (case, thermal, filament keyword
    arguments, box)."""
    t, rho, ux, uy, uz = _blank(Nx, Ny, Nz)
    t[0, :, :] = TYPE_IN
    t[-1, :, :] = TYPE_OUT
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    t[:, :, 0] = TYPE_S
    t[:, :, -1] = TYPE_S
    live = t != TYPE_S
    ux[live] = u_in
    if perturb:
        cell = np.arange(t.size, dtype=np.uint64).reshape(t.shape)
        xi = [counter_uniform(seed, 3 * cell + np.uint64(c)) for c in range(3)]
        f = t == TYPE_F
        ux[f] = u_in * (1.0 + perturb * xi[0][f])
        uy[f] = perturb * u_in * xi[1][f]
        uz[f] = perturb * u_in * xi[2][f]
    L = float(length) if length is not None else Ny / 2.0
    nu = u_in * L / Re
    tau = 0.5 + 3.0 * nu
    alpha = nu / Pr
    dT = T_hot - T_cold
    g_beta = Ra * nu * alpha / (dT * float(Ny) ** 3)
    y = np.arange(Ny + 2, dtype=np.float64)[None, :, None] - 0.5
    T = np.broadcast_to(T_cold + dT * np.exp(-y / layer), t.shape).copy()
    T[t == TYPE_S] = T_cold
    T[:, 0, :] = T_hot
    te = np.zeros(t.shape, np.int16)
    te[t == TYPE_S] = TYPE_T
    th = q7_thermal(T, te, q7_tau(alpha), boussinesq=True, g_beta=g_beta, T_ref=T_cold, alpha=alpha, H=Ny)
    a = math.radians(incline_deg)
    clamp = (Nx / 4.0, Ny / 2.0, Nz / 2.0)
    fil = dict(n=n_markers, length=L, line_density=density_ratio * L, bending=K_b * u_in ** 2 * L ** 3,
               clamp=clamp, tangent=(math.cos(a), math.sin(a), 0.0))
    reach = L / 2.0 + 4.0
    box = (max(1, int(math.floor(clamp[0] - 4))), min(Nx, int(math.ceil(clamp[0] + L + 4))),
           max(1, int(math.floor(clamp[1] - reach))), min(Ny, int(math.ceil(clamp[1] + reach))),
           max(1, int(math.floor(clamp[2] - 4))), min(Nz, int(math.ceil(clamp[2] + 4))))
    case = Case("fsi_thermal_channel", Nx, Ny, Nz, tau, t, rho, ux, uy, uz,
                params=dict(u_in=u_in, Re=Re, Pr=Pr, Ra=Ra, nu=nu, alpha=alpha, L=L, K_b=K_b,
                            density_ratio=density_ratio, T_cold=T_cold, T_hot=T_hot))
    return case, th, fil, box


def batch_cavities(count: int = 96, seed: int = 1, edges=(32, 48, 64, 96, 128),
                   u_range=(0.02, 0.08), re_range=(50.0, 1000.0)):
    """The seeded cavity ensemble: edge from `edges`, lid velocity and Re
    uniform in their ranges, all drawn with counter_uniform(seed, .)."""
    out = []
    for b in range(count):
        xi = (counter_uniform(seed, np.arange(3, dtype=np.uint64) + np.uint64(3 * b)) + 1.0) / 2.0
        e = edges[min(int(xi[0] * len(edges)), len(edges) - 1)]
        u = u_range[0] + (u_range[1] - u_range[0]) * xi[1]
        re = re_range[0] + (re_range[1] - re_range[0]) * xi[2]
        out.append(dict(N=int(e), u_lid=float(u), Re=float(re)))
    return out
