"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors, and cannot be
compiled here (lbm.cpp:12 includes Eigen 3.4.0, which is absent), so each
oracle function is checked against an independent route: textbook lattice
identities, Hermite-tensor algebra, closed-form flows (the reference's own
validation cases: Couette lbm.cpp:3195-3206, Poiseuille lbm.cpp:3355-3369),
exact conservation laws, and brute-force re-derivations of single passes on
tiny lattices.  Each test names the plausible mistake it would catch.
"""
import itertools
import math

import numpy as np
import pytest

import cases
from conftest import rel_l2_u


def _tables(oracle_mod, q=19):
    o = oracle_mod.Oracle(2, 2, 2, 0.8, q=q)
    return o.tables()


# ---- a1: velocity set ----------------------------------------------------------
@pytest.mark.parametrize("q", [19, 9])
def test_velocity_set_identities(oracle_mod, q):
    cx, cy, cz, opp, w = _tables(oracle_mod, q)
    c = np.stack([cx, cy, cz], 1).astype(float)
    # the set is exactly the textbook one: all vectors with |c|^2 <= 2 (D3Q19)
    want = {v for v in itertools.product((-1, 0, 1), repeat=3) if sum(x * x for x in v) <= 2}
    if q == 9:
        want = {v for v in itertools.product((-1, 0, 1), (-1, 0, 1), (0,))}
    assert {tuple(int(x) for x in v) for v in c} == want and len(c) == q
    # weights by speed class (Krueger et al. 2017, tables 3.1/3.2)
    c2 = (c ** 2).sum(1)
    wref = {19: {0: 1 / 3, 1: 1 / 18, 2: 1 / 36}, 9: {0: 4 / 9, 1: 1 / 9, 2: 1 / 36}}[q]
    assert np.allclose(w, [wref[int(s)] for s in c2], rtol=0, atol=1e-16)
    # opposite pairs
    assert np.array_equal(c[opp], -c)
    assert np.array_equal(opp[opp], np.arange(q))
    # isotropy: sum w = 1, sum w c = 0, sum w c c = cs2 I, 4th order
    cs2 = 1 / 3
    assert abs(w.sum() - 1) < 1e-15
    assert np.abs(w @ c).max() < 1e-16
    d = 3 if q == 19 else 2
    I = np.eye(3)
    I[2, 2] = 1.0 if q == 19 else 0.0
    assert np.allclose(np.einsum("l,la,lb->ab", w, c, c), cs2 * I, atol=1e-15)
    m4 = np.einsum("l,la,lb,lc,ld->abcd", w, c, c, c, c)
    iso = cs2 ** 2 * (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I)
                      + np.einsum("ad,bc->abcd", I, I))
    assert np.allclose(m4, iso, atol=1e-15), d


# ---- a3: cubic equilibrium --------------------------------------------------------
def test_feq_is_third_order_hermite_expansion(oracle_mod):
    """feq (lbm.cpp:628-633) equals w rho [1 + H1:u/cs2 + H2:uu/(2cs2^2) +
    H3:uuu/(6 cs2^3)] with the Hermite tensors built by einsum -- a different
    route than the closed form.  Catches a dropped term, a sign, a wrong
    power of cs2."""
    o = oracle_mod.Oracle(2, 2, 2, 0.8)
    cx, cy, cz, opp, w = o.tables()
    c = np.stack([cx, cy, cz], 1).astype(float)
    cs2 = 1 / 3
    I = np.eye(3)
    rng = np.random.default_rng(3)
    for _ in range(20):
        rho = 1 + 0.1 * rng.standard_normal()
        u = 0.1 * rng.standard_normal(3)
        for l in range(19):
            cl = c[l]
            H2 = np.outer(cl, cl) - cs2 * I
            H3 = (np.einsum("a,b,c->abc", cl, cl, cl)
                  - cs2 * (np.einsum("a,bc->abc", cl, I) + np.einsum("b,ac->abc", cl, I)
                           + np.einsum("c,ab->abc", cl, I)))
            ref = w[l] * rho * (1 + cl @ u / cs2 + np.einsum("ab,a,b", H2, u, u) / (2 * cs2 ** 2)
                                + np.einsum("abc,a,b,c", H3, u, u, u) / (6 * cs2 ** 3))
            assert o.feq(rho, *u, l) == pytest.approx(ref, rel=1e-13, abs=1e-16)


def test_feq_moments(oracle_mod):
    """sum feq = rho and sum feq c = rho u exactly (the cubic terms cancel by
    odd symmetry); second moment rho(cs2 I + u u) by 4th-order isotropy."""
    o = oracle_mod.Oracle(2, 2, 2, 0.8)
    cx, cy, cz, _, w = o.tables()
    c = np.stack([cx, cy, cz], 1).astype(float)
    rho, u = 1.03, np.array([0.04, -0.02, 0.03])
    f = np.array([o.feq(rho, *u, l) for l in range(19)])
    assert abs(f.sum() - rho) < 1e-15
    assert np.abs(f @ c - rho * u).max() < 1e-16
    assert np.allclose(np.einsum("l,la,lb->ab", f, c, c), rho * (np.eye(3) / 3 + np.outer(u, u)),
                       atol=1e-15)
    assert np.allclose([o.feq(1.0, 0, 0, 0, l) for l in range(19)], w, atol=1e-17)


# ---- a4: Init ---------------------------------------------------------------------
def test_init_sets_feq_everywhere(oracle_mod):
    case = cases.random_mask_case(4, 3, 5, seed=1)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.set_cells(**case.fields())
    o.init()
    f, fs = o.f(), o.fstar()
    assert np.array_equal(f, fs)
    for idx in [(0, 0, 0), (2, 1, 3), (5, 4, 6), (3, 2, 1)]:
        want = [o.feq(case.rho[idx], case.ux[idx], case.uy[idx], case.uz[idx], l) for l in range(19)]
        assert np.array_equal(f[idx], want)


# ---- a5: collision ----------------------------------------------------------------
def test_collision_tau_one_gives_feq_and_rest_is_fixed_point(oracle_mod):
    case = cases.random_mask_case(3, 3, 3, seed=2, tau=1.0)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, 1.0)
    o.set_cells(**case.fields())
    o.init()
    rng = np.random.default_rng(0)
    o.set_f(o.f() * (1 + 0.1 * rng.standard_normal(o.f().shape)))
    assert o.collision() == 0
    fs = o.fstar()
    fl = case.type == 0
    for idx in zip(*np.nonzero(fl)):
        want = [o.feq(case.rho[idx], case.ux[idx], case.uy[idx], case.uz[idx], l) for l in range(19)]
        assert np.array_equal(fs[idx], want)  # omega = 1: f*(1-1) + 1*feq exactly
    # rest state at tau = 0.7: f = feq is a fixed point of collision
    o2 = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, 0.7)
    o2.set_cells(**case.fields())
    o2.init()
    f0 = o2.f()
    o2.collision()
    assert np.abs(o2.fstar() - f0)[fl].max() < 4e-16  # a few ulp of f ~ 0.4


def test_guo_force_momentum_closed_form(oracle_mod):
    """Uniform force F in a fully periodic uniform box.  Collision adds
    (1-omega/2)F of momentum at step 1 (u is the setup u = 0) and exactly F
    per step afterwards, since MacroProp's u carries +F/2 (lbm.cpp:673-675).
    So after n steps u = (n - omega/2 + 1/2) F.  Catches a missing 1/2 F in
    MacroProp, a wrong (1-omega/2) prefactor or a wrong source sign."""
    tau, n = 0.8, 12
    om = 1 / tau
    F = np.array([2e-5, -1e-5, 3e-5])
    shape = (5, 4, 3)
    o = oracle_mod.Oracle(3, 2, 1, tau)
    o.set_cells(type=np.zeros(shape, np.int16), rho=1.0, ux=0.0, uy=0.0, uz=0.0,
                Fx=F[0], Fy=F[1], Fz=F[2])
    o.init()
    assert o.run(n) == 0
    rho, ux, uy, uz = o.macro()
    want = (n - om / 2 + 0.5) * F
    for comp, wv in zip((ux, uy, uz), want):
        assert np.allclose(comp, wv, rtol=1e-11, atol=0)
    assert np.allclose(rho, 1.0, atol=1e-14)


# ---- a6/a7: streaming ----------------------------------------------------------------
def test_periodic_streaming_is_the_roll_permutation(oracle_mod):
    """All-fluid box: streaming is f(x, m) = f*(x - c_m, m) with modulo-(N+2)
    wrap (lbm.cpp:197-199), i.e. numpy.roll by +c_m.  Catches a push/pull
    direction flip or a wrong wrap."""
    shape = (6, 5, 4)
    o = oracle_mod.Oracle(4, 3, 2, 0.8)
    o.set_cells(type=np.zeros(shape, np.int16))
    o.init()
    cx, cy, cz, _, _ = o.tables()
    fs = np.random.default_rng(1).random(shape + (19,))
    o.set_fstar(fs)
    o.streaming()
    f = o.f()
    for m in range(19):
        assert np.array_equal(f[..., m], np.roll(fs[..., m], (cx[m], cy[m], cz[m]), axis=(0, 1, 2)))


def _brute_stream(case, fs, f_old, cx, cy, cz, opp, w):
    """Push streaming of lbm.cpp:186-292 re-derived cell by cell."""
    t, ux, uy, uz = case.type, case.ux, case.uy, case.uz
    S = t.shape
    f = f_old.copy()
    for i, j, k in itertools.product(*map(range, S)):
        if t[i, j, k] != 0:
            continue
        nbs = [((i + cx[l]) % S[0], (j + cy[l]) % S[1], (k + cz[l]) % S[2]) for l in range(19)]
        mov = [n for n in nbs if t[n] == 1 and (ux[n] != 0 or uy[n] != 0 or uz[n] != 0)]
        uw = np.zeros(3)
        for n in mov:
            uw += (ux[n], uy[n], uz[n])
        if mov:
            uw /= len(mov)
        for l, n in enumerate(nbs):
            if t[n] == 0:
                f[n + (l,)] = fs[i, j, k, l]
            elif t[n] == 1:
                udc = uw[0] * cx[l] + uw[1] * cy[l] + uw[2] * cz[l]
                f[i, j, k, opp[l]] = fs[i, j, k, l] - 2 * w[l] * udc / (1 / 3)
    return f


def test_streaming_walls_frozen_links_brute_force(oracle_mod):
    """Random lattice with moving/stationary solids, inlet-type cells and the
    periodic ghost shell: one Streaming pass equals a brute-force re-derivation
    of the push loop, including the per-cell averaged wall velocity applied to
    every wall link (lbm.cpp:941-962) and frozen links from non-F, non-S
    neighbours.  Catches opp/sign/index slips in the bounce-back."""
    case = cases.random_mask_case(4, 3, 3, seed=5)
    case.type[0, :, :] = 0  # keep the i faces free of inlet cells (face loops)
    case.type[-1, :, :] = 0
    case.type[:, 0, :][case.type[:, 0, :] == 3] = 0
    case.type[:, -1, :][case.type[:, -1, :] == 3] = 0
    case.type[:, :, 0][case.type[:, :, 0] == 3] = 0
    case.type[:, :, -1][case.type[:, :, -1] == 3] = 0
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.set_cells(**case.fields())
    o.init()
    cx, cy, cz, opp, w = o.tables()
    rng = np.random.default_rng(9)
    fs = rng.random(case.shape + (19,))
    f_old = rng.random(case.shape + (19,))
    o.set_fstar(fs)
    o.set_f(f_old)
    o.streaming()
    got = o.f()
    want = _brute_stream(case, fs, f_old, cx, cy, cz, opp, w)
    fl = case.type == 0
    assert np.abs(got - want)[fl].max() < 1e-15
    assert np.array_equal(got[~fl], f_old[~fl])  # no face types present: non-F untouched
    # the exported pull tables describe the same pass
    src, kind = o.links()
    S = case.shape
    lin = lambda i, j, k: (i * S[1] + j) * S[2] + k  # noqa: E731
    for i, j, k in zip(*np.nonzero(fl)):
        for m in range(19):
            s = ((i - cx[m]) % S[0], (j - cy[m]) % S[1], (k - cz[m]) % S[2])
            assert src[i, j, k, m] == lin(*s)
            ts = case.type[s]
            assert kind[i, j, k, m] == (0 if ts == 0 else (1 if ts == 1 else 2))
    assert (kind[~fl] == 255).all() and (src[~fl] == -1).all()


def test_back_face_inflow_writes_the_top_layer(oracle_mod):
    """lbm.cpp:599-611: for a TYPE_IN cell at k = 0 the back-face loop
    computes feq2 from that cell's rho,u and stores it in cell (i,j,Nz+1),
    which may be fluid or solid.  Re-derived here from the 2nd-order
    equilibrium rho w (1 + 3cu + 4.5cu^2 - 1.5uu)."""
    shape = (5, 4, 5)
    t = np.zeros(shape, np.int16)
    t[2, 1, 0] = 3       # inlet cell on the back face
    t[1, 2, 0] = 3
    t[1, 2, -1] = 1      # its top partner is solid
    rng = np.random.default_rng(2)
    ux, uy, uz = (0.02 * rng.standard_normal(shape) for _ in range(3))
    rho = 1 + 0.01 * rng.standard_normal(shape)
    o = oracle_mod.Oracle(3, 2, 3, 0.7)
    o.set_cells(type=t, rho=rho, ux=ux, uy=uy, uz=uz)
    o.init()
    cx, cy, cz, _, w = o.tables()
    o.set_fstar(rng.random(shape + (19,)))
    o.streaming()
    f = o.f()
    for (i, j) in ((2, 1), (1, 2)):
        u = np.array([ux[i, j, 0], uy[i, j, 0], uz[i, j, 0]])
        cu = cx * u[0] + cy * u[1] + cz * u[2]
        want = rho[i, j, 0] * w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u))
        assert np.allclose(f[i, j, -1], want, rtol=1e-14, atol=0)


def test_face_loops_copy_and_order(oracle_mod):
    """lbm.cpp:299-625: after the push, outlet cells copy three populations
    of their inner neighbour (the D2Q9 indices of the reference's code:
    {1,5,8} at i = 0), face by face in the order i = 0, i = Nx+1, j = Ny+1,
    j = 0, k = Nz+1, k = 0.  With every shell cell an outlet:
    - a face-interior outlet at i = 0 holds its fluid neighbour's streamed
      f1, f5, f8 and its own previous values in the other 16 directions;
    - an edge cell that a later loop visits again takes, through the earlier
      loop's write, the value of the fluid cell diagonally in:
      (0,0,k).f5 = (1,1,k).f5 (i = 0 writes (0,1,k).f5, then j = 0 copies
      it), (0,Ny+1,k).f8 = (1,Ny,k).f8 and (0,j,Nz+1).f8 = (1,j,Nz).f8.
    Any order with j = Ny+1, j = 0 or k = Nz+1 before i = 0 leaves the
    previous step's values in those edge cells instead."""
    Nx, Ny, Nz = 4, 3, 3
    shape = (Nx + 2, Ny + 2, Nz + 2)
    t = np.full(shape, 2, np.int16)
    t[1:-1, 1:-1, 1:-1] = 0
    rng = np.random.default_rng(12)
    o = oracle_mod.Oracle(Nx, Ny, Nz, 0.8)
    o.set_cells(type=t)
    o.init()
    f_old = rng.random(shape + (19,))
    o.set_f(f_old)
    o.set_fstar(rng.random(shape + (19,)))
    o.streaming()
    f = o.f()
    copied = [1, 5, 8]
    kept = [m for m in range(19) if m not in copied]
    face = (slice(1, Ny + 1), slice(1, Nz + 1))
    assert np.array_equal(f[0][face][..., copied], f[1][face][..., copied])
    assert np.array_equal(f[0][face][..., kept], f_old[0][face][..., kept])
    ks, js = slice(1, Nz + 1), slice(1, Ny + 1)
    assert np.array_equal(f[0, 0, ks, 5], f[1, 1, ks, 5])
    assert np.array_equal(f[0, Ny + 1, ks, 8], f[1, Ny, ks, 8])
    assert np.array_equal(f[0, js, Nz + 1, 8], f[1, js, Nz, 8])
    # the fluid values differ from what the shell held before (the checks
    # above compare distinct numbers)
    assert not np.any(f[1, 1, ks, 5] == f_old[0, 1, ks, 5])


def test_inflow_cells_closed_forms(oracle_mod):
    """lbm.cpp:325-337 with MacroProp (lbm.cpp:668-675): an inlet's
    populations become the second-order equilibrium of its own previous
    rho, u, whose moments are rho and rho u again.  So its rho stays at the
    set-up value and its u grows by F/(2 rho) per step:
    u(n) = u0 + n F/(2 rho0), and u(n) = u0 when F = 0."""
    Nx, Ny, Nz = 4, 3, 3
    shape = (Nx + 2, Ny + 2, Nz + 2)
    t = np.zeros(shape, np.int16)
    t[0, 1:-1, 1:-1] = 3
    rng = np.random.default_rng(5)
    rho = 1 + 0.02 * rng.standard_normal(shape)
    u = 0.03 * rng.standard_normal((3,) + shape)
    F = 1e-4 * rng.standard_normal((3,) + shape)
    inl = t == 3
    for force, steps in ((np.zeros_like(F), 4), (F, 6)):
        o = oracle_mod.Oracle(Nx, Ny, Nz, 0.8)
        o.set_cells(type=t, rho=rho, ux=u[0], uy=u[1], uz=u[2], Fx=force[0], Fy=force[1], Fz=force[2])
        o.init()
        assert o.run(steps) == 0
        r, ux, uy, uz = o.macro()
        assert np.allclose(r[inl], rho[inl], rtol=1e-14, atol=0)
        for q, comp in enumerate((ux, uy, uz)):
            want = u[q][inl] + steps * 0.5 * force[q][inl] / rho[inl]
            assert np.abs(comp[inl] - want).max() < 1e-15


def test_stationary_walls_conserve_mass(oracle_mod):
    case = cases.cavity(8, u_lid=0.0, Re=100.0)
    rng = np.random.default_rng(4)
    fl = case.type == 0
    case.ux[fl] = 0.02 * rng.standard_normal(fl.sum())
    case.uy[fl] = 0.02 * rng.standard_normal(fl.sum())
    case.tau = 0.7
    o = oracle_mod.run_case(case, 0)
    m0 = o.f()[fl].sum()
    o.run(60)
    assert abs(o.f()[fl].sum() - m0) / m0 < 1e-13


# ---- a9: MacroProp --------------------------------------------------------------------
def test_macroprop_moments(oracle_mod):
    case = cases.random_mask_case(3, 4, 3, seed=3)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    rng = np.random.default_rng(2)
    F = 1e-3 * rng.standard_normal((3,) + case.shape)
    o.set_cells(type=case.type, Fx=F[0], Fy=F[1], Fz=F[2])
    o.init()
    f = rng.random(case.shape + (19,))
    o.set_f(f)
    assert o.macroprop() == 0
    cx, cy, cz, _, _ = o.tables()
    rho, ux, uy, uz = o.macro()
    assert np.allclose(rho, f.sum(-1), rtol=1e-15)
    for comp, cc, Fc in ((ux, cx, F[0]), (uy, cy, F[1]), (uz, cz, F[2])):
        assert np.allclose(comp, (f @ cc + 0.5 * Fc) / f.sum(-1), rtol=1e-12, atol=1e-15)


def test_error_codes_match_the_terminate_sites(oracle_mod):
    case = cases.cavity(4, u_lid=0.05, Re=100.0)
    o = oracle_mod.run_case(case, 0)
    f = o.f()
    f[2, 2, 2, 0] = 150.0  # density > 100 (lbm.cpp:772-775)
    o.set_f(f)
    assert o.macroprop() == 2
    case.ux[2, 2, 2] = 1.5  # |u| large enough for feq < 0 (lbm.cpp:114-117)
    o2 = oracle_mod.run_case(case, 0)
    assert o2.collision() == 1


# ---- closed-form flows (the reference's own validation cases) --------------------------
def test_couette_linear_profile(oracle_mod):
    """Steady Couette flow between a stationary wall (j=0) and a wall moving
    at U (j=Ny+1) with half-way bounce-back: u(j) = U (j - 1/2)/Ny exactly
    (the t -> inf limit of Couette_Analytic_Velocity, lbm.cpp:3195-3206).
    Catches a wrong sign or factor in the moving-wall term."""
    U = 0.01
    case = cases.couette(Nx=2, Ny=16, Nz=2, tau=0.9, u_wall=U)
    o = oracle_mod.run_case(case, 12000)
    rho, ux, uy, uz = o.macro()
    j = np.arange(1, 17)
    prof = ux[2, 1:17, 2]
    assert np.abs(prof - U * (j - 0.5) / 16).max() < 1e-9 * U
    assert np.abs(uy[:, 1:17, :]).max() < 1e-12


def test_poiseuille_parabolic_profile(oracle_mod):
    """Body-force-driven channel: u(j) = g/(2 nu) (j - 1/2)(Ny + 1/2 - j).
    With BGK and half-way bounce-back the discrete steady state equals the
    parabola up to a tau-dependent wall slip that vanishes at
    (tau - 1/2)^2 = 3/16 (Ginzburg & d'Humieres 2003); at that tau the match
    is exact.  Catches a wrong force projection or viscosity relation."""
    g = 1e-6
    tau = 0.5 + np.sqrt(3.0 / 16.0)
    nu = (tau - 0.5) / 3
    j = np.arange(1, 13)
    want = g / (2 * nu) * (j - 0.5) * (12 + 0.5 - j)
    # force on the fluid only: the parabola, exactly
    case = cases.poiseuille(Nx=2, Ny=12, Nz=2, tau=tau, g=g)
    case.Fx[case.type == 1] = 0.0
    o = oracle_mod.run_case(case, 8000)
    rho, ux, uy, uz = o.macro()
    assert np.abs(ux[2, 1:13, 2] - want).max() < 1e-9 * want.max()
    # away from the magic tau the wall slip is O(1e-3..1e-2), as theory says
    case.tau = 0.8
    o = oracle_mod.run_case(case, 12000)
    nu8 = 0.3 / 3
    want8 = g / (2 * nu8) * (j - 0.5) * (12 + 0.5 - j)
    err8 = np.abs(o.macro()[1][2, 1:13, 2] - want8).max() / want8.max()
    assert 1e-4 < err8 < 2e-2
    # force in every cell, solids included (the reference's Poiseuille sets F
    # in all cells from MacroProp, lbm.cpp:800-801): MacroProp gives the
    # solids u = F/2 (lbm.cpp:673), Determine_Wall_Velocity then treats both
    # walls as moving at g/2, and the whole profile shifts by exactly g/2.
    case = cases.poiseuille(Nx=2, Ny=12, Nz=2, tau=tau, g=g)
    o = oracle_mod.run_case(case, 8000)
    assert np.abs(o.macro()[1][2, 1:13, 2] - (want + g / 2)).max() < 1e-9 * want.max()


# ---- a11-a17: IBM pieces ------------------------------------------------------------------
def test_peskin_kernel_properties(oracle_mod):
    k = oracle_mod.kernel
    assert k(0.0) == 0.5 and k(1.0) == 0.25 and k(-1.0) == 0.25
    assert k(2.0) == 0.0 and k(-2.0) == 0.0 and k(2.5) == 0.0
    rng = np.random.default_rng(0)
    for r in rng.random(50):
        js = np.arange(-3, 4)
        v = np.array([k(r - j) for j in js])
        assert abs(v.sum() - 1) < 1e-15                    # partition of unity
        assert abs(((r - js) * v).sum()) < 1e-15           # first moment
        assert abs((v ** 2).sum() - 3 / 8) < 1e-15         # Peskin 2002, eq. (6.8)
        e = v[(js % 2) == 0].sum()
        assert abs(e - 0.5) < 1e-15                        # even/odd condition


def test_nearest_point_is_farthest_along_normal(oracle_mod):
    rng = np.random.default_rng(1)
    box = (3, 9, 4, 11, 1, 3)
    pts = np.array(list(itertools.product(range(3, 10), range(4, 12), range(1, 4))), float)
    for _ in range(40):
        m = rng.random(3) * (4, 5, 1) + (4, 6, 1.5)
        n = rng.standard_normal(3)
        n[2] = 0.0
        n /= np.linalg.norm(n)
        i, j, kk = oracle_mod.nearest(box, m, n)
        d = (pts - m) @ n
        best = np.nonzero(d == d.max())[0][-1]  # last in x,y,z loop order
        assert (i, j, kk) == tuple(int(v) for v in pts[best])


def _uniform_ibm_lattice(oracle_mod, u0, nz=1):
    shape = (16, 16, nz + 2)
    o = oracle_mod.Oracle(14, 14, nz, 0.8)
    o.set_cells(type=np.zeros(shape, np.int16), rho=1.0, ux=u0[0], uy=u0[1], uz=u0[2])
    o.init()
    return o


def test_ibm_uniform_flow_interpolates_exactly_and_forces_balance(oracle_mod):
    """Uniform flow u0 past markers held at rest, m_max = 1: interpolation of a
    uniform field returns u0 times the z-weight (0.5 in the 2D slab,
    lbm.cpp:1167 -> phi(0)), so F_b = -2 rho u0 * 0.5 and Drag = sum ds rho u0.
    With markers moving at u0 (u_des = u_fluid in 3D), the force vanishes."""
    u0 = (0.03, -0.01, 0.0)
    o = _uniform_ibm_lattice(oracle_mod, u0)
    mk = cases.cylinder_markers(7.2, 7.4, 6.0, 24, [1.0])
    err, out = o.ibm_step(mk, (0, 0, 0), (0, 0, 0), (2, 13, 2, 13, 1, 1), m_max=1)
    assert err == 0
    ub = mk["ub"].reshape(-1, 3)
    # first interpolation before any spreading is overwritten; check the force
    Fb = mk["Fb"].reshape(-1, 3)
    assert np.allclose(Fb[:, 0], -2 * 1.0 * u0[0] * 0.5, rtol=1e-12)
    assert np.allclose(Fb[:, 1], -2 * 1.0 * u0[1] * 0.5, rtol=1e-12)
    assert out[0] == pytest.approx(-(Fb[:, 0] * mk["ds"]).sum(), rel=1e-13)
    assert np.all(np.isfinite(ub))
    # 3D: markers through the full z extent, moving with the flow -> no force
    o3 = _uniform_ibm_lattice(oracle_mod, u0, nz=6)
    mk3 = cases.cylinder_markers(7.2, 7.4, 6.0, 24, [3.0, 4.0])
    err, out3 = o3.ibm_step(mk3, u0, u0, (2, 13, 2, 13, 1, 6), m_max=3)
    assert np.abs(mk3["Fb"]).max() < 1e-14
    assert abs(out3[0]) < 1e-12 and abs(out3[1]) < 1e-12


def test_cylinder_markers_geometry():
    mk = cases.cylinder_markers(51.2, 64.0, 20.0, 64, [1.0])
    r = np.hypot(mk["x"] - 51.7, mk["y"] - 64.5)
    assert np.allclose(r, 10.0, rtol=1e-9)
    # outward normals (t1 x z_hat for counter-clockwise markers)
    dot = (mk["x"] - 51.7) * mk["nx"] + (mk["y"] - 64.5) * mk["ny"]
    assert (dot > 0).all()
    assert np.allclose(np.hypot(mk["nx"], mk["ny"]), 1.0)
    assert mk["ds"].sum() == pytest.approx(64 * 20 * np.sin(np.pi / 64), rel=1e-6)


# ---- case helpers ------------------------------------------------------------------------
def test_splitmix64_known_answers():
    # splitmix64 reference sequence for seed 0 (Vigna's splitmix64.c)
    z = cases.splitmix64(np.array([0, 0x9E3779B97F4A7C15], dtype=np.uint64))
    assert int(z[0]) == 0xE220A8397B1DCDAF
    assert int(z[1]) == 0x6E789E6AA1B965F4


def test_channel_mask_precedence():
    c = cases.channel(8, 6, 4, perturb=0.01)
    assert (c.type[0, 1:-1, 1:-1] == 3).all() and (c.type[-1, 1:-1, 1:-1] == 2).all()
    assert (c.type[:, 0, :] == 1).all() and (c.type[:, :, -1] == 1).all()
    f = c.type == 0
    assert np.abs(c.ux[f] / 0.04 - 1).max() <= 0.01 and np.abs(c.uy[f]).max() <= 0.0004
    c2 = cases.channel(8, 6, 4, perturb=0.01)
    assert np.array_equal(c.ux, c2.ux)  # seeded


def test_tgv_divergence_free_and_rho():
    c = cases.taylor_green(16)
    k = c.params["k"]
    # analytic divergence: du/dx + dv/dy = u0 k (cos cos - cos cos) = 0
    assert np.abs(c.uz).max() == 0
    assert c.tau == pytest.approx(0.5 + 3 * 0.04 * 16 / (2 * np.pi) / 1600)
    assert (c.type == 0).all() and c.shape == (16, 16, 16)
    assert rel_l2_u((c.ux, c.uy, c.uz), (c.ux, c.uy, c.uz), c.type == 0) == 0.0
    assert k == pytest.approx(2 * np.pi / 16)


# ---- a10: thermal branch (ENERGY) ----------------------------------------------------------
def _thermal_oracle(oracle_mod, case, th):
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.thermal_enable(th.tau_g, th.nu, th.cp, th.j_dot_n, th.boussinesq, th.g_beta, th.T_ref)
    o.set_cells(**case.fields())
    o.set_thermal(th.T, th.type_energy)
    return o


def test_thermal_geq_moments(oracle_mod):
    """Calc_geq (lbm.cpp:646-651) = rho w T (1 + cubic): sum_l geq = rho T and
    sum_l geq c = rho T u exactly in exact arithmetic (the cubic terms have no
    zeroth or first moment on D3Q19); the solid branch with zero wall velocity
    is rho w T_w."""
    case = cases.random_mask_case(3, 3, 3, seed=2)
    rng = np.random.default_rng(9)
    T = 0.3 + rng.random(case.shape)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.thermal_enable(0.8, 0.1, 2.0)
    o.set_cells(**case.fields())
    o.set_thermal(T, None)
    cx, cy, cz, opp, w = o.tables()
    for (i, j, k) in [(1, 1, 1), (2, 3, 1), (4, 0, 2)]:
        g = np.array([o.geq(i, j, k, l) for l in range(19)])
        r, u = case.rho[i, j, k], np.array([case.ux[i, j, k], case.uy[i, j, k], case.uz[i, j, k]])
        assert g.sum() == pytest.approx(r * T[i, j, k], rel=1e-14)
        for c, ua in zip((cx, cy, cz), u):
            assert (g * c).sum() == pytest.approx(r * T[i, j, k] * ua, rel=1e-12, abs=1e-16)
    # solid branch, u_w = 0: rho w T_w with T_w = Determine_Wall_Temperature
    i, j, k = 1, 1, 1
    tw = o.wall_temperature(i, j, k)
    for l in range(19):
        assert o.geq(i, j, k, l, as_solid=True) == pytest.approx(case.rho[i, j, k] * w[l] * tw, rel=1e-15)


def test_thermal_wall_temperature_brute_force(oracle_mod):
    """Determine_Wall_Temperature (lbm.cpp:965-981): mean T over the TYPE_S
    neighbours x + c_l, wrapped, l = 0 included."""
    case, th = cases.random_thermal_case(5, 4, 3, seed=4)
    o = _thermal_oracle(oracle_mod, case, th)
    cx, cy, cz, opp, w = o.tables()
    n = case.shape
    hit = 0
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                vals = [th.T[(i + cx[l]) % n[0], (j + cy[l]) % n[1], (k + cz[l]) % n[2]]
                        for l in range(19)
                        if case.type[(i + cx[l]) % n[0], (j + cy[l]) % n[1], (k + cz[l]) % n[2]] == cases.TYPE_S]
                if not vals:
                    continue
                hit += 1
                assert o.wall_temperature(i, j, k) == pytest.approx(np.mean(vals), rel=1e-14)
    assert hit > 50


def test_thermal_face_loops_closed_forms(oracle_mod):
    """ENERGY part of the face loops (lbm.cpp:307-312, 339-345, 523-527,
    611-617): an inlet's g becomes rho w T (1 + ...) of its own rho, u, T,
    which sums to rho T, so MacroProp's T = sum g grows by the factor rho
    per step, T(n) = rho^n T(0), while rho and u stay; an outlet on the
    i = Nx+1 face holds its inner neighbour's streamed g3, g6, g7 and its
    own previous g elsewhere; the k = 0 inlet writes g of cell (i,j,Nz+1)
    with the T of (i,j,0)."""
    Nx, Ny, Nz = 4, 3, 3
    shape = (Nx + 2, Ny + 2, Nz + 2)
    t = np.zeros(shape, np.int16)
    t[0, 1:-1, 1:-1] = 3      # inlets, i = 0 face interior
    t[-1, 1:-1, 1:-1] = 2     # outlets, i = Nx+1 face interior
    t[2, 2, 0] = 3            # an inlet on the k = 0 face ...
    t[2, 2, -1] = 2           # ... under an outlet at k = Nz+1
    rng = np.random.default_rng(8)
    rho = 1 + 0.03 * rng.standard_normal(shape)
    u = 0.02 * rng.standard_normal((3,) + shape)
    T = 0.3 + 0.4 * rng.random(shape)
    te = np.zeros(shape, np.int16)
    nu = (0.8 - 0.5) / 3
    alpha, tau_g, k, cp = cases.thermal_constants(nu, 0.7)
    o = oracle_mod.Oracle(Nx, Ny, Nz, 0.8)
    o.thermal_enable(tau_g, nu, cp)
    o.set_cells(type=t, rho=rho, ux=u[0], uy=u[1], uz=u[2])
    o.set_thermal(T, te)
    o.init()
    steps = 5
    assert o.run(steps) == 0
    Tn = o.thermal()[0]
    r = o.macro()[0]
    inl = np.zeros(shape, bool)
    inl[0, 1:-1, 1:-1] = True
    assert np.allclose(r[inl], rho[inl], rtol=1e-14, atol=0)
    assert np.allclose(Tn[inl], rho[inl] ** steps * T[inl], rtol=1e-13, atol=0)
    # one more Streaming: the outlets' copies and the k = 0 write
    g_old = o.g()
    assert o.collision() == 0
    o.streaming()
    g = o.g()
    copied, kept = [3, 6, 7], [m for m in range(19) if m not in (3, 6, 7)]
    face = (slice(1, Ny + 1), slice(1, Nz + 1))
    assert np.array_equal(g[-1][face][..., copied], g[-2][face][..., copied])
    assert np.array_equal(g[-1][face][..., kept], g_old[-1][face][..., kept])
    cx, cy, cz, _, w = o.tables()
    rr, ux, uy, uz = o.macro()
    Tk = o.thermal()[0]
    uvec = np.array([ux[2, 2, 0], uy[2, 2, 0], uz[2, 2, 0]])
    cu = cx * uvec[0] + cy * uvec[1] + cz * uvec[2]
    want = rr[2, 2, 0] * w * Tk[2, 2, 0] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (uvec @ uvec))
    assert np.allclose(g[2, 2, -1], want, rtol=1e-13, atol=0)


def test_thermal_solid_temperature_after_first_macroprop(oracle_mod):
    """Solids keep their Init g = rho_s w T_w (lbm.cpp:637-641) and MacroProp
    runs over them (lbm.cpp:777-790), so after the first step a solid's T is
    rho_s times the mean set-up T of its own solid neighbours: on the walls
    of a channel that mixes in the opposite wall through the periodic wrap,
    (9 T_own + 5 T_other)/14."""
    case, th = cases.conduction_channel(Tb=1.0, Tt=0.5)
    o = _thermal_oracle(oracle_mod, case, th)
    o.init()
    assert o.run(1) == 0
    T = o.thermal()[0]
    assert np.allclose(T[:, 0, :], (9 * 1.0 + 5 * 0.5) / 14, rtol=0, atol=1e-15)
    assert np.allclose(T[:, -1, :], (9 * 0.5 + 5 * 1.0) / 14, rtol=0, atol=1e-15)


def test_thermal_dirichlet_conduction_linear_profile(oracle_mod):
    """Anti-bounce-back Dirichlet walls (lbm.cpp:277-280), fluid at rest:
    the steady state is linear between the walls' temperatures (after the
    first MacroProp) at the half-way wall positions j = 1/2, Ny + 1/2."""
    case, th = cases.conduction_channel(bottom=cases.TYPE_T, top=cases.TYPE_T)
    o = _thermal_oracle(oracle_mod, case, th)
    o.init()
    assert o.run(4000) == 0
    T = o.thermal()[0]
    Tb, Tt = th.params["Tb"], th.params["Tt"]
    Tb2, Tt2 = (9 * Tb + 5 * Tt) / 14, (9 * Tt + 5 * Tb) / 14
    eta = (np.arange(1, case.Ny + 1) - 0.5) / case.Ny
    assert np.abs(T[:, 1:-1, :] - (Tb2 + (Tt2 - Tb2) * eta)[None, :, None]).max() < 1e-12


def test_thermal_neumann_wall_is_adiabatic(oracle_mod):
    """Determine_Twall_Neumann (lbm.cpp:983-1028) with j.n = 0 sets the
    unknown populations so that the wall-normal flux vanishes: with an
    adiabatic bottom and a Dirichlet top the fluid relaxes to the top wall's
    temperature, uniform."""
    case, th = cases.conduction_channel(bottom=cases.TYPE_Q, top=cases.TYPE_T)
    o = _thermal_oracle(oracle_mod, case, th)
    o.init()
    assert o.run(6000) == 0
    T = o.thermal()[0]
    Ttop = (9 * th.params["Tt"] + 5 * th.params["Tb"]) / 14
    assert np.abs(T[:, 1:-1, :] - Ttop).max() < 2e-7


def test_thermal_couette_viscous_heating_closed_form(oracle_mod):
    """Couette flow with viscous heating (main_setup_Couette + ENERGY): the
    steady temperature is T_b' + (T_t' - T_b') eta + nu U^2/(2 cp alpha)
    eta (1 - eta), eta = (j - 1/2)/Ny, with the walls' temperatures after the
    first MacroProp.  Pins the dissipation Phi = 2 nu rho S:S of the
    non-equilibrium strain rate, the source w Phi/cp and tau_g."""
    case, th = cases.couette_thermal(Ny=12, Re=10.0)
    o = _thermal_oracle(oracle_mod, case, th)
    o.init()
    assert o.run(5000) == 0
    T = o.thermal()[0]
    p = th.params
    Tb2 = (9 * p["Tb"] + 5 * p["Tt"]) / 14
    Tt2 = (9 * p["Tt"] + 5 * p["Tb"]) / 14
    eta = (np.arange(1, case.Ny + 1) - 0.5) / case.Ny
    U = p["u_wall"]
    heat = th.nu * U ** 2 / (2 * th.cp * p["alpha"])
    Ta = Tb2 + (Tt2 - Tb2) * eta + heat * eta * (1 - eta)
    err = np.abs(T[1, 1:-1, 1] - Ta).max()
    # the heating term dominates the profile; a dropped or doubled Phi fails
    assert heat > 2 * abs(Tt2 - Tb2)
    assert err < 1e-3 * heat


def test_thermal_temperature_range_error(oracle_mod):
    """MacroProp terminates when T < 0 or T > 100 (lbm.cpp:794-797): code 3."""
    case, th = cases.conduction_channel(Tb=150.0, Tt=150.0, T_fluid=150.0)
    o = _thermal_oracle(oracle_mod, case, th)
    o.init()
    assert o.run(1) == 3


# ---- D3Q7 temperature (no reference counterpart: pinned by closed forms) ----

def _q7_oracle(oracle_mod, case, th):
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    o.q7_enable(th.tau_g, th.boussinesq, th.g_beta, th.T_ref, j_dot_n=th.j_dot_n)
    o.set_cells(**case.fields())
    o.set_q7(th.T, th.type_energy)
    o.init()
    return o


@pytest.mark.parametrize("U", [0.0, 0.04])
def test_q7_sine_mode_exact_factor(oracle_mod, U):
    """At tau_g = 1 one step maps T to sum_i w_i T(x - c_i)(1 + c_i.u/cs2):
    a sine mode in x with a uniform flow ux = U is multiplied by exactly
    lambda = 3/4 + cos(k)/4 - i U sin(k) per step (amplitude and phase).
    Pins w = 1/4, 1/8, cs2 = 1/4, the sign of c.u and the streaming
    direction of g."""
    case, th = cases.q7_sine(Np=16, U=U)
    o = _q7_oracle(oracle_mod, case, th)
    n = 25
    assert o.run(n) == 0
    T = o.q7_thermal()[0]
    p = case.params
    lam = 0.75 + 0.25 * math.cos(p["k"]) - 1j * U * math.sin(p["k"])
    x = np.arange(case.Nx + 2)
    want = p["T0"] + p["A"] * np.imag(lam ** n * np.exp(1j * p["k"] * x))
    assert np.abs(T[:, 2, 2] - want).max() < 1e-14
    # the flow stays uniform (no force): u is U in every cell
    _, ux, uy, uz = o.macro()
    assert np.abs(ux - U).max() < 1e-15 and np.abs(uy).max() < 1e-15


def test_q7_sine_mode_diffusivity(oracle_mod):
    """Any tau_g: a long-wave sine mode at rest decays at alpha k^2 per step,
    alpha = (tau_g - 1/2)/4, to O(k^2)."""
    for tau_g in (0.7, 1.3):
        case, th = cases.q7_sine(Np=64, U=0.0, tau_g=tau_g)
        o = _q7_oracle(oracle_mod, case, th)
        n = 400
        assert o.run(n) == 0
        T = o.q7_thermal()[0][:, 2, 2]
        p = case.params
        amp = np.sqrt(2.0 * np.mean((T - p["T0"]) ** 2))
        rate = -math.log(amp / p["A"]) / n
        alpha = (tau_g - 0.5) / 4.0
        assert abs(rate / (alpha * p["k"] ** 2) - 1.0) < 1e-2


def test_q7_dirichlet_conduction_linear_profile(oracle_mod):
    """Two Dirichlet walls, fluid at rest: the steady profile is linear with
    the wall temperatures at the half-way wall positions,
    T_j = Tb + (Tt - Tb)(j - 1/2)/Ny (anti-bounce-back is exact for it)."""
    for tau_g in (0.8, 1.0, 1.6):
        case, th = cases.q7_conduction(Ny=8, tau_g=tau_g, Tb=1.0, Tt=0.25)
        o = _q7_oracle(oracle_mod, case, th)
        assert o.run(6000) == 0
        T = o.q7_thermal()[0]
        j = np.arange(1, case.Ny + 1)
        want = 1.0 + (0.25 - 1.0) * (j - 0.5) / case.Ny
        assert np.abs(T[:, 1:-1, :] - want[None, :, None]).max() < 1e-10


def test_q7_relative_temperatures_and_blow_up_guard(oracle_mod):
    """D3Q7 T is relative: walls at +0.5 and -0.5 run without a range error
    to the linear profile through 0 (the ENERGY branch's T < 0 stop,
    lbm.cpp:794-797, does not apply); |T| > 100 of either sign is flagged
    at the first MacroProp."""
    case, th = cases.q7_conduction(Ny=8, tau_g=1.0, Tb=0.5, Tt=-0.5)
    o = _q7_oracle(oracle_mod, case, th)
    assert o.run(6000) == 0
    T = o.q7_thermal()[0]
    j = np.arange(1, case.Ny + 1)
    want = 0.5 - (j - 0.5) / case.Ny
    assert np.abs(T[:, 1:-1, :] - want[None, :, None]).max() < 1e-10
    for big in (150.0, -150.0):
        case, th = cases.q7_conduction(Tb=big, Tt=big, T_fluid=big)
        o = _q7_oracle(oracle_mod, case, th)
        assert o.run(1) == 3  # ORC_ERR_TEMPERATURE_RANGE


def test_q7_adiabatic_walls_conserve_heat(oracle_mod):
    """A closed box with adiabatic (bounce-back) walls and solid obstacles,
    a moving, uneven flow: the total fluid heat sum T is constant (every g
    population either streams to a fluid cell or bounces back)."""
    case = cases.random_mask_case(8, 7, 6, seed=21, p_other=0.0)
    t = case.type
    for sl in ((0, slice(None), slice(None)), (-1, slice(None), slice(None)),
               (slice(None), 0, slice(None)), (slice(None), -1, slice(None)),
               (slice(None), slice(None), 0), (slice(None), slice(None), -1)):
        t[sl] = cases.TYPE_S
    case.ux[t == cases.TYPE_S] = 0.0
    case.uy[t == cases.TYPE_S] = 0.0
    case.uz[t == cases.TYPE_S] = 0.0
    rng = np.random.default_rng(3)
    T = 0.3 + 0.4 * rng.random(case.shape)
    te = np.full(case.shape, cases.TYPE_Q, np.int16)
    th = cases.q7_thermal(T, te, 0.9)
    o = _q7_oracle(oracle_mod, case, th)
    fl = t == cases.TYPE_F
    h0 = o.q7_g()[fl].sum()
    assert o.run(60) == 0
    h1 = o.q7_thermal()[0][fl].sum()
    assert abs(h1 - h0) < 1e-12 * h0
    # and the field did change
    assert np.abs(o.q7_thermal()[0][fl] - T[fl]).max() > 1e-3


def test_q7_flux_wall_conduction_linear_profile(oracle_mod):
    """A TYPE_Q wall at j = 0 letting the heat flux q into the fluid and a
    Dirichlet wall at j = Ny+1 (T = Tt), fluid at rest: the steady profile
    carries q through every plane, alpha dT/dy = -q, with the Dirichlet wall
    temperature at its half-way position:
    T_j = Tt + (q/alpha)(Ny + 1/2 - j), alpha = (tau_g - 1/2)/4."""
    q = 2e-3
    for tau_g in (0.8, 1.0, 1.6):
        case, th = cases.q7_conduction(Ny=8, tau_g=tau_g, Tb=0.5, Tt=0.25, bottom=cases.TYPE_Q, j_dot_n=q)
        o = _q7_oracle(oracle_mod, case, th)
        assert o.run(12000) == 0
        T = o.q7_thermal()[0]
        alpha = (tau_g - 0.5) / 4.0
        j = np.arange(1, case.Ny + 1)
        want = 0.25 + q / alpha * (case.Ny + 0.5 - j)
        assert np.abs(T[:, 1:-1, :] - want[None, :, None]).max() < 1e-10


def test_q7_flux_walls_heat_balance(oracle_mod):
    """A closed box of TYPE_Q walls with solid obstacles and an uneven flow:
    every wall link of a fluid cell adds the flux j_dot_n to the heat each
    step (the rest streams or bounces back), so the total fluid heat grows
    by exactly n j_dot_n (wall links) over n steps; with the sign of the
    flux reversed it falls by as much."""
    case = cases.random_mask_case(8, 7, 6, seed=21, p_other=0.0)
    t = case.type
    for sl in ((0, slice(None), slice(None)), (-1, slice(None), slice(None)),
               (slice(None), 0, slice(None)), (slice(None), -1, slice(None)),
               (slice(None), slice(None), 0), (slice(None), slice(None), -1)):
        t[sl] = cases.TYPE_S
    case.ux[t == cases.TYPE_S] = 0.0
    case.uy[t == cases.TYPE_S] = 0.0
    case.uz[t == cases.TYPE_S] = 0.0
    fl = t == cases.TYPE_F
    # D3Q7 links from fluid cells to solids (the axis neighbours, periodic index)
    links = 0
    for ax in range(3):
        for sgn in (1, -1):
            links += int((fl & (np.roll(t, -sgn, axis=ax) == cases.TYPE_S)).sum())
    rng = np.random.default_rng(3)
    T = 0.3 + 0.4 * rng.random(case.shape)
    te = np.full(case.shape, cases.TYPE_Q, np.int16)
    for q in (1e-4, -1e-4):
        th = cases.q7_thermal(T, te, 0.9, j_dot_n=q)
        o = _q7_oracle(oracle_mod, case, th)
        h0 = o.q7_g()[fl].sum()
        n = 40
        assert o.run(n) == 0
        h1 = o.q7_thermal()[0][fl].sum()
        assert abs(h1 - h0 - n * q * links) < 1e-12 * h0


def test_q7_isothermal_dirichlet_box_is_a_fixed_point(oracle_mod):
    """Dirichlet walls and fluid at rest, all at T_w: T stays T_w exactly
    (-g* + 2 w T_w = w T_w on every wall link).  With a flow the uniform T
    would not be a fixed point unless div u = 0: the equilibrium
    w T (1 + c.u/cs2) transports T in conservative form."""
    case = cases.random_mask_case(6, 6, 6, seed=23, p_other=0.0)
    case.rho[:] = 1.0
    for u in (case.ux, case.uy, case.uz):
        u[:] = 0.0
    te = np.where(case.type == cases.TYPE_S, cases.TYPE_T, 0).astype(np.int16)
    th = cases.q7_thermal(np.full(case.shape, 0.625), te, 0.8)
    o = _q7_oracle(oracle_mod, case, th)
    assert o.run(30) == 0
    T = o.q7_thermal()[0]
    assert np.abs(T[case.type == cases.TYPE_F] - 0.625).max() < 1e-14


def test_q7_boussinesq_force_enters_velocity(oracle_mod):
    """F_y = g_beta (T - T_ref) at T's own time level, the first collision
    included: in a periodic box at rest with uniform T = T_ref + d, the Guo
    closed form rho u_y = (n - omega/2 + 1/2) g_beta d after n steps (the
    first collision starts from the set-up u = 0)."""
    case, th = cases.q7_sine(Np=8, U=0.0, A=0.0, T0=0.75)
    th.boussinesq, th.g_beta, th.T_ref = True, 1e-4, 0.5
    o = _q7_oracle(oracle_mod, case, th)
    n = 10
    assert o.run(n) == 0
    rho, ux, uy, uz = o.macro()
    d = 0.25
    omega = 1.0 / case.tau
    assert np.abs(rho * uy - (n - omega / 2 + 0.5) * th.g_beta * d).max() < 1e-15


# ---- flexible filament (no reference counterpart: closed forms) -----------

def _cantilever(oracle_mod, damping=0.0, body=(0.0, 0.0, 0.0), n=41, L=40.0, ml=1.5, EI=3.0):
    return oracle_mod.Filament(n, L, ml, EI, damping=damping, body=body, clamp=(10.0, 20.0, 5.0),
                               tangent=(1.0, 0.0, 0.0))


def test_filament_hanging_tension_exact(oracle_mod):
    """A filament clamped pointing down under its weight m_l g per unit
    length hangs straight with segment tensions m_l g (L - s_mid) (the free
    end node carries half a segment), and keeps its segment lengths."""
    n, L, ml, g = 21, 20.0, 1.5, 1e-3
    f = oracle_mod.Filament(n, L, ml, 2.0, damping=0.5, body=(0.0, -ml * g, 0.0), clamp=(0.0, 0.0, 0.0),
                            tangent=(0.0, -1.0, 0.0))
    f.step(nsteps=20)
    x, _, T = f.get()
    ds = L / (n - 1)
    s_mid = (np.arange(n - 1) + 0.5) * ds
    assert np.abs(T - ml * g * (L - s_mid)).max() < 1e-12
    assert np.abs(x[:, 0]).max() < 1e-14 and np.abs(x[:, 2]).max() < 1e-14
    assert np.abs(np.linalg.norm(np.diff(x, axis=0), axis=1) - ds).max() < 1e-12


def test_filament_cantilever_static_deflection(oracle_mod):
    """Uniform transverse load q on a clamped-free filament, damped to rest:
    tip deflection qL^4/(8 EI) (small deflection, 1% of L)."""
    L, EI, ml = 40.0, 3.0, 1.5
    q = 0.08 * EI / L ** 3
    w1 = 1.8751040687 ** 2 * math.sqrt(EI / (ml * L ** 4))
    f = _cantilever(oracle_mod, damping=2 * ml * w1, L=L, EI=EI, ml=ml)
    F = np.zeros((f.n, 3))
    F[:, 1] = -q
    for _ in range(12):
        f.step(F, nsteps=2000)
    x, xp, _ = f.get()
    tip = 20.0 - x[-1, 1]
    assert abs(tip / (q * L ** 4 / (8 * EI)) - 1) < 2e-3, tip
    assert np.abs(x - xp).max() < 1e-9  # at rest


def test_filament_cantilever_first_mode(oracle_mod):
    """Released from a small static deflection, the free tip oscillates at
    the first clamped-free beam mode, omega_1 = 1.8751^2 sqrt(EI/(m_l L^4)),
    and the segment lengths stay ds."""
    L, EI, ml = 40.0, 3.0, 1.5
    w1 = 1.8751040687 ** 2 * math.sqrt(EI / (ml * L ** 4))
    q = 0.02 * EI / L ** 3
    f = _cantilever(oracle_mod, damping=2 * ml * w1, L=L, EI=EI, ml=ml)
    F = np.zeros((f.n, 3))
    F[:, 1] = -q
    for _ in range(12):
        f.step(F, nsteps=2000)
    x0, _, _ = f.get()
    g = _cantilever(oracle_mod, L=L, EI=EI, ml=ml)
    g.set(x0, x0)
    period = 2 * math.pi / w1
    nsteps = int(3 * period)
    tip = np.empty(nsteps)
    for k in range(nsteps):
        g.step(None, nsteps=1)
        tip[k] = g.get()[0][-1, 1] - 20.0
    # upward zero crossings of the tip, interpolated
    z = np.flatnonzero((tip[:-1] < 0) & (tip[1:] >= 0))
    t = z + (-tip[z]) / (tip[z + 1] - tip[z])
    assert len(t) >= 2
    measured = 2 * math.pi / np.mean(np.diff(t))
    assert abs(measured / w1 - 1) < 1e-2, (measured, w1)
    x = g.get()[0]
    assert np.abs(np.linalg.norm(np.diff(x, axis=0), axis=1) - L / (f.n - 1)).max() < 1e-12


# ---- momentum exchange (MEA_CD / MEA_CL, lbm.cpp:3085-3163) ---------------------
def _walled(axis, N=(4, 5, 3), rho=1.2, tau=0.8):
    t, r, ux, uy, uz = cases._blank(*N)
    lo = [slice(None)] * 3
    hi = [slice(None)] * 3
    lo[axis] = 0
    hi[axis] = -1
    t[tuple(lo)] = cases.TYPE_S
    t[tuple(hi)] = cases.TYPE_S
    r[:] = rho
    return cases.Case("walled", *N, tau, t, r, ux, uy, uz)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_mea_rest_fluid_wall_pressure(oracle_mod, axis):
    """Fluid at rest at density rho between two walls normal to `axis`: every
    wall link carries f* = f = w rho, so a box holding the low wall's fluid
    layer sums 2 rho (1/18 + 4/36) = rho/3 per cell, with the sign the
    reference gives that component (-c_x for Px, +c_y for Py, +c_z for Pz);
    the whole lattice sums to zero.  Catches a dropped f or f* term (rho/6),
    the link test x + c_l in place of x - c_l (sign flip) and a wrong
    component sign."""
    rho = 1.2
    case = _walled(axis, rho=rho)
    o = oracle_mod.run_case(case, 3)
    shape = case.shape
    layer = int(np.prod([shape[d] for d in range(3) if d != axis]))
    box = []
    for d in range(3):
        box += [0, (shape[d] - 1) // 2] if d == axis else [0, shape[d] - 1]
    P = o.mea(box)
    want = (-1.0 if axis == 0 else 1.0) * rho / 3 * layer
    assert abs(P[axis] - want) < 1e-13 * layer
    assert np.abs(np.delete(P, axis)).max() < 1e-13 * layer
    full = [0, shape[0] - 1, 0, shape[1] - 1, 0, shape[2] - 1]
    assert np.abs(o.mea(full)).max() < 1e-13 * layer


def test_mea_moving_lid_second_step_closed_form(oracle_mod):
    """Fluid at rest (rho = 1.2) under a lid at j = Ny+1 moving with U along
    x.  Step 1's bounce-back sets the top cells' links from the lid to
    f_l = w rho - 6 w U c_opp(l),x (no rho factor, lbm.cpp:274): x-momentum
    U/3, density unchanged.  Step 2 relaxes them towards feq(rho, U/(3 rho)),
    whose x-momentum on those links is 6 w rho u = U/18 (the cubic terms
    cancel), and the streamed f_opp(l) come from cells still at rest.  So
    after two steps Px = -(Nx+2)(Nz+2) ((1 - omega) U/3 + omega U/18) and
    Py = 0.  Catches a missing wall-velocity term, a rho factor in the
    bounce-back, a wrong relaxation weight and a wrong sign of Px."""
    U, tau = 0.05, 0.8
    case = cases.couette(Nx=4, Ny=6, Nz=3, tau=tau, u_wall=U)
    case.rho[:] = 1.2
    o = oracle_mod.run_case(case, 2)
    om = 1.0 / tau
    cells = (case.Nx + 2) * (case.Nz + 2)
    want = -cells * ((1 - om) * U / 3 + om * U / 18)
    P = o.mea([0, case.Nx + 1, 0, case.Ny + 1, 0, case.Nz + 1])
    assert abs(P[0] - want) < 1e-13 * cells
    assert abs(P[1]) < 1e-12 * cells and abs(P[2]) < 1e-13 * cells


def test_voxel_bodies_follow_the_generators():
    """Generate_Cylinder / Generate_Sphere (lbm.cpp:1429-1450, 1759-1785):
    cell centres (i - 1/2, ...) within D/2 of the centre; the solid count
    approaches the disc area / ball volume; the MEA boxes are the int
    conversions of centre -+ D (main.cpp:246, 369)."""
    c = cases.cylinder_bb()
    solid = c.type[1:-1, 1:-1, 1] == cases.TYPE_S
    assert abs(solid.sum() / (math.pi * 10 ** 2) - 1) < 0.02
    assert c.params["mea_box"] == (0, 40, 30, 70, 1, 1)
    # later faces win (lbm.cpp:1468-1471): the j faces override the inlet's edges
    assert (c.type[0, 1:-1] == cases.TYPE_IN).all() and (c.type[:, 0, :] == cases.TYPE_OUT).all()
    s = cases.sphere_bb(80, 40, 40, D=16.0)
    vol = (s.type == cases.TYPE_S).sum()
    assert abs(vol / (4 / 3 * math.pi * 8 ** 3) - 1) < 0.03
    cx, cy, cz = s.params["center"]
    assert s.type[int(cx + 0.5), int(cy + 0.5), int(cz + 0.5)] == cases.TYPE_S
    assert s.params["mea_box"] == (0, 32, 4, 36, 4, 36)


# ---- MORE_DETAILS: stress and running statistics (lbm.cpp:681-705, 865-939) -------
def test_stress_of_equilibrium_is_isotropic(oracle_mod):
    """f = feq(rho, u) (Init, then MacroProp recovers u exactly): the second
    central moment of the D3Q19 cubic equilibrium is rho cs2 delta (the
    cubic terms are odd), so sigma = -(tau - 1/2)/tau rho/3 delta and the
    pressure is rho/3.  Catches a missing u subtraction (rho u u terms), a
    wrong (tau - 1/2)/tau factor and a wrong sign."""
    case = cases.random_mask_case(5, 4, 3, seed=8, p_solid=0.0)
    rng = np.random.default_rng(8)
    case.rho[:] = 1.0 + 0.1 * rng.standard_normal(case.shape)
    for comp in (case.ux, case.uy, case.uz):
        comp[:] = 0.04 * rng.standard_normal(case.shape)
    o = oracle_mod.run_case(case, 0)
    assert o.macroprop() == 0
    sig, p = o.stress()
    rho = o.macro()[0]
    want = -(case.tau - 0.5) / case.tau * rho / 3
    assert np.abs(sig - want[..., None, None] * np.eye(3)).max() < 1e-14
    assert np.abs(p - rho / 3).max() < 1e-15


def test_stress_in_steady_couette_is_rho_nu_shear(oracle_mod):
    """Steady Couette flow (the reference's validation case, lbm.cpp:3238-
    3251): sigma_xy = rho nu U / H with H = Ny between half-way walls, to
    1e-8 (the reference's closed form uses h = Ny + 1).  Catches a wrong
    viscosity factor and the xy/yx assignment."""
    U, tau, Ny = 0.001, 0.8, 10
    case = cases.couette(Nx=2, Ny=Ny, Nz=2, tau=tau, u_wall=U)
    o = oracle_mod.run_case(case, 8000)
    sig, _ = o.stress()
    want = (tau - 0.5) / 3 * U / Ny
    s = sig[1, 1:-1, 1]
    assert np.abs(s[:, 0, 1] / want - 1).max() < 1e-8
    assert np.array_equal(s[:, 0, 1], s[:, 1, 0])


def test_stats_of_a_steady_field_closed_forms(oracle_mod):
    """Sup_MacroProp on a field that does not change (every Nt = 10 steps,
    n = 10..60): the averages equal the field (the weights sum to one),
    the velocity rms and the averaged fluctuation products vanish, and the
    reference's forms give rms_p = p sqrt(told/(told + 1)) (pressure^2
    weighted by told) and rms_rs = |u_a u_b| (raw products against the
    averaged fluctuations)."""
    case = cases.random_mask_case(4, 3, 3, seed=5)
    rng = np.random.default_rng(5)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    rho = 1.0 + 0.05 * rng.random(case.shape)
    u = [0.03 * rng.standard_normal(case.shape) for _ in range(3)]
    o.set_cells(type=case.type, rho=rho, ux=u[0], uy=u[1], uz=u[2])
    nt = 10
    for n in range(nt, 7 * nt, nt):
        o.stats_update(n, nt)
    avg_u, rms_u, avg_p, rms_p, avg_rs, rms_rs = o.stats()
    U = np.stack(u, -1)
    told = 6 * nt - nt
    assert np.abs(avg_u - U).max() < 1e-16
    assert np.abs(rms_u).max() < 1e-16 and np.abs(avg_rs).max() < 1e-18
    assert np.abs(avg_p - rho / 3).max() < 1e-15
    assert np.abs(rms_p - rho / 3 * np.sqrt(told / (told + 1))).max() < 1e-15
    assert np.abs(rms_rs - np.abs(U[..., :, None] * U[..., None, :])).max() < 1e-17


def test_stats_running_mean_of_a_linear_signal(oracle_mod):
    """Nt = 1 (told = n - 1): u_x = n d at update n gives the arithmetic mean
    (n + 1) d / 2, and the reference's running rms (each sample against the
    average at its own time) rms_k^2 = (k - 1)(2k - 1) d^2 / 24, the same
    value as the averaged product avg_rs_xx."""
    case = cases.random_mask_case(3, 3, 3, seed=2, p_solid=0.0)
    o = oracle_mod.Oracle(case.Nx, case.Ny, case.Nz, case.tau)
    d = 0.01
    K = 9
    for n in range(1, K + 1):
        o.set_cells(type=case.type, ux=np.full(case.shape, n * d))
        o.stats_update(n, 1)
    avg_u, rms_u, _, _, avg_rs, _ = o.stats()
    assert np.abs(avg_u[..., 0] - (K + 1) * d / 2).max() < 1e-16
    want = (K - 1) * (2 * K - 1) * d * d / 24
    assert np.abs(rms_u[..., 0] - np.sqrt(want)).max() < 1e-16
    assert np.abs(avg_rs[..., 0, 0] - want).max() < 1e-18


# ---- IBM, explicit branch (lbm.cpp:1171-1258) ------------------------------------
def test_ibm_explicit_uniform_flow_cell_summed_drag(oracle_mod):
    """Uniform flow u0 past markers at rest in the 2D slab: the unforced
    interpolation gives u0 phi(0) = u0/2, F_b = -rho u0, and the spread
    force sums over the box cells to sum_n F_b ds * 1/2 (the z-weight
    again), so the cell-summed drag is u0 sum(ds)/2 -- half the implicit
    branch's marker-summed m_max = 1 value.  The spread force enters with
    + (lbm.cpp:1244-1246): each box cell's u gains F/2 rho."""
    u0 = (0.03, -0.01, 0.0)
    o = _uniform_ibm_lattice(oracle_mod, u0)
    mk = cases.cylinder_markers(7.2, 7.4, 6.0, 24, [1.0])
    err, out = o.ibm_step(mk, (0, 0, 0), (0, 0, 0), (2, 13, 2, 13, 1, 1), scheme="explicit")
    assert err == 0
    Fb = mk["Fb"].reshape(-1, 3)
    assert np.abs(Fb[:, 0] + u0[0]).max() < 1e-15 and np.abs(Fb[:, 1] + u0[1]).max() < 1e-15
    ds = mk["ds"].sum()
    assert out[0] == pytest.approx(u0[0] * ds / 2, rel=1e-12)
    assert out[1] == pytest.approx(u0[1] * ds / 2, rel=1e-12)
    Fx, _, _ = o.force()
    _, ux, _, _ = o.macro()
    box = (slice(2, 14), slice(2, 14), 1)
    assert np.abs(ux[box] - (u0[0] + Fx[box] / 2)).max() < 1e-15
    assert Fx[box].min() < -1e-4  # the force opposes the flow at the markers


def test_ibm_explicit_nearest_point_brute_force(oracle_mod):
    """The explicit branch's density point (NearestLatticeCoordinate over the
    marker's floor/ceil neighbourhood, lbm.cpp:1218-1222): with the fluid at
    rest and the body moving at V, F_b = 2 rho_near V, so F_b/2V names the
    chosen cell.  Expected: the last (x, y) in loop order with the largest
    positive dot product along the normal, from a density field whose value
    differs at every cell."""
    t, rho, ux, uy, uz = cases._blank(16, 16, 1)
    i = np.arange(18)[:, None, None]
    j = np.arange(18)[None, :, None]
    rho[:] = 1.0 + 0.001 * (i + 20 * j) + 0 * rho
    o = oracle_mod.Oracle(16, 16, 1, 0.8)
    o.set_cells(type=t, rho=rho, ux=ux, uy=uy, uz=uz)
    o.init()
    mk = cases.cylinder_markers(7.2, 7.4, 6.0, 24, [1.0])
    # marker 0 on a lattice point: no neighbour lies ahead along its normal
    # and the search keeps its start (0, 0, 1)
    mk["x"][0], mk["y"][0] = float(round(mk["x"][0])), float(round(mk["y"][0]))
    V = (0.02, 0.0, 0.0)
    err, out = o.ibm_step(mk, V, V, (2, 13, 2, 13, 1, 1), scheme="explicit")
    assert err == 0
    got = mk["Fb"].reshape(-1, 3)[:, 0] / (2 * V[0])
    for n in range(24):
        x, y = mk["x"][n], mk["y"][n]
        best, pick = 0.0, (0, 0)
        for xi in range(math.floor(x), math.ceil(x) + 1):
            for yi in range(math.floor(y), math.ceil(y) + 1):
                d = (xi - x) * mk["nx"][n] + (yi - y) * mk["ny"][n]
                if d >= best and d > 0:
                    best, pick = d, (xi, yi)
        assert got[n] == pytest.approx(rho[pick[0], pick[1], 1], rel=1e-12), n
