/*
 * lbm_oracle.h -- CPU FP64 restatement of the reference lattice Boltzmann step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline; it is never linked into, called by or shipped with
 * the product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * It follows the reference's own code (lbm.cpp / LBM.h of
 * DestinyFlow01/Lattice-Boltzmann-Method---Graduate-FSI-) operation by
 * operation: the same AoS 368-byte cell, the same three OpenMP sweeps
 * (Collision, Streaming + face loops, MacroProp), the same cubic equilibrium
 * and Guo source, the same push streaming with half-way bounce-back and the
 * same per-cell averaged wall velocity.  Compile-time #defines of the
 * reference become runtime fields of orc_lbm.
 *
 * Host array convention at this interface: cell (i,j,k) of the padded
 * (Nx+2)(Ny+2)(Nz+2) lattice sits at linear index (i*(Ny+2)+j)*(Nz+2)+k,
 * k fastest -- the order of LBM::fluid[i][j][k] (lbm.cpp:17-28).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cell-type codes, LBM.h:505-515 */
#define ORC_TYPE_F 0
#define ORC_TYPE_S 1
#define ORC_TYPE_OUT 2
#define ORC_TYPE_IN 3
#define ORC_TYPE_P 4
#define ORC_TYPE_L 5
#define ORC_TYPE_T 6
#define ORC_TYPE_Q 7

/* error codes: the reference's three std::terminate() sites */
#define ORC_OK 0
#define ORC_ERR_FEQ_NEGATIVE 1      /* lbm.cpp:114-117 */
#define ORC_ERR_DENSITY_RANGE 2     /* lbm.cpp:772-775 */
#define ORC_ERR_TEMPERATURE_RANGE 3 /* lbm.cpp:794-797 */

typedef struct orc_lbm orc_lbm;

/* LBM::LBM(Nx,Ny,Nz,tau), lbm.cpp:17-28.  q = 19 (D3Q19) or 9 (D2Q9). */
orc_lbm *orc_create(int Nx, int Ny, int Nz, double tau, int q);
void orc_destroy(orc_lbm *lb);
int64_t orc_ncells(const orc_lbm *lb);
int orc_sizeof_cell(void);

/* velocity-set tables (LBM.h:529-555) for the lattice in use */
int orc_q(const orc_lbm *lb);
void orc_tables(const orc_lbm *lb, int *cx, int *cy, int *cz, int *opp, double *w);

/* writes of lb.fluid[i][j][k].{type,density,ux,uy,uz,F} as the setups do */
void orc_set_cells(orc_lbm *lb, const int16_t *type, const double *rho,
                   const double *ux, const double *uy, const double *uz,
                   const double *Fx, const double *Fy, const double *Fz);
void orc_get_macro(const orc_lbm *lb, double *rho, double *ux, double *uy,
                   double *uz);
void orc_get_force(const orc_lbm *lb, double *Fx, double *Fy, double *Fz);
/* f and f_star as [cell][l] (l fastest, q entries per cell) */
void orc_get_f(const orc_lbm *lb, double *f);
void orc_get_fstar(const orc_lbm *lb, double *fstar);
void orc_set_f(orc_lbm *lb, const double *f);
void orc_set_fstar(orc_lbm *lb, const double *fstar);

/* LBM::Calc_feq for one (rho,u) and direction, lbm.cpp:628-633 */
double orc_feq(const orc_lbm *lb, double rho, double ux, double uy, double uz,
               int l);

/* the reference step's methods */
void orc_init(orc_lbm *lb);      /* LBM::Init, lbm.cpp:31-57 */
int orc_collision(orc_lbm *lb);  /* LBM::Collision + ConvLBM, lbm.cpp:60-179 */
void orc_streaming(orc_lbm *lb); /* LBM::Streaming, lbm.cpp:182-626 */
int orc_macroprop(orc_lbm *lb);  /* LBM::MacroProp, lbm.cpp:657-812 */
/* main.cpp:212-221: Collision -> Streaming -> MacroProp, n times */
int orc_run_plain(orc_lbm *lb, int nsteps);

/* Determine_Wall_Velocity, lbm.cpp:941-962 */
void orc_wall_velocity(const orc_lbm *lb, int i, int j, int k, double uw[3]);

/*
 * Pull form of the streaming pass (restated from lbm.cpp:186-295): for fluid
 * cell x and direction m, src = x - c_m wrapped modulo (N+2) per axis
 * (lbm.cpp:197-199) and kind = 0 (src is TYPE_F: f(x,m) = f*(src,m)),
 * 1 (src is TYPE_S: bounce-back, f(x,m) = f*(x,opp m) + 2 w_m (u_w.c_m)/cs2)
 * or 2 (any other type: f(x,m) is never rewritten).  For non-fluid cells
 * src = -1, kind = 255.  src is a linear index in this interface's order.
 */
void orc_export_links(const orc_lbm *lb, int32_t *src, uint8_t *kind);

/* number of OpenMP threads the sweeps use (0 = leave the runtime default) */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

/* ---- implicit diffuse-interface IBM, lbm.cpp:1048-1417 ----------------- */
/* Peskin 4-point kernel(r), lbm.cpp:1095-1102 */
double orc_ibm_kernel(double r);

/*
 * One IBLBM() call of the implicit scheme (lbm.cpp:1135-1417, branch
 * Implicit_Diffuse_Interface_Scheme), generalised from the reference's
 * 2D slab (zmin = zmax = 1) to a box [xmin,xmax]x[ymin,ymax]x[zmin,zmax].
 * Marker state is caller-owned and updated in place:
 *   mx,my,mz      marker positions (advanced by the trapezoid rule,
 *                 lbm.cpp:1154-1163, with vcm and vcm_before)
 *   nx,ny,nz      marker normals (Normal_Marker, lbm.cpp:1048-1069)
 *   ub            interpolated marker velocity [3*n] (out)
 *   Fb            marker force [3*n] (out)
 *   ds            marker weights (main.cpp:123-131)
 * Box limits are passed by the caller (lbm.cpp:1165-1167 computes them from
 * the body centre).  vsolid is the added-mass volume; the average density of
 * the interior cells is multiplied by it (lbm.cpp:1392-1398).
 * out[0] = Drag, out[1] = Lift (raw sums, before the C_D/C_L normalisation),
 * out[2] = the average interior density used for added mass.
 * With D3Q7 on (no reference counterpart) the step ends as MacroProp's D3Q7
 * part would: T and q of every fluid cell from g, and the Boussinesq force
 * g_beta (T - T_ref) added to F_y and half of it over rho to u_y.
 */
int orc_ibm_step(orc_lbm *lb, int nm, double *mx, double *my, double *mz,
                 const double *nx, const double *ny, const double *nz,
                 double *ub, double *Fb, const double *ds,
                 const double vcm[3], const double vcm_before[3],
                 int xmin, int xmax, int ymin, int ymax, int zmin, int zmax,
                 int m_max, double vsolid, double *out);
/* MORE_DETAILS (LBM.h:55).  orc_stress: MacroProp's pressure = rho/3 and
 * viscous stress sigma_ab = -(tau - 1/2)/tau sum_l f_l (c_la - u_a)(c_lb -
 * u_b) of every cell (lbm.cpp:681-705), sigma [cell][3][3], pressure [cell].
 * orc_stats_update: Sup_MacroProp(n) (lbm.cpp:865-939) with Nt the output
 * interval; orc_get_stats: avg_u, rms_u [cell][3], avg_p, rms_p [cell],
 * avg_rs, rms_rs [cell][3][3]. */
void orc_stress(const orc_lbm *lb, double *sigma, double *pressure);
void orc_stats_update(orc_lbm *lb, int n, int Nt);
void orc_get_stats(const orc_lbm *lb, double *avg_u, double *rms_u, double *avg_p, double *rms_p,
                   double *avg_rs, double *rms_rs);

/* MEA_CD / MEA_CL's momentum sums (lbm.cpp:3085-3163) over the inclusive
 * box {x0,x1,y0,y1,z0,z1}, evaluated where the reference does (after
 * Collision, Streaming, MacroProp: f* of the last collision, f streamed):
 * out[0] = Px = -sum (f*_l + f_opp) c_lx, out[1] = Py = +sum (...) c_ly
 * over TYPE_F cells and links whose x - c_l is TYPE_S; out[2] = Pz, the y
 * form in z (the reference has none). */
void orc_mea(const orc_lbm *lb, const int box[6], double out[3]);

/* IBLBM's explicit branch (Explicit_Diffuse_Interface_Scheme, lbm.cpp:
 * 1171-1258; the reference's default build uses the implicit one): unforced
 * moments of F/IN/OUT cells (their F kept), unforced interpolation, F_b =
 * 2 rho (u_des - u_b)/dt with rho at NearestLatticeCoordinate over the
 * marker's floor/ceil neighbourhood (starting from cell (0,0,1)), then for
 * the box's F/IN/OUT cells F = +sum F_b ds phi phi phi, u += F/(2 rho),
 * Drag and Lift summed over those cells (-F_x dx^2, -F_y dx^2), plus the
 * added mass; out as orc_ibm_step's. */
int orc_ibm_step_explicit(orc_lbm *lb, int nm, double *mx, double *my, double *mz,
                          const double *nx, const double *ny, const double *nz,
                          double *ub, double *Fb, const double *ds,
                          const double vcm[3], const double vcm_before[3],
                          int xmin, int xmax, int ymin, int ymax, int zmin, int zmax,
                          double vsolid, double *out);

/* The IBM step of a coupled filament (no reference counterpart): as
 * orc_ibm_step with the markers where they are (no rigid motion), the
 * desired velocity udes per marker (3 nm), rho at the box cell holding the
 * marker (floor(x + 1/2) per axis), no added mass; fsum (3 nm) receives
 * minus the sum over the passes of F_b, the force per unit length on the
 * filament.  D3Q7 temperature enters as in orc_ibm_step. */
int orc_fsi_ibm_step(orc_lbm *lb, int nm, const double *mx, const double *my, const double *mz,
                     const double *udes, const double *ds, double *ub, double *Fb, double *fsum,
                     int xmin, int xmax, int ymin, int ymax, int zmin, int zmax, int m_max);
/* main.cpp:195-211: IBLBM -> Collision -> Streaming (no MacroProp) */
void orc_get_fm(const orc_lbm *lb, double *Fmx, double *Fmy, double *Fmz);
/* NearestLatticeCoordinate, lbm.cpp:1074-1092 */
void orc_ibm_nearest(int xmin, int xmax, int ymin, int ymax, int zmin,
                     int zmax, double mx, double my, double mz, double nx,
                     double ny, double nz, int *i, int *j, int *k);

/* ---- thermal branch (#define ENERGY): LBM.h:590-596, lbm.cpp:121-178,
 * 263-289, 636-653, 777-806, 964-1028 ---------------------------------- */
/* Adds g, g_star, T, q, type_Energy and F_vis to every cell.  tau_g, nu, cp
 * and j_dot_n are the case constants (LBM.h:150-161); boussinesq = 1 sets
 * F[1] = g_beta (T - T_ref) in MacroProp (Rayleigh_Benard, lbm.cpp:803-806). */
int orc_thermal_enable(orc_lbm *lb, double tau_g, double nu, double cp, double j_dot_n,
                       int boussinesq, double g_beta, double T_ref);
/* writes of lb.fluid[i][j][k].T and .type_Energy as the setups do */
void orc_set_thermal(orc_lbm *lb, const double *T, const int16_t *typeE);
void orc_get_thermal(const orc_lbm *lb, double *T, double *qx, double *qy, double *qz);
/* g as [cell][l] */
void orc_get_g(const orc_lbm *lb, double *g);
void orc_set_g(orc_lbm *lb, const double *g);
/* Determine_Wall_Temperature, lbm.cpp:965-981 */
double orc_wall_temperature(const orc_lbm *lb, int i, int j, int k);
/* Calc_geq, lbm.cpp:636-653 (as_solid: the TYPE_S branch with wall velocity uw) */
double orc_geq(const orc_lbm *lb, int i, int j, int k, int l, int as_solid, double uwx,
               double uwy, double uwz);

/* ---- D3Q7 temperature (no reference counterpart; the library's
 * LBM_THERMAL_D3Q7).  The oracle's specification of it, in push form:
 *   g_i, i = 0..6 over the rest and axis directions, w = 1/4, 1/8 (x6),
 *   cs2 = 1/4;  geq_i = w_i T (1 + c_i.u/cs2);
 *   Collision (TYPE_F):  g*_i = g_i (1 - 1/tau_g) + geq_i / tau_g with the
 *     cell's T and u of the last MacroProp;
 *   Streaming (TYPE_F x, neighbour x + c_i):  TYPE_F: g_i(nb) = g*_i(x);
 *     TYPE_S with type_Energy TYPE_T: g_opp(x) = -g*_i(x) + 2 w_i T(nb);
 *     TYPE_S with type_Energy TYPE_Q: g_opp(x) = g*_i(x) + j_dot_n (the heat
 *     flux into the fluid through that wall face, orc_q7_set_flux);
 *     other TYPE_S: g_opp(x) = g*_i(x);  any other type: nothing (frozen);
 *   MacroProp (TYPE_F):  T = sum g, q = sum g c, then (boussinesq)
 *     F[1] = F_y(set-up) + g_beta (T - T_ref), before u = (sum f c + F/2)/rho;
 *     Init sets F[1] the same way from the set-up T (for the first collision).
 * A solid's T stays its set-up value. */
int orc_q7_enable(orc_lbm *lb, double tau_g, int boussinesq, double g_beta, double T_ref);
/* the heat flux through TYPE_Q wall faces (default 0: adiabatic) */
void orc_q7_set_flux(orc_lbm *lb, double j_dot_n);
/* T and type_Energy of every cell, after orc_set_cells (the set-up F_y the
 * buoyancy adds to is read here) */
void orc_q7_set(orc_lbm *lb, const double *T, const int16_t *typeE);
void orc_q7_get(const orc_lbm *lb, double *T, double *qx, double *qy, double *qz);
/* g as [cell][i], 7 per cell */
void orc_q7_get_g(const orc_lbm *lb, double *g);

/* ---- flexible filament (no reference counterpart: the reference's
 * flexible-filament structure solver is absent).  The oracle's
 * specification, lattice units:
 *   m_l X_tt = d/ds (T dX/ds) - EI d4X/ds4 + G - gamma X_t + F,  |dX/ds| = 1,
 * nodes i = 0..N-1 at spacing ds = L/(N-1); node 0 clamped at X0 with slope
 * e0 (ghost X_{-1} = X_1 - 2 ds e0); free end d2X/ds2 = d3X/ds3 = 0
 * (ghosts X_N = 2X_{N-1} - X_{N-2}, X_{N+1} = 2X_N - 2X_{N-2} + X_{N-3})
 * and T = 0; the end node carries half a segment (its tension term is
 * doubled).  One structure step of size h (K per call, h = dt/K):
 *   a_j = (X_{j+1} - X_j)/ds                         (tangents at time n)
 *   P_i = 2X_i - Xp_i + h^2/m_l (B_i + G + F_i - gamma (X_i - Xp_i)/h)
 *   X'_i = P_i + kappa c_i (T_i a_i - T_{i-1} a_{i-1}),  kappa = h^2/(m_l ds),
 *          c_i = 1 (c_{N-1} = 2), T_{N-1} = 0, X'_0 = X0;
 *   the segment tensions T_0..T_{N-2} make |X'_{j+1} - X'_j|^2 = ds^2:
 *   three Newton steps from the previous T, each a tridiagonal solve by
 *   parallel cyclic reduction (strides 1, 2, 4, ...; rows beyond the ends
 *   are zero), in this order of operations. */
typedef struct orc_filament orc_filament;
orc_filament *orc_fil_create(int n, double length, double line_density, double bending,
                             double damping, const double body[3], const double clamp[3],
                             const double tangent[3], int substeps);
void orc_fil_destroy(orc_filament *f);
int orc_fil_substeps(const orc_filament *f);
/* node-major x,y,z (3n each); x_prev may be NULL (= x_now) */
void orc_fil_set(orc_filament *f, const double *x_now, const double *x_prev);
void orc_fil_get(const orc_filament *f, double *x_now, double *x_prev, double *tension);
/* nsteps calls, each advancing dt with F (3n per unit length, or NULL) */
void orc_fil_step(orc_filament *f, const double *force, double dt, int nsteps);

#ifdef __cplusplus
}
#endif
#endif
