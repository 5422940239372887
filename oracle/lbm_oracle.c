/*
 * lbm_oracle.c -- CPU FP64 restatement of the reference lattice Boltzmann step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Nothing in the product library links or calls this file.
 *
 * Every function cites the reference lines (lbm.cpp / LBM.h / main.cpp of
 * DestinyFlow01/Lattice-Boltzmann-Method---Graduate-FSI-) it restates, and
 * is checked by tests/test_oracle_pins.py or tests/test_oracle_pins_coupling.py
 * against what the mathematics fixes (closed forms, invariants, brute-force
 * re-derivations on tiny lattices) rather than against its own formulas.  The
 * arithmetic is written in the reference's own expression order (including
 * its pow() calls) so that a gcc build without -ffast-math reproduces the
 * reference's FP64 rounding wherever the reference itself is deterministic.
 *
 * Deliberate differences from the reference, none of which changes a value:
 *  - std::terminate() becomes an error code (ORC_ERR_*); the sweep finishes.
 *  - Streaming builds no per-cell std::vector lists (lbm.cpp:207-249); the
 *    thermal Neumann wall derives the same index lists from the cell types.
 *    The wall velocity (lbm.cpp:941-962) is computed once per wall cell
 *    instead of once per wall link; it is the same value each time.
 *  - Compile-time case flags become runtime fields.
 * Where the reference itself is not deterministic or reads uninitialised
 * memory, this file fixes one behaviour (thermal branch only):
 *  - F_vis (LBM.h:595) has no initialiser; here it starts at 0.
 *  - The Neumann wall (lbm.cpp:281-287, 983-1028) reads the cell's g while
 *    other cells are still pushing into it, in OpenMP order; here every push
 *    of the pass is done first (a second sweep over the wall cells).  Its
 *    writes of the solid cells' T are not kept: MacroProp overwrites them
 *    with the moments of the solids' g before any later read.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* LBM.h:74-87: dt = dx = dy = dz = 1, cs2 = pow(dx,2)/(3*pow(dt,2)) */
#define ORC_DT 1.0
#define ORC_DX 1.0
#define ORC_DY 1.0
#define ORC_DZ 1.0

/* The reference's LATTICE for its D3Q19 build without IBM/ENERGY
 * (LBM.h:558-597): 2 + 6 (pad) + 4*8 + 19*8 + 19*8 + 3*8 = 368 bytes. */
typedef struct orc_cell {
  short type;
  double density, ux, uy, uz;
  double f[19], f_star[19];
  double F[3];
} orc_cell;

struct orc_lbm {
  int Nx, Ny, Nz;
  double tau;
  int q;
  const int *cx, *cy, *cz, *opp;
  const double *w;
  double cs2;
  orc_cell ***fluid; /* [Nx+2][Ny+2][Nz+2], lbm.cpp:17-28 */
  double *Fm;        /* implicit-IBM Fm[3] per cell (LBM.h:571-573) */
  struct orc_thermal *th; /* ENERGY fields (LBM.h:590-596), NULL when off */
  struct orc_q7 *q7;      /* D3Q7 temperature (no reference counterpart) */
  double *stats;          /* MORE_DETAILS averages and rms: 26 per cell, [cell][26] */
};

/* D3Q7 temperature (declared in lbm_oracle.h with its specification) */
typedef struct orc_q7 {
  double tau_g, g_beta, T_ref;
  double j_dot_n; /* heat flux through a TYPE_Q wall face into the fluid */
  int boussinesq;
  double *g, *g_star; /* 7 per cell */
  double *T, *q;      /* 1 and 3 per cell */
  double *F0y;        /* set-up F_y */
  int16_t *typeE;
} orc_q7;

static const int q7_cx[7] = {0, 1, -1, 0, 0, 0, 0};
static const int q7_cy[7] = {0, 0, 0, 1, -1, 0, 0};
static const int q7_cz[7] = {0, 0, 0, 0, 0, 1, -1};
static const int q7_opp[7] = {0, 2, 1, 4, 3, 6, 5};
static const double q7_w[7] = {0.25, 0.125, 0.125, 0.125, 0.125, 0.125, 0.125};

/* The ENERGY members of LATTICE (LBM.h:590-596) as arrays over LIN(i,j,k),
 * and the case constants the thermal branch reads (LBM.h:150-161,
 * 206-220, 233-241). */
typedef struct orc_thermal {
  double tau_g, nu, cp, j_dot_n;
  int boussinesq;       /* Rayleigh_Benard: F[1] = g_beta (T - T_ref) in MacroProp */
  double g_beta, T_ref; /* lbm.cpp:803-806, T_ref = (Tb + Tt)/2 */
  double *g, *g_star, *F_vis; /* q per cell */
  double *T, *q;              /* 1 and 3 per cell */
  int16_t *typeE;             /* type_Energy */
} orc_thermal;

/* D3Q19, LBM.h:542-552 */
static const int d3q19_cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
static const int d3q19_cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
static const int d3q19_cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
static const double d3q19_w[19] = {1.0 / 3, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                                   1.0 / 18, 1.0 / 18, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                   1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                   1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
static const int d3q19_opp[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

/* D2Q9, LBM.h:530-540 */
static const int d2q9_cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int d2q9_cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int d2q9_cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
static const double d2q9_w[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                                 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
static const int d2q9_opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

static int g_threads = 0;

void orc_set_threads(int nthreads) {
  g_threads = nthreads;
  if (nthreads > 0) omp_set_num_threads(nthreads);
}

int orc_get_threads(void) { return omp_get_max_threads(); }

int orc_sizeof_cell(void) { return (int)sizeof(orc_cell); }

/* LBM::LBM, lbm.cpp:17-28: one allocation per (i,j) row of Nz+2 cells.  Cell
 * defaults follow LATTICE's member initialisers (LBM.h:560-566). */
orc_lbm *orc_create(int Nx, int Ny, int Nz, double tau, int q) {
  if (Nx < 1 || Ny < 1 || Nz < 1 || (q != 19 && q != 9)) return NULL;
  orc_lbm *lb = (orc_lbm *)calloc(1, sizeof(orc_lbm));
  lb->Nx = Nx;
  lb->Ny = Ny;
  lb->Nz = Nz;
  lb->tau = tau;
  lb->q = q;
  if (q == 19) {
    lb->cx = d3q19_cx; lb->cy = d3q19_cy; lb->cz = d3q19_cz;
    lb->opp = d3q19_opp; lb->w = d3q19_w;
  } else {
    lb->cx = d2q9_cx; lb->cy = d2q9_cy; lb->cz = d2q9_cz;
    lb->opp = d2q9_opp; lb->w = d2q9_w;
  }
  lb->cs2 = pow(ORC_DX, 2) / (3 * pow(ORC_DT, 2)); /* LBM.h:86 */
  lb->fluid = (orc_cell ***)malloc(sizeof(orc_cell **) * (Nx + 2));
  for (int i = 0; i < Nx + 2; i++) {
    lb->fluid[i] = (orc_cell **)malloc(sizeof(orc_cell *) * (Ny + 2));
    for (int j = 0; j < Ny + 2; j++) {
      orc_cell *row = (orc_cell *)malloc(sizeof(orc_cell) * (Nz + 2));
      for (int k = 0; k < Nz + 2; k++) {
        memset(&row[k], 0, sizeof(orc_cell));
        row[k].type = ORC_TYPE_F;
        row[k].density = 1.0;
      }
      lb->fluid[i][j] = row;
    }
  }
  lb->Fm = (double *)calloc((size_t)3 * (Nx + 2) * (Ny + 2) * (Nz + 2), sizeof(double));
  return lb;
}

void orc_destroy(orc_lbm *lb) {
  if (!lb) return;
  for (int i = 0; i < lb->Nx + 2; i++) {
    for (int j = 0; j < lb->Ny + 2; j++) free(lb->fluid[i][j]);
    free(lb->fluid[i]);
  }
  free(lb->fluid);
  free(lb->Fm);
  free(lb->stats);
  if (lb->th) {
    free(lb->th->g); free(lb->th->g_star); free(lb->th->F_vis);
    free(lb->th->T); free(lb->th->q); free(lb->th->typeE);
    free(lb->th);
  }
  if (lb->q7) {
    free(lb->q7->g); free(lb->q7->g_star); free(lb->q7->T); free(lb->q7->q);
    free(lb->q7->F0y); free(lb->q7->typeE);
    free(lb->q7);
  }
  free(lb);
}

int64_t orc_ncells(const orc_lbm *lb) {
  return (int64_t)(lb->Nx + 2) * (lb->Ny + 2) * (lb->Nz + 2);
}

int orc_q(const orc_lbm *lb) { return lb->q; }

void orc_tables(const orc_lbm *lb, int *cx, int *cy, int *cz, int *opp, double *w) {
  for (int l = 0; l < lb->q; l++) {
    cx[l] = lb->cx[l]; cy[l] = lb->cy[l]; cz[l] = lb->cz[l];
    opp[l] = lb->opp[l]; w[l] = lb->w[l];
  }
}

#define LIN(lb, i, j, k) ((((int64_t)(i)) * ((lb)->Ny + 2) + (j)) * ((lb)->Nz + 2) + (k))

void orc_set_cells(orc_lbm *lb, const int16_t *type, const double *rho,
                   const double *ux, const double *uy, const double *uz,
                   const double *Fx, const double *Fy, const double *Fz) {
  for (int i = 0; i < lb->Nx + 2; i++)
    for (int j = 0; j < lb->Ny + 2; j++)
      for (int k = 0; k < lb->Nz + 2; k++) {
        int64_t n = LIN(lb, i, j, k);
        orc_cell *c = &lb->fluid[i][j][k];
        if (type) c->type = type[n];
        if (rho) c->density = rho[n];
        if (ux) c->ux = ux[n];
        if (uy) c->uy = uy[n];
        if (uz) c->uz = uz[n];
        if (Fx) c->F[0] = Fx[n];
        if (Fy) c->F[1] = Fy[n];
        if (Fz) c->F[2] = Fz[n];
      }
}

void orc_get_macro(const orc_lbm *lb, double *rho, double *ux, double *uy, double *uz) {
  for (int i = 0; i < lb->Nx + 2; i++)
    for (int j = 0; j < lb->Ny + 2; j++)
      for (int k = 0; k < lb->Nz + 2; k++) {
        int64_t n = LIN(lb, i, j, k);
        const orc_cell *c = &lb->fluid[i][j][k];
        if (rho) rho[n] = c->density;
        if (ux) ux[n] = c->ux;
        if (uy) uy[n] = c->uy;
        if (uz) uz[n] = c->uz;
      }
}

void orc_get_force(const orc_lbm *lb, double *Fx, double *Fy, double *Fz) {
  for (int i = 0; i < lb->Nx + 2; i++)
    for (int j = 0; j < lb->Ny + 2; j++)
      for (int k = 0; k < lb->Nz + 2; k++) {
        int64_t n = LIN(lb, i, j, k);
        const orc_cell *c = &lb->fluid[i][j][k];
        if (Fx) Fx[n] = c->F[0];
        if (Fy) Fy[n] = c->F[1];
        if (Fz) Fz[n] = c->F[2];
      }
}

void orc_get_fm(const orc_lbm *lb, double *Fmx, double *Fmy, double *Fmz) {
  int64_t nc = orc_ncells(lb);
  for (int64_t n = 0; n < nc; n++) {
    if (Fmx) Fmx[n] = lb->Fm[3 * n + 0];
    if (Fmy) Fmy[n] = lb->Fm[3 * n + 1];
    if (Fmz) Fmz[n] = lb->Fm[3 * n + 2];
  }
}

static void copy_pops(orc_lbm *lb, double *dst, const double *src, int to_host, int star) {
  int q = lb->q;
  for (int i = 0; i < lb->Nx + 2; i++)
    for (int j = 0; j < lb->Ny + 2; j++)
      for (int k = 0; k < lb->Nz + 2; k++) {
        int64_t n = LIN(lb, i, j, k);
        orc_cell *c = &lb->fluid[i][j][k];
        double *p = star ? c->f_star : c->f;
        for (int l = 0; l < q; l++) {
          if (to_host) dst[n * q + l] = p[l];
          else p[l] = src[n * q + l];
        }
      }
}

void orc_get_f(const orc_lbm *lb, double *f) { copy_pops((orc_lbm *)lb, f, NULL, 1, 0); }
void orc_get_fstar(const orc_lbm *lb, double *fs) { copy_pops((orc_lbm *)lb, fs, NULL, 1, 1); }
void orc_set_f(orc_lbm *lb, const double *f) { copy_pops(lb, NULL, f, 0, 0); }
void orc_set_fstar(orc_lbm *lb, const double *fs) { copy_pops(lb, NULL, fs, 0, 1); }

/* LBM::Calc_feq, lbm.cpp:628-633: the cubic equilibrium, written in the
 * reference's expression order. */
static inline double calc_feq(const orc_lbm *lb, const orc_cell *c, int l) {
  const double cs2 = lb->cs2;
  double u_dot_u = pow(c->ux, 2) + pow(c->uy, 2) + pow(c->uz, 2);
  double c_dot_u = c->ux * lb->cx[l] + c->uy * lb->cy[l] + c->uz * lb->cz[l];
  double f_eq = c->density * lb->w[l] *
                (1 + c_dot_u / cs2 + pow(c_dot_u, 2) / (2 * cs2 * cs2) - u_dot_u / (2 * cs2) +
                 pow(c_dot_u, 3) / (6 * pow(cs2, 3)) - c_dot_u * u_dot_u / (2 * cs2 * cs2));
  return f_eq;
}

double orc_feq(const orc_lbm *lb, double rho, double ux, double uy, double uz, int l) {
  orc_cell c;
  memset(&c, 0, sizeof c);
  c.density = rho; c.ux = ux; c.uy = uy; c.uz = uz;
  return calc_feq(lb, &c, l);
}

/* ---- thermal branch (ENERGY): LBM.h:590-596, lbm.cpp:121-178, 263-289,
 * 636-653, 777-806, 964-1028 ------------------------------------------- */

int orc_thermal_enable(orc_lbm *lb, double tau_g, double nu, double cp, double j_dot_n,
                       int boussinesq, double g_beta, double T_ref) {
  if (lb->th) return 0;
  const int64_t n = orc_ncells(lb);
  orc_thermal *th = (orc_thermal *)calloc(1, sizeof(orc_thermal));
  th->tau_g = tau_g; th->nu = nu; th->cp = cp; th->j_dot_n = j_dot_n;
  th->boussinesq = boussinesq; th->g_beta = g_beta; th->T_ref = T_ref;
  th->g = (double *)calloc((size_t)n * lb->q, sizeof(double));
  th->g_star = (double *)calloc((size_t)n * lb->q, sizeof(double));
  th->F_vis = (double *)calloc((size_t)n * lb->q, sizeof(double));
  th->T = (double *)calloc((size_t)n, sizeof(double));
  th->q = (double *)calloc((size_t)n * 3, sizeof(double));
  th->typeE = (int16_t *)calloc((size_t)n, sizeof(int16_t)); /* TYPE_F, LBM.h:594 */
  lb->th = th;
  return 0;
}

void orc_set_thermal(orc_lbm *lb, const double *T, const int16_t *typeE) {
  const int64_t n = orc_ncells(lb);
  if (!lb->th) return;
  for (int64_t c = 0; c < n; c++) {
    if (T) lb->th->T[c] = T[c];
    if (typeE) lb->th->typeE[c] = typeE[c];
  }
}

void orc_get_thermal(const orc_lbm *lb, double *T, double *qx, double *qy, double *qz) {
  const int64_t n = orc_ncells(lb);
  if (!lb->th) return;
  for (int64_t c = 0; c < n; c++) {
    if (T) T[c] = lb->th->T[c];
    if (qx) qx[c] = lb->th->q[3 * c + 0];
    if (qy) qy[c] = lb->th->q[3 * c + 1];
    if (qz) qz[c] = lb->th->q[3 * c + 2];
  }
}

void orc_get_g(const orc_lbm *lb, double *g) {
  if (!lb->th) return;
  memcpy(g, lb->th->g, sizeof(double) * (size_t)orc_ncells(lb) * lb->q);
}

void orc_set_g(orc_lbm *lb, const double *g) {
  if (!lb->th) return;
  memcpy(lb->th->g, g, sizeof(double) * (size_t)orc_ncells(lb) * lb->q);
}

/* ---- D3Q7 temperature (specification in lbm_oracle.h) ---- */
int orc_q7_enable(orc_lbm *lb, double tau_g, int boussinesq, double g_beta, double T_ref) {
  if (lb->q7 || lb->th || lb->q != 19) return 0;
  const size_t n = (size_t)orc_ncells(lb);
  orc_q7 *t = (orc_q7 *)calloc(1, sizeof(orc_q7));
  t->tau_g = tau_g;
  t->boussinesq = boussinesq;
  t->g_beta = g_beta;
  t->T_ref = T_ref;
  t->g = (double *)calloc(n * 7, sizeof(double));
  t->g_star = (double *)calloc(n * 7, sizeof(double));
  t->T = (double *)calloc(n, sizeof(double));
  t->q = (double *)calloc(n * 3, sizeof(double));
  t->F0y = (double *)calloc(n, sizeof(double));
  t->typeE = (int16_t *)calloc(n, sizeof(int16_t));
  lb->q7 = t;
  return 1;
}

void orc_q7_set_flux(orc_lbm *lb, double j_dot_n) {
  if (lb->q7) lb->q7->j_dot_n = j_dot_n;
}

void orc_q7_set(orc_lbm *lb, const double *T, const int16_t *typeE) {
  if (!lb->q7) return;
  for (int i = 0; i <= lb->Nx + 1; i++)
    for (int j = 0; j <= lb->Ny + 1; j++)
      for (int k = 0; k <= lb->Nz + 1; k++) {
        const int64_t c = LIN(lb, i, j, k);
        if (T) lb->q7->T[c] = T[c];
        if (typeE) lb->q7->typeE[c] = typeE[c];
        lb->q7->F0y[c] = lb->fluid[i][j][k].F[1];
      }
}

void orc_q7_get(const orc_lbm *lb, double *T, double *qx, double *qy, double *qz) {
  if (!lb->q7) return;
  for (int64_t c = 0; c < orc_ncells(lb); c++) {
    if (T) T[c] = lb->q7->T[c];
    if (qx) qx[c] = lb->q7->q[3 * c + 0];
    if (qy) qy[c] = lb->q7->q[3 * c + 1];
    if (qz) qz[c] = lb->q7->q[3 * c + 2];
  }
}

void orc_q7_get_g(const orc_lbm *lb, double *g) {
  if (!lb->q7) return;
  memcpy(g, lb->q7->g, sizeof(double) * (size_t)orc_ncells(lb) * 7);
}

static inline double q7_geq(int i, double T, double ux, double uy, double uz) {
  double c_dot_u = q7_cx[i] * ux + q7_cy[i] * uy + q7_cz[i] * uz;
  return q7_w[i] * T * (1 + c_dot_u / 0.25);
}

/* Determine_Wall_Temperature, lbm.cpp:965-981: the mean T of the cell's
 * TYPE_S neighbours (l = 0 included), NaN when there is none. */
double orc_wall_temperature(const orc_lbm *lb, int i, int j, int k) {
  int num = 0;
  double Twall = 0;
  for (int l = 0; l < lb->q; l++) {
    int i_nb = (i + lb->cx[l] + (lb->Nx + 2)) % (lb->Nx + 2);
    int j_nb = (j + lb->cy[l] + (lb->Ny + 2)) % (lb->Ny + 2);
    int k_nb = (k + lb->cz[l] + (lb->Nz + 2)) % (lb->Nz + 2);
    if (lb->fluid[i_nb][j_nb][k_nb].type == ORC_TYPE_S) {
      Twall += lb->th->T[LIN(lb, i_nb, j_nb, k_nb)];
      num++;
    }
  }
  Twall /= num;
  return Twall;
}

/* Calc_geq, lbm.cpp:636-653.  TYPE_S: rho of cell (i,j,k), the wall
 * temperature of (i,j,k) and the wall velocity uw; otherwise the cell's own
 * rho, u and T. */
static double calc_geq(const orc_lbm *lb, int i, int j, int k, int l, short type,
                       const double uw[3]) {
  const double cs2 = lb->cs2;
  const orc_cell *c = &lb->fluid[i][j][k];
  double g_eq;
  if (type == ORC_TYPE_S) {
    double Twall = orc_wall_temperature(lb, i, j, k);
    double u_dot_u = pow(uw[0], 2) + pow(uw[1], 2) + pow(uw[2], 2);
    double c_dot_u = uw[0] * lb->cx[l] + uw[1] * lb->cy[l] + uw[2] * lb->cz[l];
    g_eq = c->density * lb->w[l] * Twall *
           (1 + c_dot_u / cs2 + pow(c_dot_u, 2) / (2 * cs2 * cs2) - u_dot_u / (2 * cs2) +
            pow(c_dot_u, 3) / (6 * pow(cs2, 3)) - c_dot_u * u_dot_u / (2 * cs2 * cs2));
  } else {
    double u_dot_u = pow(c->ux, 2) + pow(c->uy, 2) + pow(c->uz, 2);
    double c_dot_u = c->ux * lb->cx[l] + c->uy * lb->cy[l] + c->uz * lb->cz[l];
    g_eq = c->density * lb->w[l] * lb->th->T[LIN(lb, i, j, k)] *
           (1 + c_dot_u / cs2 + pow(c_dot_u, 2) / (2 * cs2 * cs2) - u_dot_u / (2 * cs2) +
            pow(c_dot_u, 3) / (6 * pow(cs2, 3)) - c_dot_u * u_dot_u / (2 * cs2 * cs2));
  }
  return g_eq;
}

double orc_geq(const orc_lbm *lb, int i, int j, int k, int l, int as_solid, double uwx,
               double uwy, double uwz) {
  const double uw[3] = {uwx, uwy, uwz};
  return calc_geq(lb, i, j, k, l, as_solid ? ORC_TYPE_S : ORC_TYPE_F, uw);
}

/* The ENERGY part of ConvLBM, lbm.cpp:121-177: the strain-rate tensor from
 * the non-equilibrium f, the dissipation Phi, the source F_vis with its
 * time derivative, and BGK for g with relaxation time tau_g. */
static void conv_energy(const orc_lbm *lb, int i, int j, int k) {
  const orc_thermal *th = lb->th;
  const double cs2 = lb->cs2, tau = lb->tau, tau_g = th->tau_g;
  const orc_cell *c = &lb->fluid[i][j][k];
  const int64_t n = LIN(lb, i, j, k);
  const int q = lb->q;
  const int *cx = lb->cx, *cy = lb->cy, *cz = lb->cz;
  double rho = c->density, ux = c->ux, uy = c->uy, uz = c->uz;
  double S[3][3];
  for (int row = 0; row < 3; row++)
    for (int col = 0; col < 3; col++) {
      S[row][col] = 0;
      for (int l = 0; l < q; l++) {
        double f_eq = calc_feq(lb, c, l);
        if (row == 0 && col == 0) S[row][col] += cx[l] * cx[l] * (c->f[l] - f_eq);
        else if (row == 0 && col == 1) S[row][col] += cx[l] * cy[l] * (c->f[l] - f_eq);
        else if (row == 0 && col == 2) S[row][col] += cx[l] * cz[l] * (c->f[l] - f_eq);
        else if (row == 1 && col == 1) S[row][col] += cy[l] * cy[l] * (c->f[l] - f_eq);
        else if (row == 1 && col == 2) S[row][col] += cy[l] * cz[l] * (c->f[l] - f_eq);
        else if (row == 2 && col == 2) S[row][col] += cz[l] * cz[l] * (c->f[l] - f_eq);
      }
    }
  S[1][0] = S[0][1]; S[2][0] = S[0][2]; S[2][1] = S[1][2];
  for (int row = 0; row < 3; row++)
    for (int col = 0; col < 3; col++) S[row][col] = -S[row][col] / (2 * rho * tau * cs2);
  double Phi = 0;
  for (int row = 0; row < 3; row++)
    for (int col = 0; col < 3; col++) Phi += S[row][col] * S[row][col];
  Phi = 2 * th->nu * rho * Phi;
  const double zero[3] = {0, 0, 0};
  for (int l = 0; l < q; l++) {
    double c_dot_u = ux * cx[l] + uy * cy[l] + uz * cz[l];
    double F_before = th->F_vis[n * q + l];
    double F = lb->w[l] * Phi / th->cp * (1 + c_dot_u * (tau_g - 0.5) / (cs2 * tau_g));
    th->F_vis[n * q + l] = F;
    double dF_dt = (F - F_before) / ORC_DT;
    double g_eq = calc_geq(lb, i, j, k, l, c->type, zero);
    double g = th->g[n * q + l];
    th->g_star[n * q + l] = g - (g - g_eq) / tau_g + ORC_DT * F + 0.5 * pow(ORC_DT, 2) * dF_dt;
  }
}

/* Determine_Twall_Neumann, lbm.cpp:983-1028, for fluid cell (i,j,k) with
 * wall velocity uw: pos = opposite of every solid direction (in l order),
 * notpos = the other directions, n = mean of c over pos (not normalised),
 * T_w = (j.n - sum_notpos g c.n) / sum_pos eq c.n. */
static double twall_neumann(const orc_lbm *lb, int i, int j, int k, const double uw[3]) {
  const orc_thermal *th = lb->th;
  const int q = lb->q;
  const double cs2 = lb->cs2;
  const orc_cell *c = &lb->fluid[i][j][k];
  const int64_t n = LIN(lb, i, j, k);
  int pos[19], npos = 0, in_pos[19] = {0};
  for (int l = 0; l < q; l++) {
    int i_nb = (i + lb->cx[l] + (lb->Nx + 2)) % (lb->Nx + 2);
    int j_nb = (j + lb->cy[l] + (lb->Ny + 2)) % (lb->Ny + 2);
    int k_nb = (k + lb->cz[l] + (lb->Nz + 2)) % (lb->Nz + 2);
    if (lb->fluid[i_nb][j_nb][k_nb].type == ORC_TYPE_S) {
      pos[npos++] = lb->opp[l];
      in_pos[lb->opp[l]] = 1;
    }
  }
  double normal[3] = {0, 0, 0};
  for (int p = 0; p < npos; p++) {
    normal[0] += lb->cx[pos[p]]; normal[1] += lb->cy[pos[p]]; normal[2] += lb->cz[pos[p]];
  }
  normal[0] /= npos; normal[1] /= npos; normal[2] /= npos;
  double sum_cdotn_notpos = 0;
  for (int l = 0; l < q; l++) {
    if (in_pos[l]) continue;
    sum_cdotn_notpos += th->g[n * q + l] *
                        (lb->cx[l] * normal[0] + lb->cy[l] * normal[1] + lb->cz[l] * normal[2]);
  }
  double uwall_dot_uwall = uw[0] * uw[0] + uw[1] * uw[1] + uw[2] * uw[2];
  double sum_cdotn_pos = 0;
  for (int p = 0; p < npos; p++) {
    int l = pos[p];
    double uwall_dot_c = uw[0] * lb->cx[l] + uw[1] * lb->cy[l] + uw[2] * lb->cz[l];
    double eq = lb->w[l] * c->density *
                (1 + uwall_dot_c / cs2 + pow(uwall_dot_c, 2) / (2 * cs2 * cs2) -
                 uwall_dot_uwall / (2 * cs2) + pow(uwall_dot_c, 3) / (6 * pow(cs2, 3)) -
                 uwall_dot_c * uwall_dot_uwall / (2 * cs2 * cs2));
    sum_cdotn_pos += eq * (lb->cx[l] * normal[0] + lb->cy[l] * normal[1] + lb->cz[l] * normal[2]);
  }
  double numerator = th->j_dot_n - sum_cdotn_notpos;
  double denum = sum_cdotn_pos;
  return numerator / denum;
}

/* LBM::Init, lbm.cpp:31-57: f = f_star = feq for every cell, ghosts and
 * solids included (the type filter is commented out, lbm.cpp:37). */
void orc_init(orc_lbm *lb) {
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz, q = lb->q;
#pragma omp parallel for
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++)
        for (int l = 0; l < q; l++) {
          double f_eq = calc_feq(lb, &lb->fluid[i][j][k], l);
          lb->fluid[i][j][k].f[l] = f_eq;
          lb->fluid[i][j][k].f_star[l] = f_eq;
          if (lb->q7 && l < 7) {
            const int64_t n = LIN(lb, i, j, k);
            orc_cell *c = &lb->fluid[i][j][k];
            lb->q7->g[n * 7 + l] = lb->q7->g_star[n * 7 + l] = q7_geq(l, lb->q7->T[n], c->ux, c->uy, c->uz);
            /* the first collision already carries the buoyancy of the
             * set-up T, as every later one carries that of MacroProp's */
            if (l == 0 && lb->q7->boussinesq && c->type == ORC_TYPE_F)
              c->F[1] = lb->q7->F0y[n] + lb->q7->g_beta * (lb->q7->T[n] - lb->q7->T_ref);
          }
          if (lb->th) {
            /* lbm.cpp:45-50: g = g_star = Calc_geq with the cell's own type
             * and a zero wall velocity */
            const double zero[3] = {0, 0, 0};
            double g_eq = calc_geq(lb, i, j, k, l, lb->fluid[i][j][k].type, zero);
            const int64_t n = LIN(lb, i, j, k);
            lb->th->g[n * q + l] = g_eq;
            lb->th->g_star[n * q + l] = g_eq;
            lb->th->F_vis[n * q + l] = 0.0;
          }
        }
}

/* LBM::ConvLBM, lbm.cpp:75-118: BGK with the Guo source.  Returns 1 when an
 * equilibrium is negative (the reference terminates, lbm.cpp:114-117). */
static inline int conv_lbm(const orc_lbm *lb, orc_cell *c) {
  const double cs2 = lb->cs2;
  const double omega = ORC_DT / lb->tau;
  double ux = c->ux, uy = c->uy, uz = c->uz;
  double Fx = c->F[0], Fy = c->F[1], Fz = c->F[2];
  double F_dot_u = Fx * ux + Fy * uy + Fz * uz;
  int bad = 0;
  for (int l = 0; l < lb->q; l++) {
    double c_dot_u = ux * lb->cx[l] + uy * lb->cy[l] + uz * lb->cz[l];
    double f_eq = calc_feq(lb, c, l);
    double F_dot_c = Fx * lb->cx[l] + Fy * lb->cy[l] + Fz * lb->cz[l];
    double Source = (1 - omega * 0.5) * lb->w[l] *
                    (F_dot_c / cs2 + (F_dot_c) * (c_dot_u) / (cs2 * cs2) - F_dot_u / cs2);
    c->f_star[l] = c->f[l] * (1 - omega) + omega * f_eq + Source * ORC_DT;
    if (f_eq < 0) bad = 1;
  }
  return bad;
}

/* LBM::Collision, lbm.cpp:60-73: TYPE_F cells only. */
int orc_collision(orc_lbm *lb) {
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz;
  int bad = 0;
#pragma omp parallel for reduction(| : bad)
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++)
        if (lb->fluid[i][j][k].type == ORC_TYPE_F) {
          bad |= conv_lbm(lb, &lb->fluid[i][j][k]);
          if (lb->th) conv_energy(lb, i, j, k);
          if (lb->q7) {
            orc_q7 *t = lb->q7;
            const orc_cell *c = &lb->fluid[i][j][k];
            const int64_t n = LIN(lb, i, j, k);
            const double om = ORC_DT / t->tau_g;
            for (int l = 0; l < 7; l++)
              t->g_star[n * 7 + l] = t->g[n * 7 + l] * (1 - om) + om * q7_geq(l, t->T[n], c->ux, c->uy, c->uz);
          }
        }
  return bad ? ORC_ERR_FEQ_NEGATIVE : ORC_OK;
}

#define WRAP(a, n) (((a) + (n)) % (n))

/* LBM::Determine_Wall_Velocity, lbm.cpp:941-962: the mean velocity of the
 * cell's solid neighbours whose velocity is nonzero, summed in l order. */
void orc_wall_velocity(const orc_lbm *lb, int i, int j, int k, double uw[3]) {
  int num = 0;
  uw[0] = uw[1] = uw[2] = 0.0;
  for (int l = 0; l < lb->q; l++) {
    int i_nb = WRAP(i + lb->cx[l], lb->Nx + 2);
    int j_nb = WRAP(j + lb->cy[l], lb->Ny + 2);
    int k_nb = WRAP(k + lb->cz[l], lb->Nz + 2);
    const orc_cell *nb = &lb->fluid[i_nb][j_nb][k_nb];
    if (nb->type == ORC_TYPE_S) {
      if (nb->ux != 0 || nb->uy != 0 || nb->uz != 0) {
        uw[0] += nb->ux;
        uw[1] += nb->uy;
        uw[2] += nb->uz;
        num++;
      }
    }
  }
  if (num != 0) {
    uw[0] /= num; uw[1] /= num; uw[2] /= num;
  }
}

/* second-order inflow equilibrium of the face loops (lbm.cpp:325-345 etc.)
 * from the rho, u (and, with ENERGY, T) of cell s into cell d:
 * f = rho w (1 + ...), g = rho w T (1 + ...) */
static inline void inflow_feq(const orc_lbm *lb, int di, int dj, int dk, int si, int sj, int sk) {
  const double cs2 = lb->cs2;
  orc_cell *dst = &lb->fluid[di][dj][dk];
  const orc_cell *src = &lb->fluid[si][sj][sk];
  double rho = src->density, ux = src->ux, uy = src->uy, uz = src->uz;
  double u_dot_u = pow(ux, 2) + pow(uy, 2) + pow(uz, 2);
  const orc_thermal *th = lb->th;
  const double T = th ? th->T[LIN(lb, si, sj, sk)] : 0.0;
  for (int l = 0; l < lb->q; l++) {
    double c_dot_u = ux * lb->cx[l] + uy * lb->cy[l] + uz * lb->cz[l];
    double f_eq = rho * lb->w[l] *
                  (1 + c_dot_u / cs2 + pow(c_dot_u, 2) / (2 * cs2 * cs2) - u_dot_u / (2 * cs2));
    dst->f[l] = f_eq;
    if (th) {
      double g_eq = rho * lb->w[l] * T *
                    (1 + c_dot_u / cs2 + pow(c_dot_u, 2) / (2 * cs2 * cs2) - u_dot_u / (2 * cs2));
      th->g[LIN(lb, di, dj, dk) * lb->q + l] = g_eq;
    }
  }
}

/* outlet: populations a, b, c (and with ENERGY g a, b, c) of cell s into
 * cell d (lbm.cpp:303-311 etc.) */
static inline void copy3(const orc_lbm *lb, int di, int dj, int dk, int si, int sj, int sk, int a, int b,
                         int c) {
  orc_cell *dst = &lb->fluid[di][dj][dk];
  const orc_cell *src = &lb->fluid[si][sj][sk];
  dst->f[a] = src->f[a];
  dst->f[b] = src->f[b];
  dst->f[c] = src->f[c];
  if (lb->th) {
    double *g = lb->th->g;
    const int64_t d = LIN(lb, di, dj, dk) * lb->q, s = LIN(lb, si, sj, sk) * lb->q;
    g[d + a] = g[s + a];
    g[d + b] = g[s + b];
    g[d + c] = g[s + c];
  }
}

/* LBM::Streaming, lbm.cpp:182-626. */
void orc_streaming(orc_lbm *lb) {
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz, q = lb->q;
  const double cs2 = lb->cs2;
  orc_cell ***fl = lb->fluid;
  orc_thermal *th = lb->th;

  /* fluid part, lbm.cpp:183-295: push from TYPE_F cells */
#pragma omp parallel for
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++) {
        if (fl[i][j][k].type != ORC_TYPE_F) continue;
        int have_uw = 0;
        double uwall[3] = {0.0, 0.0, 0.0};
        for (int l = 0; l < q; l++) {
          int i_nb = WRAP(i + lb->cx[l], Nx + 2);
          int j_nb = WRAP(j + lb->cy[l], Ny + 2);
          int k_nb = WRAP(k + lb->cz[l], Nz + 2);
          orc_cell *nb = &fl[i_nb][j_nb][k_nb];
          if (lb->q7 && l < 7) {
            /* D3Q7 (lbm_oracle.h); its directions are D3Q19's first seven */
            orc_q7 *t = lb->q7;
            const int64_t n = LIN(lb, i, j, k), nn = LIN(lb, i_nb, j_nb, k_nb);
            if (nb->type == ORC_TYPE_F) {
              t->g[nn * 7 + l] = t->g_star[n * 7 + l];
            } else if (nb->type == ORC_TYPE_S) {
              if (t->typeE[nn] == ORC_TYPE_T)
                t->g[n * 7 + q7_opp[l]] = -t->g_star[n * 7 + l] + 2 * q7_w[l] * t->T[nn];
              else if (t->typeE[nn] == ORC_TYPE_Q)
                t->g[n * 7 + q7_opp[l]] = t->g_star[n * 7 + l] + t->j_dot_n;
              else
                t->g[n * 7 + q7_opp[l]] = t->g_star[n * 7 + l];
            }
          }
          if (nb->type == ORC_TYPE_F) {
            nb->f[l] = fl[i][j][k].f_star[l];
            if (th) th->g[LIN(lb, i_nb, j_nb, k_nb) * q + l] = th->g_star[LIN(lb, i, j, k) * q + l];
          } else if (nb->type == ORC_TYPE_S) {
            if (!have_uw) {
              orc_wall_velocity(lb, i, j, k, uwall);
              have_uw = 1;
            }
            double uwall_dot_c = uwall[0] * lb->cx[l] + uwall[1] * lb->cy[l] + uwall[2] * lb->cz[l];
            fl[i][j][k].f[lb->opp[l]] = fl[i][j][k].f_star[l] - 2 * lb->w[l] * uwall_dot_c / cs2;
            /* thermal Dirichlet wall, lbm.cpp:277-280 (anti-bounce-back);
             * the Neumann wall (TYPE_Q) runs in the second sweep below */
            if (th && th->typeE[LIN(lb, i_nb, j_nb, k_nb)] == ORC_TYPE_T) {
              double g_eq = calc_geq(lb, i, j, k, l, ORC_TYPE_S, uwall);
              const int64_t n = LIN(lb, i, j, k);
              th->g[n * q + lb->opp[l]] = -th->g_star[n * q + l] + 2 * g_eq;
            }
          }
          /* any other type: nothing is written (frozen link) */
        }
      }

  /* thermal Neumann wall, lbm.cpp:281-287: after every push of the pass */
  if (th) {
#pragma omp parallel for
    for (int i = 0; i <= Nx + 1; i++)
      for (int j = 0; j <= Ny + 1; j++)
        for (int k = 0; k <= Nz + 1; k++) {
          if (fl[i][j][k].type != ORC_TYPE_F) continue;
          int have = 0;
          double uwall[3] = {0.0, 0.0, 0.0}, Twall = 0.0;
          for (int l = 0; l < q; l++) {
            int i_nb = WRAP(i + lb->cx[l], Nx + 2);
            int j_nb = WRAP(j + lb->cy[l], Ny + 2);
            int k_nb = WRAP(k + lb->cz[l], Nz + 2);
            if (fl[i_nb][j_nb][k_nb].type != ORC_TYPE_S) continue;
            if (th->typeE[LIN(lb, i_nb, j_nb, k_nb)] != ORC_TYPE_Q) continue;
            if (!have) {
              orc_wall_velocity(lb, i, j, k, uwall);
              Twall = twall_neumann(lb, i, j, k, uwall);
              have = 1;
            }
            double uwall_dot_c = uwall[0] * lb->cx[l] + uwall[1] * lb->cy[l] + uwall[2] * lb->cz[l];
            double uwall_dot_uwall = uwall[0] * uwall[0] + uwall[1] * uwall[1] + uwall[2] * uwall[2];
            th->g[LIN(lb, i, j, k) * q + lb->opp[l]] =
                lb->w[l] * fl[i][j][k].density * Twall *
                (1 + uwall_dot_c / cs2 + pow(uwall_dot_c, 2) / (2 * cs2 * cs2) - uwall_dot_uwall / (2 * cs2));
          }
        }
  }

  /* Face loops, lbm.cpp:299-625, serial.  They use D2Q9 direction indices
   * even in D3Q19 (the reference's own code) and the back face writes its
   * inflow into k = Nz+1 (lbm.cpp:610).  With ENERGY they move g alike
   * (outlet copies, inlet g_eq = rho w T (1 + ...)).  Only that back-face
   * write can reach a TYPE_F cell. */
  /* Left (i = 0), lbm.cpp:300-352 */
  for (int j = 0; j <= Ny + 1; j++)
    for (int k = 0; k <= Nz + 1; k++) {
      short t = fl[0][j][k].type;
      if (t == ORC_TYPE_OUT) copy3(lb, 0, j, k, 1, j, k, 1, 5, 8);
      else if (t == ORC_TYPE_IN) inflow_feq(lb, 0, j, k, 0, j, k);
    }
  /* Right (i = Nx+1), lbm.cpp:354-407 */
  for (int j = 0; j <= Ny + 1; j++)
    for (int k = 0; k <= Nz + 1; k++) {
      short t = fl[Nx + 1][j][k].type;
      if (t == ORC_TYPE_OUT) copy3(lb, Nx + 1, j, k, Nx, j, k, 3, 6, 7);
      else if (t == ORC_TYPE_IN) inflow_feq(lb, Nx + 1, j, k, Nx + 1, j, k);
    }
  /* Top (j = Ny+1), lbm.cpp:409-462 */
  for (int i = 0; i <= Nx + 1; i++)
    for (int k = 0; k <= Nz + 1; k++) {
      short t = fl[i][Ny + 1][k].type;
      if (t == ORC_TYPE_OUT) copy3(lb, i, Ny + 1, k, i, Ny, k, 4, 7, 8);
      else if (t == ORC_TYPE_IN) inflow_feq(lb, i, Ny + 1, k, i, Ny + 1, k);
    }
  /* Bottom (j = 0), lbm.cpp:464-516 */
  for (int i = 0; i <= Nx + 1; i++)
    for (int k = 0; k <= Nz + 1; k++) {
      short t = fl[i][0][k].type;
      if (t == ORC_TYPE_OUT) copy3(lb, i, 0, k, i, 1, k, 2, 5, 6);
      else if (t == ORC_TYPE_IN) inflow_feq(lb, i, 0, k, i, 0, k);
    }
  /* Front (k = Nz+1), lbm.cpp:518-570 */
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++) {
      short t = fl[i][j][Nz + 1].type;
      if (t == ORC_TYPE_OUT) copy3(lb, i, j, Nz + 1, i, j, Nz, 4, 7, 8);
      else if (t == ORC_TYPE_IN) inflow_feq(lb, i, j, Nz + 1, i, j, Nz + 1);
    }
  /* Back (k = 0), lbm.cpp:572-625: reads cell k = 0, writes k = Nz+1 */
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++) {
      short t = fl[i][j][0].type;
      if (t == ORC_TYPE_OUT) copy3(lb, i, j, 0, i, j, 1, 4, 7, 8);
      else if (t == ORC_TYPE_IN) inflow_feq(lb, i, j, Nz + 1, i, j, 0);
    }
}

/* LBM::MacroProp, lbm.cpp:657-812, every cell (the type filter is commented
 * out, lbm.cpp:667). */
int orc_macroprop(orc_lbm *lb) {
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz, q = lb->q;
  int bad = 0;
#pragma omp parallel for reduction(| : bad)
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++) {
        orc_cell *c = &lb->fluid[i][j][k];
        double rho = 0, rho_ux = 0, rho_uy = 0, rho_uz = 0;
        for (int l = 0; l < q; l++) {
          rho += c->f[l];
          rho_ux += c->f[l] * lb->cx[l];
          rho_uy += c->f[l] * lb->cy[l];
          rho_uz += c->f[l] * lb->cz[l];
        }
        c->density = rho;
        if (lb->q7 && c->type == ORC_TYPE_F) {
          /* D3Q7 (lbm_oracle.h): T and the buoyancy before u */
          orc_q7 *t = lb->q7;
          const int64_t n = LIN(lb, i, j, k);
          double T = 0, qx = 0, qy = 0, qz = 0;
          for (int l = 0; l < 7; l++) {
            T += t->g[n * 7 + l];
            qx += t->g[n * 7 + l] * q7_cx[l];
            qy += t->g[n * 7 + l] * q7_cy[l];
            qz += t->g[n * 7 + l] * q7_cz[l];
          }
          t->T[n] = T;
          t->q[3 * n + 0] = qx;
          t->q[3 * n + 1] = qy;
          t->q[3 * n + 2] = qz;
          /* a D3Q7 T may be a relative temperature (negative about T_ref):
             the guard catches blow-up only, |T| > 100 or not finite */
          if (!(fabs(T) <= 100)) bad |= 2;
          if (t->boussinesq) c->F[1] = t->F0y[n] + t->g_beta * (T - t->T_ref);
        }
        c->ux = (rho_ux + 0.5 * ORC_DT * c->F[0]) / rho;
        c->uy = (rho_uy + 0.5 * ORC_DT * c->F[1]) / rho;
        c->uz = (rho_uz + 0.5 * ORC_DT * c->F[2]) / rho;
        if (c->density < 0 || c->density > 100) bad |= 1;
        if (lb->th) {
          /* lbm.cpp:777-806 */
          orc_thermal *th = lb->th;
          const int64_t n = LIN(lb, i, j, k);
          double T = 0, sum_gx = 0, sum_gy = 0, sum_gz = 0;
          for (int l = 0; l < q; l++) {
            T += th->g[n * q + l];
            sum_gx += th->g[n * q + l] * lb->cx[l];
            sum_gy += th->g[n * q + l] * lb->cy[l];
            sum_gz += th->g[n * q + l] * lb->cz[l];
          }
          th->T[n] = T;
          th->q[3 * n + 0] = sum_gx;
          th->q[3 * n + 1] = sum_gy;
          th->q[3 * n + 2] = sum_gz;
          if (T < 0 || T > 100) bad |= 2;
          if (th->boussinesq) c->F[1] = th->g_beta * (T - th->T_ref);
        }
      }
  if (bad & 1) return ORC_ERR_DENSITY_RANGE;
  return bad ? ORC_ERR_TEMPERATURE_RANGE : ORC_OK;
}

/* main.cpp:212-221 */
int orc_run_plain(orc_lbm *lb, int nsteps) {
  for (int s = 0; s < nsteps; s++) {
    int e = orc_collision(lb);
    if (e) return e;
    orc_streaming(lb);
    e = orc_macroprop(lb);
    if (e) return e;
  }
  return ORC_OK;
}

void orc_export_links(const orc_lbm *lb, int32_t *src, uint8_t *kind) {
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz, q = lb->q;
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++) {
        int64_t n = LIN(lb, i, j, k);
        int fluid = lb->fluid[i][j][k].type == ORC_TYPE_F;
        for (int m = 0; m < q; m++) {
          if (!fluid) {
            src[n * q + m] = -1;
            kind[n * q + m] = 255;
            continue;
          }
          /* the pushing cell is the one whose neighbour in direction m is
           * this cell: (i - cx[m], ...) wrapped as in lbm.cpp:197-199 */
          int is = WRAP(i - lb->cx[m], Nx + 2);
          int js = WRAP(j - lb->cy[m], Ny + 2);
          int ks = WRAP(k - lb->cz[m], Nz + 2);
          short t = lb->fluid[is][js][ks].type;
          src[n * q + m] = (int32_t)LIN(lb, is, js, ks);
          kind[n * q + m] = t == ORC_TYPE_F ? 0 : (t == ORC_TYPE_S ? 1 : 2);
        }
      }
}

/* ---------------------------------------------------------------------------
 * Implicit diffuse-interface IBM (lbm.cpp:1048-1417)
 * ------------------------------------------------------------------------- */

/* kernel(), lbm.cpp:1095-1102 */
double orc_ibm_kernel(double r) {
  double d;
  if (fabs(r) >= 0 && fabs(r) <= 1)
    d = (3 - 2 * fabs(r) + sqrt(1 + 4 * fabs(r) - 4 * pow(r, 2))) / 8;
  else if (fabs(r) >= 1 && fabs(r) <= 2)
    d = (5 - 2 * fabs(r) - sqrt(-7 + 12 * fabs(r) - 4 * pow(r, 2))) / 8;
  else
    d = 0;
  return d;
}

/* NearestLatticeCoordinate, lbm.cpp:1074-1092: the box point with the
 * largest positive projection on the marker normal (last one wins a tie). */
void orc_ibm_nearest(int xmin, int xmax, int ymin, int ymax, int zmin, int zmax,
                     double mx, double my, double mz, double nx, double ny, double nz,
                     int *i, int *j, int *k) {
  double dot = 0;
  for (int x = xmin; x <= xmax; x++)
    for (int y = ymin; y <= ymax; y++)
      for (int z = zmin; z <= zmax; z++) {
        double v0 = x - mx, v1 = y - my, v2 = z - mz;
        double dot_trial = v0 * nx + v1 * ny + v2 * nz;
        if (dot_trial >= dot && dot_trial > 0) {
          *i = x; *j = y; *k = z;
          dot = dot_trial;
        }
      }
}

static inline int ibm_cell(short t) {
  return t == ORC_TYPE_F || t == ORC_TYPE_IN || t == ORC_TYPE_OUT;
}

static void ibm_interpolate(orc_lbm *lb, int nm, const double *mx, const double *my,
                            const double *mz, double *ub, int xmin, int xmax, int ymin,
                            int ymax, int zmin, int zmax) {
#pragma omp parallel for
  for (int n = 0; n < nm; n++) {
    ub[3 * n + 0] = 0; ub[3 * n + 1] = 0; ub[3 * n + 2] = 0;
    for (int i = xmin; i <= xmax; i++)
      for (int j = ymin; j <= ymax; j++)
        for (int k = zmin; k <= zmax; k++) {
          const orc_cell *c = &lb->fluid[i][j][k];
          if (!ibm_cell(c->type)) continue;
          double rx = (i * ORC_DX - mx[n]) / ORC_DX, ry = (j * ORC_DY - my[n]) / ORC_DY,
                 rz = (k * ORC_DZ - mz[n]) / ORC_DZ;
          ub[3 * n + 0] += c->ux * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
          ub[3 * n + 1] += c->uy * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
          ub[3 * n + 2] += c->uz * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
        }
  }
}

/* MacroProp's MORE_DETAILS block, lbm.cpp:681-705: pressure = rho/3 and the
 * viscous stress of every cell from its f and u, the reference's loop and
 * product order. */
void orc_stress(const orc_lbm *lb, double *sigma, double *pressure) {
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz;
  const double tau = lb->tau;
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++) {
        const orc_cell *c = &lb->fluid[i][j][k];
        const int64_t n = LIN(lb, i, j, k);
        double sig[3][3];
        for (int alphaa = 0; alphaa < 3; alphaa++)
          for (int beta = alphaa; beta < 3; beta++) {
            const int *ca = alphaa == 0 ? lb->cx : alphaa == 1 ? lb->cy : lb->cz;
            const int *cb = beta == 0 ? lb->cx : beta == 1 ? lb->cy : lb->cz;
            const double ua = alphaa == 0 ? c->ux : alphaa == 1 ? c->uy : c->uz;
            const double ub = beta == 0 ? c->ux : beta == 1 ? c->uy : c->uz;
            double sum_sigma = 0;
            for (int l = 0; l < lb->q; l++) sum_sigma += c->f[l] * (ca[l] - ua) * (cb[l] - ub);
            sig[alphaa][beta] = -(tau - 0.5) * sum_sigma / tau;
            sig[beta][alphaa] = sig[alphaa][beta];
          }
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) sigma[n * 9 + 3 * a + b] = sig[a][b];
        pressure[n] = c->density / 3.0;
      }
}

/* LBM::Sup_MacroProp(n), lbm.cpp:865-939 (dt = 1, told = (n - Nt) dt), per
 * cell in the reference's statement order: averages first (the new average
 * of u enters the fluctuation products), then the rms values, with the
 * reference's forms kept (rms_pressure weights pressure^2 by told; the rms of
 * the Reynolds stress compares the raw products u_a u_b with the averaged
 * fluctuation products).  Layout per cell: avg_u[3], rms_u[3], avg_p,
 * rms_p, avg_rs[9], rms_rs[9].  Vorticity is never computed by the
 * reference (lbm.cpp:707-762 is commented out), so its statistics stay 0
 * and are not kept. */
void orc_stats_update(orc_lbm *lb, int n_step, int Nt) {
  const int64_t ncell = orc_ncells(lb);
  if (!lb->stats) lb->stats = (double *)calloc((size_t)ncell * 26, sizeof(double));
  const double dt = 1.0, told = (n_step - Nt) * dt;
  for (int64_t n = 0; n < ncell; n++) {
    const int64_t i = n / ((int64_t)(lb->Ny + 2) * (lb->Nz + 2));
    const int64_t j = (n / (lb->Nz + 2)) % (lb->Ny + 2), k = n % (lb->Nz + 2);
    const orc_cell *c = &lb->fluid[i][j][k];
    double *s = lb->stats + n * 26;
    double *avg_velocity = s, *rms_velocity = s + 3, *avg_pressure = s + 6, *rms_pressure = s + 7;
    double(*avg_RS)[3] = (double(*)[3])(s + 8), (*rms_RS)[3] = (double(*)[3])(s + 17);
    const double ux = c->ux, uy = c->uy, uz = c->uz, p = c->density / 3.0;
    avg_velocity[0] = (avg_velocity[0] * told + ux * dt) / (told + dt);
    avg_velocity[1] = (avg_velocity[1] * told + uy * dt) / (told + dt);
    avg_velocity[2] = (avg_velocity[2] * told + uz * dt) / (told + dt);
    *avg_pressure = (*avg_pressure * told + p * 1) / (told + 1);
    const double ux_avg = avg_velocity[0], uy_avg = avg_velocity[1], uz_avg = avg_velocity[2];
    const double d[3] = {ux - ux_avg, uy - uy_avg, uz - uz_avg};
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) avg_RS[a][b] = (avg_RS[a][b] * told + d[a] * d[b] * 1) / (told + 1);
    for (int a = 0; a < 3; a++)
      rms_velocity[a] = sqrt((pow(rms_velocity[a], 2) * told + pow(d[a], 2) * 1) / (told + 1));
    *rms_pressure = sqrt((pow(p, 2) * told + pow(p - *avg_pressure, 2) * 1) / (told + 1));
    const double uu = ux * ux, uv = ux * uy, uw = ux * uz, vv = uy * uy, vw = uy * uz, ww = uz * uz;
    const double prod[3][3] = {{uu, uv, uw}, {uv, vv, vw}, {uw, vw, ww}};
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        rms_RS[a][b] = sqrt((pow(rms_RS[a][b], 2) * told + pow(prod[a][b] - avg_RS[a][b], 2) * 1) / (told + 1));
  }
}

/* the statistics as [cell] arrays (3, 3, 1, 1, 9, 9 values per cell); zeros
 * before the first update */
void orc_get_stats(const orc_lbm *lb, double *avg_u, double *rms_u, double *avg_p, double *rms_p,
                   double *avg_rs, double *rms_rs) {
  const int64_t ncell = orc_ncells(lb);
  for (int64_t n = 0; n < ncell; n++) {
    const double *s = lb->stats ? lb->stats + n * 26 : NULL;
    for (int q = 0; q < 3; q++) {
      if (avg_u) avg_u[n * 3 + q] = s ? s[q] : 0.0;
      if (rms_u) rms_u[n * 3 + q] = s ? s[3 + q] : 0.0;
    }
    if (avg_p) avg_p[n] = s ? s[6] : 0.0;
    if (rms_p) rms_p[n] = s ? s[7] : 0.0;
    for (int q = 0; q < 9; q++) {
      if (avg_rs) avg_rs[n * 9 + q] = s ? s[8 + q] : 0.0;
      if (rms_rs) rms_rs[n * 9 + q] = s ? s[17 + q] : 0.0;
    }
  }
}

/* MEA_CD / MEA_CL, lbm.cpp:3085-3163: over the box's TYPE_F cells and the
 * directions l whose neighbour x - c_l is TYPE_S, sum (f*_l + f_opp(l)) with
 * -c_lx (drag) and +c_ly (lift), i, j, k, l serially; out[2] takes +c_lz.
 * The neighbour index wraps as the streaming's does (the reference indexes
 * x - c_l unwrapped, which stays in the lattice for boxes inside 1..N). */
void orc_mea(const orc_lbm *lb, const int box[6], double out[3]) {
  double Px = 0, Py = 0, Pz = 0;
  for (int i = box[0]; i <= box[1]; ++i)
    for (int j = box[2]; j <= box[3]; ++j)
      for (int k = box[4]; k <= box[5]; ++k) {
        const orc_cell *c = &lb->fluid[i][j][k];
        if (c->type != ORC_TYPE_F) continue;
        for (int l = 0; l < lb->q; ++l) {
          const int i_nb = WRAP(i - lb->cx[l], lb->Nx + 2), j_nb = WRAP(j - lb->cy[l], lb->Ny + 2),
                    k_nb = WRAP(k - lb->cz[l], lb->Nz + 2);
          if (lb->fluid[i_nb][j_nb][k_nb].type != ORC_TYPE_S) continue;
          Px += -(c->f_star[l] + c->f[lb->opp[l]]) * lb->cx[l];
          Py += (c->f_star[l] + c->f[lb->opp[l]]) * lb->cy[l];
          Pz += (c->f_star[l] + c->f[lb->opp[l]]) * lb->cz[l];
        }
      }
  /* pow(dx,3)/dt = 1 */
  out[0] = Px;
  out[1] = Py;
  out[2] = Pz;
}

/* D3Q7 (lbm_oracle.h) under an IBM step, which replaces MacroProp: T and q
 * of every fluid cell from its g, and the buoyancy of T added to the IBM's
 * force and velocity (half of it, as MacroProp's u). */
static void q7_ibm_buoyancy(orc_lbm *lb) {
  orc_q7 *t = lb->q7;
  if (!t) return;
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz;
#pragma omp parallel for
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++) {
        orc_cell *c = &lb->fluid[i][j][k];
        if (c->type != ORC_TYPE_F) continue;
        const int64_t n = LIN(lb, i, j, k);
        double T = 0, qx = 0, qy = 0, qz = 0;
        for (int l = 0; l < 7; l++) {
          T += t->g[n * 7 + l];
          qx += t->g[n * 7 + l] * q7_cx[l];
          qy += t->g[n * 7 + l] * q7_cy[l];
          qz += t->g[n * 7 + l] * q7_cz[l];
        }
        t->T[n] = T;
        t->q[3 * n + 0] = qx;
        t->q[3 * n + 1] = qy;
        t->q[3 * n + 2] = qz;
        if (t->boussinesq) {
          const double b = t->g_beta * (T - t->T_ref);
          c->F[1] += b;
          c->uy += 0.5 * b / c->density;
        }
      }
}

/* IBLBM, implicit branch, lbm.cpp:1135-1167 and 1260-1417. */
int orc_ibm_step(orc_lbm *lb, int nm, double *mx, double *my, double *mz,
                 const double *nx, const double *ny, const double *nz,
                 double *ub, double *Fb, const double *ds,
                 const double vcm[3], const double vcm_before[3],
                 int xmin, int xmax, int ymin, int ymax, int zmin, int zmax,
                 int m_max, double vsolid, double *out) {
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz, q = lb->q;
  double Drag = 0, Lift = 0;
  double *udes = (double *)malloc(sizeof(double) * 3 * (nm > 0 ? nm : 1));
  int err = ORC_OK;

  /* marker update, lbm.cpp:1154-1161 */
  for (int n = 0; n < nm; n++) {
    mx[n] += (vcm[0] + vcm_before[0]) * ORC_DT / 2;
    my[n] += (vcm[1] + vcm_before[1]) * ORC_DT / 2;
    mz[n] += (vcm[2] + vcm_before[2]) * ORC_DT / 2;
    udes[3 * n + 0] = vcm[0];
    udes[3 * n + 1] = vcm[1];
    udes[3 * n + 2] = vcm[2];
  }

  /* unforced moments on F/IN/OUT cells of the whole lattice, F = 0,
   * lbm.cpp:1262-1285 */
#pragma omp parallel for
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++) {
        orc_cell *c = &lb->fluid[i][j][k];
        if (!ibm_cell(c->type)) continue;
        double rho = 0, rho_u[3] = {0, 0, 0};
        for (int l = 0; l < q; l++) {
          rho += c->f[l];
          rho_u[0] += c->f[l] * lb->cx[l];
          rho_u[1] += c->f[l] * lb->cy[l];
          rho_u[2] += c->f[l] * lb->cz[l];
        }
        c->density = rho;
        c->ux = rho_u[0] / rho;
        c->uy = rho_u[1] / rho;
        c->uz = rho_u[2] / rho;
        c->F[0] = 0; c->F[1] = 0; c->F[2] = 0;
      }

  /* unforced interpolation, lbm.cpp:1288-1309 */
  ibm_interpolate(lb, nm, mx, my, mz, ub, xmin, xmax, ymin, ymax, zmin, zmax);

  for (int m = 1; m <= m_max; m++) {
    /* boundary force, serial, lbm.cpp:1314-1329 */
    for (int n = 0; n < nm; n++) {
      int i = -1, j = -1, k = 1;
      orc_ibm_nearest(xmin, xmax, ymin, ymax, zmin, zmax, mx[n], my[n], mz[n], nx[n], ny[n],
                      nz[n], &i, &j, &k);
      if (i < 0 || j < 0) { /* the reference leaves i,j uninitialised here */
        err = -1;
        i = xmin; j = ymin;
      }
      double rho = lb->fluid[i][j][k].density;
      Fb[3 * n + 0] = 2 * rho * (udes[3 * n + 0] - ub[3 * n + 0]) / ORC_DT;
      Fb[3 * n + 1] = 2 * rho * (udes[3 * n + 1] - ub[3 * n + 1]) / ORC_DT;
      Fb[3 * n + 2] = 2 * rho * (udes[3 * n + 2] - ub[3 * n + 2]) / ORC_DT;
      Drag += -Fb[3 * n + 0] * ORC_DX * ds[n];
      Lift += -Fb[3 * n + 1] * ORC_DY * ds[n];
    }

    /* spreading and velocity correction, lbm.cpp:1332-1360 */
#pragma omp parallel for
    for (int i = xmin; i <= xmax; i++)
      for (int j = ymin; j <= ymax; j++)
        for (int k = zmin; k <= zmax; k++) {
          orc_cell *c = &lb->fluid[i][j][k];
          if (!ibm_cell(c->type)) continue;
          double *Fm = &lb->Fm[3 * LIN(lb, i, j, k)];
          Fm[0] = 0; Fm[1] = 0; Fm[2] = 0;
          for (int n = 0; n < nm; n++) {
            double rx = (i * ORC_DX - mx[n]) / ORC_DX, ry = (j * ORC_DY - my[n]) / ORC_DY,
                   rz = (k * ORC_DZ - mz[n]) / ORC_DZ;
            Fm[0] += Fb[3 * n + 0] * ds[n] * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
            Fm[1] += Fb[3 * n + 1] * ds[n] * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
            Fm[2] += Fb[3 * n + 2] * ds[n] * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
          }
          c->F[0] += -Fm[0];
          c->F[1] += -Fm[1];
          c->F[2] += -Fm[2];
          c->ux += ORC_DT * Fm[0] / (2 * c->density);
          c->uy += ORC_DT * Fm[1] / (2 * c->density);
          c->uz += ORC_DT * Fm[2] / (2 * c->density);
        }

    /* re-interpolation, lbm.cpp:1363-1387 */
    ibm_interpolate(lb, nm, mx, my, mz, ub, xmin, xmax, ymin, ymax, zmin, zmax);
  }

  /* Calc_AverageDensity, lbm.cpp:1105-1132: cells of the box (any type)
   * behind every marker's normal */
  double rho_avg = 0;
  int number = 0;
  for (int i = xmin; i <= xmax; i++)
    for (int j = ymin; j <= ymax; j++)
      for (int k = zmin; k <= zmax; k++) {
        int is_dot_negative = 1;
        for (int n = 0; n < nm; n++) {
          double v0 = i * ORC_DX - mx[n], v1 = j * ORC_DY - my[n], v2 = k * ORC_DZ - mz[n];
          double dot_product = v0 * nx[n] + v1 * ny[n] + v2 * nz[n];
          if (dot_product > 0) {
            is_dot_negative = 0;
            break;
          }
        }
        if (is_dot_negative) {
          rho_avg += lb->fluid[i][j][k].density;
          number++;
        }
      }
  rho_avg /= number; /* NaN when no cell qualifies, as in the reference */

  /* added mass, lbm.cpp:1392-1398 */
  Drag += rho_avg * vsolid * (vcm[0] - vcm_before[0]) / ORC_DT;
  Lift += rho_avg * vsolid * (vcm[1] - vcm_before[1]) / ORC_DT;
  out[0] = Drag;
  out[1] = Lift;
  out[2] = rho_avg;
  free(udes);
  q7_ibm_buoyancy(lb);
  return err;
}

/* NearestLatticeCoordinate over [xmin,xmax]x[ymin,ymax]x[zmin,zmax] as the
 * explicit scheme calls it, lbm.cpp:1074-1092: the last point (in x, y, z
 * loop order) with the largest dot product > 0 along the normal; the
 * caller's start values stay when none qualifies. */
static void nearest_local(int xmin, int xmax, int ymin, int ymax, int zmin, int zmax, double mx, double my,
                          double mz, double nx, double ny, double nz, int *i, int *j, int *k) {
  double dot = 0;
  for (int x = xmin; x <= xmax; x++)
    for (int y = ymin; y <= ymax; y++)
      for (int z = zmin; z <= zmax; z++) {
        const double dot_trial = (x - mx) * nx + (y - my) * ny + (z - mz) * nz;
        if (dot_trial >= dot && dot_trial > 0) {
          *i = x; *j = y; *k = z;
          dot = dot_trial;
        }
      }
}

/* IBLBM, explicit branch (Explicit_Diffuse_Interface_Scheme), lbm.cpp:
 * 1135-1167, 1171-1258 and 1390-1398. */
int orc_ibm_step_explicit(orc_lbm *lb, int nm, double *mx, double *my, double *mz,
                          const double *nx, const double *ny, const double *nz,
                          double *ub, double *Fb, const double *ds,
                          const double vcm[3], const double vcm_before[3],
                          int xmin, int xmax, int ymin, int ymax, int zmin, int zmax,
                          double vsolid, double *out) {
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz, q = lb->q;
  double Drag = 0, Lift = 0;
  int err = ORC_OK;
  /* marker update, lbm.cpp:1154-1161 (u_des = V_CM) */
  for (int n = 0; n < nm; n++) {
    mx[n] += (vcm[0] + vcm_before[0]) * ORC_DT / 2;
    my[n] += (vcm[1] + vcm_before[1]) * ORC_DT / 2;
    mz[n] += (vcm[2] + vcm_before[2]) * ORC_DT / 2;
  }
  /* unforced moments of F/IN/OUT cells, F kept (lbm.cpp:1173-1193) */
#pragma omp parallel for
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++) {
        orc_cell *c = &lb->fluid[i][j][k];
        if (!ibm_cell(c->type)) continue;
        double rho = 0, rho_u[3] = {0, 0, 0};
        for (int l = 0; l < q; l++) {
          rho += c->f[l];
          rho_u[0] += c->f[l] * lb->cx[l];
          rho_u[1] += c->f[l] * lb->cy[l];
          rho_u[2] += c->f[l] * lb->cz[l];
        }
        c->density = rho;
        c->ux = rho_u[0] / rho;
        c->uy = rho_u[1] / rho;
        c->uz = rho_u[2] / rho;
      }
  /* unforced interpolation, lbm.cpp:1196-1212 */
  ibm_interpolate(lb, nm, mx, my, mz, ub, xmin, xmax, ymin, ymax, zmin, zmax);
  /* boundary force with rho at the local nearest point, lbm.cpp:1215-1230
   * (z range floor(z)..ceil(z): the reference's 2D call passes 1, 1 for
   * markers at z = 1) */
  for (int n = 0; n < nm; n++) {
    int i = 0, j = 0, k = 1;
    nearest_local((int)floor(mx[n]), (int)ceil(mx[n]), (int)floor(my[n]), (int)ceil(my[n]), (int)floor(mz[n]),
                  (int)ceil(mz[n]), mx[n], my[n], mz[n], nx[n], ny[n], nz[n], &i, &j, &k);
    const double rho = lb->fluid[i][j][k].density;
    Fb[3 * n + 0] = 2 * rho * (vcm[0] - ub[3 * n + 0]) / ORC_DT;
    Fb[3 * n + 1] = 2 * rho * (vcm[1] - ub[3 * n + 1]) / ORC_DT;
    Fb[3 * n + 2] = 2 * rho * (vcm[2] - ub[3 * n + 2]) / ORC_DT;
  }
  /* distribution, forced velocity and the cell-summed drag and lift,
   * lbm.cpp:1233-1256 (serial: the reference's reduction order is the
   * loop order) */
  for (int i = xmin; i <= xmax; i++)
    for (int j = ymin; j <= ymax; j++)
      for (int k = zmin; k <= zmax; k++) {
        orc_cell *c = &lb->fluid[i][j][k];
        if (!ibm_cell(c->type)) continue;
        c->F[0] = 0; c->F[1] = 0; c->F[2] = 0;
        for (int n = 0; n < nm; n++) {
          double rx = (i * ORC_DX - mx[n]) / ORC_DX, ry = (j * ORC_DY - my[n]) / ORC_DY,
                 rz = (k * ORC_DZ - mz[n]) / ORC_DZ;
          c->F[0] += Fb[3 * n + 0] * ds[n] * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
          c->F[1] += Fb[3 * n + 1] * ds[n] * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
          c->F[2] += Fb[3 * n + 2] * ds[n] * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
        }
        c->ux += ORC_DT * c->F[0] / (2 * c->density);
        c->uy += ORC_DT * c->F[1] / (2 * c->density);
        c->uz += ORC_DT * c->F[2] / (2 * c->density);
        Drag += -c->F[0] * pow(ORC_DX, 2);
        Lift += -c->F[1] * pow(ORC_DX, 2);
      }
  /* Calc_AverageDensity and added mass, lbm.cpp:1105-1132, 1392-1398 */
  double rho_avg = 0;
  int number = 0;
  for (int i = xmin; i <= xmax; i++)
    for (int j = ymin; j <= ymax; j++)
      for (int k = zmin; k <= zmax; k++) {
        int is_dot_negative = 1;
        for (int n = 0; n < nm; n++) {
          double v0 = i * ORC_DX - mx[n], v1 = j * ORC_DY - my[n], v2 = k * ORC_DZ - mz[n];
          if (v0 * nx[n] + v1 * ny[n] + v2 * nz[n] > 0) {
            is_dot_negative = 0;
            break;
          }
        }
        if (is_dot_negative) {
          rho_avg += lb->fluid[i][j][k].density;
          number++;
        }
      }
  rho_avg /= number;
  Drag += rho_avg * vsolid * (vcm[0] - vcm_before[0]) / ORC_DT;
  Lift += rho_avg * vsolid * (vcm[1] - vcm_before[1]) / ORC_DT;
  out[0] = Drag;
  out[1] = Lift;
  out[2] = rho_avg;
  q7_ibm_buoyancy(lb);
  return err;
}

/* The IBM step of a coupled filament (lbm_oracle.h): orc_ibm_step's
 * unforced moments, interpolation and m_max force/spread/interpolate passes
 * with the markers where they are, u_des per marker, rho at the box cell
 * holding the marker, no added mass; fsum = -sum of the passes' F_b. */
int orc_fsi_ibm_step(orc_lbm *lb, int nm, const double *mx, const double *my, const double *mz,
                     const double *udes, const double *ds, double *ub, double *Fb, double *fsum,
                     int xmin, int xmax, int ymin, int ymax, int zmin, int zmax, int m_max) {
  const int Nx = lb->Nx, Ny = lb->Ny, Nz = lb->Nz, q = lb->q;
  int err = ORC_OK;
#pragma omp parallel for
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++) {
        orc_cell *c = &lb->fluid[i][j][k];
        if (!ibm_cell(c->type)) continue;
        double rho = 0, rho_u[3] = {0, 0, 0};
        for (int l = 0; l < q; l++) {
          rho += c->f[l];
          rho_u[0] += c->f[l] * lb->cx[l];
          rho_u[1] += c->f[l] * lb->cy[l];
          rho_u[2] += c->f[l] * lb->cz[l];
        }
        c->density = rho;
        c->ux = rho_u[0] / rho;
        c->uy = rho_u[1] / rho;
        c->uz = rho_u[2] / rho;
        c->F[0] = 0; c->F[1] = 0; c->F[2] = 0;
      }
  ibm_interpolate(lb, nm, mx, my, mz, ub, xmin, xmax, ymin, ymax, zmin, zmax);
  for (int n = 0; n < 3 * nm; n++) fsum[n] = 0;
  for (int m = 1; m <= m_max; m++) {
    for (int n = 0; n < nm; n++) {
      const int i = (int)floor(mx[n] + 0.5), j = (int)floor(my[n] + 0.5), k = (int)floor(mz[n] + 0.5);
      double rho = NAN;
      if (i >= xmin && i <= xmax && j >= ymin && j <= ymax && k >= zmin && k <= zmax)
        rho = lb->fluid[i][j][k].density;
      else
        err = -1;
      for (int d = 0; d < 3; d++) {
        Fb[3 * n + d] = 2 * rho * (udes[3 * n + d] - ub[3 * n + d]) / ORC_DT;
        fsum[3 * n + d] = (m == 1) ? -Fb[3 * n + d] : fsum[3 * n + d] - Fb[3 * n + d];
      }
    }
#pragma omp parallel for
    for (int i = xmin; i <= xmax; i++)
      for (int j = ymin; j <= ymax; j++)
        for (int k = zmin; k <= zmax; k++) {
          orc_cell *c = &lb->fluid[i][j][k];
          if (!ibm_cell(c->type)) continue;
          double *Fm = &lb->Fm[3 * LIN(lb, i, j, k)];
          Fm[0] = 0; Fm[1] = 0; Fm[2] = 0;
          for (int n = 0; n < nm; n++) {
            double rx = (i * ORC_DX - mx[n]) / ORC_DX, ry = (j * ORC_DY - my[n]) / ORC_DY,
                   rz = (k * ORC_DZ - mz[n]) / ORC_DZ;
            Fm[0] += Fb[3 * n + 0] * ds[n] * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
            Fm[1] += Fb[3 * n + 1] * ds[n] * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
            Fm[2] += Fb[3 * n + 2] * ds[n] * orc_ibm_kernel(rx) * orc_ibm_kernel(ry) * orc_ibm_kernel(rz);
          }
          c->F[0] += -Fm[0];
          c->F[1] += -Fm[1];
          c->F[2] += -Fm[2];
          c->ux += ORC_DT * Fm[0] / (2 * c->density);
          c->uy += ORC_DT * Fm[1] / (2 * c->density);
          c->uz += ORC_DT * Fm[2] / (2 * c->density);
        }
    ibm_interpolate(lb, nm, mx, my, mz, ub, xmin, xmax, ymin, ymax, zmin, zmax);
  }
  q7_ibm_buoyancy(lb);
  return err;
}

/* ---------------------------------------------------------------------------
 * Flexible filament (specification in lbm_oracle.h).
 * ------------------------------------------------------------------------- */
struct orc_filament {
  int n, substeps;
  double ds, ml, ei, gamma, G[3], X0[3], e0[3];
  double *x, *xp, *T; /* 3n, 3n, n-1 */
};

int orc_fil_substeps(const orc_filament *f) { return f->substeps; }

orc_filament *orc_fil_create(int n, double length, double line_density, double bending,
                             double damping, const double body[3], const double clamp[3],
                             const double tangent[3], int substeps) {
  if (n < 4) return NULL;
  orc_filament *f = (orc_filament *)calloc(1, sizeof(orc_filament));
  f->n = n;
  f->ds = length / (n - 1);
  f->ml = line_density;
  f->ei = bending;
  f->gamma = damping;
  for (int d = 0; d < 3; d++) {
    f->G[d] = body[d];
    f->X0[d] = clamp[d];
    f->e0[d] = tangent[d];
  }
  if (substeps <= 0) {
    /* explicit bending: h <= ds^2 sqrt(m_l/EI)/2 (leapfrog with the largest
     * D4 eigenvalue 16/ds^4), taken with a factor 0.8 */
    double hc = bending > 0 ? f->ds * f->ds * sqrt(line_density / bending) / 2 : 1.0;
    substeps = (int)ceil(1.0 / (0.8 * hc));
    if (substeps < 1) substeps = 1;
  }
  f->substeps = substeps;
  f->x = (double *)calloc(3 * (size_t)n, sizeof(double));
  f->xp = (double *)calloc(3 * (size_t)n, sizeof(double));
  f->T = (double *)calloc((size_t)n - 1, sizeof(double));
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) f->x[3 * i + d] = f->xp[3 * i + d] = clamp[d] + i * f->ds * tangent[d];
  return f;
}

void orc_fil_destroy(orc_filament *f) {
  if (!f) return;
  free(f->x);
  free(f->xp);
  free(f->T);
  free(f);
}

void orc_fil_set(orc_filament *f, const double *x_now, const double *x_prev) {
  memcpy(f->x, x_now, sizeof(double) * 3 * f->n);
  memcpy(f->xp, x_prev ? x_prev : x_now, sizeof(double) * 3 * f->n);
}

void orc_fil_get(const orc_filament *f, double *x_now, double *x_prev, double *tension) {
  if (x_now) memcpy(x_now, f->x, sizeof(double) * 3 * f->n);
  if (x_prev) memcpy(x_prev, f->xp, sizeof(double) * 3 * f->n);
  if (tension) memcpy(tension, f->T, sizeof(double) * (f->n - 1));
}

/* Parallel cyclic reduction of a[i] x[i-1] + b[i] x[i] + c[i] x[i+1] = d[i],
 * i < m, in place (x returned in d); every row of a stride is formed from
 * the previous stride's rows. */
static void fil_pcr(int m, double *a, double *b, double *c, double *d) {
  double *na = (double *)malloc(sizeof(double) * m), *nb = (double *)malloc(sizeof(double) * m);
  double *nc = (double *)malloc(sizeof(double) * m), *nd = (double *)malloc(sizeof(double) * m);
  for (int s = 1; s < m; s <<= 1) {
    for (int i = 0; i < m; i++) {
      double al = 0, ga = 0, a2 = 0, c2 = 0, bb = b[i], dd = d[i];
      if (i - s >= 0) {
        al = -a[i] / b[i - s];
        a2 = al * a[i - s];
        bb = bb + al * c[i - s];
        dd = dd + al * d[i - s];
      }
      if (i + s < m) {
        ga = -c[i] / b[i + s];
        c2 = ga * c[i + s];
        bb = bb + ga * a[i + s];
        dd = dd + ga * d[i + s];
      }
      na[i] = a2;
      nb[i] = bb;
      nc[i] = c2;
      nd[i] = dd;
    }
    memcpy(a, na, sizeof(double) * m);
    memcpy(b, nb, sizeof(double) * m);
    memcpy(c, nc, sizeof(double) * m);
    memcpy(d, nd, sizeof(double) * m);
  }
  for (int i = 0; i < m; i++) d[i] = d[i] / b[i];
  free(na); free(nb); free(nc); free(nd);
}

static void fil_substep(orc_filament *f, const double *F, double h) {
  const int n = f->n, m = n - 1;
  const double ds = f->ds, ds4 = ds * ds * ds * ds, ml = f->ml;
  const double kap = h * h / (ml * ds);
  double *X = f->x, *Xp = f->xp;
  double *a = (double *)malloc(sizeof(double) * 3 * m);
  double *P = (double *)malloc(sizeof(double) * 3 * n);
  double *xg = (double *)malloc(sizeof(double) * 3 * (n + 3)); /* ghosts: index i+1 */
  double *ja = (double *)malloc(sizeof(double) * m), *jb = (double *)malloc(sizeof(double) * m);
  double *jc = (double *)malloc(sizeof(double) * m), *rr = (double *)malloc(sizeof(double) * m);
  double *dv = (double *)malloc(sizeof(double) * 3 * m);
  for (int j = 0; j < m; j++)
    for (int d = 0; d < 3; d++) a[3 * j + d] = (X[3 * (j + 1) + d] - X[3 * j + d]) / ds;
  /* positions with ghost nodes -1, N, N+1 */
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) xg[3 * (i + 1) + d] = X[3 * i + d];
  for (int d = 0; d < 3; d++) {
    xg[d] = X[3 + d] - 2 * ds * f->e0[d];
    xg[3 * (n + 1) + d] = 2 * X[3 * (n - 1) + d] - X[3 * (n - 2) + d];
    xg[3 * (n + 2) + d] = 2 * xg[3 * (n + 1) + d] - 2 * X[3 * (n - 2) + d] + X[3 * (n - 3) + d];
  }
  /* predictor */
  for (int d = 0; d < 3; d++) P[d] = f->X0[d];
  for (int i = 1; i < n; i++)
    for (int d = 0; d < 3; d++) {
      const double d4 = (xg[3 * (i - 1) + d] - 4 * xg[3 * i + d] + 6 * xg[3 * (i + 1) + d] - 4 * xg[3 * (i + 2) + d] +
                         xg[3 * (i + 3) + d]) / ds4;
      const double B = -f->ei * d4;
      const double Fi = F ? F[3 * i + d] : 0.0;
      const double v = (X[3 * i + d] - Xp[3 * i + d]) / h;
      P[3 * i + d] = 2 * X[3 * i + d] - Xp[3 * i + d] + h * h / ml * (B + f->G[d] + Fi - f->gamma * v);
    }
  /* Newton on |X'_{j+1} - X'_j|^2 = ds^2 */
  for (int it = 0; it < 3; it++) {
    for (int j = 0; j < m; j++) {
      /* X'_{j+1} - X'_j with X'_i = P_i + kap c_i (T_i a_i - T_{i-1} a_{i-1}) */
      const double cj1 = (j + 1 == n - 1) ? 2.0 : 1.0, cj = 1.0;
      const double Tj1 = (j + 1 < m) ? f->T[j + 1] : 0.0, Tj = f->T[j], Tjm = (j >= 1) ? f->T[j - 1] : 0.0;
      for (int d = 0; d < 3; d++) {
        const double a1 = (j + 1 < m) ? a[3 * (j + 1) + d] : 0.0, a0 = a[3 * j + d],
                     am = (j >= 1) ? a[3 * (j - 1) + d] : 0.0;
        const double xn1 = P[3 * (j + 1) + d] + kap * cj1 * (Tj1 * a1 - Tj * a0);
        const double xn0 = (j == 0) ? P[d] : P[3 * j + d] + kap * cj * (Tj * a0 - Tjm * am);
        dv[3 * j + d] = xn1 - xn0;
      }
      const double dx = dv[3 * j], dy = dv[3 * j + 1], dz = dv[3 * j + 2];
      rr[j] = -((dx * dx + dy * dy + dz * dz) - ds * ds);
      /* d(d_j)/dT_{j+1} = kap c_{j+1} a_{j+1}; /dT_j = -kap (c_{j+1} + c_j') a_j
       * (c_j' = 0 at the clamp); /dT_{j-1} = kap c_j a_{j-1} */
      const double cjp = (j == 0) ? 0.0 : cj;
      double s1 = 0, s0 = 0, sm = 0;
      for (int d = 0; d < 3; d++) {
        const double a1 = (j + 1 < m) ? a[3 * (j + 1) + d] : 0.0, a0 = a[3 * j + d],
                     am = (j >= 1) ? a[3 * (j - 1) + d] : 0.0;
        s1 += dv[3 * j + d] * a1;
        s0 += dv[3 * j + d] * a0;
        sm += dv[3 * j + d] * am;
      }
      jc[j] = (j + 1 < m) ? 2 * kap * cj1 * s1 : 0.0;
      jb[j] = -2 * kap * (cj1 + cjp) * s0;
      ja[j] = (j >= 1) ? 2 * kap * cj * sm : 0.0;
    }
    fil_pcr(m, ja, jb, jc, rr);
    for (int j = 0; j < m; j++) f->T[j] += rr[j];
  }
  /* corrector */
  for (int d = 0; d < 3; d++) {
    Xp[d] = X[d];
    X[d] = f->X0[d];
  }
  for (int i = 1; i < n; i++) {
    const double ci = (i == n - 1) ? 2.0 : 1.0;
    const double Ti = (i < m) ? f->T[i] : 0.0, Tim = f->T[i - 1];
    for (int d = 0; d < 3; d++) {
      const double ai = (i < m) ? a[3 * i + d] : 0.0, aim = a[3 * (i - 1) + d];
      const double xn = P[3 * i + d] + kap * ci * (Ti * ai - Tim * aim);
      Xp[3 * i + d] = X[3 * i + d];
      X[3 * i + d] = xn;
    }
  }
  free(a); free(P); free(xg); free(ja); free(jb); free(jc); free(rr); free(dv);
}

void orc_fil_step(orc_filament *f, const double *force, double dt, int nsteps) {
  const double h = dt / f->substeps;
  for (int s = 0; s < nsteps; s++)
    for (int k = 0; k < f->substeps; k++) fil_substep(f, force, h);
}
