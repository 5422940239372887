/*
 * lbm_hip.h -- C-ABI of the MI355X (gfx950) lattice Boltzmann step.
 *
 * The reference exposes its step as the C++ class LBM (LBM.h:620-664) plus
 * the free function IBLBM (LBM.h:687).  It has no FFI of its own; this header
 * is the boundary that class calls through.  Every entry point below names
 * the reference member or function whose work it takes over.  Plain C types
 * only: an opaque context, pointers and sizes.
 *
 * Host-array convention: cell (i,j,k) of the padded (Nx+2)(Ny+2)(Nz+2)
 * lattice is element (i*(Ny+2)+j)*(Nz+2)+k (k fastest), the order of the
 * reference's LBM::fluid[i][j][k] (lbm.cpp:17-28).  Populations of a cell are
 * 19 consecutive doubles in the reference's D3Q19 direction order
 * (LBM.h:546-552).  All host arrays are caller-owned and copied synchronously.
 *
 * Return codes: 0 = OK; 1/2/3 = the reference's three std::terminate() sites
 * (negative equilibrium lbm.cpp:114-117, density outside [0,100]
 * lbm.cpp:772-775, temperature outside [0,100] lbm.cpp:794-797); negative =
 * argument, state or HIP errors.  A context is used by one host thread; all
 * of its device work runs in order on the context's own HIP stream.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_ERR_FEQ_NEGATIVE 1
#define LBM_ERR_DENSITY_RANGE 2
#define LBM_ERR_TEMPERATURE_RANGE 3 /* ENERGY: T outside [0, 100]          *
                                     * (lbm.cpp:794-797); D3Q7: |T| > 100  *
                                     * or not finite (T relative)          */
#define LBM_ERR_ARG (-1)
#define LBM_ERR_STATE (-2)
#define LBM_ERR_HIP (-3)
#define LBM_ERR_NOMEM (-4)
#define LBM_ERR_UNSUPPORTED (-5)

/* cell-type codes, LBM.h:505-515 */
#define LBM_TYPE_F 0
#define LBM_TYPE_S 1
#define LBM_TYPE_OUT 2
#define LBM_TYPE_IN 3
#define LBM_TYPE_P 4
#define LBM_TYPE_L 5
#define LBM_TYPE_T 6
#define LBM_TYPE_Q 7

/* population storage precision (collision arithmetic is FP64 for F64 and
 * FP32 otherwise) */
#define LBM_STORE_F64 0  /* f as double                                   */
#define LBM_STORE_F32 1  /* f - w_i as float                              */
#define LBM_STORE_F16S 2 /* (f - w_i) * 2^LBM_F16S_SCALE_LOG2 as IEEE half */
#define LBM_F16S_SCALE_LOG2 12
/* f - w_i as a 16-bit float of this library's layout: sign, 4 exponent bits
 * (bias 17), 11 fraction bits, subnormals; |f - w_i| < 0.5, larger values
 * saturate.  One significant bit more than F16S at the same 2 bytes. */
#define LBM_STORE_F16C 3

/* lbm_desc.bounce_back: how the half-way bounce-back of Streaming
 * (lbm.cpp:268-275) reaches the pull.  AUTO: a face of the padded lattice
 * whose cells are all TYPE_S has its wall links resolved in the pull address
 * (the fluid cell reads its own f*(x, opp m)); every other wall link goes
 * through the solid's slot, which the wall cell writes after its collision.
 * SLOTS: every wall link through the solid's slot.  Both give the same
 * results (bit for bit at resting walls). */
#define LBM_BB_AUTO 0
#define LBM_BB_SLOTS 1

/* lbm_reduce_fp64 quantities */
#define LBM_REDUCE_MASS 0     /* sum of rho over TYPE_F cells           */
#define LBM_REDUCE_MOMENTUM 1 /* out[0..2] = sum of rho*u over TYPE_F   */
#define LBM_REDUCE_KINETIC 2  /* sum of rho*|u|^2/2 over TYPE_F         */

typedef struct lbm_ctx lbm_ctx;

typedef struct lbm_desc {
  int32_t nx, ny, nz; /* interior size, as LBM(Nx,Ny,Nz,tau) (lbm.cpp:17)   */
  int32_t lattice;    /* 19 = D3Q19 (the only lattice of this library)      */
  int32_t storage;    /* LBM_STORE_*                                        */
  int32_t device;     /* HIP device ordinal                                 */
  double tau;         /* relaxation time; omega = dt/tau (lbm.cpp:76)       */
  int32_t bounce_back; /* LBM_BB_*                                          */
  int32_t reserved[7];
} lbm_desc;

/* LBM::LBM(Nx,Ny,Nz,tau), lbm.cpp:17-28: allocates the device lattice. */
int lbm_create(const lbm_desc *desc, lbm_ctx **out);
/* The reference never frees its lattice; this releases everything. */
int lbm_destroy(lbm_ctx *ctx);

/* The setup writes to lb.fluid[i][j][k].{type,density,ux,uy,uz,F}
 * (e.g. lbm.cpp:1833-1852).  Any pointer may be NULL: type defaults to
 * TYPE_F, rho to 1, u and F to 0.  Takes effect at the next lbm_init. */
int lbm_upload_cells(lbm_ctx *ctx, const int16_t *type, const double *rho,
                     const double *ux, const double *uy, const double *uz,
                     const double *Fx, const double *Fy, const double *Fz);

/* LBM::Init, lbm.cpp:31-57: f = f_star = feq(rho,u) in every cell (with
 * ENERGY also g = g_star = Calc_geq, lbm.cpp:636-653). */
int lbm_init(lbm_ctx *ctx);
/* LBM::Collision + ConvLBM, lbm.cpp:60-179 (TYPE_F cells, BGK + Guo). */
int lbm_collide(lbm_ctx *ctx);
/* LBM::Streaming, lbm.cpp:182-625 (pull form, half-way bounce-back with
 * the per-cell wall velocity of lbm.cpp:941-962, frozen links, then the six
 * face loops of the inlet/outlet cells, lbm.cpp:299-625). */
int lbm_stream(lbm_ctx *ctx);
/* LBM::MacroProp, lbm.cpp:657-812 (every cell, inlet/outlet cells too). */
int lbm_macro(lbm_ctx *ctx);
/* main.cpp:212-221 repeated nsteps times (Collision, Streaming, MacroProp)
 * as one fused stream-collide launch per step.  emit_macro != 0 writes the
 * last step's rho,u to the device macro arrays (lbm_download_macro). */
int lbm_step(lbm_ctx *ctx, int32_t nsteps, int32_t emit_macro);
/* Waits for the context's stream; returns the sticky status (see codes). */
int lbm_sync(lbm_ctx *ctx);
int lbm_status(lbm_ctx *ctx, int64_t *first_bad_cell);

/* Reads of lb.fluid[i][j][k].{density,ux,uy,uz} as the reference holds them
 * after MacroProp (lbm.cpp:672-675).  Any pointer may be NULL. */
int lbm_download_macro(lbm_ctx *ctx, double *rho, double *ux, double *uy,
                       double *uz);
/* Reads of lb.fluid[i][j][k].f (post-streaming populations), ncells*19. */
int lbm_download_f(lbm_ctx *ctx, double *f);
/* Reads of lb.fluid[i][j][k].F (the force used by the next collision). */
int lbm_download_force(lbm_ctx *ctx, double *Fx, double *Fy, double *Fz);

/* FP64 deterministic reductions over TYPE_F cells (no reference
 * counterpart; the reference never sums mass). */
int lbm_reduce_fp64(lbm_ctx *ctx, int32_t what, double *out);

/* lbm_reduce_fp64 without allocation, for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same sums, same
 * result bit for bit; the partial sums go to one fixed device array of the
 * library (1.5 MB per device, held until the process ends) instead of the
 * context's buffers.  Lattices of up to 268,435,456 padded cells
 * (LBM_ERR_UNSUPPORTED beyond).  Calls that overlap on one device race on
 * that array. */
int lbm_reduce_fp64_noalloc(lbm_ctx *ctx, int32_t what, double *out);

/* Checkpoint and resume (no counterpart in the reference, whose only resume
 * is the interactive prompt that extends the step count, main.cpp:397-410).
 * A checkpoint holds the state that stepping changes: both population
 * buffers, the macroscopic arrays, the wall tables, the status
 * word, a per-cell force field, the inlet/outlet slots, the thermal arrays,
 * the running statistics and the context's state flags.  lbm_checkpoint_size
 * gives its byte count; lbm_checkpoint_save writes it to caller memory
 * (bytes at least that count); lbm_checkpoint_load restores it into a context
 * built by the same set-up (lbm_create, lbm_upload_cells,
 * lbm_upload_thermal, lbm_init), after which steps give the saved context's
 * results bit for bit.  LBM_ERR_ARG when the checkpoint does not match the
 * context (lattice, storage, thermal model, inlet/outlet cells, walls, force
 * field); LBM_ERR_STATE before lbm_init.  An immersed rigid body that has
 * never moved is covered between lbm_ibm_run/apply calls (its IBM steps
 * depend on the populations only; the restoring context repeats the same
 * lbm_ibm_setup); LBM_ERR_UNSUPPORTED for a body that moved, a pending
 * IBLBM result and a coupled filament. */
int lbm_checkpoint_size(lbm_ctx *ctx, uint64_t *bytes);
int lbm_checkpoint_save(lbm_ctx *ctx, void *buf, uint64_t bytes);
int lbm_checkpoint_load(lbm_ctx *ctx, const void *buf, uint64_t bytes);

/* The streaming tables the device kernels use, for bit-exact comparison with
 * lbm.cpp:186-295: per cell and direction m, src = linear index of the
 * cell x - c_m (wrapped modulo N+2 per axis, lbm.cpp:197-199) and
 * kind = 0 stream, 1 bounce-back, 2 frozen; non-fluid cells give src = -1,
 * kind = 255.  type receives the cell-type mask.  Arrays: ncells*19. */
int lbm_export_links(lbm_ctx *ctx, int16_t *type, int32_t *src,
                     uint8_t *kind);

/* ---- thermal branch (#define ENERGY) ---------------------------------------
 * Same-lattice D3Q19 energy populations g with viscous heating
 * (ConvLBM, lbm.cpp:121-177), Dirichlet (TYPE_T) and Neumann (TYPE_Q)
 * walls (lbm.cpp:276-288, 965-1028), T and q moments (lbm.cpp:777-798)
 * and the optional Boussinesq force of Rayleigh_Benard (lbm.cpp:803-806).
 * The face loops of Streaming move g as they move f (outlet copies, inlet
 * g = rho w T (1 + ...), the k = 0 inflow write into cell (i,j,Nz+1),
 * lbm.cpp:309-313, 340-346, 613-619).  Every storage: FP32 keeps g - w T0,
 * the 16-bit ones (g - w T0) in the units of their f. */
/* lbm_thermal_desc.model */
#define LBM_THERMAL_ENERGY 0 /* the reference's ENERGY branch: D3Q19 g on the *
                              * same lattice, viscous heating                  */
#define LBM_THERMAL_D3Q7 1   /* D3Q7 advection-diffusion of T (no reference   *
                              * counterpart): w0 = 1/4, w = 1/8, cs2 = 1/4,    *
                              * geq = w T (1 + c.u/cs2), alpha = (tau_g-1/2)/4; *
                              * walls: type_Energy TYPE_T Dirichlet at the     *
                              * solid's T (anti-bounce-back), TYPE_Q bounce-   *
                              * back plus j_dot_n per wall face (the heat flux *
                              * into the fluid), otherwise adiabatic           *
                              * (bounce-back); nu, cp unused; Boussinesq adds   *
                              * g_beta (T - T_ref) to F_y at T's time level;   *
                              * every storage (16-bit: g - w T0 in f's units)  */

typedef struct lbm_thermal_desc {
  double tau_g;     /* g relaxation time, 0.5 + alpha/cs2 (LBM.h:155)       */
  double nu;        /* viscosity in the dissipation Phi = 2 nu rho S:S      *
                     * (LBM.h nu; cs2 (tau - 1/2) in the reference's cases) */
  double cp;        /* specific heat, k/(rho0 alpha) (LBM.h:157)             */
  double j_dot_n;   /* Neumann wall flux (LBM.h:161)                         */
  double g_beta;    /* Boussinesq: F_y = g_beta (T - T_ref) after MacroProp  */
  double T_ref;     /* (Tb + Tt)/2 in the reference                          */
  double T0;        /* FP32 and 16-bit storage keep g - w T0 (NaN: the mean  *
                     * set-up T of the TYPE_F cells)                         */
  int32_t boussinesq;
  int32_t model;    /* LBM_THERMAL_ENERGY (0) or LBM_THERMAL_D3Q7            */
  int32_t reserved[2];
} lbm_thermal_desc;

/* Writes of lb.fluid[i][j][k].T and .type_Energy (e.g. lbm.cpp:1551-1556)
 * after lbm_upload_cells and before lbm_init; switches the context to the
 * thermal step.  type_energy may be NULL (TYPE_F everywhere). */
int lbm_upload_thermal(lbm_ctx *ctx, const lbm_thermal_desc *desc, const double *T,
                       const int16_t *type_energy);
/* Reads of lb.fluid[i][j][k].T and .q (as of the last MacroProp). */
int lbm_download_thermal(lbm_ctx *ctx, double *T, double *qx, double *qy, double *qz);
/* Reads of lb.fluid[i][j][k].g, ncells*19 (valid in the same states as
 * lbm_download_f); ncells*7 for LBM_THERMAL_D3Q7. */
int lbm_download_g(lbm_ctx *ctx, double *g);

/* ---- implicit diffuse-interface IBM (IBLBM, lbm.cpp:1135-1417) ---------- */
typedef struct lbm_ibm_desc {
  int32_t nmarkers; /* Lagrangian markers (number_nodes, LBM.h:340)           */
  int32_t m_max;    /* implicit iterations (LBM.h:352)                       */
  int32_t box[6];   /* xmin,xmax,ymin,ymax,zmin,zmax, inclusive lattice      *
                     * indices (with follow: only zmin, zmax are read)       */
  int32_t moving;   /* recompute nearest points, interior cells and the    *
                     * spreading tiles every step (they are recomputed      *
                     * anyway in every step whose vcm moves the markers)    */
  int32_t follow;   /* 1: the box is the reference's [floor(c - half),      *
                     * ceil(c + half)] in x and y around the body centre c, *
                     * which moves with the markers (lbm.cpp:1162-1166)     */
  double vsolid;    /* added-mass volume (Vsolid, LBM.h:355)                 */
  double center[2]; /* body centre (center_x, center_y of LBM.h) at set-up  */
  double half;      /* box half-width (D in lbm.cpp:1165)                   */
  int32_t scheme;   /* 0: Implicit_Diffuse_Interface_Scheme (the reference's *
                     * default, LBM.h:36; m_max passes); 1: the explicit   *
                     * branch (lbm.cpp:1171-1258: one pass, F = +spread,    *
                     * density at the marker's local nearest point, drag   *
                     * and lift summed over the box cells); explicit needs *
                     * an unforced set-up, a box that does not follow, and *
                     * every marker's density point inside the box         *
                     * (LBM_ERR_UNSUPPORTED otherwise: a marker on a       *
                     * lattice point with no neighbour ahead along its     *
                     * normal reads cell (0,0,1), lbm.cpp:1216-1225)       */
} lbm_ibm_desc;

/* Markers of the body (Generate_Cylinder_Marker, lbm.cpp:1857-1877), their
 * normals (MARKER::Normal_Marker, lbm.cpp:1048-1069) and weights ds
 * (main.cpp:123-131).  Requires lbm_init; the box may hold TYPE_F and TYPE_S
 * cells only.  From here on the context runs the IBM step (no MacroProp). */
int lbm_ibm_setup(lbm_ctx *ctx, const lbm_ibm_desc *desc, const double *x,
                  const double *y, const double *z, const double *nx,
                  const double *ny, const double *nz, const double *ds);
/* IBLBM(lb, marker, ds, Cl, Cd, V_CM, t) (lbm.cpp:1135-1417): moves the
 * markers with the trapezoid rule, runs the implicit multi-direct forcing
 * and returns out = {Drag, Lift, interior density}; Drag and Lift include
 * the added-mass term and are not yet divided by 1/2 rho0 u^2 D.  The next
 * lbm_collide uses its result. */
int lbm_ibm_apply(lbm_ctx *ctx, const double vcm[3], const double vcm_before[3],
                  double out[3]);
/* nsteps of main.cpp:199-205 (IBLBM, Collision, Streaming) as fused
 * launches.  vcm: (nsteps+1) x 3 centre-of-mass velocities (before step 1,
 * then after each step) or NULL for a body at rest.  out: nsteps x 3. */
int lbm_ibm_run(lbm_ctx *ctx, int32_t nsteps, const double *vcm, double *out);
/* marker positions (nmarkers each), velocities and forces (3 x nmarkers) */
int lbm_ibm_markers(lbm_ctx *ctx, double *x, double *y, double *z, double *ub,
                    double *Fb);

/* The current box and, with follow, the body centre (center_x, center_y
 * after the last IBLBM). */
int lbm_ibm_box(lbm_ctx *ctx, int32_t box[6], double center[2]);

/* ---- MORE_DETAILS fields (LBM.h:55) ---------------------------------------
 * MacroProp's pressure = rho/3 and viscous stress
 * sigma_ab = -(tau - 1/2)/tau sum_l f_l (c_la - u_a)(c_lb - u_b)
 * (lbm.cpp:681-705) of every cell, from the state the last MacroProp saw
 * (after lbm_step or lbm_macro).  sigma: ncells*9 ([cell][a][b]), pressure:
 * ncells; either may be NULL. */
int lbm_download_stress(lbm_ctx *ctx, double *sigma, double *pressure);
/* LBM::Sup_MacroProp(n) (lbm.cpp:865-939; called every Nt steps after the
 * step's MacroProp, main.cpp:389-391): running averages and rms of u, of
 * the pressure and of the fluctuation products, told = (n - nt) dt, on the
 * device in FP64 from the last MacroProp's rho, u.  The first call
 * allocates the 26 FP64 fields per cell (zeros, the LATTICE defaults). */
int lbm_stats_update(lbm_ctx *ctx, int32_t n, int32_t nt);
/* The statistics in fluid[i][j][k] order: avg_u, rms_u ncells*3; avg_p,
 * rms_p ncells; avg_rs, rms_rs ncells*9 ([cell][a][b]).  NULL skips a field.
 * The reference's vorticity statistics stay zero (its vorticity is never
 * computed, lbm.cpp:707-762) and are not kept.  LBM_ERR_STATE before the
 * first lbm_stats_update. */
int lbm_download_stats(lbm_ctx *ctx, double *avg_u, double *rms_u, double *avg_p, double *rms_p,
                       double *avg_rs, double *rms_rs);

/* ---- momentum exchange on bounce-back bodies ------------------------------
 * MEA_CD and MEA_CL (lbm.cpp:3085-3163, called at main.cpp:246-249 and
 * 369-372 after Collision, Streaming and MacroProp) as one FP64 reduction:
 * over the TYPE_F cells of the inclusive box {x0,x1,y0,y1,z0,z1} (reference
 * indices, within 0..N+1) and each direction l whose neighbour x - c_l is
 * TYPE_S, out[0] = Px = -sum (f*_l + f_opp(l)) c_lx, out[1] = Py =
 * +sum (f*_l + f_opp(l)) c_ly (the reference's two signs), out[2] = Pz =
 * +sum (f*_l + f_opp(l)) c_lz (no reference counterpart).  f* is the last
 * collision's output, f the streamed populations: valid after lbm_step,
 * lbm_stream or an IBM/FSI run (LBM_ERR_STATE after lbm_init alone or after
 * lbm_collide).  The reference divides by 1/2 rho0 u_max^2 D (cylinder,
 * lbm.cpp:3114) or D^2 PI/4 (sphere, lbm.cpp:3118) for C_L and C_D. */
int lbm_momentum_exchange(lbm_ctx *ctx, const int32_t box[6], double out[3]);

/* ---- flexible filament (no reference counterpart: the reference has no
 * flexible-structure solver; this is the Lagrangian structure update of a
 * coupled fluid-structure run)
 * ------------------------------------------------------------------------
 * An inextensible elastic filament in lattice units,
 *   m_l X_tt = d/ds (T dX/ds) - EI d4X/ds4 + G - gamma X_t + F,  |dX/ds| = 1,
 * on nnodes nodes (ds = length/(nnodes - 1)), node 0 clamped at `clamp`
 * along `tangent`, the last node free.  FP64 on the device: explicit
 * bending, segment tensions from the exact length constraint (Newton with a
 * tridiagonal Jacobian).  Created straight along `tangent`, at rest. */
typedef struct lbm_filament lbm_filament;
typedef struct lbm_filament_desc {
  int32_t nnodes;       /* 4 .. 256                                          */
  int32_t substeps;     /* structure steps per time step; 0: chosen below   *
                         * the explicit bending limit ds^2 sqrt(m_l/EI)/2   */
  double length;        /* L                                                */
  double line_density;  /* m_l, mass per unit length                        */
  double bending;       /* EI                                               */
  double damping;       /* gamma, force per unit length and velocity        */
  double body_force[3]; /* G per unit length (weight minus buoyancy)        */
  double clamp[3];      /* X_0                                              */
  double tangent[3];    /* unit slope at the clamp                          */
} lbm_filament_desc;
int lbm_filament_create(int32_t device, const lbm_filament_desc *desc, lbm_filament **out);
int lbm_filament_destroy(lbm_filament *f);
/* the structure steps one time step takes */
int lbm_filament_substeps(const lbm_filament *f);
/* node positions now and one time step before (3 x nnodes, node-major;
 * x_prev NULL: at rest), and the nnodes-1 segment tensions */
int lbm_filament_set(lbm_filament *f, const double *x_now, const double *x_prev);
int lbm_filament_get(lbm_filament *f, double *x_now, double *x_prev, double *tension);
/* nsteps time steps of length dt under the force F per unit length
 * (3 x nnodes, held over the steps; NULL: none) */
int lbm_filament_step(lbm_filament *f, const double *force, double dt, int32_t nsteps);

/* A filament coupled to the lattice through the implicit IBM (no reference
 * counterpart): its nodes become the markers (weights ds, half at the ends),
 * each with the desired velocity of its displacement over the last step and
 * the density of the lattice cell it sits in (a thin body has fluid on both
 * sides, so the reference's farthest-point-along-the-normal density does not
 * apply).  desc gives the box (holding the filament's reach) and m_max;
 * nmarkers, moving, follow and vsolid are set here.  Requires lbm_init; the
 * filament lives on the lattice's device and must outlive the coupling. */
int lbm_fsi_setup(lbm_ctx *ctx, lbm_filament *f, const lbm_ibm_desc *desc);
/* nsteps coupled steps (IBM with the filament's markers, Collision,
 * Streaming, then the filament step under -sum F_b per unit length, the
 * force its markers exert on the fluid taken back).  out: nsteps x 6 = the
 * total force on the filament and its free end's position, per step. */
int lbm_fsi_run(lbm_ctx *ctx, int32_t nsteps, double *out);

/* ---- macroscopic output (OutputVTK, lbm.cpp:2290-2468) ---------------------
 * The reference's .vtr file: a VTK RectilinearGrid with cells i = 0..Nx+1,
 * j = 0..Ny+1, k = 1..Nz and appended raw Float32 blocks Density, Velocity
 * ((float)u / u_scale), [Force density], [Temperature, Heat flux],
 * CellType (type as Float32), [EnergyCellType], then the coordinates.  The
 * blocks are packed on the device from the state of the last MacroProp and
 * copied to pinned host memory on a second stream, so stepping can go on
 * while the copy runs; lbm_output_vtk_end waits for the copy and writes the
 * file.  One output can be in flight per context. */
typedef struct lbm_vtk_desc {
  double u_scale;     /* velocity = (float)u / u_scale: u_max (the default  *
                       * build), u_wall/u_max (Velocity_Ratio) or 1.0       */
  double T_a, T_b;    /* T_mode 1: (float)(T - T_a) / T_b; 2: (float)T/T_a - 1 */
  int32_t velocity_name; /* 0 "Velocity", 1 "Velocity Ratio",                *
                          * 2 "Velocity (Dimensional)"                      */
  int32_t T_mode;        /* 0 (float)T "Temperature"; 1, 2 "Temperature Ratio" */
  int32_t force;         /* write "Force density" (the IBM build)            */
  int32_t thermal;       /* write Temperature, Heat flux, EnergyCellType     */
  int32_t poiseuille_coords; /* y, z coordinates (float)(j dy)/H*2-1, H=Ny+2  */
  int32_t reserved[3];
} lbm_vtk_desc;
int lbm_output_vtk_begin(lbm_ctx *ctx, const lbm_vtk_desc *desc);
int lbm_output_vtk_end(lbm_ctx *ctx, const char *path);
/* begin + end */
int lbm_output_vtk(lbm_ctx *ctx, const lbm_vtk_desc *desc, const char *path);

/* ---- ensembles and launch-shape tuning (no reference counterpart: the
 * reference runs one lattice per process) ------------------------------------
 * A batch steps many independent lattices with one launch per step: the
 * workgroup tiles of all lattices form one grid (optionally a persistent grid
 * of a few workgroups per CU, each taking a contiguous run of tiles), and the
 * launches of a step pair are captured once into a HIP graph and replayed.
 * Contexts of one batch share device, storage and force mode, have no IBM or
 * thermal branch and no cell that is neither fluid nor solid (inlet/outlet
 * cells need the face loops between steps: LBM_ERR_UNSUPPORTED), and have
 * been initialised (lbm_init). */
typedef struct lbm_batch lbm_batch;
int lbm_batch_create(lbm_ctx *const *ctxs, int32_t n, lbm_batch **out);
int lbm_batch_destroy(lbm_batch *batch);
/* nsteps reference steps (Collision, Streaming, MacroProp) of every lattice;
 * each lattice ends in the state lbm_step(ctx, nsteps, 1) leaves. */
int lbm_batch_step(lbm_batch *batch, int32_t nsteps);
/* nsteps batched fused launches timed with HIP events (after bringing every
 * lattice past its first steps); advances the lattices by nsteps + 1 or more. */
int lbm_batch_time_steps(lbm_batch *batch, int32_t nsteps, double *ms);
/* threads per workgroup (= cells per tile: 64, 128, 192 or 256) and
 * workgroups per CU of a persistent grid (0: one workgroup per tile) */
int lbm_batch_set_launch(lbm_batch *batch, int32_t block, int32_t per_cu);
int lbm_batch_get_launch(const lbm_batch *batch, int32_t *block, int32_t *per_cu);
int64_t lbm_batch_nfluid(const lbm_batch *batch);
/* Times every launch shape (block 64/128/256 x per_cu 0/1/2/4/8) over nsteps
 * batched launches and keeps the fastest.  The timed launches are dry runs:
 * each lattice's fused kernel reads its current populations and writes a
 * scratch buffer (allocated for the call and freed), so the lattices'
 * populations, macroscopic fields, status and step count are unchanged. */
int lbm_batch_autotune(lbm_batch *batch, int32_t nsteps, int32_t *block, int32_t *per_cu,
                       double *best_ms);
/* Per-lattice tuner: times nsteps dry-run launches of the fused kernel (as
 * lbm_batch_autotune: the lattice's state is unchanged) for every launch
 * shape block 64/128/256 x cells per thread 1/2/4, keeps the fastest (also
 * for every later lattice of the same size and storage in the process) and
 * sets it on the context.  Plain D3Q19 contexts (LBM_ERR_UNSUPPORTED with a
 * thermal branch, LBM_ERR_STATE under IBM).  Either pointer may be NULL. */
int lbm_autotune(lbm_ctx *ctx, int32_t nsteps, int32_t *block, int32_t *cells_per_thread);

/* sizes and identification */
int64_t lbm_ncells(const lbm_ctx *ctx);
int64_t lbm_nfluid(const lbm_ctx *ctx);
int32_t lbm_storage(const lbm_ctx *ctx);
/* Timed run of nsteps fused launches bracketed by HIP events on the
 * context's stream; *ms receives the elapsed milliseconds. */
int lbm_time_steps(lbm_ctx *ctx, int32_t nsteps, double *ms);
/* Launch shape of the fused stream kernel: threads per block (64..256, a
 * multiple of 64) and cells per thread (1, 2 or 4: each thread updates that
 * many cells, one per grid-stride pass over the lattice).  The other
 * kernels use the block size with one cell per thread.  No reference
 * counterpart: the launch-shape knob. */
int lbm_set_launch(lbm_ctx *ctx, int32_t block, int32_t cells_per_thread);
int lbm_get_launch(const lbm_ctx *ctx, int32_t *block, int32_t *cells_per_thread);
/* After lbm_init: the faces whose wall links the pull resolves in its
 * address (bit 0: i = 0, 1: i = Nx+1, 2: j = 0, 3: j = Ny+1, 4: k = 0,
 * 5: k = Nz+1; see LBM_BB_AUTO), in bits 8-13 the faces of inlet/outlet
 * cells (all cells neither fluid nor solid, bar the solids of such a face)
 * whose frozen-link values the fused launch rewrites in place, and whether
 * some wall link still goes through a solid's slot (1) or none does (0).
 * No reference counterpart: it reports how Streaming's bounce-back and
 * frozen links (lbm.cpp:252-292) are laid out. */
int lbm_bounce_back_faces(const lbm_ctx *ctx, int32_t *faces, int32_t *slot_links);
/* Bytes one fused launch moves per fluid cell (algorithmic). */
double lbm_bytes_per_cell(const lbm_ctx *ctx);
const char *lbm_error_string(int code);
/* A plain 16-byte-per-lane device copy of `bytes` bytes repeated `reps`
 * times; *gbs = (read + written bytes) / time.  No reference counterpart: the
 * practical HBM ceiling and the reference point for hardware byte counters. */
int lbm_copy_probe(int32_t device, int64_t bytes, int32_t reps, double *gbs);
/* What a population storage keeps of n values: out[i] = in[i] encoded into
 * storage format `storage` (LBM_STORE_*) by the device and decoded back,
 * with in[i] the stored quantity in lattice units (f - w_i for the shifted
 * formats, the F16S scale applied and removed inside).  Host arrays.  No
 * reference counterpart: it states the storage rounding exactly. */
int lbm_storage_roundtrip(int32_t device, int32_t storage, const double *in, double *out, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
