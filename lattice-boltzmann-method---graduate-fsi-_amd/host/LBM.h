// LBM.h -- drop-in C++ host class with the reference's interface
// (LBM.h:558-664 of the reference), backed by the gfx950 library through the
// C-ABI of lbm_hip.h.
//
// What stays the same for a caller of the reference:
//   - class LATTICE with public type/density/ux/uy/uz/f/f_star/F members;
//   - class LBM with LBM(Nx,Ny,Nz,tau), Init(), Collision(), Streaming(),
//     MacroProp(), getNx/Ny/Nz/TAU() and the public LATTICE ***fluid indexed
//     fluid[i][j][k] over the padded (Nx+2)(Ny+2)(Nz+2) lattice;
//   - setups write fluid[i][j][k].{type,density,ux,uy,uz,F} before Init();
//   - a negative equilibrium or an out-of-range density prints the
//     reference's message and calls std::terminate() (lbm.cpp:114-117,
//     772-775).
// What is different: `fluid` is a host mirror.  Init() uploads it and fills
// f/f_star; MacroProp() and Step() refresh density/ux/uy/uz; SyncMirror()
// also brings back the populations.  The lattice itself lives in HBM.
#pragma once

#include <memory>
#include <vector>

#include "lbm_hip.h"

// cell-type codes, LBM.h:505-515
#ifndef TYPE_F
#define TYPE_F 0
#define TYPE_S 1
#define TYPE_OUT 2
#define TYPE_IN 3
#define TYPE_P 4
#define TYPE_L 5
#define TYPE_T 6
#define TYPE_Q 7
#endif

namespace lbm_host {
constexpr int npop = 19;  // D3Q19, LBM.h:544
}

// LBM.h:558-597, with the ENERGY members (LBM.h:590-596) always present:
// they are read only when the lattice is thermal (LBM::SetThermal)
class LATTICE {
 public:
  short type = TYPE_F;
  double density = 1.0;
  double ux = 0.0;
  double uy = 0.0;
  double uz = 0.0;
  double f[lbm_host::npop], f_star[lbm_host::npop];
  double F[3] = {0, 0, 0};
  // ENERGY
  double g[lbm_host::npop];
  double T = 0.0;
  double q[3] = {0, 0, 0};
  short type_Energy = TYPE_F;
};

class LBM {
 private:
  int Nx, Ny, Nz = 1;
  double tau;
  struct Impl;
  std::shared_ptr<Impl> impl_;  // copies share the lattice, as the reference's
                                // by-value LBM shares its fluid pointer

 public:
  LATTICE ***fluid;  // host mirror, fluid[i][j][k]

  LBM(int Nx, int Ny, int Nz, double tau);
  // storage: LBM_STORE_F64 (default, the reference's precision),
  // LBM_STORE_F32, LBM_STORE_F16S or LBM_STORE_F16C; device: HIP ordinal
  LBM(int Nx, int Ny, int Nz, double tau, int storage, int device = 0);

  void Init();       // lbm.cpp:31-57
  void Collision();  // lbm.cpp:60-179
  void Streaming();  // lbm.cpp:182-626
  void MacroProp();  // lbm.cpp:657-863

  int getNx() const;
  int getNy() const;
  int getNz() const;
  double getTAU() const;

  // -- additions --------------------------------------------------------
  // The thermal branch (the reference's #define ENERGY and its constants,
  // LBM.h:150-161): call before Init(); Init() then also uploads
  // fluid[][][].T and .type_Energy, and MacroProp()/Step() refresh T and q.
  void SetThermal(const lbm_thermal_desc &desc);
  bool Thermal() const;
  // n reference steps (Collision, Streaming, MacroProp) as fused launches;
  // refreshes density/ux/uy/uz of the mirror.
  void Step(int n);
  // Copies density/ux/uy/uz (and f, F when asked) from the device.
  void SyncMirror(bool populations = false);
  // The reference's public per-cell helpers, evaluated on the mirror (the
  // types, density, u and T of fluid[][][] as the last MacroProp, Step or
  // SyncMirror left them): the cubic equilibria (lbm.cpp:628-653) and the
  // wall velocity and temperature a fluid cell's wall links use
  // (lbm.cpp:941-981).  ConvLBM and Determine_Twall_Neumann, the per-cell
  // bodies of Collision and Streaming, run inside the device kernels.
  double Calc_feq(int i, int j, int k, int l);
  double Calc_geq(int i, int j, int k, int l, short type, std::vector<double> uwall);
  void Determine_Wall_Velocity(int i, int j, int k, std::vector<double> &uwall);
  double Determine_Wall_Temperature(int i, int j, int k);
  // LBM::Sup_MacroProp(n) of the MORE_DETAILS build (lbm.cpp:865-939):
  // running averages and rms on the device after the step's MacroProp; Nt is
  // the output interval (the reference's #define Nt 10, LBM.h:493).
  void Sup_MacroProp(int n, int Nt = 10);
  lbm_ctx *ctx() const;
};

// MEA_CL / MEA_CD (lbm.cpp:3085-3163, LBM.h:790-792): lift and drag
// coefficients from the momentum exchange on the bounce-back links in the
// box, evaluated on the device after Streaming (as the reference calls them
// after the step's MacroProp).  The normalisation 1/2 rho0 u_max^2 ref --
// ref = D for a cylinder, 2D/sqrt(3) for the tandem triangle, D^2 PI/4 for a
// sphere -- comes from MEA_SetCase (the reference's compile-time case
// constants, lbm.cpp:3113-3120).
void MEA_SetCase(double rho0, double u_max, double ref);
double MEA_CL(LBM &lb, int xmin, int xmax, int ymin, int ymax, int zmin, int zmax);
double MEA_CD(LBM &lb, int xmin, int xmax, int ymin, int ymax, int zmin, int zmax);

// OutputVTK(int&, LBM&), lbm.cpp:2290-2468: writes ./field%06d.vtr (nout)
// from the lattice's last MacroProp.  The velocity scale, the block names
// and the optional blocks come from VTK_SetCase (the reference's compile-time
// case flags); "Force density" is written when the lattice runs the IBM and
// the thermal blocks when it is thermal.
void VTK_SetCase(const lbm_vtk_desc &desc);
void OutputVTK(int &nout, LBM &lb);
// OutputAVGVTK(int&, LBM&), lbm.cpp:2645-3030 (the MORE_DETAILS build):
// writes ./avgfield%06d.vtr (nout) with density, velocity, the averages and
// rms of Sup_MacroProp, pressure, viscous stress, the (zero) vorticity
// fields, the Reynolds-stress statistics and the cell types, scaled as the
// reference does (velocities by u_max = the VTK_SetCase velocity scale,
// Reynolds stresses by u_max^2).  The reference leaves this file's name
// unset outside its tandem-triangle case.
void OutputAVGVTK(int &nout, LBM &lb);
