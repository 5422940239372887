// example_couette_thermal.cpp -- the reference's Couette case with ENERGY
// (main_setup_Couette, lbm.cpp:1536-1584; constants LBM.h:165-220) on the
// drop-in class: setup writes lb.fluid[][][] including T and type_Energy,
// then Init() and Collision -> Streaming -> MacroProp, OutputVTK at the end.
//
// usage: lbm_example_couette_thermal Ny steps out.bin [f64|f32] [vtkdir]
// Writes density, ux, uy, uz, T of the final lattice (fluid[i][j][k] order).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "LBM.h"

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s Ny steps out.bin [f64|f32] [vtkdir]\n", argv[0]);
    return 2;
  }
  const int Nx = 2, Ny = std::atoi(argv[1]), Nz = 2;
  const int numstep = std::atoi(argv[2]);
  const int storage = argc > 4 && std::strcmp(argv[4], "f32") == 0 ? LBM_STORE_F32 : LBM_STORE_F64;

  // LBM.h:165-220 with Re = 10
  const double cs2 = 1.0 / 3.0, tau = 0.9, Re = 10.0;
  const double nu = cs2 * (tau - 0.5);
  const double u_wall = Re * nu / (Ny + 1);
  const double Pr = 0.7, Br = 10.0, Ec = Br / Pr;
  const double alpha = nu / Pr, tau_g = 0.5 + alpha / cs2, k = cs2 * tau_g, cp = k / (1.0 * alpha);
  const double Tb = 0.5, dT = u_wall * u_wall / (cp * Ec), Tt = Tb + dT;

  LBM lb(Nx, Ny, Nz, tau, storage);
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int kk = 0; kk <= Nz + 1; kk++) {
        LATTICE &c = lb.fluid[i][j][kk];
        if (j == 0 || j == Ny + 1) c.type = TYPE_S;
        if (j == Ny + 1) c.ux = u_wall;
        if (j == 0 || j == Ny + 1) {
          c.type_Energy = TYPE_T;
          lb.fluid[i][0][kk].T = Tb;
          lb.fluid[i][Ny + 1][kk].T = Tt;
        }
        if (c.type == TYPE_F) {
          c.density = 1.0;
          c.ux = c.uy = c.uz = 0.0;
          c.T = Tb;
        }
      }
  lbm_thermal_desc td{};
  td.tau_g = tau_g;
  td.nu = nu;
  td.cp = cp;
  td.T0 = NAN;
  lb.SetThermal(td);
  lb.Init();
  for (int step = 1; step <= numstep; step++) {
    lb.Collision();
    lb.Streaming();
    lb.MacroProp();
  }
  if (argc > 5) {
    // Velocity_Ratio and Temperature_Ratio (LBM.h:184-186)
    lbm_vtk_desc vd{};
    vd.u_scale = u_wall;
    vd.velocity_name = 1;
    vd.T_mode = 1;
    vd.T_a = Tb;
    vd.T_b = dT;
    VTK_SetCase(vd);
    const std::string cwd = argv[5];
    if (chdir(cwd.c_str()) != 0) return 1;
    int n = 1;
    OutputVTK(n, lb);
  }
  FILE *fp = std::fopen(argv[3], "wb");
  if (!fp) return 1;
  const int n0 = Nx + 2, n1 = Ny + 2, n2 = Nz + 2;
  std::vector<double> buf((size_t)n0 * n1 * n2);
  for (int q = 0; q < 5; q++) {
    size_t r = 0;
    for (int i = 0; i < n0; i++)
      for (int j = 0; j < n1; j++)
        for (int kk = 0; kk < n2; kk++, r++) {
          const LATTICE &c = lb.fluid[i][j][kk];
          buf[r] = q == 0 ? c.density : q == 1 ? c.ux : q == 2 ? c.uy : q == 3 ? c.uz : c.T;
        }
    std::fwrite(buf.data(), sizeof(double), buf.size(), fp);
  }
  std::fclose(fp);
  std::printf("thermal Couette Ny=%d, %d steps: T mid %.9f\n", Ny, numstep, lb.fluid[1][Ny / 2][1].T);
  return 0;
}
