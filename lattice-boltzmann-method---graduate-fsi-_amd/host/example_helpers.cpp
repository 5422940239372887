// example_helpers.cpp -- the reference's public per-cell helpers of class LBM
// (Calc_feq, Calc_geq, Determine_Wall_Velocity, Determine_Wall_Temperature;
// lbm.cpp:628-653, 941-981) on the drop-in's mirror after a few steps.
//
// usage: lbm_example_helpers cavity N steps out.bin
//        lbm_example_helpers couette Ny steps out.bin
// cavity: a lid-driven cavity (all faces solid, the j = N+1 face moving);
//   writes Calc_feq(i,j,k,l) for every cell and direction, then
//   Determine_Wall_Velocity(i,j,k) (3 values, zero where not fluid).
// couette: the thermal Couette set-up of example_couette_thermal; writes
//   Calc_geq(i,j,k,l, TYPE_F, 0) for every cell and direction, then
//   Calc_geq(i,j,k,l, TYPE_S, uwall) and Determine_Wall_Temperature(i,j,k)
//   for the fluid cells with a solid neighbour (zero elsewhere).
// All doubles, fluid[i][j][k] order.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "LBM.h"

static bool has_solid_neighbour(LBM &lb, int i, int j, int k) {
  static const int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
  static const int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
  static const int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
  const int nx = lb.getNx() + 2, ny = lb.getNy() + 2, nz = lb.getNz() + 2;
  for (int l = 0; l < 19; l++)
    if (lb.fluid[(i + cx[l] + nx) % nx][(j + cy[l] + ny) % ny][(k + cz[l] + nz) % nz].type == TYPE_S) return true;
  return false;
}

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s cavity|couette N steps out.bin\n", argv[0]);
    return 2;
  }
  const bool thermal = std::strcmp(argv[1], "couette") == 0;
  const int N = std::atoi(argv[2]), steps = std::atoi(argv[3]);
  const int Nx = thermal ? 2 : N, Ny = N, Nz = thermal ? 2 : N;
  const double cs2 = 1.0 / 3.0;
  double tau = 0.9;
  std::vector<double> out;
  if (!thermal) {
    // lid-driven cavity: u_lid 0.05, Re 100 (cases.cavity)
    const double u_lid = 0.05, nu = u_lid * N / 100.0;
    tau = 0.5 + nu / cs2;
    LBM lb(Nx, Ny, Nz, tau);
    for (int i = 0; i <= Nx + 1; i++)
      for (int j = 0; j <= Ny + 1; j++)
        for (int k = 0; k <= Nz + 1; k++) {
          LATTICE &c = lb.fluid[i][j][k];
          if (i == 0 || i == Nx + 1 || j == 0 || j == Ny + 1 || k == 0 || k == Nz + 1) c.type = TYPE_S;
          if (j == Ny + 1) c.ux = u_lid;
        }
    lb.Init();
    for (int s = 0; s < steps; s++) {
      lb.Collision();
      lb.Streaming();
      lb.MacroProp();
    }
    for (int i = 0; i <= Nx + 1; i++)
      for (int j = 0; j <= Ny + 1; j++)
        for (int k = 0; k <= Nz + 1; k++)
          for (int l = 0; l < 19; l++) out.push_back(lb.Calc_feq(i, j, k, l));
    for (int i = 0; i <= Nx + 1; i++)
      for (int j = 0; j <= Ny + 1; j++)
        for (int k = 0; k <= Nz + 1; k++) {
          std::vector<double> uw(3, 0.0);
          if (lb.fluid[i][j][k].type == TYPE_F) lb.Determine_Wall_Velocity(i, j, k, uw);
          out.insert(out.end(), uw.begin(), uw.end());
        }
  } else {
    // the thermal Couette case of example_couette_thermal (LBM.h:165-220, Re 10)
    const double Re = 10.0, nu = cs2 * (tau - 0.5), u_wall = Re * nu / (Ny + 1);
    const double Pr = 0.7, Br = 10.0, Ec = Br / Pr;
    const double alpha = nu / Pr, tau_g = 0.5 + alpha / cs2, kc = cs2 * tau_g, cp = kc / alpha;
    const double Tb = 0.5, Tt = Tb + u_wall * u_wall / (cp * Ec);
    LBM lb(Nx, Ny, Nz, tau);
    for (int i = 0; i <= Nx + 1; i++)
      for (int j = 0; j <= Ny + 1; j++)
        for (int k = 0; k <= Nz + 1; k++) {
          LATTICE &c = lb.fluid[i][j][k];
          if (j == 0 || j == Ny + 1) {
            c.type = TYPE_S;
            c.type_Energy = TYPE_T;
            c.T = j == 0 ? Tb : Tt;
          } else {
            c.T = Tb;
          }
          if (j == Ny + 1) c.ux = u_wall;
        }
    lbm_thermal_desc td{};
    td.tau_g = tau_g;
    td.nu = nu;
    td.cp = cp;
    td.T0 = NAN;
    lb.SetThermal(td);
    lb.Init();
    for (int s = 0; s < steps; s++) {
      lb.Collision();
      lb.Streaming();
      lb.MacroProp();
    }
    const std::vector<double> zero(3, 0.0);
    for (int i = 0; i <= Nx + 1; i++)
      for (int j = 0; j <= Ny + 1; j++)
        for (int k = 0; k <= Nz + 1; k++)
          for (int l = 0; l < 19; l++) out.push_back(lb.Calc_geq(i, j, k, l, TYPE_F, zero));
    for (int i = 0; i <= Nx + 1; i++)
      for (int j = 0; j <= Ny + 1; j++)
        for (int k = 0; k <= Nz + 1; k++) {
          const bool wall = lb.fluid[i][j][k].type == TYPE_F && has_solid_neighbour(lb, i, j, k);
          std::vector<double> uw(3, 0.0);
          if (wall) lb.Determine_Wall_Velocity(i, j, k, uw);
          for (int l = 0; l < 19; l++) out.push_back(wall ? lb.Calc_geq(i, j, k, l, TYPE_S, uw) : 0.0);
          out.push_back(wall ? lb.Determine_Wall_Temperature(i, j, k) : 0.0);
        }
  }
  FILE *fp = std::fopen(argv[4], "wb");
  if (!fp) return 1;
  std::fwrite(out.data(), sizeof(double), out.size(), fp);
  std::fclose(fp);
  return 0;
}
