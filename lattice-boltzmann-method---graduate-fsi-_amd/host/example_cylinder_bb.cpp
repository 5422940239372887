// example_cylinder_bb.cpp -- the reference's bounce-back cylinder driver
// (Cylinder_Basic: Generate_Cylinder and main_setup_Cylinder, lbm.cpp:1429-1487;
// the step loop and the momentum-exchange coefficients, main.cpp:195-249)
// on the drop-in classes.
//
// usage: lbm_example_cylinder_bb Nx Ny D steps Nt out.bin [f64|f32|f16s|f16c]
// Every Nt steps (after Collision, Streaming, MacroProp) records
// MEA_CD and MEA_CL over the reference's box; writes Cd[steps/Nt],
// Cl[steps/Nt], then density, ux, uy, uz of the final lattice
// ((Nx+2)(Ny+2)(Nz+2) doubles each, fluid[i][j][k] order), all doubles.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "LBM.h"

struct CylinderCase {
  int Nx, Ny, Nz;
  double D, u_max, center_x, center_y;
};

// Generate_Cylinder, lbm.cpp:1429-1450: cell centres within D/2 of the axis
static void Generate_Cylinder(LBM &lb, const CylinderCase &cs, int &count) {
  const double radius = cs.D / 2.0, dx = 1.0, dy = 1.0;
  for (int i = 1; i <= cs.Nx; i++)
    for (int j = 1; j <= cs.Ny; j++)
      for (int k = 0; k <= cs.Nz; k++) {
        const double rx = dx * i - dx * 0.5, ry = dy * j - dy * 0.5;
        if (std::sqrt(std::pow(cs.center_x - rx, 2.0) + std::pow(cs.center_y - ry, 2.0)) <= radius) {
          lb.fluid[i][j][k].type = TYPE_S;
          count++;
        }
      }
}

// main_setup_Cylinder, lbm.cpp:1453-1487
static LBM main_setup_Cylinder(const CylinderCase &cs, double tau, int storage) {
  LBM lb(cs.Nx, cs.Ny, cs.Nz, tau, storage);
  int count = 0;
  Generate_Cylinder(lb, cs, count);
  for (int i = 0; i <= cs.Nx + 1; i++)
    for (int j = 0; j <= cs.Ny + 1; j++)
      for (int k = 0; k <= cs.Nz + 1; k++) {
        LATTICE &c = lb.fluid[i][j][k];
        if (i == 0) c.type = TYPE_IN;                                   // inlet
        if (i == cs.Nx + 1) c.type = TYPE_OUT;                          // outlet
        if (j == 0 || j == cs.Ny + 1) c.type = TYPE_OUT;                // slip faces
        if (cs.Nz != 1 && (k == 0 || k == cs.Nz + 1)) c.type = TYPE_F;  // periodic
        if (c.type == TYPE_F || c.type == TYPE_IN || c.type == TYPE_OUT) {
          c.density = 1.0;
          c.ux = cs.u_max;
          c.uy = 0.0;
          c.uz = 0.0;
        }
      }
  return lb;
}

int main(int argc, char **argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s Nx Ny D steps Nt out.bin [f64|f32|f16s|f16c]\n", argv[0]);
    return 2;
  }
  CylinderCase cs;
  cs.Nx = std::atoi(argv[1]);
  cs.Ny = std::atoi(argv[2]);
  cs.Nz = 1;
  cs.D = std::atof(argv[3]);
  const int numstep = std::atoi(argv[4]), Nt = std::atoi(argv[5]);
  int storage = LBM_STORE_F64;
  if (argc > 7 && std::strcmp(argv[7], "f32") == 0) storage = LBM_STORE_F32;
  if (argc > 7 && std::strcmp(argv[7], "f16s") == 0) storage = LBM_STORE_F16S;
  if (argc > 7 && std::strcmp(argv[7], "f16c") == 0) storage = LBM_STORE_F16C;

  // case constants (the role of LBM.h:90-103)
  const double Re = 40.0, rho0 = 1.0;
  cs.u_max = 0.05;
  cs.center_x = cs.Nx / 5.0;
  cs.center_y = cs.Ny / 2.0;
  const double nu = cs.u_max * cs.D / Re, tau = 0.5 + nu / (1.0 / 3.0);
  MEA_SetCase(rho0, cs.u_max, cs.D);

  LBM lb = main_setup_Cylinder(cs, tau, storage);
  lb.Init();
  std::vector<double> CL, CD;
  for (int step = 1; step <= numstep; step++) {
    lb.Collision();
    lb.Streaming();
    lb.MacroProp();
    if (step % Nt == 0) {
      // main.cpp:246-249 (int arguments: the centre -+ D truncated)
      CL.push_back(MEA_CL(lb, cs.center_x - cs.D, cs.center_x + cs.D, cs.center_y - cs.D, cs.center_y + cs.D, 1, 1));
      CD.push_back(MEA_CD(lb, cs.center_x - cs.D, cs.center_x + cs.D, cs.center_y - cs.D, cs.center_y + cs.D, 1, 1));
    }
  }
  lb.SyncMirror(false);

  FILE *fp = std::fopen(argv[6], "wb");
  if (!fp) return 1;
  std::fwrite(CD.data(), sizeof(double), CD.size(), fp);
  std::fwrite(CL.data(), sizeof(double), CL.size(), fp);
  const int nx = cs.Nx + 2, ny = cs.Ny + 2, nz = cs.Nz + 2;
  std::vector<double> buf((size_t)nx * ny * nz);
  for (int q = 0; q < 4; q++) {
    size_t r = 0;
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++, r++) {
          const LATTICE &c = lb.fluid[i][j][k];
          buf[r] = q == 0 ? c.density : q == 1 ? c.ux : q == 2 ? c.uy : c.uz;
        }
    std::fwrite(buf.data(), sizeof(double), buf.size(), fp);
  }
  std::fclose(fp);
  if (!CD.empty())
    std::printf("cylinder BB %dx%d D=%g, %d steps: Cd %.6f Cl %.6f\n", cs.Nx, cs.Ny, cs.D, numstep, CD.back(),
                CL.back());
  return 0;
}
