// example_cylinder_ibm.cpp -- the reference's IBM driver pattern
// (main.cpp:19-59, 119-134, 195-211) on the drop-in classes: the
// Cylinder_IBM setup (lbm.cpp:1828-1855) writes lb.fluid, the markers come
// from Generate_Cylinder_Marker, and each step is IBLBM, Collision,
// Streaming.
//
// usage: lbm_example_cylinder_ibm Nx Ny D steps out.bin [f64|f32|f16s|f16c] [fixed|inline|transverse]
//        [implicit|explicit]
// Writes Cd[steps], Cl[steps], then density, ux, uy, uz of the final lattice
// ((Nx+2)(Ny+2)(Nz+2) doubles each, fluid[i][j][k] order), all doubles.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "IBM.h"
#include "LBM.h"

static LBM main_setup_Cylinder_IBM(int Nx, int Ny, int Nz, double tau, double u_max, int storage) {
  LBM lb(Nx, Ny, Nz, tau, storage);
  for (int i = 0; i <= Nx + 1; i++)
    for (int j = 0; j <= Ny + 1; j++)
      for (int k = 0; k <= Nz + 1; k++) {
        LATTICE &c = lb.fluid[i][j][k];
        if (i == 0) c.type = TYPE_IN;        // inlet
        if (i == Nx + 1) c.type = TYPE_OUT;  // outlet
        if (j == 0 || j == Ny + 1) c.type = TYPE_OUT;  // slip faces
        if (Nz != 1 && (k == 0 || k == Nz + 1)) c.type = TYPE_F;  // periodic
        if (c.type == TYPE_F || c.type == TYPE_IN || c.type == TYPE_OUT) {
          c.density = 1.0;
          c.ux = u_max;
          c.uy = 0.0;
          c.uz = 0.0;
        }
      }
  return lb;
}

int main(int argc, char **argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s Nx Ny D steps out.bin [f64|f32|f16s|f16c] [fixed|inline|transverse]\n",
                 argv[0]);
    return 2;
  }
  const int Nx = std::atoi(argv[1]), Ny = std::atoi(argv[2]), Nz = 1;
  const double D = std::atof(argv[3]);
  const int numstep = std::atoi(argv[4]);
  int storage = LBM_STORE_F64;
  if (argc > 6 && std::strcmp(argv[6], "f32") == 0) storage = LBM_STORE_F32;
  if (argc > 6 && std::strcmp(argv[6], "f16s") == 0) storage = LBM_STORE_F16S;
  if (argc > 6 && std::strcmp(argv[6], "f16c") == 0) storage = LBM_STORE_F16C;
  const char *motion = argc > 7 ? argv[7] : "fixed";
  const bool expl = argc > 8 && std::strcmp(argv[8], "explicit") == 0;

  // case constants (the role of LBM.h:318-356)
  const double PI = 3.14159265, Re = 20.0, u_max = 0.04;
  const double nu = u_max * D / Re, tau = 0.5 + 3.0 * nu;
  IBM_CASE cs;
  cs.D = D;
  cs.u_max = u_max;
  cs.center_x = Nx / 5.0;
  cs.center_y = Ny / 2.0;
  cs.number_nodes = (int)std::lround(PI * D);
  cs.m_max = 20;
  cs.Vsolid = PI * D * D * 0.25;
  cs.scheme = expl ? 1 : 0;  // Explicit_ / Implicit_Diffuse_Interface_Scheme (LBM.h:35-36)
  if (std::strcmp(motion, "inline") == 0) {
    cs.motion = IBM_CASE::INLINE_OSCILLATING;
    cs.frequency = 0.01;
    cs.coef_length = D + 1.0;
  } else if (std::strcmp(motion, "transverse") == 0) {
    cs.motion = IBM_CASE::TRANSVERSE_OSCILLATION;
    cs.frequency_e = 0.05;
    cs.amplitude = 0.1;
  }
  IBM_SetCase(cs);

  LBM lb = main_setup_Cylinder_IBM(Nx, Ny, Nz, tau, u_max, storage);
  std::vector<MARKER> marker(cs.number_nodes);
  Generate_Cylinder_Marker(marker);
  lb.Init();

  // main.cpp:123-131
  std::vector<double> ds;
  const int number_nodes = (int)marker.size();
  for (int i = 0; i < number_nodes; i++) {
    const int j = (i + 1) % number_nodes;
    const double deltax = marker[i].x - marker[j].x, deltay = marker[i].y - marker[j].y,
                 deltaz = marker[i].z - marker[j].z;
    ds.push_back(std::sqrt(deltax * deltax + deltay * deltay + deltaz * deltaz));
  }
  std::vector<double> CL, CD;
  double V_CM[3] = {0, 0, 0};
  const double dt = 1.0;
  for (int step = 1; step <= numstep; step++) {
    IBLBM(lb, marker, ds, CL, CD, V_CM, step * dt);
    lb.Collision();
    lb.Streaming();
  }
  lb.SyncMirror(false);
  // main.cpp:236 and 254-257: the marker file and the coefficient series
  int time = numstep;
  OutputVTK_Marker(time, marker);
  OutputCSV(CD, "Cd.csv", dt * u_max / D);
  OutputCSV(CL, "Cl.csv", dt * u_max / D);

  FILE *fp = std::fopen(argv[5], "wb");
  if (!fp) return 1;
  std::fwrite(CD.data(), sizeof(double), CD.size(), fp);
  std::fwrite(CL.data(), sizeof(double), CL.size(), fp);
  const int nx = Nx + 2, ny = Ny + 2, nz = Nz + 2;
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
  std::printf("cylinder IBM %dx%d D=%g, %d steps (%s): Cd %.6f Cl %.6f\n", Nx, Ny, D, numstep, motion,
              CD.back(), CL.back());
  return 0;
}
