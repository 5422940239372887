// example_cavity.cpp -- the reference's driver pattern (main.cpp:10-67,
// 195-221) on the drop-in LBM class: a setup function writes lb.fluid, then
// Init() and the Collision -> Streaming -> MacroProp loop.
//
// usage: lbm_example_cavity N steps out.bin [fused|details] [f64|f32|f16s|f16c]
// Writes density, ux, uy, uz (each (N+2)^3 doubles in fluid[i][j][k] order).
// "details": the MORE_DETAILS loop (Sup_MacroProp every Nt = 10 steps after
// MacroProp, main.cpp:389-391) and OutputAVGVTK(1) into ./avgfield000001.vtr.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "LBM.h"

// Lid-driven cavity in the pattern of main_setup_Couette (lbm.cpp:1536-1584):
// every ghost face TYPE_S, the j = Ny+1 face moving with ux = u_lid.
static LBM main_setup_Cavity(int N, double u_lid, double Re, int storage) {
  const double nu = u_lid * N / Re;
  const double tau = 0.5 + 3.0 * nu;
  LBM lb(N, N, N, tau, storage);
  for (int i = 0; i <= N + 1; i++)
    for (int j = 0; j <= N + 1; j++)
      for (int k = 0; k <= N + 1; k++) {
        LATTICE &c = lb.fluid[i][j][k];
        if (i == 0 || i == N + 1 || j == 0 || j == N + 1 || k == 0 || k == N + 1) c.type = TYPE_S;
        if (j == N + 1) c.ux = u_lid;
        if (c.type == TYPE_F) {
          c.density = 1.0;
          c.ux = 0.0;
          c.uy = 0.0;
          c.uz = 0.0;
        }
      }
  return lb;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s N steps out.bin [fused] [f64|f32|f16s|f16c]\n", argv[0]);
    return 2;
  }
  const int N = std::atoi(argv[1]);
  const int numstep = std::atoi(argv[2]);
  const bool fused = argc > 4 && std::strcmp(argv[4], "fused") == 0;
  const bool details = argc > 4 && std::strcmp(argv[4], "details") == 0;
  int storage = LBM_STORE_F64;
  if (argc > 5 && std::strcmp(argv[5], "f32") == 0) storage = LBM_STORE_F32;
  if (argc > 5 && std::strcmp(argv[5], "f16s") == 0) storage = LBM_STORE_F16S;
  if (argc > 5 && std::strcmp(argv[5], "f16c") == 0) storage = LBM_STORE_F16C;

  LBM lb = main_setup_Cavity(N, 0.05, 100.0, storage);
  lb.Init();
  if (fused) {
    lb.Step(numstep);
  } else {
    for (int step = 1; step <= numstep; step++) {
      lb.Collision();
      lb.Streaming();
      lb.MacroProp();
      if (details && step % 10 == 0) lb.Sup_MacroProp(step);
    }
  }
  if (details) {
    lbm_vtk_desc vd{};
    vd.u_scale = 0.05;  // u_max: the lid velocity
    VTK_SetCase(vd);
    int nout = 1;
    OutputAVGVTK(nout, lb);
  }
  FILE *fp = std::fopen(argv[3], "wb");
  if (!fp) return 1;
  const int n = N + 2;
  std::vector<double> buf((size_t)n * n * n);
  for (int q = 0; q < 4; q++) {
    size_t r = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++, r++) {
          const LATTICE &c = lb.fluid[i][j][k];
          buf[r] = q == 0 ? c.density : q == 1 ? c.ux : q == 2 ? c.uy : c.uz;
        }
    std::fwrite(buf.data(), sizeof(double), buf.size(), fp);
  }
  std::fclose(fp);
  std::printf("cavity %d^3, %d steps (%s) done\n", N, numstep, fused ? "fused" : "reference calls");
  return 0;
}
