// LBM.cpp -- the drop-in LBM class over the C-ABI (see LBM.h).
#include "LBM.h"

#include <cmath>
#include <cstdio>
#include <exception>
#include <iostream>
#include <stdexcept>
#include <string>

struct LBM::Impl {
  lbm_ctx *ctx = nullptr;
  LATTICE ***fluid = nullptr;
  int nx = 0, ny = 0, nz = 0;  // padded extents
  bool thermal = false;
  lbm_thermal_desc td{};
  ~Impl() {
    if (ctx) lbm_destroy(ctx);
    if (fluid) {
      for (int i = 0; i < nx; i++) {
        for (int j = 0; j < ny; j++) delete[] fluid[i][j];
        delete[] fluid[i];
      }
      delete[] fluid;
    }
  }
  size_t n() const { return (size_t)nx * ny * nz; }
};

namespace {

// The reference's three terminate sites print a line and call
// std::terminate() (lbm.cpp:114-117, 772-775, 794-797).
void check(int code, const char *what) {
  if (code == LBM_OK) return;
  if (code == LBM_ERR_FEQ_NEGATIVE) {
    std::cout << "Terminating in Collision function" << std::endl;
    std::terminate();
  }
  if (code == LBM_ERR_DENSITY_RANGE) {
    std::cout << "Terminating in MacroProp function for density" << std::endl;
    std::terminate();
  }
  if (code == LBM_ERR_TEMPERATURE_RANGE) {
    std::cout << "Terminating in MacroProp function for temperature" << std::endl;
    std::terminate();
  }
  throw std::runtime_error(std::string(what) + ": " + lbm_error_string(code));
}

}  // namespace

LBM::LBM(int Nx, int Ny, int Nz, double tau) : LBM(Nx, Ny, Nz, tau, LBM_STORE_F64, 0) {}

LBM::LBM(int Nx_, int Ny_, int Nz_, double tau_, int storage, int device)
    : Nx(Nx_), Ny(Ny_), Nz(Nz_), tau(tau_), impl_(std::make_shared<Impl>()) {
  Impl &m = *impl_;
  m.nx = Nx + 2;
  m.ny = Ny + 2;
  m.nz = Nz + 2;
  // lbm.cpp:17-28: one row of Nz+2 cells per (i,j)
  m.fluid = new LATTICE **[m.nx];
  for (int i = 0; i < m.nx; i++) {
    m.fluid[i] = new LATTICE *[m.ny];
    for (int j = 0; j < m.ny; j++) m.fluid[i][j] = new LATTICE[m.nz];
  }
  fluid = m.fluid;
  lbm_desc d{};
  d.nx = Nx;
  d.ny = Ny;
  d.nz = Nz;
  d.lattice = 19;
  d.storage = storage;
  d.device = device;
  d.tau = tau;
  check(lbm_create(&d, &m.ctx), "lbm_create");
}

lbm_ctx *LBM::ctx() const { return impl_->ctx; }

void LBM::Init() {
  Impl &m = *impl_;
  const size_t n = m.n();
  std::vector<int16_t> type(n);
  std::vector<double> rho(n), ux(n), uy(n), uz(n), Fx(n), Fy(n), Fz(n);
  size_t r = 0;
  for (int i = 0; i < m.nx; i++)
    for (int j = 0; j < m.ny; j++)
      for (int k = 0; k < m.nz; k++, r++) {
        const LATTICE &c = fluid[i][j][k];
        type[r] = c.type;
        rho[r] = c.density;
        ux[r] = c.ux;
        uy[r] = c.uy;
        uz[r] = c.uz;
        Fx[r] = c.F[0];
        Fy[r] = c.F[1];
        Fz[r] = c.F[2];
      }
  check(lbm_upload_cells(m.ctx, type.data(), rho.data(), ux.data(), uy.data(), uz.data(),
                         Fx.data(), Fy.data(), Fz.data()),
        "lbm_upload_cells");
  if (m.thermal) {
    std::vector<double> T(n);
    std::vector<int16_t> te(n);
    r = 0;
    for (int i = 0; i < m.nx; i++)
      for (int j = 0; j < m.ny; j++)
        for (int k = 0; k < m.nz; k++, r++) {
          T[r] = fluid[i][j][k].T;
          te[r] = fluid[i][j][k].type_Energy;
        }
    check(lbm_upload_thermal(m.ctx, &m.td, T.data(), te.data()), "lbm_upload_thermal");
  }
  check(lbm_init(m.ctx), "lbm_init");
  // the reference leaves f = f_star = feq (and g = geq) in every cell
  std::vector<double> f(n * lbm_host::npop), g;
  check(lbm_download_f(m.ctx, f.data()), "lbm_download_f");
  if (m.thermal) {
    g.resize(n * lbm_host::npop);
    check(lbm_download_g(m.ctx, g.data()), "lbm_download_g");
  }
  r = 0;
  for (int i = 0; i < m.nx; i++)
    for (int j = 0; j < m.ny; j++)
      for (int k = 0; k < m.nz; k++, r++)
        for (int l = 0; l < lbm_host::npop; l++) {
          fluid[i][j][k].f[l] = f[r * lbm_host::npop + l];
          fluid[i][j][k].f_star[l] = f[r * lbm_host::npop + l];
          if (m.thermal) fluid[i][j][k].g[l] = g[r * lbm_host::npop + l];
        }
}

void LBM::SetThermal(const lbm_thermal_desc &desc) {
  impl_->thermal = true;
  impl_->td = desc;
}

bool LBM::Thermal() const { return impl_->thermal; }

void LBM::Collision() {
  check(lbm_collide(impl_->ctx), "lbm_collide");
  check(lbm_status(impl_->ctx, nullptr), "Collision");
}

void LBM::Streaming() { check(lbm_stream(impl_->ctx), "lbm_stream"); }

void LBM::MacroProp() {
  check(lbm_macro(impl_->ctx), "lbm_macro");
  check(lbm_status(impl_->ctx, nullptr), "MacroProp");
  SyncMirror(false);
}

void LBM::Step(int n) {
  check(lbm_step(impl_->ctx, n, 1), "lbm_step");
  check(lbm_sync(impl_->ctx), "Step");
  SyncMirror(false);
}

void LBM::SyncMirror(bool populations) {
  Impl &m = *impl_;
  const size_t n = m.n();
  std::vector<double> rho(n), ux(n), uy(n), uz(n);
  check(lbm_download_macro(m.ctx, rho.data(), ux.data(), uy.data(), uz.data()),
        "lbm_download_macro");
  std::vector<double> f, g, Fx, Fy, Fz, T, qx, qy, qz;
  if (populations) {
    f.resize(n * lbm_host::npop);
    Fx.resize(n);
    Fy.resize(n);
    Fz.resize(n);
    check(lbm_download_f(m.ctx, f.data()), "lbm_download_f");
    check(lbm_download_force(m.ctx, Fx.data(), Fy.data(), Fz.data()), "lbm_download_force");
    if (m.thermal) {
      g.resize(n * lbm_host::npop);
      check(lbm_download_g(m.ctx, g.data()), "lbm_download_g");
    }
  }
  if (m.thermal) {
    T.resize(n);
    qx.resize(n);
    qy.resize(n);
    qz.resize(n);
    check(lbm_download_thermal(m.ctx, T.data(), qx.data(), qy.data(), qz.data()), "lbm_download_thermal");
  }
  size_t r = 0;
  for (int i = 0; i < m.nx; i++)
    for (int j = 0; j < m.ny; j++)
      for (int k = 0; k < m.nz; k++, r++) {
        LATTICE &c = fluid[i][j][k];
        c.density = rho[r];
        c.ux = ux[r];
        c.uy = uy[r];
        c.uz = uz[r];
        if (populations) {
          for (int l = 0; l < lbm_host::npop; l++) c.f[l] = f[r * lbm_host::npop + l];
          c.F[0] = Fx[r];
          c.F[1] = Fy[r];
          c.F[2] = Fz[r];
          if (m.thermal)
            for (int l = 0; l < lbm_host::npop; l++) c.g[l] = g[r * lbm_host::npop + l];
        }
        if (m.thermal) {
          c.T = T[r];
          c.q[0] = qx[r];
          c.q[1] = qy[r];
          c.q[2] = qz[r];
        }
      }
}

void LBM::Sup_MacroProp(int n, int Nt) { check(lbm_stats_update(impl_->ctx, n, Nt), "Sup_MacroProp"); }

// ---- per-cell helpers on the mirror (lbm.cpp:628-653, 941-981) ---------------
namespace {
// D3Q19, LBM.h:546-551
const int kCx[lbm_host::npop] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
const int kCy[lbm_host::npop] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
const int kCz[lbm_host::npop] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
const double kW[lbm_host::npop] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                                   1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                   1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
const double kCs2 = 1.0 / 3.0;  // dx^2/(3 dt^2), LBM.h:86

// the bracket of the cubic equilibrium, lbm.cpp:631
double cubic(double c_dot_u, double u_dot_u) {
  return 1 + c_dot_u / kCs2 + std::pow(c_dot_u, 2) / (2 * kCs2 * kCs2) - u_dot_u / (2 * kCs2) +
         std::pow(c_dot_u, 3) / (6 * std::pow(kCs2, 3)) - c_dot_u * u_dot_u / (2 * kCs2 * kCs2);
}
}  // namespace

double LBM::Calc_feq(int i, int j, int k, int l) {
  const LATTICE &c = fluid[i][j][k];
  const double u_dot_u = std::pow(c.ux, 2) + std::pow(c.uy, 2) + std::pow(c.uz, 2);
  const double c_dot_u = c.ux * kCx[l] + c.uy * kCy[l] + c.uz * kCz[l];
  return c.density * kW[l] * cubic(c_dot_u, u_dot_u);
}

double LBM::Calc_geq(int i, int j, int k, int l, short type, std::vector<double> uwall) {
  const LATTICE &c = fluid[i][j][k];
  if (type == TYPE_S) {
    const double Twall = Determine_Wall_Temperature(i, j, k);
    const double u_dot_u = std::pow(uwall[0], 2) + std::pow(uwall[1], 2) + std::pow(uwall[2], 2);
    const double c_dot_u = uwall[0] * kCx[l] + uwall[1] * kCy[l] + uwall[2] * kCz[l];
    return c.density * kW[l] * Twall * cubic(c_dot_u, u_dot_u);
  }
  const double u_dot_u = std::pow(c.ux, 2) + std::pow(c.uy, 2) + std::pow(c.uz, 2);
  const double c_dot_u = c.ux * kCx[l] + c.uy * kCy[l] + c.uz * kCz[l];
  return c.density * kW[l] * c.T * cubic(c_dot_u, u_dot_u);
}

void LBM::Determine_Wall_Velocity(int i, int j, int k, std::vector<double> &uwall) {
  int num = 0;
  for (int l = 0; l < lbm_host::npop; l++) {
    const LATTICE &n = fluid[(i + kCx[l] + Nx + 2) % (Nx + 2)][(j + kCy[l] + Ny + 2) % (Ny + 2)]
                            [(k + kCz[l] + Nz + 2) % (Nz + 2)];
    if (n.type == TYPE_S && (n.ux != 0 || n.uy != 0 || n.uz != 0)) {
      uwall[0] += n.ux;
      uwall[1] += n.uy;
      uwall[2] += n.uz;
      num++;
    }
  }
  if (num != 0) {
    uwall[0] /= num;
    uwall[1] /= num;
    uwall[2] /= num;
  }
}

double LBM::Determine_Wall_Temperature(int i, int j, int k) {
  int num = 0;
  double Twall = 0;
  for (int l = 0; l < lbm_host::npop; l++) {
    const LATTICE &n = fluid[(i + kCx[l] + Nx + 2) % (Nx + 2)][(j + kCy[l] + Ny + 2) % (Ny + 2)]
                            [(k + kCz[l] + Nz + 2) % (Nz + 2)];
    if (n.type == TYPE_S) {
      Twall += n.T;
      num++;
    }
  }
  Twall /= num;  // no solid neighbour: 0/0, as the reference
  return Twall;
}

int LBM::getNx() const { return Nx; }
int LBM::getNy() const { return Ny; }
int LBM::getNz() const { return Nz; }
double LBM::getTAU() const { return tau; }

// ---- OutputVTK -------------------------------------------------------------
namespace {
lbm_vtk_desc g_vtk = {1.0, 0.0, 1.0, 0, 0, 0, 0, 0, {0, 0, 0}};
}

void VTK_SetCase(const lbm_vtk_desc &desc) { g_vtk = desc; }

namespace {
double g_mea_norm = 0.5 * 1.0 * 0.1 * 0.1 * 20.0;  // Cylinder_Basic: rho0 1, u_max 0.1, D 20

double mea_sum(LBM &lb, int xmin, int xmax, int ymin, int ymax, int zmin, int zmax, int comp) {
  const int32_t box[6] = {xmin, xmax, ymin, ymax, zmin, zmax};
  double P[3];
  check(lbm_momentum_exchange(lb.ctx(), box, P), "lbm_momentum_exchange");
  return P[comp];
}
}  // namespace

void MEA_SetCase(double rho0, double u_max, double ref) { g_mea_norm = 0.5 * rho0 * u_max * u_max * ref; }

double MEA_CL(LBM &lb, int xmin, int xmax, int ymin, int ymax, int zmin, int zmax) {
  return mea_sum(lb, xmin, xmax, ymin, ymax, zmin, zmax, 1) / g_mea_norm;
}

double MEA_CD(LBM &lb, int xmin, int xmax, int ymin, int ymax, int zmin, int zmax) {
  return mea_sum(lb, xmin, xmax, ymin, ymax, zmin, zmax, 0) / g_mea_norm;
}

void OutputVTK(int &nout, LBM &lb) {
  lbm_vtk_desc d = g_vtk;
  d.force = lbm_ibm_box(lb.ctx(), nullptr, nullptr) == LBM_OK ? 1 : 0;  // the IBM build
  d.thermal = lb.Thermal() ? 1 : 0;                                      // the ENERGY build
  char filename[128];
  std::snprintf(filename, sizeof filename, "./field%06d.vtr", nout);
  check(lbm_output_vtk(lb.ctx(), &d, filename), "OutputVTK");
}

void OutputAVGVTK(int &nout, LBM &lb) {
  const int Nx = lb.getNx(), Ny = lb.getNy(), Nz = lb.getNz();
  const size_t nx = Nx + 2, ny = Ny + 2, nz = Nz + 2, n = nx * ny * nz;
  std::vector<double> rho(n), ux(n), uy(n), uz(n), sig(9 * n), p(n);
  std::vector<double> au(3 * n, 0.0), ru(3 * n, 0.0), ap(n, 0.0), rp(n, 0.0), ars(9 * n, 0.0), rrs(9 * n, 0.0);
  lbm_ctx *c = lb.ctx();
  check(lbm_download_macro(c, rho.data(), ux.data(), uy.data(), uz.data()), "OutputAVGVTK");
  check(lbm_download_stress(c, sig.data(), p.data()), "OutputAVGVTK");
  const int e = lbm_download_stats(c, au.data(), ru.data(), ap.data(), rp.data(), ars.data(), rrs.data());
  if (e != LBM_ERR_STATE) check(e, "OutputAVGVTK");  // before any Sup_MacroProp: the LATTICE zeros
  const double u_max = g_vtk.u_scale, D = 1.0, dx = 1.0, dy = 1.0, dz = 1.0;
  const int zcomp = Nz;
  char filename[128];
  std::snprintf(filename, sizeof filename, "./avgfield%06d.vtr", nout);
  FILE *fp = std::fopen(filename, "wb");
  if (!fp) throw std::runtime_error(std::string("OutputAVGVTK: cannot open ") + filename);
  const long cells = (long)(Nx + 2) * (Ny + 2) * zcomp;
  unsigned long offset = 0;
  std::fprintf(fp, "<?xml version=\"1.0\"?>\n");
  std::fprintf(fp, "<VTKFile type=\"RectilinearGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n");
  std::fprintf(fp, "  <RectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\">\n", Nx + 2, Ny + 2, zcomp);
  std::fprintf(fp, "  <Piece Extent=\"0 %d 0 %d 0 %d\">\n", Nx + 2, Ny + 2, zcomp);
  std::fprintf(fp, "    <PointData>\n");
  std::fprintf(fp, "    </PointData>\n");
  std::fprintf(fp, "    <CellData>\n");
  struct Block {
    const char *name;
    int comps;
  };
  const Block blocks[] = {{"Density", 1},          {"Velocity", 3},         {"Avg velocity", 3},
                          {"Rms velocity", 3},     {"Pressure", 1},         {"Avg pressure", 1},
                          {"Rms pressure", 1},     {"Viscous Stress", 9},   {"Vorticity", 3},
                          {"Avg vorticity", 3},    {"Rms vorticity", 3},    {"Avg reynolds Stress", 9},
                          {"Rms reynolds Stress", 9}};
  for (const Block &b : blocks) {
    if (b.comps == 1)
      std::fprintf(fp, "      <DataArray type=\"Float32\" Name=\"%s\" format=\"appended\" offset=\"%ld\" />\n", b.name,
                   (long)offset);
    else
      std::fprintf(fp,
                   "      <DataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"appended\" "
                   "offset=\"%ld\" />\n",
                   b.name, b.comps, (long)offset);
    offset += 4 + b.comps * 4 * cells;
  }
  std::fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CellType\" format=\"appended\" offset=\"%ld\" />\n", (long)offset);
  offset += 4 + 1 * 4 * cells;
  std::fprintf(fp, "    </CellData>\n");
  std::fprintf(fp, "    <Coordinates>\n");
  std::fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CoordinateX\" format=\"appended\" offset=\"%ld\" />\n", (long)offset);
  offset += 4 + 1 * 4 * (Nx + 3);
  std::fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CoordinateY\" format=\"appended\" offset=\"%ld\" />\n", (long)offset);
  offset += 4 + 1 * 4 * (Ny + 3);
  std::fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CoordinateZ\" format=\"appended\" offset=\"%ld\" />\n", (long)offset);
  std::fprintf(fp, "    </Coordinates>\n");
  std::fprintf(fp, "  </Piece>\n");
  std::fprintf(fp, "  </RectilinearGrid>\n");
  std::fprintf(fp, "  <AppendedData encoding=\"raw\">");
  std::fprintf(fp, "_");
  // one block: its byte count, then comps values per cell for k = 1..zcomp,
  // j, i (x fastest), each value as the reference rounds it
  auto block = [&](int comps, auto value) {
    const unsigned int array_size = (unsigned int)(comps * 4 * cells);
    std::fwrite(&array_size, sizeof(int), 1, fp);
    for (int k = 1; k <= zcomp; k++)
      for (int j = 0; j < Ny + 2; j++)
        for (int i = 0; i < Nx + 2; i++) {
          const size_t r = ((size_t)i * ny + j) * nz + k;
          for (int q = 0; q < comps; q++) {
            const float val32 = value(r, q);
            std::fwrite(&val32, sizeof(float), 1, fp);
          }
        }
  };
  const double *uvec[3] = {ux.data(), uy.data(), uz.data()};
  block(1, [&](size_t r, int) { return (float)rho[r]; });
  block(3, [&](size_t r, int q) { return (float)((float)uvec[q][r] / u_max); });
  block(3, [&](size_t r, int q) { return (float)((float)au[3 * r + q] / u_max); });
  block(3, [&](size_t r, int q) { return (float)((float)ru[3 * r + q] / u_max); });
  block(1, [&](size_t r, int) { return (float)p[r]; });
  block(1, [&](size_t r, int) { return (float)ap[r]; });
  block(1, [&](size_t r, int) { return (float)rp[r]; });
  block(9, [&](size_t r, int q) { return (float)sig[9 * r + q]; });
  for (int v = 0; v < 3; v++)  // vorticity, its average and rms: never computed by the reference
    block(3, [&](size_t, int) { return (float)((float)0.0 / (u_max / D)); });
  block(9, [&](size_t r, int q) { return (float)((float)ars[9 * r + q] / (std::pow(u_max, 2))); });
  block(9, [&](size_t r, int q) { return (float)((float)rrs[9 * r + q] / (std::pow(u_max, 2))); });
  block(1, [&](size_t r, int) {
    const int i = (int)(r / (ny * nz)), j = (int)((r / nz) % ny), k = (int)(r % nz);
    return (float)(int)lb.fluid[i][j][k].type;
  });
  auto coords = [&](int count, double d) {
    const unsigned int array_size = (unsigned int)(1 * 4 * count);
    std::fwrite(&array_size, sizeof(int), 1, fp);
    for (int i = 0; i < count; i++) {
      const float val32 = (float)(i * d);
      std::fwrite(&val32, sizeof(float), 1, fp);
    }
  };
  coords(Nx + 3, dx);
  coords(Ny + 3, dy);
  coords(zcomp + 1, dz);
  std::fprintf(fp, "  </AppendedData>\n");
  std::fprintf(fp, "</VTKFile>\n");
  std::fclose(fp);
}
