// LBM.cpp -- the drop-in LBM class over the C-ABI (see LBM.h).
#include "LBM.h"

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
