// IBM.cpp -- the reference's IBM free functions over the C-ABI (IBM.h).
#include "IBM.h"

#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

namespace {

// the literal the reference uses for pi (LBM.h:74, lbm.cpp:1859)
constexpr double kPI = 3.14159265;
constexpr double kDT = 1.0;

IBM_CASE g_case;

void check(int code, const char *what) {
  if (code != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_error_string(code));
}

}  // namespace

void IBM_SetCase(const IBM_CASE &c) { g_case = c; }
IBM_CASE &IBM_Case() { return g_case; }

void MARKER::Normal_Marker(int, double x1, double x2, double y1, double y2, double z1, double z2) {
  const double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
  const double norm = std::sqrt(dx * dx + dy * dy + dz * dz);
  const double t1[3] = {dx / norm, dy / norm, dz / norm};
  // t1 x (0, 0, 1)
  normal[0] = t1[1];
  normal[1] = -t1[0];
  normal[2] = 0.0;
}

double kernel(double r) {
  const double a = std::fabs(r);
  if (a <= 1) return (3 - 2 * a + std::sqrt(1 + 4 * a - 4 * (r * r))) / 8;
  if (a <= 2) return (5 - 2 * a - std::sqrt(-7 + 12 * a - 4 * (r * r))) / 8;
  return 0.0;
}

void Generate_Cylinder_Marker(std::vector<MARKER> &marker) {
  const IBM_CASE &c = g_case;
  const int N = c.number_nodes;
  if ((int)marker.size() != N * c.level_marker)
    throw std::runtime_error("Generate_Cylinder_Marker: marker.size() != number_nodes * level_marker");
  const double radius = c.D / 2.0;
  for (int l = 0; l < c.level_marker; l++) {
    MARKER *ring = marker.data() + (size_t)l * N;
    for (int i = 0; i < N; i++) {
      ring[i].x = radius * std::cos(2 * kPI / N * i) + c.center_x + 0.5;
      ring[i].y = radius * std::sin(2 * kPI / N * i) + c.center_y + 0.5;
      ring[i].z = c.zmin + l;
    }
    for (int n = 0; n < N; n++) {
      const int n2 = (n + 1) % N;
      ring[n].Normal_Marker(n, ring[n].x, ring[n2].x, ring[n].y, ring[n2].y, ring[n].z, ring[n2].z);
    }
  }
}

void IBLBM(LBM &lb, std::vector<MARKER> &marker, std::vector<double> &ds, std::vector<double> &Cl,
           std::vector<double> &Cd, double V_CM[], double t) {
  IBM_CASE &c = g_case;
  const double V_CM_before[3] = {V_CM[0], V_CM[1], V_CM[2]};
  switch (c.motion) {
    case IBM_CASE::FIXED:
      V_CM[0] = 0; V_CM[1] = 0; V_CM[2] = 0;
      break;
    case IBM_CASE::INLINE_OSCILLATING:
      V_CM[0] = -c.u_max * std::cos(2 * kPI * c.frequency * t);
      V_CM[1] = 0; V_CM[2] = 0;
      break;
    case IBM_CASE::TRANSVERSE_OSCILLATION:
      V_CM[0] = 0; V_CM[2] = 0;
      V_CM[1] = -2 * kPI * c.frequency_e * c.amplitude * std::sin(2 * kPI * c.frequency_e * t);
      break;
    case IBM_CASE::MOVING_CONSTANT:
      V_CM[0] = -c.u_max;
      V_CM[1] = 0; V_CM[2] = 0;
      break;
  }
  lbm_ctx *ctx = lb.ctx();
  const size_t n = marker.size();
  if (ds.size() < n) throw std::runtime_error("IBLBM: ds.size() < marker.size()");
  // first call on this lattice: hand the body to the device
  if (lbm_ibm_box(ctx, nullptr, nullptr) == LBM_ERR_STATE) {
    std::vector<double> x(n), y(n), z(n), nx(n), ny(n), nz(n);
    for (size_t i = 0; i < n; i++) {
      x[i] = marker[i].x;
      y[i] = marker[i].y;
      z[i] = marker[i].z;
      nx[i] = marker[i].normal[0];
      ny[i] = marker[i].normal[1];
      nz[i] = marker[i].normal[2];
    }
    lbm_ibm_desc d{};
    d.nmarkers = (int32_t)n;
    d.m_max = c.m_max;
    d.box[4] = c.zmin;
    d.box[5] = c.zmax;
    d.follow = 1;
    d.vsolid = c.Vsolid;
    d.center[0] = c.center_x;
    d.center[1] = c.center_y;
    d.half = c.D;
    d.scheme = c.scheme;
    if (c.scheme == 1 && c.motion == IBM_CASE::FIXED) {
      // a body at rest: the box the reference recomputes every step never
      // moves (lbm.cpp:1165-1167), so it is placed once
      d.follow = 0;
      d.box[0] = (int)std::floor(c.center_x - c.D);
      d.box[1] = (int)std::ceil(c.center_x + c.D);
      d.box[2] = (int)std::floor(c.center_y - c.D);
      d.box[3] = (int)std::ceil(c.center_y + c.D);
    }
    check(lbm_ibm_setup(ctx, &d, x.data(), y.data(), z.data(), nx.data(), ny.data(), nz.data(),
                        ds.data()),
          "lbm_ibm_setup");
  }
  double out[3];
  check(lbm_ibm_apply(ctx, V_CM, V_CM_before, out), "lbm_ibm_apply");
  // the centre moves as in lbm.cpp:1162-1163
  c.center_x += (V_CM[0] + V_CM_before[0]) * kDT / 2;
  c.center_y += (V_CM[1] + V_CM_before[1]) * kDT / 2;
  std::vector<double> x(n), y(n), z(n), ub(3 * n), Fb(3 * n);
  check(lbm_ibm_markers(ctx, x.data(), y.data(), z.data(), ub.data(), Fb.data()), "lbm_ibm_markers");
  for (size_t i = 0; i < n; i++) {
    MARKER &m = marker[i];
    m.x = x[i];
    m.y = y[i];
    m.z = z[i];
    for (int a = 0; a < 3; a++) {
      m.u_boundary[a] = ub[3 * i + a];
      m.F_boundary[a] = Fb[3 * i + a];
      m.u_boundary_des[a] = V_CM[a];
    }
  }
  // lbm.cpp:1401-1416
  const double L = c.coef_length > 0 ? c.coef_length : c.D;
  const double q = 0.5 * c.rho0 * c.u_max * c.u_max * L;
  Cl.push_back(out[1] / q);
  Cd.push_back(out[0] / q);
}

void OutputVTK_Marker(int &time, std::vector<MARKER> &marker) {
  const IBM_CASE &c = g_case;
  const int number_nodes = c.number_nodes;
  std::stringstream output_filename;
  output_filename << "wall_t " << int(time / c.Nt) << ".vtk";
  std::ofstream out(output_filename.str().c_str());
  out << "# vtk DataFile Version 3.0\n";
  out << "wall_state\n";
  out << "ASCII\n";
  out << "DATASET POLYDATA\n";
  out << "POINTS " << number_nodes << " float\n";
  for (int n = 0; n < number_nodes; n++) out << marker[n].x << " " << marker[n].y << " " << marker[n].z << " \n";
  out << "LINES " << number_nodes << " " << 3 * number_nodes << "\n";
  for (int n = 0; n < number_nodes; ++n) out << "2 " << n % number_nodes << " " << (n + 1) % number_nodes << "\n";
  out << "VERTICES 1 " << number_nodes + 1 << "\n";
  out << number_nodes << " ";
  for (int n = 0; n < number_nodes; n++) out << n << " ";
}

void OutputCSV(std::vector<double> &values, const std::string &name, double time_per_sample) {
  std::ofstream ofs(name);
  ofs << "time,error\n";
  for (size_t i = 0; i < values.size(); i++) ofs << (i + 1) * time_per_sample << "," << values[i] << "\n";
}
