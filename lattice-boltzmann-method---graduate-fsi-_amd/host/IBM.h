// IBM.h -- drop-in immersed-boundary interface of the reference (class
// MARKER, LBM.h:600-616; the IBM free functions, LBM.h:670-688;
// Generate_Cylinder_Marker, lbm.cpp:1857-1877) backed by the gfx950
// implicit diffuse-interface IBM of lbm_hip.h.
//
// A reference driver keeps its loop (main.cpp:195-211):
//     IBLBM(lb, marker, ds, CL, CD, V_CM, step * dt);
//     lb.Collision();
//     lb.Streaming();
// The case constants the reference takes from compile-time #defines
// (LBM.h:318-470: D, u_max, center_x/center_y, number_nodes, m_max, Vsolid,
// the motion law) come from an IBM_CASE set once with IBM_SetCase().
//
// What differs: the lattice stays in HBM, so IBLBM does not refresh
// lb.fluid (call lb.SyncMirror() before reading it); the markers' x, y, z,
// u_boundary, u_boundary_des and F_boundary are refreshed every call.
// ds and the marker normals are read at the first IBLBM call on a lattice.
#pragma once

#include <string>
#include <vector>

#include "LBM.h"

// LBM.h:600-616 (implicit scheme)
class MARKER {
 public:
  double x = 0, y = 0, z = 1;              // current position
  double u_boundary[3] = {0, 0, 0};        // interpolated fluid velocity
  double u_boundary_des[3] = {0, 0, 0};    // desired velocity (V_CM)
  double F_boundary[3] = {0, 0, 0};        // Lagrangian force
  double normal[3] = {0, 0, 0};
  double Fm_boundary[3] = {0, 0, 0};

  // lbm.cpp:1048-1069: n = t1 x z_hat with t1 the unit vector from
  // (x1,y1,z1) to (x2,y2,z2)
  void Normal_Marker(int n, double x1, double x2, double y1, double y2, double z1, double z2);
};

// The reference's compile-time IBM case parameters, at run time.
struct IBM_CASE {
  enum Motion {
    FIXED = 0,               // Cylinder_IBM: V_CM = 0 (lbm.cpp:1140-1141)
    INLINE_OSCILLATING,      // V_CM = (-u_max cos(2 pi f t), 0, 0) (lbm.cpp:1142-1144)
    TRANSVERSE_OSCILLATION,  // V_CM = (0, -2 pi f_e A sin(2 pi f_e t), 0) (lbm.cpp:1145-1147)
    MOVING_CONSTANT,         // Cylinder_IBM_Moving_Arbie: V_CM = (-u_max, 0, 0) (lbm.cpp:1148-1150)
  };
  Motion motion = FIXED;
  double D = 20;              // cylinder diameter; the box is centre +- D
  double u_max = 0.1;
  double rho0 = 1.0;          // reference density of the coefficients
  double center_x = 0, center_y = 0;  // body centre, moved by IBLBM
  int number_nodes = 200;     // markers per ring
  int level_marker = 1;       // rings, at z = zmin .. zmin + level_marker - 1
  int m_max = 20;             // implicit iterations
  double Vsolid = 0;          // added-mass volume
  double frequency = 0;       // INLINE_OSCILLATING
  double frequency_e = 0;     // TRANSVERSE_OSCILLATION
  double amplitude = 0;       // TRANSVERSE_OSCILLATION
  double coef_length = 0;     // C_L, C_D length: D (0 selects D), D + dx for inline oscillation
  int zmin = 1, zmax = 1;     // box z range (the reference's 2D slab: 1..1)
  int Nt = 10;                // output interval (steps), names the marker files
  int scheme = 0;             // 0 Implicit_Diffuse_Interface_Scheme, 1 the explicit branch
};

void IBM_SetCase(const IBM_CASE &c);
IBM_CASE &IBM_Case();

// kernel(r), lbm.cpp:1095-1102: the 4-point Peskin function
double kernel(double r);

// lbm.cpp:1857-1877 for every ring: x = r cos(2 pi i/N) + center_x + 0.5,
// y = r sin(2 pi i/N) + center_y + 0.5, counter-clockwise; normals from the
// next marker of the same ring.  marker.size() must be
// number_nodes * level_marker.
void Generate_Cylinder_Marker(std::vector<MARKER> &marker);

// IBLBM, lbm.cpp:1135-1417 (implicit scheme): updates V_CM for time t,
// moves the markers and the box, runs the multi-direct forcing on the
// device and appends C_L and C_D.  The next lb.Collision() uses its result.
void IBLBM(LBM &lb, std::vector<MARKER> &marker, std::vector<double> &ds, std::vector<double> &Cl,
           std::vector<double> &Cd, double V_CM[], double t);

// OutputVTK_Marker, lbm.cpp:3033-3077: "wall_t <time/Nt>.vtk" in the working
// directory, ASCII POLYDATA with the first number_nodes markers as POINTS, a
// closed LINES loop and one VERTICES cell (the positions IBLBM last moved).
void OutputVTK_Marker(int &time, std::vector<MARKER> &marker);

// OutputCSV, lbm.cpp:3169-3186: "time,error" then one row per value, the
// time column (i + 1) * time_per_sample.  The reference chooses that factor by
// case (Nt dt for the analytic cases, dt u_max/D for the IBM cylinder,
// dt frequency for the inline oscillation, Nt dt u_max/D otherwise).
void OutputCSV(std::vector<double> &values, const std::string &name, double time_per_sample);
