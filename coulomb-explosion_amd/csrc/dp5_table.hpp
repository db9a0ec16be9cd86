// Dormand-Prince 5(4) coefficients for the engine's per-lane integrator.
//
// The tableau, error weights and controller constants are the method's
// published values, as the reference prints them (src/integrators.cpp:203-226;
// include/coulomb/batched_integrator.hpp:196-211); they have to be the same
// numbers for the engine to be the same method.  The tables of the
// Runge-Kutta-Nystrom arrangement the kernel uses (dx/dt = v, so stage positions
// need only the acceleration stages) follow from them and are written below as
// exact rationals, each rounded once to double:
//   c[s]     = sum_l a[s][l]                       (node of stage s)
//   aa[s][l] = sum_{l<j<s} a[s][j] a[j][l]         (A*A)
//   b[l]     = a[6][l]                             (5th-order weights)
//   ea[l]    = sum_{s>l} e[s] a[s][l]              (E*A)
// so that, with k_l the stage accelerations,
//   x_s = x + h (c_s v + h sum_l aa[s][l] k_l),   v_new = v + h sum_l b_l k_l,
//   err_x = h^2 sum_l ea_l k_l,                     err_v = h sum_s e_s k_s.
#pragma once

namespace ce {

struct DP5Tables {
  double a[7][7];
  double e[7];
  double c[7];
  double aa[7][7];
  double b[7];
  double ea[7];
};

constexpr DP5Tables make_dp5_tables() {
  DP5Tables t{};
  t.a[1][0] = 1.0 / 5;
  t.a[2][0] = 3.0 / 40;
  t.a[2][1] = 9.0 / 40;
  t.a[3][0] = 44.0 / 45;
  t.a[3][1] = -56.0 / 15;
  t.a[3][2] = 32.0 / 9;
  t.a[4][0] = 19372.0 / 6561;
  t.a[4][1] = -25360.0 / 2187;
  t.a[4][2] = 64448.0 / 6561;
  t.a[4][3] = -212.0 / 729;
  t.a[5][0] = 9017.0 / 3168;
  t.a[5][1] = -355.0 / 33;
  t.a[5][2] = 46732.0 / 5247;
  t.a[5][3] = 49.0 / 176;
  t.a[5][4] = -5103.0 / 18656;
  t.a[6][0] = 35.0 / 384;
  t.a[6][1] = 0.0;
  t.a[6][2] = 500.0 / 1113;
  t.a[6][3] = 125.0 / 192;
  t.a[6][4] = -2187.0 / 6784;
  t.a[6][5] = 11.0 / 84;
  t.e[0] = 71.0 / 57600;
  t.e[1] = 0.0;
  t.e[2] = -71.0 / 16695;
  t.e[3] = 71.0 / 1920;
  t.e[4] = -17253.0 / 339200;
  t.e[5] = 22.0 / 525;
  t.e[6] = -1.0 / 40;
  // Nodes as exact rationals (0, 1/5, 3/10, 4/5, 8/9, 1, 1).
  t.c[0] = 0.0;
  t.c[1] = 1.0 / 5;
  t.c[2] = 3.0 / 10;
  t.c[3] = 4.0 / 5;
  t.c[4] = 8.0 / 9;
  t.c[5] = 1.0;
  t.c[6] = 1.0;
  // A*A and E*A evaluated exactly in rational arithmetic from the tableau above
  // (a double-precision sum leaves rounding residue, e.g. -5.6e-17 for the exact
  // zero (E*A)[1]).
  t.aa[2][0] = 9.0 / 200;
  t.aa[3][0] = -12.0 / 25;
  t.aa[3][1] = 4.0 / 5;
  t.aa[4][0] = -12248.0 / 6561;
  t.aa[4][1] = 7208.0 / 2187;
  t.aa[4][2] = -6784.0 / 6561;
  t.aa[5][0] = -533.0 / 264;
  t.aa[5][1] = 91.0 / 22;
  t.aa[5][2] = -56.0 / 33;
  t.aa[5][3] = 7.0 / 88;
  t.aa[6][0] = 35.0 / 384;
  t.aa[6][2] = 50.0 / 159;
  t.aa[6][3] = 25.0 / 192;
  t.aa[6][4] = -243.0 / 6784;
  t.ea[0] = 611.0 / 230400;
  t.ea[2] = -514.0 / 83475;
  t.ea[3] = 391.0 / 38400;
  t.ea[4] = -4617.0 / 1356800;
  t.ea[5] = -11.0 / 3360;
  for (int l = 0; l < 7; ++l) t.b[l] = t.a[6][l];
  return t;
}

inline constexpr DP5Tables kDP5 = make_dp5_tables();

// Controller constants (scipy RK45 / src/integrators.cpp:220-226).
inline constexpr double kSafety = 0.9;
inline constexpr double kMinFactor = 0.2;
inline constexpr double kMaxFactor = 10.0;
inline constexpr double kExponent = -0.2;

}  // namespace ce
