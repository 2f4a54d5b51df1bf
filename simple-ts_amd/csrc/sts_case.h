// Internal C++ view of a SIMPLE-TS case (host setup), shared by the host
// library (sts_host.cpp) and the driver (sts_driver.cpp).
#pragma once
#include <string>
#include <vector>

#include "simple_ts.h"

namespace sts {

// One line of EnteredData.txt: the reference reads the 126 lines positionally
// into globals of these C++ types (S:5241-5561).
enum class Kind { Double, Unsigned, Int, Bool };

struct Entry {
  const char *name;
  Kind kind;
};

extern const Entry kEntries[];
extern const int kNumEntries;

struct Polyhedron {
  std::vector<double> x, y;
  double p1x = 0, p1y = 0, p2x = 1, p2y = 1;
  double u_body = 0, v_body = 0, p_body = 0, T_body = 0;
  bool is_slip = false, is_jump = false, drag = false;
  double F_slip = 1.1466, w_slip = 0.1, F_jump = 2.1904, dTdn = 0;
  // state the point test keeps between calls (Polyhedrons_Objects.h:40-60)
  bool rect_defined = false;
  double x_rect_min = 0, x_rect_max = 0, y_rect_min = 0, y_rect_max = 0;
  double kpi = 0, bpi = 0;
  bool inside(double x, double y);
};

struct Circle {
  double x_c = 0, y_c = 0, r_c = 0;
  double u_body = 0, v_body = 0, p_body = 0, T_body = 0;
  bool is_slip = false, is_jump = false, drag = false;
  double F_slip = 1.1466, w_slip = 0.1, F_jump = 2.1904, dTdn = 0;
  bool inside(double x, double y) const;
};

struct Case {
  std::string dir;
  std::vector<double> val;          // entry values, index = line - 1
  std::vector<std::string> comment; // "//..." text after the value, as read
  // derived (S:2013-2071)
  double rho_gas_nd = 0, cT = 0, cT1 = 0, cT2 = 0, cPch = 0, c_Fr_v = 0, u_given_xb_in_it = 0;
  int Nx = 0, Ny = 0, Iter_T = 0;
  std::vector<double> x_f, y_f, hx, hy, x_v, y_v;
  std::vector<unsigned> vol;
  std::vector<Circle> circles;
  std::vector<Polyhedron> polys;
  std::vector<sts_body> bodies;     // circles first, then polyhedrons
  int Nx_u[2], Ny_u[2], Nx_v[2], Ny_v[2], Nx_o[2], Ny_o[2];

  int idx(const char *name) const;
  double &at(const char *name);
  double get(const char *name) const;
  bool flag(const char *name) const { return get(name) != 0.0; }
  void fill_grid(sts_grid &g) const;
};

// Returns "" on success, else the error message.
std::string load_case(const std::string &dir, Case &c);
std::string write_config(const Case &c, const std::string &path);

}  // namespace sts
