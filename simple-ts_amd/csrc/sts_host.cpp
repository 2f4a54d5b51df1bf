// Host setup of a SIMPLE-TS case: EnteredData.txt and body files, constants,
// mesh, body mask and solve boxes.  Restates the reference's setup code
// (SIMPLE_TS_MPI.cpp "S:n", Polyhedrons_Objects.cpp, Circles_Objects.h,
// Objects_for_Solving.h); no GPU involved.  Built with -ffp-contract=off so
// the mesh and mask are the reference's to the last bit.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "simple_ts_host.h"
#include "sts_case.h"

namespace sts {

// The 126 positional entries of EnteredData.txt in file order with the C++
// type of the global each is read into (S:5241-5561; types from the
// declarations S:245-1792).
const Entry kEntries[] = {
    {"u_gas_nd", Kind::Double},
    {"v_gas_nd", Kind::Double},
    {"T_gas_nd", Kind::Double},
    {"p_gas_nd", Kind::Double},
    {"Ma", Kind::Double},
    {"Pr", Kind::Double},
    {"gamma1", Kind::Double},
    {"c_mu", Kind::Double},
    {"Kn", Kind::Double},
    {"Fr", Kind::Double},
    {"SolveGravityField", Kind::Bool},
    {"Re", Kind::Double},
    {"TypeOfNonDimensiolization", Kind::Unsigned},
    {"Nt", Kind::Double},
    {"Nx", Kind::Unsigned},
    {"Ny", Kind::Unsigned},
    {"N_I", Kind::Unsigned},
    {"N_I_p_c", Kind::Unsigned},
    {"N_I_T", Kind::Unsigned},
    {"ht", Kind::Double},
    {"x_b", Kind::Double},
    {"x_e", Kind::Double},
    {"y_b", Kind::Double},
    {"y_e", Kind::Double},
    {"kind_of_mesh", Kind::Unsigned},
    {"hxmin", Kind::Double},
    {"hxmax", Kind::Double},
    {"xfmin", Kind::Double},
    {"xfmax", Kind::Double},
    {"xbmin", Kind::Double},
    {"xbmax", Kind::Double},
    {"xmid", Kind::Double},
    {"hymin", Kind::Double},
    {"hymax", Kind::Double},
    {"ytmin", Kind::Double},
    {"ytmax", Kind::Double},
    {"ybmin", Kind::Double},
    {"ybmax", Kind::Double},
    {"ymid", Kind::Double},
    {"MaxError_Velocities", Kind::Double},
    {"MaxError_p_c", Kind::Double},
    {"MaxError_T", Kind::Double},
    {"MaxError_to_stop_program", Kind::Double},
    {"max_time_for_solve", Kind::Double},
    {"solve_is_finished", Kind::Bool},
    {"Nt_save_solved_data", Kind::Int},
    {"Nt_save_DragCoefficient", Kind::Unsigned},
    {"Nt_save_temp_data", Kind::Int},
    {"N_I_check", Kind::Int},
    {"WriteDataInShortFormat_kind", Kind::Int},
    {"ToReadSolvedDataFromFile", Kind::Bool},
    {"ToReadSolvedDataFromBinaryFile", Kind::Bool},
    {"ToWriteSolvedDataToBinaryFile", Kind::Bool},
    {"ToWriteSolvedDataToNewFiles", Kind::Bool},
    {"ToContinueFromInterpolation", Kind::Bool},
    {"ToImport_data_for_circles_from_file_b", Kind::Bool},
    {"ToImport_data_for_polyhedrons_from_file_b", Kind::Bool},
    {"ToImport_data_for_circles_from_file_p", Kind::Bool},
    {"ToImport_data_for_polyhedrons_from_file_p", Kind::Bool},
    {"ToImport_data_for_circles_from_file_V", Kind::Bool},
    {"ToImport_data_for_polyhedrons_from_file_V", Kind::Bool},
    {"Given_velocity_on_xb", Kind::Bool},
    {"Given_velocity_on_xe", Kind::Bool},
    {"Periodic_boundary_conditions_about_OX", Kind::Bool},
    {"Pressure_BC", Kind::Bool},
    {"ToStartFromExactSolutionForChannelFlowWithPressureBC", Kind::Bool},
    {"p_BC_xb_correction_method", Kind::Unsigned},
    {"Pressure_ratio_correct", Kind::Bool},
    {"GivenReKn", Kind::Bool},
    {"correct_p_BC_xb", Kind::Bool},
    {"u_given_xb", Kind::Double},
    {"u_given_xb_error", Kind::Double},
    {"p_correction_auto", Kind::Bool},
    {"p_correction_min", Kind::Double},
    {"p_correction_max", Kind::Double},
    {"p_correction", Kind::Double},
    {"p_BC_xb", Kind::Double},
    {"T_BC_xb", Kind::Double},
    {"p_BC_xe", Kind::Double},
    {"T_BC_xe", Kind::Double},
    {"dudx_0_BC_xb", Kind::Bool},
    {"durhodx_0_BC_xb", Kind::Bool},
    {"drhodt_durhodx_dvrhody_0_BC_xb", Kind::Bool},
    {"ToCalculate_ui_1j_from_equation_for_uij_at_x_b", Kind::Bool},
    {"dudx_0_BC_xe", Kind::Bool},
    {"durhodx_0_BC_xe", Kind::Bool},
    {"drhodt_durhodx_dvrhody_0_BC_xe", Kind::Bool},
    {"ToCalculate_ui1j_from_equation_for_uij_at_x_e", Kind::Bool},
    {"u_uMin_Mout_0_BC_xe", Kind::Bool},
    {"dudy_0_BC_yb", Kind::Bool},
    {"dudy_0_BC_ye", Kind::Bool},
    {"dvdx_0_BC_xb", Kind::Bool},
    {"ToCalculate_vi_1j_from_equation_for_vij_at_x_b", Kind::Bool},
    {"dvdx_0_BC_xe", Kind::Bool},
    {"ToCalculate_vi1j_from_equation_for_vij_at_x_e", Kind::Bool},
    {"dvdy_0_BC_yb", Kind::Bool},
    {"dvdy_0_BC_ye", Kind::Bool},
    {"dpdx_0_BC_xb", Kind::Bool},
    {"ToCalculate_pi_1j_from_equation_for_pij_at_x_b", Kind::Bool},
    {"dpdx_0_BC_xe", Kind::Bool},
    {"ToCalculate_pi1j_from_equation_for_pij_at_x_e", Kind::Bool},
    {"dpdy_0_BC_yb", Kind::Bool},
    {"dpdy_0_BC_ye", Kind::Bool},
    {"dTdx_0_BC_xb", Kind::Bool},
    {"ToCalculate_Temperi_1j_from_equation_for_Temperij_at_x_b", Kind::Bool},
    {"dTdx_0_BC_xe", Kind::Bool},
    {"ToCalculate_Temperi1j_from_equation_for_Temperij_at_x_e", Kind::Bool},
    {"dTdy_0_BC_yb", Kind::Bool},
    {"dTdy_0_BC_ye", Kind::Bool},
    {"To_Use_Kn_local_in_wall_BC", Kind::Bool},
    {"ToSolveContinuityEquation", Kind::Bool},
    {"use_interpolation_for_u_BC_by_time", Kind::Bool},
    {"t_BC_b", Kind::Double},
    {"t_BC_e", Kind::Double},
    {"u_BC_xb_t_BC_b", Kind::Double},
    {"u_BC_xb_t_BC_e", Kind::Double},
    {"w_u", Kind::Double},
    {"w_v", Kind::Double},
    {"w_p", Kind::Double},
    {"w_Temper", Kind::Double},
    {"N_SubDomains_x", Kind::Unsigned},
    {"N_SubDomains_y", Kind::Unsigned},
    {"ToReadSolvedDataFromSeparateFileForEachSubDomain", Kind::Bool},
    {"ToWriteSolvedDataToFileForEntireComputationalDomain", Kind::Bool},
    {"ToWriteSolvedDataToSeparateFileForEachSubDomain", Kind::Bool},
    {"ToOnlyWriteDataFor_Nx_Ny_ForAllSubDomains_AndExit", Kind::Bool},
};
const int kNumEntries = (int)(sizeof(kEntries) / sizeof(kEntries[0]));

int Case::idx(const char *name) const {
  for (int k = 0; k < kNumEntries; k++)
    if (std::strcmp(kEntries[k].name, name) == 0) return k;
  return -1;
}
double &Case::at(const char *name) {
  int k = idx(name);
  if (k < 0) {
    std::fprintf(stderr, "simple-ts: unknown EnteredData entry %s\n", name);
    std::abort();
  }
  return val[(size_t)k];
}
double Case::get(const char *name) const {
  int k = idx(name);
  return k < 0 ? std::numeric_limits<double>::quiet_NaN() : val[(size_t)k];
}

// ---------------------------------------------------------------------------
// Text input with the reference's ">> value; getline(rest)" discipline.
// ---------------------------------------------------------------------------
namespace {

struct Reader {
  std::istringstream in;
  bool ok = true;
  explicit Reader(const std::string &text) : in(text) {}
  double num() {
    double d = 0;
    if (!(in >> d)) ok = false;
    return d;
  }
  std::string rest() {
    std::string s;
    std::getline(in, s);
    return s;
  }
};

bool read_file(const std::string &path, std::string &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

std::string strip_cr(std::string s) {
  while (!s.empty() && (s.back() == '\r' || s.back() == '\n')) s.pop_back();
  return s;
}

// Parses one entry the way operator>> into the entry's type would; bools
// accept only 0 and 1 (anything else sets failbit in the reference).
bool parse_entry(const std::string &line, Kind kind, double &v, std::string &comment) {
  const char *s = line.c_str();
  char *end = nullptr;
  while (*s == ' ' || *s == '\t') s++;
  if (kind == Kind::Double) {
    v = std::strtod(s, &end);
  } else {
    long long x = std::strtoll(s, &end, 10);
    if (kind == Kind::Bool && x != 0 && x != 1) return false;
    if (kind == Kind::Unsigned && x < 0) return false;
    v = (double)x;
  }
  if (end == s) return false;
  std::string r(end);
  size_t p = r.find_first_not_of(" \t");
  comment = (p == std::string::npos) ? std::string() : strip_cr(r.substr(p));
  return true;
}

// ---------------------------------------------------------------------------
// Mesh step functions, S:4736-4952
// ---------------------------------------------------------------------------
struct MeshFn {
  unsigned kind = 5;
  double hmin = 0, hmax = 0, fmin = 0, fmax = 0, bmin = 0, bmax = 0, mid = 0, hmid = 0;
  double af0 = 0, af1 = 0, af2 = 0, af3 = 0, ab0 = 0, ab1 = 0, ab2 = 0, ab3 = 0;

  double front(double x) const { return af0 + af1 * x + af2 * x * x; }
  double back(double x) const { return ab0 + ab1 * x + ab2 * x * x + ab3 * x * x * x; }
  double front3(double x) const { return af0 + af1 * x + af2 * x * x + af3 * x * x * x; }
  double par(double x) const {
    if (x < fmax) return hmax;
    else if ((fmax <= x) && (x <= mid)) return front(x);
    else if ((mid < x) && (x <= bmax)) return back(x);
    return hmax;
  }
  double par_flat(double x) const {
    if (x < fmax) return hmax;
    else if ((fmax <= x) && (x <= fmin)) return front(x);
    else if ((fmin < x) && (x <= bmin)) return hmin;
    else if ((bmin < x) && (x <= bmax)) return back(x);
    return hmax;
  }
  double linear(double x) const {
    if (x < fmax) return hmax;
    else if ((fmax <= x) && (x <= fmin)) return ((hmin - hmax) / (fmin - fmax)) * (x - fmax) + hmax;
    else if ((fmin < x) && (x <= mid)) return ((hmid - hmin) / (mid - fmin)) * (x - fmin) + hmin;
    else if ((mid < x) && (x <= bmin)) return ((hmin - hmid) / (bmin - mid)) * (x - mid) + hmid;
    else if ((bmin < x) && (x <= bmax)) return ((hmax - hmin) / (bmax - bmin)) * (x - bmin) + hmin;
    return hmax;
  }
  double linear_flat(double x) const {
    if (x < fmax) return hmax;
    else if ((fmax <= x) && (x <= fmin)) return ((hmin - hmax) / (fmin - fmax)) * (x - fmax) + hmax;
    else if ((fmin < x) && (x <= bmin)) return hmin;
    else if ((bmin < x) && (x <= bmax)) return ((hmax - hmin) / (bmax - bmin)) * (x - bmin) + hmin;
    return hmax;
  }
  double poly3(double x) const {
    if (x < fmax) return hmax;
    else if ((fmax <= x) && (x <= fmin)) return front3(x);
    else if ((fmin < x) && (x <= bmin)) return hmin;
    else if ((bmin < x) && (x <= bmax)) return back(x);
    return hmax;
  }
  // hx_all / hy_all: the step for the kind, clamped to [hmin, hmax]
  double operator()(double x) const {
    double h;
    if (kind == 1) h = linear(x);
    else if (kind == 2) h = par(x);
    else if (kind == 3) h = linear_flat(x);
    else if (kind == 4) h = par_flat(x);
    else h = poly3(x);
    if (hmax < h) h = hmax;
    else if (h < hmin) h = hmin;
    return h;
  }
};

// S:2078-2148: coefficients of the Par and Polynom3 meshes.  The Polynom3 y
// coefficients use the x step limits hxmin/hxmax, as the reference does.
void mesh_coefficients(MeshFn &X, MeshFn &Y, double hxmin_, double hxmax_) {
  if (X.kind == 2 || X.kind == 4) {
    X.mid = X.fmin + 0.5 * (X.bmin - X.fmin);
    X.af2 = (X.hmax - X.hmin) / ((X.fmax - X.fmin) * (X.fmax - X.fmin));
    X.af1 = -2.0 * X.af2 * X.fmin;
    X.af0 = X.hmin + X.af2 * X.fmin * X.fmin;
    X.ab2 = (X.hmax - X.hmin) / ((X.bmax - X.bmin) * (X.bmax - X.bmin));
    X.ab1 = -2.0 * X.ab2 * X.bmin;
    X.ab0 = X.hmin + X.ab2 * X.bmin * X.bmin;
    X.hmid = X.par(X.mid);
    Y.mid = Y.fmin + 0.5 * (Y.bmin - Y.fmin);  // ymid = ybmin + 0.5*(ytmin - ybmin)
    Y.af2 = (Y.hmax - Y.hmin) / ((Y.fmax - Y.fmin) * (Y.fmax - Y.fmin));
    Y.af1 = -2.0 * Y.af2 * Y.fmin;
    Y.af0 = Y.hmin + Y.af2 * Y.fmin * Y.fmin;
    Y.ab2 = (Y.hmax - Y.hmin) / ((Y.bmax - Y.bmin) * (Y.bmax - Y.bmin));
    Y.ab1 = -2.0 * Y.ab2 * Y.bmin;
    Y.ab0 = Y.hmin + Y.ab2 * Y.bmin * Y.bmin;
    Y.hmid = Y.par(Y.mid);
  } else if (X.kind == 5) {
    const double xfmin = X.fmin, xfmax = X.fmax, xbmin = X.bmin, xbmax = X.bmax;
    const double hmin = hxmin_, hmax = hxmax_;
    double t = 1.0 / (3.0 * xfmax * pow(xfmin, 2) - pow(xfmin, 3) - 3.0 * pow(xfmax, 2) * xfmin + pow(xfmax, 3));
    X.af3 = 2.0 * (hmin - hmax) * t;
    X.af2 = -3.0 * (xfmax + xfmin) * (hmin - hmax) * t;
    X.af1 = 6.0 * xfmin * xfmax * (hmin - hmax) * t;
    X.af0 = (-3.0 * hmin * pow(xfmax, 2) * xfmin + hmin * pow(xfmax, 3) + 3.0 * hmax * pow(xfmin, 2) * xfmax
             - hmax * pow(xfmin, 3)) * t;
    t = 1.0 / (3.0 * xbmin * pow(xbmax, 2) - pow(xbmax, 3) - 3.0 * pow(xbmin, 2) * xbmax + pow(xbmin, 3));
    X.ab3 = 2.0 * (hmax - hmin) * t;
    X.ab2 = -3.0 * (xbmin + xbmax) * (hmax - hmin) * t;
    X.ab1 = 6.0 * xbmax * xbmin * (hmax - hmin) * t;
    X.ab0 = (-3.0 * hmax * pow(xbmin, 2) * xbmax + hmax * pow(xbmin, 3) + 3.0 * hmin * pow(xbmax, 2) * xbmin
             - hmin * pow(xbmax, 3)) * t;
    // y: bottom plays "front" (ybmin/ybmax), top plays "back" (ytmin/ytmax)
    const double ybmin = Y.fmin, ybmax = Y.fmax, ytmin = Y.bmin, ytmax = Y.bmax;
    t = 1.0 / (3.0 * ybmax * pow(ybmin, 2) - pow(ybmin, 3) - 3.0 * pow(ybmax, 2) * ybmin + pow(ybmax, 3));
    Y.af3 = 2.0 * (hmin - hmax) * t;
    Y.af2 = -3.0 * (ybmax + ybmin) * (hmin - hmax) * t;
    Y.af1 = 6.0 * ybmin * ybmax * (hmin - hmax) * t;
    Y.af0 = (-3.0 * hmin * pow(ybmax, 2) * ybmin + hmin * pow(ybmax, 3) + 3.0 * hmax * pow(ybmin, 2) * ybmax
             - hmax * pow(ybmin, 3)) * t;
    t = 1.0 / (3.0 * ytmin * pow(ytmax, 2) - pow(ytmax, 3) - 3.0 * pow(ytmin, 2) * ytmax + pow(ytmin, 3));
    Y.ab3 = 2.0 * (hmax - hmin) * t;
    Y.ab2 = -3.0 * (ytmin + ytmax) * (hmax - hmin) * t;
    Y.ab1 = 6.0 * ytmax * ytmin * (hmax - hmin) * t;
    Y.ab0 = (-3.0 * hmax * pow(ytmin, 2) * ytmax + hmax * pow(ytmin, 3) + 3.0 * hmin * pow(ytmax, 2) * ytmin
             - hmin * pow(ytmax, 3)) * t;
  }
}

// define_Nx_Ny, S:4955-5041 (the y tolerance uses hxmax, as there).
int adjust_count(int N, double b, double e, double hxmax, double hmax, const MeshFn &F) {
  const double TOL = 5 * hxmax;
  for (;;) {
    double t = b;
    for (int c = 0; c < N; c++) t = t + F(t);
    double d = t - e;
    if (std::fabs(d) > TOL) {
      if (d < 0) N = N + int(std::fabs(d / hmax));
      if (d > 0) {
        if ((N - int(std::fabs(d / hmax))) > 2) N = N - int(std::fabs(d / hmax));
        else N = int(0.75 * N);
      }
    } else {
      return N;
    }
  }
}

// ---------------------------------------------------------------------------
// Bodies
// ---------------------------------------------------------------------------
// Segment intersection test of the polygon walk, Polyhedrons_Objects.cpp:327-501.
bool segments_cross(double x11, double y11, double x12, double y12, double x21, double y21, double x22, double y22) {
  bool hit = false;
  double xi = 0, yi = 0;
  const double x1min = x11 < x12 ? x11 : x12, x1max = x11 > x12 ? x11 : x12;
  const double y1min = y11 < y12 ? y11 : y12, y1max = y11 > y12 ? y11 : y12;
  const double x2min = x21 < x22 ? x21 : x22, x2max = x21 > x22 ? x21 : x22;
  const double y2min = y21 < y22 ? y21 : y22, y2max = y21 > y22 ? y21 : y22;
  double k1, b1, k2, b2;
  if ((x1min != x1max) && (y1min != y1max) && (x2min != x2max) && (y2min != y2max)) {
    k1 = (y11 - y12) / (x11 - x12);
    b1 = y11 - k1 * x11;
    k2 = (y21 - y22) / (x21 - x22);
    b2 = y21 - k2 * x21;
    if (k1 != k2) {
      xi = (b1 - b2) / (k2 - k1);
      yi = k1 * xi + b1;
      hit = true;
    }
  } else if (y1min == y1max && x2min == x2max) {
    xi = x2min;
    yi = y1min;
    hit = true;
  } else if ((y1min == y1max) && (x2min != x2max) && (y2min != y2max)) {
    yi = y1min;
    k2 = (y21 - y22) / (x21 - x22);
    b2 = y21 - k2 * x21;
    xi = (yi - b2) / k2;
    hit = true;
  } else if (x1min == x1max && y2min == y2max) {
    xi = x1min;
    yi = y2min;
    hit = true;
  } else if ((x1min == x1max) && (x2min != x2max) && (y2min != y2max)) {
    xi = x1min;
    k2 = (y21 - y22) / (x21 - x22);
    b2 = y21 - k2 * x21;
    yi = k2 * xi + b2;
    hit = true;
  } else if ((x1min != x1max) && (y1min != y1max) && (x2min == x2max)) {
    xi = x2min;
    k1 = (y11 - y12) / (x11 - x12);
    b1 = y11 - k1 * x11;
    yi = k1 * xi + b1;
    hit = true;
  } else if ((x1min != x1max) && (y1min != y1max) && (y2min == y2max)) {
    yi = y2min;
    k1 = (y11 - y12) / (x11 - x12);
    b1 = y11 - k1 * x11;
    xi = (yi - b1) / k1;
    hit = true;
  }
  if (hit && !(x1min <= xi && xi <= x1max && x2min <= xi && xi <= x2max && y1min <= yi && yi <= y1max &&
               y2min <= yi && yi <= y2max))
    hit = false;
  return hit;
}

}  // namespace

// Point-in-polygon of Polyhedrons_Objects.cpp:190-324: count crossings of
// the segment from an inside point to (x, y) with the polygon's sides; an
// even count means inside.  The inside point slides along the segment
// between the two given inside points until it is not collinear with a
// vertex.
bool Polyhedron::inside(double x, double y) {
  const int n = (int)this->x.size();
  if (!rect_defined) {
    x_rect_min = x_rect_max = this->x[0];
    y_rect_min = y_rect_max = this->y[0];
    for (int k = 0; k < n; k++) {
      x_rect_min = x_rect_min < this->x[k] ? x_rect_min : this->x[k];
      x_rect_max = x_rect_max > this->x[k] ? x_rect_max : this->x[k];
      y_rect_min = y_rect_min < this->y[k] ? y_rect_min : this->y[k];
      y_rect_max = y_rect_max > this->y[k] ? y_rect_max : this->y[k];
    }
    rect_defined = true;
  }
  bool in = (x_rect_min <= x && x <= x_rect_max && y_rect_min <= y && y <= y_rect_max);
  if (!in) return false;
  const double xpmin = p1x < p2x ? p1x : p2x, xpmax = p1x > p2x ? p1x : p2x;
  const double ypmin = p1y < p2y ? p1y : p2y, ypmax = p1y > p2y ? p1y : p2y;
  if ((xpmin != xpmax) && (ypmin != ypmax)) {
    kpi = (p2y - p1y) / (p2x - p1x);
    bpi = p1y - kpi * p1x;
  }
  double px = 0, py = 0;
  bool again;
  int div = 1;
  do {
    again = false;
    if (xpmin == xpmax) {
      px = xpmin;
      py = ypmin + (ypmax - ypmin) / div;
    } else if (ypmin == ypmax) {
      px = xpmin + (xpmax - xpmin) / div;
      py = ypmin;
    } else {
      px = xpmin + (xpmax - xpmin) / div;
      py = kpi * px + bpi;
    }
    if (px == x && py == y) again = true;
    if (!again) {
      int k = 0;
      do {
        if (px == x && x == this->x[k]) again = true;
        else if (py == y && y == this->y[k]) again = true;
        else if (this->y[k] == (py - y) * (this->x[k] - px) / (px - x) + py) again = true;
        k++;
      } while (k < n && !again);
    }
    div++;
  } while (div < 5000 && again);
  if (again) {
    std::cout << "simple-ts: no usable inside point for a polyhedron; the point is taken as inside" << std::endl;
    return in;
  }
  double crossings = 0;
  for (int k = 0; k < n; k++) {
    const int k1 = (k < n - 1) ? k + 1 : 0;
    crossings = crossings + segments_cross(px, py, x, y, this->x[k], this->y[k], this->x[k1], this->y[k1]);
  }
  return (crossings / 2.0) == std::floor(crossings / 2.0);
}

// Circles_Objects.h:118-126
bool Circle::inside(double x, double y) const {
  return ((pow((x - x_c), 2) + pow((y - y_c), 2)) <= pow(r_c, 2));
}

namespace {

std::string read_bodies(Case &c) {
  std::string text;
  if (c.flag("ToImport_data_for_circles_from_file_b")) {
    if (!read_file(c.dir + "/solve_circles_b.txt", text)) return "cannot open " + c.dir + "/solve_circles_b.txt";
    Reader r(text);
    unsigned n = (unsigned)r.num();
    r.rest();
    for (unsigned k = 0; k < n && r.ok; k++) {
      Circle b;
      b.x_c = r.num();
      b.y_c = r.num();
      b.r_c = r.num();
      r.rest();
      b.u_body = r.num();
      b.v_body = r.num();
      r.rest();
      b.p_body = r.num(); r.rest();
      b.T_body = r.num(); r.rest();
      b.is_slip = r.num() != 0; r.rest();
      b.F_slip = r.num(); r.rest();
      b.w_slip = r.num(); r.rest();
      b.is_jump = r.num() != 0; r.rest();
      b.F_jump = r.num(); r.rest();
      b.dTdn = r.num(); r.rest();
      b.drag = r.num() != 0; r.rest();
      c.circles.push_back(b);
    }
    if (!r.ok) return "cannot parse solve_circles_b.txt";
  }
  if (c.flag("ToImport_data_for_polyhedrons_from_file_b")) {
    if (!read_file(c.dir + "/solve_polyhedrons_b.txt", text))
      return "cannot open " + c.dir + "/solve_polyhedrons_b.txt";
    Reader r(text);
    unsigned n = (unsigned)r.num();
    r.rest();
    for (unsigned k = 0; k < n && r.ok; k++) {
      Polyhedron b;
      unsigned nv = (unsigned)r.num();
      r.rest();
      b.p1x = r.num();
      b.p1y = r.num();
      r.rest();
      b.p2x = r.num();
      b.p2y = r.num();
      r.rest();
      b.u_body = r.num();
      b.v_body = r.num();
      r.rest();
      b.p_body = r.num(); r.rest();
      b.T_body = r.num(); r.rest();
      b.is_slip = r.num() != 0; r.rest();
      b.F_slip = r.num(); r.rest();
      b.w_slip = r.num(); r.rest();
      b.is_jump = r.num() != 0; r.rest();
      b.F_jump = r.num(); r.rest();
      b.dTdn = r.num(); r.rest();
      b.drag = r.num() != 0; r.rest();
      for (unsigned v = 0; v < nv; v++) {
        b.x.push_back(r.num());
        b.y.push_back(r.num());
      }
      if (nv < 3) return "polyhedron with fewer than 3 vertices";
      c.polys.push_back(b);
    }
    if (!r.ok) return "cannot parse solve_polyhedrons_b.txt";
  }
  return "";
}

}  // namespace

std::string load_case(const std::string &dir, Case &c) {
  c = Case();
  c.dir = dir;
  std::string text;
  const std::string path = dir + "/EnteredData.txt";
  if (!read_file(path, text)) return "cannot open " + path;
  std::istringstream in(text);
  c.val.assign((size_t)kNumEntries, 0.0);
  c.comment.assign((size_t)kNumEntries, std::string());
  std::string line;
  for (int k = 0; k < kNumEntries; k++) {
    if (!std::getline(in, line)) return path + ": expected 126 lines, got " + std::to_string(k);
    if (!parse_entry(line, kEntries[k].kind, c.val[(size_t)k], c.comment[(size_t)k]))
      return path + ": cannot parse line " + std::to_string(k + 1) + " (" + kEntries[k].name + ")";
  }
  // Non-dimensional constants, S:2013-2071
  const unsigned type = (unsigned)c.get("TypeOfNonDimensiolization");
  c.rho_gas_nd = c.get("p_gas_nd") / c.get("T_gas_nd");
  if (type == 1 || type == 2) {
    if (c.flag("GivenReKn")) c.at("u_given_xb") = c.get("Re") * c.get("Kn") * sqrt(15.0 * M_PI / 128.0);
    c.u_given_xb_in_it = c.get("u_given_xb");
    c.cT = (1.0 / c.get("Re")) * (c.u_given_xb_in_it * c.get("p_BC_xb") / c.get("T_BC_xb"));
    const double Ma = c.get("Ma"), g = c.get("gamma1");
    c.cPch = (type == 1) ? 0.5 : c.u_given_xb_in_it * c.u_given_xb_in_it / (g * Ma * Ma);
  } else if (type == 3) {
    c.at("Ma") = sqrt(6.0 / 5.0);
    c.at("Pr") = 2.0 / 3.0;
    c.at("gamma1") = 5.0 / 3.0;
    c.at("c_mu") = (5.0 / 16.0) * sqrt(2.0 * M_PI / c.get("gamma1"));
    c.cT = c.get("Kn") * 5.0 * sqrt(M_PI) / 16.0;
    c.cPch = 0.5;
  } else {
    return "TypeOfNonDimensiolization must be 1, 2 or 3";
  }
  const double gamma1 = c.get("gamma1"), Ma = c.get("Ma"), Kn = c.get("Kn"), c_mu = c.get("c_mu");
  c.c_Fr_v = c.flag("SolveGravityField") ? 2.0 / (gamma1 * c.get("Fr") * Ma * Ma) : 0;
  c.cT1 = gamma1 * c_mu * Kn / (Ma * c.get("Pr"));
  c.cT2 = gamma1 * (gamma1 - 1.0) * c_mu * Kn * Ma;

  // Mesh, S:2078-2248
  MeshFn X, Y;
  X.kind = Y.kind = (unsigned)c.get("kind_of_mesh");
  X.hmin = c.get("hxmin"); X.hmax = c.get("hxmax");
  X.fmin = c.get("xfmin"); X.fmax = c.get("xfmax"); X.bmin = c.get("xbmin"); X.bmax = c.get("xbmax");
  X.mid = c.get("xmid");
  Y.hmin = c.get("hymin"); Y.hmax = c.get("hymax");
  Y.fmin = c.get("ybmin"); Y.fmax = c.get("ybmax"); Y.bmin = c.get("ytmin"); Y.bmax = c.get("ytmax");
  Y.mid = c.get("ymid");
  mesh_coefficients(X, Y, X.hmin, X.hmax);
  if (X.kind == 2 || X.kind == 4) {
    c.at("xmid") = X.mid;
    c.at("ymid") = Y.mid;
  }
  int Nx = (int)c.get("Nx"), Ny = (int)c.get("Ny");
  if (!c.flag("ToReadSolvedDataFromFile")) {
    Nx = adjust_count(Nx, c.get("x_b"), c.get("x_e"), X.hmax, X.hmax, X);
    Ny = adjust_count(Ny, c.get("y_b"), c.get("y_e"), X.hmax, Y.hmax, Y);
    c.at("Nx") = Nx;
    c.at("Ny") = Ny;
  }
  if (Nx < 4 || Ny < 4) return "grid must be at least 4x4";
  c.Nx = Nx;
  c.Ny = Ny;
  c.x_f.assign((size_t)Nx + 1, 0.0);
  c.y_f.assign((size_t)Ny + 1, 0.0);
  c.hx.assign((size_t)Nx, 0.0);
  c.hy.assign((size_t)Ny, 0.0);
  c.x_v.assign((size_t)Nx, 0.0);
  c.y_v.assign((size_t)Ny, 0.0);
  const double x_b = c.get("x_b"), y_b = c.get("y_b");
  c.x_f[0] = x_b - 0.5 * X.hmax;
  c.x_f[1] = x_b + 0.5 * X.hmax;
  for (int i = 1; i < Nx; i++) {
    double t = c.x_f[i] + 0.5 * X(c.x_f[i]);
    c.x_f[i + 1] = c.x_f[i] + X(t);
  }
  c.y_f[0] = y_b - 0.5 * Y.hmax;
  c.y_f[1] = y_b + 0.5 * Y.hmax;
  for (int j = 1; j < Ny; j++) {
    double t = c.y_f[j] + 0.5 * Y(c.y_f[j]);
    c.y_f[j + 1] = c.y_f[j] + Y(t);
  }
  auto centres = [&]() {
    for (int i = 0; i < Nx; i++) {
      c.hx[i] = c.x_f[i + 1] - c.x_f[i];
      c.x_v[i] = c.x_f[i] + 0.5 * c.hx[i];
    }
    for (int j = 0; j < Ny; j++) {
      c.hy[j] = c.y_f[j + 1] - c.y_f[j];
      c.y_v[j] = c.y_f[j] + 0.5 * c.hy[j];
    }
  };
  centres();
  const int N_I_T = (int)c.get("N_I_T");
  if (N_I_T > 0) {  // mesh smoothing; N_I_T doubles as its count (S:2217-2246)
    c.Iter_T = 0;
    do {
      for (int i = 1; i < Nx; i++) c.x_f[i + 1] = c.x_f[i] + X(c.x_v[i]);
      for (int j = 1; j < Ny; j++) c.y_f[j + 1] = c.y_f[j] + Y(c.y_v[j]);
      centres();
      c.Iter_T++;
    } while (c.Iter_T < N_I_T);
  }

  // Bodies and mask, S:6269-6345
  std::string err = read_bodies(c);
  if (!err.empty()) return err;
  const size_t Na = (size_t)Nx * (size_t)Ny;
  c.vol.assign(Na, 0u);
  if (c.flag("ToImport_data_for_circles_from_file_b") || c.flag("ToImport_data_for_polyhedrons_from_file_b")) {
    for (size_t k = 0; k < c.circles.size(); k++)
      for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++)
          if (c.circles[k].inside(c.x_v[i], c.y_v[j])) c.vol[(size_t)i + (size_t)j * Nx] = 1u + (unsigned)k;
    for (size_t k = 0; k < c.polys.size(); k++)
      for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++)
          if (c.polys[k].inside(c.x_v[i], c.y_v[j]))
            c.vol[(size_t)i + (size_t)j * Nx] = 1u + (unsigned)(c.circles.size() + k);
  }
  for (const Circle &b : c.circles)
    c.bodies.push_back(sts_body{b.u_body, b.v_body, b.p_body, b.T_body, b.is_slip, b.F_slip, b.is_jump, b.F_jump,
                                b.dTdn});
  for (const Polyhedron &b : c.polys)
    c.bodies.push_back(sts_body{b.u_body, b.v_body, b.p_body, b.T_body, b.is_slip, b.F_slip, b.is_jump, b.F_jump,
                                b.dTdn});

  // Solve boxes, S:6009-6267 (serial build): periodic about OX spans every
  // column (S:6055-6058, S:6098-6102); y limits are those of the other cases
  // (S:6140-6144, S:6173-6177).
  const bool pbc = c.flag("Pressure_BC"), gxb = c.flag("Given_velocity_on_xb"), gxe = c.flag("Given_velocity_on_xe");
  if (c.flag("Periodic_boundary_conditions_about_OX")) {
    c.Nx_u[0] = c.Nx_v[0] = c.Nx_o[0] = 0;
    c.Nx_u[1] = c.Nx_v[1] = c.Nx_o[1] = Nx;
  } else if (pbc) {
    c.Nx_u[0] = c.Nx_v[0] = c.Nx_o[0] = 1;
    c.Nx_u[1] = Nx - 1;
    c.Nx_v[1] = Nx - 2;
    c.Nx_o[1] = Nx - 2;
  } else {
    c.Nx_u[0] = gxb ? 2 : 1;
    c.Nx_v[0] = 1;
    c.Nx_o[0] = 1;
    c.Nx_u[1] = Nx - 1;
    c.Nx_v[1] = Nx - 2;
    c.Nx_o[1] = gxe ? Nx - 1 : Nx - 2;
  }
  c.Ny_u[0] = 1;
  c.Ny_v[0] = 1;
  c.Ny_o[0] = 1;
  c.Ny_u[1] = Ny - 2;
  c.Ny_v[1] = Ny - 1;
  c.Ny_o[1] = Ny - 2;
  return "";
}

void Case::fill_grid(sts_grid &g) const {
  std::memset(&g, 0, sizeof g);
  g.Nx = Nx;
  g.Ny = Ny;
  g.hx = hx.data();
  g.hy = hy.data();
  g.vol_inf_fb = vol.data();
  g.nbodies = (int)bodies.size();
  g.bodies = bodies.empty() ? nullptr : bodies.data();
  for (int k = 0; k < 2; k++) {
    g.Nx_u[k] = Nx_u[k]; g.Ny_u[k] = Ny_u[k]; g.Nx_v[k] = Nx_v[k];
    g.Ny_v[k] = Ny_v[k]; g.Nx_o[k] = Nx_o[k]; g.Ny_o[k] = Ny_o[k];
  }
  g.ht = get("ht");
  g.gamma1 = get("gamma1");
  g.Kn = get("Kn");
  g.cT = cT;
  g.cT1 = cT1;
  g.cT2 = cT2;
  g.cPch = cPch;
  g.c_Fr_v = c_Fr_v;
  g.given_xb = flag("Given_velocity_on_xb");
  g.given_xe = flag("Given_velocity_on_xe");
  g.kn_local = flag("To_Use_Kn_local_in_wall_BC");
  g.dudx_0_xb = flag("dudx_0_BC_xb");
  g.durhodx_0_xb = flag("durhodx_0_BC_xb");
  g.drhodt_xb = flag("drhodt_durhodx_dvrhody_0_BC_xb");
  g.dudx_0_xe = flag("dudx_0_BC_xe");
  g.durhodx_0_xe = flag("durhodx_0_BC_xe");
  g.drhodt_xe = flag("drhodt_durhodx_dvrhody_0_BC_xe");
  g.u_Min_Mout_xe = flag("u_uMin_Mout_0_BC_xe");
  g.dudy_0_yb = flag("dudy_0_BC_yb");
  g.dudy_0_ye = flag("dudy_0_BC_ye");
  g.dvdx_0_xb = flag("dvdx_0_BC_xb");
  g.dvdx_0_xe = flag("dvdx_0_BC_xe");
  g.dvdy_0_yb = flag("dvdy_0_BC_yb");
  g.dvdy_0_ye = flag("dvdy_0_BC_ye");
  g.dpdx_0_xb = flag("dpdx_0_BC_xb");
  g.dpdx_0_xe = flag("dpdx_0_BC_xe");
  g.dpdy_0_yb = flag("dpdy_0_BC_yb");
  g.dpdy_0_ye = flag("dpdy_0_BC_ye");
  g.dTdx_0_xb = flag("dTdx_0_BC_xb");
  g.dTdx_0_xe = flag("dTdx_0_BC_xe");
  g.dTdy_0_yb = flag("dTdy_0_BC_yb");
  g.dTdy_0_ye = flag("dTdy_0_BC_ye");
  g.periodic_x = flag("Periodic_boundary_conditions_about_OX");
}

// WriteEnteredDataToFile, S:5044-5238: std::setprecision(15), default float
// format; integers and bools as integers.
std::string write_config(const Case &c, const std::string &path) {
  std::ofstream o(path);
  if (!o) return "cannot write " + path;
  o << std::setprecision(15);
  for (int k = 0; k < kNumEntries; k++) {
    const double v = c.val[(size_t)k];
    if (kEntries[k].kind == Kind::Double) o << v;
    else if (kEntries[k].kind == Kind::Int) o << (int)v;
    else o << (unsigned)v;
    o << "\t" << c.comment[(size_t)k] << std::endl;
  }
  return "";
}

}  // namespace sts

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct sts_case {
  sts::Case c;
};

namespace {
thread_local std::string g_host_err;
}

extern "C" const char *sts_host_last_error(void) { return g_host_err.c_str(); }

extern "C" int sts_case_load(const char *dir, sts_case **out) {
  if (!dir || !out) {
    g_host_err = "null argument";
    return STS_ERR_ARG;
  }
  *out = nullptr;
  sts_case *h = new sts_case();
  std::string err = sts::load_case(dir, h->c);
  if (!err.empty()) {
    delete h;
    g_host_err = err;
    return STS_ERR_IO;
  }
  *out = h;
  return STS_OK;
}

extern "C" void sts_case_free(sts_case *c) { delete c; }

extern "C" int sts_case_dims(const sts_case *c, int *Nx, int *Ny) {
  if (!c) return STS_ERR_ARG;
  if (Nx) *Nx = c->c.Nx;
  if (Ny) *Ny = c->c.Ny;
  return STS_OK;
}

extern "C" int sts_case_grid(const sts_case *c, sts_grid *g) {
  if (!c || !g) return STS_ERR_ARG;
  c->c.fill_grid(*g);
  return STS_OK;
}

extern "C" int sts_case_mesh(const sts_case *c, double *x_f, double *y_f, double *hx, double *hy, double *x_v,
                             double *y_v) {
  if (!c) return STS_ERR_ARG;
  const sts::Case &k = c->c;
  if (x_f) std::memcpy(x_f, k.x_f.data(), sizeof(double) * k.x_f.size());
  if (y_f) std::memcpy(y_f, k.y_f.data(), sizeof(double) * k.y_f.size());
  if (hx) std::memcpy(hx, k.hx.data(), sizeof(double) * k.hx.size());
  if (hy) std::memcpy(hy, k.hy.data(), sizeof(double) * k.hy.size());
  if (x_v) std::memcpy(x_v, k.x_v.data(), sizeof(double) * k.x_v.size());
  if (y_v) std::memcpy(y_v, k.y_v.data(), sizeof(double) * k.y_v.size());
  return STS_OK;
}

extern "C" int sts_case_vol(const sts_case *c, unsigned int *vol) {
  if (!c || !vol) return STS_ERR_ARG;
  std::memcpy(vol, c->c.vol.data(), sizeof(unsigned) * c->c.vol.size());
  return STS_OK;
}

extern "C" double sts_case_value(const sts_case *c, const char *name) {
  if (!c || !name) return std::numeric_limits<double>::quiet_NaN();
  const sts::Case &k = c->c;
  const std::string n(name);
  if (n == "cT") return k.cT;
  if (n == "cT1") return k.cT1;
  if (n == "cT2") return k.cT2;
  if (n == "cPch") return k.cPch;
  if (n == "c_Fr_v") return k.c_Fr_v;
  if (n == "rho_gas_nd") return k.rho_gas_nd;
  if (n == "Iter_T") return k.Iter_T;
  return k.get(name);
}

extern "C" int sts_case_write_config(const sts_case *c, const char *path) {
  if (!c || !path) return STS_ERR_ARG;
  std::string err = sts::write_config(c->c, path);
  if (!err.empty()) {
    g_host_err = err;
    return STS_ERR_IO;
  }
  return STS_OK;
}
