// sts_gpu: the SIMPLE-TS program on one MI355X.
//
// Mirrors main() of the reference (SIMPLE_TS_MPI.cpp, "S:n"): it reads the
// same EnteredData.txt and body files from the case directory, performs the
// same setup sequence (S:2476-3226), runs loops 1-3 (S:3230-4691) with the
// same convergence and stop logic, and writes the same output files (mesh
// text files, u/v/p/Temper/vol_inf_fb fields, EnteredData.txt, it_begin.txt,
// Tmp.txt, Iterations_in_loops.txt, ContinuityEquation*).  Every field
// operation runs on the GPU through the C ABI of libsimple_ts_hip.so; the
// host only evaluates the loop-control scalars, as the reference's rank 0
// does.
//
// Usage: sts_gpu [case_dir] [--timing file.json] (--help lists the other options)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <chrono>
#include <sstream>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "simple_ts.h"
#include "sts_case.h"

namespace {

[[noreturn]] void die(const std::string &msg) {
  std::cerr << "sts_gpu: " << msg << std::endl;
  std::exit(2);
}

void ck(int rc, const char *what) {
  if (rc != STS_OK) die(std::string(what) + ": " + sts_last_error());
}

// WriteMassiveToFile_1D (WriteReadData.h:139-173): scientific, 15 digits.
template <typename T>
void write_text(const std::string &path, const T *a, size_t n) {
  std::ofstream o(path);
  if (!o) die("cannot write " + path);
  for (size_t k = 0; k < n; k++) o << std::setiosflags(std::ios::scientific) << std::setprecision(15) << a[k] << std::endl;
}

template <typename T>
void write_bin(const std::string &path, const T *a, size_t n) {
  std::ofstream o(path, std::ios::out | std::ios::trunc | std::ios::binary);
  if (!o) die("cannot write " + path);
  o.write((const char *)a, (std::streamsize)(sizeof(T) * n));
}

template <typename T>
void read_field(const std::string &path, T *a, size_t n, bool binary) {
  std::ifstream in(binary ? std::ifstream(path, std::ios::in | std::ios::binary) : std::ifstream(path));
  if (!in) die("cannot read " + path);
  if (binary) {
    in.read((char *)a, (std::streamsize)(sizeof(T) * n));
    if ((size_t)in.gcount() != sizeof(T) * n) die("short file " + path);
  } else {
    for (size_t k = 0; k < n; k++)
      if (!(in >> a[k])) die("cannot parse " + path);
  }
}

std::string num(double v) {  // toString (stringConv.h): default stream format
  std::ostringstream s;
  s << v;
  return s.str();
}

struct Driver {
  sts::Case c;
  std::string dir;
  std::string norms_file;  // --continuity-norms: norms of every continuity evaluation
  sts_ctx *ctx = nullptr;
  sts_grid grid{};
  int Nx = 0, Ny = 0;
  size_t Na = 0;
  std::vector<unsigned char> mask;

  // loop state (S:245-1792 globals of the same names)
  double it = 0, it_begin = 0, ht = 0;
  unsigned it_counter = 0, Iter = 0, Iter_pr1 = 0, Iter_pr2 = 0, N_I = 1, N_I_p_c = 2, Iter_p_c = 0;
  bool continue_iter = true, countinue_time_step = true, solve_is_finished = false;
  std::vector<double> max_res_p, max_res_T, max_res_p0, max_res_T0;
  std::vector<unsigned> max_Iter_p_c;
  double max_res_u = 0, max_res_v = 0, max_res_u0 = 0, max_res_v0 = 0;
  double u_given_xb_in_it = 0, p_BC_xb = 0;
  bool tmp_first = true, iters_first = true;
  time_t time_solve_start = 0;
  double time_solve_stop = 0;
  long long outer_total = 0;

  std::string path(const std::string &f) const { return dir + "/" + f; }

  void host_fields(std::vector<double> &u, std::vector<double> &v, std::vector<double> &p, std::vector<double> &T,
                   std::vector<double> &rho) {
    for (std::vector<double> *a : {&u, &v, &p, &T, &rho}) a->assign(Na, 0.0);
    for (size_t k = 0; k < Na; k++) {
      u[k] = c.get("u_gas_nd");
      v[k] = c.get("v_gas_nd");
      p[k] = c.get("p_gas_nd");
      rho[k] = c.rho_gas_nd;
      T[k] = c.get("T_gas_nd");
    }
  }

  void get(int f, std::vector<double> &a) {
    a.resize(Na);
    ck(sts_get_field(ctx, f, a.data()), "sts_get_field");
  }
  void set(int f, const std::vector<double> &a) { ck(sts_set_field(ctx, f, a.data()), "sts_set_field"); }

  // S:2547-2562, S:2908-2930: fixed p (and T, rho) in the columns outside
  // the solve box when the pressure is given on both ends.
  void pressure_columns(bool with_rho) {
    std::vector<double> p, T, rho;
    get(STS_P, p);
    get(STS_T, T);
    if (with_rho) get(STS_RHO, rho);
    const double pb = c.get("p_BC_xb"), pe = c.get("p_BC_xe"), Tb = c.get("T_BC_xb"), Te = c.get("T_BC_xe");
    for (int j = 0; j < Ny; j++) {
      for (int i = 0; i < c.Nx_o[0]; i++) {
        size_t n = (size_t)i + (size_t)j * Nx;
        if (!mask[n]) continue;
        p[n] = pb;
        T[n] = Tb;
        if (with_rho) rho[n] = pb / Tb;
      }
      for (int i = c.Nx_o[1]; i < Nx; i++) {
        size_t n = (size_t)i + (size_t)j * Nx;
        if (!mask[n]) continue;
        p[n] = pe;
        T[n] = Te;
        if (with_rho) rho[n] = pe / Te;
      }
    }
    set(STS_P, p);
    set(STS_T, T);
    if (with_rho) set(STS_RHO, rho);
  }

  // The channel's fluid rows at column 1: jb the last body row below it,
  // je the first body row above it (S:2569-2580, S:2719-2730; rows keep their
  // previous values when the column has no such transition).
  unsigned j_channel_b = 0, j_channel_e = 0;
  void channel_rows() {
    for (int j = 1; j < Ny; j++) {
      const size_t node = 1 + (size_t)j * Nx;
      if (!mask[node - Nx] && mask[node]) j_channel_b = (unsigned)j - 1;
      if (mask[node - Nx] && !mask[node]) j_channel_e = (unsigned)j;
    }
  }

  // Slip coefficient of the exact solutions: polyhedron 0's (S:2586-2592).
  const sts::Polyhedron &slip_body() const {
    if (c.polys.empty()) die("ToStartFromExactSolutionForChannelFlowWithPressureBC needs a polyhedron body 0");
    return c.polys[0];
  }

  // ToStartFromExactSolutionForChannelFlowWithPressureBC with Pressure_BC
  // (S:2566-2689): p on the solve columns and u on every fluid cell from the
  // exact solution of slip flow in a channel between the box-end pressures.
  void exact_start_pressure_bc() {
    std::vector<double> u, p;
    get(STS_U, u);
    get(STS_P, p);
    channel_rows();
    const double *x_v = c.x_v.data(), *y_v = c.y_v.data();
    const double h_channel = y_v[j_channel_e] - y_v[j_channel_b];
    const double L_channel = x_v[c.Nx_o[1]] - x_v[c.Nx_o[0] - 1];
    const double Kn = c.get("Kn"), pb = c.get("p_BC_xb"), pe = c.get("p_BC_xe");
    const sts::Polyhedron &b0 = slip_body();
    const double sigma = (double)b0.is_slip * b0.F_slip, Px = pb / pe;
    std::vector<double> xa(Nx), ya(Ny);
    for (int i = 0; i < Nx; i++) xa[i] = (x_v[i] - x_v[0]) / L_channel;
    for (int j = 0; j < Ny; j++) {
      if ((unsigned)j <= j_channel_b) ya[j] = -0.5 * h_channel;
      else if (j_channel_e <= (unsigned)j) ya[j] = 0.5 * h_channel;
      else ya[j] = y_v[(unsigned)j + j_channel_b] - (y_v[j_channel_b] + 0.5 * h_channel);
    }
    // the square root of the pressure profile, (p/pe + 6 sigma Kn)
    auto root = [&](int i) {
      return sqrt(pow((6.0 * sigma * Kn), 2) + (1.0 + 12.0 * sigma * Kn) * xa[i] +
                  (Px + 12.0 * sigma * Kn) * Px * (1.0 - xa[i]));
    };
    for (int j = 0; j < Ny; j++)
      for (int i = c.Nx_o[0]; i < c.Nx_o[1]; i++) {
        const size_t node = (size_t)i + (size_t)j * Nx;
        if (mask[node]) p[node] = pe * (-6.0 * sigma * Kn + root(i));
      }
    const unsigned type = (unsigned)c.get("TypeOfNonDimensiolization");
    double nd = 0;
    if (type == 1 || type == 2) nd = c.get("Re") * h_channel * h_channel / (4.0 * c.get("u_given_xb"));
    else if (type == 3) nd = h_channel / (5.0 * sqrt(M_PI * c.get("T_BC_xb")) * Kn * L_channel);
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < Nx; i++) {
        const size_t node = (size_t)i + (size_t)j * Nx;
        if (!mask[node]) continue;
        u[node] = -nd * ((pe * (1.0 - Px * Px + 12.0 * sigma * Kn * (1.0 - Px))) / (2.0 * root(i))) *
                  (1.0 - 4.0 * ya[j] * ya[j] / (h_channel * h_channel) + 4.0 * sigma * Kn / p[node]);
      }
    set(STS_U, u);
    set(STS_P, p);
  }

  // ToStartFromExactSolutionForChannelFlowWithPressureBC with a given inflow
  // velocity (S:2694-2806): the inflow columns take p_BC_xb, one box-end
  // pressure is set so that the channel's centre moves at u_given_xb, and
  // the inflow columns i < Nx_u[0] take the parabolic slip profile.
  void exact_start_given_velocity() {
    std::vector<double> u, p;
    get(STS_U, u);
    get(STS_P, p);
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < c.Nx_o[0] + 1; i++) {
        const size_t node = (size_t)i + (size_t)j * Nx;
        if (mask[node]) p[node] = p_BC_xb;
      }
    channel_rows();
    const double *x_v = c.x_v.data(), *y_v = c.y_v.data();
    const double h_channel = c.y_f[j_channel_e + 1] - c.y_f[j_channel_b];  // j = Ny > 0 after the scan
    const double L_channel = x_v[c.Nx_o[1]] - x_v[c.Nx_o[0] - 1];
    const double yc = y_v[j_channel_b] + 0.5 * h_channel;
    const double ug = c.get("u_given_xb"), Re = c.get("Re"), Kn = c.get("Kn");
    // the reference's centre-line factor, its operator precedence kept
    const double q = yc - y_v[j_channel_b] / h_channel;
    const double dp = 4.0 * L_channel * ug * ug / (Re * h_channel * h_channel * (q - q * q));
    if (c.flag("correct_p_BC_xb")) p_BC_xb = c.get("p_BC_xe") + dp;
    else c.at("p_BC_xe") = p_BC_xb - dp;
    c.at("p_BC_xb") = p_BC_xb;
    const double dpdx = (c.get("p_BC_xe") - p_BC_xb) / L_channel;
    const unsigned type = (unsigned)c.get("TypeOfNonDimensiolization");
    double nd = 0;
    if (type == 1 || type == 2) nd = Re * h_channel * h_channel / (4.0 * ug);
    else if (type == 3) nd = 4.0 * h_channel * h_channel / (5.0 * sqrt(M_PI * c.get("T_BC_xb")) * Kn);
    const sts::Polyhedron &b0 = slip_body();
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < c.Nx_u[0]; i++) {
        const size_t node = (size_t)i + (size_t)j * Nx;
        if (!mask[node]) continue;
        const double eta = (y_v[j] - y_v[j_channel_b]) / h_channel;
        u[node] = -nd * dpdx * (eta - eta * eta + Kn * b0.F_slip * (double)b0.is_slip / h_channel);
      }
    set(STS_U, u);
    set(STS_P, p);
  }

  // ... with a given outflow velocity (S:2808-2830): the outflow columns
  // take p_BC_xe.
  void exact_start_given_outflow() {
    std::vector<double> p;
    get(STS_P, p);
    for (int j = 0; j < Ny; j++)
      for (int i = c.Nx_o[1]; i < Nx; i++) {
        const size_t node = (size_t)i + (size_t)j * Nx;
        if (mask[node]) p[node] = c.get("p_BC_xe");
      }
    set(STS_P, p);
  }

  // WriteSolvedDataToFile, S:5564-5731 (serial build).
  void write_solved() {
    ck(sts_jump_overwrite(ctx), "sts_jump_overwrite");  // S:5566-5567
    std::vector<double> u, v, p, T;
    get(STS_U, u);
    get(STS_V, v);
    get(STS_P, p);
    get(STS_T, T);
    const bool bin = c.flag("ToWriteSolvedDataToBinaryFile");
    const std::string ext = bin ? ".bin" : ".txt";
    auto out = [&](const std::string &name) {
      if (bin) {
        write_bin(path(name + ".bin"), name == "u" ? u.data() : name == "v" ? v.data() : name == "p" ? p.data() : T.data(),
                  Na);
      } else {
        write_text(path(name + ".txt"), name == "u" ? u.data() : name == "v" ? v.data() : name == "p" ? p.data() : T.data(),
                   Na);
      }
    };
    for (const char *n : {"u", "v", "p", "Temper"}) out(n);
    if (bin) write_bin(path("vol_inf_fb.bin"), c.vol.data(), Na);
    else write_text(path("vol_inf_fb.txt"), c.vol.data(), Na);
    if (c.flag("ToWriteSolvedDataToNewFiles")) {
      const std::string s = num(it);
      if (bin) {
        write_bin(path("u." + s + ".bin"), u.data(), Na);
        write_bin(path("v." + s + ".bin"), v.data(), Na);
        write_bin(path("p." + s + ".bin"), p.data(), Na);
        write_bin(path("Temper." + s + ".bin"), T.data(), Na);
        write_bin(path("vol_inf_fb." + s + ".bin"), c.vol.data(), Na);
      } else {
        write_text(path("u." + s + ".txt"), u.data(), Na);
        write_text(path("v." + s + ".txt"), v.data(), Na);
        write_text(path("p." + s + ".txt"), p.data(), Na);
        write_text(path("Temper." + s + ".txt"), T.data(), Na);
        write_text(path("vol_inf_fb.txt." + s + ".txt"), c.vol.data(), Na);  // the reference's name
      }
    }
    write_config();
    std::ofstream o(path("it_begin.txt"));  // Write_it_ToFile, S:6742-6760
    o << it << std::endl;
  }

  void write_config() {
    c.at("p_BC_xb") = p_BC_xb;
    c.at("solve_is_finished") = solve_is_finished;
    std::string err = sts::write_config(c, path("EnteredData.txt"));
    if (!err.empty()) die(err);
  }

  // SolveContinuityEquation, S:10621-10714
  void continuity() {
    std::vector<double> cs(Na), cf(Na);
    double norms[3];
    ck(sts_continuity(ctx, cs.data(), cf.data(), norms), "sts_continuity");
    if (!norms_file.empty()) {  // not a reference output: one line per evaluation
      std::ofstream o(norms_file, std::ios::app);
      o << std::setprecision(17) << it_counter << ' ' << norms[0] << ' ' << norms[1] << ' ' << norms[2] << '\n';
    }
    if (c.flag("ToWriteSolvedDataToBinaryFile")) {
      write_bin(path("ContinuityEquation_stationary.bin"), cs.data(), Na);
      write_bin(path("ContinuityEquation.bin"), cf.data(), Na);
      if (c.flag("ToWriteSolvedDataToNewFiles")) {
        write_bin(path("ContinuityEquation_stationary." + num(it) + ".bin"), cs.data(), Na);
        write_bin(path("ContinuityEquation." + num(it) + ".bin"), cf.data(), Na);
      }
    } else {
      write_text(path("ContinuityEquation_stationary.txt"), cs.data(), Na);
      write_text(path("ContinuityEquation.txt"), cf.data(), Na);
      if (c.flag("ToWriteSolvedDataToNewFiles")) {
        write_text(path("ContinuityEquation_stationary." + num(it) + ".txt"), cs.data(), Na);
        write_text(path("ContinuityEquation." + num(it) + ".txt"), cf.data(), Na);
      }
    }
  }

  // WriteTempDataToFile, S:5838-5942
  void write_tmp() {
    std::ofstream o;
    if (tmp_first && !c.flag("ToReadSolvedDataFromFile")) o.open(path("Tmp.txt"));
    else o.open(path("Tmp.txt"), std::ios::out | std::ios::app);
    tmp_first = false;
    if (!o) die("cannot write Tmp.txt");
    o << "it = " << it << "\tit_counter = " << it_counter << "\tIter = " << Iter << "\tIter_p_c = " << Iter_p_c
      << "\tIter_T = " << c.Iter_T << std::endl;
    unsigned sum = 0;
    for (unsigned k = 0; k < Iter; k++) {
      o << "\t" << k << "\t" << max_Iter_p_c[k] << std::setiosflags(std::ios::scientific) << std::setprecision(15)
        << "\t" << max_res_p0[k] << "\t" << max_res_p[k] << "\t" << max_res_T0[k] << "\t" << max_res_T[k] << std::endl;
      sum += max_Iter_p_c[k];
    }
    o << "\t\t" << sum << std::endl;
    o << std::setiosflags(std::ios::scientific) << std::setprecision(15) << "\t" << max_res_u0 << "\t" << max_res_u
      << std::endl
      << "\t" << max_res_v0 << "\t" << max_res_v << std::endl;
  }

  // WriteDragCoefficient, S:8147-8430: for every body marked
  // ToSolveDragCoefficientForThisBody (circles first), the face sums from the
  // GPU, the nondimensional factors of the case and one line of CD.<k>.txt.
  // As in the reference, only the first file written in a run is truncated
  // and given the header line.
  bool cd_first = true;
  void write_drag() {
    const size_t nc = c.circles.size(), nb = nc + c.polys.size();
    const unsigned type = (unsigned)c.get("TypeOfNonDimensiolization");
    const bool restart = c.flag("ToReadSolvedDataFromFile");
    double CD_fr_x_last = 0, CD_p_x_last = 0;  // the last body written (S:8149-8150, S:8337-8338)
    for (size_t k = 0; k < nb; k++) {
      if (!(k < nc ? c.circles[k].drag : c.polys[k - nc].drag)) continue;
      double s[12];
      ck(sts_drag_sums(ctx, (int)k, s), "sts_drag_sums");
      double c_CD_p = 0, c_CD_fr = 0;
      const double rho_b = p_BC_xb / c.get("T_BC_xb");
      if (type == 1) {
        c_CD_p = 1.0 / (u_given_xb_in_it * u_given_xb_in_it * rho_b);
        c_CD_fr = 2.0 / (c.get("Re") * u_given_xb_in_it);
      } else if (type == 2) {
        c_CD_p = 2.0 / (c.get("gamma1") * c.get("Ma") * c.get("Ma") * rho_b);
        c_CD_fr = 2.0 / (c.get("Re") * u_given_xb_in_it);
      } else if (type == 3) {
        c_CD_p = 1.0 / (u_given_xb_in_it * u_given_xb_in_it * rho_b);
        c_CD_fr = (5.0 / 8.0) * sqrt(M_PI) * c.get("Kn") / (u_given_xb_in_it * u_given_xb_in_it * rho_b);
      }
      const double fr_bottom = c_CD_fr * s[0], p_bottom = c_CD_p * s[1], fr_front = c_CD_fr * s[3],
                   p_front = c_CD_p * s[4], fr_top = c_CD_fr * s[6], p_top = c_CD_p * s[7],
                   fr_behind = c_CD_fr * s[9], p_behind = c_CD_p * s[10];
      const double S_bottom = s[2], S_front = s[5], S_top = s[8], S_behind = s[11];
      const double CD_fr_x = (fr_top / S_top) + (fr_bottom / S_bottom);
      const double CD_p_x = (p_front / S_front) - (p_behind / S_behind);
      CD_fr_x_last = CD_fr_x;
      CD_p_x_last = CD_p_x;
      const std::string name = path("CD." + std::to_string(k) + ".txt");
      std::ofstream o;
      if (cd_first && !restart) o.open(name);
      else o.open(name, std::ios::out | std::ios::app);
      if (!o) {
        std::cout << "The program can not open file " << "CD." << k << ".txt" << " to write data." << std::endl;
        continue;
      }
      const int w = 30;
      if (!restart && cd_first) {
        for (const char *h : {"//it (number of step by time) 00", "ht (step by time) 01", "CD_fr_x (CD from friction) 02",
                              "CD_p_x (CD from pressure) 03", "S_for_CD_bottom 04", "S_for_CD_front 05",
                              "S_for_CD_top 06", "S_for_CD_behind 07", "CD_fr_bottom 08", "CD_p_bottom 09",
                              "CD_fr_front 10", "CD_p_front 11", "CD_fr_top 12", "CD_p_top 13", "CD_fr_behind 14",
                              "CD_p_behind 15"})
          o << std::setw(w) << h;
        o << std::endl;
      }
      o << std::setw(w) << it;
      o << std::setiosflags(std::ios::scientific) << std::setprecision(10);
      for (double x : {ht, CD_fr_x, CD_p_x, S_bottom, S_front, S_top, S_behind, fr_bottom, p_bottom, fr_front,
                       p_front, fr_top, p_top, fr_behind, p_behind})
        o << std::setw(w) << x;
      o << std::endl;
      cd_first = false;
    }
    if ((int)c.get("WriteDataInShortFormat_kind") == 1) write_short(CD_fr_x_last, CD_p_x_last);
    if (c.flag("Pressure_BC")) write_p_u_xb();
  }

  // DataInShortFormat.txt, S:8438-8500: Re, Kn, CD, CD_p, u_gas_nd, the mass
  // flows through column 0 and column Nx_u[1] (sum over j of rho u hy on fluid
  // cells), p_BC_xb, p_BC_xe; rewritten at every drag-coefficient write.  A
  // node past the arrays (Nx_u[1] = Nx in a periodic case, last row) counts
  // as body.
  void write_short(double CD_fr_x, double CD_p_x) {
    std::vector<double> rho(Na), u(Na);
    ck(sts_get_field(ctx, STS_RHO, rho.data()), "sts_get_field");
    ck(sts_get_field(ctx, STS_U, u.data()), "sts_get_field");
    auto flow = [&](unsigned i) {
      double q = 0;
      for (int j = 0; j < Ny; j++) {
        const size_t node = (size_t)i + (size_t)j * (size_t)Nx;
        if (node < Na && mask[node]) q += rho[node] * u[node] * c.hy[j];
      }
      return q;
    };
    const double q_xb = flow(0), q_xe = flow((unsigned)c.Nx_u[1]);
    std::ofstream o(path("DataInShortFormat.txt"));
    if (!o) {
      std::cout << "The program can not open file DataInShortFormat.txt to write data." << std::endl;
      return;
    }
    const int w = 30;
    o << std::setiosflags(std::ios::scientific) << std::setprecision(15);
    for (double x : {c.get("Re"), c.get("Kn"), CD_p_x + CD_fr_x, CD_p_x, c.get("u_gas_nd"), q_xb, q_xe, p_BC_xb,
                     c.get("p_BC_xe")})
      o << std::setw(w) << x;
    o << std::endl;
  }

  // p_and_u_xb_it.txt, S:8504-8565: the inflow pressure and velocities at every
  // drag-coefficient write; the first write of a run (not a restart) truncates
  // the file and writes the header.
  bool pu_first = true;
  void write_p_u_xb() {
    const bool restart = c.flag("ToReadSolvedDataFromFile");
    std::ofstream o;
    if (pu_first && !restart) o.open(path("p_and_u_xb_it.txt"));
    else o.open(path("p_and_u_xb_it.txt"), std::ios::out | std::ios::app);
    if (!o) {
      std::cout << "The program can not open file p_and_u_xb_it.txt to write data." << std::endl;
      return;
    }
    const int w = 30;
    if (!restart && pu_first)
      for (const char *h : {"//it (number of step by time)", "ht (step by time)", "p_BC_xb", "u_given_xb_in_it",
                            "u_given_xb_in_it_max", "u_given_xb_in_it_mean", "p_correction"})
        o << std::setw(w) << h;
    if (!restart && pu_first) o << std::endl;
    o << std::setw(w) << it;
    o << std::setiosflags(std::ios::scientific) << std::setprecision(15);
    for (double x : {ht, p_BC_xb, u_given_xb_in_it, u_xb_max, u_xb_mean, c.get("p_correction")}) o << std::setw(w) << x;
    o << std::endl;
    pu_first = false;
  }

  // WriteIterationsInLoopsToFile, S:5945-6006
  void write_iters() {
    std::ofstream o;
    if (iters_first) {
      if (c.flag("ToReadSolvedDataFromFile")) {
        o.open(path("Iterations_in_loops.txt"), std::ios::out | std::ios::app);
      } else {
        o.open(path("Iterations_in_loops.txt"));
        o << "it" << "\tit_counter" << "\tIter" << "\tsum_max_Iter_p_c" << std::endl;
      }
    } else {
      o.open(path("Iterations_in_loops.txt"), std::ios::out | std::ios::app);
    }
    iters_first = false;
    unsigned sum = 0;
    for (unsigned k = 0; k < Iter; k++) sum += max_Iter_p_c[k];
    o << it << "\t" << it_counter << "\t\t" << Iter << "\t" << sum << std::endl;
  }

  // S:2013-2071 / S:4347-4361 for the current inflow bookkeeping
  void scalars() {
    const unsigned type = (unsigned)c.get("TypeOfNonDimensiolization");
    double cT = c.cT, cPch = c.cPch;
    if (type == 1 || type == 2) {
      cT = (1.0 / c.get("Re")) * (u_given_xb_in_it * p_BC_xb / c.get("T_BC_xb"));
      if (type == 2) cPch = u_given_xb_in_it * u_given_xb_in_it / (c.get("gamma1") * c.get("Ma") * c.get("Ma"));
      else cPch = 0.5;
    } else {
      cT = c.get("Kn") * 5.0 * sqrt(M_PI) / 16.0;
      cPch = 0.5;
    }
    ck(sts_set_scalars(ctx, cT, cPch), "sts_set_scalars");
  }

  // BoundaryConditions_Pressure_begin_iteration_for_it, S:7430-7647.  The
  // largest u and the hy-weighted mean u over the fluid cells of column
  // Nx_u[0] (S:7452-7486) are formed when something reads them:
  // p_and_u_xb_it.txt (Pressure_BC), or p_BC_xb_correction_method 1 / 2,
  // which take them as the inflow velocity of the step's bookkeeping (cT of
  // types 1 and 2, the drag coefficients).  With Pressure_BC and
  // Pressure_ratio_correct, the pressure at one box end follows that
  // velocity (pressure_ratio_correct).
  double u_xb_max = 0, u_xb_mean = 0;
  void pressure_begin(double probe) {
    if (c.flag("Periodic_boundary_conditions_about_OX")) return;  // S:7436
    const unsigned method = (unsigned)c.get("p_BC_xb_correction_method");
    if (c.flag("Pressure_BC") || method != 0) {
      std::vector<double> u(Na);
      ck(sts_get_field(ctx, STS_U, u.data()), "sts_get_field");
      const unsigned i = (unsigned)c.Nx_u[0];
      u_xb_max = u[0];
      double mean = 0, h = 0;
      for (int j = 0; j < Ny; j++) {
        const size_t node = (size_t)i + (size_t)j * (size_t)Nx;
        if (!mask[node]) continue;
        u_xb_max = (u_xb_max > u[node]) ? u_xb_max : u[node];
        mean += u[node] * c.hy[j];
        h += c.hy[j];
      }
      u_xb_mean = mean / h;
    }
    if (c.flag("GivenReKn")) c.at("u_given_xb") = c.get("Re") * c.get("Kn") * sqrt(15.0 * M_PI / 128.0);
    if (method == 0) {
      u_given_xb_in_it = c.get("u_given_xb");
      if (c.flag("Given_velocity_on_xb")) p_BC_xb = probe;
    } else if (method == 1) {
      u_given_xb_in_it = u_xb_max;
    } else if (method == 2) {
      u_given_xb_in_it = u_xb_mean;
    }
    if (c.flag("Pressure_BC") && c.flag("Pressure_ratio_correct")) pressure_ratio_correct();
  }

  // S:7525-7640: p_correction adapts to the last correction's velocity
  // residuals (p_correction_auto), the inflow (correct_p_BC_xb) or outflow
  // pressure moves by p_correction times the squared miss of u_given_xb when
  // the inflow velocity is outside u_given_xb (1 +- u_given_xb_error), and
  // that end's fluid columns take the pressure and its density.
  void pressure_ratio_correct() {
    double &pc = c.at("p_correction");
    if (c.flag("p_correction_auto")) {
      const double r = sqrt(max_res_u * max_res_u + max_res_v * max_res_v);
      double t;
      if (Iter_pr1 <= Iter_pr2 && Iter_pr1 < (N_I - 1)) t = pc * (1.0 + r);
      else t = pc * (1.0 - r);
      if (t < c.get("p_correction_min")) t = c.get("p_correction_min");
      else if (c.get("p_correction_max") < t) t = c.get("p_correction_max");
      pc = t;
    }
    const double ug = c.get("u_given_xb"), err = c.get("u_given_xb_error"), ui = u_given_xb_in_it;
    const bool at_xb = c.flag("correct_p_BC_xb");
    double &pe = c.at("p_BC_xe");
    if (ui < (ug * (1.0 - err))) {
      if (at_xb) {
        p_BC_xb = p_BC_xb * (1.0 + pc * (ug - ui) * (ug - ui));
      } else {
        const double t = pe * (1.0 - pc * (ug - ui) * (ug - ui));
        pe = (t < 0.0) ? 0.0 : t;
      }
    } else if ((ug * (1.0 + err)) < ui) {
      if (at_xb) {
        const double t = p_BC_xb * (1.0 - pc * (ug - ui) * (ug - ui));
        p_BC_xb = (t > pe) ? t : pe;
      } else {
        const double t = p_BC_xb * (1.0 + pc * (ug - ui) * (ug - ui));
        pe = (t > p_BC_xb) ? p_BC_xb : t;
      }
    }
    c.at("p_BC_xb") = p_BC_xb;
    std::vector<double> p, rho;
    get(STS_P, p);
    get(STS_RHO, rho);
    const int i0 = at_xb ? 0 : c.Nx_o[1], i1 = at_xb ? c.Nx_o[0] : Nx;
    const double pv = at_xb ? p_BC_xb : pe, rv = pv / c.get(at_xb ? "T_BC_xb" : "T_BC_xe");
    for (int j = 0; j < Ny; j++)
      for (int i = i0; i < i1; i++) {
        const size_t node = (size_t)i + (size_t)j * Nx;
        if (!mask[node]) continue;
        p[node] = pv;
        rho[node] = rv;
      }
    set(STS_P, p);
    set(STS_RHO, rho);
  }

  void run() {
    time_solve_start = time(nullptr);
    Nx = c.Nx;
    Ny = c.Ny;
    Na = (size_t)Nx * (size_t)Ny;
    mask.resize(Na);
    for (size_t k = 0; k < Na; k++) mask[k] = c.vol[k] == 0;
    ht = c.get("ht");
    N_I = (unsigned)c.get("N_I");
    N_I_p_c = (unsigned)c.get("N_I_p_c");
    if (N_I < 1) N_I = 1;
    if (N_I_p_c < 1 || N_I_p_c > 62) die("N_I_p_c must be 1..62");
    p_BC_xb = c.get("p_BC_xb");
    u_given_xb_in_it = c.u_given_xb_in_it;
    time_solve_stop = (double)time_solve_start + c.get("max_time_for_solve") * 3600.0;
    solve_is_finished = false;
    const double Nt = c.get("Nt");

    // mesh files written by the serial build, S:2249-2259
    write_text(path("x_f.txt"), c.x_f.data(), c.x_f.size());
    write_text(path("hx.txt"), c.hx.data(), c.hx.size());
    write_text(path("x_v.txt"), c.x_v.data(), c.x_v.size());
    write_text(path("y_f.txt"), c.y_f.data(), c.y_f.size());
    write_text(path("hy.txt"), c.hy.data(), c.hy.size());
    write_text(path("y_v.txt"), c.y_v.data(), c.y_v.size());
    if (!c.flag("ToReadSolvedDataFromFile")) write_config();

    c.fill_grid(grid);
    ck(sts_create(&grid, mixed ? 32 : 64, &ctx), "sts_create");
    max_res_p.assign(N_I, 0.0);
    max_res_T.assign(N_I, 0.0);
    max_res_p0.assign(N_I, 0.0);
    max_res_T0.assign(N_I, 0.0);
    max_Iter_p_c.assign(N_I, 0u);

    std::vector<double> u, v, p, T, rho;
    host_fields(u, v, p, T, rho);
    if (c.flag("ToReadSolvedDataFromFile")) {  // ReadSolvedDataFromFile, S:5734-5835
      const bool bin = c.flag("ToReadSolvedDataFromBinaryFile");
      const std::string ext = bin ? ".bin" : ".txt";
      read_field(path("u" + ext), u.data(), Na, bin);
      read_field(path("v" + ext), v.data(), Na, bin);
      read_field(path("p" + ext), p.data(), Na, bin);
      read_field(path("Temper" + ext), T.data(), Na, bin);
      for (size_t k = 0; k < Na; k++) rho[k] = mask[k] ? p[k] / T[k] : 0;
      set(STS_U, u);
      set(STS_V, v);
      set(STS_P, p);
      set(STS_T, T);
      set(STS_RHO, rho);
      ck(sts_rho_in_middle_points(ctx), "rho_in_middle_points");
      std::ifstream itf(path("it_begin.txt"));  // Read_it_FromFile, S:6762-6787
      if (!(itf >> it_begin)) die("cannot read it_begin.txt");
      it = it_begin;
    } else {
      set(STS_U, u);
      set(STS_V, v);
      set(STS_P, p);
      set(STS_T, T);
      set(STS_RHO, rho);
      ck(sts_apply_bodies(ctx), "apply_bodies");  // S:2507
      if (c.flag("Pressure_BC")) pressure_columns(true);
      const bool exact = c.flag("ToStartFromExactSolutionForChannelFlowWithPressureBC");
      if (exact && c.flag("Pressure_BC")) exact_start_pressure_bc();             // S:2566-2689
      if (exact && c.flag("Given_velocity_on_xb")) exact_start_given_velocity();  // S:2694-2806
      if (exact && c.flag("Given_velocity_on_xe")) exact_start_given_outflow();   // S:2808-2830
      ck(sts_save_previous(ctx), "save_previous");  // S:2838-2845
      ck(sts_rho_in_middle_points(ctx), "rho_in_middle_points");
      ck(sts_apply_bodies(ctx), "apply_bodies");
      ck(sts_boundary_conditions(ctx, STS_BC_VELOCITIES, nullptr), "boundary_conditions");
      ck(sts_save_previous(ctx), "save_previous");
      ck(sts_rho_in_middle_points(ctx), "rho_in_middle_points");
      it_begin = 0;
      it = it_begin;
      write_solved();
    }
    // S:2870-2930: coefficient arrays start at zero
    const std::vector<double> zero(Na, 0.0);
    for (int f : {STS_D_P_C_12, STS_D_P_C_34, STS_APX, STS_APY, STS_BPX, STS_BPY, STS_GAMMA_YF, STS_GAMMA_XF,
                  STS_FX, STS_FY, STS_U_PSEUDO, STS_V_PSEUDO, STS_D_UX, STS_D_UY, STS_D_VX, STS_D_VY, STS_DTX,
                  STS_DTY, STS_SCT1, STS_RHO_U, STS_RHO_V})
      set(f, zero);
    Iter = 0;
    Iter_pr1 = Iter;
    Iter_pr2 = Iter_pr1;
    if (c.flag("Pressure_BC") && !c.flag("ToReadSolvedDataFromFile")) pressure_columns(false);
    if (c.flag("ToContinueFromInterpolation")) {  // S:2982-2991: the wall conditions on the state read in
      ck(sts_apply_bodies(ctx), "apply_bodies");
      ck(sts_boundary_conditions(ctx, STS_BC_SLIP_U | STS_BC_SLIP_V | STS_BC_JUMP, nullptr), "bc");
      ck(sts_jump_overwrite(ctx), "jump_overwrite");
    }
    ck(sts_properties(ctx), "properties");  // S:2948-2953
    ck(sts_rho_in_middle_points(ctx), "rho_in_middle_points");
    it = it_begin;
    it_counter = 0;
    countinue_time_step = true;
    ck(sts_save_previous(ctx), "save_previous");  // S:2975-2983
    scalars();                                    // S:3043-3057
    ck(sts_rho_in_middle_points(ctx), "rho_in_middle_points");
    double probe = 0;
    ck(sts_boundary_conditions(ctx, STS_BC_SLIP_U | STS_BC_SLIP_V | STS_BC_VELOCITIES, &probe), "bc");
    pressure_begin(probe);
    ck(sts_boundary_conditions(ctx, STS_BC_JUMP | STS_BC_TEMPERATURE | STS_BC_PRESSURE, nullptr), "bc");
    ck(sts_fluxes_diffusion(ctx), "fluxes_diffusion");

    // loop 2 on the device when nothing the host computes changes inside it:
    // cT is constant (type 3) and no box-end pressure follows the inflow
    // velocity from one outer iteration to the next
    const bool ratio = c.flag("Pressure_BC") && c.flag("Pressure_ratio_correct") &&
                       !c.flag("Periodic_boundary_conditions_about_OX");
    const bool fused = (unsigned)c.get("TypeOfNonDimensiolization") == 3 && !ratio;
    auto t_timed = std::chrono::steady_clock::now();
    long long outer_at_start = 0;
    if (warmup_steps == 0) start_timing(t_timed, outer_at_start);
    int save_ctr = 0, temp_ctr = 0, drag_ctr = 0;
    const int Nt_drag = (int)c.get("Nt_save_DragCoefficient");
    const int Nt_save = (int)c.get("Nt_save_solved_data"), Nt_temp = (int)c.get("Nt_save_temp_data");
    const auto t0 = std::chrono::steady_clock::now();
    time_solve_start = time(nullptr);
    const double eV = c.get("MaxError_Velocities"), ep = c.get("MaxError_p_c"), eT = c.get("MaxError_T");
    const double es = c.get("MaxError_to_stop_program");
    if (mc.N_I32 < 0) mc.N_I32 = (int)N_I;
    std::vector<double> lg(6 * (size_t)N_I), lg32(6 * ((size_t)N_I + (size_t)mc.N_I32));
    do {  // loop 1
      if (!fused) ck(sts_save_previous(ctx), "save_previous");
      for (unsigned k = 0; k < Iter; k++) {
        max_res_p[k] = max_res_T[k] = max_res_p0[k] = max_res_T0[k] = 0;
        max_Iter_p_c[k] = 0;
      }
      it += ht;
      it_counter++;
      Iter_pr2 = Iter_pr1;
      Iter_pr1 = Iter;
      Iter = 0;
      continue_iter = true;
      if (fused) {
        // loop 2 with the test on the GPU (same test, same order of work)
        sts_loop_control lc{(int)N_I, (int)N_I_p_c, eV, ep, eT, 0};
        int iters = 0;
        if (mixed) {
          // FP32 phase first; the FP64 phase's iterations are the reference's Iter
          int n32 = 0;
          ck(sts_time_step_mixed(ctx, &lc, &mc, &n32, &iters, lg32.data(), &probe), "time_step_mixed");
          std::copy(lg32.begin() + 6 * (size_t)n32, lg32.begin() + 6 * (size_t)(n32 + iters), lg.begin());
          outer_total32 += n32;
        } else {
          ck(sts_time_step(ctx, &lc, &iters, lg.data(), &probe), "time_step");
        }
        outer_total += iters;
        Iter_p_c = N_I_p_c;
        for (int k = 0; k < iters; k++) {
          max_res_p[k] = lg[6 * k + 2];
          max_res_T[k] = lg[6 * k + 3];
          max_res_p0[k] = lg[6 * k + 4];
          max_res_T0[k] = lg[6 * k + 5];
          max_Iter_p_c[k] = Iter_p_c;
        }
        Iter = (unsigned)iters;
        max_res_u = lg[6 * (iters - 1) + 0];
        max_res_v = lg[6 * (iters - 1) + 1];
        max_res_u0 = lg[0];
        max_res_v0 = lg[1];
        countinue_time_step =
            (it_counter < Nt && (double)time(nullptr) < time_solve_stop &&
             ((!(Iter <= 1 && Iter_p_c <= 2)) || es < max_res_u || es < max_res_v || es < max_res_p0[Iter - 1] ||
              es < max_res_T0[Iter - 1]));
        pressure_begin(probe);
        if (c.flag("ToSolveContinuityEquation")) continuity();
      } else {
        do {  // loop 2
          double r[4], r0[2];
          // cT of the diffusion terms is recomputed after the correction
          // (S:4347) from the inflow values of the previous iteration.
          ck(sts_assemble(ctx), "assemble");
          std::vector<double> rs(2 * N_I_p_c);
          ck(sts_pt_jacobi(ctx, (int)N_I_p_c, rs.data()), "pt_jacobi");
          double ruv[2];
          ck(sts_correct(ctx, ruv), "correct");
          max_res_u = ruv[0];
          max_res_v = ruv[1];
          scalars();
          ck(sts_boundary_conditions(ctx, STS_BC_SLIP_U | STS_BC_SLIP_V | STS_BC_VELOCITIES, &probe), "bc");
          pressure_begin(probe);  // S:4509, between the velocity and the temperature-jump conditions
          ck(sts_boundary_conditions(ctx, STS_BC_JUMP | STS_BC_TEMPERATURE | STS_BC_PRESSURE, nullptr), "bc");
          ck(sts_fluxes_diffusion(ctx), "fluxes_diffusion");
          r[0] = ruv[0];
          r[1] = ruv[1];
          r[2] = rs[2 * (N_I_p_c - 1)];
          r[3] = rs[2 * (N_I_p_c - 1) + 1];
          r0[0] = rs[0];
          r0[1] = rs[1];
          outer_total++;
          Iter_p_c = N_I_p_c;
          max_res_u = r[0];
          max_res_v = r[1];
          max_res_p[Iter] = r[2];
          max_res_T[Iter] = r[3];
          max_res_p0[Iter] = r0[0];
          max_res_T0[Iter] = r0[1];
          max_Iter_p_c[Iter] = Iter_p_c;
          Iter++;
          if (Iter == 1) {
            max_res_u0 = max_res_u;
            max_res_v0 = max_res_v;
          }
          continue_iter = (Iter < N_I && (eV < max_res_u || eV < max_res_v || ep < max_res_p[Iter - 1] ||
                                          eT < max_res_T[Iter - 1]));
          countinue_time_step =
              (it_counter < Nt && (double)time(nullptr) < time_solve_stop &&
               ((!(Iter <= 1 && Iter_p_c <= 2)) || es < max_res_u || es < max_res_v ||
                es < max_res_p0[Iter - 1] || es < max_res_T0[Iter - 1]));
          if (c.flag("ToSolveContinuityEquation")) continuity();
        } while (continue_iter);
      }
      if (c.flag("use_interpolation_for_u_BC_by_time") &&
          !c.flag("Periodic_boundary_conditions_about_OX")) {  // S:8102-8144 (S:8105)
        const double tb = c.get("t_BC_b"), te = c.get("t_BC_e"), ub = c.get("u_BC_xb_t_BC_b"),
                     ue = c.get("u_BC_xb_t_BC_e");
        double ug;
        if (it <= tb) ug = ub;
        else if (tb < it && it <= te) ug = ((ue - ub) / (te - tb)) * (it - tb) + ub;
        else ug = ue;
        c.at("u_gas_nd") = ug;
        c.at("u_given_xb") = ug;
        ck(sts_set_inflow_u(ctx, ug), "set_inflow_u");
      }
      if (++save_ctr >= Nt_save) {
        write_solved();
        if (c.flag("ToSolveContinuityEquation")) continuity();
        save_ctr = 0;
      }
      if (++drag_ctr >= Nt_drag) {  // S:4653-4661
        write_drag();
        drag_ctr = 0;
      }
      if (++temp_ctr >= Nt_temp) {
        write_tmp();
        write_iters();
        temp_ctr = 0;
      }
      if (it_counter == warmup_steps) start_timing(t_timed, outer_at_start);
    } while (countinue_time_step);
    const auto t_end = std::chrono::steady_clock::now();
    const double loop_s = std::chrono::duration<double>(t_end - t0).count();
    timed_seconds = std::chrono::duration<double>(t_end - t_timed).count();
    timed_steps = it_counter - warmup_steps;
    timed_outer = outer_total - outer_at_start;
    for (size_t q = 0; q < time_kernels.size(); q++)
      ck(sts_kernel_time(ctx, time_kernels[q], &kernel_launches[q], &kernel_ms[q]), "sts_kernel_time");
    solve_is_finished = Iter <= 1 && Iter_p_c <= 1 && max_res_u < c.get("MaxError_to_stop_program") &&
                        max_res_v < c.get("MaxError_to_stop_program") &&
                        max_res_p[Iter - 1] < c.get("MaxError_to_stop_program") &&
                        max_res_T[Iter - 1] < c.get("MaxError_to_stop_program");
    c.at("max_time_for_solve") = (double)(time(nullptr) - time_solve_start) / 3600.0;
    write_solved();
    if (c.flag("ToSolveContinuityEquation")) continuity();
    write_drag();  // S:4713
    write_tmp();
    loop_seconds = loop_s;
    ck(sts_destroy(ctx), "sts_destroy");
  }
  double loop_seconds = 0;
  // mixed-precision time steps (sts_time_step_mixed) when set
  bool mixed = false;
  sts_mixed_control mc{0, 1e-6, 1e-6, 1e-6, 4, 0.7, 0, 0, 0, 0, 0, 0};
  long long outer_total32 = 0, outer32_at_start = 0;
  // timing of the steps after the warm-up ones (benchmarking)
  unsigned warmup_steps = 0;
  std::vector<int> time_kernels;  // kernel ids timed with HIP events
  std::vector<double> kernel_ms;
  std::vector<long long> kernel_launches;
  double timed_seconds = 0;
  unsigned timed_steps = 0;
  long long timed_outer = 0;

  void start_timing(std::chrono::steady_clock::time_point &t, long long &outer0) {
    ck(sts_time_kernel(ctx, -1), "sts_time_kernel");
    for (int k : time_kernels) ck(sts_time_kernel(ctx, k), "sts_time_kernel");
    kernel_ms.assign(time_kernels.size(), 0.0);
    kernel_launches.assign(time_kernels.size(), 0);
    t = std::chrono::steady_clock::now();
    outer0 = outer_total;
    outer32_at_start = outer_total32;
  }
};

}  // namespace

int main(int argc, char **argv) {
  std::string dir = ".", timing, norms_file;
  unsigned warmup = 0;
  std::vector<int> kernels;
  const char *knames[2 * STS_K_COUNT] = {
      "assemble",   "sweep",      "correct", "wall",      "edges",  "flux_diff", "assemble_inl",   "aa_store",
      "aa_mix",     "assemble32", "sweep32", "correct32", "wall32", "edges32",   "flux_diff32", "assemble_inl32",
      "aa_store32", "aa_mix32"};
  bool mixed = false;
  sts_mixed_control mc{0, 1e-6, 1e-6, 1e-6, 4, 0.7, 0, 0, 0, 0, 0, 0};
  int n32 = -1;
  for (int a = 1; a < argc; a++) {
    std::string s = argv[a];
    if (s == "--timing" && a + 1 < argc) timing = argv[++a];
    else if (s == "--warmup" && a + 1 < argc) warmup = (unsigned)std::atoi(argv[++a]);
    else if (s == "--time-kernel" && a + 1 < argc) {
      std::stringstream list(argv[++a]);
      std::string k;
      while (std::getline(list, k, ',')) {
        int id = -1;
        for (int q = 0; q < 2 * STS_K_COUNT; q++)
          if (k == knames[q]) id = q;
        if (id < 0) die("unknown kernel " + k);
        kernels.push_back(id);
      }
    } else if (s == "--mixed") {
      mixed = true;
    } else if (s == "--n32" && a + 1 < argc) {
      n32 = std::atoi(argv[++a]);
    } else if (s == "--extrapolate" && a + 1 < argc) {
      mc.extrapolate = std::atoi(argv[++a]);
    } else if (s == "--fp32-above" && a + 1 < argc) {
      mc.fp32_above = std::atoi(argv[++a]);
    } else if (s == "--guard" && a + 1 < argc) {
      mc.guard = std::atoi(argv[++a]);
    } else if (s == "--anderson" && a + 1 < argc) {
      mc.anderson = std::atoi(argv[++a]);
    } else if (s == "--anderson32" && a + 1 < argc) {
      mc.anderson32 = std::atoi(argv[++a]);
    } else if (s == "--anderson-every" && a + 1 < argc) {
      mc.anderson_every = std::atoi(argv[++a]);
    } else if (s == "--continuity-norms" && a + 1 < argc) {
      norms_file = argv[++a];
    } else if (s == "--tol32" && a + 1 < argc) {
      mc.tol32_uv = mc.tol32_p = mc.tol32_T = std::atof(argv[++a]);
    } else if (s == "--stall" && a + 2 < argc) {
      mc.stall_w = std::atoi(argv[++a]);
      mc.stall_ratio = std::atof(argv[++a]);
    } else if (s == "-h" || s == "--help") {
      std::cout << "usage: sts_gpu [case_dir] [--timing file.json] [--warmup steps] [--time-kernel "
                   "name[,name...]] (assemble, assemble_inl, sweep, correct, wall, edges, flux_diff, aa_store, "
                   "aa_mix; FP32 instances with suffix 32) [--extrapolate degree] [--mixed [--n32 cap] "
                   "[--tol32 tol] [--stall window ratio] [--fp32-above iterations]] [--guard max_pause] "
                   "[--anderson depth [--anderson-every iterations]] [--anderson32 depth] [--continuity-norms file]"
                << std::endl;
      return 0;
    } else dir = s;
  }
  Driver d;
  d.dir = dir;
  d.norms_file = norms_file;
  d.warmup_steps = warmup;
  d.time_kernels = kernels;
  d.kernel_ms.assign(kernels.size(), 0.0);
  d.kernel_launches.assign(kernels.size(), 0);
  d.mixed = mixed;
  mc.N_I32 = mixed ? n32 : 0;
  d.mixed = mixed || mc.extrapolate > 0 || mc.anderson > 0;
  d.mc = mc;
  std::string err = sts::load_case(dir, d.c);
  if (!err.empty()) die(err);
  d.run();
  if (!timing.empty()) {
    std::ofstream o(timing);
    o << std::setprecision(17) << "{\"Nx\": " << d.Nx << ", \"Ny\": " << d.Ny << ", \"steps\": " << d.it_counter
      << ", \"outer_iterations\": " << d.outer_total << ", \"loop_seconds\": " << d.loop_seconds
      << ", \"warmup_steps\": " << d.warmup_steps << ", \"timed_steps\": " << d.timed_steps
      << ", \"timed_outer_iterations\": " << d.timed_outer << ", \"timed_seconds\": " << d.timed_seconds
      << ", \"kernels\": {";
    for (size_t q = 0; q < kernels.size(); q++)
      o << (q ? ", " : "") << "\"" << knames[kernels[q]] << "\": {\"launches\": " << d.kernel_launches[q]
        << ", \"ms\": " << d.kernel_ms[q] << "}";
    o << "}, \"mixed\": " << (mixed ? 1 : 0)
      << ", \"fp32_iterations\": " << d.outer_total32
      << ", \"timed_fp32_iterations\": " << d.outer_total32 - d.outer32_at_start << "}" << std::endl;
  }
  return 0;
}
