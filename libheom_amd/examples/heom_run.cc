// heom_run: a reference-style C++ caller of the drop-in headers.
//
// Reads one model from a text file, fills the qme_base members exactly as a
// libheom caller does (n_level, n_level_2, H, alloc_noises, V/len_gamma/gamma/
// phi_0/sigma/S/A/s_delta, then set_param), propagates with qme_solver::solve
// and prints rho_S at every output time (what the callback sees).
//
// Input (whitespace separated, complex numbers as "re im"):
//   form(0 hilb|1 liou|2 ado|3 redfield_hilb|4 redfield_liou) N max_depth n_noise
//   H[N*N]
//   per noise: len V[N*N] gamma[len*len] phi_0[len] sigma[len] S[len*len] A[len*len] s_delta
//   solver(0 rk4|1 rkdp|2 lsrk4) dt atol rtol n_t t_list[n_t]
//   rho0[main_size]
// Output: one line per output time: t followed by rho_S (N*N complex, re im).
//
// heom_run model.txt calc_diff: instead of solving, print one line holding
//   qme_solver::calc_diff(drho, rho0), i.e. drho = -D(rho0) (main_size complex).
// heom_run model.txt reparam: solve, then halve H, call set_param again on the
//   same qme object and solve again with the same qme_solver (both runs'
//   output lines are printed, the second after a line "reparam").
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "libheom_amd.hpp"

using namespace libheom;
using dtype = complex128;
using engine = hip;

static dtype rd(std::istream& in) {
  double re, im;
  in >> re >> im;
  return dtype(re, im);
}

static void fill(std::istream& in, lil_matrix<dtype>& m, int rows, int cols) {
  m.set_shape(rows, cols);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      dtype v = rd(in);
      if (v != dtype(0.0)) m.push(i, j, v);
    }
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: heom_run model.txt [calc_diff|reparam]\n");
    return 2;
  }
  const std::string mode = argc > 2 ? argv[2] : "";
  std::ifstream in(argv[1]);
  int form, N, depth, n_noise;
  in >> form >> N >> depth >> n_noise;
  std::unique_ptr<qme_base<dtype, row_major, engine>> qme;
  if (form == 0)
    qme.reset(new heom_hilb<dynamic, dtype, dense_matrix, row_major, engine>(depth, 1, 1));
  else if (form == 1)
    qme.reset(new heom_liou<dynamic, dtype, sparse_matrix, row_major, row_major, engine>(depth, 1, 1));
  else if (form == 2)
    qme.reset(new heom_ado<dynamic, dtype, sparse_matrix, row_major, row_major, engine>(depth, 1, 1));
  else if (form == 3)
    qme.reset(new redfield_hilb<dynamic, dtype, dense_matrix, row_major, engine>());
  else
    qme.reset(new redfield_liou<dynamic, dtype, sparse_matrix, row_major, row_major, engine>());
  qme->n_level = N;
  qme->n_level_2 = N * N;
  fill(in, qme->H, N, N);
  qme->alloc_noises(n_noise);
  for (int u = 0; u < n_noise; ++u) {
    int len;
    in >> len;
    fill(in, qme->V[u], N, N);
    qme->len_gamma[u] = len;
    fill(in, qme->gamma[u], len, len);
    qme->phi_0[u].resize(len);
    for (auto& x : qme->phi_0[u]) x = rd(in);
    qme->sigma[u].resize(len);
    for (auto& x : qme->sigma[u]) x = rd(in);
    fill(in, qme->S[u], len, len);
    fill(in, qme->A[u], len, len);
    qme->s_delta[u] = rd(in);
  }
  int solver_kind, n_t;
  double dt, atol, rtol;
  in >> solver_kind >> dt >> atol >> rtol >> n_t;
  std::vector<double> t_list(n_t);
  for (auto& t : t_list) in >> t;

  engine eng(0);
  try {
    qme->set_param(&eng);
    std::vector<dtype> rho(qme->main_size());
    for (auto& x : rho) x = rd(in);
    std::unique_ptr<solver_base<dtype, row_major, engine>> solver;
    if (solver_kind == 0)
      solver.reset(new rk4<dtype, row_major, engine>());
    else if (solver_kind == 2)
      solver.reset(new lsrk4<dtype, row_major, engine>());
    else
      solver.reset(new rkdp<dtype, row_major, engine>());
    qme_solver<dtype, row_major, engine> qs(&eng, qme.get(), solver.get());
    if (mode == "calc_diff") {
      std::vector<dtype> drho(rho.size());
      qs.calc_diff(drho.data(), rho.data());
      for (size_t k = 0; k < drho.size(); ++k)
        std::printf("%s%.17g %.17g", k ? " " : "", drho[k].real(), drho[k].imag());
      std::printf("\n");
      return 0;
    }
    const std::vector<dtype> rho0 = rho;
    kwargs_t kw{{"dt", dt}, {"atol", atol}, {"rtol", rtol}};
    auto print = [&](double t) {
      std::printf("%.17g", t);
      for (int k = 0; k < N * N; ++k) std::printf(" %.17g %.17g", rho[k].real(), rho[k].imag());
      std::printf("\n");
    };
    qs.solve(rho.data(), t_list.data(), n_t, print, kw);
    if (mode == "reparam") {
      for (auto& r : qme->H.data)
        for (auto& c : r.second) c.second *= 0.5;
      qme->set_param(&eng);
      rho = rho0;
      std::printf("reparam\n");
      qs.solve(rho.data(), t_list.data(), n_t, print, kw);
    }
  } catch (heom_error& e) {
    std::fprintf(stderr, "heom_error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
