// pointint_align: headless counterpart of the reference's alignment
// application (cpp/app/align/main.cpp:160-230 without the viewer).  Loads two
// OFF meshes, builds the Lebedev x Gauss-Legendre k-grid, computes the moment
// warm start and runs the Newton optimizer on -log(similarity).
//
//   pointint_align mesh1.off mesh2.off [lebedev_number] [Nrad] [options]
//     --lebedev-dir DIR   directory of lebedev_NNN.txt (default: data/lebedev
//                         beside this executable)
//     --measure ochiai|jaccard|dice   (default dice)
//     --max-iter N  --tol X  --damping X
//     --no-warm-start     start from the identity transform
//     --dofs d0 .. d11    start (or, with --eval-only, evaluate) at these DoFs
//     --eval-only         print the similarity, gradient and Hessian at the
//                         start DoFs and stop
//     --warm-start-only   print the start DoFs and stop (host only, no GPU)
//
// Lines starting with '@' are the machine-readable results.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "pointint/align.hpp"
#include "pointint_hip.h"

using namespace PoIntInt;

namespace {

// OFF reader: header "OFF", counts "nv nf ne", nv vertex lines, nf polygon
// lines "k i0 .. ik-1" (fan-triangulated); '#' starts a comment.
bool read_off(const std::string& path, MatrixXd& V, MatrixXi& F) {
  std::ifstream in(path);
  if (!in) return false;
  std::vector<std::string> tok;
  std::string line;
  while (std::getline(in, line)) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream ls(line);
    std::string t;
    while (ls >> t) tok.push_back(t);
  }
  // every token must parse whole
  auto num = [](const std::string& t, double& x) {
    char* end = nullptr;
    x = std::strtod(t.c_str(), &end);
    return end && *end == '\0' && std::isfinite(x);
  };
  auto integer = [](const std::string& t, long& x) {
    char* end = nullptr;
    x = std::strtol(t.c_str(), &end, 10);
    return end && *end == '\0';
  };
  size_t p = 0;
  if (p < tok.size() && tok[p] == "OFF") ++p;
  long nv = 0, nf = 0, ne = 0;
  if (p + 3 > tok.size() || !integer(tok[p], nv) || !integer(tok[p + 1], nf) ||
      !integer(tok[p + 2], ne))
    return false;
  p += 3;
  if (nv < 0 || nf < 0 || p + 3 * (size_t)nv > tok.size()) return false;
  V = MatrixXd((int)nv, 3);
  for (int i = 0; i < nv; ++i)
    for (int c = 0; c < 3; ++c)
      if (!num(tok[p++], V(i, c))) return false;
  std::vector<std::array<int, 3>> tris;
  for (long f = 0; f < nf; ++f) {
    long k = 0;
    if (p >= tok.size() || !integer(tok[p++], k)) return false;
    if (k < 3 || p + (size_t)k > tok.size()) return false;
    std::vector<int> idx(k);
    for (long j = 0; j < k; ++j) {
      long v = 0;
      if (!integer(tok[p++], v) || v < 0 || v >= nv) return false;
      idx[j] = (int)v;
    }
    for (long j = 1; j + 1 < k; ++j) tris.push_back({idx[0], idx[j], idx[j + 1]});
  }
  F = MatrixXi((int)tris.size(), 3);
  for (size_t f = 0; f < tris.size(); ++f)
    for (int c = 0; c < 3; ++c) F((int)f, c) = tris[f][c];
  return true;
}

std::string exe_dir(const char* argv0) {
  std::string s(argv0);
  const size_t k = s.rfind('/');
  return k == std::string::npos ? std::string(".") : s.substr(0, k);
}

void print_vec(const char* tag, const VectorXd& v) {
  std::printf("%s", tag);
  for (double x : v) std::printf(" %.17g", x);
  std::printf("\n");
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  std::string leb_dir, measure_name = "dice";
  int max_iter = 50;
  double tol = 1e-6, damping = 0.1;
  bool warm = true, eval_only = false, warm_only = false;
  VectorXd start;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto need = [&](int k) {
      if (i + k >= argc) {
        std::fprintf(stderr, "missing value after %s\n", a.c_str());
        std::exit(2);
      }
    };
    if (a == "--lebedev-dir") { need(1); leb_dir = argv[++i]; }
    else if (a == "--measure") { need(1); measure_name = argv[++i]; }
    else if (a == "--max-iter") { need(1); max_iter = std::atoi(argv[++i]); }
    else if (a == "--tol") { need(1); tol = std::atof(argv[++i]); }
    else if (a == "--damping") { need(1); damping = std::atof(argv[++i]); }
    else if (a == "--no-warm-start") warm = false;
    else if (a == "--eval-only") eval_only = true;
    else if (a == "--warm-start-only") warm_only = true;
    else if (a == "--dofs") {
      need(12);
      start.resize(12);
      for (int k = 0; k < 12; ++k) start[k] = std::atof(argv[++i]);
    } else pos.push_back(a);
  }
  if (pos.size() < 2) {
    std::fprintf(stderr,
                 "usage: %s mesh1.off mesh2.off [lebedev_number] [Nrad] [--lebedev-dir D] "
                 "[--measure ochiai|jaccard|dice] [--max-iter N] [--tol X] [--damping X] "
                 "[--no-warm-start] [--dofs d0..d11] [--eval-only] [--warm-start-only]\n",
                 argv[0]);
    return 2;
  }
  const int leb = pos.size() > 2 ? std::atoi(pos[2].c_str()) : 29;
  const int nrad = pos.size() > 3 ? std::atoi(pos[3].c_str()) : 32;
  if (leb < 3 || leb > 131 || leb % 2 == 0) {
    std::fprintf(stderr, "Error: lebedev_number must be an odd number in range [3, 131]\n");
    return 2;
  }
  SimilarityMeasure measure = SIM_DICE;
  if (measure_name == "ochiai") measure = SIM_OCHIAI;
  else if (measure_name == "jaccard") measure = SIM_JACCARD;
  else if (measure_name != "dice") {
    std::fprintf(stderr, "unknown measure %s\n", measure_name.c_str());
    return 2;
  }
  MatrixXd V1, V2;
  MatrixXi F1, F2;
  if (!read_off(pos[0], V1, F1) || !read_off(pos[1], V2, F2)) {
    std::fprintf(stderr, "Failed to load meshes\n");
    return 1;
  }
  const Geometry g1 = make_triangle_mesh(V1, F1), g2 = make_triangle_mesh(V2, F2);
  auto aff = std::make_shared<AffineDoF>();
  if (leb_dir.empty()) leb_dir = exe_dir(argv[0]) + "/data/lebedev";
  char name[32];
  std::snprintf(name, sizeof(name), "/lebedev_%03d.txt", leb);
  const LebedevGrid L = load_lebedev_txt(leb_dir + name);
  if (L.dirs.empty()) {
    std::fprintf(stderr, "Failed to load %s%s\n", leb_dir.c_str(), name);
    return 1;
  }
  const KGrid kg = build_kgrid(L.dirs, L.weights, nrad);
  std::printf("@kgrid %zu\n", kg.kmag.size());

  VectorXd dofs = start;
  if (dofs.empty()) {
    dofs = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (warm) dofs = compute_warm_start(g1, g2, aff);
  }
  print_vec("@start", dofs);
  if (warm_only) return 0;
  // V1 from the k-space self intersection, as the application does
  // (main.cpp:222-229)
  const double Vol1 = compute_intersection_volume_cuda(g1, g1, kg);
  std::printf("@V1 %.17g\n", Vol1);
  if (eval_only) {
    const SimilarityResult r = compute_similarity(g1, g2, aff, dofs, kg, measure, Vol1, SIM_ALL);
    std::printf("@eval %.17g %.17g\n", r.value, r.similarity);
    print_vec("@grad", r.gradient);
    print_vec("@hess", r.hessian.data);
    return 0;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const OptimizationState st =
      optimize_alignment(g1, g2, aff, dofs, kg, measure, Vol1, tol, max_iter, damping);
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::printf("@final %d %d %.17g %.17g %.3f\n", st.converged ? 1 : 0, st.iteration, st.similarity,
              st.similarity_loss, ms);
  print_vec("@dofs", st.dofs);
  std::printf("@status %s\n", st.status.c_str());
  return 0;
}
