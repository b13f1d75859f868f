// Alignment of two geometries by a volumetric similarity (pointint/align.hpp):
// the similarity measures and their derivatives on the matrix driver, a damped
// Newton optimizer with Armijo backtracking, and the moment warm start.
// Host code on the C ABI; every volume, gradient and Hessian comes from the
// device through the matrix driver.
#include "pointint/align.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>

#include "cxx_internal.hpp"

namespace PoIntInt {

namespace {

const VectorXd kIdentityDofs = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};

// ---- small dense linear algebra (n <= 12) ----

// Symmetric eigen-decomposition by cyclic Jacobi rotations (Golub & Van Loan,
// Matrix Computations, 4th ed., Alg. 8.5.3).  a: n x n row-major.  Returns
// eigenvalues ascending and eigenvectors as the columns of Q (row-major),
// each column with its largest-magnitude component positive (first such
// component on ties).  false for non-finite input or no convergence.
bool sym_eigen(int n, std::vector<double> a, std::vector<double>& evals, std::vector<double>& Q) {
  for (double x : a)
    if (!std::isfinite(x)) return false;
  std::vector<double> V((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  auto A = [&](int r, int c) -> double& { return a[(size_t)r * n + c]; };
  double fro = 0.0;
  for (double x : a) fro += x * x;
  fro = std::sqrt(fro);
  bool done = fro == 0.0;
  for (int sweep = 0; sweep < 100 && !done; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += 2.0 * A(p, q) * A(p, q);
    if (std::sqrt(off) <= 1e-15 * fro) {
      done = true;
      break;
    }
    for (int p = 0; p < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = A(p, q);
        if (apq == 0.0) continue;
        const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {  // V <- V J
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
    }
  }
  if (!done) return false;
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return A(x, x) < A(y, y); });
  evals.assign(n, 0.0);
  Q.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    const int src = idx[j];
    evals[j] = A(src, src);
    int big = 0;
    for (int k = 1; k < n; ++k)
      if (std::fabs(V[(size_t)k * n + src]) > std::fabs(V[(size_t)big * n + src])) big = k;
    const double sg = V[(size_t)big * n + src] < 0.0 ? -1.0 : 1.0;
    for (int k = 0; k < n; ++k) Q[(size_t)k * n + j] = sg * V[(size_t)k * n + src];
  }
  return true;
}

using M3 = std::array<double, 9>;  // row-major 3 x 3

M3 mul(const M3& x, const M3& y) {
  M3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) r[3 * i + j] += x[3 * i + k] * y[3 * k + j];
  return r;
}

M3 transpose(const M3& x) {
  M3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[3 * i + j] = x[3 * j + i];
  return r;
}

double det3(const M3& m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// M = U S V^T with singular values descending (the order of a Jacobi SVD):
// V from the eigenvectors of M^T M, U's columns M v / sigma, completed to an
// orthonormal basis where sigma vanishes.
bool svd3(const M3& M, M3& U, M3& V) {
  const M3 MtM = mul(transpose(M), M);
  std::vector<double> ev, Qv;
  if (!sym_eigen(3, std::vector<double>(MtM.begin(), MtM.end()), ev, Qv)) return false;
  double sig[3];
  for (int j = 0; j < 3; ++j) {  // descending
    sig[j] = std::sqrt(std::max(ev[2 - j], 0.0));
    for (int k = 0; k < 3; ++k) V[3 * k + j] = Qv[3 * k + (2 - j)];
  }
  U = M3{};
  int good = 0;
  for (int j = 0; j < 3; ++j) {
    if (!(sig[j] > 1e-12 * std::max(sig[0], 1e-300))) break;
    double u[3] = {0, 0, 0};
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) u[r] += M[3 * r + k] * V[3 * k + j];
    for (int r = 0; r < 3; ++r) U[3 * r + j] = u[r] / sig[j];
    ++good;
  }
  for (int j = good; j < 3; ++j) {  // complete the basis
    double best[3] = {0, 0, 0};
    double bn = -1.0;
    for (int e = 0; e < 3; ++e) {
      double u[3] = {0, 0, 0};
      u[e] = 1.0;
      for (int p = 0; p < j; ++p) {
        const double d = U[3 * 0 + p] * u[0] + U[3 * 1 + p] * u[1] + U[3 * 2 + p] * u[2];
        for (int r = 0; r < 3; ++r) u[r] -= d * U[3 * r + p];
      }
      const double nn = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      if (nn > bn) {
        bn = nn;
        for (int r = 0; r < 3; ++r) best[r] = u[r] / nn;
      }
    }
    for (int r = 0; r < 3; ++r) U[3 * r + j] = best[r];
  }
  return true;
}

// ---- similarity from the matrix driver's result ----

struct PairTerms {
  double V_int = 0.0, V2 = 0.0;
  VectorXd gV_int, gV2;   // 12
  MatrixXd HV_int, HV2;   // 12 x 12
};

// cpp/app/align/similarity_measure.cpp:45-64: V_int = V(0,1), V2 = V(1,1),
// dV_int = grad_matrix[1][0], dV2 = 2 grad_matrix[1][1], H_int =
// hessian_jj[0][1], H_2 = hessian_ii[1][1] + hessian_ij[1][1] +
// hessian_ij[1][1]^T + hessian_jj[1][1].
bool pair_terms(const IntersectionVolumeMatrixResult& r, bool grad, bool hess, PairTerms& t) {
  if (r.volume_matrix.rows() != 2) return false;
  t.V_int = r.volume_matrix(0, 1);
  t.V2 = r.volume_matrix(1, 1);
  if (grad) {
    if (r.grad_matrix.size() != 2 || r.grad_matrix[1][0].size() != 12 ||
        r.grad_matrix[1][1].size() != 12)
      return false;
    t.gV_int = r.grad_matrix[1][0];
    t.gV2 = r.grad_matrix[1][1];
    for (double& x : t.gV2) x *= 2.0;
  }
  if (hess) {
    if (r.hessian_ij.size() != 2 || r.hessian_ii.size() != 2 || r.hessian_jj.size() != 2)
      return false;
    t.HV_int = r.hessian_jj[0][1];
    t.HV2 = MatrixXd(12, 12);
    for (int a = 0; a < 12; ++a)
      for (int b = 0; b < 12; ++b)
        t.HV2(a, b) = r.hessian_ii[1][1](a, b) + r.hessian_ij[1][1](a, b) +
                      r.hessian_ij[1][1](b, a) + r.hessian_jj[1][1](a, b);
  }
  return true;
}

// cpp/app/align/similarity_measure.cpp:66-155
SimilarityResult similarity_from_terms(const PairTerms& t, SimilarityMeasure measure, double V1,
                                       int flags) {
  const bool need_g = (flags & SIM_GRADIENT) != 0, need_h = (flags & SIM_HESSIAN) != 0;
  const double Vi = t.V_int, V2 = t.V2;
  double s = 0.0;
  VectorXd gs(12, 0.0);
  MatrixXd Hs = MatrixXd::Zero(12, 12);
  const VectorXd& gi = t.gV_int;
  const VectorXd& g2 = t.gV2;
  // Hs += a * X + b * (u v^T + v u^T) + c * w w^T, built term by term below
  auto add_mat = [&](double a, const MatrixXd& X) {
    for (int p = 0; p < 12; ++p)
      for (int q = 0; q < 12; ++q) Hs(p, q) += a * X(p, q);
  };
  auto add_outer = [&](double a, const VectorXd& u, const VectorXd& v) {
    for (int p = 0; p < 12; ++p)
      for (int q = 0; q < 12; ++q) Hs(p, q) += a * u[p] * v[q];
  };
  if (measure == SIM_OCHIAI) {
    const double inv = 1.0 / std::sqrt(V1 * V2);
    s = inv * Vi;
    if (need_g) {
      const double coeff = -0.5 * inv * inv * inv * Vi * V1;
      for (int p = 0; p < 12; ++p) gs[p] = inv * gi[p] + coeff * g2[p];
    }
    if (need_h) {
      add_mat(inv, t.HV_int);
      add_mat(-0.5 * inv * Vi / V2, t.HV2);
      add_outer(0.75 * inv * Vi / (V2 * V2), g2, g2);
      add_outer(-0.5 * inv / V2, gi, g2);
      add_outer(-0.5 * inv / V2, g2, gi);
    }
  } else if (measure == SIM_JACCARD) {
    const double den = V1 + V2 - Vi;
    s = Vi / den;
    const double inv = 1.0 / den, inv2 = 1.0 / (den * den);
    VectorXd dD(12, 0.0);
    if (need_g || need_h)
      for (int p = 0; p < 12; ++p) dD[p] = g2[p] - gi[p];
    if (need_g)
      for (int p = 0; p < 12; ++p) gs[p] = inv * gi[p] - inv2 * Vi * dD[p];
    if (need_h) {
      add_mat(inv + inv2 * Vi, t.HV_int);  // inv H_int - inv2 Vi (H_2 - H_int)
      add_mat(-inv2 * Vi, t.HV2);
      add_outer(-inv2, gi, dD);
      add_outer(-inv2, dD, gi);
      add_outer(2.0 * inv2 * inv * Vi, dD, dD);
    }
  } else {  // SIM_DICE
    const double den = V1 + V2;
    s = 2.0 * Vi / den;
    const double inv = 2.0 / den, coeff = -2.0 * Vi / (den * den);
    if (need_g)
      for (int p = 0; p < 12; ++p) gs[p] = inv * gi[p] + coeff * g2[p];
    if (need_h) {
      add_mat(inv, t.HV_int);
      add_mat(coeff, t.HV2);
      add_outer(-inv / den, gi, g2);
      add_outer(-inv / den, g2, gi);
      add_outer(-2.0 * coeff / den, g2, g2);
    }
  }
  SimilarityResult out;
  out.similarity = s;
  const double ss = std::max(s, 1e-10);
  out.value = -std::log(ss);
  if (need_g) {
    out.gradient.assign(12, 0.0);
    for (int p = 0; p < 12; ++p) out.gradient[p] = -gs[p] / ss;
  }
  if (need_h) {
    out.hessian = MatrixXd(12, 12);
    for (int p = 0; p < 12; ++p)
      for (int q = 0; q < 12; ++q)
        out.hessian(p, q) = gs[p] * gs[q] / (ss * ss) - Hs(p, q) / ss;
  }
  return out;
}

ComputationFlags volume_flags(int flags) {
  ComputationFlags f = ComputationFlags::VOLUME_ONLY;
  if (flags & SIM_GRADIENT) f = f | ComputationFlags::GRADIENT;
  if (flags & SIM_HESSIAN) f = f | ComputationFlags::HESSIAN;
  return f;
}

SimilarityResult failed_similarity(int flags) {
  SimilarityResult r;
  r.similarity = 0.0;
  r.value = -std::log(1e-10);
  if (flags & SIM_GRADIENT) r.gradient.assign(12, 0.0);
  if (flags & SIM_HESSIAN) r.hessian = MatrixXd::Zero(12, 12);
  return r;
}

// Geometry and k-grid uploaded once for the optimizer's many evaluations.
struct ResidentPair {
  pii_context* ctx = nullptr;
  pii_geometry* g[2] = {nullptr, nullptr};
  pii_kgrid* kg = nullptr;
  int Q = 0;
  bool ok = false;
  ResidentPair(const Geometry& g1, const Geometry& g2, const KGrid& kgrid) {
    ctx = detail::context();
    if (!ctx) return;
    kg = detail::upload(ctx, kgrid);
    g[0] = detail::upload(ctx, g1);
    g[1] = detail::upload(ctx, g2);
    Q = (int)kgrid.kmag.size();
    ok = kg && g[0] && g[1];
  }
  ~ResidentPair() {
    for (pii_geometry* x : g)
      if (x) pii_geometry_destroy(x);
    if (kg) pii_kgrid_destroy(kg);
  }
  ResidentPair(const ResidentPair&) = delete;
  ResidentPair& operator=(const ResidentPair&) = delete;

  SimilarityResult similarity(const VectorXd& dofs2, SimilarityMeasure measure, double V1,
                              int flags) const {
    if (!ok) return failed_similarity(flags);
    const auto r = detail::volume_matrix_resident(ctx, {g[0], g[1]}, {PII_DOF_AFFINE, PII_DOF_AFFINE},
                                                  {kIdentityDofs, dofs2}, kg, Q,
                                                  volume_flags(flags), false);
    PairTerms t;
    if (!pair_terms(r, (flags & SIM_GRADIENT) != 0, (flags & SIM_HESSIAN) != 0, t))
      return failed_similarity(flags);
    return similarity_from_terms(t, measure, V1, flags);
  }
};

double norm(const VectorXd& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

}  // namespace

SimilarityResult compute_similarity(const Geometry& ref_geom1, const Geometry& ref_geom2,
                                    const std::shared_ptr<AffineDoF>& affine_dof,
                                    const VectorXd& dofs2, const KGrid& kgrid,
                                    SimilarityMeasure measure, double V1, int flags) {
  if (dofs2.size() != 12) {
    std::cerr << "compute_similarity: expected 12 DoFs, got " << dofs2.size() << std::endl;
    return failed_similarity(flags);
  }
  const std::vector<Geometry> geoms = {ref_geom1, ref_geom2};
  const std::vector<std::shared_ptr<DoFParameterization>> dofs = {std::make_shared<AffineDoF>(),
                                                                   affine_dof};
  const std::vector<VectorXd> dv = {kIdentityDofs, dofs2};
  const auto r = compute_intersection_volume_matrix_cuda(geoms, dofs, dv, kgrid,
                                                         volume_flags(flags), 256, false);
  PairTerms t;
  if (!pair_terms(r, (flags & SIM_GRADIENT) != 0, (flags & SIM_HESSIAN) != 0, t))
    return failed_similarity(flags);
  return similarity_from_terms(t, measure, V1, flags);
}

OptimizationState optimize_alignment(const Geometry& ref_geom1, const Geometry& ref_geom2,
                                     const std::shared_ptr<AffineDoF>& /*affine_dof*/,
                                     const VectorXd& initial_dofs, const KGrid& kgrid,
                                     SimilarityMeasure measure, double V1, double tolerance,
                                     int max_iterations, double damping) {
  OptimizationState state;
  state.dofs = initial_dofs;
  if (initial_dofs.size() != 12) {
    state.status = "Expected 12 affine DoFs";
    return state;
  }
  const ResidentPair pair(ref_geom1, ref_geom2, kgrid);
  if (!pair.ok) {
    state.status = "Device upload failed";
    return state;
  }
  // cpp/app/align/optimizer.cpp:25-29
  const double alpha_init = 1.0, alpha_min = 1e-10, c_armijo = 0.1, rho = 0.5;
  const int max_ls_iter = 20;
  for (int iter = 0; iter < max_iterations; ++iter) {
    state.iteration = iter;
    const SimilarityResult sr = pair.similarity(state.dofs, measure, V1, SIM_ALL);
    state.similarity = sr.similarity;
    state.similarity_loss = sr.value;
    if (norm(sr.gradient) < tolerance) {
      state.converged = true;
      state.status = "Converged (gradient norm < tolerance)";
      break;
    }
    // PSD projection of the Hessian, then damping (optimizer.cpp:55-74); the
    // Newton system (H_psd + damping I) delta = -g is solved in the same
    // eigenbasis.
    std::vector<double> ev, Q;
    if (!sym_eigen(12, sr.hessian.data, ev, Q)) {
      state.status = "Hessian eigendecomposition failed";
      break;
    }
    VectorXd delta(12, 0.0);
    bool solved = true;
    for (int j = 0; j < 12; ++j) {
      const double lam = std::max(ev[j], 0.0) + damping;
      if (!(lam > 0.0) || !std::isfinite(lam)) {
        solved = false;
        break;
      }
      double qg = 0.0;
      for (int k = 0; k < 12; ++k) qg += Q[(size_t)k * 12 + j] * sr.gradient[k];
      for (int k = 0; k < 12; ++k) delta[k] -= Q[(size_t)k * 12 + j] * qg / lam;
    }
    if (!solved) {
      state.status = "Hessian solve failed";
      break;
    }
    double descent = 0.0;
    for (int k = 0; k < 12; ++k) descent -= delta[k] * sr.gradient[k];
    std::cout << "  Iteration " << iter << ": -d.g = " << descent << std::endl;
    // Armijo backtracking, function values only (optimizer.cpp:84-118)
    const double f0 = sr.value;
    double alpha = alpha_init;
    VectorXd trial(12);
    auto step = [&](double a) {
      for (int k = 0; k < 12; ++k) trial[k] = state.dofs[k] + a * delta[k];
      return pair.similarity(trial, measure, V1, SIM_FUNC_ONLY).value;
    };
    double f_new = step(alpha);
    std::printf("Backtracking line search.\nTrying alpha = %g, f = %g.\n", alpha, f_new);
    int ls_iter = 0;
    while (f_new > f0 - c_armijo * alpha * descent && alpha > alpha_min && ls_iter < max_ls_iter) {
      alpha *= rho;
      f_new = step(alpha);
      std::printf("Trying alpha = %g, f = %g.\n", alpha, f_new);
      ++ls_iter;
    }
    if (alpha <= alpha_min) {
      state.status = "Line search failed (step size too small)";
      break;
    }
    std::cout << "    Line search: alpha = " << alpha << ", f_new = " << f_new << ", f0 = " << f0
              << std::endl;
    state.dofs = trial;
  }
  // As in the reference (optimizer.cpp:127-129), every non-converged exit
  // reports "Reached max iterations".
  if (!state.converged) state.status = "Reached max iterations";
  return state;
}

Moments compute_moments(const Geometry& geom, const std::shared_ptr<AffineDoF>& affine_dof,
                        const VectorXd& dofs) {
  Moments m;
  if (dofs.size() != 12) {
    std::cerr << "compute_moments: expected 12 DoFs, got " << dofs.size() << std::endl;
    return m;
  }
  const double* A = dofs.data() + 3;
  const double* t = dofs.data();
  auto xf = [&](double x, double y, double z, double out[3]) {
    for (int r = 0; r < 3; ++r) out[r] = A[3 * r] * x + A[3 * r + 1] * y + A[3 * r + 2] * z + t[r];
  };
  if (geom.type == GEOM_TRIANGLE) {
    // warm_start.cpp:22-70: tetrahedra (0, a, b, c); second moment from
    // the tetrahedron centroids
    double V = 0.0, first[3] = {0, 0, 0}, second[9] = {0};
    for (const TriPacked& tr : geom.tris) {
      double a[3], b[3], c[3];
      xf(tr.a.x, tr.a.y, tr.a.z, a);
      xf(tr.a.x + tr.e1.x, tr.a.y + tr.e1.y, tr.a.z + tr.e1.z, b);
      xf(tr.a.x + tr.e2.x, tr.a.y + tr.e2.y, tr.a.z + tr.e2.z, c);
      const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      const double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                            u[0] * v[1] - u[1] * v[0]};
      const double vol = (a[0] * cr[0] + a[1] * cr[1] + a[2] * cr[2]) / 6.0;
      V += vol;
      double ct[3];
      for (int r = 0; r < 3; ++r) ct[r] = (a[r] + b[r] + c[r]) / 4.0;
      for (int r = 0; r < 3; ++r) {
        first[r] += vol * ct[r];
        for (int q = 0; q < 3; ++q) second[3 * r + q] += vol * ct[r] * ct[q];
      }
    }
    if (std::fabs(V) > 1e-10) {
      m.volume = V;
      for (int r = 0; r < 3; ++r) m.centroid[r] = first[r] / V;
      for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q)
          m.covariance(r, q) = second[3 * r + q] / V - m.centroid[r] * m.centroid[q];
    }
    return m;
  }
  // warm_start.cpp:71-128: disks / Gaussians
  m.volume = affine_dof ? affine_dof->compute_volume(geom, dofs) : AffineDoF().compute_volume(geom, dofs);
  const bool disk = geom.type == GEOM_DISK;
  const size_t ne = disk ? geom.disks.size() : geom.gaussians.size();
  std::vector<double> cw(3 * ne), wt(ne), s2(ne, 0.0);
  double sum[3] = {0, 0, 0}, tw = 0.0;
  for (size_t e = 0; e < ne; ++e) {
    const double3 c = disk ? geom.disks[e].c : geom.gaussians[e].c;
    xf(c.x, c.y, c.z, &cw[3 * e]);
    wt[e] = disk ? geom.disks[e].area : geom.gaussians[e].w;
    if (!disk) s2[e] = geom.gaussians[e].sigma * geom.gaussians[e].sigma;
    for (int r = 0; r < 3; ++r) sum[r] += wt[e] * cw[3 * e + r];
    tw += wt[e];
  }
  if (tw > 1e-10)
    for (int r = 0; r < 3; ++r) m.centroid[r] = sum[r] / tw;
  for (size_t e = 0; e < ne; ++e) {
    double d[3];
    for (int r = 0; r < 3; ++r) d[r] = cw[3 * e + r] - m.centroid[r];
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q)
        m.covariance(r, q) += wt[e] * (d[r] * d[q] + (r == q ? s2[e] : 0.0));
  }
  if (tw > 1e-10)
    for (double& x : m.covariance.data) x /= tw;
  return m;
}

VectorXd compute_warm_start(const Geometry& geom1, const Geometry& geom2,
                            const std::shared_ptr<AffineDoF>& affine_dof) {
  const Moments m1 = compute_moments(geom1, affine_dof, kIdentityDofs);
  const Moments m2 = compute_moments(geom2, affine_dof, kIdentityDofs);
  std::cout << "Warm start: Computing affine transform to match moments" << std::endl;
  std::cout << "  Geometry 1: V=" << m1.volume << ", c=[" << m1.centroid[0] << " "
            << m1.centroid[1] << " " << m1.centroid[2] << "]" << std::endl;
  std::cout << "  Geometry 2: V=" << m2.volume << ", c=[" << m2.centroid[0] << " "
            << m2.centroid[1] << " " << m2.centroid[2] << "]" << std::endl;
  M3 R = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double s = 1.0;
  std::vector<double> e1, e2, q1, q2;
  if (!sym_eigen(3, m1.covariance.data, e1, q1) || !sym_eigen(3, m2.covariance.data, e2, q2)) {
    std::cout << "Warning: Covariance eigendecomposition failed, using identity transform"
              << std::endl;
  } else {
    // principal axes as columns, made right-handed (warm_start.cpp:163-175)
    M3 P1, P2;
    std::copy(q1.begin(), q1.end(), P1.begin());
    std::copy(q2.begin(), q2.end(), P2.begin());
    if (det3(P1) < 0)
      for (int r = 0; r < 3; ++r) P1[3 * r + 2] = -P1[3 * r + 2];
    if (det3(P2) < 0)
      for (int r = 0; r < 3; ++r) P2[3 * r + 2] = -P2[3 * r + 2];
    // Procrustes: R = U V^T for P1 P2^T = U S V^T, reflection removed
    // (warm_start.cpp:177-190)
    M3 U, V;
    if (svd3(mul(P1, transpose(P2)), U, V)) {
      R = mul(U, transpose(V));
      if (det3(R) < 0) {
        for (int r = 0; r < 3; ++r) U[3 * r + 2] = -U[3 * r + 2];
        R = mul(U, transpose(V));
      }
    }
    // uniform scale from the volume ratio (warm_start.cpp:192-215)
    if (std::fabs(m2.volume) > 1e-10) {
      s = std::cbrt(m1.volume / m2.volume);
    } else {
      double mean = 0.0;
      int cnt = 0;
      for (int i = 0; i < 3; ++i)
        if (e2[i] > 1e-6) {
          mean += std::sqrt(e1[i] / e2[i]);
          ++cnt;
        }
      s = cnt > 0 ? mean / cnt : 1.0;
    }
    s = std::max(0.1, std::min(10.0, s));
  }
  VectorXd out(12, 0.0);
  for (int r = 0; r < 3; ++r) {
    double Ac2 = 0.0;
    for (int q = 0; q < 3; ++q) {
      out[3 + 3 * r + q] = s * R[3 * r + q];
      Ac2 += s * R[3 * r + q] * m2.centroid[q];
    }
    out[r] = m1.centroid[r] - Ac2;
  }
  std::cout << "  Similarity transform: scale = " << s << std::endl;
  std::cout << "  Rotation determinant: " << det3(R) << std::endl;
  std::cout << "  Warm start transform computed" << std::endl;
  return out;
}

}  // namespace PoIntInt
