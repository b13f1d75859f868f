// pointint/align.hpp -- rigid/affine alignment of two geometries by
// maximising a volumetric similarity, on the MI355X intersection-volume
// library (libpointint_cxx.so).
//
// Mirrors the reference's alignment application (paths relative to the
// reference tree):
//   similarity measures   cpp/app/align/similarity_measure.hpp:10-49
//   Newton optimizer      cpp/app/align/optimizer.hpp:11-37
//   moment warm start     cpp/app/align/warm_start.hpp:9-38
// with the stand-in dense types of pointint.hpp in place of Eigen's.  The
// interactive viewer of the reference application (cpp/app/align/main.cpp)
// is not part of the library; pointint_align is its headless counterpart.
#pragma once

#include <memory>
#include <string>

#include "pointint/pointint.hpp"

namespace PoIntInt {

// cpp/app/align/similarity_measure.hpp:10-15
enum SimilarityMeasure { SIM_OCHIAI, SIM_JACCARD, SIM_DICE };

// cpp/app/align/similarity_measure.hpp:17-24
enum SimilarityComputationFlags {
  SIM_FUNC_ONLY = 0x01,
  SIM_GRADIENT = 0x02,
  SIM_HESSIAN = 0x04,
  SIM_FUNC_GRAD = SIM_FUNC_ONLY | SIM_GRADIENT,
  SIM_ALL = SIM_FUNC_ONLY | SIM_GRADIENT | SIM_HESSIAN
};

// The objective is -log(similarity) (cpp/app/align/similarity_measure.hpp:26-33).
struct SimilarityResult {
  double value = 0.0;   // -log(max(similarity, 1e-10))
  VectorXd gradient;    // d value / d dofs2 (12), empty unless SIM_GRADIENT
  MatrixXd hessian;     // 12 x 12, empty unless SIM_HESSIAN
  double similarity = 0.0;
};

// Similarity of ref_geom1 (identity DoFs, fixed) and ref_geom2 under the
// AffineDoF vector dofs2 ([t, A row-major]); V1 is the volume of geometry 1.
// Ochiai V12 / sqrt(V1 V2), Jaccard V12 / (V1 + V2 - V12), Dice
// 2 V12 / (V1 + V2), with V12 and V2 = V22 from the matrix driver; gradient
// from its grad_matrix, Hessian from its Gauss-Newton blocks.
// cpp/app/align/similarity_measure.cpp:9-158
SimilarityResult compute_similarity(const Geometry& ref_geom1, const Geometry& ref_geom2,
                                    const std::shared_ptr<AffineDoF>& affine_dof,
                                    const VectorXd& dofs2, const KGrid& kgrid,
                                    SimilarityMeasure measure, double V1,
                                    int flags = SIM_FUNC_GRAD);

// cpp/app/align/optimizer.hpp:11-19
struct OptimizationState {
  VectorXd dofs;
  double similarity = 0.0;
  double similarity_loss = 0.0;
  int iteration = 0;
  bool converged = false;
  std::string status;
};

// Damped Newton on -log(similarity): the Hessian's negative eigenvalues are
// clamped to zero, damping * I is added, and the step is accepted by Armijo
// backtracking (c = 0.1, halving, at most 20 trials, alpha >= 1e-10).
// Converges when |gradient| < tolerance.
// cpp/app/align/optimizer.cpp:10-134
OptimizationState optimize_alignment(const Geometry& ref_geom1, const Geometry& ref_geom2,
                                     const std::shared_ptr<AffineDoF>& affine_dof,
                                     const VectorXd& initial_dofs, const KGrid& kgrid,
                                     SimilarityMeasure measure, double V1,
                                     double tolerance = 1e-6, int max_iterations = 50,
                                     double damping = 0.1);

// cpp/app/align/warm_start.hpp:9-14
struct Moments {
  double volume = 0.0;
  Vector3d centroid{0.0, 0.0, 0.0};
  MatrixXd covariance = MatrixXd::Zero(3, 3);
};

// Volume, centroid and covariance of a geometry under AffineDoF dofs.
// Triangle meshes: signed tetrahedra against the origin, second moment from
// the tetrahedron centroids; disks / Gaussians: element-weighted centres
// (Gaussians add sigma^2 I).  cpp/app/align/warm_start.cpp:8-131
Moments compute_moments(const Geometry& geom, const std::shared_ptr<AffineDoF>& affine_dof,
                        const VectorXd& dofs);

// Similarity transform A = s R, t = c1 - A c2 that maps geom2's centroid and
// principal axes onto geom1's, s = cbrt(V1 / V2) clamped to [0.1, 10].
// Returns the 12 AffineDoF parameters.  cpp/app/align/warm_start.cpp:133-245
VectorXd compute_warm_start(const Geometry& geom1, const Geometry& geom2,
                            const std::shared_ptr<AffineDoF>& affine_dof);

}  // namespace PoIntInt
