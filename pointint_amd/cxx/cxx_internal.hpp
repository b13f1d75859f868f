// Shared host plumbing of the C++ API (libpointint_cxx.so): the process-wide
// device context and the uploads of Geometry / KGrid through the C ABI.
#pragma once

#include "pointint/pointint.hpp"
#include "pointint_hip.h"

#include <vector>

namespace PoIntInt {
namespace detail {

// One context per process on device 0, created on first use (null on failure).
pii_context* context();
// true for PII_OK; otherwise prints the library's message to std::cerr.
bool report(int rc, const char* where);
pii_geometry* upload(pii_context* ctx, const Geometry& g);
pii_kgrid* upload(pii_context* ctx, const KGrid& kg);
// The matrix driver on geometries and k-grid already resident on the device
// (Q = number of k-nodes, for the profile printout).
IntersectionVolumeMatrixResult volume_matrix_resident(
    pii_context* ctx, const std::vector<const pii_geometry*>& geoms, const std::vector<int>& kinds,
    const std::vector<VectorXd>& dof_vectors, const pii_kgrid* kg, int Q, ComputationFlags flags,
    bool enable_profiling);
// Current form-factor precision and Gram mode (set_precision).
int ff_precision();
int gram_mode();

}  // namespace detail
}  // namespace PoIntInt
