// Host-side helpers of libpointint_hip.so (no device code).
#pragma once

#include <cstdint>
#include <vector>

namespace pii {

const char* last_error();
void gauss_legendre(int n, double a, double b, double* x, double* w);
int lebedev_load(const char* path, std::vector<double>& dirs, std::vector<double>& w);
void kgrid_build(const double* ldirs, const double* lw, int M, int nrad, double* dirs,
                 double* kmag, double* w);
bool kgrid_pairs(int Q, const double* dirs, const double* kmag, const double* w,
                 std::vector<int>& anti, std::vector<int>& rep);
void kgrid_order_reps(int Q, const double* dirs, const double* kmag, const std::vector<int>& anti,
                      std::vector<int>& rep);
void build_patches(int nv, int nf, const int32_t* F, std::vector<int>& hdr,
                   std::vector<int>& pverts, std::vector<uint32_t>& pfaces,
                   std::vector<uint32_t>& ploc);

}  // namespace pii
