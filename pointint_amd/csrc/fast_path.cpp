// Fast-path drivers: form factors over the representative k-nodes (one per
// antipodal pair when the grid is symmetric), then the Gram.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pii_capi_internal.hpp"

namespace pii {

void launch_form_factors_fast(hipStream_t s, const Batch& b, const KGridDev& kg,
                              int out_mode, double2* d_A, int ldA, float2* d_X, int ldX,
                              unsigned* d_colmax, Workspace& w, Context* prof) {
  const double* krep = kg.rep.as<double>();
  const int N = b.N;
  bool tri_series = false, surfel_series = false, sh_series = false;
  int ser_tri_faces = 0, ser_surf_elems = 0, max_patches = 0, max_surf = 0;
  for (const ObjDesc& d : b.h_objs) {
    const bool tri = d.type == PII_GEOM_TRIANGLE;
    if (d.nser > 0 && d.shc) {
      sh_series = true;
    } else if (d.nser > 0) {
      (tri ? tri_series : surfel_series) = true;
      (tri ? ser_tri_faces : ser_surf_elems) = std::max(tri ? ser_tri_faces : ser_surf_elems, d.n);
    }
    if (tri) max_patches = std::max(max_patches, d.npatch);
    else max_surf = std::max(max_surf, d.n);
    if (!prof) continue;
    prof->work[tri ? 0 : 3] += (double)d.n * (kg.Qh - d.qbeg);
    if (d.nser > 0) {
      if (!d.shc) prof->work[tri ? 1 : 4] += (double)d.n * b.Mh;
      prof->work[2] += (double)d.qbeg;
    }
  }
  // splits and their partial-sum buffer (one buffer, used in launch order)
  const int dirblk = (b.Mh + 63) / 64;
  const int s_tri = tri_series ? split_rounds(dirblk * N, series_tri_slots(),
                                              max_split(ser_tri_faces, 256)) : 1;
  const int s_surf = surfel_series ? split_rounds(dirblk * N, series_surfel_slots(),
                                                  max_split(ser_surf_elems, 256)) : 1;
  // workgroups of the per-element kernel that have nodes (objects start at
  // their own first direct node)
  const int span = (out_mode == 0 ? kg.Qh : ldX) - b.qbeg_min;
  int ff_wgs = 0;
  for (const ObjDesc& d : b.h_objs)
    ff_wgs += std::max(0, ((out_mode == 0 ? kg.Qh : ldX) - d.qbeg + 63) / 64);
  const int s_ff = split_rounds(ff_wgs, ff_fast_slots(),
                                std::max(max_split(max_patches, 2), max_split(max_surf, 256)));
  const size_t need = std::max({series_split_doubles(N, b.Mh, s_tri),
                                series_split_doubles(N, b.Mh, s_surf),
                                ff_split_doubles(N, std::max(span, 0), s_ff)});
  if (need) w.split.ensure(sizeof(double) * need);
  double* part = need ? w.split.as<double>() : nullptr;
  if (tri_series || surfel_series || sh_series) {
    if (prof) prof->phase_begin(4);
    if (sh_series)
      launch_series_sh(s, b.d_objs, N, krep, kg.Qh, b.Mh, out_mode, d_A, ldA, d_X, ldX,
                       d_colmax);
    if (tri_series)
      launch_series_tri(s, b.d_objs, N, krep, kg.Qh, b.Mh, out_mode, d_A, ldA, d_X, ldX,
                        d_colmax, s_tri, part);
    if (surfel_series)
      launch_series_surfel(s, b.d_objs, N, krep, kg.Qh, b.Mh, out_mode, d_A, ldA, d_X, ldX,
                           d_colmax, s_surf, part);
    if (prof) prof->phase_end(4);
    PII_HIP(hipGetLastError());
  }
  if (prof) prof->phase_begin(5);
  launch_ff_fast(s, b.d_objs, N, krep, kg.Qh, b.qbeg_min, out_mode, d_A, ldA, d_X, ldX,
                 d_colmax, s_ff, part);
  if (prof) prof->phase_end(5);
  PII_HIP(hipGetLastError());
}

void gram_split_run(hipStream_t s, Workspace& w, int fp16, const float2* d_X, int ldX, int ncols,
                    const unsigned* d_colmax, std::vector<GramQuery>& qs, double* d_V) {
  GramPlanHeader h{};
  gram_plan_layout(qs, h);
  const int K = 2 * ldX;
  w.planes.ensure(gram_plane_bytes(fp16, K, ncols));
  w.part.ensure(sizeof(double) * (size_t)std::max<long long>(h.part_doubles, 1));
  const size_t plan_bytes = sizeof(GramPlanHeader) + sizeof(GramQuery) * qs.size();
  w.gplan_h.ensure(plan_bytes);
  w.gplan.ensure(plan_bytes);
  std::memcpy(w.gplan_h.p, &h, sizeof(h));
  std::memcpy(static_cast<char*>(w.gplan_h.p) + sizeof(h), qs.data(),
              sizeof(GramQuery) * qs.size());
  PII_HIP(hipMemcpyAsync(w.gplan.p, w.gplan_h.p, plan_bytes, hipMemcpyHostToDevice, s));
  launch_split_planes(s, fp16, reinterpret_cast<const float*>(d_X), K, ncols, d_colmax,
                      w.planes.p, w.gplan.as<GramPlanHeader>());
  launch_gram_planes(s, fp16, w.planes.p, K, ncols, w.gplan.as<GramPlanHeader>(),
                     gram_planes_grid(h.nitems, 1 << 20), h.nentries, w.part.as<double>(),
                     d_colmax, d_V);
}

void form_factor_fast_full(Context& c, const Batch& b, const KGridDev& kg, double* out,
                           hipMemcpyKind kind) {
  hipStream_t s = c.stream;
  const int N = b.N, Qh = kg.Qh, Q = kg.Q;
  const int ldA = (int)align_up((size_t)Qh, 64);
  c.w.X.ensure(sizeof(double2) * (size_t)ldA * N);
  c.w.J.ensure(sizeof(double2) * (size_t)Q * N);
  c.phase_begin(1);
  launch_form_factors_fast(s, b, kg, 0, c.w.X.as<double2>(), ldA, nullptr, 0, nullptr, c.w, &c);
  c.phase_end(1);
  PII_HIP(hipGetLastError());
  launch_expand(s, c.w.X.as<double2>(), ldA, kg.rep_index.as<int>(), kg.anti_index.as<int>(), Qh,
                Q, N, c.w.J.as<double2>());
  PII_HIP(hipGetLastError());
  PII_HIP(hipMemcpyAsync(out, c.w.J.p, sizeof(double2) * (size_t)Q * N, kind, s));
  PII_HIP(hipStreamSynchronize(s));
}

void volume_matrix_fast(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                        double* V_host) {
  hipStream_t s = c.stream;
  const int N = b.N, Qh = kg.Qh;
  const int ldX = (int)align_up((size_t)Qh, 64);
  c.w.X.ensure(sizeof(float2) * (size_t)ldX * N);
  c.w.colmax.ensure(sizeof(unsigned) * (size_t)N);
  unsigned* colmax = c.w.colmax.as<unsigned>();
  PII_HIP(hipMemsetAsync(colmax, 0, sizeof(unsigned) * N, s));
  c.phase_begin(1);
  launch_form_factors_fast(s, b, kg, 1, nullptr, 0, c.w.X.as<float2>(), ldX, colmax, c.w, &c);
  c.phase_end(1);
  PII_HIP(hipGetLastError());
  c.w.V.ensure(sizeof(double) * (size_t)N * N);
  c.phase_begin(2);
  if (gram_mode == PII_GRAM_FP64) {
    const int nsplit = gram_fp64_splits(Qh, N);
    c.w.part.ensure(sizeof(double) * (size_t)nsplit * N * N);
    launch_gram_fp64_x(s, c.w.X.as<float2>(), Qh, N, ldX, c.w.part.as<double>(), nsplit,
                       c.w.V.as<double>());
  } else if (gram_mode == PII_GRAM_BF16X3 || gram_mode == PII_GRAM_FP16X3) {
    std::vector<GramQuery> qs{gram_query(0, N, 2 * ldX)};
    gram_split_run(s, c.w, gram_mode == PII_GRAM_FP16X3 ? 1 : 0, c.w.X.as<float2>(), ldX, N,
                   colmax, qs, c.w.V.as<double>());
  } else {
    throw Error(PII_ERR_INVALID, "unknown Gram mode");
  }
  c.phase_end(2);
  PII_HIP(hipGetLastError());
  // V_host may be device memory (the device-output entries)
  PII_HIP(hipMemcpyAsync(V_host, c.w.V.p, sizeof(double) * (size_t)N * N, hipMemcpyDefault, s));
  PII_HIP(hipStreamSynchronize(s));
}

// Gram of the D FP32 columns in c.w.Xt (leading dimension ldX, column
// maxima colmax) into c.w.H (D x D, FP64): entry (a, b) = (1/8 pi^3)
// sum_q w_q Re(x_a conj x_b) over the full grid.
static void column_gram(Context& c, const KGridDev& kg, int gram_mode, int D, int ldX,
                        unsigned* colmax) {
  hipStream_t s = c.stream;
  c.w.H.ensure(sizeof(double) * (size_t)D * D);
  c.phase_begin(2);
  if (gram_mode == PII_GRAM_FP64) {
    const int nsplit = gram_fp64_splits(kg.Qh, D);
    c.w.part.ensure(sizeof(double) * (size_t)nsplit * D * D);
    launch_gram_fp64_x(s, c.w.Xt.as<float2>(), kg.Qh, D, ldX, c.w.part.as<double>(), nsplit,
                       c.w.H.as<double>());
  } else {
    std::vector<GramQuery> qs{gram_query(0, D, 2 * ldX)};
    gram_split_run(s, c.w, gram_mode == PII_GRAM_FP16X3 ? 1 : 0, c.w.Xt.as<float2>(), ldX, D,
                   colmax, qs, c.w.H.as<double>());
  }
  c.phase_end(2);
  PII_HIP(hipGetLastError());
}

// Gram of the columns [sqrt(w) A_0..A_{N-1} | sqrt(w) dA_i/dtheta_t] of
// AffineDoF objects: the first N columns (when with_A) come from the fast
// form-factor path, the 12 N derivative columns (12 col + t) from the
// forward-mode kernel.  Leaves the (D x D) result, D = (with_A ? 13 : 12) N,
// in c.w.H.  Entry (a, b) is (1/8 pi^3) sum_q w_q Re(x_a conj x_b) over the
// full grid, so the A-A block is V, the dA-A block holds dV_ij/dtheta_i
// (i != j; half of it for i == j) and the dA-dA block the Gauss-Newton H.
static int affine_gram(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                       bool with_A) {
  const int N = b.N, Qh = kg.Qh;
  const int na = with_A ? N : 0;
  if ((long long)na + 12LL * N > kMaxDerivativeColumns)
    throw Error(PII_ERR_INVALID, "affine derivative Gram: " + std::to_string(na + 12LL * N) +
                                     " columns exceed the limit of " +
                                     std::to_string(kMaxDerivativeColumns));
  const int D = na + 12 * N;
  for (const ObjDesc& d : b.h_objs)
    if (d.dof_kind != PII_DOF_AFFINE)
      throw Error(PII_ERR_UNSUPPORTED, "affine derivatives support AffineDoF objects only");
  if (gram_mode != PII_GRAM_FP64 && gram_mode != PII_GRAM_BF16X3 && gram_mode != PII_GRAM_FP16X3)
    throw Error(PII_ERR_INVALID, "unknown Gram mode");
  hipStream_t s = c.stream;
  const int ldX = (int)align_up((size_t)Qh, 64);
  c.w.Xt.ensure(sizeof(float2) * (size_t)ldX * D);
  c.w.cmax2.ensure(sizeof(unsigned) * (size_t)D);
  float2* X = c.w.Xt.as<float2>();
  unsigned* colmax = c.w.cmax2.as<unsigned>();
  PII_HIP(hipMemsetAsync(colmax, 0, sizeof(unsigned) * D, s));
  if (with_A) {
    c.phase_begin(1);
    launch_form_factors_fast(s, b, kg, 1, nullptr, 0, X, ldX, colmax, c.w, &c);
    c.phase_end(1);
    PII_HIP(hipGetLastError());
  }
  // per-face records of triangle objects
  std::vector<long long> off(N, 0);
  size_t tot = 0;
  for (int slot = 0; slot < N; ++slot) {
    off[slot] = (long long)tot;
    if (b.h_objs[slot].type == PII_GEOM_TRIANGLE) tot += align_up((size_t)b.h_objs[slot].n * 18, 64);
  }
  const size_t rec_bytes = align_up(sizeof(float) * std::max<size_t>(tot, 1), 256);
  c.w.dth.ensure(rec_bytes + sizeof(long long) * N);
  float* d_drec = c.w.dth.as<float>();
  long long* d_off = reinterpret_cast<long long*>(static_cast<char*>(c.w.dth.p) + rec_bytes);
  PII_HIP(hipMemcpyAsync(d_off, off.data(), sizeof(long long) * N, hipMemcpyHostToDevice, s));
  launch_dtheta_records(s, b.d_objs, b.d_eprefix, N, b.total_e, d_drec, d_off);
  PII_HIP(hipGetLastError());
  // Derivative series (objects with dser): records, offsets of their records
  // and shell values, then the three kernels.
  std::vector<long long> ds_off(2 * (size_t)N, 0);  // record offsets, value offsets
  size_t ds_rec = 0, ds_val = 0;
  int max_qbeg = 0, max_dfaces = 0;
  for (int slot = 0; slot < N; ++slot) {
    const ObjDesc& d = b.h_objs[slot];
    if (!d.dser) continue;
    ds_off[slot] = (long long)ds_rec;
    ds_off[N + slot] = (long long)ds_val;
    ds_rec += align_up(dseries_record_doubles(d.n), 32);
    ds_val += dseries_value_count(d.qbeg);
    max_qbeg = std::max(max_qbeg, d.qbeg);
    max_dfaces = std::max(max_dfaces, d.n);
  }
  const bool dseries = max_qbeg > 0 && b.Mh > 0;
  int ds_split = 1;
  if (dseries) {
    const int dirblk = (b.Mh + 63) / 64;
    ds_split = split_rounds(dirblk * N * 5, dseries_slots(), max_split(max_dfaces, 256));
    const size_t rec_b = align_up(sizeof(double) * ds_rec, 256);
    c.w.dsrec.ensure(rec_b + sizeof(long long) * 2 * N);
    c.w.dspart.ensure(sizeof(double) * dseries_part_doubles(N, b.Mh, ds_split));
    c.w.dsval.ensure(sizeof(double2) * ds_val);
    PII_HIP(hipMemcpyAsync(static_cast<char*>(c.w.dsrec.p) + rec_b, ds_off.data(),
                           sizeof(long long) * 2 * N, hipMemcpyHostToDevice, s));
  }
  // per-element kernel: split so that the workgroups that have nodes
  // (objects with a derivative series start at qbeg) fill their rounds
  int max_patches = 0, max_surf = 0, base = 0;
  for (const ObjDesc& d : b.h_objs) {
    if (d.type == PII_GEOM_TRIANGLE) max_patches = std::max(max_patches, d.npatch);
    else max_surf = std::max(max_surf, d.n);
    base += (ldX - (d.type == PII_GEOM_TRIANGLE && d.dser ? d.qbeg : 0) + 63) / 64;
  }
  const int nsplit = split_rounds(base, dtheta_slots(),
                                  std::max(max_split(max_patches, 2), max_split(max_surf, 256)));
  const size_t need = dtheta_split_doubles(N, ldX, nsplit);
  if (need) c.w.split.ensure(sizeof(double) * need);
  c.phase_begin(3);
  if (dseries) {
    const size_t rec_b = align_up(sizeof(double) * ds_rec, 256);
    const long long* d_dsoff =
        reinterpret_cast<const long long*>(static_cast<char*>(c.w.dsrec.p) + rec_b);
    launch_dseries_records(s, b.d_objs, b.d_eprefix, N, b.total_e, c.w.dsrec.as<double>(),
                           d_dsoff);
    launch_dseries(s, b.d_objs, N, kg.rep.as<double>(), Qh, b.Mh, c.w.dsrec.as<double>(), d_dsoff,
                   c.w.dspart.as<double>(), ds_split, c.w.dsval.as<double2>(), d_dsoff + N,
                   max_qbeg, X + (size_t)na * ldX, ldX, colmax + na);
    PII_HIP(hipGetLastError());
  }
  launch_dtheta_affine(s, b.d_objs, N, kg.rep.as<double>(), Qh, d_drec, d_off,
                       X + (size_t)na * ldX, ldX, colmax + na, nsplit,
                       need ? c.w.split.as<double>() : nullptr);
  c.phase_end(3);
  PII_HIP(hipGetLastError());
  column_gram(c, kg, gram_mode, D, ldX, colmax);
  return D;
}

// Gram of the derivative columns for any mix of DoF kinds: the same column
// set as affine_gram ([sqrt(w) A | sqrt(w) dA_i/dtheta_i], object i's D_i =
// ndofs[i] columns in caller order), with every derivative column from the
// FP64 Jacobian kernels (the reference's per-node element derivatives) on the
// representative nodes.  This is the route for TriangleMeshDoF objects
// (D_i = 3 x DoF vertices); all-affine batches take affine_gram.
static int general_gram(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                        bool with_A, const int* ndofs) {
  const int N = b.N, Qh = kg.Qh;
  const int na = with_A ? N : 0;
  if (gram_mode != PII_GRAM_FP64 && gram_mode != PII_GRAM_BF16X3 && gram_mode != PII_GRAM_FP16X3)
    throw Error(PII_ERR_INVALID, "unknown Gram mode");
  std::vector<long long> first(N);  // first derivative column, caller order
  long long D = na;
  int maxd = 0;
  for (int i = 0; i < N; ++i) {
    first[i] = D;
    D += ndofs[i];
    maxd = std::max(maxd, ndofs[i]);
  }
  if (D > kMaxDerivativeColumns)
    throw Error(PII_ERR_INVALID, "derivative Gram: " + std::to_string(D) +
                                     " columns exceed the limit of " +
                                     std::to_string(kMaxDerivativeColumns));
  hipStream_t s = c.stream;
  const int ldX = (int)align_up((size_t)Qh, 64);
  c.w.Xt.ensure(sizeof(float2) * (size_t)ldX * D);
  c.w.cmax2.ensure(sizeof(unsigned) * (size_t)D);
  float2* X = c.w.Xt.as<float2>();
  unsigned* colmax = c.w.cmax2.as<unsigned>();
  PII_HIP(hipMemsetAsync(colmax, 0, sizeof(unsigned) * D, s));
  if (with_A) {
    c.phase_begin(1);
    launch_form_factors_fast(s, b, kg, 1, nullptr, 0, X, ldX, colmax, c.w, &c);
    c.phase_end(1);
    PII_HIP(hipGetLastError());
  }
  c.w.J.ensure(sizeof(double2) * (size_t)ldX * std::max(maxd, 1));
  c.phase_begin(3);
  for (int slot = 0; slot < N; ++slot) {
    const ObjDesc& d = b.h_objs[slot];
    const long long col = first[d.col];
    launch_jacobian_columns(s, b.d_objs + slot, d.dof_kind, d.n, kg.rep.as<double>(), Qh,
                            ndofs[d.col], c.w.J.as<double2>(), ldX, X + (size_t)col * ldX, ldX,
                            colmax + col);
    PII_HIP(hipGetLastError());
  }
  c.phase_end(3);
  column_gram(c, kg, gram_mode, (int)D, ldX, colmax);
  return (int)D;
}

void volume_matrix_affine_all(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                              double* V_host, double* grad_host, double* H_host) {
  const int N = b.N;
  const bool with_A = V_host || grad_host;
  const int na = with_A ? N : 0;
  const int D = affine_gram(c, b, kg, gram_mode, with_A);
  std::vector<double> G((size_t)D * D);
  PII_HIP(hipMemcpyAsync(G.data(), c.w.H.p, sizeof(double) * G.size(), hipMemcpyDeviceToHost,
                         c.stream));
  PII_HIP(hipStreamSynchronize(c.stream));
  auto at = [&](int r, int col) { return G[(size_t)r * D + col]; };
  if (V_host)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) V_host[(size_t)i * N + j] = at(i, j);
  if (grad_host)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j)
        for (int t = 0; t < 12; ++t)
          grad_host[((size_t)i * N + j) * 12 + t] = at(na + 12 * i + t, j);
  if (H_host) {
    const int DH = 12 * N;
    for (int r = 0; r < DH; ++r)
      for (int col = 0; col < DH; ++col) H_host[(size_t)r * DH + col] = at(na + r, na + col);
  }
}

void volume_matrix_derivatives(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                               const int* ndofs, double* V_host, double* grad_host,
                               double* H_host) {
  bool all_affine = true;
  for (const ObjDesc& d : b.h_objs) all_affine = all_affine && d.dof_kind == PII_DOF_AFFINE;
  if (all_affine) {  // same layouts with 12 DoFs per object
    volume_matrix_affine_all(c, b, kg, gram_mode, V_host, grad_host, H_host);
    return;
  }
  const int N = b.N;
  const bool with_A = V_host || grad_host;
  const int na = with_A ? N : 0;
  const int D = general_gram(c, b, kg, gram_mode, with_A, ndofs);
  const int DH = D - na;
  const double* G = c.w.H.as<double>();
  if (H_host)
    PII_HIP(hipMemcpy2DAsync(H_host, sizeof(double) * DH, G + (size_t)na * D + na,
                             sizeof(double) * D, sizeof(double) * DH, DH, hipMemcpyDeviceToHost,
                             c.stream));
  if (V_host)
    PII_HIP(hipMemcpy2DAsync(V_host, sizeof(double) * N, G, sizeof(double) * D,
                             sizeof(double) * N, N, hipMemcpyDeviceToHost, c.stream));
  if (grad_host) {
    // rows na.. (derivative columns), columns 0..N-1 (A columns): row r of
    // the copy is dV_(i, .)/dtheta_(i, t) for derivative column r = o_i + t
    std::vector<double> B((size_t)DH * N);
    PII_HIP(hipMemcpy2DAsync(B.data(), sizeof(double) * N, G + (size_t)na * D, sizeof(double) * D,
                             sizeof(double) * N, DH, hipMemcpyDeviceToHost, c.stream));
    PII_HIP(hipStreamSynchronize(c.stream));
    size_t o = 0;
    for (int i = 0; i < N; ++i) {
      const int Di = ndofs[i];
      double* gi = grad_host + o * N;  // (N x D_i) block of object i
      for (int j = 0; j < N; ++j)
        for (int t = 0; t < Di; ++t) gi[(size_t)j * Di + t] = B[(o + t) * N + j];
      o += Di;
    }
  }
  PII_HIP(hipStreamSynchronize(c.stream));
}

void volume_matrix_hessian(Context& c, const Batch& b, const KGridDev& kg, int gram_mode,
                           const int* ndofs, double* H_host, double* V_host) {
  bool all_affine = true;
  for (const ObjDesc& d : b.h_objs) all_affine = all_affine && d.dof_kind == PII_DOF_AFFINE;
  if (all_affine) {
    volume_matrix_affine_all(c, b, kg, gram_mode, V_host, nullptr, H_host);
    return;
  }
  const int N = b.N;
  const int na = V_host ? N : 0;
  const int D = general_gram(c, b, kg, gram_mode, V_host != nullptr, ndofs);
  const int DH = D - na;
  const double* G = c.w.H.as<double>();
  PII_HIP(hipMemcpy2DAsync(H_host, sizeof(double) * DH, G + (size_t)na * D + na,
                           sizeof(double) * D, sizeof(double) * DH, DH, hipMemcpyDeviceToHost,
                           c.stream));
  if (V_host)
    PII_HIP(hipMemcpy2DAsync(V_host, sizeof(double) * N, G, sizeof(double) * D,
                             sizeof(double) * N, N, hipMemcpyDeviceToHost, c.stream));
  PII_HIP(hipStreamSynchronize(c.stream));
}

void volume_matrix_vjp(Context& c, const Batch& b, const GeometryDev* const* geoms,
                       const KGridDev& kg, const double* G, double* const* grads,
                       double* V_host, bool grads_on_device) {
  hipStream_t s = c.stream;
  const int N = b.N, Qh = kg.Qh;
  const int ldX = (int)align_up((size_t)Qh, 64);
  for (int i = 0; i < N; ++i)
    if (geoms[i]->type != PII_GEOM_TRIANGLE)
      throw Error(PII_ERR_UNSUPPORTED, "adjoint supports triangle geometry only");
  c.w.X.ensure(sizeof(float2) * (size_t)ldX * N);
  c.w.B.ensure(sizeof(float2) * (size_t)ldX * N);
  c.w.colmax.ensure(sizeof(unsigned) * (size_t)N);
  unsigned* colmax = c.w.colmax.as<unsigned>();
  PII_HIP(hipMemsetAsync(colmax, 0, sizeof(unsigned) * N, s));
  c.phase_begin(1);
  launch_form_factors_fast(s, b, kg, 1, nullptr, 0, c.w.X.as<float2>(), ldX, colmax, c.w, &c);
  c.phase_end(1);
  PII_HIP(hipGetLastError());
  if (V_host) {
    const int nsplit = gram_fp64_splits(Qh, N);
    c.w.part.ensure(sizeof(double) * (size_t)nsplit * N * N);
    c.w.V.ensure(sizeof(double) * (size_t)N * N);
    c.phase_begin(2);
    launch_gram_fp64_x(s, c.w.X.as<float2>(), Qh, N, ldX, c.w.part.as<double>(), nsplit,
                       c.w.V.as<double>());
    c.phase_end(2);
    PII_HIP(hipGetLastError());
  }
  // C = G + G^T, indexed [j][i] by caller column.
  std::vector<double> C((size_t)N * N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) C[(size_t)i * N + j] = G[(size_t)i * N + j] + G[(size_t)j * N + i];
  c.w.C.ensure(sizeof(double) * C.size());
  PII_HIP(hipMemcpyAsync(c.w.C.p, C.data(), sizeof(double) * C.size(), hipMemcpyHostToDevice, s));
  // World-vertex gradient slices per slot.
  std::vector<int> goff(N + 1, 0);
  int max_patches = 0;
  for (int slot = 0; slot < N; ++slot) {
    goff[slot + 1] = goff[slot] + 3 * b.h_objs[slot].nv;
    max_patches = std::max(max_patches, b.h_objs[slot].npatch);
  }
  c.w.gw.ensure(sizeof(double) * std::max(goff[N], 1) + sizeof(int) * (N + 1) + 256);
  double* d_gw = c.w.gw.as<double>();
  int* d_goff = reinterpret_cast<int*>(reinterpret_cast<char*>(c.w.gw.p) +
                                       align_up(sizeof(double) * std::max(goff[N], 1), 256));
  PII_HIP(hipMemcpyAsync(d_goff, goff.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, s));
  PII_HIP(hipMemsetAsync(d_gw, 0, sizeof(double) * std::max(goff[N], 1), s));
  c.phase_begin(3);
  launch_mix_columns(s, c.w.X.as<float2>(), N, ldX, Qh, c.w.C.as<double>(), c.w.B.as<float2>());
  if (b.nseries > 0) {
    int max_faces = 0;
    for (const ObjDesc& d : b.h_objs)
      if (d.nser > 0) max_faces = std::max(max_faces, d.n);
    c.w.gam.ensure(sizeof(double) * (size_t)N * b.Mh * kGammaStride);
    launch_series_adjoint_tri(s, b.d_objs, N, max_faces, kg.rep.as<double>(), Qh, b.Mh,
                              c.w.B.as<float2>(), ldX, c.w.gam.as<double>(), d_gw, d_goff);
  }
  // enough workgroups to fill the chip: split the k-range when objects are few
  const int wgs = std::max(1, max_patches * N);
  const int ksplit = std::max(1, std::min(64, (4096 + wgs - 1) / wgs));
  launch_adjoint_tri(s, b.d_objs, N, max_patches, kg.rep.as<double>(), Qh, b.qbeg_min,
                     c.w.B.as<float2>(), ldX, ksplit, d_gw, d_goff);
  c.phase_end(3);
  PII_HIP(hipGetLastError());
  if (grads_on_device) {
    // Gradients into device memory: vertex DoFs of indexed meshes are the
    // world-vertex gradients themselves, affine DoFs their moments (a
    // reduction kernel); vertex DoFs with shared ids are summed on the host.
    std::vector<AffineGradJob> jobs;
    bool host_map = false;
    for (int slot = 0; slot < N; ++slot) {
      const int i = b.order[slot];
      const GeometryDev& g = *geoms[i];
      const ObjDesc& d = b.h_objs[slot];
      if (d.dof_kind == PII_DOF_TRIMESH && g.dof_identity)
        PII_HIP(hipMemcpyAsync(grads[i], d_gw + goff[slot], sizeof(double) * 3 * (size_t)g.nv,
                               hipMemcpyDeviceToDevice, s));
      else if (d.dof_kind == PII_DOF_AFFINE)
        jobs.push_back({d_gw + goff[slot], g.verts.as<double>(), g.nv, 0, grads[i]});
      else
        host_map = true;
    }
    if (!jobs.empty()) {
      c.w.C.ensure(sizeof(AffineGradJob) * jobs.size());
      PII_HIP(hipMemcpyAsync(c.w.C.p, jobs.data(), sizeof(AffineGradJob) * jobs.size(),
                             hipMemcpyHostToDevice, s));
      launch_affine_dof_grads(s, c.w.C.as<AffineGradJob>(), (int)jobs.size());
      PII_HIP(hipGetLastError());
    }
    if (V_host)
      PII_HIP(hipMemcpyAsync(V_host, c.w.V.p, sizeof(double) * (size_t)N * N, hipMemcpyDefault,
                             s));
    if (host_map) {
      std::vector<double> gw((size_t)std::max(goff[N], 1)), out;
      PII_HIP(hipMemcpyAsync(gw.data(), d_gw, sizeof(double) * goff[N], hipMemcpyDeviceToHost,
                             s));
      PII_HIP(hipStreamSynchronize(s));
      for (int slot = 0; slot < N; ++slot) {
        const int i = b.order[slot];
        const GeometryDev& g = *geoms[i];
        const ObjDesc& d = b.h_objs[slot];
        if (d.dof_kind != PII_DOF_TRIMESH || g.dof_identity) continue;
        out.assign(3 * (size_t)g.ndof_verts, 0.0);
        for (int v = 0; v < g.nv; ++v) {
          const int dv = g.h_dofvid[v];
          if (dv < 0) continue;
          for (int cc = 0; cc < 3; ++cc) out[3 * dv + cc] += gw[goff[slot] + 3 * v + cc];
        }
        PII_HIP(hipMemcpyAsync(grads[i], out.data(), sizeof(double) * out.size(),
                               hipMemcpyHostToDevice, s));
        PII_HIP(hipStreamSynchronize(s));
      }
    }
    PII_HIP(hipStreamSynchronize(s));
    return;
  }
  // world-vertex gradients come back through pinned staging (one DMA copy)
  c.w.hout.ensure(sizeof(double) * std::max(goff[N], 1));
  const double* gw = c.w.hout.as<double>();
  PII_HIP(hipMemcpyAsync(c.w.hout.p, d_gw, sizeof(double) * goff[N], hipMemcpyDeviceToHost, s));
  if (V_host)
    PII_HIP(hipMemcpyAsync(V_host, c.w.V.p, sizeof(double) * (size_t)N * N,
                           hipMemcpyDeviceToHost, s));
  PII_HIP(hipStreamSynchronize(s));
  // World-vertex gradients -> DoF gradients.
  for (int slot = 0; slot < N; ++slot) {
    const int i = b.order[slot];
    const GeometryDev& g = *geoms[i];
    const ObjDesc& d = b.h_objs[slot];
    const double* gv = gw + goff[slot];
    double* out = grads[i];
    if (d.dof_kind == PII_DOF_TRIMESH && g.dof_identity) {
      std::memcpy(out, gv, sizeof(double) * 3 * (size_t)g.nv);
    } else if (d.dof_kind == PII_DOF_TRIMESH) {
      std::fill(out, out + 3 * g.ndof_verts, 0.0);
      for (int v = 0; v < g.nv; ++v) {
        const int dv = g.h_dofvid[v];
        if (dv < 0) continue;
        for (int cc = 0; cc < 3; ++cc) out[3 * dv + cc] += gv[3 * v + cc];
      }
    } else {
      // W = A V + t:  dL/dt = sum_v gW_v,  dL/dA_rc = sum_v gW_v,r V_v,c
      std::fill(out, out + 12, 0.0);
      for (int v = 0; v < g.nv; ++v) {
        for (int r = 0; r < 3; ++r) {
          out[r] += gv[3 * v + r];
          for (int cc = 0; cc < 3; ++cc) out[3 + 3 * r + cc] += gv[3 * v + r] * g.h_verts[3 * v + cc];
        }
      }
    }
  }
}

}  // namespace pii
