// FP64 CPU oracle for the k-space intersection-volume path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++17 with
// std::thread, the reference's FP64 CPU algorithm so that the HIP library in
// pointint_amd/ can be checked against it.  Only tests/, the smoke check and
// the cpu_baseline leg of bench.py load the library built from it; the
// product never links or calls it.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference tree).  The reference's CPU path cannot be built here (it
// needs Eigen 3.4.0 fetched at configure time, TBB headers and
// <cuda_runtime.h>), so this restatement is the oracle; what pins it is listed
// in the test files (closed forms, invariants, brute-force quadrature, finite
// differences).
//
// Layouts (all row-major, FP64 unless stated):
//   triangles  : tri[T][12] = {a.xyz, e1.xyz, e2.xyz, S.xyz}, vid[T][3] int32
//   disks      : disk[D][8] = {c.xyz, n.xyz, rho, area}
//   gaussians  : gau[G][8]  = {c.xyz, n.xyz, sigma, w}
//   affine DoFs: 12 = {t.xyz, A row-major}
//   mesh DoFs  : 3V vertex coordinates
//   complex    : interleaved (re, im)

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace {

using cd = std::complex<double>;
const cd I1(0.0, 1.0);
const double kPi = 3.14159265358979323846;

int g_threads = 0;
int n_threads() {
  if (g_threads > 0) return g_threads;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

// Static block partition of [0, n) over the worker threads, like the
// reference's tbb::parallel_for over k-nodes / pairs.
template <class F>
void parallel_for(int64_t n, F&& body) {
  int nt = (int)std::min<int64_t>(n_threads(), std::max<int64_t>(n, 1));
  if (nt <= 1) {
    for (int64_t i = 0; i < n; ++i) body(i, 0);
    return;
  }
  std::vector<std::thread> th;
  std::atomic<int64_t> next(0);
  const int64_t chunk = std::max<int64_t>(1, n / (nt * 16));
  for (int t = 0; t < nt; ++t) {
    th.emplace_back([&, t]() {
      for (;;) {
        int64_t b = next.fetch_add(chunk);
        if (b >= n) break;
        int64_t e = std::min(n, b + chunk);
        for (int64_t i = b; i < e; ++i) body(i, t);
      }
    });
  }
  for (auto& x : th) x.join();
}

struct V3 { double x, y, z; };
inline V3 v3(const double* p) { return {p[0], p[1], p[2]}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

struct M3 { double m[3][3]; };
inline V3 mul(const M3& A, V3 v) {
  return {A.m[0][0] * v.x + A.m[0][1] * v.y + A.m[0][2] * v.z,
          A.m[1][0] * v.x + A.m[1][1] * v.y + A.m[1][2] * v.z,
          A.m[2][0] * v.x + A.m[2][1] * v.y + A.m[2][2] * v.z};
}
inline V3 mulT(const M3& A, V3 v) {
  return {A.m[0][0] * v.x + A.m[1][0] * v.y + A.m[2][0] * v.z,
          A.m[0][1] * v.x + A.m[1][1] * v.y + A.m[2][1] * v.z,
          A.m[0][2] * v.x + A.m[1][2] * v.y + A.m[2][2] * v.z};
}
// cpp/src/dof/affine_dof.cpp:17-29 -- t = dofs[0:3], A = dofs[3:12] row-major.
inline void affine_from_dofs(const double* d, M3& A, V3& t) {
  t = {d[0], d[1], d[2]};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A.m[r][c] = d[3 + 3 * r + c];
}

// ---------------------------------------------------------------------------
// Element functions: cpp/src/element_functions.cpp:17-120
// ---------------------------------------------------------------------------
cd E_func(double z) {  // :17-28
  if (std::abs(z) < 1e-4) {
    double z2 = z * z, z4 = z2 * z2;
    return cd(1.0 - z2 / 6.0 + z4 / 120.0, z * 0.5 - z * z2 / 24.0 + z4 * z / 720.0);
  }
  return cd(std::sin(z) / z, (1.0 - std::cos(z)) / z);
}
cd E_prime(double z) {  // :30-44
  if (std::abs(z) < 1e-4) {
    double z2 = z * z, z3 = z2 * z, z4 = z2 * z2;
    return cd(-z / 3.0 + z3 / 30.0, 0.5 - z2 / 8.0 + z4 / 144.0);
  }
  double s = std::sin(z), c = std::cos(z), z2 = z * z;
  return cd((z * c - s) / z2, (z * s - (1.0 - c)) / z2);
}
cd E_double_prime(double z) {  // :46-60
  double z2 = z * z;
  if (std::abs(z) < 1e-4) return cd(-1.0 / 3.0 + z2 / 10.0, -z / 4.0 + z2 * z / 36.0);
  double s = std::sin(z), c = std::cos(z), z3 = z * z * z;
  return cd((2.0 * s - 2.0 * z * c - z2 * s) / z3,
            (2.0 * (1.0 - c) - 2.0 * z * s + z2 * c) / z3);
}
cd Phi_ab(double a, double b) {  // :63-70
  double d = b - a;
  if (std::abs(d) < 1e-3) return -2.0 * I1 * E_prime(0.5 * (a + b));
  return -2.0 * I1 * (E_func(b) - E_func(a)) / d;
}
void Phi_ab_gradient(double a, double b, cd& dA, cd& dB) {  // :73-93
  double d = b - a;
  if (std::abs(d) < 1e-3) {
    cd g = -I1 * E_double_prime(0.5 * (a + b));
    dA = g; dB = g;
    return;
  }
  double d2 = d * d;
  cd t1 = (E_func(b) - E_func(a)) / d2;
  dA = -2.0 * I1 * (t1 - E_prime(a) / d);
  dB = 2.0 * I1 * (t1 - E_prime(b) / d);
}
double J1_over_x(double x) {  // :100-109
  if (std::abs(x) < 1e-5) {
    double x2 = x * x;
    return 0.5 - x2 / 16.0 + x2 * x2 / 384.0;
  }
  return std::cyl_bessel_j(1.0, x) / x;
}
double J1_over_x_prime(double x) {  // :111-120
  if (std::abs(x) < 1e-5) return -x / 8.0 + (x * x * x) / 96.0;
  return -std::cyl_bessel_j(2.0, x) / x;
}

// ---------------------------------------------------------------------------
// A(k) per DoF kind: cpp/src/dof/affine_dof.cpp:55-170 (affine: triangles,
// disks, Gaussians) and cpp/src/dof/triangle_mesh_dof.cpp:18-74 (vertex DoFs).
// The reference sums with a TBB parallel_reduce over elements; here the
// element sum is serial (the k-loop above it is the parallel one).
// ---------------------------------------------------------------------------
enum { GT_TRI = 0, GT_DISK = 1, GT_GAUSS = 2 };
enum { DOF_AFFINE = 0, DOF_MESH = 1 };

struct Geo {
  int type;
  int n;
  const double* el;   // 12 (tri) or 8 (disk/gaussian) doubles per element
  const int32_t* vid; // tri only
};

cd A_affine(const Geo& g, V3 k, const double* dofs) {
  M3 A; V3 t;
  affine_from_dofs(dofs, A, t);
  double kmag = norm(k);
  if (kmag < 1e-10) return cd(0, 0);
  V3 khat = scl(k, 1.0 / kmag);
  cd sum(0, 0);
  if (g.type == GT_TRI) {  // affine_dof.cpp:69-102
    for (int i = 0; i < g.n; ++i) {
      const double* e = g.el + 12 * (size_t)i;
      V3 ap = add(mul(A, v3(e)), t);
      V3 e1p = mul(A, v3(e + 3)), e2p = mul(A, v3(e + 6));
      V3 Sp = scl(cross(e1p, e2p), 0.5);
      double al = dot(k, e1p), be = dot(k, e2p), ga = dot(khat, Sp);
      cd phi = Phi_ab(al, be);
      cd ph = std::exp(I1 * dot(k, ap));
      sum += ph * phi * ga;
    }
  } else if (g.type == GT_DISK) {  // affine_dof.cpp:103-135
    V3 kp = mulT(A, k);
    for (int i = 0; i < g.n; ++i) {
      const double* e = g.el + 8 * (size_t)i;
      V3 cp = add(mul(A, v3(e)), t);
      V3 n = v3(e + 3);
      double rho = e[6], area = e[7];
      double kdn = dot(kp, n);
      double r2 = std::max(0.0, dot(kp, kp) - kdn * kdn);
      double r = std::sqrt(r2);
      double rr = rho * r;
      double S = (rr < 1e-10) ? area : 2.0 * area * J1_over_x(rr);
      double Ap = S * kdn / kmag;
      sum += std::exp(dot(k, cp) * I1) * Ap;
    }
  } else if (g.type == GT_GAUSS) {  // affine_dof.cpp:136-167 (no clamp of r2)
    V3 kp = mulT(A, k);
    for (int i = 0; i < g.n; ++i) {
      const double* e = g.el + 8 * (size_t)i;
      V3 cp = add(mul(A, v3(e)), t);
      V3 n = v3(e + 3);
      double sig = e[6], w = e[7];
      double kdn = dot(kp, n);
      double r2 = dot(kp, kp) - kdn * kdn;
      double S = w * std::exp(-0.5 * sig * sig * r2);
      double Ap = S * kdn / kmag;
      sum += std::exp(dot(k, cp) * I1) * Ap;
    }
  }
  return sum;
}

cd A_mesh(const Geo& g, V3 k, const double* V) {  // triangle_mesh_dof.cpp:18-74
  double kn = norm(k);
  if (kn < 1e-10) return cd(0, 0);
  V3 khat = scl(k, 1.0 / kn);
  cd sum(0, 0);
  for (int f = 0; f < g.n; ++f) {
    const int32_t* vi = g.vid + 3 * (size_t)f;
    if (vi[0] < 0 || vi[1] < 0 || vi[2] < 0) continue;
    V3 v0 = v3(V + 3 * (size_t)vi[0]), v1 = v3(V + 3 * (size_t)vi[1]), v2 = v3(V + 3 * (size_t)vi[2]);
    V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    V3 sn = scl(cross(e1, e2), 0.5);
    double kap = dot(khat, sn);
    cd T = std::exp(I1 * dot(k, v0));
    cd psi = Phi_ab(dot(k, e1), dot(k, e2));
    sum += kap * T * psi;
  }
  return sum;
}

// dA/dθ: affine_dof.cpp:172-322 and triangle_mesh_dof.cpp:76-143.
void A_grad_affine(const Geo& g, V3 k, const double* dofs, cd* grad /*12*/) {
  M3 A; V3 t;
  affine_from_dofs(dofs, A, t);
  for (int i = 0; i < 12; ++i) grad[i] = 0;
  double kmag = norm(k);
  if (kmag < 1e-10) return;
  V3 khat = scl(k, 1.0 / kmag);
  double kv[3] = {k.x, k.y, k.z};
  // gA is a complex 3x3 matrix; grad[3 + 3r + c] += gA^T stored column-major,
  // i.e. the reference maps gA.transpose() into the column-major Eigen buffer
  // at offset 3, which lands gA(r,c) at index 3 + 3r + c (affine_dof.cpp:225).
  auto add_gA = [&](cd gA[3][3]) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) grad[3 + 3 * r + c] += gA[r][c];
  };
  if (g.type == GT_TRI) {
    for (int i = 0; i < g.n; ++i) {
      const double* e = g.el + 12 * (size_t)i;
      V3 a = v3(e), e1 = v3(e + 3), e2 = v3(e + 6), S = v3(e + 9);
      V3 ap = add(mul(A, a), t);
      V3 e1p = mul(A, e1), e2p = mul(A, e2);
      V3 Sp = scl(cross(e1p, e2p), 0.5);
      double al = dot(k, e1p), be = dot(k, e2p), ga = dot(khat, Sp);
      cd phi = Phi_ab(al, be);
      cd ph = std::exp(I1 * dot(k, ap));
      cd Atri = ph * phi * ga;
      cd dPa, dPb;
      Phi_ab_gradient(al, be, dPa, dPb);
      for (int d = 0; d < 3; ++d) grad[d] += I1 * Atri * kv[d];
      double av[3] = {a.x, a.y, a.z};
      double e1v[3] = {e1.x, e1.y, e1.z}, e2v[3] = {e2.x, e2.y, e2.z};
      cd gA[3][3];
      // term 1: A_tri * i * k * a^T ; term 2: gamma*phase * k (dPa e1 + dPb e2)^T
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          gA[r][c] = Atri * I1 * kv[r] * av[c] + ga * ph * kv[r] * (dPa * e1v[c] + dPb * e2v[c]);
      // term 3: - [khat]x * A * [S]x * phase * phi   (affine_dof.cpp:224)
      double Kx[3][3] = {{0, -khat.z, khat.y}, {khat.z, 0, -khat.x}, {-khat.y, khat.x, 0}};
      double Sx[3][3] = {{0, -S.z, S.y}, {S.z, 0, -S.x}, {-S.y, S.x, 0}};
      double KA[3][3], KAS[3][3];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double s = 0;
          for (int m = 0; m < 3; ++m) s += Kx[r][m] * A.m[m][c];
          KA[r][c] = s;
        }
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double s = 0;
          for (int m = 0; m < 3; ++m) s += KA[r][m] * Sx[m][c];
          KAS[r][c] = s;
        }
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) gA[r][c] -= KAS[r][c] * ph * phi;
      add_gA(gA);
    }
  } else {
    V3 kp = mulT(A, k);
    for (int i = 0; i < g.n; ++i) {
      const double* e = g.el + 8 * (size_t)i;
      V3 c = v3(e), n = v3(e + 3);
      V3 cp = add(mul(A, c), t);
      double kdn = dot(kp, n);
      double S, dS;
      if (g.type == GT_DISK) {  // affine_dof.cpp:240-269
        double rho = e[6], area = e[7];
        double r2 = std::max(0.0, dot(kp, kp) - kdn * kdn);
        double r = std::sqrt(r2), rr = rho * r;
        S = (rr < 1e-10) ? area : 2.0 * area * J1_over_x(rr);
        dS = (rr < 1e-10) ? -0.25 * rho : 2.0 * J1_over_x_prime(rr) / r;
        dS *= rho * area;
      } else {  // affine_dof.cpp:284-311
        double sig = e[6], w = e[7];
        double r2 = dot(kp, kp) - kdn * kdn;
        S = w * std::exp(-0.5 * sig * sig * r2);
        dS = -sig * sig * S;
      }
      double Ap = S * kdn / kmag;
      cd ph = std::exp(dot(k, cp) * I1);
      cd Ael = ph * Ap;
      for (int d = 0; d < 3; ++d) grad[d] += I1 * Ael * kv[d];
      double cv[3] = {c.x, c.y, c.z}, nv[3] = {n.x, n.y, n.z}, khv[3] = {khat.x, khat.y, khat.z};
      V3 kperp = sub(kp, scl(n, kdn));
      double kpv[3] = {kperp.x, kperp.y, kperp.z};
      cd gA[3][3];
      for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc)
          gA[r][cc] = Ael * I1 * kv[r] * cv[cc] + khv[r] * nv[cc] * ph * S +
                      ph * (kdn / kmag * dS) * kv[r] * kpv[cc];
      add_gA(gA);
    }
  }
}

void A_grad_mesh(const Geo& g, V3 k, const double* V, int ndofs, cd* grad) {
  for (int i = 0; i < ndofs; ++i) grad[i] = 0;
  double kn = norm(k);
  if (kn < 1e-10) return;
  V3 khat = scl(k, 1.0 / kn);
  for (int f = 0; f < g.n; ++f) {
    const int32_t* vi = g.vid + 3 * (size_t)f;
    if (vi[0] < 0 || vi[1] < 0 || vi[2] < 0) continue;
    V3 v0 = v3(V + 3 * (size_t)vi[0]), v1 = v3(V + 3 * (size_t)vi[1]), v2 = v3(V + 3 * (size_t)vi[2]);
    V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    V3 sn = scl(cross(e1, e2), 0.5);
    double kap = dot(khat, sn);
    V3 dk[3] = {scl(cross(khat, sub(e2, e1)), 0.5), scl(cross(khat, e2), -0.5),
                scl(cross(khat, e1), 0.5)};
    cd T = std::exp(I1 * dot(k, v0));
    double ke1 = dot(k, e1), ke2 = dot(k, e2);
    cd psi = Phi_ab(ke1, ke2);
    cd dPa, dPb;
    Phi_ab_gradient(ke1, ke2, dPa, dPb);
    double kv[3] = {k.x, k.y, k.z};
    cd coef[3] = {psi * I1 - dPa - dPb, dPa, dPb};
    for (int j = 0; j < 3; ++j) {
      double dkv[3] = {dk[j].x, dk[j].y, dk[j].z};
      for (int d = 0; d < 3; ++d)
        grad[3 * (size_t)vi[j] + d] += T * (psi * dkv[d] + kap * kv[d] * coef[j]);
    }
  }
}

Geo make_geo(int type, int n, const double* el, const int32_t* vid) {
  Geo g;
  g.type = type; g.n = n; g.el = el; g.vid = vid;
  return g;
}

cd A_any(const Geo& g, int dof_kind, const double* dofs, V3 k) {
  return dof_kind == DOF_MESH ? A_mesh(g, k, dofs) : A_affine(g, k, dofs);
}

}  // namespace

extern "C" {

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads() { return n_threads(); }

// cpp/include/quadrature/gauss_legendre.hpp:16-62 -- Newton on P_n from the
// Tricomi-style initial guess cos(pi (i+0.75)/(n+0.5)), |dz| <= 1e-14, nodes
// mapped to [a,b] ascending (x[i] = xm - xl*z_i).
void orc_gauss_legendre(int n, double a, double b, double* x, double* w) {
  int m = (n + 1) / 2;
  for (int i = 0; i < m; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), z1;
    do {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      double pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (std::abs(z - z1) > 1e-14);
    double p1 = 1.0, p2 = 0.0;
    for (int j = 1; j <= n; ++j) {
      double p3 = p2;
      p2 = p1;
      p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
    }
    double pp = n * (z * p1 - p2) / (z * z - 1.0);
    double wi = 2.0 / ((1.0 - z * z) * pp * pp);
    double xm = 0.5 * (b + a), xl = 0.5 * (b - a);
    x[i] = xm - xl * z;
    x[n - 1 - i] = xm + xl * z;
    w[i] = xl * wi;
    w[n - 1 - i] = xl * wi;
  }
}

// cpp/include/quadrature/lebedev_io.hpp:24-59.  Returns the number of
// directions (or -1 on open failure).  With dirs == NULL only counts.
int orc_load_lebedev(const char* path, double* dirs, double* w, int cap) {
  std::ifstream in(path);
  if (!in) return -1;
  std::string line;
  std::vector<double> D, W;
  double wsum = 0;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    double th, ph, ww;
    if (!(ss >> th >> ph >> ww)) continue;
    th *= kPi / 180.0;
    ph *= kPi / 180.0;
    D.push_back(std::cos(th) * std::sin(ph));
    D.push_back(std::sin(th) * std::sin(ph));
    D.push_back(std::cos(ph));
    W.push_back(ww);
    wsum += ww;
  }
  if (std::abs(wsum - 1.0) < 0.1)
    for (auto& x : W) x *= 4.0 * kPi;
  int M = (int)W.size();
  if (dirs && w) {
    int n = std::min(M, cap);
    std::memcpy(dirs, D.data(), sizeof(double) * 3 * n);
    std::memcpy(w, W.data(), sizeof(double) * n);
  }
  return M;
}

// cpp/include/quadrature/kgrid.hpp:31-47 -- radial-major q = ir*M + j,
// k = tan t, w = w_leb * (w_t * sec^2 t).
void orc_build_kgrid(const double* ldirs, const double* lw, int M, int nrad,
                     double* dirs, double* kmag, double* w) {
  std::vector<double> t(nrad), wt(nrad);
  orc_gauss_legendre(nrad, 0.0, 0.5 * kPi, t.data(), wt.data());
  for (int ir = 0; ir < nrad; ++ir) {
    double sec2 = 1.0 / std::cos(t[ir]) / std::cos(t[ir]);
    double k = std::tan(t[ir]);
    for (int j = 0; j < M; ++j) {
      size_t q = (size_t)ir * M + j;
      dirs[3 * q] = ldirs[3 * j];
      dirs[3 * q + 1] = ldirs[3 * j + 1];
      dirs[3 * q + 2] = ldirs[3 * j + 2];
      kmag[q] = k;
      w[q] = lw[j] * (wt[ir] * sec2);
    }
  }
}

// cpp/src/geometry/packing.cpp:10-28
void orc_pack_tris(const double* V, const int32_t* F, int nf, double* tri, int32_t* vid) {
  for (int i = 0; i < nf; ++i) {
    V3 a = v3(V + 3 * (size_t)F[3 * i]), b = v3(V + 3 * (size_t)F[3 * i + 1]),
       c = v3(V + 3 * (size_t)F[3 * i + 2]);
    V3 e1 = sub(b, a), e2 = sub(c, a), S = scl(cross(e1, e2), 0.5);
    double* o = tri + 12 * (size_t)i;
    o[0] = a.x; o[1] = a.y; o[2] = a.z;
    o[3] = e1.x; o[4] = e1.y; o[5] = e1.z;
    o[6] = e2.x; o[7] = e2.y; o[8] = e2.z;
    o[9] = S.x; o[10] = S.y; o[11] = S.z;
    vid[3 * i] = F[3 * i]; vid[3 * i + 1] = F[3 * i + 1]; vid[3 * i + 2] = F[3 * i + 2];
  }
}

// cpp/src/geometry/packing.cpp:30-63 (disks) and :84-111 (Gaussians)
void orc_pack_disks(const double* P, const double* N, const double* r, int n, int is_radius,
                    double* out) {
  for (int i = 0; i < n; ++i) {
    V3 nn = v3(N + 3 * (size_t)i);
    double l = norm(nn);
    if (l > 0) nn = scl(nn, 1.0 / l);
    double rho, area;
    if (is_radius) { rho = r[i]; area = kPi * rho * rho; }
    else { area = r[i]; rho = std::sqrt(area / kPi); }
    double* o = out + 8 * (size_t)i;
    o[0] = P[3 * i]; o[1] = P[3 * i + 1]; o[2] = P[3 * i + 2];
    o[3] = nn.x; o[4] = nn.y; o[5] = nn.z;
    o[6] = rho; o[7] = area;
  }
}
void orc_pack_gaussians(const double* P, const double* N, const double* sig, const double* wt,
                        int n, double* out) {
  for (int i = 0; i < n; ++i) {
    V3 nn = v3(N + 3 * (size_t)i);
    double l = norm(nn);
    if (l > 0) nn = scl(nn, 1.0 / l);
    double* o = out + 8 * (size_t)i;
    o[0] = P[3 * i]; o[1] = P[3 * i + 1]; o[2] = P[3 * i + 2];
    o[3] = nn.x; o[4] = nn.y; o[5] = nn.z;
    o[6] = sig[i]; o[7] = wt[i];
  }
}

// Element functions exported for the branch/continuity checks.
void orc_E(double z, double* out) { cd v = E_func(z); out[0] = v.real(); out[1] = v.imag(); }
void orc_Ep(double z, double* out) { cd v = E_prime(z); out[0] = v.real(); out[1] = v.imag(); }
void orc_Epp(double z, double* out) { cd v = E_double_prime(z); out[0] = v.real(); out[1] = v.imag(); }
void orc_Phi(double a, double b, double* out) { cd v = Phi_ab(a, b); out[0] = v.real(); out[1] = v.imag(); }
void orc_Phi_grad(double a, double b, double* out) {
  cd x, y;
  Phi_ab_gradient(a, b, x, y);
  out[0] = x.real(); out[1] = x.imag(); out[2] = y.real(); out[3] = y.imag();
}
double orc_J1_over_x(double x) { return J1_over_x(x); }
double orc_J1_over_x_prime(double x) { return J1_over_x_prime(x); }

// A(k) at one k-vector.
void orc_compute_A(int type, int n, const double* el, const int32_t* vid, int dof_kind,
                   const double* dofs, const double* k, double* out) {
  Geo g = make_geo(type, n, el, vid);
  cd v = A_any(g, dof_kind, dofs, v3(k));
  out[0] = v.real(); out[1] = v.imag();
}

// A(k_q) for all Q nodes of a grid, parallel over q like
// cpp/src/compute_intersection_volume_multi_object.cpp:74-93.
void orc_compute_A_grid(int type, int n, const double* el, const int32_t* vid, int dof_kind,
                        const double* dofs, int Q, const double* dirs, const double* kmag,
                        double* out) {
  Geo g = make_geo(type, n, el, vid);
  parallel_for(Q, [&](int64_t q, int) {
    V3 k = scl(v3(dirs + 3 * q), kmag[q]);
    cd v = A_any(g, dof_kind, dofs, k);
    out[2 * q] = v.real();
    out[2 * q + 1] = v.imag();
  });
}

// dA/dθ at one k (out: 2*ndofs doubles).
void orc_compute_A_gradient(int type, int n, const double* el, const int32_t* vid, int dof_kind,
                            const double* dofs, int ndofs, const double* k, double* out) {
  Geo g = make_geo(type, n, el, vid);
  std::vector<cd> gr(ndofs);
  if (dof_kind == DOF_MESH) A_grad_mesh(g, v3(k), dofs, ndofs, gr.data());
  else A_grad_affine(g, v3(k), dofs, gr.data());
  for (int i = 0; i < ndofs; ++i) { out[2 * i] = gr[i].real(); out[2 * i + 1] = gr[i].imag(); }
}

// V = J^T D J / (8 pi^3): cpp/src/compute_intersection_volume_multi_object.cpp:98-149.
// J row-major Q x N complex (J[(q*N+obj)]), parallel over upper-triangle pairs.
void orc_gram(const double* J, int Q, int N, const double* w, double* V) {
  int64_t npairs = (int64_t)N * (N + 1) / 2;
  std::vector<int> pi(npairs), pj(npairs);
  int64_t idx = 0;
  for (int i = 0; i < N; ++i)
    for (int j = i; j < N; ++j) { pi[idx] = i; pj[idx] = j; ++idx; }
  const double scale = 8.0 * kPi * kPi * kPi;
  parallel_for(npairs, [&](int64_t p, int) {
    int i = pi[p], j = pj[p];
    double s = 0;
    for (int q = 0; q < Q; ++q) {
      const double* a = J + 2 * ((size_t)q * N + i);
      const double* b = J + 2 * ((size_t)q * N + j);
      s += w[q] * (a[0] * b[0] + a[1] * b[1]);
    }
    double v = s / scale;
    V[(size_t)i * N + j] = v;
    V[(size_t)j * N + i] = v;
  });
}

// Vector-Jacobian product of one object's A(k):
//   out[d] = (1/8 pi^3) sum_q w_q Re(dA(k_q)/dθ_d * conj(B_q))
// which equals sum_j G_ij-weighted grad_matrix rows of the reference
// (cpp/src/compute_intersection_volume_multi_object.cpp:186-233) by linearity.
void orc_vjp(int type, int n, const double* el, const int32_t* vid, int dof_kind,
             const double* dofs, int ndofs, int Q, const double* dirs, const double* kmag,
             const double* w, const double* B, double* out) {
  Geo g = make_geo(type, n, el, vid);
  int nt = n_threads();
  std::vector<std::vector<double>> acc(nt, std::vector<double>(ndofs, 0.0));
  parallel_for(Q, [&](int64_t q, int t) {
    V3 k = scl(v3(dirs + 3 * q), kmag[q]);
    std::vector<cd> gr(ndofs);
    if (dof_kind == DOF_MESH) A_grad_mesh(g, k, dofs, ndofs, gr.data());
    else A_grad_affine(g, k, dofs, gr.data());
    cd b(B[2 * q], B[2 * q + 1]);
    auto& a = acc[t];
    for (int d = 0; d < ndofs; ++d) a[d] += w[q] * (gr[d].real() * b.real() + gr[d].imag() * b.imag());
  });
  const double scale = 8.0 * kPi * kPi * kPi;
  for (int d = 0; d < ndofs; ++d) {
    double s = 0;
    for (int t = 0; t < nt; ++t) s += acc[t][d];
    out[d] = s / scale;
  }
}

// Self volume by the divergence theorem: affine_dof.cpp:324-384 and
// triangle_mesh_dof.cpp:145-179.
double orc_compute_volume(int type, int n, const double* el, const int32_t* vid, int dof_kind,
                          const double* dofs) {
  if (dof_kind == DOF_MESH) {
    double s = 0;
    for (int f = 0; f < n; ++f) {
      const int32_t* vi = vid + 3 * (size_t)f;
      if (vi[0] < 0 || vi[1] < 0 || vi[2] < 0) continue;
      V3 a = v3(dofs + 3 * (size_t)vi[0]), b = v3(dofs + 3 * (size_t)vi[1]), c = v3(dofs + 3 * (size_t)vi[2]);
      s += dot(a, cross(b, c));
    }
    return s / 6.0;
  }
  M3 A; V3 t;
  affine_from_dofs(dofs, A, t);
  double det = A.m[0][0] * (A.m[1][1] * A.m[2][2] - A.m[1][2] * A.m[2][1]) -
               A.m[0][1] * (A.m[1][0] * A.m[2][2] - A.m[1][2] * A.m[2][0]) +
               A.m[0][2] * (A.m[1][0] * A.m[2][1] - A.m[1][1] * A.m[2][0]);
  double vb = 0;
  if (type == GT_TRI) {
    for (int i = 0; i < n; ++i) { const double* e = el + 12 * (size_t)i; vb += dot(v3(e), v3(e + 9)); }
  } else {
    for (int i = 0; i < n; ++i) { const double* e = el + 8 * (size_t)i; vb += dot(v3(e), v3(e + 3)) * e[7]; }
  }
  vb /= 3.0;
  return vb * std::abs(det);
}

// Divergence-theorem volume gradient for vertex DoFs: triangle_mesh_dof.cpp:181-218.
void orc_compute_volume_gradient_mesh(int n, const int32_t* vid, const double* dofs, int ndofs,
                                      double* out) {
  for (int i = 0; i < ndofs; ++i) out[i] = 0;
  for (int f = 0; f < n; ++f) {
    const int32_t* vi = vid + 3 * (size_t)f;
    if (vi[0] < 0 || vi[1] < 0 || vi[2] < 0) continue;
    V3 a = v3(dofs + 3 * (size_t)vi[0]), b = v3(dofs + 3 * (size_t)vi[1]), c = v3(dofs + 3 * (size_t)vi[2]);
    V3 g0 = cross(b, c), g1 = cross(c, a), g2 = cross(a, b);
    double* o0 = out + 3 * (size_t)vi[0]; double* o1 = out + 3 * (size_t)vi[1]; double* o2 = out + 3 * (size_t)vi[2];
    o0[0] += g0.x / 6; o0[1] += g0.y / 6; o0[2] += g0.z / 6;
    o1[0] += g1.x / 6; o1[1] += g1.y / 6; o1[2] += g1.z / 6;
    o2[0] += g2.x / 6; o2[1] += g2.y / 6; o2[2] += g2.z / 6;
  }
}

// cpp/src/form_factor_helpers.cpp:145-179
double orc_exact_cube_ff(double kx, double ky, double kz) {
  auto sinc = [](double x) {
    if (std::abs(x) < 1e-4) return 1.0 - x * x / 6.0 + x * x * x * x / 120.0;
    return std::sin(x) / x;
  };
  return sinc(0.5 * kx) * sinc(0.5 * ky) * sinc(0.5 * kz);
}
double orc_exact_sphere_ff(double k, double R) {
  if (std::abs(k) < 1e-4) {
    double R3 = R * R * R;
    return 4.0 * kPi * R3 / 3.0 * (1.0 - k * k * R * R / 10.0 + k * k * k * k * R * R * R * R / 280.0);
  }
  double kR = k * R;
  return 4.0 * kPi * (std::sin(kR) - kR * std::cos(kR)) / (k * k * k);
}

// cpp/src/analytical_intersection.cpp:15-60
double orc_box_box(const double* c1, const double* h1, const double* c2, const double* h2) {
  double v = 1.0;
  for (int i = 0; i < 3; ++i) {
    double lo = std::max(c1[i] - h1[i], c2[i] - h2[i]);
    double hi = std::min(c1[i] + h1[i], c2[i] + h2[i]);
    v *= std::max(0.0, hi - lo);
  }
  return v;
}
double orc_sphere_sphere(const double* c1, double r1, const double* c2, double r2) {
  double d = norm(sub(v3(c2), v3(c1)));
  if (d >= r1 + r2) return 0.0;
  if (d <= std::abs(r1 - r2)) {
    double rm = std::min(r1, r2);
    return (4.0 * kPi / 3.0) * rm * rm * rm;
  }
  double t1 = (r1 + r2 - d) * (r1 + r2 - d);
  double t2 = d * d + 2.0 * d * (r1 + r2) - 3.0 * (r1 * r1 + r2 * r2) + 6.0 * r1 * r2;
  return kPi * t1 * t2 / (12.0 * d);
}

// cpp/src/geometry/geometry_helpers.cpp:16-38
void orc_unit_cube(double* V, int32_t* F) {
  static const double v[24] = {-0.5, -0.5, -0.5, 0.5, -0.5, -0.5, 0.5, 0.5, -0.5, -0.5, 0.5, -0.5,
                               -0.5, -0.5, 0.5,  0.5, -0.5, 0.5,  0.5, 0.5, 0.5,  -0.5, 0.5, 0.5};
  static const int32_t f[36] = {0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
                                2, 7, 6, 2, 3, 7, 0, 4, 7, 0, 7, 3, 1, 6, 5, 1, 2, 6};
  std::memcpy(V, v, sizeof(v));
  std::memcpy(F, f, sizeof(f));
}

// cpp/src/geometry/geometry_helpers.cpp:40-81 (Fibonacci sphere of unit radius)
void orc_sphere_pointcloud(int n, double* P, double* N, double* radii) {
  const double ga = kPi * (3.0 - std::sqrt(5.0));
  const double area = 4.0 * kPi / n;
  const double rs = std::sqrt(area / kPi);
  for (int i = 0; i < n; ++i) {
    double y = 1.0 - (2.0 * i) / (n - 1.0);
    double ry = std::sqrt(1.0 - y * y);
    double th = ga * i;
    double x = ry * std::cos(th), z = ry * std::sin(th);
    P[3 * i] = x; P[3 * i + 1] = y; P[3 * i + 2] = z;
    N[3 * i] = x; N[3 * i + 1] = y; N[3 * i + 2] = z;
    radii[i] = rs;
  }
}

}  // extern "C"
