// Host-side utilities of libpointint_hip.so: error state, Gauss-Legendre
// nodes, Lebedev rule files, k-grid construction and the triangle-patch
// builder used by the shared-vertex form-factor kernel.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "pii_internal.hpp"
#include "pii_host.hpp"

namespace pii {

namespace {
thread_local std::string t_last_error;
}

void set_error(const std::string& msg) { t_last_error = msg; }
void clear_error() { t_last_error.clear(); }
const char* last_error() { return t_last_error.c_str(); }

// Gauss-Legendre on [a,b]: Newton iteration on P_n from the standard initial
// guess cos(pi (i + 3/4) / (n + 1/2)), stopping at |dz| <= 1e-14, weights
// 2 / ((1 - z^2) P_n'(z)^2), nodes ascending.  Same node/weight definition as
// cpp/include/quadrature/gauss_legendre.hpp:16-62.
void gauss_legendre(int n, double a, double b, double* x, double* w) {
  const int half = (n + 1) / 2;
  const double mid = 0.5 * (a + b), rad = 0.5 * (b - a);
  for (int i = 0; i < half; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      // Three-term recurrence for P_n(z) and P_{n-1}(z).
      double pn = 1.0, pn1 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double pn2 = pn1;
        pn1 = pn;
        pn = ((2.0 * j - 1.0) * z * pn1 - (j - 1.0) * pn2) / j;
      }
      dp = n * (z * pn - pn1) / (z * z - 1.0);
      double dz = pn / dp;
      z -= dz;
      if (std::abs(dz) <= 1e-14) {
        // refresh the derivative at the converged root
        pn = 1.0; pn1 = 0.0;
        for (int j = 1; j <= n; ++j) {
          double pn2 = pn1;
          pn1 = pn;
          pn = ((2.0 * j - 1.0) * z * pn1 - (j - 1.0) * pn2) / j;
        }
        dp = n * (z * pn - pn1) / (z * z - 1.0);
        break;
      }
    }
    const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    x[i] = mid - rad * z;
    x[n - 1 - i] = mid + rad * z;
    w[i] = rad * wi;
    w[n - 1 - i] = rad * wi;
  }
}

int lebedev_load(const char* path, std::vector<double>& dirs, std::vector<double>& w) {
  std::ifstream in(path);
  if (!in) return PII_ERR_IO;
  dirs.clear();
  w.clear();
  std::string line;
  double sum = 0.0;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    double theta, phi, wt;
    if (!(ss >> theta >> phi >> wt)) continue;
    const double th = theta * (kPi / 180.0), ph = phi * (kPi / 180.0);
    dirs.push_back(std::cos(th) * std::sin(ph));
    dirs.push_back(std::sin(th) * std::sin(ph));
    dirs.push_back(std::cos(ph));
    w.push_back(wt);
    sum += wt;
  }
  if (w.empty()) return PII_ERR_IO;
  // Rules normalised to unit total weight are rescaled to the sphere area.
  if (std::abs(sum - 1.0) < 0.1)
    for (double& x : w) x *= 4.0 * kPi;
  return PII_OK;
}

void kgrid_build(const double* ldirs, const double* lw, int M, int nrad, double* dirs,
                 double* kmag, double* w) {
  std::vector<double> t(nrad), wt(nrad);
  gauss_legendre(nrad, 0.0, 0.5 * kPi, t.data(), wt.data());
  for (int ir = 0; ir < nrad; ++ir) {
    const double c = std::cos(t[ir]);
    const double sec2 = 1.0 / c / c;
    const double k = std::tan(t[ir]);
    for (int j = 0; j < M; ++j) {
      const size_t q = (size_t)ir * M + j;
      dirs[3 * q + 0] = ldirs[3 * j + 0];
      dirs[3 * q + 1] = ldirs[3 * j + 1];
      dirs[3 * q + 2] = ldirs[3 * j + 2];
      kmag[q] = k;
      w[q] = lw[j] * (wt[ir] * sec2);
    }
  }
}

// Antipodal pairing of a grid.  Nodes pair within one shell (equal kmag) when
// their directions sum to < 1e-12 and their weights agree to 1e-13 relative.
// Returns true when every node has a partner; rep lists one node per pair in
// ascending full-grid order.
bool kgrid_pairs(int Q, const double* dirs, const double* kmag, const double* w,
                 std::vector<int>& anti, std::vector<int>& rep) {
  anti.assign(Q, -1);
  rep.clear();
  // Group nodes by shell: consecutive runs of identical kmag.
  int beg = 0;
  bool ok = true;
  while (beg < Q) {
    int end = beg + 1;
    while (end < Q && kmag[end] == kmag[beg]) ++end;
    // Sort shell members by (x, y, z) so antipodes can be found by a
    // lexicographic lookup of the negated direction.
    std::vector<int> idx(end - beg);
    for (int i = beg; i < end; ++i) idx[i - beg] = i;
    // Exact lexicographic order (a strict weak ordering, as std::sort
    // requires); the 1e-12 tolerance applies only to the antipode match.
    std::sort(idx.begin(), idx.end(), [&](int a, int b) {
      const double* p = dirs + 3 * (size_t)a;
      const double* q = dirs + 3 * (size_t)b;
      if (p[0] != q[0]) return p[0] < q[0];
      if (p[1] != q[1]) return p[1] < q[1];
      return p[2] < q[2];
    });
    for (int i = beg; i < end && ok; ++i) {
      if (anti[i] >= 0) continue;
      const double neg[3] = {-dirs[3 * (size_t)i], -dirs[3 * (size_t)i + 1],
                             -dirs[3 * (size_t)i + 2]};
      // candidates: x within 1e-12 of -x_i, a contiguous run of the order
      auto it = std::lower_bound(idx.begin(), idx.end(), neg[0] - 1e-12,
                                 [&](int a, double x) { return dirs[3 * (size_t)a] < x; });
      int found = -1;
      for (; it != idx.end() && dirs[3 * (size_t)*it] <= neg[0] + 1e-12; ++it) {
        const double* d = dirs + 3 * (size_t)*it;
        double s = std::abs(d[0] - neg[0]) + std::abs(d[1] - neg[1]) + std::abs(d[2] - neg[2]);
        if (s < 1e-12 && *it != i && anti[*it] < 0) { found = *it; break; }
      }
      if (found < 0 || std::abs(w[found] - w[i]) > 1e-13 * std::abs(w[i])) {
        ok = false;
        break;
      }
      anti[i] = found;
      anti[found] = i;
    }
    if (!ok) break;
    beg = end;
  }
  if (!ok) {
    anti.assign(Q, -1);
    rep.resize(Q);
    for (int q = 0; q < Q; ++q) rep[q] = q;
    return false;
  }
  for (int q = 0; q < Q; ++q)
    if (q < anti[q]) rep.push_back(q);
  return true;
}

// Morton key of a unit direction on the octahedral map (full sphere; the
// lower hemisphere folds onto the square's corners), 16 bits per axis.
static uint32_t direction_key(const double* d) {
  const double n = std::abs(d[0]) + std::abs(d[1]) + std::abs(d[2]);
  double u = n > 0 ? d[0] / n : 0.0, v = n > 0 ? d[1] / n : 0.0;
  if (d[2] < 0) {
    const double fu = (1.0 - std::abs(v)) * (u >= 0 ? 1.0 : -1.0);
    const double fv = (1.0 - std::abs(u)) * (v >= 0 ? 1.0 : -1.0);
    u = fu;
    v = fv;
  }
  auto q16 = [](double x) {
    const double t = std::min(1.0, std::max(0.0, 0.5 * (x + 1.0)));
    return (uint32_t)std::min(65535.0, std::floor(t * 65536.0));
  };
  uint32_t a = q16(u), b = q16(v), key = 0;
  for (int bit = 0; bit < 16; ++bit) {
    key |= ((a >> bit) & 1u) << (2 * bit);
    key |= ((b >> bit) & 1u) << (2 * bit + 1);
  }
  return key;
}

// Representatives of a paired grid: the member of each pair on the upper
// hemisphere (z > 0, then y > 0, then x > 0), ordered inside each shell along
// the Morton curve of the direction, so that the 64 nodes of a wave lie close
// together on the sphere.  Unpaired grids keep every node, ordered the same
// way.  The order depends only on the directions, so shells with the same
// direction set get the same order.
void kgrid_order_reps(int Q, const double* dirs, const double* kmag, const std::vector<int>& anti,
                      std::vector<int>& rep) {
  auto upper = [&](int q) {
    const double* d = dirs + 3 * (size_t)q;
    if (d[2] != 0.0) return d[2] > 0.0;
    if (d[1] != 0.0) return d[1] > 0.0;
    return d[0] >= 0.0;
  };
  for (int& q : rep)
    if (anti[q] >= 0 && !upper(q)) q = anti[q];
  size_t beg = 0;
  while (beg < rep.size()) {
    size_t end = beg + 1;
    while (end < rep.size() && kmag[rep[end]] == kmag[rep[beg]]) ++end;
    std::vector<std::pair<uint32_t, int>> keyed(end - beg);
    for (size_t i = beg; i < end; ++i)
      keyed[i - beg] = {direction_key(dirs + 3 * (size_t)rep[i]), rep[i]};
    // equal keys fall back to the node index (shell-major, so the same order
    // in every shell of a product grid)
    std::sort(keyed.begin(), keyed.end());
    for (size_t i = beg; i < end; ++i) rep[i] = keyed[i - beg].second;
    beg = end;
  }
  (void)Q;
}

// Greedy patch builder.  Faces are visited in breadth-first order over
// shared vertices; each patch grows from an unassigned seed by repeatedly
// taking the frontier face that adds the fewest new vertices, until either
// kPatchVerts vertices or kPatchFaces faces would be exceeded.
void build_patches(int nv, int nf, const int32_t* F, std::vector<int>& hdr,
                   std::vector<int>& pverts, std::vector<uint32_t>& pfaces,
                   std::vector<uint32_t>& ploc) {
  hdr.clear(); pverts.clear(); pfaces.clear(); ploc.clear();
  if (nf == 0) return;
  // vertex -> faces (CSR)
  std::vector<int> vstart(nv + 1, 0);
  for (int f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) vstart[F[3 * f + c] + 1]++;
  for (int v = 0; v < nv; ++v) vstart[v + 1] += vstart[v];
  std::vector<int> vfaces(vstart[nv]);
  {
    std::vector<int> fill(vstart.begin(), vstart.end() - 1);
    for (int f = 0; f < nf; ++f)
      for (int c = 0; c < 3; ++c) vfaces[fill[F[3 * f + c]]++] = f;
  }
  // Global BFS order supplies the seeds.
  std::vector<int> order;
  order.reserve(nf);
  {
    std::vector<char> seen(nf, 0);
    for (int s = 0; s < nf; ++s) {
      if (seen[s]) continue;
      size_t head = order.size();
      order.push_back(s);
      seen[s] = 1;
      while (head < order.size()) {
        int f = order[head++];
        for (int c = 0; c < 3; ++c) {
          int v = F[3 * f + c];
          for (int i = vstart[v]; i < vstart[v + 1]; ++i) {
            int g = vfaces[i];
            if (!seen[g]) { seen[g] = 1; order.push_back(g); }
          }
        }
      }
    }
  }
  std::vector<char> assigned(nf, 0);
  std::vector<int> vslot(nv, -1);     // vertex -> local slot in the current patch
  std::vector<int> cand;              // frontier faces of the current patch
  std::vector<int> in_cand(nf, -1);   // patch id that queued the face
  int pid = 0;
  for (int seed : order) {
    if (assigned[seed]) continue;
    const int vbeg = (int)pverts.size(), fbeg = (int)pfaces.size();
    std::vector<int> local_verts;
    cand.clear();
    auto new_verts = [&](int f) {
      int n = 0;
      for (int c = 0; c < 3; ++c) {
        int v = F[3 * f + c];
        bool dup = false;
        for (int d = 0; d < c; ++d) dup |= (F[3 * f + d] == v);
        if (vslot[v] < 0 && !dup) ++n;
      }
      return n;
    };
    auto add_face = [&](int f) {
      uint32_t loc = 0;
      for (int c = 0; c < 3; ++c) {
        int v = F[3 * f + c];
        if (vslot[v] < 0) {
          vslot[v] = (int)local_verts.size();
          local_verts.push_back(v);
        }
        loc |= (uint32_t)vslot[v] << (8 * c);
      }
      assigned[f] = 1;
      pfaces.push_back((uint32_t)f);
      ploc.push_back(loc);
      for (int c = 0; c < 3; ++c) {
        int v = F[3 * f + c];
        for (int i = vstart[v]; i < vstart[v + 1]; ++i) {
          int g = vfaces[i];
          if (!assigned[g] && in_cand[g] != pid) { in_cand[g] = pid; cand.push_back(g); }
        }
      }
    };
    add_face(seed);
    for (;;) {
      if ((int)pfaces.size() - fbeg >= kPatchFaces) break;
      int best = -1, best_new = 4, best_pos = -1;
      for (int i = 0; i < (int)cand.size(); ++i) {
        int g = cand[i];
        if (assigned[g]) continue;
        int nn = new_verts(g);
        if (nn < best_new) { best_new = nn; best = g; best_pos = i; if (nn == 0) break; }
      }
      if (best < 0) break;
      if ((int)local_verts.size() + best_new > kPatchVerts) break;
      cand[best_pos] = cand.back();
      cand.pop_back();
      add_face(best);
    }
    for (int v : local_verts) { pverts.push_back(v); vslot[v] = -1; }
    hdr.push_back(vbeg);
    hdr.push_back((int)local_verts.size());
    hdr.push_back(fbeg);
    hdr.push_back((int)pfaces.size() - fbeg);
    ++pid;
  }
}

}  // namespace pii
