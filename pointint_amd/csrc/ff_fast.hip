// Fast form factor for gfx950: FP64 dot products, phases and range reduction;
// the hardware FP32 sine/cosine on the reduced phases, FP32 for the complex
// combination and the small-span Taylor branch; FP64 accumulation across
// patches.
//
// Triangles (meshes and triangle soups) use a cancellation-free form of the
// reference's element transform (cpp/src/element_functions.cpp:63-70 with the
// accumulation of cpp/src/dof/triangle_mesh_dof.cpp:36-71):
//
//   A_tri(k) = gamma e^{i k.a} Phi(k.e1, k.e2) = gamma * 2 g[u0, u1, u2]
//
// where u_j = k.v_j, g(z) = -e^{iz} and g[...] is the second divided
// difference.  With half phasors Q_j = e^{i u_j / 2}, m the node whose u lies
// between the other two (p, q):
//
//   g[p, m, q] = -i Q_m (Q_q sinc(h_q) - Q_p sinc(h_p)) / (u_q - u_p),
//   h_p = (u_m - u_p)/2,  h_q = (u_q - u_m)/2,
//
// and sin(h) = Im(Q_b conj Q_a) comes from the phasors themselves.  On a lane
// whose span u_q - u_p is below kLaneTaylor the divided difference is summed
// from its Taylor series instead:  Phi(a, b) = sum_n 2 i^n h_n(a, b) /
// (n+2)!, h_n the complete homogeneous symmetric polynomial.
//
// A workgroup (4 waves) owns 64 k-nodes of one object; for triangles each
// wave owns 16 of them and works alone (no barrier inside the patch loop).
// Faces are grouped in patches of <= 64 vertices.  Per patch a wave fills its
// own LDS table of the vertices' half phasors for its 16 nodes, then takes
// the patch's faces 16 at a time: the dot products k.v (vertex phases) and
// alpha = k.e1, beta = k.e2 (FP64) and gamma = khat.2S (FP32) come from the
// matrix cores, 16 faces x 16 nodes per product with K = (x, y, z, pad); each
// lane then evaluates 4 (face, node) terms, its node and one face of each
// group of 4 faces, and reads its three phasors from the table.
#include "device_common.hpp"
#include "j1_modphase.hpp"
#include "pii_internal.hpp"

namespace pii {

namespace {

constexpr int kLanes = 64;
constexpr int kWaves = 4;
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
// Lanes with a span below kLaneTaylor take the Taylor sum (7 terms suffice
// there); the others, spans >= 0.3, take the main formula, whose FP32 phasor
// error over the span stays below 1e-6 of the term.  (A wave-uniform Taylor
// path for faces whose whole wave has a small span, chosen per face from a
// span bound in the record, cost 8-10 % on its compare-and-branch alone at
// the high |k| this kernel now serves, where it almost never applies.)
constexpr float kLaneTaylor = 0.3f;
constexpr float kSincPoly = 0.5f;     // sinc(h) by polynomial below |h| = 0.5

// e^{i 2 pi r} for r in revolutions (|r| <= 0.5), FP32 result, on the
// hardware sine/cosine (revolution input).  Measured on MI355X over
// [-0.5, 0.5] revolutions: max abs error 1.25e-7, the same with or without a
// quadrant reduction to |f| <= 1/8 revolution first, so none is made.
__device__ __forceinline__ float2 phasor_rev_hw(double r) {
  const float f = (float)r;
  return make_float2(__builtin_amdgcn_cosf(f), __builtin_amdgcn_sinf(f));
}

// e^{i u} (full phasor) of an FP64 phase.
__device__ __forceinline__ float2 full_phasor_hw(double u) {
  const double t = u * 0.15915494309189535;  // 1/(2 pi)
  return phasor_rev_hw(t - rint(t));
}

// J1(x)/x = sum_m (-1)^m y^m / (2 m! (m+1)!), y = x^2/4, Horner in FP32 with
// terms m <= 10: truncation < 1e-10 for x <= 4; rounding ~1e-7 absolute
// (the value itself is at most 1/2).
__device__ __forceinline__ float j1ox_small(float x) {
  const float y = 0.25f * x * x;
  float s = 3.451844740057556e-15f;           // 1/(2 10! 11!)
  s = s * y - 3.7970292140633116e-13f;         // 1/(2 9! 10!)
  s = s * y + 3.417326292656981e-11f;          // 1/(2 8! 9!)
  s = s * y - 2.460474930713026e-09f;          // 1/(2 7! 8!)
  s = s * y + 1.3778659611992944e-07f;         // 1/(2 6! 7!)
  s = s * y - 5.787037037037037e-06f;          // 1/(2 5! 6!)
  s = s * y + 1.7361111111111112e-04f;         // 1/(2 4! 5!)
  s = s * y - 3.472222222222222e-03f;          // 1/(2 3! 4!)
  s = s * y + 4.1666666666666664e-02f;         // 1/(2 2! 3!)
  s = s * y - 0.25f;                           // 1/(2 1! 2!)
  return s * y + 0.5f;
}

// sinc(h) / 2 from w = 2 h and sin(h): polynomial below |h| = kSincPoly,
// sin(h) / w above (both evaluated, then selected, so lanes do not diverge).
// The per-face terms are accumulated at half scale (their factor 2 is
// applied once per patch), so the half spans need no multiplication by 1/2.
__device__ __forceinline__ float half_sinc_from(float w, float sin_h) {
  // (1 - h^2/6 + h^4/120 - h^6/5040) / 2 with h = w/2: the first omitted
  // term, h^8/9!, is below 1.1e-8 for |h| < 0.5
  const float w2 = w * w;
  float p = w2 * (-1.0f / 645120.0f) + (1.0f / 3840.0f);
  p = p * w2 - (1.0f / 48.0f);
  p = p * w2 + 0.5f;
  const float r = sin_h * __builtin_amdgcn_rcpf(w);
  return fabsf(w) < 2.0f * kSincPoly ? p : r;
}

// Phi(a, b) = sum_{n<NT} c_n h_n(a, b), c_n = 2 i^n / (n+2)!, in FP32 from
// the FP32-rounded alpha, beta (every term is positive-weighted and the sum
// is near 1 for spans below kLaneTaylor, so FP32 carries it to 1e-7), times
// (g2/2) e^{i u0} with e^{i u0} = Q0^2, accumulated at half scale.
// |h_n| <= (n+1) H^n for span H, so the truncation error is bounded by the
// first omitted term 2(NT+1) H^NT / (NT+2)!.
template <int NT>
__device__ __forceinline__ void tri_taylor(float a, float b, float g2, float2 Q0, float2& acc) {
  constexpr float c[12] = {1.0f, 1.0f / 3.0f, -1.0f / 12.0f, -1.0f / 60.0f, 1.0f / 360.0f,
                           1.0f / 2520.0f, -1.0f / 20160.0f, -1.0f / 181440.0f,
                           1.0f / 1814400.0f, 1.0f / 19958400.0f, -1.0f / 239500800.0f,
                           -1.0f / 3113510400.0f};
  float an = 1.0f, hn = 1.0f, re = 1.0f, im = 0.0f;
#pragma unroll
  for (int n = 1; n < NT; ++n) {
    an *= a;
    hn = fmaf(b, hn, an);
    if (n & 1) im = fmaf(hn, c[n], im);
    else re = fmaf(hn, c[n], re);
  }
  const float px = Q0.x * Q0.x - Q0.y * Q0.y;
  const float py = 2.0f * Q0.x * Q0.y;
  const float tr = re * (0.25f * g2), ti = im * (0.25f * g2);
  acc.x = fmaf(px, tr, fmaf(-py, ti, acc.x));
  acc.y = fmaf(px, ti, fmaf(py, tr, acc.y));
}

// Main branch: middle node m and the other two (p, q); see the file comment.
// The formula is symmetric in p and q, so only the middle node matters: u_m
// is the median of (0, a, b) (v_med3), so m is node 0 when that median is 0
// (0 lies between a and b), else the one of nodes 1, 2 whose phase is
// smaller in magnitude.  Q_p = Q1 when m = 0 (else Q0), Q_q = Q1 when m = 2
// (else Q2), each read from the phasor table through its vertex's row
// pointer (r0, r1, r2).  Half spans and the span are formed in FP32 from the
// FP32-rounded alpha, beta: their absolute error, 6e-8 |k| diam, changes the
// result by a relative 6e-8 at most, because the span D itself is at least
// that large here.
__device__ __forceinline__ void tri_main(double al, double be, float g2, const float2* r0,
                                         const float2* r1, const float2* r2, float2& acc) {
  const float a = (float)al, b = (float)be;
  const float um = __builtin_amdgcn_fmed3f(0.0f, a, b);
  const bool m0 = um == 0.0f;
  const bool m2 = (!m0) & (fabsf(b) <= fabsf(a));
  const float up = m0 ? a : 0.0f;
  const float uq = m2 ? a : b;
  // the three phasors read in (p, m, q) order: the selects act on the
  // LDS addresses, not on the phasors
  const float2* rm = m0 ? r0 : (m2 ? r2 : r1);
  const float2* rp = m0 ? r1 : r0;
  const float2* rq = m2 ? r1 : r2;
  const float2 Qm = *rm, Qp = *rp, Qq = *rq;
  const float wp = um - up;                        // 2 hp
  const float wq = uq - um;                        // 2 hq
  const float D = uq - up;
  const float sin_p = Qm.y * Qp.x - Qm.x * Qp.y;   // Im(Qm conj Qp) = sin(hp)
  const float sin_q = Qq.y * Qm.x - Qq.x * Qm.y;   // Im(Qq conj Qm) = sin(hq)
  const float sp = half_sinc_from(wp, sin_p);      // sinc(hp) / 2
  const float sq = half_sinc_from(wq, sin_q);
  const float nx = Qq.x * sq - Qp.x * sp;
  const float ny = Qq.y * sq - Qp.y * sp;
  const float tx = Qm.x * nx - Qm.y * ny;
  const float ty = Qm.x * ny + Qm.y * nx;
  const float f = g2 * __builtin_amdgcn_rcpf(D);
  // -i * (tx + i ty) = ty - i tx
  acc.x = fmaf(ty, f, acc.x);
  acc.y = fmaf(-tx, f, acc.y);
}

// gamma * e^{i u0} * Phi(alpha, beta) with a per-lane branch on the span
// max(0, a, b) - min(0, a, b) = max(|a|, |b|, |a - b|), taken in FP32 (both
// sides of the threshold are accurate near it).  Lanes below kLaneTaylor
// take the 7-term Taylor sum (first omitted term < 1e-8 there).  The two
// branches keep separate accumulators, so neither copies the other's.
__device__ __forceinline__ void tri_term(double al, double be, float g2, const float2* r0,
                                         const float2* r1, const float2* r2, float2& tacc,
                                         float2& macc) {
  const float a = (float)al, b = (float)be;
  const float H = fmaxf(fmaxf(fabsf(a), fabsf(b)), fabsf(a - b));
  if (H < kLaneTaylor) tri_taylor<7>(a, b, g2, *r0, tacc);
  else tri_main(al, be, g2, r0, r1, r2, macc);
}

// J1(x)/x for x >= 4 in FP32 from the modulus and phase of J1 + i Y1:
// J1(x) = m(y) sqrt(y) cos(x - 3 pi/4 + 2 pi p(y)), y = 1/x, with m and p the
// degree-7 fits of j1_modphase.hpp (m relative 2.3e-9, p 6e-11 revolutions).
// The phase in revolutions, x/(2 pi) - 3/8 + p, takes x/(2 pi) as the FP32
// product plus its rounding error (an FMA) plus x times the low part of
// 1/(2 pi): after the reduction to [-1/2, 1/2] it carries only the rounding
// of that last sum (6e-8 revolutions), whatever the size of x.  Against
// scipy.special.j1 in an FP32 emulation: 3.6e-8 absolute on [4, 400].
__device__ __forceinline__ float j1ox_large(float x) {
  constexpr float kInv2PiHi = 0.15915493667125702f;  // fl32(1/(2 pi))
  constexpr float kInv2PiLo = 6.420638327e-09f;      // 1/(2 pi) - kInv2PiHi
  const float y = __builtin_amdgcn_rcpf(x);
  const float sy = __builtin_amdgcn_rsqf(x);           // sqrt(y)
  float m = kJ1Mod[kJ1ModPhaseDeg];
  float p = kJ1Phase[kJ1ModPhaseDeg];
#pragma unroll
  for (int i = kJ1ModPhaseDeg - 1; i >= 0; --i) {
    m = fmaf(m, y, kJ1Mod[i]);
    p = fmaf(p, y, kJ1Phase[i]);
  }
  const float r1 = x * kInv2PiHi;
  const float e1 = fmaf(x, kInv2PiHi, -r1);  // rounding error of r1
  float f = (r1 - rintf(r1)) + (fmaf(x, kInv2PiLo, e1) + (p - 0.375f));
  f -= rintf(f);
  return m * (y * sy) * __builtin_amdgcn_cosf(f);
}

// Per-lane operands of one block of 16 faces for lane (fq, nl): component
// min(fq, 2) of face nl's e1, e2 (rows of the FP64 products, whose element r
// in lane (fq, nl) is row fq + 4 r; the fourth row of the B operand is 0), of
// face (nl >> 2) + 4 (nl & 3)'s 2S (rows of the FP32 product, whose element
// r is row 4 fq + r, which must hold face fq + 4 r), and the table offsets
// of the faces fq + 4 r the lane evaluates.  The loads go through a buffer
// resource over the block's records that ends at the object's last record:
// rows past the patch read the next patch's records, rows past the object
// read 0 (the hardware's range check); the kernel masks both rows' terms.
// A value at a 32-bit byte offset from a uniform base (saddr + voffset form).
template <class T>
__device__ __forceinline__ T ldg(const PII_GLOBAL_AS char* base, uint32_t off) {
  return *reinterpret_cast<const PII_GLOBAL_AS T*>(base + off);
}

struct FaceOffsets {
  uint32_t a, g, o[4];  // byte offsets in a block's records
  __device__ __forceinline__ FaceOffsets(int fq, int nl) {
    constexpr uint32_t rb = kRecFastTri * sizeof(double);
    const uint32_t c = (uint32_t)min(fq, 2);
    a = (uint32_t)nl * rb + 8u * c;
    g = (uint32_t)((nl >> 2) + 4 * (nl & 3)) * rb + 48u + 4u * c;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = (uint32_t)(4 * r + fq) * rb + 64u;
  }
};
typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
// Buffer resource over records [first, min(first + 16, n)) of an object's
// fast-triangle records rec (uniform; 0x00020000: the gfx9 raw-buffer format
// word).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t block_rsrc(const double* rec, int first,
                                                             int n) {
  constexpr int kBlockBytes = 16 * kRecFastTri * (int)sizeof(double);
  // every input is uniform; readfirstlane keeps the descriptor in SGPRs
  const int rem = __builtin_amdgcn_readfirstlane(n - first);
  const int bytes = rem >= 16 ? kBlockBytes : max(rem, 0) * kRecFastTri * (int)sizeof(double);
  const uint64_t base = reinterpret_cast<uint64_t>(rec + (size_t)kRecFastTri * first);
  const uint64_t ub = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(base >> 32)) << 32) |
                      (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base);
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(ub), (short)0, bytes,
                                           0x00020000);
}
struct FaceBlock {
  double e1, e2;
  float sg;
  uint32_t o[4][3];
  __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t rs, const FaceOffsets& f) {
    e1 = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, f.a, 0, 0));
    e2 = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, f.a, 24, 0));
    sg = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, f.g, 0, 0));
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const u32x3 v =
          __builtin_bit_cast(u32x3, __builtin_amdgcn_raw_buffer_load_b96(rs, f.o[r], 0, 0));
      o[r][0] = v.x;
      o[r][1] = v.y;
      o[r][2] = v.z;
    }
  }
};

// Five waves per SIMD: the table (32 KB per workgroup) allows five workgroups
// per compute unit, and registers are held to that.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
k_ff_fast(const ObjDesc* __restrict__ objs,
                                                 const double* __restrict__ krep, int Qh,
                                                 int out_mode, double2* __restrict__ Aout,
                                                 int ldA, float2* __restrict__ Xout, int ldX,
                                                 unsigned* __restrict__ colmax,
                                                 double2* __restrict__ part, int ldP) {
  // Per-wave vertex phasor tables (triangles); after the patch loop its first
  // 4 KB hold the cross-wave reduction (behind a barrier).
  __shared__ float2 tab[kWaves][kPatchVerts][kWaveNodes];
  static_assert(sizeof(double2) * kWaves * kLanes <= sizeof(tab),
                "reduction buffer must fit in the phasor table");
  static_assert(kWaves * kWaveNodes == kLanes, "the waves' nodes make the block's");
  double2 (*red)[kLanes] = reinterpret_cast<double2 (*)[kLanes]>(&tab[0][0][0]);
  const BlockObj bo = xcd_block_map();
  const ObjDesc& d = objs[bo.obj];
  const int lane = threadIdx.x & (kLanes - 1);
  // wave index as an SGPR value: loop indices derived from it stay uniform,
  // so element records are fetched with scalar loads
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // This object's nodes start at qbeg (the moment series wrote those below).
  // Both exits are uniform over the workgroup.
  const int q0 = d.qbeg + bo.blk * kLanes;
  if (q0 >= (out_mode == 0 ? Qh : ldX)) return;
  const int q = q0 + lane;
  if (q0 >= Qh) {  // padding columns of X only (the finishing kernel's, when split)
    if (!part && wave == 0 && q < ldX) Xout[(size_t)d.col * ldX + q] = make_float2(0.f, 0.f);
    return;
  }
  const int qc = q < Qh ? q : Qh - 1;
  double accx = 0.0, accy = 0.0;

  if (d.type == PII_GEOM_TRIANGLE) {
    // Lane = (face group fq, node nl): node q0 + 16 wave + nl, and in every
    // round the face fb + 4 r + fq of the current 16-face block.
    const int fq = lane >> 4, nl = lane & (kWaveNodes - 1);
    const int qt = min(q0 + wave * kWaveNodes + nl, Qh - 1);
    // B operands of the products: component fq of this lane's node (0 in
    // the fourth row, so the A operands' fourth column is free)
    const double bk = fq < 3 ? krep[(size_t)fq * Qh + qt] : 0.0;
    const float bh = fq < 3 ? (float)krep[(size_t)(3 + fq) * Qh + qt] : 0.0f;
    const dbl4 z4 = {0.0, 0.0, 0.0, 0.0};
    const flt4 z4f = {0.0f, 0.0f, 0.0f, 0.0f};
    const float2* tw = &tab[wave][0][nl];  // this lane's node column of the wave's table
    const FaceOffsets fo(fq, nl);
    const auto* hdr = as_const(d.patch_hdr);
    const auto* pverts = reinterpret_cast<const PII_GLOBAL_AS char*>(as_global(d.patch_verts));
    const auto* world = reinterpret_cast<const PII_GLOBAL_AS char*>(as_global(d.world));
    const uint32_t wc = 8u * (uint32_t)min(fq, 2);  // byte offset of the lane's component
    // patches [pb, pe) of this workgroup's split (all when not split)
    const int pchunk = (d.npatch + (int)gridDim.z - 1) / (int)gridDim.z;
    const int pb = (int)blockIdx.z * pchunk, pe = min(d.npatch, pb + pchunk);
    // Loads run one step ahead: the next patch's vertex ids, and the next
    // face block's operands (past a patch's last block: the next patch's
    // first, since the patches' faces are consecutive in the records).  The
    // face blocks alternate between two sets of registers.
    int gvn[4];
    FaceBlock fa, fz;
    if (pb < pe) {
      const int vbeg = hdr[4 * pb], vcnt = hdr[4 * pb + 1];
#pragma unroll
      for (int vb = 0; vb < 4; ++vb)
        gvn[vb] = ldg<int>(pverts, 4u * (uint32_t)(vbeg + min(vb * 16 + nl, vcnt - 1)));
      fa.load(block_rsrc(d.rec, hdr[4 * pb + 2], d.n), fo);
    }
    for (int p = pb; p < pe; ++p) {
      const int vcnt = hdr[4 * p + 1];
      const int fbeg = hdr[4 * p + 2], fcnt = hdr[4 * p + 3];
      // this patch's vertex coordinates, then the next patch's vertex ids
      double wv[4];
#pragma unroll
      for (int vb = 0; vb < 4; ++vb) wv[vb] = ldg<double>(world, 24u * (uint32_t)gvn[vb] + wc);
      if (p + 1 < pe) {
        const int vbeg1 = hdr[4 * p + 4], vcnt1 = hdr[4 * p + 5];
#pragma unroll
        for (int vb = 0; vb < 4; ++vb)
          gvn[vb] = ldg<int>(pverts, 4u * (uint32_t)(vbeg1 + min(vb * 16 + nl, vcnt1 - 1)));
      }
      // Vertex phases t = u / (4 pi), u = k.v, 16 vertices x 16 nodes per
      // product; lane (fq, nl) gets vertex 16 vb + fq + 4 r of node nl in
      // element r.  The half phasor e^{i u/2} is e^{2 pi i t}.
#pragma unroll
      for (int vb = 0; vb < 4; ++vb) {
        if (vb * 16 < vcnt) {
          const dbl4 t = __builtin_amdgcn_mfma_f64_16x16x4f64(wv[vb] * 0.07957747154594767, bk,
                                                              z4, 0, 0, 0);
#pragma unroll
          for (int r = 0; r < 4; ++r)
            tab[wave][vb * 16 + fq + 4 * r][nl] = phasor_rev_hw(t[r] - rint(t[r]));
        }
      }
      float2 tacc = make_float2(0.f, 0.f), macc = make_float2(0.f, 0.f);
      // one block of 16 faces (m of them in the patch, uniform): the three
      // products, then the four rounds; in the patch's last block, lanes
      // whose face lies past the patch sit out
      auto face_block = [&](const FaceBlock& cur, int m) {
        const dbl4 A1 = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.e1, bk, z4, 0, 0, 0);
        const dbl4 A2 = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.e2, bk, z4, 0, 0, 0);
        const flt4 G = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.sg, bh, z4f, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (4 * r < m && 4 * r + fq < m)
            tri_term(A1[r], A2[r], G[r], tw + cur.o[r][0], tw + cur.o[r][1], tw + cur.o[r][2],
                     tacc, macc);
        }
      };
      // the block after face fb's: the patch's next, or the next patch's first
      auto after = [&](int fb) { return block_rsrc(d.rec, fbeg + min(fb + 16, fcnt), d.n); };
      for (int fb = 0; fb < fcnt; fb += 32) {
        fz.load(after(fb), fo);
        face_block(fa, fcnt - fb);
        if (fb + 16 >= fcnt) {  // odd block count: the next patch's first is in fz
          fa = fz;
          break;
        }
        fa.load(after(fb + 16), fo);
        face_block(fz, fcnt - fb - 16);
      }
      // the terms' factor 2 (half-scale sums)
      accx += 2.0 * ((double)tacc.x + (double)macc.x);
      accy += 2.0 * ((double)tacc.y + (double)macc.y);
    }
    // node totals: sum the four face groups; lane l of wave w then keeps
    // node l when l lies in w's 16 (the cross-wave sum below adds zeros)
    accx += __shfl_xor(accx, 16);
    accy += __shfl_xor(accy, 16);
    accx += __shfl_xor(accx, 32);
    accy += __shfl_xor(accy, 32);
    if ((lane >> 4) != wave) { accx = 0.0; accy = 0.0; }
    __syncthreads();  // every wave is done with its table (red aliases it)
  } else {
    // Disks / Gaussians, one node per lane (q0 + lane) and a quarter of the
    // elements per wave: A = S(|k'_perp|) (k'.n)/|k| e^{i k.c'}, k' = A^T k,
    // with c' and A n precomputed (cpp/src/dof/affine_dof.cpp:103-167).
    const double kx = krep[qc], ky = krep[(size_t)Qh + qc], kz = krep[2 * (size_t)Qh + qc];
    const auto* p = as_const(d.dofs);
    const double kpx = p[3] * kx + p[6] * ky + p[9] * kz;
    const double kpy = p[4] * kx + p[7] * ky + p[10] * kz;
    const double kpz = p[5] * kx + p[8] * ky + p[11] * kz;
    const double kp2 = kpx * kpx + kpy * kpy + kpz * kpz;
    const double km = krep[6 * (size_t)Qh + qc];
    const double inv_km = 1.0 / km;
    const bool disk = (d.type == PII_GEOM_DISK);
    // |k'| <= kmax ||A||_F for every lane: bounds rho |k'_perp| per disk.
    double kmax = km;
    for (int off = 32; off > 0; off >>= 1) kmax = fmax(kmax, __shfl_xor(kmax, off));
    double an2 = 0.0;
    for (int i = 3; i < 12; ++i) an2 += p[i] * p[i];
    const float kbound = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(
        (float)(kmax * sqrt(an2) * (1.0 + 1e-6)))));
    float px = 0.f, py = 0.f;
    int cnt = 0;
    const auto* rec = as_const(d.rec);
    const int echunk = (d.n + (int)gridDim.z - 1) / (int)gridDim.z;
    const int eb = (int)blockIdx.z * echunk, ee = min(d.n, eb + echunk);
    for (int e = eb + wave; e < ee; e += kWaves) {
      const auto* r = rec + (size_t)kRecSurfel * e;
      const double kdn = kx * r[3] + ky * r[4] + kz * r[5];
      const double r2 = kp2 - kdn * kdn;
      float S;
      if (disk) {
        const float rho = (float)r[6];
        const float x = rho * sqrtf((float)fmax(0.0, r2));
        float j;
        if (rho * kbound < 4.0f) {
          j = j1ox_small(x);          // every lane has x < 4
        } else if (x < 4.0f) {
          j = j1ox_small(x);
        } else {
          j = j1ox_large(x);
        }
        S = 2.0f * (float)r[7] * j;
      } else {
        // no clamp of r2, as on the reference's CPU path (affine_dof.cpp:152)
        S = (float)r[7] * __expf((float)(-0.5 * r[6] * r[6] * r2));
      }
      const float Ap = S * (float)(kdn * inv_km);
      const float2 ph = full_phasor_hw(kx * r[0] + ky * r[1] + kz * r[2]);
      px += Ap * ph.x;
      py += Ap * ph.y;
      if (++cnt == 64) {
        accx += (double)px; accy += (double)py;
        px = py = 0.f; cnt = 0;
      }
    }
    accx += (double)px;
    accy += (double)py;
  }
  // |k| = 0 (32-bit offsets from the grid's base: no 64-bit index stays live
  // across the patch loop)
  const auto* kg = reinterpret_cast<const PII_GLOBAL_AS char*>(as_global(krep));
  if (ldg<double>(kg, 8u * (6u * (uint32_t)Qh + (uint32_t)qc)) < 1e-10) { accx = 0.0; accy = 0.0; }

  // Sum the four waves' partial sums of the same 64 nodes.
  red[wave][lane] = make_double2(accx, accy);
  __syncthreads();
  if (wave == 0) {
    double sx = 0.0, sy = 0.0;
    for (int w = 0; w < kWaves; ++w) { sx += red[w][lane].x; sy += red[w][lane].y; }
    if (part) {  // split: partial sums, finished by k_ff_finish
      part[((size_t)blockIdx.z * gridDim.y + bo.obj) * ldP + (size_t)bo.blk * kLanes + lane] =
          make_double2(sx, sy);
      return;
    }
    if (out_mode == 0) {
      if (q < Qh) Aout[(size_t)d.col * ldA + q] = make_double2(sx, sy);
    } else {
      const double sw = sqrt(ldg<double>(kg, 8u * (7u * (uint32_t)Qh + (uint32_t)qc)));
      const float2 x = make_float2((float)(sx * sw), (float)(sy * sw));
      if (q < ldX) Xout[(size_t)d.col * ldX + q] = (q < Qh) ? x : make_float2(0.f, 0.f);
      // column maximum of |X| (all 64 lanes take part in the shuffles)
      float m = (q < Qh) ? fmaxf(fabsf(x.x), fabsf(x.y)) : 0.0f;
      for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
      if (lane == 0) atomicMax(colmax + d.col, __float_as_uint(m));
    }
  }
}

// Finishes a split launch: node i of object obj (q = qbeg + i) sums the
// nsplit partials in split order, then writes A or X and the column maximum
// as the unsplit epilogue of k_ff_fast.
__global__ void __launch_bounds__(256) k_ff_finish(const ObjDesc* __restrict__ objs,
                                                   const double* __restrict__ krep, int Qh,
                                                   int out_mode, double2* __restrict__ Aout,
                                                   int ldA, float2* __restrict__ Xout, int ldX,
                                                   unsigned* __restrict__ colmax,
                                                   const double2* __restrict__ part, int ldP,
                                                   int nsplit) {
  const int obj = blockIdx.y;
  const ObjDesc& d = objs[obj];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = d.qbeg + i;
  const int limit = out_mode == 0 ? Qh : ldX;
  const bool valid = i < ldP && q < limit;
  double sx = 0.0, sy = 0.0;
  if (valid && q < Qh) {
    const double2* p = part + (size_t)obj * ldP + i;
    const size_t stride = (size_t)gridDim.y * ldP;
    for (int z = 0; z < nsplit; ++z) {
      const double2 v = p[(size_t)z * stride];
      sx += v.x;
      sy += v.y;
    }
  }
  if (out_mode == 0) {
    if (valid && q < Qh) Aout[(size_t)d.col * ldA + q] = make_double2(sx, sy);
    return;
  }
  float m = 0.0f;
  if (valid) {
    float2 x = make_float2(0.f, 0.f);
    if (q < Qh) {
      const double sw = sqrt(krep[7 * (size_t)Qh + q]);
      x = make_float2((float)(sx * sw), (float)(sy * sw));
      m = fmaxf(fabsf(x.x), fabsf(x.y));
    }
    Xout[(size_t)d.col * ldX + q] = x;
  }
  // column maximum over the wave (every lane takes part in the shuffles)
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(colmax + d.col, __float_as_uint(m));
}

// Full-grid expansion of representative-node values: A(-k) = -conj(A(k)).
__global__ void __launch_bounds__(256) k_expand(const double2* __restrict__ Arep, int ldA,
                                                const int* __restrict__ rep_index,
                                                const int* __restrict__ anti, int Qh, int Q,
                                                int N, double2* __restrict__ Afull) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int col = blockIdx.y;
  if (i >= Qh || col >= N) return;
  const double2 a = Arep[(size_t)col * ldA + i];
  const int q = rep_index[i];
  Afull[(size_t)col * Q + q] = a;
  const int qa = anti[q];
  if (qa >= 0 && qa != q) Afull[(size_t)col * Q + qa] = make_double2(-a.x, a.y);
}

}  // namespace

size_t ff_split_doubles(int N, int span, int nsplit) {
  const size_t ldP = (size_t)(span + kLanes - 1) / kLanes * kLanes;
  return nsplit > 1 ? 2 * (size_t)nsplit * N * ldP : 0;
}

int ff_fast_slots() {
  static std::atomic<int> cache[16];
  return resident_workgroups(k_ff_fast, kLanes * kWaves, cache);
}

void launch_ff_fast(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep, int Qh,
                    int qbeg_min, int out_mode, double2* d_A, int ldA, float2* d_X, int ldX,
                    unsigned* d_colmax, int nsplit, double* d_part) {
  const int span = (out_mode == 0 ? Qh : ldX) - qbeg_min;
  if (N <= 0 || Qh <= 0 || span <= 0) return;
  const int nblk = (span + kLanes - 1) / kLanes;
  const int ldP = nblk * kLanes;
  const bool split = nsplit > 1 && d_part;
  double2* part = split ? reinterpret_cast<double2*>(d_part) : nullptr;
  hipLaunchKernelGGL(k_ff_fast, dim3(nblk, N, split ? nsplit : 1), dim3(kLanes * kWaves), 0, s,
                     d_objs, d_krep, Qh, out_mode, d_A, ldA, d_X, ldX, d_colmax, part, ldP);
  if (split)
    hipLaunchKernelGGL(k_ff_finish, dim3((ldP + 255) / 256, N), dim3(256), 0, s, d_objs, d_krep,
                       Qh, out_mode, d_A, ldA, d_X, ldX, d_colmax, part, ldP, nsplit);
}

void launch_expand(hipStream_t s, const double2* d_Arep, int ldA, const int* d_rep,
                   const int* d_anti, int Qh, int Q, int N, double2* d_Afull) {
  if (N <= 0 || Qh <= 0) return;
  hipLaunchKernelGGL(k_expand, dim3((Qh + 255) / 256, N), dim3(256), 0, s, d_Arep, ldA, d_rep,
                     d_anti, Qh, Q, N, d_Afull);
}

}  // namespace pii
