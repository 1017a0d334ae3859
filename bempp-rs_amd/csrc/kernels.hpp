// kernels.hpp -- HIP kernels (gfx950) for dense boundary-operator assembly.
//
// Regular (non-adjacent) pairs: src/boundary_assemblers.rs:689-773 with the
// per-pair sum of src/boundary_assemblers/cell_pair_assemblers.rs:283-339.
// Adjacent pairs (Sauter-Schwab/Duffy rules): src/boundary_assemblers.rs:610-685
// with src/boundary_assemblers/cell_pair_assemblers.rs:109-152.
//
// Per quadrature-point pair (x on the test cell, y on the trial cell) the kernel
// scalar K is
//   SL : G                        DL : grad_y G . n_y
//   ADL: -grad_y G . n_x          HS : G (curl-curl and, for Helmholtz, -k^2 n_x.n_y term)
// with G = 1/(4 pi r) (Laplace) or e^{ikr}/(4 pi r) (Helmholtz) and
// grad_y G = G (1 - ikr) (x - y) / r^2. The local matrix is
//   M_ab = sum_p sum_q K_pq (phi_a(p) w_p J_x) (psi_b(q) w_q J_y)
// and the hypersingular curl-curl term is (curlJ_a . curlJ_b)/(J_x J_y) * sum_ab M_ab
// for flat P1 cells (the curls are constant and the basis sums to one).
#pragma once
#include <hip/hip_runtime.h>

#include "devmath.hpp"

namespace bempp_amd {
namespace dev {

constexpr double kInv4Pi = 0.079577471545947667884441881686257181;  // 1/(4 pi)
// per-cell record: v0 v1 v2 (9) | n (3) | jdet (1) | curlJ_b (9) | pad (2) |
//                  centroid (3) at 24 | radius (max centroid-vertex distance) at 27 | pad (4)
constexpr int kGeoStride = 32;
constexpr int kGeoCentroid = 24;
constexpr int kGeoRadius = 27;
// Regular rules of up to kMaxQ points are compiled with the point count as a template
// parameter (points held in registers); any other rule runs the same kernels instantiated
// with Q = 0, which take the point count at run time (RegularArgs::nq) and stream the
// points from memory.
constexpr int kMaxQ = 7;
constexpr int kTile = 64;   // cells per tile side (lanes of a wavefront)
constexpr int kChunk = 8;   // trial cells per chunk of the P0 tile kernel

enum OutType { kOutF64 = 0, kOutC128 = 1, kOutF32 = 2, kOutC64 = 3 };
enum Mode { kStore = 0, kAccumulate = 1, kAtomic = 2, kCoo = 3 };

struct RegularArgs {
  const double* test_pts;   // [n_test][Q][4]: x, y, z, w_p * jdet
  const double* trial_pts;  // [n_trial][Q][4]
  const float* test_pts32;  // [n_test][Q][4]: x - centroid (3), w_p * jdet, FP32 (mixed precision)
  const float* trial_pts32; // [n_trial][Q][4]
  // [n_trial][3][4] FP32, mixed-precision P0 tiles: V_0 - c, |V_0 - c|^2; V_1 - c, |V_1 - c|^2;
  // w_class * jdet for weight classes 0..2, |V_2 - c|^2 (c the centroid, V_k the vertices)
  const float* trial_vtx32;
  // [n][12] FP64, the same layout (FP64 P0 x P0 tiles; orbit_form: the rule is in orbit form)
  const double* test_vtx64;
  const double* trial_vtx64;
  int orbit_form;
  const double* test_geo;   // [n_test][kGeoStride]
  const double* trial_geo;  // [n_trial][kGeoStride]
  const int* test_vid;      // [n_test][4]
  const int* trial_vid;     // [n_trial][4]
  const int* test_dofs;     // [n_test][nbt]
  const int* trial_dofs;    // [n_trial][nbs]
  // DOFs of 64-cell blocks (tiled kernel with a P1 side): unique DOFs of block t are
  // blk_dofs[blk_off[t] .. blk_off[t+1]) in ascending order; ldof[c][a] is the position
  // of cell c's a-th DOF in its block's list
  const int* test_blk_off;
  const int* test_blk_dofs;
  const int* test_ldof;     // [n_test][nbt]
  const int* trial_blk_off;
  const int* trial_blk_dofs;
  const int* trial_ldof;    // [n_trial][nbs]
  int n_test, n_trial;
  int nbt, nbs;
  int same_grid;
  int trial_chunk;
  double k;
  double phi[kMaxQ][3];     // basis values at the regular points
  // the regular rule, any point count: [nq][4] = P1 basis values phi_0..phi_2 (the
  // barycentric coordinates) and the weight at each point
  const double* rule_tab;
  int nq;
  // rule weight x basis value per side (a P0 side: {w, 0, 0}), FP32 (mixed-precision P1 tiles)
  float wphi_test[kMaxQ][3];
  float wphi_trial[kMaxQ][3];
  // the trial side's in FP64 (the streamed FP64 P1 cell-pair sums apply the trial Jacobian once
  // per pair)
  double wphi_trial64[kMaxQ][3];
  // mixed precision, one grid: far_tiles[a * far_tiles_ld + b] != 0 when every cell pair of tile
  // (a, b) is far (then no pair of it is adjacent or near); null when not computed
  const unsigned char* far_tiles;
  int far_tiles_ld;
  void* out;
  long long ld;
  int out_type;
  int mode;
};

struct SingularArgs {
  const int4* pairs;        // test cell, trial cell, shared count, permutation code
  int npairs;
  const double* rule;       // 5 arrays of rule_total: test xi0, test xi1, trial xi0, trial xi1, weight
  const double* rule_aos;   // the same rule, one record of 6 per point: a0 a1 b0 b1 w pad
  int rule_total;
  int rule_off[4];          // indexed by shared count 1..3
  int rule_n[4];
  const double* test_geo;
  const double* trial_geo;
  const int* test_dofs;
  const int* trial_dofs;
  int nbt, nbs;
  double k;
  void* out;                // dense output, or COO values [npairs][nbt*nbs] (mode kCoo)
  long long ld;
  int out_type;
  int mode;
};

struct cplx {
  double re, im;
};

// Kernel scalar (without the 1/(4 pi) factor) from r^2 = |x - y|^2 and, for the
// double-layer kernels, dn = (x - y).n_y (DL) or -(x - y).n_x (ADL). CENTRED (Helmholtz,
// phase tier 1 or 2): the phase is taken relative to kR, i.e. the result lacks the factor
// e^{ikR} that the caller applies once per cell pair; requires |k r - kR| <= 0.5 (tier 1)
// or <= 0.25 (tier 2). CENTRED = 0: full argument reduction.
template <int OP, bool HELM, int CENTRED = 0>
__device__ __forceinline__ cplx kernel_scalar_r2(double r2, double dn, double k, double kR = 0.0) {
  const double rinv = rsqrt_fast(r2);
  cplx g;
  double kr = 0.0, s = 0.0, c = 1.0;
  if constexpr (HELM) {
    const double r = r2 * rinv;
    if constexpr (CENTRED) {
      sincos_small<CENTRED>(fma(k, r, -kR), &s, &c);
      if constexpr (OP == 1 || OP == 2) kr = k * r;
    } else {
      kr = k * r;
      sincos_cw(kr, &s, &c);
    }
    g = {c * rinv, s * rinv};
  } else {
    g = {rinv, 0.0};
  }
  if constexpr (OP == 0 || OP == 3) {
    return g;
  } else {
    const double f = dn * rinv * rinv * rinv;
    if constexpr (HELM) {
      // e^{ikr}(1 - ikr) = (c + kr s) + i (s - kr c)
      return {(c + kr * s) * f, (s - kr * c) * f};
    } else {
      return {f, 0.0};
    }
  }
}

// The same for displacement d = x - y; nx: test normal, ny: trial normal.
template <int OP, bool HELM, int CENTRED = 0>
__device__ __forceinline__ cplx kernel_scalar(double dx, double dy, double dz, double k, const double* nx,
                                              const double* ny, double kR = 0.0) {
  double dn = 0.0;
  if constexpr (OP == 1) dn = dx * ny[0] + dy * ny[1] + dz * ny[2];
  if constexpr (OP == 2) dn = -(dx * nx[0] + dy * nx[1] + dz * nx[2]);
  return kernel_scalar_r2<OP, HELM, CENTRED>(dx * dx + dy * dy + dz * dz, dn, k, kR);
}

// acc += w G for the single-layer kernel (without 1/(4 pi)); the weight is folded into
// 1/r so that the complex update costs one product and two fused multiply-adds.
// WEIGHTED = false: w is ignored and the unweighted kernel value is added
template <bool HELM, int CENTRED, bool WEIGHTED = true>
__device__ __forceinline__ void green_accumulate_r2(double r2, double k, double kR, double w, cplx& acc) {
  const double rinv = rsqrt_fast(r2);
  if constexpr (HELM) {
    const double r = r2 * rinv;
    double s, c;
    if constexpr (CENTRED)
      sincos_small<CENTRED>(fma(k, r, -kR), &s, &c);
    else
      sincos_cw(k * r, &s, &c);
    const double rw = WEIGHTED ? rinv * w : rinv;
    acc.re = fma(c, rw, acc.re);
    acc.im = fma(s, rw, acc.im);
  } else {
    acc.re = fma(rinv, w, acc.re);
  }
}

template <bool HELM, int CENTRED>
__device__ __forceinline__ void green_accumulate(double dx, double dy, double dz, double k, double kR, double w,
                                                 cplx& acc) {
  green_accumulate_r2<HELM, CENTRED>(dx * dx + dy * dy + dz * dz, k, kR, w, acc);
}

// Overwriting stores of the two complex output types take a short path; every other
// combination of output type and mode goes through store_value. kind: 1 complex64 store,
// 2 complex128 store, 0 other (wave-uniform).
__device__ __forceinline__ int fast_store_kind(int out_type, int mode) {
  return mode != kStore ? 0 : out_type == kOutC64 ? 1 : out_type == kOutC128 ? 2 : 0;
}
__device__ __forceinline__ void store_value(void* out, long long idx, cplx v, int out_type, int mode);
__device__ __forceinline__ void store_entry(void* out, long long idx, cplx v, int kind, int out_type, int mode) {
  if (kind == 1)
    ((float2*)out)[idx] = make_float2((float)v.re, (float)v.im);
  else if (kind == 2)
    ((double2*)out)[idx] = make_double2(v.re, v.im);
  else
    store_value(out, idx, v, out_type, mode);
}

__device__ __forceinline__ void store_value(void* out, long long idx, cplx v, int out_type, int mode) {
  if (mode == kAtomic) {
    switch (out_type) {
      case kOutF64: unsafeAtomicAdd((double*)out + idx, v.re); break;
      case kOutC128:
        unsafeAtomicAdd((double*)out + 2 * idx, v.re);
        unsafeAtomicAdd((double*)out + 2 * idx + 1, v.im);
        break;
      case kOutF32: unsafeAtomicAdd((float*)out + idx, (float)v.re); break;
      default:
        unsafeAtomicAdd((float*)out + 2 * idx, (float)v.re);
        unsafeAtomicAdd((float*)out + 2 * idx + 1, (float)v.im);
        break;
    }
    return;
  }
  const bool acc = mode == kAccumulate;
  switch (out_type) {
    case kOutF64: {
      double* o = (double*)out + idx;
      *o = acc ? *o + v.re : v.re;
      break;
    }
    case kOutC128: {
      double2* o = (double2*)out + idx;
      double2 x = acc ? *o : make_double2(0.0, 0.0);
      *o = make_double2(x.x + v.re, x.y + v.im);
      break;
    }
    case kOutF32: {
      float* o = (float*)out + idx;
      *o = acc ? *o + (float)v.re : (float)v.re;
      break;
    }
    default: {
      float2* o = (float2*)out + idx;
      float2 x = acc ? *o : make_float2(0.f, 0.f);
      *o = make_float2(x.x + (float)v.re, x.y + (float)v.im);
      break;
    }
  }
}

// The hypersingular curl-curl entries are (curl_a . curl_b) msum / (J_x J_y): callers scale
// msum by 1 / (J_x J_y) once per cell pair with this, before the entries (local_entry).
template <int OP, typename TG, typename SG>
__device__ __forceinline__ cplx hs_msum(cplx msum, TG tgeo, SG sgeo) {
  if constexpr (OP == 3) {
    const double s = 1.0 / (tgeo[12] * sgeo[12]);
    return {msum.re * s, msum.im * s};
  }
  return msum;
}

// Epilogue shared by both loops: M (nbt x nbs local sums without 1/(4pi)) ->
// local matrix entries. tgeo/sgeo: geometry records; lt/ls: local basis index
// of each row/column of M; msum: the sum of M scaled by hs_msum.
template <int OP, bool HELM, typename TG, typename SG>
__device__ __forceinline__ cplx local_entry(const cplx M[3][3], int a, int b, cplx msum, int nbt, int nbs,
                                            TG tgeo, SG sgeo, int la, int lb, double k) {
  cplx v = M[a][b];
  if constexpr (OP == 3) {
    cplx out = {0.0, 0.0};
    if (nbt == 3 && nbs == 3) {
      const auto ct = tgeo + 13 + 3 * la;
      const auto cs = sgeo + 13 + 3 * lb;
      const double cc = ct[0] * cs[0] + ct[1] * cs[1] + ct[2] * cs[2];
      out = {cc * msum.re, cc * msum.im};
    }
    if constexpr (HELM) {
      const double nn = -k * k * (tgeo[9] * sgeo[9] + tgeo[10] * sgeo[10] + tgeo[11] * sgeo[11]);
      out.re += nn * v.re;
      out.im += nn * v.im;
    }
    v = out;
  }
  return {v.re * kInv4Pi, v.im * kInv4Pi};
}

// Read-only view in the constant address space: lets the compiler fetch
// wave-uniform data (the trial cell) with scalar loads into SGPRs.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* constant_view(const T* p) {
  return (const __attribute__((address_space(4))) T*)p;
}

// Wave-uniform copy of a small struct from read-only device memory with scalar loads
// (a struct cannot be copied from the constant address space directly).
template <typename T>
__device__ __forceinline__ T load_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "struct size must be a multiple of 4 bytes");
  constexpr int kWords = sizeof(T) / 4;
  const auto w = constant_view((const int*)p);
  int words[kWords];
#pragma unroll
  for (int i = 0; i < kWords; ++i) words[i] = w[i];
  T v;
  __builtin_memcpy(&v, words, sizeof(T));
  return v;
}

// e^{ikR} for the centroid distance R of a cell pair (CENTRED Helmholtz pairs).
__device__ __forceinline__ cplx pair_phase(const double* ct, const double* cs, double k, double* kR) {
  const double dx = ct[0] - cs[0], dy = ct[1] - cs[1], dz = ct[2] - cs[2];
  const double R2 = dx * dx + dy * dy + dz * dz;
  const double R = R2 * rsqrt_nr(R2);
  *kR = k * R;
  cplx e;
  sincos_cw(*kR, &e.im, &e.re);
  return e;
}

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}

// Applies the trial Jacobian (and, CENTRED, the pair phase e^{ikR}) to the local block and
// forms msum = sum_ab M_ab: with e scaled by J_y first, a centred pair pays two products.
template <int CENTRED>
__device__ __forceinline__ void scale_by_trial_jacobian(cplx (&M)[3][3], cplx& msum, cplx e, double jy) {
  if constexpr (CENTRED) {
    e.re *= jy;
    e.im *= jy;
  }
  msum = {0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if constexpr (CENTRED) {
        M[a][b] = cmul(M[a][b], e);
      } else {
        M[a][b].re *= jy;
        M[a][b].im *= jy;
      }
      msum.re += M[a][b].re;
      msum.im += M[a][b].im;
    }
}

// Local block M_ab (without 1/(4 pi)) of one regular cell pair with a P1 side:
// M_ab = sum_p sum_q K_pq (phi_a(p) w_p J_x)(psi_b(q) w_q J_y), summed over q first
// into T_b(p); msum = sum_ab M_ab (for the hypersingular curl-curl term).
template <int Q, int OP, bool HELM, int CENTRED, typename Args, typename SG>
__device__ __forceinline__ void p1_local(Args& A, const double (&xp)[Q][3], const double (&wp)[Q],
                                         const double (&nx)[3], const double (&ct)[3], const double (&sp)[Q * 4],
                                         SG sgeo, cplx (&M)[3][3], cplx& msum) {
  const double ny[3] = {sgeo[9], sgeo[10], sgeo[11]};
  double kR = 0.0;
  cplx e = {1.0, 0.0};
  if constexpr (CENTRED) {
    const double cs[3] = {sgeo[kGeoCentroid], sgeo[kGeoCentroid + 1], sgeo[kGeoCentroid + 2]};
    e = pair_phase(ct, cs, A.k, &kR);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) M[a][b] = {0.0, 0.0};
#pragma unroll
  for (int p = 0; p < Q; ++p) {
    cplx T[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const cplx kv = kernel_scalar<OP, HELM, CENTRED>(xp[p][0] - sp[4 * q], xp[p][1] - sp[4 * q + 1],
                                                       xp[p][2] - sp[4 * q + 2], A.k, nx, ny, kR);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double f = (A.nbs == 3 ? A.phi[q][b] : (b == 0 ? 1.0 : 0.0)) * sp[4 * q + 3];
        T[b].re = fma(kv.re, f, T[b].re);
        T[b].im = fma(kv.im, f, T[b].im);
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double f = (A.nbt == 3 ? A.phi[p][a] : (a == 0 ? 1.0 : 0.0)) * wp[p];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        M[a][b].re = fma(f, T[b].re, M[a][b].re);
        M[a][b].im = fma(f, T[b].im, M[a][b].im);
      }
    }
  }
  msum = {0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if constexpr (CENTRED) M[a][b] = cmul(M[a][b], e);
      msum.re += M[a][b].re;
      msum.im += M[a][b].im;
    }
}

// p1_local with the test points read per test point p (loop not unrolled) instead of held
// in registers for the whole cell pair: the Helmholtz tile kernel's register footprint
// drops from ~300 to ~165 VGPRs (the lane's points stay in the L1 cache across its trial
// cells).
template <int Q, int OP, bool HELM, int CENTRED, typename Args, typename SG>
__device__ __forceinline__ void p1_local_streamed(Args& A, const double* tpts, const double (&nx)[3],
                                                  const double (&ct)[3], const double (&sp)[Q * 4], SG sgeo,
                                                  cplx (&M)[3][3], cplx& msum) {
  const double ny[3] = {sgeo[9], sgeo[10], sgeo[11]};
  double kR = 0.0;
  cplx e = {1.0, 0.0};
  if constexpr (CENTRED) {
    const double cs[3] = {sgeo[kGeoCentroid], sgeo[kGeoCentroid + 1], sgeo[kGeoCentroid + 2]};
    e = pair_phase(ct, cs, A.k, &kR);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) M[a][b] = {0.0, 0.0};
#pragma unroll 1
  for (int p = 0; p < Q; ++p) {
    const double x0 = tpts[4 * p], x1 = tpts[4 * p + 1], x2 = tpts[4 * p + 2], w = tpts[4 * p + 3];
    cplx T[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const cplx kv = kernel_scalar<OP, HELM, CENTRED>(x0 - sp[4 * q], x1 - sp[4 * q + 1], x2 - sp[4 * q + 2], A.k,
                                                       nx, ny, kR);
#pragma unroll
      for (int b = 0; b < 3; ++b) {  // trial weight x basis value; the trial Jacobian comes last
        T[b].re = fma(kv.re, A.wphi_trial64[q][b], T[b].re);
        T[b].im = fma(kv.im, A.wphi_trial64[q][b], T[b].im);
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double f = (A.nbt == 3 ? A.phi[p][a] : (a == 0 ? 1.0 : 0.0)) * w;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        M[a][b].re = fma(f, T[b].re, M[a][b].re);
        M[a][b].im = fma(f, T[b].im, M[a][b].im);
      }
    }
  }
  scale_by_trial_jacobian<CENTRED>(M, msum, e, sgeo[12]);
}

// p1_local for a rule of any point count (A.nq): test points streamed from tpts, trial
// points and basis values (A.rule_tab) read with wave-uniform loads point by point.
template <int OP, bool HELM, int CENTRED, typename Args, typename SP, typename SG>
__device__ __forceinline__ void p1_local_generic(Args& A, const double* tpts, const double (&nx)[3],
                                                 const double (&ct)[3], SP spp, SG sgeo, cplx (&M)[3][3],
                                                 cplx& msum) {
  const double ny[3] = {sgeo[9], sgeo[10], sgeo[11]};
  double kR = 0.0;
  cplx e = {1.0, 0.0};
  if constexpr (CENTRED) {
    const double cs[3] = {sgeo[kGeoCentroid], sgeo[kGeoCentroid + 1], sgeo[kGeoCentroid + 2]};
    e = pair_phase(ct, cs, A.k, &kR);
  }
  const auto tab = constant_view(A.rule_tab);
  const int nq = A.nq;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) M[a][b] = {0.0, 0.0};
#pragma unroll 1
  for (int p = 0; p < nq; ++p) {
    const double x0 = tpts[4 * p], x1 = tpts[4 * p + 1], x2 = tpts[4 * p + 2], w = tpts[4 * p + 3];
    cplx T[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll 2
    for (int q = 0; q < nq; ++q) {
      const cplx kv = kernel_scalar<OP, HELM, CENTRED>(x0 - spp[4 * q], x1 - spp[4 * q + 1], x2 - spp[4 * q + 2],
                                                       A.k, nx, ny, kR);
      const double wq = spp[4 * q + 3];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double f = (A.nbs == 3 ? tab[4 * q + b] : (b == 0 ? 1.0 : 0.0)) * wq;
        T[b].re = fma(kv.re, f, T[b].re);
        T[b].im = fma(kv.im, f, T[b].im);
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double f = (A.nbt == 3 ? tab[4 * p + a] : (a == 0 ? 1.0 : 0.0)) * w;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        M[a][b].re = fma(f, T[b].re, M[a][b].re);
        M[a][b].im = fma(f, T[b].im, M[a][b].im);
      }
    }
  }
  msum = {0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if constexpr (CENTRED) M[a][b] = cmul(M[a][b], e);
      msum.re += M[a][b].re;
      msum.im += M[a][b].im;
    }
}

// ---------------------------------------------------------------------------
// Regular pairs with a P1 space on either side. One thread per test cell; the
// block's trial cells are walked in order with wave-uniform (scalar) loads of
// the trial geometry; the local block is added with atomics at the DOF indices.
// ---------------------------------------------------------------------------
template <int Q, int OP, bool HELM, bool CENTRED>
__global__ void __launch_bounds__(256) regular_kernel(const RegularArgs A) {
  constexpr bool GEN = Q == 0;       // rule of any point count, points streamed
  constexpr int QR = GEN ? 1 : Q;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool i_ok = i < A.n_test;
  const int il = i_ok ? i : A.n_test - 1;  // clamped: loads stay in bounds
  double xp[QR][3], wp[QR];
  if constexpr (!GEN) {
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const double* s = A.test_pts + ((size_t)il * Q + p) * 4;
      xp[p][0] = s[0];
      xp[p][1] = s[1];
      xp[p][2] = s[2];
      wp[p] = s[3];
    }
  }
  const double* tpts = A.test_pts + (size_t)il * (GEN ? A.nq : Q) * 4;
  const double* tgeo = A.test_geo + (size_t)il * kGeoStride;
  const double nx[3] = {tgeo[9], tgeo[10], tgeo[11]};
  const double ct[3] = {tgeo[kGeoCentroid], tgeo[kGeoCentroid + 1], tgeo[kGeoCentroid + 2]};
  const int tv0 = A.test_vid[4 * il], tv1 = A.test_vid[4 * il + 1], tv2 = A.test_vid[4 * il + 2];
  const auto trial_pts = constant_view(A.trial_pts);
  const auto trial_geo = constant_view(A.trial_geo);
  const auto trial_vid = constant_view(A.trial_vid);
  const int j0 = blockIdx.y * A.trial_chunk;
  const int j1 = min(j0 + A.trial_chunk, A.n_trial);
  for (int j = j0; j < j1; ++j) {
    const auto spp = trial_pts + (size_t)j * (GEN ? A.nq : Q) * 4;
    double sp[QR * 4];
    if constexpr (!GEN) {
#pragma unroll
      for (int m = 0; m < Q * 4; ++m) sp[m] = spp[m];
    }
    const auto sg = trial_geo + (size_t)j * kGeoStride;
    const int sv0 = trial_vid[4 * j], sv1 = trial_vid[4 * j + 1], sv2 = trial_vid[4 * j + 2];
    const bool adjacent = A.same_grid && (tv0 == sv0 || tv0 == sv1 || tv0 == sv2 || tv1 == sv0 || tv1 == sv1 ||
                                          tv1 == sv2 || tv2 == sv0 || tv2 == sv1 || tv2 == sv2);
    if (adjacent || !i_ok) continue;
    cplx M[3][3], msum;
    if constexpr (GEN)
      p1_local_generic<OP, HELM, CENTRED>(A, tpts, nx, ct, spp, sg, M, msum);
    else
      p1_local<Q, OP, HELM, CENTRED>(A, xp, wp, nx, ct, sp, sg, M, msum);
    msum = hs_msum<OP>(msum, tgeo, sg);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if (b >= A.nbs) continue;  // uniform; constant indices keep M in registers
      const long long col = A.trial_dofs[(size_t)j * A.nbs + b];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (a >= A.nbt) continue;
        const long long row = A.test_dofs[(size_t)i * A.nbt + a];
        const cplx v = local_entry<OP, HELM>(M, a, b, msum, A.nbt, A.nbs, tgeo, sg, a, b, A.k);
        store_value(A.out, row + A.ld * col, v, A.out_type, kAtomic);
      }
    }
  }
}

// Weight classes of the symmetric triangle rules (triangle_rule): the points of one orbit
// share a weight, so the mixed-precision far-field point-pair sums are taken per (test
// class, trial class) in FP32, unweighted, and weighted once per cell pair. Q = 1, 3: one class;
// Q = 6: {0,1,2} {3,4,5}; Q = 7: {0} {1,2,3} {4,5,6}. The host checks a rule against
// these classes before launching a kernel that relies on them (rule_has_weight_classes).
__host__ __device__ constexpr int weight_classes(int Q) { return Q == 6 ? 2 : Q == 7 ? 3 : 1; }
template <int Q>
constexpr int kWeightClasses = weight_classes(Q);
__host__ __device__ constexpr int weight_class(int Q, int p) {
  return Q == 6 ? (p >= 3 ? 1 : 0) : Q == 7 ? (p == 0 ? 0 : p <= 3 ? 1 : 2) : 0;
}
// first point of class c
__host__ __device__ constexpr int class_point(int Q, int c) { return Q == 7 ? (c == 0 ? 0 : 3 * c - 2) : 3 * c; }

// Orbit form of the same rules: point q is c + t (V_k - c) (c the centroid, V_k a vertex).
// An S21 orbit with repeated barycentric coordinate a has t = 1 - 3a and its three points
// take k = 0, 1, 2 in order; the centroid has t = 0. The host checks a rule against this
// table before launching the P0 tiles, which take the rule points in this form.
__host__ __device__ constexpr double orbit_t(int Q, int q) {
  return Q == 3   ? 0.5
         : Q == 6 ? (q < 3 ? -0.33784547274789466 : 0.72527135947068777)
         : Q == 7 ? (q == 0 ? 0.0 : q <= 3 ? 0.69614047802963098 : -0.41042619231534527)
                  : 0.0;
}
__host__ __device__ constexpr int orbit_vertex(int Q, int q) { return Q == 7 ? (q == 0 ? 0 : (q - 1) % 3) : q % 3; }

// One regular P0 cell pair (any operator; the hypersingular one is the single-layer sum)
// from point coordinates: test points xp (w_p J_x in wp) held in registers, trial points
// sp (x, y, z, w_q J_y per point). The near-pair kernels and the matrix-free apply use it.
template <int Q, int OP, bool HELM, int CENTRED>
__device__ __forceinline__ cplx p0_pair(const double (&xp)[Q][3], const double (&wp)[Q], const double* nx,
                                        const double (&sp)[Q * 4], const double* ny, double k, double kR) {
  cplx acc = {0.0, 0.0};
#pragma unroll
  for (int p = 0; p < Q; ++p) {
    cplx ap = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if constexpr (OP == 0 || OP == 3) {  // single layer (and the P0 hypersingular n.n term)
        green_accumulate<HELM, CENTRED>(xp[p][0] - sp[4 * q], xp[p][1] - sp[4 * q + 1], xp[p][2] - sp[4 * q + 2], k,
                                        kR, sp[4 * q + 3], ap);
      } else {
        const cplx kv = kernel_scalar<OP, HELM, CENTRED>(xp[p][0] - sp[4 * q], xp[p][1] - sp[4 * q + 1],
                                                         xp[p][2] - sp[4 * q + 2], k, nx, ny, kR);
        ap.re = fma(kv.re, sp[4 * q + 3], ap.re);
        ap.im = fma(kv.im, sp[4 * q + 3], ap.im);
      }
    }
    acc.re = fma(ap.re, wp[p], acc.re);
    acc.im = fma(ap.im, wp[p], acc.im);
  }
  return acc;
}

// One regular P0 cell pair (all four operators; the hypersingular one is the single-layer
// sum, scaled by -k^2 n_x.n_y by the caller) with both rules in orbit form, x_p = c_t + tau_p d'_k'(p), y_q = c_s + t_q d_k(q) (d', d the
// vertex offsets of the test and trial cell from their centroids, D = c_t - c_s):
//   r^2 = (R^2 + t_q^2 |d_k|^2) + A_p + t_q W_pk,
//   A_p = tau_p 2 D.d'_k' + tau_p^2 |d'_k'|^2,  W_pk = -2 D.d_k - 2 tau_p d'_k'.d_k,
// two instructions per point pair instead of six (three differences and a dot product);
// W_p2 = -W_p0 - W_p1 and the d'_2 products follow the same way, the offsets of a cell
// summing to zero. tau2g[p] = tau_p^2 |d'_k'(p)|^2; sv: the trial cell's 12-value record.
// The centred Helmholtz single-layer (and P0 hypersingular) orbit sums of rules with at most
// two weight classes run in scaled lengths: the caller multiplies the test cell's vertex
// offsets and tau_p^2 |d'_k'|^2 by k^2 / 2 (p0_pair_orbit scales the rest).
__host__ __device__ constexpr bool orbit_scaled(int Q, int OP, bool HELM, int CENTRED) {
  return HELM && (OP == 0 || OP == 3) && CENTRED > 0 && weight_classes(Q) <= 2;
}

template <int Q, int OP, bool HELM, int CENTRED, typename SV>
__device__ __forceinline__ cplx p0_pair_orbit(const double (&td)[2][3], const double (&tau2g)[Q],
                                              const double (&twc)[3], const SV& sv, const double (&D)[3], double R2,
                                              const double (&nx)[3], const double (&ny)[3], double k, double kR) {
  // double layer: (x - y).n_y = D.n_y + tau_p d'_k'.n_y (o_q lies in the trial plane);
  // adjoint double layer: -(x - y).n_x = -D.n_x + t_q d_k.n_x (o_p lies in the test plane)
  double dnp[Q], dnq[Q];
  if constexpr (OP == 1) {
    const double Dn = D[0] * ny[0] + D[1] * ny[1] + D[2] * ny[2];
    double dkn[3];
    dkn[0] = td[0][0] * ny[0] + td[0][1] * ny[1] + td[0][2] * ny[2];
    dkn[1] = td[1][0] * ny[0] + td[1][1] * ny[1] + td[1][2] * ny[2];
    dkn[2] = -(dkn[0] + dkn[1]);
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const double tau = orbit_t(Q, p);
      dnp[p] = tau == 0.0 ? Dn : fma(tau, dkn[orbit_vertex(Q, p)], Dn);
    }
  }
  if constexpr (OP == 2) {
    const double Dn = -(D[0] * nx[0] + D[1] * nx[1] + D[2] * nx[2]);
    double dkn[3];
    dkn[0] = sv[0] * nx[0] + sv[1] * nx[1] + sv[2] * nx[2];
    dkn[1] = sv[4] * nx[0] + sv[5] * nx[1] + sv[6] * nx[2];
    dkn[2] = -(dkn[0] + dkn[1]);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double t = orbit_t(Q, q);
      dnq[q] = t == 0.0 ? Dn : fma(t, dkn[orbit_vertex(Q, q)], Dn);
    }
  }
  // orbit_scaled: td and tau2g arrive multiplied by sc = k^2 / 2; E, R^2 and the |d_k|^2
  // terms are scaled here, so that every r^2 below is sc r^2
  constexpr bool SCALED = orbit_scaled(Q, OP, HELM, CENTRED);
  const double sc = SCALED ? 0.5 * k * k : 1.0;
  const double E0 = (SCALED ? -k * k : -2.0) * (D[0] * sv[0] + D[1] * sv[1] + D[2] * sv[2]);
  const double E1 = (SCALED ? -k * k : -2.0) * (D[0] * sv[4] + D[1] * sv[5] + D[2] * sv[6]);
  double N[3][2], F[3];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    N[m][0] = td[m][0] * sv[0] + td[m][1] * sv[1] + td[m][2] * sv[2];
    N[m][1] = td[m][0] * sv[4] + td[m][1] * sv[5] + td[m][2] * sv[6];
    F[m] = 2.0 * (D[0] * td[m][0] + D[1] * td[m][1] + D[2] * td[m][2]);
  }
  N[2][0] = -(N[0][0] + N[1][0]);
  N[2][1] = -(N[0][1] + N[1][1]);
  F[2] = -(F[0] + F[1]);
  const double gk[3] = {sv[3], sv[7], sv[11]};
  const double R2s = SCALED ? sc * R2 : R2;
  double RC[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double t = orbit_t(Q, q);
    RC[q] = t == 0.0 ? R2s : fma(SCALED ? t * t * sc : t * t, gk[orbit_vertex(Q, q)], R2s);
  }
  if constexpr ((OP == 0 || OP == 3) && HELM && kWeightClasses<Q> <= 2) {
    // unweighted sums per (test weight class, trial weight class), the class weights applied
    // once per cell pair: no weight multiply per point pair or per test point
    constexpr int NC = kWeightClasses<Q>;
    const double kRs = SCALED ? 1.4142135623730951 * fabs(kR) : 0.0;  // sqrt(2) |k| R
    cplx cacc[NC][NC];
#pragma unroll
    for (int a = 0; a < NC; ++a)
#pragma unroll
      for (int b = 0; b < NC; ++b) cacc[a][b] = {0.0, 0.0};
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const int kp = orbit_vertex(Q, p);
      const double tau = orbit_t(Q, p);
      double W[3];
      W[0] = tau == 0.0 ? E0 : fma(-2.0 * tau, N[kp][0], E0);
      W[1] = tau == 0.0 ? E1 : fma(-2.0 * tau, N[kp][1], E1);
      W[2] = -(W[0] + W[1]);
      const double Ap = fma(tau, F[kp], tau2g[p]);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const double t = orbit_t(Q, q);
        const double base = tau == 0.0 ? RC[q] : Ap + RC[q];
        const double r2 = t == 0.0 ? base : fma(t, W[orbit_vertex(Q, q)], base);
        cplx& ca = cacc[weight_class(Q, p)][weight_class(Q, q)];
        if constexpr (SCALED) {
          // h = sc r^2, y ~ h^(-1/2): one Newton step y (3 - h y^2) without its factor 1/2 gives
          // rz = 2 sqrt(2) / (|k| r), and h rz = sqrt(2) |k| r (the phase in one FMA)
          const double y = __builtin_amdgcn_rsq(r2);
          const double rz = y * fma(-(r2 * y), y, 3.0);
          double sn, cs;
          sincos_small<CENTRED, true>(fma(r2, rz, -kRs), &sn, &cs);
          ca.re = fma(cs, rz, ca.re);
          ca.im = fma(sn, rz, ca.im);
        } else {
          green_accumulate_r2<HELM, CENTRED, false>(r2, k, kR, 1.0, ca);
        }
      }
    }
    cplx acc = {0.0, 0.0};
#pragma unroll
    for (int a = 0; a < NC; ++a)
#pragma unroll
      for (int b = 0; b < NC; ++b) {
        const double w = twc[a] * sv[8 + b];
        acc.re = fma(cacc[a][b].re, w, acc.re);
        acc.im = fma(cacc[a][b].im, w, acc.im);
      }
    if constexpr (SCALED) {
      // the sums hold 2 sqrt(2) cos(a) / (|k| r) and 4 sin(a) / (|k| r), a = |k| (r - R);
      // sin(k (r - R)) = sign(k) sin a
      acc.re *= fabs(k) * (0.25 * 1.4142135623730951);
      acc.im *= 0.25 * k;
    }
    return acc;
  }
  cplx acc = {0.0, 0.0};
#pragma unroll
  for (int p = 0; p < Q; ++p) {
    const int kp = orbit_vertex(Q, p);
    const double tau = orbit_t(Q, p);
    double W[3];
    W[0] = tau == 0.0 ? E0 : fma(-2.0 * tau, N[kp][0], E0);
    W[1] = tau == 0.0 ? E1 : fma(-2.0 * tau, N[kp][1], E1);
    W[2] = -(W[0] + W[1]);
    const double Ap = fma(tau, F[kp], tau2g[p]);
    cplx ap = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double t = orbit_t(Q, q);
      const double base = tau == 0.0 ? RC[q] : Ap + RC[q];
      const double r2 = t == 0.0 ? base : fma(t, W[orbit_vertex(Q, q)], base);
      if constexpr (OP == 0 || OP == 3) {
        green_accumulate_r2<HELM, CENTRED>(r2, k, kR, sv[8 + weight_class(Q, q)], ap);
      } else {
        const cplx kv = kernel_scalar_r2<OP, HELM, CENTRED>(r2, OP == 1 ? dnp[p] : dnq[q], k, kR);
        ap.re = fma(kv.re, sv[8 + weight_class(Q, q)], ap.re);
        ap.im = fma(kv.im, sv[8 + weight_class(Q, q)], ap.im);
      }
    }
    acc.re = fma(ap.re, twc[weight_class(Q, p)], acc.re);
    acc.im = fma(ap.im, twc[weight_class(Q, p)], acc.im);
  }
  return acc;
}

// Adds a tile's LDS block acc[trial DOF][test DOF] (block-local indices) to the output
// with global atomics. A mirrored tile (symmetric operator, off-diagonal tile) also adds
// the transpose, in a second pass whose consecutive threads take consecutive trial DOFs:
// the transposed entries of one test DOF are then neighbours in memory, so a wave's
// atomics fall into a few cache lines instead of one line per lane.
template <typename Args>
__device__ __forceinline__ void flush_tile(const Args& A, const double2* acc, int t0, int ndt, int s0, int nds,
                                           bool mirror) {
  // the blocks' DOF numbers copied to LDS first (at most 256 a side; larger blocks read them
  // from global memory): each entry's row and column then cost an LDS read, not a global load
  // and its wait
  __shared__ int frow[256], fcol[256];
  const bool staged = ndt <= 256 && nds <= 256;  // block-uniform
  if (staged) {
    for (int t = threadIdx.x; t < ndt; t += blockDim.x) frow[t] = A.test_blk_dofs[t0 + t];
    for (int t = threadIdx.x; t < nds; t += blockDim.x) fcol[t] = A.trial_blk_dofs[s0 + t];
    __syncthreads();
  }
  const int nacc = ndt * nds;
  for (int e = threadIdx.x; e < nacc; e += blockDim.x) {
    const double2 v = acc[e];
    if (v.x == 0.0 && v.y == 0.0) continue;
    const int lsi = e / ndt, lti = e - lsi * ndt;
    const long long row = staged ? frow[lti] : A.test_blk_dofs[t0 + lti];
    const long long col = staged ? fcol[lsi] : A.trial_blk_dofs[s0 + lsi];
    store_value(A.out, row + A.ld * col, {v.x, v.y}, A.out_type, kAtomic);
  }
  if (!mirror) return;
  for (int e = threadIdx.x; e < nacc; e += blockDim.x) {
    const int lti = e / nds, lsi = e - lti * nds;
    const double2 v = acc[lsi * ndt + lti];
    if (v.x == 0.0 && v.y == 0.0) continue;
    const long long row = staged ? frow[lti] : A.test_blk_dofs[t0 + lti];
    const long long col = staged ? fcol[lsi] : A.trial_blk_dofs[s0 + lsi];
    store_value(A.out, col + A.ld * row, {v.x, v.y}, A.out_type, kAtomic);
  }
}

// ---------------------------------------------------------------------------
// Regular pairs with a P1 side, tiled: one workgroup (4 wavefronts) per tile of 64 test
// x 64 trial cells (lane = test cell, 16 trial cells per wavefront). The local blocks of
// the tile's cell pairs are summed in LDS at (block-local test DOF, block-local trial DOF)
// -- a compact 64-cell patch of a P1 space has ~45 vertices, so the tile touches ~45 x 45
// entries instead of 64 x 64 x 9 -- and each non-zero LDS entry is then added to the
// output with one global atomic. symmetric (single layer / hypersingular on one space):
// tiles a <= b only; an off-diagonal tile also adds its entries at the transposed place.
// Dynamic LDS: (max block DOFs)^2 double2.
// ---------------------------------------------------------------------------
template <int Q, int OP, bool HELM, int CENTRED>
__global__ void __launch_bounds__(256) regular_tiled_lds(const RegularArgs A, const int2* __restrict__ tiles,
                                                         int symmetric) {
  extern __shared__ double2 acc[];
  const int2 tile = tiles[blockIdx.x];
  const int t0 = A.test_blk_off[tile.x], ndt = A.test_blk_off[tile.x + 1] - t0;
  const int s0 = A.trial_blk_off[tile.y], nds = A.trial_blk_off[tile.y + 1] - s0;
  const int nacc = ndt * nds;
  for (int e = threadIdx.x; e < nacc; e += blockDim.x) acc[e] = make_double2(0.0, 0.0);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = tile.x * kTile + lane;
  const bool i_ok = i < A.n_test;
  const int il = i_ok ? i : A.n_test - 1;
  // Helmholtz: test points read per point (fewer registers, more waves); Laplace: held in
  // registers (its kernel is short enough that the register version runs at higher occupancy)
  constexpr bool GEN = Q == 0;  // rule of any point count, points streamed
  constexpr int QR = GEN ? 1 : Q;
  const double* tpts = A.test_pts + (size_t)il * (GEN ? A.nq : Q) * 4;
  // Laplace hypersingular: the P1 surface curls are constant on a cell, so a cell pair needs
  // only sum_pq (w_p J_x)(w_q J_y) G_pq -- the single-layer P0 sum, taken in orbit form as
  // the P0 tiles take it (p0_pair_orbit) when the rule has that form
  constexpr bool LHS = OP == 3 && !HELM && !GEN;
  double xp[QR][3], wp[QR];
  if constexpr (!HELM && !GEN && !LHS) {
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      xp[p][0] = tpts[4 * p];
      xp[p][1] = tpts[4 * p + 1];
      xp[p][2] = tpts[4 * p + 2];
      wp[p] = tpts[4 * p + 3];
    }
  }
  const double* tgeo = A.test_geo + (size_t)il * kGeoStride;
  const double nx[3] = {tgeo[9], tgeo[10], tgeo[11]};
  const double ct[3] = {tgeo[kGeoCentroid], tgeo[kGeoCentroid + 1], tgeo[kGeoCentroid + 2]};
  const int tv0 = A.test_vid[4 * il], tv1 = A.test_vid[4 * il + 1], tv2 = A.test_vid[4 * il + 2];
  const bool lhs_orbit = LHS && A.orbit_form;
  double td[2][3], tau2g[QR], twc[3];
  if constexpr (LHS) {
    if (lhs_orbit) {
      const double* tv = A.test_vtx64 + (size_t)il * 12;
#pragma unroll
      for (int m = 0; m < 2; ++m)
        for (int d = 0; d < 3; ++d) td[m][d] = tv[4 * m + d];
      const double tg[3] = {tv[3], tv[7], tv[11]};
#pragma unroll
      for (int p = 0; p < QR; ++p) {
        const double tau = orbit_t(Q, p);
        tau2g[p] = tau * tau * tg[orbit_vertex(Q, p)];
      }
      for (int c = 0; c < 3; ++c) twc[c] = tv[8 + c];
    }
  }
  const auto trial_vtx = constant_view(A.trial_vtx64);
  int lt[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (a < A.nbt) lt[a] = A.test_ldof[(size_t)il * A.nbt + a];
  const auto trial_pts = constant_view(A.trial_pts);
  const auto trial_geo = constant_view(A.trial_geo);
  const auto trial_vid = constant_view(A.trial_vid);
  const auto trial_ldof = constant_view(A.trial_ldof);
#pragma unroll 1
  for (int jj = 0; jj < 16; ++jj) {
    const int j = tile.y * kTile + wave * 16 + jj;
    if (j >= A.n_trial) break;  // wave-uniform
    const auto spp = trial_pts + (size_t)j * (GEN ? A.nq : Q) * 4;
    double sp[QR * 4];
    if constexpr (!GEN && !LHS) {  // Laplace HS reads the trial points only off the orbit form
#pragma unroll
      for (int m = 0; m < Q * 4; ++m) sp[m] = spp[m];
    }
    const auto sg = trial_geo + (size_t)j * kGeoStride;
    const int sv0 = trial_vid[4 * j], sv1 = trial_vid[4 * j + 1], sv2 = trial_vid[4 * j + 2];
    int ls[3] = {0, 0, 0};
#pragma unroll
    for (int b = 0; b < 3; ++b)
      if (b < A.nbs) ls[b] = trial_ldof[(size_t)j * A.nbs + b];
    const bool adjacent = A.same_grid && (tv0 == sv0 || tv0 == sv1 || tv0 == sv2 || tv1 == sv0 || tv1 == sv1 ||
                                          tv1 == sv2 || tv2 == sv0 || tv2 == sv1 || tv2 == sv2);
    if (adjacent || !i_ok) continue;
    cplx M[3][3], msum;
    if (LHS && lhs_orbit) {
      if constexpr (LHS) {
        const auto svp = trial_vtx + (size_t)j * 12;
        double sv[12];
#pragma unroll
        for (int m = 0; m < 12; ++m) sv[m] = svp[m];
        const double D[3] = {ct[0] - sg[kGeoCentroid], ct[1] - sg[kGeoCentroid + 1], ct[2] - sg[kGeoCentroid + 2]};
        const double R2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
        const double ny[3] = {sg[9], sg[10], sg[11]};
        msum = p0_pair_orbit<Q, 0, false, 0>(td, tau2g, twc, sv, D, R2, nx, ny, 0.0, 0.0);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b) M[a][b] = {0.0, 0.0};  // unused by the curl-curl entries
      }
    } else if constexpr (GEN) {
      p1_local_generic<OP, HELM, CENTRED>(A, tpts, nx, ct, spp, sg, M, msum);
    } else if constexpr (HELM || LHS) {  // LHS here: a rule off the orbit form
      if constexpr (LHS) {
#pragma unroll
        for (int m = 0; m < Q * 4; ++m) sp[m] = spp[m];
      }
      p1_local_streamed<Q, OP, HELM, CENTRED>(A, tpts, nx, ct, sp, sg, M, msum);
    } else {
      p1_local<Q, OP, HELM, CENTRED>(A, xp, wp, nx, ct, sp, sg, M, msum);
    }
    msum = hs_msum<OP>(msum, tgeo, sg);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (a >= A.nbt || b >= A.nbs) continue;  // uniform; constant indices keep M in registers
        const cplx v = local_entry<OP, HELM>(M, a, b, msum, A.nbt, A.nbs, tgeo, sg, a, b, A.k);
        double2* slot = acc + (ls[b] * ndt + lt[a]);
        atomicAdd(&slot->x, v.re);
        if constexpr (HELM) atomicAdd(&slot->y, v.im);
      }
  }
  __syncthreads();
  flush_tile(A, acc, t0, ndt, s0, nds, symmetric && tile.x != tile.y);
}

// ---------------------------------------------------------------------------
// P0 x P0 regular pairs, tiled. One workgroup (4 wavefronts) per 64 x 64 tile
// (tile row a: test cells 64a + lane; tile column b: trial cells, 16 per
// wavefront in two chunks of 8). The tile is written column-major with
// coalesced stores along the test index. When `symmetric` is set (single
// layer / hypersingular with identical test and trial space), only tiles with
// a <= b are launched and an off-diagonal tile is also written transposed at
// (b, a): A[j][i] sums the same quadrature terms as A[i][j] because G is
// symmetric and both cells use the same rule. The transpose goes through LDS
// so the mirrored stores are 128-byte contiguous segments.
// ---------------------------------------------------------------------------

// One regular P0 cell pair for a rule of any point count (nq): the lane's test points
// streamed from tpts (x, y, z, w J per point), the wave-uniform trial cell's points read
// with scalar loads point by point.
template <int OP, bool HELM, int CENTRED, typename SP>
__device__ __forceinline__ cplx p0_pair_generic(const double* tpts, int nq, const double (&nx)[3], SP spp,
                                                const double (&ny)[3], double k, double kR) {
  cplx acc = {0.0, 0.0};
#pragma unroll 1
  for (int p = 0; p < nq; ++p) {
    const double x0 = tpts[4 * p], x1 = tpts[4 * p + 1], x2 = tpts[4 * p + 2], w = tpts[4 * p + 3];
    cplx ap = {0.0, 0.0};
#pragma unroll 4
    for (int q = 0; q < nq; ++q) {
      if constexpr (OP == 0 || OP == 3) {
        green_accumulate<HELM, CENTRED>(x0 - spp[4 * q], x1 - spp[4 * q + 1], x2 - spp[4 * q + 2], k, kR,
                                        spp[4 * q + 3], ap);
      } else {
        const cplx kv = kernel_scalar<OP, HELM, CENTRED>(x0 - spp[4 * q], x1 - spp[4 * q + 1], x2 - spp[4 * q + 2],
                                                         k, nx, ny, kR);
        ap.re = fma(kv.re, spp[4 * q + 3], ap.re);
        ap.im = fma(kv.im, spp[4 * q + 3], ap.im);
      }
    }
    acc.re = fma(ap.re, w, acc.re);
    acc.im = fma(ap.im, w, acc.im);
  }
  return acc;
}

// NW: wavefronts per tile (each covers kTile / NW trial cells; the workgroup has 64 NW
// threads); CH: trial cells per chunk (a wavefront covers its trial cells in chunks of CH);
// the mirror transpose stores CH contiguous entries per segment.
// Args: RegularArgs, possibly qualified with the constant address space (batched
// launches read each item's arguments in place, with scalar loads).
template <int Q, int OP, bool HELM, int CENTRED, int CH = kChunk, int NW = 4, typename Args>
__device__ __forceinline__ void p0_tile(Args& A, const int2 tile, int symmetric,
                                        double2 (&lds)[NW][CH][kTile]) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: lets trial loads be scalar
  const int i = tile.x * kTile + lane;
  const bool i_ok = i < A.n_test;
  const int il = i_ok ? i : A.n_test - 1;  // clamped: loads stay in bounds
  // Q = 0: a rule of any point count (A.nq), points streamed (p0_pair_generic)
  constexpr bool GEN = Q == 0;
  constexpr int QR = GEN ? 1 : Q;
  const double* tpts = A.test_pts + (size_t)il * A.nq * 4;
  // rule points in orbit form (p0_pair_orbit): the test cell's vertex offsets
  double td[2][3], tau2g[QR], twc[3];
  if constexpr (!GEN) {
    const double* tv = A.test_vtx64 + (size_t)il * 12;
#pragma unroll
    for (int m = 0; m < 2; ++m)
      for (int d = 0; d < 3; ++d) td[m][d] = tv[4 * m + d];
    const double tg[3] = {tv[3], tv[7], tv[11]};
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const double tau = orbit_t(Q, p);
      tau2g[p] = tau * tau * tg[orbit_vertex(Q, p)];
    }
    for (int c = 0; c < 3; ++c) twc[c] = tv[8 + c];
    if constexpr (orbit_scaled(Q, OP, HELM, CENTRED)) {  // lengths scaled by |k| / sqrt 2
      const double sc = 0.5 * A.k * A.k;
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) td[m][d] *= sc;
#pragma unroll
      for (int p = 0; p < Q; ++p) tau2g[p] *= sc;
    }
  }
  const double* tgeo = A.test_geo + (size_t)il * kGeoStride;
  const double nx[3] = {tgeo[9], tgeo[10], tgeo[11]};
  const double ct[3] = {tgeo[kGeoCentroid], tgeo[kGeoCentroid + 1], tgeo[kGeoCentroid + 2]};
  const int tv0 = A.test_vid[4 * il], tv1 = A.test_vid[4 * il + 1], tv2 = A.test_vid[4 * il + 2];
  const bool mirror = symmetric && tile.x != tile.y;
  const int skind = fast_store_kind(A.out_type, A.mode);
  const auto trial_vtx = constant_view(A.trial_vtx64);
  const auto trial_pts = constant_view(A.trial_pts);
  const auto trial_geo = constant_view(A.trial_geo);
  const auto trial_vid = constant_view(A.trial_vid);
  constexpr int TPW = kTile / NW;  // trial cells per wavefront
  static_assert((NW == 2 || NW == 4 || NW == 8) && (CH == 4 || CH == 8 || CH == 16) && TPW % CH == 0,
                "chunk must divide the trial cells of a wavefront");
  for (int c = 0; c < TPW / CH; ++c) {
    const int jbase = tile.y * kTile + wave * TPW + c * CH;
#pragma unroll 1
    for (int jj = 0; jj < CH; ++jj) {
      const int j = jbase + jj;
      if (j >= A.n_trial) {  // wave-uniform
        lds[wave][jj][lane] = make_double2(0.0, 0.0);
        continue;
      }
      // wave-uniform trial cell: loaded unconditionally (scalar loads) before any per-lane branch
      double sp[12];
      if constexpr (!GEN) {
        const auto spp = trial_vtx + (size_t)j * 12;
#pragma unroll
        for (int m = 0; m < 12; ++m) sp[m] = spp[m];
      }
      const auto gpp = trial_pts + (size_t)j * A.nq * 4;
      const auto sg = trial_geo + (size_t)j * kGeoStride;
      const int sv0 = trial_vid[4 * j], sv1 = trial_vid[4 * j + 1], sv2 = trial_vid[4 * j + 2];
      const bool adjacent = A.same_grid && (tv0 == sv0 || tv0 == sv1 || tv0 == sv2 || tv1 == sv0 || tv1 == sv1 ||
                                            tv1 == sv2 || tv2 == sv0 || tv2 == sv1 || tv2 == sv2);
      const double ny[3] = {sg[9], sg[10], sg[11]};
      cplx acc = {0.0, 0.0};
      if (!adjacent) {
        const double cs[3] = {sg[kGeoCentroid], sg[kGeoCentroid + 1], sg[kGeoCentroid + 2]};
        const double D[3] = {ct[0] - cs[0], ct[1] - cs[1], ct[2] - cs[2]};
        const double R2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
        if constexpr (GEN) {
          if constexpr (CENTRED) {
            double kR;
            const cplx e = pair_phase(ct, cs, A.k, &kR);
            acc = cmul(p0_pair_generic<OP, HELM, CENTRED>(tpts, A.nq, nx, gpp, ny, A.k, kR), e);
          } else {
            acc = p0_pair_generic<OP, HELM, 0>(tpts, A.nq, nx, gpp, ny, A.k, 0.0);
          }
        } else if constexpr (CENTRED) {
          double kR;
          const cplx e = pair_phase(ct, cs, A.k, &kR);
          acc = cmul(p0_pair_orbit<Q, OP, HELM, CENTRED>(td, tau2g, twc, sp, D, R2, nx, ny, A.k, kR), e);
        } else {
          acc = p0_pair_orbit<Q, OP, HELM, 0>(td, tau2g, twc, sp, D, R2, nx, ny, A.k, 0.0);
        }
      }
      if constexpr (OP == 3) {  // P0 hypersingular: only the Helmholtz -k^2 n_x.n_y term survives
        const double nn = HELM ? -A.k * A.k * (nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2]) : 0.0;
        acc = {nn * acc.re, nn * acc.im};
      }
      // adjacent pairs get 0 here (their Duffy value is added by the singular kernel)
      const cplx v = {acc.re * kInv4Pi, acc.im * kInv4Pi};
      if (i_ok) store_entry(A.out, (long long)i + A.ld * (long long)j, v, skind, A.out_type, A.mode);
      if (mirror) lds[wave][jj][lane] = make_double2(v.re, v.im);
    }
    if (mirror) {  // block-uniform
      __syncthreads();
      const int row = lane % CH;
      const int jr = jbase + row;  // row of the mirrored entry (a trial cell of this tile)
#pragma unroll
      for (int s = 0; s < CH; ++s) {
        const int col = s * (kTile / CH) + lane / CH;
        const int ic = tile.x * kTile + col;  // column of the mirrored entry (a test cell)
        const double2 val = lds[wave][row][col];
        if (jr < A.n_trial && ic < A.n_test)
          store_entry(A.out, (long long)jr + A.ld * (long long)ic, {val.x, val.y}, skind, A.out_type, A.mode);
      }
      __syncthreads();
    }
  }
}

// the scaled orbit sums fit three waves per SIMD (at most 168 VGPRs)
template <int Q, int OP, bool HELM, int CENTRED>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(orbit_scaled(Q, OP, HELM, CENTRED) ? 3 : 1)))
    regular_p0_tiled(const RegularArgs A, const int2* __restrict__ tiles, int symmetric) {
  __shared__ double2 lds[4][kChunk][kTile];
  p0_tile<Q, OP, HELM, CENTRED>(A, tiles[blockIdx.x], symmetric, lds);
}

// ---------------------------------------------------------------------------
// Batched P0 x P0 assembly: many independent meshes/wavenumbers in one persistent
// launch. Tiles of all items form one list {item, tile row, tile column, _};
// workgroups take tiles from a global counter until the list is exhausted (every
// workgroup exits once the counter passes the end, so the grid always drains).
// The per-item arguments live in device memory and are read with scalar loads.
// Items are grouped by phase mode: one launch per CENTRED value, so that each
// kernel carries the registers of one tile variant only.
// ---------------------------------------------------------------------------
struct BatchTile {
  int item, a, b, pad;
};

template <int Q, int OP, bool HELM, int CENTRED, int CH, int NW = 4>
__global__ void __launch_bounds__(64 * NW) batched_p0_persistent(const RegularArgs* __restrict__ items,
                                                                 const BatchTile* __restrict__ tiles, int ntiles,
                                                                 int* counter, int symmetric) {
  __shared__ double2 lds[NW][CH][kTile];
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();  // everyone has read `next` before thread 0 overwrites it
    if (t >= ntiles) return;
    const BatchTile bt = load_uniform(tiles + t);
    p0_tile<Q, OP, HELM, CENTRED, CH, NW>(constant_view(items)[bt.item], make_int2(bt.a, bt.b), symmetric, lds);
  }
}

template <typename T = int>
__global__ void reset_counter_kernel(T* counter, int n) {
  if (blockIdx.x == 0 && (int)threadIdx.x < n) counter[threadIdx.x] = 0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---------------------------------------------------------------------------
// Adjacent pairs: one wavefront per pair, lanes stride over the rule points.
// The canonical rule (shared vertices at canonical positions) is mapped to each
// cell by permuting its vertices; normals, Jacobian determinants and curls come
// from the cell's own vertex order, and canonical basis function k of a P1 cell
// is its local basis function perm[k]. P0P0: both spaces P0 (one accumulator).
// SMALL (Helmholtz): |k| |x - y| <= 0.5 for every point pair of every adjacent
// cell pair, so e^{ikr} uses the small-argument series without reduction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int perm_digit(int code, int k) { return (code >> (2 * k)) & 3; }

// A wave-uniform double moved to SGPRs (so that per-point FMAs can take it as their one
// scalar operand instead of holding a VGPR copy).
__device__ __forceinline__ double to_sgpr(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

template <int OP, bool HELM, bool P0P0, bool SMALL, typename Args>
__device__ __forceinline__ void singular_pair(Args& A, const int pair, const int lane) {
  const auto pw = constant_view((const int*)A.pairs) + 4 * (size_t)pair;
  const int tc = pw[0], sc = pw[1], m = pw[2], code = pw[3];
  const auto tg = constant_view(A.test_geo) + (size_t)tc * kGeoStride;
  const auto sg = constant_view(A.trial_geo) + (size_t)sc * kGeoStride;
  // x - y = D0 + E1 a0 + E2 a1 - F1 b0 - F2 b1: first-vertex difference (VGPRs) and edge
  // vectors (SGPRs) of the canonically ordered test and trial cells
  double D0[3], E1[3], E2[3], F1[3], F2[3];
  {
    const int t0 = perm_digit(code, 0), t1 = perm_digit(code, 1), t2 = perm_digit(code, 2);
    const int s0 = perm_digit(code >> 8, 0), s1 = perm_digit(code >> 8, 1), s2 = perm_digit(code >> 8, 2);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double x0 = tg[3 * t0 + d], x1 = tg[3 * t1 + d], x2 = tg[3 * t2 + d];
      const double y0 = sg[3 * s0 + d], y1 = sg[3 * s1 + d], y2 = sg[3 * s2 + d];
      D0[d] = x0 - y0;
      E1[d] = to_sgpr(x1 - x0);
      E2[d] = to_sgpr(x2 - x0);
      F1[d] = to_sgpr(y1 - y0);
      F2[d] = to_sgpr(y2 - y0);
    }
  }
  const double nx[3] = {tg[9], tg[10], tg[11]};
  const double ny[3] = {sg[9], sg[10], sg[11]};
  // selects rather than a dynamic index, so a struct copy of A stays in registers
  const int off = m == 1 ? A.rule_off[1] : (m == 2 ? A.rule_off[2] : A.rule_off[3]);
  const int n = m == 1 ? A.rule_n[1] : (m == 2 ? A.rule_n[2] : A.rule_n[3]);
  constexpr int NB = P0P0 ? 1 : 3;
  cplx M[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) M[a][b] = {0.0, 0.0};
  // one pointer per lane into the interleaved rule: the five loads share it (immediate offsets)
  const double* rp = A.rule_aos + (size_t)(off + lane) * 6;
  for (int pt = lane; pt < n; pt += 64, rp += 6 * 64) {
    const double2 ta = *(const double2*)rp, tb = *(const double2*)(rp + 2);
    const double a0 = ta.x, a1 = ta.y, b0 = tb.x, b1 = tb.y;
    const double w = rp[4];  // the Jacobians are applied once per pair below
    double d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = fma(E1[c], a0, fma(E2[c], a1, fma(-F1[c], b0, fma(-F2[c], b1, D0[c]))));
    if constexpr (P0P0 && (OP == 0 || OP == 3)) {
      green_accumulate<HELM, SMALL>(d[0], d[1], d[2], A.k, 0.0, w, M[0][0]);
    } else {
      const cplx kv = kernel_scalar<OP, HELM, SMALL>(d[0], d[1], d[2], A.k, nx, ny, 0.0);
      if constexpr (P0P0) {
        M[0][0].re = fma(kv.re, w, M[0][0].re);
        M[0][0].im = fma(kv.im, w, M[0][0].im);
      } else {
        const double lt[3] = {1.0 - a0 - a1, a0, a1};
        const double ls[3] = {1.0 - b0 - b1, b0, b1};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const double fa = (A.nbt == 3 ? lt[a] : (a == 0 ? 1.0 : 0.0)) * w;
          const cplx u = {kv.re * fa, kv.im * fa};
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            const double fb = A.nbs == 3 ? ls[b] : (b == 0 ? 1.0 : 0.0);
            M[a][b].re = fma(u.re, fb, M[a][b].re);
            M[a][b].im = fma(u.im, fb, M[a][b].im);
          }
        }
      }
    }
  }
  const double jj = tg[12] * sg[12];
  cplx Ml[3][3];
  cplx msum = {0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) Ml[a][b] = {0.0, 0.0};
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      Ml[a][b].re = wave_sum(M[a][b].re) * jj;
      Ml[a][b].im = HELM ? wave_sum(M[a][b].im) * jj : 0.0;
      msum.re += Ml[a][b].re;
      msum.im += Ml[a][b].im;
    }
  if (lane != 0) return;
  msum = hs_msum<OP>(msum, tg, sg);
  const int nbt = P0P0 ? 1 : A.nbt, nbs = P0P0 ? 1 : A.nbs;
  // loops unrolled over the largest block with uniform skips: constant indices keep Ml in
  // registers (a runtime bound would index it dynamically, through per-lane stack memory)
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b >= nbs) continue;
    const int lb = nbs == 3 ? perm_digit(code >> 8, b) : 0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      if (a >= nbt) continue;
      const int la = nbt == 3 ? perm_digit(code, a) : 0;
      const cplx v = local_entry<OP, HELM>(Ml, a, b, msum, nbt, nbs, A.test_geo + (size_t)tc * kGeoStride,
                                           A.trial_geo + (size_t)sc * kGeoStride, la, lb, A.k);
      if (A.mode == kCoo) {
        const long long e = (long long)pair * nbt * nbs + la + nbt * lb;
        double* o = (double*)A.out;
        o[2 * e] = v.re;
        o[2 * e + 1] = v.im;
      } else {
        const long long row = A.test_dofs[(size_t)tc * nbt + la];
        const long long col = A.trial_dofs[(size_t)sc * nbs + lb];
        store_value(A.out, row + A.ld * col, v, A.out_type, A.mode);
      }
    }
  }
}

template <int OP, bool HELM, bool P0P0, bool SMALL>
__global__ void __launch_bounds__(256) singular_kernel(const SingularArgs A) {
  const int pair = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6));
  if (pair >= A.npairs) return;  // uniform per wavefront
  singular_pair<OP, HELM, P0P0, SMALL>(A, pair, threadIdx.x & 63);
}

// Adjacent P0 x P0 pairs of a batch: one wavefront per pair of the concatenated
// pair lists (item i owns pairs [first[i], first[i + 1])); SMALL is chosen per item.
template <int OP, bool HELM>
__global__ void __launch_bounds__(256) batched_singular_kernel(const SingularArgs* __restrict__ items,
                                                               const int* __restrict__ small,
                                                               const int* __restrict__ first, int nitems) {
  const int g = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6));
  const auto fv = constant_view(first);
  if (g >= fv[nitems]) return;  // uniform per wavefront
  int lo = 0, hi = nitems - 1;  // last item with first[item] <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fv[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const auto& A = constant_view(items)[lo];
  const int lane = threadIdx.x & 63;
  if (HELM && constant_view(small)[lo])
    singular_pair<OP, HELM, true, true>(A, g - fv[lo], lane);
  else
    singular_pair<OP, HELM, true, false>(A, g - fv[lo], lane);
}

// Singular correction: the regular-rule sum over each adjacent cell pair (the part of
// a regular-rule evaluation over all point pairs -- e.g. by a fast multipole method --
// that the Duffy rules replace), point pairs at zero distance (the same point of a
// coincident pair) contributing nothing. One thread per adjacent pair; COO output of
// nbt x nbs entries per pair (cells' own local basis order), interleaved complex.
template <int Q, int OP, bool HELM>
__global__ void __launch_bounds__(256) correction_kernel(const RegularArgs A, const int4* __restrict__ pairs,
                                                         int npairs, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  const int4 pr = pairs[t];
  const int c = pr.x, d = pr.y;
  const double* tgeo = A.test_geo + (size_t)c * kGeoStride;
  const double* sgeo = A.trial_geo + (size_t)d * kGeoStride;
  const double nx[3] = {tgeo[9], tgeo[10], tgeo[11]};
  const double ny[3] = {sgeo[9], sgeo[10], sgeo[11]};
  cplx M[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) M[a][b] = {0.0, 0.0};
  // Q = 0: a rule of any point count (A.nq), basis values from A.rule_tab
  const int nq = Q > 0 ? Q : A.nq;
  auto phi = [&](int p, int a) { return Q > 0 ? A.phi[p][a] : A.rule_tab[4 * p + a]; };
  for (int p = 0; p < nq; ++p) {
    const double* xs = A.test_pts + ((size_t)c * nq + p) * 4;
    for (int q = 0; q < nq; ++q) {
      const double* ys = A.trial_pts + ((size_t)d * nq + q) * 4;
      const double dx = xs[0] - ys[0], dy = xs[1] - ys[1], dz = xs[2] - ys[2];
      if (dx == 0.0 && dy == 0.0 && dz == 0.0) continue;
      const cplx kv = kernel_scalar<OP, HELM, 0>(dx, dy, dz, A.k, nx, ny);
      const double w = xs[3] * ys[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double fa = (A.nbt == 3 ? phi(p, a) : (a == 0 ? 1.0 : 0.0)) * w;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const double f = fa * (A.nbs == 3 ? phi(q, b) : (b == 0 ? 1.0 : 0.0));
          M[a][b].re = fma(kv.re, f, M[a][b].re);
          M[a][b].im = fma(kv.im, f, M[a][b].im);
        }
      }
    }
  }
  cplx msum = {0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      msum.re += M[a][b].re;
      msum.im += M[a][b].im;
    }
  msum = hs_msum<OP>(msum, tgeo, sgeo);
  const int nbl = A.nbt * A.nbs;
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a >= A.nbt || b >= A.nbs) continue;
      const cplx v = local_entry<OP, HELM>(M, a, b, msum, A.nbt, A.nbs, tgeo, sgeo, a, b, A.k);
      const long long e = (long long)t * nbl + a + A.nbt * b;
      out[2 * e] = v.re;
      out[2 * e + 1] = v.im;
    }
}

// ---------------------------------------------------------------------------
// Potential (off-surface) operators: P[j][i] = sum_cells sum_q K(x_j, y_q) psi_i(y_q) w_q,
// K = G (single layer) or grad_y G . n_y (double layer). One thread per evaluation point
// (lane), the block walks a chunk of trial cells with wave-uniform scalar loads.
// ---------------------------------------------------------------------------
struct PotentialArgs {
  const double* x;          // [np][3] evaluation points
  long long np;
  const double* trial_pts;  // [n_trial][Q][4]
  const double* trial_geo;  // [n_trial][kGeoStride]
  const int* trial_dofs;    // [n_trial][nbs]
  int n_trial, nbs, trial_chunk;
  double k;
  double phi[kMaxQ][3];
  const double* rule_tab;   // [nq][4] basis values and weights (rules of any point count)
  int nq;
  void* out;                // [np x ndofs] column-major
  int out_type;
  int mode;                 // kAccumulate (P0: one thread per entry) or kAtomic (P1)
};

template <int Q, int OP, bool HELM>
__global__ void __launch_bounds__(256) potential_kernel(const PotentialArgs A) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const bool ok = t < A.np;
  const long long tl = ok ? t : A.np - 1;
  const double x[3] = {A.x[3 * tl], A.x[3 * tl + 1], A.x[3 * tl + 2]};
  const auto trial_pts = constant_view(A.trial_pts);
  const auto trial_geo = constant_view(A.trial_geo);
  const auto trial_dofs = constant_view(A.trial_dofs);
  const int j0 = blockIdx.y * A.trial_chunk;
  const int j1 = min(j0 + A.trial_chunk, A.n_trial);
  // Q = 0: a rule of any point count (A.nq), basis values from A.rule_tab
  const int nq = Q > 0 ? Q : A.nq;
  const auto tab = constant_view(A.rule_tab);
  for (int j = j0; j < j1; ++j) {
    const auto sp = trial_pts + (size_t)j * nq * 4;
    const auto sg = trial_geo + (size_t)j * kGeoStride;
    const double ny[3] = {sg[9], sg[10], sg[11]};
    cplx T[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int q = 0; q < (Q > 0 ? Q : nq); ++q) {
      const cplx kv = kernel_scalar<OP, HELM, 0>(x[0] - sp[4 * q], x[1] - sp[4 * q + 1], x[2] - sp[4 * q + 2], A.k,
                                                 ny, ny);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double ph = Q > 0 ? A.phi[q < kMaxQ ? q : 0][b] : tab[4 * q + b];
        const double f = (A.nbs == 3 ? ph : (b == 0 ? 1.0 : 0.0)) * sp[4 * q + 3];
        T[b].re = fma(kv.re, f, T[b].re);
        T[b].im = fma(kv.im, f, T[b].im);
      }
    }
    if (!ok) continue;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if (b >= A.nbs) continue;
      const long long col = trial_dofs[(size_t)j * A.nbs + b];
      store_value(A.out, t + A.np * col, {T[b].re * kInv4Pi, T[b].im * kInv4Pi}, A.out_type, A.mode);
    }
  }
}

// (a template so that every translation unit including this header may instantiate it)
template <typename T = double>
__global__ void fill_zero_kernel(T* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

}  // namespace dev

// Device inputs of the matrix-free kernels (apply, shape gradient), filled from the
// assembly caches by gradient_inputs() in device.hip.
struct GradientInputs {
  dev::RegularArgs R;
  const int4* pairs;
  int npairs;
  const double* rule;
  int rule_total;
  int rule_off[4];
  int rule_n[4];
  bool centred;
  int Q;
  int device;
  void* stream;
};

struct Assembler;      // core.hpp
struct FunctionSpace;  // core.hpp

// Device inputs of one dense assembly, filled by prepare() in device.hip from the
// grid/space caches (geometry, points, DOFs, adjacent pairs, rule tables).
struct Prepared {
  int device;
  bool centred;         // Helmholtz phases taken relative to the cell-pair centroid distance
  int tier;             // 0: not centred; 1: |k| (rad_i + rad_j) <= 0.5; 2: <= 0.25
  bool small_adjacent;  // |k| |x - y| <= 0.5 on every adjacent cell pair
  bool mixed;           // FP32 far pairs + FP64 near-pair list (P0 x P0 on one grid)
  const int2* near;
  int nnear;
  const int4* near_segs;  // near pairs grouped by (test block, trial block): tx, ty, first, count
  int nsegs;
  hipStream_t stream;
  int Q;
  bool same_grid;
  dev::RegularArgs R;
  dev::SingularArgs S;
  int npairs;
};

// stream: the launch stream; NULL selects the assembler's own stream unless caller_stream
// is set (then NULL is the null stream).
Prepared prepare(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, void* out,
                 void* stream, bool caller_stream = false);

}  // namespace bempp_amd
