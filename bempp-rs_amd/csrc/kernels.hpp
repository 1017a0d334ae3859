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

namespace bempp_amd {
namespace dev {

constexpr double kInv4Pi = 0.079577471545947667884441881686257181;  // 1/(4 pi)
constexpr int kGeoStride = 24;  // v0 v1 v2 (9) | n (3) | jdet (1) | curlJ_b (9) | pad (2)
constexpr int kMaxQ = 7;

enum OutType { kOutF64 = 0, kOutC128 = 1, kOutF32 = 2, kOutC64 = 3 };
enum Mode { kStore = 0, kAccumulate = 1, kAtomic = 2, kCoo = 3 };

struct RegularArgs {
  const double* test_pts;   // [n_test][Q][4]: x, y, z, w_p * jdet
  const double* trial_pts;  // [n_trial][Q][4]
  const double* test_geo;   // [n_test][kGeoStride]
  const double* trial_geo;  // [n_trial][kGeoStride]
  const int* test_vid;      // [n_test][4]
  const int* trial_vid;     // [n_trial][4]
  const int* test_dofs;     // [n_test][nbt]
  const int* trial_dofs;    // [n_trial][nbs]
  int n_test, n_trial;
  int nbt, nbs;
  int same_grid;
  int trial_chunk;
  double k;
  double phi[kMaxQ][3];     // basis values at the regular points
  void* out;
  long long ld;
  int out_type;
  int mode;
};

struct SingularArgs {
  const int4* pairs;        // test cell, trial cell, shared count, permutation code
  int npairs;
  const double* rule;       // 5 arrays of rule_total: test xi0, test xi1, trial xi0, trial xi1, weight
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

// Kernel scalar (without the 1/(4 pi) factor) for displacement d = x - y.
// nx: test normal, ny: trial normal.
template <int OP, bool HELM>
__device__ __forceinline__ cplx kernel_scalar(double dx, double dy, double dz, double k, const double* nx,
                                              const double* ny) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double rinv = 1.0 / sqrt(r2);
  cplx g;
  double kr = 0.0, s = 0.0, c = 1.0;
  if constexpr (HELM) {
    kr = k * (r2 * rinv);
    sincos(kr, &s, &c);
    g = {c * rinv, s * rinv};
  } else {
    g = {rinv, 0.0};
  }
  if constexpr (OP == 0 || OP == 3) {
    return g;
  } else {
    double dn;
    if constexpr (OP == 1)
      dn = dx * ny[0] + dy * ny[1] + dz * ny[2];
    else
      dn = -(dx * nx[0] + dy * nx[1] + dz * nx[2]);
    const double f = dn * rinv * rinv * rinv;
    if constexpr (HELM) {
      // e^{ikr}(1 - ikr) = (c + kr s) + i (s - kr c)
      return {(c + kr * s) * f, (s - kr * c) * f};
    } else {
      return {f, 0.0};
    }
  }
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

// Epilogue shared by both loops: M (nbt x nbs local sums without 1/(4pi)) ->
// local matrix entries. tgeo/sgeo: geometry records; lt/ls: local basis index
// of each row/column of M.
template <int OP, bool HELM>
__device__ __forceinline__ cplx local_entry(const cplx M[3][3], int a, int b, cplx msum, int nbt, int nbs,
                                            const double* tgeo, const double* sgeo, int la, int lb, double k) {
  cplx v = M[a][b];
  if constexpr (OP == 3) {
    cplx out = {0.0, 0.0};
    if (nbt == 3 && nbs == 3) {
      const double* ct = tgeo + 13 + 3 * la;
      const double* cs = sgeo + 13 + 3 * lb;
      const double cc = (ct[0] * cs[0] + ct[1] * cs[1] + ct[2] * cs[2]) / (tgeo[12] * sgeo[12]);
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

// ---------------------------------------------------------------------------
// Regular pairs. One thread per test cell; the block's trial cells are walked
// in order with wave-uniform (scalar) loads of the trial geometry. P0P0:
// both spaces P0, the thread owns output row i and writes column j directly
// (coalesced along the test index). Otherwise the local block is added with
// atomics at the DOF indices.
// ---------------------------------------------------------------------------
template <int Q, int OP, bool HELM, bool P0P0>
__global__ void __launch_bounds__(256) regular_kernel(const RegularArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n_test) return;
  double xp[Q][3], wp[Q];
#pragma unroll
  for (int p = 0; p < Q; ++p) {
    const double* s = A.test_pts + ((size_t)i * Q + p) * 4;
    xp[p][0] = s[0];
    xp[p][1] = s[1];
    xp[p][2] = s[2];
    wp[p] = s[3];
  }
  const double* tgeo = A.test_geo + (size_t)i * kGeoStride;
  const double nx[3] = {tgeo[9], tgeo[10], tgeo[11]};
  const int tv0 = A.test_vid[4 * i], tv1 = A.test_vid[4 * i + 1], tv2 = A.test_vid[4 * i + 2];
  const int j0 = blockIdx.y * A.trial_chunk;
  const int j1 = min(j0 + A.trial_chunk, A.n_trial);
  for (int j = j0; j < j1; ++j) {
    const double* sp = A.trial_pts + (size_t)j * Q * 4;
    const double* sgeo = A.trial_geo + (size_t)j * kGeoStride;
    const int sv0 = A.trial_vid[4 * j], sv1 = A.trial_vid[4 * j + 1], sv2 = A.trial_vid[4 * j + 2];
    const bool adjacent = A.same_grid && (tv0 == sv0 || tv0 == sv1 || tv0 == sv2 || tv1 == sv0 || tv1 == sv1 ||
                                          tv1 == sv2 || tv2 == sv0 || tv2 == sv1 || tv2 == sv2);
    const double ny[3] = {sgeo[9], sgeo[10], sgeo[11]};
    if constexpr (P0P0) {
      cplx acc = {0.0, 0.0};
      if (!adjacent) {
#pragma unroll
        for (int p = 0; p < Q; ++p) {
          cplx ap = {0.0, 0.0};
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const cplx kv = kernel_scalar<OP, HELM>(xp[p][0] - sp[4 * q], xp[p][1] - sp[4 * q + 1],
                                                     xp[p][2] - sp[4 * q + 2], A.k, nx, ny);
            ap.re += kv.re * sp[4 * q + 3];
            ap.im += kv.im * sp[4 * q + 3];
          }
          acc.re += ap.re * wp[p];
          acc.im += ap.im * wp[p];
        }
      }
      if (adjacent && A.mode == kAccumulate) continue;
      cplx M[3][3];
      M[0][0] = acc;
      const cplx v = local_entry<OP, HELM>(M, 0, 0, acc, 1, 1, tgeo, sgeo, 0, 0, A.k);
      store_value(A.out, (long long)i + A.ld * (long long)j, v, A.out_type, A.mode);
    } else {
      if (adjacent) continue;
      cplx M[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) M[a][b] = {0.0, 0.0};
#pragma unroll
      for (int p = 0; p < Q; ++p) {
        cplx T[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const cplx kv = kernel_scalar<OP, HELM>(xp[p][0] - sp[4 * q], xp[p][1] - sp[4 * q + 1],
                                                   xp[p][2] - sp[4 * q + 2], A.k, nx, ny);
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            const double f = (A.nbs == 3 ? A.phi[q][b] : (b == 0 ? 1.0 : 0.0)) * sp[4 * q + 3];
            T[b].re += kv.re * f;
            T[b].im += kv.im * f;
          }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const double f = (A.nbt == 3 ? A.phi[p][a] : (a == 0 ? 1.0 : 0.0)) * wp[p];
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            M[a][b].re += f * T[b].re;
            M[a][b].im += f * T[b].im;
          }
        }
      }
      cplx msum = {0.0, 0.0};
      for (int a = 0; a < A.nbt; ++a)
        for (int b = 0; b < A.nbs; ++b) {
          msum.re += M[a][b].re;
          msum.im += M[a][b].im;
        }
      for (int b = 0; b < A.nbs; ++b) {
        const long long col = A.trial_dofs[(size_t)j * A.nbs + b];
        for (int a = 0; a < A.nbt; ++a) {
          const long long row = A.test_dofs[(size_t)i * A.nbt + a];
          const cplx v = local_entry<OP, HELM>(M, a, b, msum, A.nbt, A.nbs, tgeo, sgeo, a, b, A.k);
          store_value(A.out, row + A.ld * col, v, A.out_type, kAtomic);
        }
      }
    }
  }
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
// is its local basis function perm[k].
// ---------------------------------------------------------------------------
template <int OP, bool HELM>
__global__ void __launch_bounds__(256) singular_kernel(const SingularArgs A) {
  const int pair = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (pair >= A.npairs) return;  // uniform per wavefront
  const int4 P = A.pairs[pair];
  const int tc = P.x, sc = P.y, m = P.z, code = P.w;
  int tperm[3], sperm[3];
#pragma unroll
  for (int kk = 0; kk < 3; ++kk) {
    tperm[kk] = (code >> (2 * kk)) & 3;
    sperm[kk] = (code >> (8 + 2 * kk)) & 3;
  }
  const double* tgeo = A.test_geo + (size_t)tc * kGeoStride;
  const double* sgeo = A.trial_geo + (size_t)sc * kGeoStride;
  double tv[3][3], sv[3][3];
#pragma unroll
  for (int kk = 0; kk < 3; ++kk)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      tv[kk][d] = tgeo[3 * tperm[kk] + d];
      sv[kk][d] = sgeo[3 * sperm[kk] + d];
    }
  const double nx[3] = {tgeo[9], tgeo[10], tgeo[11]};
  const double ny[3] = {sgeo[9], sgeo[10], sgeo[11]};
  const double jj = tgeo[12] * sgeo[12];
  const int off = A.rule_off[m], n = A.rule_n[m], tot = A.rule_total;
  cplx M[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) M[a][b] = {0.0, 0.0};
  for (int pt = lane; pt < n; pt += 64) {
    const double a0 = A.rule[off + pt], a1 = A.rule[tot + off + pt];
    const double b0 = A.rule[2 * tot + off + pt], b1 = A.rule[3 * tot + off + pt];
    const double w = A.rule[4 * tot + off + pt] * jj;
    const double lt[3] = {1.0 - a0 - a1, a0, a1};
    const double ls[3] = {1.0 - b0 - b1, b0, b1};
    double x[3], y[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      x[d] = tv[0][d] * lt[0] + tv[1][d] * lt[1] + tv[2][d] * lt[2];
      y[d] = sv[0][d] * ls[0] + sv[1][d] * ls[1] + sv[2][d] * ls[2];
    }
    cplx kv = kernel_scalar<OP, HELM>(x[0] - y[0], x[1] - y[1], x[2] - y[2], A.k, nx, ny);
    kv.re *= w;
    kv.im *= w;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double fa = A.nbt == 3 ? lt[a] : (a == 0 ? 1.0 : 0.0);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double f = fa * (A.nbs == 3 ? ls[b] : (b == 0 ? 1.0 : 0.0));
        M[a][b].re += kv.re * f;
        M[a][b].im += kv.im * f;
      }
    }
  }
  cplx msum = {0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      M[a][b].re = wave_sum(M[a][b].re);
      M[a][b].im = HELM ? wave_sum(M[a][b].im) : 0.0;
      msum.re += M[a][b].re;
      msum.im += M[a][b].im;
    }
  if (lane != 0) return;
  for (int b = 0; b < A.nbs; ++b) {
    const int lb = A.nbs == 3 ? sperm[b] : 0;
    for (int a = 0; a < A.nbt; ++a) {
      const int la = A.nbt == 3 ? tperm[a] : 0;
      const cplx v = local_entry<OP, HELM>(M, a, b, msum, A.nbt, A.nbs, tgeo, sgeo, la, lb, A.k);
      if (A.mode == kCoo) {
        const long long e = (long long)pair * A.nbt * A.nbs + la + A.nbt * lb;
        double* o = (double*)A.out;
        o[2 * e] = v.re;
        o[2 * e + 1] = v.im;
      } else {
        const long long row = A.test_dofs[(size_t)tc * A.nbt + la];
        const long long col = A.trial_dofs[(size_t)sc * A.nbs + lb];
        store_value(A.out, row + A.ld * col, v, A.out_type, A.mode);
      }
    }
  }
}

__global__ void fill_zero_kernel(double* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

}  // namespace dev
}  // namespace bempp_amd
