// gradient.hip -- matrix-free operator application y = A u and the shape
// gradient d(v^T A(X) u)/dX of the single-layer operator (P0 x P0, one grid),
// both by recomputation (no assembled matrix).
//
// Shape gradient. With x(xi) = sum_k lambda_k(xi) V_k on a flat cell, J = |(V1-V0) x (V2-V0)|
// and dJ/dV_k = n x (V_{k-1} - V_{k+1}) (indices mod 3), a cell-pair integral
// A_cj = sum_{p,q} (w_p J_c)(w_q J_j) G(x_p, y_q) has the test-side derivative
//   dA_cj/dV_k(c) = (dJ_c/dV_k / J_c) A_cj + sum_{p,q} lambda_k(p) (w_p J_c)(w_q J_j) grad_x G(x_p, y_q).
// For the regular rule A_cj(X) = A_jc(X) as functions of the geometry (same rule on
// both cells, G symmetric), so d(v^T A u)/d(cell c) = sum_j (v_c u_j + u_c v_j) dA_cj/d(cell c):
// one thread per test cell, no cross-thread reduction. Adjacent pairs use Duffy
// rules that are not symmetric, so both cells are differentiated explicitly.
// grad_x G = G'(r) (x - y)/r with G' = e^{ikr}(ikr - 1)/(4 pi r^2) (Laplace: k = 0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>
#include <string>
#include <vector>

#include "core.hpp"
#include "kernels.hpp"

namespace bempp_amd {

#define HIP_CHECK_G(expr)                                                                         \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) throw Error(std::string(#expr) + " failed: " + hipGetErrorString(_e)); \
  } while (0)

namespace dev {

struct GradArgs {
  RegularArgs R;       // geometry, points, k (the output fields are unused)
  const double* u;     // [n] complex interleaved (Laplace: imaginary parts zero)
  const double* v;     // [n]
  double* gcell;       // [n_cells][3 vertices][3 comps][2] accumulated (atomics)
  const int4* pairs;   // adjacent pairs (Duffy)
  int npairs;
  const double* rule;
  int rule_total;
  int rule_off[4];
  int rule_n[4];
  double lam[kMaxQ][3];  // barycentric coordinates of the regular points
};

// G (value) and F with grad_x G = F (x - y), both without 1/(4 pi); CENTRED:
// without the factor e^{ikR}.
template <bool HELM, bool CENTRED>
__device__ __forceinline__ void green_and_grad(double dx, double dy, double dz, double k, double kR, cplx* G,
                                               cplx* F) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double rinv = rsqrt_nr(r2);
  const double rinv3 = rinv * rinv * rinv;
  if constexpr (HELM) {
    const double r = r2 * rinv;
    double s, c;
    if constexpr (CENTRED)
      sincos_small(fma(k, r, -kR), &s, &c);
    else
      sincos_cw(k * r, &s, &c);
    const double kr = k * r;
    *G = {c * rinv, s * rinv};
    // e^{ikr}(ikr - 1) = -[(c + kr s) + i (s - kr c)]
    *F = {-(c + kr * s) * rinv3, -(s - kr * c) * rinv3};
  } else {
    *G = {rinv, 0.0};
    *F = {-rinv3, 0.0};
  }
}

__device__ __forceinline__ void atomic_add_c(double* p, cplx v) {
  unsafeAtomicAdd(p, v.re);
  unsafeAtomicAdd(p + 1, v.im);
}

// dJ/dV_k / J for a flat cell: n x (V_{k-1} - V_{k+1}) / J.
__device__ __forceinline__ void djac_over_j(const double* geo, int k, double out[3]) {
  const double* vm = geo + 3 * ((k + 2) % 3);
  const double* vp = geo + 3 * ((k + 1) % 3);
  const double e[3] = {vm[0] - vp[0], vm[1] - vp[1], vm[2] - vp[2]};
  const double* n = geo + 9;
  const double jinv = 1.0 / geo[12];
  out[0] = (n[1] * e[2] - n[2] * e[1]) * jinv;
  out[1] = (n[2] * e[0] - n[0] * e[2]) * jinv;
  out[2] = (n[0] * e[1] - n[1] * e[0]) * jinv;
}

// Regular pairs: one thread per test cell c, a chunk of trial cells j per block.
// Q = 0: a rule of any point count (A.nq), test points streamed, barycentric coordinates
// from A.rule_tab.
template <int Q, bool HELM, bool CENTRED>
__global__ void __launch_bounds__(256) grad_regular_kernel(const GradArgs G) {
  const RegularArgs& A = G.R;
  constexpr bool GEN = Q == 0;
  constexpr int QR = GEN ? 1 : Q;
  const int nq = GEN ? A.nq : Q;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const bool c_ok = c < A.n_test;
  const int cl = c_ok ? c : A.n_test - 1;
  const double* tpts = A.test_pts + (size_t)cl * nq * 4;
  double xp[QR][3], wp[QR];
  if constexpr (!GEN) {
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const double* s = tpts + 4 * p;
      xp[p][0] = s[0];
      xp[p][1] = s[1];
      xp[p][2] = s[2];
      wp[p] = s[3];
    }
  }
  const auto tab = constant_view(A.rule_tab);
  const double* tgeo = A.test_geo + (size_t)cl * kGeoStride;
  const double ct[3] = {tgeo[kGeoCentroid], tgeo[kGeoCentroid + 1], tgeo[kGeoCentroid + 2]};
  const int tv0 = A.test_vid[4 * cl], tv1 = A.test_vid[4 * cl + 1], tv2 = A.test_vid[4 * cl + 2];
  const cplx uc = {G.u[2 * cl], G.u[2 * cl + 1]}, vc = {G.v[2 * cl], G.v[2 * cl + 1]};
  cplx Pacc = {0.0, 0.0};
  cplx Qacc[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) Qacc[k][d] = {0.0, 0.0};
  const auto trial_pts = constant_view(A.trial_pts);
  const auto trial_geo = constant_view(A.trial_geo);
  const auto trial_vid = constant_view(A.trial_vid);
  const auto uu = constant_view(G.u);
  const auto vv = constant_view(G.v);
  const int j0 = blockIdx.y * A.trial_chunk;
  const int j1 = min(j0 + A.trial_chunk, A.n_trial);
  for (int j = j0; j < j1; ++j) {
    const auto spp = trial_pts + (size_t)j * nq * 4;
    double sp[QR * 4];
    if constexpr (!GEN) {
#pragma unroll
      for (int m = 0; m < Q * 4; ++m) sp[m] = spp[m];
    }
    const auto sg = trial_geo + (size_t)j * kGeoStride;
    const double cs[3] = {sg[kGeoCentroid], sg[kGeoCentroid + 1], sg[kGeoCentroid + 2]};
    const int sv0 = trial_vid[4 * j], sv1 = trial_vid[4 * j + 1], sv2 = trial_vid[4 * j + 2];
    const cplx uj = {uu[2 * j], uu[2 * j + 1]}, vj = {vv[2 * j], vv[2 * j + 1]};
    const bool adjacent = tv0 == sv0 || tv0 == sv1 || tv0 == sv2 || tv1 == sv0 || tv1 == sv1 || tv1 == sv2 ||
                          tv2 == sv0 || tv2 == sv1 || tv2 == sv2;
    if (adjacent || !c_ok) continue;
    double kR = 0.0;
    cplx e = {1.0, 0.0};
    if constexpr (CENTRED) e = pair_phase(ct, cs, A.k, &kR);
    cplx P = {0.0, 0.0};
    cplx T[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) T[k][d] = {0.0, 0.0};
#pragma unroll
    for (int p = 0; p < (GEN ? nq : Q); ++p) {
      const double x0 = GEN ? tpts[4 * p] : xp[p][0], x1 = GEN ? tpts[4 * p + 1] : xp[p][1];
      const double x2 = GEN ? tpts[4 * p + 2] : xp[p][2], wpp = GEN ? tpts[4 * p + 3] : wp[p];
      cplx Pp = {0.0, 0.0};
      cplx Tp[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
      for (int q = 0; q < (GEN ? nq : Q); ++q) {
        const double y0 = GEN ? spp[4 * q] : sp[4 * q], y1 = GEN ? spp[4 * q + 1] : sp[4 * q + 1];
        const double y2 = GEN ? spp[4 * q + 2] : sp[4 * q + 2], wq = GEN ? spp[4 * q + 3] : sp[4 * q + 3];
        const double d[3] = {x0 - y0, x1 - y1, x2 - y2};
        cplx g, f;
        green_and_grad<HELM, CENTRED>(d[0], d[1], d[2], A.k, kR, &g, &f);
        Pp.re = fma(g.re, wq, Pp.re);
        Pp.im = fma(g.im, wq, Pp.im);
        const cplx fw = {f.re * wq, f.im * wq};
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) {
          Tp[dd].re = fma(fw.re, d[dd], Tp[dd].re);
          Tp[dd].im = fma(fw.im, d[dd], Tp[dd].im);
        }
      }
      P.re = fma(Pp.re, wpp, P.re);
      P.im = fma(Pp.im, wpp, P.im);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double lw = (GEN ? tab[4 * p + k] : G.lam[GEN ? 0 : p][k]) * wpp;
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) {
          T[k][dd].re = fma(lw, Tp[dd].re, T[k][dd].re);
          T[k][dd].im = fma(lw, Tp[dd].im, T[k][dd].im);
        }
      }
    }
    // weight (v_c u_j + u_c v_j), times e^{ikR}
    cplx W = {vc.re * uj.re - vc.im * uj.im + uc.re * vj.re - uc.im * vj.im,
              vc.re * uj.im + vc.im * uj.re + uc.re * vj.im + uc.im * vj.re};
    W = cmul(W, e);
    const cplx wP = cmul(W, P);
    Pacc.re += wP.re;
    Pacc.im += wP.im;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int dd = 0; dd < 3; ++dd) {
        const cplx t = cmul(W, T[k][dd]);
        Qacc[k][dd].re += t.re;
        Qacc[k][dd].im += t.im;
      }
  }
  if (!c_ok) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double dj[3];
    djac_over_j(tgeo, k, dj);
#pragma unroll
    for (int dd = 0; dd < 3; ++dd) {
      const cplx gsum = {(dj[dd] * Pacc.re + Qacc[k][dd].re) * kInv4Pi, (dj[dd] * Pacc.im + Qacc[k][dd].im) * kInv4Pi};
      atomic_add_c(G.gcell + (((size_t)c * 3 + k) * 3 + dd) * 2, gsum);
    }
  }
}

// Matrix-free y = A u for P0 x P0, regular pairs: one thread per test cell, a
// chunk of trial cells per block, y_i += sum_j A_ij u_j (atomics across chunks).
template <int Q, int OP, bool HELM, bool CENTRED>
__global__ void __launch_bounds__(256) apply_regular_kernel(const GradArgs G, double* y) {
  const RegularArgs& A = G.R;
  constexpr bool GEN = Q == 0;  // a rule of any point count (A.nq), points streamed
  constexpr int QR = GEN ? 1 : Q;
  const int nq = GEN ? A.nq : Q;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool i_ok = i < A.n_test;
  const int il = i_ok ? i : A.n_test - 1;
  const double* tpts = A.test_pts + (size_t)il * nq * 4;
  double xp[QR][3], wp[QR];
  if constexpr (!GEN) {
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const double* s = tpts + 4 * p;
      xp[p][0] = s[0];
      xp[p][1] = s[1];
      xp[p][2] = s[2];
      wp[p] = s[3];
    }
  }
  const double* tgeo = A.test_geo + (size_t)il * kGeoStride;
  const double nx[3] = {tgeo[9], tgeo[10], tgeo[11]};
  const double ct[3] = {tgeo[kGeoCentroid], tgeo[kGeoCentroid + 1], tgeo[kGeoCentroid + 2]};
  const int tv0 = A.test_vid[4 * il], tv1 = A.test_vid[4 * il + 1], tv2 = A.test_vid[4 * il + 2];
  const auto trial_pts = constant_view(A.trial_pts);
  const auto trial_geo = constant_view(A.trial_geo);
  const auto trial_vid = constant_view(A.trial_vid);
  const auto uu = constant_view(G.u);
  cplx acc = {0.0, 0.0};
  const int j0 = blockIdx.y * A.trial_chunk;
  const int j1 = min(j0 + A.trial_chunk, A.n_trial);
  for (int j = j0; j < j1; ++j) {
    const auto spp = trial_pts + (size_t)j * nq * 4;
    double sp[QR * 4];
    if constexpr (!GEN) {
#pragma unroll
      for (int m = 0; m < Q * 4; ++m) sp[m] = spp[m];
    }
    const auto sgp = trial_geo + (size_t)j * kGeoStride;
    double sg[kGeoStride];
#pragma unroll
    for (int m = 0; m < kGeoStride; ++m) sg[m] = sgp[m];
    const int sv0 = trial_vid[4 * j], sv1 = trial_vid[4 * j + 1], sv2 = trial_vid[4 * j + 2];
    const cplx uj = {uu[2 * j], uu[2 * j + 1]};
    const bool adjacent = A.same_grid && (tv0 == sv0 || tv0 == sv1 || tv0 == sv2 || tv1 == sv0 || tv1 == sv1 ||
                                          tv1 == sv2 || tv2 == sv0 || tv2 == sv1 || tv2 == sv2);
    if (adjacent || !i_ok) continue;
    const double ny[3] = {sg[9], sg[10], sg[11]};
    cplx v;
    double kR = 0.0;
    cplx e = {1.0, 0.0};
    if constexpr (CENTRED) e = pair_phase(ct, sg + kGeoCentroid, A.k, &kR);
    if constexpr (GEN)
      v = p0_pair_generic<OP, HELM, CENTRED>(tpts, nq, nx, spp, ny, A.k, kR);
    else
      v = p0_pair<Q, OP, HELM, CENTRED>(xp, wp, nx, sp, ny, A.k, kR);
    if constexpr (CENTRED) v = cmul(v, e);
    if constexpr (OP == 3) {
      const double nn = HELM ? -A.k * A.k * (nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2]) : 0.0;
      v = {nn * v.re, nn * v.im};
    }
    const cplx t = cmul(v, uj);
    acc.re += t.re;
    acc.im += t.im;
  }
  if (i_ok) atomic_add_c(y + 2 * (size_t)i, {acc.re * kInv4Pi, acc.im * kInv4Pi});
}

// Adjacent pairs: one wavefront per pair, both cells differentiated (weight v_i u_j).
template <bool HELM>
__global__ void __launch_bounds__(256) grad_singular_kernel(const GradArgs G) {
  const RegularArgs& A = G.R;
  const int pair = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6));
  const int lane = threadIdx.x & 63;
  if (pair >= G.npairs) return;
  const auto pw = constant_view((const int*)G.pairs) + 4 * (size_t)pair;
  const int tc = pw[0], sc = pw[1], m = pw[2], code = pw[3];
  const double* tg = A.test_geo + (size_t)tc * kGeoStride;
  const double* sg = A.trial_geo + (size_t)sc * kGeoStride;
  double tv[3][3], sv[3][3];
#pragma unroll
  for (int kk = 0; kk < 3; ++kk) {
    const int pt = (code >> (2 * kk)) & 3, ps = (code >> (8 + 2 * kk)) & 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      tv[kk][d] = tg[3 * pt + d];
      sv[kk][d] = sg[3 * ps + d];
    }
  }
  const double jj = tg[12] * sg[12];
  const int off = G.rule_off[m], n = G.rule_n[m], tot = G.rule_total;
  cplx P = {0.0, 0.0};
  cplx Tt[3][3], Ts[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) Tt[k][d] = Ts[k][d] = {0.0, 0.0};
  for (int pt = lane; pt < n; pt += 64) {
    const double a0 = G.rule[off + pt], a1 = G.rule[tot + off + pt];
    const double b0 = G.rule[2 * tot + off + pt], b1 = G.rule[3 * tot + off + pt];
    const double w = G.rule[4 * tot + off + pt] * jj;
    const double lt[3] = {1.0 - a0 - a1, a0, a1};
    const double ls[3] = {1.0 - b0 - b1, b0, b1};
    double d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      d[c] = (tv[0][c] * lt[0] + tv[1][c] * lt[1] + tv[2][c] * lt[2]) -
             (sv[0][c] * ls[0] + sv[1][c] * ls[1] + sv[2][c] * ls[2]);
    cplx g, f;
    green_and_grad<HELM, false>(d[0], d[1], d[2], A.k, 0.0, &g, &f);
    P.re = fma(g.re, w, P.re);
    P.im = fma(g.im, w, P.im);
    const cplx fw = {f.re * w, f.im * w};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int dd = 0; dd < 3; ++dd) {
        const cplx fd = {fw.re * d[dd], fw.im * d[dd]};
        Tt[k][dd].re = fma(lt[k], fd.re, Tt[k][dd].re);
        Tt[k][dd].im = fma(lt[k], fd.im, Tt[k][dd].im);
        Ts[k][dd].re = fma(-ls[k], fd.re, Ts[k][dd].re);  // grad_y G = -grad_x G
        Ts[k][dd].im = fma(-ls[k], fd.im, Ts[k][dd].im);
      }
  }
  P.re = wave_sum(P.re);
  P.im = wave_sum(P.im);
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int dd = 0; dd < 3; ++dd) {
      Tt[k][dd].re = wave_sum(Tt[k][dd].re);
      Tt[k][dd].im = wave_sum(Tt[k][dd].im);
      Ts[k][dd].re = wave_sum(Ts[k][dd].re);
      Ts[k][dd].im = wave_sum(Ts[k][dd].im);
    }
  if (lane != 0) return;
  const cplx W = cmul({G.v[2 * tc], G.v[2 * tc + 1]}, {G.u[2 * sc], G.u[2 * sc + 1]});
  const cplx WP = cmul(W, P);
  for (int k = 0; k < 3; ++k) {
    const int lt_ = (code >> (2 * k)) & 3, ls_ = (code >> (8 + 2 * k)) & 3;
    double djt[3], djs[3];
    djac_over_j(tg, lt_, djt);
    djac_over_j(sg, ls_, djs);
    for (int dd = 0; dd < 3; ++dd) {
      const cplx at = cmul(W, Tt[k][dd]);
      const cplx as = cmul(W, Ts[k][dd]);
      atomic_add_c(G.gcell + (((size_t)tc * 3 + lt_) * 3 + dd) * 2,
                   {(djt[dd] * WP.re + at.re) * kInv4Pi, (djt[dd] * WP.im + at.im) * kInv4Pi});
      atomic_add_c(G.gcell + (((size_t)sc * 3 + ls_) * 3 + dd) * 2,
                   {(djs[dd] * WP.re + as.re) * kInv4Pi, (djs[dd] * WP.im + as.im) * kInv4Pi});
    }
  }
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Filled by device.hip (shares its caches).
GradientInputs gradient_inputs(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test,
                               void* stream = nullptr, bool caller_stream = false);

namespace {

template <int Q, bool HELM, bool CENTRED>
void launch_grad_regular(dim3 grid, hipStream_t s, const dev::GradArgs& G) {
  dev::grad_regular_kernel<Q, HELM, CENTRED><<<grid, 256, 0, s>>>(G);
}

void launch_grad(int Q, bool helm, bool centred, dim3 grid, hipStream_t s, const dev::GradArgs& G) {
#define BEMPP_G(QQ)                                                          \
  if (!helm) launch_grad_regular<QQ, false, false>(grid, s, G);              \
  else if (centred) launch_grad_regular<QQ, true, true>(grid, s, G);         \
  else launch_grad_regular<QQ, true, false>(grid, s, G);
  switch (Q) {
    case 1: BEMPP_G(1); break;
    case 3: BEMPP_G(3); break;
    case 6: BEMPP_G(6); break;
    case 7: BEMPP_G(7); break;
    default: BEMPP_G(0); break;  // any other rule: point count at run time
  }
#undef BEMPP_G
  HIP_CHECK_G(hipGetLastError());
}

}  // namespace

namespace {

// dtype array -> interleaved complex128
std::vector<double> to_c128(const void* x, DType dt, int64_t n) {
  std::vector<double> h(2 * n, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    switch (dt) {
      case kF64: h[2 * i] = ((const double*)x)[i]; break;
      case kF32: h[2 * i] = ((const float*)x)[i]; break;
      case kC64:
        h[2 * i] = ((const double*)x)[2 * i];
        h[2 * i + 1] = ((const double*)x)[2 * i + 1];
        break;
      default:
        h[2 * i] = ((const float*)x)[2 * i];
        h[2 * i + 1] = ((const float*)x)[2 * i + 1];
        break;
    }
  }
  return h;
}

template <bool HELM, bool CENTRED>
void launch_apply_op(int op, int Q, dim3 grid, hipStream_t s, const dev::GradArgs& G, double* y) {
#define BEMPP_A(QQ)                                                                                   \
  switch (op) {                                                                                       \
    case kSL: dev::apply_regular_kernel<QQ, kSL, HELM, CENTRED><<<grid, 256, 0, s>>>(G, y); break;   \
    case kDL: dev::apply_regular_kernel<QQ, kDL, HELM, CENTRED><<<grid, 256, 0, s>>>(G, y); break;   \
    case kADL: dev::apply_regular_kernel<QQ, kADL, HELM, CENTRED><<<grid, 256, 0, s>>>(G, y); break; \
    case kHS: dev::apply_regular_kernel<QQ, kHS, HELM, CENTRED><<<grid, 256, 0, s>>>(G, y); break;   \
    default: throw Error("operator not supported");                                                  \
  }
  switch (Q) {
    case 1: BEMPP_A(1); break;
    case 3: BEMPP_A(3); break;
    case 6: BEMPP_A(6); break;
    case 7: BEMPP_A(7); break;
    default: BEMPP_A(0); break;  // any other rule: point count at run time
  }
#undef BEMPP_A
  HIP_CHECK_G(hipGetLastError());
}

}  // namespace

// y = A u without assembling A (P0 x P0): regular pairs recomputed on the device,
// adjacent pairs through the Duffy kernel's COO values.
void apply(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, const void* u, void* y,
           bool transpose) {
  if (test.nb != 1 || trial.nb != 1) throw Error("matrix-free apply is implemented for P0 spaces");
  // A^T u: the regular part of the single-layer and hypersingular operators on one
  // space is symmetric (same rule on both cells), so only the Duffy part is transposed.
  if (transpose && !((a.op == kSL || a.op == kHS) && test.grid == trial.grid))
    throw Error("transposed apply is implemented for the single-layer and hypersingular operators on one space");
  GradientInputs in = gradient_inputs(a, trial, test);
  hipStream_t s = (hipStream_t)in.stream;
  const int64_t nt = test.size, ns = trial.size;
  std::vector<double> hu = to_c128(u, a.dtype, ns), hy(2 * nt, 0.0);
  double *du = nullptr, *dy = nullptr;
  try {
    HIP_CHECK_G(hipMalloc((void**)&du, hu.size() * sizeof(double)));
    HIP_CHECK_G(hipMalloc((void**)&dy, hy.size() * sizeof(double)));
    HIP_CHECK_G(hipMemcpyAsync(du, hu.data(), hu.size() * sizeof(double), hipMemcpyHostToDevice, s));
    dev::fill_zero_kernel<<<256, 256, 0, s>>>(dy, 2 * nt);
    HIP_CHECK_G(hipGetLastError());
    dev::GradArgs G{};
    G.R = in.R;
    G.u = du;
    const int bx = (int)((nt + 255) / 256);
    const int chunk = (int)std::max<int64_t>(16, std::min<int64_t>(256, (ns + 63) / 64));
    G.R.trial_chunk = chunk;
    dim3 grid(bx, (unsigned)((ns + chunk - 1) / chunk));
    if (a.kind == kLaplace) launch_apply_op<false, false>(a.op, in.Q, grid, s, G, dy);
    else if (in.centred) launch_apply_op<true, true>(a.op, in.Q, grid, s, G, dy);
    else launch_apply_op<true, false>(a.op, in.Q, grid, s, G, dy);
    HIP_CHECK_G(hipMemcpyAsync(hy.data(), dy, hy.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK_G(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(du);
    (void)hipFree(dy);
    throw;
  }
  (void)hipFree(du);
  (void)hipFree(dy);
  std::vector<int64_t> rows, cols;
  std::vector<double> vals;
  assemble_singular_coo(a, trial, test, rows, cols, vals);
  const bool cv = is_complex(a.dtype);
  for (size_t e = 0; e < rows.size(); ++e) {
    const int64_t r = transpose ? cols[e] : rows[e], c = transpose ? rows[e] : cols[e];
    const double vr = cv ? vals[2 * e] : vals[e], vi = cv ? vals[2 * e + 1] : 0.0;
    const double ur = hu[2 * c], ui = hu[2 * c + 1];
    hy[2 * r] += vr * ur - vi * ui;
    hy[2 * r + 1] += vr * ui + vi * ur;
  }
  for (int64_t i = 0; i < nt; ++i) {
    switch (a.dtype) {
      case kF64: ((double*)y)[i] = hy[2 * i]; break;
      case kF32: ((float*)y)[i] = (float)hy[2 * i]; break;
      case kC64:
        ((double*)y)[2 * i] = hy[2 * i];
        ((double*)y)[2 * i + 1] = hy[2 * i + 1];
        break;
      default:
        ((float*)y)[2 * i] = (float)hy[2 * i];
        ((float*)y)[2 * i + 1] = (float)hy[2 * i + 1];
        break;
    }
  }
}

namespace {

void check_shape_gradient(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test) {
  if (a.op != kSL) throw Error("shape gradient is implemented for the single-layer operator");
  if (test.grid != trial.grid) throw Error("shape gradient needs test and trial spaces on one grid");
  if (test.nb != 1 || trial.nb != 1) throw Error("shape gradient is implemented for P0 spaces");
}

// Enqueues the regular-pair and adjacent-pair gradient kernels on s. G.u, G.v (complex128,
// nc entries each) and G.gcell (nc x 18 doubles, zeroed) are set by the caller; the
// geometry, rule and pair fields come from `in`.
void enqueue_gradient_kernels(const Assembler& a, const GradientInputs& in, int64_t nc, hipStream_t s,
                              dev::GradArgs& G) {
  G.R = in.R;
  G.pairs = in.pairs;
  G.npairs = in.npairs;
  G.rule = in.rule;
  G.rule_total = in.rule_total;
  for (int mm = 0; mm < 4; ++mm) {
    G.rule_off[mm] = in.rule_off[mm];
    G.rule_n[mm] = in.rule_n[mm];
  }
  std::vector<double> xi, w;
  triangle_rule(in.Q, xi, w);
  for (int p = 0; p < dev::kMaxQ; ++p)
    for (int k = 0; k < 3; ++k) G.lam[p][k] = 0.0;
  for (int p = 0; p < std::min(in.Q, dev::kMaxQ); ++p) {
    G.lam[p][0] = 1.0 - xi[2 * p] - xi[2 * p + 1];
    G.lam[p][1] = xi[2 * p];
    G.lam[p][2] = xi[2 * p + 1];
  }
  const int bx = (int)((nc + 255) / 256);
  const int chunk = (int)std::max<int64_t>(16, std::min<int64_t>(256, (nc + 63) / 64));
  G.R.trial_chunk = chunk;
  launch_grad(in.Q, a.kind == kHelmholtz, in.centred, dim3(bx, (unsigned)((nc + chunk - 1) / chunk)), s, G);
  if (in.npairs > 0) {
    dim3 sgrid((unsigned)((in.npairs + 3) / 4));
    if (a.kind == kHelmholtz) dev::grad_singular_kernel<true><<<sgrid, 256, 0, s>>>(G);
    else dev::grad_singular_kernel<false><<<sgrid, 256, 0, s>>>(G);
    HIP_CHECK_G(hipGetLastError());
  }
}

}  // namespace

void shape_gradient(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, const void* u,
                    const void* v, void* grad) {
  check_shape_gradient(a, trial, test);
  GradientInputs in = gradient_inputs(a, trial, test);
  hipStream_t s = (hipStream_t)in.stream;
  const Grid& g = *test.grid;
  const int64_t nc = g.nc(), nv = g.nv();
  const bool cplx_in = is_complex(a.dtype);
  // u, v in the assembler dtype -> complex128 on the device
  std::vector<double> hu = to_c128(u, a.dtype, nc), hv = to_c128(v, a.dtype, nc);
  double *du = nullptr, *dv = nullptr, *dg = nullptr;
  const size_t gbytes = (size_t)nc * 18 * sizeof(double);
  std::vector<double> hg(nc * 18);
  auto cleanup = [&]() {
    (void)hipFree(du);
    (void)hipFree(dv);
    (void)hipFree(dg);
  };
  try {
    HIP_CHECK_G(hipMalloc((void**)&du, hu.size() * sizeof(double)));
    HIP_CHECK_G(hipMalloc((void**)&dv, hv.size() * sizeof(double)));
    HIP_CHECK_G(hipMalloc((void**)&dg, gbytes));
    HIP_CHECK_G(hipMemcpyAsync(du, hu.data(), hu.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK_G(hipMemcpyAsync(dv, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice, s));
    dev::fill_zero_kernel<<<512, 256, 0, s>>>(dg, (long long)nc * 18);
    HIP_CHECK_G(hipGetLastError());
    dev::GradArgs G{};
    G.u = du;
    G.v = dv;
    G.gcell = dg;
    enqueue_gradient_kernels(a, in, nc, s, G);
    HIP_CHECK_G(hipMemcpyAsync(hg.data(), dg, gbytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK_G(hipStreamSynchronize(s));
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
  // cell-vertex gradients -> vertex gradients
  std::vector<double> gv(nv * 6, 0.0);
  for (int64_t c = 0; c < nc; ++c)
    for (int k = 0; k < 3; ++k) {
      const int64_t vtx = g.cells[3 * c + k];
      for (int d = 0; d < 3; ++d) {
        gv[(vtx * 3 + d) * 2] += hg[((c * 3 + k) * 3 + d) * 2];
        gv[(vtx * 3 + d) * 2 + 1] += hg[((c * 3 + k) * 3 + d) * 2 + 1];
      }
    }
  if (cplx_in) {
    std::copy(gv.begin(), gv.end(), (double*)grad);
  } else {
    for (int64_t i = 0; i < nv * 3; ++i) ((double*)grad)[i] = gv[2 * i];
  }
}


// ---------------------------------------------------------------------------
// Shape gradient that allocates no device memory. Its partial results (u and v
// as complex128, the per-cell-vertex gradients) live in `scratch`, one static
// copy per GPU for the life of the process: [u: 2 nc][v: 2 nc][gcell: 18 nc] doubles.
// Every call on a GPU reuses that GPU's copy and clears what must start at zero;
// calls that overlap in time on one GPU race on it.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace dev {

// x in the assembler dtype -> interleaved complex128
__global__ void to_c128_kernel(const void* __restrict__ x, int dtype, long long n, double* __restrict__ out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double re, im = 0.0;
    switch (dtype) {
      case kF64: re = ((const double*)x)[i]; break;
      case kF32: re = ((const float*)x)[i]; break;
      case kC64:
        re = ((const double*)x)[2 * i];
        im = ((const double*)x)[2 * i + 1];
        break;
      default:
        re = ((const float*)x)[2 * i];
        im = ((const float*)x)[2 * i + 1];
        break;
    }
    out[2 * i] = re;
    out[2 * i + 1] = im;
  }
}

// Cell-vertex gradients [nc][3 vertices][3 comps] (complex) -> vertex gradients
// grad[vertex][comp] (complex128, or float64 real parts), added with atomics into a
// zeroed output.
__global__ void scatter_vertex_gradient_kernel(const double* __restrict__ gcell, const int* __restrict__ vid,
                                               long long nc, double* __restrict__ grad, int complex_out) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < 9 * nc;
       t += (long long)gridDim.x * blockDim.x) {
    const long long c = t / 9;
    const int kd = (int)(t - 9 * c);
    const long long e = (long long)vid[4 * c + kd / 3] * 3 + kd % 3;
    if (complex_out) {
      unsafeAtomicAdd(grad + 2 * e, gcell[2 * t]);
      unsafeAtomicAdd(grad + 2 * e + 1, gcell[2 * t + 1]);
    } else {
      unsafeAtomicAdd(grad + e, gcell[2 * t]);
    }
  }
}

}  // namespace dev

namespace {

// Device address of `scratch` on `device` (looked up once per device, with it current).
double* noalloc_scratch(int device) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* ptrs[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) throw Error("device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!ptrs[device]) {
    void* p = nullptr;
    HIP_CHECK_G(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error("hipGetSymbolAddress returned a null scratch address");
    ptrs[device] = p;
  }
  return (double*)ptrs[device];
}

}  // namespace

void shape_gradient_noalloc(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, const void* u,
                            const void* v, void* grad, void* stream) {
  check_shape_gradient(a, trial, test);
  const Grid& g = *test.grid;
  const int64_t nc = g.nc(), nv = g.nv();
  if (nc > kNoAllocMaxCells)
    throw Error("shape gradient without allocation: " + std::to_string(nc) + " cells exceed its static scratch buffer (" +
                std::to_string(kNoAllocScratchBytes) + " bytes per GPU, at most " + std::to_string(kNoAllocMaxCells) +
                " cells)");
  if (device_count() <= 0)
    throw Error("no HIP device available: assembly requires an MI355X GPU (there is no CPU fallback)");
  int device = 0;
  HIP_CHECK_G(hipGetDevice(&device));
  double* sp = noalloc_scratch(device);
  GradientInputs in;
  {
    NoDeviceAllocScope guard;  // the caches must exist already: nothing is uploaded here
    in = gradient_inputs(a, trial, test, stream, true);
  }
  hipStream_t s = (hipStream_t)stream;
  double* du = sp;
  double* dv = sp + 2 * nc;
  double* dg = sp + 4 * nc;
  const bool cplx_out = is_complex(a.dtype);
  // fill kernels, not memset nodes: a caller may capture the stream in a graph
  dev::to_c128_kernel<<<256, 256, 0, s>>>(u, (int)a.dtype, nc, du);
  dev::to_c128_kernel<<<256, 256, 0, s>>>(v, (int)a.dtype, nc, dv);
  dev::fill_zero_kernel<<<512, 256, 0, s>>>(dg, (long long)nc * 18);
  dev::fill_zero_kernel<<<256, 256, 0, s>>>((double*)grad, (long long)nv * 3 * (cplx_out ? 2 : 1));
  HIP_CHECK_G(hipGetLastError());
  dev::GradArgs G{};
  G.u = du;
  G.v = dv;
  G.gcell = dg;
  enqueue_gradient_kernels(a, in, nc, s, G);
  dev::scatter_vertex_gradient_kernel<<<512, 256, 0, s>>>(dg, in.R.test_vid, nc, (double*)grad, cplx_out ? 1 : 0);
  HIP_CHECK_G(hipGetLastError());
}

}  // namespace bempp_amd
