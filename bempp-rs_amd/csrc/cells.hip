// cells.hip -- dense assembly, and the sparse singular part and correction, on grids with
// quadrilateral cells (single-element quadrilateral grids and mixed triangle/quadrilateral
// grids of degree-1 geometry), FP64.
//
// The reference's drivers take any cell type (src/boundary_assemblers.rs:396-521 loops over
// the grid's entity_types(2), :524-562 picks a Duffy rule per (test type, trial type)).
// Quadrilateral geometry is bilinear, x(xi, eta) = V0 + xi (V1 - V0) + eta (V2 - V0)
// + xi eta (V3 - V1 - V2 + V0), so the Jacobian and normal vary over a cell; P1 on a
// quadrilateral is Q1 ((1-xi)(1-eta), xi(1-eta), (1-xi)eta, xi eta); the regular rule is
// m x m Gauss (ndelement's quadrilateral rule). Adjacent pairs use the pair rules of
// cell_pair_rule (host.cpp): each quadrilateral split into two reference triangles so that
// every half pair shares the pair's common vertex, each half pair taking the triangle
// Sauter-Schwab rule through the composite map.
//
// Regular pairs: one thread per test cell, trial cells wave-uniform (scalar loads); each
// cell's points are precomputed records (point, w J, normal, w, basis values, J-scaled
// surface curls t_eta d_xi(phi) - t_xi d_eta(phi)). Adjacent pairs: one wavefront per pair,
// the geometry evaluated at each rule point, wavefront reductions. Both add into the output
// with atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "core.hpp"
#include "kernels.hpp"

namespace bempp_amd {

#define HIP_CHECK_Q(expr)                                                                         \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) throw Error(std::string(#expr) + " failed: " + hipGetErrorString(_e)); \
  } while (0)

namespace dev {

constexpr int kRec = 24;  // doubles per point record

struct CellsArgs {
  const double* trec;  // test point records [points][kRec]
  const double* srec;
  const int* tfirst;   // first record of each test cell [n_test + 1]
  const int* sfirst;
  const int* tvid;     // vertex ids [cells][4] (-1: none)
  const int* svid;
  const int* tdofs;    // DOFs [cells][4] (-1: none)
  const int* sdofs;
  const double* tcv;   // vertex coordinates [cells][4][3]
  const double* scv;
  int n_test, n_trial, same_grid, trial_chunk;
  double k;
  void* out;
  long long ld;
  int out_type;
  const double* rule;  // adjacent-pair rules, records (test xi, eta, trial xi, eta, weight)
  const int4* pairs;   // test cell, trial cell, first rule record, record count
  int npairs;
  int correction;      // adjacent pairs under the cells' regular rules instead of their pair rules
  double* coo;         // non-null: the adjacent pairs' values in COO layout instead of `out`
};

// Point, normal, Jacobian, basis values and J-scaled curls of a cell at (x, y); V: the
// vertex coordinates (three for a triangle, four for a quadrilateral).
template <int NB>
__device__ __forceinline__ void cell_point(const double* V, bool quad, double x, double y, double* X, double* n,
                                           double& J, double* phi, double (*curl)[3]) {
  double tx[3], ty[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double e1 = V[3 + d] - V[d], e2 = V[6 + d] - V[d];
    if (quad) {
      const double tw = V[9 + d] - V[3 + d] - V[6 + d] + V[d];
      tx[d] = fma(y, tw, e1);
      ty[d] = fma(x, tw, e2);
      X[d] = V[d] + x * e1 + y * e2 + x * y * tw;
    } else {
      tx[d] = e1;
      ty[d] = e2;
      X[d] = V[d] + x * e1 + y * e2;
    }
  }
  const double c[3] = {tx[1] * ty[2] - tx[2] * ty[1], tx[2] * ty[0] - tx[0] * ty[2], tx[0] * ty[1] - tx[1] * ty[0]};
  J = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  const double ji = 1.0 / J;
  n[0] = c[0] * ji, n[1] = c[1] * ji, n[2] = c[2] * ji;
  if constexpr (NB == 1) {
    phi[0] = 1.0;
    curl[0][0] = curl[0][1] = curl[0][2] = 0.0;
  } else {
    double dx[4], dy[4];
    if (quad) {
      phi[0] = (1 - x) * (1 - y), phi[1] = x * (1 - y), phi[2] = (1 - x) * y, phi[3] = x * y;
      dx[0] = -(1 - y), dx[1] = 1 - y, dx[2] = -y, dx[3] = y;
      dy[0] = -(1 - x), dy[1] = -x, dy[2] = 1 - x, dy[3] = x;
    } else {
      phi[0] = 1 - x - y, phi[1] = x, phi[2] = y, phi[3] = 0.0;
      dx[0] = -1, dx[1] = 1, dx[2] = 0, dx[3] = 0;
      dy[0] = -1, dy[1] = 0, dy[2] = 1, dy[3] = 0;
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int d = 0; d < 3; ++d) curl[b][d] = ty[d] * dx[b] - tx[d] * dy[b];
  }
}

// adds the integrand of one point pair (without 1/(4 pi)) to M:
// wpq = (w_p J_p)(w_q J_q), ww = w_p w_q (the curl-curl term's weight, its Jacobians cancel)
// TIER > 0: the phase relative to kR (the caller applies e^{ikR} once per cell pair).
template <int OP, bool HELM, int NBT, int NBS, int TIER = 0, typename PP, typename PC, typename QN, typename QP,
          typename QC>
__device__ __forceinline__ void add_point_pair(const double* xp, const double* np, const PP& php, const PC& cp,
                                               const double* xq, const QN& nq, const QP& phq, const QC& cq,
                                               double wpq, double ww, double k, cplx (&M)[NBT][NBS],
                                               double kR = 0.0) {
  const double dx = xp[0] - xq[0], dy = xp[1] - xq[1], dz = xp[2] - xq[2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  double dn = 0.0;
  if constexpr (OP == 1) dn = dx * nq[0] + dy * nq[1] + dz * nq[2];
  if constexpr (OP == 2) dn = -(dx * np[0] + dy * np[1] + dz * np[2]);
  const cplx g = kernel_scalar_r2<OP, HELM, TIER>(r2, dn, k, kR);
  if constexpr (OP != 3) {
#pragma unroll
    for (int a = 0; a < NBT; ++a)
#pragma unroll
      for (int b = 0; b < NBS; ++b) {
        const double f = wpq * php[a] * phq[b];
        M[a][b].re = fma(g.re, f, M[a][b].re);
        if constexpr (HELM) M[a][b].im = fma(g.im, f, M[a][b].im);
      }
  } else {
    const double nn = HELM ? -k * k * (np[0] * nq[0] + np[1] * nq[1] + np[2] * nq[2]) * wpq : 0.0;
#pragma unroll
    for (int a = 0; a < NBT; ++a)
#pragma unroll
      for (int b = 0; b < NBS; ++b) {
        double f = ww * (cp[a][0] * cq[b][0] + cp[a][1] * cq[b][1] + cp[a][2] * cq[b][2]);
        if constexpr (HELM) f = fma(nn, php[a] * phq[b], f);
        M[a][b].re = fma(g.re, f, M[a][b].re);
        if constexpr (HELM) M[a][b].im = fma(g.im, f, M[a][b].im);
      }
  }
}

// centroid of a cell from its vertex coordinates (the bilinear centre x(1/2, 1/2) of a
// quadrilateral is its vertex mean too)
template <typename V>
__device__ __forceinline__ void cell_centroid(const V& cv, bool quad, double* c) {
#pragma unroll
  for (int d = 0; d < 3; ++d)
    c[d] = quad ? 0.25 * (cv[d] + cv[3 + d] + cv[6 + d] + cv[9 + d]) : (cv[d] + cv[3 + d] + cv[6 + d]) * (1.0 / 3.0);
}

// TIER (Helmholtz): 0 full argument reduction per point pair; 1, 2 the phase relative to the
// centroid distance R, |k (r - R)| <= 0.5 / 0.25 over the whole grid pair.
template <int OP, bool HELM, int NBT, int NBS, int TIER>
__global__ void __launch_bounds__(256) cells_regular_kernel(const CellsArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool i_ok = i < A.n_test;
  const int il = i_ok ? i : A.n_test - 1;
  int tv[4], td[NBT];
#pragma unroll
  for (int m = 0; m < 4; ++m) tv[m] = A.tvid[4 * il + m];
  double ct[3];
  if constexpr (TIER > 0) cell_centroid(A.tcv + 12 * (size_t)il, tv[3] >= 0, ct);
  const auto scv = constant_view(A.scv);
#pragma unroll
  for (int a = 0; a < NBT; ++a) td[a] = A.tdofs[4 * il + a];
  const int p0 = A.tfirst[il], p1 = A.tfirst[il + 1];
  const auto svid = constant_view(A.svid);
  const auto sdofs = constant_view(A.sdofs);
  const auto sfirst = constant_view(A.sfirst);
  const auto srec = constant_view(A.srec);
  const int j0 = blockIdx.y * A.trial_chunk;
  const int j1 = min(j0 + A.trial_chunk, A.n_trial);
  for (int j = j0; j < j1; ++j) {
    bool adjacent = false;
    if (A.same_grid)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int sv = svid[4 * j + b];
#pragma unroll
        for (int a = 0; a < 4; ++a) adjacent = adjacent || (sv >= 0 && sv == tv[a]);
      }
    if (adjacent || !i_ok) continue;
    cplx M[NBT][NBS];
#pragma unroll
    for (int a = 0; a < NBT; ++a)
#pragma unroll
      for (int b = 0; b < NBS; ++b) M[a][b] = {0.0, 0.0};
    double kR = 0.0;
    cplx e = {1.0, 0.0};
    if constexpr (TIER > 0) {
      double cs[3];
      cell_centroid(scv + 12 * (size_t)j, svid[4 * j + 3] >= 0, cs);
      e = pair_phase(ct, cs, A.k, &kR);
    }
    const int q0 = sfirst[j], q1 = sfirst[j + 1];
    for (int p = p0; p < p1; ++p) {
      const double* P = A.trec + (size_t)p * kRec;
      const double xp[3] = {P[0], P[1], P[2]}, np[3] = {P[4], P[5], P[6]};
      const double wJp = P[3], wp = P[7];
      double php[NBT], cp[NBT][3];
#pragma unroll
      for (int a = 0; a < NBT; ++a) {
        php[a] = P[8 + a];
        if constexpr (OP == 3)
#pragma unroll
          for (int d = 0; d < 3; ++d) cp[a][d] = P[12 + 3 * a + d];
      }
      for (int q = q0; q < q1; ++q) {
        const auto Q = srec + (size_t)q * kRec;
        const double xq[3] = {Q[0], Q[1], Q[2]}, nq[3] = {Q[4], Q[5], Q[6]};
        double phq[NBS], cq[NBS][3];
#pragma unroll
        for (int b = 0; b < NBS; ++b) {
          phq[b] = Q[8 + b];
          if constexpr (OP == 3)
#pragma unroll
            for (int d = 0; d < 3; ++d) cq[b][d] = Q[12 + 3 * b + d];
        }
        add_point_pair<OP, HELM, NBT, NBS, TIER>(xp, np, php, cp, xq, nq, phq, cq, wJp * Q[3], wp * Q[7], A.k, M, kR);
      }
    }
    if constexpr (TIER > 0)
#pragma unroll
      for (int a = 0; a < NBT; ++a)
#pragma unroll
        for (int b = 0; b < NBS; ++b) M[a][b] = cmul(M[a][b], e);
#pragma unroll
    for (int b = 0; b < NBS; ++b) {
      const int col = sdofs[4 * j + b];
      if (col < 0) continue;  // uniform: the trial cell's basis count
#pragma unroll
      for (int a = 0; a < NBT; ++a)
        if (td[a] >= 0)
          store_value(A.out, (long long)td[a] + A.ld * col, {M[a][b].re * kInv4Pi, M[a][b].im * kInv4Pi}, A.out_type,
                      kAtomic);
    }
  }
}

template <int OP, bool HELM, int NBT, int NBS>
__global__ void __launch_bounds__(256) cells_singular_kernel(const CellsArgs A) {
  const int pair = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6));
  if (pair >= A.npairs) return;  // uniform per wavefront
  const int lane = threadIdx.x & 63;
  const int4 pr = A.pairs[pair];
  const int tc = pr.x, sc = pr.y;
  double tV[12], sV[12];
#pragma unroll
  for (int m = 0; m < 12; ++m) {
    tV[m] = A.tcv[12 * (size_t)tc + m];
    sV[m] = A.scv[12 * (size_t)sc + m];
  }
  const bool tq = A.tvid[4 * tc + 3] >= 0, sq = A.svid[4 * sc + 3] >= 0;
  cplx M[NBT][NBS];
#pragma unroll
  for (int a = 0; a < NBT; ++a)
#pragma unroll
    for (int b = 0; b < NBS; ++b) M[a][b] = {0.0, 0.0};
  if (A.correction) {
    // the regular rule of each cell: every (test point, trial point) pair of their records
    const int p0 = A.tfirst[tc], q0 = A.sfirst[sc], nqs = A.sfirst[sc + 1] - q0;
    const int n = (A.tfirst[tc + 1] - p0) * nqs;
    for (int i = lane; i < n; i += 64) {
      const int ip = i / nqs;
      const double* P = A.trec + (size_t)(p0 + ip) * kRec;
      const double* Q = A.srec + (size_t)(q0 + i - ip * nqs) * kRec;
      const double xp[3] = {P[0], P[1], P[2]}, np[3] = {P[4], P[5], P[6]};
      const double xq[3] = {Q[0], Q[1], Q[2]}, nq[3] = {Q[4], Q[5], Q[6]};
      // coincident points (a cell with itself) are left out, as a point-to-point fast
      // multipole evaluation does
      if (xp[0] == xq[0] && xp[1] == xq[1] && xp[2] == xq[2]) continue;
      double php[NBT], cp[NBT][3], phq[NBS], cq[NBS][3];
#pragma unroll
      for (int a = 0; a < NBT; ++a) {
        php[a] = P[8 + a];
#pragma unroll
        for (int d = 0; d < 3; ++d) cp[a][d] = P[12 + 3 * a + d];
      }
#pragma unroll
      for (int b = 0; b < NBS; ++b) {
        phq[b] = Q[8 + b];
#pragma unroll
        for (int d = 0; d < 3; ++d) cq[b][d] = Q[12 + 3 * b + d];
      }
      add_point_pair<OP, HELM, NBT, NBS>(xp, np, php, cp, xq, nq, phq, cq, P[3] * Q[3], P[7] * Q[7], A.k, M);
    }
  } else {
    for (int i = lane; i < pr.w; i += 64) {
      const double* r = A.rule + 5 * (size_t)(pr.z + i);
      double xp[3], np[3], Jp, php[4], cp[4][3], xq[3], nq[3], Jq, phq[4], cq[4][3];
      cell_point<NBT>(tV, tq, r[0], r[1], xp, np, Jp, php, cp);
      cell_point<NBS>(sV, sq, r[2], r[3], xq, nq, Jq, phq, cq);
      add_point_pair<OP, HELM, NBT, NBS>(xp, np, php, cp, xq, nq, phq, cq, r[4] * Jp * Jq, r[4], A.k, M);
    }
  }
#pragma unroll
  for (int a = 0; a < NBT; ++a)
#pragma unroll
    for (int b = 0; b < NBS; ++b) {
      M[a][b].re = wave_sum(M[a][b].re);
      if constexpr (HELM) M[a][b].im = wave_sum(M[a][b].im);
    }
  if (lane != 0) return;
  if (A.coo) {
    // position pair * NBT * NBS + a + NBT * b, (re, im); entries of basis functions a cell
    // does not have (a triangle's fourth) are written as zeros and dropped by their DOF -1
    double* o = A.coo + 2 * (size_t)pair * NBT * NBS;
#pragma unroll
    for (int b = 0; b < NBS; ++b)
#pragma unroll
      for (int a = 0; a < NBT; ++a) {
        o[2 * (a + NBT * b)] = M[a][b].re * kInv4Pi;
        o[2 * (a + NBT * b) + 1] = HELM ? M[a][b].im * kInv4Pi : 0.0;
      }
    return;
  }
#pragma unroll
  for (int b = 0; b < NBS; ++b) {
    const int col = A.sdofs[4 * sc + b];
    if (col < 0) continue;
#pragma unroll
    for (int a = 0; a < NBT; ++a) {
      const int row = A.tdofs[4 * tc + a];
      if (row >= 0)
        store_value(A.out, (long long)row + A.ld * col, {M[a][b].re * kInv4Pi, M[a][b].im * kInv4Pi}, A.out_type,
                    kAtomic);
    }
  }
}

}  // namespace dev

namespace {

int cells_out_type(DType d) {
  switch (d) {
    case kF64: return dev::kOutF64;
    case kC64: return dev::kOutC128;
    case kF32: return dev::kOutF32;
    default: return dev::kOutC64;
  }
}

struct DeviceBuf {
  void* p = nullptr;
  ~DeviceBuf() {
    if (p) (void)hipFree(p);
  }
};

template <typename T>
T* upload_to(std::vector<DeviceBuf>& keep, const std::vector<T>& h, hipStream_t s) {
  keep.emplace_back();
  HIP_CHECK_Q(hipMalloc(&keep.back().p, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) HIP_CHECK_Q(hipMemcpyAsync(keep.back().p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return (T*)keep.back().p;
}

// Point records of every cell of a space's grid for its regular rule (dev::kRec doubles each).
void point_records(const FunctionSpace& sp, const Options& o, std::vector<double>& rec, std::vector<int>& first) {
  const Grid& g = *sp.grid;
  std::vector<double> txi, tw, qxi, qw;
  triangle_rule((int)o.quadrature_degrees.at(kTriangle), txi, tw);
  bool any_quad = false;
  for (int64_t c = 0; c < g.nc(); ++c) any_quad = any_quad || g.cell_nv(c) == 4;
  if (any_quad) quadrilateral_rule((int)o.quadrature_degrees.at(kQuadrilateral), qxi, qw);
  rec.clear();
  first.assign(g.nc() + 1, 0);
  for (int64_t c = 0; c < g.nc(); ++c) {
    const bool quad = g.cell_nv(c) == 4;
    const std::vector<double>& xi = quad ? qxi : txi;
    const std::vector<double>& w = quad ? qw : tw;
    double V[4][3];
    for (int m = 0; m < g.cell_nv(c); ++m)
      for (int d = 0; d < 3; ++d) V[m][d] = g.points[3 * g.vertex(c, m) + d];
    for (size_t p = 0; p < w.size(); ++p) {
      const double x = xi[2 * p], y = xi[2 * p + 1];
      double tx[3], ty[3], X[3];
      for (int d = 0; d < 3; ++d) {
        const double e1 = V[1][d] - V[0][d], e2 = V[2][d] - V[0][d];
        const double tw4 = quad ? V[3][d] - V[1][d] - V[2][d] + V[0][d] : 0.0;
        tx[d] = e1 + y * tw4;
        ty[d] = e2 + x * tw4;
        X[d] = V[0][d] + x * e1 + y * e2 + x * y * tw4;
      }
      const double cr[3] = {tx[1] * ty[2] - tx[2] * ty[1], tx[2] * ty[0] - tx[0] * ty[2], tx[0] * ty[1] - tx[1] * ty[0]};
      const double J = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
      double phi[4] = {1, 0, 0, 0}, dx[4] = {0, 0, 0, 0}, dy[4] = {0, 0, 0, 0};
      if (sp.degree == 1 && quad) {
        phi[0] = (1 - x) * (1 - y), phi[1] = x * (1 - y), phi[2] = (1 - x) * y, phi[3] = x * y;
        dx[0] = -(1 - y), dx[1] = 1 - y, dx[2] = -y, dx[3] = y;
        dy[0] = -(1 - x), dy[1] = -x, dy[2] = 1 - x, dy[3] = x;
      } else if (sp.degree == 1) {
        phi[0] = 1 - x - y, phi[1] = x, phi[2] = y;
        dx[0] = -1, dx[1] = 1;
        dy[0] = -1, dy[2] = 1;
      }
      double r[dev::kRec] = {X[0], X[1], X[2], w[p] * J, cr[0] / J, cr[1] / J, cr[2] / J, w[p]};
      for (int b = 0; b < 4; ++b) {
        r[8 + b] = phi[b];
        for (int d = 0; d < 3; ++d) r[12 + 3 * b + d] = ty[d] * dx[b] - tx[d] * dy[b];
      }
      rec.insert(rec.end(), r, r + dev::kRec);
    }
    first[c + 1] = first[c] + (int)w.size();
  }
}

void cell_tables(const FunctionSpace& sp, std::vector<int>& vid, std::vector<int>& dofs, std::vector<double>& cv) {
  const Grid& g = *sp.grid;
  vid.assign(4 * g.nc(), -1);
  dofs.assign(4 * g.nc(), -1);
  cv.assign(12 * g.nc(), 0.0);
  for (int64_t c = 0; c < g.nc(); ++c) {
    for (int m = 0; m < g.cell_nv(c); ++m) {
      vid[4 * c + m] = (int)g.vertex(c, m);
      for (int d = 0; d < 3; ++d) cv[12 * c + 3 * m + d] = g.points[3 * g.vertex(c, m) + d];
    }
    for (int b = 0; b < sp.cell_nb(c); ++b) dofs[4 * c + b] = (int)sp.dofs[(size_t)sp.nb * c + b];
  }
}

size_t singular_points(const Options& o, uint8_t t, uint8_t s) {
  auto it = o.singular_quadrature_degrees.find({t, s});
  if (it == o.singular_quadrature_degrees.end()) throw Error("no singular quadrature degree set for this cell pair type");
  return it->second;
}

// Adjacent ordered cell pairs, each once (at its smallest shared vertex, as
// src/boundary_assemblers.rs:775-792 classifies them), with their pair rules.
void adjacent_cell_pairs(const Grid& g, const Options& o, std::vector<int4>& pairs, std::vector<double>& rule) {
  pairs.clear();
  rule.clear();
  std::map<std::tuple<int, int, int, std::vector<std::pair<int, int>>>, std::pair<int, int>> configs;
  std::vector<double> pts5;
  for (int64_t v = 0; v < g.nv(); ++v)
    for (int64_t ia = g.vc_offsets[v]; ia < g.vc_offsets[v + 1]; ++ia) {
      const int64_t tc = g.vc_cells[ia];
      for (int64_t ib = g.vc_offsets[v]; ib < g.vc_offsets[v + 1]; ++ib) {
        const int64_t sc = g.vc_cells[ib];
        std::vector<std::pair<int, int>> shared;
        bool smallest = true;
        for (int si = 0; si < g.cell_nv(sc) && smallest; ++si)
          for (int ti = 0; ti < g.cell_nv(tc); ++ti)
            if (g.vertex(tc, ti) == g.vertex(sc, si)) {
              if (g.vertex(tc, ti) < v) smallest = false;
              shared.push_back({ti, si});
            }
        if (!smallest || shared.empty()) continue;
        const uint8_t tt = g.cell_nv(tc) == 4 ? kQuadrilateral : kTriangle;
        const uint8_t st = g.cell_nv(sc) == 4 ? kQuadrilateral : kTriangle;
        const int n = (int)singular_points(o, tt, st);
        auto key = std::make_tuple((int)tt, (int)st, n, tc == sc ? std::vector<std::pair<int, int>>{} : shared);
        auto it = configs.find(key);
        if (it == configs.end()) {
          cell_pair_rule(g, tc, sc, n, pts5);
          it = configs.emplace(key, std::make_pair((int)(rule.size() / 5), (int)(pts5.size() / 5))).first;
          rule.insert(rule.end(), pts5.begin(), pts5.end());
        }
        pairs.push_back(make_int4((int)tc, (int)sc, it->second.first, it->second.second));
      }
    }
}

// tier < 0: the adjacent pairs only (no regular kernel)
template <int OP, bool HELM, int NBT, int NBS>
void launch_cells(int tier, dim3 rg, dim3 sg, hipStream_t s, const dev::CellsArgs& A) {
  if (HELM && tier == 2) dev::cells_regular_kernel<OP, HELM, NBT, NBS, 2><<<rg, 256, 0, s>>>(A);
  else if (HELM && tier == 1) dev::cells_regular_kernel<OP, HELM, NBT, NBS, 1><<<rg, 256, 0, s>>>(A);
  else if (tier >= 0) dev::cells_regular_kernel<OP, HELM, NBT, NBS, 0><<<rg, 256, 0, s>>>(A);
  if (A.npairs > 0) dev::cells_singular_kernel<OP, HELM, NBT, NBS><<<sg, 256, 0, s>>>(A);
}

template <int OP, bool HELM>
void launch_cells_nb(int nbt, int nbs, int tier, dim3 rg, dim3 sg, hipStream_t s, const dev::CellsArgs& A) {
  if (nbt == 1 && nbs == 1) launch_cells<OP, HELM, 1, 1>(tier, rg, sg, s, A);
  else if (nbt == 1) launch_cells<OP, HELM, 1, 4>(tier, rg, sg, s, A);
  else if (nbs == 1) launch_cells<OP, HELM, 4, 1>(tier, rg, sg, s, A);
  else launch_cells<OP, HELM, 4, 4>(tier, rg, sg, s, A);
}

template <bool HELM>
void launch_cells_op(int op, int nbt, int nbs, int tier, dim3 rg, dim3 sg, hipStream_t s, const dev::CellsArgs& A) {
  switch (op) {
    case kSL: launch_cells_nb<0, HELM>(nbt, nbs, tier, rg, sg, s, A); break;
    case kDL: launch_cells_nb<1, HELM>(nbt, nbs, tier, rg, sg, s, A); break;
    case kADL: launch_cells_nb<2, HELM>(nbt, nbs, tier, rg, sg, s, A); break;
    case kHS: launch_cells_nb<3, HELM>(nbt, nbs, tier, rg, sg, s, A); break;
    default: throw Error("operator not supported");
  }
}

// largest centroid-to-vertex distance over a grid's cells: every point of a cell lies within
// it of the centroid (a bilinear cell lies in the convex hull of its vertices)
double max_cell_radius(const Grid& g) {
  double rmax = 0.0;
  for (int64_t c = 0; c < g.nc(); ++c) {
    const int n = g.cell_nv(c);
    double cen[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m < n; ++m)
      for (int d = 0; d < 3; ++d) cen[d] += g.points[3 * g.vertex(c, m) + d] / n;
    for (int m = 0; m < n; ++m) {
      double r2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double x = g.points[3 * g.vertex(c, m) + d] - cen[d];
        r2 += x * x;
      }
      rmax = std::max(rmax, std::sqrt(r2));
    }
  }
  return rmax;
}

// The cells' point records, vertex and DOF tables and (same grid) adjacent pairs with their
// rules, uploaded into `keep`; `A` points at them.
void upload_cells(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, hipStream_t s,
                  std::vector<DeviceBuf>& keep, dev::CellsArgs& A) {
  std::vector<double> trec, srec, tcv, scv, rule;
  std::vector<int> tfirst, sfirst, tvid, svid, tdofs, sdofs;
  std::vector<int4> pairs;
  point_records(test, a.options, trec, tfirst);
  point_records(trial, a.options, srec, sfirst);
  cell_tables(test, tvid, tdofs, tcv);
  cell_tables(trial, svid, sdofs, scv);
  const bool same = test.grid == trial.grid;
  if (same) adjacent_cell_pairs(*test.grid, a.options, pairs, rule);
  keep.reserve(16);
  A = dev::CellsArgs{};
  A.trec = upload_to(keep, trec, s);
  A.srec = upload_to(keep, srec, s);
  A.tfirst = upload_to(keep, tfirst, s);
  A.sfirst = upload_to(keep, sfirst, s);
  A.tvid = upload_to(keep, tvid, s);
  A.svid = upload_to(keep, svid, s);
  A.tdofs = upload_to(keep, tdofs, s);
  A.sdofs = upload_to(keep, sdofs, s);
  A.tcv = upload_to(keep, tcv, s);
  A.scv = upload_to(keep, scv, s);
  A.rule = upload_to(keep, rule, s);
  A.pairs = upload_to(keep, pairs, s);
  A.npairs = (int)pairs.size();
  A.n_test = (int)test.grid->nc();
  A.n_trial = (int)trial.grid->nc();
  A.same_grid = same ? 1 : 0;
  A.k = a.k;
  A.ld = (long long)test.size;
  A.out_type = cells_out_type(a.dtype);
}

}  // namespace

void coo_to_csr_device(const double* vals, const int4* pairs, int npairs, const int* test_dofs, int nbt, int tstride,
                       const int* trial_dofs, int nbs, int sstride, int64_t nrows, int64_t ncols, DType dtype,
                       hipStream_t s, CsrHost& out);  // csr.hip

void assemble_cells_singular_csr(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test,
                                 bool correction, void* stream, CsrHost& out) {
  const hipStream_t s = (hipStream_t)stream;
  out.indptr.assign((size_t)test.size + 1, 0);
  out.indices.clear();
  out.data.clear();
  if (test.grid != trial.grid || test.grid->nc() == 0) return;  // no neighbouring cells across grids
  std::vector<DeviceBuf> keep;
  dev::CellsArgs A;
  upload_cells(a, trial, test, s, keep, A);
  if (A.npairs == 0) return;
  const int nbt = test.degree == 0 ? 1 : 4, nbs = trial.degree == 0 ? 1 : 4;
  keep.emplace_back();
  HIP_CHECK_Q(hipMalloc(&keep.back().p, (size_t)A.npairs * nbt * nbs * 2 * sizeof(double)));
  A.coo = (double*)keep.back().p;
  A.correction = correction ? 1 : 0;
  const dim3 sg((unsigned)((A.npairs + 3) / 4));
  if (a.kind == kHelmholtz)
    launch_cells_op<true>(a.op, nbt, nbs, -1, dim3(1), sg, s, A);
  else
    launch_cells_op<false>(a.op, nbt, nbs, -1, dim3(1), sg, s, A);
  HIP_CHECK_Q(hipGetLastError());
  coo_to_csr_device(A.coo, A.pairs, A.npairs, A.tdofs, nbt, 4, A.sdofs, nbs, 4, test.size, trial.size, a.dtype, s, out);
  // coo_to_csr_device ends with a stream synchronisation: the uploads may be freed
}

void assemble_cells_device(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, void* dev_out,
                           bool accumulate, void* stream) {
  const hipStream_t s = (hipStream_t)stream;
  const long long n_out = (long long)test.size * (long long)trial.size * (is_complex(a.dtype) ? 2 : 1);
  if (!accumulate) {
    if (a.dtype == kF32 || a.dtype == kC32)
      dev::fill_zero_kernel<<<2048, 256, 0, s>>>((float*)dev_out, n_out);
    else
      dev::fill_zero_kernel<<<2048, 256, 0, s>>>((double*)dev_out, n_out);
    HIP_CHECK_Q(hipGetLastError());
  }
  if (test.grid->nc() == 0 || trial.grid->nc() == 0) return;
  std::vector<DeviceBuf> keep;
  dev::CellsArgs A;
  upload_cells(a, trial, test, s, keep, A);
  A.out = dev_out;
  const int bx = (A.n_test + 255) / 256;
  long long chunk = ((long long)bx * A.n_trial + 16383) / 16384;
  chunk = std::max(4LL, std::min(64LL, chunk));
  chunk = std::max(chunk, ((long long)A.n_trial + 65534) / 65535);
  A.trial_chunk = (int)chunk;
  const dim3 rg(bx, (unsigned)((A.n_trial + chunk - 1) / chunk));
  const dim3 sg((unsigned)((A.npairs + 3) / 4));
  const int nbt = test.degree == 0 ? 1 : 4, nbs = trial.degree == 0 ? 1 : 4;
  // phase tier from the grid-wide bound |k (r - R)| <= |k| (rad_test + rad_trial), with a
  // margin for rounding (the tiers' polynomials hold to 0.5 and 0.25 exactly)
  int tier = 0;
  if (a.kind == kHelmholtz && a.k != 0.0) {
    const double bound = std::fabs(a.k) * (max_cell_radius(*test.grid) + max_cell_radius(*trial.grid)) * (1.0 + 1e-12);
    tier = bound <= 0.25 ? 2 : (bound <= 0.5 ? 1 : 0);
  }
  if (a.kind == kHelmholtz)
    launch_cells_op<true>(a.op, nbt, nbs, tier, rg, sg, s, A);
  else
    launch_cells_op<false>(a.op, nbt, nbs, 0, rg, sg, s, A);
  HIP_CHECK_Q(hipGetLastError());
  // the uploads above are freed when `keep` goes out of scope: wait for the kernels
  HIP_CHECK_Q(hipStreamSynchronize(s));
}

}  // namespace bempp_amd
