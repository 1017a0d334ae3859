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
// Regular pairs: one thread per test cell; each cell's points are precomputed records (point,
// w J, normal, w, basis values, J-scaled surface curls t_eta d_xi(phi) - t_xi d_eta(phi)); the
// workgroup's trial cells are staged in LDS and read back as broadcasts; symmetric operators
// on one space evaluate pairs j > i and store each block also transposed; entries with one
// writer (no DOFs shared between cells) are stored plainly, the others with atomics.
// Adjacent pairs: one wavefront per pair, the geometry evaluated at each rule point,
// wavefront reductions, atomics. The records, tables and pair rules are built on the host on
// first use and cached on the device with the grid and spaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>
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
  const double* txw;   // the same points' (x, y, z, w J) [points][4] (the fields every operator reads)
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
  int symmetric;        // single layer / hypersingular on one space: pairs j > i only, mirrored
  int store_mode;       // regular entries: kStore / kAccumulate (one writer each) or kAtomic
  int max_trial_points;  // host: the most regular points of a trial cell (picks the staged kernel)
  int adjacent_tier;     // host: phase tier of the adjacent pairs (picks the pair kernel)
  double radius_sum;     // host: largest test plus largest trial centroid-to-vertex distance
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
  const double c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  const double ji = rsqrt_fast(c2);  // 1 / J and J = c2 / J without an IEEE square root or division
  J = c2 * ji;
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
  if constexpr (OP == 0 && NBT == 1 && NBS == 1) {  // P0 single layer: the kernel times w_p J_p w_q J_q
    green_accumulate_r2<HELM, TIER>(r2, k, kR, wpq, M[0][0]);
    return;
  }
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

// records of trial cells staged in LDS per workgroup (24 KB); cells with more points take
// the unstaged kernel
constexpr int kStageRecords = 128;

// TIER (Helmholtz): 0 full argument reduction per point pair; 1, 2 the phase relative to the
// centroid distance R, |k (r - R)| <= 0.5 / 0.25 over the whole grid pair.
// STAGED: the workgroup's four wavefronts walk the same trial cells (blockIdx.y's chunk), whose
// records are contiguous: runs of whole cells (at most kStageRecords records) are copied into
// LDS once and read back as broadcasts, instead of a scalar load (and wait) per trial point.
template <int OP, bool HELM, int NBT, int NBS, int TIER, bool STAGED>
__global__ void __launch_bounds__(256) cells_regular_kernel(const CellsArgs A) {
  __shared__ double stage[STAGED ? kStageRecords * kRec : 1];
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
  const int j0 = blockIdx.y * A.trial_chunk;
  const int j1 = min(j0 + A.trial_chunk, A.n_trial);
  // symmetric: a block whose trial cells all precede its test cells has nothing to do
  // (uniform over the block, before any barrier)
  if (A.symmetric && j1 <= (int)(blockIdx.x * blockDim.x)) return;
  constexpr bool kScaledSL = OP == 0 && HELM && TIER > 0 && NBT == 1 && NBS == 1;
  // trial cell j, its nq point records from Qj on
  auto trial_cell = [&](int j, const auto* Qj, int nq) {
    if (A.symmetric && j <= i) return;  // j < i: the mirror of pair (j, i); j = i: adjacent
    bool adjacent = false;
    if (A.same_grid)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int sv = svid[4 * j + b];
#pragma unroll
        for (int a = 0; a < 4; ++a) adjacent = adjacent || (sv >= 0 && sv == tv[a]);
      }
    if (adjacent || !i_ok) return;
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
    // test points two at a time (PP = 2; one left over: PP = 1) when the local block is at most
    // 4 entries: each trial point's record is read once for both, and their two point-pair
    // chains interleave (the 4 x 4 blocks would need twice their registers)
    auto test_points = [&](auto pp, int p) {
      constexpr int PP = decltype(pp)::value;
      double xp[PP][3], np[PP][3], wJp[PP], wp[PP], php[PP][NBT], cp[PP][NBT][3];
#pragma unroll
      for (int u = 0; u < PP; ++u) {
        const double* P = A.trec + (size_t)(p + u) * kRec;
        const double* XW = A.txw + 4 * (size_t)(p + u);  // re-read per trial cell: 32 B, not a record
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          xp[u][d] = XW[d];
          np[u][d] = P[4 + d];
        }
        wJp[u] = XW[3];
        wp[u] = P[7];
#pragma unroll
        for (int a = 0; a < NBT; ++a) {
          php[u][a] = NBT == 1 ? 1.0 : P[8 + a];  // P0: the basis value 1 at compile time
          if constexpr (OP == 3)
#pragma unroll
            for (int d = 0; d < 3; ++d) cp[u][a][d] = P[12 + 3 * a + d];
        }
      }
      if constexpr (kScaledSL) {
        // P0 single layer, centred tiers: lengths scaled by |k| / sqrt 2 (h = k^2 r^2 / 2), one
        // Newton step y (3 - h y^2) of y = rsq(h) giving rz = 2 sqrt(2) / (|k| r) and the phase
        // h rz - sqrt(2) |k| R in one FMA; the trial weight with rz, the test weight per point
        const double sc = 0.5 * A.k * A.k, kRs = 1.4142135623730951 * fabs(kR);
        cplx ap[PP];
#pragma unroll
        for (int u = 0; u < PP; ++u) ap[u] = {0.0, 0.0};
        for (int q = 0; q < nq; ++q) {
          const auto Q = Qj + (size_t)q * kRec;
          const double xq[3] = {Q[0], Q[1], Q[2]}, wq = Q[3];
#pragma unroll
          for (int u = 0; u < PP; ++u) {
            const double dx = xp[u][0] - xq[0], dy = xp[u][1] - xq[1], dz = xp[u][2] - xq[2];
            const double h = sc * (dx * dx + dy * dy + dz * dz);
            const double y = __builtin_amdgcn_rsq(h);
            const double rz = y * fma(-(h * y), y, 3.0);
            double sn, cs;
            sincos_small<TIER, true>(fma(h, rz, -kRs), &sn, &cs);
            const double w = rz * wq;
            ap[u].re = fma(cs, w, ap[u].re);
            ap[u].im = fma(sn, w, ap[u].im);
          }
        }
#pragma unroll
        for (int u = 0; u < PP; ++u) {
          M[0][0].re = fma(ap[u].re, wJp[u], M[0][0].re);
          M[0][0].im = fma(ap[u].im, wJp[u], M[0][0].im);
        }
      } else {
        for (int q = 0; q < nq; ++q) {
          const auto Q = Qj + (size_t)q * kRec;
          const double xq[3] = {Q[0], Q[1], Q[2]}, nq3[3] = {Q[4], Q[5], Q[6]};
          double phq[NBS], cq[NBS][3];
#pragma unroll
          for (int b = 0; b < NBS; ++b) {
            phq[b] = NBS == 1 ? 1.0 : Q[8 + b];
            if constexpr (OP == 3)
#pragma unroll
              for (int d = 0; d < 3; ++d) cq[b][d] = Q[12 + 3 * b + d];
          }
#pragma unroll
          for (int u = 0; u < PP; ++u)
            add_point_pair<OP, HELM, NBT, NBS, TIER>(xp[u], np[u], php[u], cp[u], xq, nq3, phq, cq, wJp[u] * Q[3],
                                                     wp[u] * Q[7], A.k, M, kR);
        }
      }
    };
    if constexpr (NBT * NBS <= 4) {
      int p = p0;
      for (; p + 1 < p1; p += 2) test_points(std::integral_constant<int, 2>{}, p);
      if (p < p1) test_points(std::integral_constant<int, 1>{}, p);
    } else {
      for (int p = p0; p < p1; ++p) {
        const double* P = A.trec + (size_t)p * kRec;
        const double* XW = A.txw + 4 * (size_t)p;
        const double xp[3] = {XW[0], XW[1], XW[2]}, np[3] = {P[4], P[5], P[6]};
        const double wJp = XW[3], wp = P[7];
        double php[NBT], cp[NBT][3];
#pragma unroll
        for (int a = 0; a < NBT; ++a) {
          php[a] = P[8 + a];
          if constexpr (OP == 3)
#pragma unroll
            for (int d = 0; d < 3; ++d) cp[a][d] = P[12 + 3 * a + d];
        }
        for (int q = 0; q < nq; ++q) {
          const auto Q = Qj + (size_t)q * kRec;
          const double xq[3] = {Q[0], Q[1], Q[2]}, nq3[3] = {Q[4], Q[5], Q[6]};
          double phq[NBS], cq[NBS][3];
#pragma unroll
          for (int b = 0; b < NBS; ++b) {
            phq[b] = Q[8 + b];
            if constexpr (OP == 3)
#pragma unroll
              for (int d = 0; d < 3; ++d) cq[b][d] = Q[12 + 3 * b + d];
          }
          add_point_pair<OP, HELM, NBT, NBS, TIER>(xp, np, php, cp, xq, nq3, phq, cq, wJp * Q[3], wp * Q[7], A.k, M,
                                                   kR);
        }
      }
    }
    if constexpr (kScaledSL) {
      // the sums hold 2 sqrt(2) cos(a) / (|k| r) and 4 sin(a) / (|k| r), a = |k| (r - R);
      // sin(k (r - R)) = sign(k) sin a
      M[0][0].re *= fabs(A.k) * (0.25 * 1.4142135623730951);
      M[0][0].im *= 0.25 * A.k;
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
        if (td[a] >= 0) {
          const cplx v = {M[a][b].re * kInv4Pi, M[a][b].im * kInv4Pi};
          store_value(A.out, (long long)td[a] + A.ld * col, v, A.out_type, A.store_mode);
          // the transposed pair (j, i) adds the same block transposed (symmetric kernel, one space)
          if (A.symmetric) store_value(A.out, (long long)col + A.ld * td[a], v, A.out_type, A.store_mode);
        }
    }
  };
  if constexpr (STAGED) {
    // every lane (those past the last test cell included) takes part in the copies and barriers
    for (int j = j0; j < j1;) {
      const int base = sfirst[j];
      int jend = j + 1;  // the host guarantees that one cell's records fit
      while (jend < j1 && sfirst[jend + 1] - base <= kStageRecords) ++jend;
      const int n = (sfirst[jend] - base) * kRec;
      const double* src = A.srec + (size_t)base * kRec;
      __syncthreads();
      for (int t = threadIdx.x; t < n; t += blockDim.x) stage[t] = src[t];
      __syncthreads();
      for (; j < jend; ++j) trial_cell(j, stage + (sfirst[j] - base) * kRec, sfirst[j + 1] - sfirst[j]);
    }
  } else {
    const auto srec = constant_view(A.srec);
    for (int j = j0; j < j1; ++j) trial_cell(j, srec + (size_t)sfirst[j] * kRec, sfirst[j + 1] - sfirst[j]);
  }
}

// TIER (Helmholtz): 0 full argument reduction; 1, 2 the small-argument sine/cosine of k r
// directly, |k r| <= 0.5 / 0.25 for every point pair of two cells sharing a vertex.
template <int OP, bool HELM, int NBT, int NBS, int TIER = 0>
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
      add_point_pair<OP, HELM, NBT, NBS, TIER>(xp, np, php, cp, xq, nq, phq, cq, P[3] * Q[3], P[7] * Q[7], A.k, M);
    }
  } else {
    for (int i = lane; i < pr.w; i += 64) {
      const double* r = A.rule + 5 * (size_t)(pr.z + i);
      double xp[3], np[3], Jp, php[4], cp[4][3], xq[3], nq[3], Jq, phq[4], cq[4][3];
      cell_point<NBT>(tV, tq, r[0], r[1], xp, np, Jp, php, cp);
      cell_point<NBS>(sV, sq, r[2], r[3], xq, nq, Jq, phq, cq);
      add_point_pair<OP, HELM, NBT, NBS, TIER>(xp, np, php, cp, xq, nq, phq, cq, r[4] * Jp * Jq, r[4], A.k, M);
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

// owns one device allocation; movable (vectors of them reallocate), not copyable
struct DeviceBuf {
  void* p = nullptr;
  DeviceBuf() = default;
  DeviceBuf(const DeviceBuf&) = delete;
  DeviceBuf& operator=(const DeviceBuf&) = delete;
  DeviceBuf(DeviceBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
  ~DeviceBuf() {
    if (p) (void)hipFree(p);
  }
};

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
template <int OP, bool HELM, int NBT, int NBS, bool STAGED>
void launch_regular_cells(int tier, dim3 rg, hipStream_t s, const dev::CellsArgs& A) {
  if (HELM && tier == 2) dev::cells_regular_kernel<OP, HELM, NBT, NBS, 2, STAGED><<<rg, 256, 0, s>>>(A);
  else if (HELM && tier == 1) dev::cells_regular_kernel<OP, HELM, NBT, NBS, 1, STAGED><<<rg, 256, 0, s>>>(A);
  else dev::cells_regular_kernel<OP, HELM, NBT, NBS, 0, STAGED><<<rg, 256, 0, s>>>(A);
}

template <int OP, bool HELM, int NBT, int NBS>
void launch_cells(int tier, dim3 rg, dim3 sg, hipStream_t s, const dev::CellsArgs& A) {
  if (tier >= 0) {
    if (A.max_trial_points <= dev::kStageRecords)
      launch_regular_cells<OP, HELM, NBT, NBS, true>(tier, rg, s, A);
    else
      launch_regular_cells<OP, HELM, NBT, NBS, false>(tier, rg, s, A);
  }
  if (A.npairs > 0) {
    if (HELM && A.adjacent_tier == 2) dev::cells_singular_kernel<OP, HELM, NBT, NBS, 2><<<sg, 256, 0, s>>>(A);
    else if (HELM && A.adjacent_tier == 1) dev::cells_singular_kernel<OP, HELM, NBT, NBS, 1><<<sg, 256, 0, s>>>(A);
    else dev::cells_singular_kernel<OP, HELM, NBT, NBS, 0><<<sg, 256, 0, s>>>(A);
  }
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

// Device data built on first use and cached: per grid the point records for a basis degree and
// the regular rules, the vertex tables and the adjacent pairs with their rules for the
// singular rules; per space the DOF table. Each field is set by the key that builds it.
struct CellsBlob : DeviceBlob {
  std::vector<DeviceBuf> bufs;
  const double* rec = nullptr;  // point records (dev::kRec doubles each)
  const double* xw = nullptr;   // their (x, y, z, w J), 4 doubles each
  const int* first = nullptr;   // first record of each cell [nc + 1]
  int max_points = 0;
  const int* vid = nullptr;     // vertex ids [nc][4]
  const double* cv = nullptr;   // vertex coordinates [nc][4][3]
  double max_radius = 0.0;      // largest centroid-to-vertex distance
  const int4* pairs = nullptr;  // adjacent pairs and their rules
  const double* rule = nullptr;
  int npairs = 0;
  const int* dofs = nullptr;    // DOFs [nc][4]

  template <typename T>
  const T* upload(const std::vector<T>& h) {
    bufs.emplace_back();
    HIP_CHECK_Q(hipMalloc(&bufs.back().p, std::max<size_t>(h.size(), 1) * sizeof(T)));
    if (!h.empty()) HIP_CHECK_Q(hipMemcpy(bufs.back().p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return (const T*)bufs.back().p;
  }
};

template <typename Build>
const CellsBlob& cells_cached(std::mutex& m, DeviceBlobCache& cache, const std::string& key, Build build) {
  int dev = 0;
  HIP_CHECK_Q(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(m);
  std::unique_ptr<DeviceBlob>& slot = cache[{dev, key}];
  if (!slot) {
    auto b = std::make_unique<CellsBlob>();
    build(*b);
    slot = std::move(b);
  }
  return static_cast<const CellsBlob&>(*slot);
}

const CellsBlob& cached_records(const FunctionSpace& sp, const Options& o) {
  auto nq = [&](uint8_t t) {
    auto it = o.quadrature_degrees.find(t);
    return it == o.quadrature_degrees.end() ? -1L : (long)it->second;
  };
  const std::string key = "rec/" + std::to_string(sp.degree) + "/" + std::to_string(nq(kTriangle)) + "/" +
                          std::to_string(nq(kQuadrilateral));
  Grid& g = const_cast<Grid&>(*sp.grid);
  return cells_cached(g.cache_mutex, g.cells_cache, key, [&](CellsBlob& b) {
    std::vector<double> rec;
    std::vector<int> first;
    point_records(sp, o, rec, first);
    b.rec = b.upload(rec);
    std::vector<double> xw(4 * (rec.size() / dev::kRec));
    for (size_t r = 0; r < xw.size() / 4; ++r)
      for (int m = 0; m < 4; ++m) xw[4 * r + m] = rec[dev::kRec * r + m];
    b.xw = b.upload(xw);
    b.first = b.upload(first);
    for (size_t c = 0; c + 1 < first.size(); ++c) b.max_points = std::max(b.max_points, first[c + 1] - first[c]);
  });
}

const CellsBlob& cached_tables(const Grid& grid) {
  Grid& g = const_cast<Grid&>(grid);
  return cells_cached(g.cache_mutex, g.cells_cache, "tab", [&](CellsBlob& b) {
    std::vector<int> vid(4 * g.nc(), -1);
    std::vector<double> cv(12 * g.nc(), 0.0);
    for (int64_t c = 0; c < g.nc(); ++c)
      for (int m = 0; m < g.cell_nv(c); ++m) {
        vid[4 * c + m] = (int)g.vertex(c, m);
        for (int d = 0; d < 3; ++d) cv[12 * c + 3 * m + d] = g.points[3 * g.vertex(c, m) + d];
      }
    b.vid = b.upload(vid);
    b.cv = b.upload(cv);
    b.max_radius = max_cell_radius(g);
  });
}

const CellsBlob& cached_pairs(const Grid& grid, const Options& o) {
  std::string key = "adj";
  for (const auto& kv : o.singular_quadrature_degrees)
    key += "/" + std::to_string(kv.first.first) + "," + std::to_string(kv.first.second) + "," +
           std::to_string(kv.second);
  Grid& g = const_cast<Grid&>(grid);
  return cells_cached(g.cache_mutex, g.cells_cache, key, [&](CellsBlob& b) {
    std::vector<int4> pairs;
    std::vector<double> rule;
    adjacent_cell_pairs(g, o, pairs, rule);
    b.pairs = b.upload(pairs);
    b.rule = b.upload(rule);
    b.npairs = (int)pairs.size();
  });
}

const CellsBlob& cached_dofs(const FunctionSpace& space) {
  FunctionSpace& sp = const_cast<FunctionSpace&>(space);
  return cells_cached(sp.cache_mutex, sp.cells_cache, "dofs", [&](CellsBlob& b) {
    std::vector<int> dofs(4 * sp.grid->nc(), -1);
    for (int64_t c = 0; c < sp.grid->nc(); ++c)
      for (int k = 0; k < sp.cell_nb(c); ++k) dofs[4 * c + k] = (int)sp.dofs[(size_t)sp.nb * c + k];
    b.dofs = b.upload(dofs);
  });
}

// The kernels' arguments from the cached device data (no allocation once cached).
void cells_args(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, dev::CellsArgs& A) {
  const CellsBlob& tr = cached_records(test, a.options);
  const CellsBlob& sr = cached_records(trial, a.options);
  const CellsBlob& tt = cached_tables(*test.grid);
  const CellsBlob& st = cached_tables(*trial.grid);
  const bool same = test.grid == trial.grid;
  A = dev::CellsArgs{};
  A.trec = tr.rec;
  A.txw = tr.xw;
  A.srec = sr.rec;
  A.tfirst = tr.first;
  A.sfirst = sr.first;
  A.tvid = tt.vid;
  A.svid = st.vid;
  A.tcv = tt.cv;
  A.scv = st.cv;
  A.tdofs = cached_dofs(test).dofs;
  A.sdofs = cached_dofs(trial).dofs;
  if (same) {
    const CellsBlob& pr = cached_pairs(*test.grid, a.options);
    A.pairs = pr.pairs;
    A.rule = pr.rule;
    A.npairs = pr.npairs;
  }
  A.n_test = (int)test.grid->nc();
  A.n_trial = (int)trial.grid->nc();
  A.same_grid = same ? 1 : 0;
  A.k = a.k;
  A.ld = (long long)test.size;
  A.out_type = cells_out_type(a.dtype);
  A.store_mode = dev::kAtomic;  // safe for any space; assemble_cells_device relaxes it when it can
  A.max_trial_points = sr.max_points;
  A.radius_sum = tt.max_radius + st.max_radius;
  // points of two cells sharing a vertex are at most 2 (rad_i + rad_j) apart: |k r| <= |k| 2 A.radius_sum
  // (with a margin for rounding; the tiers' polynomials hold to 0.5 and 0.25 exactly)
  A.adjacent_tier = 0;
  if (a.kind == kHelmholtz && same) {
    const double bound = std::fabs(a.k) * 2.0 * A.radius_sum * (1.0 + 1e-12);
    A.adjacent_tier = bound <= 0.25 ? 2 : (bound <= 0.5 ? 1 : 0);
  }
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
  dev::CellsArgs A;
  cells_args(a, trial, test, A);
  if (A.npairs == 0) return;
  const int nbt = test.degree == 0 ? 1 : 4, nbs = trial.degree == 0 ? 1 : 4;
  DeviceBuf coo;
  HIP_CHECK_Q(hipMalloc(&coo.p, (size_t)A.npairs * nbt * nbs * 2 * sizeof(double)));
  A.coo = (double*)coo.p;
  A.correction = correction ? 1 : 0;
  const dim3 sg((unsigned)((A.npairs + 3) / 4));
  if (a.kind == kHelmholtz)
    launch_cells_op<true>(a.op, nbt, nbs, -1, dim3(1), sg, s, A);
  else
    launch_cells_op<false>(a.op, nbt, nbs, -1, dim3(1), sg, s, A);
  HIP_CHECK_Q(hipGetLastError());
  coo_to_csr_device(A.coo, A.pairs, A.npairs, A.tdofs, nbt, 4, A.sdofs, nbs, 4, test.size, trial.size, a.dtype, s, out);
  // coo_to_csr_device ends with a stream synchronisation: the COO buffer may be freed
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
  dev::CellsArgs A;
  cells_args(a, trial, test, A);
  A.out = dev_out;
  // the single-layer and hypersingular kernels are symmetric in (x, y): on one space the pair
  // (j, i) gives the transpose of the block of (i, j)
  A.symmetric = (a.op == kSL || a.op == kHS) && &test == &trial ? 1 : 0;
  // without DOFs shared between cells (P0, discontinuous P1 on both sides) every regular entry
  // has one writer (adjacent entries belong to the pair kernel): plain stores after the zero
  // fill, or a read-add-write when accumulating, instead of atomics
  const bool one_writer = (test.degree == 0 || !test.continuous) && (trial.degree == 0 || !trial.continuous);
  A.store_mode = one_writer ? (accumulate ? dev::kAccumulate : dev::kStore) : dev::kAtomic;
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
    const double bound = std::fabs(a.k) * A.radius_sum * (1.0 + 1e-12);
    tier = bound <= 0.25 ? 2 : (bound <= 0.5 ? 1 : 0);
  }
  if (a.kind == kHelmholtz)
    launch_cells_op<true>(a.op, nbt, nbs, tier, rg, sg, s, A);
  else
    launch_cells_op<false>(a.op, nbt, nbs, 0, rg, sg, s, A);
  HIP_CHECK_Q(hipGetLastError());
}

}  // namespace bempp_amd
