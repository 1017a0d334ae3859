// kernels_mixed.hpp -- mixed-precision regular-pair assembly (P0 x P0).
//
// gfx950 issues FP32 FMAs and FP32 transcendentals at twice the FP64 rate, so far
// cell pairs are evaluated in FP32 while everything that needs FP64 stays FP64:
//   per cell pair (FP64): D = c_test - c_trial (centroids), R = |D|, kR, e^{ikR};
//   per quadrature-point pair (FP32), with offsets o = x - c of each point from its
//   cell centroid (stored in FP32) and e = o_p - o_q:
//     s = e . (2D + e) = r^2 - R^2        (no cancellation: |e| << R)
//     r^2 = R^2 + s, 1/r = rsq(r^2)
//     delta = r - R = s / (r + R),  phase k delta (|k delta| <= 0.5 or 0.25: short
//     minimax series; otherwise reduced by multiples of pi/2 in FP32 first)
//   the 6 trial points of a test point are summed in FP32, the rows in FP64, and the
//   cell-pair sum is rotated by e^{ikR} in FP64.
// Far means R >= kNearFactor (rad_test + rad_trial): then |e| <= R/3, r >= 2R/3 and the
// FP32 terms carry relative errors of a few 1e-7. Nearer (and adjacent) pairs are
// written as zero here and assembled in FP64 by near_pair_kernel (and the Duffy kernel).
#pragma once
#include "kernels.hpp"

namespace bempp_amd {
namespace dev {

constexpr double kNearFactor = 3.0;


// Far-pair predicate shared by the mixed tile kernel and the near-pair kernel.
// Written with explicit fma so both evaluate it bitwise identically: every
// non-adjacent pair is taken by exactly one of them.
__device__ __forceinline__ bool far_pair(const double* ct, const double* cs, double trad, double srad,
                                         double* R2_out) {
  const double dx = ct[0] - cs[0], dy = ct[1] - cs[1], dz = ct[2] - cs[2];
  const double R2 = fma(dx, dx, fma(dy, dy, dz * dz));
  const double near = kNearFactor * (trad + srad);
  *R2_out = R2;
  return R2 >= near * near;
}

struct cplxf {
  float re, im;
};

// sin/cos in FP32 without reduction: TIER 1 for |x| <= 0.5, TIER 2 for |x| <= 0.25.
// sin x = x + x^3 P(x^2), cos x = 1 + x^2 C(x^2) with weighted minimax P, C of degrees
// 2/3 (tier 1) and 1/1 (tier 2) (Lawson's iteration at 50 digits); float32 Horner
// errors below 1.8e-8 (sin) and 3.7e-8 (cos) for tier 1, 7.8e-9 and 4.3e-8 for tier 2.
template <int TIER>
__device__ __forceinline__ void sincos_small_f32(float x, float* s, float* c) {
  static_assert(TIER == 1 || TIER == 2, "phase tiers: 1 (|x| <= 0.5), 2 (|x| <= 0.25)");
  const float u = x * x;
  float p, q;
  if constexpr (TIER == 1) {
    p = fmaf(u, -1.9700505e-4f, 8.3331093e-3f);
    p = fmaf(p, u, -1.6666666e-1f);
    q = fmaf(u, 2.4634197e-5f, -1.3888538e-3f);
    q = fmaf(q, u, 4.1666664e-2f);
    q = fmaf(q, u, -0.5f);
  } else {
    p = fmaf(u, 8.3148913e-3f, -1.6666627e-1f);
    q = fmaf(u, 4.15459625e-2f, -4.99997667e-1f);
  }
  *s = fmaf(x * u, p, x);
  *c = fmaf(u, q, 1.0f);
}

// sin/cos in FP32 for moderate |x| (the far-pair phase k delta, |k delta| <= |k| (rad_i + rad_j)):
// x = n pi/2 + t with pi/2 split in two FP32 parts (each n C fused exactly), |t| <= pi/4,
// Taylor series to t^9 / t^10 (truncation below 2e-9), quadrant by n mod 4.
__device__ __forceinline__ void sincos_f32_quadrant(float t, int q, float* s, float* c);

__device__ __forceinline__ void sincos_f32_wide(float x, float* s, float* c) {
  const float n = __builtin_rintf(x * 0.63661977f);
  float t = fmaf(-n, 1.5707964f, x);
  t = fmaf(-n, -4.3711388e-8f, t);
  sincos_f32_quadrant(t, (int)n, s, c);
}

// sin/cos of n pi/2 + t for |t| <= pi/4 given t and the quadrant count n: Taylor series to
// t^9 / t^10 (truncation below 2e-9), quadrant by n mod 4.
__device__ __forceinline__ void sincos_f32_quadrant(float t, int q, float* s, float* c) {
  const float t2 = t * t;
  float ps = fmaf(t2, 2.7557319e-6f, -1.9841270e-4f);
  ps = fmaf(ps, t2, 8.3333333e-3f);
  ps = fmaf(ps, t2, -1.6666667e-1f);
  const float st = fmaf(ps * t2, t, t);
  float pc = fmaf(t2, -2.7557319e-7f, 2.4801587e-5f);
  pc = fmaf(pc, t2, -1.3888889e-3f);
  pc = fmaf(pc, t2, 4.1666667e-2f);
  pc = fmaf(pc, t2, -0.5f);
  const float ct = fmaf(pc, t2, 1.0f);
  // quadrant q = n mod 4: sin x = (s, c, -s, -c)[q], cos x = (c, -s, -c, s)[q]; the
  // negations flip the sign bit (bit 1 of q for sin, of q + 1 for cos)
  const bool swap = q & 1;
  const float sv = swap ? ct : st, cv = swap ? st : ct;
  *s = __int_as_float(__float_as_int(sv) ^ ((q & 2) << 30));
  *c = __int_as_float(__float_as_int(cv) ^ (((q + 1) & 2) << 30));
}

// acc += w K for one far quadrature-point pair in FP32 (without e^{ikR} and 1/(4 pi)),
// given s = r^2 - R^2 and, for the double-layer kernels, dn = (x - y).n_y (DL) or
// -(x - y).n_x (ADL). PH: phase of k (r - R): 0 range-reduced, 1 or 2 small tiers.
template <int OP, bool HELM, int PH>
__device__ __forceinline__ void far_accumulate_f32(float s, float dn, float R2, float R, float k, float w,
                                                   cplxf& acc) {
  const float r2 = R2 + s;
  const float rinv = __builtin_amdgcn_rsqf(r2);
  if constexpr (HELM) {
    const float r = r2 * rinv;
    const float kdelta = k * (s * __builtin_amdgcn_rcpf(r + R));  // k (r - R) without cancellation
    float sn, cs;
    if constexpr (PH == 0)
      sincos_f32_wide(kdelta, &sn, &cs);
    else
      sincos_small_f32<PH>(kdelta, &sn, &cs);
    if constexpr (OP == 0 || OP == 3) {
      const float rw = rinv * w;
      acc.re = fmaf(cs, rw, acc.re);
      acc.im = fmaf(sn, rw, acc.im);
    } else {
      // e^{ik(r-R)} (1 - ikr) dn / r^3
      const float kr = k * r;
      const float f = dn * (rinv * rinv) * (rinv * w);
      acc.re = fmaf(fmaf(kr, sn, cs), f, acc.re);
      acc.im = fmaf(fmaf(-kr, cs, sn), f, acc.im);
    }
  } else {
    if constexpr (OP == 0 || OP == 3)
      acc.re = fmaf(rinv, w, acc.re);
    else
      acc.re = fmaf(dn * (rinv * rinv), rinv * w, acc.re);
  }
}

// Helmholtz far quadrature-point pair with lengths scaled by kappa = |k| (P0 tiles): given
// sh = kappa^2 s / 2, R2s = (kappa R)^2 and Rs = kappa R, adds e^{ix} / r' (SL, HS) or
// e^{ix} (1 - i r') dn / r'^3 (DL, ADL) to acc, with r' = kappa r and x = r' - Rs = kappa (r - R)
// (the caller multiplies by kappa or kappa^3 and conjugates for k < 0). x starts as
// r'^2 rsq(r'^2) - Rs (absolute error ~1e-7 r') and takes one Newton step on
// x (2 Rs + x) = 2 sh, with the derivative 2 (Rs + x) = 2 r' taken from the same rsq, which
// leaves a relative error of ~1e-7 in x without a division.
template <int OP, int PH>
__device__ __forceinline__ void far_accumulate_scaled(float sh, float dn, float R2s, float Rs, cplxf& acc) {
  const float r2 = fmaf(2.0f, sh, R2s);
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float x0 = fmaf(r2, rinv, -Rs);
  const float f = fmaf(x0, fmaf(0.5f, x0, Rs), -sh);
  const float x = fmaf(-f, rinv, x0);
  float sn, cs;
  if constexpr (PH == 0)
    sincos_f32_wide(x, &sn, &cs);
  else
    sincos_small_f32<PH>(x, &sn, &cs);
  if constexpr (OP == 0 || OP == 3) {
    acc.re = fmaf(cs, rinv, acc.re);
    acc.im = fmaf(sn, rinv, acc.im);
  } else {
    const float rp = Rs + x;
    const float f3 = dn * (rinv * rinv) * rinv;
    acc.re = fmaf(fmaf(rp, sn, cs), f3, acc.re);
    acc.im = fmaf(fmaf(-rp, cs, sn), f3, acc.im);
  }
}

// Mixed-precision P0 x P0 tiles: same tiling, symmetry and LDS mirror as
// p0_tile; near and adjacent pairs are written as zero. With offsets o_p, o_q of the
// points from their cell centroids and D = c_test - c_trial,
//   s_pq = r^2 - R^2 = a_p - b_q - 2 o_p.o_q,  a_p = o_p.(2D + o_p),  b_q = o_q.(2D - o_q),
// so a cell pair costs 2Q short dot products and a point pair one 3-term dot product.
// The double-layer normal products split the same way: (x - y).n = (D + o_p).n - o_q.n.
template <int Q, int OP, bool HELM, int PH, int CH = kChunk, int NW = 4, typename Args>
__device__ __forceinline__ void p0_tile_mixed(Args& A, const int2 tile, int symmetric,
                                              double2 (&lds)[NW][CH][kTile]) {
  if constexpr (HELM) {
    // k = 0: the Helmholtz kernels reduce to the Laplace ones (and the lengths cannot be scaled by k)
    if (A.k == 0.0) {
      p0_tile_mixed<Q, OP, false, PH, CH, NW>(A, tile, symmetric, lds);
      return;
    }
  }
  constexpr bool kDN = OP == 1 || OP == 2;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = tile.x * kTile + lane;
  const bool i_ok = i < A.n_test;
  const int il = i_ok ? i : A.n_test - 1;
  constexpr int NC = kWeightClasses<Q>;
  // Helmholtz: lengths scaled by kappa = |k| and s carried as sigma s, sigma = kappa^2 / 2
  // (far_accumulate_scaled); Laplace: unscaled, sigma = 1.
  const double kappa = HELM ? fabs(A.k) : 1.0;
  const float sig = HELM ? (float)(0.5 * kappa * kappa) : 1.0f;
  // test point offsets o_p kept as m_p = -2 sigma o_p and sigma |o_p|^2
  float mp[Q][3], oo[Q];
#pragma unroll
  for (int p = 0; p < Q; ++p) {
    const float4 o = ((const float4*)A.test_pts32)[(size_t)il * Q + p];
    mp[p][0] = -2.0f * sig * o.x;
    mp[p][1] = -2.0f * sig * o.y;
    mp[p][2] = -2.0f * sig * o.z;
    oo[p] = sig * fmaf(o.x, o.x, fmaf(o.y, o.y, o.z * o.z));
  }
  float wt[NC];  // test weight of each class (rule weight x Jacobian)
#pragma unroll
  for (int c = 0; c < NC; ++c) wt[c] = ((const float4*)A.test_pts32)[(size_t)il * Q + class_point(Q, c)].w;
  const double* tgeo = A.test_geo + (size_t)il * kGeoStride;
  const float nx[3] = {(float)tgeo[9], (float)tgeo[10], (float)tgeo[11]};
  const double ct[3] = {tgeo[kGeoCentroid], tgeo[kGeoCentroid + 1], tgeo[kGeoCentroid + 2]};
  const double trad = tgeo[kGeoRadius];
  const int tv0 = A.test_vid[4 * il], tv1 = A.test_vid[4 * il + 1], tv2 = A.test_vid[4 * il + 2];
  const bool mirror = symmetric && tile.x != tile.y;
  const int skind = fast_store_kind(A.out_type, A.mode);
  // every cell pair of the tile far (host bound): no adjacency or near test per pair
  const bool tile_far = A.far_tiles != nullptr && A.far_tiles[(size_t)tile.x * A.far_tiles_ld + tile.y] != 0;
  // The FP32 trial data are read with ordinary (vector) loads so that they live in
  // VGPRs: gfx950 dual-issues FP32 VALU operations only without SGPR operands.
  const float4* trial_vtx32 = (const float4*)A.trial_vtx32;
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
      if (j >= A.n_trial) {
        lds[wave][jj][lane] = make_double2(0.0, 0.0);
        continue;
      }
      // trial vertex offsets d_0, d_1 (d_2 = -d_0 - d_1), |d_k|^2, class weights
      const float4 ve0 = trial_vtx32[3 * (size_t)j], ve1 = trial_vtx32[3 * (size_t)j + 1],
                   ve2 = trial_vtx32[3 * (size_t)j + 2];
      const float dk[2][3] = {{ve0.x, ve0.y, ve0.z}, {ve1.x, ve1.y, ve1.z}};
      const float gk[3] = {ve0.w, ve1.w, ve2.w};
      const float wcl[3] = {ve2.x, ve2.y, ve2.z};
      const auto sg = trial_geo + (size_t)j * kGeoStride;
      const double cs[3] = {sg[kGeoCentroid], sg[kGeoCentroid + 1], sg[kGeoCentroid + 2]};
      const double srad = sg[kGeoRadius];
      const float ny[3] = {(float)sg[9], (float)sg[10], (float)sg[11]};
      // cell-pair quantities in FP64
      const double Dx = ct[0] - cs[0], Dy = ct[1] - cs[1], Dz = ct[2] - cs[2];
      double R2;
      bool far;
      if (tile_far) {
        R2 = fma(Dx, Dx, fma(Dy, Dy, Dz * Dz));
        far = true;
      } else {
        const int sv0 = trial_vid[4 * j], sv1 = trial_vid[4 * j + 1], sv2 = trial_vid[4 * j + 2];
        const bool adjacent = A.same_grid && (tv0 == sv0 || tv0 == sv1 || tv0 == sv2 || tv1 == sv0 ||
                                              tv1 == sv1 || tv1 == sv2 || tv2 == sv0 || tv2 == sv1 || tv2 == sv2);
        far = far_pair(ct, cs, trad, srad, &R2) && !adjacent;
      }
      cplx acc = {0.0, 0.0};
      if (far) {
        const double R = R2 * rsqrt_fast(R2);
        const float D[3] = {(float)Dx, (float)Dy, (float)Dz};
        const float D2[3] = {2.0f * D[0], 2.0f * D[1], 2.0f * D[2]};
        const float R2f = (float)(HELM ? kappa * kappa * R2 : R2), Rf = (float)(kappa * R);
        const float* nrm = OP == 1 ? ny : nx;
        const float sgn = OP == 2 ? -1.0f : 1.0f;
        // o_p.n = -m_p.n / (2 sigma) (times the sign of the adjoint double layer)
        const float hsc = -0.5f * sgn / sig;
        const float hn[3] = {hsc * nrm[0], hsc * nrm[1], hsc * nrm[2]};
        // D.n in FP64: (x - y).n ~ r^2 / 2 on smooth surfaces, well below |D| |n|
        double Dn64 = 0.0;
        if constexpr (OP == 1) Dn64 = Dx * sg[9] + Dy * sg[10] + Dz * sg[11];
        if constexpr (OP == 2) Dn64 = -(Dx * tgeo[9] + Dy * tgeo[10] + Dz * tgeo[11]);
        const float Dn = (float)Dn64;
        // Trial points o_q = t_q d_k(q) (orbit form): with w_pk = (m_p - 2 sigma D).d_k,
        //   sigma s_pq = sigma a_p + sigma t_q^2 |d_k|^2 + t_q w_pk,
        // and w_p2 = -w_p0 - w_p1 since the d_k sum to zero.
        const float Dm[3] = {-sig * D2[0], -sig * D2[1], -sig * D2[2]};
        const float Dd[2] = {fmaf(Dm[0], dk[0][0], fmaf(Dm[1], dk[0][1], Dm[2] * dk[0][2])),
                             fmaf(Dm[0], dk[1][0], fmaf(Dm[1], dk[1][1], Dm[2] * dk[1][2]))};
        float cq[Q], nq[Q];  // sigma t_q^2 |d_k|^2 and (DL/ADL) o_q.n
        float dkn[3] = {0.0f, 0.0f, 0.0f};
        if constexpr (kDN) {
          dkn[0] = sgn * fmaf(dk[0][0], nrm[0], fmaf(dk[0][1], nrm[1], dk[0][2] * nrm[2]));
          dkn[1] = sgn * fmaf(dk[1][0], nrm[0], fmaf(dk[1][1], nrm[1], dk[1][2] * nrm[2]));
          dkn[2] = -(dkn[0] + dkn[1]);
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const float t = orbit_t(Q, q);
          cq[q] = t == 0.0f ? 0.0f : (sig * (t * t)) * gk[orbit_vertex(Q, q)];
          if constexpr (kDN) nq[q] = t * dkn[orbit_vertex(Q, q)];
        }
        cplxf cls[NC][NC];  // unweighted sums per (test class, trial class)
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
          for (int b = 0; b < NC; ++b) cls[a][b] = {0.0f, 0.0f};
#pragma unroll
        for (int p = 0; p < Q; ++p) {
          // sigma a_p = sigma o_p.(2D + o_p) = sigma |o_p|^2 - D.m_p
          const float mx = mp[p][0], my = mp[p][1], mz = mp[p][2];
          const float ap = fmaf(-D[0], mx, fmaf(-D[1], my, fmaf(-D[2], mz, oo[p])));
          float dnp = 0.0f;
          if constexpr (kDN) dnp = fmaf(mx, hn[0], fmaf(my, hn[1], fmaf(mz, hn[2], Dn)));
          const float w0 = fmaf(mx, dk[0][0], fmaf(my, dk[0][1], fmaf(mz, dk[0][2], Dd[0])));
          const float w1 = fmaf(mx, dk[1][0], fmaf(my, dk[1][1], fmaf(mz, dk[1][2], Dd[1])));
          const float wk[3] = {w0, w1, -(w0 + w1)};
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const float t = orbit_t(Q, q);
            const float s = t == 0.0f ? ap : fmaf(t, wk[orbit_vertex(Q, q)], ap + cq[q]);
            cplxf& cl = cls[weight_class(Q, p)][weight_class(Q, q)];
            if constexpr (HELM)
              far_accumulate_scaled<OP, PH>(s, kDN ? dnp - nq[q] : 0.0f, R2f, Rf, cl);
            else
              far_accumulate_f32<OP, false, PH>(s, kDN ? dnp - nq[q] : 0.0f, R2f, Rf, 0.0f, 1.0f, cl);
          }
        }
        // class weights, length scaling and e^{ikR} in FP32 (relative error ~1e-7 per entry);
        // only the reduction kR = n pi/2 + t needs FP64
        cplxf af = {0.0f, 0.0f};
#pragma unroll
        for (int a = 0; a < NC; ++a) {
          cplxf row = {0.0f, 0.0f};
#pragma unroll
          for (int b = 0; b < NC; ++b) {
            const float ws = wcl[b];
            row.re = fmaf(cls[a][b].re, ws, row.re);
            row.im = fmaf(cls[a][b].im, ws, row.im);
          }
          af.re = fmaf(row.re, wt[a], af.re);
          af.im = fmaf(row.im, wt[a], af.im);
        }
        if constexpr (HELM) {
          // undo the length scaling (1/r = kappa / r'), conjugate for k < 0, rotate by e^{ikR}
          const double kR = A.k * R;
          const double n = rint(kR * 0.63661977236758134);
          const double t = fma(-n, 6.123233995736766e-17, fma(-n, 1.5707963267948966, kR));
          float sn, cs;
          sincos_f32_quadrant((float)t, (int)n, &sn, &cs);
          const float ks = (float)(kDN ? kappa * kappa * kappa : kappa);
          const float im = A.k < 0.0 ? -af.im : af.im;
          af = {fmaf(af.re, cs, -im * sn) * ks, fmaf(af.re, sn, im * cs) * ks};
        }
        if constexpr (OP == 3) {
          const float nn = HELM ? (float)(-A.k * A.k * (tgeo[9] * sg[9] + tgeo[10] * sg[10] + tgeo[11] * sg[11])) : 0.0f;
          af = {nn * af.re, nn * af.im};
        }
        acc = {(double)af.re, (double)af.im};
      }
      const cplx v = {acc.re * kInv4Pi, acc.im * kInv4Pi};
      if (i_ok) store_entry(A.out, (long long)i + A.ld * (long long)j, v, skind, A.out_type, A.mode);
      if (mirror) lds[wave][jj][lane] = make_double2(v.re, v.im);
    }
    if (mirror) {
      __syncthreads();
      const int row = lane % CH;
      const int jr = jbase + row;
#pragma unroll
      for (int s = 0; s < CH; ++s) {
        const int col = s * (kTile / CH) + lane / CH;
        const int ic = tile.x * kTile + col;
        const double2 val = lds[wave][row][col];
        if (jr < A.n_trial && ic < A.n_test)
          store_entry(A.out, (long long)jr + A.ld * (long long)ic, {val.x, val.y}, skind, A.out_type, A.mode);
      }
      __syncthreads();
    }
  }
}

// 4 waves per SIMD (<= 128 VGPRs): FP32 dual issue needs several resident waves.
template <int Q, int OP, bool HELM, int PH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) regular_p0_tiled_mixed(const RegularArgs A, const int2* __restrict__ tiles,
                                                              int symmetric) {
  __shared__ double2 lds[4][kChunk][kTile];
  p0_tile_mixed<Q, OP, HELM, PH>(A, tiles[blockIdx.x], symmetric, lds);
}

// ---------------------------------------------------------------------------
// Mixed precision with a P1 side: the LDS-accumulating tiles of regular_tiled_lds with
// far cell pairs evaluated in FP32, with the point-pair arithmetic of p0_tile_mixed. With
// W_b(q) = rule weight x basis value (wphi_trial; a P0 side has {w, 0, 0}), the trial sums
// T_b(p) = sum_q W_b(q) K_pq and the test sums M_ab = sum_p W_a(p) T_b(p) are taken in
// FP32; the Jacobians, the length scaling and e^{ikR} are applied once per cell pair in
// FP64. The Laplace hypersingular operator needs only sum_pq w_p w_q K_pq (its P1 surface
// curls are constant on a cell). Near and adjacent pairs contribute nothing here: near
// pairs come from near_pair_p1_tiled in FP64, adjacent pairs from the Duffy kernel.
// ---------------------------------------------------------------------------

// FP32 kernel value of a far point pair without weights, e^{ikR} and 1/(4 pi). Helmholtz:
// lengths scaled by kappa as in far_accumulate_scaled (sh = kappa^2 s / 2); Laplace:
// unscaled (sh = s, R2s = R^2).
template <int OP, bool HELM, int PH>
__device__ __forceinline__ cplxf far_value(float sh, float dn, float R2s, float Rs) {
  if constexpr (HELM) {
    const float r2 = fmaf(2.0f, sh, R2s);
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float x0 = fmaf(r2, rinv, -Rs);
    const float f = fmaf(x0, fmaf(0.5f, x0, Rs), -sh);
    const float x = fmaf(-f, rinv, x0);
    float sn, cs;
    if constexpr (PH == 0)
      sincos_f32_wide(x, &sn, &cs);
    else
      sincos_small_f32<PH>(x, &sn, &cs);
    if constexpr (OP == 0 || OP == 3) {
      return {cs * rinv, sn * rinv};
    } else {
      const float rp = Rs + x;
      const float f3 = dn * (rinv * rinv) * rinv;
      return {fmaf(rp, sn, cs) * f3, fmaf(-rp, cs, sn) * f3};
    }
  } else {
    const float rinv = __builtin_amdgcn_rsqf(R2s + sh);
    if constexpr (OP == 0 || OP == 3)
      return {rinv, 0.0f};
    else
      return {dn * (rinv * rinv) * rinv, 0.0f};
  }
}

template <int Q, int OP, bool HELM, int PH>
__device__ __forceinline__ void p1_tile_mixed(const RegularArgs& A, const int2 tile, int symmetric, double2* acc) {
  if constexpr (HELM) {
    // k = 0: the Helmholtz kernels reduce to the Laplace ones (and the lengths cannot be scaled by k)
    if (A.k == 0.0) {
      p1_tile_mixed<Q, OP, false, PH>(A, tile, symmetric, acc);
      return;
    }
  }
  constexpr bool kDN = OP == 1 || OP == 2;
  constexpr bool kSum = OP == 3 && !HELM;
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
  const double kappa = HELM ? fabs(A.k) : 1.0;
  const float sig = HELM ? (float)(0.5 * kappa * kappa) : 1.0f;
  float mp[Q][3], mq[Q];  // -2 sigma o_p and sigma |o_p|^2
#pragma unroll
  for (int p = 0; p < Q; ++p) {
    const float4 o = ((const float4*)A.test_pts32)[(size_t)il * Q + p];
    mp[p][0] = -2.0f * sig * o.x;
    mp[p][1] = -2.0f * sig * o.y;
    mp[p][2] = -2.0f * sig * o.z;
    mq[p] = sig * fmaf(o.x, o.x, fmaf(o.y, o.y, o.z * o.z));
  }
  // trial weights x basis values in VGPRs: an SGPR operand would stop FP32 dual issue
  float ws[Q][3];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int b = 0; b < 3; ++b) asm volatile("v_mov_b32 %0, %1" : "=v"(ws[q][b]) : "s"(A.wphi_trial[q][b]));
  const double* tgeo = A.test_geo + (size_t)il * kGeoStride;
  const float nx[3] = {(float)tgeo[9], (float)tgeo[10], (float)tgeo[11]};
  const double ct[3] = {tgeo[kGeoCentroid], tgeo[kGeoCentroid + 1], tgeo[kGeoCentroid + 2]};
  const double trad = tgeo[kGeoRadius];
  const int tv0 = A.test_vid[4 * il], tv1 = A.test_vid[4 * il + 1], tv2 = A.test_vid[4 * il + 2];
  int lt[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (a < A.nbt) lt[a] = A.test_ldof[(size_t)il * A.nbt + a];
  const float4* trial_pts32 = (const float4*)A.trial_pts32;
  const auto trial_geo = constant_view(A.trial_geo);
  const auto trial_vid = constant_view(A.trial_vid);
  const auto trial_ldof = constant_view(A.trial_ldof);
#pragma unroll 1
  for (int jj = 0; jj < 16; ++jj) {
    const int j = tile.y * kTile + wave * 16 + jj;
    if (j >= A.n_trial) break;  // wave-uniform
    float sp[Q][3];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const float4 t = trial_pts32[(size_t)j * Q + q];
      sp[q][0] = t.x;
      sp[q][1] = t.y;
      sp[q][2] = t.z;
    }
    const auto sg = trial_geo + (size_t)j * kGeoStride;
    const double cs[3] = {sg[kGeoCentroid], sg[kGeoCentroid + 1], sg[kGeoCentroid + 2]};
    const double srad = sg[kGeoRadius];
    const float ny[3] = {(float)sg[9], (float)sg[10], (float)sg[11]};
    const int sv0 = trial_vid[4 * j], sv1 = trial_vid[4 * j + 1], sv2 = trial_vid[4 * j + 2];
    int ls[3] = {0, 0, 0};
#pragma unroll
    for (int b = 0; b < 3; ++b)
      if (b < A.nbs) ls[b] = trial_ldof[(size_t)j * A.nbs + b];
    const bool adjacent = A.same_grid && (tv0 == sv0 || tv0 == sv1 || tv0 == sv2 || tv1 == sv0 || tv1 == sv1 ||
                                          tv1 == sv2 || tv2 == sv0 || tv2 == sv1 || tv2 == sv2);
    double R2;
    const bool far = far_pair(ct, cs, trad, srad, &R2) && !adjacent && i_ok;
    if (!far) continue;
    const double Dx = ct[0] - cs[0], Dy = ct[1] - cs[1], Dz = ct[2] - cs[2];
    const double R = R2 * rsqrt_nr(R2);
    const float D[3] = {(float)Dx, (float)Dy, (float)Dz};
    const float D2[3] = {2.0f * D[0], 2.0f * D[1], 2.0f * D[2]};
    const float R2s = (float)(HELM ? kappa * kappa * R2 : R2), Rs = (float)(kappa * R);
    const float* nrm = OP == 1 ? ny : nx;
    const float sgn = OP == 2 ? -1.0f : 1.0f;
    double Dn64 = 0.0;
    if constexpr (OP == 1) Dn64 = Dx * sg[9] + Dy * sg[10] + Dz * sg[11];
    if constexpr (OP == 2) Dn64 = -(Dx * tgeo[9] + Dy * tgeo[10] + Dz * tgeo[11]);
    const float Dn = (float)Dn64;
    // o_p.n = -m_p.n / (2 sigma) (times the sign of the adjoint double layer)
    const float hsc = -0.5f * sgn / sig;
    const float hn[3] = {hsc * nrm[0], hsc * nrm[1], hsc * nrm[2]};
    float bq[Q], nq[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const float ox = sp[q][0], oy = sp[q][1], oz = sp[q][2];
      bq[q] = sig * fmaf(D2[0] - ox, ox, fmaf(D2[1] - oy, oy, (D2[2] - oz) * oz));
      nq[q] = kDN ? sgn * fmaf(ox, nrm[0], fmaf(oy, nrm[1], oz * nrm[2])) : 0.0f;
    }
    cplxf M[3][3];
    float S = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) M[a][b] = {0.0f, 0.0f};
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const float ap = fmaf(-mp[p][0], D[0], fmaf(-mp[p][1], D[1], fmaf(-mp[p][2], D[2], mq[p])));
      const float dnp = kDN ? fmaf(mp[p][0], hn[0], fmaf(mp[p][1], hn[1], fmaf(mp[p][2], hn[2], Dn))) : 0.0f;
      cplxf T[3] = {{0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}};
      float T0 = 0.0f;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const float s = fmaf(mp[p][0], sp[q][0], fmaf(mp[p][1], sp[q][1], fmaf(mp[p][2], sp[q][2], ap - bq[q])));
        const cplxf kv = far_value<OP, HELM, PH>(s, dnp - nq[q], R2s, Rs);
        if constexpr (kSum) {
          T0 = fmaf(kv.re, ws[q][0] + ws[q][1] + ws[q][2], T0);
        } else {
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            T[b].re = fmaf(kv.re, ws[q][b], T[b].re);
            if constexpr (HELM) T[b].im = fmaf(kv.im, ws[q][b], T[b].im);
          }
        }
      }
      if constexpr (kSum) {
        S = fmaf(A.wphi_test[p][0] + A.wphi_test[p][1] + A.wphi_test[p][2], T0, S);
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            M[a][b].re = fmaf(A.wphi_test[p][a], T[b].re, M[a][b].re);
            if constexpr (HELM) M[a][b].im = fmaf(A.wphi_test[p][a], T[b].im, M[a][b].im);
          }
      }
    }
    // Jacobians, the length scaling (1/r = kappa / r'), k < 0 (conjugate) and e^{ikR} in FP32;
    // only the reduction kR = n pi/2 + t needs FP64
    double scale = tgeo[12] * sg[12];
    if constexpr (HELM) scale *= kDN ? kappa * kappa * kappa : kappa;
    float er = (float)scale, ei = 0.0f;
    if constexpr (HELM) {
      const double kR = A.k * R;
      const double n = rint(kR * 0.63661977236758134);
      const double t = fma(-n, 6.123233995736766e-17, fma(-n, 1.5707963267948966, kR));
      float sn, cs;
      sincos_f32_quadrant((float)t, (int)n, &sn, &cs);
      ei = er * sn;
      er = er * cs;
    }
    const float sk = HELM && A.k < 0.0 ? -1.0f : 1.0f;
    cplx M64[3][3];
    cplx msum = {0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const float mr = M[a][b].re, mi = sk * M[a][b].im;
        M64[a][b] = {(double)fmaf(mr, er, -mi * ei), (double)fmaf(mr, ei, mi * er)};
        msum.re += M64[a][b].re;
        msum.im += M64[a][b].im;
      }
    if constexpr (kSum) msum = {(double)S * scale, 0.0};
    msum = hs_msum<OP>(msum, tgeo, sg);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (a >= A.nbt || b >= A.nbs) continue;
        const cplx v = local_entry<OP, HELM>(M64, a, b, msum, A.nbt, A.nbs, tgeo, sg, a, b, A.k);
        double2* slot = acc + (ls[b] * ndt + lt[a]);
        atomicAdd(&slot->x, v.re);
        if constexpr (HELM) atomicAdd(&slot->y, v.im);
      }
  }
  __syncthreads();
  flush_tile(A, acc, t0, ndt, s0, nds, symmetric && tile.x != tile.y);
}

template <int Q, int OP, bool HELM, int PH>
__global__ void __launch_bounds__(256) regular_tiled_lds_mixed(const RegularArgs A, const int2* __restrict__ tiles,
                                                               int symmetric) {
  extern __shared__ double2 acc[];
  p1_tile_mixed<Q, OP, HELM, PH>(A, tiles[blockIdx.x], symmetric, acc);
}

// ---------------------------------------------------------------------------
// The near-pair list of one grid, built on the device: for every tile (blocks a, b of
// kTile cells) not flagged all-far, its non-adjacent cell pairs that far_pair rejects --
// the same predicate, evaluated the same way, as the mixed tiles use, so the list holds
// exactly the pairs the tiles leave out. Two passes over the candidate tiles (count, then
// fill at the offsets the host computed from the counts); thread t of a tile's 256 takes
// test cell kTile a + t % 64 and trial cells kTile b + 16 (t / 64) + 0..15, and writes its
// pairs in that order after the pairs of threads < t (deterministic order).
// ---------------------------------------------------------------------------
struct NearBuildArgs {
  const double* geo;   // [n][kGeoStride]
  const int* vid;      // [n][4]
  int n;
  const int2* tiles;   // candidate tiles (a, b)
  int* counts;         // [ntiles] pairs per tile (count pass)
  const int* offsets;  // [ntiles] first pair of each tile (fill pass)
  int2* out;
};

__device__ __forceinline__ int near_build_thread(const NearBuildArgs& B, int2 tile, int t, int2* out) {
  const int i = tile.x * kTile + (t & 63);
  if (i >= B.n) return 0;
  const double* tg = B.geo + (size_t)i * kGeoStride;
  const double ct[3] = {tg[kGeoCentroid], tg[kGeoCentroid + 1], tg[kGeoCentroid + 2]};
  const double trad = tg[kGeoRadius];
  const int tv0 = B.vid[4 * i], tv1 = B.vid[4 * i + 1], tv2 = B.vid[4 * i + 2];
  int cnt = 0;
  for (int jj = 0; jj < 16; ++jj) {
    const int j = tile.y * kTile + (t >> 6) * 16 + jj;
    if (j >= B.n) break;
    const double* sg = B.geo + (size_t)j * kGeoStride;
    const int sv0 = B.vid[4 * j], sv1 = B.vid[4 * j + 1], sv2 = B.vid[4 * j + 2];
    const bool adjacent = tv0 == sv0 || tv0 == sv1 || tv0 == sv2 || tv1 == sv0 || tv1 == sv1 || tv1 == sv2 ||
                          tv2 == sv0 || tv2 == sv1 || tv2 == sv2;
    double R2;
    if (!adjacent && !far_pair(ct, sg + kGeoCentroid, trad, sg[kGeoRadius], &R2)) {
      if (out) out[cnt] = make_int2(i, j);
      ++cnt;
    }
  }
  return cnt;
}

// (templates so that every translation unit including this header may instantiate them)
template <typename Args = NearBuildArgs>
__global__ void __launch_bounds__(256) near_count_kernel(const Args B) {
  __shared__ int total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  const int c = near_build_thread(B, B.tiles[blockIdx.x], threadIdx.x, nullptr);
  if (c) atomicAdd(&total, c);
  __syncthreads();
  if (threadIdx.x == 0) B.counts[blockIdx.x] = total;
}

template <typename Args = NearBuildArgs>
__global__ void __launch_bounds__(256) near_fill_kernel(const Args B) {
  __shared__ int scan[256];
  const int2 tile = B.tiles[blockIdx.x];
  const int c = near_build_thread(B, tile, threadIdx.x, nullptr);
  scan[threadIdx.x] = c;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {  // inclusive prefix sum over the block
    const int v = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0;
    __syncthreads();
    scan[threadIdx.x] += v;
    __syncthreads();
  }
  if (c) near_build_thread(B, tile, threadIdx.x, B.out + B.offsets[blockIdx.x] + scan[threadIdx.x] - c);
}

// FP64 evaluation of the near (non-adjacent) cell pairs with a P1 side, grouped by
// (test block, trial block) of kTile cells: one workgroup per group sums its pairs' local
// blocks in LDS at block-local DOF indices (as regular_tiled_lds does) and adds the touched
// entries to the output with one global atomic each.
template <int Q, int OP, bool HELM, int CENTRED>
__global__ void __launch_bounds__(256) near_pair_p1_tiled(const RegularArgs A, const int2* __restrict__ pairs,
                                                          const int4* __restrict__ segs) {
  extern __shared__ double2 acc[];
  const int4 sgm = segs[blockIdx.x];
  const int t0 = A.test_blk_off[sgm.x], ndt = A.test_blk_off[sgm.x + 1] - t0;
  const int s0 = A.trial_blk_off[sgm.y], nds = A.trial_blk_off[sgm.y + 1] - s0;
  for (int e = threadIdx.x; e < ndt * nds; e += blockDim.x) acc[e] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int e = threadIdx.x; e < sgm.w; e += blockDim.x) {
    const int2 pr = pairs[sgm.z + e];
    const int i = pr.x, j = pr.y;
    const double* tg = A.test_geo + (size_t)i * kGeoStride;
    const double* sg = A.trial_geo + (size_t)j * kGeoStride;
    double R2;
    if (far_pair(tg + kGeoCentroid, sg + kGeoCentroid, tg[kGeoRadius], sg[kGeoRadius], &R2)) continue;
    double xp[Q][3], wp[Q], sp[Q * 4];
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const double* s = A.test_pts + ((size_t)i * Q + p) * 4;
      xp[p][0] = s[0];
      xp[p][1] = s[1];
      xp[p][2] = s[2];
      wp[p] = s[3];
    }
#pragma unroll
    for (int m = 0; m < Q * 4; ++m) sp[m] = A.trial_pts[(size_t)j * Q * 4 + m];
    const double nx[3] = {tg[9], tg[10], tg[11]};
    const double ct[3] = {tg[kGeoCentroid], tg[kGeoCentroid + 1], tg[kGeoCentroid + 2]};
    cplx M[3][3], msum;
    p1_local<Q, OP, HELM, CENTRED>(A, xp, wp, nx, ct, sp, sg, M, msum);
    msum = hs_msum<OP>(msum, tg, sg);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (a >= A.nbt || b >= A.nbs) continue;
        const cplx v = local_entry<OP, HELM>(M, a, b, msum, A.nbt, A.nbs, tg, sg, a, b, A.k);
        double2* slot = acc + (A.trial_ldof[(size_t)j * A.nbs + b] * ndt + A.test_ldof[(size_t)i * A.nbt + a]);
        atomicAdd(&slot->x, v.re);
        if constexpr (HELM) atomicAdd(&slot->y, v.im);
      }
  }
  __syncthreads();
  flush_tile(A, acc, t0, ndt, s0, nds, false);
}

// FP64 evaluation of a list of near (non-adjacent) P0 x P0 cell pairs, one
// thread per pair, added to their (unique) entries.
template <int Q, int OP, bool HELM, bool CENTRED, typename Args>
__device__ __forceinline__ void near_pair(Args& A, const int2 pr) {
  const int i = pr.x, j = pr.y;
  const double* tg = A.test_geo + (size_t)i * kGeoStride;
  const double* sg = A.trial_geo + (size_t)j * kGeoStride;
  double R2;
  // re-check with the tiles' predicate (the list holds exactly the pairs they leave out)
  if (far_pair(tg + kGeoCentroid, sg + kGeoCentroid, tg[kGeoRadius], sg[kGeoRadius], &R2)) return;
  double xp[Q][3], wp[Q], sp[Q * 4];
#pragma unroll
  for (int p = 0; p < Q; ++p) {
    const double* s = A.test_pts + ((size_t)i * Q + p) * 4;
    xp[p][0] = s[0];
    xp[p][1] = s[1];
    xp[p][2] = s[2];
    wp[p] = s[3];
  }
#pragma unroll
  for (int m = 0; m < Q * 4; ++m) sp[m] = A.trial_pts[(size_t)j * Q * 4 + m];
  const double nx[3] = {tg[9], tg[10], tg[11]};
  const double ny[3] = {sg[9], sg[10], sg[11]};
  cplx acc;
  if constexpr (CENTRED) {
    double kR;
    const cplx e = pair_phase(tg + kGeoCentroid, sg + kGeoCentroid, A.k, &kR);
    acc = cmul(p0_pair<Q, OP, HELM, CENTRED>(xp, wp, nx, sp, ny, A.k, kR), e);
  } else {
    acc = p0_pair<Q, OP, HELM, 0>(xp, wp, nx, sp, ny, A.k, 0.0);
  }
  if constexpr (OP == 3) {
    const double nn = HELM ? -A.k * A.k * (nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2]) : 0.0;
    acc = {nn * acc.re, nn * acc.im};
  }
  store_value(A.out, (long long)i + A.ld * (long long)j, {acc.re * kInv4Pi, acc.im * kInv4Pi}, A.out_type,
              kAccumulate);
}

template <int Q, int OP, bool HELM, bool CENTRED>
__global__ void __launch_bounds__(256) near_pair_kernel(const RegularArgs A, const int2* __restrict__ pairs,
                                                        int npairs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  near_pair<Q, OP, HELM, CENTRED>(A, pairs[t]);
}

// ---------------------------------------------------------------------------
// Batched mixed precision: persistent launches over the tiles of all items (FP32
// far field; one launch per phase tier PH of the items), then the near-pair lists of
// all items concatenated (item i owns near pairs [first[i], first[i + 1])), one
// thread per pair.
// ---------------------------------------------------------------------------
// Waves per SIMD the persistent tile kernel asks for: 4 (128 VGPRs) where its registers fit
// without spilling (SL/HS, 4-entry mirror segments), otherwise 3.
constexpr int batched_mixed_waves(int op, int ch) { return (op == kSL || op == kHS) && ch == 4 ? 4 : 3; }

template <int Q, int OP, bool HELM, int PH, int CH, int NW = 4>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(batched_mixed_waves(OP, CH))))
batched_p0_mixed_persistent(const RegularArgs* __restrict__ items, const BatchTile* __restrict__ tiles, int ntiles,
                            int* counter, int symmetric) {
  __shared__ double2 lds[NW][CH][kTile];
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();
    if (t >= ntiles) return;
    const BatchTile bt = load_uniform(tiles + t);
    p0_tile_mixed<Q, OP, HELM, PH, CH, NW>(constant_view(items)[bt.item], make_int2(bt.a, bt.b), symmetric, lds);
  }
}

template <int Q, int OP, bool HELM>
__global__ void __launch_bounds__(256) batched_near_kernel(const RegularArgs* __restrict__ items,
                                                           const int* __restrict__ centred,
                                                           const int2* __restrict__ pairs,
                                                           const int* __restrict__ first, int nitems) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= first[nitems]) return;
  int lo = 0, hi = nitems - 1;  // last item with first[item] <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const RegularArgs& A = items[lo];
  if (HELM && centred[lo])
    near_pair<Q, OP, HELM, true>(A, pairs[g]);
  else
    near_pair<Q, OP, HELM, false>(A, pairs[g]);
}

}  // namespace dev

// p1mixed.hip
void launch_mixed_lds(int op, bool helm, int tier, int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& A,
                      const int2* tiles, int sym, size_t lds, const int2* near, const int4* near_segs, int nsegs);

}  // namespace bempp_amd
