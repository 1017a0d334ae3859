// device.hip -- device-side runtime of the assembly engine: per-grid geometry
// caches in HBM, rule tables, kernel launches, and the host entry points
// (dense into device or host memory, singular COO).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "core.hpp"
#include "kernels.hpp"
#include "kernels_mixed.hpp"

namespace bempp_amd {

#define HIP_CHECK(expr)                                                                           \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) throw Error(std::string(#expr) + " failed: " + hipGetErrorString(_e)); \
  } while (0)

// ---------------------------------------------------------------------------
// Device caches
// ---------------------------------------------------------------------------
struct DeviceGridCache {
  int device = -1;
  double* geo = nullptr;          // [nc][kGeoStride]
  int* vid = nullptr;             // [nc][4]
  std::map<int, double*> pts;     // regular point count -> [nc][Q][4]
  int4* pairs = nullptr;          // adjacent pairs
  int npairs = -1;
  std::vector<AdjacentPair> host_pairs;
  double max_radius = 0.0;        // largest centroid-to-vertex distance of any cell
  std::map<int, float*> pts32;    // regular point count -> [nc][Q][4] FP32 offsets from the centroid, w*jdet
  std::map<int, float*> vtx32;    // regular point count -> [nc][3][4] FP32 vertex offsets, class weights
  std::map<int, double*> vtx64;   // the same in FP64
  int2* near = nullptr;           // non-adjacent pairs the mixed-precision far test rejects,
  int nnear = -1;                 // grouped by (test block, trial block) of 64 cells
  int4* near_segs = nullptr;      // per block pair with near pairs: test block, trial block, first, count
  int nsegs = 0;
  unsigned char* far_tiles = nullptr;  // [nblocks][nblocks]: every cell pair of the tile is far
  std::vector<unsigned char> far_tiles_host;
  int far_tiles_ld = 0;
};

struct DeviceSpaceCache {
  int device = -1;
  int* dofs = nullptr;  // [nc][nb] int32
  // 64-cell blocks (tiled kernel with a P1 side), built on first use
  int* blk_off = nullptr;   // [nblocks + 1]
  int* blk_dofs = nullptr;  // unique DOFs per block, ascending
  int* ldof = nullptr;      // [nc][nb] position in the block's list
  int max_nd = -1;          // largest block DOF count
};

void free_device_grid_cache(DeviceGridCache* c) {
  if (!c) return;
  int cur = 0;
  if (hipGetDevice(&cur) == hipSuccess && hipSetDevice(c->device) == hipSuccess) {
    (void)hipFree(c->geo);
    (void)hipFree(c->vid);
    for (auto& kv : c->pts) (void)hipFree(kv.second);
    for (auto& kv : c->pts32) (void)hipFree(kv.second);
    for (auto& kv : c->vtx32) (void)hipFree(kv.second);
    for (auto& kv : c->vtx64) (void)hipFree(kv.second);
    (void)hipFree(c->near);
    (void)hipFree(c->near_segs);
    (void)hipFree(c->far_tiles);
    (void)hipFree(c->pairs);
    (void)hipSetDevice(cur);
  }
  delete c;
}

void free_device_space_cache(DeviceSpaceCache* c) {
  if (!c) return;
  int cur = 0;
  if (hipGetDevice(&cur) == hipSuccess && hipSetDevice(c->device) == hipSuccess) {
    (void)hipFree(c->dofs);
    if (c->blk_off) (void)hipFree(c->blk_off);
    if (c->blk_dofs) (void)hipFree(c->blk_dofs);
    if (c->ldof) (void)hipFree(c->ldof);
    (void)hipSetDevice(cur);
  }
  delete c;
}

int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

namespace {

int current_device() {
  int n = device_count();
  if (n <= 0) throw Error("no HIP device available: assembly requires an MI355X GPU (there is no CPU fallback)");
  int d = 0;
  HIP_CHECK(hipGetDevice(&d));
  return d;
}

// Set while an entry point that must not allocate device memory gathers its inputs:
// a cache that is not built yet then raises instead of being uploaded.
thread_local int t_no_device_alloc = 0;

// Uninitialised device array of n elements (at least one).
template <typename T>
T* device_array(size_t n) {
  if (t_no_device_alloc)
    throw Error("this entry point allocates no device memory, but the device data of this grid, space or rule "
                "is not built yet: make one ordinary call with the same assembler and spaces first");
  T* d = nullptr;
  HIP_CHECK(hipMalloc((void**)&d, std::max<size_t>(n, 1) * sizeof(T)));
  return d;
}

template <typename T>
T* upload(const std::vector<T>& h) {
  if (t_no_device_alloc)
    throw Error("this entry point allocates no device memory, but the device data of this grid, space or rule "
                "is not built yet: make one ordinary call with the same assembler and spaces first");
  T* d = nullptr;
  HIP_CHECK(hipMalloc((void**)&d, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// Per-cell geometry record: vertices, unit normal, jdet = |J0 x J1|, and
// curlJ_b = J1 d0(phi_b) - J0 d1(phi_b) for the P1 basis (constant on a flat cell).
void cell_record(const Grid& g, int64_t c, double* r) {
  const double* v[3];
  for (int m = 0; m < 3; ++m) v[m] = &g.points[3 * g.cells[3 * c + m]];
  double J0[3], J1[3];
  for (int d = 0; d < 3; ++d) {
    r[d] = v[0][d];
    r[3 + d] = v[1][d];
    r[6 + d] = v[2][d];
    J0[d] = v[1][d] - v[0][d];
    J1[d] = v[2][d] - v[0][d];
  }
  const double n[3] = {J0[1] * J1[2] - J0[2] * J1[1], J0[2] * J1[0] - J0[0] * J1[2], J0[0] * J1[1] - J0[1] * J1[0]};
  const double jdet = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (!(jdet > 0.0)) throw Error("cell " + std::to_string(c) + " is degenerate (zero area)");
  for (int d = 0; d < 3; ++d) {
    r[9 + d] = n[d] / jdet;
    r[13 + d] = J0[d] - J1[d];  // basis 0: grad (-1,-1)
    r[16 + d] = J1[d];          // basis 1: grad (1,0)
    r[19 + d] = -J0[d];         // basis 2: grad (0,1)
  }
  r[12] = jdet;
  r[22] = r[23] = 0.0;
  double rad2 = 0.0;
  for (int d = 0; d < 3; ++d) r[dev::kGeoCentroid + d] = (v[0][d] + v[1][d] + v[2][d]) / 3.0;
  for (int m = 0; m < 3; ++m) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d) s += (v[m][d] - r[dev::kGeoCentroid + d]) * (v[m][d] - r[dev::kGeoCentroid + d]);
    rad2 = std::max(rad2, s);
  }
  // slightly enlarged so that rounding never lets a quadrature point escape the bound
  r[dev::kGeoRadius] = std::sqrt(rad2) * (1.0 + 1e-12);
  for (int m = dev::kGeoRadius + 1; m < dev::kGeoStride; ++m) r[m] = 0.0;
}

DeviceGridCache* grid_cache(Grid& g, int device) {
  std::lock_guard<std::mutex> lk(g.cache_mutex);
  auto it = g.device_cache.find(device);
  if (it != g.device_cache.end()) return it->second;
  auto* c = new DeviceGridCache();
  c->device = device;
  const int64_t nc = g.nc();
  std::vector<double> geo(nc * dev::kGeoStride);
  std::vector<int> vid(nc * 4);
  for (int64_t i = 0; i < nc; ++i) {
    cell_record(g, i, &geo[i * dev::kGeoStride]);
    c->max_radius = std::max(c->max_radius, geo[i * dev::kGeoStride + dev::kGeoRadius]);
    for (int m = 0; m < 3; ++m) vid[4 * i + m] = (int)g.cells[3 * i + m];
    vid[4 * i + 3] = -1;
  }
  try {
    c->geo = upload(geo);
    c->vid = upload(vid);
  } catch (...) {
    free_device_grid_cache(c);
    throw;
  }
  g.device_cache[device] = c;
  return c;
}

const double* grid_points_for_rule(Grid& g, DeviceGridCache* c, int Q, const std::vector<double>& xi,
                                   const std::vector<double>& w) {
  std::lock_guard<std::mutex> lk(g.cache_mutex);
  auto it = c->pts.find(Q);
  if (it != c->pts.end()) return it->second;
  const int64_t nc = g.nc();
  std::vector<double> pts(nc * Q * 4);
  for (int64_t i = 0; i < nc; ++i) {
    const double* v0 = &g.points[3 * g.cells[3 * i]];
    const double* v1 = &g.points[3 * g.cells[3 * i + 1]];
    const double* v2 = &g.points[3 * g.cells[3 * i + 2]];
    double J0[3], J1[3];
    for (int d = 0; d < 3; ++d) {
      J0[d] = v1[d] - v0[d];
      J1[d] = v2[d] - v0[d];
    }
    const double cr[3] = {J0[1] * J1[2] - J0[2] * J1[1], J0[2] * J1[0] - J0[0] * J1[2],
                          J0[0] * J1[1] - J0[1] * J1[0]};
    const double jdet = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    for (int p = 0; p < Q; ++p) {
      double* o = &pts[(i * Q + p) * 4];
      for (int d = 0; d < 3; ++d) o[d] = v0[d] + J0[d] * xi[2 * p] + J1[d] * xi[2 * p + 1];
      o[3] = w[p] * jdet;
    }
  }
  double* d = upload(pts);
  c->pts[Q] = d;
  return d;
}

// FP32 copies of the regular points as offsets from the cell centroid (mixed precision).
const float* grid_points32_for_rule(Grid& g, DeviceGridCache* c, int Q, const std::vector<double>& xi,
                                    const std::vector<double>& w) {
  std::lock_guard<std::mutex> lk(g.cache_mutex);
  auto it = c->pts32.find(Q);
  if (it != c->pts32.end()) return it->second;
  const int64_t nc = g.nc();
  std::vector<float> p32(nc * Q * 4);
  for (int64_t i = 0; i < nc; ++i) {
    const double* v0 = &g.points[3 * g.cells[3 * i]];
    const double* v1 = &g.points[3 * g.cells[3 * i + 1]];
    const double* v2 = &g.points[3 * g.cells[3 * i + 2]];
    double J0[3], J1[3], cen[3];
    for (int d = 0; d < 3; ++d) {
      J0[d] = v1[d] - v0[d];
      J1[d] = v2[d] - v0[d];
      cen[d] = (v0[d] + v1[d] + v2[d]) / 3.0;
    }
    const double cr[3] = {J0[1] * J1[2] - J0[2] * J1[1], J0[2] * J1[0] - J0[0] * J1[2],
                          J0[0] * J1[1] - J0[1] * J1[0]};
    const double jdet = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    for (int p = 0; p < Q; ++p) {
      float* o = &p32[(i * Q + p) * 4];
      for (int d = 0; d < 3; ++d) o[d] = (float)((v0[d] + J0[d] * xi[2 * p] + J1[d] * xi[2 * p + 1]) - cen[d]);
      o[3] = (float)(w[p] * jdet);
    }
  }
  float* d = upload(p32);
  c->pts32[Q] = d;
  return d;
}

// Per cell, for the P0 tiles (rule points in orbit form c + t (V_k - c)): d_0 = V_0 - c,
// |d_0|^2; d_1 = V_1 - c, |d_1|^2; the class weights w * jdet, |d_2|^2 (12 values of type T).
template <typename T>
std::vector<T> vertex_records(const Grid& g, int Q, const std::vector<double>& w) {
  const int64_t nc = g.nc();
  std::vector<T> out(nc * 12, T(0));
  for (int64_t i = 0; i < nc; ++i) {
    const double* v[3] = {&g.points[3 * g.cells[3 * i]], &g.points[3 * g.cells[3 * i + 1]],
                          &g.points[3 * g.cells[3 * i + 2]]};
    double cen[3], d[3][3], J0[3], J1[3];
    for (int k = 0; k < 3; ++k) {
      cen[k] = (v[0][k] + v[1][k] + v[2][k]) / 3.0;
      J0[k] = v[1][k] - v[0][k];
      J1[k] = v[2][k] - v[0][k];
    }
    for (int m = 0; m < 3; ++m)
      for (int k = 0; k < 3; ++k) d[m][k] = v[m][k] - cen[k];
    const double cr[3] = {J0[1] * J1[2] - J0[2] * J1[1], J0[2] * J1[0] - J0[0] * J1[2],
                          J0[0] * J1[1] - J0[1] * J1[0]};
    const double jdet = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    T* o = &out[i * 12];
    for (int m = 0; m < 2; ++m)
      for (int k = 0; k < 3; ++k) o[4 * m + k] = (T)d[m][k];
    for (int m = 0; m < 3; ++m)
      o[m == 2 ? 11 : 4 * m + 3] = (T)(d[m][0] * d[m][0] + d[m][1] * d[m][1] + d[m][2] * d[m][2]);
    for (int cl = 0; cl < dev::weight_classes(Q); ++cl) o[8 + cl] = (T)(w[dev::class_point(Q, cl)] * jdet);
  }
  return out;
}

const float* grid_vertex32_for_rule(Grid& g, DeviceGridCache* c, int Q, const std::vector<double>& w) {
  std::lock_guard<std::mutex> lk(g.cache_mutex);
  auto it = c->vtx32.find(Q);
  if (it != c->vtx32.end()) return it->second;
  float* d = upload(vertex_records<float>(g, Q, w));
  c->vtx32[Q] = d;
  return d;
}

const double* grid_vertex64_for_rule(Grid& g, DeviceGridCache* c, int Q, const std::vector<double>& w) {
  std::lock_guard<std::mutex> lk(g.cache_mutex);
  auto it = c->vtx64.find(Q);
  if (it != c->vtx64.end()) return it->second;
  double* d = upload(vertex_records<double>(g, Q, w));
  c->vtx64[Q] = d;
  return d;
}

// True when every point of the rule is c + t (V_k - c) with the tabulated t, k (orbit_t,
// orbit_vertex) and the weights follow the weight classes.
bool rule_in_orbit_form(int Q, const std::vector<double>& xi, const std::vector<double>& w) {
  for (int p = 0; p < Q; ++p) {
    if (w[p] != w[dev::class_point(Q, dev::weight_class(Q, p))]) return false;
    const double lam[3] = {1.0 - xi[2 * p] - xi[2 * p + 1], xi[2 * p], xi[2 * p + 1]};
    const double t = dev::orbit_t(Q, p);
    const int kv = dev::orbit_vertex(Q, p);
    for (int m = 0; m < 3; ++m)
      if (std::fabs(lam[m] - (m == kv ? (1.0 + 2.0 * t) / 3.0 : (1.0 - t) / 3.0)) > 1e-15) return false;
  }
  return true;
}

// Tiles (64-cell blocks a, b) whose cell pairs are all far for the mixed-precision split:
// with block centre C (mean centroid), rho (largest centroid distance from C) and rmax
// (largest cell radius), every pair has R >= |C_a - C_b| - rho_a - rho_b; the tile is
// flagged when that bound exceeds the far limit kNearFactor (rmax_a + rmax_b) by 0.1 %, far
// beyond the rounding of the per-pair test it replaces (no flagged pair can be near or
// adjacent, adjacent cells being closer than rad_i + rad_j).
const unsigned char* grid_far_tiles(Grid& g, DeviceGridCache* c, int* ld) {
  std::lock_guard<std::mutex> lk(g.cache_mutex);
  if (!c->far_tiles) {
    const int64_t nc = g.nc();
    const int nb = (int)((nc + dev::kTile - 1) / dev::kTile);
    std::vector<double> cen(3 * nc), rad(nc), bc(3 * nb, 0.0), rho(nb, 0.0), rmax(nb, 0.0);
    std::vector<double> rec(dev::kGeoStride);
    for (int64_t i = 0; i < nc; ++i) {
      cell_record(g, i, rec.data());
      for (int d = 0; d < 3; ++d) cen[3 * i + d] = rec[dev::kGeoCentroid + d];
      rad[i] = rec[dev::kGeoRadius];
      const int b = (int)(i / dev::kTile);
      for (int d = 0; d < 3; ++d) bc[3 * b + d] += cen[3 * i + d];
      rmax[b] = std::max(rmax[b], rad[i]);
    }
    for (int b = 0; b < nb; ++b) {
      const int64_t n = std::min<int64_t>(dev::kTile, nc - (int64_t)b * dev::kTile);
      for (int d = 0; d < 3; ++d) bc[3 * b + d] /= (double)n;
    }
    for (int64_t i = 0; i < nc; ++i) {
      const int b = (int)(i / dev::kTile);
      double s = 0.0;
      for (int d = 0; d < 3; ++d) s += (cen[3 * i + d] - bc[3 * b + d]) * (cen[3 * i + d] - bc[3 * b + d]);
      rho[b] = std::max(rho[b], std::sqrt(s));
    }
    std::vector<unsigned char> flags((size_t)nb * nb, 0);
    for (int a = 0; a < nb; ++a)
      for (int b = 0; b < nb; ++b) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) s += (bc[3 * a + d] - bc[3 * b + d]) * (bc[3 * a + d] - bc[3 * b + d]);
        const double lower = std::sqrt(s) - rho[a] - rho[b];
        flags[(size_t)a * nb + b] = lower >= dev::kNearFactor * 1.001 * (rmax[a] + rmax[b]) ? 1 : 0;
      }
    c->far_tiles = upload(flags);
    c->far_tiles_host = std::move(flags);
    c->far_tiles_ld = nb;
  }
  *ld = c->far_tiles_ld;
  return c->far_tiles;
}

// The mixed-precision near-pair list of a grid (non-adjacent pairs the far test rejects),
// built on the device from the candidate tiles (those not flagged all-far; see
// near_count_kernel / near_fill_kernel), grouped by tile: segs[s] = {test block, trial
// block, first pair, pair count} for each tile with near pairs.
const int2* grid_near_pairs(Grid& g, DeviceGridCache* c, int* n, const int4** segs, int* nsegs) {
  int ld = 0;
  (void)grid_far_tiles(g, c, &ld);
  std::lock_guard<std::mutex> lk(g.cache_mutex);
  if (c->nnear < 0) {
    const int nb = c->far_tiles_ld;
    std::vector<int2> cand;
    for (int a = 0; a < nb; ++a)
      for (int b = 0; b < nb; ++b)
        if (!c->far_tiles_host[(size_t)a * nb + b]) cand.push_back(make_int2(a, b));
    std::vector<int> counts(cand.size()), offsets(cand.size());
    std::vector<int4> sg;
    long long total = 0;
    int2* tiles = nullptr;
    int* dcounts = nullptr;
    int* doffsets = nullptr;
    int2* out = nullptr;
    auto cleanup = [&]() {
      (void)hipFree(tiles);
      (void)hipFree(dcounts);
      (void)hipFree(doffsets);
    };
    try {
      if (!cand.empty()) {
        tiles = upload(cand);
        dcounts = upload(counts);
        dev::NearBuildArgs B{};
        B.geo = c->geo;
        B.vid = c->vid;
        B.n = (int)g.nc();
        B.tiles = tiles;
        B.counts = dcounts;
        dev::near_count_kernel<><<<(unsigned)cand.size(), 256, 0, nullptr>>>(B);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpy(counts.data(), dcounts, counts.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t t = 0; t < cand.size(); ++t) {
          offsets[t] = (int)total;
          if (counts[t] > 0) sg.push_back(make_int4(cand[t].x, cand[t].y, (int)total, counts[t]));
          total += counts[t];
          if (total > INT32_MAX) throw Error("too many near pairs for the mixed-precision split");
        }
        if (total > 0) {
          doffsets = upload(offsets);
          out = device_array<int2>((size_t)total);
          B.offsets = doffsets;
          B.out = out;
          dev::near_fill_kernel<><<<(unsigned)cand.size(), 256, 0, nullptr>>>(B);
          HIP_CHECK(hipGetLastError());
          HIP_CHECK(hipStreamSynchronize(nullptr));
        }
      }
      c->near_segs = upload(sg);
    } catch (...) {
      cleanup();
      (void)hipFree(out);
      throw;
    }
    cleanup();
    c->near = out ? out : device_array<int2>(1);
    c->nnear = (int)total;
    c->nsegs = (int)sg.size();
  }
  *n = c->nnear;
  if (segs) *segs = c->near_segs;
  if (nsegs) *nsegs = c->nsegs;
  return c->near;
}

const int4* grid_pairs(Grid& g, DeviceGridCache* c, int* npairs) {
  std::lock_guard<std::mutex> lk(g.cache_mutex);
  if (c->npairs < 0) {
    adjacent_pairs(g, c->host_pairs);
    // group by adjacency type: consecutive wavefronts then read the same rule table
    std::stable_sort(c->host_pairs.begin(), c->host_pairs.end(),
                     [](const AdjacentPair& x, const AdjacentPair& y) { return x.nshared < y.nshared; });
    std::vector<int4> p(c->host_pairs.size());
    for (size_t i = 0; i < p.size(); ++i) {
      const AdjacentPair& a = c->host_pairs[i];
      p[i] = make_int4(a.test, a.trial, a.nshared, a.perm);
    }
    c->pairs = upload(p);
    c->npairs = (int)p.size();
  }
  *npairs = c->npairs;
  return c->pairs;
}

const int* space_dofs(FunctionSpace& s, int device) {
  std::lock_guard<std::mutex> lk(s.cache_mutex);
  auto it = s.device_cache.find(device);
  if (it != s.device_cache.end()) return it->second->dofs;
  if (s.size > INT32_MAX) throw Error("function space too large for 32-bit DOF indices");
  std::vector<int> d(s.dofs.begin(), s.dofs.end());
  auto* c = new DeviceSpaceCache();
  c->device = device;
  c->dofs = upload(d);
  s.device_cache[device] = c;
  return c->dofs;
}

// Block DOF lists of a space for the tiled kernel with a P1 side: cells [64 t, 64 t + 64)
// form block t (consecutive cells of a refined mesh are a compact patch).
const DeviceSpaceCache* space_blocks(FunctionSpace& s, int device) {
  space_dofs(s, device);
  std::lock_guard<std::mutex> lk(s.cache_mutex);
  DeviceSpaceCache* c = s.device_cache.at(device);
  if (c->max_nd >= 0) return c;
  const int64_t nc = s.grid->nc();
  const int nb = s.nb;
  const int64_t nblk = (nc + dev::kTile - 1) / dev::kTile;
  std::vector<int> off(nblk + 1, 0), bd, ldof((size_t)nc * nb);
  int max_nd = 0;
  std::vector<int64_t> tmp;
  for (int64_t t = 0; t < nblk; ++t) {
    const int64_t c0 = t * dev::kTile, c1 = std::min(nc, c0 + dev::kTile);
    tmp.assign(s.dofs.begin() + c0 * nb, s.dofs.begin() + c1 * nb);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    for (int64_t cc = c0; cc < c1; ++cc)
      for (int a = 0; a < nb; ++a)
        ldof[cc * nb + a] = (int)(std::lower_bound(tmp.begin(), tmp.end(), s.dofs[cc * nb + a]) - tmp.begin());
    for (int64_t d : tmp) bd.push_back((int)d);
    off[t + 1] = (int)bd.size();
    max_nd = std::max(max_nd, (int)tmp.size());
  }
  c->blk_off = upload(off);
  c->blk_dofs = upload(bd);
  c->ldof = upload(ldof);
  c->max_nd = max_nd;
  return c;
}

// Canonical Sauter-Schwab rules for one Gauss order n, per device (kept for the
// life of the process: a few hundred KB).
struct RuleTable {
  double* d = nullptr;    // 5 arrays of total points
  double* aos = nullptr;  // one record of 6 per point (a0 a1 b0 b1 w pad)
  int total = 0;
  int off[4] = {0, 0, 0, 0};
  int n[4] = {0, 0, 0, 0};
};

const RuleTable& rule_table(int device, int n) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, RuleTable> tables;
  std::lock_guard<std::mutex> lk(mu);
  auto it = tables.find({device, n});
  if (it != tables.end()) return it->second;
  RuleTable t;
  std::vector<double> parts[4];
  for (int m = 1; m <= 3; ++m) {
    canonical_pair_rule(m, n, parts[m]);
    t.off[m] = t.total;
    t.n[m] = (int)(parts[m].size() / 5);
    t.total += t.n[m];
  }
  std::vector<double> soa(5 * (size_t)t.total);
  for (int m = 1; m <= 3; ++m)
    for (int i = 0; i < t.n[m]; ++i)
      for (int f = 0; f < 5; ++f) soa[(size_t)f * t.total + t.off[m] + i] = parts[m][5 * i + f];
  t.d = upload(soa);
  std::vector<double> aos(6 * (size_t)t.total, 0.0);
  for (int m = 1; m <= 3; ++m)
    for (int i = 0; i < t.n[m]; ++i)
      for (int f = 0; f < 5; ++f) aos[6 * (size_t)(t.off[m] + i) + f] = parts[m][5 * i + f];
  t.aos = upload(aos);
  return tables.emplace(std::make_pair(device, n), t).first->second;
}

// Regular rule as a device table [Q][4] = barycentric coordinates (the P1 basis values)
// and weight of each point, per device and point count; read by the kernels that take
// the point count at run time.
const double* regular_rule_table(int device, int Q, const std::vector<double>& xi, const std::vector<double>& w) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, double*> tables;
  std::lock_guard<std::mutex> lk(mu);
  auto it = tables.find({device, Q});
  if (it != tables.end()) return it->second;
  std::vector<double> t(4 * (size_t)Q);
  for (int p = 0; p < Q; ++p) {
    t[4 * p] = 1.0 - xi[2 * p] - xi[2 * p + 1];
    t[4 * p + 1] = xi[2 * p];
    t[4 * p + 2] = xi[2 * p + 1];
    t[4 * p + 3] = w[p];
  }
  double* d = upload(t);
  tables[{device, Q}] = d;
  return d;
}

// Template point count of the kernels for a rule of Q points: Q itself for the rules
// compiled with their point count (held in registers), 0 (point count at run time) otherwise.
int kernel_q(int Q) { return (Q == 1 || Q == 3 || Q == 6 || Q == 7) ? Q : 0; }

hipStream_t assembler_stream(const Assembler& a, int device) {
  auto& m = const_cast<Assembler&>(a);
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (!m.stream || m.device != device) {
    if (m.stream) {
      int cur = 0;
      (void)hipGetDevice(&cur);
      (void)hipSetDevice(m.device);
      (void)hipStreamDestroy((hipStream_t)m.stream);
      (void)hipSetDevice(cur);
    }
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    m.stream = s;
    m.device = device;
  }
  return (hipStream_t)m.stream;
}

int out_type(DType d) {
  switch (d) {
    case kF64: return dev::kOutF64;
    case kC64: return dev::kOutC128;
    case kF32: return dev::kOutF32;
    default: return dev::kOutC64;
  }
}

int regular_points(const Options& o) {
  auto it = o.quadrature_degrees.find(kTriangle);
  if (it == o.quadrature_degrees.end()) throw Error("no regular quadrature degree set for triangles");
  return (int)it->second;
}

int singular_points(const Options& o) {
  auto it = o.singular_quadrature_degrees.find({kTriangle, kTriangle});
  if (it == o.singular_quadrature_degrees.end()) throw Error("no singular quadrature degree set for triangle pairs");
  return (int)it->second;
}

// Dynamic LDS budget of the LDS-accumulating tile kernel (block DOF lists of a compact
// 64-cell P1 patch: ~45 x 45 x 16 B = 32 KB).
constexpr size_t kMaxTileLds = 64 * 1024;

// Tile coordinates (tile row = test block, tile column = trial block) for the
// P0 x P0 kernel: all tiles, or the upper triangle a <= b when symmetric.
// Cached per device and shape for the life of the process.
const int2* tile_list(int device, int nta, int ntb, bool symmetric, int* count) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, bool>, std::pair<int2*, int>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(device, nta, ntb, symmetric);
  auto it = cache.find(key);
  if (it == cache.end()) {
    std::vector<int2> t;
    for (int b = 0; b < ntb; ++b)
      for (int a = 0; a < (symmetric ? std::min(b + 1, nta) : nta); ++a) t.push_back(make_int2(a, b));
    it = cache.emplace(key, std::make_pair(upload(t), (int)t.size())).first;
  }
  *count = it->second.second;
  return it->second.first;
}

// ---- kernel dispatch: tiled P0 x P0 or general, by operator, kernel and phase mode ----
// CENTRED: phase tier (0: full reduction, 1: |k (r - R)| <= 0.5, 2: <= 0.25; tier 2
// for the tile kernels, the general kernel runs tier 1 there).
// mode: 0 general kernel (atomics per cell pair), 1 P0 x P0 tiles, 2 tiles with LDS
// accumulation (a P1 side; lds = dynamic LDS bytes).
template <int OP, bool HELM, int CENTRED>
void launch_q(int mode, int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& A, const int2* tiles, int sym,
              size_t lds) {
  constexpr int GC = CENTRED > 1 ? 1 : CENTRED;
#define BEMPP_LAUNCH(QQ)                                                                                \
  if (mode == 1) dev::regular_p0_tiled<QQ, OP, HELM, CENTRED><<<grid, 256, 0, s>>>(A, tiles, sym);       \
  else if (mode == 2) dev::regular_tiled_lds<QQ, OP, HELM, CENTRED><<<grid, 256, lds, s>>>(A, tiles, sym); \
  else dev::regular_kernel<QQ, OP, HELM, GC><<<grid, 256, 0, s>>>(A);
  switch (kernel_q(Q)) {
    case 1: BEMPP_LAUNCH(1); break;
    case 3: BEMPP_LAUNCH(3); break;
    case 6: BEMPP_LAUNCH(6); break;
    case 7: BEMPP_LAUNCH(7); break;
    default: BEMPP_LAUNCH(0); break;  // point count at run time
  }
#undef BEMPP_LAUNCH
}

template <bool HELM, int CENTRED>
void launch_op(int op, int mode, int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& A, const int2* tiles,
               int sym, size_t lds) {
  switch (op) {
    case kSL: launch_q<kSL, HELM, CENTRED>(mode, Q, grid, s, A, tiles, sym, lds); break;
    case kDL: launch_q<kDL, HELM, CENTRED>(mode, Q, grid, s, A, tiles, sym, lds); break;
    case kADL: launch_q<kADL, HELM, CENTRED>(mode, Q, grid, s, A, tiles, sym, lds); break;
    case kHS: launch_q<kHS, HELM, CENTRED>(mode, Q, grid, s, A, tiles, sym, lds); break;
    default: throw Error("operator not supported by the dense assembler");
  }
}

// Regular pairs: mode 1 the P0 x P0 tile kernel, mode 2 the LDS-accumulating tile kernel
// (grid = tile count), mode 0 the general kernel (grid = test blocks x trial chunks).
void launch_regular(int op, bool helm, int tier, int mode, int Q, dim3 grid, hipStream_t s,
                    const dev::RegularArgs& A, const int2* tiles = nullptr, int sym = 0, size_t lds = 0) {
  if (!helm) launch_op<false, 0>(op, mode, Q, grid, s, A, tiles, sym, lds);
  else if (tier == 2 && (mode == 1 || mode == 2)) launch_op<true, 2>(op, mode, Q, grid, s, A, tiles, sym, lds);
  else if (tier >= 1) launch_op<true, 1>(op, mode, Q, grid, s, A, tiles, sym, lds);
  else launch_op<true, 0>(op, mode, Q, grid, s, A, tiles, sym, lds);
  HIP_CHECK(hipGetLastError());
}

// Mixed precision (P0 x P0, one grid): FP32 far-pair tiles, then the FP64 near-pair list.
// TIER: phase tier of the cell pairs (0: range-reduced FP32 phase and FP64 near pairs
// without centring; 1, 2: centred small-argument tiers).
template <int OP, bool HELM, int TIER>
void launch_mixed_q(int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& A, const int2* tiles, int sym,
                    const int2* near, int nnear) {
  dim3 ng((unsigned)((nnear + 255) / 256));
  constexpr int NC = HELM && TIER > 0 ? 1 : 0;  // the near-pair kernel runs tier 1 when centred
#define BEMPP_LAUNCH(QQ)                                                                      \
  dev::regular_p0_tiled_mixed<QQ, OP, HELM, TIER><<<grid, 256, 0, s>>>(A, tiles, sym);        \
  if (nnear > 0) dev::near_pair_kernel<QQ, OP, HELM, NC><<<ng, 256, 0, s>>>(A, near, nnear);
  switch (Q) {
    case 1: BEMPP_LAUNCH(1); break;
    case 3: BEMPP_LAUNCH(3); break;
    case 6: BEMPP_LAUNCH(6); break;
    case 7: BEMPP_LAUNCH(7); break;
    default: throw Error("unsupported regular quadrature point count");
  }
#undef BEMPP_LAUNCH
}

template <bool HELM, int TIER>
void launch_mixed_op(int op, int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& A, const int2* tiles, int sym,
                     const int2* near, int nnear) {
  switch (op) {
    case kSL: launch_mixed_q<kSL, HELM, TIER>(Q, grid, s, A, tiles, sym, near, nnear); break;
    case kDL: launch_mixed_q<kDL, HELM, TIER>(Q, grid, s, A, tiles, sym, near, nnear); break;
    case kADL: launch_mixed_q<kADL, HELM, TIER>(Q, grid, s, A, tiles, sym, near, nnear); break;
    case kHS: launch_mixed_q<kHS, HELM, TIER>(Q, grid, s, A, tiles, sym, near, nnear); break;
    default: throw Error("operator not supported by the dense assembler");
  }
}

void launch_mixed(int op, bool helm, int tier, int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& A,
                  const int2* tiles, int sym, const int2* near, int nnear) {
  if (!helm) launch_mixed_op<false, 1>(op, Q, grid, s, A, tiles, sym, near, nnear);
  else if (tier == 2) launch_mixed_op<true, 2>(op, Q, grid, s, A, tiles, sym, near, nnear);
  else if (tier == 1) launch_mixed_op<true, 1>(op, Q, grid, s, A, tiles, sym, near, nnear);
  else launch_mixed_op<true, 0>(op, Q, grid, s, A, tiles, sym, near, nnear);
  HIP_CHECK(hipGetLastError());
}

template <bool HELM, bool P0P0, bool SMALL>
void launch_singular_t(int op, dim3 grid, hipStream_t s, const dev::SingularArgs& A) {
  switch (op) {
    case kSL: dev::singular_kernel<kSL, HELM, P0P0, SMALL><<<grid, 256, 0, s>>>(A); break;
    case kDL: dev::singular_kernel<kDL, HELM, P0P0, SMALL><<<grid, 256, 0, s>>>(A); break;
    case kADL: dev::singular_kernel<kADL, HELM, P0P0, SMALL><<<grid, 256, 0, s>>>(A); break;
    case kHS: dev::singular_kernel<kHS, HELM, P0P0, SMALL><<<grid, 256, 0, s>>>(A); break;
    default: throw Error("operator not supported by the dense assembler");
  }
}

void launch_singular(int op, bool helm, bool p0p0, bool small, dim3 grid, hipStream_t s,
                     const dev::SingularArgs& A) {
  if (!helm) {
    if (p0p0) launch_singular_t<false, true, false>(op, grid, s, A);
    else launch_singular_t<false, false, false>(op, grid, s, A);
  } else if (small) {
    if (p0p0) launch_singular_t<true, true, true>(op, grid, s, A);
    else launch_singular_t<true, false, true>(op, grid, s, A);
  } else {
    if (p0p0) launch_singular_t<true, true, false>(op, grid, s, A);
    else launch_singular_t<true, false, false>(op, grid, s, A);
  }
  HIP_CHECK(hipGetLastError());
}

}  // namespace

NoDeviceAllocScope::NoDeviceAllocScope() { ++t_no_device_alloc; }
NoDeviceAllocScope::~NoDeviceAllocScope() { --t_no_device_alloc; }

void check_assembler(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test) {
  if (a.op != kSL && a.op != kDL && a.op != kADL && a.op != kHS)
    throw Error("operator type not supported for dense assembly");
  if ((a.kind == kLaplace) == is_complex(a.dtype)) throw Error("dtype does not match the kernel");
  (void)trial;
  (void)test;
}

Prepared prepare(const Assembler& a, const FunctionSpace& trial_c, const FunctionSpace& test_c, void* out,
                 void* stream, bool caller_stream) {
  check_assembler(a, trial_c, test_c);
  require_triangles(test_c, "this entry point");
  require_triangles(trial_c, "this entry point");
  auto& trial = const_cast<FunctionSpace&>(trial_c);
  auto& test = const_cast<FunctionSpace&>(test_c);
  auto& tgrid = const_cast<Grid&>(*test.grid);
  auto& sgrid = const_cast<Grid&>(*trial.grid);
  Prepared P{};
  P.device = current_device();
  P.stream = (stream || caller_stream) ? (hipStream_t)stream : assembler_stream(a, P.device);
  P.Q = regular_points(a.options);
  std::vector<double> xi, w;
  triangle_rule(P.Q, xi, w);
  const int nsing = singular_points(a.options);
  if (nsing < 1 || nsing > 64) throw Error("singular quadrature degree must be in [1, 64]");
  P.same_grid = test.grid == trial.grid;
  // the device sine/cosine reduces k r exactly only while k r <= 1e8
  if (a.kind == kHelmholtz) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (const Grid* g : {test.grid, trial.grid})
      for (int64_t v = 0; v < g->nv(); ++v)
        for (int d = 0; d < 3; ++d) {
          lo[d] = std::min(lo[d], g->points[3 * v + d]);
          hi[d] = std::max(hi[d], g->points[3 * v + d]);
        }
    double diag = 0.0;
    for (int d = 0; d < 3; ++d) diag += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    if (!(std::fabs(a.k) * std::sqrt(diag) <= 1e8))
      throw Error("|k| times the mesh diameter exceeds 1e8 (or is not finite)");
  }

  DeviceGridCache* tc = grid_cache(tgrid, P.device);
  DeviceGridCache* sc = P.same_grid ? tc : grid_cache(sgrid, P.device);
  // |k (r - R)| <= |k| (rad_test + rad_trial) for every quadrature-point pair of a
  // regular cell pair; the small-argument sine/cosine needs this bound <= 0.5.
  const double phase_bound = std::fabs(a.k) * (tc->max_radius + sc->max_radius);
  // (|k| below 1e-100, k = 0 included, takes the full reduction: the centred tile sums work
  // in lengths scaled by |k|)
  P.centred = a.kind == kHelmholtz && std::fabs(a.k) >= 1e-100 && phase_bound <= 0.5;
  P.tier = !P.centred ? 0 : (phase_bound <= 0.25 ? 2 : 1);
  // points of two cells sharing a vertex are at most 2 (rad_i + rad_j) apart
  P.small_adjacent = a.kind == kHelmholtz && P.same_grid && std::fabs(a.k) * 4.0 * tc->max_radius <= 0.5;
  dev::RegularArgs& R = P.R;
  R.test_pts = grid_points_for_rule(tgrid, tc, P.Q, xi, w);
  R.trial_pts = grid_points_for_rule(sgrid, sc, P.Q, xi, w);
  R.nq = P.Q;
  R.rule_tab = regular_rule_table(P.device, P.Q, xi, w);
  // the FP64 P0 x P0 tiles compiled for a point count take both rules' points in orbit
  // form; other rules run the tiles with the point count at run time
  const bool compiled_q = kernel_q(P.Q) != 0;
  R.orbit_form = compiled_q && rule_in_orbit_form(P.Q, xi, w) ? 1 : 0;
  if (compiled_q && !R.orbit_form && test.nb == 1 && trial.nb == 1)
    throw Error("P0 tiles: quadrature rule not in the tabulated orbit form");
  R.test_vtx64 = R.orbit_form ? grid_vertex64_for_rule(tgrid, tc, P.Q, w) : nullptr;
  R.trial_vtx64 = R.orbit_form ? grid_vertex64_for_rule(sgrid, sc, P.Q, w) : nullptr;
  // mixed precision: one grid (P0 x P0: p0 mixed tiles; a P1 side: LDS mixed tiles), rules
  // compiled for their point count; other rules assemble in FP64. The Laplace hypersingular
  // operator on P1 x P1 also stays in FP64: its cell-pair sum is the single-layer sum, which
  // the FP64 tiles take in orbit form, and with the P1 scatter around it that runs faster
  // than the FP32 point pairs of the mixed tiles (FP32 without packed math issues at the
  // FP64 rate)
  const bool laplace_hs_p1 = a.kind == kLaplace && a.op == kHS && test.nb == 3 && trial.nb == 3 && R.orbit_form;
  P.mixed = a.options.precision == kMixed && P.same_grid && compiled_q && !laplace_hs_p1;
  P.near = nullptr;
  P.nnear = 0;
  P.near_segs = nullptr;
  P.nsegs = 0;
  R.test_pts32 = R.trial_pts32 = nullptr;
  R.trial_vtx32 = nullptr;
  R.far_tiles = nullptr;
  R.far_tiles_ld = 0;
  if (P.mixed) {
    // the mixed P0 tiles sum point pairs per weight class and take the trial points in
    // orbit form (kernels_mixed.hpp)
    if (!rule_in_orbit_form(P.Q, xi, w))
      throw Error("mixed precision: quadrature rule not in the tabulated orbit form");
    R.test_pts32 = R.trial_pts32 = grid_points32_for_rule(tgrid, tc, P.Q, xi, w);
    R.trial_vtx32 = grid_vertex32_for_rule(tgrid, tc, P.Q, w);
    R.far_tiles = grid_far_tiles(tgrid, tc, &R.far_tiles_ld);
    P.near = grid_near_pairs(tgrid, tc, &P.nnear, &P.near_segs, &P.nsegs);
  }
  R.test_geo = tc->geo;
  R.trial_geo = sc->geo;
  R.test_vid = tc->vid;
  R.trial_vid = sc->vid;
  R.test_dofs = space_dofs(test, P.device);
  R.trial_dofs = space_dofs(trial, P.device);
  R.n_test = (int)tgrid.nc();
  R.n_trial = (int)sgrid.nc();
  R.nbt = test.nb;
  R.nbs = trial.nb;
  R.same_grid = P.same_grid ? 1 : 0;
  R.k = a.k;
  for (int p = 0; p < dev::kMaxQ; ++p)
    for (int b = 0; b < 3; ++b) R.phi[p][b] = 0.0;
  for (int p = 0; p < std::min(P.Q, dev::kMaxQ); ++p) {
    R.phi[p][0] = 1.0 - xi[2 * p] - xi[2 * p + 1];
    R.phi[p][1] = xi[2 * p];
    R.phi[p][2] = xi[2 * p + 1];
  }
  for (int p = 0; p < dev::kMaxQ; ++p)
    for (int b = 0; b < 3; ++b) {
      const bool on = p < P.Q;
      R.wphi_test[p][b] = on ? (float)(w[p] * (test.nb == 3 ? R.phi[p][b] : (b == 0 ? 1.0 : 0.0))) : 0.0f;
      R.wphi_trial[p][b] = on ? (float)(w[p] * (trial.nb == 3 ? R.phi[p][b] : (b == 0 ? 1.0 : 0.0))) : 0.0f;
      R.wphi_trial64[p][b] = on ? w[p] * (trial.nb == 3 ? R.phi[p][b] : (b == 0 ? 1.0 : 0.0)) : 0.0;
    }
  R.out = out;
  R.ld = (long long)test.size;
  R.out_type = out_type(a.dtype);

  P.npairs = 0;
  if (P.same_grid) {
    const RuleTable& rt = rule_table(P.device, nsing);
    dev::SingularArgs& S = P.S;
    S.pairs = grid_pairs(tgrid, tc, &P.npairs);
    S.npairs = P.npairs;
    S.rule = rt.d;
    S.rule_aos = rt.aos;
    S.rule_total = rt.total;
    for (int m = 0; m < 4; ++m) {
      S.rule_off[m] = rt.off[m];
      S.rule_n[m] = rt.n[m];
    }
    S.test_geo = tc->geo;
    S.trial_geo = tc->geo;
    S.test_dofs = R.test_dofs;
    S.trial_dofs = R.trial_dofs;
    S.nbt = test.nb;
    S.nbs = trial.nb;
    S.k = a.k;
    S.out = out;
    S.ld = R.ld;
    S.out_type = R.out_type;
  }
  return P;
}

// Geometry, rule tables and adjacent pairs for the matrix-free kernels (gradient.hip).
GradientInputs gradient_inputs(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test,
                               void* stream, bool caller_stream) {
  Prepared P = prepare(a, trial, test, nullptr, stream, caller_stream);
  GradientInputs g{};
  g.R = P.R;
  g.pairs = P.S.pairs;
  g.npairs = P.npairs;
  g.rule = P.S.rule;
  g.rule_total = P.S.rule_total;
  for (int m = 0; m < 4; ++m) {
    g.rule_off[m] = P.S.rule_off[m];
    g.rule_n[m] = P.S.rule_n[m];
  }
  g.centred = P.centred;
  g.Q = P.Q;
  g.device = P.device;
  g.stream = P.stream;
  return g;
}

void assemble_device(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, void* dev_out,
                     bool accumulate, void* stream) {
  if (!dev_out) throw Error("output pointer is null");
  if (test.grid->general() || trial.grid->general()) {
    // grids with quadrilateral cells: the general-cell kernels (FP64 for every precision mode)
    check_assembler(a, trial, test);
    (void)current_device();
    assemble_cells_device(a, trial, test, dev_out, accumulate, stream);
    return;
  }
  Prepared P = prepare(a, trial, test, dev_out, stream);
  P.stream = (hipStream_t)stream;  // device entry point: the caller's stream, NULL = the null stream
  const bool p0p0 = test.nb == 1 && trial.nb == 1;
  const long long n_out = (long long)test.size * (long long)trial.size * (is_complex(a.dtype) ? 2 : 1);
  if (!p0p0 && !accumulate) {
    // a fill kernel, not a memset node: the caller may capture this stream in a graph
    if (a.dtype == kF32 || a.dtype == kC32)
      dev::fill_zero_kernel<<<2048, 256, 0, P.stream>>>((float*)dev_out, n_out);
    else
      dev::fill_zero_kernel<<<2048, 256, 0, P.stream>>>((double*)dev_out, n_out);
    HIP_CHECK(hipGetLastError());
  }
  // regular pairs
  dev::RegularArgs R = P.R;
  R.mode = p0p0 ? (accumulate ? dev::kAccumulate : dev::kStore) : dev::kAtomic;
  const int bx = (R.n_test + 255) / 256;
  // trial chunk: enough blocks to fill 256 CUs several times over
  long long chunk = ((long long)bx * R.n_trial + 16383) / 16384;
  chunk = std::max(4LL, std::min(128LL, chunk));
  chunk = std::max(chunk, ((long long)R.n_trial + 65534) / 65535);
  R.trial_chunk = (int)chunk;
  dim3 grid(bx, (unsigned)((R.n_trial + chunk - 1) / chunk));
  auto& am = const_cast<Assembler&>(a);
  if (a.profiling) {
    for (int e = 0; e < 4; ++e)
      if (!am.events[e]) {
        hipEvent_t ev;
        HIP_CHECK(hipEventCreate(&ev));
        am.events[e] = ev;
      }
    am.events_recorded[0] = am.events_recorded[1] = false;
  }
  if (R.n_test > 0 && R.n_trial > 0) {
    if (a.profiling) HIP_CHECK(hipEventRecord((hipEvent_t)am.events[0], P.stream));
    if (p0p0) {
      // symmetric operators on one space: upper-triangle tiles, mirrored writes
      // (a P0 space is determined by its grid, so same grid means same space)
      const bool sym = (a.op == kSL || a.op == kHS) && P.same_grid;
      const int nta = (R.n_test + dev::kTile - 1) / dev::kTile, ntb = (R.n_trial + dev::kTile - 1) / dev::kTile;
      int ntiles = 0;
      const int2* tiles = tile_list(P.device, nta, ntb, sym, &ntiles);
      if (P.mixed) {
        launch_mixed(a.op, a.kind == kHelmholtz, P.tier, P.Q, dim3(ntiles), P.stream, R, tiles, sym ? 1 : 0,
                     P.near, P.nnear);
      } else {
        launch_regular(a.op, a.kind == kHelmholtz, P.tier, 1, P.Q, dim3(ntiles), P.stream, R, tiles,
                       sym ? 1 : 0);
      }
    } else {
      // a P1 side: tiles with LDS accumulation when the blocks' DOF lists fit in LDS
      const DeviceSpaceCache* tb = space_blocks(const_cast<FunctionSpace&>(test), P.device);
      const DeviceSpaceCache* sb = space_blocks(const_cast<FunctionSpace&>(trial), P.device);
      const size_t lds = (size_t)tb->max_nd * (size_t)sb->max_nd * sizeof(double2);
      // + the flush's static copies of the blocks' DOF numbers (2 x 256 ints)
      if (lds + 2 * 256 * sizeof(int) <= kMaxTileLds) {
        R.test_blk_off = tb->blk_off;
        R.test_blk_dofs = tb->blk_dofs;
        R.test_ldof = tb->ldof;
        R.trial_blk_off = sb->blk_off;
        R.trial_blk_dofs = sb->blk_dofs;
        R.trial_ldof = sb->ldof;
        // symmetric operators on one space (the DOF numbering is fixed by grid and family)
        const bool sym = (a.op == kSL || a.op == kHS) && P.same_grid && test.nb == trial.nb &&
                         test.degree == trial.degree && test.continuous == trial.continuous;
        const int nta = (R.n_test + dev::kTile - 1) / dev::kTile, ntb = (R.n_trial + dev::kTile - 1) / dev::kTile;
        int ntiles = 0;
        const int2* tiles = tile_list(P.device, nta, ntb, sym, &ntiles);
        if (P.mixed)
          launch_mixed_lds(a.op, a.kind == kHelmholtz, P.tier, P.Q, dim3(ntiles), P.stream, R, tiles, sym ? 1 : 0,
                           lds, P.near, P.near_segs, P.nsegs);
        else
          launch_regular(a.op, a.kind == kHelmholtz, P.tier, 2, P.Q, dim3(ntiles), P.stream, R, tiles,
                         sym ? 1 : 0, lds);
      } else {
        launch_regular(a.op, a.kind == kHelmholtz, P.tier, 0, P.Q, grid, P.stream, R);
      }
    }
    if (a.profiling) {
      HIP_CHECK(hipEventRecord((hipEvent_t)am.events[1], P.stream));
      am.events_recorded[0] = true;
    }
  }
  // adjacent pairs
  if (P.npairs > 0) {
    dev::SingularArgs S = P.S;
    S.mode = p0p0 ? dev::kAccumulate : dev::kAtomic;
    dim3 sg((unsigned)((P.npairs + 3) / 4));
    if (a.profiling) HIP_CHECK(hipEventRecord((hipEvent_t)am.events[2], P.stream));
    launch_singular(a.op, a.kind == kHelmholtz, p0p0, P.small_adjacent, sg, P.stream, S);
    if (a.profiling) {
      HIP_CHECK(hipEventRecord((hipEvent_t)am.events[3], P.stream));
      am.events_recorded[1] = true;
    }
  }
}

void kernel_times(const Assembler& a, float* ms) {
  for (int k = 0; k < 2; ++k) {
    ms[k] = -1.0f;
    if (!a.events_recorded[k]) continue;
    HIP_CHECK(hipEventSynchronize((hipEvent_t)a.events[2 * k + 1]));
    HIP_CHECK(hipEventElapsedTime(&ms[k], (hipEvent_t)a.events[2 * k], (hipEvent_t)a.events[2 * k + 1]));
  }
}

// Refuses a device staging buffer larger than the free device memory with a clear error.
static void check_device_memory(size_t bytes, const char* what) {
  size_t free_b = 0, total_b = 0;
  HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  if (bytes > free_b)
    throw Error(std::string(what) + " needs " + std::to_string(bytes) + " bytes of device memory, " +
                std::to_string(free_b) + " of " + std::to_string(total_b) + " are free");
}

void assemble_host(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, void* host_out) {
  if (!host_out) throw Error("output pointer is null");
  const int device = current_device();
  (void)device;
  const size_t bytes = (size_t)test.size * (size_t)trial.size * dtype_size(a.dtype);
  check_device_memory(bytes, "assembly into host memory (staged on the device)");
  void* d = nullptr;
  HIP_CHECK(hipMalloc(&d, std::max<size_t>(bytes, 1)));
  try {
    hipStream_t s = assembler_stream(a, device);
    if (bytes) HIP_CHECK(hipMemcpyAsync(d, host_out, bytes, hipMemcpyHostToDevice, s));
    assemble_device(a, trial, test, d, true, s);
    if (bytes) HIP_CHECK(hipMemcpyAsync(host_out, d, bytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(d);
    throw;
  }
  HIP_CHECK(hipFree(d));
}

template <int OP, bool HELM>
void launch_correction_q(int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& R, const int4* pairs, int n,
                         double* out) {
  switch (Q) {
    case 1: dev::correction_kernel<1, OP, HELM><<<grid, 256, 0, s>>>(R, pairs, n, out); break;
    case 3: dev::correction_kernel<3, OP, HELM><<<grid, 256, 0, s>>>(R, pairs, n, out); break;
    case 6: dev::correction_kernel<6, OP, HELM><<<grid, 256, 0, s>>>(R, pairs, n, out); break;
    case 7: dev::correction_kernel<7, OP, HELM><<<grid, 256, 0, s>>>(R, pairs, n, out); break;
    default: dev::correction_kernel<0, OP, HELM><<<grid, 256, 0, s>>>(R, pairs, n, out); break;
  }
}

template <bool HELM>
void launch_correction(int op, int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& R, const int4* pairs, int n,
                       double* out) {
  switch (op) {
    case kSL: launch_correction_q<kSL, HELM>(Q, grid, s, R, pairs, n, out); break;
    case kDL: launch_correction_q<kDL, HELM>(Q, grid, s, R, pairs, n, out); break;
    case kADL: launch_correction_q<kADL, HELM>(Q, grid, s, R, pairs, n, out); break;
    case kHS: launch_correction_q<kHS, HELM>(Q, grid, s, R, pairs, n, out); break;
    default: throw Error("operator not supported");
  }
}

// Values of the adjacent pairs in COO layout [npairs][nbt * nbs] (re, im), position
// pair * nbt * nbs + t + nbt * b: the Duffy rules (correction = false) or the regular rule
// (the singular correction). Returns a new device buffer; the caller frees it.
static double* adjacent_coo_values(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test,
                                   const Prepared& P, bool correction) {
  const size_t nvals = (size_t)P.npairs * test.nb * trial.nb;
  double* d = nullptr;
  HIP_CHECK(hipMalloc((void**)&d, nvals * 2 * sizeof(double)));
  try {
    if (correction) {
      const dim3 grid((unsigned)((P.npairs + 255) / 256));
      if (a.kind == kHelmholtz) launch_correction<true>(a.op, P.Q, grid, P.stream, P.R, P.S.pairs, P.npairs, d);
      else launch_correction<false>(a.op, P.Q, grid, P.stream, P.R, P.S.pairs, P.npairs, d);
    } else {
      dev::SingularArgs S = P.S;
      S.mode = dev::kCoo;
      S.out = d;
      dim3 sg((unsigned)((P.npairs + 3) / 4));
      launch_singular(a.op, a.kind == kHelmholtz, test.nb == 1 && trial.nb == 1, P.small_adjacent, sg, P.stream, S);
    }
    HIP_CHECK(hipGetLastError());
  } catch (...) {
    (void)hipFree(d);
    throw;
  }
  return d;
}

// COO triplets on the host (cells' own local basis order, the adjacent-pair order).
static void adjacent_coo(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, bool correction,
                         std::vector<int64_t>& rows, std::vector<int64_t>& cols, std::vector<double>& vals) {
  rows.clear();
  cols.clear();
  vals.clear();
  if (test.grid != trial.grid) return;  // no neighbouring cells across grids (:240-243)
  Prepared P = prepare(a, trial, test, nullptr, nullptr);
  if (P.npairs == 0) return;
  const int nbl = test.nb * trial.nb;
  const size_t nvals = (size_t)P.npairs * nbl;
  std::vector<double> h(nvals * 2);
  double* d = adjacent_coo_values(a, trial, test, P, correction);
  try {
    HIP_CHECK(hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, P.stream));
    HIP_CHECK(hipStreamSynchronize(P.stream));
  } catch (...) {
    (void)hipFree(d);
    throw;
  }
  HIP_CHECK(hipFree(d));
  const auto& pairs = const_cast<Grid&>(*test.grid).device_cache.at(P.device)->host_pairs;
  rows.resize(nvals);
  cols.resize(nvals);
  const bool cplx = is_complex(a.dtype);
  vals.resize(nvals * (cplx ? 2 : 1));
  for (size_t p = 0; p < pairs.size(); ++p)
    for (int b = 0; b < trial.nb; ++b)
      for (int t = 0; t < test.nb; ++t) {
        const size_t e = p * nbl + t + test.nb * b;
        rows[e] = test.dofs[(size_t)pairs[p].test * test.nb + t];
        cols[e] = trial.dofs[(size_t)pairs[p].trial * trial.nb + b];
        if (cplx) {
          vals[2 * e] = h[2 * e];
          vals[2 * e + 1] = h[2 * e + 1];
        } else {
          vals[e] = h[2 * e];
        }
      }
}

void assemble_singular_coo(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test,
                           std::vector<int64_t>& rows, std::vector<int64_t>& cols, std::vector<double>& vals) {
  adjacent_coo(a, trial, test, false, rows, cols, vals);
}

void assemble_singular_correction_coo(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test,
                                      std::vector<int64_t>& rows, std::vector<int64_t>& cols,
                                      std::vector<double>& vals) {
  adjacent_coo(a, trial, test, true, rows, cols, vals);
}

void coo_to_csr_device(const double* vals, const int4* pairs, int npairs, const int* test_dofs, int nbt, int tstride,
                       const int* trial_dofs, int nbs, int sstride, int64_t nrows, int64_t ncols, DType dtype,
                       hipStream_t s, CsrHost& out);  // csr.hip

void assemble_singular_csr(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, bool correction,
                           CsrHost& out) {
  if (test.grid->general() || trial.grid->general()) {
    check_assembler(a, trial, test);
    assemble_cells_singular_csr(a, trial, test, correction, assembler_stream(a, current_device()), out);
    return;
  }
  out.indptr.assign((size_t)test.size + 1, 0);
  out.indices.clear();
  out.data.clear();
  if (test.grid != trial.grid) return;  // no neighbouring cells across grids (:240-243)
  Prepared P = prepare(a, trial, test, nullptr, nullptr);
  if (P.npairs == 0) return;
  double* d = adjacent_coo_values(a, trial, test, P, correction);
  try {
    coo_to_csr_device(d, P.S.pairs, P.npairs, P.S.test_dofs, test.nb, test.nb, P.S.trial_dofs, trial.nb, trial.nb,
                      test.size, trial.size, a.dtype, P.stream, out);
  } catch (...) {
    (void)hipFree(d);
    throw;
  }
  HIP_CHECK(hipFree(d));
}

// ---- potential (off-surface) operators ----
template <int OP, bool HELM>
void launch_potential_q(int Q, dim3 grid, hipStream_t s, const dev::PotentialArgs& P) {
  switch (Q) {
    case 1: dev::potential_kernel<1, OP, HELM><<<grid, 256, 0, s>>>(P); break;
    case 3: dev::potential_kernel<3, OP, HELM><<<grid, 256, 0, s>>>(P); break;
    case 6: dev::potential_kernel<6, OP, HELM><<<grid, 256, 0, s>>>(P); break;
    case 7: dev::potential_kernel<7, OP, HELM><<<grid, 256, 0, s>>>(P); break;
    default: dev::potential_kernel<0, OP, HELM><<<grid, 256, 0, s>>>(P); break;
  }
}

void assemble_potential_device(const Assembler& a, const FunctionSpace& space_c, const double* dev_points,
                               int64_t npoints, void* dev_out, bool accumulate, void* stream) {
  require_triangles(space_c, "potential assembly");
  if (a.op != kSL && a.op != kDL) throw Error("potential operators exist for the single and double layer only");
  if ((a.kind == kLaplace) == is_complex(a.dtype)) throw Error("dtype does not match the kernel");
  if (npoints < 0) throw Error("negative point count");
  if (npoints == 0) return;
  if (!dev_points || !dev_out) throw Error("points or output pointer is null");
  auto& space = const_cast<FunctionSpace&>(space_c);
  auto& grid = const_cast<Grid&>(*space.grid);
  const int device = current_device();
  hipStream_t s = (hipStream_t)stream;
  const int Q = regular_points(a.options);
  std::vector<double> xi, w;
  triangle_rule(Q, xi, w);
  DeviceGridCache* gc = grid_cache(grid, device);
  dev::PotentialArgs P{};
  P.x = dev_points;
  P.np = npoints;
  P.trial_pts = grid_points_for_rule(grid, gc, Q, xi, w);
  P.trial_geo = gc->geo;
  P.trial_dofs = space_dofs(space, device);
  P.n_trial = (int)grid.nc();
  P.nbs = space.nb;
  P.k = a.k;
  for (int p = 0; p < dev::kMaxQ; ++p)
    for (int b = 0; b < 3; ++b) P.phi[p][b] = 0.0;
  for (int p = 0; p < std::min(Q, dev::kMaxQ); ++p) {
    P.phi[p][0] = 1.0 - xi[2 * p] - xi[2 * p + 1];
    P.phi[p][1] = xi[2 * p];
    P.phi[p][2] = xi[2 * p + 1];
  }
  P.nq = Q;
  P.rule_tab = regular_rule_table(device, Q, xi, w);
  P.out = dev_out;
  P.out_type = out_type(a.dtype);
  P.mode = space.nb == 1 ? dev::kAccumulate : dev::kAtomic;
  if (!accumulate) {
    const long long n_out = (long long)npoints * (long long)space.size * (is_complex(a.dtype) ? 2 : 1);
    // a fill kernel, not a memset node: the caller may capture this stream in a graph
    if (a.dtype == kF32 || a.dtype == kC32) dev::fill_zero_kernel<<<2048, 256, 0, s>>>((float*)dev_out, n_out);
    else dev::fill_zero_kernel<<<2048, 256, 0, s>>>((double*)dev_out, n_out);
    HIP_CHECK(hipGetLastError());
  }
  const long long bx = (npoints + 255) / 256;
  if (bx > INT32_MAX) throw Error("too many evaluation points");
  long long chunk = (bx * P.n_trial + 16383) / 16384;
  chunk = std::max(4LL, std::min(128LL, chunk));
  chunk = std::max(chunk, ((long long)P.n_trial + 65534) / 65535);
  P.trial_chunk = (int)chunk;
  const dim3 g((unsigned)bx, (unsigned)((P.n_trial + chunk - 1) / chunk));
  const bool helm = a.kind == kHelmholtz;
  if (a.op == kSL) {
    if (helm) launch_potential_q<kSL, true>(Q, g, s, P);
    else launch_potential_q<kSL, false>(Q, g, s, P);
  } else {
    if (helm) launch_potential_q<kDL, true>(Q, g, s, P);
    else launch_potential_q<kDL, false>(Q, g, s, P);
  }
  HIP_CHECK(hipGetLastError());
}

void assemble_potential_host(const Assembler& a, const FunctionSpace& space, const double* points, int64_t npoints,
                             void* host_out) {
  require_triangles(space, "potential assembly");
  if (npoints == 0) return;
  if (!points || !host_out) throw Error("points or output pointer is null");
  if (a.kind == kHelmholtz) {  // the device sine/cosine reduction is exact while k r <= 1e8
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    auto grow = [&](const double* x) {
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], x[d]);
        hi[d] = std::max(hi[d], x[d]);
      }
    };
    for (int64_t v = 0; v < space.grid->nv(); ++v) grow(&space.grid->points[3 * v]);
    for (int64_t j = 0; j < npoints; ++j) grow(points + 3 * j);
    double diag = 0.0;
    for (int d = 0; d < 3; ++d) diag += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    if (!(std::fabs(a.k) * std::sqrt(diag) <= 1e8)) throw Error("|k| times the extent of mesh and points exceeds 1e8");
  }
  const int device = current_device();
  const size_t bytes = (size_t)npoints * (size_t)space.size * dtype_size(a.dtype);
  check_device_memory(bytes + (size_t)npoints * 3 * sizeof(double), "potential assembly into host memory");
  void* d = nullptr;
  double* dp = nullptr;
  HIP_CHECK(hipMalloc(&d, std::max<size_t>(bytes, 1)));
  try {
    HIP_CHECK(hipMalloc((void**)&dp, (size_t)npoints * 3 * sizeof(double)));
    hipStream_t s = assembler_stream(a, device);
    HIP_CHECK(hipMemcpyAsync(dp, points, (size_t)npoints * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    if (bytes) HIP_CHECK(hipMemcpyAsync(d, host_out, bytes, hipMemcpyHostToDevice, s));
    assemble_potential_device(a, space, dp, npoints, d, true, s);
    if (bytes) HIP_CHECK(hipMemcpyAsync(host_out, d, bytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(d);
    if (dp) (void)hipFree(dp);
    throw;
  }
  HIP_CHECK(hipFree(d));
  HIP_CHECK(hipFree(dp));
}

Assembler::~Assembler() {
  for (void* e : events)
    if (e) (void)hipEventDestroy((hipEvent_t)e);
  if (stream) {
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess) {
      (void)hipSetDevice(device);
      (void)hipStreamDestroy((hipStream_t)stream);
      (void)hipSetDevice(cur);
    }
  }
}

}  // namespace bempp_amd
