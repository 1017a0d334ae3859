// host.cpp -- grids, function spaces, quadrature rules and adjacent-pair
// classification for the MI355X assembly engine (host side, no HIP calls).
#include <algorithm>
#include <cmath>
#include <map>

#include "core.hpp"

namespace bempp_amd {

namespace {
thread_local std::string g_last_error;
constexpr double kPi = 3.14159265358979323846;
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Grid
// ---------------------------------------------------------------------------
Grid::Grid(std::vector<double> pts, std::vector<int64_t> cls) : points(std::move(pts)), cells(std::move(cls)) {
  if (points.size() % 3 || cells.size() % 3) throw Error("grid arrays must have 3 entries per point / cell");
  const int64_t n_v = nv(), n_c = nc();
  if (n_v > INT32_MAX || n_c > INT32_MAX) throw Error("grid too large for 32-bit cell/vertex indices");
  std::vector<int64_t> count(n_v + 1, 0);
  for (int64_t c = 0; c < n_c; ++c) {
    const int64_t* cv = &cells[3 * c];
    for (int m = 0; m < 3; ++m)
      if (cv[m] < 0 || cv[m] >= n_v) throw Error("cell " + std::to_string(c) + " has a vertex index out of range");
    if (cv[0] == cv[1] || cv[0] == cv[2] || cv[1] == cv[2])
      throw Error("cell " + std::to_string(c) + " repeats a vertex");
    for (int m = 0; m < 3; ++m) count[cv[m] + 1]++;
  }
  for (int64_t v = 0; v < n_v; ++v) count[v + 1] += count[v];
  vc_offsets = count;
  vc_cells.assign(count[n_v], 0);
  std::vector<int64_t> fill(count.begin(), count.end() - 1);
  for (int64_t c = 0; c < n_c; ++c)
    for (int m = 0; m < 3; ++m) vc_cells[fill[cells[3 * c + m]]++] = c;  // ascending c
}

Grid::Grid(std::vector<double> pts, std::vector<int64_t> cell_vertices, std::vector<uint8_t> cell_types)
    : points(std::move(pts)) {
  if (points.size() % 3) throw Error("grid points must have 3 coordinates each");
  const int64_t n_v = nv(), n_c = (int64_t)cell_types.size();
  if (n_v > INT32_MAX || n_c > INT32_MAX) throw Error("grid too large for 32-bit cell/vertex indices");
  std::vector<int64_t> off(n_c + 1, 0);
  bool any_quad = false;
  for (int64_t c = 0; c < n_c; ++c) {
    if (cell_types[c] != kTriangle && cell_types[c] != kQuadrilateral)
      throw Error("cell " + std::to_string(c) + ": only triangle (2) and quadrilateral (3) cells are supported");
    any_quad = any_quad || cell_types[c] == kQuadrilateral;
    off[c + 1] = off[c] + (cell_types[c] == kTriangle ? 3 : 4);
  }
  if ((int64_t)cell_vertices.size() != off[n_c])
    throw Error("cell vertex list has " + std::to_string(cell_vertices.size()) + " entries, the cell types need " +
                std::to_string(off[n_c]));
  for (int64_t c = 0; c < n_c; ++c) {
    const int nvv = (int)(off[c + 1] - off[c]);
    for (int m = 0; m < nvv; ++m) {
      const int64_t v = cell_vertices[off[c] + m];
      if (v < 0 || v >= n_v) throw Error("cell " + std::to_string(c) + " has a vertex index out of range");
      for (int q = 0; q < m; ++q)
        if (cell_vertices[off[c] + q] == v) throw Error("cell " + std::to_string(c) + " repeats a vertex");
    }
  }
  if (!any_quad) {
    cells = std::move(cell_vertices);
  } else {
    types = std::move(cell_types);
    cell_verts = std::move(cell_vertices);
    cell_off = std::move(off);
  }
  std::vector<int64_t> count(n_v + 1, 0);
  for (int64_t c = 0; c < n_c; ++c)
    for (int m = 0; m < cell_nv(c); ++m) count[vertex(c, m) + 1]++;
  for (int64_t v = 0; v < n_v; ++v) count[v + 1] += count[v];
  vc_offsets = count;
  vc_cells.assign(count[n_v], 0);
  std::vector<int64_t> fill(count.begin(), count.end() - 1);
  for (int64_t c = 0; c < n_c; ++c)
    for (int m = 0; m < cell_nv(c); ++m) vc_cells[fill[vertex(c, m)]++] = c;  // ascending c
}

Grid::~Grid() {
  for (auto& kv : device_cache) free_device_grid_cache(kv.second);
}

void Grid::set_points(const double* pts) {
  std::lock_guard<std::mutex> lk(cache_mutex);
  std::copy(pts, pts + points.size(), points.begin());
  ++version;
  for (auto& kv : device_cache) free_device_grid_cache(kv.second);
  device_cache.clear();
  cells_cache.clear();
}

// Sphere by repeated midpoint refinement with projection to radius 1
// (src/shapes.rs:59-101): children [c0,e2,e1], [c1,e0,e2], [c2,e1,e0],
// [e0,e1,e2] with e0=mid(c1,c2), e1=mid(c0,c2), e2=mid(c0,c1); midpoints are
// numbered on first creation.
Grid* make_sphere(int seed, uint32_t level) {
  std::vector<double> p;
  std::vector<int64_t> c;
  if (seed == 0) {  // octahedron, src/shapes.rs:38-56
    p = {0, 0, 1, 1, 0, 0, 0, 1, 0, -1, 0, 0, 0, -1, 0, 0, 0, -1};
    c = {0, 1, 2, 0, 2, 3, 0, 3, 4, 0, 4, 1, 5, 2, 1, 5, 3, 2, 5, 4, 3, 5, 1, 4};
  } else if (seed == 1) {  // icosahedron with outward-oriented faces
    const double t = 0.5 * (1.0 + std::sqrt(5.0));
    const double raw[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                               {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& v : raw) {
      const double s = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int d = 0; d < 3; ++d) p.push_back(v[d] / s);
    }
    c = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
         3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  } else {
    throw Error("sphere seed must be 0 (octahedron) or 1 (icosahedron)");
  }
  if (level > 9) throw Error("refinement level too large");
  static const int kEdge[3][2] = {{1, 2}, {0, 2}, {0, 1}};
  for (uint32_t l = 0; l < level; ++l) {
    std::map<std::pair<int64_t, int64_t>, int64_t> mid;
    std::vector<int64_t> next;
    next.reserve(c.size() * 4);
    for (size_t t = 0; t < c.size() / 3; ++t) {
      const int64_t cv[3] = {c[3 * t], c[3 * t + 1], c[3 * t + 2]};
      int64_t e[3];
      for (int m = 0; m < 3; ++m) {
        const int64_t a = cv[kEdge[m][0]], b = cv[kEdge[m][1]];
        const auto key = std::make_pair(std::min(a, b), std::max(a, b));
        auto it = mid.find(key);
        if (it == mid.end()) {
          double q[3];
          for (int d = 0; d < 3; ++d) q[d] = 0.5 * (p[3 * a + d] + p[3 * b + d]);
          const double s = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
          const int64_t id = (int64_t)p.size() / 3;
          for (int d = 0; d < 3; ++d) p.push_back(q[d] / s);
          it = mid.emplace(key, id).first;
        }
        e[m] = it->second;
      }
      const int64_t kids[12] = {cv[0], e[2], e[1], cv[1], e[0], e[2], cv[2], e[1], e[0], e[0], e[1], e[2]};
      next.insert(next.end(), kids, kids + 12);
    }
    c.swap(next);
  }
  return new Grid(std::move(p), std::move(c));
}

// ---------------------------------------------------------------------------
// Function space
// ---------------------------------------------------------------------------
FunctionSpace::FunctionSpace(const Grid* g, int deg, bool cont) : grid(g), degree(deg) {
  if (deg != 0 && deg != 1) throw Error("only Lagrange degree 0 and 1 are supported");
  continuous = cont && deg == 1;  // degree 0 DOFs always sit on the cell interior
  nb = deg == 0 ? 1 : (g->general() ? 4 : 3);
  const int64_t n_c = g->nc();
  dofs.assign(n_c * nb, -1);
  if (continuous) {
    std::vector<int64_t> vdof(g->nv(), -1);
    for (int64_t c = 0; c < n_c; ++c)
      for (int m = 0; m < g->cell_nv(c); ++m) {
        int64_t& d = vdof[g->vertex(c, m)];
        if (d < 0) d = size++;
        dofs[nb * c + m] = d;
      }
  } else {
    for (int64_t c = 0; c < n_c; ++c)
      for (int b = 0; b < cell_nb(c); ++b) dofs[nb * c + b] = size++;
  }
}

void require_triangles(const FunctionSpace& s, const char* what) {
  if (s.grid->general())
    throw Error(std::string(what) +
                " is not supported on grids with quadrilateral cells (dense assembly and the sparse singular part and "
                "correction are)");
}

FunctionSpace::~FunctionSpace() {
  for (auto& kv : device_cache) free_device_space_cache(kv.second);
}

// src/function.rs:281-351: walk cells in order; give each the first colour not
// held by any of its DOF-carrying entities (vertices for continuous P1, the
// cell itself otherwise).
std::vector<std::vector<int64_t>> FunctionSpace::colouring() const {
  std::vector<std::vector<int64_t>> colours;
  const int64_t n_c = grid->nc();
  std::vector<std::vector<int>> held(continuous ? grid->nv() : n_c);
  for (int64_t c = 0; c < n_c; ++c) {
    int64_t ents[4];
    int ne = 0;
    if (continuous)
      for (int m = 0; m < grid->cell_nv(c); ++m) ents[ne++] = grid->vertex(c, m);
    else
      ents[ne++] = c;
    size_t col = 0;
    for (; col < colours.size(); ++col) {
      bool taken = false;
      for (int m = 0; m < ne && !taken; ++m)
        taken = std::find(held[ents[m]].begin(), held[ents[m]].end(), (int)col) != held[ents[m]].end();
      if (!taken) break;
    }
    if (col == colours.size()) colours.emplace_back();
    colours[col].push_back(c);
    for (int m = 0; m < ne; ++m) held[ents[m]].push_back((int)col);
  }
  return colours;
}

// ---------------------------------------------------------------------------
// Quadrature
// ---------------------------------------------------------------------------
namespace {
// fully symmetric orbit {(a,a),(1-2a,a),(a,1-2a)} with weight w each
void orbit21(std::vector<double>& xi, std::vector<double>& w, double a, double wt) {
  const double pts[6] = {a, a, 1.0 - 2.0 * a, a, a, 1.0 - 2.0 * a};
  xi.insert(xi.end(), pts, pts + 6);
  w.insert(w.end(), 3, wt);
}

// orbit of the barycentric point (a, b, 1 - a - b) with three distinct coordinates: its
// six permutations (l0, l1, l2), as reference coordinates (xi0, xi1) = (l1, l2)
void orbit111(std::vector<double>& xi, std::vector<double>& w, double a, double b, double wt) {
  const double l[3] = {a, b, 1.0 - a - b};
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (const auto& p : perm) {
    xi.push_back(l[p[1]]);
    xi.push_back(l[p[2]]);
  }
  w.insert(w.end(), 6, wt);
}

// Orbit tables {type, a, b, weight}: type 0 the centroid, 1 an S21 orbit (a, a, 1 - 2a),
// 2 an S111 orbit (a, b, 1 - a - b). Derived from the moment equations by
// triangle_rules.py in this directory (positive weights, interior points; the
// 37-point rule is the degree-13 solution of smallest truncation error).
struct RuleOrbit {
  int type;
  double a, b, w;
};
// 16 points, degree 8: centroid, 3 S21, 1 S111 (Dunavant's degree-8 rule)
const RuleOrbit kRule16[] = {
    {0, 0.333333333333333333333, 0.0, 0.0721578038388935841255},
    {1, 0.0505472283170309754584, 0.0, 0.0162292488115990401555},
    {1, 0.170569307751760206622, 0.0, 0.0516086852673591251409},
    {1, 0.459292588292723156029, 0.0, 0.0475458171336423123969},
    {2, 0.00839477740995760533721, 0.263112829634638113422, 0.0136151570872174971324},
};
// 37 points, degree 13: centroid, 4 S21, 4 S111
const RuleOrbit kRule37[] = {
    {0, 0.333333333333333333333, 0.0, 0.0333326559198216055272},
    {1, 0.124457822063999260968, 0.0, 0.0162728885505310468142},
    {1, 0.225840072867902017418, 0.0, 0.0285184397656679532653},
    {1, 0.429159612543110726445, 0.0, 0.028185691589537655882},
    {1, 0.487428645297366234185, 0.0, 0.013523851440530054372},
    {2, 0.00493532348954305527152, 0.286214753544342015383, 0.00456921940718551612603},
    {2, 0.0163507805075914482666, 0.0328542486808598094305, 0.00197048267071738041324},
    {2, 0.0267328097943362938005, 0.124525415851328238663, 0.00875670102545966496751},
    {2, 0.0712744715119110429253, 0.284520764019818232792, 0.0192310519035338157386},
};

template <size_t N>
void expand_orbits(const RuleOrbit (&t)[N], std::vector<double>& xi, std::vector<double>& w) {
  for (const RuleOrbit& o : t) {
    if (o.type == 0) {
      xi.push_back(1.0 / 3.0);
      xi.push_back(1.0 / 3.0);
      w.push_back(o.w);
    } else if (o.type == 1) {
      orbit21(xi, w, o.a, o.w);
    } else {
      orbit111(xi, w, o.a, o.b, o.w);
    }
  }
}
}  // namespace

// Symmetric triangle rules (ndelement's rules are called at src/boundary_assemblers.rs:419-464):
// 1 point degree 1; 3 points degree 2; 6 points degree 4 (two S21 orbits with the
// closed-form abscissae a = (8 - sqrt10 +- sqrt(38 - 44 sqrt(2/5)))/18); 7 points
// degree 5 (Radon: centroid + orbits a = (6 -+ sqrt15)/21) -- these four coincide with the
// Witherden-Vincent rules of the same point counts; 16 points degree 8 and 37 points
// degree 13 (the reference's default) from the orbit tables above.
void triangle_rule(int npoints, std::vector<double>& xi, std::vector<double>& w) {
  xi.clear();
  w.clear();
  if (npoints == 1) {
    xi = {1.0 / 3.0, 1.0 / 3.0};
    w = {0.5};
  } else if (npoints == 3) {
    orbit21(xi, w, 1.0 / 6.0, 1.0 / 6.0);
  } else if (npoints == 6) {
    const double root10 = std::sqrt(10.0);
    const double disc = std::sqrt(38.0 - 44.0 * std::sqrt(2.0 / 5.0));
    const double wdisc = std::sqrt(213125.0 - 53320.0 * root10);
    orbit21(xi, w, (8.0 - root10 + disc) / 18.0, (620.0 + wdisc) / 7440.0);
    orbit21(xi, w, (8.0 - root10 - disc) / 18.0, (620.0 - wdisc) / 7440.0);
  } else if (npoints == 7) {
    const double root15 = std::sqrt(15.0);
    xi = {1.0 / 3.0, 1.0 / 3.0};
    w = {9.0 / 80.0};
    orbit21(xi, w, (6.0 - root15) / 21.0, (155.0 - root15) / 2400.0);
    orbit21(xi, w, (6.0 + root15) / 21.0, (155.0 + root15) / 2400.0);
  } else if (npoints == 16) {
    expand_orbits(kRule16, xi, w);
  } else if (npoints == 37) {
    expand_orbits(kRule37, xi, w);
  } else {
    throw Error("unsupported regular quadrature point count " + std::to_string(npoints) +
                " for triangles (supported: 1, 3, 6, 7, 16, 37)");
  }
}

void gauss01(int n, std::vector<double>& x, std::vector<double>& w) {
  if (n < 1 || n > 64) throw Error("singular quadrature degree must be in [1, 64]");
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 64; ++it) {
      double p = 1.0, pm = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double pn = ((2 * j - 1) * z * p - (j - 1) * pm) / j;
        pm = p;
        p = pn;
      }
      dp = n * (z * p - pm) / (z * z - 1.0);
      const double step = p / dp;
      z -= step;
      if (std::fabs(step) <= 1e-17) break;
    }
    double p = 1.0, pm = 0.0;
    for (int j = 1; j <= n; ++j) {
      const double pn = ((2 * j - 1) * z * p - (j - 1) * pm) / j;
      pm = p;
      p = pn;
    }
    dp = n * (z * p - pm) / (z * z - 1.0);
    const double wt = 1.0 / ((1.0 - z * z) * dp * dp);
    x[i] = 0.5 * (1.0 - z);
    x[n - 1 - i] = 0.5 * (1.0 + z);
    w[i] = wt;
    w[n - 1 - i] = wt;
  }
}

// Sauter & Schwab, Boundary Element Methods (2011), section 5.2: the 4-D
// integral over tau x tau (tau = {0 <= x2 <= x1 <= 1}) is split into regions
// (6 identical panels, 5 common edge, 2 common vertex), each mapped from
// [0,1]^4 with a polynomial Jacobian that cancels the 1/r singularity. The
// result is pulled to the reference triangle by (x1,x2) -> (x1 - x2, x2).
void canonical_pair_rule(int nshared, int n, std::vector<double>& pts5) {
  if (nshared < 1 || nshared > 3) throw Error("shared vertex count must be 1, 2 or 3");
  std::vector<double> g, gw;
  gauss01(n, g, gw);
  pts5.clear();
  const int regions = nshared == 3 ? 6 : nshared == 2 ? 5 : 2;
  pts5.reserve((size_t)regions * n * n * n * n * 5);
  auto emit = [&](double x1, double x2, double y1, double y2, double wt) {
    const double row[5] = {x1 - x2, x2, y1 - y2, y2, wt};
    pts5.insert(pts5.end(), row, row + 5);
  };
  for (int r = 0; r < regions; ++r)
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b)
        for (int c = 0; c < n; ++c)
          for (int d = 0; d < n; ++d) {
            const double s = g[a], e1 = g[b], e2 = g[c], e3 = g[d];
            const double base = gw[a] * gw[b] * gw[c] * gw[d] * s * s * s;
            if (nshared == 3) {
              const double wt = base * e1 * e1 * e2;
              const double p = s * (1 - e1 + e1 * e2), q = s * (1 - e1 * e2 * e3), u = s * (1 - e1);
              const double h1 = s * e1 * (1 - e2 + e2 * e3), h2 = s * (1 - e1 * e2), h3 = s * e1 * (1 - e2);
              const double h4 = s * e1 * (1 - e2 * e3);
              switch (r) {
                case 0: emit(s, p, q, u, wt); break;
                case 1: emit(q, u, s, p, wt); break;
                case 2: emit(s, h1, h2, h3, wt); break;
                case 3: emit(h2, h3, s, h1, wt); break;
                case 4: emit(q, h4, s, h3, wt); break;
                default: emit(s, h3, q, h4, wt); break;
              }
            } else if (nshared == 2) {
              const double w1 = base * e1 * e1, w2 = w1 * e2;
              switch (r) {
                case 0: emit(s, s * e1 * e3, s * (1 - e1 * e2), s * e1 * (1 - e2), w1); break;
                case 1: emit(s, s * e1, s * (1 - e1 * e2 * e3), s * e1 * e2 * (1 - e3), w2); break;
                case 2: emit(s * (1 - e1 * e2), s * e1 * (1 - e2), s, s * e1 * e2 * e3, w2); break;
                case 3: emit(s * (1 - e1 * e2 * e3), s * e1 * e2 * (1 - e3), s, s * e1, w2); break;
                default: emit(s * (1 - e1 * e2 * e3), s * e1 * (1 - e2 * e3), s, s * e1 * e2, w2); break;
              }
            } else {
              const double wt = base * e2;
              if (r == 0) emit(s, s * e1, s * e2, s * e2 * e3, wt);
              else emit(s * e2, s * e2 * e3, s, s * e1, wt);
            }
          }
}

// Each ordered pair of cells sharing at least one vertex, found once at its
// smallest shared vertex (make_cell_blocks + get_pairs_if_smallest,
// src/boundary_assemblers.rs:564-606,775-792). Shared vertex pairs are ordered by
// trial-local index; the canonical->local permutation of each cell puts the
// shared vertices first in that order and the rest ascending.
void adjacent_pairs(const Grid& g, std::vector<AdjacentPair>& out) {
  out.clear();
  const int64_t n_v = g.nv();
  for (int64_t v = 0; v < n_v; ++v) {
    for (int64_t ia = g.vc_offsets[v]; ia < g.vc_offsets[v + 1]; ++ia) {
      const int64_t tc = g.vc_cells[ia];
      const int64_t* tv = &g.cells[3 * tc];
      for (int64_t ib = g.vc_offsets[v]; ib < g.vc_offsets[v + 1]; ++ib) {
        const int64_t sc = g.vc_cells[ib];
        const int64_t* sv = &g.cells[3 * sc];
        int tl[3], sl[3], m = 0;
        bool smallest = true;
        for (int si = 0; si < 3 && smallest; ++si)
          for (int ti = 0; ti < 3; ++ti)
            if (tv[ti] == sv[si]) {
              if (tv[ti] < v) { smallest = false; break; }
              tl[m] = ti;
              sl[m] = si;
              ++m;
            }
        if (!smallest || m == 0) continue;
        int tp[3], sp[3];
        bool tu[3] = {false, false, false}, su[3] = {false, false, false};
        for (int k = 0; k < m; ++k) { tp[k] = tl[k]; sp[k] = sl[k]; tu[tl[k]] = su[sl[k]] = true; }
        for (int x = 0, kt = m, ks = m; x < 3; ++x) {
          if (!tu[x]) tp[kt++] = x;
          if (!su[x]) sp[ks++] = x;
        }
        int code = 0;
        for (int k = 0; k < 3; ++k) code |= (tp[k] << (2 * k)) | (sp[k] << (8 + 2 * k));
        out.push_back({(int32_t)tc, (int32_t)sc, m, code});
      }
    }
  }
}

// m x m Gauss-Legendre on the reference square (ndelement's quadrilateral rule)
void quadrilateral_rule(int npoints, std::vector<double>& xi, std::vector<double>& w) {
  int m = 1;
  while (m * m < npoints) ++m;
  if (m * m != npoints)
    throw Error("a quadrilateral regular rule needs a square number of points (m x m Gauss), got " +
                std::to_string(npoints));
  std::vector<double> g, gw;
  gauss01(m, g, gw);
  xi.clear();
  w.clear();
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      xi.push_back(g[i]);
      xi.push_back(g[j]);
      w.push_back(gw[i] * gw[j]);
    }
}

namespace {
// Corners of the reference cells: the triangle's are the square's first three.
const double kRefCorner[4][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}, {1.0, 1.0}};

// The triangles a cell is integrated over for a singular pair, as local vertex triples: a
// triangle is itself; a quadrilateral is split along a diagonal so that both halves hold the
// pair's first shared vertex s (and, for a shared edge (s, s2), one half holds the edge):
// shared vertex: the diagonal through s; shared edge: the diagonal through s, halves
// (s, s2, opposite(s)) and (s, opposite(s), the fourth vertex); coincident: the diagonal 0-3.
// Every (test half, trial half) pair then shares s, so the triangle Sauter-Schwab rules
// apply to it through the composite map.
int split_cell(int nvv, const int* shared, int nshared, int (*tri)[3]) {
  if (nvv == 3) {
    tri[0][0] = 0, tri[0][1] = 1, tri[0][2] = 2;
    return 1;
  }
  if (nshared >= 3) {
    const int t[2][3] = {{0, 1, 3}, {0, 3, 2}};
    std::copy(&t[0][0], &t[0][0] + 6, &tri[0][0]);
    return 2;
  }
  const int s = shared[0], o = 3 - s;  // tensor order: 0-3 and 1-2 are the diagonals
  int rest[2], nr = 0;
  for (int v = 0; v < 4; ++v)
    if (v != s && v != o) rest[nr++] = v;
  if (nshared == 1) {
    const int t[2][3] = {{s, rest[0], o}, {s, o, rest[1]}};
    std::copy(&t[0][0], &t[0][0] + 6, &tri[0][0]);
    return 2;
  }
  const int s2 = shared[1];
  if (s2 == o) throw Error("two quadrilaterals share opposite vertices without sharing an edge");
  const int t4 = rest[0] == s2 ? rest[1] : rest[0];
  const int t[2][3] = {{s, s2, o}, {s, o, t4}};
  std::copy(&t[0][0], &t[0][0] + 6, &tri[0][0]);
  return 2;
}
}  // namespace

void cell_pair_rule(const Grid& g, int64_t tc, int64_t sc, int n, std::vector<double>& pts5) {
  pts5.clear();
  const int tn = g.cell_nv(tc), sn = g.cell_nv(sc);
  int tsh[4], ssh[4], m = 0;
  if (tc == sc) {
    for (int v = 0; v < tn; ++v) tsh[m] = ssh[m] = v, ++m;
  } else {
    for (int si = 0; si < sn; ++si)  // shared vertices in trial-local order
      for (int ti = 0; ti < tn; ++ti)
        if (g.vertex(tc, ti) == g.vertex(sc, si)) tsh[m] = ti, ssh[m] = si, ++m;
  }
  if (m == 0) throw Error("cell pair without a common vertex");
  int ttri[2][3], stri[2][3];
  const int nt = split_cell(tn, tsh, m, ttri), ns = split_cell(sn, ssh, m, stri);
  std::vector<double> canon[4];
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < ns; ++b) {
      // shared vertices of the two halves, in the trial half's local order
      int tl[3], sl[3], k = 0;
      for (int bi = 0; bi < 3; ++bi)
        for (int ai = 0; ai < 3; ++ai)
          if (g.vertex(tc, ttri[a][ai]) == g.vertex(sc, stri[b][bi])) tl[k] = ai, sl[k] = bi, ++k;
      if (k == 0) throw Error("sub-triangle pair without a common vertex");
      if (canon[k].empty()) canonical_pair_rule(k, n, canon[k]);
      // canonical vertex j -> half-local vertex perm[j]: shared ones first, then the others
      int tp[3], sp[3];
      bool tu[3] = {false, false, false}, su[3] = {false, false, false};
      for (int j = 0; j < k; ++j) tp[j] = tl[j], sp[j] = sl[j], tu[tl[j]] = su[sl[j]] = true;
      for (int x = 0, jt = k, js = k; x < 3; ++x) {
        if (!tu[x]) tp[jt++] = x;
        if (!su[x]) sp[js++] = x;
      }
      // reference coordinates of the half's vertices in the cell's reference element
      double tv[3][2], sv[3][2];
      for (int j = 0; j < 3; ++j)
        for (int d = 0; d < 2; ++d) tv[j][d] = kRefCorner[ttri[a][j]][d], sv[j][d] = kRefCorner[stri[b][j]][d];
      const double tdet = std::fabs((tv[1][0] - tv[0][0]) * (tv[2][1] - tv[0][1]) -
                                    (tv[2][0] - tv[0][0]) * (tv[1][1] - tv[0][1]));
      const double sdet = std::fabs((sv[1][0] - sv[0][0]) * (sv[2][1] - sv[0][1]) -
                                    (sv[2][0] - sv[0][0]) * (sv[1][1] - sv[0][1]));
      const std::vector<double>& C = canon[k];
      for (size_t i = 0; i < C.size() / 5; ++i) {
        const double* r = &C[5 * i];
        // barycentric coordinates of the canonical point, moved to the half's local vertices
        const double tl3[3] = {1.0 - r[0] - r[1], r[0], r[1]}, sl3[3] = {1.0 - r[2] - r[3], r[2], r[3]};
        double tb[3], sb[3];
        for (int j = 0; j < 3; ++j) tb[tp[j]] = tl3[j], sb[sp[j]] = sl3[j];
        for (int d = 0; d < 2; ++d) pts5.push_back(tb[0] * tv[0][d] + tb[1] * tv[1][d] + tb[2] * tv[2][d]);
        for (int d = 0; d < 2; ++d) pts5.push_back(sb[0] * sv[0][d] + sb[1] * sv[1][d] + sb[2] * sv[2][d]);
        pts5.push_back(r[4] * tdet * sdet);
      }
    }
}

}  // namespace bempp_amd
