// capi.cpp -- extern "C" boundary of the assembly engine (declared in
// include/bempp_amd.h). Mirrors src/bindings/boundary_assemblers.rs and
// src/bindings/matrices.rs of the reference: opaque boxed handles, caller-owned
// output memory. Errors become a non-zero status / NULL plus bempp_last_error()
// where the reference panics.
#include <algorithm>
#include <complex>
#include <cstring>
#include <memory>
#include <numeric>

#include "../../include/bempp_amd.h"
#include "core.hpp"

using namespace bempp_amd;

struct BoundaryAssemblerOptionsT {
  Options o;
};
struct BoundaryAssemblerT {
  Assembler* a;
};
struct GridT {
  Grid* g;
  // topology vertex index -> point index (grids built from ndgrid data; empty: identity)
  std::vector<int64_t> topo_vertex_point;
};
struct SpaceT {
  std::shared_ptr<FunctionSpace> s;
  const GridT* grid = nullptr;
  uint8_t dtype = BEMPP_DTYPE_F64;
  std::vector<int64_t> point_dof;  // continuous P1: point -> DOF (built with the handle)
};
struct ElementFamilyT {
  size_t degree;
  uint8_t continuity;
  uint8_t dtype;
};
struct MatrixT {
  bool csr = false;
  DType dtype = kF64;
  size_t shape0 = 0, shape1 = 0;
  std::vector<unsigned char> data;  // dense column-major, or CSR values
  std::vector<uintptr_t> indices, indptr;
};

namespace {

template <typename F>
auto guarded(F&& f, decltype(f()) fail) -> decltype(f()) {
  try {
    return f();
  } catch (const std::exception& e) {
    set_last_error(e.what());
  } catch (...) {
    set_last_error("unknown error");
  }
  return fail;
}

void need(const void* p, const char* what) {
  if (!p) throw Error(std::string(what) + " is null");
}

}  // namespace

extern "C" {

const char* bempp_last_error(void) { return last_error(); }
const char* bempp_version(void) { return "bempp-amd 0.1.0 (gfx950)"; }
int bempp_device_count(void) { return device_count(); }

// ---- options ----
BoundaryAssemblerOptionsT* boundary_assembler_options_new(void) { return new BoundaryAssemblerOptionsT(); }
void boundary_assembler_options_t_free(BoundaryAssemblerOptionsT* o) { delete o; }

void boundary_assembler_options_set_regular_quadrature_degree(BoundaryAssemblerOptionsT* o, uint8_t cell,
                                                               size_t npoints) {
  if (!o) return;
  auto it = o->o.quadrature_degrees.find(cell);
  if (it != o->o.quadrature_degrees.end()) it->second = npoints;
}
size_t boundary_assembler_options_get_regular_quadrature_degree(const BoundaryAssemblerOptionsT* o, uint8_t cell) {
  if (!o) return 0;
  auto it = o->o.quadrature_degrees.find(cell);
  return it == o->o.quadrature_degrees.end() ? 0 : it->second;
}
int boundary_assembler_options_has_regular_quadrature_degree(const BoundaryAssemblerOptionsT* o, uint8_t cell) {
  return o && o->o.quadrature_degrees.count(cell) ? 1 : 0;
}
void boundary_assembler_options_set_singular_quadrature_degree(BoundaryAssemblerOptionsT* o, uint8_t c0, uint8_t c1,
                                                                size_t npoints) {
  if (!o) return;
  auto it = o->o.singular_quadrature_degrees.find({c0, c1});
  if (it != o->o.singular_quadrature_degrees.end()) it->second = npoints;
}
size_t boundary_assembler_options_get_singular_quadrature_degree(const BoundaryAssemblerOptionsT* o, uint8_t c0,
                                                                 uint8_t c1) {
  if (!o) return 0;
  auto it = o->o.singular_quadrature_degrees.find({c0, c1});
  return it == o->o.singular_quadrature_degrees.end() ? 0 : it->second;
}
int boundary_assembler_options_has_singular_quadrature_degree(const BoundaryAssemblerOptionsT* o, uint8_t c0,
                                                              uint8_t c1) {
  return o && o->o.singular_quadrature_degrees.count({c0, c1}) ? 1 : 0;
}
void boundary_assembler_options_set_batch_size(BoundaryAssemblerOptionsT* o, size_t b) {
  if (o) o->o.batch_size = b;
}
size_t boundary_assembler_options_get_batch_size(const BoundaryAssemblerOptionsT* o) {
  return o ? o->o.batch_size : 0;
}
int boundary_assembler_options_set_precision(BoundaryAssemblerOptionsT* o, uint8_t p) {
  return guarded(
      [&]() {
        need(o, "options");
        if (p != kFP64 && p != kMixed) throw Error("precision must be 0 (fp64) or 1 (mixed)");
        o->o.precision = (Precision)p;
        return 0;
      },
      1);
}
uint8_t boundary_assembler_options_get_precision(const BoundaryAssemblerOptionsT* o) {
  return o ? (uint8_t)o->o.precision : 0;
}

// ---- grids ----
GridT* grid_new(const double* points, size_t npoints, const int64_t* cells, size_t ncells) {
  return guarded(
      [&]() -> GridT* {
        need(points, "points");
        need(cells, "cells");
        std::vector<double> p(points, points + 3 * npoints);
        std::vector<int64_t> c(cells, cells + 3 * ncells);
        return new GridT{new Grid(std::move(p), std::move(c))};
      },
      (GridT*)nullptr);
}
GridT* grid_new_sphere(int seed, uint32_t level) {
  return guarded([&]() -> GridT* { return new GridT{make_sphere(seed, level)}; }, (GridT*)nullptr);
}
void grid_t_free(GridT* g) {
  if (!g) return;
  delete g->g;
  delete g;
}
size_t grid_npoints(const GridT* g) { return g ? (size_t)g->g->nv() : 0; }
size_t grid_ncells(const GridT* g) { return g ? (size_t)g->g->nc() : 0; }
void grid_points_copy(const GridT* g, double* p) {
  if (g && p) std::copy(g->g->points.begin(), g->g->points.end(), p);
}
void grid_cells_copy(const GridT* g, int64_t* c) {
  if (g && c) std::copy(g->g->cells.begin(), g->g->cells.end(), c);
}
GridT* grid_new_cells(const double* points, size_t npoints, const int64_t* cell_vertices, const uint8_t* cell_types,
                      size_t ncells) {
  return guarded(
      [&]() -> GridT* {
        need(points, "points");
        need(cell_types, "cell types");
        size_t nverts = 0;
        for (size_t c = 0; c < ncells; ++c) nverts += cell_types[c] == BEMPP_CELL_QUADRILATERAL ? 4 : 3;
        if (nverts) need(cell_vertices, "cell vertices");
        std::vector<double> p(points, points + 3 * npoints);
        std::vector<int64_t> v(cell_vertices, cell_vertices + nverts);
        std::vector<uint8_t> t(cell_types, cell_types + ncells);
        return new GridT{new Grid(std::move(p), std::move(v), std::move(t))};
      },
      (GridT*)nullptr);
}
size_t grid_cell_vertices_size(const GridT* g) {
  if (!g) return 0;
  return g->g->general() ? g->g->cell_verts.size() : g->g->cells.size();
}
void grid_cell_vertices_copy(const GridT* g, int64_t* v) {
  if (!g || !v) return;
  const auto& src = g->g->general() ? g->g->cell_verts : g->g->cells;
  std::copy(src.begin(), src.end(), v);
}
void grid_cell_types_copy(const GridT* g, uint8_t* t) {
  if (!g || !t) return;
  for (int64_t c = 0; c < g->g->nc(); ++c) t[c] = g->g->cell_nv(c) == 4 ? BEMPP_CELL_QUADRILATERAL : BEMPP_CELL_TRIANGLE;
}
int grid_set_points(GridT* g, const double* p) {
  return guarded(
      [&]() {
        need(g, "grid");
        need(p, "points");
        g->g->set_points(p);
        return 0;
      },
      1);
}

// Grid from the arrays of an ndgrid single-element grid (src/bindings/ndgrid.rs:16-56).
// Flat triangles only (tdim 2, gdim 3, geometry degree 1, 3 points per cell). The arrays are
// copied. The topology arrays are optional; when given, the cells' topology vertices must be a
// relabelling of their geometry points with the same local order, and the relabelling is kept
// for entity-indexed queries (space_get_local_dof_numbers with entity_dim 0).
GridT* single_element_grid_borrowed_create_bempp(size_t tdim, const size_t*, const size_t* const*,
                                                 const uint8_t* entity_types, const size_t* entity_counts,
                                                 const size_t* const* const* downward_connectivity,
                                                 const size_t* const* downward_connectivity_shape0,
                                                 const size_t* const* const* const*, const size_t* const* const*,
                                                 const void* points, size_t gdim, size_t npoints, uint8_t dtype,
                                                 const size_t* cells, size_t points_per_cell, size_t ncells,
                                                 size_t geometry_degree) {
  return guarded(
      [&]() -> GridT* {
        need(points, "points");
        need(cells, "cells");
        if (tdim != 2) throw Error("only grids of topological dimension 2 are supported, got " + std::to_string(tdim));
        if (gdim != 3) throw Error("only grids in 3 dimensions are supported, got " + std::to_string(gdim));
        const uint8_t ctype = points_per_cell == 4 ? BEMPP_CELL_QUADRILATERAL : BEMPP_CELL_TRIANGLE;
        if (geometry_degree != 1 || (points_per_cell != 3 && points_per_cell != 4))
          throw Error("only degree-1 geometry is supported (3 points per triangle, 4 per quadrilateral)");
        if (entity_types && entity_types[2] != ctype)
          throw Error("cell type " + std::to_string(entity_types[2]) + " does not match " +
                      std::to_string(points_per_cell) + " points per cell (triangle or quadrilateral cells only)");
        if (dtype != BEMPP_DTYPE_F32 && dtype != BEMPP_DTYPE_F64)
          throw Error("grid points must be f32 or f64, got dtype " + std::to_string(dtype));
        if (entity_counts && entity_counts[2] != ncells)
          throw Error("entity_counts[2] does not match the number of cells");
        std::vector<double> p(3 * npoints);
        for (size_t i = 0; i < 3 * npoints; ++i)
          p[i] = dtype == BEMPP_DTYPE_F64 ? ((const double*)points)[i] : (double)((const float*)points)[i];
        const size_t ppc = points_per_cell;
        std::vector<int64_t> c(ppc * ncells);
        for (size_t i = 0; i < ppc * ncells; ++i) {
          if (cells[i] >= npoints) throw Error("cell " + std::to_string(i / ppc) + " refers to a missing point");
          c[i] = (int64_t)cells[i];
        }
        std::vector<int64_t> topo;
        const size_t* tc = downward_connectivity && downward_connectivity[2] ? downward_connectivity[2][0] : nullptr;
        if (tc) {
          if (downward_connectivity_shape0 && downward_connectivity_shape0[2] &&
              downward_connectivity_shape0[2][0] != ppc)
            throw Error("cell-to-vertex connectivity must have " + std::to_string(ppc) + " vertices per cell");
          const size_t nvert = entity_counts ? entity_counts[0] : npoints;
          topo.assign(nvert, -1);
          std::vector<int64_t> point_topo(npoints, -1);
          for (size_t i = 0; i < ppc * ncells; ++i) {
            const size_t v = tc[i];
            if (v >= nvert) throw Error("cell " + std::to_string(i / ppc) + " refers to a missing vertex");
            if ((topo[v] >= 0 && topo[v] != c[i]) || (point_topo[c[i]] >= 0 && point_topo[c[i]] != (int64_t)v))
              throw Error("topology and geometry cells differ by more than a relabelling of the vertices");
            topo[v] = c[i];
            point_topo[c[i]] = (int64_t)v;
          }
        }
        Grid* grid = ppc == 3 ? new Grid(std::move(p), std::move(c))
                              : new Grid(std::move(p), std::move(c), std::vector<uint8_t>(ncells, ctype));
        auto* g = new GridT{grid, std::move(topo)};
        return g;
      },
      (GridT*)nullptr);
}
void grid_t_free_bempp(GridT* g) { grid_t_free(g); }

// ---- element families (src/bindings/ndelement.rs:8-44) ----
ElementFamilyT* create_lagrange_family_bempp(size_t degree, uint8_t continuity, uint8_t dtype) {
  return guarded(
      [&]() -> ElementFamilyT* {
        if (continuity > 1) throw Error("continuity must be 0 (standard) or 1 (discontinuous)");
        if (dtype > BEMPP_DTYPE_C64) throw Error("invalid dtype " + std::to_string(dtype));
        return new ElementFamilyT{degree, continuity, dtype};
      },
      (ElementFamilyT*)nullptr);
}
ElementFamilyT* create_raviart_thomas_family_bempp(size_t, uint8_t, uint8_t) {
  set_last_error("Raviart-Thomas spaces are not supported (no operator of this library uses them)");
  return nullptr;
}
ElementFamilyT* create_nedelec_family_bempp(size_t, uint8_t, uint8_t) {
  set_last_error("Nedelec spaces are not supported (no operator of this library uses them)");
  return nullptr;
}
void element_family_t_free_bempp(ElementFamilyT* f) { delete f; }

// ---- spaces ----
SpaceT* function_space_new(const GridT* g, size_t degree, uint8_t continuity) {
  return guarded(
      [&]() -> SpaceT* {
        need(g, "grid");
        if (continuity > 1) throw Error("continuity must be 0 (standard) or 1 (discontinuous)");
        auto* sp = new SpaceT{std::make_shared<FunctionSpace>(g->g, (int)degree, continuity == BEMPP_CONTINUITY_STANDARD)};
        sp->grid = g;
        const FunctionSpace& fs = *sp->s;
        if (fs.continuous && fs.degree == 1) {
          sp->point_dof.assign(fs.grid->nv(), -1);
          for (int64_t c = 0; c < fs.grid->nc(); ++c)
            for (int m = 0; m < fs.cell_nb(c); ++m) sp->point_dof[fs.grid->vertex(c, m)] = fs.dofs[(size_t)fs.nb * c + m];
        }
        return sp;
      },
      (SpaceT*)nullptr);
}
SpaceT* function_space(const GridT* g, const ElementFamilyT* family) {
  return guarded(
      [&]() -> SpaceT* {
        need(family, "element family");
        SpaceT* sp = function_space_new(g, family->degree, family->continuity);
        if (!sp) throw Error(last_error());
        sp->dtype = family->dtype;
        return sp;
      },
      (SpaceT*)nullptr);
}
SpaceT* space_local_space(const SpaceT* s) {
  return guarded(
      [&]() -> SpaceT* {
        need(s, "space");
        return new SpaceT{s->s, s->grid, s->dtype, s->point_dof};
      },
      (SpaceT*)nullptr);
}
void space_t_free(SpaceT* s) { delete s; }
uint8_t space_dtype(const SpaceT* s) { return s ? s->dtype : BEMPP_DTYPE_F64; }

namespace {
// DOFs of an entity (src/function.rs:258-260 over assign_dofs, :365-445): a continuous P1
// space has one DOF per vertex; P0 and discontinuous P1 spaces have theirs on the cell.
const int64_t* entity_dofs(const SpaceT* s, size_t dim, size_t index, size_t* n) {
  *n = 0;
  if (!s) return nullptr;
  const FunctionSpace& fs = *s->s;
  if (fs.continuous && fs.degree == 1) {
    if (dim != 0) return nullptr;
    int64_t point = (int64_t)index;
    if (s->grid && !s->grid->topo_vertex_point.empty()) {
      if (index >= s->grid->topo_vertex_point.size()) return nullptr;
      point = s->grid->topo_vertex_point[index];
    }
    if (point < 0 || point >= (int64_t)s->point_dof.size() || s->point_dof[point] < 0) return nullptr;
    *n = 1;
    return &s->point_dof[point];
  }
  if (dim != 2 || index >= (size_t)fs.grid->nc()) return nullptr;
  *n = (size_t)fs.cell_nb((int64_t)index);
  return &fs.dofs[index * fs.nb];
}
}  // namespace

size_t space_get_local_dof_numbers_size(const SpaceT* s, size_t entity_dim, size_t entity_index) {
  size_t n;
  entity_dofs(s, entity_dim, entity_index, &n);
  return n;
}
void space_get_local_dof_numbers(const SpaceT* s, size_t entity_dim, size_t entity_index, uintptr_t* dofs) {
  size_t n;
  const int64_t* d = entity_dofs(s, entity_dim, entity_index, &n);
  if (dofs)
    for (size_t i = 0; i < n; ++i) dofs[i] = (uintptr_t)d[i];
}
// Serial spaces own every DOF (src/function.rs:355-357); ghost queries get rank 0 and the index.
int space_is_owned(const SpaceT* s, size_t local) { return s && local < (size_t)s->s->size ? 1 : 0; }
size_t space_ownership_process(const SpaceT*, size_t) { return 0; }
size_t space_ownership_index(const SpaceT*, size_t local) { return local; }
size_t space_global_size(const SpaceT* s) { return s ? (size_t)s->s->size : 0; }
size_t space_local_size(const SpaceT* s) { return s ? (size_t)s->s->size : 0; }
int space_is_serial(const SpaceT* s) { return s ? 1 : 0; }
int space_has_cell_dofs(const SpaceT* s, size_t cell) { return s && cell < (size_t)s->s->grid->nc() ? 1 : 0; }
size_t space_cell_dofs_size(const SpaceT* s, size_t cell) {
  return space_has_cell_dofs(s, cell) ? (size_t)s->s->cell_nb((int64_t)cell) : 0;
}
void space_cell_dofs(const SpaceT* s, size_t cell, uintptr_t* dofs) {
  if (!space_has_cell_dofs(s, cell) || !dofs) return;
  for (int b = 0; b < s->s->cell_nb((int64_t)cell); ++b) dofs[b] = (uintptr_t)s->s->dofs[cell * s->s->nb + b];
}
size_t space_global_dof_index(const SpaceT*, size_t local) { return local; }
size_t space_cell_colouring(const SpaceT* s, int64_t* colour) {
  if (!s) return 0;
  auto cols = s->s->colouring();
  if (colour)
    for (size_t c = 0; c < cols.size(); ++c)
      for (int64_t cell : cols[c]) colour[cell] = (int64_t)c;
  return cols.size();
}

// ---- assemblers ----
BoundaryAssemblerT* laplace_boundary_assembler_new(const BoundaryAssemblerOptionsT* o, uint8_t op, uint8_t dtype) {
  return guarded(
      [&]() -> BoundaryAssemblerT* {
        need(o, "options");
        if (op > kHS) throw Error("Invalid operator for Laplace: " + std::to_string(op));
        if (dtype != kF32 && dtype != kF64) throw Error("Invalid dtype for Laplace: " + std::to_string(dtype));
        auto* a = new Assembler();
        a->kind = kLaplace;
        a->op = (Op)op;
        a->dtype = (DType)dtype;
        a->options = o->o;
        return new BoundaryAssemblerT{a};
      },
      (BoundaryAssemblerT*)nullptr);
}
BoundaryAssemblerT* helmholtz_boundary_assembler_new(const void* k, const BoundaryAssemblerOptionsT* o, uint8_t op,
                                                     uint8_t dtype) {
  return guarded(
      [&]() -> BoundaryAssemblerT* {
        need(o, "options");
        need(k, "wavenumber");
        if (op > kHS) throw Error("Invalid operator for Helmholtz: " + std::to_string(op));
        if (dtype != kC32 && dtype != kC64) throw Error("Invalid dtype for Helmholtz: " + std::to_string(dtype));
        auto* a = new Assembler();
        a->kind = kHelmholtz;
        a->op = (Op)op;
        a->dtype = (DType)dtype;
        a->k = dtype == kC32 ? (double)*(const float*)k : *(const double*)k;
        a->options = o->o;
        return new BoundaryAssemblerT{a};
      },
      (BoundaryAssemblerT*)nullptr);
}
void boundary_assembler_t_free(BoundaryAssemblerT* a) {
  if (!a) return;
  delete a->a;
  delete a;
}
uint8_t boundary_assembler_dtype(const BoundaryAssemblerT* a) { return a ? (uint8_t)a->a->dtype : 255; }

static void check_spaces(const BoundaryAssemblerT* a, const SpaceT* trial, const SpaceT* test) {
  need(a, "assembler");
  need(trial, "trial space");
  need(test, "test space");
}

int boundary_assembler_assemble_into_memory(const BoundaryAssemblerT* a, const SpaceT* trial, const SpaceT* test,
                                            void* data) {
  return guarded(
      [&]() {
        check_spaces(a, trial, test);
        assemble_host(*a->a, *trial->s, *test->s, data);
        return 0;
      },
      1);
}

int boundary_assembler_assemble_into_device_memory(const BoundaryAssemblerT* a, const SpaceT* trial,
                                                   const SpaceT* test, void* data, int accumulate, void* stream) {
  return guarded(
      [&]() {
        check_spaces(a, trial, test);
        assemble_device(*a->a, *trial->s, *test->s, data, accumulate != 0, stream);
        return 0;
      },
      1);
}

MatrixT* boundary_assembler_assemble(const BoundaryAssemblerT* a, const SpaceT* trial, const SpaceT* test) {
  return guarded(
      [&]() -> MatrixT* {
        check_spaces(a, trial, test);
        auto* m = new MatrixT();
        m->dtype = a->a->dtype;
        m->shape0 = (size_t)test->s->size;
        m->shape1 = (size_t)trial->s->size;
        m->data.assign(m->shape0 * m->shape1 * dtype_size(m->dtype), 0);
        try {
          assemble_host(*a->a, *trial->s, *test->s, m->data.data());
        } catch (...) {
          delete m;
          throw;
        }
        return m;
      },
      (MatrixT*)nullptr);
}

namespace {

// The device-built CSR (src/boundary_assemblers.rs:122-155) as a MatrixT.
MatrixT* singular_csr_matrix(const BoundaryAssemblerT* a, const SpaceT* trial, const SpaceT* test, bool correction) {
  CsrHost c;
  assemble_singular_csr(*a->a, *trial->s, *test->s, correction, c);
  auto* m = new MatrixT();
  m->csr = true;
  m->dtype = a->a->dtype;
  m->shape0 = (size_t)test->s->size;
  m->shape1 = (size_t)trial->s->size;
  m->indptr.assign(c.indptr.begin(), c.indptr.end());
  m->indices.assign(c.indices.begin(), c.indices.end());
  m->data = std::move(c.data);
  return m;
}

}  // namespace

MatrixT* boundary_assembler_assemble_singular(const BoundaryAssemblerT* a, const SpaceT* trial, const SpaceT* test) {
  return guarded(
      [&]() -> MatrixT* {
        check_spaces(a, trial, test);
        return singular_csr_matrix(a, trial, test, false);
      },
      (MatrixT*)nullptr);
}

int potential_assemble_into_memory(const BoundaryAssemblerT* a, const SpaceT* space, const double* points,
                                   size_t npoints, void* data) {
  return guarded(
      [&]() {
        need(a, "assembler");
        need(space, "space");
        assemble_potential_host(*a->a, *space->s, points, (int64_t)npoints, data);
        return 0;
      },
      1);
}

int potential_assemble_into_device_memory(const BoundaryAssemblerT* a, const SpaceT* space, const double* points,
                                          size_t npoints, void* data, int accumulate, void* stream) {
  return guarded(
      [&]() {
        need(a, "assembler");
        need(space, "space");
        assemble_potential_device(*a->a, *space->s, points, (int64_t)npoints, data, accumulate != 0, stream);
        return 0;
      },
      1);
}

MatrixT* boundary_assembler_assemble_singular_correction(const BoundaryAssemblerT* a, const SpaceT* trial,
                                                         const SpaceT* test) {
  return guarded(
      [&]() -> MatrixT* {
        check_spaces(a, trial, test);
        return singular_csr_matrix(a, trial, test, true);
      },
      (MatrixT*)nullptr);
}

int boundary_assembler_shape_gradient(const BoundaryAssemblerT* a, const SpaceT* trial, const SpaceT* test,
                                      const void* u, const void* v, void* grad) {
  return guarded(
      [&]() {
        check_spaces(a, trial, test);
        need(u, "u");
        need(v, "v");
        need(grad, "grad");
        shape_gradient(*a->a, *trial->s, *test->s, u, v, grad);
        return 0;
      },
      1);
}

int boundary_assembler_shape_gradient_noalloc(const BoundaryAssemblerT* a, const SpaceT* trial, const SpaceT* test,
                                              const void* u, const void* v, void* grad, void* stream) {
  return guarded(
      [&]() {
        check_spaces(a, trial, test);
        need(u, "u");
        need(v, "v");
        need(grad, "grad");
        shape_gradient_noalloc(*a->a, *trial->s, *test->s, u, v, grad, stream);
        return 0;
      },
      1);
}

int boundary_assembler_apply(const BoundaryAssemblerT* a, const SpaceT* trial, const SpaceT* test, const void* u,
                              void* y, int transpose) {
  return guarded(
      [&]() {
        check_spaces(a, trial, test);
        need(u, "u");
        need(y, "y");
        apply(*a->a, *trial->s, *test->s, u, y, transpose != 0);
        return 0;
      },
      1);
}

int boundary_assembler_set_profiling(BoundaryAssemblerT* a, int enable) {
  return guarded(
      [&]() {
        need(a, "assembler");
        a->a->profiling = enable != 0;
        return 0;
      },
      1);
}

int boundary_assembler_kernel_times(const BoundaryAssemblerT* a, float* ms) {
  return guarded(
      [&]() {
        need(a, "assembler");
        need(ms, "ms");
        kernel_times(*a->a, ms);
        return 0;
      },
      1);
}

// ---- batched assembly ----
struct BoundaryAssemblerBatchT {
  Batch* b;
};

BoundaryAssemblerBatchT* boundary_assembler_batch_new(const BoundaryAssemblerT* const* assemblers,
                                                      const SpaceT* const* spaces, void* const* outputs, size_t n) {
  return guarded(
      [&]() -> BoundaryAssemblerBatchT* {
        need(assemblers, "assemblers");
        need(spaces, "spaces");
        need(outputs, "outputs");
        std::vector<const Assembler*> a(n);
        std::vector<const FunctionSpace*> s(n);
        std::vector<void*> o(outputs, outputs + n);
        for (size_t i = 0; i < n; ++i) {
          need(assemblers[i], "assembler");
          need(spaces[i], "space");
          a[i] = assemblers[i]->a;
          s[i] = spaces[i]->s.get();
        }
        Batch* b = batch_new(a, s, o);
        return new BoundaryAssemblerBatchT{b};
      },
      nullptr);
}

void boundary_assembler_batch_t_free(BoundaryAssemblerBatchT* batch) {
  if (!batch) return;
  batch_free(batch->b);
  delete batch;
}

int boundary_assembler_batch_assemble(BoundaryAssemblerBatchT* batch, void* stream) {
  return guarded(
      [&]() {
        need(batch, "batch");
        batch_assemble(*batch->b, stream);
        return 0;
      },
      1);
}

int boundary_assembler_batch_set_graph(BoundaryAssemblerBatchT* batch, int enable) {
  return guarded(
      [&]() {
        need(batch, "batch");
        batch_set_graph(*batch->b, enable != 0);
        return 0;
      },
      1);
}

int boundary_assembler_batch_set_launch(BoundaryAssemblerBatchT* batch, int blocks_per_cu, int chunk, int order) {
  return guarded(
      [&]() {
        need(batch, "batch");
        batch_set_launch(*batch->b, blocks_per_cu, chunk, order);
        return 0;
      },
      1);
}

int boundary_assembler_batch_get_launch(const BoundaryAssemblerBatchT* batch, int* blocks_per_cu, int* chunk,
                                        int* order) {
  return guarded(
      [&]() {
        need(batch, "batch");
        need(blocks_per_cu, "blocks_per_cu");
        need(chunk, "chunk");
        need(order, "order");
        batch_get_launch(*batch->b, blocks_per_cu, chunk, order);
        return 0;
      },
      1);
}

int boundary_assembler_batch_set_tile_waves(BoundaryAssemblerBatchT* batch, int waves) {
  return guarded(
      [&]() {
        need(batch, "batch");
        batch_set_tile_waves(*batch->b, waves);
        return 0;
      },
      1);
}

int boundary_assembler_batch_get_tile_waves(const BoundaryAssemblerBatchT* batch, int* waves) {
  return guarded(
      [&]() {
        need(batch, "batch");
        need(waves, "waves");
        *waves = batch_get_tile_waves(*batch->b);
        return 0;
      },
      1);
}

int boundary_assembler_batch_counts(const BoundaryAssemblerBatchT* batch, int64_t* counts) {
  return guarded(
      [&]() {
        need(batch, "batch");
        need(counts, "counts");
        long long c[4];
        batch_counts(*batch->b, c);
        for (int i = 0; i < 4; ++i) counts[i] = (int64_t)c[i];
        return 0;
      },
      1);
}

double boundary_assembler_batch_autotune(BoundaryAssemblerBatchT* batch, const char* cache_path, int reps,
                                         int* from_cache) {
  return guarded(
      [&]() {
        need(batch, "batch");
        int fc = 0;
        const double ms = batch_autotune(*batch->b, cache_path, reps, &fc);
        if (from_cache) *from_cache = fc;
        return ms;
      },
      -1.0);
}

int boundary_assembler_batch_set_profiling(BoundaryAssemblerBatchT* batch, int enable) {
  return guarded(
      [&]() {
        need(batch, "batch");
        batch_set_profiling(*batch->b, enable != 0);
        return 0;
      },
      1);
}

int boundary_assembler_batch_kernel_times(const BoundaryAssemblerBatchT* batch, float* ms) {
  return guarded(
      [&]() {
        need(batch, "batch");
        need(ms, "ms");
        batch_kernel_times(*batch->b, ms);
        return 0;
      },
      1);
}

// ---- matrices ----
size_t matrix_shape0(const MatrixT* m) { return m ? m->shape0 : 0; }
size_t matrix_shape1(const MatrixT* m) { return m ? m->shape1 : 0; }
int matrix_is_csr(const MatrixT* m) { return m && m->csr ? 1 : 0; }
uint8_t matrix_dtype(const MatrixT* m) { return m ? (uint8_t)m->dtype : 255; }
const void* matrix_dense_data(const MatrixT* m) { return m && !m->csr ? m->data.data() : nullptr; }
void matrix_dense_data_copy(const MatrixT* m, void* d) {
  if (m && !m->csr && d) std::memcpy(d, m->data.data(), m->data.size());
}
size_t matrix_csr_data_size(const MatrixT* m) { return m && m->csr ? m->data.size() / dtype_size(m->dtype) : 0; }
size_t matrix_csr_indices_size(const MatrixT* m) { return m && m->csr ? m->indices.size() : 0; }
size_t matrix_csr_indptr_size(const MatrixT* m) { return m && m->csr ? m->indptr.size() : 0; }
void matrix_csr_data_copy(const MatrixT* m, void* d) {
  if (m && m->csr && d) std::memcpy(d, m->data.data(), m->data.size());
}
void matrix_csr_indices_copy(const MatrixT* m, uintptr_t* d) {
  if (m && m->csr && d) std::copy(m->indices.begin(), m->indices.end(), d);
}
void matrix_csr_indptr_copy(const MatrixT* m, uintptr_t* d) {
  if (m && m->csr && d) std::copy(m->indptr.begin(), m->indptr.end(), d);
}
void matrix_t_free(MatrixT* m) { delete m; }

// Quadrature tables used by the engine (no counterpart in the reference; they
// expose what ndelement::quadrature::simplex_rule and
// bempp_quadrature::duffy::triangle_duffy provide to the reference).
int bempp_triangle_rule(int npoints, double* xi, double* w) {
  return guarded(
      [&]() {
        std::vector<double> a, b;
        triangle_rule(npoints, a, b);
        if (xi) std::copy(a.begin(), a.end(), xi);
        if (w) std::copy(b.begin(), b.end(), w);
        return (int)b.size();
      },
      -1);
}
int bempp_canonical_pair_rule(int nshared, int n, double* pts5) {
  return guarded(
      [&]() {
        std::vector<double> r;
        canonical_pair_rule(nshared, n, r);
        if (pts5) std::copy(r.begin(), r.end(), pts5);
        return (int)(r.size() / 5);
      },
      -1);
}

}  // extern "C"
