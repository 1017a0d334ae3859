// core.hpp -- host-side types of the MI355X boundary-element assembly engine.
//
// Mirrors the reference's Rust API for the dense/singular assembly path:
//   BoundaryAssemblerOptions      src/boundary_assemblers.rs:32-100
//   BoundaryAssembler             src/boundary_assemblers.rs:102-231
//   FunctionSpace (serial part)   src/function.rs:240-445
//   laplace/helmholtz factories   src/laplace.rs:34-78, src/helmholtz.rs:44-107
//   regular_sphere                src/shapes.rs:19-114
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace bempp_amd {

// ndelement ReferenceCellType numbering
enum CellType : uint8_t { kPoint = 0, kInterval = 1, kTriangle = 2, kQuadrilateral = 3 };
enum DType : uint8_t { kF32 = 0, kF64 = 1, kC32 = 2, kC64 = 3 };
enum Op : uint8_t { kSL = 0, kDL = 1, kADL = 2, kHS = 3, kEF = 4, kMF = 5 };
enum Kind : uint8_t { kLaplace = 0, kHelmholtz = 1 };
enum Precision : uint8_t { kFP64 = 0, kMixed = 1 };

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& msg);
const char* last_error();

// ---------------------------------------------------------------------------
// Options (src/boundary_assemblers.rs:32-100). Defaults: 37 regular points per
// triangle/quadrilateral, 4 Gauss points per Duffy dimension, batch size 128.
// set_* only change keys that already exist (.entry().and_modify, :61-65,78-81).
// ---------------------------------------------------------------------------
struct Options {
  std::map<uint8_t, size_t> quadrature_degrees{{kTriangle, 37}, {kQuadrilateral, 37}};
  std::map<std::pair<uint8_t, uint8_t>, size_t> singular_quadrature_degrees{
      {{kTriangle, kTriangle}, 4}, {{kQuadrilateral, kQuadrilateral}, 4},
      {{kQuadrilateral, kTriangle}, 4}, {{kTriangle, kQuadrilateral}, 4}};
  size_t batch_size = 128;
  Precision precision = kFP64;
};

struct DeviceGridCache;   // device.cpp
struct DeviceSpaceCache;  // device.cpp
void free_device_grid_cache(DeviceGridCache*);
void free_device_space_cache(DeviceSpaceCache*);

// Device data of the kernels for grids with quadrilaterals, cached per grid or space and device
// under a key naming what it was built from; freed with its owner.
struct DeviceBlob {
  virtual ~DeviceBlob() = default;
};
using DeviceBlobCache = std::map<std::pair<int, std::string>, std::unique_ptr<DeviceBlob>>;

// ---------------------------------------------------------------------------
// Grid of flat triangles: vertex coordinates and cell -> vertex ids, plus the
// vertex -> cells adjacency the singular pair classification walks
// (src/boundary_assemblers.rs:575-593).
// ---------------------------------------------------------------------------
struct Grid {
  std::vector<double> points;  // [nv][3]
  std::vector<int64_t> cells;  // [nc][3] (triangle grids; empty for grids with quadrilaterals)
  std::vector<int64_t> vc_offsets, vc_cells;  // vertex -> cells, ascending
  uint64_t version = 0;  // bumped when points change
  // Grids with quadrilateral cells (single-element or mixed, degree-1 geometry): per cell its
  // type (kTriangle or kQuadrilateral) and its vertices back to back from cell_off[c]
  // (quadrilaterals in tensor order: reference (0,0), (1,0), (0,1), (1,1)). Empty for
  // triangle grids; a grid of triangles only is stored as a triangle grid.
  std::vector<uint8_t> types;
  std::vector<int64_t> cell_verts, cell_off;

  Grid(std::vector<double> pts, std::vector<int64_t> cls);
  Grid(std::vector<double> pts, std::vector<int64_t> cell_vertices, std::vector<uint8_t> cell_types);
  ~Grid();
  int64_t nv() const { return (int64_t)points.size() / 3; }
  int64_t nc() const { return general() ? (int64_t)types.size() : (int64_t)cells.size() / 3; }
  bool general() const { return !types.empty(); }
  int cell_nv(int64_t c) const { return general() && types[c] == kQuadrilateral ? 4 : 3; }
  int64_t vertex(int64_t c, int m) const { return general() ? cell_verts[cell_off[c] + m] : cells[3 * c + m]; }
  void set_points(const double* pts);

  std::mutex cache_mutex;
  std::map<int, DeviceGridCache*> device_cache;  // device id -> cache
  DeviceBlobCache cells_cache;                   // grids with quadrilaterals (cells.hip)
};

Grid* make_sphere(int seed, uint32_t level);

// ---------------------------------------------------------------------------
// Function space (serial part of src/function.rs): Lagrange degree 0 or 1,
// continuous or discontinuous. DOFs are numbered in first-encounter order over
// cells (assign_dofs, :365-445); colouring is the greedy scheme of :281-351.
// ---------------------------------------------------------------------------
struct FunctionSpace {
  const Grid* grid;
  int degree;
  bool continuous;
  int nb;                        // basis functions per cell (the largest: 4 on quadrilaterals)
  std::vector<int64_t> dofs;     // [nc][nb]; a triangle's fourth entry is -1
  int64_t size = 0;

  FunctionSpace(const Grid* g, int degree, bool continuous);
  int cell_nb(int64_t c) const { return degree == 0 ? 1 : grid->cell_nv(c); }
  ~FunctionSpace();
  std::vector<std::vector<int64_t>> colouring() const;

  std::mutex cache_mutex;
  std::map<int, DeviceSpaceCache*> device_cache;
  DeviceBlobCache cells_cache;  // grids with quadrilaterals (cells.hip)
};

// ---------------------------------------------------------------------------
// Quadrature (host side; tables are uploaded to the device).
// ---------------------------------------------------------------------------
// Symmetric rule on the reference triangle (0,0),(1,0),(0,1), weights sum 1/2.
void triangle_rule(int npoints, std::vector<double>& xi, std::vector<double>& w);
// Gauss-Legendre on [0,1].
void gauss01(int n, std::vector<double>& x, std::vector<double>& w);
// m x m Gauss-Legendre rule on the reference square (npoints = m^2), weights sum 1.
void quadrilateral_rule(int npoints, std::vector<double>& xi, std::vector<double>& w);
// Duffy rule of an adjacent cell pair of a grid with quadrilaterals, in the two cells' own
// reference coordinates: records (test xi0, test xi1, trial xi0, trial xi1, weight).
void cell_pair_rule(const Grid& g, int64_t test_cell, int64_t trial_cell, int n, std::vector<double>& pts5);
// Canonical Sauter-Schwab rule with `nshared` (1,2,3) shared vertices at the
// canonical positions V0 (vertex), V0,V1 (edge), all (coincident); each point
// is (test xi0, test xi1, trial xi0, trial xi1, weight); weights sum to 1/4.
void canonical_pair_rule(int nshared, int n, std::vector<double>& pts5);

// Adjacent (singular) pair: cells sharing >= 1 vertex, with the canonical ->
// local vertex permutations of both cells (shared vertices first, ordered by
// the trial-local index as src/boundary_assemblers.rs:775-792 lists them).
struct AdjacentPair {
  int32_t test, trial;
  int32_t nshared;
  int32_t perm;  // test perm in bits 0..5 (2 bits each), trial perm in bits 8..13
};
void adjacent_pairs(const Grid& g, std::vector<AdjacentPair>& out);

// ---------------------------------------------------------------------------
// Assembler (src/boundary_assemblers.rs:105-116 + the factories).
// ---------------------------------------------------------------------------
struct Assembler {
  Kind kind;
  Op op;
  DType dtype;
  double k = 0.0;
  Options options;  // copied: the reference borrows it for the assembler's lifetime
  int device = -1;
  void* stream = nullptr;  // hipStream_t, created lazily
  // Kernel timing with HIP events on the launch stream (regular, singular).
  bool profiling = false;
  void* events[4] = {nullptr, nullptr, nullptr, nullptr};
  bool events_recorded[2] = {false, false};
  ~Assembler();
};

Assembler* make_assembler(Kind kind, Op op, DType dtype, double k, const Options& opts);

// Dense assembly into device memory on the current device (accumulate or overwrite).
void assemble_device(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, void* dev_out,
                     bool accumulate, void* stream);
// Dense assembly on a grid with quadrilateral cells (FP64, any output dtype).
void assemble_cells_device(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, void* dev_out,
                           bool accumulate, void* stream);
// Refuses grids with quadrilateral cells in the paths specialised to triangles.
void require_triangles(const FunctionSpace& s, const char* what);
// Dense assembly accumulated into a host buffer.
void assemble_host(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, void* host_out);
// Singular part as COO triplets (rows = test DOFs, cols = trial DOFs); values
// in the assembler dtype (interleaved for complex).
void assemble_singular_coo(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test,
                           std::vector<int64_t>& rows, std::vector<int64_t>& cols, std::vector<double>& vals);
// Regular-rule contributions of the adjacent pairs (the "singular correction") as COO.
void assemble_singular_correction_coo(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test,
                                      std::vector<int64_t>& rows, std::vector<int64_t>& cols,
                                      std::vector<double>& vals);
// CSR matrix in host memory: row pointers [nrows + 1], column indices and values (in the
// assembler dtype, interleaved for complex) of the nnz stored entries.
struct CsrHost {
  std::vector<int64_t> indptr, indices;
  std::vector<unsigned char> data;
};
// The singular part (correction = false) or the singular correction (true) as CSR, built on
// the device: columns ascending within each row, duplicate entries summed in COO order, and
// no stored entries when no |entry| exceeds 1e-10.
void assemble_singular_csr(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, bool correction,
                           CsrHost& out);
// The same on a grid with quadrilateral cells, on `stream`.
void assemble_cells_singular_csr(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test,
                                 bool correction, void* stream, CsrHost& out);
// Potential operator [npoints x ndofs] (column-major) of a single/double-layer assembler
// on `space` at the given points; device variant (points and output in device memory,
// on `stream`) and host variant (accumulates into a host buffer).
void assemble_potential_device(const Assembler& a, const FunctionSpace& space, const double* dev_points,
                               int64_t npoints, void* dev_out, bool accumulate, void* stream);
void assemble_potential_host(const Assembler& a, const FunctionSpace& space, const double* points, int64_t npoints,
                             void* host_out);
// y = A u (or A^T u) without assembling A (P0 x P0 spaces); u, y in the assembler dtype.
void apply(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, const void* u, void* y,
           bool transpose = false);
// Shape gradient of v^T A(X) u with respect to the grid's vertex coordinates.
void shape_gradient(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, const void* u,
                    const void* v, void* grad);
// The same with u, v and grad in device memory, enqueued on `stream` (NULL: the null stream),
// allocating no device memory: its partial results live in a static device buffer of
// kNoAllocScratchBytes per GPU, so grids of at most kNoAllocMaxCells cells; the grid's,
// space's and rule's device data must already exist (an ordinary call builds them).
constexpr long long kNoAllocScratchBytes = 16LL << 20;
constexpr long long kNoAllocBytesPerCell = 176;  // u, v as complex128 (32 B) + cell-vertex gradient (144 B)
constexpr long long kNoAllocMaxCells = kNoAllocScratchBytes / kNoAllocBytesPerCell;
void shape_gradient_noalloc(const Assembler& a, const FunctionSpace& trial, const FunctionSpace& test, const void* u,
                            const void* v, void* grad, void* stream);
// While alive, the device caches refuse to allocate (an Error instead of an upload).
struct NoDeviceAllocScope {
  NoDeviceAllocScope();
  ~NoDeviceAllocScope();
  NoDeviceAllocScope(const NoDeviceAllocScope&) = delete;
  NoDeviceAllocScope& operator=(const NoDeviceAllocScope&) = delete;
};

int device_count();
// Elapsed milliseconds of the regular and singular kernels of the last
// profiled assembly (waits for them); -1 when not recorded.
void kernel_times(const Assembler& a, float* ms2);

// Batched dense assembly (batch.hip): items (assembler, P0 space, device output) share
// kernel, operator, dtype, precision and regular rule; each output is overwritten.
struct Batch;
Batch* batch_new(const std::vector<const Assembler*>& asms, const std::vector<const FunctionSpace*>& spaces,
                 const std::vector<void*>& outputs);
void batch_free(Batch* b);
void batch_assemble(Batch& b, void* stream);
void batch_set_graph(Batch& b, bool enable);
void batch_set_profiling(Batch& b, bool enable);
void batch_set_launch(Batch& b, int blocks_per_cu, int chunk, int order);
void batch_get_launch(const Batch& b, int* blocks_per_cu, int* chunk, int* order);
void batch_set_tile_waves(Batch& b, int waves);
int batch_get_tile_waves(const Batch& b);
void batch_counts(const Batch& b, long long* out4);
void batch_kernel_times(const Batch& b, float* ms2);
double batch_autotune(Batch& b, const char* cache_path, int reps, int* from_cache);

inline bool is_complex(DType d) { return d == kC32 || d == kC64; }
inline size_t dtype_size(DType d) { return d == kF32 ? 4 : (d == kF64 || d == kC32) ? 8 : 16; }

}  // namespace bempp_amd
