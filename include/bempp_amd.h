/* bempp_amd.h -- C ABI of the MI355X boundary-element assembly engine.
 *
 * The entry points mirror the reference's C bindings for dense and singular
 * boundary-operator assembly (src/bindings/boundary_assemblers.rs,
 * src/bindings/matrices.rs; those bindings are disabled in the reference,
 * src/lib.rs:5, and the shape is kept). Each declaration cites the reference
 * function it replaces. Additions with no counterpart in the reference say so.
 *
 * Conventions
 *  - Handles are opaque and owned by the caller; every *_new/_create has a
 *    matching *_t_free that accepts NULL.
 *  - Dense matrices are column-major (Fortran order): entry (test dof i,
 *    trial dof j) lives at data[i + n_test * j]; the trial space is passed
 *    before the test space, as in the reference (src/boundary_assemblers.rs:158-214).
 *  - assemble_into_memory ACCUMULATES (+=) into the caller's buffer
 *    (src/boundary_assemblers.rs:176-214).
 *  - Where the reference panics, these functions return a non-zero status
 *    (or NULL) and set a message readable with bempp_last_error().
 *  - Assembly runs on the current HIP device; there is no CPU fallback: with
 *    no usable GPU the assembly entry points fail with an error.
 */
#ifndef BEMPP_AMD_H
#define BEMPP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ndelement ReferenceCellType numbering (python/bempp/function_space.py:211-222). */
enum {
  BEMPP_CELL_POINT = 0,
  BEMPP_CELL_INTERVAL = 1,
  BEMPP_CELL_TRIANGLE = 2,
  BEMPP_CELL_QUADRILATERAL = 3,
  BEMPP_CELL_TETRAHEDRON = 4,
  BEMPP_CELL_HEXAHEDRON = 5,
  BEMPP_CELL_PRISM = 6,
  BEMPP_CELL_PYRAMID = 7
};

/* DType (python/bempp/assembly/boundary.py:13-18; F32=0, F64=1 at function_space.py:13-16). */
enum { BEMPP_DTYPE_F32 = 0, BEMPP_DTYPE_F64 = 1, BEMPP_DTYPE_C32 = 2, BEMPP_DTYPE_C64 = 3 };

/* OperatorType, #[repr(u8)] (src/bindings/boundary_assemblers.rs:15-24). */
enum {
  BEMPP_OP_SINGLE_LAYER = 0,
  BEMPP_OP_DOUBLE_LAYER = 1,
  BEMPP_OP_ADJOINT_DOUBLE_LAYER = 2,
  BEMPP_OP_HYPERSINGULAR = 3,
  BEMPP_OP_ELECTRIC_FIELD = 4,
  BEMPP_OP_MAGNETIC_FIELD = 5
};

/* Continuity (ndelement; python/bempp/function_space.py:335-338). */
enum { BEMPP_CONTINUITY_STANDARD = 0, BEMPP_CONTINUITY_DISCONTINUOUS = 1 };

/* Precision modes (no counterpart in the reference). FP64: every pair in FP64.
 * MIXED: regular far-field pairs evaluated in FP32 on pair-recentred
 * coordinates with FP64 coordinate differences and phase reduction; Duffy and
 * near pairs stay FP64; accumulation FP64. */
enum { BEMPP_PRECISION_FP64 = 0, BEMPP_PRECISION_MIXED = 1 };

typedef struct BoundaryAssemblerOptionsT BoundaryAssemblerOptionsT;
typedef struct BoundaryAssemblerT BoundaryAssemblerT;
typedef struct GridT GridT;
typedef struct SpaceT SpaceT;
typedef struct MatrixT MatrixT;
typedef struct ElementFamilyT ElementFamilyT;

/* Last error message of the calling thread (no counterpart: the reference panics). */
const char* bempp_last_error(void);

/* Library version string (no counterpart in the reference). */
const char* bempp_version(void);

/* Number of visible HIP devices; 0 when no GPU is usable (no counterpart). */
int bempp_device_count(void);

/* ---- options: src/bindings/boundary_assemblers.rs:34-125 ---- */
/* replaces boundary_assembler_options_new (:34-40) */
BoundaryAssemblerOptionsT* boundary_assembler_options_new(void);
/* replaces boundary_assembler_options_t_free (:42-50) */
void boundary_assembler_options_t_free(BoundaryAssemblerOptionsT* options);
/* replaces boundary_assembler_options_set_regular_quadrature_degree (:52-60); like
 * the reference it only changes cell types that already have an entry (:61-65). */
void boundary_assembler_options_set_regular_quadrature_degree(BoundaryAssemblerOptionsT* options, uint8_t cell_type,
                                                               size_t npoints);
/* replaces ..._get_regular_quadrature_degree (:61-68); returns 0 when absent (reference panics). */
size_t boundary_assembler_options_get_regular_quadrature_degree(const BoundaryAssemblerOptionsT* options,
                                                                uint8_t cell_type);
/* replaces ..._has_regular_quadrature_degree (:69-76) */
int boundary_assembler_options_has_regular_quadrature_degree(const BoundaryAssemblerOptionsT* options,
                                                             uint8_t cell_type);
/* replaces ..._set_singular_quadrature_degree (:78-87) */
void boundary_assembler_options_set_singular_quadrature_degree(BoundaryAssemblerOptionsT* options,
                                                                uint8_t cell_type0, uint8_t cell_type1,
                                                                size_t npoints);
/* replaces ..._get_singular_quadrature_degree (:88-98); returns 0 when absent. */
size_t boundary_assembler_options_get_singular_quadrature_degree(const BoundaryAssemblerOptionsT* options,
                                                                 uint8_t cell_type0, uint8_t cell_type1);
/* replaces ..._has_singular_quadrature_degree (:99-109) */
int boundary_assembler_options_has_singular_quadrature_degree(const BoundaryAssemblerOptionsT* options,
                                                              uint8_t cell_type0, uint8_t cell_type1);
/* replaces ..._set_batch_size (:111-118) */
void boundary_assembler_options_set_batch_size(BoundaryAssemblerOptionsT* options, size_t batch_size);
/* replaces ..._get_batch_size (:119-125) */
size_t boundary_assembler_options_get_batch_size(const BoundaryAssemblerOptionsT* options);
/* Precision mode (no counterpart in the reference); returns non-zero on a bad mode. */
int boundary_assembler_options_set_precision(BoundaryAssemblerOptionsT* options, uint8_t precision);
uint8_t boundary_assembler_options_get_precision(const BoundaryAssemblerOptionsT* options);

/* ---- grids and spaces (the reference binds ndgrid/ndelement objects,
 *      src/bindings/ndgrid.rs:16-56, src/bindings/space.rs:20-42; here a grid is
 *      built from plain arrays of flat triangles) ---- */
/* A grid of flat triangles: points [npoints][3] f64, cells [ncells][3] vertex ids. */
GridT* grid_new(const double* points, size_t npoints, const int64_t* cells, size_t ncells);
/* Octahedron sphere of src/shapes.rs:19-114 (seed 0) or an icosahedron refined
 * by the same scheme (seed 1). */
GridT* grid_new_sphere(int seed, uint32_t refinement_level);
void grid_t_free(GridT* grid);
size_t grid_npoints(const GridT* grid);
size_t grid_ncells(const GridT* grid);
void grid_points_copy(const GridT* grid, double* points);
/* Triangle grids: cells [ncells][3]. (A grid with quadrilateral cells copies nothing here;
 * use grid_cell_vertices_copy.) */
void grid_cells_copy(const GridT* grid, int64_t* cells);
/* A grid of triangle and quadrilateral cells (ndgrid's single-element quadrilateral grids and
 * mixed grids, degree-1 geometry; no counterpart binding in the reference, which builds them
 * through ndgrid's MixedGridBuilder, tests/mixed_assembly.rs.bak:15-31): points [npoints][3]
 * f64; cell_types[c] BEMPP_CELL_TRIANGLE (3 vertices) or BEMPP_CELL_QUADRILATERAL (4
 * vertices, tensor order: reference (0,0), (1,0), (0,1), (1,1)); cell_vertices the cells'
 * vertices back to back. Dense assembly (assemble_into_memory, assemble_into_device_memory)
 * and the sparse singular part and correction (assemble_singular,
 * assemble_singular_correction) support these grids in FP64 (the regular rule of
 * quadrilaterals is m x m Gauss: its point count must be a square); the other entry points
 * refuse them. */
GridT* grid_new_cells(const double* points, size_t npoints, const int64_t* cell_vertices, const uint8_t* cell_types,
                      size_t ncells);
/* Number of entries of the cells' vertex lists, the lists back to back, and the cell types. */
size_t grid_cell_vertices_size(const GridT* grid);
void grid_cell_vertices_copy(const GridT* grid, int64_t* cell_vertices);
void grid_cell_types_copy(const GridT* grid, uint8_t* cell_types);
/* Replace the vertex coordinates (same topology); drops cached device geometry. */
int grid_set_points(GridT* grid, const double* points);

/* The entry points the reference's python layer binds for grids, element families and
 * spaces (python/bempp/function_space.py:55-64,143-191), under the same names. */
/* replaces single_element_grid_borrowed_create_bempp (src/bindings/ndgrid.rs:16-56): a grid
 * from the arrays of an ndgrid single-element grid: triangles (3 points per cell,
 * entity_types[2] triangle) or quadrilaterals (4 points per cell, entity_types[2]
 * quadrilateral), tdim 2, gdim 3, geometry degree 1; points f32 or f64 [npoints][3], cells
 * [ncells][points_per_cell] point indices. Identifier and upward-connectivity arrays
 * are not used; the cell-to-vertex topology (downward_connectivity[2][0], optional) must be a
 * relabelling of the cells' points. The arrays are copied. NULL on error. */
GridT* single_element_grid_borrowed_create_bempp(size_t tdim, const size_t* id_sizes, const size_t* const* id_pointers,
                                                 const uint8_t* entity_types, const size_t* entity_counts,
                                                 const size_t* const* const* downward_connectivity,
                                                 const size_t* const* downward_connectivity_shape0,
                                                 const size_t* const* const* const* upward_connectivity,
                                                 const size_t* const* const* upward_connectivity_lens,
                                                 const void* points, size_t gdim, size_t npoints, uint8_t dtype,
                                                 const size_t* cells, size_t points_per_cell, size_t ncells,
                                                 size_t geometry_degree);
/* replaces grid_t_free_bempp (src/bindings/ndgrid.rs:6-14) */
void grid_t_free_bempp(GridT* grid);
/* replaces create_lagrange_family_bempp (src/bindings/ndelement.rs:18-25); degree 0
 * (discontinuous) or 1 are the spaces function_space accepts. */
ElementFamilyT* create_lagrange_family_bempp(size_t degree, uint8_t continuity, uint8_t dtype);
/* replace create_raviart_thomas_family_bempp / create_nedelec_family_bempp
 * (src/bindings/ndelement.rs:27-44): vector-valued spaces are not supported; NULL. */
ElementFamilyT* create_raviart_thomas_family_bempp(size_t degree, uint8_t continuity, uint8_t dtype);
ElementFamilyT* create_nedelec_family_bempp(size_t degree, uint8_t continuity, uint8_t dtype);
/* replaces element_family_t_free_bempp (src/bindings/ndelement.rs:8-16) */
void element_family_t_free_bempp(ElementFamilyT* family);
/* replaces function_space(grid, family) (src/bindings/space.rs:20-42); the space borrows
 * the grid (free the grid after its spaces). */
SpaceT* function_space(const GridT* grid, const ElementFamilyT* family);
/* function_space with the Lagrange family given by degree (0 or 1) and continuity. */
SpaceT* function_space_new(const GridT* grid, size_t degree, uint8_t continuity);
/* replaces space_local_space (src/bindings/space.rs:44-56): a serial space is its own
 * local space (a new handle to the same space). */
SpaceT* space_local_space(const SpaceT* space);
void space_t_free(SpaceT* space);
/* replaces space_get_local_dof_numbers_size / space_get_local_dof_numbers
 * (src/bindings/space.rs:69-107): DOFs of entity (entity_dim, entity_index) -- a vertex of a
 * continuous P1 space, a cell of a P0 or discontinuous P1 space; none otherwise. */
size_t space_get_local_dof_numbers_size(const SpaceT* space, size_t entity_dim, size_t entity_index);
void space_get_local_dof_numbers(const SpaceT* space, size_t entity_dim, size_t entity_index, uintptr_t* dofs);
/* replaces space_is_owned (src/bindings/space.rs:178-187): a serial space owns every DOF.
 * space_ownership_process / space_ownership_index are what python/bempp/function_space.py:96-103
 * calls for a ghost DOF (the reference's Rust bindings do not define them): rank 0, the index. */
int space_is_owned(const SpaceT* space, size_t local_dof_index);
size_t space_ownership_process(const SpaceT* space, size_t local_dof_index);
size_t space_ownership_index(const SpaceT* space, size_t local_dof_index);
/* replaces space_dtype (src/bindings/space.rs:189-198): the element family's dtype (F64 for
 * function_space_new). */
uint8_t space_dtype(const SpaceT* space);
/* Space accessors (python/bempp/function_space.py:256-311) */
size_t space_global_size(const SpaceT* space);
size_t space_local_size(const SpaceT* space);
int space_is_serial(const SpaceT* space);
size_t space_cell_dofs_size(const SpaceT* space, size_t cell);
int space_has_cell_dofs(const SpaceT* space, size_t cell);
void space_cell_dofs(const SpaceT* space, size_t cell, uintptr_t* dofs);
size_t space_global_dof_index(const SpaceT* space, size_t local_dof_index);
/* Greedy cell colouring of src/function.rs:281-351: colour of each cell, and count. */
size_t space_cell_colouring(const SpaceT* space, int64_t* cell_colour);

/* ---- assemblers: src/bindings/boundary_assemblers.rs:127-347 ---- */
/* replaces laplace_boundary_assembler_new (:127-166); dtype F32 or F64. */
BoundaryAssemblerT* laplace_boundary_assembler_new(const BoundaryAssemblerOptionsT* options, uint8_t operator_type,
                                                   uint8_t dtype);
/* replaces helmholtz_boundary_assembler_new (:168-210); k points to a float
 * (C32) or double (C64). */
BoundaryAssemblerT* helmholtz_boundary_assembler_new(const void* k, const BoundaryAssemblerOptionsT* options,
                                                     uint8_t operator_type, uint8_t dtype);
/* replaces boundary_assembler_t_free (generated by #[cfuncs], :26-27) */
void boundary_assembler_t_free(BoundaryAssemblerT* assembler);
/* replaces boundary_assembler_dtype (:212-233) */
uint8_t boundary_assembler_dtype(const BoundaryAssemblerT* assembler);

/* replaces boundary_assembler_assemble_into_memory (:309-347): host buffer of
 * n_test*n_trial elements of the assembler's dtype, Fortran order, accumulated into. */
int boundary_assembler_assemble_into_memory(const BoundaryAssemblerT* assembler, const SpaceT* trial_space,
                                            const SpaceT* test_space, void* data);

/* Device-resident variant (no counterpart in the reference): `data` is device
 * memory on the current HIP device. accumulate != 0 adds into it, otherwise it
 * is overwritten. `stream` is a hipStream_t (NULL: the device's null stream).
 * The call returns once the work is enqueued; synchronize the stream to wait. */
int boundary_assembler_assemble_into_device_memory(const BoundaryAssemblerT* assembler, const SpaceT* trial_space,
                                                   const SpaceT* test_space, void* data, int accumulate,
                                                   void* stream);

/* replaces boundary_assembler_assemble (:272-307): dense matrix handle. */
MatrixT* boundary_assembler_assemble(const BoundaryAssemblerT* assembler, const SpaceT* trial_space,
                                     const SpaceT* test_space);
/* replaces boundary_assembler_assemble_singular (:235-270): CSR matrix of the
 * adjacent-pair (Duffy) contributions with duplicates summed; empty when every
 * |entry| <= 1e-10 (src/boundary_assemblers.rs:122-155). */
MatrixT* boundary_assembler_assemble_singular(const BoundaryAssemblerT* assembler, const SpaceT* trial_space,
                                              const SpaceT* test_space);

/* Singular correction (python/bempp/assembly/boundary.py:157-169 binds
 * boundary_assembler_assemble_singular_correction; the Rust side has no implementation):
 * CSR matrix of the regular-rule contributions of the adjacent pairs, i.e. what a
 * regular-rule evaluation over all point pairs (an FMM, tests/fmm.rs.bak:22-95) adds at
 * the pairs the Duffy rules integrate, so that dense = regular(all) - correction +
 * singular. Point pairs at zero distance contribute nothing. Same CSR conventions as
 * assemble_singular. */
MatrixT* boundary_assembler_assemble_singular_correction(const BoundaryAssemblerT* assembler,
                                                         const SpaceT* trial_space, const SpaceT* test_space);

/* Potential (off-surface) operators -- the PotentialAssembler of the reference's commented
 * tests (tests/compare_to_bempp_cl.rs:208-320; src/boundary_evaluators.rs is empty there):
 * P[j][i] = integral over the grid of K(x_j, y) psi_i(y), K = G (single layer) or
 * grad_y G . n_y (double layer), with the assembler's kernel, wavenumber, dtype and regular
 * rule; psi_i the basis of `space`; points [npoints][3] f64. The matrix is npoints x ndofs,
 * Fortran order (entry (j, i) at data[j + npoints * i]).
 * into_memory accumulates into a host buffer; requires |k| times the extent of mesh and
 * points <= 1e8. */
int potential_assemble_into_memory(const BoundaryAssemblerT* assembler, const SpaceT* space, const double* points,
                                   size_t npoints, void* data);
/* Device variant (no counterpart): points and data in device memory on the current device,
 * accumulate != 0 adds, otherwise overwrites; enqueued on `stream` (NULL: null stream). */
int potential_assemble_into_device_memory(const BoundaryAssemblerT* assembler, const SpaceT* space,
                                          const double* points, size_t npoints, void* data, int accumulate,
                                          void* stream);

/* Matrix-free operator application y = A u (no counterpart in the reference; the
 * FMM prototype of tests/fmm.rs.bak applies operators this way): A is recomputed on
 * the device and never stored. P0 spaces; u has n_trial and y n_test entries of the
 * assembler's dtype; y is overwritten. transpose != 0 applies A^T (single-layer and
 * hypersingular operators on one space). */
int boundary_assembler_apply(const BoundaryAssemblerT* assembler, const SpaceT* trial_space,
                             const SpaceT* test_space, const void* u, void* y, int transpose);

/* Shape gradient (no counterpart in the reference): grad[v][d] = d(v^T A(X) u)/dX[v][d]
 * for the single-layer operator of `assembler` on P0 spaces of one grid, computed
 * matrix-free on the device with recomputation. u has n_trial,
 * v n_test entries of the assembler's dtype; grad is real f64 [npoints][3] for
 * Laplace and complex128 [npoints][3] for Helmholtz (d/dX of the complex bilinear form). */
int boundary_assembler_shape_gradient(const BoundaryAssemblerT* assembler, const SpaceT* trial_space,
                                      const SpaceT* test_space, const void* u, const void* v, void* grad);

/* Shape gradient that allocates no device memory. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The same result as boundary_assembler_shape_gradient, with u, v
 * (assembler dtype) and grad (f64 or complex128 [npoints][3], overwritten) in device memory
 * on the current device; the work is enqueued on `stream` (a hipStream_t; NULL: the null
 * stream) and the call returns without waiting. Its partial results live in one static
 * 16 MiB device buffer per GPU (176 B per cell): grids of at most 95325 cells, larger ones
 * are rejected with an error. Calls on one GPU share that buffer, so calls that overlap in
 * time on one GPU race: order them on one stream. The grid's, space's and rule's device
 * data must exist already (any ordinary assembly, apply or shape-gradient call with the same
 * assembler and spaces builds them); otherwise the call fails instead of allocating. */
int boundary_assembler_shape_gradient_noalloc(const BoundaryAssemblerT* assembler, const SpaceT* trial_space,
                                              const SpaceT* test_space, const void* u, const void* v, void* grad,
                                              void* stream);

/* Kernel timing (no counterpart in the reference): with profiling enabled, each
 * assembly records HIP events around its regular-pair and adjacent-pair kernels
 * on the launch stream; kernel_times waits for them and writes the elapsed
 * milliseconds (ms[0] regular, ms[1] adjacent; -1 when a kernel did not run). */
int boundary_assembler_set_profiling(BoundaryAssemblerT* assembler, int enable);
int boundary_assembler_kernel_times(const BoundaryAssemblerT* assembler, float* ms);

/* ---- batched assembly (no counterpart in the reference: it assembles one operator
 *      per call) ----
 * n items, item i = assemblers[i] on the P0 space spaces[i] (test = trial), written to
 * the device buffer outputs[i] (n_i x n_i, Fortran order, the assembler's dtype, on the
 * current device). Items share kernel, operator, dtype, precision and regular rule;
 * wavenumbers, grids and singular rules may differ; every grid holds triangles only
 * (an item on a grid with quadrilateral cells is refused). All items run in one persistent
 * launch (plus one launch for the adjacent pairs, and one for the near pairs in mixed
 * precision). The batch keeps device pointers into the grids' caches: free it before
 * its grids, spaces or assemblers; after grid_set_points on one of its grids,
 * assembling fails until a new batch is built. NULL on error. */
typedef struct BoundaryAssemblerBatchT BoundaryAssemblerBatchT;
BoundaryAssemblerBatchT* boundary_assembler_batch_new(const BoundaryAssemblerT* const* assemblers,
                                                      const SpaceT* const* spaces, void* const* outputs, size_t n);
void boundary_assembler_batch_t_free(BoundaryAssemblerBatchT* batch);
/* Overwrites every output. stream: a hipStream_t (NULL: the device's null stream).
 * Returns once enqueued. Launches of one batch must not overlap in time (they share
 * its tile counter): order them on one stream. */
int boundary_assembler_batch_assemble(BoundaryAssemblerBatchT* batch, void* stream);
/* enable != 0: capture the launch sequence into a hipGraph on the next assemble and
 * replay it afterwards (recaptured when the launch shape changes). */
int boundary_assembler_batch_set_graph(BoundaryAssemblerBatchT* batch, int enable);
/* Launch shape: blocks_per_cu workgroups per compute unit (0: one workgroup per tile),
 * trial chunk 4 or 8, tile order 0 (item-major) or 1 (largest items first). */
int boundary_assembler_batch_set_launch(BoundaryAssemblerBatchT* batch, int blocks_per_cu, int chunk, int order);
int boundary_assembler_batch_get_launch(const BoundaryAssemblerBatchT* batch, int* blocks_per_cu, int* chunk,
                                        int* order);
/* Tile split: 2, 4 (default) or 8 wavefronts per 64 x 64 pair tile (workgroups of 64 x waves
 * threads, each wavefront covering 64 / waves trial cells); 2 and 8 need trial chunk 8. */
int boundary_assembler_batch_set_tile_waves(BoundaryAssemblerBatchT* batch, int waves);
int boundary_assembler_batch_get_tile_waves(const BoundaryAssemblerBatchT* batch, int* waves);
/* counts[4] = items, pair tiles, adjacent pairs, near pairs */
int boundary_assembler_batch_counts(const BoundaryAssemblerBatchT* batch, int64_t* counts);
/* Times candidate launch shapes (best of `reps` launches each; outputs are
 * overwritten) and keeps the fastest; with cache_path, a JSON-lines file of earlier
 * results keyed by device and batch signature is read first and appended to.
 * Returns the best time in ms (the cached one when *from_cache is set), -1 on error. */
double boundary_assembler_batch_autotune(BoundaryAssemblerBatchT* batch, const char* cache_path, int reps,
                                         int* from_cache);
/* Profiling: HIP events around the last eager launch: ms[0] tiles (and near pairs),
 * ms[1] adjacent pairs; for a graph replay ms[0] is the whole replay and ms[1] -1. */
int boundary_assembler_batch_set_profiling(BoundaryAssemblerBatchT* batch, int enable);
int boundary_assembler_batch_kernel_times(const BoundaryAssemblerBatchT* batch, float* ms);

/* ---- matrices: src/bindings/matrices.rs:6-138 ---- */
size_t matrix_shape0(const MatrixT* matrix);
size_t matrix_shape1(const MatrixT* matrix);
/* 0 dense, 1 CSR (no counterpart: the reference dispatches on the boxed type) */
int matrix_is_csr(const MatrixT* matrix);
uint8_t matrix_dtype(const MatrixT* matrix);
const void* matrix_dense_data(const MatrixT* matrix);
void matrix_dense_data_copy(const MatrixT* matrix, void* data);
size_t matrix_csr_data_size(const MatrixT* matrix);
size_t matrix_csr_indices_size(const MatrixT* matrix);
size_t matrix_csr_indptr_size(const MatrixT* matrix);
void matrix_csr_data_copy(const MatrixT* matrix, void* data);
void matrix_csr_indices_copy(const MatrixT* matrix, uintptr_t* indices);
void matrix_csr_indptr_copy(const MatrixT* matrix, uintptr_t* indptr);
void matrix_t_free(MatrixT* matrix);

/* ---- quadrature tables (no counterpart in the reference's bindings; they expose
 *      what ndelement::quadrature::simplex_rule and
 *      bempp_quadrature::duffy::triangle_duffy provide to src/boundary_assemblers.rs:419-464,524-562) ---- */
/* Symmetric rule on the reference triangle: xi [npoints][2], w [npoints]; returns
 * npoints, or -1 for an unsupported count (supported: 1, 3, 6, 7). */
int bempp_triangle_rule(int npoints, double* xi, double* w);
/* Canonical Sauter-Schwab rule, nshared = 1 (vertex), 2 (edge), 3 (coincident),
 * n Gauss points per dimension: pts5 [npts][5] = test xi0, xi1, trial xi0, xi1,
 * weight. Returns npts (pass pts5 = NULL to query it). */
int bempp_canonical_pair_rule(int nshared, int n, double* pts5);

#ifdef __cplusplus
}
#endif

#endif /* BEMPP_AMD_H */
