"""Boundary assembly (python/bempp/assembly/boundary.py of the reference).

Same names, argument meaning and error behaviour: options are edited through
setters that raise ValueError for cell types without an entry, become read-only
once an assembler uses them, `assemble(trial_space, test_space)` returns a dense
Fortran-ordered matrix test x trial, `assemble_singular` a scipy CSR matrix of
the adjacent-pair part. Helmholtz assemblers (absent from the reference's python
layer, present in its Rust API src/helmholtz.rs) are created with
`create_helmholtz_assembler`. All assembly runs on the GPU.
"""
import ctypes
import typing
from enum import Enum

import numpy as np
from scipy.sparse import csr_matrix

from .. import _lib as L
from ..element import ReferenceCellType
from ..function_space import FunctionSpace

DTYPE_F32, DTYPE_F64, DTYPE_C32, DTYPE_C64 = 0, 1, 2, 3
_rtypes = {np.float32: DTYPE_F32, np.float64: DTYPE_F64, np.complex64: DTYPE_C32, np.complex128: DTYPE_C64}
_dtypes = {j: i for i, j in _rtypes.items()}


def _rtype(dtype):
    return _rtypes[np.dtype(dtype).type]


class OperatorType(Enum):
    """Operator type (src/bindings/boundary_assemblers.rs:15-24)."""

    SingleLayer = 0
    DoubleLayer = 1
    AdjointDoubleLayer = 2
    Hypersingular = 3
    ElectricField = 4
    MagneticField = 5


class Precision(Enum):
    """Arithmetic of the device kernels (no counterpart in the reference)."""

    FP64 = 0
    Mixed = 1


def _convert_to_scipy(m, dtype) -> csr_matrix:
    """Convert a native CSR matrix handle to scipy (python/bempp/assembly/boundary.py:22-32)."""
    lib = L.lib()
    try:
        shape = (lib.matrix_shape0(m), lib.matrix_shape1(m))
        data = np.empty(lib.matrix_csr_data_size(m), dtype=dtype)
        indices = np.empty(lib.matrix_csr_indices_size(m), dtype=np.uintp)
        indptr = np.empty(lib.matrix_csr_indptr_size(m), dtype=np.uintp)
        lib.matrix_csr_data_copy(m, L.ptr(data))
        lib.matrix_csr_indices_copy(m, L.ptr(indices))
        lib.matrix_csr_indptr_copy(m, L.ptr(indptr))
    finally:
        lib.matrix_t_free(m)
    return csr_matrix((data, indices.astype(np.int64), indptr.astype(np.int64)), shape=shape)


class BoundaryAssemblerOptions:
    """Boundary assembler options (python/bempp/assembly/boundary.py:46-107)."""

    def __init__(self):
        self._read_only = False
        self._rs_options = L.check_ptr(L.lib().boundary_assembler_options_new())

    def __del__(self, _release=L.release):
        _release("boundary_assembler_options_t_free", getattr(self, "_rs_options", None))
        self._rs_options = None

    def set_regular_quadrature_degree(self, cell: ReferenceCellType, degree: int):
        """Set the (non-singular) quadrature degree (number of points) for a cell type."""
        if self._read_only:
            raise ValueError("Cannot edit options already being used by an assembler")
        if not L.lib().boundary_assembler_options_has_regular_quadrature_degree(self._rs_options, int(cell)):
            raise ValueError(f"Invalid cell type: {cell}")
        L.lib().boundary_assembler_options_set_regular_quadrature_degree(self._rs_options, int(cell), degree)

    def regular_quadrature_degree(self, cell: ReferenceCellType) -> int:
        if not L.lib().boundary_assembler_options_has_regular_quadrature_degree(self._rs_options, int(cell)):
            raise ValueError(f"Invalid cell type: {cell}")
        return int(L.lib().boundary_assembler_options_get_regular_quadrature_degree(self._rs_options, int(cell)))

    def set_singular_quadrature_degree(self, cell0: ReferenceCellType, cell1: ReferenceCellType, degree: int):
        """Set the singular quadrature degree (Gauss points per dimension) for a cell pair."""
        if self._read_only:
            raise ValueError("Cannot edit options already being used by an assembler")
        if not L.lib().boundary_assembler_options_has_singular_quadrature_degree(self._rs_options, int(cell0),
                                                                                int(cell1)):
            raise ValueError(f"Invalid cell pair: {cell0} and {cell1}")
        L.lib().boundary_assembler_options_set_singular_quadrature_degree(self._rs_options, int(cell0), int(cell1),
                                                                          degree)

    def singular_quadrature_degree(self, cell0: ReferenceCellType, cell1: ReferenceCellType) -> int:
        if not L.lib().boundary_assembler_options_has_singular_quadrature_degree(self._rs_options, int(cell0),
                                                                                int(cell1)):
            raise ValueError(f"Invalid cell pair: {cell0} and {cell1}")
        return int(L.lib().boundary_assembler_options_get_singular_quadrature_degree(self._rs_options, int(cell0),
                                                                                    int(cell1)))

    def set_batch_size(self, batch_size: int):
        if self._read_only:
            raise ValueError("Cannot edit options already being used by an assembler")
        L.lib().boundary_assembler_options_set_batch_size(self._rs_options, batch_size)

    def batch_size(self) -> int:
        return int(L.lib().boundary_assembler_options_get_batch_size(self._rs_options))

    def set_precision(self, precision: Precision):
        """FP64 (default) or Mixed (FP32 far field, FP64 near/singular pairs)."""
        if self._read_only:
            raise ValueError("Cannot edit options already being used by an assembler")
        L.check(L.lib().boundary_assembler_options_set_precision(self._rs_options, Precision(precision).value))

    def precision(self) -> Precision:
        return Precision(L.lib().boundary_assembler_options_get_precision(self._rs_options))


def _torch_stream_handle(stream):
    if stream is not None:
        return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))
    import torch

    return ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))


class BoundaryAssembler:
    """Boundary assembler (python/bempp/assembly/boundary.py:110-174)."""

    def __init__(self, options: BoundaryAssemblerOptions, rs_assembler, owned: bool = True):
        self.options = options
        self.options._read_only = True
        self._rs_assembler = rs_assembler
        self._owned = owned

    def __del__(self, _release=L.release):
        if getattr(self, "_owned", False):
            _release("boundary_assembler_t_free", getattr(self, "_rs_assembler", None))
        self._rs_assembler = None

    @property
    def dtype(self):
        return _dtypes[L.lib().boundary_assembler_dtype(self._rs_assembler)]

    def _check_spaces(self, trial_space, test_space):
        for s in (trial_space, test_space):
            if not isinstance(s, FunctionSpace):
                raise ValueError("spaces must be bempp_amd FunctionSpace objects")
            if s.dtype is not None and np.dtype(s.dtype) != np.dtype(self.dtype):
                raise ValueError(f"space dtype {s.dtype} does not match assembler dtype {self.dtype}")

    def assemble(self, trial_space: FunctionSpace, test_space: FunctionSpace) -> np.ndarray:
        """Assemble the operator into a dense test x trial matrix (Fortran order)."""
        self._check_spaces(trial_space, test_space)
        output = np.zeros((test_space.global_size, trial_space.global_size), dtype=self.dtype, order="F")
        self.assemble_into_memory(trial_space, test_space, output)
        return output

    def assemble_into_memory(self, trial_space, test_space, output: np.ndarray):
        """Accumulate the operator into a Fortran-ordered host array (src/boundary_assemblers.rs:176-214)."""
        self._check_spaces(trial_space, test_space)
        if (output.dtype != np.dtype(self.dtype) or not output.flags.f_contiguous
                or output.shape != (test_space.global_size, trial_space.global_size)):
            raise ValueError("output must be a Fortran-ordered test x trial array of the assembler dtype")
        L.check(L.lib().boundary_assembler_assemble_into_memory(self._rs_assembler, trial_space._rs_space,
                                                                 test_space._rs_space, L.ptr(output)))
        return output

    def assemble_into_device(self, trial_space, test_space, out, accumulate=False, stream=None):
        """Assemble into a device tensor (torch, on the current GPU): `out` has shape
        (n_test, n_trial) with strides (1, n_test) (column-major), or is a flat
        buffer of n_test*n_trial elements. Work is enqueued on `stream` (default:
        torch's current stream)."""
        import torch

        self._check_spaces(trial_space, test_space)
        nt, ns = test_space.global_size, trial_space.global_size
        tdt = {np.float64: torch.float64, np.float32: torch.float32, np.complex128: torch.complex128,
               np.complex64: torch.complex64}[self.dtype]
        if not out.is_cuda or out.dtype != tdt:
            raise ValueError(f"out must be a {tdt} tensor on the GPU")
        if out.dim() == 2:
            if tuple(out.shape) != (nt, ns) or (nt > 1 and out.stride(0) != 1) or (ns > 1 and out.stride(1) != nt):
                raise ValueError("out must be a column-major (n_test, n_trial) tensor")
        elif out.dim() != 1 or out.numel() != nt * ns or not out.is_contiguous():
            raise ValueError("out must be a contiguous flat buffer of n_test*n_trial elements")
        L.check(L.lib().boundary_assembler_assemble_into_device_memory(
            self._rs_assembler, trial_space._rs_space, test_space._rs_space, ctypes.c_void_p(out.data_ptr()),
            1 if accumulate else 0, _torch_stream_handle(stream)))
        return out

    def assemble_device(self, trial_space, test_space, device=None, stream=None):
        """Allocate a column-major device tensor and assemble into it."""
        import torch

        nt, ns = test_space.global_size, trial_space.global_size
        tdt = {np.float64: torch.float64, np.float32: torch.float32, np.complex128: torch.complex128,
               np.complex64: torch.complex64}[self.dtype]
        buf = torch.empty((ns, nt), dtype=tdt, device=device or "cuda").t()
        return self.assemble_into_device(trial_space, test_space, buf, accumulate=False, stream=stream)

    def set_profiling(self, enable=True):
        """Record HIP events around the device kernels of each assembly."""
        L.check(L.lib().boundary_assembler_set_profiling(self._rs_assembler, 1 if enable else 0))

    def kernel_times(self):
        """(regular_ms, adjacent_ms) of the last profiled assembly (waits for the kernels)."""
        ms = np.zeros(2, dtype=np.float32)
        L.check(L.lib().boundary_assembler_kernel_times(self._rs_assembler, L.ptr(ms)))
        return float(ms[0]), float(ms[1])

    def assemble_singular(self, trial_space: FunctionSpace, test_space: FunctionSpace) -> csr_matrix:
        """Assemble the singular part of an operator into a CSR matrix."""
        self._check_spaces(trial_space, test_space)
        m = L.check_ptr(L.lib().boundary_assembler_assemble_singular(self._rs_assembler, trial_space._rs_space,
                                                                      test_space._rs_space))
        return _convert_to_scipy(m, self.dtype)

    def assemble_singular_correction(self, trial_space: FunctionSpace, test_space: FunctionSpace) -> csr_matrix:
        """Regular-rule contributions of the adjacent cell pairs as a CSR matrix
        (python/bempp/assembly/boundary.py:157-169): dense = regular(all pairs) - this + singular."""
        self._check_spaces(trial_space, test_space)
        m = L.check_ptr(L.lib().boundary_assembler_assemble_singular_correction(
            self._rs_assembler, trial_space._rs_space, test_space._rs_space))
        return _convert_to_scipy(m, self.dtype)

    def apply(self, trial_space, test_space, u, transpose=False):
        """y = A u (A^T u with transpose) without assembling A (P0 spaces): the operator is
        recomputed on the GPU."""
        self._check_spaces(trial_space, test_space)
        u = np.ascontiguousarray(u, dtype=self.dtype)
        if u.shape != (trial_space.global_size,):
            raise ValueError("u must have n_trial entries")
        y = np.zeros(test_space.global_size, dtype=self.dtype)
        L.check(L.lib().boundary_assembler_apply(self._rs_assembler, trial_space._rs_space, test_space._rs_space,
                                                  L.ptr(u), L.ptr(y), 1 if transpose else 0))
        return y

    def shape_gradient(self, trial_space, test_space, u, v):
        """d(v^T A(X) u)/dX for the grid's vertex coordinates X, shape [npoints, 3]
        (real for Laplace, complex for Helmholtz). Matrix-free on the GPU."""
        self._check_spaces(trial_space, test_space)
        u = np.ascontiguousarray(u, dtype=self.dtype)
        v = np.ascontiguousarray(v, dtype=self.dtype)
        if u.shape != (trial_space.global_size,) or v.shape != (test_space.global_size,):
            raise ValueError("u must have n_trial and v n_test entries")
        npts = L.lib().grid_npoints(test_space.grid._handle)
        gdt = np.float64 if np.dtype(self.dtype).kind == "f" else np.complex128
        grad = np.zeros((npts, 3), dtype=gdt)
        L.check(L.lib().boundary_assembler_shape_gradient(self._rs_assembler, trial_space._rs_space,
                                                           test_space._rs_space, L.ptr(u), L.ptr(v), L.ptr(grad)))
        return grad


def create_laplace_assembler(operator_type: OperatorType, options: typing.Optional[BoundaryAssemblerOptions] = None,
                             dtype=np.float64):
    """Create a Laplace assembler (python/bempp/assembly/boundary.py:177-188)."""
    options = BoundaryAssemblerOptions() if options is None else options
    h = L.check_ptr(L.lib().laplace_boundary_assembler_new(options._rs_options, OperatorType(operator_type).value,
                                                           _rtype(dtype)))
    return BoundaryAssembler(options, h)


def create_helmholtz_assembler(wavenumber: float, operator_type: OperatorType,
                               options: typing.Optional[BoundaryAssemblerOptions] = None, dtype=np.complex128):
    """Create a Helmholtz assembler (Rust: helmholtz::assembler::*(k, &options), src/helmholtz.rs:44-107)."""
    options = BoundaryAssemblerOptions() if options is None else options
    rt = _rtype(dtype)
    k = np.array([wavenumber], dtype=np.float32 if rt == DTYPE_C32 else np.float64)
    h = L.check_ptr(L.lib().helmholtz_boundary_assembler_new(L.ptr(k), options._rs_options,
                                                             OperatorType(operator_type).value, rt))
    return BoundaryAssembler(options, h)
