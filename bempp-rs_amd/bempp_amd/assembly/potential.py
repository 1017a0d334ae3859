"""Potential (off-surface) operators: P[j, i] = integral of K(x_j, y) psi_i(y) over the grid,
K = G (single layer) or grad_y G . n_y (double layer).

Mirrors the PotentialAssembler of the reference's commented tests
(tests/compare_to_bempp_cl.rs:208-320: `assembler.assemble_into_dense(&mut matrix, &space,
&points)` with matrix [npoints, ndofs]); the reference's src/boundary_evaluators.rs is
empty. The kernel, wavenumber, dtype and regular rule come from a boundary assembler.
"""
import ctypes
import typing

import numpy as np

from .. import _lib as L
from .boundary import (BoundaryAssembler, BoundaryAssemblerOptions, OperatorType, _torch_stream_handle,
                       create_helmholtz_assembler, create_laplace_assembler)


class PotentialAssembler:
    def __init__(self, boundary_assembler: BoundaryAssembler):
        self._a = boundary_assembler

    @property
    def dtype(self):
        return self._a.dtype

    def assemble(self, space, points) -> np.ndarray:
        """Dense potential matrix [npoints, ndofs] (Fortran order)."""
        points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((len(points), space.global_size), dtype=self.dtype, order="F")
        return self.assemble_into_memory(space, points, out)

    def assemble_into_memory(self, space, points, output: np.ndarray):
        """Accumulate into a Fortran-ordered [npoints, ndofs] host array."""
        points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        if (output.dtype != np.dtype(self.dtype) or not output.flags.f_contiguous
                or output.shape != (len(points), space.global_size)):
            raise ValueError("output must be a Fortran-ordered [npoints, ndofs] array of the assembler dtype")
        L.check(L.lib().potential_assemble_into_memory(self._a._rs_assembler, space._rs_space, L.ptr(points),
                                                       len(points), L.ptr(output)))
        return output

    def assemble_into_device(self, space, points, out, accumulate=False, stream=None):
        """Device variant: `points` a contiguous [npoints, 3] float64 GPU tensor, `out` a
        column-major [npoints, ndofs] GPU tensor of the assembler dtype."""
        import torch

        tdt = {np.float64: torch.float64, np.float32: torch.float32, np.complex128: torch.complex128,
               np.complex64: torch.complex64}[self.dtype]
        if not points.is_cuda or points.dtype != torch.float64 or not points.is_contiguous() or points.dim() != 2 \
                or points.shape[1] != 3:
            raise ValueError("points must be a contiguous [npoints, 3] float64 GPU tensor")
        n, m = points.shape[0], space.global_size
        if not out.is_cuda or out.dtype != tdt or tuple(out.shape) != (n, m) or (n > 1 and out.stride(0) != 1) \
                or (m > 1 and out.stride(1) != n):
            raise ValueError(f"out must be a column-major ({n}, {m}) {tdt} GPU tensor")
        L.check(L.lib().potential_assemble_into_device_memory(
            self._a._rs_assembler, space._rs_space, ctypes.c_void_p(points.data_ptr()), n,
            ctypes.c_void_p(out.data_ptr()), 1 if accumulate else 0, _torch_stream_handle(stream)))
        return out


def create_laplace_potential_assembler(operator_type: OperatorType,
                                       options: typing.Optional[BoundaryAssemblerOptions] = None, dtype=np.float64):
    """Laplace single- or double-layer potential."""
    return PotentialAssembler(create_laplace_assembler(operator_type, options, dtype))


def create_helmholtz_potential_assembler(wavenumber: float, operator_type: OperatorType,
                                         options: typing.Optional[BoundaryAssemblerOptions] = None,
                                         dtype=np.complex128):
    """Helmholtz single- or double-layer potential."""
    return PotentialAssembler(create_helmholtz_assembler(wavenumber, operator_type, options, dtype))
