"""Batched dense assembly: many independent P0 x P0 operators (each with its own grid,
wavenumber and output) in one persistent GPU launch, optionally replayed from a
captured hipGraph, with an autotuned launch shape.

The reference assembles one operator per call (src/boundary_assemblers.rs:158-214);
this module has no counterpart there. Each item is an assembler (sharing kernel,
operator, dtype, precision and regular rule with the others) with a DP0 space; the
outputs are column-major device tensors that every launch overwrites.
"""
import ctypes

import numpy as np

from .. import _lib as L
from .boundary import BoundaryAssembler, _torch_stream_handle
from ..function_space import FunctionSpace


def _torch_dtype(dt):
    import torch

    return {np.float64: torch.float64, np.float32: torch.float32, np.complex128: torch.complex128,
            np.complex64: torch.complex64}[dt]


class BatchAssembler:
    """Batch of dense assemblies. `assemblers[i]` on `spaces[i]` (trial = test) into
    `outputs[i]`; outputs default to new column-major tensors on the current GPU.

    The batch keeps references to its assemblers, spaces and outputs. It reads the
    grids' device geometry in place, so after `grid.set_points` a new batch must be
    built (assembling the old one raises)."""

    def __init__(self, assemblers, spaces, outputs=None, device=None):
        import torch

        if len(assemblers) != len(spaces) or len(assemblers) == 0:
            raise ValueError("need the same positive number of assemblers and spaces")
        for a, s in zip(assemblers, spaces):
            if not isinstance(a, BoundaryAssembler) or not isinstance(s, FunctionSpace):
                raise ValueError("items must be (BoundaryAssembler, FunctionSpace)")
        if outputs is None:
            outputs = []
            for a, s in zip(assemblers, spaces):
                n = s.global_size
                outputs.append(torch.empty((n, n), dtype=_torch_dtype(a.dtype), device=device or "cuda").t())
        if len(outputs) != len(assemblers):
            raise ValueError("need one output per item")
        for a, s, o in zip(assemblers, spaces, outputs):
            n = s.global_size
            if not o.is_cuda or o.dtype != _torch_dtype(a.dtype):
                raise ValueError("outputs must be GPU tensors of the assemblers' dtype")
            if tuple(o.shape) != (n, n) or (n > 1 and (o.stride(0) != 1 or o.stride(1) != n)):
                raise ValueError("outputs must be column-major (n, n) tensors")
        self.assemblers, self.spaces, self.outputs = list(assemblers), list(spaces), list(outputs)
        n = len(assemblers)
        ah = (ctypes.c_void_p * n)(*[a._rs_assembler for a in assemblers])
        sh = (ctypes.c_void_p * n)(*[s._rs_space for s in spaces])
        oh = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outputs])
        self._h = L.check_ptr(L.lib().boundary_assembler_batch_new(ah, sh, oh, n))

    def __del__(self, _release=L.release):
        _release("boundary_assembler_batch_t_free", getattr(self, "_h", None))
        self._h = None

    def assemble(self, stream=None):
        """Enqueue the batch on `stream` (default: torch's current stream); returns the outputs."""
        L.check(L.lib().boundary_assembler_batch_assemble(self._h, _torch_stream_handle(stream)))
        return self.outputs

    def set_graph(self, enable=True):
        """Capture the launch sequence into a hipGraph on the next assemble and replay it."""
        L.check(L.lib().boundary_assembler_batch_set_graph(self._h, 1 if enable else 0))

    def set_launch(self, blocks_per_cu=2, chunk=8, order=0, waves=4):
        """Launch shape: workgroups per CU (0: one per tile), trial chunk (4 or 8), tile order
        (0 item-major, 1 largest items first), wavefronts per 64 x 64 tile (2, 4 or 8; 2 and
        8 with chunk 8)."""
        L.check(L.lib().boundary_assembler_batch_set_tile_waves(self._h, 4))
        L.check(L.lib().boundary_assembler_batch_set_launch(self._h, blocks_per_cu, chunk, order))
        L.check(L.lib().boundary_assembler_batch_set_tile_waves(self._h, waves))

    def launch(self):
        """(blocks_per_cu, chunk, order, waves) of the current launch shape."""
        v = [ctypes.c_int(0) for _ in range(3)]
        L.check(L.lib().boundary_assembler_batch_get_launch(self._h, *[ctypes.byref(x) for x in v]))
        w = ctypes.c_int(0)
        L.check(L.lib().boundary_assembler_batch_get_tile_waves(self._h, ctypes.byref(w)))
        return tuple(x.value for x in v) + (w.value,)

    def counts(self):
        """dict: items, tiles, adjacent pairs, near pairs."""
        c = np.zeros(4, dtype=np.int64)
        L.check(L.lib().boundary_assembler_batch_counts(self._h, L.ptr(c)))
        return dict(items=int(c[0]), tiles=int(c[1]), adjacent_pairs=int(c[2]), near_pairs=int(c[3]))

    def autotune(self, cache_path=None, reps=3):
        """Pick the fastest launch shape (outputs are overwritten); returns (ms, from_cache)."""
        fc = ctypes.c_int(0)
        path = None if cache_path is None else str(cache_path).encode()
        ms = L.lib().boundary_assembler_batch_autotune(self._h, path, int(reps), ctypes.byref(fc))
        if ms < 0:
            raise ValueError(L.last_error())
        return float(ms), bool(fc.value)

    def set_profiling(self, enable=True):
        L.check(L.lib().boundary_assembler_batch_set_profiling(self._h, 1 if enable else 0))

    def kernel_times(self):
        """(tiles_ms, adjacent_ms) of the last profiled launch (graph replay: (total, -1))."""
        ms = np.zeros(2, dtype=np.float32)
        L.check(L.lib().boundary_assembler_batch_kernel_times(self._h, L.ptr(ms)))
        return float(ms[0]), float(ms[1])
