"""bempp_amd: MI355X-native dense boundary-operator assembly with the bempp-rs API.

Mirrors python/bempp of the reference (assembly.boundary, function_space) plus
the grid and element helpers the reference takes from ndgrid/ndelement.
"""
from . import assembly, element, function_space, grid, shapes, workloads  # noqa: F401
from ._lib import device_count, lib_path, version  # noqa: F401

__all__ = ["assembly", "element", "function_space", "grid", "shapes", "workloads", "device_count", "version", "lib_path"]
