"""Perturbed-sphere meshes and random vectors for the shape-gradient checks (the workload's
definitions live in bempp_amd.workloads, shared with the benchmark)."""
from bempp_amd.workloads import complex_normal, perturb_radially, real_sph_harm  # noqa: F401
