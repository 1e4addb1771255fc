"""MI355X-native per-face finite-volume update for moving-mesh Voronoi hydrodynamics.

A drop-in for the hot path of the reference crate mvmm_hydro (yuyttenhove/hydro_rs): the
FiniteVolumeSolver / RiemannFluxSolver surface and the runner step, executed by hand-written gfx950 HIP
kernels behind the C-ABI of include/hfv.h.
"""
from . import _lib  # noqa: F401
