"""sfhip -- MI355X (gfx950) implementation of the StokesFlow.jl per-timestep hot path.

Host-side mirror of the reference API on top of the C ABI in include/sf_hip.h
(libsfhip.so, hand-written HIP kernels).  See :mod:`sfhip.api`.
"""
from ._lib import EXPORTS, LIB_PATH, SFError, load  # noqa: F401
from .api import (  # noqa: F401
    Context, Ensemble, FlowState, StokesFlow2D, autotune_pt, bilinearMarkerToGrid, context, fzeros, from_numpy, grid_origins,
    init_markers, marker_cell_keys, markers_to_grids3, moveMarkersRK4, probe_dram, plume_properties, pt_iterations, rt_properties, solveStokes, solveStokes_batched,
    saveGrid, saveMarkers, saveStats, set_pt_fusion, set_pt_tuning, step, to_numpy, time_passes, pass_times,
    Materials, autotune_batched, clear_batched_tuning, initializeMarkers, lattice_markers, marker_properties, setInitialMarkerProperties,
    StepGraph, set_grid_wait_limit,
)

__version__ = "0.1.0"
