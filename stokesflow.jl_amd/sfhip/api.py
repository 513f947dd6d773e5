"""Host-side mirror of the reference's hot-path API, running on libsfhip (gfx950).

Functions keep the reference's names, argument order and meaning (Julia's trailing ``!``
is dropped; arrays are still modified in place):

* :func:`solveStokes`           -- scripts/StokesSolver_multixpu.jl:14-96
* :func:`bilinearMarkerToGrid`  -- scripts/MarkerToGrid.jl:18-38
* :func:`moveMarkersRK4`        -- scripts/MoveMarkers.jl:15-23
* :func:`StokesFlow2D`          -- the timestep loop of scripts/StokesFlow2D_multixpu.jl:55-235
                                   at one rank (no plotting / file output)

Grid arrays are float64 CUDA (HIP) tensors in Julia's layout: shape ``(n1, n2)`` with
strides ``(1, n1)`` (column-major, x fastest), so ``A[i, j]`` is the reference's
``A[i+1, j+1]``.  :func:`fzeros` makes such arrays.  Marker arrays are 1-D float64 CUDA
tensors.  Errors follow the reference: mismatched marker array lengths raise
``AssertionError`` (its ``@assert``); data problems it only prints about are printed once
with a count and returned in ``stats``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

from . import _lib
from ._lib import BatchInstance, BatchParams, M2GOut, RefineOut, RefineParams, RK4Out, SolveOut, StepArgs, check


# --------------------------------------------------------------------------------------
# arrays

def fzeros(n1, n2, device="cuda"):
    """Zeroed float64 (n1, n2) array in column-major (Julia) layout on ``device``."""
    return torch.zeros((n2, n1), dtype=torch.float64, device=device).t()


def from_numpy(a, device="cuda"):
    """Copy a numpy array to the device, keeping Julia layout for 2-D arrays."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        return torch.from_numpy(a.copy()).to(device)
    f = np.asfortranarray(a)
    return torch.from_numpy(np.ascontiguousarray(f.T)).to(device).t()


def to_numpy(t):
    """Device array -> numpy (Fortran order for 2-D arrays)."""
    a = t.detach().cpu()
    if a.dim() == 1:
        return a.numpy().copy()
    return np.asfortranarray(a.t().contiguous().numpy().T)


def _grid_ptr(t, shape, name):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.numel() > 1 and t.stride() != (1, shape[0]):
        raise ValueError(f"{name} must be column-major (strides (1, {shape[0]})), got {t.stride()}")
    return t.data_ptr()


def _grid_shape(t, name):
    if not isinstance(t, torch.Tensor) or t.dim() != 2:
        raise TypeError(f"{name} must be a 2-D tensor")
    return tuple(t.shape)


def _marker_ptr(t, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
        raise TypeError(f"{name} must be a float64 CUDA tensor")
    if t.dim() != 1 or (t.numel() > 1 and t.stride(0) != 1):
        raise ValueError(f"{name} must be a contiguous 1-D tensor")
    return t.data_ptr()


# --------------------------------------------------------------------------------------
# contexts

class Context:
    """Owns an ``sf_ctx``: device scratch for an Nx x Ny grid (0 x 0: markers only)."""

    def __init__(self, nx=0, ny=0, stream=None):
        self.lib = _lib.load()
        self.nx, self.ny = nx, ny
        h = C.c_void_p()
        check(self.lib.sf_ctx_create(C.byref(h), nx, ny, stream))
        self.handle = h

    def use_current_stream(self):
        s = torch.cuda.current_stream().cuda_stream
        check(self.lib.sf_ctx_set_stream(self.handle, C.c_void_p(s)))
        return self.handle

    def close(self):
        if self.handle:
            self.lib.sf_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = {}


def autotune_pt(nx, ny, iters=20, state=None, g_y=1.0, dx=None, dy=None):
    """Time every PT launch shape on an nx x ny grid (sf_ctx_autotune_pt) and keep the fastest
    for later solves of that size.  ``state`` (a :class:`FlowState`) gives the problem to time
    on (copied, not modified); without it a synthetic problem is used.  Returns (best, table)."""
    nx, ny = int(nx), int(ny)
    ctx = context(nx, ny)
    best = _lib.PTTuning()
    table = (_lib.PTTuning * 64)()
    n = C.c_int(0)
    if state is not None:
        s = state
        ptrs = [_grid_ptr(s.P, (nx - 1, ny - 1), "P"), _grid_ptr(s.Vx_small, (nx, ny + 1), "Vx_small"),
                _grid_ptr(s.Vy_small, (nx + 1, ny), "Vy_small"), _grid_ptr(s.dVxdtau, (nx - 2, ny - 1), "dVxdtau"),
                _grid_ptr(s.dVydtau, (nx - 1, ny - 2), "dVydtau"), _grid_ptr(s.rho_vy, (nx + 1, ny), "rho_vy"),
                _grid_ptr(s.mu_b, (nx, ny), "mu_b"), _grid_ptr(s.mu_p, (nx - 1, ny - 1), "mu_p")]
    else:
        ptrs = [None] * 8
    ddx = float(dx) if dx is not None else 1.0 / (nx - 1)
    ddy = float(dy) if dy is not None else 1.0 / (ny - 1)
    check(ctx.lib.sf_ctx_autotune_pt(ctx.handle, int(iters), *ptrs, float(g_y), ddx, ddy, C.byref(best), table, 64,
                                     C.byref(n)))
    return best.as_dict(), [table[i].as_dict() for i in range(n.value)]


def probe_dram(nbytes=1 << 30):
    """DRAM ceiling (sf_probe_dram): best copy and 3-read + 1-write bandwidth in GB/s."""
    ctx = context(0, 0)
    cp, tr = C.c_double(), C.c_double()
    check(ctx.lib.sf_probe_dram(ctx.handle, int(nbytes), C.byref(cp), C.byref(tr)))
    return {"copy_gbs": cp.value, "r3w1_gbs": tr.value}


def set_pt_fusion(nx, ny, iters_per_launch=0, rows=0):
    """PT iterations per launch for the nx x ny context (sf_ctx_set_pt_fusion): 2, 3 or 4 runs
    that many plain iterations as one pass, 1 one per launch, 0 the size default; the FP32 sweeps
    of mixed_precision solves also take 5.  Results are identical."""
    ctx = context(nx, ny)
    check(ctx.lib.sf_ctx_set_pt_fusion(ctx.handle, int(iters_per_launch), int(rows)))


def time_passes(nx, ny, enable=True):
    """Per-launch timing of the pass kernels of the nx x ny context (sf_ctx_time_passes): from now
    on every FP64 multi-iteration pass and every FP32 correction pass is bracketed by two HIP
    events on the context's stream; the totals restart at zero."""
    ctx = context(nx, ny)
    check(ctx.lib.sf_ctx_time_passes(ctx.handle, int(bool(enable))))


def pass_times(nx, ny):
    """Totals of the timed pass launches of the nx x ny context (sf_ctx_pass_times): a list of
    dicts with precision (0 FP64, 1 FP32), iters_per_launch, launches and seconds."""
    ctx = context(nx, ny)
    buf = (_lib.PassTime * 16)()
    n = C.c_int(0)
    check(ctx.lib.sf_ctx_pass_times(ctx.handle, buf, 16, C.byref(n)))
    return [buf[i].as_dict() for i in range(min(n.value, 16))]


def set_pt_tuning(nx, ny, rows=0, depth=0, max_waves=-1):
    """Launch shape of the one-iteration PT kernel for the nx x ny context (sf_ctx_set_pt_tuning)."""
    ctx = context(nx, ny)
    check(ctx.lib.sf_ctx_set_pt_tuning(ctx.handle, int(rows), int(depth), int(max_waves)))


def context(nx=0, ny=0):
    key = (nx, ny, torch.cuda.current_device())
    ctx = _contexts.get(key)
    if ctx is None:
        ctx = Context(nx, ny)
        _contexts[key] = ctx
    ctx.use_current_stream()
    return ctx


# --------------------------------------------------------------------------------------
# solveStokes!

def solveStokes(P, Vx_pad, Vy_pad, rho_vy, mu_b, mu_p, txx, tyy, txy, divV, dtauPt, Rx, Ry, dVxdtau, dVydtau,
                dtauVx, dtauVy, Vx_s, Vy_s, g_y, dx, dy, Nx, Ny, dt, maxdisp, comm=None, *,
                use_free_surface_stabilization=True, eps=1e-8, print_info=True, iterMax=100000, stats=None,
                mixed_precision=False, refine=None, no_alloc=False):
    """Pseudo-transient Stokes solve (StokesSolver_multixpu.jl:14-96).  Returns (dt, T_eff, runtime).

    ``comm`` is accepted for signature compatibility and must be None (one rank).  The
    scratch arrays txx, tyy, txy, divV, dtauPt, Rx, Ry, dtauVx, dtauVy may be None.

    ``mixed_precision=True`` solves with FP32 PT sweeps inside FP64 defect correction
    (``sf_solve_stokes_refined``); ``refine`` may hold ``rebase_interval`` and ``verbose``;
    ``iterMax`` caps the FP32 iterations.  Then T_eff is 0, runtime is the wall time and
    ``stats`` also receives the refinement counters.

    ``no_alloc=True`` runs the FP64 solve through ``sf_solve_stokes_noalloc``: no context and no
    device allocation (its scratch is a static device array of the library), on torch's current
    stream; grids whose scratch does not fit raise ``SFError`` (SF_ERR_SIZE).
    """
    if comm is not None:
        raise ValueError("one-rank build: comm must be None")
    nx, ny = int(Nx), int(Ny)
    sh = {
        "P": (nx - 1, ny - 1), "Vx_pad": (nx + 2, ny + 1), "Vy_pad": (nx + 1, ny + 2), "rho_vy": (nx + 1, ny),
        "mu_b": (nx, ny), "mu_p": (nx - 1, ny - 1), "txx": (nx - 1, ny - 1), "tyy": (nx - 1, ny - 1),
        "txy": (nx, ny), "divV": (nx - 1, ny - 1), "dtauPt": (nx - 1, ny - 1), "Rx": (nx - 2, ny - 1),
        "Ry": (nx - 1, ny - 2), "dVxdtau": (nx - 2, ny - 1), "dVydtau": (nx - 1, ny - 2),
        "dtauVx": (nx - 2, ny - 1), "dtauVy": (nx - 1, ny - 2), "Vx_s": (nx, ny + 1), "Vy_s": (nx + 1, ny),
    }
    args = dict(P=P, Vx_pad=Vx_pad, Vy_pad=Vy_pad, rho_vy=rho_vy, mu_b=mu_b, mu_p=mu_p, txx=txx, tyy=tyy, txy=txy,
                divV=divV, dtauPt=dtauPt, Rx=Rx, Ry=Ry, dVxdtau=dVxdtau, dVydtau=dVydtau, dtauVx=dtauVx,
                dtauVy=dtauVy, Vx_s=Vx_s, Vy_s=Vy_s)
    ptrs = [_grid_ptr(args[k], sh[k], k) for k in sh]
    out = SolveOut()
    if no_alloc:
        if mixed_precision:
            raise ValueError("no_alloc applies to the FP64 solve only")
        lib = _lib.load()
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(lib.sf_solve_stokes_noalloc(stream, *ptrs, float(g_y), float(dx), float(dy), nx, ny, float(dt),
                                          float(maxdisp), int(bool(use_free_surface_stabilization)), float(eps),
                                          int(iterMax), int(print_info), C.byref(out)))
        if stats is not None:
            stats.update(out.as_dict())
        return out.dt_new, out.t_eff, out.runtime
    ctx = context(nx, ny)
    if mixed_precision:
        prm = RefineParams()
        prm.iter_max = int(iterMax)
        for key, val in (refine or {}).items():
            setattr(prm, key, val)
        rout = RefineOut()
        check(ctx.lib.sf_solve_stokes_refined(ctx.handle, *ptrs, float(g_y), float(dx), float(dy), nx, ny, float(dt),
                                              float(maxdisp), int(bool(use_free_surface_stabilization)), float(eps),
                                              C.byref(prm), C.byref(out), C.byref(rout)))
        if print_info:
            print(f"Total steps = {out.niter} ({rout.n_fp32_iter} FP32 + {rout.n_fp64_iter} FP64, "
                  f"{rout.n_outer} corrections), err = {out.err:1.3e}, time = {out.runtime:1.3e} sec")
        if stats is not None:
            stats.update(out.as_dict())
            stats.update(rout.as_dict())
        return out.dt_new, out.t_eff, out.runtime
    check(ctx.lib.sf_solve_stokes(ctx.handle, *ptrs, float(g_y), float(dx), float(dy), nx, ny, float(dt),
                                  float(maxdisp), int(bool(use_free_surface_stabilization)), float(eps),
                                  int(iterMax), int(print_info), C.byref(out)))
    if stats is not None:
        stats.update(out.as_dict())
    return out.dt_new, out.t_eff, out.runtime


_BATCH_ARRAYS = ("P", "Vx_pad", "Vy_pad", "dVxdtau", "dVydtau", "Vx_s", "Vy_s", "rho_vy", "mu_b", "mu_p")


_BATCH_MODES = {"auto": _lib.SF_BATCH_AUTO, "persistent": _lib.SF_BATCH_PERSISTENT, "wide": _lib.SF_BATCH_WIDE,
                "grid": _lib.SF_BATCH_GRID}


def solveStokes_batched(problems, *, mode="auto", rows=0, depth=0, threads=0, max_blocks=0, budget_mb=0.0,
                        mixed_precision=False):
    """Many independent solveStokes! problems in one persistent launch (sf_solve_stokes_batched).

    ``problems`` is a sequence of dicts with the arrays P, Vx_pad, Vy_pad, dVxdtau, dVydtau,
    Vx_s, Vy_s, rho_vy, mu_b, mu_p (shapes as for :func:`solveStokes`) and the scalars Nx, Ny,
    dx, dy, g_y, dt, maxdisp, and optionally eps (1e-8), iterMax (100000) and
    use_free_surface_stabilization (True).  Grids may differ in size.  Arrays are updated in
    place; returns one stats dict per problem (dt_new, err, niter, converged, dt_last, and
    runtime = wall time of the whole batch).

    ``mode`` "persistent": one launch, one workgroup per instance from a device work queue;
    "wide" (= "auto"): one to four PT iterations per launch over an active set of instances
    whose working sets fit ``budget_mb`` (default 1536 MiB; 1024 MiB with mixed precision),
    refilled as instances converge; "grid": one instance after another, each solved by one
    launch over the whole chip with the PT loop on the device (``depth`` = PT iterations per
    pass, 1 to 3, default 3).
    ``mixed_precision=True`` (wide mode): every instance is solved as
    ``solveStokes(..., mixed_precision=True)`` solves it (FP32 sweeps inside FP64 defect
    correction), one to three FP32 sweeps per launch of all active instances.
    """
    n = len(problems)
    arr = _batch_instances(problems)
    modes = _BATCH_MODES
    prm = BatchParams(int(rows), int(depth), int(threads), int(max_blocks), modes[mode], int(bool(mixed_precision)),
                      float(budget_mb))
    outs = (SolveOut * max(n, 1))()
    ctx = context(0, 0)
    check(ctx.lib.sf_solve_stokes_batched(ctx.handle, arr, n, C.byref(prm), outs))
    keys = ("dt_new", "err", "niter", "converged", "dt_last", "runtime")
    return [{k: getattr(outs[i], k) for k in keys} for i in range(n)]


def autotune_batched(problems, *, budget_mb=0.0, iters=30):
    """Autotuner of the batched solver's wide passes (sf_batch_autotune): per grid size among
    ``problems`` (as for :func:`solveStokes_batched`; their arrays are copied, not modified), the
    fastest iterations per pass and tile rows, kept for later batched solves and ensembles.
    Returns the table of every timed candidate."""
    n = len(problems)
    arr = _batch_instances(problems)
    prm = BatchParams(0, 0, 0, 0, _lib.SF_BATCH_WIDE, 0, float(budget_mb))
    table = (_lib.BatchTuning * 512)()
    nt = C.c_int(0)
    ctx = context(0, 0)
    check(ctx.lib.sf_batch_autotune(ctx.handle, arr, n, C.byref(prm), int(iters), table, 512, C.byref(nt)))
    return [table[i].as_dict() for i in range(min(nt.value, 512))]


def clear_batched_tuning():
    """Back to the size-default shapes of the wide passes (sf_batch_tuning_clear)."""
    ctx = context(0, 0)
    check(ctx.lib.sf_batch_tuning_clear(ctx.handle))


def _batch_instances(problems):
    n = len(problems)
    arr = (BatchInstance * max(n, 1))()
    for i, pb in enumerate(problems):
        nx, ny = int(pb["Nx"]), int(pb["Ny"])
        sh = {"P": (nx - 1, ny - 1), "Vx_pad": (nx + 2, ny + 1), "Vy_pad": (nx + 1, ny + 2),
              "dVxdtau": (nx - 2, ny - 1), "dVydtau": (nx - 1, ny - 2), "Vx_s": (nx, ny + 1), "Vy_s": (nx + 1, ny),
              "rho_vy": (nx + 1, ny), "mu_b": (nx, ny), "mu_p": (nx - 1, ny - 1)}
        b = arr[i]
        for k in _BATCH_ARRAYS:
            setattr(b, k, _grid_ptr(pb[k], sh[k], f"problem {i} {k}"))
        b.nx, b.ny = nx, ny
        b.dx, b.dy, b.g_y = float(pb["dx"]), float(pb["dy"]), float(pb["g_y"])
        b.dt, b.maxdisp = float(pb["dt"]), float(pb["maxdisp"])
        b.eps = float(pb.get("eps", 1e-8))
        b.iter_max = int(pb.get("iterMax", 100000))
        b.use_fss = int(bool(pb.get("use_free_surface_stabilization", True)))
    return arr


def pt_iterations(n_iter, P, Vx_s, Vy_s, dVxdtau, dVydtau, rho_vy, mu_b, mu_p, g_y, dx, dy, Nx, Ny, dt):
    """``n_iter`` PT loop bodies at a fixed stabilisation dt (no dt refresh, no check)."""
    nx, ny = int(Nx), int(Ny)
    ptrs = [
        _grid_ptr(P, (nx - 1, ny - 1), "P"), _grid_ptr(Vx_s, (nx, ny + 1), "Vx_s"),
        _grid_ptr(Vy_s, (nx + 1, ny), "Vy_s"), _grid_ptr(dVxdtau, (nx - 2, ny - 1), "dVxdtau"),
        _grid_ptr(dVydtau, (nx - 1, ny - 2), "dVydtau"), _grid_ptr(rho_vy, (nx + 1, ny), "rho_vy"),
        _grid_ptr(mu_b, (nx, ny), "mu_b"), _grid_ptr(mu_p, (nx - 1, ny - 1), "mu_p"),
    ]
    ctx = context(nx, ny)
    check(ctx.lib.sf_pt_iterations(ctx.handle, int(n_iter), *ptrs, float(g_y), float(dx), float(dy), nx, ny,
                                   float(dt)))


# --------------------------------------------------------------------------------------
# bilinearMarkerToGrid!

def bilinearMarkerToGrid(x_grid_min, y_grid_min, val_grid, x_m, y_m, val_m, dx, dy, val_wt_sum=None,
                         wt_sum=None, grid=None, *, stats=None, print_info=True):
    """Marker -> grid interpolation (MarkerToGrid.jl:18-38), deterministic.

    ``val_wt_sum`` / ``wt_sum`` may be None; ``grid`` (ImplicitGlobalGrid) must be None.
    """
    if grid is not None:
        raise ValueError("one-rank build: grid must be None")
    gnx, gny = _grid_shape(val_grid, "val_grid")
    nm = x_m.numel()
    assert y_m.numel() == nm and val_m.numel() == nm
    pv = _grid_ptr(val_grid, (gnx, gny), "val_grid")
    if (val_wt_sum is None) != (wt_sum is None):
        raise ValueError("pass both helper arrays or neither")
    ws_nx = ws_ny = 0
    pvw = pw = None
    if wt_sum is not None:
        ws_nx, ws_ny = _grid_shape(wt_sum, "wt_sum")
        pw = _grid_ptr(wt_sum, (ws_nx, ws_ny), "wt_sum")
        pvw = _grid_ptr(val_wt_sum, (ws_nx, ws_ny), "val_wt_sum")
    ctx = context()
    out = M2GOut()
    check(ctx.lib.sf_marker_to_grid(ctx.handle, float(x_grid_min), float(y_grid_min), pv, gnx, gny,
                                    _marker_ptr(x_m, "x_m"), _marker_ptr(y_m, "y_m"), _marker_ptr(val_m, "val_m"),
                                    nm, float(dx), float(dy), pvw, pw, ws_nx, ws_ny, C.byref(out)))
    if print_info and out.n_bad_markers:
        print(f"ATTENTION: erroneous marker position ({out.n_bad_markers} markers)")
    if print_info and out.n_empty_nodes:
        print(f"PROBLEM: no markers close enough for value interpolation at {out.n_empty_nodes} nodes")
    if stats is not None:
        stats.update(n_bad_markers=out.n_bad_markers, n_empty_nodes=out.n_empty_nodes)


# --------------------------------------------------------------------------------------
# moveMarkersRK4!

def markers_to_grids3(state, x_m, y_m, rho_m, mu_m, dx, dy):
    """The three marker -> grid passes of a timestep into ``state`` (a :class:`FlowState`):
    rho_vy from rho_m, mu_b and mu_p from mu_m (sf_markers_to_grids3).  Returns the three
    passes' (n_bad_markers, n_empty_nodes)."""
    s = state
    nx, ny = s.nx, s.ny
    nm = x_m.numel()
    assert y_m.numel() == nm and rho_m.numel() == nm and mu_m.numel() == nm
    out = (M2GOut * 3)()
    ctx = context(0, 0)
    check(ctx.lib.sf_markers_to_grids3(ctx.handle, _grid_ptr(s.rho_vy, (nx + 1, ny), "rho_vy"),
                                       _grid_ptr(s.mu_b, (nx, ny), "mu_b"), _grid_ptr(s.mu_p, (nx - 1, ny - 1), "mu_p"),
                                       nx, ny, _marker_ptr(x_m, "x_m"), _marker_ptr(y_m, "y_m"),
                                       _marker_ptr(rho_m, "rho_m"), _marker_ptr(mu_m, "mu_m"), nm, float(dx),
                                       float(dy), out))
    return [(o.n_bad_markers, o.n_empty_nodes) for o in out]


def marker_cell_keys(x_grid_min, y_grid_min, gnx, gny, x_m, y_m, dx, dy):
    """Integer cell keys of the marker -> grid pass (sf_marker_cell_keys): returns (keys, counts,
    n_bad) as int32 CUDA tensors (keys[m] = iy*(gnx+1)+ix of the 1-based top-left node, -1 if
    outside) and the number of markers outside the grid range."""
    px, py = _marker_ptr(x_m, "x_m"), _marker_ptr(y_m, "y_m")
    assert x_m.numel() == y_m.numel()
    keys = torch.empty(x_m.numel(), dtype=torch.int32, device=x_m.device)
    counts = torch.empty((int(gnx) + 1) * (int(gny) + 1), dtype=torch.int32, device=x_m.device)
    out = M2GOut()
    ctx = context(0, 0)
    check(ctx.lib.sf_marker_cell_keys(ctx.handle, float(x_grid_min), float(y_grid_min), int(gnx), int(gny), px, py,
                                      x_m.numel(), float(dx), float(dy), keys.data_ptr(), counts.data_ptr(),
                                      C.byref(out)))
    return keys, counts, out.n_bad_markers


def moveMarkersRK4(x_m, y_m, Vx, Vy, x_vx_min, y_vx_min, x_vy_min, y_vy_min, dt, lx, ly, dx, dy, *, stats=None,
                   print_info=True):
    """RK4 marker advection (MoveMarkers.jl:15-69), positions updated in place."""
    nm = x_m.numel()
    assert nm == y_m.numel(), "x and y Marker coords have not same length"
    n1x, n2x = _grid_shape(Vx, "Vx")
    n1y, n2y = _grid_shape(Vy, "Vy")
    ctx = context()
    out = RK4Out()
    check(ctx.lib.sf_move_markers_rk4(ctx.handle, _marker_ptr(x_m, "x_m"), _marker_ptr(y_m, "y_m"), nm,
                                      _grid_ptr(Vx, (n1x, n2x), "Vx"), n1x, n2x, _grid_ptr(Vy, (n1y, n2y), "Vy"),
                                      n1y, n2y, float(x_vx_min), float(y_vx_min), float(x_vy_min), float(y_vy_min),
                                      float(dt), float(lx), float(ly), float(dx), float(dy), C.byref(out)))
    if print_info and out.n_clamped:
        print(f"WARNING: V-interpolation index failsafe used {out.n_clamped} times")
    if stats is not None:
        stats.update(n_clamped=out.n_clamped)


# --------------------------------------------------------------------------------------
# markers

def init_markers(Nx, Ny, md, dx, dy, jitter_frac=0.25, seed=0, device="cuda"):
    """Seeded md x md markers per cell (x outer / y inner) with +-jitter_frac cell jitter."""
    n = md * (Nx - 1) * md * (Ny - 1)
    x = torch.empty(n, dtype=torch.float64, device=device)
    y = torch.empty(n, dtype=torch.float64, device=device)
    ctx = context()
    check(ctx.lib.sf_init_markers(ctx.handle, x.data_ptr(), y.data_ptr(), int(Nx), int(Ny), int(md), float(dx),
                                  float(dy), float(jitter_frac), int(seed) & 0xFFFFFFFFFFFFFFFF))
    return x, y


def plume_properties(x_m, y_m, air_h, cx, cy, r, rho_air, rho_plume, rho_bg, mu_air, mu_plume, mu_bg):
    """Density/viscosity of the plume setups (air layer, disc, background) per marker."""
    mat = (C.c_double * 10)(air_h, cx, cy, r, rho_air, rho_plume, rho_bg, mu_air, mu_plume, mu_bg)
    rho = torch.empty_like(x_m)
    mu = torch.empty_like(x_m)
    ctx = context()
    check(ctx.lib.sf_plume_properties(ctx.handle, _marker_ptr(x_m, "x_m"), _marker_ptr(y_m, "y_m"), x_m.numel(),
                                      mat, rho.data_ptr(), mu.data_ptr()))
    return rho, mu


def lattice_markers(Nx, Ny, md, lx, ly, dx, dy, rand_pos=False, seed=0, device="cuda"):
    """The reference's marker lattice on the device (sf_lattice_markers, common.jl:40-87 at one
    rank): LinRange(dxm/2, lx - dxm/2, md (Nx-1)) x LinRange(dym/2, ly - dym/2, md (Ny-1)), x outer,
    bit for bit as Julia 1.8 evaluates it; rand_pos adds (u - 0.5) dxm / dym from the build's
    seeded generator."""
    n = md * (Nx - 1) * md * (Ny - 1)
    x = torch.empty(n, dtype=torch.float64, device=device)
    y = torch.empty(n, dtype=torch.float64, device=device)
    ctx = context()
    check(ctx.lib.sf_lattice_markers(ctx.handle, x.data_ptr(), y.data_ptr(), int(Nx), int(Ny), int(md), float(lx),
                                     float(ly), float(dx), float(dy), int(bool(rand_pos)),
                                     int(seed) & 0xFFFFFFFFFFFFFFFF))
    return x, y


_SHAPE_KINDS = {"y_below": _lib.SF_SHAPE_Y_BELOW, "y_above": _lib.SF_SHAPE_Y_ABOVE, "circle": _lib.SF_SHAPE_CIRCLE,
                "rect": _lib.SF_SHAPE_RECT, "sine_below": _lib.SF_SHAPE_SINE_BELOW}


class Materials:
    """Material descriptor in place of the reference's density(x, y) / viscosity(x, y) callbacks:
    an ordered list of shapes, each with its (rho, mu); the first shape that contains a point
    decides, else the background.  Shapes: ("y_below", h), ("y_above", h), ("circle", cx, cy, r),
    ("rect", x0, x1, y0, y1), ("sine_below", y0, amp, k, lx), each followed by rho, mu:
    ``Materials([("y_below", 2.0, 1e-3, 1e-2), ("circle", 5, 5, 2, 3.2, 0.1)], rho_bg=3.3, mu_bg=1)``
    is the plume of StokesFlow2D_multixpu.jl:387-404."""

    def __init__(self, shapes, rho_bg, mu_bg):
        if len(shapes) > _lib.SF_MAX_SHAPES:
            raise ValueError(f"at most {_lib.SF_MAX_SHAPES} shapes")
        self.shapes = [tuple(sh) for sh in shapes]
        self.rho_bg, self.mu_bg = float(rho_bg), float(mu_bg)

    def as_array(self):
        arr = (_lib.Shape * max(len(self.shapes), 1))()
        for i, sh in enumerate(self.shapes):
            kind, *vals = sh
            params, (rho, mu) = vals[:-2], vals[-2:]
            if kind not in _SHAPE_KINDS:
                raise ValueError(f"unknown shape {kind!r}")
            arr[i].kind = _SHAPE_KINDS[kind]
            for name, v in zip("abcde", params):
                setattr(arr[i], name, float(v))
            arr[i].rho, arr[i].mu = float(rho), float(mu)
        return arr


def marker_properties(x_m, y_m, materials, x0=0.0, y0=0.0):
    """rho_m, mu_m of every marker from a :class:`Materials` descriptor at global positions
    (x0 + x_m, y0 + y_m) (sf_marker_properties)."""
    rho = torch.empty_like(x_m)
    mu = torch.empty_like(x_m)
    setInitialMarkerProperties(x_m, y_m, rho, mu, x0, y0, materials)
    return rho, mu


def initializeMarkers(marker_density, lx, ly, dx, dy, Nx, Ny, RAND_MARKER_POS=False, seed=0):
    """initializeMarkersCPU (common.jl:40-100) at one rank, on the device: the marker lattice and
    zeroed rho_m, mu_m.  Returns (x_m, y_m, rho_m, mu_m)."""
    x, y = lattice_markers(Nx, Ny, marker_density, lx, ly, dx, dy, rand_pos=RAND_MARKER_POS, seed=seed)
    return x, y, torch.zeros_like(x), torch.zeros_like(x)


def setInitialMarkerProperties(x_m, y_m, rho_m, mu_m, x0, y0, materials):
    """setInitialMarkerProperties! (common.jl:104-110) with a :class:`Materials` descriptor in
    place of the density / viscosity callbacks; rho_m, mu_m are filled in place."""
    nm = x_m.numel()
    assert y_m.numel() == nm and rho_m.numel() == nm and mu_m.numel() == nm
    arr = materials.as_array()
    ctx = context()
    check(ctx.lib.sf_marker_properties(ctx.handle, _marker_ptr(x_m, "x_m"), _marker_ptr(y_m, "y_m"), nm, float(x0),
                                       float(y0), arr, len(materials.shapes), materials.rho_bg, materials.mu_bg,
                                       _marker_ptr(rho_m, "rho_m"), _marker_ptr(mu_m, "mu_m")))


def rt_properties(x_m, y_m, y0, amp, k, lx, rho_up, mu_up, rho_low, mu_low):
    """Two-layer Rayleigh-Taylor material: upper layer where y < y0 + amp sin(2 pi k x / lx)."""
    mat = (C.c_double * 8)(y0, amp, k, lx, rho_up, mu_up, rho_low, mu_low)
    rho = torch.empty_like(x_m)
    mu = torch.empty_like(x_m)
    ctx = context()
    check(ctx.lib.sf_rt_properties(ctx.handle, _marker_ptr(x_m, "x_m"), _marker_ptr(y_m, "y_m"), x_m.numel(), mat,
                                   rho.data_ptr(), mu.data_ptr()))
    return rho, mu


# --------------------------------------------------------------------------------------
# timestep loop

def grid_origins(dx, dy):
    """Grid origins of StokesFlow2D_multixpu.jl:117-126."""
    return {
        "rho": (0 * dx - 0.5 * dx, 0 * dy), "mu_b": (0.0, 0.0), "mu_p": (0 * dx + 0.5 * dx, 0 * dy + 0.5 * dy),
        "vx": (-1 * dx, 0 * dy - 0.5 * dy), "vy": (0 * dx - 0.5 * dx, -1 * dy),
    }


class FlowState:
    """Device arrays of StokesFlow2D_multixpu.jl:89-109 (the persistent solver state)."""

    def __init__(self, nx, ny, device="cuda"):
        self.nx, self.ny = nx, ny
        z = lambda a, b: fzeros(a, b, device)  # noqa: E731
        self.P = z(nx - 1, ny - 1)
        self.Vx = z(nx + 2, ny + 1)
        self.Vy = z(nx + 1, ny + 2)
        self.rho_vy = z(nx + 1, ny)
        self.mu_b = z(nx, ny)
        self.mu_p = z(nx - 1, ny - 1)
        self.dVxdtau = z(nx - 2, ny - 1)
        self.dVydtau = z(nx - 1, ny - 2)
        self.Vx_small = z(nx, ny + 1)
        self.Vy_small = z(nx + 1, ny)


def _step_args(state, x_m, y_m, rho_m, mu_m, dx, dy, dt, g_y, maxdisp, eps, iterMax, use_fss):
    nm = x_m.numel()
    assert y_m.numel() == nm and rho_m.numel() == nm and mu_m.numel() == nm
    s = state
    a = StepArgs()
    a.nx, a.ny, a.dx, a.dy = s.nx, s.ny, float(dx), float(dy)
    a.g_y, a.maxdisp, a.eps, a.iter_max = float(g_y), float(maxdisp), float(eps), int(iterMax)
    a.use_fss, a.dt = int(bool(use_fss)), float(dt)
    nx, ny = s.nx, s.ny
    a.P = _grid_ptr(s.P, (nx - 1, ny - 1), "P")
    a.Vx_pad = _grid_ptr(s.Vx, (nx + 2, ny + 1), "Vx")
    a.Vy_pad = _grid_ptr(s.Vy, (nx + 1, ny + 2), "Vy")
    a.rho_vy = _grid_ptr(s.rho_vy, (nx + 1, ny), "rho_vy")
    a.mu_b = _grid_ptr(s.mu_b, (nx, ny), "mu_b")
    a.mu_p = _grid_ptr(s.mu_p, (nx - 1, ny - 1), "mu_p")
    a.dVxdtau = _grid_ptr(s.dVxdtau, (nx - 2, ny - 1), "dVxdtau")
    a.dVydtau = _grid_ptr(s.dVydtau, (nx - 1, ny - 2), "dVydtau")
    a.Vx_s = _grid_ptr(s.Vx_small, (nx, ny + 1), "Vx_small")
    a.Vy_s = _grid_ptr(s.Vy_small, (nx + 1, ny), "Vy_small")
    a.x_m, a.y_m = _marker_ptr(x_m, "x_m"), _marker_ptr(y_m, "y_m")
    a.rho_m, a.mu_m = _marker_ptr(rho_m, "rho_m"), _marker_ptr(mu_m, "mu_m")
    a.nm = nm
    return a


def step(state, x_m, y_m, rho_m, mu_m, dx, dy, dt, *, g_y=9.81, maxdisp=0.5, eps=1e-8, iterMax=100000,
         use_free_surface_stabilization=True):
    """One timestep (StokesFlow2D_multixpu.jl:164-181): M2G x3 -> solve -> RK4.  Returns (dt_new, stats)."""
    a = _step_args(state, x_m, y_m, rho_m, mu_m, dx, dy, dt, g_y, maxdisp, eps, iterMax,
                   use_free_surface_stabilization)
    ctx = context(state.nx, state.ny)
    so = SolveOut()
    mo = (M2GOut * 3)()
    ro = RK4Out()
    check(ctx.lib.sf_step(ctx.handle, C.byref(a), C.byref(so), mo, C.byref(ro)))
    stats = so.as_dict()
    stats["m2g"] = [(m.n_bad_markers, m.n_empty_nodes) for m in mo]
    stats["n_clamped"] = ro.n_clamped
    return a.dt, stats


def set_grid_wait_limit(seconds):
    """Time limit of one grid-barrier wait in the grid mode (sf_set_grid_wait_limit; default 5 s)."""
    check(_lib.load().sf_set_grid_wait_limit(float(seconds)))


class StepGraph:
    """The captured timestep of one StokesFlow2D problem (sf_graph_*): the loop body of
    StokesFlow2D_multixpu.jl:164-181 as one hipGraph, the solve with its PT loop on the device
    (one chip-wide launch), RK4 with the new dt read on the device.  Arguments as for
    :func:`step` (``dt``: the first step's input dt); ``iters_per_pass`` 1 to 3 (0: 3).  The
    tensors must stay alive while the graph exists.  Bit-identical to repeated :func:`step` calls
    (err to rounding)."""

    def __init__(self, state, x_m, y_m, rho_m, mu_m, dx, dy, dt=0.0, *, g_y=9.81, maxdisp=0.5, eps=1e-8,
                 iterMax=100000, use_free_surface_stabilization=True, iters_per_pass=0):
        self.lib = _lib.load()
        self._keep = (state, x_m, y_m, rho_m, mu_m)
        a = _step_args(state, x_m, y_m, rho_m, mu_m, dx, dy, dt, g_y, maxdisp, eps, iterMax,
                       use_free_surface_stabilization)
        self._ctx = context(0, 0)
        self._ctx.use_current_stream()
        self.handle = C.c_void_p()
        check(self.lib.sf_graph_create(self._ctx.handle, C.byref(a), int(iters_per_pass), C.byref(self.handle)))

    def step(self, n_steps=1):
        """Replay n_steps timesteps; returns (dt, stats of the last step's solve)."""
        self._ctx.use_current_stream()
        so = SolveOut()
        dt = C.c_double()
        check(self.lib.sf_graph_step(self.handle, int(n_steps), C.byref(so), C.byref(dt)))
        return dt.value, so.as_dict()

    def close(self):
        if self.handle:
            self.lib.sf_graph_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Many independent StokesFlow2D problems stepped together (sf_ensemble_*).

    ``members`` is a sequence of dicts with keys state (a :class:`FlowState`), x_m, y_m, rho_m,
    mu_m, dx, dy and optionally dt (0), g_y (9.81), maxdisp (0.5), eps (1e-8), iterMax (100000),
    use_free_surface_stabilization (True).  One :meth:`step` runs, for every member, the three
    marker -> grid passes, then one batched solve of all members, then RK4 with each member's
    new dt.  ``mode`` "persistent": the whole step is one hipGraph replay (small members);
    "grid": the same with one chip-wide grid-mode solve per member;
    "wide" (= "auto"): marker -> grid and RK4 are hipGraph replays around the wide batched
    solve; ``mixed_precision=True`` (wide form) solves with FP32 sweeps inside FP64 defect
    correction.  The members' tensors must stay alive while the ensemble exists.
    """

    def __init__(self, members, *, use_graph=True, mode="auto", rows=0, threads=0, max_blocks=0, budget_mb=0.0,
                 mixed_precision=False):
        self.lib = _lib.load()
        self.n = len(members)
        self._keep = list(members)
        arr = (StepArgs * max(self.n, 1))()
        for i, m in enumerate(members):
            arr[i] = _step_args(m["state"], m["x_m"], m["y_m"], m["rho_m"], m["mu_m"], m["dx"], m["dy"],
                                m.get("dt", 0.0), m.get("g_y", 9.81), m.get("maxdisp", 0.5), m.get("eps", 1e-8),
                                m.get("iterMax", 100000), m.get("use_free_surface_stabilization", True))
        ctx = context(0, 0)
        self._ctx = ctx
        self.handle = C.c_void_p()
        prm = BatchParams(int(rows), 0, int(threads), int(max_blocks), _BATCH_MODES[mode], int(bool(mixed_precision)),
                          float(budget_mb))
        check(self.lib.sf_ensemble_create(ctx.handle, arr, self.n, C.byref(prm), int(bool(use_graph)),
                                          C.byref(self.handle)))

    def step(self, n_steps=1):
        """Advance every member n_steps timesteps; returns per-member stats of the last step."""
        self._ctx.use_current_stream()
        outs = (SolveOut * max(self.n, 1))()
        dts = (C.c_double * max(self.n, 1))()
        check(self.lib.sf_ensemble_step(self.handle, int(n_steps), outs, dts))
        keys = ("dt_new", "err", "niter", "converged", "dt_last", "runtime")
        return [{k: getattr(outs[i], k) for k in keys} for i in range(self.n)]

    def close(self):
        if self.handle:
            self.lib.sf_ensemble_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _path(path):
    return os.fsencode(os.fspath(path))


def saveStats(path, dims, nt):
    """saveStats (StokesFlow2D_multixpu.jl:212-223) through sf_save_stats: "dims" (3 Int64) and
    "nt" in a MAT-file level 5 file."""
    d = (C.c_int64 * 3)(*[int(v) for v in dims])
    check(_lib.load().sf_save_stats(_path(path), d, int(nt)))


def saveMarkers(path, x_m, y_m, rho_m, x_shift=0.0, y_shift=0.0):
    """saveMarkers (:230-240) through sf_save_markers: Float32 x_m + x_shift, y_m + y_shift, rho_m
    from device arrays (x_shift = (lx - dx) * coords[1] in the reference; 0 at one rank)."""
    nm = x_m.numel()
    for a in (y_m, rho_m):
        if a.numel() != nm:
            raise ValueError("saveMarkers: marker arrays differ in length")
    ptrs = [_marker_ptr(a, n) for a, n in ((x_m, "x_m"), (y_m, "y_m"), (rho_m, "rho_m"))]
    ctx = context(0, 0)
    check(ctx.lib.sf_save_markers(ctx.handle, _path(path), *ptrs, int(nm), float(x_shift), float(y_shift)))


def saveGrid(path, Nx, Ny, dx, dy, mu_b, P, Vx, Vy):
    """saveGrid (:247-299) through sf_save_grid: Float32 node coordinates (:117-124) and mu_b, P
    and the padded Vx, Vy."""
    ptrs = [_grid_ptr(mu_b, (Nx, Ny), "mu_b"), _grid_ptr(P, (Nx - 1, Ny - 1), "P"),
            _grid_ptr(Vx, (Nx + 2, Ny + 1), "Vx"), _grid_ptr(Vy, (Nx + 1, Ny + 2), "Vy")]
    ctx = context(0, 0)
    check(ctx.lib.sf_save_grid(ctx.handle, _path(path), int(Nx), int(Ny), float(dx), float(dy), *ptrs))


def StokesFlow2D(Nt, Nx, Ny, Lx_glob, Ly_glob, x_m, y_m, rho_m, mu_m, *, stokes_eps=1e-5, g_y=9.81, maxdisp=0.5,
                 iterMax=100000, print_info=False, state=None, save_to_file=False, out_dir="viz_out", captured=False):
    """Timestep loop of StokesFlow2D (StokesFlow2D_multixpu.jl:55-235) at one rank.

    The reference evaluates density/viscosity callbacks on a marker lattice it builds
    itself; here the caller passes the marker arrays (see :func:`init_markers` and the
    material helpers).  Returns (Vx, Vy, t_tot) with Vx, Vy the padded velocity arrays.
    save_to_file writes the reference's snapshot files into out_dir (stats.mat, and
    markers_%04d_0000.mat / grid_%04d_0000.mat for step 0 and after every step, :137-141,
    :196-200), read by viz.jl.  captured=True replays each timestep as one hipGraph with the PT loop
    on the device (:class:`StepGraph`); the results are the same bits.
    """
    dx = Lx_glob / (Nx - 1)
    dy = Ly_glob / (Ny - 1)
    st = state if state is not None else FlowState(Nx, Ny, device=x_m.device)

    def snapshot(t):
        saveStats(os.path.join(out_dir, "stats.mat"), (1, 1, 1), t)
        saveMarkers(os.path.join(out_dir, f"markers_{t:04d}_0000.mat"), x_m, y_m, rho_m)
        saveGrid(os.path.join(out_dir, f"grid_{t:04d}_0000.mat"), Nx, Ny, dx, dy, st.mu_b, st.P, st.Vx, st.Vy)

    if save_to_file:
        os.makedirs(out_dir, exist_ok=True)
        snapshot(0)
    dt, t_tot = 0.0, 0.0
    g = (StepGraph(st, x_m, y_m, rho_m, mu_m, dx, dy, 0.0, g_y=g_y, maxdisp=maxdisp, eps=stokes_eps, iterMax=iterMax)
         if captured else None)
    try:
        for t in range(Nt):
            if print_info:
                print(f"(t, t_tot) = ({t + 1}, {t_tot})")
            if g is not None:
                dt, _ = g.step(1)
            else:
                dt, _ = step(st, x_m, y_m, rho_m, mu_m, dx, dy, dt, g_y=g_y, maxdisp=maxdisp, eps=stokes_eps,
                             iterMax=iterMax)
            if save_to_file:
                snapshot(t + 1)
            t_tot += dt
    finally:
        if g is not None:
            g.close()
    return st.Vx, st.Vy, t_tot
