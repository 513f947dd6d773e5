"""ctypes bindings of the CPU oracle (``libsf_oracle.so``) -- test infrastructure only.

Arrays are numpy float64 in Fortran order so that ``A[i, j]`` (0-based) is the
reference's ``A[i+1, j+1]`` and the memory layout is Julia's column-major one.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsf_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class SolveOut(C.Structure):
    _fields_ = [
        ("dt_new", C.c_double),
        ("err", C.c_double),
        ("t_eff", C.c_double),
        ("runtime", C.c_double),
        ("niter", C.c_int64),
        ("converged", C.c_int32),
        ("reserved", C.c_int32),
        ("dt_last", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off, OpenMP)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        i, d, i64, u64 = C.c_int, C.c_double, C.c_int64, C.c_uint64
        L.sfo_solve_stokes.argtypes = [_dp] * 19 + [d, d, d, i, i, d, d, i, d, i64, C.POINTER(SolveOut)]
        L.sfo_solve_stokes.restype = i
        L.sfo_compute_timesteps.argtypes = [i, i, _dp, _dp, _dp, _dp, d, d, d, i]
        L.sfo_pt_iterations.argtypes = [i, i, i64] + [_dp] * 15 + [d, d, d, d]
        L.sfo_compute_dt.argtypes = [i, i, _dp, _dp, d, d, d]
        L.sfo_compute_dt.restype = d
        L.sfo_compute_residuals.argtypes = [i, i] + [_dp] * 9 + [d, d, d, d]
        L.sfo_compute_P_tau.argtypes = [i, i] + [_dp] * 10 + [d, d]
        L.sfo_compute_err.argtypes = [i, i, _dp, _dp, _dp]
        L.sfo_compute_err.restype = d
        L.sfo_mean_abs.argtypes = [_dp, i64]
        L.sfo_mean_abs.restype = d
        L.sfo_topleft_index.argtypes = [d, d, d, d, d, d, _i64p, _i64p, _dp, _dp]
        L.sfo_marker_to_grid.argtypes = [d, d, _dp, i, i, _dp, _dp, _dp, i64, d, d, _dp, _dp, i, i, _i64p, _i64p]
        L.sfo_marker_to_grid.restype = i
        L.sfo_cell_keys.argtypes = [d, d, i, i, _dp, _dp, i64, d, d, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    _i64p]
        L.sfo_interpolate_v.argtypes = [d, d, _dp, i, i, _dp, i, i, d, d, d, d, d, d, _dp, _dp, _i64p]
        L.sfo_interpolate_v.restype = i
        L.sfo_move_markers_rk4.argtypes = [_dp, _dp, i64, _dp, i, i, _dp, i, i, d, d, d, d, d, d, d, d, d, _i64p]
        L.sfo_move_markers_rk4.restype = i
        L.sfo_lattice_markers.argtypes = [i, i, i, d, d, d, d, _dp, _dp]
        L.sfo_lattice_markers.restype = i64
        L.sfo_seeded_markers.argtypes = [i, i, i, d, d, d, u64, _dp, _dp]
        L.sfo_uniform.argtypes = [u64, u64]
        L.sfo_uniform.restype = d
        L.sfo_plume_properties.argtypes = [_dp, _dp, i64, d, d, d, d, d, d, d, d, d, d, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), "float64 Fortran arrays"
    return a.ctypes.data_as(_dp)


def zeros(*shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def solver_arrays(nx, ny):
    """Zeroed arrays of every solveStokes! argument (StokesFlow2D_multixpu.jl:89-109)."""
    return {
        "P": zeros(nx - 1, ny - 1), "Vx_pad": zeros(nx + 2, ny + 1), "Vy_pad": zeros(nx + 1, ny + 2),
        "rho_vy": zeros(nx + 1, ny), "mu_b": zeros(nx, ny), "mu_p": zeros(nx - 1, ny - 1),
        "txx": zeros(nx - 1, ny - 1), "tyy": zeros(nx - 1, ny - 1), "txy": zeros(nx, ny),
        "divV": zeros(nx - 1, ny - 1), "dtauP": zeros(nx - 1, ny - 1),
        "Rx": zeros(nx - 2, ny - 1), "Ry": zeros(nx - 1, ny - 2),
        "dVx": zeros(nx - 2, ny - 1), "dVy": zeros(nx - 1, ny - 2),
        "dtauVx": zeros(nx - 2, ny - 1), "dtauVy": zeros(nx - 1, ny - 2),
        "Vx_s": zeros(nx, ny + 1), "Vy_s": zeros(nx + 1, ny),
    }


_SOLVE_ORDER = ["P", "Vx_pad", "Vy_pad", "rho_vy", "mu_b", "mu_p", "txx", "tyy", "txy", "divV", "dtauP",
                "Rx", "Ry", "dVx", "dVy", "dtauVx", "dtauVy", "Vx_s", "Vy_s"]


def solve_stokes(S, g_y, dx, dy, nx, ny, dt, maxdisp, use_fss=True, eps=1e-8, iter_max=100000):
    """solveStokes! (StokesSolver_multixpu.jl:14-96) on the arrays of ``S`` (in place)."""
    out = SolveOut()
    rc = lib().sfo_solve_stokes(*[_p(S[k]) for k in _SOLVE_ORDER], g_y, dx, dy, nx, ny, dt, maxdisp,
                                int(bool(use_fss)), eps, int(iter_max), C.byref(out))
    if rc != 0:
        raise ValueError(f"sfo_solve_stokes failed ({rc})")
    return out.as_dict()


def compute_timesteps(S, nx, ny, dx, dy):
    m = min(dx, dy)
    lib().sfo_compute_timesteps(nx, ny, _p(S["mu_p"]), _p(S["dtauVx"]), _p(S["dtauVy"]), _p(S["dtauP"]),
                                0.45, 0.5, m * m, max(nx, ny))


def pt_iterations(S, nx, ny, n_iter, g_y, dx, dy, dt):
    """n_iter loop bodies at fixed dt (no dt refresh, no check); dtau* must be set."""
    lib().sfo_pt_iterations(nx, ny, int(n_iter), _p(S["P"]), _p(S["Vx_s"]), _p(S["Vy_s"]), _p(S["dVx"]),
                            _p(S["dVy"]), _p(S["rho_vy"]), _p(S["mu_b"]), _p(S["mu_p"]), _p(S["txx"]),
                            _p(S["tyy"]), _p(S["txy"]), _p(S["divV"]), _p(S["dtauP"]), _p(S["dtauVx"]),
                            _p(S["dtauVy"]), g_y, dx, dy, dt)


def compute_dt(Vx_s, Vy_s, maxdisp, dx, dy):
    nx = Vx_s.shape[0]
    ny = Vy_s.shape[1]
    return lib().sfo_compute_dt(nx, ny, _p(Vx_s), _p(Vy_s), maxdisp, dx, dy)


def mean_abs(a):
    a = np.asfortranarray(a, dtype=np.float64)
    return lib().sfo_mean_abs(_p(a), a.size)


def topleft_index(xmin, ymin, x, y, dx, dy):
    ix, iy = C.c_int64(), C.c_int64()
    dxij, dyij = C.c_double(), C.c_double()
    lib().sfo_topleft_index(xmin, ymin, x, y, dx, dy, C.byref(ix), C.byref(iy), C.byref(dxij), C.byref(dyij))
    return ix.value, iy.value, dxij.value, dyij.value


def marker_to_grid(xmin, ymin, val_grid, x_m, y_m, val_m, dx, dy, val_wt_sum, wt_sum):
    """bilinearMarkerToGrid! (MarkerToGrid.jl:18-38, CPU path) in place; returns (n_bad, n_empty)."""
    assert x_m.shape == y_m.shape == val_m.shape
    nb, ne = C.c_int64(), C.c_int64()
    gnx, gny = val_grid.shape
    wnx, wny = wt_sum.shape
    rc = lib().sfo_marker_to_grid(xmin, ymin, _p(val_grid), gnx, gny, _p(x_m), _p(y_m), _p(val_m), x_m.size,
                                  dx, dy, _p(val_wt_sum), _p(wt_sum), wnx, wny, C.byref(nb), C.byref(ne))
    if rc != 0:
        raise ValueError(f"sfo_marker_to_grid failed ({rc})")
    return nb.value, ne.value


def cell_keys(xmin, ymin, gnx, gny, x_m, y_m, dx, dy):
    """Per-marker cell keys and per-cell counts of the M2G CPU path; returns (keys, counts, n_bad)."""
    assert x_m.shape == y_m.shape
    keys = np.empty(x_m.size, dtype=np.int32)
    counts = np.empty((gnx + 1) * (gny + 1), dtype=np.int32)
    nb = C.c_int64()
    rc = lib().sfo_cell_keys(xmin, ymin, gnx, gny, _p(x_m), _p(y_m), x_m.size, dx, dy,
                             keys.ctypes.data_as(C.POINTER(C.c_int32)), counts.ctypes.data_as(C.POINTER(C.c_int32)),
                             C.byref(nb))
    if rc != 0:
        raise ValueError("non-finite marker coordinate")
    return keys, counts, nb.value


def interpolate_v(x, y, Vx, Vy, x_vx_min, y_vx_min, x_vy_min, y_vy_min, dx, dy):
    vx, vy, nc = C.c_double(), C.c_double(), C.c_int64(0)
    rc = lib().sfo_interpolate_v(x, y, _p(Vx), Vx.shape[0], Vx.shape[1], _p(Vy), Vy.shape[0], Vy.shape[1],
                                 x_vx_min, y_vx_min, x_vy_min, y_vy_min, dx, dy, C.byref(vx), C.byref(vy),
                                 C.byref(nc))
    if rc != 0:
        raise ValueError("non-finite position")
    return vx.value, vy.value, nc.value


def move_markers_rk4(x_m, y_m, Vx, Vy, x_vx_min, y_vx_min, x_vy_min, y_vy_min, dt, lx, ly, dx, dy):
    """moveMarkersRK4! (MoveMarkers.jl:15-69) in place; returns the failsafe clamp count."""
    assert x_m.shape == y_m.shape
    nc = C.c_int64(0)
    rc = lib().sfo_move_markers_rk4(_p(x_m), _p(y_m), x_m.size, _p(Vx), Vx.shape[0], Vx.shape[1], _p(Vy),
                                    Vy.shape[0], Vy.shape[1], x_vx_min, y_vx_min, x_vy_min, y_vy_min, dt,
                                    lx, ly, dx, dy, C.byref(nc))
    if rc != 0:
        raise ValueError("non-finite marker position")
    return nc.value


def lattice_markers(nx, ny, md, lx, ly, dx, dy):
    n = lib().sfo_lattice_markers(nx, ny, md, lx, ly, dx, dy, None, None)
    x = np.empty(n)
    y = np.empty(n)
    lib().sfo_lattice_markers(nx, ny, md, lx, ly, dx, dy, _p(x), _p(y))
    return x, y


def seeded_markers(nx, ny, md, dx, dy, jitter_frac, seed):
    n = md * (nx - 1) * md * (ny - 1)
    x = np.empty(n)
    y = np.empty(n)
    lib().sfo_seeded_markers(nx, ny, md, dx, dy, jitter_frac, seed, _p(x), _p(y))
    return x, y


def uniform(seed, counter):
    return lib().sfo_uniform(seed, counter)


def plume_properties(x_m, y_m, air_h, cx, cy, r, rho_air, rho_plume, rho_bg, mu_air, mu_plume, mu_bg):
    rho = np.empty_like(x_m)
    mu = np.empty_like(x_m)
    lib().sfo_plume_properties(_p(x_m), _p(y_m), x_m.size, air_h, cx, cy, r, rho_air, rho_plume, rho_bg,
                               mu_air, mu_plume, mu_bg, _p(rho), _p(mu))
    return rho, mu


def grid_origins(dx, dy):
    """Grid origins of StokesFlow2D_multixpu.jl:117-126 (x[1], x_p[1], x_vx[1], x_vy[1], ...)."""
    return {
        "rho": (0 * dx - 0.5 * dx, 0 * dy),            # x_rho[1] = x_vy[1], y_rho[1] = y_vy[2]
        "mu_b": (0.0, 0.0),                            # x[1], y[1]
        "mu_p": (0 * dx + 0.5 * dx, 0 * dy + 0.5 * dy),  # x_p[1], y_p[1]
        "vx": (-1 * dx, 0 * dy - 0.5 * dy),            # x_vx[1], y_vx[1]
        "vy": (0 * dx - 0.5 * dx, -1 * dy),            # x_vy[1], y_vy[1]
    }


def stokesflow2d(nt, nx, ny, lx_glob, ly_glob, x_m, y_m, rho_m, mu_m, g_y=9.81, maxdisp=0.5, eps=1e-5,
                 iter_max=100000, S=None, callback=None):
    """Timestep loop of StokesFlow2D (StokesFlow2D_multixpu.jl:76-131,155-218) at one rank.

    Markers are modified in place.  Returns (S, t_tot, per-step solve outputs).
    """
    dx = lx_glob / (nx - 1)
    dy = ly_glob / (ny - 1)
    lx = (nx - 1) * dx
    ly = (ny - 1) * dy
    if S is None:
        S = solver_arrays(nx, ny)
    wnx, wny = nx + 1, ny  # max.(size(rho_vy), size(mu_b), size(mu_p))
    vws, ws = zeros(wnx, wny), zeros(wnx, wny)
    o = grid_origins(dx, dy)
    dt, t_tot = 0.0, 0.0
    outs = []
    for t in range(nt):
        marker_to_grid(*o["rho"], S["rho_vy"], x_m, y_m, rho_m, dx, dy, vws, ws)
        marker_to_grid(*o["mu_b"], S["mu_b"], x_m, y_m, mu_m, dx, dy, vws, ws)
        marker_to_grid(*o["mu_p"], S["mu_p"], x_m, y_m, mu_m, dx, dy, vws, ws)
        res = solve_stokes(S, g_y, dx, dy, nx, ny, dt, maxdisp, True, eps, iter_max)
        dt = res["dt_new"]
        move_markers_rk4(x_m, y_m, S["Vx_pad"], S["Vy_pad"], *o["vx"], *o["vy"], dt, lx, ly, dx, dy)
        outs.append(res)
        if callback is not None:
            callback(t, S, x_m, y_m)
        t_tot += dt
    return S, t_tot, outs
