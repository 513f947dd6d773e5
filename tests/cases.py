"""Input builders shared by the tests (closed-form fields of the reference tests, seeded
random fields and marker sets)."""
import numpy as np

from oracle import sf_oracle as O


def stokes_solver_inputs(g):
    """Closed-form rho_vy, mu_b, mu_p of test/test_StokesSolver.jl:62-82 at one rank."""
    nx, ny, dx, dy = g["nx"], g["ny"], g["dx"], g["dy"]
    lxg = dx * (nx - 1)
    lyg = dy * (ny - 1)
    xmid, ymid = 0.5 * lxg, 0.5 * lyg
    x = np.array([(ix - 1) * dx for ix in range(1, nx + 1)]) + 0.0
    y = np.array([(iy - 1) * dy for iy in range(1, ny + 1)]) + 0.0
    xp = np.array([(ix - 1) * dx + 0.5 * dx for ix in range(1, nx)]) + 0.0
    yp = np.array([(iy - 1) * dy + 0.5 * dy for iy in range(1, ny)]) + 0.0
    xr = np.array([(ix - 1) * dx - 0.5 * dx for ix in range(1, nx + 2)]) + 0.0
    S = O.solver_arrays(nx, ny)
    S["rho_vy"][:] = 3.3 * np.exp(-0.01 * ((xr[:, None] - 1.1 * xmid) ** 2 + (y[None, :] - 1.3 * ymid) ** 2))
    S["mu_b"][:] = np.exp(-0.025 * ((x[:, None] - 1.1 * xmid) ** 2 + (y[None, :] - 1.3 * ymid) ** 2))
    S["mu_p"][:] = np.exp(-0.025 * ((xp[:, None] - 1.1 * xmid) ** 2 + (yp[None, :] - 1.3 * ymid) ** 2))
    return S


def random_state(nx, ny, seed, scale=1.0, contrast=2.0):
    """Seeded smooth-ish material fields and a random PT state (for per-iteration parity)."""
    rng = np.random.default_rng(seed)
    S = O.solver_arrays(nx, ny)
    S["rho_vy"][:] = 1.0 + 0.2 * rng.random((nx + 1, ny))
    S["mu_b"][:] = 10.0 ** (contrast * rng.random((nx, ny)) - contrast / 2)
    S["mu_p"][:] = 10.0 ** (contrast * rng.random((nx - 1, ny - 1)) - contrast / 2)
    for k in ("P", "Vx_s", "Vy_s", "dVx", "dVy"):
        S[k][:] = scale * (rng.random(S[k].shape) - 0.5)
    return S


def unit_square(n):
    """dx, dy, lx, ly of an n x n node grid on [0, 1]^2."""
    dx = 1.0 / (n - 1)
    return dx, dx, (n - 1) * dx, (n - 1) * dx


def plume_markers(nx, ny, md, seed, jitter=0.25, mu_plume=0.1, mu_air=0.01):
    """Plume set-up of the benchmark workloads on the host (oracle generator).

    mu_air defaults to 1e-2 (the reference's exampleCall value): with 1e-3 the reference PT
    scheme diverges (see test_gpu_pt.py::test_divergence_reproduced)."""
    dx = 1.0 / (nx - 1)
    dy = 1.0 / (ny - 1)
    x, y = O.seeded_markers(nx, ny, md, dx, dy, jitter, seed)
    rho, mu = O.plume_properties(x, y, 0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, mu_air, mu_plume, 1.0)
    return dx, dy, x, y, rho, mu


def smooth_state(nx, ny, seed, contrast=1.0):
    """Smooth (Gaussian-bump) material fields on which the PT scheme converges, plus a small
    random initial state."""
    rng = np.random.default_rng(seed)
    S = O.solver_arrays(nx, ny)
    cx, cy = 0.3 + 0.4 * rng.random(), 0.3 + 0.4 * rng.random()

    def bump(n1, n2, ox, oy):
        X, Y = np.meshgrid((np.arange(n1) + ox) / (nx - 1), (np.arange(n2) + oy) / (ny - 1), indexing="ij")
        return np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / 0.05)

    S["rho_vy"][:] = 1.0 + 0.5 * bump(nx + 1, ny, -0.5, 0.0)
    S["mu_b"][:] = 10.0 ** (-contrast * bump(nx, ny, 0.0, 0.0))
    S["mu_p"][:] = 10.0 ** (-contrast * bump(nx - 1, ny - 1, 0.5, 0.5))
    for k in ("P", "Vx_s", "Vy_s", "dVx", "dVy"):
        S[k][:] = 1e-3 * (rng.random(S[k].shape) - 0.5)
    return S
