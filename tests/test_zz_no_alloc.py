"""sf_solve_stokes_noalloc (no context, scratch in a static device array) against sf_solve_stokes
on the same inputs, through the C ABI: two calls in a row, the second on a smaller grid (the
static scratch keeps the first call's contents), each bitwise equal to the context solve.  Also
the size limit, which is checked on the host before anything runs."""
import ctypes as C

import numpy as np
import pytest
import torch

import sfhip
from sfhip import _lib
from oracle import sf_oracle as O

import cases

pytestmark = pytest.mark.gpu

STATE = ("P", "Vx_pad", "Vy_pad", "dVx", "dVy", "Vx_s", "Vy_s")
SCRATCH = ("txx", "tyy", "txy", "divV", "dtauP", "Rx", "Ry", "dtauVx", "dtauVy")


def plume_problem(n, seed):
    dx, dy, x, y, rho, mu = cases.plume_markers(n, n, 4, seed=seed)
    S = O.solver_arrays(n, n)
    vws, ws = O.zeros(n + 1, n), O.zeros(n + 1, n)
    o = O.grid_origins(dx, dy)
    O.marker_to_grid(*o["rho"], S["rho_vy"], x, y, rho, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_b"], S["mu_b"], x, y, mu, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_p"], S["mu_p"], x, y, mu, dx, dy, vws, ws)
    return S, dx, dy


def solve(S, n, dx, dy, no_alloc, scratch, iter_max):
    G = {k: sfhip.from_numpy(S[k]) for k in S}
    opt = (lambda k: G[k]) if scratch else (lambda k: None)
    st = {}
    res = sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], opt("txx"),
                            opt("tyy"), opt("txy"), opt("divV"), opt("dtauP"), opt("Rx"), opt("Ry"), G["dVx"],
                            G["dVy"], opt("dtauVx"), opt("dtauVy"), G["Vx_s"], G["Vy_s"], 1.0, dx, dy, n, n, 0.0, 0.5,
                            eps=1e-8, print_info=False, iterMax=iter_max, stats=st, no_alloc=no_alloc)
    torch.cuda.synchronize()
    return {k: sfhip.to_numpy(G[k]) for k in G}, res, st


@pytest.mark.parametrize("scratch", [False, True])
def test_noalloc_matches_context_solve_two_calls(gpu, scratch):
    for n, seed, iter_max in ((200, 31, 100000), (61, 32, 100000)):  # second call: smaller grid
        print(f"noalloc solve n={n} seed={seed} scratch={scratch}", flush=True)
        S, dx, dy = plume_problem(n, seed)
        a, (dt_a, _, _), sa = solve(S, n, dx, dy, True, scratch, iter_max)
        b, (dt_b, _, _), sb = solve(S, n, dx, dy, False, scratch, iter_max)
        assert sa["niter"] == sb["niter"] and sa["converged"] == sb["converged"], (n, sa["niter"], sb["niter"])
        assert dt_a == dt_b and sa["err"] == sb["err"] and sa["dt_last"] == sb["dt_last"], n
        assert sa["runtime"] > 0.0 and sa["t_eff"] > 0.0
        for k in STATE + (SCRATCH if scratch else ()):
            assert np.array_equal(a[k], b[k]), (n, k)


def test_noalloc_iteration_cap_and_oracle(gpu):
    """An odd cap (the last iterate sits in the static scratch and must be copied back) against the
    oracle directly."""
    n = 47
    S, dx, dy = plume_problem(n, 5)
    a, (dt_a, _, _), sa = solve(S, n, dx, dy, True, True, 333)
    exp = O.solve_stokes(S, 1.0, dx, dy, n, n, 0.0, 0.5, True, 1e-8, 333)
    assert sa["niter"] == exp["niter"] == 333 and dt_a == exp["dt_new"]
    for k in STATE + ("txx", "tyy", "txy", "divV", "Rx", "Ry"):
        assert np.array_equal(a[k], S[k]), k


def test_noalloc_rejects_grids_beyond_the_static_scratch(gpu):
    """Checked on the host from the grid size: nothing is launched."""
    lib = _lib.load()
    n = 2000  # about 88 B per node with every scratch argument NULL: 352 MB > 64 MiB
    null = [None] * 19
    out = _lib.SolveOut()
    rc = lib.sf_solve_stokes_noalloc(None, *null, 1.0, 1.0 / (n - 1), 1.0 / (n - 1), n, n, 0.0, 0.5, 1, 1e-8, 10, 0,
                                     C.byref(out))
    assert rc == _lib.SF_ERR_SIZE
    assert b"static" in lib.sf_last_error()
