"""PT solver on the GPU (C ABI through sfhip) against the CPU oracle and the reference's
golden vectors.  FP64 iterations are compared bit for bit: the kernels evaluate the
reference's expressions in the same order with no FMA contraction."""
import numpy as np
import pytest
import torch

import sfhip
from oracle import sf_oracle as O

import cases

pytestmark = pytest.mark.gpu

STATE = ("P", "Vx_s", "Vy_s", "dVx", "dVy")


def to_gpu(S, keys):
    return {k: sfhip.from_numpy(S[k]) for k in keys}


def run_gpu_iterations(S, nx, ny, n_iter, g_y, dx, dy, dt):
    G = to_gpu(S, STATE + ("rho_vy", "mu_b", "mu_p"))
    sfhip.pt_iterations(n_iter, G["P"], G["Vx_s"], G["Vy_s"], G["dVx"], G["dVy"], G["rho_vy"], G["mu_b"], G["mu_p"],
                        g_y, dx, dy, nx, ny, dt)
    torch.cuda.synchronize()
    return {k: sfhip.to_numpy(G[k]) for k in STATE}


def oracle_iterations(S, nx, ny, n_iter, g_y, dx, dy, dt):
    O.compute_timesteps(S, nx, ny, dx, dy)
    O.pt_iterations(S, nx, ny, n_iter, g_y, dx, dy, dt)
    return {k: S[k] for k in STATE}


@pytest.mark.parametrize("n_iter", [1, 2, 7, 60])
def test_iterations_bitwise_golden_inputs(gpu, golden, n_iter):
    g = golden["stokes_solver"]
    nx, ny = g["nx"], g["ny"]
    S = cases.stokes_solver_inputs(g)
    rng = np.random.default_rng(n_iter)
    for k in STATE:  # a non-trivial state so every term is exercised
        S[k][:] = rng.random(S[k].shape) - 0.5
    got = run_gpu_iterations(S, nx, ny, n_iter, g["g_y"], g["dx"], g["dy"], 0.013)
    exp = oracle_iterations(S, nx, ny, n_iter, g["g_y"], g["dx"], g["dy"], 0.013)
    for k in STATE:
        assert np.array_equal(got[k], exp[k]), (k, np.max(np.abs(got[k] - exp[k])))


# sizes around the strip width (62 output columns) and the 64-row segments, minimal grids
@pytest.mark.parametrize("nx,ny", [(3, 3), (4, 7), (5, 3), (62, 10), (63, 65), (64, 64), (65, 129), (125, 66),
                                   (200, 31)])
def test_iterations_bitwise_sizes(gpu, nx, ny):
    S = cases.random_state(nx, ny, seed=nx * 1000 + ny)
    got = run_gpu_iterations(S, nx, ny, 5, 1.7, 0.11, 0.07, 0.02)
    exp = oracle_iterations(S, nx, ny, 5, 1.7, 0.11, 0.07, 0.02)
    for k in STATE:
        assert np.array_equal(got[k], exp[k]), (nx, ny, k)


# two-iteration passes (sf_ctx_set_pt_fusion): 60-column strips, tiles of `rows` rows; sizes
# around the strip width and tile heights, minimal grids, odd iteration counts (one plain
# iteration after the pairs)
@pytest.mark.parametrize("nx,ny,rows", [(3, 3, 0), (4, 4, 1), (3, 9, 2), (59, 5, 3), (60, 8, 8), (61, 9, 8),
                                        (62, 17, 16), (63, 31, 0), (121, 12, 5), (122, 40, 8), (200, 7, 4)])
@pytest.mark.parametrize("n_iter", [2, 5])
def test_pair_pass_bitwise(gpu, nx, ny, rows, n_iter):
    S = cases.random_state(nx, ny, seed=nx * 7919 + ny * 31 + rows)
    try:
        sfhip.set_pt_fusion(nx, ny, 2, rows)
        got = run_gpu_iterations(S, nx, ny, n_iter, 1.3, 0.09, 0.05, 0.017)
    finally:
        sfhip.set_pt_fusion(nx, ny, 0, 0)
    exp = oracle_iterations(S, nx, ny, n_iter, 1.3, 0.09, 0.05, 0.017)
    for k in STATE:
        assert np.array_equal(got[k], exp[k]), (nx, ny, rows, n_iter, k, np.max(np.abs(got[k] - exp[k])))


# three iterations per pass: 58-column strips; n_iter 7 = 3 + 3 + 1, 5 = 3 + 2, 4 = 3 + 1 launches
# the tall tiles of the size defaults (24-64 rows): several tiles per strip, a partial last one
@pytest.mark.parametrize("nx,ny,rows", [(3, 3, 0), (4, 5, 1), (5, 9, 2), (57, 6, 3), (58, 8, 8), (59, 13, 16),
                                        (60, 4, 4), (117, 21, 5), (118, 40, 16), (175, 9, 0), (175, 150, 24),
                                        (175, 150, 40), (175, 150, 48), (176, 203, 64)])
@pytest.mark.parametrize("n_iter", [3, 4, 5, 7])
def test_triple_pass_bitwise(gpu, nx, ny, rows, n_iter):
    S = cases.random_state(nx, ny, seed=nx * 7001 + ny * 37 + rows)
    try:
        sfhip.set_pt_fusion(nx, ny, 3, rows)
        got = run_gpu_iterations(S, nx, ny, n_iter, 1.1, 0.07, 0.06, 0.013)
    finally:
        sfhip.set_pt_fusion(nx, ny, 0, 0)
    exp = oracle_iterations(S, nx, ny, n_iter, 1.1, 0.07, 0.06, 0.013)
    for k in STATE:
        assert np.array_equal(got[k], exp[k]), (nx, ny, rows, n_iter, k, np.max(np.abs(got[k] - exp[k])))


# four iterations per pass: 56-column strips; n_iter 4, 5 = 4 + 1, 7 = 4 + 3, 10 = 4 + 4 + 2 launches
@pytest.mark.parametrize("nx,ny,rows", [(3, 3, 0), (4, 5, 1), (5, 9, 2), (55, 6, 3), (56, 8, 8), (57, 13, 16),
                                        (58, 4, 4), (113, 21, 5), (114, 40, 16), (169, 9, 0), (169, 150, 24),
                                        (169, 150, 40), (170, 203, 56)])
@pytest.mark.parametrize("n_iter", [4, 5, 7, 10])
def test_quad_pass_bitwise(gpu, nx, ny, rows, n_iter):
    S = cases.random_state(nx, ny, seed=nx * 6007 + ny * 41 + rows)
    try:
        sfhip.set_pt_fusion(nx, ny, 4, rows)
        got = run_gpu_iterations(S, nx, ny, n_iter, 1.2, 0.08, 0.05, 0.011)
    finally:
        sfhip.set_pt_fusion(nx, ny, 0, 0)
    exp = oracle_iterations(S, nx, ny, n_iter, 1.2, 0.08, 0.05, 0.011)
    for k in STATE:
        assert np.array_equal(got[k], exp[k]), (nx, ny, rows, n_iter, k, np.max(np.abs(got[k] - exp[k])))


def test_quad_pass_solve_golden(gpu, golden):
    """A full solve with four iterations per pass is bitwise the oracle's (checks, dt refreshes and
    the timing mark split the passes into four, three, two and one)."""
    g = golden["stokes_solver"]
    nx, ny = g["nx"], g["ny"]
    S = cases.stokes_solver_inputs(g)
    try:
        sfhip.set_pt_fusion(nx, ny, 4, 0)
        R, _, st = gpu_solve(S, nx, ny, g["g_y"], g["dx"], g["dy"], 0.0, g["maxdisp"], True, g["eps"], 100000)
    finally:
        sfhip.set_pt_fusion(nx, ny, 0, 0)
    exp = O.solve_stokes(S, g["g_y"], g["dx"], g["dy"], nx, ny, 0.0, g["maxdisp"], True, g["eps"])
    assert st["niter"] == exp["niter"]
    for k in S:
        assert np.array_equal(R[k], S[k]), k


def test_triple_pass_solve_golden(gpu, golden):
    """A full solve with three iterations per pass is bitwise the oracle's (checks, dt refreshes
    and the timing mark split the passes)."""
    g = golden["stokes_solver"]
    nx, ny = g["nx"], g["ny"]
    S = cases.stokes_solver_inputs(g)
    try:
        sfhip.set_pt_fusion(nx, ny, 3, 0)
        R, _, st = gpu_solve(S, nx, ny, g["g_y"], g["dx"], g["dy"], 0.0, g["maxdisp"], True, g["eps"], 100000)
    finally:
        sfhip.set_pt_fusion(nx, ny, 0, 0)
    exp = O.solve_stokes(S, g["g_y"], g["dx"], g["dy"], nx, ny, 0.0, g["maxdisp"], True, g["eps"])
    assert st["niter"] == exp["niter"]
    for k in S:
        assert np.array_equal(R[k], S[k]), k


def test_pair_and_single_launch_agree(gpu):
    """Forcing one iteration per launch gives the same bits as the size-default passes (two
    iterations each at this size)."""
    nx, ny = 97, 83
    S = cases.random_state(nx, ny, seed=5)
    try:
        sfhip.set_pt_fusion(nx, ny, 1, 0)
        one = run_gpu_iterations(S, nx, ny, 9, 0.8, 0.01, 0.012, 0.003)
    finally:
        sfhip.set_pt_fusion(nx, ny, 0, 0)
    two = run_gpu_iterations(S, nx, ny, 9, 0.8, 0.01, 0.012, 0.003)
    for k in STATE:
        assert np.array_equal(one[k], two[k]), k


def gpu_solve(S, nx, ny, g_y, dx, dy, dt, maxdisp, use_fss, eps, iter_max, scratch=True):
    keys = list(S.keys())
    G = to_gpu(S, keys)
    st = {}
    opt = (lambda k: G[k]) if scratch else (lambda k: None)
    res = sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], opt("txx"),
                            opt("tyy"), opt("txy"), opt("divV"), opt("dtauP"), opt("Rx"), opt("Ry"), G["dVx"],
                            G["dVy"], opt("dtauVx"), opt("dtauVy"), G["Vx_s"], G["Vy_s"], g_y, dx, dy, nx, ny, dt,
                            maxdisp, use_free_surface_stabilization=use_fss, eps=eps, print_info=False,
                            iterMax=iter_max, stats=st)
    torch.cuda.synchronize()
    return {k: sfhip.to_numpy(G[k]) for k in keys}, res, st


def test_solve_golden_and_oracle(gpu, golden):
    g = golden["stokes_solver"]
    nx, ny = g["nx"], g["ny"]
    S = cases.stokes_solver_inputs(g)
    R, (dt, t_eff, runtime), st = gpu_solve(S, nx, ny, g["g_y"], g["dx"], g["dy"], 0.0, g["maxdisp"], True, g["eps"],
                                            100000)
    exp = O.solve_stokes(S, g["g_y"], g["dx"], g["dy"], nx, ny, 0.0, g["maxdisp"], True, g["eps"])
    assert st["converged"] and st["niter"] == exp["niter"]
    assert st["err"] == pytest.approx(exp["err"], rel=1e-12)
    assert dt == exp["dt_new"]
    # every array the reference leaves filled is bit-identical to the oracle's
    for k in S:
        assert np.array_equal(R[k], S[k]), k
    inds = np.asarray(g["test_inds"]) - 1
    vx = R["Vx_pad"].ravel(order="F")[inds]
    vy = R["Vy_pad"].ravel(order="F")[inds]
    for got, ref in ((vx, np.array(g["vx_ref"])), (vy, np.array(g["vy_ref"]))):
        assert np.linalg.norm(got - ref) <= g["rtol"] * max(np.linalg.norm(got), np.linalg.norm(ref))
    Vx, Vy = R["Vx_pad"], R["Vy_pad"]
    assert np.all(Vx[0, :] == -Vx[2, :]) and np.all(Vx[1, :] == 0.0) and np.all(Vx[-2, :] == 0.0)
    assert np.all(Vy[:, 0] == -Vy[:, 2]) and np.all(Vy[:, 1] == 0.0) and np.all(Vy[:, -2] == 0.0)
    assert np.all(Vx[:, 0] == Vx[:, 1]) and np.all(Vy[0, :] == Vy[1, :])


@pytest.mark.parametrize("iter_max,use_fss", [(37, True), (0, True), (400, False)])
def test_solve_iter_cap_and_no_stabilisation(gpu, iter_max, use_fss):
    nx, ny = 40, 33
    S = cases.smooth_state(nx, ny, seed=5)
    R, (dt, _, _), st = gpu_solve(S, nx, ny, 1.0, 0.03, 0.025, 0.004, 0.5, use_fss, 1e-14, iter_max)
    exp = O.solve_stokes(S, 1.0, 0.03, 0.025, nx, ny, 0.004, 0.5, use_fss, 1e-14, iter_max)
    assert np.isfinite(exp["err"]) or iter_max < 165
    assert st["niter"] == exp["niter"] == iter_max and not st["converged"]
    assert dt == exp["dt_new"]
    for k in ("P", "Vx_s", "Vy_s", "dVx", "dVy", "Vx_pad", "Vy_pad", "txx", "tyy", "txy", "divV", "dtauP", "dtauVx",
              "dtauVy"):
        assert np.array_equal(R[k], S[k]), k
    if iter_max >= 200:  # at least one check happened: residual arrays match too
        assert np.array_equal(R["Rx"], S["Rx"]) and np.array_equal(R["Ry"], S["Ry"])


def test_solve_without_scratch_arrays(gpu):
    nx, ny = 30, 41
    S = cases.smooth_state(nx, ny, seed=8, contrast=2.0)
    R, (dt, _, _), st = gpu_solve(S, nx, ny, 1.0, 1 / 29, 1 / 40, 0.0, 0.5, True, 1e-8, 100000, scratch=False)
    exp = O.solve_stokes(S, 1.0, 1 / 29, 1 / 40, nx, ny, 0.0, 0.5, True, 1e-8)
    assert st["niter"] == exp["niter"] and st["converged"]
    for k in ("P", "Vx_s", "Vy_s", "dVx", "dVy", "Vx_pad", "Vy_pad"):
        assert np.array_equal(R[k], S[k]), k


def test_solve_rejects_bad_shapes(gpu):
    nx, ny = 10, 12
    S = cases.random_state(nx, ny, seed=1)
    G = to_gpu(S, list(S.keys()))
    with pytest.raises(ValueError):
        sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], None, None, None, None,
                          None, None, None, G["dVx"], G["dVy"], None, None, G["Vx_s"], G["Vy_s"], 1.0, 0.1, 0.1,
                          nx + 1, ny, 0.0, 0.5, print_info=False)
    with pytest.raises(ValueError):
        sfhip.solveStokes(G["P"].t().contiguous().t().contiguous(), G["Vx_pad"], G["Vy_pad"], G["rho_vy"],
                          G["mu_b"], G["mu_p"], None, None, None, None, None, None, None, G["dVx"], G["dVy"], None,
                          None, G["Vx_s"], G["Vy_s"], 1.0, 0.1, 0.1, nx, ny, 0.0, 0.5, print_info=False)


def test_divergence_reproduced(gpu):
    """Air viscosity 1e-3 against a matrix of 1 (63 x 63 cells): the reference PT scheme blows
    up.  The GPU path stops at the same check iteration with the same non-finite err."""
    nx = ny = 64
    dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=1001, mu_air=1e-3)
    S = O.solver_arrays(nx, ny)
    vws, ws = O.zeros(nx + 1, ny), O.zeros(nx + 1, ny)
    o = O.grid_origins(dx, dy)
    O.marker_to_grid(*o["rho"], S["rho_vy"], x, y, rho, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_b"], S["mu_b"], x, y, mu, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_p"], S["mu_p"], x, y, mu, dx, dy, vws, ws)
    R, (dt, _, _), st = gpu_solve(S, nx, ny, 1.0, dx, dy, 0.0, 0.5, True, 1e-8, 100000)
    exp = O.solve_stokes(S, 1.0, dx, dy, nx, ny, 0.0, 0.5, True, 1e-8)
    assert st["niter"] == exp["niter"] == 960
    assert np.isnan(st["err"]) and np.isnan(exp["err"]) and not st["converged"]
    assert np.isnan(dt) and np.isnan(exp["dt_new"])


# ------------------------------------------------------------------ FP32 sweeps in FP64 defect correction

def test_refined_matches_fp64_solver_golden(gpu, golden):
    """The refined solver takes the FP64 solver's iterations (the FP32 windows continue one PT
    iteration), meets the reference err test in FP64, and its solution agrees with the FP64
    oracle's to FP32 rounding levels."""
    g = golden["stokes_solver"]
    nx, ny = g["nx"], g["ny"]
    S = cases.stokes_solver_inputs(g)
    keys = list(S.keys())
    G = to_gpu(S, keys)
    st = {}
    sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], None, None, None, None,
                      None, None, None, G["dVx"], G["dVy"], None, None, G["Vx_s"], G["Vy_s"], g["g_y"], g["dx"],
                      g["dy"], nx, ny, 0.0, g["maxdisp"], eps=g["eps"], print_info=False, stats=st,
                      mixed_precision=True)
    torch.cuda.synchronize()
    exp = O.solve_stokes(S, g["g_y"], g["dx"], g["dy"], nx, ny, 0.0, g["maxdisp"], True, g["eps"])
    assert st["converged"] and st["err_fp64"] <= g["eps"]
    assert st["n_fp32_iter"] == exp["niter"] and st["n_fp64_iter"] == 1
    R = {k: sfhip.to_numpy(G[k]) for k in ("Vx_pad", "Vy_pad", "P")}
    for k in R:  # tolerance of the defect-corrected solution against the FP64 oracle
        rel = np.linalg.norm(R[k] - S[k]) / np.linalg.norm(S[k])
        assert rel <= 1e-6, (k, rel)
    inds = np.asarray(g["test_inds"]) - 1
    for a, ref in ((R["Vx_pad"], g["vx_ref"]), (R["Vy_pad"], g["vy_ref"])):
        got = a.ravel(order="F")[inds]
        assert np.linalg.norm(got - np.array(ref)) <= 1e-6 * np.linalg.norm(ref)
    Vx, Vy = R["Vx_pad"], R["Vy_pad"]  # the corrections keep the free-slip identities exactly
    assert np.all(Vx[1, :] == 0.0) and np.all(Vx[0, :] == -Vx[2, :]) and np.all(Vx[:, 0] == Vx[:, 1])
    assert np.all(Vy[:, 1] == 0.0) and np.all(Vy[:, 0] == -Vy[:, 2]) and np.all(Vy[0, :] == Vy[1, :])


@pytest.mark.parametrize("n,rebase", [(64, 0), (96, 97)])
def test_refined_plume(gpu, n, rebase):
    """Plume set-up (air mu 1e-2): same iteration count as the FP64 oracle (to the
    re-base granularity), reference err <= 1e-8, solution close to the oracle's."""
    dx, dy, x, y, rho, mu = cases.plume_markers(n, n, 4, seed=77)
    S = O.solver_arrays(n, n)
    vws, ws = O.zeros(n + 1, n), O.zeros(n + 1, n)
    o = O.grid_origins(dx, dy)
    O.marker_to_grid(*o["rho"], S["rho_vy"], x, y, rho, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_b"], S["mu_b"], x, y, mu, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_p"], S["mu_p"], x, y, mu, dx, dy, vws, ws)
    G = to_gpu(S, list(S.keys()))
    st = {}
    sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], None, None, None, None,
                      None, None, None, G["dVx"], G["dVy"], None, None, G["Vx_s"], G["Vy_s"], 1.0, dx, dy, n, n, 0.0,
                      0.5, eps=1e-8, print_info=False, stats=st, mixed_precision=True,
                      refine={"rebase_interval": rebase})
    torch.cuda.synchronize()
    exp = O.solve_stokes(S, 1.0, dx, dy, n, n, 0.0, 0.5, True, 1e-8)
    assert st["converged"] and st["err_fp64"] <= 1e-8
    # convergence is tested at multiples of ncheck, as in the reference; FP32 rounding may move
    # the first passing test by one ncheck
    ncheck = min(5 * n, 2000)
    assert st["n_fp32_iter"] % ncheck == 0
    assert abs(st["n_fp32_iter"] - exp["niter"]) <= ncheck
    for k in ("Vx_pad", "Vy_pad", "P"):
        a = sfhip.to_numpy(G[k])
        assert np.linalg.norm(a - S[k]) / np.linalg.norm(S[k]) <= 1e-4, k


@pytest.mark.parametrize("n,rebase", [(61, 0), (96, 97), (130, 45)])
def test_refined_pair_sweeps_bitwise(gpu, n, rebase):
    """FP32 correction sweeps one to five per pass give the same bits: windows of odd and even
    length (rebase 97 / 45 and ncheck), strips of 54, 56, 58, 60 and 62 columns."""
    dx, dy, x, y, rho, mu = cases.plume_markers(n, n + 3, 4, seed=78)
    S = O.solver_arrays(n, n + 3)
    vws, ws = O.zeros(n + 1, n + 3), O.zeros(n + 1, n + 3)
    o = O.grid_origins(dx, dy)
    O.marker_to_grid(*o["rho"], S["rho_vy"], x, y, rho, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_b"], S["mu_b"], x, y, mu, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_p"], S["mu_p"], x, y, mu, dx, dy, vws, ws)
    out = []
    for fuse in (1, 2, 3, 4, 5):
        G = to_gpu(S, list(S.keys()))
        st = {}
        try:
            sfhip.set_pt_fusion(n, n + 3, fuse, 0)
            sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], None, None, None,
                              None, None, None, None, G["dVx"], G["dVy"], None, None, G["Vx_s"], G["Vy_s"], 1.0, dx, dy,
                              n, n + 3, 0.0, 0.5, eps=1e-8, print_info=False, stats=st, mixed_precision=True,
                              refine={"rebase_interval": rebase})
        finally:
            sfhip.set_pt_fusion(n, n + 3, 0, 0)
        torch.cuda.synchronize()
        out.append((st, {k: sfhip.to_numpy(G[k]) for k in ("Vx_pad", "Vy_pad", "P", "dVx", "dVy")}))
    s1, a1 = out[0]
    for s2, a2 in out[1:]:
        assert s1["n_fp32_iter"] == s2["n_fp32_iter"] and s1["err_fp64"] == s2["err_fp64"]
        for k in a1:
            assert np.array_equal(a1[k], a2[k]), k


@pytest.mark.parametrize("rows", [24, 40])
def test_refined_sweeps_tall_tiles_bitwise(gpu, rows):
    """Four and five FP32 sweeps per pass with the tile heights of the size defaults (40 rows at
    2048^2 and 4096^2) on a grid several tiles tall: the same bits as one sweep per launch."""
    nx, ny = 130, 197
    dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=79)
    S = O.solver_arrays(nx, ny)
    vws, ws = O.zeros(nx + 1, ny), O.zeros(nx + 1, ny)
    o = O.grid_origins(dx, dy)
    O.marker_to_grid(*o["rho"], S["rho_vy"], x, y, rho, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_b"], S["mu_b"], x, y, mu, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_p"], S["mu_p"], x, y, mu, dx, dy, vws, ws)
    out = []
    for fuse in (1, 4, 5):
        G = to_gpu(S, list(S.keys()))
        st = {}
        try:
            sfhip.set_pt_fusion(nx, ny, fuse, rows if fuse > 1 else 0)
            sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], None, None, None,
                              None, None, None, None, G["dVx"], G["dVy"], None, None, G["Vx_s"], G["Vy_s"], 1.0, dx, dy,
                              nx, ny, 0.0, 0.5, eps=1e-8, print_info=False, iterMax=2999, stats=st,
                              mixed_precision=True, refine={"rebase_interval": 333})
        finally:
            sfhip.set_pt_fusion(nx, ny, 0, 0)
        torch.cuda.synchronize()
        out.append((st, {k: sfhip.to_numpy(G[k]) for k in ("Vx_pad", "Vy_pad", "P", "dVx", "dVy")}))
    s1, a1 = out[0]
    for s2, a2 in out[1:]:
        assert s1["n_fp32_iter"] == s2["n_fp32_iter"] and s1["err_fp64"] == s2["err_fp64"]
        for k in a1:
            assert np.array_equal(a1[k], a2[k]), k


def test_pass_timer_counts_launches_and_keeps_results(gpu):
    """sf_ctx_time_passes brackets each pass launch with events; the totals count the launches of
    each kind, and timing does not change a bit."""
    nx, ny = 150, 97
    S = cases.random_state(nx, ny, seed=12)
    exp = oracle_iterations({k: v.copy(order="F") for k, v in S.items()}, nx, ny, 23, 1.0, 0.01, 0.012, 0.003)
    try:
        sfhip.set_pt_fusion(nx, ny, 3, 8)
        sfhip.time_passes(nx, ny, True)
        got = run_gpu_iterations(S, nx, ny, 23, 1.0, 0.01, 0.012, 0.003)  # 7 passes of 3, one of 2
        t = sfhip.pass_times(nx, ny)
        assert sfhip.pass_times(nx, ny) == t  # reading twice gives the same totals
    finally:
        sfhip.time_passes(nx, ny, False)
        sfhip.set_pt_fusion(nx, ny, 0, 0)
    assert [(e["precision"], e["iters_per_launch"], e["launches"]) for e in t] == [(0, 2, 1), (0, 3, 7)]
    assert all(e["seconds"] > 0.0 for e in t)
    for k in STATE:
        assert np.array_equal(got[k], exp[k]), k


def test_autotune_keeps_results(gpu, golden):
    """The autotuner only picks launch shapes (one-iteration kernel and 2- to 4-iteration passes):
    iterations after tuning are bitwise the oracle's."""
    try:
        best, table = sfhip.autotune_pt(63, 47, iters=5)
        singles = [t for t in table if t["iters_per_launch"] == 1]
        passes = [t for t in table if t["iters_per_launch"] > 1]
        assert len(singles) == 20 and {t["iters_per_launch"] for t in passes} == {2, 3, 4}
        assert all(t["rows"] <= 47 for t in passes)
        assert best["us_per_iter"] == min(t["us_per_iter"] for t in table)
        S = cases.smooth_state(63, 47, seed=4)
        exp = {k: v.copy(order="F") for k, v in S.items()}
        got = run_gpu_iterations(S, 63, 47, 9, 1.0, 1.0 / 62, 1.0 / 46, 0.01)
        ref = oracle_iterations(exp, 63, 47, 9, 1.0, 1.0 / 62, 1.0 / 46, 0.01)
        for k in STATE:
            assert np.array_equal(got[k], ref[k]), k
    finally:
        sfhip.set_pt_tuning(63, 47, 0, 0, -1)
        sfhip.set_pt_fusion(63, 47, 0, 0)


def test_dram_probe(gpu):
    r = sfhip.probe_dram(1 << 26)
    assert 500.0 < r["copy_gbs"] < 9000.0 and 500.0 < r["r3w1_gbs"] < 9000.0
