"""Batched ensemble solver (sf_solve_stokes_batched) against per-instance solves.

Each instance of the persistent, wide and grid modes runs the same PT tile code as the
single-instance solver, so its iterates, iteration count and dt must be bit-identical to
``solveStokes`` on the same inputs; err is a sum in another fixed order (rtol 1e-12).
One instance is also checked against the CPU oracle directly."""
import numpy as np
import pytest
import torch

import sfhip
from oracle import sf_oracle as O

import cases

pytestmark = pytest.mark.gpu

ARRAYS = ("P", "Vx_pad", "Vy_pad", "rho_vy", "mu_b", "mu_p", "dVx", "dVy", "Vx_s", "Vy_s")
OUT = ("P", "Vx_pad", "Vy_pad", "dVx", "dVy", "Vx_s", "Vy_s")


def plume_instance(nx, ny, seed, mu_plume=0.1):
    dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=seed, mu_plume=mu_plume)
    S = O.solver_arrays(nx, ny)
    vws, ws = O.zeros(nx + 1, ny), O.zeros(nx + 1, ny)
    o = O.grid_origins(dx, dy)
    O.marker_to_grid(*o["rho"], S["rho_vy"], x, y, rho, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_b"], S["mu_b"], x, y, mu, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_p"], S["mu_p"], x, y, mu, dx, dy, vws, ws)
    return S, dx, dy


def problem(S, nx, ny, dx, dy, **kw):
    G = {k: sfhip.from_numpy(S[k]) for k in ARRAYS}
    pb = {"P": G["P"], "Vx_pad": G["Vx_pad"], "Vy_pad": G["Vy_pad"], "dVxdtau": G["dVx"], "dVydtau": G["dVy"],
          "Vx_s": G["Vx_s"], "Vy_s": G["Vy_s"], "rho_vy": G["rho_vy"], "mu_b": G["mu_b"], "mu_p": G["mu_p"],
          "Nx": nx, "Ny": ny, "dx": dx, "dy": dy, "g_y": 1.0, "dt": 0.0, "maxdisp": 0.5}
    pb.update(kw)
    return pb, G


def single(S, nx, ny, dx, dy, eps=1e-8, iterMax=100000):
    G = {k: sfhip.from_numpy(S[k]) for k in ARRAYS}
    st = {}
    sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], None, None, None, None,
                      None, None, None, G["dVx"], G["dVy"], None, None, G["Vx_s"], G["Vy_s"], 1.0, dx, dy, nx, ny,
                      0.0, 0.5, eps=eps, iterMax=iterMax, print_info=False, stats=st)
    torch.cuda.synchronize()
    return st, {k: sfhip.to_numpy(G[k]) for k in OUT}


SIZES = [(24, 24), (33, 47), (40, 29), (64, 64), (17, 90)]


@pytest.mark.parametrize("prm", [{"mode": "persistent"}, {"mode": "persistent", "max_blocks": 1},
                                 {"mode": "persistent", "threads": 64},
                                 {"mode": "persistent", "threads": 320, "rows": 7},
                                 {"mode": "wide"}, {"mode": "wide", "budget_mb": 1e-3}, {"mode": "wide", "rows": 7},
                                 {"mode": "grid"}, {"mode": "grid", "depth": 2}, {"mode": "grid", "depth": 1}])
def test_batched_matches_single_solves(gpu, prm):
    insts = [(nx, ny, *plume_instance(nx, ny, seed=10 + k)) for k, (nx, ny) in enumerate(SIZES)]
    problems, gs = [], []
    for nx, ny, S, dx, dy in insts:
        pb, G = problem(S, nx, ny, dx, dy)
        problems.append(pb)
        gs.append(G)
    outs = sfhip.solveStokes_batched(problems, **prm)
    torch.cuda.synchronize()
    for (nx, ny, S, dx, dy), o, G in zip(insts, outs, gs):
        st, exp = single(S, nx, ny, dx, dy)
        assert o["converged"] and st["converged"], (nx, ny)
        assert o["niter"] == st["niter"], (nx, ny, o["niter"], st["niter"])
        assert o["dt_new"] == st["dt_new"] and o["dt_last"] == st["dt_last"], (nx, ny)
        assert o["err"] == pytest.approx(st["err"], rel=1e-12, abs=0.0)
        for k in OUT:
            assert np.array_equal(sfhip.to_numpy(G[k]), exp[k]), (nx, ny, k)


@pytest.mark.parametrize("mode", ["persistent", "wide", "grid"])
def test_batched_against_oracle(gpu, mode):
    nx, ny = 40, 29
    S, dx, dy = plume_instance(nx, ny, seed=5)
    pb, G = problem(S, nx, ny, dx, dy)
    (o,) = sfhip.solveStokes_batched([pb], mode=mode)
    torch.cuda.synchronize()
    exp = O.solve_stokes(S, 1.0, dx, dy, nx, ny, 0.0, 0.5, True, 1e-8)
    assert o["niter"] == exp["niter"] and o["dt_new"] == exp["dt_new"]
    for k in OUT:
        assert np.array_equal(sfhip.to_numpy(G[k]), S[k]), k


@pytest.mark.parametrize("mode", ["persistent", "wide", "grid"])
@pytest.mark.parametrize("cap", [37, 250])
def test_batched_iteration_cap(gpu, cap, mode):
    """An odd cap leaves the last iterate in the scratch buffers: it must be copied back."""
    insts = [(nx, ny, *plume_instance(nx, ny, seed=3 + k)) for k, (nx, ny) in enumerate([(24, 24), (33, 47)])]
    problems, gs = [], []
    for nx, ny, S, dx, dy in insts:
        pb, G = problem(S, nx, ny, dx, dy, iterMax=cap)
        problems.append(pb)
        gs.append(G)
    outs = sfhip.solveStokes_batched(problems, mode=mode)
    torch.cuda.synchronize()
    for (nx, ny, S, dx, dy), o, G in zip(insts, outs, gs):
        st, exp = single(S, nx, ny, dx, dy, iterMax=cap)
        assert o["niter"] == st["niter"] == cap and not o["converged"]
        for k in OUT:
            assert np.array_equal(sfhip.to_numpy(G[k]), exp[k]), (nx, ny, k)


@pytest.mark.parametrize("mode", ["persistent", "wide", "grid"])
def test_batched_no_stabilisation_and_mixed_eps(gpu, mode):
    nx, ny = 33, 47
    S, dx, dy = plume_instance(nx, ny, seed=8)
    pa, Ga = problem(S, nx, ny, dx, dy, use_free_surface_stabilization=False)
    pb, Gb = problem(S, nx, ny, dx, dy, eps=1e-5)
    oa, ob = sfhip.solveStokes_batched([pa, pb], mode=mode)
    torch.cuda.synchronize()
    Sa = {k: v.copy(order="F") for k, v in S.items()}
    xa = O.solve_stokes(Sa, 1.0, dx, dy, nx, ny, 0.0, 0.5, False, 1e-8)
    Sb = {k: v.copy(order="F") for k, v in S.items()}
    xb = O.solve_stokes(Sb, 1.0, dx, dy, nx, ny, 0.0, 0.5, True, 1e-5)
    assert oa["niter"] == xa["niter"] and ob["niter"] == xb["niter"] and ob["niter"] < xa["niter"]
    for k in OUT:
        assert np.array_equal(sfhip.to_numpy(Ga[k]), Sa[k]), k
        assert np.array_equal(sfhip.to_numpy(Gb[k]), Sb[k]), k


def test_batched_errors(gpu):
    assert sfhip.solveStokes_batched([]) == []
    S, dx, dy = plume_instance(24, 24, seed=1)
    pb, _ = problem(S, 24, 24, dx, dy)
    with pytest.raises(ValueError):
        sfhip.solveStokes_batched([dict(pb, Nx=25)])  # shape check on the host
    with pytest.raises(sfhip.SFError):
        sfhip.solveStokes_batched([pb], mode="persistent", threads=100)


def test_wide_refills_beyond_active_set_limit(gpu):
    """More instances than one wide launch can hold (64 slots): the active set must retire and
    refill many times, and every instance still matches its single solve."""
    sizes = [(12 + (k % 5), 10 + (k % 7)) for k in range(70)]
    insts = [(nx, ny, *plume_instance(nx, ny, seed=100 + k)) for k, (nx, ny) in enumerate(sizes)]
    problems, gs = [], []
    for nx, ny, S, dx, dy in insts:
        pb, G = problem(S, nx, ny, dx, dy)
        problems.append(pb)
        gs.append(G)
    outs = sfhip.solveStokes_batched(problems, mode="wide")
    torch.cuda.synchronize()
    for (nx, ny, S, dx, dy), o, G in zip(insts, outs, gs):
        st, exp = single(S, nx, ny, dx, dy)
        assert o["niter"] == st["niter"] and o["dt_new"] == st["dt_new"], (nx, ny)
        for k in OUT:
            assert np.array_equal(sfhip.to_numpy(G[k]), exp[k]), (nx, ny, k)


def single_refined(S, nx, ny, dx, dy, eps=1e-8, iterMax=100000):
    G = {k: sfhip.from_numpy(S[k]) for k in ARRAYS}
    st = {}
    sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], None, None, None, None,
                      None, None, None, G["dVx"], G["dVy"], None, None, G["Vx_s"], G["Vy_s"], 1.0, dx, dy, nx, ny,
                      0.0, 0.5, eps=eps, iterMax=iterMax, print_info=False, stats=st, mixed_precision=True)
    torch.cuda.synchronize()
    return st, {k: sfhip.to_numpy(G[k]) for k in OUT}


@pytest.mark.parametrize("prm", [{}, {"budget_mb": 1e-3}, {"rows": 7}])
def test_wide_refined_matches_single_refined(gpu, prm):
    """FP32 sweeps in FP64 defect correction, batched: every instance runs the algorithm of the
    single refined solver with the same kernels' arithmetic, so results are bitwise equal."""
    insts = [(nx, ny, *plume_instance(nx, ny, seed=60 + k)) for k, (nx, ny) in enumerate(SIZES)]
    problems, gs = [], []
    for nx, ny, S, dx, dy in insts:
        pb, G = problem(S, nx, ny, dx, dy)
        problems.append(pb)
        gs.append(G)
    outs = sfhip.solveStokes_batched(problems, mode="wide", mixed_precision=True, **prm)
    torch.cuda.synchronize()
    for (nx, ny, S, dx, dy), o, G in zip(insts, outs, gs):
        st, exp = single_refined(S, nx, ny, dx, dy)
        assert o["converged"] and st["converged"] and o["err"] <= 1e-8, (nx, ny)
        assert o["niter"] == st["niter"] and o["dt_new"] == st["dt_new"] and o["err"] == st["err"], (nx, ny)
        for k in OUT:
            assert np.array_equal(sfhip.to_numpy(G[k]), exp[k]), (nx, ny, k)


def test_wide_refined_iteration_cap_and_errors(gpu):
    nx, ny = 33, 47
    S, dx, dy = plume_instance(nx, ny, seed=9)
    pb, G = problem(S, nx, ny, dx, dy, iterMax=777)
    (o,) = sfhip.solveStokes_batched([pb], mode="wide", mixed_precision=True)
    torch.cuda.synchronize()
    st, exp = single_refined(S, nx, ny, dx, dy, iterMax=777)
    assert not o["converged"] and o["niter"] == st["niter"]
    for k in OUT:
        assert np.array_equal(sfhip.to_numpy(G[k]), exp[k]), k
    with pytest.raises(sfhip.SFError):
        sfhip.solveStokes_batched([pb], mode="persistent", mixed_precision=True)
    with pytest.raises(sfhip.SFError):
        sfhip.solveStokes_batched([pb], mode="grid", mixed_precision=True)


@pytest.mark.parametrize("mode", ["persistent", "wide", "grid"])
def test_batched_minimal_grids(gpu, mode):
    """The smallest grids the solver accepts, batched with a larger one (tile edge cases)."""
    sizes = [(3, 3), (4, 7), (5, 3), (63, 65)]
    insts = []
    for k, (nx, ny) in enumerate(sizes):
        S = cases.smooth_state(nx, ny, seed=200 + k)
        insts.append((nx, ny, S, 1.0 / (nx - 1), 1.0 / (ny - 1)))
    problems, gs = [], []
    for nx, ny, S, dx, dy in insts:
        pb, G = problem(S, nx, ny, dx, dy, iterMax=500)
        problems.append(pb)
        gs.append(G)
    outs = sfhip.solveStokes_batched(problems, mode=mode)
    torch.cuda.synchronize()
    for (nx, ny, S, dx, dy), o, G in zip(insts, outs, gs):
        st, exp = single(S, nx, ny, dx, dy, iterMax=500)
        assert o["niter"] == st["niter"], (nx, ny)
        for k in OUT:
            assert np.array_equal(sfhip.to_numpy(G[k]), exp[k]), (nx, ny, k)


def test_wide_refined_refills_beyond_active_set_limit(gpu):
    sizes = [(12 + (k % 5), 10 + (k % 7)) for k in range(70)]
    insts = [(nx, ny, *plume_instance(nx, ny, seed=400 + k)) for k, (nx, ny) in enumerate(sizes)]
    problems, gs = [], []
    for nx, ny, S, dx, dy in insts:
        pb, G = problem(S, nx, ny, dx, dy)
        problems.append(pb)
        gs.append(G)
    outs = sfhip.solveStokes_batched(problems, mode="wide", mixed_precision=True)
    torch.cuda.synchronize()
    for (nx, ny, S, dx, dy), o, G in zip(insts, outs, gs):
        st, exp = single_refined(S, nx, ny, dx, dy)
        assert o["niter"] == st["niter"] and o["dt_new"] == st["dt_new"], (nx, ny)
        for k in OUT:
            assert np.array_equal(sfhip.to_numpy(G[k]), exp[k]), (nx, ny, k)


def test_batch_autotune_keeps_results(gpu):
    """sf_batch_autotune times the wide passes per grid size (two to four iterations per pass,
    several tile heights) and keeps the fastest; a batched solve afterwards is still bitwise equal
    to one solve per instance, and the caller's arrays are untouched by the tuning."""
    sizes = [(24, 24), (33, 47), (24, 24), (64, 64)]
    insts = [(nx, ny, *plume_instance(nx, ny, seed=500 + k)) for k, (nx, ny) in enumerate(sizes)]
    problems, gs = [], []
    for nx, ny, S, dx, dy in insts:
        pb, G = problem(S, nx, ny, dx, dy)
        problems.append(pb)
        gs.append(G)
    before = [{k: v.clone() for k, v in G.items()} for G in gs]
    try:
        table = sfhip.autotune_batched(problems, iters=12)
        torch.cuda.synchronize()
        for G, B in zip(gs, before):
            assert all(torch.equal(G[k], B[k]) for k in G)
        assert {(t["nx"], t["ny"]) for t in table} == {(24, 24), (33, 47), (64, 64)}
        assert {t["iters_per_launch"] for t in table} == {2, 3, 4}
        assert all(t["ns_per_node_iter"] > 0 for t in table)
        assert [t["instances"] for t in table if (t["nx"], t["ny"]) == (24, 24)][0] == 2
        outs = sfhip.solveStokes_batched(problems, mode="wide")
        torch.cuda.synchronize()
    finally:
        sfhip.clear_batched_tuning()
    for (nx, ny, S, dx, dy), o, G in zip(insts, outs, gs):
        st, exp = single(S, nx, ny, dx, dy)
        assert o["niter"] == st["niter"] and o["dt_new"] == st["dt_new"], (nx, ny)
        for k in OUT:
            assert np.array_equal(sfhip.to_numpy(G[k]), exp[k]), (nx, ny, k)
