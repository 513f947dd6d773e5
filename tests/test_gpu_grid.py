"""Grid mode (the whole PT loop of one problem in one chip-wide launch) and the captured
timestep of one problem (sf_graph_*).

Every pass runs the single solver's tile code, so fields, iteration counts and dt must be
bit-identical to ``solveStokes`` / ``step``; err is a sum in another order (rtol 1e-12).  The
larger grids here have more tiles per pass than the launch has waves, so each wave runs
several tiles per pass, and cross checks, dt refreshes and the iteration cap."""
import numpy as np
import pytest
import torch

import sfhip
from oracle import sf_oracle as O

import cases
from test_gpu_batch import OUT, plume_instance, problem, single
from test_gpu_ensemble import member, same

pytestmark = pytest.mark.gpu


def check_against_single(insts, outs, gs, **kw):
    for (nx, ny, S, dx, dy), o, G in zip(insts, outs, gs):
        st, exp = single(S, nx, ny, dx, dy, **kw)
        assert o["niter"] == st["niter"], (nx, ny, o["niter"], st["niter"])
        assert o["converged"] == st["converged"], (nx, ny)
        assert o["dt_new"] == st["dt_new"] and o["dt_last"] == st["dt_last"], (nx, ny)
        assert o["err"] == pytest.approx(st["err"], rel=1e-12, abs=0.0)
        for k in OUT:
            assert np.array_equal(sfhip.to_numpy(G[k]), exp[k]), (nx, ny, k)


@pytest.mark.parametrize("depth", [3, 2])
def test_grid_mode_converged_plumes(gpu, depth):
    sizes = [(200, 150), (301, 257)]
    insts = [(nx, ny, *plume_instance(nx, ny, seed=600 + k)) for k, (nx, ny) in enumerate(sizes)]
    problems, gs = [], []
    for nx, ny, S, dx, dy in insts:
        pb, G = problem(S, nx, ny, dx, dy)
        problems.append(pb)
        gs.append(G)
    print(f"grid mode, depth {depth}, sizes {sizes}", flush=True)
    outs = sfhip.solveStokes_batched(problems, mode="grid", depth=depth)
    torch.cuda.synchronize()
    assert all(o["converged"] for o in outs)
    check_against_single(insts, outs, gs)


@pytest.mark.parametrize("cap", [1201, 2000])
def test_grid_mode_cap_with_checks_and_dt_refresh(gpu, cap):
    """600 x 520: ncheck 2000, ndtupdate 5200; a 1201-iteration cap ends on a single iteration
    after an odd number of passes."""
    nx, ny = 600, 520
    S = cases.smooth_state(nx, ny, seed=610)
    insts = [(nx, ny, S, 1.0 / (nx - 1), 1.0 / (ny - 1))]
    pb, G = problem(S, nx, ny, insts[0][3], insts[0][4], iterMax=cap)
    print(f"grid mode 600x520 cap {cap}", flush=True)
    outs = sfhip.solveStokes_batched([pb], mode="grid")
    torch.cuda.synchronize()
    assert outs[0]["niter"] == cap
    check_against_single(insts, outs, [G], iterMax=cap)


def test_grid_mode_more_tiles_than_waves(gpu):
    """3000 x 2900 has more tiles per three-iteration pass than the chip holds waves."""
    nx, ny = 3000, 2900
    S = cases.smooth_state(nx, ny, seed=620)
    insts = [(nx, ny, S, 1.0 / (nx - 1), 1.0 / (ny - 1))]
    pb, G = problem(S, nx, ny, insts[0][3], insts[0][4], iterMax=14)
    print("grid mode 3000x2900, 14 iterations", flush=True)
    outs = sfhip.solveStokes_batched([pb], mode="grid")
    torch.cuda.synchronize()
    check_against_single(insts, outs, [G], iterMax=14)


SIZES = [(63, 47), (130, 97)]


@pytest.mark.parametrize("ipp", [0, 1])
def test_step_graph_matches_steps(gpu, ipp):
    """Replays of the captured step, with eager work in between, against sf_step calls."""
    for k, (nx, ny) in enumerate(SIZES):
        g_m, r_m = member(nx, ny, 700 + k), member(nx, ny, 700 + k)
        g = sfhip.StepGraph(g_m["state"], g_m["x_m"], g_m["y_m"], g_m["rho_m"], g_m["mu_m"], g_m["dx"], g_m["dy"],
                            0.0, g_y=1.0, eps=1e-8, iters_per_pass=ipp)
        dt = 0.0
        for t in range(3):
            gdt, gst = g.step(1)
            torch.cuda.synchronize()
            dt, st = sfhip.step(r_m["state"], r_m["x_m"], r_m["y_m"], r_m["rho_m"], r_m["mu_m"], r_m["dx"],
                                r_m["dy"], dt, g_y=1.0, eps=1e-8)
            torch.cuda.synchronize()
            assert gst["niter"] == st["niter"] and gdt == dt and gst["dt_new"] == st["dt_new"], (nx, ny, t)
            assert gst["err"] == pytest.approx(st["err"], rel=1e-12, abs=0.0)
            same(g_m, r_m)
        g.close()


def test_step_graph_multi_step_and_oracle_loop(gpu):
    """Two timesteps in one call against the oracle's own timestep loop
    (StokesFlow2D_multixpu.jl:155-218 at one rank)."""
    nx, ny = 40, 29
    m = member(nx, ny, 90)
    dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=90)
    g = sfhip.StepGraph(m["state"], m["x_m"], m["y_m"], m["rho_m"], m["mu_m"], m["dx"], m["dy"], 0.0, g_y=1.0,
                        eps=1e-8)
    gdt, gst = g.step(2)
    torch.cuda.synchronize()
    g.close()
    S, _, outs = O.stokesflow2d(2, nx, ny, 1.0, 1.0, x, y, rho, mu, g_y=1.0, eps=1e-8)  # dx = 1 / (nx - 1)
    assert gst["niter"] == outs[1]["niter"] and gdt == outs[1]["dt_new"]
    st = m["state"]
    for f, o in (("Vx", "Vx_pad"), ("Vy", "Vy_pad"), ("P", "P"), ("rho_vy", "rho_vy"), ("mu_b", "mu_b"),
                 ("mu_p", "mu_p"), ("dVxdtau", "dVx"), ("dVydtau", "dVy")):
        assert np.array_equal(sfhip.to_numpy(getattr(st, f)), S[o]), f
    assert np.array_equal(sfhip.to_numpy(m["x_m"]), x) and np.array_equal(sfhip.to_numpy(m["y_m"]), y)


def test_step_graph_errors(gpu):
    m = member(24, 24, 3)
    with pytest.raises(sfhip.SFError):
        sfhip.StepGraph(m["state"], m["x_m"], m["y_m"], m["rho_m"], m["mu_m"], m["dx"], m["dy"], iters_per_pass=4)
    m["x_m"][5] = float("nan")
    g = sfhip.StepGraph(m["state"], m["x_m"], m["y_m"], m["rho_m"], m["mu_m"], m["dx"], m["dy"])
    with pytest.raises(sfhip.SFError):
        g.step(1)
    g.close()


def test_stokesflow2d_captured_matches_eager(gpu):
    """The timestep loop with each step one hipGraph replay equals the eager loop bit for bit."""
    nx, ny = 48, 33
    outs = []
    for captured in (False, True):
        dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=77)
        X, Y, R, M = (sfhip.from_numpy(v) for v in (x, y, rho, mu))
        vx, vy, t_tot = sfhip.StokesFlow2D(3, nx, ny, 1.0, 1.0, X, Y, R, M, stokes_eps=1e-8, g_y=1.0,
                                           captured=captured)
        torch.cuda.synchronize()
        outs.append((sfhip.to_numpy(vx), sfhip.to_numpy(vy), t_tot, sfhip.to_numpy(X), sfhip.to_numpy(Y)))
    (a_vx, a_vy, a_t, a_x, a_y), (b_vx, b_vy, b_t, b_x, b_y) = outs
    assert a_t == b_t
    assert np.array_equal(a_vx, b_vx) and np.array_equal(a_vy, b_vy)
    assert np.array_equal(a_x, b_x) and np.array_equal(a_y, b_y)


def test_grid_wait_limit_error_path(gpu):
    """A zero wait limit makes the first barrier wait that finds its barrier incomplete fail: every
    workgroup leaves the kernel and the call returns an error; the default limit then solves
    normally (the failure is forced from the host; no kernel waits without bound)."""
    nx, ny = 600, 520
    S = cases.smooth_state(nx, ny, seed=630)
    dx, dy = 1.0 / (nx - 1), 1.0 / (ny - 1)
    pb, _ = problem(S, nx, ny, dx, dy, iterMax=10)
    try:
        sfhip.set_grid_wait_limit(0.0)
        with pytest.raises(sfhip.SFError, match="exceeded 0 s"):
            sfhip.solveStokes_batched([pb], mode="grid")
    finally:
        sfhip.set_grid_wait_limit(5.0)
    with pytest.raises(sfhip.SFError):
        sfhip.set_grid_wait_limit(-1.0)
    pb, G = problem(S, nx, ny, dx, dy, iterMax=10)
    outs = sfhip.solveStokes_batched([pb], mode="grid")
    torch.cuda.synchronize()
    check_against_single([(nx, ny, S, dx, dy)], outs, [G], iterMax=10)
