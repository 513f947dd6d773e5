"""Timestep loop (marker->grid x3 -> PT solve -> RK4) on the GPU against the oracle's loop,
the reference's end-to-end golden vectors, and size-independent properties at the
511 x 511-cell workload."""
import numpy as np
import pytest
import torch

import sfhip
from oracle import sf_oracle as O

import cases

pytestmark = pytest.mark.gpu


def run_both(nt, nx, ny, lx_glob, ly_glob, x, y, rho, mu, g_y, eps, device_markers=None):
    """Run nt steps on the GPU (sf_step) and in the oracle; compare after every step.  The GPU
    runs on device_markers (x, y, rho, mu already on the device) when given, else on copies of
    the host arrays."""
    dx, dy = lx_glob / (nx - 1), ly_glob / (ny - 1)
    if device_markers is not None:
        xg, yg, rg, mg = device_markers
    else:
        xg, yg, rg, mg = (sfhip.from_numpy(a) for a in (x, y, rho, mu))
    st = sfhip.FlowState(nx, ny)
    S = O.solver_arrays(nx, ny)
    vws, ws = O.zeros(nx + 1, ny), O.zeros(nx + 1, ny)
    o = O.grid_origins(dx, dy)
    xh, yh = x.copy(), y.copy()
    dt_g = dt_o = 0.0
    worst = 0.0
    for t in range(nt):
        # step-0 cell keys are bit-exact by construction (identical inputs); later steps carry V
        dt_g, stats = sfhip.step(st, xg, yg, rg, mg, dx, dy, dt_g, g_y=g_y, eps=eps)
        O.marker_to_grid(*o["rho"], S["rho_vy"], xh, yh, rho, dx, dy, vws, ws)
        O.marker_to_grid(*o["mu_b"], S["mu_b"], xh, yh, mu, dx, dy, vws, ws)
        O.marker_to_grid(*o["mu_p"], S["mu_p"], xh, yh, mu, dx, dy, vws, ws)
        res = O.solve_stokes(S, g_y, dx, dy, nx, ny, dt_o, 0.5, True, eps)
        dt_o = res["dt_new"]
        O.move_markers_rk4(xh, yh, S["Vx_pad"], S["Vy_pad"], *o["vx"], *o["vy"], dt_o, (nx - 1) * dx,
                           (ny - 1) * dy, dx, dy)
        torch.cuda.synchronize()
        assert stats["niter"] == res["niter"], (t, stats["niter"], res["niter"])
        for name, ga, ca in (("Vx", st.Vx, S["Vx_pad"]), ("Vy", st.Vy, S["Vy_pad"]), ("P", st.P, S["P"]),
                             ("rho", st.rho_vy, S["rho_vy"]), ("mu_b", st.mu_b, S["mu_b"]),
                             ("mu_p", st.mu_p, S["mu_p"])):
            a = sfhip.to_numpy(ga)
            rel = np.linalg.norm(a - ca) / max(np.linalg.norm(ca), 1e-300)
            worst = max(worst, rel)
            assert rel <= 1e-10, (t, name, rel)
        assert np.array_equal(sfhip.to_numpy(xg), xh) and np.array_equal(sfhip.to_numpy(yg), yh), t
        assert dt_g == dt_o
    return st, S, worst


def test_flow_golden_pipeline(gpu, golden):
    f = golden["stokes_flow_2d"]
    nx, ny, L = f["nx"], f["ny"], f["lx"]
    dx, dy = L / (nx - 1), f["ly"] / (ny - 1)
    x, y = O.lattice_markers(nx, ny, f["marker_density"], (nx - 1) * dx, (ny - 1) * dy, dx, dy)
    rho, mu = O.plume_properties(x, y, f["air_height"], f["plume_x"], f["plume_y"], f["plume_r"], f["rho_air"],
                                 f["rho_plume"], f["rho_matrix"], f["mu_air"], f["mu_plume"], f["mu_matrix"])
    st, S, worst = run_both(f["nt"], nx, ny, L, f["ly"], x, y, rho, mu, f["g_y"], f["eps"])
    inds = np.asarray(f["inds"]) - 1
    # test/test_StokesFlow2D.jl:107-108: isapprox(..., atol=1e-6) per element
    np.testing.assert_allclose(sfhip.to_numpy(st.Vx).ravel(order="F")[inds], f["vx_ref"], rtol=0, atol=f["atol"])
    np.testing.assert_allclose(sfhip.to_numpy(st.Vy).ravel(order="F")[inds], f["vy_ref"], rtol=0, atol=f["atol"])


def test_flow_golden_pipeline_device_markers(gpu, golden):
    """The end-to-end golden run with the markers set up on the device: the reference's lattice
    (initializeMarkersCPU) and its plume callbacks as a material descriptor
    (setInitialMarkerProperties!), then 5 timesteps, against the oracle loop bit for bit and the
    reference's vectors at its own tolerance."""
    f = golden["stokes_flow_2d"]
    nx, ny, L = f["nx"], f["ny"], f["lx"]
    dx, dy = L / (nx - 1), f["ly"] / (ny - 1)
    lx, ly = (nx - 1) * dx, (ny - 1) * dy
    xg, yg, rg, mg = sfhip.initializeMarkers(f["marker_density"], lx, ly, dx, dy, nx, ny)
    mats = sfhip.Materials([("y_below", f["air_height"], f["rho_air"], f["mu_air"]),
                            ("circle", f["plume_x"], f["plume_y"], f["plume_r"], f["rho_plume"], f["mu_plume"])],
                           rho_bg=f["rho_matrix"], mu_bg=f["mu_matrix"])
    sfhip.setInitialMarkerProperties(xg, yg, rg, mg, 0.0, 0.0, mats)
    x, y = O.lattice_markers(nx, ny, f["marker_density"], lx, ly, dx, dy)
    rho, mu = O.plume_properties(x, y, f["air_height"], f["plume_x"], f["plume_y"], f["plume_r"], f["rho_air"],
                                 f["rho_plume"], f["rho_matrix"], f["mu_air"], f["mu_plume"], f["mu_matrix"])
    for a, b in ((xg, x), (yg, y), (rg, rho), (mg, mu)):
        assert np.array_equal(sfhip.to_numpy(a), b)
    st, S, worst = run_both(f["nt"], nx, ny, L, f["ly"], x, y, rho, mu, f["g_y"], f["eps"],
                            device_markers=(xg, yg, rg, mg))
    inds = np.asarray(f["inds"]) - 1
    np.testing.assert_allclose(sfhip.to_numpy(st.Vx).ravel(order="F")[inds], f["vx_ref"], rtol=0, atol=f["atol"])
    np.testing.assert_allclose(sfhip.to_numpy(st.Vy).ravel(order="F")[inds], f["vy_ref"], rtol=0, atol=f["atol"])


def test_plume_64_two_steps(gpu):
    """The 63 x 63-cell plume set-up (16 markers per cell, seeded jitter, g = 1, air mu 1e-2), 2 steps."""
    nx = ny = 64
    dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=1001)
    run_both(2, nx, ny, 1.0, 1.0, x, y, rho, mu, 1.0, 1e-8)


def test_plume_512_properties(gpu):
    """511 x 511 cells, 16 markers per cell: one step; properties that hold at any size."""
    nx = ny = 512
    dx = dy = 1.0 / (nx - 1)
    x, y = sfhip.init_markers(nx, ny, 4, dx, dy, 0.25, seed=2001)
    rho, mu = sfhip.plume_properties(x, y, 0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, 0.01, 0.1, 1.0)
    assert x.numel() == 4_177_936
    x0, y0 = x.clone(), y.clone()
    st = sfhip.FlowState(nx, ny)
    dt, stats = sfhip.step(st, x, y, rho, mu, dx, dy, 0.0, g_y=1.0, eps=1e-8, iterMax=4000)
    # marker count conserved, positions inside the domain, displacement <= maxdisp cells
    assert torch.all((x >= 0) & (x <= 1.0)) and torch.all((y >= 0) & (y <= 1.0))
    disp = torch.maximum((x - x0).abs() / dx, (y - y0).abs() / dy).max().item()
    assert disp <= 0.75  # RK4-averaged interpolated velocity may exceed the nodal maximum slightly
    assert stats["m2g"] == [(0, 0), (0, 0), (0, 0)]
    # deterministic: the same step from the same inputs is bit-identical
    x2, y2 = x0.clone(), y0.clone()
    st2 = sfhip.FlowState(nx, ny)
    dt2, stats2 = sfhip.step(st2, x2, y2, rho, mu, dx, dy, 0.0, g_y=1.0, eps=1e-8, iterMax=4000)
    assert dt2 == dt and stats2["niter"] == stats["niter"]
    assert torch.equal(x, x2) and torch.equal(y, y2) and torch.equal(st.Vx, st2.Vx) and torch.equal(st.P, st2.P)
    # free-slip identities on the padded output
    Vx, Vy = st.Vx, st.Vy
    assert torch.equal(Vx[1, :], torch.zeros_like(Vx[1, :])) and torch.equal(Vx[0, :], -Vx[2, :])
    assert torch.equal(Vy[:, 1], torch.zeros_like(Vy[:, 1])) and torch.equal(Vy[:, 0], -Vy[:, 2])
    # the returned dt is the CFL value of the final V
    vmax_x = st.Vx_small.abs().max().item()
    vmax_y = st.Vy_small.abs().max().item()
    assert dt == 0.5 * min(dx / vmax_x, dy / vmax_y)
