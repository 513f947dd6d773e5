"""Marker -> grid and RK4 advection on the GPU against the CPU oracle (bit for bit), the
reference's known answers, and the seeded marker generator."""
import numpy as np
import pytest
import torch

import sfhip
from oracle import sf_oracle as O

import cases

pytestmark = pytest.mark.gpu


def m2g_gpu(xmin, ymin, grid0, x, y, v, dx, dy, helper_shape=None, helper_fill=0.0):
    vg = sfhip.from_numpy(grid0)
    vws = ws = None
    if helper_shape is not None:
        vws = sfhip.from_numpy(np.full(helper_shape, helper_fill, order="F"))
        ws = sfhip.from_numpy(np.full(helper_shape, helper_fill, order="F"))
    st = {}
    sfhip.bilinearMarkerToGrid(xmin, ymin, vg, sfhip.from_numpy(x), sfhip.from_numpy(y), sfhip.from_numpy(v), dx, dy,
                               vws, ws, stats=st, print_info=False)
    torch.cuda.synchronize()
    h = (sfhip.to_numpy(vws), sfhip.to_numpy(ws)) if ws is not None else None
    return sfhip.to_numpy(vg), h, st


def test_m2g_known_answer(gpu, golden):
    k = golden["marker_to_grid_kat"]
    vg, (vws, ws), st = m2g_gpu(k["x_grid_min"], k["y_grid_min"], np.zeros(k["grid_shape"], order="F"),
                                np.array(k["x_m"]), np.array(k["y_m"]), np.array(k["val_m"]), k["dx"], k["dy"],
                                helper_shape=k["grid_shape"], helper_fill=k["helper_fill"])
    np.testing.assert_allclose(vg, np.array(k["reference_columns"]).T, rtol=1.5e-8)
    assert st == {"n_bad_markers": 0, "n_empty_nodes": 0}


@pytest.mark.parametrize("origin", ["rho", "mu_b", "mu_p"])
@pytest.mark.parametrize("nx,ny,md,seed", [(9, 7, 4, 1), (64, 64, 4, 2), (130, 67, 3, 3)])
def test_m2g_bitwise_vs_oracle(gpu, origin, nx, ny, md, seed):
    dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, md, seed)
    rng = np.random.default_rng(seed)
    val = rho * (1.0 + rng.random(x.size))
    # a few markers outside every grid, and a shuffled order (changes the summation order)
    x = np.concatenate([x, [-3 * dx, 0.5, 2.0]])
    y = np.concatenate([y, [0.5, -4 * dy, 2.0]])
    val = np.concatenate([val, [7.0, 7.0, 7.0]])
    perm = rng.permutation(x.size)
    x, y, val = x[perm], y[perm], val[perm]
    o = O.grid_origins(dx, dy)[origin]
    shape = {"rho": (nx + 1, ny), "mu_b": (nx, ny), "mu_p": (nx - 1, ny - 1)}[origin]
    grid0 = np.asfortranarray(rng.random(shape))
    vg, (vws, ws), st = m2g_gpu(o[0], o[1], grid0, x, y, val, dx, dy, helper_shape=(nx + 1, ny), helper_fill=3.0)
    eg = grid0.copy(order="F")
    evws, ews = O.zeros(nx + 1, ny), O.zeros(nx + 1, ny)
    nb, ne = O.marker_to_grid(o[0], o[1], eg, x, y, val, dx, dy, evws, ews)
    assert np.array_equal(vg, eg)
    assert np.array_equal(vws, evws) and np.array_equal(ws, ews)
    assert (st["n_bad_markers"], st["n_empty_nodes"]) == (nb, ne)
    assert nb >= 2


def test_m2g_empty_nodes_keep_old_values(gpu):
    grid0 = np.asfortranarray(np.arange(20.0).reshape(4, 5, order="F"))
    x = np.array([0.01, 0.02])
    y = np.array([0.01, 0.015])
    vg, _, st = m2g_gpu(0.0, 0.0, grid0, x, y, np.array([2.0, 4.0]), 0.1, 0.1)
    eg = grid0.copy(order="F")
    nb, ne = O.marker_to_grid(0.0, 0.0, eg, x, y, np.array([2.0, 4.0]), 0.1, 0.1, O.zeros(4, 5), O.zeros(4, 5))
    assert np.array_equal(vg, eg) and st["n_empty_nodes"] == ne == 16
    # no markers at all: everything keeps its value
    vg, _, st = m2g_gpu(0.0, 0.0, grid0, np.zeros(0), np.zeros(0), np.zeros(0), 0.1, 0.1)
    assert np.array_equal(vg, grid0) and st["n_empty_nodes"] == 20


def test_m2g_errors(gpu):
    vg = sfhip.fzeros(4, 4)
    x = sfhip.from_numpy(np.array([0.1, np.nan]))
    y = sfhip.from_numpy(np.array([0.1, 0.1]))
    v = sfhip.from_numpy(np.array([1.0, 1.0]))
    with pytest.raises(sfhip.SFError) as e:
        sfhip.bilinearMarkerToGrid(0.0, 0.0, vg, x, y, v, 0.1, 0.1, print_info=False)
    assert e.value.code == -3
    with pytest.raises(AssertionError):
        sfhip.bilinearMarkerToGrid(0.0, 0.0, vg, x, y[:1], v, 0.1, 0.1, print_info=False)


def test_m2g_run_twice_identical(gpu):
    nx = ny = 200
    dx, dy = 1.0 / (nx - 1), 1.0 / (ny - 1)
    x, y = sfhip.init_markers(nx, ny, 4, dx, dy, 0.25, seed=99)
    rho, mu = sfhip.plume_properties(x, y, 0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, 0.001, 0.1, 1.0)
    a, b = sfhip.fzeros(nx + 1, ny), sfhip.fzeros(nx + 1, ny)
    sfhip.bilinearMarkerToGrid(-0.5 * dx, 0.0, a, x, y, rho, dx, dy, print_info=False)
    sfhip.bilinearMarkerToGrid(-0.5 * dx, 0.0, b, x, y, rho, dx, dy, print_info=False)
    assert torch.equal(a, b)


# ------------------------------------------------------------------------------ RK4

def rk4_gpu(x, y, Vx, Vy, dx, dy, dt, lx, ly):
    xg, yg = sfhip.from_numpy(x), sfhip.from_numpy(y)
    st = {}
    sfhip.moveMarkersRK4(xg, yg, sfhip.from_numpy(Vx), sfhip.from_numpy(Vy), -dx, -0.5 * dy, -0.5 * dx, -dy, dt, lx,
                         ly, dx, dy, stats=st, print_info=False)
    torch.cuda.synchronize()
    return sfhip.to_numpy(xg), sfhip.to_numpy(yg), st


@pytest.mark.parametrize("nx,ny,seed", [(5, 4, 0), (33, 47, 1), (128, 128, 2)])
def test_rk4_bitwise_vs_oracle(gpu, nx, ny, seed):
    rng = np.random.default_rng(seed)
    dx, dy = 1.0 / (nx - 1), 1.3 / (ny - 1)
    lx, ly = (nx - 1) * dx, (ny - 1) * dy
    Vx = np.asfortranarray(rng.random((nx + 2, ny + 1)) - 0.5)
    Vy = np.asfortranarray(rng.random((nx + 1, ny + 2)) - 0.5)
    n = 5000
    x = rng.random(n) * lx
    y = rng.random(n) * ly
    x[:4] = [0.0, lx, 0.0, lx]  # corners: the failsafe clamps are exercised
    y[:4] = [0.0, 0.0, ly, ly]
    dt = 0.5 * min(dx / np.max(np.abs(Vx)), dy / np.max(np.abs(Vy)))
    gx, gy, st = rk4_gpu(x, y, Vx, Vy, dx, dy, dt, lx, ly)
    ex, ey = x.copy(), y.copy()
    nc = O.move_markers_rk4(ex, ey, Vx, Vy, -dx, -0.5 * dy, -0.5 * dx, -dy, dt, lx, ly, dx, dy)
    assert np.array_equal(gx, ex) and np.array_equal(gy, ey)
    assert st["n_clamped"] == nc and nc > 0


def test_rk4_constant_velocity(gpu, golden):
    c = golden["move_markers_constant"]
    nx, ny, dx, dy = c["nx"], c["ny"], c["dx"], c["dy"]
    lx, ly = (nx - 1) * dx, (ny - 1) * dy
    rng = np.random.default_rng(4)
    n = c["markers_per_cell"] * (nx - 1) * (ny - 1)
    x = rng.random(n) * (lx - dx) + 0.5 * dx
    y = rng.random(n) * (ly - dy) + 0.5 * dy
    Vx = np.full((nx + 2, ny + 1), c["vx"], order="F")
    Vy = np.full((nx + 1, ny + 2), c["vy"], order="F")
    dt = 0.5 * min(dx / 1.0, dy / 2.0)
    gx, gy, _ = rk4_gpu(x, y, Vx, Vy, dx, dy, dt, lx, ly)
    np.testing.assert_allclose(gx, x + dt, rtol=1.5e-8)
    np.testing.assert_allclose(gy, y - 2.0 * dt, rtol=1.5e-8)


def test_rk4_errors(gpu):
    x = sfhip.from_numpy(np.array([0.5, np.inf]))
    y = sfhip.from_numpy(np.array([0.5, 0.5]))
    Vx, Vy = sfhip.fzeros(6, 5), sfhip.fzeros(5, 6)
    with pytest.raises(sfhip.SFError) as e:
        sfhip.moveMarkersRK4(x, y, Vx, Vy, -0.25, -0.125, -0.125, -0.25, 0.1, 1.0, 1.0, 0.25, 0.25, print_info=False)
    assert e.value.code == -3
    with pytest.raises(AssertionError):
        sfhip.moveMarkersRK4(x, y[:1], Vx, Vy, -0.25, -0.125, -0.125, -0.25, 0.1, 1.0, 1.0, 0.25, 0.25)


# ------------------------------------------------------------------------------ set-up

@pytest.mark.parametrize("nx,ny,md,jit,seed", [(9, 7, 4, 0.25, 42), (64, 33, 4, 0.25, 1001), (20, 26, 5, 0.0, 3)])
def test_init_markers_and_properties_bitwise(gpu, nx, ny, md, jit, seed):
    dx, dy = 1.0 / (nx - 1), 1.0 / (ny - 1)
    xg, yg = sfhip.init_markers(nx, ny, md, dx, dy, jit, seed=seed)
    ex, ey = O.seeded_markers(nx, ny, md, dx, dy, jit, seed)
    assert np.array_equal(sfhip.to_numpy(xg), ex) and np.array_equal(sfhip.to_numpy(yg), ey)
    rg, mg = sfhip.plume_properties(xg, yg, 0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, 0.001, 0.1, 1.0)
    er, em = O.plume_properties(ex, ey, 0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, 0.001, 0.1, 1.0)
    assert np.array_equal(sfhip.to_numpy(rg), er) and np.array_equal(sfhip.to_numpy(mg), em)


def test_rt_properties(gpu):
    nx = ny = 33
    dx = dy = 1.0 / 32
    xg, yg = sfhip.init_markers(nx, ny, 4, dx, dy, 0.25, seed=5)
    rg, mg = sfhip.rt_properties(xg, yg, 0.5, 0.02, 3.0, 1.0, 1.1, 1e3, 1.0, 1.0)
    x, y = sfhip.to_numpy(xg), sfhip.to_numpy(yg)
    yi = 0.5 + 0.02 * np.sin(2 * np.pi * 3.0 * x / 1.0)
    far = np.abs(y - yi) > 1e-12
    up = y < yi
    assert np.array_equal(sfhip.to_numpy(rg)[far], np.where(up, 1.1, 1.0)[far])
    assert np.array_equal(sfhip.to_numpy(mg)[far], np.where(up, 1e3, 1.0)[far])


@pytest.mark.parametrize("origin", ["rho", "mu_b", "mu_p"])
@pytest.mark.parametrize("nx,ny,md,seed", [(9, 7, 4, 1), (130, 67, 3, 3)])
def test_cell_keys_bitwise_vs_oracle(gpu, origin, nx, ny, md, seed):
    """Per-marker integer cell keys and per-cell counts equal the oracle's, including markers
    exactly on grid lines and outside the grid."""
    dx, dy, x, y, _, _ = cases.plume_markers(nx, ny, md, seed)
    o = O.grid_origins(dx, dy)[origin]
    k = np.arange(5, dtype=np.float64)
    x = np.concatenate([x, o[0] + k * dx, [o[0] - 1.5 * dx, 0.5, 1e3], o[0] + 0.5 * dx + 0 * k])
    y = np.concatenate([y, o[1] + 0.5 * dy + 0 * k, [0.5, o[1] - 2.5 * dy, 0.5], o[1] + k * dy])
    gnx, gny = {"rho": (nx + 1, ny), "mu_b": (nx, ny), "mu_p": (nx - 1, ny - 1)}[origin]
    keys, counts, nb = sfhip.marker_cell_keys(o[0], o[1], gnx, gny, sfhip.from_numpy(x), sfhip.from_numpy(y), dx, dy)
    ek, ec, enb = O.cell_keys(o[0], o[1], gnx, gny, x, y, dx, dy)
    assert np.array_equal(keys.cpu().numpy(), ek)
    assert np.array_equal(counts.cpu().numpy(), ec)
    assert nb == enb >= 2


def test_cell_keys_empty(gpu):
    e = torch.empty(0, dtype=torch.float64, device="cuda")
    keys, counts, nb = sfhip.marker_cell_keys(0.0, 0.0, 4, 3, e, e, 0.25, 0.5)
    assert keys.numel() == 0 and nb == 0 and int(counts.sum()) == 0



def test_markers_to_grids3_matches_three_calls(gpu):
    nx, ny = 40, 29
    dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=12)
    xg, yg, rg, mg = (sfhip.from_numpy(a) for a in (x, y, rho, mu))
    st = sfhip.FlowState(nx, ny)
    counts = sfhip.markers_to_grids3(st, xg, yg, rg, mg, dx, dy)
    S = O.solver_arrays(nx, ny)
    vws, ws = O.zeros(nx + 1, ny), O.zeros(nx + 1, ny)
    o = O.grid_origins(dx, dy)
    exp = [O.marker_to_grid(*o["rho"], S["rho_vy"], x, y, rho, dx, dy, vws, ws),
           O.marker_to_grid(*o["mu_b"], S["mu_b"], x, y, mu, dx, dy, vws, ws),
           O.marker_to_grid(*o["mu_p"], S["mu_p"], x, y, mu, dx, dy, vws, ws)]
    assert counts == [tuple(e) for e in exp]
    for name, k in (("rho_vy", "rho_vy"), ("mu_b", "mu_b"), ("mu_p", "mu_p")):
        assert np.array_equal(sfhip.to_numpy(getattr(st, name)), S[k]), name


# ------------------------------------------------------------------ marker set-up (common.jl:40-110)

@pytest.mark.parametrize("nx,ny,md", [(20, 26, 5), (64, 48, 4), (33, 17, 3), (7, 9, 1), (3, 2, 2)])
def test_lattice_bitwise_vs_oracle(gpu, nx, ny, md):
    """initializeMarkersCPU's LinRange lattice on the device, bit for bit the oracle's (which the
    reference's end-to-end golden vectors pin)."""
    lx, ly = 10.0, 7.5
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    x, y = sfhip.lattice_markers(nx, ny, md, (nx - 1) * dx, (ny - 1) * dy, dx, dy)
    xo, yo = O.lattice_markers(nx, ny, md, (nx - 1) * dx, (ny - 1) * dy, dx, dy)
    assert x.numel() == xo.size == md * md * (nx - 1) * (ny - 1)
    assert np.array_equal(sfhip.to_numpy(x), xo) and np.array_equal(sfhip.to_numpy(y), yo)


def test_lattice_random_positions(gpu):
    """RAND_MARKER_POS: each marker moves by less than half a marker spacing from its lattice
    position, stays in the domain, and the seed decides the positions."""
    nx, ny, md = 40, 30, 4
    dx, dy = 1.0 / (nx - 1), 0.75 / (ny - 1)
    lx, ly = (nx - 1) * dx, (ny - 1) * dy
    x0, y0 = sfhip.lattice_markers(nx, ny, md, lx, ly, dx, dy)
    x1, y1 = sfhip.lattice_markers(nx, ny, md, lx, ly, dx, dy, rand_pos=True, seed=9)
    x2, y2 = sfhip.lattice_markers(nx, ny, md, lx, ly, dx, dy, rand_pos=True, seed=9)
    x3, _ = sfhip.lattice_markers(nx, ny, md, lx, ly, dx, dy, rand_pos=True, seed=10)
    assert torch.equal(x1, x2) and torch.equal(y1, y2) and not torch.equal(x1, x3)
    assert ((x1 - x0).abs() <= 0.5 * dx / md).all() and ((y1 - y0).abs() <= 0.5 * dy / md).all()
    assert (x1 >= 0).all() and (x1 <= lx).all() and (y1 >= 0).all() and (y1 <= ly).all()


def test_materials_descriptor(gpu):
    """The material descriptor gives the plume and Rayleigh-Taylor helpers' values bit for bit, and
    rectangles and half planes as numpy evaluates them; the first containing shape wins."""
    nx = ny = 50
    dx = dy = 1.0 / (nx - 1)
    x, y = sfhip.init_markers(nx, ny, 4, dx, dy, 0.25, seed=4)
    rho_p, mu_p = sfhip.plume_properties(x, y, 0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, 0.01, 0.1, 1.0)
    mats = sfhip.Materials([("y_below", 0.1, 0.001, 0.01), ("circle", 0.5, 0.8, 0.1, 0.9, 0.1)], rho_bg=1.0, mu_bg=1.0)
    rho, mu = sfhip.marker_properties(x, y, mats)
    assert torch.equal(rho, rho_p) and torch.equal(mu, mu_p)
    rho_r, mu_r = sfhip.rt_properties(x, y, 0.5, 0.02, 3, 1.0, 1.1, 10.0, 1.0, 1.0)
    rho, mu = sfhip.marker_properties(x, y, sfhip.Materials([("sine_below", 0.5, 0.02, 3, 1.0, 1.1, 10.0)], 1.0, 1.0))
    assert torch.equal(rho, rho_r) and torch.equal(mu, mu_r)
    mats = sfhip.Materials([("rect", 0.2, 0.4, 0.3, 0.6, 2.0, 20.0), ("y_above", 0.5, 3.0, 30.0)], rho_bg=1.0,
                           mu_bg=10.0)
    rho, mu = sfhip.marker_properties(x, y, mats, x0=0.25, y0=-0.1)
    xh, yh = sfhip.to_numpy(x) + 0.25, sfhip.to_numpy(y) + (-0.1)
    inr = (xh >= 0.2) & (xh < 0.4) & (yh >= 0.3) & (yh < 0.6)
    above = yh >= 0.5
    exp_rho = np.where(inr, 2.0, np.where(above, 3.0, 1.0))
    exp_mu = np.where(inr, 20.0, np.where(above, 30.0, 10.0))
    assert inr.any() and (above & ~inr).any()
    assert np.array_equal(sfhip.to_numpy(rho), exp_rho) and np.array_equal(sfhip.to_numpy(mu), exp_mu)
    # an unknown shape kind is rejected by the C ABI before anything runs
    from sfhip import _lib
    bad = (_lib.Shape * 1)()
    bad[0].kind = 99
    L = sfhip.load()
    ctx = sfhip.context()
    rc = L.sf_marker_properties(ctx.handle, x.data_ptr(), y.data_ptr(), x.numel(), 0.0, 0.0, bad, 1, 1.0, 1.0,
                                rho.data_ptr(), mu.data_ptr())
    assert rc == _lib.SF_ERR_ARG and b"unknown kind" in L.sf_last_error()
