"""Pins the CPU oracle to the reference's golden vectors, known answers and to what the
mathematics fixes.  CPU only."""
import numpy as np
import pytest

from oracle import sf_oracle as O

import cases


def isapprox_vec(a, b, rtol):
    """Julia's isapprox on vectors: norm(a-b) <= rtol * max(norm(a), norm(b))."""
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def linear(a, inds):
    return a.ravel(order="F")[np.asarray(inds) - 1]


# ---------------------------------------------------------------------------- solver

@pytest.fixture(scope="module")
def solved(golden):
    g = golden["stokes_solver"]
    S = cases.stokes_solver_inputs(g)
    out = O.solve_stokes(S, g["g_y"], g["dx"], g["dy"], g["nx"], g["ny"], g["dt"], g["maxdisp"], True, g["eps"])
    return g, S, out


def test_solver_golden_vectors(solved):
    g, S, out = solved
    assert out["converged"]
    assert isapprox_vec(linear(S["Vx_pad"], g["test_inds"]), np.array(g["vx_ref"]), g["rtol"])
    assert isapprox_vec(linear(S["Vy_pad"], g["test_inds"]), np.array(g["vy_ref"]), g["rtol"])


def test_solver_bc_identities_and_dt(solved):
    """test/test_StokesSolver.jl:116,128-140 (exact equalities)."""
    g, S, out = solved
    Vx, Vy = S["Vx_pad"], S["Vy_pad"]
    assert np.all(Vx[0, :] == -Vx[2, :]) and np.all(Vx[1, :] == 0.0)
    assert np.all(Vx[-1, :] == -Vx[-3, :]) and np.all(Vx[-2, :] == 0.0)
    assert np.all(Vx[:, 0] == Vx[:, 1]) and np.all(Vx[:, -1] == Vx[:, -2])
    assert np.all(Vy[:, 0] == -Vy[:, 2]) and np.all(Vy[:, 1] == 0.0)
    assert np.all(Vy[:, -1] == -Vy[:, -3]) and np.all(Vy[:, -2] == 0.0)
    assert np.all(Vy[0, :] == Vy[1, :]) and np.all(Vy[-1, :] == Vy[-2, :])
    dt_ref = g["maxdisp"] * min(g["dx"] / np.max(np.abs(Vx[1:-1, :])), g["dy"] / np.max(np.abs(Vy[:, 1:-1])))
    assert out["dt_new"] == pytest.approx(dt_ref, rel=1e-15)


def test_solver_golden_rejects_planted_mistakes(golden):
    """The golden vectors discriminate: dropping the free-surface term or flipping gravity fails."""
    g = golden["stokes_solver"]
    for kwargs in ({"use_fss": False, "g_y": g["g_y"]}, {"use_fss": True, "g_y": -g["g_y"]}):
        S = cases.stokes_solver_inputs(g)
        O.solve_stokes(S, kwargs["g_y"], g["dx"], g["dy"], g["nx"], g["ny"], 0.0, 0.5, kwargs["use_fss"], 1e-8)
        assert not isapprox_vec(linear(S["Vx_pad"], g["test_inds"]), np.array(g["vx_ref"]), g["rtol"])


def test_hydrostatic_invariant():
    """Uniform density: V stays ~0 and dP/dy = rho g (momentum balance with tau = 0)."""
    nx, ny = 17, 21
    dx, dy = 0.1, 0.05
    S = O.solver_arrays(nx, ny)
    S["rho_vy"][:] = 2.5
    S["mu_b"][:] = 1.0
    S["mu_p"][:] = 1.0
    out = O.solve_stokes(S, 3.0, dx, dy, nx, ny, 0.0, 0.5, True, 1e-11)
    assert out["converged"]
    assert np.max(np.abs(S["Vx_s"])) < 1e-9 and np.max(np.abs(S["Vy_s"])) < 1e-9
    dPdy = (S["P"][:, 1:] - S["P"][:, :-1]) / dy
    np.testing.assert_allclose(dPdy, 2.5 * 3.0, rtol=1e-8)


def test_mirror_symmetry():
    """A plume centred at x = L/2 on an unjittered lattice: Vx antisymmetric, Vy symmetric."""
    nx = ny = 21
    dx, dy, lx, ly = cases.unit_square(nx)
    x, y = O.lattice_markers(nx, ny, 4, lx, ly, dx, dy)
    rho, mu = O.plume_properties(x, y, 0.1, 0.5, 0.7, 0.15, 0.001, 0.9, 1.0, 0.01, 0.1, 1.0)
    S = O.solver_arrays(nx, ny)
    vws, ws = O.zeros(nx + 1, ny), O.zeros(nx + 1, ny)
    o = O.grid_origins(dx, dy)
    O.marker_to_grid(*o["rho"], S["rho_vy"], x, y, rho, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_b"], S["mu_b"], x, y, mu, dx, dy, vws, ws)
    O.marker_to_grid(*o["mu_p"], S["mu_p"], x, y, mu, dx, dy, vws, ws)
    O.solve_stokes(S, 1.0, dx, dy, nx, ny, 0.0, 0.5, True, 1e-10)
    Vx, Vy = S["Vx_s"], S["Vy_s"]
    scale = np.max(np.abs(Vy))
    assert scale > 1e-4
    np.testing.assert_allclose(Vx, -Vx[::-1, :], atol=1e-7 * scale)
    np.testing.assert_allclose(Vy, Vy[::-1, :], atol=1e-7 * scale)


def test_mean_abs_exact_on_integers():
    a = np.arange(-5000, 7001, dtype=np.float64)
    assert O.mean_abs(a) == np.sum(np.abs(a)) / a.size


# ---------------------------------------------------------------------------- marker -> grid

def test_marker_to_grid_known_answer(golden):
    k = golden["marker_to_grid_kat"]
    vg = O.zeros(*k["grid_shape"])
    vws = np.full(k["grid_shape"], k["helper_fill"], order="F")
    ws = np.full(k["grid_shape"], k["helper_fill"], order="F")
    nb, ne = O.marker_to_grid(k["x_grid_min"], k["y_grid_min"], vg, np.array(k["x_m"]), np.array(k["y_m"]),
                              np.array(k["val_m"]), k["dx"], k["dy"], vws, ws)
    ref = np.array(k["reference_columns"]).T
    assert (nb, ne) == (0, 0)
    np.testing.assert_allclose(vg, ref, rtol=1.5e-8)


def _brute_force_m2g(xmin, ymin, gnx, gny, x, y, v, dx, dy):
    """Direct per-marker, per-node restatement of MarkerToGrid.jl:110-172 in plain Python."""
    wt = np.zeros((gnx, gny))
    vw = np.zeros((gnx, gny))
    for m in range(x.size):
        ix = int(np.floor((x[m] - xmin) / dx)) + 1
        iy = int(np.floor((y[m] - ymin) / dy)) + 1
        fx = (x[m] - (xmin + (ix - 1) * dx)) / dx
        fy = (y[m] - (ymin + (iy - 1) * dy)) / dy
        if not (0 <= ix <= gnx and 0 <= iy <= gny):
            continue
        for (a, b, w) in ((ix, iy, (1 - fx) * (1 - fy)), (ix + 1, iy, fx * (1 - fy)),
                          (ix, iy + 1, (1 - fx) * fy), (ix + 1, iy + 1, fx * fy)):
            if 1 <= a <= gnx and 1 <= b <= gny:
                wt[a - 1, b - 1] += w
                vw[a - 1, b - 1] += w * v[m]
    return wt, vw


@pytest.mark.parametrize("origin", ["rho", "mu_b", "mu_p"])
def test_marker_to_grid_brute_force(origin):
    rng = np.random.default_rng(3)
    nx, ny = 7, 6
    dx, dy = 0.3, 0.4
    o = O.grid_origins(dx, dy)[origin]
    shape = {"rho": (nx + 1, ny), "mu_b": (nx, ny), "mu_p": (nx - 1, ny - 1)}[origin]
    n = 400
    x = rng.random(n) * (nx - 1) * dx * 1.1 - 0.05
    y = rng.random(n) * (ny - 1) * dy * 1.1 - 0.05
    v = rng.random(n) + 1.0
    vg = np.full(shape, -7.0, order="F")
    vws, ws = O.zeros(nx + 1, ny), O.zeros(nx + 1, ny)
    nb, ne = O.marker_to_grid(o[0], o[1], vg, x, y, v, dx, dy, vws, ws)
    wt, vw = _brute_force_m2g(o[0], o[1], shape[0], shape[1], x, y, v, dx, dy)
    assert np.array_equal(ws[: shape[0], : shape[1]], wt)
    assert np.array_equal(vws[: shape[0], : shape[1]], vw)
    exp = np.where(wt != 0, vw / np.where(wt != 0, wt, 1), -7.0)
    assert np.array_equal(vg, exp)
    assert ne == int(np.sum(wt == 0))


def test_marker_to_grid_constant_and_linear():
    nx, ny = 12, 10
    dx, dy = 0.25, 0.2
    lx, ly = (nx - 1) * dx, (ny - 1) * dy
    x, y = O.lattice_markers(nx, ny, 4, lx, ly, dx, dy)
    vg = O.zeros(nx, ny)
    vws, ws = O.zeros(nx + 1, ny), O.zeros(nx + 1, ny)
    O.marker_to_grid(0.0, 0.0, vg, x, y, np.full(x.size, 2.75), dx, dy, vws, ws)
    # sum(w * c) / sum(w) over up to 36 weights: a few ulp of rounding
    np.testing.assert_allclose(vg, 2.75, rtol=1e-15)
    # a linear field is reproduced at interior nodes of a symmetric lattice
    O.marker_to_grid(0.0, 0.0, vg, x, y, 1.5 * x - 0.5 * y + 3.0, dx, dy, vws, ws)
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
    np.testing.assert_allclose(vg[1:-1, 1:-1], (1.5 * X - 0.5 * Y + 3.0)[1:-1, 1:-1], rtol=1e-13)


def test_marker_to_grid_counts_bad_and_empty():
    vg = np.full((4, 4), 9.0, order="F")
    vws, ws = O.zeros(4, 4), O.zeros(4, 4)
    x = np.array([0.05, -5.0, 0.15])
    y = np.array([0.05, 0.05, 9.0])
    nb, ne = O.marker_to_grid(0.0, 0.0, vg, x, y, np.array([1.0, 2.0, 3.0]), 0.1, 0.1, vws, ws)
    assert nb == 2 and ne == 12
    assert vg[0, 0] == 1.0 and vg[3, 3] == 9.0
    with pytest.raises(ValueError):
        O.marker_to_grid(0.0, 0.0, vg, np.array([np.nan]), np.array([0.0]), np.array([1.0]), 0.1, 0.1, vws, ws)


# ---------------------------------------------------------------------------- RK4

def test_rk4_constant_velocity(golden):
    """test/test_MoveMarkers.jl:17-47 with the oracle's seeded markers."""
    c = golden["move_markers_constant"]
    nx, ny, dx, dy = c["nx"], c["ny"], c["dx"], c["dy"]
    lx, ly = (nx - 1) * dx, (ny - 1) * dy
    rng = np.random.default_rng(11)
    n = c["markers_per_cell"] * (nx - 1) * (ny - 1)
    x = rng.random(n) * (lx - dx) + 0.5 * dx
    y = rng.random(n) * (ly - dy) + 0.5 * dy
    x0, y0 = x.copy(), y.copy()
    Vx = np.full((nx + 2, ny + 1), c["vx"], order="F")
    Vy = np.full((nx + 1, ny + 2), c["vy"], order="F")
    dt = 0.5 * min(dx / 1.0, dy / 2.0)
    O.move_markers_rk4(x, y, Vx, Vy, -dx, -0.5 * dy, -0.5 * dx, -dy, dt, lx, ly, dx, dy)
    np.testing.assert_allclose(x, x0 + 1.0 * dt, rtol=1.5e-8)
    np.testing.assert_allclose(y, y0 - 2.0 * dt, rtol=1.5e-8)


def test_interpolate_v_quadrants(golden):
    """test/test_MoveMarkers.jl:54-152: indices and the closed-form value in every quadrant."""
    q = golden["move_markers_quadrants"]
    dx, dy = q["dx"], q["dy"]
    rng = np.random.default_rng(5)
    Vx = np.asfortranarray(rng.random((4, 3)))
    Vy = np.asfortranarray(rng.random((3, 4)))
    A = lambda M, i, j: M[i - 1, j - 1]  # noqa: E731  1-based like the reference test
    for k, (ox, oy) in enumerate(q["quadrant_offsets"]):
        x = (0.99 * rng.random() + 0.005 + ox) * 0.5 * dx
        y = (0.99 * rng.random() + 0.005 + oy) * 0.5 * dy
        ix, iy, fx, fy = O.topleft_index(-dx, -0.5 * dy, x, y, dx, dy)
        assert [ix, iy] == q["vx_index"][k]
        s = +1 if fx >= 0.5 else -1
        b = ix + 2 if s > 0 else ix - 1
        lo, mid, hi = (ix, ix + 1, b) if s > 0 else (b, ix, ix + 1)
        vxref = ((1 - fx) * (1 - fy) * A(Vx, ix, iy) + fx * (1 - fy) * A(Vx, ix + 1, iy) + fx * fy * A(Vx, ix + 1, iy + 1)
                 + (1 - fx) * fy * A(Vx, ix, iy + 1) + 0.5 * (fx - 0.5) ** 2 * (
                     (1 - fy) * (A(Vx, lo, iy) - 2 * A(Vx, mid, iy) + A(Vx, hi, iy))
                     + fy * (A(Vx, lo, iy + 1) - 2 * A(Vx, mid, iy + 1) + A(Vx, hi, iy + 1))))
        ix, iy, fx, fy = O.topleft_index(-0.5 * dx, -dy, x, y, dx, dy)
        assert [ix, iy] == q["vy_index"][k]
        s = +1 if fy >= 0.5 else -1
        b = iy + 2 if s > 0 else iy - 1
        lo, mid, hi = (iy, iy + 1, b) if s > 0 else (b, iy, iy + 1)
        vyref = ((1 - fx) * (1 - fy) * A(Vy, ix, iy) + fx * (1 - fy) * A(Vy, ix + 1, iy) + fx * fy * A(Vy, ix + 1, iy + 1)
                 + (1 - fx) * fy * A(Vy, ix, iy + 1) + 0.5 * (fy - 0.5) ** 2 * (
                     (1 - fx) * (A(Vy, ix, lo) - 2 * A(Vy, ix, mid) + A(Vy, ix, hi))
                     + fx * (A(Vy, ix + 1, lo) - 2 * A(Vy, ix + 1, mid) + A(Vy, ix + 1, hi))))
        vx, vy, _ = O.interpolate_v(x, y, Vx, Vy, -dx, -0.5 * dy, -0.5 * dx, -dy, dx, dy)
        assert vx == pytest.approx(vxref, rel=1.5e-8) and vy == pytest.approx(vyref, rel=1.5e-8)


def test_interpolate_v_linear_field_exact():
    """The correction vanishes on a linear field, so interpolation is exact."""
    nx, ny, dx, dy = 9, 8, 0.2, 0.3
    xv = (np.arange(nx + 2) - 1) * dx
    yv = np.arange(ny + 1) * dy - 0.5 * dy
    Vx = np.asfortranarray(2.0 * xv[:, None] + 0.5 * yv[None, :] - 1.0)
    xw = np.arange(nx + 1) * dx - 0.5 * dx
    yw = (np.arange(ny + 2) - 1) * dy
    Vy = np.asfortranarray(-0.75 * xw[:, None] + 1.25 * yw[None, :] + 0.5)
    rng = np.random.default_rng(9)
    for _ in range(50):
        x = rng.random() * (nx - 1) * dx
        y = rng.random() * (ny - 1) * dy
        vx, vy, nc = O.interpolate_v(x, y, Vx, Vy, -dx, -0.5 * dy, -0.5 * dx, -dy, dx, dy)
        assert nc == 0
        assert vx == pytest.approx(2.0 * x + 0.5 * y - 1.0, abs=1e-13)
        assert vy == pytest.approx(-0.75 * x + 1.25 * y + 0.5, abs=1e-13)


# ---------------------------------------------------------------------------- set-up and pipeline

def test_lattice_matches_reference_lattice():
    """test/test_InitialMarkerCoords.jl:101-131 at one rank."""
    nx, ny, md, L = 11, 9, 5, 10.0
    dx, dy = L / (nx - 1), L / (ny - 1)
    x, y = O.lattice_markers(nx, ny, md, (nx - 1) * dx, (ny - 1) * dy, dx, dy)
    nmx, nmy = md * (nx - 1), md * (ny - 1)
    assert x.size == nmx * nmy
    xs = np.linspace(0.5 * dx / md, L - 0.5 * dx / md, nmx)
    ys = np.linspace(0.5 * dy / md, L - 0.5 * dy / md, nmy)
    np.testing.assert_allclose(x.reshape(nmx, nmy), np.repeat(xs[:, None], nmy, 1), rtol=1e-14)
    np.testing.assert_allclose(y.reshape(nmx, nmy), np.repeat(ys[None, :], nmx, 0), rtol=1e-14)


def test_seeded_markers_lattice_and_jitter():
    nx, ny, md = 9, 7, 4
    dx, dy = 1.0 / 8, 1.0 / 6
    x, y = O.seeded_markers(nx, ny, md, dx, dy, 0.25, 42)
    x0, y0 = O.seeded_markers(nx, ny, md, dx, dy, 0.0, 42)
    assert x.size == md * md * (nx - 1) * (ny - 1)
    assert np.all(np.abs(x - x0) <= 0.25 * dx + 1e-15) and np.all(np.abs(y - y0) <= 0.25 * dy + 1e-15)
    assert np.all((x >= 0) & (x <= 1.0)) and np.all((y >= 0) & (y <= 1.0))
    u = np.array([O.uniform(42, k) for k in range(20000)])
    assert abs(u.mean() - 0.5) < 0.01 and u.min() >= 0.0 and u.max() < 1.0


def test_flow_golden_vectors(golden):
    """test/test_StokesFlow2D.jl:86-88 (5 timesteps) at the reference test's own bound,
    isapprox(atol=1e-6) per element (:107-108)."""
    f = golden["stokes_flow_2d"]
    S = flow_golden_run(f)
    np.testing.assert_allclose(linear(S["Vx_pad"], f["inds"]), f["vx_ref"], rtol=0, atol=f["atol"])
    np.testing.assert_allclose(linear(S["Vy_pad"], f["inds"]), f["vy_ref"], rtol=0, atol=f["atol"])


def flow_golden_run(f, flip=None):
    """The golden set-up of test/test_StokesFlow2D.jl:53-98 through the oracle's timestep loop.
    flip: marker indices whose material is swapped between plume and matrix before the run."""
    nx, ny, L = f["nx"], f["ny"], f["lx"]
    dx, dy = L / (nx - 1), f["ly"] / (ny - 1)
    x, y = O.lattice_markers(nx, ny, f["marker_density"], (nx - 1) * dx, (ny - 1) * dy, dx, dy)
    rho, mu = O.plume_properties(x, y, f["air_height"], f["plume_x"], f["plume_y"], f["plume_r"], f["rho_air"],
                                 f["rho_plume"], f["rho_matrix"], f["mu_air"], f["mu_plume"], f["mu_matrix"])
    for k in flip or ():
        plume = rho[k] == f["rho_plume"]
        rho[k] = f["rho_matrix"] if plume else f["rho_plume"]
        mu[k] = f["mu_matrix"] if plume else f["mu_plume"]
    S, t_tot, outs = O.stokesflow2d(f["nt"], nx, ny, L, f["ly"], x, y, rho, mu, g_y=f["g_y"], eps=f["eps"])
    assert all(o["converged"] for o in outs)
    return S


def test_flow_golden_lattice_on_circle(golden):
    """Four markers of the md = 5 lattice lie exactly on the plume circle ((x-5)^2 + (y-5)^2 = 4
    has the lattice points (3,5), (7,5), (5,3), (5,7)).  Julia 1.8's LinRange leaves all four
    at distance^2 >= r^2, so the strict '<' of :67 makes them matrix; one plume marker more
    moves the golden Vx by ~5e-4, far outside atol 1e-6.  This pins the lattice to the ulp."""
    f = golden["stokes_flow_2d"]
    nx, ny, L = f["nx"], f["ny"], f["lx"]
    dx, dy = L / (nx - 1), f["ly"] / (ny - 1)
    x, y = O.lattice_markers(nx, ny, f["marker_density"], (nx - 1) * dx, (ny - 1) * dy, dx, dy)
    d2 = (x - f["plume_x"]) ** 2 + (y - f["plume_y"]) ** 2
    on = np.flatnonzero(np.abs(d2 - f["plume_r"] ** 2) < 1e-9)
    assert on.size == 4 and np.all(d2[on] >= f["plume_r"] ** 2)
    S = flow_golden_run(f, flip=on[:1])
    assert np.max(np.abs(linear(S["Vx_pad"], f["inds"]) - np.array(f["vx_ref"]))) > 100 * f["atol"]


def test_cell_keys_lattice_closed_form():
    """An unjittered md x md lattice puts md^2 markers in every cell of the mu_b grid (origin
    (0,0)) and md^2/2 in the half cells at the x ends of the rho grid (origin (-dx/2, 0));
    markers left of the first node by more than a cell are counted as erroneous."""
    n, md = 12, 4
    dx = dy = 1.0 / (n - 1)
    x, y = O.lattice_markers(n, n, md, (n - 1) * dx, (n - 1) * dy, dx, dy)
    o = O.grid_origins(dx, dy)
    keys, counts, nb = O.cell_keys(*o["mu_b"], n, n, x, y, dx, dy)
    c = counts.reshape(n + 1, n + 1)  # [iy, ix]
    assert nb == 0 and np.all(c[1:n, 1:n] == md * md) and c.sum() == x.size
    assert np.array_equal(np.bincount(keys, minlength=c.size), counts)
    keys, counts, nb = O.cell_keys(*o["rho"], n + 1, n, x, y, dx, dy)
    c = counts.reshape(n + 1, n + 2)
    assert nb == 0 and np.all(c[1:n, 2:n] == md * md)
    assert np.all(c[1:n, 1] == md * md // 2) and np.all(c[1:n, n] == md * md // 2)
    xb = np.concatenate([x, [-1.5 * dx, -0.5 * dx]])
    yb = np.concatenate([y, [0.5, 0.5]])
    keys, counts, nb = O.cell_keys(*o["mu_b"], n, n, xb, yb, dx, dy)
    assert nb == 1 and keys[-2] == -1 and keys[-1] % (n + 1) == 0  # ix = 0 is inside the accepted range
    g, vws, ws = O.zeros(n, n), O.zeros(n + 1, n), O.zeros(n + 1, n)
    assert O.marker_to_grid(*o["mu_b"], g, xb, yb, np.ones_like(xb), dx, dy, vws, ws)[0] == nb
