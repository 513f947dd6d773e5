"""Snapshot files from device arrays (saveMarkers, saveGrid; StokesFlow2D_multixpu.jl:230-299) read
back with scipy's MAT reader: every value equals numpy's FP64 -> FP32 rounding of the same data."""
import numpy as np
import pytest
import scipy.io
import torch

import sfhip

import cases

pytestmark = pytest.mark.gpu


def test_save_markers_chunks_and_shift(gpu, tmp_path):
    """More markers than one 2^23-float staging chunk, a coordinate shift, signed zeros and
    values that round differently in FP32."""
    nm = (1 << 23) + 12345
    g = torch.Generator(device="cuda").manual_seed(9)
    x = torch.rand(nm, generator=g, device="cuda", dtype=torch.float64) * 3.7
    y = torch.rand(nm, generator=g, device="cuda", dtype=torch.float64) - 0.5
    rho = torch.rand(nm, generator=g, device="cuda", dtype=torch.float64) + 1.0
    x[0], y[0] = -0.0, 1.0 + 2.0 ** -30
    path = tmp_path / "markers_0003_0000.mat"
    sfhip.saveMarkers(path, x, y, rho, x_shift=0.25, y_shift=0.0)
    m = scipy.io.loadmat(str(path))
    xh, yh, rh = (sfhip.to_numpy(a) for a in (x, y, rho))
    for name, ref in (("x_m", (xh + 0.25).astype(np.float32)), ("y_m", (yh + 0.0).astype(np.float32)),
                      ("rho_m", rh.astype(np.float32))):
        got = m[name]
        assert got.dtype == np.float32 and got.shape == (nm, 1), name
        assert np.array_equal(got.ravel().view(np.uint32), ref.view(np.uint32)), name


def test_save_grid_coordinates_and_fields(gpu, tmp_path):
    nx, ny = 23, 17
    dx, dy = 10.0 / (nx - 1), 10.0 / (ny - 1)
    S = cases.random_state(nx, ny, seed=31)
    mu_b = sfhip.from_numpy(S["mu_b"])
    P = sfhip.from_numpy(S["P"])
    Vx = sfhip.from_numpy(np.random.default_rng(1).random((nx + 2, ny + 1)) - 0.5)
    Vy = sfhip.from_numpy(np.random.default_rng(2).random((nx + 1, ny + 2)) - 0.5)
    path = tmp_path / "grid_0001_0000.mat"
    sfhip.saveGrid(path, nx, ny, dx, dy, mu_b, P, Vx, Vy)
    m = scipy.io.loadmat(str(path))
    # StokesFlow2D_multixpu.jl:117-124 (1-based i)
    coords = {"x": [(i - 1) * dx for i in range(1, nx + 1)], "y": [(i - 1) * dy for i in range(1, ny + 1)],
              "x_p": [(i - 1) * dx + 0.5 * dx for i in range(1, nx)],
              "y_p": [(i - 1) * dy + 0.5 * dy for i in range(1, ny)],
              "x_vx": [(i - 2) * dx for i in range(1, nx + 3)],
              "y_vx": [(i - 1) * dy - 0.5 * dy for i in range(1, ny + 2)],
              "x_vy": [(i - 1) * dx - 0.5 * dx for i in range(1, nx + 2)],
              "y_vy": [(i - 2) * dy for i in range(1, ny + 3)]}
    for name, ref in coords.items():
        ref = np.asarray(ref, dtype=np.float64).astype(np.float32)
        assert m[name].dtype == np.float32 and m[name].shape == (ref.size, 1), name
        assert np.array_equal(m[name].ravel(), ref), name
    for name, t in (("mu_b", mu_b), ("P", P), ("Vx", Vx), ("Vy", Vy)):
        ref = sfhip.to_numpy(t).astype(np.float32)
        assert m[name].dtype == np.float32 and m[name].shape == ref.shape, name
        assert np.array_equal(m[name], ref), name


def test_stokesflow2d_writes_reference_files(gpu, tmp_path):
    """save_to_file writes stats.mat and one markers/grid pair per step plus step 0, named as the
    reference names them, and the last grid file holds the returned velocities."""
    nx = ny = 21
    dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=55)
    xg, yg, rg, mg = (sfhip.from_numpy(a) for a in (x, y, rho, mu))
    out = tmp_path / "viz_out"
    Vx, Vy, t_tot = sfhip.StokesFlow2D(2, nx, ny, 1.0, 1.0, xg, yg, rg, mg, g_y=1.0, stokes_eps=1e-6,
                                       save_to_file=True, out_dir=str(out))
    names = sorted(p.name for p in out.iterdir())
    assert names == ["grid_0000_0000.mat", "grid_0001_0000.mat", "grid_0002_0000.mat", "markers_0000_0000.mat",
                     "markers_0001_0000.mat", "markers_0002_0000.mat", "stats.mat"]
    st = scipy.io.loadmat(str(out / "stats.mat"))
    assert st["dims"].ravel().tolist() == [1, 1, 1] and int(st["nt"][0, 0]) == 2
    g = scipy.io.loadmat(str(out / "grid_0002_0000.mat"))
    assert np.array_equal(g["Vx"], sfhip.to_numpy(Vx).astype(np.float32))
    assert np.array_equal(g["Vy"], sfhip.to_numpy(Vy).astype(np.float32))
    mk = scipy.io.loadmat(str(out / "markers_0002_0000.mat"))
    assert np.array_equal(mk["x_m"].ravel(), sfhip.to_numpy(xg).astype(np.float32))
    assert t_tot > 0.0
