"""Graph-captured ensemble timestep (sf_ensemble_*) against per-member sf_step calls.

Every member's step is the same marker -> grid passes, the same PT iterations (batched
kernel, bit-identical to the single solver) and the same RK4, so grids, markers and dt must
be bit-identical to stepping the members one by one.  The graph is replayed several times,
with eager work and synchronisations in between, and checked after every replay."""
import numpy as np
import pytest
import torch

import sfhip
from oracle import sf_oracle as O

import cases

pytestmark = pytest.mark.gpu

SIZES = [(24, 24), (33, 47), (40, 29)]
FIELDS = ("Vx", "Vy", "P", "rho_vy", "mu_b", "mu_p", "dVxdtau", "dVydtau")


def member(nx, ny, seed):
    dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=seed)
    return {"state": sfhip.FlowState(nx, ny), "x_m": sfhip.from_numpy(x), "y_m": sfhip.from_numpy(y),
            "rho_m": sfhip.from_numpy(rho), "mu_m": sfhip.from_numpy(mu), "dx": dx, "dy": dy, "g_y": 1.0,
            "eps": 1e-8}


def same(a, b):
    for f in FIELDS:
        assert torch.equal(getattr(a["state"], f), getattr(b["state"], f)), f
    assert torch.equal(a["x_m"], b["x_m"]) and torch.equal(a["y_m"], b["y_m"])


@pytest.mark.parametrize("mode", ["persistent", "wide", "grid"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_ensemble_matches_single_steps(gpu, use_graph, mode):
    ens_m = [member(nx, ny, 20 + k) for k, (nx, ny) in enumerate(SIZES)]
    ref_m = [member(nx, ny, 20 + k) for k, (nx, ny) in enumerate(SIZES)]
    ens = sfhip.Ensemble(ens_m, use_graph=use_graph, mode=mode)
    dts = [0.0] * len(SIZES)
    for t in range(3):
        outs = ens.step(1)
        torch.cuda.synchronize()
        for k, m in enumerate(ref_m):  # eager work between replays
            dts[k], st = sfhip.step(m["state"], m["x_m"], m["y_m"], m["rho_m"], m["mu_m"], m["dx"], m["dy"], dts[k],
                                    g_y=1.0, eps=1e-8)
            assert outs[k]["niter"] == st["niter"], (t, k)
            assert outs[k]["dt_new"] == dts[k], (t, k)
        torch.cuda.synchronize()
        for a, b in zip(ens_m, ref_m):
            same(a, b)
    ens.close()


@pytest.mark.parametrize("mode", ["persistent", "wide", "grid"])
def test_ensemble_multi_step_replay(gpu, mode):
    """n_steps replays in one call equal n single-step calls."""
    a = [member(nx, ny, 40 + k) for k, (nx, ny) in enumerate(SIZES)]
    b = [member(nx, ny, 40 + k) for k, (nx, ny) in enumerate(SIZES)]
    ea, eb = sfhip.Ensemble(a, mode=mode), sfhip.Ensemble(b, mode=mode)
    oa = ea.step(3)
    for _ in range(3):
        ob = eb.step(1)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        same(x, y)
    assert [o["dt_new"] for o in oa] == [o["dt_new"] for o in ob]
    ea.close()
    eb.close()


@pytest.mark.parametrize("mode", ["persistent", "wide", "grid"])
def test_ensemble_nonfinite_marker(gpu, mode):
    m = member(24, 24, 3)
    m["x_m"][5] = float("nan")
    ens = sfhip.Ensemble([m], mode=mode)
    with pytest.raises(sfhip.SFError):
        ens.step(1)
    ens.close()


def step_refined(m, dt):
    """One StokesFlow2D step with the refined solver, from the reference-named calls."""
    st, dx, dy = m["state"], m["dx"], m["dy"]
    nx, ny = st.nx, st.ny
    o = sfhip.grid_origins(dx, dy)
    sfhip.bilinearMarkerToGrid(*o["rho"], st.rho_vy, m["x_m"], m["y_m"], m["rho_m"], dx, dy, print_info=False)
    sfhip.bilinearMarkerToGrid(*o["mu_b"], st.mu_b, m["x_m"], m["y_m"], m["mu_m"], dx, dy, print_info=False)
    sfhip.bilinearMarkerToGrid(*o["mu_p"], st.mu_p, m["x_m"], m["y_m"], m["mu_m"], dx, dy, print_info=False)
    stats = {}
    dt, _, _ = sfhip.solveStokes(st.P, st.Vx, st.Vy, st.rho_vy, st.mu_b, st.mu_p, None, None, None, None, None, None,
                                 None, st.dVxdtau, st.dVydtau, None, None, st.Vx_small, st.Vy_small, 1.0, dx, dy, nx,
                                 ny, dt, 0.5, eps=1e-8, print_info=False, stats=stats, mixed_precision=True)
    sfhip.moveMarkersRK4(m["x_m"], m["y_m"], st.Vx, st.Vy, *o["vx"], *o["vy"], dt, (nx - 1) * dx, (ny - 1) * dy, dx,
                         dy, print_info=False)
    return dt, stats


@pytest.mark.parametrize("use_graph", [True, False])
def test_ensemble_refined_matches_single_refined_steps(gpu, use_graph):
    ens_m = [member(nx, ny, 70 + k) for k, (nx, ny) in enumerate(SIZES)]
    ref_m = [member(nx, ny, 70 + k) for k, (nx, ny) in enumerate(SIZES)]
    ens = sfhip.Ensemble(ens_m, use_graph=use_graph, mixed_precision=True)
    dts = [0.0] * len(SIZES)
    for t in range(2):
        outs = ens.step(1)
        torch.cuda.synchronize()
        for k, m in enumerate(ref_m):
            dts[k], st = step_refined(m, dts[k])
            assert outs[k]["niter"] == st["niter"] and outs[k]["dt_new"] == dts[k], (t, k)
        torch.cuda.synchronize()
        for a, b in zip(ens_m, ref_m):
            same(a, b)
    ens.close()


@pytest.mark.parametrize("mode", ["persistent", "wide", "grid"])
def test_ensemble_graph_against_oracle_loop(gpu, mode):
    """The captured ensemble step against the oracle's own timestep loop (StokesFlow2D_multixpu.jl:
    155-218 at one rank) per member: 2 FP64 steps, fields and markers bitwise, same iterations and
    dt."""
    ms = [member(nx, ny, 90 + k) for k, (nx, ny) in enumerate(SIZES)]
    host = []
    for k, (nx, ny) in enumerate(SIZES):
        dx, dy, x, y, rho, mu = cases.plume_markers(nx, ny, 4, seed=90 + k)
        host.append((nx, ny, dx, dy, x, y, rho, mu))
    ens = sfhip.Ensemble(ms, use_graph=True, mode=mode)
    got = [ens.step(1) for _ in range(2)]
    torch.cuda.synchronize()
    ens.close()
    for k, (nx, ny, dx, dy, x, y, rho, mu) in enumerate(host):
        S, _, outs = O.stokesflow2d(2, nx, ny, 1.0, 1.0, x, y, rho, mu, g_y=1.0, eps=1e-8)  # dx = 1 / (nx - 1)
        for t in range(2):
            assert got[t][k]["niter"] == outs[t]["niter"] and got[t][k]["dt_new"] == outs[t]["dt_new"], (k, t)
        st = ms[k]["state"]
        for f, o in (("Vx", "Vx_pad"), ("Vy", "Vy_pad"), ("P", "P"), ("rho_vy", "rho_vy"), ("mu_b", "mu_b"),
                     ("mu_p", "mu_p"), ("dVxdtau", "dVx"), ("dVydtau", "dVy")):
            assert np.array_equal(sfhip.to_numpy(getattr(st, f)), S[o]), (k, f)
        assert np.array_equal(sfhip.to_numpy(ms[k]["x_m"]), x) and np.array_equal(sfhip.to_numpy(ms[k]["y_m"]), y)
