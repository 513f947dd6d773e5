"""The benchmark workloads at their own shapes: the kernels the bench times, checked where it runs
them.

* 4095^2 cells (4096^2 nodes): the four- and three-iteration FP64 passes with the size-default
  tile heights (52 and 48 rows), bitwise against the oracle; the five-sweep FP32 pass of the refined solver
  bitwise against one sweep per launch, and its state against the FP64 solver's.
* 256 mixed-size plume instances (64..512 nodes per side): the batched solver (FP64 and refined)
  bitwise against one solve per instance, and one instance per size up to 256^2 against the
  oracle.
* 64 Rayleigh-Taylor members at 1023^2 cells (16.7 M markers each): one hipGraph-captured,
  refined ensemble timestep against per-member refined timesteps.
* The configurations as first specified (plume viscosity 1e-2 with air 1e-4; log10 plume
  contrast 4; Rayleigh-Taylor contrast 1e3 and 1e6) are beyond the reference PT scheme: the
  oracle diverges or stalls on them at 64^2 and the GPU reproduces it exactly.

Iteration caps keep every case to seconds; the caps fall after at least one convergence check
of every size."""
import numpy as np
import pytest
import torch

import sfhip
from oracle import sf_oracle as O

import cases

pytestmark = pytest.mark.gpu

STATE = ("P", "Vx_s", "Vy_s", "dVx", "dVy")


def device_plume(n, seed, cx=0.5, cy=0.8, r=0.1, mu_plume=0.1, mu_air=0.01):
    """Markers (16 per cell) and the three material grids of a plume instance, on the device."""
    dx = 1.0 / (n - 1)
    x, y = sfhip.init_markers(n, n, 4, dx, dx, 0.25, seed=seed)
    rho, mu = sfhip.plume_properties(x, y, 0.1, cx, cy, r, 0.001, 0.9, 1.0, mu_air, mu_plume, 1.0)
    st = sfhip.FlowState(n, n)
    assert sfhip.markers_to_grids3(st, x, y, rho, mu, dx, dx) == [(0, 0)] * 3
    del x, y, rho, mu
    return st, dx


# ------------------------------------------------------------------------------ 4095^2 cells

def test_config3_fp64_passes_4096_bitwise(gpu):
    """11 PT iterations at 4096^2 nodes with the library's default launch plan (two four-iteration
    passes of 52-row tiles, then a three-iteration pass of 48-row tiles), bitwise against the
    oracle."""
    n = 4096
    print("config 3: four- and three-iteration passes at 4096^2", flush=True)
    S = cases.random_state(n, n, seed=4096, scale=1e-3, contrast=4.0)
    G = {k: sfhip.from_numpy(S[k]) for k in STATE + ("rho_vy", "mu_b", "mu_p")}
    dx = 1.0 / (n - 1)
    sfhip.set_pt_fusion(n, n, 0, 0)
    sfhip.time_passes(n, n, True)
    try:
        sfhip.pt_iterations(11, G["P"], G["Vx_s"], G["Vy_s"], G["dVx"], G["dVy"], G["rho_vy"], G["mu_b"], G["mu_p"],
                            1.0, dx, dx, n, n, 0.01)
        torch.cuda.synchronize()
        launches = sfhip.pass_times(n, n)
    finally:
        sfhip.time_passes(n, n, False)
    assert sorted((t["precision"], t["iters_per_launch"], t["launches"]) for t in launches) == [(0, 3, 1), (0, 4, 2)]
    got = {k: sfhip.to_numpy(G[k]) for k in STATE}
    del G
    torch.cuda.empty_cache()
    O.compute_timesteps(S, n, n, dx, dx)
    O.pt_iterations(S, n, n, 11, 1.0, dx, dx, 0.01)
    for k in STATE:
        assert np.array_equal(got[k], S[k]), (k, np.max(np.abs(got[k] - S[k])))


def _solve(st, n, dx, iter_max, refined):
    stats = {}
    sfhip.solveStokes(st.P, st.Vx, st.Vy, st.rho_vy, st.mu_b, st.mu_p, None, None, None, None, None, None, None,
                      st.dVxdtau, st.dVydtau, None, None, st.Vx_small, st.Vy_small, 1.0, dx, dx, n, n, 0.0, 0.5,
                      eps=1e-8, print_info=False, iterMax=iter_max, stats=stats, mixed_precision=refined)
    torch.cuda.synchronize()
    return stats


def _reset(st):
    for t in (st.P, st.Vx, st.Vy, st.dVxdtau, st.dVydtau, st.Vx_small, st.Vy_small):
        t.zero_()


def test_config3_five_sweep_pass_4096(gpu):
    """The refined solver's FP32 window at 4096^2 on the plume workload (268 M markers): five
    sweeps per pass (the size default) bitwise equal to one sweep per launch, and the state after
    2000 iterations within 1e-4 relative L2 of the FP64 solver's after the same 2000 (FP32
    rounding carried through one window)."""
    n = 4096
    print("config 3: five-sweep FP32 pass at 4096^2", flush=True)
    st, dx = device_plume(n, seed=3000)
    out = {}
    for fuse in (0, 1):
        _reset(st)
        sfhip.set_pt_fusion(n, n, fuse, 0)
        sfhip.time_passes(n, n, True)
        try:
            s = _solve(st, n, dx, 2000, True)
            kinds = {(t["precision"], t["iters_per_launch"]): t["launches"] for t in sfhip.pass_times(n, n)}
        finally:
            sfhip.time_passes(n, n, False)
            sfhip.set_pt_fusion(n, n, 0, 0)
        assert s["n_fp32_iter"] == 2000 and s["n_outer"] == 1
        out[fuse] = (kinds, st.Vx.clone(), st.Vy.clone(), st.P.clone(), st.dVxdtau.clone(), st.dVydtau.clone())
    assert out[0][0] == {(1, 5): 400}, out[0][0]
    assert out[1][0] == {(1, 1): 2000}, out[1][0]
    for a, b in zip(out[0][1:], out[1][1:]):
        assert torch.equal(a, b)
    _reset(st)
    s64 = _solve(st, n, dx, 2000, False)
    assert s64["niter"] == 2000
    for name, a, b in zip(("Vx", "Vy", "P"), out[0][1:4], (st.Vx, st.Vy, st.P)):
        rel = ((a - b).norm() / b.norm()).item()
        print(f"  {name}: refined vs FP64 after 2000 iterations, rel L2 {rel:.2e}", flush=True)
        assert rel <= 1e-4, (name, rel)  # FP32 rounding carried through one 2000-sweep window


# ------------------------------------------------------------------------------ 256 instances

def _ensemble_specs(n_inst=256, seed=4000):
    rng = np.random.default_rng(seed)
    return [{"n": int(rng.choice((64, 128, 256, 512))), "cx": float(rng.uniform(0.2, 0.8)),
             "cy": float(rng.uniform(0.6, 0.9)), "r": float(rng.uniform(0.05, 0.15)),
             "c": float(rng.uniform(0.5, 1.25)), "seed": seed + 1 + k} for k in range(n_inst)]


def test_config4_256_instances(gpu):
    """256 plume instances, sizes uniform over 64..512 nodes per side, capped at 2001 iterations
    (after the first check of every size): batched FP64 and refined solves bitwise equal to one
    solve per instance; one instance per size up to 256^2 against the oracle."""
    specs = _ensemble_specs()
    print("config 4: 256 instances", flush=True)
    insts = [device_plume(s["n"], s["seed"], s["cx"], s["cy"], s["r"], 10.0 ** (-s["c"])) + (s["n"],) for s in specs]
    mats = [(st.rho_vy.clone(), st.mu_b.clone(), st.mu_p.clone()) for st, _, _ in insts]
    cap = 2001

    def problems():
        return [{"P": st.P, "Vx_pad": st.Vx, "Vy_pad": st.Vy, "dVxdtau": st.dVxdtau, "dVydtau": st.dVydtau,
                 "Vx_s": st.Vx_small, "Vy_s": st.Vy_small, "rho_vy": st.rho_vy, "mu_b": st.mu_b, "mu_p": st.mu_p,
                 "Nx": n, "Ny": n, "dx": dx, "dy": dx, "g_y": 1.0, "dt": 0.0, "maxdisp": 0.5, "eps": 1e-8,
                 "iterMax": cap} for st, dx, n in insts]

    def snapshot():
        return [(st.Vx.clone(), st.Vy.clone(), st.P.clone(), st.dVxdtau.clone()) for st, _, _ in insts]

    for refined in (False, True):
        for st, _, _ in insts:
            _reset(st)
        outs = sfhip.solveStokes_batched(problems(), mode="wide", mixed_precision=refined)
        torch.cuda.synchronize()
        batched = snapshot()
        loop = []
        for st, dx, n in insts:
            _reset(st)
            loop.append(_solve(st, n, dx, cap, refined))
        looped = snapshot()
        for i, (o, s) in enumerate(zip(outs, loop)):
            assert o["niter"] == s["niter"] and o["dt_new"] == s["dt_new"], (refined, i, specs[i]["n"])
            assert all(torch.equal(a, b) for a, b in zip(batched[i], looped[i])), (refined, i, specs[i]["n"])
        print(f"config 4 refined={refined}: {sum(o['converged'] for o in outs)} of 256 converged within {cap}",
              flush=True)
    # the FP64 loop's state is in place now for the refined run; redo one instance per size in FP64
    done = set()
    for (st, dx, n), (rho, mub, mup), spec in zip(insts, mats, specs):
        if n > 256 or n in done:
            continue
        done.add(n)
        S = O.solver_arrays(n, n)
        S["rho_vy"][:], S["mu_b"][:], S["mu_p"][:] = (sfhip.to_numpy(t) for t in (rho, mub, mup))
        exp = O.solve_stokes(S, 1.0, dx, dx, n, n, 0.0, 0.5, True, 1e-8, cap)
        _reset(st)
        s = _solve(st, n, dx, cap, False)
        assert s["niter"] == exp["niter"] and s["dt_new"] == exp["dt_new"], n
        for k, g in (("Vx_pad", st.Vx), ("Vy_pad", st.Vy), ("P", st.P), ("dVx", st.dVxdtau), ("dVy", st.dVydtau)):
            assert np.array_equal(sfhip.to_numpy(g), S[k]), (n, k)
    assert done == {64, 128, 256}


# ------------------------------------------------------------------------------ 64 x 1023^2 cells

def test_config5_captured_refined_step_64x1024(gpu):
    """One timestep of 64 Rayleigh-Taylor members at 1024^2 nodes (16.7 M markers each, 1.07 G in
    all) as one hipGraph-captured refined ensemble step, against the same step run member by
    member with the refined solver; solves capped at 2000 iterations."""
    n, n_mem, cap = 1024, 64, 2000
    dx = 1.0 / (n - 1)
    rng = np.random.default_rng(5000)
    specs = [{"k": int(rng.integers(1, 9)), "c": float(rng.uniform(0.0, 1.0)), "seed": 5001 + i} for i in range(n_mem)]
    print("config 5: 64 x 1024^2 captured refined step", flush=True)

    def member(s):
        x, y = sfhip.init_markers(n, n, 4, dx, dx, 0.25, seed=s["seed"])
        rho, mu = sfhip.rt_properties(x, y, 0.5, 0.02, s["k"], 1.0, 1.1, 10.0 ** s["c"], 1.0, 1.0)
        return {"state": sfhip.FlowState(n, n), "x_m": x, "y_m": y, "rho_m": rho, "mu_m": mu, "dx": dx, "dy": dx,
                "g_y": 1.0, "eps": 1e-8, "iterMax": cap}

    ens_m = [member(s) for s in specs]
    ens = sfhip.Ensemble(ens_m, use_graph=True, mode="wide", mixed_precision=True)
    outs = ens.step(1)
    torch.cuda.synchronize()
    ens.close()
    for i, s in enumerate(specs):
        m = member(s)
        st = m["state"]
        o = sfhip.grid_origins(dx, dx)
        sfhip.bilinearMarkerToGrid(*o["rho"], st.rho_vy, m["x_m"], m["y_m"], m["rho_m"], dx, dx, print_info=False)
        sfhip.bilinearMarkerToGrid(*o["mu_b"], st.mu_b, m["x_m"], m["y_m"], m["mu_m"], dx, dx, print_info=False)
        sfhip.bilinearMarkerToGrid(*o["mu_p"], st.mu_p, m["x_m"], m["y_m"], m["mu_m"], dx, dx, print_info=False)
        stats = {}
        dt, _, _ = sfhip.solveStokes(st.P, st.Vx, st.Vy, st.rho_vy, st.mu_b, st.mu_p, None, None, None, None, None,
                                     None, None, st.dVxdtau, st.dVydtau, None, None, st.Vx_small, st.Vy_small, 1.0, dx,
                                     dx, n, n, 0.0, 0.5, eps=1e-8, print_info=False, iterMax=cap, stats=stats,
                                     mixed_precision=True)
        sfhip.moveMarkersRK4(m["x_m"], m["y_m"], st.Vx, st.Vy, *o["vx"], *o["vy"], dt, (n - 1) * dx, (n - 1) * dx,
                             dx, dx, print_info=False)
        torch.cuda.synchronize()
        e = ens_m[i]
        assert outs[i]["niter"] == stats["niter"] and outs[i]["dt_new"] == dt, i
        for f in ("Vx", "Vy", "P", "rho_vy", "mu_b", "mu_p", "dVxdtau", "dVydtau"):
            assert torch.equal(getattr(e["state"], f), getattr(st, f)), (i, f)
        assert torch.equal(e["x_m"], m["x_m"]) and torch.equal(e["y_m"], m["y_m"]), i
        del m, st
        ens_m[i] = None
        torch.cuda.empty_cache()


# ------------------------------------------------------------------------------ configurations as first written

def _oracle_grids(n, x, y, rho, mu):
    dx = 1.0 / (n - 1)
    S = O.solver_arrays(n, n)
    vws, ws = O.zeros(n + 1, n), O.zeros(n + 1, n)
    o = O.grid_origins(dx, dx)
    O.marker_to_grid(*o["rho"], S["rho_vy"], x, y, rho, dx, dx, vws, ws)
    O.marker_to_grid(*o["mu_b"], S["mu_b"], x, y, mu, dx, dx, vws, ws)
    O.marker_to_grid(*o["mu_p"], S["mu_p"], x, y, mu, dx, dx, vws, ws)
    return S, dx


@pytest.mark.parametrize("name,mats,niter,outcome", [
    ("plume mu 1e-2, air mu 1e-4 (4095^2 workload as first written)", ("plume", 1e-4, 1e-2), 100000, "huge"),
    ("plume log10 contrast 4 (ensemble upper end)", ("plume", 1e-2, 1e-4), 320, "nan"),
    ("Rayleigh-Taylor contrast 1e6 (ensemble upper end)", ("rt", 6.0), 320, "nan"),
    ("Rayleigh-Taylor contrast 1e3 (ensemble lower end)", ("rt", 3.0), 640, "nan"),
])
def test_configs_as_written_beyond_the_scheme(gpu, name, mats, niter, outcome):
    """The reference PT scheme (StokesSolver_multixpu.jl:17-27: Vsc 0.45, damping 1 - 4/n,
    stabilisation from the CFL dt) on these materials at 63^2 cells: the oracle either blows up
    (NaN at a check) or grows without bound to iterMax; the GPU solve takes the same iterations
    to the same err."""
    n = 64
    dx = 1.0 / (n - 1)
    x, y = O.seeded_markers(n, n, 4, dx, dx, 0.25, 1003)
    if mats[0] == "rt":
        up = y < 0.5 + 0.02 * np.sin(2 * np.pi * 3 * x / 1.0)
        rho, mu = np.where(up, 1.1, 1.0), np.where(up, 10.0 ** mats[1], 1.0)
    else:
        rho, mu = O.plume_properties(x, y, 0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, mats[1], mats[2], 1.0)
    S, dx = _oracle_grids(n, x, y, rho, mu)
    G = {k: sfhip.from_numpy(S[k]) for k in S}
    st = {}
    sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], None, None, None, None,
                      None, None, None, G["dVx"], G["dVy"], None, None, G["Vx_s"], G["Vy_s"], 1.0, dx, dx, n, n, 0.0,
                      0.5, eps=1e-8, print_info=False, iterMax=100000, stats=st)
    torch.cuda.synchronize()
    exp = O.solve_stokes(S, 1.0, dx, dx, n, n, 0.0, 0.5, True, 1e-8, 100000)
    assert exp["niter"] == st["niter"] == niter and not exp["converged"] and not st["converged"], name
    if outcome == "nan":
        assert np.isnan(exp["err"]) and np.isnan(st["err"])
    else:
        assert exp["err"] > 1e100 and st["err"] == pytest.approx(exp["err"], rel=1e-9)
    for k in ("Vx_pad", "Vy_pad", "P"):
        assert np.array_equal(sfhip.to_numpy(G[k]), S[k], equal_nan=True), k


@pytest.mark.parametrize("name,mats,expect", [
    ("plume mu 0.1, air mu 1e-2 (the substituted 4095^2 materials)", ("plume", 1e-2, 0.1), 8640),
    ("plume log10 contrast 1.25 (upper end of the substituted ensemble range)", ("plume", 1e-2, 10 ** -1.25), 9920),
])
def test_substituted_configs_converge(gpu, name, mats, expect):
    """The substituted materials converge at 63^2 cells in the oracle and on the GPU alike."""
    n = 64
    dx = 1.0 / (n - 1)
    x, y = O.seeded_markers(n, n, 4, dx, dx, 0.25, 1003)
    rho, mu = O.plume_properties(x, y, 0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, mats[1], mats[2], 1.0)
    S, dx = _oracle_grids(n, x, y, rho, mu)
    G = {k: sfhip.from_numpy(S[k]) for k in S}
    st = {}
    sfhip.solveStokes(G["P"], G["Vx_pad"], G["Vy_pad"], G["rho_vy"], G["mu_b"], G["mu_p"], None, None, None, None,
                      None, None, None, G["dVx"], G["dVy"], None, None, G["Vx_s"], G["Vy_s"], 1.0, dx, dx, n, n, 0.0,
                      0.5, eps=1e-8, print_info=False, iterMax=100000, stats=st)
    torch.cuda.synchronize()
    exp = O.solve_stokes(S, 1.0, dx, dx, n, n, 0.0, 0.5, True, 1e-8, 100000)
    assert exp["converged"] and st["converged"] and exp["niter"] == st["niter"] == expect, name
