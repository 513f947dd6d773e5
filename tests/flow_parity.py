"""Long-run parity of the full timestep loop: the plume on 511 x 511 cells (4,177,936 markers,
16 per cell, seed 2000, air viscosity 1e-2), 20 steps at eps 1e-8, GPU against the oracle.

Not collected by pytest (the oracle needs about 400 s on 8 threads).  Usage, from the
repository root:

    python tests/flow_parity.py oracle OUTDIR     # CPU oracle (StokesFlow2D loop at one rank)
    python tests/flow_parity.py gpu OUTDIR        # libsfhip on the GPU (sf_step per step)
    python tests/flow_parity.py compare DIR_A DIR_B

Each run writes step-0 and last-step fields (Vx, Vy, P, rho_vy, mu_b, mu_p), a marker sample
and the per-step iteration counts and dt; compare prints relative L2 differences and
whether everything is bitwise equal.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for _p in (ROOT, os.path.join(ROOT, "stokesflow.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

from oracle import sf_oracle as O  # noqa: E402

N, NT, SEED = 512, 20, 2000


def setup():
    dx = dy = 1.0 / (N - 1)
    x, y = O.seeded_markers(N, N, 4, dx, dy, 0.25, SEED)
    rho, mu = O.plume_properties(x, y, 0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, 0.01, 0.1, 1.0)
    return dx, dy, x, y, rho, mu


def save(out, t, fields, x, y):
    np.savez(os.path.join(out, f"step{t}.npz"), xs=x[::97], ys=y[::97], xsum=x.sum(), ysum=y.sum(), **fields)


def run_oracle(out):
    dx, dy, x, y, rho, mu = setup()
    steps, t0 = [], time.perf_counter()

    def cb(t, S, xm, ym):
        steps.append(time.perf_counter() - t0)
        print(f"step {t} {steps[-1]:.1f} s", flush=True)
        if t in (0, NT - 1):
            save(out, t, {"Vx": S["Vx_pad"], "Vy": S["Vy_pad"], "P": S["P"], "rho": S["rho_vy"], "mu_b": S["mu_b"],
                          "mu_p": S["mu_p"]}, xm, ym)

    _, _, outs = O.stokesflow2d(NT, N, N, 1.0, 1.0, x, y, rho, mu, g_y=1.0, eps=1e-8, callback=cb)
    return {"seconds": time.perf_counter() - t0, "threads": os.cpu_count(),
            "steps": [{"niter": o["niter"], "dt_new": o["dt_new"], "err": o["err"]} for o in outs]}


def run_gpu(out):
    import torch

    import sfhip

    dx, dy, x, y, rho, mu = setup()
    xg, yg, rg, mg = (sfhip.from_numpy(a) for a in (x, y, rho, mu))
    st = sfhip.FlowState(N, N)
    dt, steps = 0.0, []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(NT):
        dt, s = sfhip.step(st, xg, yg, rg, mg, dx, dy, dt, g_y=1.0, eps=1e-8)
        torch.cuda.synchronize()
        steps.append({"niter": s["niter"], "dt_new": dt, "err": s["err"]})
        print(f"step {t} {time.perf_counter() - t0:.2f} s", flush=True)
        if t in (0, NT - 1):
            save(out, t, {k: sfhip.to_numpy(v) for k, v in (("Vx", st.Vx), ("Vy", st.Vy), ("P", st.P),
                                                              ("rho", st.rho_vy), ("mu_b", st.mu_b),
                                                              ("mu_p", st.mu_p))},
                 sfhip.to_numpy(xg), sfhip.to_numpy(yg))
    return {"seconds": time.perf_counter() - t0, "steps": steps}


def compare(a, b):
    ja, jb = (json.load(open(os.path.join(d, "run.json"))) for d in (a, b))
    same = [sa["niter"] == sb["niter"] and sa["dt_new"] == sb["dt_new"] for sa, sb in zip(ja["steps"], jb["steps"])]
    print("iterations and dt equal at every step:", all(same))
    ok = all(same)
    for t in (0, NT - 1):
        A, B = np.load(os.path.join(a, f"step{t}.npz")), np.load(os.path.join(b, f"step{t}.npz"))
        for k in ("Vx", "Vy", "P", "rho", "mu_b", "mu_p", "xs", "ys"):
            rel = np.linalg.norm(A[k] - B[k]) / np.linalg.norm(A[k])
            eq = np.array_equal(A[k], B[k])
            ok &= eq
            print(f"step {t} {k}: rel L2 {rel:.2e} bitwise {eq}")
    print("seconds:", ja["seconds"], jb["seconds"])
    return ok


if __name__ == "__main__":
    cmd = sys.argv[1]
    if cmd == "compare":
        sys.exit(0 if compare(sys.argv[2], sys.argv[3]) else 1)
    out = sys.argv[2]
    os.makedirs(out, exist_ok=True)
    rec = run_oracle(out) if cmd == "oracle" else run_gpu(out)
    json.dump(rec, open(os.path.join(out, "run.json"), "w"), indent=1)
