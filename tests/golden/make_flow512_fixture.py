"""Writes tests/golden/flow512_oracle.json: the oracle's StokesFlow2D loop on the 511 x 511-cell
plume (4,177,936 markers, 16 per cell, seed 2000, air viscosity 1e-2, g = 1), 20 timesteps at
eps 1e-8.  The GPU test (tests/test_gpu_flow512.py) runs the same 20 steps through sf_step and
compares against this file: per-step iteration count, dt and err (exact float64 bits), and the
SHA-256 of the final fields and marker arrays.

The loop is the oracle's restatement of StokesFlow2D_multixpu.jl:155-218 (marker -> grid x3,
solveStokes!, moveMarkersRK4!).  The markers come from the oracle's seeded generator, which is
bit-identical to sf_init_markers / sf_plume_properties on the device.

About 6-7 minutes on 8 threads.  Usage, from the repository root:

    python tests/golden/make_flow512_fixture.py
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import sf_oracle as O  # noqa: E402

N, NT, SEED, EPS = 512, 20, 2000, 1e-8
PLUME = (0.1, 0.5, 0.8, 0.1, 0.001, 0.9, 1.0, 0.01, 0.1, 1.0)  # air_h, cx, cy, r, rho air/plume/bg, mu air/plume/bg
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flow512_oracle.json")


def digest(a):
    """SHA-256 of an array's float64 values in column-major (Julia) order."""
    return hashlib.sha256(np.asarray(a, dtype=np.float64).ravel(order="F").tobytes()).hexdigest()


def main():
    dx = dy = 1.0 / (N - 1)
    x, y = O.seeded_markers(N, N, 4, dx, dy, 0.25, SEED)
    rho, mu = O.plume_properties(x, y, *PLUME)
    rec = {"n": N, "steps": NT, "seed": SEED, "eps": EPS, "g_y": 1.0, "maxdisp": 0.5, "md": 4, "jitter": 0.25,
           "plume": list(PLUME), "markers": int(x.size),
           "markers_initial_sha256": {"x_m": digest(x), "y_m": digest(y), "rho_m": digest(rho), "mu_m": digest(mu)}}
    t0 = time.perf_counter()
    fields = {}

    def cb(t, S, xm, ym):
        print(f"step {t} {time.perf_counter() - t0:.1f} s", flush=True)
        if t == 0:
            fields["step1"] = {k: digest(S[s]) for k, s in (("Vx", "Vx_pad"), ("Vy", "Vy_pad"), ("P", "P"),
                                                             ("rho_vy", "rho_vy"), ("mu_b", "mu_b"),
                                                             ("mu_p", "mu_p"))}
            fields["step1"]["x_m"], fields["step1"]["y_m"] = digest(xm), digest(ym)

    S, t_tot, outs = O.stokesflow2d(NT, N, N, 1.0, 1.0, x, y, rho, mu, g_y=1.0, eps=EPS, callback=cb)
    rec["per_step"] = [{"niter": int(o["niter"]), "dt_new": float(o["dt_new"]).hex(), "err": float(o["err"]).hex(),
                        "converged": int(o["converged"])} for o in outs]
    rec["t_tot"] = float(t_tot).hex()
    rec["sha256_after_step1"] = fields["step1"]
    rec["sha256_final"] = {k: digest(S[s]) for k, s in (("Vx", "Vx_pad"), ("Vy", "Vy_pad"), ("P", "P"),
                                                         ("rho_vy", "rho_vy"), ("mu_b", "mu_b"), ("mu_p", "mu_p"),
                                                         ("dVxdtau", "dVx"), ("dVydtau", "dVy"))}
    rec["sha256_final"]["x_m"], rec["sha256_final"]["y_m"] = digest(x), digest(y)
    # a few plain numbers so that a mismatch can be read without the hashes
    rec["final_summary"] = {"Vx_l2": float(np.linalg.norm(S["Vx_pad"])), "Vy_l2": float(np.linalg.norm(S["Vy_pad"])),
                            "P_l2": float(np.linalg.norm(S["P"])), "x_sum": float(x.sum()), "y_sum": float(y.sum())}
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
    print("wrote", OUT, f"({time.perf_counter() - t0:.0f} s)")


if __name__ == "__main__":
    main()
