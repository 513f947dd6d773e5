"""The 511 x 511-cell plume timestep loop (4,177,936 markers, 16 per cell, 20 steps at eps 1e-8)
on the GPU against the oracle's run of the same loop, committed as fixtures by
tests/golden/make_flow512_fixture.py: every step's iteration count and dt bit for bit, err to
1e-12 relative (a mean summed in another fixed order), and the SHA-256 of the fields and marker
arrays after step 1 and after step 20.  Markers and
materials are generated on the device (bit-identical to the oracle's generator, which the
fixture's initial-marker hashes pin)."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

import sfhip

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "flow512_oracle.json")


def digest(t):
    """SHA-256 of a device array's float64 values in column-major (Julia) order."""
    a = sfhip.to_numpy(t)
    return hashlib.sha256(np.asarray(a, dtype=np.float64).ravel(order="F").tobytes()).hexdigest()


def field_digests(st, x, y):
    return {"Vx": digest(st.Vx), "Vy": digest(st.Vy), "P": digest(st.P), "rho_vy": digest(st.rho_vy),
            "mu_b": digest(st.mu_b), "mu_p": digest(st.mu_p), "x_m": digest(x), "y_m": digest(y)}


def test_flow512_twenty_steps_bitwise_vs_oracle_fixture(gpu):
    f = json.load(open(FIXTURE))
    n, nt = f["n"], f["steps"]
    dx = dy = 1.0 / (n - 1)
    x, y = sfhip.init_markers(n, n, f["md"], dx, dy, f["jitter"], seed=f["seed"])
    rho, mu = sfhip.plume_properties(x, y, *f["plume"])
    assert x.numel() == f["markers"] == 4_177_936
    init = {"x_m": digest(x), "y_m": digest(y), "rho_m": digest(rho), "mu_m": digest(mu)}
    assert init == f["markers_initial_sha256"]
    st = sfhip.FlowState(n, n)
    dt, t_tot = 0.0, 0.0
    for t in range(nt):
        dt, s = sfhip.step(st, x, y, rho, mu, dx, dy, dt, g_y=f["g_y"], maxdisp=f["maxdisp"], eps=f["eps"])
        ref = f["per_step"][t]
        print(f"step {t + 1}: niter {s['niter']} (oracle {ref['niter']}) dt {dt!r}", flush=True)
        assert s["niter"] == ref["niter"], t
        assert dt == float.fromhex(ref["dt_new"]), t
        # err is a mean in another fixed summation order (tree on the device): equal to rounding
        assert s["err"] == pytest.approx(float.fromhex(ref["err"]), rel=1e-12, abs=0.0), t
        assert bool(s["converged"]) == bool(ref["converged"]), t
        assert s["m2g"] == [(0, 0), (0, 0), (0, 0)], t
        t_tot += dt
        if t == 0:
            torch.cuda.synchronize()
            got = field_digests(st, x, y)
            assert got == f["sha256_after_step1"], {k: got[k] == f["sha256_after_step1"][k] for k in got}
    torch.cuda.synchronize()
    got = field_digests(st, x, y)
    got["dVxdtau"], got["dVydtau"] = digest(st.dVxdtau), digest(st.dVydtau)
    assert got == f["sha256_final"], {k: got[k] == f["sha256_final"][k] for k in got}
    assert t_tot == float.fromhex(f["t_tot"])
