"""Accuracy of reduced-precision population storage against FP64.

Runs the same Taylor-Green vortex (u0 = 0.05, tau, no noise, no force, rho0 = 1)
with FP64 populations and with FP32 / FP16 deviation storage (f - w_s; FP32
collision, FP64 moment sums), then compares density and velocity field by
field:

    rel-L2(rho) = ||rho_X - rho_64|| / ||rho_64 - 1||   (deviation from rho0)
    rel-L2(u)   = ||u_X - u_64|| / ||u_64||

and reports the MLUPS of each run.  Default: D3Q27, AA, 512^3, tau = 0.51,
2000 steps (the configuration the FP16 storage is specified for).

    python lbmminiapp_amd/precision_study.py [--n 512] [--q 27] [--tau 0.51] [--steps 2000]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from lbmminiapp_amd.capi import DenseLattice  # noqa: E402


def run(n, q, tau, steps, storage, u0, noise, f16_scale):
    with DenseLattice(n, n, n, q=q, pattern="aa", storage=storage, omega=1.0 / tau,
                      force=(0.0, 0.0, 0.0), f16_scale=f16_scale) as lat:
        lat.init_tgv(u0, noise, 1)
        m0 = lat.mass()
        lat.synchronize()
        t0 = time.perf_counter()
        ms = lat.run_timed(steps)
        wall = time.perf_counter() - t0
        m1 = lat.mass()
        rho, u = lat.moments()
        nf = lat.info()["n_fluid"]
    return {"rho": rho, "u": u, "mlups": nf * steps / wall / 1e6, "device_ms": ms,
            "mass_drift": (m1 - m0) / m0}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--q", type=int, default=27)
    ap.add_argument("--tau", type=float, default=0.51)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--u0", type=float, default=0.05)
    ap.add_argument("--noise", type=float, default=0.0)
    ap.add_argument("--f16-scale", type=float, default=0.0)
    ap.add_argument("--storages", default="f32s,f16s")
    args = ap.parse_args(argv)
    ref = run(args.n, args.q, args.tau, args.steps, "f64", args.u0, args.noise, 0.0)
    drho = np.linalg.norm(ref["rho"] - 1.0)
    nu = np.linalg.norm(ref["u"])
    report = {"config": {"q": args.q, "n": args.n, "tau": args.tau, "steps": args.steps,
                         "u0": args.u0, "noise": args.noise, "pattern": "AA"},
              "f64": {"mlups": ref["mlups"], "mass_drift": ref["mass_drift"]}}
    print(f"# f64: {ref['mlups']:.0f} MLUPS", flush=True)
    for st in [s for s in args.storages.split(",") if s]:
        r = run(args.n, args.q, args.tau, args.steps, st, args.u0, args.noise, args.f16_scale)
        e_rho = float(np.linalg.norm(r["rho"] - ref["rho"]) / drho)
        e_u = float(np.linalg.norm(r["u"] - ref["u"]) / nu)
        e_umax = float(np.abs(r["u"] - ref["u"]).max() / np.abs(ref["u"]).max())
        report[st] = {"mlups": r["mlups"], "speedup_vs_f64": r["mlups"] / ref["mlups"],
                      "rel_l2_u": e_u, "rel_l2_rho_dev": e_rho, "rel_max_u": e_umax,
                      "mass_drift": r["mass_drift"]}
        print(f"# {st}: {r['mlups']:.0f} MLUPS, u rel-L2 {e_u:.3e}, rho-deviation rel-L2 "
              f"{e_rho:.3e}, mass drift {r['mass_drift']:.2e}", flush=True)
        del r
    print(json.dumps(report))


if __name__ == "__main__":
    main()
