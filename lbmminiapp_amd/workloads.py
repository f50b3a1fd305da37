"""Measured runs of the workloads beside the headline 512^3 D3Q19 AA line.
Each prints one JSON line per run.

  tgv64      D3Q19 AA SOA FP64 64^3 periodic Taylor-Green (u0 = 0.05 plus U(-1e-4, 1e-4)
             noise, tau = 0.6), F = 0, 1000 steps.
  decks      The reference's benchmark decks: x-periodic cylinder pipe, simulationSize 2 / 4 / 8,
             D3Q19 SOA FP64 AA / AB push / AB pull, the deck's own step count.  Run through the
             drop-in driver (bin/lbm-proxy-app: reference fluid order and adjacency, sparse
             kernels) and again with `engine dense` on the same geometry.  Compare with the
             reference's published MLUPS for these decks.
  channel    D3Q19 AB pull, SOA, FP64, 256^3 channel (solid layers y = 0 and y = 255, periodic
             x and z), body force 1e-6 along x, tau = 0.8, seeded spheres of radius 4-12 added
             until 5% of the cells are solid (seed 2), rest start, 5000 steps.
  porous     D3Q19 AB push, FP32 deviation storage, 384^3 packed spheres of radius 6-16 added
             until 35% solid (seed 6), periodic, inlet/outlet pressure drop 1e-4 along x (the
             pressure-drop boundary: density jump 3e-4 across the periodic x boundary), 3000
             steps; SOA and AOS, and an FP64 SOA run for the field-by-field error of FP32.
  precision  D3Q27 AA 512^3 Taylor-Green (u0 = 0.05, tau = 0.51), 2000 steps, FP64 vs FP32 / FP16
             deviation storage field by field (precision_study.py).

MLUPS = fluid nodes * steps / (1e6 * seconds of the timed steps); "achieved_gbs" counts the
algorithmic 2 Q b bytes per fluid node.

  porous_adjoint  the adjoint of the outlet flux with respect to tau on the porous workload.

    python -m lbmminiapp_amd.workloads tgv64|decks|channel|porous|porous_adjoint|precision [--steps K]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from lbmminiapp_amd import capi  # noqa: E402

DRIVER = os.path.join(ROOT, "lbmminiapp_amd", "bin", "lbm-proxy-app")
DECKS = os.path.join(ROOT, "tests", "decks")
PEAK_HBM_GBS = 8000.0

# The reference's published MLUPS for its decks (its image.png: an Intel GPU; best of its SYCL
# and Kokkos-SYCL builds), keyed by (simulationSize, pattern).  Two tiles are the whole GPU.
PUBLISHED_1TILE = {
    (2, "aa"): 1017.2, (4, "aa"): 913.3, (8, "aa"): 489.2,
    (2, "ab"): 689.1, (4, "ab"): 695.5, (8, "ab"): 401.5,
    (2, "ab_pull"): 882.7, (4, "ab_pull"): 861.6, (8, "ab_pull"): 459.9,
}
PUBLISHED_2TILES = {
    (2, "aa"): 1388.7, (4, "aa"): 1212.9, (8, "aa"): 956.4,
    (2, "ab"): 708.4, (4, "ab"): 909.1, (8, "ab"): 775.1,
    (2, "ab_pull"): 1696.7, (4, "ab_pull"): 1559.0, (8, "ab_pull"): 1040.1,
}


def _emit(rec):
    print(json.dumps(rec), flush=True)


def run_deck(size, pattern, engine, steps=None, timeout=600):
    src = os.path.join(DECKS, f"{size}_input_{pattern}_unroll.txt")
    with open(src) as fh:
        text = fh.read()
    if steps is not None:
        text = re.sub(r"(?m)^timesteps\s+\d+", f"timesteps {steps}", text)
    if engine == "dense":
        text += "\nengine dense\n"
    with tempfile.TemporaryDirectory() as d:
        deck = os.path.join(d, "deck.txt")
        with open(deck, "w") as fh:
            fh.write(text)
        out = subprocess.run([DRIVER, deck], cwd=d, capture_output=True, text=True, timeout=timeout)
    if out.returncode != 0:
        raise RuntimeError(f"driver failed ({out.returncode}): {out.stderr[-2000:]}")
    m = re.search(r"Total ([0-9.eE+-]+) MFLUPS", out.stdout)
    if not m:
        raise RuntimeError("no MFLUPS line in the driver output:\n" + out.stdout[-2000:])
    nsteps = int(re.search(r"(?m)^timesteps\s+(\d+)", text).group(1))
    return float(m.group(1)), nsteps, out.stdout


def decks(args):
    for size in (2, 4, 8):
        for pattern in ("aa", "ab", "ab_pull"):
            for engine in ("sparse", "dense"):
                mlups, nsteps, _ = run_deck(size, pattern, engine, args.steps)
                pub = PUBLISHED_2TILES[(size, pattern)]
                _emit({"workload": f"reference deck {size}_input_{pattern}_unroll "
                                   f"(cylinder, D3Q19 SOA FP64)",
                       "engine": engine, "steps": nsteps, "mlups": round(mlups, 1),
                       "published_1tile_mlups": PUBLISHED_1TILE[(size, pattern)],
                       "published_2tiles_mlups": pub,
                       "speedup_vs_published_gpu": round(mlups / pub, 2)})


def _time_dense(lat, steps, warmup):
    lat.step(warmup)
    lat.synchronize()
    m0 = lat.mass()
    t0 = time.perf_counter()
    ms = lat.run_timed(steps)
    lat.synchronize()
    wall = time.perf_counter() - t0
    m1 = lat.mass()
    info = lat.info()
    nf = info["n_fluid"]
    return {"mlups": round(nf * steps / wall / 1e6, 1),
            "achieved_gbs": round(info["bytes_per_lup"] * nf * steps / (ms * 1e-3) / 1e9, 1),
            "frac_of_8tbs": round(info["bytes_per_lup"] * nf * steps / (ms * 1e-3) / 1e9
                                  / PEAK_HBM_GBS, 4),
            "fluid_nodes": nf, "steps": steps, "mass_drift": (m1 - m0) / m0}


def tgv64(args):
    """D3Q19 AA SOA FP64 64^3 periodic Taylor-Green (u0 = 0.05 plus seeded
    U(-1e-4, 1e-4) noise, tau = 0.6), F = 0, 1000 steps."""
    n = 64
    steps = args.steps or 1000
    with capi.DenseLattice(n, n, n, q=19, pattern="aa", storage="f64", omega=1 / 0.6) as lat:
        lat.init_tgv(0.05, 1e-4, 1)
        r = _time_dense(lat, steps, 10)
    r.update({"workload": "D3Q19 AA SOA FP64 64^3 periodic Taylor-Green, tau = 0.6, F = 0"})
    _emit(r)


def channel(args):
    n = 256
    solid = capi.channel(n, n, n)
    capi.add_spheres(solid, n, n, n, 4.0, 12.0, 0.05, 2, periodic=(1, 0, 1))
    steps = args.steps or 5000
    with capi.DenseLattice(n, n, n, q=19, pattern="ab", pull=True, layout="soa", storage="f64",
                           periodic=(1, 0, 1), omega=1 / 0.8, force=(1e-6, 0, 0),
                           solid=solid) as lat:
        lat.init_tgv(0.0, 0.0, 1)  # rest: rho = 1, u = 0
        r = _time_dense(lat, steps, 10)
        rho, u = lat.moments()
    fluid = solid == 0
    r.update({"workload": "D3Q19 AB pull SOA FP64 256^3 channel, 5% spheres r 4-12 (seed 2), "
                          "F = 1e-6, tau = 0.8",
              "solid_fraction": round(float(1 - fluid.mean()), 4),
              "mean_ux": float(u[fluid, 0].mean())})
    _emit(r)


def porous(args):
    n = 384
    solid = np.zeros(n ** 3, np.uint8)
    capi.add_spheres(solid, n, n, n, 6.0, 16.0, 0.35, 6, periodic=(1, 1, 1))
    steps = args.steps or 3000
    dp = 1e-4
    fluid = solid == 0
    x = np.arange(n ** 3) % n
    fields = {}
    for layout, storage in (("soa", "f32s"), ("aos", "f32s"), ("soa", "f64")):
        with capi.DenseLattice(n, n, n, q=19, pattern="ab", layout=layout, storage=storage,
                               periodic=(1, 1, 1), omega=1.0, solid=solid,
                               density_jump_x=3.0 * dp) as lat:
            lat.init_tgv(0.0, 0.0, 1)
            r = _time_dense(lat, steps, 10)
            rho, u = lat.moments()
        r.update({"workload": f"D3Q19 AB push {layout.upper()} "
                              f"{'FP32-deviation' if storage == 'f32s' else 'FP64'} 384^3 porous, "
                              f"35% spheres r 6-16 (seed 6), inlet/outlet pressure drop 1e-4 along x",
                  "layout": layout, "storage": storage,
                  "solid_fraction": round(float((solid != 0).mean()), 4),
                  "mean_ux": float(u[fluid, 0].mean()),
                  "rho_inlet_minus_outlet": float(rho[fluid & (x == 0)].mean()
                                                  - rho[fluid & (x == n - 1)].mean())})
        if layout == "soa":
            fields[storage] = (rho[fluid], u[fluid])
        if storage == "f64":
            (r32, u32), (r64, u64) = fields["f32s"], fields["f64"]
            r["f32s_vs_f64"] = {"rel_l2_u": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
                                "rel_l2_rho_dev": float(np.linalg.norm(r32 - r64)
                                                        / np.linalg.norm(r64 - 1.0))}
        _emit(r)


def porous_adjoint(args):
    """Adjoint of the outlet mass flux (plane x = nx - 1) with respect to a per-node tau field
    in [0.55, 1.0] (smooth seeded field, sigma 4, seed 7) on the porous workload, FP32 deviation
    storage, AB push SOA.  Nested checkpoints every 250 / 25 / 5 steps keep 25 states on the
    device instead of ~110 for one level."""
    n = 384
    solid = np.zeros(n ** 3, np.uint8)
    capi.add_spheres(solid, n, n, n, 6.0, 16.0, 0.35, 6, periodic=(1, 1, 1))
    tau = 0.55 + 0.45 * capi.smooth_random_field(n, n, n, 4.0, 7)
    steps = args.steps or 3000
    nf = int((solid == 0).sum())
    with capi.Adjoint(n, n, n, q=19, storage="f32s", omega=1.0, solid=solid, tau=tau,
                      objective="outlet_flux", outlet_x=n - 1, checkpoint_every=250,
                      checkpoint_inner=(25, 5), density_jump_x=3e-4) as adj:
        t0 = time.perf_counter()
        J = adj.forward(steps)
        fwd = time.perf_counter() - t0
        _, gt, ms = adj.backward(gamma=False)
        mem = adj.device_bytes()
    fluid = solid == 0
    _emit({"workload": "adjoint of outlet flux wrt tau, D3Q19 AB push FP32-deviation 384^3 porous "
                       "(35% spheres r 6-16, seed 6), tau in [0.55, 1] (sigma 4, seed 7), inlet/"
                       "outlet pressure drop 1e-4 along x",
           "steps": steps, "checkpoints": [250, 25, 5], "objective": J,
           "forward_s": round(fwd, 2), "backward_ms": round(ms, 1),
           "adjoint_mlups": round(nf * steps / (ms * 1e3), 1), "device_bytes": mem,
           "grad_l2": float(np.linalg.norm(gt[fluid])), "grad_nonzero": int((gt != 0).sum()),
           "fluid_nodes": nf})


def precision(args):
    from lbmminiapp_amd import precision_study
    extra = ["--steps", str(args.steps)] if args.steps else []
    precision_study.main(extra)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["tgv64", "decks", "channel", "porous",
                                         "porous_adjoint", "precision"])
    ap.add_argument("--steps", type=int, default=None, help="override the step count")
    args = ap.parse_args(argv)
    {"tgv64": tgv64, "decks": decks, "channel": channel, "porous": porous,
     "porous_adjoint": porous_adjoint, "precision": precision}[args.workload](args)


if __name__ == "__main__":
    main()
