#!/usr/bin/env python3
"""Benchmark of the collide-stream hot path on one MI355X.

Workload (the headline metric): D3Q19 BGK, AA in-place propagation, SOA
x-row layout, 512^3 fully periodic box, tau = 0.6, rho0 = 1, Taylor-Green
vortex u0 = 0.05 plus seeded U(-1e-4, 1e-4) noise, populations resident in HBM
before the timed region.  One "step" is one AA time step over the whole box
(the even and odd kernels alternate).  Legs:

  f64   populations in FP64 (the reference arithmetic)       -> `value`
  f16s  FP16 deviations f - w_s, FP32 collision, FP64 moments -> extra field
  f32s  FP32 deviations                                        -> extra field

value = fluid nodes * K / (1e6 * wall seconds of the K timed steps), bracketed
by device synchronisation on both sides.  roofline.achieved = 2 Q b bytes per
node * nodes / mean kernel duration from HIP events recorded on the engine's
stream around the same K launches.

copy_kernel: the measured copy ceiling of the same GPU (a non-temporal
16-byte-per-lane copy of 4 GiB), and the fraction of it the headline leg
reaches.

cpu_baseline: the reference's own AA kernel (KernelAA_SOA_Unroll, the
`unroll 1` kernel of its benchmark decks) compiled from its sources into
oracle/_ref/, one process per host core (the reference's one MPI rank per
core: every core of the affinity mask, within the job's CPU share) on its own
128 x 128 x 8 z-slab of the same workload; if that library is absent, the C
oracle (OpenMP) instead.  The line states the cores used, the affinity mask
and nproc.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 512] [--legs f64,f16s,f32s]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
TAU = 0.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--q", type=int, default=19)
    ap.add_argument("--legs", default="f64,f16s,f32s,f64_force")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


def gpu_leg(leg, n, q, steps, warmup, keep_fields=False):
    """leg: f64 | f32s | f16s (Taylor-Green, F = 0) or f64_force (the reference's
    body force F = (1e-6, 0, 0) switched on, same initial state)."""
    from lbmminiapp_amd.capi import DenseLattice

    storage = leg.split("_")[0]
    force = (1e-6, 0.0, 0.0) if leg.endswith("_force") else (0.0, 0.0, 0.0)
    lat = DenseLattice(n, n, n, q=q, pattern="aa", storage=storage, omega=1.0 / TAU, force=force)
    try:
        lat.init_tgv(0.05, 1e-4, 1)
        lat.step(warmup)
        lat.synchronize()
        # per-kernel durations (even / odd launches), HIP events around single launches
        probe = {"even": [], "odd": []}
        for i in range(4):
            kind = "even" if lat.info()["time_step"] % 2 == 0 else "odd"
            probe[kind].append(lat.run_timed(1))
        if lat.info()["time_step"] % 2:
            lat.step(1)
        lat.synchronize()
        m0 = lat.mass()
        t0 = time.perf_counter()
        ev_ms = lat.run_timed(steps)
        lat.synchronize()
        wall = time.perf_counter() - t0
        m1 = lat.mass()
        info = lat.info()
        fields = lat.moments() if keep_fields else None
    finally:
        lat.close()
    nf = info["n_fluid"]
    bpl = info["bytes_per_lup"]
    launch_s = ev_ms * 1e-3 / steps
    achieved = bpl * nf / launch_s / 1e9
    return {
        "storage": storage,
        "force": force[0],
        "mlups": nf * steps / wall / 1e6,
        "ms_per_step": wall * 1e3 / steps,
        "event_ms_per_launch": ev_ms / steps,
        "even_kernel_ms": sum(probe["even"]) / max(1, len(probe["even"])),
        "odd_kernel_ms": sum(probe["odd"]) / max(1, len(probe["odd"])),
        "bytes_per_lup": bpl,
        "achieved_gbs": achieved,
        "frac": achieved / PEAK_HBM_GBS,
        "n_fluid": nf,
        "mass_drift": (m1 - m0) / m0,
        "time_step": info["time_step"],
        "fields": fields,
    }


def adjoint_leg(n=128, steps=500, every=50, storage="f64"):
    """Adjoint MLUPS on the adjoint configuration: D3Q19 n^3 channel (walls
    y = 0, n-1), Brinkman field gamma = smoothed seeded noise (sigma 4, seed 4)
    in [0, 1], F = (1e-6, 0, 0), tau = 0.8, J = kinetic energy after `steps`
    steps, checkpoint every `every`.  Value = fluid nodes * steps / backward
    device time (recomputation included)."""
    from lbmminiapp_amd import capi

    solid = capi.channel(n, n, n)
    gamma = capi.smooth_random_field(n, n, n, 4.0, 4)
    nf = int((solid == 0).sum())
    with capi.Adjoint(n, n, n, q=19, storage=storage, omega=1 / 0.8, force=(1e-6, 0, 0),
                      solid=solid, gamma=gamma, objective="kinetic_energy",
                      checkpoint_every=every) as adj:
        t0 = time.perf_counter()
        J = adj.forward(steps)
        fwd_s = time.perf_counter() - t0
        _, _, ms = adj.backward(tau=False)
        mem = adj.device_bytes()
    return {"value": round(nf * steps / (ms * 1e3), 1), "unit": "MLUPS", "storage": storage,
            "backward_ms": round(ms, 2), "forward_ms": round(fwd_s * 1e3, 2),
            "objective": J, "device_bytes": mem,
            "config": f"D3Q19 {n}^3 channel, Brinkman gamma field, {steps} steps, "
                      f"checkpoint every {every}, {storage}"}


def copy_peak(nbytes=4 << 30, reps=10):
    """The measured copy ceiling of this GPU: the library's non-temporal
    16-byte copy kernel between two nbytes buffers (nbytes read + nbytes
    written per dispatch, reps timed dispatches after one untimed)."""
    from lbmminiapp_amd import capi

    gbs = capi.copy_bandwidth(nbytes, reps)
    return {"value": round(gbs, 1), "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
            "bytes_per_dispatch": 2 * nbytes,
            "note": "non-temporal 16-byte-per-lane copy kernel, 4 GiB read + 4 GiB written per "
                    "dispatch; "
                    "frac_reached = the headline leg's achieved bandwidth / this"}


SLAB_PLANES = 8  # z-planes of 128 x 128 cells per reference rank


def _cpu_rank(rank, nranks, budget_s, barrier, probes, out):
    """One reference 'rank': the reference's KernelAA_SOA_Unroll (oracle/_ref, one thread) on its
    128 x 128 x SLAB_PLANES z-slab of the periodic Taylor-Green sample.  Every rank runs the same
    step count (sized from the slowest rank's probe) from a common start."""
    import numpy as np

    from oracle import oracle as O

    q, n = 19, 128
    omega = 1.0 / TAU
    geom = O.DenseGeometry(n, n, SLAB_PLANES)
    u = O.tgv_velocity(geom, u0=0.05, noise=1e-4, seed=1 + rank)
    fc = O.equilibrium(q, np.ones(geom.nfluid), u)
    adj = geom.adjacency(q, O.SOA, O.ADJ_PULL)
    f = O.to_storage(fc, q, O.SOA, aa_even_ready=True)
    nf = geom.nfluid
    barrier.wait()
    probes[rank] = O.ref_run("aa_soa_unroll", q, "f64", omega, nf, f.copy(), adj, 2)
    barrier.wait()
    steps = max(2, 2 * int(budget_s / max(probes[:])))
    barrier.wait()
    t = O.ref_run("aa_soa_unroll", q, "f64", omega, nf, f.copy(), adj, steps)
    out.put((rank, nf, steps, t))


def host_cores():
    """(cores to use, cores in this process's affinity mask, nproc).  The cores used are the
    affinity mask, capped at the host's CPU share for this job when OMP_NUM_THREADS states one
    (a one-GPU lease allots 16 CPUs of a larger host)."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    use = min(affinity, share) if share > 0 else affinity
    return max(1, use), affinity, nproc


def cpu_baseline(budget_s):
    """The reference's own AA kernel on the host cores, one process per core as its MPI ranks
    run (each on its own 128 x 128 x SLAB_PLANES z-slab, no halo exchange, so an upper bound for
    the reference).  Runs before any GPU work: the worker processes are spawned from a process that
    has not initialised the GPU.  Falls back to the C oracle with OpenMP when oracle/_ref is
    absent."""
    import multiprocessing as mp

    from oracle import oracle as O

    omega = 1.0 / TAU
    ranks, affinity, nproc = host_cores()
    host = {"host_affinity_cores": affinity, "host_nproc": nproc}
    why = "the reference build (oracle/_ref) is absent"
    res = None
    if O.ref_available(19, "f64", omega):
        import queue

        ctx = mp.get_context("spawn")
        barrier, out, probes = ctx.Barrier(ranks), ctx.Queue(), ctx.Array("d", ranks)
        procs = [ctx.Process(target=_cpu_rank, args=(r, ranks, budget_s, barrier, probes, out))
                 for r in range(ranks)]
        for p in procs:
            p.start()
        res, deadline = [], time.time() + 10 * budget_s + 120
        while len(res) < ranks and time.time() < deadline:
            try:
                res.append(out.get(timeout=1.0))
            except queue.Empty:
                if any(p.exitcode not in (None, 0) for p in procs):
                    break
        for p in procs:
            if len(res) < ranks and p.is_alive():
                p.terminate()
            p.join()
        if len(res) < ranks:
            why = f"the reference run failed (exit codes {[p.exitcode for p in procs]})"
            res = None
    if res:
        lups = sum(nf * steps for _, nf, steps, _ in res)
        wall = max(t for *_, t in res)
        one = res[0]
        return {"value": lups / wall / 1e6, "unit": "MLUPS", "cores": ranks, "kind": "reference",
                **host,
                "sample": f"D3Q19 AA SOA FP64 periodic Taylor-Green 128x128x{SLAB_PLANES * ranks} "
                          f"split into {ranks} z-slabs of 128x128x{SLAB_PLANES}, one process per "
                          f"slab and core (the reference's one-rank-per-core model, without its "
                          f"halo exchange), each running the reference's KernelAA_SOA_Unroll "
                          f"(unroll 1) compiled from its sources with g++ -O2, omega = 1/{TAU}, "
                          f"~{budget_s:.0f} s; one process: "
                          f"{one[1] * one[2] / one[3] / 1e6:.2f} MLUPS while all ran"}
    import numpy as np

    q, n = 19, 128
    geom = O.DenseGeometry(n, n, n)
    u = O.tgv_velocity(geom, u0=0.05, noise=1e-4, seed=1)
    fc = O.equilibrium(q, np.ones(geom.nfluid), u)
    adj = geom.adjacency(q, O.SOA, O.ADJ_PULL)
    f = O.to_storage(fc, q, O.SOA, aa_even_ready=True)
    nf = geom.nfluid
    os.environ.setdefault("OMP_NUM_THREADS", str(ranks))
    g = f.copy()
    t0 = time.perf_counter()
    O.run_aa(q, O.SOA, nf, g, adj, 2, omega, (0.0, 0.0, 0.0))
    dt = time.perf_counter() - t0
    steps = max(2, 2 * int(budget_s / dt))
    g = f.copy()
    t0 = time.perf_counter()
    O.run_aa(q, O.SOA, nf, g, adj, steps, omega, (0.0, 0.0, 0.0))
    dt = time.perf_counter() - t0
    return {"value": nf * steps / dt / 1e6, "unit": "MLUPS", "cores": ranks, "kind": "port", **host,
            "sample": f"D3Q19 AA SOA FP64 periodic Taylor-Green box {n}^3 ({nf} nodes) x {steps} "
                      f"steps, C oracle (restatement of KernelAA_SOA_Base) with OpenMP, "
                      f"{ranks} threads; {why}"}


def load_traffic(workload_key):
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh)
    entry = data.get(workload_key)
    if not entry:
        return None
    return entry.get("bytes_per_launch")


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: the engine runs on exactly one GPU; the whole domain is one GPU's "
              f"subdomain (multi-GPU halo exchange is a design, not a measured path). "
              f"--gpus {args.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)
    if args.steps % 2:
        args.steps += 1  # whole AA even/odd pairs
    legs = [s for s in args.legs.split(",") if s]
    n, q = args.size, args.q
    # the CPU leg first, while this process has not touched the GPU (it spawns workers)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    results = {}
    # density and velocity of the F = 0 legs after the same steps from the same state: the
    # reduced-precision legs are compared with the FP64 leg (itself bit-identical to the oracle)
    keep = "f64" in legs and any(l in ("f32s", "f16s") for l in legs)
    for leg in legs:
        results[leg] = gpu_leg(leg, n, q, args.steps, args.warmup,
                               keep_fields=keep and leg in ("f64", "f32s", "f16s"))
        print(f"# {leg}: {results[leg]['mlups']:.1f} MLUPS, "
              f"{results[leg]['achieved_gbs']:.0f} GB/s ({100 * results[leg]['frac']:.1f}% of "
              f"{PEAK_HBM_GBS:.0f} GB/s)", file=sys.stderr, flush=True)
    head = results.get("f64") or results[legs[0]]
    copy = copy_peak()
    print(f"# copy kernel: {copy['value']:.0f} GB/s ({100 * copy['frac']:.1f}% of "
          f"{PEAK_HBM_GBS:.0f} GB/s)", file=sys.stderr, flush=True)
    workload = f"D3Q{q} AA SOA {n}^3 periodic Taylor-Green, tau={TAU}, {head['storage']}"
    traffic = load_traffic(f"d3q{q}_aa_{n}_{head['storage']}")
    line = {
        "metric": "MLUPS & % HBM roofline, D3Q19 AA 512^3: FP64 vs FP16-shifted; adjoint MLUPS",
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32s": "f32", "f16s": "f16"}[head["storage"]],
        "data": "synthetic (Taylor-Green vortex + seeded noise, generated on the device)",
        "config": {"workload": workload, "nx": n, "ny": n, "nz": n, "q": q, "pattern": "AA",
                   "layout": "SOA", "storage": head["storage"], "tau": TAU,
                   "fluid_nodes": head["n_fluid"]},
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(head["frac"], 4),
                     "traffic": traffic},
        "kernels": {"even_ms": round(head["even_kernel_ms"], 4),
                    "odd_ms": round(head["odd_kernel_ms"], 4),
                    "bytes_per_lup": head["bytes_per_lup"]},
        "mass_drift": head["mass_drift"],
        "copy_kernel": dict(copy, frac_reached=round(head["achieved_gbs"] / copy["value"], 4)),
    }
    for leg in legs:
        if results[leg] is head:
            continue
        r = results[leg]
        line[f"{leg}_storage"] = {
            "value": round(r["mlups"], 1), "unit": "MLUPS", "ms_per_step": round(r["ms_per_step"], 4),
            "achieved_gbs": round(r["achieved_gbs"], 1), "frac": round(r["frac"], 4),
            "bytes_per_lup": r["bytes_per_lup"], "speedup_vs_f64": round(r["mlups"] / head["mlups"], 3),
            "mass_drift": r["mass_drift"]}
        if r["fields"] is not None and head["fields"] is not None:
            import numpy as np
            assert r["time_step"] == head["time_step"]
            (rho, u), (rho64, u64) = r["fields"], head["fields"]
            line[f"{leg}_storage"].update({
                "steps_compared": int(head["time_step"]),
                "rel_l2_u_vs_f64": float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
                "rel_l2_rho_dev_vs_f64": float(np.linalg.norm(rho - rho64)
                                               / np.linalg.norm(rho64 - 1.0))})
    if not args.no_adjoint:
        line["adjoint"] = adjoint_leg(storage="f64")
        line["adjoint_f32"] = adjoint_leg(storage="f32s")
        print(f"# adjoint: {line['adjoint']['value']} MLUPS (FP64), "
              f"{line['adjoint_f32']['value']} MLUPS (FP32)", file=sys.stderr, flush=True)
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
