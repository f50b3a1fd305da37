"""Writes the golden fixtures in tests/golden/*.npz from the reference's own
kernels, compiled from /root/reference/src into oracle/_ref/ by
oracle/Makefile.ref (run `make -C oracle` first).  Run in the container that
holds the reference:

    python tests/golden/gen_golden.py

Each fixture holds the inputs (grid, solid mask, periodic flags, omega, the
initial storage array, the adjacency) and the reference's output after
`steps` steps of one kernel variant.  Initial states are f^eq(rho, u) with rho
and u drawn from the counter-based splitmix64 generator in oracle/oracle.py
(integer arithmetic then +,*,/ only, so they are reproducible bit for bit on
any host).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

NX, NY, NZ = 8, 7, 6


def seeded_state(geom, q, seed):
    n = geom.nfluid
    cell = geom.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.02 * (O.uniform01(seed, np.uint64(4) * cell) - 0.5)
    u = np.stack([0.05 * (2.0 * O.uniform01(seed, np.uint64(4) * cell + np.uint64(k + 1)) - 1.0)
                  for k in range(3)], axis=1)
    return O.equilibrium(q, rho, u)


def seeded_solid(seed, frac=0.12):
    r = O.uniform01(seed, np.arange(NX * NY * NZ, dtype=np.uint64))
    return (r < frac).astype(np.uint8)


CASES = [
    # name, q, prec, omega, variant, layout, periodic, steps
    ("aa_soa_base_q19_f64_w0p6", 19, "f64", 1 / 0.6, "aa_soa_base", O.SOA, (1, 1, 1), 9),
    ("aa_aos_base_q19_f64_w1", 19, "f64", 1.0, "aa_aos_base", O.AOS, (1, 0, 1), 8),
    ("aa_soa_unroll_q19_f64_w0p8", 19, "f64", 1 / 0.8, "aa_soa_unroll", O.SOA, (1, 0, 1), 8),
    ("ab_soa_base_q19_f64_w0p8", 19, "f64", 1 / 0.8, "ab_soa_base", O.SOA, (1, 0, 1), 7),
    ("ab_aos_base_q19_f64_w0p6", 19, "f64", 1 / 0.6, "ab_aos_base", O.AOS, (1, 1, 1), 7),
    ("ab_soa_pull_q19_f64_w0p8", 19, "f64", 1 / 0.8, "ab_soa_pull", O.SOA, (1, 0, 1), 7),
    ("ab_soa_unroll_q19_f64_w1", 19, "f64", 1.0, "ab_soa_unroll", O.SOA, (1, 0, 1), 6),
    ("ab_soa_unroll_pull_q19_f64_w0p51", 19, "f64", 1 / 0.51, "ab_soa_unroll_pull", O.SOA,
     (1, 1, 1), 6),
    ("aa_soa_base_q27_f64_w0p51", 27, "f64", 1 / 0.51, "aa_soa_base", O.SOA, (1, 1, 1), 9),
    ("ab_soa_base_q27_f64_w1", 27, "f64", 1.0, "ab_soa_base", O.SOA, (1, 0, 1), 6),
    ("aa_soa_base_q19_f32_w1", 19, "f32", 1.0, "aa_soa_base", O.SOA, (1, 0, 1), 8),
    ("ab_soa_pull_q27_f32_w1", 27, "f32", 1.0, "ab_soa_pull", O.SOA, (1, 1, 1), 6),
]


def ref_geometry_fixture(q, deck_text, name):
    """Fluid points, border/bulk ranges and adjacency from the reference's own
    GeometryAAAB pipeline (oracle/ref_harness.cpp: ref_geometry) on a deck."""
    import ctypes
    import tempfile
    L = O.ref_lib(q, "f64", 1.0)
    L.ref_geometry.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 4
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as fh:
        fh.write(deck_text)
        deck = fh.name
    n = np.zeros(1, np.int32)
    ranges = np.zeros(4, np.int32)
    L.ref_geometry(deck.encode(), O._ptr(n), O._ptr(ranges), None, None)
    pts = np.zeros(3 * int(n[0]), np.int32)
    adj = np.zeros(q * int(n[0]), np.int32)
    L.ref_geometry(deck.encode(), O._ptr(n), O._ptr(ranges), O._ptr(pts), O._ptr(adj))
    os.unlink(deck)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), q=q, deck=deck_text, ranges=ranges,
                        fluid_pts=pts.reshape(-1, 3), adjacency=adj)
    print(name, int(n[0]), "fluid nodes")


GEOMETRY_CASES = [
    ("geometry_cylinder_s0p5_aa_q19", 19, "simulationSize 0.5\npropPattern 1\npull 0\n"),
    ("geometry_cylinder_s0p5_ab_q19", 19, "simulationSize 0.5\npropPattern 0\npull 0\n"),
    ("geometry_cylinder_s0p5_pull_q27", 27, "simulationSize 0.5\npropPattern 0\npull 1\n"),
]


def main():
    for name, q, deck in GEOMETRY_CASES:
        ref_geometry_fixture(q, deck, name)
    for k, (name, q, prec, omega, variant, layout, periodic, steps) in enumerate(CASES):
        if not O.ref_available(q, prec, omega):
            raise SystemExit(f"reference build for q={q} {prec} omega={omega} missing: make -C oracle")
        solid = seeded_solid(100 + k)
        geom = O.DenseGeometry(NX, NY, NZ, solid, periodic=periodic, order="reference")
        n = geom.nfluid
        fc = seeded_state(geom, q, 200 + k)
        aa = variant.startswith("aa")
        mode = O.ADJ_PULL if (aa or "pull" in variant) else O.ADJ_PUSH
        adj = geom.adjacency(q, layout, mode)
        dt = np.float64 if prec == "f64" else np.float32
        f0 = O.to_storage(fc, q, layout, aa_even_ready=aa).astype(dt)
        f = f0.copy()
        O.ref_run(variant, q, prec, omega, n, f, adj, steps)
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), q=q, prec=prec, omega=omega, variant=variant,
            layout=layout, periodic=np.array(periodic, np.int32), steps=steps,
            dims=np.array([NX, NY, NZ], np.int32), solid=solid,
            cell_of_fluid=geom.cell_of_fluid, adjacency=adj, f0=f0, f_ref=f)
        print(name, n, "fluid nodes")


if __name__ == "__main__":
    main()
