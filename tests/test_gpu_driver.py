"""End to end on the GPU: the deck-driven driver (lbm-proxy-app, the
Main.cpp-style C++ host layer over the C ABI) on the reference's own benchmark
decks (cylinder pipe, simulationSize 2, 153,216 fluid nodes, 1000 steps).
The dumped populations equal the oracle's Base-kernel run on the driver's own
fluid ordering bit for bit, and stay within 1e-12 of the reference's unrolled
AB kernels (the `unroll 1` the decks ask for).  The dense engine on the same
cylinder mask equals the sparse drop-in kernels cell by cell."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "lbmminiapp_amd", "bin", "lbm-proxy-app")
DECKS = os.path.join(ROOT, "tests", "decks")


def read_dump(path):
    raw = open(path, "rb").read()
    assert raw[:4] == b"LBMD"
    version, kind, q, layout, pdf = np.frombuffer(raw[4:24], np.int32)
    n, steps = np.frombuffer(raw[24:40], np.int64)
    off = 40
    out = dict(kind=int(kind), q=int(q), layout=int(layout), n=int(n), steps=int(steps))
    dt = np.float64 if pdf == 8 else np.float32
    if kind in (0, 1):
        out["pts"] = np.frombuffer(raw[off:off + 12 * n], np.int32).reshape(n, 3)
        off += 12 * n
        out["adj"] = np.frombuffer(raw[off:off + 4 * n * q], np.int32)
        off += 4 * n * q
    out["f"] = np.frombuffer(raw[off:off + pdf * n * q], dt)
    return out


def run_driver(deck_text, tmp_path, name):
    deck = tmp_path / f"{name}.txt"
    dump = tmp_path / f"{name}.bin"
    deck.write_text(deck_text + f"\ndump {dump}\n")
    print(f"[driver] {name}", flush=True)
    r = subprocess.run([DRIVER, str(deck)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "MFLUPS" in r.stdout and "Passed baseline validation" in r.stdout
    return read_dump(dump), r.stdout


@pytest.mark.parametrize("deck_name", ["2_input_aa_unroll.txt", "2_input_ab_unroll.txt",
                                       "2_input_ab_pull_unroll.txt"])
def test_reference_deck_bitwise(deck_name, tmp_path):
    text = open(os.path.join(DECKS, deck_name)).read()
    d, stdout = run_driver(text, tmp_path, deck_name[:-4])
    n, q = d["n"], d["q"]
    assert n == 153216 and q == 19 and d["steps"] == 1000
    adj = d["adj"].copy()
    f0 = np.zeros(q * n)
    if d["kind"] == 0:
        O.lib().orc_init_rest_aa_f64(q, O.SOA, n, O._ptr(f0))
        want = f0.copy()
        O.run_aa(q, O.SOA, n, want, adj, 1000, 1.0, (1e-6, 0.0, 0.0))
    else:
        O.lib().orc_init_rest_ab_f64(q, O.SOA, n, O._ptr(f0))
        pull = "pull" in deck_name
        want = O.run_ab(q, O.SOA, n, f0, adj, 1000, 1.0, pull=pull, force=(1e-6, 0.0, 0.0))
        unrolled = O.run_unroll19("ab_pull" if pull else "ab_push", n, f0, adj, 1000, 1.0)
        assert np.abs(d["f"] - unrolled).max() <= 1e-12 * np.abs(unrolled).max()
    assert np.array_equal(d["f"], want)


def test_dense_engine_on_reference_cylinder(tmp_path):
    """`engine dense` on the same cylinder (x-periodic, solid wall cells): the
    structured-grid kernels with the link mask reproduce the sparse kernels'
    populations (AA, even step count) cell by cell, bit for bit."""
    base = ("simulationSize 2\npropPattern 1\npull 0\nunroll 1\ntimesteps 200\n")
    sp, _ = run_driver(base, tmp_path, "sparse")
    de, out = run_driver(base + "engine dense\n", tmp_path, "dense")
    n, q = sp["n"], sp["q"]
    _, opp, _ = O.lattice(q)
    nx, ny = 168, 35
    cell = sp["pts"][:, 0] + nx * (sp["pts"][:, 1] + ny * sp["pts"][:, 2].astype(np.int64))
    sparse_canon = O.aa_canonical(sp["f"], q, n, O.SOA, sp["adj"], 200)
    dense = de["f"].reshape(q, -1)
    assert np.array_equal(dense[:, cell], sparse_canon)


def read_vtk(path):
    """Legacy VTK (ASCII or BINARY) as written by the driver: dims and the rho,
    id, velocity and nodes arrays, read in file order."""
    raw = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = raw.index(b"\n", pos)
        text = raw[pos:end]
        pos = end + 1
        return text

    hdr = [line() for _ in range(8)]
    binary = hdr[2].strip() == b"BINARY"
    dims = tuple(int(v) for v in hdr[4].split()[1:4])
    npts = int(np.prod(dims))
    assert hdr[7].split()[1] == str(npts).encode()
    out = {"dims": dims, "binary": binary}
    for name, comp, dt in (("rho", 1, "f8"), ("id", 1, "i4"), ("velocity", 3, "f8"),
                           ("nodes", 3, "i4")):
        h = line()
        assert h.split()[1].decode() == name, h
        if h.startswith(b"SCALARS"):
            assert line().startswith(b"LOOKUP_TABLE")
        if binary:
            nbytes = npts * comp * int(dt[1])
            arr = np.frombuffer(raw[pos:pos + nbytes], ">" + dt)
            pos += nbytes
            assert line() == b""
        else:
            arr = np.array(b" ".join(line() for _ in range(npts)).split(), dtype=dt)
        out[name] = arr.reshape(npts, comp).squeeze()
    return out


def test_vtk_binary_matches_ascii(tmp_path):
    """writevtk 2 (legacy binary) holds the same fields as writevtk 1 (the
    reference's ASCII format): ids and coordinates equal, rho and velocity
    equal to the 6 significant digits the ASCII stream prints."""
    base = "simulationSize 1\npropPattern 1\npull 0\nunroll 1\ntimesteps 50\n"
    res = {}
    for mode in (1, 2):
        d = tmp_path / f"vtk{mode}"
        d.mkdir()
        (d / "deck.txt").write_text(base + f"writevtk {mode}\n")
        print(f"[driver] writevtk {mode}", flush=True)
        r = subprocess.run([DRIVER, "deck.txt"], cwd=d, capture_output=True, text=True,
                           timeout=240)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        res[mode] = read_vtk(d / "outs.00000.vtk")
    a, b = res[1], res[2]
    assert not a["binary"] and b["binary"] and a["dims"] == b["dims"]
    assert np.array_equal(a["id"], b["id"]) and np.array_equal(a["nodes"], b["nodes"])
    fluid = b["id"] >= 0
    assert fluid.sum() > 0 and np.abs(b["rho"][fluid] - 1).max() < 1e-2
    assert np.allclose(a["rho"], b["rho"], rtol=1e-5, atol=0)
    assert np.allclose(a["velocity"], b["velocity"], rtol=1e-5, atol=1e-12)
    assert np.abs(b["velocity"][fluid, 0]).max() > 0


def _run_deck(text, tmp_path, name):
    deck = tmp_path / f"{name}.txt"
    dump = tmp_path / f"{name}.bin"
    deck.write_text(text + f"\ndump {dump}\n")
    print(f"[driver] {name}", flush=True)
    r = subprocess.run([DRIVER, str(deck)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return read_dump(dump)


@pytest.mark.parametrize("engine", ["sparse", "dense"])
def test_mask_file_reproduces_generated_geometry(engine, tmp_path):
    """`writemask` stores the solid mask a run used; a deck that reads it back
    (`maskfile`, dimensions from the file) runs the same fluid nodes, adjacency
    and populations bit for bit.  The file holds the generator's mask (seeded
    sphere pack), cell x + nx*(y + ny*z)."""
    from lbmminiapp_amd import capi

    nx, ny, nz = 48, 40, 32
    physics = "timesteps 40\npropPattern 0\npull 0\nperiodic 1 1 1\n" + (
        "engine dense\n" if engine == "dense" else "")
    mask = tmp_path / "porous.mask"
    gen = _run_deck(physics + f"geometry porous\nsize {nx} {ny} {nz}\nsolidfraction 0.2\n"
                              f"radiusmin 3\nradiusmax 6\nseed 9\nwritemask {mask}\n",
                    tmp_path, f"gen_{engine}")
    dims, solid = capi.read_mask(mask)
    want = np.zeros(nx * ny * nz, np.uint8)
    capi.add_spheres(want, nx, ny, nz, 3.0, 6.0, 0.2, 9, periodic=(1, 1, 1))
    assert dims == (nx, ny, nz) and np.array_equal(solid, (want != 0).astype(np.uint8))
    assert 0.15 < solid.mean() < 0.3
    again = _run_deck(physics + f"maskfile {mask}\n", tmp_path, f"file_{engine}")
    assert again["n"] == gen["n"] and again["kind"] == gen["kind"]
    if "adj" in gen:
        assert np.array_equal(again["adj"], gen["adj"])
        assert np.array_equal(again["pts"], gen["pts"])
    assert np.array_equal(again["f"], gen["f"])


def _vtk_origin(path):
    with open(path, "rb") as fh:
        for _ in range(6):
            ln = fh.readline()
    assert ln.startswith(b"ORIGIN")
    return tuple(int(v) for v in ln.split()[1:4])


@pytest.mark.parametrize("pattern,steps", [("aa", 41), ("aa", 42), ("push", 41), ("push", 42),
                                           ("pull", 41), ("pull", 42)])
def test_readout_vtk_vs_oracle_moments(pattern, steps, tmp_path):
    """The host readout behind writeVTK (PropagationAA::getRho/getVelocity,
    parity-aware: slot opp(s) after even step counts, through the adjacency
    after odd ones; PropagationAB: local slots of the current array) on the
    cylinder deck (simulationSize 1), after odd and even step counts.  Every
    fluid point's rho and velocity in the binary VTK equal the oracle's
    moments (reference operation order) of the oracle's own state, bit for
    bit; solid points hold zeros."""
    pp = {"aa": (1, 0), "push": (0, 0), "pull": (0, 1)}[pattern]
    deck = (f"simulationSize 1\npropPattern {pp[0]}\npull {pp[1]}\nunroll 1\n"
            f"timesteps {steps}\nwritevtk 2\n")
    (tmp_path / "deck.txt").write_text(deck + f"dump {tmp_path / 'd.bin'}\n")
    print(f"[driver] readout {pattern} {steps}", flush=True)
    r = subprocess.run([DRIVER, "deck.txt"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    d = read_dump(tmp_path / "d.bin")
    vtk = read_vtk(tmp_path / "outs.00000.vtk")
    x0, y0, z0 = _vtk_origin(tmp_path / "outs.00000.vtk")
    n, q = d["n"], d["q"]
    adj = d["adj"].copy()
    f0 = np.zeros(q * n)
    if pattern == "aa":
        O.lib().orc_init_rest_aa_f64(q, O.SOA, n, O._ptr(f0))
        want = f0.copy()
        O.run_aa(q, O.SOA, n, want, adj, steps, 1.0, (1e-6, 0.0, 0.0))
        canon = O.aa_canonical(want, q, n, O.SOA, adj, steps)
    else:
        O.lib().orc_init_rest_ab_f64(q, O.SOA, n, O._ptr(f0))
        want = O.run_ab(q, O.SOA, n, f0, adj, steps, 1.0, pull=pattern == "pull",
                        force=(1e-6, 0.0, 0.0))
        canon = O.from_storage(want, q, n, O.SOA)  # local slots, as the AB readout reads them
    assert np.array_equal(d["f"], want)
    rho, u = O.moments(q, canon)
    nx, ny, _ = vtk["dims"]
    pts = d["pts"].astype(np.int64)
    idx = (pts[:, 0] - x0) + nx * ((pts[:, 1] - y0) + ny * (pts[:, 2] - z0))
    assert np.array_equal(vtk["id"][idx], np.arange(n))
    assert np.array_equal(vtk["rho"][idx], rho)
    assert np.array_equal(vtk["velocity"][idx], u)
    solid = vtk["id"] < 0
    assert solid.any() and not vtk["rho"][solid].any() and not vtk["velocity"][solid].any()
    assert np.abs(u[:, 0]).max() > 0
