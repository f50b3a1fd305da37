"""The C++ host layer's geometry (GeometryAAAB over the engine's generators)
against the reference's own GeometryAAAB pipeline (fixtures written by
tests/golden/gen_golden.py from the reference sources): same fluid points in
the same order, same border/bulk ranges, same adjacency, bit for bit.  Also
the deck parser's error behaviour.  No GPU: the driver's --dump-geometry mode
stops before any device call."""
import glob
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "lbmminiapp_amd", "bin", "lbm-proxy-app")
GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "geometry_*.npz")))


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        from lbmminiapp_amd import build
        build.build()
    return DRIVER


def read_geometry_dump(path):
    raw = open(path, "rb").read()
    assert raw[:4] == b"LBMG"
    version, q, layout, mode = np.frombuffer(raw[4:20], np.int32)
    n = int(np.frombuffer(raw[20:28], np.int64)[0])
    dims = np.frombuffer(raw[28:40], np.int32)
    ranges = np.frombuffer(raw[40:56], np.int32)
    off = 56
    pts = np.frombuffer(raw[off:off + 12 * n], np.int32).reshape(n, 3)
    off += 12 * n
    adj = np.frombuffer(raw[off:off + 4 * n * q], np.int32)
    return dict(q=int(q), layout=int(layout), mode=int(mode), n=n, dims=dims, ranges=ranges,
                pts=pts, adj=adj)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_cylinder_geometry_matches_reference(driver, path, tmp_path):
    z = np.load(path, allow_pickle=False)
    q = int(z["q"])
    deck = tmp_path / "deck.txt"
    deck.write_text(str(z["deck"]) + f"lattice {q}\n")
    out = tmp_path / "geom.bin"
    r = subprocess.run([driver, str(deck), "--dump-geometry", str(out)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    g = read_geometry_dump(out)
    assert g["q"] == q
    assert np.array_equal(g["pts"], z["fluid_pts"])
    assert np.array_equal(g["ranges"], z["ranges"])
    assert np.array_equal(g["adj"], z["adjacency"])


def test_reference_deck_geometry_s2(driver, tmp_path):
    """The reference's own benchmark deck (simulationSize 2): 153,216 fluid
    nodes, the count its binary prints; adjacency equals the oracle's restated
    rule on the same fluid ordering."""
    from oracle import oracle as O
    deck = os.path.join(ROOT, "tests", "decks", "2_input_aa_unroll.txt")
    out = tmp_path / "geom.bin"
    r = subprocess.run([driver, deck, "--dump-geometry", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    g = read_geometry_dump(out)
    assert g["n"] == 153216
    nx, ny, nz = (int(d) for d in g["dims"])
    assert (nx, ny, nz) == (168, 35, 35)
    cell = g["pts"][:, 0] + nx * (g["pts"][:, 1] + ny * g["pts"][:, 2].astype(np.int64))
    solid = np.ones(nx * ny * nz, np.uint8)
    solid[cell] = 0
    geom = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 0, 0))
    geom.cell_of_fluid = cell.astype(np.int32)
    geom.fluid_of_cell = np.full(nx * ny * nz, -1, np.int32)
    geom.fluid_of_cell[cell] = np.arange(g["n"], dtype=np.int32)
    assert np.array_equal(g["adj"], geom.adjacency(19, O.SOA, O.ADJ_PULL))


def test_deck_errors(driver, tmp_path):
    r = subprocess.run([driver, str(tmp_path / "missing.txt")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open input deck" in r.stderr
    bad = tmp_path / "bad.txt"
    bad.write_text("lattice 21\n")
    r = subprocess.run([driver, str(bad), "--dump-geometry", str(tmp_path / "g.bin")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "lattice" in r.stderr
    warn = tmp_path / "warn.txt"
    warn.write_text("simulationSize 0.5\nfrobnicate 3\n")
    r = subprocess.run([driver, str(warn), "--dump-geometry", str(tmp_path / "g.bin")],
                       capture_output=True, text=True)
    assert r.returncode == 0 and "unknown deck key 'frobnicate'" in r.stderr


def test_deck_rejects_unimplemented_pattern(driver, tmp_path):
    """propPattern 2 (MR) is named in the reference decks' comments but has no
    implementation (its Main.cpp leaves the propagation unset): an error here."""
    deck = tmp_path / "mr.txt"
    deck.write_text("simulationSize 0.5\npropPattern 2\n")
    r = subprocess.run([driver, str(deck), "--dump-geometry", str(tmp_path / "g.bin")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "propPattern" in r.stderr


def test_mask_file_geometry(driver, tmp_path):
    """`maskfile`: the fluid points are exactly the mask's fluid cells, in the
    reference's order (border first, then bulk, x-major within each); the
    dimensions come from the file; `writemask` writes the mask back unchanged;
    a file with another magic is an error."""
    from lbmminiapp_amd import capi

    nx, ny, nz = 12, 9, 7
    cell = np.arange(nx * ny * nz)
    x, y, z = cell % nx, (cell // nx) % ny, cell // (nx * ny)
    solid = (((x - 6) ** 2 + (y - 4) ** 2 < 6) | (z == 3) & (x < 4)).astype(np.uint8)
    mask = tmp_path / "m.mask"
    capi.write_mask(mask, solid, (nx, ny, nz))
    back = tmp_path / "back.mask"
    deck = tmp_path / "file.txt"
    deck.write_text(f"maskfile {mask}\nwritemask {back}\nperiodic 1 0 1\npropPattern 1\n")
    out = tmp_path / "geom.bin"
    r = subprocess.run([driver, str(deck), "--dump-geometry", str(out)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    g = read_geometry_dump(out)
    assert tuple(int(d) for d in g["dims"]) == (nx, ny, nz)
    fluid = set(np.nonzero(solid == 0)[0].tolist())
    got = g["pts"][:, 0] + nx * (g["pts"][:, 1] + ny * g["pts"][:, 2])
    assert g["n"] == len(fluid) and set(got.tolist()) == fluid
    assert int(g["ranges"][0]) == 0  # border start; ranges = border start, count, bulk start, count
    nb = int(g["ranges"][1])
    pts = g["pts"]
    on_face = ((pts[:, 0] == 0) | (pts[:, 0] == nx - 1) | (pts[:, 1] == 0) | (pts[:, 1] == ny - 1)
               | (pts[:, 2] == 0) | (pts[:, 2] == nz - 1))
    assert on_face[:nb].all() and not on_face[nb:].any()
    for part in (pts[:nb], pts[nb:]):
        key = part[:, 0] * ny * nz + part[:, 1] * nz + part[:, 2]
        assert np.all(np.diff(key) > 0)
    back_dims, back_solid = capi.read_mask(back)
    assert back_dims == (nx, ny, nz) and np.array_equal(back_solid, solid)
    bad = tmp_path / "bad.mask"
    bad.write_bytes(b"NOTAMASK" + mask.read_bytes()[8:])
    deck.write_text(f"maskfile {bad}\n")
    r = subprocess.run([driver, str(deck), "--dump-geometry", str(out)], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "not a mask file" in r.stderr
