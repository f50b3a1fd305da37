"""The C oracle against the golden fixtures written by the reference's own
kernels (tests/golden/gen_golden.py): bit-for-bit equality for every variant
(AA / AB push / AB pull, SOA / AOS, Base / unrolled, Q19 / Q27, FP64 / FP32)."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith("geometry_"))


def load_case(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


def run_oracle_variant(c):
    q = int(c["q"])
    variant = str(c["variant"])
    layout = int(c["layout"])
    omega = float(c["omega"])
    steps = int(c["steps"])
    adj = c["adjacency"]
    f = c["f0"].copy()
    n = f.size // q
    force = (1e-6, 0.0, 0.0)
    if variant in ("aa_soa_base", "aa_aos_base"):
        O.run_aa(q, layout, n, f, adj, steps, omega, force)
        return f
    if variant == "aa_soa_unroll":
        return O.run_unroll19("aa", n, f, adj, steps, omega)
    if variant in ("ab_soa_base", "ab_aos_base"):
        return O.run_ab(q, layout, n, f, adj, steps, omega, pull=False, force=force)
    if variant == "ab_soa_pull":
        return O.run_ab(q, layout, n, f, adj, steps, omega, pull=True, force=force)
    if variant == "ab_soa_unroll":
        return O.run_unroll19("ab_push", n, f, adj, steps, omega)
    if variant == "ab_soa_unroll_pull":
        return O.run_unroll19("ab_pull", n, f, adj, steps, omega)
    raise ValueError(variant)


def test_golden_present():
    assert len(GOLDEN) >= 12


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_fixture(path):
    c = load_case(path)
    out = run_oracle_variant(c)
    assert out.dtype == c["f_ref"].dtype
    assert np.array_equal(out, c["f_ref"]), "oracle differs from the reference kernels' output"


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_fixture_adjacency_rebuilds(path):
    """The fixture's adjacency equals orc_build_adjacency on the stored mask
    (the GeometryAAAB::setupAdjacency rule with the stored periodic flags)."""
    c = load_case(path)
    nx, ny, nz = (int(v) for v in c["dims"])
    g = O.DenseGeometry(nx, ny, nz, c["solid"], periodic=tuple(c["periodic"]), order="reference")
    assert np.array_equal(g.cell_of_fluid, c["cell_of_fluid"])
    variant = str(c["variant"])
    mode = O.ADJ_PULL if (variant.startswith("aa") or "pull" in variant) else O.ADJ_PUSH
    assert np.array_equal(g.adjacency(int(c["q"]), int(c["layout"]), mode), c["adjacency"])
