"""GPU parity of the drop-in sparse kernels (lbm_sparse_*, through the C ABI)
with the reference: bit-for-bit against the golden fixtures written by the
reference's own Base kernels, and against the oracle on larger random cases
(random solid masks, partially periodic boxes, border/bulk split launches)."""
import glob
import os

import numpy as np
import pytest

from lbmminiapp_amd.capi import DeviceBuffer, SparseKernels
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith("geometry_"))


def _case(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


def run_gpu(q, layout, prec, omega, adj, f0, steps, pattern, splits=None, pull=False,
            adjacency_mode=1):
    n = f0.size // q
    k = SparseKernels(n, q, adj, omega, layout=layout, storage=prec, adjacency_mode=adjacency_mode)
    ranges = [(0, n)] if splits is None else list(zip([0] + splits, splits + [n]))
    try:
        if pattern == "aa":
            buf = DeviceBuffer.from_array(f0)
            for t in range(steps):
                for a, b in ranges:
                    (k.aa_even if t % 2 == 0 else k.aa_odd)(buf, a, b - a)
            k.synchronize()
            return buf.to_array(f0.dtype)
        src, tgt = DeviceBuffer.from_array(f0), DeviceBuffer.from_array(f0)
        for _ in range(steps):
            for a, b in ranges:
                k.ab(src, tgt, a, b - a, pull=pull)
            src, tgt = tgt, src
        k.synchronize()
        return src.to_array(f0.dtype)
    finally:
        k.close()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_sparse_matches_reference_fixture(path):
    c = _case(path)
    q, layout, omega, steps = int(c["q"]), int(c["layout"]), float(c["omega"]), int(c["steps"])
    variant = str(c["variant"])
    prec = str(c["prec"])
    print(f"[sparse] {os.path.basename(path)}", flush=True)
    pattern = "aa" if variant.startswith("aa") else "ab"
    pull = "pull" in variant
    out = run_gpu(q, "soa" if layout == 0 else "aos", prec, omega, c["adjacency"], c["f0"], steps,
                  pattern, pull=pull, adjacency_mode=1 if (pattern == "aa" or pull) else 0)
    ref = c["f_ref"]
    if "unroll" not in variant:
        assert np.array_equal(out, ref), "GPU differs from the reference Base kernel bit for bit"
        return
    # unrolled reference kernels: same update, different association (and the
    # AA kernel's rest-population force weight W1 instead of W2).  The GPU
    # follows the Base arithmetic: equal to the oracle Base run bit for bit,
    # and within 1e-12 relative of the unrolled reference for AB.
    n = out.size // q
    base = c["f0"].copy()
    if pattern == "aa":
        O.run_aa(q, layout, n, base, c["adjacency"], steps, omega)
    else:
        base = O.run_ab(q, layout, n, base, c["adjacency"], steps, omega, pull=pull)
    assert np.array_equal(out, base)
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    if pattern == "ab":
        assert rel < 1e-12
    else:
        assert 0 < rel < 1e-6  # the documented AA-unroll rest-weight deviation


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("layout", ["soa", "aos"])
def test_sparse_random_case_vs_oracle(q, prec, layout):
    nx, ny, nz = 24, 20, 16
    solid = (O.uniform01(77, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.1).astype(np.uint8)
    g = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 0, 1), order="reference")
    n = g.nfluid
    cell = g.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.02 * (O.uniform01(5, 4 * cell) - 0.5)
    u = np.stack([0.05 * (2 * O.uniform01(5, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    fc = O.equilibrium(q, rho, u)
    L = O.SOA if layout == "soa" else O.AOS
    dt = np.float64 if prec == "f64" else np.float32
    omega = 1 / 0.6
    print(f"[sparse] random q={q} {prec} {layout}", flush=True)
    # AA with a border/bulk split (two launches per step, as PropagationAA::run)
    adj = g.adjacency(q, L, O.ADJ_PULL)
    f0 = O.to_storage(fc, q, L, aa_even_ready=True).astype(dt)
    want = f0.copy()
    O.run_aa(q, L, n, want, adj, 11, omega)
    got = run_gpu(q, layout, prec, omega, adj, f0, 11, "aa", splits=[n // 3])
    assert np.array_equal(got, want)
    # AB push
    adjp = g.adjacency(q, L, O.ADJ_PUSH)
    f0 = O.to_storage(fc, q, L).astype(dt)
    want = O.run_ab(q, L, n, f0, adjp, 7, omega)
    got = run_gpu(q, layout, prec, omega, adjp, f0, 7, "ab", adjacency_mode=0)
    assert np.array_equal(got, want)
    if layout == "soa":
        want = O.run_ab(q, L, n, f0, adj, 7, omega, pull=True)
        got = run_gpu(q, layout, prec, omega, adj, f0, 7, "ab", pull=True, splits=[n // 2])
        assert np.array_equal(got, want)


def test_sparse_init_rest_matches_reference_initialisation():
    g = O.DenseGeometry(6, 5, 4)
    for q in (19, 27):
        adj = g.adjacency(q, O.SOA, O.ADJ_PULL)
        k = SparseKernels(g.nfluid, q, adj, 1.0)
        for aa in (True, False):
            buf = DeviceBuffer(8 * q * g.nfluid)
            k.init_rest(buf, aa)
            k.synchronize()
            got = buf.to_array(np.float64)
            want = np.zeros(q * g.nfluid)
            fn = O.lib().orc_init_rest_aa_f64 if aa else O.lib().orc_init_rest_ab_f64
            fn(q, O.SOA, g.nfluid, O._ptr(want))
            assert np.array_equal(got, want)
        k.close()


@pytest.mark.parametrize("xcd", ["0", "3", "16", None])
def test_sparse_xcd_block_mapping_vs_oracle(xcd, monkeypatch):
    """Node blocks remapped to runs per XCD (LBM_SPARSE_XCD: off, runs of 3
    and 16 blocks, and the default by footprint) on a grid of 240 blocks with
    a border/bulk split: AA, AB push and AB pull equal the oracle bit for bit
    (every node still updated exactly once)."""
    if xcd is None:
        monkeypatch.delenv("LBM_SPARSE_XCD", raising=False)
    else:
        monkeypatch.setenv("LBM_SPARSE_XCD", xcd)
    nx, ny, nz, q = 48, 40, 32, 19
    solid = (O.uniform01(33, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.05).astype(np.uint8)
    g = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 0, 1), order="reference")
    n = g.nfluid
    cell = g.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.02 * (O.uniform01(6, 4 * cell) - 0.5)
    u = np.stack([0.05 * (2 * O.uniform01(6, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    fc = O.equilibrium(q, rho, u)
    omega = 1 / 0.6
    print(f"[sparse] xcd={xcd} n={n}", flush=True)
    adj = g.adjacency(q, O.SOA, O.ADJ_PULL)
    f0 = O.to_storage(fc, q, O.SOA, aa_even_ready=True)
    want = f0.copy()
    O.run_aa(q, O.SOA, n, want, adj, 4, omega)
    got = run_gpu(q, "soa", "f64", omega, adj, f0, 4, "aa", splits=[n // 7])
    assert np.array_equal(got, want)
    adjp = g.adjacency(q, O.SOA, O.ADJ_PUSH)
    f0 = O.to_storage(fc, q, O.SOA)
    want = O.run_ab(q, O.SOA, n, f0, adjp, 3, omega)
    got = run_gpu(q, "soa", "f64", omega, adjp, f0, 3, "ab", adjacency_mode=0, splits=[n // 5])
    assert np.array_equal(got, want)
    want = O.run_ab(q, O.SOA, n, f0, adj, 3, omega, pull=True)
    got = run_gpu(q, "soa", "f64", omega, adj, f0, 3, "ab", pull=True)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("prec,layout", [("f32", "soa"), ("f64", "aos"), ("f32", "aos")])
def test_sparse_xcd_block_mapping_other_storage(prec, layout, monkeypatch):
    """The XCD node-block runs (LBM_SPARSE_XCD=5, a run length that does not
    divide the 240 blocks) with FP32 storage and the AOS layout: AA and AB
    push equal the oracle bit for bit (AOS has no pull kernel in the
    reference)."""
    monkeypatch.setenv("LBM_SPARSE_XCD", "5")
    nx, ny, nz, q = 48, 40, 32, 19
    solid = (O.uniform01(35, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.05).astype(np.uint8)
    g = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 0, 1), order="reference")
    n = g.nfluid
    cell = g.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.02 * (O.uniform01(8, 4 * cell) - 0.5)
    u = np.stack([0.05 * (2 * O.uniform01(8, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    fc = O.equilibrium(q, rho, u)
    L = O.SOA if layout == "soa" else O.AOS
    dt = np.float64 if prec == "f64" else np.float32
    omega = 1 / 0.7
    print(f"[sparse] xcd runs {prec} {layout} n={n}", flush=True)
    adj = g.adjacency(q, L, O.ADJ_PULL)
    f0 = O.to_storage(fc, q, L, aa_even_ready=True).astype(dt)
    want = f0.copy()
    O.run_aa(q, L, n, want, adj, 4, omega)
    got = run_gpu(q, layout, prec, omega, adj, f0, 4, "aa", splits=[n // 3])
    assert np.array_equal(got, want)
    adjp = g.adjacency(q, L, O.ADJ_PUSH)
    f0 = O.to_storage(fc, q, L).astype(dt)
    want = O.run_ab(q, L, n, f0, adjp, 3, omega)
    got = run_gpu(q, layout, prec, omega, adjp, f0, 3, "ab", adjacency_mode=0)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("q,layout,prec", [(19, "soa", "f64"), (27, "aos", "f64"),
                                           (19, "soa", "f32")])
def test_sparse_mass_reduction(q, layout, prec):
    """lbm_sparse_mass: the FP64 sum of the stored values of a fluid range,
    against the exact (math.fsum) sum of the same values; bit-identical on a
    repeated call; unchanged to round-off by AA steps (mass conservation);
    LBM_ERR_NUMERIC on a NaN; ranges longer than one partial-sum chunk and
    ragged ones."""
    import math

    from lbmminiapp_amd import capi

    geom = O.DenseGeometry(40, 30, 20, (O.uniform01(5, np.arange(24000, dtype=np.uint64)) < 0.2)
                           .astype(np.uint8))
    n = geom.nfluid
    cell = geom.cell_of_fluid.astype(np.uint64)
    u = np.stack([0.03 * (2 * O.uniform01(6, 4 * cell + np.uint64(k)) - 1) for k in range(3)], 1)
    fc = O.equilibrium(q, 1.0 + 0.01 * O.uniform01(7, cell), u)
    lay = O.SOA if layout == "soa" else O.AOS
    adj = geom.adjacency(q, lay, O.ADJ_PULL)
    dt = np.float64 if prec == "f64" else np.float32
    f0 = O.to_storage(fc, q, lay, aa_even_ready=True).astype(dt)
    k = SparseKernels(n, q, adj, 1.2, layout=layout, storage=prec)
    try:
        buf = DeviceBuffer.from_array(f0)
        vals = O.from_storage(f0.astype(np.float64), q, n, lay)
        for start, count in ((0, n), (0, 4096), (1234, 9000), (n - 17, 17), (3, 0)):
            m = k.mass(buf, start, count)
            exact = math.fsum(vals[:, start:start + count].reshape(-1))
            assert abs(m - exact) <= 1e-14 * max(abs(exact), 1.0), (start, count, m, exact)
            assert k.mass(buf, start, count) == m
        m0 = k.mass(buf)
        for t in range(6):
            (k.aa_even if t % 2 == 0 else k.aa_odd)(buf)
        m1 = k.mass(buf)
        assert abs(m1 - m0) <= (1e-13 if prec == "f64" else 1e-5) * m0
        bad = f0.copy()
        bad[7] = np.nan
        buf.upload(bad)
        with pytest.raises(capi.LbmError) as ei:
            k.mass(buf)
        assert ei.value.status == capi.ERR_NUMERIC
        buf.free()
    finally:
        k.close()
