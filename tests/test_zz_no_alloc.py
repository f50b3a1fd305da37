"""lbm_sparse_mass_noalloc (the allocation-free FP64 mass reduction, partial
sums in a fixed per-GPU scratch array) against lbm_sparse_mass (the same
kernels with an allocated workspace), through the C ABI: equal bit for bit on
two calls in a row, the second over a smaller range, and within 1e-14 of the
exact sum; a range beyond the scratch's capacity and a host result pointer are
rejected with LBM_ERR_ARG before any launch."""
import ctypes
import math

import numpy as np
import pytest

from lbmminiapp_amd import capi
from lbmminiapp_amd.capi import DeviceBuffer, SparseKernels
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(q=19, nx=48, ny=40, nz=36, seed=5):
    solid = (O.uniform01(seed, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.2).astype(np.uint8)
    geom = O.DenseGeometry(nx, ny, nz, solid)
    cell = geom.cell_of_fluid.astype(np.uint64)
    u = np.stack([0.03 * (2 * O.uniform01(seed + 1, 4 * cell + np.uint64(k)) - 1)
                  for k in range(3)], 1)
    fc = O.equilibrium(q, 1.0 + 0.01 * O.uniform01(seed + 2, cell), u)
    adj = geom.adjacency(q, O.SOA, O.ADJ_PULL)
    return geom, fc, adj


def test_noalloc_mass_equals_allocating_mass_twice():
    q = 19
    geom, fc, adj = _case(q)
    n = geom.nfluid
    f0 = O.to_storage(fc, q, O.SOA, aa_even_ready=True)
    vals = O.from_storage(f0, q, n, O.SOA)
    k = SparseKernels(n, q, adj, 1.2)
    buf = DeviceBuffer.from_array(f0)
    out = DeviceBuffer(8)
    try:
        print("[no-alloc] two calls, the second over a smaller range", flush=True)
        for start, count in ((0, n), (1000, n // 3)):
            want = k.mass(buf, start, count)
            k.mass_noalloc(buf, out, start, count)
            k.synchronize()
            got = float(out.to_array(np.float64)[0])
            exact = math.fsum(vals[:, start:start + count].reshape(-1))
            assert got == want, (start, count, got, want)
            assert abs(got - exact) <= 1e-14 * exact
        # after AA steps (state in the odd-step convention), again equal
        k.aa_even(buf)
        k.mass_noalloc(buf, out)
        k.synchronize()
        assert float(out.to_array(np.float64)[0]) == k.mass(buf)
        # a host result pointer is refused, not written through
        host = ctypes.c_double(-1.0)
        st = capi.load().lbm_sparse_mass_noalloc(k.h, buf.ptr, 0, n, ctypes.byref(host))
        assert st == capi.ERR_ARG and host.value == -1.0
    finally:
        out.free()
        buf.free()
        k.close()


def test_noalloc_mass_rejects_ranges_beyond_the_scratch():
    """A node list of 16,777,217 nodes (one past 4096 chunks of 4096): the
    allocation-free call over the whole list returns LBM_ERR_ARG with a message
    naming the limit, over 16,777,216 nodes it runs and equals lbm_sparse_mass."""
    q, n = 19, 4096 * 4096 + 1
    idx = np.arange(n, dtype=np.int64)
    adj = np.empty(q * n, np.int32)  # every slot streams to itself: a valid adjacency
    for s in range(q):
        adj[s * n:(s + 1) * n] = s * n + idx
    print("[no-alloc] 16,777,217-node list", flush=True)
    k = SparseKernels(n, q, adj, 1.0)
    del adj
    buf = DeviceBuffer(q * n * 8)
    out = DeviceBuffer(8)
    try:
        k.init_rest(buf, aa=True)
        with pytest.raises(capi.LbmError) as ei:
            k.mass_noalloc(buf, out, 0, n)
        assert ei.value.status == capi.ERR_ARG and "16777216" in str(ei.value)
        k.mass_noalloc(buf, out, 1, n - 1)
        k.synchronize()
        got = float(out.to_array(np.float64)[0])
        assert got == k.mass(buf, 1, n - 1)
        assert abs(got - (n - 1)) <= 1e-12 * n
    finally:
        out.free()
        buf.free()
        k.close()
