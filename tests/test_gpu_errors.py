"""Clean failures through the C ABI: an allocation larger than the device
returns LBM_ERR_OOM and leaves the library usable; a NaN or Inf in the
populations makes the FP64 mass reduction return LBM_ERR_NUMERIC; bad
arguments return LBM_ERR_ARG.  The reference has only asserts here
(src/geometries/GeometryAAAB.cpp:86)."""
import ctypes

import numpy as np
import pytest

from lbmminiapp_amd import capi
from lbmminiapp_amd.capi import DenseLattice, LbmError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_oversized_create_is_oom_and_library_stays_usable():
    """1024^3 D3Q19 AB FP64 needs two population arrays of 163 GB each (326 GB,
    more than the device holds): lbm_dense_create returns LBM_ERR_OOM with a
    message, frees what it had allocated, and a small handle then runs and
    matches the oracle."""
    n = 1024
    print("[errors] oversized create", flush=True)
    with pytest.raises(LbmError) as ei:
        DenseLattice(n, n, n, q=19, pattern="ab", storage="f64", omega=1.0)
    assert ei.value.status == capi.ERR_OOM
    assert "hipMalloc" in str(ei.value)
    geom = O.DenseGeometry(16, 8, 8)
    fc = O.equilibrium(19, np.ones(geom.nfluid), np.full((geom.nfluid, 3), 0.01))
    with DenseLattice(16, 8, 8, q=19, pattern="aa", storage="f64", omega=1.2) as lat:
        lat.set_state(fc)
        lat.step(3)
        got = lat.get_state()
    adj = geom.adjacency(19, O.SOA, O.ADJ_PULL)
    f = O.to_storage(fc, 19, O.SOA, aa_even_ready=True)
    O.run_aa(19, O.SOA, geom.nfluid, f, adj, 3, 1.2, (0.0, 0.0, 0.0))
    assert np.array_equal(got, O.aa_canonical(f, 19, geom.nfluid, O.SOA, adj, 3))


def test_oversized_device_alloc_is_oom():
    p = ctypes.c_void_p()
    st = capi.load().lbm_device_alloc(ctypes.c_size_t(1 << 50), ctypes.byref(p))
    assert st == capi.ERR_OOM and not p.value


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
@pytest.mark.parametrize("storage", ["f64", "f32s"])
def test_non_finite_population_is_numeric_error(bad, storage):
    """A non-finite population set through set_state, or carried into the
    state by a time step, makes lbm_dense_mass return LBM_ERR_NUMERIC; a
    finite state set afterwards is accepted again."""
    geom = O.DenseGeometry(32, 8, 6)
    fc = O.equilibrium(19, np.ones(geom.nfluid), np.full((geom.nfluid, 3), 0.01))
    with DenseLattice(32, 8, 6, q=19, pattern="aa", storage=storage, omega=1.2) as lat:
        lat.set_state(fc)
        assert abs(lat.mass() - geom.nfluid) < 1e-9 * geom.nfluid
        bad_state = fc.copy()
        bad_state[5, 123] = bad
        lat.set_state(bad_state)
        with pytest.raises(LbmError) as ei:
            lat.mass()
        assert ei.value.status == capi.ERR_NUMERIC
        lat.step(2)  # the non-finite value spreads through the collision
        with pytest.raises(LbmError) as ei:
            lat.mass()
        assert ei.value.status == capi.ERR_NUMERIC
        lat.set_state(fc)
        assert abs(lat.mass() - geom.nfluid) < 1e-9 * geom.nfluid


def test_bad_arguments_are_arg_errors():
    L = capi.load()
    out = ctypes.c_void_p()
    assert L.lbm_dense_create(None, ctypes.byref(out)) == capi.ERR_ARG
    for kw in ({"q": 15}, {"omega": 2.5}, {"storage": 7}):
        with pytest.raises(LbmError) as ei:
            DenseLattice(8, 8, 8, **kw)
        assert ei.value.status == capi.ERR_ARG
    with DenseLattice(8, 8, 8) as lat:
        assert L.lbm_dense_step(lat.h, -1) == capi.ERR_ARG
        assert L.lbm_dense_graph_create(lat.h, 3, ctypes.byref(out)) == capi.ERR_ARG
