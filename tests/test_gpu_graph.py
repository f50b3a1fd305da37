"""Streams and graphs of the structured-grid engine: time steps captured as a
HIP graph and replayed on a caller's stream equal the same steps enqueued one
by one, bit for bit, for every replay, with other device work, eager steps and
synchronisations between the replays."""
import ctypes

import numpy as np
import pytest

from lbmminiapp_amd import capi
from lbmminiapp_amd.capi import DenseLattice, DeviceBuffer, SparseKernels
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NX, NY, NZ = 64, 32, 24

CASES = {
    # periodic FP64 AA: the LDS-staged odd kernel on plain rows, and the odd
    # step on interleaved rows
    "aa_f64_staged": dict(pattern="aa", storage="f64", irows="0"),
    "aa_f64_irows": dict(pattern="aa", storage="f64", irows="1", force=(1e-5, 0, 0)),
    "aa_f16s_irows_jump": dict(pattern="aa", storage="f16s", irows="1", jump=2e-4),
    "aa_f32s_masked_force": dict(pattern="aa", storage="f32s", solid=True, force=(1e-5, 0, 0)),
    "ab_pull_f16s_masked": dict(pattern="ab", pull=True, storage="f16s", solid=True),
    "ab_push_f64_masked": dict(pattern="ab", pull=False, storage="f64", solid=True, force=(0, 1e-5, 0)),
    "ab_push_as_gather_f64": dict(pattern="ab", pull=False, storage="f64", solid=True, env="1"),
    "aa_aos_f64_jump": dict(pattern="aa", storage="f64", layout="aos", solid=True, jump=2e-4),
    "q27_aa_f16s": dict(q=27, pattern="aa", storage="f16s"),
    "q27_ab_pull_f64_masked": dict(q=27, pattern="ab", pull=True, storage="f64", solid=True),
}


def _lattice(case):
    solid = None
    if case.get("solid"):
        solid = np.zeros(NX * NY * NZ, np.uint8)
        capi.add_spheres(solid, NX, NY, NZ, 2.0, 4.0, 0.15, 3, periodic=(1, 1, 1))
    lat = DenseLattice(NX, NY, NZ, q=case.get("q", 19), pattern=case["pattern"], pull=case.get("pull", False),
                       layout=case.get("layout", "soa"), storage=case["storage"], omega=1 / 0.7,
                       force=case.get("force", (0.0, 0.0, 0.0)), solid=solid,
                       density_jump_x=case.get("jump", 0.0))
    lat.init_tgv(0.03, 1e-3, 5)
    return lat


class _Stream:
    """A non-blocking stream of the HIP runtime the library is bound to (found by
    its soname once the library is loaded, so that the handle is valid there)."""

    def __init__(self):
        capi.load()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipStreamCreateWithFlags.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        self.hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.s = ctypes.c_void_p()
        assert self.hip.hipStreamCreateWithFlags(ctypes.byref(self.s), 1) == 0  # hipStreamNonBlocking

    def synchronize(self):
        assert self.hip.hipStreamSynchronize(self.s) == 0

    def capture(self, enqueue):
        """Record what enqueue() puts on this stream as an instantiated graph."""
        h = self.hip
        h.hipStreamBeginCapture.argtypes = [ctypes.c_void_p, ctypes.c_int]
        h.hipStreamEndCapture.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.hipGraphInstantiate.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_size_t]
        assert h.hipStreamBeginCapture(self.s, 1) == 0  # hipStreamCaptureModeThreadLocal
        try:
            enqueue()
        finally:
            graph = ctypes.c_void_p()
            assert h.hipStreamEndCapture(self.s, ctypes.byref(graph)) == 0
        exe = ctypes.c_void_p()
        assert h.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, 0) == 0
        return graph, exe

    def launch(self, exe):
        self.hip.hipGraphLaunch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        assert self.hip.hipGraphLaunch(exe, self.s) == 0

    def destroy_graph(self, graph, exe):
        self.hip.hipGraphExecDestroy.argtypes = [ctypes.c_void_p]
        self.hip.hipGraphDestroy.argtypes = [ctypes.c_void_p]
        assert self.hip.hipGraphExecDestroy(exe) == 0
        assert self.hip.hipGraphDestroy(graph) == 0

    def close(self):
        if self.s:
            assert self.hip.hipStreamDestroy(self.s) == 0
            self.s = ctypes.c_void_p()


@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_replays_equal_eager_steps(name, monkeypatch):
    case = CASES[name]
    if "env" in case:
        monkeypatch.setenv("LBM_DENSE_PUSH_GATHER", case["env"])
    if "irows" in case:
        monkeypatch.setenv("LBM_DENSE_IROWS", case["irows"])
    print(f"[graph] {name} {NX}x{NY}x{NZ}", flush=True)
    stream = _Stream()
    try:
        with _lattice(case) as a, _lattice(case) as b:
            # both handles on the caller's stream: a's eager steps are other
            # device work ordered between b's replays
            a.set_stream(stream.s.value)
            b.set_stream(stream.s.value)
            with b.graph(4) as g:
                # capture runs nothing and leaves the handle where it was
                assert b.info()["time_step"] == 0
                assert np.array_equal(b.get_state(), a.get_state())
                for r in range(3):
                    g.launch(1)
                    a.step(4)
                    assert np.array_equal(b.get_state(), a.get_state()), f"replay {r}"
                    b.step(2)
                    a.step(2)
                    stream.synchronize()
                g.launch(2)
                a.step(8)
                assert np.array_equal(b.get_state(), a.get_state())
                assert b.info()["time_step"] == a.info()["time_step"] == 26
                # an odd number of steps away from the capture: refused, nothing enqueued
                b.step(1)
                a.step(1)
                with pytest.raises(capi.LbmError):
                    g.launch(1)
                # a graph captured in the other phase
                with b.graph(2) as g2:
                    g2.launch(3)
                    a.step(6)
                    assert np.array_equal(b.get_state(), a.get_state())
                b.step(1)
                a.step(1)
                g.launch(1)
                a.step(4)
                assert np.array_equal(b.get_state(), a.get_state())
                assert abs(b.mass() - a.mass()) == 0.0
            a.set_stream(None)
            b.set_stream(None)
            b.step(2)
            a.step(2)
            assert np.array_equal(b.get_state(), a.get_state())
    finally:
        stream.close()


def test_graph_rejects_odd_step_counts():
    with _lattice(CASES["aa_f64_staged"]) as lat:
        for n in (0, 1, 3, -2):
            with pytest.raises(capi.LbmError):
                lat.graph(n)


def test_handle_outlives_its_graphs():
    """lbm_dense_destroy refuses a handle whose graphs are alive (they hold its
    device pointers), and succeeds once they are destroyed."""
    lat = _lattice(CASES["aa_f64_staged"])
    g = lat.graph(2)
    with pytest.raises(capi.LbmError):
        lat.close()
    g.launch(1)
    g.close()
    lat.close()


@pytest.mark.parametrize("pattern", ["aa", "pull"])
def test_sparse_graph_replays_equal_oracle(pattern):
    """The drop-in kernels on a caller's stream: two time steps of the
    reference's loop (border and bulk launches per step) captured as a HIP
    graph, replayed with eager steps and synchronisations between the replays;
    every replay equals the oracle bit for bit (D3Q19 FP64 SOA, 10% solid)."""
    nx, ny, nz, q = 24, 20, 16, 19
    solid = (O.uniform01(91, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.1).astype(np.uint8)
    g = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 0, 1), order="reference")
    n = g.nfluid
    cell = g.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.02 * (O.uniform01(9, 4 * cell) - 0.5)
    u = np.stack([0.05 * (2 * O.uniform01(9, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    fc = O.equilibrium(q, rho, u)
    omega = 1 / 0.6
    adj = g.adjacency(q, O.SOA, O.ADJ_PULL)
    ranges = [(0, n // 3), (n // 3, n)]
    print(f"[graph] sparse {pattern} n={n}", flush=True)
    stream = _Stream()
    k = SparseKernels(n, q, adj, omega)
    try:
        k.set_stream(stream.s.value)
        if pattern == "aa":
            f0 = O.to_storage(fc, q, O.SOA, aa_even_ready=True)
            want = f0.copy()
            buf = DeviceBuffer.from_array(f0)

            def two_steps():
                for step in (k.aa_even, k.aa_odd):
                    for a, b in ranges:
                        step(buf, a, b - a)

            def advance(steps):
                O.run_aa(q, O.SOA, n, want, adj, steps, omega)

            def result():
                return buf.to_array(np.float64)
        else:
            f0 = O.to_storage(fc, q, O.SOA)
            state = {"want": f0.copy()}
            src, tgt = DeviceBuffer.from_array(f0), DeviceBuffer.from_array(f0)

            def two_steps():
                for s_, t_ in ((src, tgt), (tgt, src)):
                    for a, b in ranges:
                        k.ab(s_, t_, a, b - a, pull=True)

            def advance(steps):
                state["want"] = O.run_ab(q, O.SOA, n, state["want"], adj, steps, omega, pull=True)

            def result():
                return src.to_array(np.float64)

        graph, exe = stream.capture(two_steps)
        try:
            stream.synchronize()
            # capture ran nothing
            assert np.array_equal(result(), f0)
            for r in range(3):
                stream.launch(exe)
                advance(2)
                k.synchronize()
                assert np.array_equal(result(), want if pattern == "aa" else state["want"]), f"replay {r}"
                two_steps()  # eager steps on the same stream between replays
                advance(2)
                k.synchronize()
        finally:
            stream.destroy_graph(graph, exe)
        k.set_stream(None)
        two_steps()
        advance(2)
        k.synchronize()
        assert np.array_equal(result(), want if pattern == "aa" else state["want"])
    finally:
        k.close()
        stream.close()
