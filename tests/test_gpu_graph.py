"""Streams and graphs of the structured-grid engine: time steps captured as a
HIP graph and replayed on a caller's stream equal the same steps enqueued one
by one, bit for bit, for every replay, with other device work, eager steps and
synchronisations between the replays."""
import ctypes

import numpy as np
import pytest

from lbmminiapp_amd import capi
from lbmminiapp_amd.capi import DenseLattice

pytestmark = pytest.mark.gpu

NX, NY, NZ = 64, 32, 24

CASES = {
    # periodic FP64 AA: the LDS-staged odd kernel
    "aa_f64_staged": dict(pattern="aa", storage="f64"),
    "aa_f32s_masked_force": dict(pattern="aa", storage="f32s", solid=True, force=(1e-5, 0, 0)),
    "ab_pull_f16s_masked": dict(pattern="ab", pull=True, storage="f16s", solid=True),
    "ab_push_f64_masked": dict(pattern="ab", pull=False, storage="f64", solid=True, force=(0, 1e-5, 0)),
    "ab_push_as_gather_f64": dict(pattern="ab", pull=False, storage="f64", solid=True, env="1"),
    "aa_aos_f64_jump": dict(pattern="aa", storage="f64", layout="aos", solid=True, jump=2e-4),
}


def _lattice(case):
    solid = None
    if case.get("solid"):
        solid = np.zeros(NX * NY * NZ, np.uint8)
        capi.add_spheres(solid, NX, NY, NZ, 2.0, 4.0, 0.15, 3, periodic=(1, 1, 1))
    lat = DenseLattice(NX, NY, NZ, q=19, pattern=case["pattern"], pull=case.get("pull", False),
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

    def close(self):
        if self.s:
            assert self.hip.hipStreamDestroy(self.s) == 0
            self.s = ctypes.c_void_p()


@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_replays_equal_eager_steps(name, monkeypatch):
    case = CASES[name]
    if "env" in case:
        monkeypatch.setenv("LBM_DENSE_PUSH_GATHER", case["env"])
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
