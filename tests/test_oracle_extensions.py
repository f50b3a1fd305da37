"""Closed-form pins of the two per-node extensions of the forward model
(no reference counterpart): relaxation rate omega(x) = 1/tau(x) and Brinkman
velocity scaling u = (1 - gamma(x)) (m + F/2) / rho.

* a uniform tau field equals the scalar omega bit for bit; gamma = 0 equals
  the reference model bit for bit;
* gamma = 1 relaxes every node to the rest equilibrium:
  f* = (1 - w) f + w w_s rho + (1 - w/2) w_s 3 c_s.F;
* uniform gamma in a periodic box with body force F reaches the steady
  momentum m = F (1 - w gamma / 2) / (w gamma)  (from sum_s c_s f*_s =
  (1 - w) m + w rho u + (1 - w/2) F with rho u = (1 - gamma)(m + F/2)).
A dropped Guo term, a wrong sign on gamma or a missing F/2 each change the
steady state, so the last check fails for those mistakes."""
import numpy as np
import pytest

from oracle import oracle as O


def box(q, n=6, seed=2):
    g = O.DenseGeometry(n, n - 1, n - 2)
    cell = g.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.02 * (O.uniform01(seed, 4 * cell) - 0.5)
    u = np.stack([0.04 * (2 * O.uniform01(seed, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    return g, O.equilibrium(q, rho, u)


@pytest.mark.parametrize("q", [19, 27])
def test_uniform_fields_reduce_to_reference(q):
    g, fc = box(q)
    adj = g.adjacency(q, O.SOA, O.ADJ_PUSH)
    f = np.ascontiguousarray(fc.reshape(-1))
    F = (1e-5, 2e-6, 0.0)
    ref = O.run_ab(q, O.SOA, g.nfluid, f, adj, 5, 1 / 0.7, force=F)
    tau = O.run_ab(q, O.SOA, g.nfluid, f, adj, 5, 1.3, force=F,
                   omega_node=np.full(g.nfluid, 1 / 0.7))
    gam = O.run_ab(q, O.SOA, g.nfluid, f, adj, 5, 1 / 0.7, force=F,
                   gamma_node=np.zeros(g.nfluid))
    assert np.array_equal(ref, tau)
    assert np.array_equal(ref, gam)


@pytest.mark.parametrize("q", [19, 27])
def test_gamma_one_relaxes_to_rest(q):
    g, fc = box(q)
    c, opp, w = O.lattice(q)
    n = g.nfluid
    # AB push with a purely bounce-back adjacency is one collision (slots swapped to opp)
    adj = (n * opp[:, None] + np.arange(n)[None, :]).astype(np.int32).reshape(-1)
    omega, F = 1.4, np.array([3e-5, -1e-5, 2e-5])
    out = O.run_ab(q, O.SOA, n, np.ascontiguousarray(fc.reshape(-1)), adj, 1, omega, force=F,
                   gamma_node=np.ones(n)).reshape(q, n)[opp]
    rho = fc.sum(0)
    cf = 3.0 * (c.astype(float) @ F)
    want = (1 - omega) * fc + omega * w[:, None] * rho + (1 - omega / 2) * (w * cf)[:, None]
    assert np.abs(out - want).max() < 1e-15


@pytest.mark.parametrize("gamma", [0.05, 0.3, 0.8])
def test_uniform_brinkman_steady_state(gamma):
    q = 19
    g = O.DenseGeometry(3, 3, 3)
    n = g.nfluid
    adj = g.adjacency(q, O.SOA, O.ADJ_PUSH)
    omega, F = 1.2, np.array([1e-5, -4e-6, 2e-6])
    f0 = O.equilibrium(q, np.ones(n), np.zeros((n, 3)))
    out = O.run_ab(q, O.SOA, n, np.ascontiguousarray(f0.reshape(-1)), adj, 3000, omega, force=F,
                   gamma_node=np.full(n, gamma)).reshape(q, n)
    m = (O.lattice(q)[0].astype(float).T @ out).T
    want = F * (1 - omega * gamma / 2) / (omega * gamma)
    assert np.abs(m - want).max() < 1e-10 * np.abs(want).max()
