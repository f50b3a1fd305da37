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


# ---------------------------------------------------------------- pressure drop
def _channel(nx, ny, nz=1):
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    return O.DenseGeometry(nx, ny, nz, solid.reshape(-1), periodic=(1, 0, 1))


@pytest.mark.parametrize("pattern", ["aa", "push", "pull"])
def test_pressure_jump_conserves_mass(pattern):
    """Every forward crossing (+w_s jump) is paired with the backward crossing
    of the opposite direction on the same link (-w_s jump): the total mass is
    unchanged to round-off, on a channel with walls and random solid cells."""
    q = 19
    nx, ny, nz = 6, 7, 5
    solid = (O.uniform01(3, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.15).astype(np.uint8)
    g = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 0, 1))
    cell = g.cell_of_fluid.astype(np.uint64)
    u = np.stack([0.03 * (2 * O.uniform01(9, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    fc = O.equilibrium(q, np.ones(g.nfluid), u)
    m0 = fc.sum()
    jump = 3e-3
    if pattern == "aa":
        adj = g.adjacency(q, O.SOA, O.ADJ_PULL)
        f = O.to_storage(fc, q, O.SOA, aa_even_ready=True)
        O.run_aa_jump(q, O.SOA, g, f, adj, 9, 1 / 0.7, jump)
        m1 = O.gather_canonical_jump(f, q, g, O.SOA, adj, jump).sum()
    elif pattern == "push":
        adj = g.adjacency(q, O.SOA, O.ADJ_PUSH)
        f = O.run_ab_jump(q, O.SOA, g, np.ascontiguousarray(fc.reshape(-1)), adj, 9, 1 / 0.7, jump)
        m1 = f.sum()
    else:
        adj = g.adjacency(q, O.SOA, O.ADJ_PULL)
        f = O.run_ab_jump(q, O.SOA, g, np.ascontiguousarray(fc.reshape(-1)), adj, 9, 1 / 0.7, jump,
                          pull=True)
        m1 = O.gather_canonical_jump(f, q, g, O.SOA, adj, jump).sum()
    assert abs(m1 - m0) <= 1e-13 * m0


def _steady_centre_velocity(jump, force, sign_flip=False, no_weight=False):
    q, nx, ny = 19, 8, 12
    g = _channel(nx, ny)
    adj = g.adjacency(q, O.SOA, O.ADJ_PULL)
    f = np.ascontiguousarray(O.equilibrium(q, np.ones(g.nfluid), np.zeros((g.nfluid, 3))).reshape(-1))
    j = -jump if sign_flip else jump
    if no_weight:  # planted mistake: the jump without the lattice weight
        _, _, w = O.lattice(q)
        j = jump / w[0]
    f = O.run_ab_jump(q, O.SOA, g, f, adj, 4000, 1 / 0.8, j, pull=True, force=force)
    fcan = O.gather_canonical_jump(f, q, g, O.SOA, adj, j)
    c = O.lattice(q)[0].astype(float)
    ux = (c[:, 0] @ fcan) / fcan.sum(0)
    x, y, _ = g.coords()
    return ux[(y == ny // 2) & (x == 3)].mean()


def test_pressure_jump_drives_the_body_force_flow():
    """A channel (walls y = 0, 11) driven by the density jump drho across the
    periodic x boundary reaches the steady velocity of the same channel driven
    by the equivalent body force F = (drho / 3) / nx (pressure drop over the
    length), within 1%; a flipped sign reverses the flow and a jump without
    the weight w_s is off by an order of magnitude."""
    drho, nx = 3e-5, 8
    u_jump = _steady_centre_velocity(drho, (0.0, 0.0, 0.0))
    u_force = _steady_centre_velocity(0.0, (drho / 3 / nx, 0.0, 0.0))
    assert u_force > 0 and abs(u_jump - u_force) < 1e-2 * u_force, (u_jump, u_force)
    assert _steady_centre_velocity(drho, (0.0, 0.0, 0.0), sign_flip=True) < 0
    u_bad = _steady_centre_velocity(drho, (0.0, 0.0, 0.0), no_weight=True)
    assert abs(u_bad - u_force) > 0.5 * u_force
