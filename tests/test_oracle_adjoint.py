"""Pins the oracle's discrete adjoint (orc_adjoint) with central finite
differences of the oracle forward model, for both parameter fields (Brinkman
gamma, per-node tau) and both objectives (final kinetic energy, outlet
x-flux).  The forward model is the reference collide-stream step (bit-for-bit
pinned elsewhere) with the two per-node extensions; it reduces to the
reference when both fields are absent.

The finite differences use the 4th-order stencil
  (8 [J(p+h) - J(p-h)] - [J(p+2h) - J(p-2h)]) / (12 h)
on a small channel, so truncation error (~h^4) and round-off (~eps J / h)
are both far below the 1e-6 relative tolerance."""
import numpy as np
import pytest

from oracle import oracle as O


def channel_case(q, seed=3):
    nx, ny, nz = 8, 7, 5
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    solid[2, 3, 4] = 1  # one obstacle cell
    g = O.DenseGeometry(nx, ny, nz, solid.reshape(-1), periodic=(1, 1, 1))
    cell = g.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.01 * (O.uniform01(seed, 4 * cell) - 0.5)
    u = np.stack([0.03 * (2 * O.uniform01(seed, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    f0 = O.equilibrium(q, rho, u)
    n = g.nfluid
    gamma = 0.2 + 0.6 * O.uniform01(seed + 10, np.arange(n, dtype=np.uint64))
    tau = 0.55 + 0.45 * O.uniform01(seed + 20, np.arange(n, dtype=np.uint64))
    return g, f0, gamma, tau


def fd(fun, p, i, h):
    def J(delta):
        pp = p.copy()
        pp[i] += delta
        return fun(pp)
    return (8 * (J(h) - J(-h)) - (J(2 * h) - J(-2 * h))) / (12 * h)


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("objective", [0, 1])
def test_adjoint_gradients_match_finite_differences(q, objective):
    g, f0, gamma, tau = channel_case(q)
    steps, omega, F = 12, 1.0, (2e-5, 0.0, 0.0)
    J, gg, gt, _ = O.adjoint(q, g, f0, steps, omega, F, tau_node=tau, gamma_node=gamma,
                             objective=objective, outlet_x=g.nx - 2)
    n = g.nfluid
    rng_nodes = (O.uniform01(99, np.arange(6, dtype=np.uint64)) * n).astype(int)

    def J_gamma(gm):
        return O.adjoint(q, g, f0, steps, omega, F, tau_node=tau, gamma_node=gm,
                         objective=objective, outlet_x=g.nx - 2)[0]

    def J_tau(tu):
        return O.adjoint(q, g, f0, steps, omega, F, tau_node=tu, gamma_node=gamma,
                         objective=objective, outlet_x=g.nx - 2)[0]

    for i in rng_nodes:
        d_g = fd(J_gamma, gamma, i, 1e-3)
        d_t = fd(J_tau, tau, i, 1e-3)
        scale_g = np.abs(gg).max()
        scale_t = np.abs(gt).max()
        assert abs(gg[i] - d_g) <= 1e-6 * max(abs(d_g), 1e-3 * scale_g), (i, gg[i], d_g)
        assert abs(gt[i] - d_t) <= 1e-6 * max(abs(d_t), 1e-3 * scale_t), (i, gt[i], d_t)


def test_adjoint_without_fields_is_the_reference_model():
    """With no tau / gamma fields the adjoint's forward pass is the reference
    push step: its final state equals run_ab bit for bit."""
    q = 19
    g, f0, _, _ = channel_case(q)
    adj = g.adjacency(q, O.SOA, O.ADJ_PUSH)
    _, _, _, ff = O.adjoint(q, g, f0, 7, 1 / 0.7, (1e-6, 0, 0))
    want = O.run_ab(q, O.SOA, g.nfluid, np.ascontiguousarray(f0.reshape(-1)), adj, 7, 1 / 0.7,
                    force=(1e-6, 0, 0))
    assert np.array_equal(ff.reshape(-1), want)


def test_objective_closed_forms():
    q = 19
    g = O.DenseGeometry(6, 5, 4)
    u = np.zeros((g.nfluid, 3))
    u[:, 0] = 0.02
    u[:, 2] = -0.01
    f = O.equilibrium(q, np.full(g.nfluid, 1.1), u)
    ke = O.objective(q, g, f, 0)
    assert abs(ke - g.nfluid * 0.5 * 1.1 * (0.02 ** 2 + 0.01 ** 2)) < 1e-15 * g.nfluid
    flux = O.objective(q, g, f, 1, outlet_x=3)
    assert abs(flux - (g.nfluid / 6) * 1.1 * 0.02) < 1e-15 * g.nfluid


@pytest.mark.parametrize("objective", [0, 1])
def test_adjoint_lambda0_matches_finite_differences(objective):
    """lambda^0 = dJ/df^0: perturbing one initial population f_s(x) changes J by
    lambda^0_s(x) times the perturbation (4th-order central differences), for
    seeded (s, x) pairs including a population next to the obstacle."""
    q = 19
    g, f0, gamma, tau = channel_case(q, seed=4)
    steps, omega, F = 9, 1.0, (2e-5, 0.0, 0.0)
    J, _, _, _, lam0 = O.adjoint(q, g, f0, steps, omega, F, tau_node=tau, gamma_node=gamma,
                                 objective=objective, outlet_x=g.nx - 2, with_lambda0=True)
    n = g.nfluid
    assert lam0.shape == (q, n)

    def J_f(ff):
        return O.adjoint(q, g, ff.reshape(q, n), steps, omega, F, tau_node=tau,
                         gamma_node=gamma, objective=objective, outlet_x=g.nx - 2)[0]

    flat = np.ascontiguousarray(f0.reshape(-1))
    picks = (O.uniform01(77, np.arange(8, dtype=np.uint64)) * flat.size).astype(int)
    scale = np.abs(lam0).max()
    assert scale > 0
    for i in picks:
        d = fd(J_f, flat, i, 1e-4)
        got = lam0.reshape(-1)[i]
        assert abs(got - d) <= 1e-6 * max(abs(d), 1e-3 * scale), (i, got, d)
