"""GPU discrete adjoint (lbm_adjoint_*, through the C ABI).

* small channel: objective, final state and both gradients against the FP64
  oracle adjoint (itself pinned by finite differences in test_oracle_adjoint);
* checkpoint interval k does not change the result (bit-for-bit);
* D3Q19 128^3 channel with a smooth seeded Brinkman field (sigma = 4), body
  force 1e-6, kinetic energy after 500 steps, checkpoint every 50: FP64
  gradient against 4th-order central finite differences on 32 seeded nodes,
  relative tolerance 1e-6; FP32 against FP64;
* outlet-flux objective with a per-node tau field."""
import numpy as np
import pytest

from lbmminiapp_amd import capi
from lbmminiapp_amd.capi import Adjoint
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_channel(q, seed=3, nx=16):
    ny, nz = 9, 6
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    solid[3, 4, 7] = 1
    solid = solid.reshape(-1)
    g = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 1, 1))
    cells = np.arange(nx * ny * nz, dtype=np.uint64)
    gamma = 0.2 + 0.6 * O.uniform01(seed + 10, cells)
    tau = 0.55 + 0.45 * O.uniform01(seed + 20, cells)
    cell = g.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.01 * (O.uniform01(seed, 4 * cell) - 0.5)
    u = np.stack([0.03 * (2 * O.uniform01(seed, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    fc = O.equilibrium(q, rho, u)
    full = np.zeros((q, g.ncell))
    full[:, g.cell_of_fluid] = fc
    return g, solid, gamma, tau, fc, full


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("objective", [0, 1])
def test_adjoint_matches_oracle(q, objective):
    g, solid, gamma, tau, fc, full = small_channel(q)
    steps, F = 10, (2e-5, 0.0, 0.0)
    out_x = g.nx - 3
    print(f"[adjoint] q={q} objective={objective}", flush=True)
    with Adjoint(g.nx, g.ny, g.nz, q=q, storage="f64", omega=1.0, force=F, solid=solid, tau=tau,
                 gamma=gamma, objective=objective, outlet_x=out_x, checkpoint_every=4) as adj:
        adj.set_initial_state(full)
        J = adj.forward(steps)
        fT = adj.final_state()
        gg, gt, _ = adj.backward()
    Jo, ggo, gto, ffo = O.adjoint(q, g, fc, steps, 1.0, F, tau_node=tau[g.cell_of_fluid],
                                  gamma_node=gamma[g.cell_of_fluid], objective=objective,
                                  outlet_x=out_x)
    assert np.array_equal(fT[:, g.cell_of_fluid], ffo)
    assert abs(J - Jo) <= 1e-13 * abs(Jo)
    for got, want in ((gg[g.cell_of_fluid], ggo), (gt[g.cell_of_fluid], gto)):
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    assert not gg[solid != 0].any() and not gt[solid != 0].any()


def test_adjoint_keeps_plain_rows(monkeypatch):
    """The adjoint's kernels index its arrays by cell, so its engine keeps plain
    x-rows even when interleaved rows are forced (LBM_DENSE_IROWS=1) on a box
    whose rows are a multiple of 64: objective, final state and gradients
    still equal the oracle adjoint."""
    monkeypatch.setenv("LBM_DENSE_IROWS", "1")
    q = 19
    g, solid, gamma, tau, fc, full = small_channel(q, seed=5, nx=64)
    steps, F = 8, (2e-5, 0.0, 0.0)
    with Adjoint(g.nx, g.ny, g.nz, q=q, storage="f64", omega=1.0, force=F, solid=solid, tau=tau,
                 gamma=gamma, objective=0, checkpoint_every=3) as adj:
        adj.set_initial_state(full)
        J = adj.forward(steps)
        fT = adj.final_state()
        gg, gt, _ = adj.backward()
    Jo, ggo, gto, ffo = O.adjoint(q, g, fc, steps, 1.0, F, tau_node=tau[g.cell_of_fluid],
                                  gamma_node=gamma[g.cell_of_fluid], objective=0)
    assert np.array_equal(fT[:, g.cell_of_fluid], ffo)
    assert abs(J - Jo) <= 1e-13 * abs(Jo)
    for got, want in ((gg[g.cell_of_fluid], ggo), (gt[g.cell_of_fluid], gto)):
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


def test_checkpoint_interval_does_not_change_gradients():
    """One level (k = 1, 3, 5, 17) and nested levels ((k1, k2, k3) with ragged
    last segments, k2 = 1, and intervals longer than the run) give the same
    gradients bit for bit."""
    q = 19
    g, solid, gamma, tau, fc, full = small_channel(q, seed=5)
    res = []
    plans = [(1, (0, 0)), (3, (0, 0)), (5, (0, 0)), (17, (0, 0)), (17, (5, 2)), (8, (3, 0)),
             (7, (1, 0)), (16, (4, 4)), (40, (20, 3))]
    for every, inner in plans:
        print(f"[adjoint] checkpoints every {every}, inner {inner}", flush=True)
        with Adjoint(g.nx, g.ny, g.nz, q=q, storage="f64", omega=1.0, force=(1e-5, 0, 0),
                     solid=solid, tau=tau, gamma=gamma, checkpoint_every=every,
                     checkpoint_inner=inner) as adj:
            adj.set_initial_state(full)
            adj.forward(17)
            res.append(adj.backward()[:2])
    for gg, gt in res[1:]:
        assert np.array_equal(gg, res[0][0]) and np.array_equal(gt, res[0][1])


def brinkman_channel_case(n=128):
    solid = capi.channel(n, n, n)
    gamma = capi.smooth_random_field(n, n, n, 4.0, 4)
    return solid, gamma


def test_channel_128_brinkman_gradient_vs_finite_differences():
    """D3Q19 128^3 channel (walls y = 0, 127), gamma from seeded white noise
    (seed 4) smoothed with a periodic Gaussian (sigma = 4) and rescaled to
    [0, 1], F = (1e-6, 0, 0), tau = 0.8, J = kinetic energy after 500 steps,
    checkpoint every 50.  FP64 adjoint gradient vs 6th-order central finite
    differences at 32 seeded fluid nodes (seed 5): relative error <= 1e-6.
    Step h = 4e-2: the round-off in J(+kh) - J(-kh) divided by h reaches
    ~2e-6 of the smallest sampled gradients (|g| ~ 1.7e-12) at h = 1e-2 and
    2e-2, while 4th- and 6th-order differences agree to 1e-7 at h = 4e-2
    (truncation negligible); measured with the CPU oracle at this size."""
    n, steps = 128, 500
    solid, gamma = brinkman_channel_case(n)
    print("[adjoint] Brinkman channel 128^3 forward/backward", flush=True)
    with Adjoint(n, n, n, q=19, storage="f64", omega=1 / 0.8, force=(1e-6, 0, 0), solid=solid,
                 gamma=gamma, objective="kinetic_energy", checkpoint_every=50) as adj:
        J = adj.forward(steps)
        gg, _, ms = adj.backward(tau=False)
        print(f"[adjoint] J = {J:.6e}, backward {ms:.1f} ms, "
              f"{n ** 3 * steps / (ms * 1e3):.0f} adjoint MLUPS", flush=True)
        fluid = np.nonzero(solid == 0)[0]
        picks = fluid[(O.uniform01(5, np.arange(32, dtype=np.uint64)) * fluid.size).astype(int)]
        worst = 0.0
        for i, cell in enumerate(picks):
            d = adj.finite_difference("gamma", int(cell), 4e-2, steps, order=6)
            rel = abs(gg[cell] - d) / abs(d)
            worst = max(worst, rel)
            if i < 4 or rel > 1e-6:
                print(f"[adjoint] cell {cell}: adjoint {gg[cell]:.9e} fd {d:.9e} rel {rel:.2e}",
                      flush=True)
    print(f"[adjoint] worst relative error over 32 nodes: {worst:.3e}")
    assert worst <= 1e-6


def test_channel_128_brinkman_fp32_adjoint_vs_fp64():
    """The Brinkman channel case at its full extent in FP32 deviation storage (FP32 forward and
    adjoint arithmetic, FP64 gradient accumulation): D3Q19 128^3 channel,
    smooth gamma field, kinetic energy after 500 steps, checkpoint every 50.
    J within 1e-7 relative and the gamma gradient within 1e-6 rel-L2 of the
    FP64 adjoint (itself pinned to finite differences at 1e-6); measured on
    the MI355X: J 1.2e-8, gradient 3.4e-7."""
    n, steps = 128, 500
    solid, gamma = brinkman_channel_case(n)
    out = {}
    for st in ("f64", "f32s"):
        print(f"[adjoint] Brinkman channel 128^3 {st}", flush=True)
        with Adjoint(n, n, n, q=19, storage=st, omega=1 / 0.8, force=(1e-6, 0, 0), solid=solid,
                     gamma=gamma, checkpoint_every=50) as adj:
            J = adj.forward(steps)
            gg, _, ms = adj.backward(tau=False)
            out[st] = (J, gg)
    J64, g64 = out["f64"]
    J32, g32 = out["f32s"]
    fluid = solid == 0
    rel = np.linalg.norm(g32[fluid] - g64[fluid]) / np.linalg.norm(g64[fluid])
    print(f"[adjoint] FP32 vs FP64: J rel {abs(J32 - J64) / J64:.3e}, gradient rel-L2 {rel:.3e}")
    assert abs(J32 - J64) / J64 < 1e-7
    assert rel < 1e-6


def test_outlet_flux_tau_gradient_vs_finite_differences():
    """Outlet mass flux objective with a per-node tau field in [0.55, 1.0]
    (smooth seeded field, seed 7) on a small porous channel; 6th-order
    differences with h = 1e-2 (J is strongly nonlinear in tau: the 4th-order
    difference at h = 2e-2 is off by 6e-6 from truncation alone)."""
    nx, ny, nz = 48, 24, 16
    solid = capi.channel(nx, ny, nz)
    capi.add_spheres(solid, nx, ny, nz, 2.0, 3.0, 0.15, 6, periodic=(1, 0, 1))
    tau = 0.55 + 0.45 * capi.smooth_random_field(nx, ny, nz, 2.0, 7)
    steps = 60
    with Adjoint(nx, ny, nz, q=19, storage="f64", omega=1.0, force=(1e-5, 0, 0), solid=solid,
                 tau=tau, objective="outlet_flux", outlet_x=nx - 1, checkpoint_every=16) as adj:
        adj.forward(steps)
        _, gt, _ = adj.backward(gamma=False)
        fluid = np.nonzero(solid == 0)[0]
        # nodes within reach of the outlet plane in `steps` steps
        near = fluid[(fluid % nx) > nx - 20]
        picks = near[(O.uniform01(8, np.arange(8, dtype=np.uint64)) * near.size).astype(int)]
        for cell in picks:
            d = adj.finite_difference("tau", int(cell), 1e-2, steps, order=6)
            assert abs(gt[cell] - d) <= 1e-6 * abs(d) + 1e-12 * np.abs(gt).max(), (cell, gt[cell], d)


def test_pressure_drop_adjoint_vs_finite_differences():
    """The forward model with the pressure-drop boundary (density jump 3e-3
    across the periodic x boundary, no body force): the jump is additive on
    the crossing links, so the transposed step is unchanged and the tau
    gradient of the outlet flux still matches 6th-order finite differences."""
    nx, ny, nz = 48, 24, 16
    solid = capi.channel(nx, ny, nz)
    capi.add_spheres(solid, nx, ny, nz, 2.0, 3.0, 0.15, 6, periodic=(1, 0, 1))
    tau = 0.55 + 0.45 * capi.smooth_random_field(nx, ny, nz, 2.0, 7)
    steps = 60
    print("[adjoint] pressure drop, outlet flux vs tau", flush=True)
    with Adjoint(nx, ny, nz, q=19, storage="f64", periodic=(1, 0, 1), omega=1.0, tau=tau,
                 solid=solid, objective="outlet_flux", outlet_x=nx - 1, checkpoint_every=16,
                 density_jump_x=3e-3) as adj:
        J = adj.forward(steps)
        _, gt, _ = adj.backward(gamma=False)
        assert J > 0  # the jump drives flow through the outlet in +x
        fluid = np.nonzero(solid == 0)[0]
        near = fluid[((fluid % nx) > nx - 20) | ((fluid % nx) < 10)]
        picks = near[(O.uniform01(9, np.arange(8, dtype=np.uint64)) * near.size).astype(int)]
        for cell in picks:
            d = adj.finite_difference("tau", int(cell), 1e-2, steps, order=6)
            assert abs(gt[cell] - d) <= 1e-6 * abs(d) + 1e-12 * np.abs(gt).max(), (cell, gt[cell], d)


@pytest.mark.parametrize("nx", [160, 100])
def test_adjoint_non_power_of_two_rows(nx):
    """Rows of 160 and 100 cells run the seed and difference reductions in
    workgroups of 192 and 128 lanes (rounded up to whole wavefronts).  J, the
    final state and the gradients equal the FP64 oracle; two seeded central
    differences agree with the adjoint gradient."""
    q, ny, nz = 19, 6, 4
    solid = capi.channel(nx, ny, nz)
    g = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 1, 1))
    cells = np.arange(g.ncell, dtype=np.uint64)
    gamma = 0.1 + 0.5 * O.uniform01(61, cells)
    cell = g.cell_of_fluid.astype(np.uint64)
    u = np.stack([0.03 * (2 * O.uniform01(62, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    fc = O.equilibrium(q, np.ones(g.nfluid), u)
    full = np.zeros((q, g.ncell))
    full[:, g.cell_of_fluid] = fc
    steps, F = 12, (1e-5, 0.0, 0.0)
    print(f"[adjoint] rows of {nx}", flush=True)
    with Adjoint(nx, ny, nz, q=q, storage="f64", omega=1 / 0.7, force=F, solid=solid,
                 gamma=gamma, checkpoint_every=5) as adj:
        adj.set_initial_state(full)
        J = adj.forward(steps)
        gg, _, _ = adj.backward(tau=False)
        picks = g.cell_of_fluid[(O.uniform01(63, np.arange(2, dtype=np.uint64)) * g.nfluid).astype(int)]
        fds = [adj.finite_difference("gamma", int(c), 1e-3, steps, order=6) for c in picks]
    Jo, ggo, _, _ = O.adjoint(q, g, fc, steps, 1 / 0.7, F, gamma_node=gamma[g.cell_of_fluid])
    assert abs(J - Jo) <= 1e-13 * abs(Jo), (J, Jo)
    assert np.abs(gg[g.cell_of_fluid] - ggo).max() <= 1e-10 * np.abs(ggo).max()
    for c, d in zip(picks, fds):
        assert abs(gg[c] - d) <= 1e-6 * abs(d) + 1e-9 * np.abs(gg).max(), (c, gg[c], d)


@pytest.mark.parametrize("steps", [10, 11])
def test_adjoint_lambda0_vs_oracle(steps):
    """lambda^0 = dJ/df^0 read back after even and odd step counts (the two
    ping-pong arrays) equals the oracle's (pinned by finite differences in the
    initial populations) in every fluid slot, and is zero in every solid cell."""
    q = 19
    g, solid, gamma, tau, fc, full = small_channel(q, seed=9)
    F = (2e-5, 0.0, 0.0)
    with Adjoint(g.nx, g.ny, g.nz, q=q, storage="f64", omega=1.0, force=F, solid=solid, tau=tau,
                 gamma=gamma, checkpoint_every=4) as adj:
        adj.set_initial_state(full)
        adj.forward(steps)
        adj.backward()
        lam = adj.lambda0()
    _, _, _, _, lam_o = O.adjoint(q, g, fc, steps, 1.0, F, tau_node=tau[g.cell_of_fluid],
                                  gamma_node=gamma[g.cell_of_fluid], with_lambda0=True)
    assert np.abs(lam[:, g.cell_of_fluid] - lam_o).max() <= 1e-12 * np.abs(lam_o).max()
    assert not lam[:, solid != 0].any()


@pytest.mark.parametrize("param", ["gamma", "tau"])
def test_directional_difference_matches_adjoint(param):
    """lbm_adjoint_fd_direction: the 6th-order central difference of J along a
    seeded whole-field direction equals sum_x grad(x) v(x) from the GPU adjoint
    (and from the oracle adjoint) to 1e-7 relative, for both fields and the
    kinetic-energy objective on the small channel."""
    q = 19
    g, solid, gamma, tau, fc, full = small_channel(q, seed=7)
    steps, F = 12, (2e-5, 0.0, 0.0)
    v = O.uniform01(91, np.arange(g.ncell, dtype=np.uint64)) - 0.5
    with Adjoint(g.nx, g.ny, g.nz, q=q, storage="f64", omega=1.0, force=F, solid=solid, tau=tau,
                 gamma=gamma, checkpoint_every=5) as adj:
        adj.set_initial_state(full)
        adj.forward(steps)
        gg, gt, _ = adj.backward()
        d = adj.finite_difference_direction(param, v, 1e-3, steps, order=6)
        J_after = adj.forward(steps)  # the field is restored after the differences
    _, ggo, gto, _ = O.adjoint(q, g, fc, steps, 1.0, F, tau_node=tau[g.cell_of_fluid],
                               gamma_node=gamma[g.cell_of_fluid])
    fl = g.cell_of_fluid
    grad, grad_o = (gg, ggo) if param == "gamma" else (gt, gto)
    pred = float(np.dot(grad[fl], v[fl]))
    assert abs(pred - float(np.dot(grad_o, v[fl]))) <= 1e-10 * abs(pred)
    assert abs(pred - d) <= 1e-7 * abs(d), (pred, d)
    Jo = O.adjoint(q, g, fc, steps, 1.0, F, tau_node=tau[fl], gamma_node=gamma[fl])[0]
    assert abs(J_after - Jo) <= 1e-13 * abs(Jo)
