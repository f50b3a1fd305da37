"""The benchmark configurations at their full sizes, through properties that do
not depend on the size: bit-for-bit agreement with the oracle over a few steps
(the OpenMP oracle runs the full grid), mass conservation, flow direction, the
imposed pressure drop, and reduced-precision error against FP64."""
import numpy as np
import pytest

from lbmminiapp_amd import capi
from lbmminiapp_amd.capi import Adjoint, DenseLattice
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_channel_256_spheres_pull_full_size():
    """D3Q19 AB pull FP64, 256^3 channel (walls y = 0, 255), 5% spheres of
    radius 4-12 (seed 2), F = 1e-6 x, tau = 0.8: equal to the oracle bit for
    bit after 4 steps from a seeded state; then mass conserved to 1e-12 and
    flow along +x after 300 more steps."""
    n, q = 256, 19
    solid = capi.channel(n, n, n)
    capi.add_spheres(solid, n, n, n, 4.0, 12.0, 0.05, 2, periodic=(1, 0, 1))
    geom = O.DenseGeometry(n, n, n, solid, periodic=(1, 0, 1))
    cell = geom.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.01 * (O.uniform01(3, 4 * cell) - 0.5)
    u = np.stack([0.02 * (2 * O.uniform01(3, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    fc = O.equilibrium(q, rho, u)
    del rho, u, cell
    omega, F = 1 / 0.8, (1e-6, 0.0, 0.0)
    print("[channel] 256^3 dense vs oracle, 4 steps", flush=True)
    full = np.zeros((q, geom.ncell))
    full[:, geom.cell_of_fluid] = fc
    with DenseLattice(n, n, n, q=q, pattern="ab", pull=True, storage="f64", periodic=(1, 0, 1),
                      omega=omega, force=F, solid=solid) as lat:
        lat.set_state(full)
        del full
        lat.step(4)
        got = lat.get_state()[:, geom.cell_of_fluid]
        m0 = lat.mass()
        lat.step(300)
        m1 = lat.mass()
        _, vel = lat.moments()
    adj = geom.adjacency(q, O.SOA, O.ADJ_PULL)
    adjm = O.from_storage(adj, q, geom.nfluid, O.SOA)
    flat = np.empty(q * geom.nfluid)
    for s in range(q):
        flat[adjm[s]] = fc[s]
    out = O.run_ab(q, O.SOA, geom.nfluid, flat, adj, 4, omega, pull=True, force=F)
    want = O.aa_canonical(out, q, geom.nfluid, O.SOA, adj, 1)
    assert np.array_equal(got, want)
    assert abs(m1 - m0) <= 1e-12 * m0
    fluid = solid == 0
    assert np.isfinite(vel).all() and vel[fluid, 0].mean() > 0


def _tgv_fields(n, q, storage, tau, steps, noise):
    with DenseLattice(n, n, n, q=q, pattern="aa", storage=storage, omega=1 / tau) as lat:
        lat.init_tgv(0.05, noise, 1)
        m0 = lat.mass()
        lat.step(steps)
        m1 = lat.mass()
        rho, u = lat.moments()
    return rho, u, m0, m1


def _precision_vs_f64(n, q, tau, steps, noise, gates):
    """Runs FP64, FP32 and FP16 deviation storage from the same Taylor-Green
    state and checks each reduced storage against FP64 field by field:
    velocity rel-L2, density rel-L2 relative to rho itself and relative to the
    deviation rho - 1, and mass drift."""
    rho64, u64, m0, m1 = _tgv_fields(n, q, "f64", tau, steps, noise)
    assert abs(m1 - m0) <= 1e-12 * m0
    nu, nr, nd = np.linalg.norm(u64), np.linalg.norm(rho64), np.linalg.norm(rho64 - 1.0)
    for st, (g_u, g_rho, g_dev) in gates.items():
        print(f"[precision] D3Q{q} {n}^3 {st} x {steps}", flush=True)
        rho, u, s0, s1 = _tgv_fields(n, q, st, tau, steps, noise)
        e_u = np.linalg.norm(u - u64) / nu
        e_rho = np.linalg.norm(rho - rho64) / nr
        e_dev = np.linalg.norm(rho - rho64) / nd
        print(f"[precision] {st}: u rel-L2 {e_u:.3e}, rho rel-L2 {e_rho:.3e}, "
              f"(rho - 1) rel-L2 {e_dev:.3e}, mass drift {(s1 - s0) / s0:.2e}", flush=True)
        assert e_u <= g_u and e_rho <= g_rho and e_dev <= g_dev, (st, e_u, e_rho, e_dev)
        assert abs(s1 - s0) <= 1e-5 * s0
        del rho, u


def test_d3q27_512_reduced_precision_2000_steps():
    """D3Q27 AA 512^3 periodic Taylor-Green (u0 = 0.05, tau = 0.51, no noise, F = 0)
    over its 2000 steps: FP32 and FP16 deviation storage (FP32 collision, FP64
    moment sums) against the FP64 run, field by field.  Gates: velocity rel-L2
    <= 1e-4 (FP32) and <= 1e-3 (FP16); density rel-L2 relative to rho <= 3e-7
    (FP32) and <= 2e-5 (FP16), relative to rho - 1 <= 1e-4 (FP32) and <= 1e-2 (FP16).
    Measured on the MI355X: FP32 u 3.2e-5, rho 8.1e-8, rho - 1 4.5e-5; FP16 u
    4.3e-4, rho 9.4e-6, rho - 1 5.2e-3."""
    _precision_vs_f64(512, 27, 0.51, 2000, 0.0,
                      {"f32s": (1e-4, 3e-7, 1e-4), "f16s": (1e-3, 2e-5, 1e-2)})


def test_headline_d3q19_512_fp16_2000_steps():
    """The headline workload (D3Q19 AA 512^3 periodic Taylor-Green, u0 = 0.05
    plus seeded U(-1e-4, 1e-4) noise, tau = 0.6) over 2000 steps: FP16
    deviation storage within velocity rel-L2 1e-3 of FP64, FP32 within 1e-4;
    density rel-L2 relative to rho <= 2e-5 / 3e-7 and relative to rho - 1
    <= 1e-2 / 2e-4.  Measured on the MI355X: FP16 u 7.6e-4, rho 6.5e-6, rho - 1
    3.8e-3; FP32 u 3.7e-5, rho 9.8e-8, rho - 1 5.6e-5."""
    _precision_vs_f64(512, 19, 0.6, 2000, 1e-4,
                      {"f32s": (1e-4, 3e-7, 2e-4), "f16s": (1e-3, 2e-5, 1e-2)})


def _porous_384():
    n = 384
    solid = np.zeros(n ** 3, np.uint8)
    capi.add_spheres(solid, n, n, n, 6.0, 16.0, 0.35, 6, periodic=(1, 1, 1))
    return n, solid


def test_porous_384_aos_and_soa_3000_steps():
    """D3Q19 AB push FP32 deviations, 384^3 packed spheres (radius 6-16, 35%
    solid, seed 6), inlet/outlet pressure drop 1e-4 (density jump 3e-4 across
    the periodic x boundary), its 3000 steps, in the SOA and in the AOS layout
    (the AOS arm of the comparison).  For each layout: the inlet plane's mean
    density exceeds the outlet plane's by 3e-4 within 5%, the mean flow runs
    along +x, the mass is unchanged to 1e-9.  The two layouts agree field by
    field: velocity rel-L2 <= 1e-6, max |rho difference| <= 1e-7."""
    n, solid = _porous_384()
    fluid = solid == 0
    x = np.arange(n ** 3) % n
    fields = {}
    for layout in ("soa", "aos"):
        print(f"[porous] 384^3 porous {layout} x 3000", flush=True)
        with DenseLattice(n, n, n, q=19, pattern="ab", layout=layout, storage="f32s", omega=1.0,
                          solid=solid, density_jump_x=3e-4) as lat:
            lat.init_tgv(0.0, 0.0, 1)
            m0 = lat.mass()
            lat.step(3000)
            m1 = lat.mass()
            rho, vel = lat.moments()
        drho = rho[fluid & (x == 0)].mean() - rho[fluid & (x == n - 1)].mean()
        print(f"[porous] {layout}: inlet - outlet {drho:.4e}, mean ux {vel[fluid, 0].mean():.3e}, "
              f"mass drift {(m1 - m0) / m0:.2e}", flush=True)
        assert abs(drho - 3e-4) < 0.05 * 3e-4, (layout, drho)
        assert vel[fluid, 0].mean() > 0
        assert abs(m1 - m0) <= 1e-9 * m0
        fields[layout] = (rho[fluid], vel[fluid])
        del rho, vel
    (rs, us), (ra, ua) = fields["soa"], fields["aos"]
    e_u = np.linalg.norm(ua - us) / np.linalg.norm(us)
    e_r = np.abs(ra - rs).max()
    print(f"[porous] AOS vs SOA: u rel-L2 {e_u:.3e}, max |drho| {e_r:.3e}, "
          f"bitwise {np.array_equal(ua, us) and np.array_equal(ra, rs)}", flush=True)
    assert e_u <= 1e-6 and e_r <= 1e-7


def test_porous_384_tau_adjoint_nested_checkpoints():
    """Adjoint of the outlet mass flux (plane x = nx - 1) with respect to a
    per-node tau field in [0.55, 1.0] (smooth seeded field, sigma 4, seed 7)
    on the 384^3 porous medium with the pressure drop, nested checkpoints
    every 250 / 25 / 5 steps.  Reduced extent: 500 steps instead of 3000 (the
    device holds 19 FP64 states of 8.6 GB).

    FP64: along a smooth seeded whole-field direction v (sigma 4, seed 13,
    values in [-1/2, 1/2]) the adjoint's sum_x dJ/dtau(x) v(x) agrees with the
    4th-order central difference of J (h = 1e-3) to 1e-6 relative.  Single
    cells near the outlet agree with 6th-order differences (h = 1e-2) to 1e-4:
    one cell moves J = 3.4 by about 4e-9 per 1e-2 of tau, and the FP64
    round-off of the two 500-step runs behind each difference scatters the
    quotient by 1e-5 to 8e-5 (measured over h = 1.25e-3 .. 1e-2, orders 2 to
    6, without converging), which is the resolution of a single-cell
    difference at this size.  FP32 deviations (FP32 forward and adjoint
    arithmetic, FP64 accumulation): the gradient is within 5e-2 rel-L2 of the
    FP64 gradient (measured 2.5e-2: the porous flow's 500 FP32 forward steps
    and their transposes), J within 1e-5."""
    n, solid = _porous_384()
    tau = 0.55 + 0.45 * capi.smooth_random_field(n, n, n, 4.0, 7)
    v = capi.smooth_random_field(n, n, n, 4.0, 13) - 0.5
    steps = 500
    fluid = np.nonzero(solid == 0)[0]
    near = fluid[(fluid % n) > n - 12]
    picks = near[(O.uniform01(12, np.arange(2, dtype=np.uint64)) * near.size).astype(int)]
    grads, objs = {}, {}
    for st in ("f64", "f32s"):
        print(f"[porous-adjoint] 384^3 {st} x {steps}", flush=True)
        with Adjoint(n, n, n, q=19, storage=st, omega=1.0, solid=solid, tau=tau,
                     objective="outlet_flux", outlet_x=n - 1, checkpoint_every=250,
                     checkpoint_inner=(25, 5), density_jump_x=3e-4) as adj:
            J = adj.forward(steps)
            _, gt, ms = adj.backward(gamma=False)
            print(f"[porous-adjoint] {st}: J = {J:.6e}, backward {ms:.0f} ms, "
                  f"{fluid.size * steps / (ms * 1e3):.0f} adjoint MLUPS, "
                  f"{adj.device_bytes() / 1e9:.1f} GB", flush=True)
            if st == "f64":
                pred = float(np.dot(gt[fluid], v[fluid]))
                d = adj.finite_difference_direction("tau", v, 1e-3, steps, order=4)
                rel = abs(pred - d) / abs(d)
                print(f"[porous-adjoint] direction: adjoint {pred:.12e} fd {d:.12e} "
                      f"rel {rel:.2e}", flush=True)
                assert rel <= 1e-6, (pred, d)
                for cell in picks:
                    d = adj.finite_difference("tau", int(cell), 1e-2, steps, order=6)
                    rel = abs(gt[cell] - d) / abs(d)
                    print(f"[porous-adjoint] cell {cell}: adjoint {gt[cell]:.9e} fd {d:.9e} "
                          f"rel {rel:.2e}", flush=True)
                    assert rel <= 1e-4, (cell, gt[cell], d)
        grads[st] = gt[fluid]
        objs[st] = J
        del gt
    e = np.linalg.norm(grads["f32s"] - grads["f64"]) / np.linalg.norm(grads["f64"])
    ej = abs(objs["f32s"] - objs["f64"]) / abs(objs["f64"])
    print(f"[porous-adjoint] FP32 vs FP64: gradient rel-L2 {e:.3e}, J rel {ej:.3e}", flush=True)
    assert e <= 5e-2 and ej <= 1e-5
