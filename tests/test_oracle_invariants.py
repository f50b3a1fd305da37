"""Closed-form pins of the oracle that do not rely on the reference build:
lattice moment identities, conservation laws of the BGK + Guo collision,
the equilibrium fixed point, cross-pattern equivalence, the AA in-place
partition property, brute-force adjacency, Taylor-Green viscous decay and the
bounce-back Poiseuille profile.

Each check is shown to be discriminating: a numpy restatement of the
collision with a planted mistake (dropped Guo term, flipped sign, missing F/2
shift, the wrong rest-population force weight) fails at least one of them,
while the correct restatement passes and equals the oracle."""
import numpy as np
import pytest

from oracle import oracle as O


# --------------------------------------------------------------------------- lattice
@pytest.mark.parametrize("q", [19, 27])
def test_lattice_moment_identities(q):
    c, opp, w = O.lattice(q)
    c = c.astype(np.float64)
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.abs(w @ c).max() < 1e-15
    assert np.abs(np.einsum("s,sa,sb->ab", w, c, c) - np.eye(3) / 3).max() < 1e-15
    assert np.abs(np.einsum("s,sa,sb,sc->abc", w, c, c, c)).max() < 1e-15
    # fourth-order isotropy: sum w c_a c_b c_c c_d = (d_ab d_cd + d_ac d_bd + d_ad d_bc) / 9
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
           + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.abs(np.einsum("s,sa,sb,sc,sd->abcd", w, c, c, c, c) - iso).max() < 1e-15
    assert np.array_equal(opp[opp], np.arange(q))
    assert np.array_equal(c[opp], -c)
    assert (c[-1] == 0).all() and opp[-1] == q - 1
    # direction order of the reference: +-x, +-y, +-z first
    assert np.array_equal(c[:6], [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                                  [0, 0, -1]])


# --------------------------------------------------------------------------- collision
def collide_np(f, omega, F, mutate=None):
    """BGK + Guo collision in numpy (f[q, n]); `mutate` plants one mistake."""
    q = f.shape[0]
    c, _, w = O.lattice(q)
    c = c.astype(np.float64)
    F = np.asarray(F, np.float64)
    rho = f.sum(0)
    m = c.T @ f
    u = m / rho if mutate == "no_half_force" else (m + 0.5 * F[:, None]) / rho
    uu = 3.0 * (u * u).sum(0)
    fu = 3.0 * (F @ u)
    cu = 3.0 * (c @ u)
    cf = 3.0 * (c @ F)
    cu_eq = -cu if mutate == "sign_cu" else cu
    eq = w[:, None] * rho * (1.0 + cu_eq + 0.5 * cu_eq * cu_eq - 0.5 * uu)
    Fs = w[:, None] * (cf[:, None] - fu + cu * cf[:, None])
    if mutate == "drop_guo":
        Fs = np.zeros_like(Fs)
    if mutate == "wrong_rest_weight":  # 1/36 instead of 1/3 for the rest population
        Fs[-1] *= (1.0 / 36.0) / (1.0 / 3.0)
    return (1.0 - omega) * f + omega * eq + (1.0 - 0.5 * omega) * Fs


def invariant_errors(f, out, F):
    q = f.shape[0]
    c = O.lattice(q)[0].astype(np.float64)
    mass = np.abs(out.sum(0) - f.sum(0)).max()
    mom = np.abs(c.T @ out - (c.T @ f + np.asarray(F)[:, None])).max()
    return mass, mom


def random_state(q, n, seed, amp=0.05):
    g = O.DenseGeometry(n, 1, 1)
    cell = np.arange(n, dtype=np.uint64)
    rho = 1.0 + 0.05 * (O.uniform01(seed, 4 * cell) - 0.5)
    u = np.stack([amp * (2 * O.uniform01(seed, 4 * cell + np.uint64(k + 1)) - 1) for k in range(3)], 1)
    feq = O.equilibrium(q, rho, u)
    # add a non-equilibrium part so that the state is not a fixed point
    noise = 1e-3 * (O.uniform01(seed + 1, np.arange(q * n, dtype=np.uint64)).reshape(q, n) - 0.5)
    noise -= noise.mean(0)  # keep rho; momentum changes slightly, which is fine
    return feq + noise * O.lattice(q)[2][:, None], g


def oracle_collide(f, omega, F):
    """One collision through the oracle: AB push with an identity adjacency
    (every link bounced back onto the node's own opposite slot = no streaming)
    followed by undoing the opposite-slot swap."""
    q, n = f.shape
    _, opp, _ = O.lattice(q)
    adj = (n * opp[:, None] + np.arange(n)[None, :]).astype(np.int32).reshape(-1)
    out = O.run_ab(q, O.SOA, n, np.ascontiguousarray(f.reshape(-1)), adj, 1, omega, force=F)
    out = out.reshape(q, n)
    return out[opp]


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("omega", [1.0, 1 / 0.6, 1 / 0.51])
def test_collision_conservation_and_planted_mistakes(q, omega):
    F = (1e-4, -3e-5, 2e-5)
    f, _ = random_state(q, 64, seed=5)
    good = collide_np(f, omega, F)
    orc = oracle_collide(f, omega, F)
    assert np.abs(orc - good).max() < 1e-15
    for out in (good, orc):
        mass, mom = invariant_errors(f, out, F)
        assert mass < 1e-15 and mom < 1e-15
    for mutant in ("drop_guo", "sign_cu", "no_half_force", "wrong_rest_weight"):
        bad = collide_np(f, omega, F, mutate=mutant)
        mass, mom = invariant_errors(f, bad, F)
        assert mass > 1e-12 or mom > 1e-12, f"invariants miss the planted mistake {mutant}"


@pytest.mark.parametrize("q", [19, 27])
def test_equilibrium_fixed_point(q):
    g = O.DenseGeometry(50, 1, 1)
    cell = np.arange(50, dtype=np.uint64)
    rho = 1.0 + 0.1 * (O.uniform01(3, cell) - 0.5)
    u = np.stack([0.08 * (2 * O.uniform01(3, 3 * cell + np.uint64(k)) - 1) for k in range(3)], 1)
    feq = O.equilibrium(q, rho, u)
    for omega in (1.0, 1.5, 1.9):
        out = oracle_collide(feq, omega, (0.0, 0.0, 0.0))
        assert np.abs(out - feq).max() < 4 * np.finfo(float).eps * np.abs(feq).max()
    # rest state with zero force is invariant under a full AA step
    geom = O.DenseGeometry(5, 4, 3)
    adj = geom.adjacency(q, O.SOA, O.ADJ_PULL)
    f = np.zeros(q * geom.nfluid)
    O.lib()
    getattr(O.lib(), "orc_init_rest_aa_f64")(q, O.SOA, geom.nfluid, O._ptr(f))
    f0 = f.copy()
    O.run_aa(q, O.SOA, geom.nfluid, f, adj, 2, 1.3, force=(0.0, 0.0, 0.0))
    assert np.abs(f - f0).max() < 4 * np.finfo(float).eps


# --------------------------------------------------------------------------- patterns
@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("layout", [O.SOA, O.AOS])
def test_aa_equals_push_equals_pull(q, layout):
    """After an even number of steps AA, AB push and AB pull hold the same
    pre-collision state bit for bit (halfway bounce-back on a random mask)."""
    nx, ny, nz = 7, 6, 5
    solid = (O.uniform01(21, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.15).astype(np.uint8)
    g = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 0, 1))
    n = g.nfluid
    fc, _ = random_state(q, n, 9)
    omega, F = 1 / 0.7, (2e-5, 0.0, -1e-5)
    adj_pull = g.adjacency(q, layout, O.ADJ_PULL)
    adj_push = g.adjacency(q, layout, O.ADJ_PUSH)
    steps = 6
    a = O.to_storage(fc, q, layout, aa_even_ready=True)
    O.run_aa(q, layout, n, a, adj_pull, steps, omega, F)
    aa_c = O.aa_canonical(a, q, n, layout, adj_pull, steps)
    push = O.from_storage(O.run_ab(q, layout, n, O.to_storage(fc, q, layout), adj_push, steps,
                                   omega, force=F), q, n, layout)
    assert np.array_equal(aa_c, push)
    if layout == O.SOA:
        # pull stores post-collision values; seed it with the inverse streaming of fc
        adjm = O.from_storage(adj_pull, q, n, layout)
        flat = np.empty(q * n)
        for s in range(q):
            flat[adjm[s]] = fc[s]
        out = O.run_ab(q, layout, n, flat, adj_pull, steps, omega, pull=True, force=F)
        pull_c = O.aa_canonical(out, q, n, layout, adj_pull, 1)  # gather = odd readout
        assert np.array_equal(pull_c, push)


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("mode", [O.ADJ_PUSH, O.ADJ_PULL])
def test_adjacency_is_a_partition(q, mode):
    """Every slot is the target of exactly one (node, direction) link: the AA
    read set of a node equals its write set and node sets are disjoint, so the
    in-place update is race free (src/kernels/KernelAA_SOA_Base.cpp:330-333)."""
    nx, ny, nz = 6, 5, 4
    solid = (O.uniform01(31, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.2).astype(np.uint8)
    for per in ((1, 1, 1), (1, 0, 1), (0, 0, 0)):
        g = O.DenseGeometry(nx, ny, nz, solid, periodic=per)
        for layout in (O.SOA, O.AOS):
            adj = g.adjacency(q, layout, mode)
            assert adj.min() >= 0 and adj.max() < q * g.nfluid
            assert np.array_equal(np.sort(adj), np.arange(q * g.nfluid))


@pytest.mark.parametrize("q", [19, 27])
def test_adjacency_brute_force(q):
    """orc_build_adjacency against a direct restatement of the rule in
    src/geometries/GeometryAAAB.cpp:44-85 on a tiny box."""
    nx, ny, nz = 4, 3, 3
    solid = (O.uniform01(41, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.25).astype(np.uint8)
    c, opp, _ = O.lattice(q)
    for per in ((1, 0, 0), (1, 1, 1), (0, 1, 0)):
        g = O.DenseGeometry(nx, ny, nz, solid, periodic=per, order="reference")
        n = g.nfluid
        for mode in (O.ADJ_PUSH, O.ADJ_PULL):
            adj = g.adjacency(q, O.SOA, mode).reshape(q, n)
            x, y, z = g.coords()
            for i in range(n):
                for s in range(q):
                    d = c[s] if mode == O.ADJ_PUSH else c[opp[s]]
                    p = [int(x[i]) + d[0], int(y[i]) + d[1], int(z[i]) + d[2]]
                    dims = (nx, ny, nz)
                    inside = True
                    for k in range(3):
                        if per[k]:
                            p[k] %= dims[k]
                        elif not 0 <= p[k] < dims[k]:
                            inside = False
                    j = g.fluid_of_cell[p[0] + nx * (p[1] + ny * p[2])] if inside else -1
                    want = n * opp[s] + i if j < 0 else n * s + j
                    assert adj[s, i] == want


# --------------------------------------------------------------------------- physics
def test_taylor_green_viscous_decay():
    """2-D Taylor-Green vortex (nz = 1, periodic): kinetic energy decays as
    exp(-4 nu k^2 t); the fitted nu must be (1/omega - 1/2) / 3 within 1 %."""
    q, nx, tau = 19, 32, 0.8
    omega = 1.0 / tau
    nu = (tau - 0.5) / 3.0
    g = O.DenseGeometry(nx, nx, 1)
    x, y, _ = g.coords()
    k = 2 * np.pi / nx
    u = np.zeros((g.nfluid, 3))
    u[:, 0] = 0.01 * np.sin(k * x) * np.cos(k * y)
    u[:, 1] = -0.01 * np.cos(k * x) * np.sin(k * y)
    fc = O.equilibrium(q, np.ones(g.nfluid), u)
    adj = g.adjacency(q, O.SOA, O.ADJ_PUSH)
    c = O.lattice(q)[0].astype(np.float64)

    def energy(f):
        m = c.T @ f
        return (m ** 2 / f.sum(0)).sum()

    t0, t1 = 20, 120
    f = O.to_storage(fc, q, O.SOA)
    f_a = O.run_ab(q, O.SOA, g.nfluid, f, adj, t0, omega, force=(0.0, 0.0, 0.0))
    f_b = O.run_ab(q, O.SOA, g.nfluid, f_a, adj, t1 - t0, omega, force=(0.0, 0.0, 0.0))
    e_a = energy(O.from_storage(f_a, q, g.nfluid, O.SOA))
    e_b = energy(O.from_storage(f_b, q, g.nfluid, O.SOA))
    nu_fit = -np.log(e_b / e_a) / (4 * k * k * (t1 - t0))
    assert abs(nu_fit / nu - 1.0) < 1e-2


def test_poiseuille_bounce_back_channel():
    """Body-force-driven channel between solid layers y = 0 and y = ny-1 with
    halfway bounce-back: the steady profile is the parabola
    u(y) = F (y - 1/2)(ny - 3/2 - y) / (2 nu) (walls half-way between nodes).
    For BGK the bounce-back wall sits exactly half-way at tau = 1/2 + sqrt(3/16)
    (He, Zou, Luo, Dembo, J. Stat. Phys. 87, 1997), so the profile is exact there."""
    q, nx, ny, nz = 19, 2, 12, 2
    tau = 0.5 + np.sqrt(3.0 / 16.0)
    omega, nu = 1 / tau, (tau - 0.5) / 3
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    g = O.DenseGeometry(nx, ny, nz, solid.reshape(-1), periodic=(1, 1, 1))
    F = 1e-6
    fc = O.equilibrium(q, np.ones(g.nfluid), np.zeros((g.nfluid, 3)))
    adj = g.adjacency(q, O.SOA, O.ADJ_PUSH)
    out = O.from_storage(O.run_ab(q, O.SOA, g.nfluid, O.to_storage(fc, q, O.SOA), adj, 8000,
                                  omega, force=(F, 0.0, 0.0)), q, g.nfluid, O.SOA)
    rho = out.sum(0)
    ux = (O.lattice(q)[0][:, 0] @ out + 0.5 * F) / rho
    _, y, _ = g.coords()
    h_lo, h_hi = 0.5, ny - 1.5
    expect = F * (y - h_lo) * (h_hi - y) / (2 * nu)
    assert np.abs(ux - expect).max() / expect.max() < 1e-9
    # mass is conserved to round-off by the Base kernels with bounce-back
    assert abs(out.sum() - fc.sum()) / fc.sum() < 1e-12


# --------------------------------------------------------------------------- readout
@pytest.mark.parametrize("q", [19, 27])
def test_moments_readout_closed_forms(q):
    """orc_moments (the reference readout, u = m / rho without the F/2 shift):
    moments of f^eq(rho, u) give back rho and u (sum f^eq = rho, sum f^eq c =
    rho u); swapping two opposite populations flips the momentum they carry;
    the rest population carries mass only."""
    n = 64
    cell = np.arange(n, dtype=np.uint64)
    rho = 1.0 + 0.05 * (O.uniform01(5, cell) - 0.5)
    u = np.stack([0.04 * (2 * O.uniform01(6 + k, cell) - 1) for k in range(3)], 1)
    f = O.equilibrium(q, rho, u)
    r, v = O.moments(q, f)
    assert np.abs(r - rho).max() < 1e-15 and np.abs(v - u).max() < 1e-15
    c, opp, _ = O.lattice(q)
    g = f.copy()
    g[[0, opp[0]]] = g[[opp[0], 0]]  # swap f_+x and f_-x
    r2, v2 = O.moments(q, g)
    assert np.array_equal(r2, r)
    dx = (f[0] - f[opp[0]]) / rho
    assert np.abs(v2[:, 0] - (v[:, 0] - 2 * dx)).max() < 1e-15
    h = f.copy()
    h[q - 1] += 0.25
    r3, v3 = O.moments(q, h)
    assert np.abs(r3 - (r + 0.25)).max() < 1e-15
    assert np.abs(v3 * r3[:, None] - v * r[:, None]).max() < 1e-15
