"""The C oracle against the reference kernels built from /root/reference/src
into oracle/_ref/ (only where that build exists: this container).  Bit-for-bit
equality on seeded states, random solid masks and both fluid orderings."""
import numpy as np
import pytest

from oracle import oracle as O

CASES = [(q, prec, om) for q in (19, 27) for prec in ("f64", "f32")
         for om in (1.0, 1 / 0.6, 1 / 0.8, 1 / 0.51)]


def _seeded(geom, q, seed):
    cell = geom.cell_of_fluid.astype(np.uint64)
    rho = 1.0 + 0.02 * (O.uniform01(seed, np.uint64(4) * cell) - 0.5)
    u = np.stack([0.05 * (2 * O.uniform01(seed, np.uint64(4) * cell + np.uint64(k + 1)) - 1)
                  for k in range(3)], axis=1)
    return O.equilibrium(q, rho, u)


@pytest.mark.parametrize("q,prec,omega", CASES)
def test_oracle_equals_reference(q, prec, omega):
    if not O.ref_available(q, prec, omega):
        pytest.skip("reference build not present (needs /root/reference and make -C oracle)")
    dt = np.float64 if prec == "f64" else np.float32
    nx, ny, nz = 9, 7, 6
    solid = (O.uniform01(7, np.arange(nx * ny * nz, dtype=np.uint64)) < 0.1).astype(np.uint8)
    for order in ("x_fastest", "reference"):
        g = O.DenseGeometry(nx, ny, nz, solid, periodic=(1, 0, 1), order=order)
        n = g.nfluid
        fc = _seeded(g, q, 11)
        for layout in (O.SOA, O.AOS):
            L = "soa" if layout == O.SOA else "aos"
            adj = g.adjacency(q, layout, O.ADJ_PULL)
            f0 = O.to_storage(fc, q, layout, aa_even_ready=True).astype(dt)
            a = f0.copy()
            O.run_aa(q, layout, n, a, adj, 5, omega)
            b = f0.copy()
            O.ref_run(f"aa_{L}_base", q, prec, omega, n, b, adj, 5)
            assert np.array_equal(a, b), ("aa", L, order)
            adjp = g.adjacency(q, layout, O.ADJ_PUSH)
            f0 = O.to_storage(fc, q, layout).astype(dt)
            a = O.run_ab(q, layout, n, f0, adjp, 4, omega)
            b = f0.copy()
            O.ref_run(f"ab_{L}_base", q, prec, omega, n, b, adjp, 4)
            assert np.array_equal(a, b), ("ab push", L, order)
            if layout == O.SOA:
                a = O.run_ab(q, layout, n, f0, adj, 4, omega, pull=True)
                b = f0.copy()
                O.ref_run("ab_soa_pull", q, prec, omega, n, b, adj, 4)
                assert np.array_equal(a, b), ("ab pull", order)
                if q == 19:
                    a = O.run_unroll19("ab_push", n, f0, adjp, 4, omega)
                    b = f0.copy()
                    O.ref_run("ab_soa_unroll", q, prec, omega, n, b, adjp, 4)
                    assert np.array_equal(a, b), ("ab unroll", order)
                    a = O.run_unroll19("ab_pull", n, f0, adj, 4, omega)
                    b = f0.copy()
                    O.ref_run("ab_soa_unroll_pull", q, prec, omega, n, b, adj, 4)
                    assert np.array_equal(a, b), ("ab unroll pull", order)
                    fa = O.to_storage(fc, q, layout, aa_even_ready=True).astype(dt)
                    a = fa.copy()
                    O.run_unroll19("aa", n, a, adj, 5, omega)
                    b = fa.copy()
                    O.ref_run("aa_soa_unroll", q, prec, omega, n, b, adj, 5)
                    assert np.array_equal(a, b), ("aa unroll", order)


def test_reference_lattice_tables():
    import ctypes
    for q in (19, 27):
        if not O.ref_available(q, "f64", 1.0):
            pytest.skip("reference build not present")
        L = O.ref_lib(q, "f64", 1.0)
        dirs = np.zeros(3 * q, np.int32)
        opp = np.zeros(q, np.int32)
        w = np.zeros(q, np.float64)
        assert L.ref_lattice(dirs.ctypes.data_as(ctypes.c_void_p),
                             opp.ctypes.data_as(ctypes.c_void_p),
                             w.ctypes.data_as(ctypes.c_void_p)) == q
        d2, o2, w2 = O.lattice(q)
        assert np.array_equal(dirs.reshape(q, 3), d2)
        assert np.array_equal(opp, o2)
        assert np.array_equal(w, w2)
