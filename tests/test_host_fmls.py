"""The segmenter's host tables (mrtrix3_amd/csrc/host/fmls.cpp through include/mrfixel.h, no GPU
needed) against independent constructions: the direction adjacency against scipy's convex hull of
the directions and their antipodes, the integration weights against numpy's least-squares QR,
the FastLookupSet grid and the Legendre table against the oracle, and the tessellation's own
invariants."""
import numpy as np
import pytest

from mrtrix3_amd import fixel
from oracle import orc


@pytest.fixture(scope="module")
def tables():
    return fixel.DirectionSet()


@pytest.fixture(scope="module")
def fm():
    return orc.FmlsOracle(8)


def test_unit_vectors_of_the_tessellation(tables, fm):
    assert tables.n == 1281
    np.testing.assert_allclose(tables.unit, fm.unit, rtol=0, atol=1e-15)
    np.testing.assert_allclose(np.linalg.norm(tables.unit, axis=1), 1.0, atol=1e-15)


def test_adjacency_matches_the_convex_hull(tables, fm):
    assert np.array_equal(tables.adj, fm.adj)


def test_integration_weights(tables, fm):
    np.testing.assert_allclose(tables.weights, fm.weights, rtol=1e-9, atol=0)
    assert abs(tables.weights.sum() - 4 * np.pi) < 1e-7   # the least-squares residual


def test_lookup_grid_matches_the_oracle(tables, fm):
    naz, nel, start, idx = fm.lookup_tables()
    assert (tables.n_az, tables.n_el) == (naz, nel)
    assert np.array_equal(tables.cell_start, start)
    assert np.array_equal(tables.cell_dirs, idx)


def test_legendre_table_matches_the_oracle():
    for lmax in (2, 8, 12):
        got = fixel.legendre_table(lmax, 2562)
        lib = orc.lib()
        import ctypes as C
        lib.orc_fmls_al_table.restype = C.c_int
        lib.orc_fmls_al_table.argtypes = [C.c_int, C.c_int, orc.D]
        ref = np.zeros_like(got)
        lib.orc_fmls_al_table(lmax, 2562, orc._d(ref))
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-15)
        # row 0 is the north pole: only m = 0 survives, P_l^0(1) = sqrt((2l+1) / 4 pi)
        for l in range(0, lmax + 1, 2):
            assert abs(got[0, l * l // 4] - np.sqrt((2 * l + 1) / (4 * np.pi))) < 1e-13


def test_custom_direction_set():
    # a 300-direction set: the tables are built for any spread set (the hull must give 5-6
    # neighbours, weights must integrate the constant)
    azel = np.loadtxt(orc.os.path.join(orc.os.path.dirname(orc.HERE), "mrtrix3_amd", "data",
                                       "electrostatic_repulsion_300.txt"))
    t = fixel.DirectionSet(azel)
    o = orc.FmlsOracle(8, azel=azel)
    assert t.n == 300
    assert np.array_equal(t.adj, o.adj)
    np.testing.assert_allclose(t.weights, o.weights, rtol=1e-8)
    # fewer directions, a coarser least-squares fit: the constant integrates to 4 pi within 2e-4
    assert abs(t.weights.sum() - 4 * np.pi) < 2e-4 * 4 * np.pi and np.all(t.weights > 0)
