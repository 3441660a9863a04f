"""Pins the oracle's periodic geometry and neighbour lists
(structure.cpp:7-165): the literal O(N^2 * images) loop against an
independent numpy brute force, the cell-list build against the literal
loop bit for bit, pair symmetry, wrapping (ctests/test_structure.cpp:3-22)
and the sweep rule for small cells."""
import numpy as np
import pytest

from helpers import fcc, random_cell


def numpy_neighbors(cell, pos, cutoff, sweep):
    """Brute force in numpy with its own wrapping: sets of (i, j, s) keys."""
    frac = pos @ np.linalg.inv(cell)
    frac -= np.floor(frac)
    w = frac @ cell
    out = set()
    rng = range(-sweep, sweep + 1)
    for i in range(len(pos)):
        for j in range(len(pos)):
            for s1 in rng:
                for s2 in rng:
                    for s3 in rng:
                        im = w[j] - w[i] + s1 * cell[0] + s2 * cell[1] + s3 * cell[2]
                        d = np.linalg.norm(im)
                        if d < cutoff - 1e-9 and d > 1e-12:
                            out.add((i, j, s1, s2, s3))
                        elif abs(d - cutoff) <= 1e-9:
                            out.add(("edge", i, j, s1, s2, s3))
    return out


CASES = [
    ("fcc108", lambda: fcc(4.05, (3, 3, 3), 0.05, 1000), 5.0),
    ("triclinic", lambda: random_cell(30, 4, nos=3, scale=8.0, skew=1.5), 4.5),
    ("small_cell_sweep2", lambda: random_cell(3, 7, nos=2, scale=3.0, skew=0.2), 5.0),
    ("slab", lambda: (np.diag([9.0, 9.0, 25.0]),
                      np.zeros(20, np.int32),
                      np.random.default_rng(2).uniform(0, 1, (20, 3)) * [9, 9, 6]), 5.0),
]


@pytest.mark.parametrize("name,make,cutoff", CASES)
def test_literal_matches_numpy_bruteforce(oracle, name, make, cutoff):
    cell, species, pos = make()
    nb = oracle.Neighbors(cell, species, pos, cutoff, "literal")
    ref = numpy_neighbors(cell, pos, cutoff, nb.sweep)
    cum = nb.cumulative_neighbor_count
    got = set()
    img = nb.images
    for i in range(len(pos)):
        for q in range(cum[i], cum[i + 1]):
            got.add((i, int(nb.structure_indices[q]), *map(int, img[q])))
    strict = {k for k in ref if k[0] != "edge"}
    edge = {k[1:] for k in ref if k[0] == "edge"}
    assert strict <= got
    assert got - strict <= edge


@pytest.mark.parametrize("name,make,cutoff", CASES)
def test_cells_equal_literal_bitwise(oracle, name, make, cutoff):
    cell, species, pos = make()
    a = oracle.Neighbors(cell, species, pos, cutoff, "literal")
    b = oracle.Neighbors(cell, species, pos, cutoff, "cells")
    assert a.n_neighbors == b.n_neighbors and a.sweep == b.sweep
    for f in ("neighbor_count", "cumulative_neighbor_count", "structure_indices",
              "neighbor_species"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
    assert np.array_equal(a.relative_positions.view(np.int64), b.relative_positions.view(np.int64))


def test_list_order_and_contents(oracle):
    cell, species, pos = random_cell(12, 9, nos=2, scale=6.0)
    nb = oracle.Neighbors(cell, species, pos, 5.0, "literal")
    cum, si, img = nb.cumulative_neighbor_count, nb.structure_indices, nb.images
    rel = nb.relative_positions
    w = nb.wrapped_positions
    span = 2 * nb.sweep + 1
    for i in range(len(pos)):
        keys = [(si[q] * span + img[q, 0] + nb.sweep) * span ** 2
                + (img[q, 1] + nb.sweep) * span + img[q, 2] + nb.sweep
                for q in range(cum[i], cum[i + 1])]
        assert keys == sorted(keys)  # order j, s1, s2, s3 (structure.cpp:67-71)
        for q in range(cum[i], cum[i + 1]):
            im = w[si[q]] - w[i] + img[q] @ cell
            np.testing.assert_allclose(rel[q, 1:], im, atol=1e-12)
            assert rel[q, 0] == pytest.approx(np.linalg.norm(im), abs=1e-12)
            assert 0 < rel[q, 0] < 5.0
        assert np.all(nb.neighbor_species[cum[i]:cum[i + 1]] == species[si[cum[i]:cum[i + 1]]])


def test_pair_symmetry(oracle):
    cell, species, pos = random_cell(25, 13, nos=3, scale=7.5, skew=1.0)
    nb = oracle.Neighbors(cell, species, pos, 5.5, "cells")
    cum, si, img = nb.cumulative_neighbor_count, nb.structure_indices, nb.images
    pairs = set()
    for i in range(len(pos)):
        for q in range(cum[i], cum[i + 1]):
            pairs.add((i, int(si[q]), *map(int, img[q])))
    for (i, j, a, b, c) in pairs:
        assert (j, i, -a, -b, -c) in pairs


def test_wrapping(oracle):
    # ctests/test_structure.cpp:3-22: wrapped - original is a lattice vector.
    rng = np.random.default_rng(1)
    cell = np.array([[5.0, 0.2, 0.1], [0.3, 6.0, 0.0], [0.1, -0.4, 5.5]])
    pos = rng.uniform(-20, 20, (40, 3))
    nb = oracle.Neighbors(cell, np.zeros(40, np.int32), pos, 3.0, "literal")
    w = nb.wrapped_positions
    frac = w @ np.linalg.inv(cell)
    assert np.all(frac > -1e-12) and np.all(frac < 1 + 1e-12)
    d = (w - pos) @ np.linalg.inv(cell)
    np.testing.assert_allclose(d, np.round(d), atol=1e-10)


def test_sweep_and_volume(oracle):
    cell = np.diag([3.0, 3.5, 4.0])
    nb = oracle.Neighbors(cell, np.zeros(2, np.int32), np.array([[0, 0, 0], [1.5, 1.5, 2.0]]),
                          7.0, "literal")
    assert nb.single_sweep_cutoff == pytest.approx(3.0)
    assert nb.sweep == 3  # ceil(7 / 3)
    assert nb.volume == pytest.approx(42.0)
