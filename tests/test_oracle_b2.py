"""Pins the oracle's B2 descriptor (b2.cpp:76-399): the dimer closed form
B_{n1 n2 l} = g_n1 g_n2 (2l+1)/(4 pi), rotation invariance
(ctests/test_descriptor.cpp:20-71), the dense derivatives against finite
differences, and the species-sorted packing."""
import numpy as np
import pytest

from helpers import random_cell


def b2_index(nr, lmax):
    return [(n1, n2, l) for n1 in range(nr) for n2 in range(n1, nr) for l in range(lmax + 1)]


def test_dimer_closed_form(oracle):
    # Two atoms of different species in a large box: each sees one bond.
    cell = np.eye(3) * 30.0
    pos = np.array([[5.0, 5.0, 5.0], [6.1, 5.7, 4.2]])
    species = np.array([0, 1], np.int32)
    nos, N, lmax, rc = 2, 5, 4, 4.0
    nb = oracle.Neighbors(cell, species, pos, rc, "literal")
    d = oracle.B2Values(nb, nos, N, lmax, np.full((2, 2), rc))
    v = pos[1] - pos[0]
    g = oracle.radial(*v, rc, N)[0]
    for s, atom in ((0, 0), (1, 1)):
        B = d.descriptors[s][0]
        nbr_species = 1 - s
        idx = b2_index(nos * N, lmax)
        expect = np.zeros(len(idx))
        for k, (n1, n2, l) in enumerate(idx):
            s1, r1 = divmod(n1, N)
            s2, r2 = divmod(n2, N)
            if s1 == nbr_species and s2 == nbr_species:
                expect[k] = g[r1] * g[r2] * (2 * l + 1) / (4 * np.pi)
        np.testing.assert_allclose(B, expect, atol=1e-14, rtol=1e-12)
        assert d.descriptor_norms[s][0] == pytest.approx(np.linalg.norm(expect), rel=1e-13)


def rotation(a, b, c):
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


def test_rotation_invariance(oracle):
    cell, species, pos = random_cell(10, 21, nos=3, scale=10.0, skew=0.0)
    R = rotation(1.28, -3.21, 0.42)  # ctests/test_descriptor.cpp:20-31
    args = (3, 3, 3, np.full((3, 3), 5.0))
    d1 = oracle.B2Values(oracle.Neighbors(cell, species, pos, 5.0), *args)
    d2 = oracle.B2Values(oracle.Neighbors(cell @ R.T, species, pos @ R.T, 5.0), *args)
    for s in range(3):
        np.testing.assert_allclose(d1.descriptors[s], d2.descriptors[s], atol=1e-10)


def test_force_dervs_fd(oracle):
    cell, species, pos = random_cell(8, 3, nos=2, scale=6.0)
    cut = np.array([[4.5, 4.0], [4.0, 4.2]])
    args = (2, 3, 3, cut)

    def desc(p):
        return oracle.B2Values(oracle.Neighbors(cell, species, p, 4.5, "literal"), *args)

    d = desc(pos)
    h = 1e-5
    # dB_i / dr_j for atom i = atom_indices[s][0] and each of its bonds.
    for s in range(2):
        if d.n_clusters_by_type[s] == 0:
            continue
        i = d.atom_indices[s][0]
        nn = d.neighbor_counts[s][0]
        c0 = d.cumulative_neighbor_counts[s][0]
        dB = d.descriptor_force_dervs[s]
        for b in range(min(nn, 3)):
            j = d.neighbor_indices[s][c0 + b]
            if j == i:
                continue
            # Sum the derivative rows of every bond of atom i that points at j
            # (several images may); moving j moves all of them.
            rows = [c0 + q for q in range(nn) if d.neighbor_indices[s][c0 + q] == j]
            for comp in range(3):
                pp = pos.copy()
                pp[j, comp] += h
                pm = pos.copy()
                pm[j, comp] -= h
                Bp = desc(pp).descriptors[s][0]
                Bm = desc(pm).descriptors[s][0]
                fd = (Bp - Bm) / (2 * h)
                an = sum(dB[3 * r + comp] for r in rows)
                np.testing.assert_allclose(an, fd, atol=1e-6 * max(1, np.abs(fd).max()))


def test_force_dots_and_packing(oracle):
    cell, species, pos = random_cell(15, 8, nos=3, scale=7.0)
    d = oracle.B2Values(oracle.Neighbors(cell, species, pos, 5.0), 3, 2, 2, np.full((3, 3), 5.0))
    for s in range(3):
        dB, B = d.descriptor_force_dervs[s], d.descriptors[s]
        fdots = d.descriptor_force_dots[s]
        for r in range(d.n_clusters_by_type[s]):
            c0, nn = d.cumulative_neighbor_counts[s][r], d.neighbor_counts[s][r]
            np.testing.assert_allclose(fdots[3 * c0:3 * (c0 + nn)], dB[3 * c0:3 * (c0 + nn)] @ B[r],
                                       rtol=1e-12, atol=1e-14)
        ai = d.atom_indices[s]
        assert list(ai) == sorted(ai)
        assert np.all(species[ai] == s)
        cnc = d.cumulative_neighbor_counts[s]
        assert np.all(np.diff(cnc) == d.neighbor_counts[s][:-1])
    assert d.cumulative_type_count == [0] + list(np.cumsum(d.n_clusters_by_type))
