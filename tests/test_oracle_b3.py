"""Pins the oracle's B3 descriptor (b3.cpp:80-443): its Wigner 3j symbols
against sympy (the generator of wigner3j.cpp, dev/derivations/
compute_wigner.py), the count of (l1, l2, l3) triples against the table of
b3.cpp:200-211, the complex harmonics against scipy (Condon-Shortley phase,
y_grad.cpp:24-665 ordering) with finite-difference gradients, the dimer
closed form of the bispectrum, rotation invariance, and the dense
derivatives against finite differences."""
import numpy as np
import pytest

from helpers import random_cell


def test_wigner3j_matches_sympy(oracle):
    from sympy.physics.wigner import wigner_3j
    L = 4
    n = 0
    for l1 in range(L + 1):
        for l2 in range(L + 1):
            for l3 in range(L + 1):
                for m1 in range(-l1, l1 + 1):
                    for m2 in range(-l2, l2 + 1):
                        m3 = -m1 - m2
                        if abs(m3) > l3:
                            continue
                        ref = float(wigner_3j(l1, l2, l3, m1, m2, m3))
                        assert oracle.wigner3j(l1, l2, l3, m1, m2, m3) == pytest.approx(
                            ref, abs=1e-14)
                        n += 1
    assert n > 1000
    # Selection rules.
    assert oracle.wigner3j(1, 1, 1, 1, 1, -1) == 0.0
    assert oracle.wigner3j(1, 1, 3, 0, 0, 0) == 0.0


def test_number_of_l_triples(oracle):
    # b3.cpp:200-211 (n_ls for lmax = 0..4).
    assert [oracle.b3_n_ls(l) for l in range(5)] == [1, 5, 15, 34, 65]


def test_complex_harmonics_match_scipy(oracle):
    from scipy.special import sph_harm_y
    rng = np.random.default_rng(3)
    L = 4
    for _ in range(20):
        v = rng.normal(size=3) * rng.uniform(0.5, 3.0)
        r = np.linalg.norm(v)
        theta, phi = np.arccos(v[2] / r), np.arctan2(v[1], v[0])
        Y = oracle.complex_sph_harm(*v, L)[0]
        for l in range(L + 1):
            for m in range(-l, l + 1):
                ref = sph_harm_y(l, m, theta, phi)
                assert abs(Y[l * l + m + l] - ref) < 1e-13, (l, m)
    # y_grad.cpp:63-75: index 1 is sqrt(3 / 2 pi) (x - i y) / (2 r).
    x, y, z = 0.3, -0.7, 1.1
    r = np.sqrt(x * x + y * y + z * z)
    Y = oracle.complex_sph_harm(x, y, z, 1)[0]
    assert Y[1] == pytest.approx(np.sqrt(3 / (2 * np.pi)) * (x - 1j * y) / (2 * r), abs=1e-15)
    assert Y[3] == pytest.approx(-np.sqrt(3 / (2 * np.pi)) * (x + 1j * y) / (2 * r), abs=1e-15)


def test_complex_harmonic_gradients_fd(oracle):
    p = np.array([0.4, -1.2, 0.9])
    Y, Yx, Yy, Yz = oracle.complex_sph_harm(*p, 4)
    h = 1e-6
    for comp, G in enumerate((Yx, Yy, Yz)):
        pp, pm = p.copy(), p.copy()
        pp[comp] += h
        pm[comp] -= h
        fd = (oracle.complex_sph_harm(*pp, 4)[0] - oracle.complex_sph_harm(*pm, 4)[0]) / (2 * h)
        np.testing.assert_allclose(G, fd, atol=1e-8)


def b3_index(nr, lmax):
    ls = [(l1, l2, l3) for l1 in range(lmax + 1) for l2 in range(lmax + 1)
          for l3 in range(lmax + 1) if abs(l1 - l2) <= l3 <= l1 + l2]
    return [(n1, n2, n3, t) for n1 in range(nr) for n2 in range(n1, nr)
            for n3 in range(n2, nr) for t in ls]


def test_dimer_closed_form(oracle):
    # One bond: c[n, lm] = g_n Y_lm(u), and the rotation to u = z leaves only
    # m = 0, so B3 = g_n1 g_n2 g_n3 (l1 l2 l3; 0 0 0) prod sqrt((2 l + 1) / 4 pi).
    cell = np.eye(3) * 30.0
    pos = np.array([[5.0, 5.0, 5.0], [6.1, 5.7, 4.2]])
    species = np.array([0, 1], np.int32)
    nos, N, lmax, rc = 2, 3, 3, 4.0
    d = oracle.B3Values(oracle.Neighbors(cell, species, pos, rc, "literal"), nos, N, lmax,
                        np.full((2, 2), rc))
    g = oracle.radial(*(pos[1] - pos[0]), rc, N)[0]
    idx = b3_index(nos * N, lmax)
    assert d.n_descriptors == len(idx)
    for s in (0, 1):
        other = 1 - s
        expect = np.zeros(len(idx))
        for k, (n1, n2, n3, (l1, l2, l3)) in enumerate(idx):
            sp = [n // N for n in (n1, n2, n3)]
            if sp != [other] * 3:
                continue
            gg = g[n1 % N] * g[n2 % N] * g[n3 % N]
            w = oracle.wigner3j(l1, l2, l3, 0, 0, 0)
            expect[k] = gg * w * np.prod([np.sqrt((2 * l + 1) / (4 * np.pi))
                                          for l in (l1, l2, l3)])
        np.testing.assert_allclose(d.descriptors[s][0], expect, atol=1e-14, rtol=1e-12)


def rotation(a, b, c):
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


def test_rotation_invariance(oracle):
    cell, species, pos = random_cell(10, 21, nos=2, scale=9.0, skew=0.0)
    R = rotation(1.28, -3.21, 0.42)
    args = (2, 2, 3, np.full((2, 2), 4.5))
    d1 = oracle.B3Values(oracle.Neighbors(cell, species, pos, 4.5), *args, with_dervs=False)
    d2 = oracle.B3Values(oracle.Neighbors(cell @ R.T, species, pos @ R.T, 4.5), *args,
                         with_dervs=False)
    for s in range(2):
        np.testing.assert_allclose(d1.descriptors[s], d2.descriptors[s], atol=1e-11)


def test_force_dervs_fd(oracle):
    cell, species, pos = random_cell(7, 3, nos=2, scale=6.0)
    cut = np.array([[4.5, 4.0], [4.0, 4.2]])
    args = (2, 2, 2, cut)

    def desc(p):
        return oracle.B3Values(oracle.Neighbors(cell, species, p, 4.5, "literal"), *args)

    d = desc(pos)
    h = 1e-5
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
            rows = [c0 + q for q in range(nn) if d.neighbor_indices[s][c0 + q] == j]
            for comp in range(3):
                pp, pm = pos.copy(), pos.copy()
                pp[j, comp] += h
                pm[j, comp] -= h
                fd = (desc(pp).descriptors[s][0] - desc(pm).descriptors[s][0]) / (2 * h)
                an = sum(dB[3 * r + comp] for r in rows)
                np.testing.assert_allclose(an, fd, atol=1e-6 * max(1, np.abs(fd).max()))
        fd_dots = d.descriptor_force_dots[s][:3 * nn]
        np.testing.assert_allclose(fd_dots, dB[:3 * nn] @ d.descriptors[s][0], rtol=1e-12,
                                   atol=1e-14)
