"""The Lennard-Jones labeller of the 108-atom training frames: forces and
stress are the finite-difference derivatives of its energy (forces
-dE/dx; FLARE stress -(1/V) dE/d(strain) with the strain cell(:, a) +=
cell(:, b) e), the forces sum to zero, and the dimer energy is the
textbook shifted potential."""
import numpy as np
import pytest

from flare_amd.workloads import fcc_al_108, lennard_jones


@pytest.fixture(scope="module")
def frame():
    cell, species, pos = fcc_al_108(frame=3)
    return cell, pos, lennard_jones(cell, pos)


def test_forces_fd(frame):
    cell, pos, (E, F, S) = frame
    h = 1e-5
    for a, c in ((0, 0), (17, 1), (80, 2)):
        pp, pm = pos.copy(), pos.copy()
        pp[a, c] += h
        pm[a, c] -= h
        fd = -(lennard_jones(cell, pp)[0] - lennard_jones(cell, pm)[0]) / (2 * h)
        assert F[a, c] == pytest.approx(fd, abs=1e-6)
    assert np.abs(F.sum(0)).max() < 1e-10


def test_stress_fd(frame):
    cell, pos, (E, F, S) = frame
    V = abs(np.linalg.det(cell))
    h = 1e-6
    for q, (a, b) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        es = []
        for t in (h, -h):
            c, p = cell.copy(), pos.copy()
            c[:, a] += c[:, b] * t
            p[:, a] += p[:, b] * t
            es.append(lennard_jones(c, p)[0])
        assert S[q] * V == pytest.approx(-(es[0] - es[1]) / (2 * h), rel=1e-6, abs=1e-6)


def test_dimer():
    cell = np.eye(3) * 30.0
    pos = np.array([[0.0, 0, 0], [3.1, 0, 0]])
    E, F, S = lennard_jones(cell, pos)
    x = 2.6 / 3.1
    shift = 4 * 0.4 * ((2.6 / 7.8) ** 12 - (2.6 / 7.8) ** 6)
    assert E == pytest.approx(4 * 0.4 * (x ** 12 - x ** 6) - shift, rel=1e-14)
    dphi = -24 * 0.4 * (2 * x ** 12 - x ** 6) / 3.1
    assert F[1, 0] == pytest.approx(-dphi, rel=1e-13)
    assert S[0] == pytest.approx(-dphi * 3.1 / 30.0 ** 3, rel=1e-13)
