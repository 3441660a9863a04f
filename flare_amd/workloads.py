"""Seeded synthetic structures and models for the benchmark workloads and
tests (there are no datasets in this environment).

  fcc_al_108      108-atom single-species FCC (a = 4.05 A), Gaussian
                  displacements sigma = 0.05 A, rng seed 1000 + frame.
  fcc_alloy_32k   32,000-atom FCC (a = 4.0 A, 20^3 cells), two species 50/50
                  by random permutation (seed 2), displacements 0.08 A.
  lj_training_frames  20 of the 108-atom frames labelled by Lennard-Jones
                  (eps 0.4 eV, sigma 2.6 A, cutoff 7.8 A, shifted to 0).
"""
import numpy as np

FCC_BASIS = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], dtype=np.float64)
BCC_BASIS = np.array([[0, 0, 0], [0.5, 0.5, 0.5]], dtype=np.float64)


def lattice(basis, a, reps):
    nx, ny, nz = reps
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    g = g.reshape(-1, 1, 3)
    pos = (g + basis[None, :, :]).reshape(-1, 3) * a
    cell = np.diag([nx * a, ny * a, nz * a]).astype(np.float64)
    return cell, pos


def fcc_al_108(frame=0, sigma=0.05):
    """108-atom FCC aluminium frames: seed 1000 + frame."""
    cell, pos = lattice(FCC_BASIS, 4.05, (3, 3, 3))
    rng = np.random.default_rng(1000 + frame)
    pos = pos + rng.normal(scale=sigma, size=pos.shape)
    return cell, np.zeros(len(pos), np.int32), pos


def fcc_alloy_32k(sigma=0.08):
    """32,000-atom two-species FCC alloy, species by permutation (seed 2)."""
    cell, pos = lattice(FCC_BASIS, 4.0, (20, 20, 20))
    n = len(pos)
    rng = np.random.default_rng(2)
    species = np.zeros(n, np.int32)
    species[rng.permutation(n)[: n // 2]] = 1
    pos = pos + np.random.default_rng(22).normal(scale=sigma, size=pos.shape)
    return cell, species, pos


def random_alloy(n_cells, a=4.0, nos=2, sigma=0.08, seed=0, basis=FCC_BASIS):
    cell, pos = lattice(basis, a, n_cells)
    rng = np.random.default_rng(seed)
    species = rng.integers(0, nos, len(pos)).astype(np.int32)
    pos = pos + rng.normal(scale=sigma, size=pos.shape)
    return cell, species, pos


def lennard_jones(cell, positions, eps=0.4, sig=2.6, rc=7.8):
    """Shifted Lennard-Jones energy, forces and stress (the labels of the
    108-atom training frames), brute force over the 27 nearest images of
    the minimum-image vector (cells wider than rc). The stress is in FLARE's
    order and sign (xx, xy, xz, yy, yz, zz) = -(1/V) dE/d(strain), i.e.
    minus ASE's stress (otf.py:633-647 converts between them)."""
    pos = np.asarray(positions)
    n = len(pos)
    inv = np.linalg.inv(cell)
    shift = 4 * eps * ((sig / rc) ** 12 - (sig / rc) ** 6)
    E = 0.0
    F = np.zeros_like(pos)
    W = np.zeros((3, 3))
    reps = [np.array([i, j, k]) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)]
    for i in range(n):
        d = pos - pos[i]
        d = d - np.round(d @ inv) @ cell
        for r in reps:
            dd = d + r @ cell
            rr = np.linalg.norm(dd, axis=1)
            m = (rr > 1e-9) & (rr < rc)
            x = sig / rr[m]
            E += 0.5 * np.sum(4 * eps * (x ** 12 - x ** 6) - shift)
            fmag = 24 * eps * (2 * x ** 12 - x ** 6) / rr[m] ** 2
            F[i] -= np.sum(fmag[:, None] * dd[m], axis=0)
            W += 0.5 * (fmag[:, None] * dd[m]).T @ dd[m]
    V = abs(np.linalg.det(cell))
    S = W[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]] / V
    return E, F, S


def lj_training_frames(n_frames=20):
    """The 108-atom FCC training frames (seeds 1000 + k) with their LJ
    energy, forces and stress: [(cell, species, positions, E, F, S)]."""
    out = []
    for k in range(n_frames):
        cell, species, pos = fcc_al_108(frame=k)
        E, F, S = lennard_jones(cell, pos)
        out.append((cell, species, pos, E, F, S))
    return out


def batch_structures(n_structures=512, seed=5, n_min=16, n_max=2048, nos=3):
    """Batch of structures of uneven sizes: atom counts log-uniform over [n_min, n_max];
    species uniform over nos; density 0.05-0.09 atoms/A^3; a 50% coin flip
    between bulk (random triclinic cell) and slab (10-15 A of vacuum along
    the third lattice vector). Sites come from an FCC lattice scaled to the
    density with random vacancies down to N atoms, displaced by 0.1 A."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_structures):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        rho = rng.uniform(0.05, 0.09)
        # FCC supercell with at least n sites (4 per conventional cell).
        reps = np.ones(3, int)
        while 4 * reps.prod() < n * rng.uniform(1.0, 1.05) + 1e-9:
            reps[np.argmin(reps)] += 1
        n_sites = 4 * int(reps.prod())
        a = (4.0 / rho) ** (1.0 / 3.0)
        cell, pos = lattice(FCC_BASIS, a, tuple(reps))
        keep = np.sort(rng.choice(n_sites, n, replace=False))
        pos = pos[keep]
        # Random triclinic strain (shear up to 10%).
        strain = np.eye(3) + np.triu(rng.uniform(-0.1, 0.1, (3, 3)), 1)
        cell = cell @ strain
        pos = pos @ strain
        if rng.uniform() < 0.5:
            c = cell[2] / np.linalg.norm(cell[2])
            cell = cell.copy()
            cell[2] = cell[2] + c * rng.uniform(10.0, 15.0)
        pos = pos + rng.normal(scale=0.1, size=pos.shape)
        species = rng.integers(0, nos, n).astype(np.int32)
        out.append((cell, species, pos))
    return out
