"""fm_mapped_compute_noalloc, the allocation-free pair_style flare
evaluation: through the C-ABI its energies, forces and virials equal
fm_mapped_compute's bit for bit on two calls in a row (a two-structure
batch, then a smaller single structure, so the second call reuses the
scratch the first left behind), with device outputs as well as host ones,
and a structure beyond the scratch size is refused with FM_ERR_CAPACITY."""
import ctypes as C

import numpy as np
import pytest

from helpers import random_cell

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def lib():
    from flare_amd import _lib
    return _lib


def potential(power):
    """A mapped potential from a small two-species model (B2 [2, 4, 3])."""
    from flare_amd import _C_flare as F
    from flare_amd.mapping import MappedPotential, mapping_coefficients
    cut = np.array([[4.5, 4.0], [4.0, 4.2]])
    cell, species, pos = random_cell(30, 12, nos=2, scale=7.5, skew=0.5)
    calc = F.B2("chebyshev", "quadratic", [0.0, 4.5], [], [2, 4, 3], cut)
    s = F.Structure(cell, species, pos, 4.5, [calc])
    sgp = F.SparseGP([F.NormalizedDotProduct(1.5, power)], 0.01, 0.05, 0.005)
    d = s.descriptors[0]
    atoms = [int(d.atom_indices[t][r]) for t in range(2)
             for r in range(0, d.n_clusters_by_type[t], 3)]
    sgp.add_specific_environments(s, atoms)
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(3).normal(scale=0.1, size=sgp.n_sparse))
    beta = mapping_coefficients(sgp)
    return MappedPotential(power=int(power), descriptor_settings=[2, 4, 3], cutoffs=cut,
                           beta=beta), calc


def run(lib, fn, pot, h, n, n_struct):
    e, f, v = np.zeros(n), np.zeros((n, 3)), np.zeros(6 * n_struct)
    lib.check(fn(pot.h, h, lib.ptr(e), lib.ptr(f), lib.ptr(v)))
    return e, f, v


@pytest.mark.parametrize("power", [1.0, 2.0])
def test_noalloc_matches_mapped_compute(lib, power):
    from flare_amd.batch import Batch
    from flare_amd import _C_flare as F
    L = lib.load()
    pot, calc = potential(power)
    big = [random_cell(n, 40 + n, nos=2, scale=9.0, skew=0.4) for n in (180, 95)]
    small = random_cell(23, 7, nos=2, scale=6.0, skew=0.3)
    b = Batch(big, 4.5, [calc], form=lib.FORM_FP64)
    s = F.Structure(*small, 4.5, [calc])
    for label, h, n, ns in (("batch of 180 + 95 atoms", b._h, 275, 2),
                            ("single 23 atoms", s._h, 23, 1)):
        say(f"power {power}: {label}")
        want = run(lib, L.fm_mapped_compute, pot, h, n, ns)
        got = run(lib, L.fm_mapped_compute_noalloc, pot, h, n, ns)
        for w, g in zip(want, got):
            assert np.array_equal(w, g)
        assert np.all(np.isfinite(want[1])) and np.abs(want[1]).max() > 0


def test_noalloc_device_outputs(lib):
    """Outputs in device memory (allocated through the HIP runtime the
    library itself links)."""
    from flare_amd import _C_flare as F
    L = lib.load()
    hip = C.CDLL("libamdhip64.so.7")
    pot, calc = potential(2.0)
    cell, species, pos = random_cell(60, 9, nos=2, scale=8.0, skew=0.3)
    s = F.Structure(cell, species, pos, 4.5, [calc])
    want = run(lib, L.fm_mapped_compute, pot, s._h, 60, 1)
    sizes = [60 * 8, 180 * 8, 6 * 8]
    dev = [C.c_void_p() for _ in sizes]
    try:
        for p, nb in zip(dev, sizes):
            assert hip.hipMalloc(C.byref(p), C.c_size_t(nb)) == 0
        say("device outputs, 60 atoms")
        lib.check(L.fm_mapped_compute_noalloc(pot.h, s._h, *dev))
        got = [np.zeros(nb // 8) for nb in sizes]
        for g, p, nb in zip(got, dev, sizes):
            assert hip.hipMemcpy(C.c_void_p(g.ctypes.data), p, C.c_size_t(nb), 2) == 0
    finally:
        for p in dev:
            if p.value:
                hip.hipFree(p)
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1], want[1].reshape(-1))
    assert np.array_equal(got[2], want[2])


def test_noalloc_rejects_oversize(lib):
    from flare_amd import _C_flare as F
    from flare_amd.workloads import fcc_alloy_32k
    L = lib.load()
    pot, calc = potential(2.0)
    cell, species, pos = fcc_alloy_32k()
    s = F.Structure(cell, species, pos, 4.5, [calc])
    say("32,000 atoms: beyond the scratch array")
    with pytest.raises(lib.FlareMIError, match="allocation-free entry point holds"):
        run(lib, L.fm_mapped_compute_noalloc, pot, s._h, len(pos), 1)
    # The allocating entry point still evaluates it.
    e, f, v = run(lib, L.fm_mapped_compute, pot, s._h, len(pos), 1)
    assert np.all(np.isfinite(f))
