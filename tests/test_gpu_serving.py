"""Serving with continuous batching: requests arriving as a Poisson stream
are admitted into batches as workers free up; every result equals the
structure evaluated alone (bit for bit), and latencies are recorded."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workers,cap", [(2, None), (3, 1200)])
def test_serving_results_equal_alone(workers, cap):
    from flare_amd import _C_flare as C
    from flare_amd.serving import Server, latency_stats, poisson_run
    from flare_amd.workloads import batch_structures
    structs = batch_structures(n_structures=24, seed=23, n_min=32, n_max=400, nos=3)
    cut = np.full((3, 3), 5.5)
    calc = C.B2("chebyshev", "quadratic", [0.0, 5.5], [], [3, 4, 3], cut)
    first = C.Structure(*structs[0], 5.5, [calc], form=1)
    sgp = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    sgp.add_all_environments(first)
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(6).normal(scale=0.1, size=sgp.n_sparse))
    # A small budget forces several batches per burst.
    srv = Server(sgp, calc, 5.5, form=1, hbm_budget=200 << 20, n_workers=workers,
                 max_batch_atoms=cap)
    try:
        reqs, wall = poisson_run(srv, structs, rate=400.0, seed=1)
    finally:
        srv.close()
    stats = latency_stats(reqs)
    print("serving:", stats, "batches", len(srv.batches), flush=True)
    assert stats["n"] == len(structs)
    assert len(srv.batches) >= 2
    if cap is not None:  # a batch exceeds the cap only when it is a single request
        assert all(b[1] <= cap or b[0] == 1 for b in srv.batches)
    for r, (cell, species, pos) in zip(reqs, structs):
        mean, var = r.future.result()
        s = C.Structure(cell, species, pos, 5.5, [calc], form=1)
        sgp.predict_local_uncertainties(s)
        assert np.array_equal(mean, s.mean_efs)
        assert np.array_equal(var, s.local_uncertainties[0])


def test_oversized_request_is_rejected():
    from flare_amd import _C_flare as C
    from flare_amd import _lib
    from flare_amd.serving import Server
    from flare_amd.workloads import batch_structures
    structs = batch_structures(n_structures=2, seed=3, n_min=64, n_max=64, nos=1)
    calc = C.B2("chebyshev", "quadratic", [0.0, 5.0], [], [1, 3, 2], np.full((1, 1), 5.0))
    s = C.Structure(*structs[0], 5.0, [calc], form=1)
    sgp = C.SparseGP([C.NormalizedDotProduct(1.0, 2.0)], 0.01, 0.05, 0.005)
    sgp.add_all_environments(s)
    sgp.update_L_inverse()
    sgp.set_alpha(np.ones(sgp.n_sparse) * 0.1)
    srv = Server(sgp, calc, 5.0, form=1, hbm_budget=1 << 16, n_workers=1)
    try:
        r = srv.submit(*structs[1])
        with pytest.raises(_lib.FlareMIError):
            r.future.result(timeout=30)
    finally:
        srv.close()
