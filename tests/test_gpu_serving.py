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
    assert srv.peak_inflight <= 200 << 20  # global admission across workers
    if cap is not None:  # a batch exceeds the cap only when it is a single request
        assert all(b[1] <= cap or b[0] == 1 for b in srv.batches)
    for r, (cell, species, pos) in zip(reqs, structs):
        mean, var = r.future.result()
        s = C.Structure(cell, species, pos, 5.5, [calc], form=1)
        sgp.predict_local_uncertainties(s)
        assert np.array_equal(mean, s.mean_efs)
        assert np.array_equal(var[0], s.local_uncertainties[0])


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


@pytest.mark.parametrize("form", [0, 1])
def test_native_server_results_equal_alone(form):
    """fm_submit / fm_poll (the library's own continuous-batching server):
    requests submitted from two host threads, polled without waiting and
    then waiting; every result equals the structure evaluated alone, for a
    two-kernel model (calculator k feeds kernel k); a request larger than
    the budget is refused with FM_ERR_CAPACITY."""
    import threading
    from flare_amd import _C_flare as C
    from flare_amd import _lib
    from flare_amd.serving import NativeServer
    from flare_amd.workloads import batch_structures
    print("native server, form", form, flush=True)
    structs = batch_structures(n_structures=20, seed=29, n_min=32, n_max=300, nos=3)
    cut = np.full((3, 3), 5.5)
    calcs = [C.B2("chebyshev", "quadratic", [0.0, 5.5], [], [3, 4, 3], cut),
             C.B2("chebyshev", "quadratic", [0.0, 4.0], [], [3, 3, 2], np.full((3, 3), 4.0))]
    first = C.Structure(*structs[0], 5.5, calcs, form=form)
    sgp = C.SparseGP([C.NormalizedDotProduct(2.0, 2.0), C.NormalizedDotProduct(1.1, 1.0)],
                     0.01, 0.05, 0.005)
    sgp.add_all_environments(first)
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(6).normal(scale=0.1, size=sgp.n_sparse))
    srv = NativeServer(sgp, calcs, 5.5, form=form, hbm_budget=300 << 20, n_workers=3,
                       max_batch_atoms=1500)
    try:
        ids = [None] * len(structs)

        def submit(lo, hi):
            for i in range(lo, hi):
                ids[i] = srv.submit(*structs[i])
        th = [threading.Thread(target=submit, args=(0, 10)),
              threading.Thread(target=submit, args=(10, 20))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        results = {}
        r0 = srv.poll(ids[0], timeout_ms=0)  # may or may not be done; does not block
        if r0 is not None:
            results[0] = r0
        for i in range(len(structs)):
            if i in results:
                continue
            r = srv.poll(ids[i], timeout_ms=-1)
            assert r is not None
            results[i] = r
        from flare_amd.workloads import random_alloy
        big = random_alloy((20, 20, 20), a=4.0, nos=3, sigma=0.05, seed=1)  # 32,000 atoms
        with pytest.raises(_lib.FlareMIError):
            srv.submit(*big)
    finally:
        srv.close()
    for i, (cell, species, pos) in enumerate(structs):
        s = C.Structure(cell, species, pos, 5.5, calcs, form=form)
        sgp.predict_local_uncertainties(s)
        mean, var = results[i]
        assert np.array_equal(mean, s.mean_efs), i
        for k in range(2):
            assert np.array_equal(var[k], s.local_uncertainties[k]), (i, k)
