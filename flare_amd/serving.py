"""Serving: a request loop with continuous batching on one GPU.

Requests (one structure each) are queued as they arrive. Each worker owns a
device context (its own HIP stream) with the model resident (alpha, the
sparse set and L_inv stay on the device for the life of the server) and
repeatedly admits the oldest waiting requests whose estimated HBM footprint
fits its budget, evaluates them as one batch, and completes their futures.
With two workers a batch is admitted and built while the previous one is
still on the GPU, so their work overlaps on the device. A request's result
does not depend on which other requests share its batch (see
flare_amd/batch.py), so it equals evaluating the structure alone.
"""
import collections
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import _lib
from .batch import Batch


def footprint_bytes(n_atoms, n_descriptors, max_sparse, neighbors_per_atom=64, esz=4):
    """Device bytes a batch of n_atoms needs: neighbour and bond lists,
    descriptors, the w and k matrices, G, bond forces and partials."""
    per_atom = (40 * neighbors_per_atom + 8 * neighbors_per_atom + esz * 2 * n_descriptors
                + (esz + 8) * max_sparse + 3 * esz * neighbors_per_atom + 200)
    return int(n_atoms * per_atom)


class Request:
    __slots__ = ("structure", "arrival", "future", "started", "finished")

    def __init__(self, structure, arrival):
        self.structure = structure
        self.arrival = arrival
        self.future = Future()
        self.started = None
        self.finished = None


class Server:
    def __init__(self, sgp, calc, cutoff, form=_lib.FORM_CHOSEN, hbm_budget=16 << 30,
                 n_workers=2, device=0, local=True, max_batch_atoms=None):
        self.sgp = sgp
        self.calc = calc
        self.cutoff = cutoff
        self.form = form
        self.local = local
        self.budget = hbm_budget
        # Optional cap on the atoms of one batch (latency against throughput:
        # smaller batches start sooner; the HBM budget still applies).
        self.max_batch_atoms = max_batch_atoms
        nd = calc.n_descriptors
        self.max_sparse = max(max(c.n_clusters_by_type) for c in sgp.sparse_descriptors)
        self._fp = lambda n: footprint_bytes(n, nd, self.max_sparse,
                                             esz=8 if form == _lib.FORM_FP64 else 4)
        self._queue = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self._ctxs = [_lib.Context(device) for _ in range(n_workers)]
        for ctx in self._ctxs:  # resident model per stream
            for k in range(sgp.n_kernels):
                sgp._model(k, ctx)
        self.batches = []
        self._workers = [threading.Thread(target=self._work, args=(ctx,), daemon=True)
                         for ctx in self._ctxs]
        for w in self._workers:
            w.start()

    def submit(self, cell, species, positions):
        req = Request((cell, species, positions), time.perf_counter())
        n = len(species)
        if self._fp(n) > self.budget:
            req.future.set_exception(_lib.FlareMIError(
                _lib.FM_ERR_CAPACITY, f"request of {n} atoms exceeds the HBM budget"))
            return req
        with self._cv:
            self._queue.append(req)
            self._cv.notify()
        return req

    def _admit(self):
        """The oldest waiting requests that fit the budget (at least one)."""
        batch, used, atoms = [], 0, 0
        cap = self.max_batch_atoms
        while self._queue:
            n = len(self._queue[0].structure[1])
            need = self._fp(n)
            if batch and (used + need > self.budget or (cap is not None and atoms + n > cap)):
                break
            batch.append(self._queue.popleft())
            used += need
            atoms += n
        return batch

    def _work(self, ctx):
        while True:
            with self._cv:
                while not self._queue and not self._stop:
                    self._cv.wait()
                if self._stop and not self._queue:
                    return
                reqs = self._admit()
            t0 = time.perf_counter()
            for r in reqs:
                r.started = t0
            try:
                b = Batch([r.structure for r in reqs], self.cutoff, self.calc, form=self.form,
                          context=ctx)
                means, vars_ = b.predict(self.sgp, local=self.local)
                del b
                t1 = time.perf_counter()
                for i, r in enumerate(reqs):
                    r.finished = t1
                    r.future.set_result((means[i], vars_[i] if vars_ is not None else None))
                self.batches.append((len(reqs), int(sum(len(r.structure[1]) for r in reqs)),
                                     t1 - t0))
            except Exception as e:  # deliver the failure to every request
                for r in reqs:
                    r.finished = time.perf_counter()
                    r.future.set_exception(e)

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for w in self._workers:
            w.join()


def poisson_run(server, structures, rate, seed=0):
    """Submit structures at Poisson arrival times (rate per second); wait for
    all; return (requests, wall seconds)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(structures))
    reqs = []
    t0 = time.perf_counter()
    t_next = t0
    for s, g in zip(structures, gaps):
        t_next += g
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        reqs.append(server.submit(*s))
    for r in reqs:
        r.future.exception()
    return reqs, time.perf_counter() - t0


def latency_stats(reqs):
    lat = np.array([r.finished - r.arrival for r in reqs if r.finished is not None])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "n": int(lat.size)}
