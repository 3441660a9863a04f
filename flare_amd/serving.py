"""Serving: a request loop with continuous batching on one GPU.

Requests (one structure each) are queued as they arrive. Each worker owns a
device context (its own HIP stream) with the model resident (alpha, the
sparse set and L_inv stay on the device for the life of the server) and
repeatedly admits the oldest waiting requests whose HBM footprint fits the
budget left beside the batches in flight on every worker, evaluates them as
one batch, and completes their futures.
With two workers a batch is admitted and built while the previous one is
still on the GPU, so their work overlaps on the device. A request's result
does not depend on which other requests share its batch (see
flare_amd/batch.py), so it equals evaluating the structure alone.
"""
import collections
import ctypes as C
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import _lib
from .batch import Batch


def footprint_bytes(n_atoms, n_neighbors, n_descriptors, n_rad_h, max_sparse, esz=4):
    """Device bytes a batch needs: neighbour list (40 B per entry), bond list
    and bond forces, and per atom the descriptors and the gradient rows, the
    density coefficients and their adjoint, the int8 slices of the
    descriptors, w (BF16 pairs) and k (FP64) against the largest species
    block of the sparse set, and small per-atom partials."""
    per_nb = 40 + 8 + 3 * esz + 8
    per_atom = (2 * esz * n_descriptors + 2 * esz * n_rad_h + 5 * n_descriptors
                + (4 + 8) * max_sparse + 256)
    return int(n_neighbors * per_nb + n_atoms * per_atom)


class Request:
    __slots__ = ("structure", "arrival", "future", "started", "finished", "need")

    def __init__(self, structure, arrival, need):
        self.structure = structure
        self.arrival = arrival
        self.future = Future()
        self.started = None
        self.finished = None
        self.need = need


class Server:
    """Continuous batching over n_workers contexts (HIP streams). Admission
    is global: the bytes of every batch in flight on any worker count
    against ``hbm_budget``. A request is sized at submission from its atom
    count and the neighbours within the largest cutoff at a density of at
    least ``max_density`` atoms / A^3 (or its own, if higher); once its batch
    is built the estimate is replaced by the measured neighbour count."""

    def __init__(self, sgp, calcs, cutoff, form=_lib.FORM_CHOSEN, hbm_budget=16 << 30,
                 n_workers=2, device=0, local=True, max_batch_atoms=None, max_density=0.1):
        self.sgp = sgp
        self.calcs = list(calcs) if isinstance(calcs, (list, tuple)) else [calcs]
        self.cutoff = cutoff
        self.form = form
        self.local = local
        self.budget = hbm_budget
        self.max_density = max_density
        # Optional cap on the atoms of one batch (latency against throughput:
        # smaller batches start sooner; the HBM budget still applies).
        self.max_batch_atoms = max_batch_atoms
        self._esz = 8 if form == _lib.FORM_FP64 else 4
        self._per_kernel = []
        for k, calc in enumerate(self.calcs[:sgp.n_kernels]):
            nos, N, lmax = calc.descriptor_settings
            ms = max(sgp.sparse_descriptors[k].n_clusters_by_type or [0])
            self._per_kernel.append((calc.n_descriptors, nos * N * (lmax + 1) ** 2, ms))
        self._queue = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self._inflight = 0
        self.peak_inflight = 0
        self._ctxs = [_lib.Context(device) for _ in range(n_workers)]
        for ctx in self._ctxs:  # resident model per stream
            for k in range(sgp.n_kernels):
                sgp._model(k, ctx)
        self.batches = []
        self._workers = [threading.Thread(target=self._work, args=(ctx,), daemon=True)
                         for ctx in self._ctxs]
        for w in self._workers:
            w.start()

    def _bytes(self, n_atoms, n_neighbors):
        return sum(footprint_bytes(n_atoms, n_neighbors, d, nh, ms, self._esz)
                   for d, nh, ms in self._per_kernel)

    def _estimate(self, cell, n):
        vol = abs(float(np.linalg.det(np.asarray(cell, dtype=np.float64).reshape(3, 3))))
        rho = max(self.max_density, n / vol if vol > 0 else 0.0)
        nb = rho * 4.0 / 3.0 * np.pi * self.cutoff ** 3
        return self._bytes(n, int(np.ceil(n * nb)))

    def submit(self, cell, species, positions):
        n = len(species)
        req = Request((cell, species, positions), time.perf_counter(), self._estimate(cell, n))
        if req.need > self.budget:
            req.future.set_exception(_lib.FlareMIError(
                _lib.FM_ERR_CAPACITY, f"request of {n} atoms exceeds the HBM budget"))
            return req
        with self._cv:
            self._queue.append(req)
            self._cv.notify()
        return req

    def _admit(self):
        """The oldest waiting requests that fit the budget beside the batches
        in flight (none when the oldest does not fit yet)."""
        batch, used, atoms = [], 0, 0
        cap = self.max_batch_atoms
        while self._queue:
            r = self._queue[0]
            n = len(r.structure[1])
            if self._inflight + used + r.need > self.budget:
                break
            if batch and cap is not None and atoms + n > cap:
                break
            batch.append(self._queue.popleft())
            used += r.need
            atoms += n
        return batch, used

    def _measured(self, b):
        """Bytes of a built batch from its neighbour count."""
        L = _lib.load()
        sweep, nnb = C.c_int(), C.c_int64()
        vol, ssc = C.c_double(), C.c_double()
        _lib.check(L.fm_structure_info(b._h, C.byref(sweep), C.byref(nnb), C.byref(vol),
                                       C.byref(ssc)))
        return self._bytes(int(b.offsets[-1]), int(nnb.value))

    def _work(self, ctx):
        while True:
            with self._cv:
                while True:
                    if self._stop and not self._queue:
                        return
                    reqs, need = self._admit()
                    if reqs:
                        break
                    self._cv.wait()
                self._inflight += need
                self.peak_inflight = max(self.peak_inflight, self._inflight)
            t0 = time.perf_counter()
            for r in reqs:
                r.started = t0
            try:
                b = Batch([r.structure for r in reqs], self.cutoff, self.calcs, form=self.form,
                          context=ctx)
                real = self._measured(b)
                with self._cv:
                    self._inflight += real - need
                    self.peak_inflight = max(self.peak_inflight, self._inflight)
                need = real
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
            finally:
                with self._cv:
                    self._inflight -= need
                    self._cv.notify_all()

    def close(self):
        """Stops the workers once the queue is drained, then releases the
        server's models and contexts."""
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for w in self._workers:
            w.join()
        for ctx in self._ctxs:
            for key in [k for k, m in self.sgp._models.items() if m.ctx is ctx]:
                del self.sgp._models[key]
            ctx.close()
        self._ctxs = []


class NativeServer:
    """The library's own continuous-batching server (fm_server_create,
    fm_submit, fm_poll in include/flare_mi.h): worker threads, contexts and
    the admission queue live in the native library; Python only submits
    and polls. Same arguments and results as Server."""

    def __init__(self, sgp, calcs, cutoff, form=_lib.FORM_CHOSEN, hbm_budget=16 << 30,
                 n_workers=2, device=0, local=True, max_batch_atoms=0, max_density=0.1):
        from ._C_flare import _DeviceModel
        sgp._check_state()
        calcs = list(calcs) if isinstance(calcs, (list, tuple)) else [calcs]
        K = sgp.n_kernels
        if len(calcs) < K:
            raise ValueError("one descriptor calculator per kernel is needed")
        descs, keep = [], []
        for k in range(K):
            d, kp = _DeviceModel.describe(sgp, k)
            descs.append(d)
            keep.append(kp)
        settings = []
        for calc in calcs[:K]:
            st, kp = calc._settings()
            settings.append(st)
            keep.append(kp)
        kinds = _lib.i32([_lib.DESCRIPTOR_B3 if c.descriptor_name == "B3" else _lib.DESCRIPTOR_B2
                          for c in calcs[:K]])
        models = (_lib.ModelDesc * K)(*descs)
        sets = (_lib.B2Settings * K)(*settings)
        flags = _lib.PREDICT_MEAN | (_lib.PREDICT_LOCAL if local else 0)
        d = _lib.ServerDesc(K, models, sets, kinds.ctypes.data_as(C.POINTER(C.c_int32)),
                            float(cutoff), int(form), flags, int(n_workers), int(hbm_budget),
                            int(max_batch_atoms or 0), float(max_density))
        h = C.c_void_p()
        _lib.check(_lib.load().fm_server_create(int(device), C.byref(d), C.byref(h)))
        self.h = h
        self.n_kernels = K
        self.local = local

    def submit(self, cell, species, positions):
        """Queues a structure; returns its request id."""
        cell = _lib.f64(np.asarray(cell).reshape(9))
        sp = _lib.i32(species)
        pos = _lib.f64(np.asarray(positions).reshape(-1))
        rid = C.c_int64()
        _lib.check(_lib.load().fm_submit(self.h, _lib.ptr(cell), len(sp), _lib.ptr(sp),
                                         _lib.ptr(pos), C.byref(rid)))
        return rid.value, len(sp)

    def poll(self, request, timeout_ms=-1):
        """(mean_efs, [local variances per kernel]) once done, else None."""
        rid, n = request
        mean = np.zeros(1 + 3 * n + 6)
        var = np.zeros(self.n_kernels * n) if self.local else None
        done = C.c_int()
        _lib.check(_lib.load().fm_poll(self.h, rid, int(timeout_ms), C.byref(done),
                                       _lib.ptr(mean), _lib.ptr(var)))
        if not done.value:
            return None
        return mean, ([var[k * n:(k + 1) * n] for k in range(self.n_kernels)]
                      if self.local else None)

    def close(self):
        if self.h is not None:
            _lib.load().fm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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
