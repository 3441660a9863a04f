"""Many structures in one pass: a batch holds structures of uneven sizes and
cells in one device object per descriptor calculator, so every stage
(neighbour list, descriptors, kernel contraction, adjoint, variance) is one
launch for the whole batch. Kernel k of the model acts on the descriptors
of calculator k (sparse_gp.cpp:651-660); means add up over kernels and the
local variances are kept per kernel (sparse_gp.cpp:731-757). Each
structure's results equal those of evaluating it alone, bit for bit: rows
of the contractions sum in the same order whatever their neighbours in a
tile, and per-structure reductions use chunks aligned to the structure.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FORM_FP64


class Batch:
    """Structures (cell, species, positions) with the model's descriptor
    calculators (one per kernel; a single calculator may be passed alone)."""

    def __init__(self, structures, cutoff, calcs, form=FORM_FP64, context=None):
        self._ctx = context or _lib.default_context()
        self.form = form
        if not isinstance(calcs, (list, tuple)):
            calcs = [calcs]
        self.calcs = list(calcs)
        self.n_atoms = np.array([len(s[1]) for s in structures], np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.n_atoms)]).astype(np.int64)
        cells = _lib.f64(np.stack([np.asarray(s[0], np.float64).reshape(9) for s in structures]))
        species = _lib.i32(np.concatenate([np.asarray(s[1]) for s in structures]))
        pos = _lib.f64(np.concatenate([np.asarray(s[2], np.float64).reshape(-1, 3)
                                       for s in structures]))
        L = _lib.load()
        self._hs = []
        for calc in self.calcs:
            h = C.c_void_p()
            _lib.check(L.fm_batch_create(self._ctx.h, len(structures), _lib.ptr(cells),
                                         _lib.ptr(self.n_atoms), _lib.ptr(species),
                                         _lib.ptr(pos), float(cutoff), C.byref(h)))
            self._hs.append(h)
            settings, keep = calc._settings()
            fn = (L.fm_structure_compute_b3 if calc.descriptor_name == "B3"
                  else L.fm_structure_compute_b2)
            _lib.check(fn(h, C.byref(settings), int(form)))
            del keep

    @property
    def _h(self):
        return self._hs[0]

    def __len__(self):
        return len(self.n_atoms)

    def predict(self, sgp, local=True):
        """Per-structure mean_efs (list) and local variances: a list over
        structures of lists over kernels, each in the structure's
        species-sorted order (None without local)."""
        sgp._check_state()
        if len(self._hs) < sgp.n_kernels:
            raise ValueError("the batch has fewer descriptor calculators than the model has "
                             "kernels")
        flags = _lib.PREDICT_MEAN | (_lib.PREDICT_LOCAL if local else 0)
        n = int(self.offsets[-1])
        total = int(np.sum(1 + 3 * self.n_atoms.astype(np.int64) + 6))
        mean = np.zeros(total)
        var = []
        L = _lib.load()
        for k in range(sgp.n_kernels):
            m = sgp._model(k, self._ctx)
            out = np.zeros(total)
            v = np.zeros(n) if local else None
            _lib.check(L.fm_predict(m.h, self._hs[k], flags, int(self.form), _lib.ptr(out),
                                    _lib.ptr(v), None))
            mean += out
            var.append(v)
        means, vars_ = [], []
        o = 0
        for i, na in enumerate(self.n_atoms):
            means.append(mean[o:o + 1 + 3 * na + 6])
            o += 1 + 3 * na + 6
            if local:
                a, b = self.offsets[i], self.offsets[i + 1]
                vars_.append([v[a:b] for v in var])
        return means, (vars_ if local else None)

    def __del__(self):
        try:
            for h in self._hs:
                _lib.load().fm_structure_destroy(h)
        except Exception:
            pass
