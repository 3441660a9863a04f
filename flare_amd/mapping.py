"""The mapped model and its files: ``write_mapping_coefficients``
(sparse_gp.cpp:1200-1264), ``write_L_inverse`` and
``write_sparse_descriptors`` (sparse_gp.cpp:1341-1466), and the
pair_style flare potential read from a coefficient file and evaluated on
the GPU (pair_flare.cpp:68-184, 262-419).
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from ._C_flare import B2, FORM_FP64, Structure

_CUT_NAMES = {_lib.CUTOFF_QUADRATIC: "quadratic", _lib.CUTOFF_COSINE: "cosine"}


def _date():
    # ctime() without its newline, as sparse_gp.cpp:1208-1211.
    return time.ctime()


def _fmt(v):
    return "%.16e" % v


def _descriptor_header(calc, coeff_size):
    """B2::write_to_file (b2.cpp:52-74): cutoffs with 2 decimals."""
    nos, N, lmax = calc.descriptor_settings
    lines = [calc.radial_basis + "\n", f"{nos} {N} {lmax} {coeff_size}\n",
             calc.cutoff_function + "\n",
             "".join("%.2f " % c for c in calc.cutoffs.reshape(-1)) + "\n"]
    return "".join(lines)


def mapping_coefficients(sgp, kernel_index=0, context=None):
    """Per-species beta (n_species x beta_size) of a NormalizedDotProduct
    kernel, computed on the GPU."""
    ctx = context or _lib.default_context()
    m = sgp._model(kernel_index, ctx)
    L = _lib.load()
    bs = C.c_int64()
    _lib.check(L.fm_model_mapping_coefficients(m.h, None, C.byref(bs)))
    nos = sgp.sparse_descriptors[kernel_index].n_types
    beta = np.zeros((nos, bs.value))
    _lib.check(L.fm_model_mapping_coefficients(m.h, _lib.ptr(beta), C.byref(bs)))
    return beta


def write_mapping_coefficients(sgp, file_name, contributor, kernel_index, calc, date=None):
    """SparseGP::write_mapping_coefficients (sparse_gp.cpp:1200-1264)."""
    beta = mapping_coefficients(sgp, kernel_index)
    kern = sgp.kernels[kernel_index]
    out = [f"DATE: {date or _date()} CONTRIBUTOR: {contributor}\n",
           "%.0f %s\n" % (kern.power, kern.kernel_name),
           _descriptor_header(calc, beta.shape[1])]
    count = 0
    body = []
    for row in beta:
        if count != 0:
            body.append("\n")
        for v in row:
            body.append(("  " if v > 0 else " ") + _fmt(v))
            count += 1
            if count == 5:
                count = 0
                body.append("\n")
    with open(file_name, "w") as f:
        f.write("".join(out) + "".join(body))


def write_L_inverse(sgp, file_name, contributor, calcs, date=None):
    """SparseGP::write_L_inverse (sparse_gp.cpp:1341-1402)."""
    kern = sgp.kernels[0]
    hyps = sgp.hyperparameters
    out = [f"L_inverse_block file DATE: {date or _date()} CONTRIBUTOR: {contributor}\n",
           "%.0f %s\n" % (kern.power, kern.kernel_name), f"{len(hyps)}\n",
           "".join(_fmt(h) + " " for h in hyps) + "\n"]
    off = 0
    for i, cd in enumerate(sgp.sparse_descriptors):
        out.append(_descriptor_header(calcs[i], sgp.n_kernels))
        n = cd.n_clusters
        Lb = sgp.L_inv[off:off + n, off:off + n]
        off += n
        out.append(f"{n}\n")
        count = 1
        for j in range(n):
            for k in range(j + 1):
                out.append(_fmt(Lb[j, k]) + " ")
                if count % 5 == 0:
                    out.append("\n")
                count += 1
        if count % 5 != 0:
            out.append("\n")
    with open(file_name, "w") as f:
        f.write("".join(out))


def write_sparse_descriptors(sgp, file_name, contributor, date=None):
    """SparseGP::write_sparse_descriptors (sparse_gp.cpp:1404-1466)."""
    out = [f"sparse_descriptors file DATE: {date or _date()} CONTRIBUTOR: {contributor}\n",
           f"{sgp.n_kernels}\n"]
    for i, cd in enumerate(sgp.sparse_descriptors):
        out.append(f"{i} {cd.n_clusters} {cd.n_types}\n")
        normalize = "NormalizedDotProduct" in sgp.kernels[i].kernel_name
        count = 1
        for s in range(cd.n_types):
            out.append(f"{cd.n_clusters_by_type[s]}\n")
            for j in range(cd.n_clusters_by_type[s]):
                nrm = cd.descriptor_norms[s][j]
                for k in range(cd.n_descriptors):
                    if nrm < 1e-8:
                        v = 0.0
                    elif normalize:
                        v = cd.descriptors[s][j, k] / nrm
                    else:
                        v = cd.descriptors[s][j, k]
                    out.append(_fmt(v) + " ")
                    if count % 5 == 0:
                        out.append("\n")
                    count += 1
            if count % 5 != 1:
                out.append("\n")
    with open(file_name, "w") as f:
        f.write("".join(out))


class MappedPotential:
    """pair_style flare: a coefficient file (or arrays) resident on the
    GPU, evaluated on periodic structures."""

    def __init__(self, file_name=None, power=None, normalized=True, descriptor_settings=None,
                 cutoff_function="quadratic", cutoffs=None, beta=None, context=None):
        self._ctx = context or _lib.default_context()
        L = _lib.load()
        h = C.c_void_p()
        if file_name is not None:
            _lib.check(L.fm_mapped_read_file(self._ctx.h, str(file_name).encode(), C.byref(h)))
        else:
            nos, N, lmax = descriptor_settings
            cut = _lib.f64(np.asarray(cutoffs).reshape(-1))
            b = _lib.f64(np.asarray(beta).reshape(-1))
            P = C.POINTER(C.c_double)
            d = _lib.MappedDesc(int(power), 1 if normalized else 0, nos, N, lmax,
                                {"quadratic": 0, "cosine": 1}[cutoff_function],
                                cut.ctypes.data_as(P), b.ctypes.data_as(P))
            _lib.check(L.fm_mapped_create(self._ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        vals = [C.c_int() for _ in range(6)]
        _lib.check(L.fm_mapped_info(h, *[C.byref(v) for v in vals], None))
        self.power, self.normalized, nos, N, lmax, cut = [v.value for v in vals]
        cuts = np.zeros(nos * nos)
        _lib.check(L.fm_mapped_info(h, None, None, None, None, None, None, _lib.ptr(cuts)))
        self.cutoffs = cuts.reshape(nos, nos)
        self.descriptor_settings = [nos, N, lmax]
        self.cutoff_function = _CUT_NAMES[cut]
        self.cutoff = float(self.cutoffs.max())
        self.descriptor = B2("chebyshev", self.cutoff_function, [0.0, self.cutoff], [],
                             self.descriptor_settings, self.cutoffs)

    def compute(self, cell, species, positions):
        """Per-atom energies, forces (N x 3) and the virial
        sum_bonds -f (x) r (xx, xy, xz, yy, yz, zz)."""
        s = Structure(cell, species, positions, self.cutoff, [self.descriptor], form=FORM_FP64,
                      context=self._ctx)
        n = s.noa
        e = np.zeros(n)
        F = np.zeros((n, 3))
        v = np.zeros(6)
        _lib.check(_lib.load().fm_mapped_compute(self.h, s._h, _lib.ptr(e), _lib.ptr(F),
                                                 _lib.ptr(v)))
        return e, F, v

    def __del__(self):
        try:
            _lib.load().fm_mapped_destroy(self.h)
        except Exception:
            pass
