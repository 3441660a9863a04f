"""The mapped model and its files: ``write_mapping_coefficients``
(sparse_gp.cpp:1200-1264), ``write_varmap_coefficients``
(sparse_gp.cpp:1266-1339, normalized_dot_product.cpp:769-830),
``write_L_inverse`` and ``write_sparse_descriptors``
(sparse_gp.cpp:1341-1466); the pair_style flare potential read from a
coefficient file and evaluated on the GPU (pair_flare.cpp:68-184, 262-419);
and compute flare/std/atom (compute_flare_std_atom.cpp:95-228, 329-686):
the readers of the L_inv / sparse-descriptor and variance-map files and the
per-atom sigma they define, evaluated on the GPU.
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


# ---------------------------------------------------------------------------
# compute flare/std/atom: files and per-atom sigma
# ---------------------------------------------------------------------------
def _numbers(f, n):
    """The next n whitespace-separated numbers of an open file (the grab()
    of compute_flare_std_atom.cpp, which reads across line breaks)."""
    out = []
    while len(out) < n:
        line = f.readline()
        if not line:
            raise ValueError(f"file ended after {len(out)} of {n} values")
        out.extend(float(t) for t in line.split())
    if len(out) != n:
        raise ValueError(f"expected {n} values, found {len(out)} on the last lines")
    return np.array(out)


def _header_after_hyps(f):
    """radial basis, 'n_species n_max l_max size', cutoff function and the
    cutoff matrix (b2.cpp:52-74; parse_cutoff_matrix,
    compute_flare_std_atom.cpp:329-358)."""
    radial = f.readline().split()[0]
    nos, nmax, lmax, size = (int(t) for t in f.readline().split()[:4])
    cutoff_name = f.readline().split()[0]
    cutoffs = _numbers(f, nos * nos).reshape(nos, nos)
    return dict(radial=radial, n_species=nos, n_max=nmax, l_max=lmax, size=size,
                cutoff_function=cutoff_name, cutoffs=cutoffs)


def read_L_inverse(file_name):
    """ComputeFlareStdAtom::read_L_inverse (compute_flare_std_atom.cpp:484-611):
    power and kernel name, hyperparameters, the descriptor header (its size
    field is the number of kernels), the number of sparse environments and
    the lower triangle of L_inv row by row."""
    with open(file_name) as f:
        f.readline()
        t = f.readline().split()
        power, kernel_name = int(float(t[0])), t[1]
        n_hyps = int(f.readline().split()[0])
        hyps = _numbers(f, n_hyps)
        h = _header_after_hyps(f)
        n = int(f.readline().split()[0])
        tri = _numbers(f, n * (n + 1) // 2)
    L = np.zeros((n, n))
    L[np.tril_indices(n)] = tri  # row-major lower triangle, as the reader fills it
    h.update(power=power, kernel_name=kernel_name, hyperparameters=hyps, n_kernels=h.pop("size"),
             n_clusters=n, L_inv=L)
    return h


def read_sparse_descriptors(file_name, n_descriptors):
    """ComputeFlareStdAtom::read_sparse_descriptors
    (compute_flare_std_atom.cpp:613-686): per kernel, the number of sparse
    environments and of types, then per type its count and its rows
    (n_descriptors values each). Returns one list of per-type arrays per
    kernel."""
    out = []
    with open(file_name) as f:
        f.readline()
        n_kernels = int(f.readline().split()[0])
        for _ in range(n_kernels):
            _, n_clusters, n_types = (int(t) for t in f.readline().split()[:3])
            per_type = []
            for _ in range(n_types):
                m = int(f.readline().split()[0])
                per_type.append(_numbers(f, m * n_descriptors).reshape(m, n_descriptors))
            if sum(len(a) for a in per_type) != n_clusters:
                raise ValueError("sparse descriptor counts do not add up to n_clusters")
            out.append(per_type)
    return out


def varmap_coefficients(sgp, kernel_index=0):
    """NormalizedDotProduct::compute_varmap_coefficients
    (normalized_dot_product.cpp:769-830), power 1: per species s, the first
    d rows of the Householder Q of A = [X_hat_s^T; sqrt(jitter)/sigma I]
    (d + M_s x M_s), stored as a d x d matrix whose columns past M_s are
    zero. Then |beta_s^T B_hat|^2 = sigma^2 |L_s^-1 k|^2 for k = X_hat_s
    B_hat, so 1 - |beta_s^T B_hat|^2 is sigma^2 / sigma_f^2 of the power-1
    model. A one-time FP64 producer on the host (LAPACK Householder QR, as
    the reference's Eigen HouseholderQR)."""
    kern = sgp.kernels[kernel_index]
    if kern.kernel_name != "NormalizedDotProduct":
        raise ValueError("variance maps are built for NormalizedDotProduct kernels")
    if float(kern.power) != 1.0:
        raise ValueError("variance maps are implemented for power 1 only "
                         "(normalized_dot_product.cpp:775-781)")
    cd = sgp.sparse_descriptors[kernel_index]
    d = cd.n_descriptors
    jr = np.sqrt(sgp.Kuu_jitter)
    beta = np.zeros((cd.n_types, d * d))
    for s in range(cd.n_types):
        X = np.array(cd.descriptors[s], dtype=np.float64, copy=True)
        nrm = np.asarray(cd.descriptor_norms[s])
        m = X.shape[0]
        if m == 0:
            continue
        keep = nrm >= 1e-8
        X[keep] /= nrm[keep, None]
        A = np.vstack([X.T, jr * np.eye(m) / kern.sigma])
        Q, _ = np.linalg.qr(A, mode="reduced")
        b = np.zeros((d, d))
        b[:, :m] = Q[:d, :]
        beta[s] = b.reshape(-1)
    return beta


def write_varmap_coefficients(sgp, file_name, contributor, kernel_index, calc, date=None):
    """SparseGP::write_varmap_coefficients (sparse_gp.cpp:1266-1339)."""
    beta = varmap_coefficients(sgp, kernel_index)
    kern = sgp.kernels[kernel_index]
    hyps = sgp.hyperparameters
    out = [f"DATE: {date or _date()} CONTRIBUTOR: {contributor}\n", f"{len(hyps)}\n",
           "".join(_fmt(h) + " " for h in hyps) + "\n", kern.kernel_name + "\n",
           _descriptor_header(calc, beta.shape[1])]
    count = 0
    body = []
    for row in beta:
        if count != 0:
            body.append("\n")
            count = 0
        for v in row:
            body.append(("  " if v > 0 else " ") + _fmt(v))
            count += 1
            if count == 5:
                count = 0
                body.append("\n")
    with open(file_name, "w") as f:
        f.write("".join(out) + "".join(body))


def read_varmap_coefficients(file_name):
    """ComputeFlareStdAtom::read_file (compute_flare_std_atom.cpp:361-482):
    hyperparameters, kernel name, the descriptor header (size d^2) and one
    d x d beta matrix per species."""
    with open(file_name) as f:
        f.readline()
        n_hyps = int(f.readline().split()[0])
        hyps = _numbers(f, n_hyps)
        kernel_name = f.readline().split()[0]
        h = _header_after_hyps(f)
        nr = h["n_species"] * h["n_max"]
        d = nr * (nr + 1) // 2 * (h["l_max"] + 1)
        if h["size"] != d * d:
            raise ValueError("Beta size doesn't match the number of descriptors.")
        beta = _numbers(f, h["n_species"] * d * d).reshape(h["n_species"], d, d)
    h.update(kernel_name=kernel_name, hyperparameters=hyps, n_descriptors=d, beta=beta)
    return h


class _SparseSet:
    """The fields of a ClusterDescriptor that a device model reads."""

    def __init__(self, per_type):
        self.descriptors = [np.ascontiguousarray(a, dtype=np.float64) for a in per_type]
        self.descriptor_norms = [np.ones(len(a)) for a in per_type]
        self.n_clusters_by_type = [len(a) for a in per_type]
        self.n_types = len(per_type)
        self.n_clusters = sum(self.n_clusters_by_type)
        self.n_descriptors = per_type[0].shape[1] if per_type else 0


class StdAtom:
    """compute flare/std/atom (compute_flare_std_atom.cpp:95-228): the
    per-atom sigma / |sigma_f| of a sparse GP from its files, on the GPU.

    StdAtom(L_inv_file, sparse_descriptors_file): sigma^2 = 1 -
    sigma_f^2 |L_inv k|^2 with k = (x_hat_j . B_hat)^power over the atom's own
    species (the normalised sparse rows of the file). StdAtom(varmap_file):
    sigma^2 = 1 - |beta_s^T B_hat|^2. Both run the device model of the
    sparse GP's local variance: rows with unit norms (so x_hat is the file's
    row), alpha = 0, and for the variance map the columns of beta_s as rows
    with L_inv = I, sigma = 1, power 1. Signed square root, 0 for an empty
    environment (|B|^2 < 1e-8), in atom order. NormalizedDotProduct only."""

    def __init__(self, *files, context=None):
        from ._C_flare import _DeviceModel
        self._ctx = context or _lib.default_context()
        if len(files) == 2:
            info = read_L_inverse(files[0])
            nr = info["n_species"] * info["n_max"]
            d = nr * (nr + 1) // 2 * (info["l_max"] + 1)
            per_type = read_sparse_descriptors(files[1], d)[0]
            if sum(len(a) for a in per_type) != info["n_clusters"]:
                raise ValueError("n_clusters in sparse_descriptors and L_inv not match")
            sigma, power, Linv = float(info["hyperparameters"][0]), info["power"], info["L_inv"]
            self.scale = sigma * sigma  # the model's variance is sigma^2 times the file's
        elif len(files) == 1:
            info = read_varmap_coefficients(files[0])
            per_type = []
            for b in info["beta"]:
                nz = np.flatnonzero(np.any(b != 0.0, axis=0))
                m = int(nz[-1]) + 1 if nz.size else 0
                per_type.append(np.ascontiguousarray(b[:, :m].T))
            sigma, power, self.scale = 1.0, 1, 1.0
            M = sum(len(a) for a in per_type)
            Linv = np.eye(M)
        else:
            raise ValueError("StdAtom takes (L_inv_file, sparse_descriptors_file) or (varmap_file)")
        if info["kernel_name"] != "NormalizedDotProduct":
            raise ValueError("only NormalizedDotProduct files are evaluated on the GPU")
        if info["radial"] != "chebyshev":
            raise ValueError("Please use chebyshev radial basis function.")
        self.cutoffs = info["cutoffs"]
        self.cutoff = float(self.cutoffs.max())
        self.descriptor = B2("chebyshev", info["cutoff_function"], [0.0, self.cutoff], [],
                             [info["n_species"], info["n_max"], info["l_max"]], self.cutoffs)
        cd = _SparseSet(per_type)
        self._model = _DeviceModel.__new__(_DeviceModel)
        self._model._create(self._ctx, sigma, float(power), cd, np.zeros(cd.n_clusters), Linv)

    def compute(self, cell, species, positions):
        """sigma / |sigma_f| per atom (signed square root of the variance)."""
        s = Structure(cell, species, positions, self.cutoff, [self.descriptor], form=FORM_FP64,
                      context=self._ctx)
        n = s.noa
        var = np.zeros(n)
        _lib.check(_lib.load().fm_predict(self._model.h, s._h, _lib.PREDICT_LOCAL, FORM_FP64, None,
                                          _lib.ptr(var), None))
        var = var / self.scale
        # Species-sorted rows back to atom order (atom order kept within a
        # species); empty environments as compute_flare_std_atom.cpp:176-177.
        sp = np.asarray(species)
        order = np.argsort(sp, kind="stable")
        v = np.zeros(n)
        v[order] = var
        # |B|^2 per atom: norms and atom indices only (no descriptor rows).
        dv = s.descriptors[0]
        nn = np.zeros(n)
        for t in range(dv.n_types):
            ns = dv.n_clusters_by_type[t]
            nrm, ai = np.zeros(ns), np.zeros(ns, np.int32)
            _lib.check(_lib.load().fm_structure_get_b2(s._h, t, None, _lib.ptr(nrm), None, None,
                                                       _lib.ptr(ai), None, None))
            nn[ai] = nrm * nrm
        std = np.sign(v) * np.sqrt(np.abs(v))
        std[nn < 1e-8] = 0.0
        return std
