"""ctypes bindings of the FP64 oracle (oracle/libref_cpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's cpu_baseline leg, never by the product package. Each
function restates the reference algorithm named in oracle/ref_cpu.hpp.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

CUTOFF_KIND = {"quadratic": 0, "cosine": 1}


def lib(fast=False):
    name = "libref_cpu_fast.so" if fast else "libref_cpu.so"
    if not fast and os.environ.get("ORACLE_SANITIZED") == "1":
        name = os.path.join("_asan", "libref_cpu.so")
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise OSError(f"oracle library {path} is not built (run make -C oracle)")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.ref_last_error.restype = C.c_char_p
    L.ref_geom_new.restype = vp
    L.ref_geom_new.argtypes = [_dp, C.c_int, _ip, _dp, C.c_double, C.c_int]
    L.ref_geom_free.argtypes = [vp]
    L.ref_geom_info.argtypes = [vp, _ip, _dp]
    L.ref_geom_get.argtypes = [vp, C.c_int, C.c_void_p]
    L.ref_desc_new.restype = vp
    L.ref_desc_new.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int]
    L.ref_desc_new_b3.restype = vp
    L.ref_desc_new_b3.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp,
                                  C.c_int]
    L.ref_wigner3j.restype = C.c_double
    L.ref_wigner3j.argtypes = [C.c_int] * 6
    L.ref_b3_n_ls.restype = C.c_int
    L.ref_b3_n_ls.argtypes = [C.c_int]
    L.ref_complex_sph_harm.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _dp]
    L.ref_desc_free.argtypes = [vp]
    L.ref_desc_sizes.argtypes = [vp, C.c_int, _ip]
    L.ref_desc_get.argtypes = [vp, C.c_int, C.c_int, C.c_void_p]
    L.ref_model_new.restype = vp
    L.ref_model_new.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, _ip, _dp, _dp, _dp,
                                vp, vp, vp, vp]
    L.ref_model_free.argtypes = [vp]
    L.ref_envs_struc.argtypes = [vp, vp, _dp]
    L.ref_envs_envs.argtypes = [vp, vp, _dp]
    L.ref_self_kernel.argtypes = [vp, vp, _dp]
    L.ref_predict.argtypes = [vp, vp, C.c_int, _dp, C.c_void_p]
    L.ref_predict_scalable.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                       _dp, _dp, _dp, _dp]
    L.ref_mapping.argtypes = [vp, C.c_void_p, C.POINTER(C.c_int)]
    L.ref_coeff_text.restype = C.c_long
    L.ref_coeff_text.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_char_p,
                                 C.c_char_p, C.c_int, C.c_char_p, C.c_long]
    L.ref_pair_flare.argtypes = [C.c_char_p, vp, _dp, _dp, _dp]
    L.ref_sph_harm.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _dp]
    L.ref_radial.argtypes = [C.c_double] * 5 + [C.c_int, C.c_double, C.c_int, _dp]
    L.ref_gemm_abt.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    _LIBS[name] = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Neighbors:
    """Geometry + neighbour list (structure.cpp:7-165). mode 'literal' is the
    O(N^2 * images) loop, 'cells' the cell-list build with identical output."""

    def __init__(self, cell, species, positions, cutoff, mode="literal", fast=False):
        self._L = lib(fast)
        cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
        species = np.ascontiguousarray(species, dtype=np.int32)
        positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        self.noa = len(species)
        self.h = self._L.ref_geom_new(cell, self.noa, species, positions, float(cutoff),
                                      0 if mode == "literal" else 1)
        ints = np.zeros(4, np.int32)
        dbls = np.zeros(4)
        self._L.ref_geom_info(self.h, ints, dbls)
        self.sweep = int(ints[1])
        self.n_neighbors = int(ints[2])
        self.volume = float(dbls[0])
        self.single_sweep_cutoff = float(dbls[1])

    def _get(self, field, n, dtype):
        out = np.zeros(n, dtype=dtype)
        if n:
            self._L.ref_geom_get(self.h, field, out.ctypes.data_as(C.c_void_p))
        return out

    @property
    def wrapped_positions(self):
        return self._get(0, 3 * self.noa, np.float64).reshape(-1, 3)

    @property
    def neighbor_count(self):
        return self._get(1, self.noa, np.int32)

    @property
    def cumulative_neighbor_count(self):
        return self._get(2, self.noa + 1, np.int32)

    @property
    def structure_indices(self):
        return self._get(3, self.n_neighbors, np.int32)

    @property
    def neighbor_species(self):
        return self._get(4, self.n_neighbors, np.int32)

    @property
    def relative_positions(self):
        return self._get(5, 4 * self.n_neighbors, np.float64).reshape(-1, 4)

    @property
    def images(self):
        return self._get(6, 3 * self.n_neighbors, np.int32).reshape(-1, 3)

    def __del__(self):
        try:
            self._L.ref_geom_free(self.h)
        except Exception:
            pass


class B2Values:
    """DescriptorValues of the B2 descriptor (b2.cpp:76-399), per species."""

    FIELDS = {"descriptors": 0, "descriptor_force_dervs": 1, "neighbor_coordinates": 2,
              "descriptor_norms": 3, "descriptor_force_dots": 4, "cutoff_values": 5,
              "neighbor_counts": 6, "cumulative_neighbor_counts": 7, "atom_indices": 8,
              "neighbor_indices": 9}

    _new = "ref_desc_new"

    def __init__(self, nbrs, nos, N, lmax, cutoffs, cutoff_function="quadratic", r1=0.0,
                 with_dervs=True):
        self._L = nbrs._L
        self.nbrs = nbrs
        self.settings = (nos, N, lmax)
        self.cutoffs = np.ascontiguousarray(cutoffs, dtype=np.float64).reshape(nos * nos)
        self.cutoff_kind = CUTOFF_KIND[cutoff_function]
        self.r1 = r1
        self.h = getattr(self._L, self._new)(nbrs.h, nos, N, lmax, self.cutoff_kind, r1,
                                             self.cutoffs, 1 if with_dervs else 0)
        sz = np.zeros(8, np.int32)
        self._L.ref_desc_sizes(self.h, 0, sz)
        self.n_descriptors, self.n_types, self.n_atoms, self.n_clusters = map(int, sz[:4])
        self.has_dervs = bool(sz[7])
        self.n_clusters_by_type, self.n_neighbors_by_type, self.cumulative_type_count = [], [], [0]
        for s in range(self.n_types):
            self._L.ref_desc_sizes(self.h, s, sz)
            self.n_clusters_by_type.append(int(sz[4]))
            self.n_neighbors_by_type.append(int(sz[5]))
            self.cumulative_type_count.append(self.cumulative_type_count[-1] + int(sz[4]))
        self.volume = nbrs.volume

    def get(self, name):
        f = self.FIELDS[name]
        nd = self.n_descriptors
        out = []
        for s in range(self.n_types):
            ns, nb = self.n_clusters_by_type[s], self.n_neighbors_by_type[s]
            shape, dt = {
                0: ((ns, nd), np.float64), 1: ((3 * nb, nd), np.float64),
                2: ((nb, 3), np.float64), 3: ((ns,), np.float64), 4: ((3 * nb,), np.float64),
                5: ((ns,), np.float64), 6: ((ns,), np.int32), 7: ((ns,), np.int32),
                8: ((ns,), np.int32), 9: ((nb,), np.int32)}[f]
            if f in (1, 4) and not self.has_dervs:
                out.append(np.zeros((0,) + shape[1:], dt))
                continue
            a = np.zeros(shape, dtype=dt)
            if a.size:
                self._L.ref_desc_get(self.h, f, s, a.ctypes.data_as(C.c_void_p))
            out.append(a)
        return out

    def __getattr__(self, name):
        if name in B2Values.FIELDS:
            return self.get(name)
        raise AttributeError(name)

    def __del__(self):
        try:
            self._L.ref_desc_free(self.h)
        except Exception:
            pass


class B3Values(B2Values):
    """DescriptorValues of the B3 descriptor (b3.cpp:80-289), per species;
    lmax <= 4."""

    _new = "ref_desc_new_b3"


def wigner3j(j1, j2, j3, m1, m2, m3):
    return lib().ref_wigner3j(j1, j2, j3, m1, m2, m3)


def b3_n_ls(lmax):
    return lib().ref_b3_n_ls(lmax)


def complex_sph_harm(x, y, z, lmax):
    """Complex Y_l^m (Condon-Shortley phase, index l^2 + m + l) and its
    Cartesian gradient: arrays Y, Yx, Yy, Yz of (lmax + 1)^2 complex."""
    H = (lmax + 1) ** 2
    out = np.zeros(8 * H)
    lib().ref_complex_sph_harm(x, y, z, lmax, out)
    z_ = out[0::2] + 1j * out[1::2]
    return z_[:H], z_[H:2 * H], z_[2 * H:3 * H], z_[3 * H:]


class Model:
    """Sparse-GP prediction state (sparse_gp.h): NDP kernel, sparse set (raw
    descriptors and norms per species, type-major), alpha, L_inv, Kuu^-1,
    Sigma, R_inv (all FP64, M x M row-major)."""

    def __init__(self, sigma, power, sparse_descriptors, sparse_norms, alpha, L_inv=None,
                 Kuu_inverse=None, Sigma=None, R_inv=None, fast=False):
        self._L = lib(fast)
        self.sigma, self.power = float(sigma), float(power)
        nos = len(sparse_descriptors)
        nd = sparse_descriptors[0].shape[1]
        self.n_by_type = np.array([len(x) for x in sparse_descriptors], np.int32)
        desc = np.ascontiguousarray(np.concatenate(sparse_descriptors, axis=0), dtype=np.float64)
        norms = np.ascontiguousarray(np.concatenate(sparse_norms), dtype=np.float64)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                      for a in (alpha, L_inv, Kuu_inverse, Sigma, R_inv)]
        a, Li, Ki, Sg, Ri = self._keep
        self.M = int(self.n_by_type.sum())
        self.h = self._L.ref_model_new(self.sigma, self.power, nos, nd, self.n_by_type, desc.reshape(-1),
                                       norms, a, _ptr(Li), _ptr(Ki), _ptr(Sg), _ptr(Ri))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._L.ref_last_error().decode())

    def envs_struc(self, desc):
        out = np.zeros(self.M * (1 + 3 * desc.n_atoms + 6))
        self._check(self._L.ref_envs_struc(self.h, desc.h, out))
        return out.reshape(self.M, -1)

    def envs_envs(self, desc):
        out = np.zeros(desc.n_clusters * self.M)
        self._check(self._L.ref_envs_envs(self.h, desc.h, out))
        return out.reshape(desc.n_clusters, self.M)

    def self_kernel(self, desc):
        out = np.zeros(1 + 3 * desc.n_atoms + 6)
        self._check(self._L.ref_self_kernel(self.h, desc.h, out))
        return out

    def predict(self, desc, kind="mean"):
        k = {"mean": 0, "local": 1, "dtc": 2, "sor": 3}[kind]
        n_out = 1 + 3 * desc.n_atoms + 6
        mean = np.zeros(n_out)
        var = np.zeros(n_out if k >= 2 else desc.n_atoms)
        self._check(self._L.ref_predict(self.h, desc.h, k, mean, var.ctypes.data_as(C.c_void_p)))
        return (mean, var) if k else (mean, None)

    def predict_scalable(self, nbrs, nos, N, lmax, cutoffs, cutoff_function="quadratic", r1=0.0):
        n = nbrs.noa
        mean = np.zeros(1 + 3 * n + 6)
        var = np.zeros(n)
        le = np.zeros(n)
        cut = np.ascontiguousarray(cutoffs, dtype=np.float64).reshape(-1)
        self._check(self._L.ref_predict_scalable(self.h, nbrs.h, nos, N, lmax,
                                                 CUTOFF_KIND[cutoff_function], r1, cut, mean, var, le))
        return mean, var, le

    def mapping(self):
        bs = C.c_int(0)
        self._check(self._L.ref_mapping(self.h, None, C.byref(bs)))
        nos = len(self.n_by_type)
        out = np.zeros(nos * bs.value)
        self._check(self._L.ref_mapping(self.h, out.ctypes.data_as(C.c_void_p), C.byref(bs)))
        return out.reshape(nos, bs.value)

    def file_text(self, nos, N, lmax, cutoffs, cutoff_function="quadratic", which="beta",
                  date="Thu Jan  1 00:00:00 1970", contributor="oracle"):
        w = {"beta": 0, "L_inv": 1, "sparse_desc": 2}[which]
        cut = np.ascontiguousarray(cutoffs, dtype=np.float64).reshape(-1)
        args = (self.h, nos, N, lmax, CUTOFF_KIND[cutoff_function], cut, date.encode(),
                contributor.encode(), w)
        n = self._L.ref_coeff_text(*args, None, 0)
        buf = C.create_string_buffer(n + 1)
        self._L.ref_coeff_text(*args, buf, n + 1)
        return buf.value.decode()

    def __del__(self):
        try:
            self._L.ref_model_free(self.h)
        except Exception:
            pass


def pair_flare(text, nbrs):
    L = nbrs._L
    e = np.zeros(nbrs.noa)
    f = np.zeros(3 * nbrs.noa)
    v = np.zeros(6)
    rc = L.ref_pair_flare(text.encode(), nbrs.h, e, f, v)
    if rc:
        raise RuntimeError(L.ref_last_error().decode())
    return e, f.reshape(-1, 3), v


def sph_harm(x, y, z, lmax):
    H = (lmax + 1) ** 2
    out = np.zeros(4 * H)
    lib().ref_sph_harm(x, y, z, lmax, out)
    return out.reshape(4, H)


def radial(x, y, z, rc, N, r1=0.0, cutoff_function="quadratic"):
    r = float(np.sqrt(x * x + y * y + z * z))
    out = np.zeros(4 * N)
    lib().ref_radial(x, y, z, r, rc, N, r1, CUTOFF_KIND[cutoff_function], out)
    return out.reshape(4, N)


def gemm_abt(A, B, fast=False):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    out = np.zeros((A.shape[0], B.shape[0]))
    lib(fast).ref_gemm_abt(A.shape[0], B.shape[0], A.shape[1], A.reshape(-1), B.reshape(-1),
                           out.reshape(-1))
    return out


def refsrc():
    """The reference's own radial.cpp/cutoffs.cpp (oracle/_ref), or None."""
    path = os.path.join(_HERE, "_ref", "libref_radial.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.refsrc_chebyshev.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double, _dp]
    L.refsrc_cutoff.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
    L.refsrc_radial.argtypes = [C.c_double] * 5 + [C.c_int, C.c_double, C.c_int, _dp]
    return L
