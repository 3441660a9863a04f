"""Benchmark: atom-evals/s of the sparse-GP force evaluation (E, F, stress
and local sigma per atom) on one MI355X, chosen form vs the FP64 drop-in.

Workload (default): a 32,000-atom two-species random 50/50 FCC alloy
(a = 4.0 A, 20^3 cells, displacements 0.08 A); B2 with 12 radial functions
per species, l_max 6, cutoff 6.0 A, quadratic cutoff (d = 2100); 2,000
sparse environments drawn from the structure (seed 3); alpha ~ N(0, 1e-2)
(seed 4); NormalizedDotProduct sigma 2, power 2; L_inv from the FP64
Cholesky of Kuu + 1e-8 I.

One step = neighbour list + descriptors + kernel contraction + adjoint
forces/stress + local variances, from positions already resident in HBM to
results on the host. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-evals/s (E,F,σ) chosen vs FP64 drop-in at stated error; p50/p99 latency"
PEAK = {"bf16": 2500.0, "fp64": 78.6}  # MI355X dense TFLOP/s


class Hip:
    """hipMalloc / hipMemcpy through ctypes, for inputs resident in HBM."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def upload(self, a):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_model(C_, cell, species, pos, n_sparse, seed_sparse=3, seed_alpha=4):
    from flare_amd import _C_flare as CF
    calc = CF.B2("chebyshev", "quadratic", [0.0, 6.0], [], [2, 12, 6], np.full((2, 2), 6.0))
    s = CF.Structure(cell, species, pos, 6.0, [calc])
    rng = np.random.default_rng(seed_sparse)
    atoms = sorted(rng.choice(len(pos), n_sparse, replace=False).tolist())
    sgp = CF.SparseGP([CF.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    sgp.add_specific_environments(s, atoms)
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(seed_alpha).normal(scale=0.1, size=sgp.n_sparse))
    return calc, sgp


def run_form(lib, ctx, model_h, calc, cell, sp_dev, pos_dev, n, form, steps, warmup,
             cutoff=6.0, compute="fm_structure_compute_b2"):
    from flare_amd import _lib
    settings, keep = calc._settings()
    cellv = _lib.f64(cell.reshape(-1))
    flags = _lib.PREDICT_MEAN | _lib.PREDICT_LOCAL
    mean = np.zeros(1 + 3 * n + 6)
    var = np.zeros(n)
    lat = []
    stage = {k: [0.0, 0] for k in _lib.Context.STAGES}
    kern = {k: [0.0, 0] for k in _lib.Context.KERNELS}
    nnb = C.c_int64()

    def step():
        h = C.c_void_p()
        _lib.check(lib.fm_structure_create_device(ctx.h, _lib.ptr(cellv), n, sp_dev, pos_dev,
                                                  cutoff, C.byref(h)))
        _lib.check(getattr(lib, compute)(h, C.byref(settings), form))
        _lib.check(lib.fm_predict(model_h, h, flags, form, _lib.ptr(mean), _lib.ptr(var), None))
        sw, vol, ssc = C.c_int(), C.c_double(), C.c_double()
        _lib.check(lib.fm_structure_info(h, C.byref(sw), C.byref(nnb), C.byref(vol),
                                         C.byref(ssc)))
        lib.fm_structure_destroy(h)

    for _ in range(warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        step()
        lat.append(time.perf_counter() - a)
        for k, (ms, nl) in ctx.timings().items():
            stage[k][0] += ms
            stage[k][1] += nl
        for k, (ms, nl) in ctx.kernel_timings().items():
            kern[k][0] += ms
            kern[k][1] += nl
    ctx.synchronize()
    total = time.perf_counter() - t0
    del keep
    return dict(total=total, lat=np.array(lat), stage=stage, kernels=kern, mean=mean.copy(),
                var=var.copy(), n_neighbors=int(nnb.value))


# MI355X dense peaks. The chosen form's kernel contraction runs 15 int8
# slice products per FP64-accurate product, its gradient contraction 3 BF16
# products per FP32 product: their ceilings in algorithmic flops are the
# MFMA peaks divided by those counts.
PEAK_INT8 = 5000.0  # TOP/s
I8_PRODUCTS = 15
HBM_PEAK = 8000.0  # GB/s


def kernel_rooflines(r, steps, n, n_by, M_by, d, nrad_h, f64):
    """Per-kernel rooflines from the live kernel timings of run_form:
    achieved = algorithmic work per step / kernel time per step (the
    average launch's work over its average duration). Algorithmic work per
    step (sums over species launches; nb = neighbours per atom, s = 8 B in
    FP64 and 4 B in the chosen form):
      kernel contraction   2 n_s M_s d flops
      gradient contraction 2 n_s M_s d flops
      variance contraction M_s^2 n_s flops (lower-triangular L_s^-1 times K^T)
      descriptor kernel    n (40 nb + s (nrad H + d) + 16) bytes: neighbour
                           entries in, c and B rows and norms out; the
                           chosen form also writes the int8 slices of the
                           row and its scale, n (5 d + 8) bytes
      adjoint dE/dc        n (s (2 d + 2 nrad H) + 16) bytes: G, B, c in, u out
      bond forces          n (nb (40 + 3 s) + s nrad H + 48) bytes
      row slicing          n (4 d + 5 d + 8) bytes
      force gather         n (6 s nb + 72) bytes"""
    s = 8 if f64 else 4
    nb = r["n_neighbors"] / max(n, 1)
    kf = sum(2.0 * a * b * d for a, b in zip(n_by, M_by))
    vf = sum(float(b) * b * a for a, b in zip(n_by, M_by))
    work = {
        "kernel_contraction": ("mfma", kf, "TFLOP/s",
                               PEAK["fp64"] if f64 else PEAK_INT8 / I8_PRODUCTS),
        "gradient_contraction": ("mfma", kf, "TFLOP/s", PEAK["fp64"] if f64 else PEAK["bf16"] / 3),
        "variance_contraction": ("mfma", vf, "TFLOP/s", PEAK["fp64"]),
        "descriptor": ("hbm", n * (40 * nb + s * (nrad_h + d) + 16 + (0 if f64 else 5 * d + 8)),
                       "GB/s", HBM_PEAK),
        "adjoint_dEdc": ("hbm", n * (s * (2 * d + 2 * nrad_h) + 16), "GB/s", HBM_PEAK),
        "bond_forces": ("hbm", n * (nb * (40 + 3 * s) + s * nrad_h + 48), "GB/s", HBM_PEAK),
        "slice": ("hbm", n * (9 * d + 8), "GB/s", HBM_PEAK),
        "gather": ("hbm", n * (6 * s * nb + 72), "GB/s", HBM_PEAK),
    }
    out = {}
    for k, (bound, per_step, unit, peak) in work.items():
        ms, nl = r["kernels"][k]
        if nl == 0 or ms <= 0:
            continue
        scale = 1e-12 if unit == "TFLOP/s" else 1e-9
        achieved = per_step * steps / (ms * 1e-3) * scale
        e = {"bound": bound, "achieved": round(achieved, 2), "peak": round(peak, 1),
             "unit": unit, "frac": round(achieved / peak, 4),
             "avg_launch_ms": round(ms / nl, 4), "launches_per_step": nl // steps,
             "ms_per_step": round(ms / steps, 4),
             ("flops_per_step" if bound == "mfma" else "bytes_per_step"): per_step}
        if k == "kernel_contraction" and not f64:
            e["note"] = ("peak = int8 MFMA 5.0 POP/s / 15 slice products per "
                         "FP64-accurate product")
        if k == "gradient_contraction" and not f64:
            e["note"] = "peak = BF16 MFMA 2.5 PFLOP/s / 3 products per split product"
        out[k] = e
    return out


def batched_leg(lib, ctx, hip, steps, warmup):
    """512 structures (16-2048 atoms, 3 species, bulk/slab cells)
    in one batch per step; B2 [3, 8, 4] rc 5.5 (d = 1500); 1500 sparse
    environments from the first structures (seed 6); FP64 vs chosen."""
    from flare_amd import _C_flare as CF
    from flare_amd import _lib
    from flare_amd.workloads import batch_structures
    structs = batch_structures(512, seed=5)
    cut = np.full((3, 3), 5.5)
    calc = CF.B2("chebyshev", "quadratic", [0.0, 5.5], [], [3, 8, 4], cut)
    sgp = CF.SparseGP([CF.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    rng = np.random.default_rng(6)
    need = 1500
    for st in structs:
        if need <= 0:
            break
        s = CF.Structure(*st, 5.5, [calc])
        take = min(need, len(st[1]))
        sgp.add_specific_environments(s, sorted(rng.choice(len(st[1]), take, replace=False)))
        need -= take
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(4).normal(scale=0.1, size=sgp.n_sparse))
    n_atoms = np.array([len(s[1]) for s in structs], np.int32)
    cells = np.stack([s[0].reshape(9) for s in structs]).astype(np.float64)
    sp_dev = hip.upload(np.ascontiguousarray(np.concatenate([s[1] for s in structs]).astype(np.int32)))
    pos_dev = hip.upload(np.ascontiguousarray(np.concatenate([s[2] for s in structs]).reshape(-1)))
    hip.L.hipDeviceSynchronize()
    n = int(n_atoms.sum())
    total = int(np.sum(1 + 3 * n_atoms.astype(np.int64) + 6))
    settings, keep = calc._settings()
    model_h = sgp._model(0, ctx).h
    flags = _lib.PREDICT_MEAN | _lib.PREDICT_LOCAL
    out = {}
    res = {}
    for name, form in (("fp64", _lib.FORM_FP64), ("chosen", _lib.FORM_CHOSEN)):
        mean = np.zeros(total)
        var = np.zeros(n)

        def step():
            h = C.c_void_p()
            _lib.check(lib.fm_batch_create_device(ctx.h, len(structs), _lib.ptr(cells),
                                                  _lib.ptr(n_atoms), sp_dev, pos_dev, 5.5,
                                                  C.byref(h)))
            _lib.check(lib.fm_structure_compute_b2(h, C.byref(settings), form))
            _lib.check(lib.fm_predict(model_h, h, flags, form, _lib.ptr(mean), _lib.ptr(var),
                                      None))
            lib.fm_structure_destroy(h)

        for _ in range(warmup):
            step()
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            a = time.perf_counter()
            step()
            lat.append(time.perf_counter() - a)
        dt = time.perf_counter() - t0
        res[name] = (mean.copy(), var.copy())
        out[name] = {"atom_evals_per_s": n * steps / dt, "ms_per_step": dt / steps * 1e3,
                     "p50_ms": float(np.percentile(lat, 50) * 1e3),
                     "p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "stage_ms": {k: round(v[0], 4) for k, v in ctx.timings().items()}}
    # The FP64 drop-in as a flare_pp caller uses it: one structure per
    # evaluation (same positions, same model), one warm-up pass, one timed
    # pass over the 512 structures.
    mean1 = np.zeros(int(1 + 3 * n_atoms.max() + 6))
    var1 = np.zeros(int(n_atoms.max()))
    offs = np.concatenate([[0], np.cumsum(n_atoms)])

    def one_pass():
        lat = []
        for k, st in enumerate(structs):
            a = time.perf_counter()
            h = C.c_void_p()
            na = int(n_atoms[k])
            _lib.check(lib.fm_structure_create_device(
                ctx.h, _lib.ptr(cells[k]), na, C.c_void_p(sp_dev.value + 4 * int(offs[k])),
                C.c_void_p(pos_dev.value + 24 * int(offs[k])), 5.5, C.byref(h)))
            _lib.check(lib.fm_structure_compute_b2(h, C.byref(settings), _lib.FORM_FP64))
            _lib.check(lib.fm_predict(model_h, h, flags, _lib.FORM_FP64, _lib.ptr(mean1),
                                      _lib.ptr(var1), None))
            lib.fm_structure_destroy(h)
            lat.append(time.perf_counter() - a)
        return lat

    one_pass()
    t0 = time.perf_counter()
    lat = one_pass()
    dt = time.perf_counter() - t0
    out["fp64_structure_by_structure"] = {
        "atom_evals_per_s": n / dt, "ms_per_pass": dt * 1e3,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3)}
    del keep
    # Errors of the chosen form against FP64 over all forces / variances.
    o = 0
    dF = dE = 0.0
    for na in n_atoms:
        m64, mch = res["fp64"][0][o:o + 1 + 3 * na + 6], res["chosen"][0][o:o + 1 + 3 * na + 6]
        dE = max(dE, abs(m64[0] - mch[0]) / na)
        dF = max(dF, float(np.abs(m64[1:1 + 3 * na] - mch[1:1 + 3 * na]).max()))
        o += 1 + 3 * na + 6
    out["workload"] = (f"512 structures, {n} atoms (16-2048 each), 3 species, bulk/slab; "
                       f"B2 [3,8,4] rc 5.5 (d=1500); {sgp.n_sparse} sparse; one batch per step")
    out["speedup_vs_fp64"] = out["chosen"]["atom_evals_per_s"] / out["fp64"]["atom_evals_per_s"]
    out["speedup_vs_fp64_structure_by_structure"] = (
        out["chosen"]["atom_evals_per_s"] / out["fp64_structure_by_structure"]["atom_evals_per_s"])
    out["error_vs_fp64"] = {"energy_meV_per_atom": dE * 1e3, "force_max_meV_per_A": dF * 1e3,
                            **sigma_errors(res["fp64"][1], res["chosen"][1],
                                           sgp.kernels[0].sigma)}
    return out, (sgp, calc)


def large_leg(lib, ctx, hip, steps, warmup):
    """The 1M-atom B2 workload: a 1,048,576-atom three-species BCC alloy
    (a = 3.2 A, 128 x 64 x 64 cells, displacements 0.1 A, seed 8); B2
    [3, 10, 6] rc 5.5 (d = 3255); NDP sigma 2, p 2; 4096 sparse environments
    (seed 9) drawn from a 16^3-cell sample of the same alloy (seed 11; the
    sparse rows of the full structure would be read back whole, 27 GB).
    One step = one full evaluation of the 1M atoms; FP64 vs chosen."""
    from flare_amd import _C_flare as CF
    from flare_amd import _lib
    from flare_amd.workloads import BCC_BASIS, random_alloy
    cut = np.full((3, 3), 5.5)
    calc = CF.B2("chebyshev", "quadratic", [0.0, 5.5], [], [3, 10, 6], cut)
    cs, ss, ps = random_alloy((16, 16, 16), a=3.2, nos=3, sigma=0.1, seed=11, basis=BCC_BASIS)
    small = CF.Structure(cs, ss, ps, 5.5, [calc])
    sgp = CF.SparseGP([CF.NormalizedDotProduct(2.0, 2.0)], 0.01, 0.05, 0.005)
    picks = np.random.default_rng(9).choice(len(ss), 4096, replace=False)
    sgp.add_specific_environments(small, sorted(int(a) for a in picks))
    del small
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(4).normal(scale=0.1, size=sgp.n_sparse))
    cell, species, pos = random_alloy((128, 64, 64), a=3.2, nos=3, sigma=0.1, seed=8,
                                      basis=BCC_BASIS)
    n = len(pos)
    pos_dev = hip.upload(np.ascontiguousarray(pos.reshape(-1)))
    sp_dev = hip.upload(np.ascontiguousarray(species.astype(np.int32)))
    hip.L.hipDeviceSynchronize()
    model_h = sgp._model(0, ctx).h
    out, res = {}, {}
    for name, form in (("fp64", _lib.FORM_FP64), ("chosen", _lib.FORM_CHOSEN)):
        log(f"  1M atoms, form {name} ...")
        r = run_form(lib, ctx, model_h, calc, cell, sp_dev, pos_dev, n, form, steps, warmup,
                     cutoff=5.5)
        res[name] = r
        out[name] = {"atom_evals_per_s": n * steps / r["total"],
                     "ms_per_step": r["total"] / steps * 1e3,
                     "stage_ms": {k: round(v[0] / steps, 3) for k, v in r["stage"].items()}}
    m64, mch = res["fp64"]["mean"], res["chosen"]["mean"]
    out["workload"] = (f"{n} atoms, 3-species BCC alloy a = 3.2 A (128x64x64 cells); "
                       f"B2 [3,10,6] rc 5.5 (d=3255); {sgp.n_sparse} sparse; NDP sigma 2 p 2")
    out["speedup_vs_fp64"] = out["chosen"]["atom_evals_per_s"] / out["fp64"]["atom_evals_per_s"]
    out["error_vs_fp64"] = {"energy_meV_per_atom": abs(m64[0] - mch[0]) / n * 1e3,
                            "force_max_meV_per_A": float(np.abs(m64[1:-6] - mch[1:-6]).max()) * 1e3,
                            **sigma_errors(res["fp64"]["var"], res["chosen"]["var"],
                                           sgp.kernels[0].sigma)}
    hip.L.hipFree(pos_dev)
    hip.L.hipFree(sp_dev)
    return out


def run_model_forms(lib, ctx, parts, cell, sp_dev, pos_dev, n, form, steps, warmup, cutoff):
    """One step = every kernel of a model on the same atoms: per kernel
    (calculator, model, compute entry point) a device structure is built,
    its descriptors computed and the kernel's prediction made; means add,
    local variances are per kernel."""
    from flare_amd import _lib
    flags = _lib.PREDICT_MEAN | _lib.PREDICT_LOCAL
    cellv = _lib.f64(cell.reshape(-1))
    sets = [calc._settings() for calc, _, _ in parts]
    outs = [np.zeros(1 + 3 * n + 6) for _ in parts]
    vars_ = [np.zeros(n) for _ in parts]
    stage = [{k: [0.0, 0] for k in _lib.Context.STAGES} for _ in parts]
    kernel_s = [0.0 for _ in parts]

    def step(timed):
        for k, (calc, model_h, compute) in enumerate(parts):
            a = time.perf_counter()
            h = C.c_void_p()
            _lib.check(lib.fm_structure_create_device(ctx.h, _lib.ptr(cellv), n, sp_dev,
                                                      pos_dev, cutoff, C.byref(h)))
            _lib.check(getattr(lib, compute)(h, C.byref(sets[k][0]), form))
            _lib.check(lib.fm_predict(model_h, h, flags, form, _lib.ptr(outs[k]),
                                      _lib.ptr(vars_[k]), None))
            lib.fm_structure_destroy(h)
            if timed:
                kernel_s[k] += time.perf_counter() - a
                for name, (ms, nl) in ctx.timings().items():
                    stage[k][name][0] += ms
                    stage[k][name][1] += nl

    for _ in range(warmup):
        step(False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    ctx.synchronize()
    total = time.perf_counter() - t0
    return dict(total=total, kernel_s=kernel_s, stage=stage, mean=sum(outs),
                var=[v.copy() for v in vars_])


def large_b3_leg(lib, ctx, hip, steps, warmup):
    """The 1M-atom B2 + B3 workload as one model: the 1,048,576-atom three-species BCC alloy of
    the B2 leg (a = 3.2 A, 128 x 64 x 64 cells, displacements 0.1 A, seed 8)
    with two kernels, NDP sigma 2, p 2 on B2 [3, 10, 6] rc 5.5 (d = 3255)
    and NDP sigma 2, p 2 on B3 [3, 6, 4] rc 5.5 (d = 74,100); each kernel's
    sparse set is every atom of a 16 x 16 x 8-cell sample (seed 11), 4096
    environments. The B3 rows alone would take 296 GB in FP32, so its
    prediction forms them in row tiles (automatic above 32 GB)."""
    from flare_amd import _C_flare as CF
    from flare_amd import _lib
    from flare_amd.workloads import BCC_BASIS, random_alloy
    cut = np.full((3, 3), 5.5)
    b2 = CF.B2("chebyshev", "quadratic", [0.0, 5.5], [], [3, 10, 6], cut)
    b3 = CF.B3("chebyshev", "quadratic", [0.0, 5.5], [], [3, 6, 4], cut)
    cs, ss, ps = random_alloy((16, 16, 8), a=3.2, nos=3, sigma=0.1, seed=11, basis=BCC_BASIS)
    small = CF.Structure(cs, ss, ps, 5.5, [b2, b3])
    sgp = CF.SparseGP([CF.NormalizedDotProduct(2.0, 2.0), CF.NormalizedDotProduct(2.0, 2.0)],
                      0.01, 0.05, 0.005)
    sgp.add_specific_environments(small, list(range(len(ss))))
    del small
    sgp.update_L_inverse()
    sgp.set_alpha(np.random.default_rng(4).normal(scale=0.1, size=sgp.n_sparse))
    cell, species, pos = random_alloy((128, 64, 64), a=3.2, nos=3, sigma=0.1, seed=8,
                                      basis=BCC_BASIS)
    n = len(pos)
    pos_dev = hip.upload(np.ascontiguousarray(pos.reshape(-1)))
    sp_dev = hip.upload(np.ascontiguousarray(species.astype(np.int32)))
    hip.L.hipDeviceSynchronize()
    parts = [(b2, sgp._model(0, ctx).h, "fm_structure_compute_b2"),
             (b3, sgp._model(1, ctx).h, "fm_structure_compute_b3")]
    out, res = {}, {}
    for name, form in (("fp64", _lib.FORM_FP64), ("chosen", _lib.FORM_CHOSEN)):
        log(f"  B2 + B3 model, {n} atoms, form {name} ...")
        r = run_model_forms(lib, ctx, parts, cell, sp_dev, pos_dev, n, form, steps, warmup,
                            cutoff=5.5)
        res[name] = r
        out[name] = {"atom_evals_per_s": n * steps / r["total"],
                     "ms_per_step": r["total"] / steps * 1e3,
                     "ms_per_step_by_kernel": {"b2": r["kernel_s"][0] / steps * 1e3,
                                               "b3": r["kernel_s"][1] / steps * 1e3},
                     "stage_ms_b3": {k: round(v[0] / steps, 3)
                                     for k, v in r["stage"][1].items()}}
    m64, mch = res["fp64"]["mean"], res["chosen"]["mean"]
    out["workload"] = (f"{n} atoms, 3-species BCC alloy a = 3.2 A (128x64x64 cells); two "
                       f"kernels: B2 [3,10,6] (d={b2.n_descriptors}) + B3 [3,6,4] "
                       f"(d={b3.n_descriptors}), rc 5.5; {sgp.n_sparse} sparse; NDP sigma 2 p 2")
    out["speedup_vs_fp64"] = out["chosen"]["atom_evals_per_s"] / out["fp64"]["atom_evals_per_s"]
    out["error_vs_fp64"] = {"energy_meV_per_atom": abs(m64[0] - mch[0]) / n * 1e3,
                            "force_max_meV_per_A": float(np.abs(m64[1:-6] - mch[1:-6]).max()) * 1e3}
    for k, tag in enumerate(("b2", "b3")):
        out["error_vs_fp64"][f"sigma_{tag}"] = sigma_errors(res["fp64"]["var"][k],
                                                            res["chosen"]["var"][k],
                                                            sgp.kernels[k].sigma)
    hip.L.hipFree(pos_dev)
    hip.L.hipFree(sp_dev)
    return out


def serving_leg(sgp, calc, batched, n_requests=2000, n_fp64=200):
    """Served stream: Poisson stream of requests (sizes log-uniform 32-8192
    atoms, the batched model), served with continuous batching in the
    chosen form at 0.7x the batched throughput (three workers, batches of
    at most 32,768 atoms) after an untimed warm-up stream; FP64 drop-in
    evaluated request by request on a prefix of the same stream."""
    from flare_amd import _C_flare as CF
    from flare_amd import _lib
    from flare_amd.serving import Server, latency_stats, poisson_run
    from flare_amd.workloads import batch_structures
    reqs_s = batch_structures(n_requests, seed=41, n_min=32, n_max=8192)
    mean_n = float(np.mean([len(s[1]) for s in reqs_s]))
    rate = 0.7 * batched["chosen"]["atom_evals_per_s"] / mean_n
    # Three workers (contexts) and batches of at most 32,768 atoms (the best
    # p50/p99 of the worker counts and caps tried). At 0.7x the large-batch
    # rate the server runs near saturation: batches fill to the cap and
    # their fixed cost (launches, synchronisations, small-kernel tails) and
    # the contention between overlapping batches set the latency.
    srv = Server(sgp, calc, 5.5, form=_lib.FORM_CHOSEN, hbm_budget=24 << 30, n_workers=3,
                 max_batch_atoms=32768)
    try:
        # Untimed warm-up stream of the same length and distribution (other
        # structures): the contraction tiles of each batch shape are
        # autotuned on first use, and the batch shapes that only form under
        # load would otherwise stall the timed run.
        poisson_run(srv, batch_structures(n_requests, seed=43, n_min=32, n_max=8192),
                    rate=rate, seed=4)
        srv.batches.clear()
        reqs, wall = poisson_run(srv, reqs_s, rate=rate, seed=3)
    finally:
        srv.close()
    st = latency_stats(reqs)
    atoms = sum(len(s[1]) for s in reqs_s)
    # Batch wall time against its atoms: the intercept is the fixed cost of a
    # batch (host work, launches and synchronisations), the slope its rate.
    bt = np.array([(b[1], b[2]) for b in srv.batches], np.float64)
    slope, icpt = np.polyfit(bt[:, 0], bt[:, 1], 1) if len(bt) > 2 else (float("nan"),) * 2
    # FP64 drop-in, request by request, on the first n_fp64 requests.
    lat64 = []
    t0 = time.perf_counter()
    for cell, sp, pos in reqs_s[:n_fp64]:
        a = time.perf_counter()
        s = CF.Structure(cell, sp, pos, 5.5, [calc], form=_lib.FORM_FP64)
        sgp.predict_local_uncertainties(s)
        lat64.append(time.perf_counter() - a)
    dt64 = time.perf_counter() - t0
    atoms64 = sum(len(s[1]) for s in reqs_s[:n_fp64])
    return {"workload": f"{n_requests} requests, 32-8192 atoms (mean {mean_n:.0f}), Poisson "
                        f"rate {rate:.1f}/s (0.7x batched chosen throughput); 3 workers, "
                        f"batches of at most 32768 atoms",
            "atom_evals_per_s": atoms / wall, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
            "batches": len(srv.batches),
            "mean_batch_atoms": float(bt[:, 0].mean()),
            "batch_time_fit": {"fixed_ms": float(icpt * 1e3), "ns_per_atom": float(slope * 1e9)},
            "fp64_request_by_request": {"atom_evals_per_s": atoms64 / dt64,
                                        "p50_ms": float(np.percentile(lat64, 50) * 1e3),
                                        "p99_ms": float(np.percentile(lat64, 99) * 1e3),
                                        "requests": n_fp64}}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sgp, threads, cell, species, pos, reps=5):
    """The oracle's scalable mode (cell list + per-atom adjoint with K.alpha,
    FP64, OpenMP) on the full 32,000-atom workload and model: one warm-up
    evaluation, then the median of `reps` timed ones."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import ref_cpu as R
    cd = sgp.sparse_descriptors[0]
    m = R.Model(2.0, 2.0, cd.descriptors, cd.descriptor_norms, sgp.alpha, L_inv=sgp.L_inv,
                fast=True)

    def once():
        t0 = time.perf_counter()
        nb = R.Neighbors(cell, species, pos, 6.0, "cells", fast=True)
        m.predict_scalable(nb, 2, 12, 6, np.full((2, 2), 6.0))
        return time.perf_counter() - t0

    once()
    times = sorted(once() for _ in range(reps))
    dt = times[len(times) // 2]
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"value": len(pos) / dt, "unit": "atom-evals/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus_available": avail,
            "sample": f"the full {len(pos)}-atom workload and model (FP64; neighbours, "
                      f"descriptors, K.alpha, adjoint forces, local sigma), median of {reps} "
                      f"evaluations after one warm-up: {dt:.2f} s "
                      f"(range {times[0]:.2f}-{times[-1]:.2f} s)"}


def measured_traffic():
    """HBM bytes per launch of each kernel, per form, as measured with
    rocprofv3 FETCH_SIZE + WRITE_SIZE (calibrated against a plain copy) and
    recorded in profiles/roofline_traffic.md as lines
    `traffic_bytes_per_launch <kernel>_<form> <bytes>`; empty when absent."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "roofline_traffic.md")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            m = re.match(r"^traffic_bytes_per_launch (\w+) ([0-9.eE+-]+)\s*$", line.strip())
            if m:
                out[m.group(1)] = float(m.group(2))
    return out


def sigma_errors(v64, vch, sig_f):
    """Local-uncertainty errors of the chosen form against FP64. FLARE reports
    sigma / |sigma_f| (calculator.py:108-122), so errors are given in units
    of sigma_f; the per-atom relative error is also given, over all atoms and
    over atoms with sigma >= 0.01 and >= 0.1 sigma_f (near the sparse set
    sigma^2 is a cancellation of terms of size sigma_f^2, so its relative
    error is not bounded even in FP64)."""
    s64, sch = np.sqrt(np.abs(v64)), np.sqrt(np.abs(vch))
    rel = np.abs(sch - s64) / np.maximum(s64, 1e-300)
    out = {"sigma_err_max_over_sigma_f": float(np.max(np.abs(sch - s64)) / sig_f),
           "var_err_max_over_sigma_f2": float(np.max(np.abs(vch - v64)) / sig_f ** 2),
           "sigma_rel_max": float(rel.max())}
    for t in ("0.01", "0.1"):
        big = s64 >= float(t) * sig_f
        out[f"sigma_rel_max_sigma_ge_{t}"] = float(rel[big].max()) if big.any() else None
        out[f"fraction_sigma_ge_{t}"] = float(big.mean())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sparse", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workloads", default="all",
                    choices=["single", "batched", "serving", "large", "large_b3", "all"],
                    help="single: the 32k-atom structure only; all: also the batched, "
                         "served, 1M-atom B2 and 1M-atom B2 + B3 workloads; batched / "
                         "large / large_b3: that workload alone (its own JSON line), "
                         "serving: the batched workload and the served stream at its rate")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 512 MiB buffer three times (the known-size kernel "
                         "that FETCH_SIZE / WRITE_SIZE counters are scaled by)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported: "
              "this build evaluates one GPU's share; sharding is a design only",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        a = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": a.numel() * 4, "copies": 3}), flush=True)
        return

    from flare_amd import _lib
    from flare_amd.workloads import fcc_alloy_32k
    lib = _lib.load()
    ctx = _lib.default_context()
    if args.workloads in ("batched", "serving"):
        log("batched workload ...")
        out, (bsgp, bcalc) = batched_leg(lib, ctx, Hip(), max(3, args.steps // 2), 1)
        res = {"batched": out}
        if args.workloads == "serving":
            log("serving workload ...")
            res["serving"] = serving_leg(bsgp, bcalc, out)
        print(json.dumps(res), flush=True)
        return
    if args.workloads == "large":
        log("1M-atom workload ...")
        print(json.dumps({"large_b2": large_leg(lib, ctx, Hip(), 2, 1)}), flush=True)
        return
    if args.workloads == "large_b3":
        log("B3 workload ...")
        print(json.dumps({"large_b3": large_b3_leg(lib, ctx, Hip(), 1, 1)}), flush=True)
        return
    cell, species, pos = fcc_alloy_32k()
    n = len(pos)
    log(f"workload: {n} atoms, building model ...")
    calc, sgp = build_model(None, cell, species, pos, args.sparse)
    hip = Hip()
    pos_dev = hip.upload(np.ascontiguousarray(pos.reshape(-1)))
    sp_dev = hip.upload(np.ascontiguousarray(species.astype(np.int32)))
    hip.L.hipDeviceSynchronize()
    model_h = sgp._model(0, ctx).h

    res = {}
    for name, form in (("fp64", _lib.FORM_FP64), ("chosen", _lib.FORM_CHOSEN)):
        log(f"timing form {name} ...")
        res[name] = run_form(lib, ctx, model_h, calc, cell, sp_dev, pos_dev, n, form,
                             args.steps, args.warmup)
        r = res[name]
        log(f"  {name}: {r['total'] / args.steps * 1e3:.2f} ms/step, "
            + ", ".join(f"{k} {v[0] / args.steps:.3f} ms" for k, v in r['stage'].items()))

    f64, ch = res["fp64"], res["chosen"]
    dE = abs(ch["mean"][0] - f64["mean"][0]) / n * 1e3
    dF = np.abs(ch["mean"][1:1 + 3 * n] - f64["mean"][1:1 + 3 * n]).max() * 1e3
    d = 2100
    M_by = [len(x) for x in sgp.sparse_descriptors[0].descriptor_norms]
    n_by = [int((species == t).sum()) for t in range(2)]

    traffic = measured_traffic()
    nrad_h = 2 * 12 * 49

    def summary(r, form_name):
        kr = kernel_rooflines(r, args.steps, n, n_by, M_by, d, nrad_h, form_name == "fp64")
        for k, e in kr.items():
            e["traffic"] = traffic.get(f"{k}_{form_name}")
        # The headline roofline is that of the kernel taking most of the step.
        dom = max(kr, key=lambda k: kr[k]["ms_per_step"])
        return {"atom_evals_per_s": n * args.steps / r["total"],
                "ms_per_step": r["total"] / args.steps * 1e3,
                "p50_ms": float(np.percentile(r["lat"], 50) * 1e3),
                "p99_ms": float(np.percentile(r["lat"], 99) * 1e3),
                "stage_ms": {k: round(v[0] / args.steps, 4) for k, v in r["stage"].items()},
                "roofline": dict(kernel=dom, **kr[dom]), "kernels": kr,
                "neighbors_per_atom": r["n_neighbors"] / n}

    chosen = summary(ch, "chosen")
    fp64 = summary(f64, "fp64")
    out = {
        "metric": METRIC,
        "value": chosen["atom_evals_per_s"],
        "unit": "atom-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int8x15+bf16x3+fp64",
        "data": "synthetic",
        "config": {"workload": "32,000-atom 2-species FCC alloy, B2 [2,12,6] rc 6.0 (d=2100), "
                               "2,000 sparse envs, NDP sigma 2 p 2; E/F/stress/local sigma",
                   "form": "chosen: descriptors computed in FP64, kept as FP32 rows for "
                           "the gradient contraction and adjoint and as int8 slices (5 per "
                           "operand) for the kernel contraction (15 products, exact int32 "
                           "sums, FP64 epilogue); gradient contraction split-BF16x3 (FP32 "
                           "accumulate); FP32 adjoint; FP64 variance contraction",
                   "n_atoms": n, "n_sparse": int(sum(M_by)), "n_descriptors": d},
        "p50_ms": chosen["p50_ms"],
        "p99_ms": chosen["p99_ms"],
        "roofline": chosen["roofline"],
        "kernels": chosen["kernels"],
        "stage_ms": chosen["stage_ms"],
        "fp64_dropin": fp64,
        "speedup_vs_fp64": chosen["atom_evals_per_s"] / fp64["atom_evals_per_s"],
        "error_vs_fp64": {"energy_meV_per_atom": dE, "force_max_meV_per_A": dF,
                          **sigma_errors(f64["var"], ch["var"], sgp.kernels[0].sigma)},
    }
    if args.workloads == "all":
        log("batched workload ...")
        out["batched"], (bsgp, bcalc) = batched_leg(lib, ctx, hip, max(3, args.steps // 2),
                                                    1)
        log("serving workload ...")
        out["serving"] = serving_leg(bsgp, bcalc, out["batched"])
        log("1M-atom workload ...")
        out["large_b2"] = large_leg(lib, ctx, hip, 2, 1)
        log("1M-atom B2 + B3 workload ...")
        out["large_b3"] = large_b3_leg(lib, ctx, hip, 1, 1)
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads ...")
        out["cpu_baseline"] = cpu_baseline(sgp, threads, cell, species, pos)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
