// The C ABI of libflare_mi (include/flare_mi.h): contexts, structures,
// models and predictions. Host-side setup (3x3 cell algebra, model upload)
// lives here; every per-atom and per-bond stage runs on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <memory>
#include <vector>

#include "fm_gemm.hpp"
#include "fm_internal.hpp"

namespace {
thread_local std::string g_last_error;

template <typename F>
int guarded(F &&f) {
  try {
    f();
    return FM_OK;
  } catch (const fm::Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return FM_ERR_INVALID;
  }
}

}  // namespace

namespace fm {
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace fm

namespace {

// Cofactor inverse and first-row determinant of a 3x3 matrix, the cell
// algebra of structure.cpp:16-21 (cell_transpose_inverse, cell_dot_inverse,
// volume = |det cell|). This translation unit is built with
// -ffp-contract=off so these match the oracle's expressions bit for bit.
double det3(const double *m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

void inverse3(const double *m, double *o) {
  double c00 = m[4] * m[8] - m[5] * m[7];
  double c01 = m[5] * m[6] - m[3] * m[8];
  double c02 = m[3] * m[7] - m[4] * m[6];
  double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  double inv = 1.0 / det;
  o[0] = c00 * inv;
  o[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
  o[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
  o[3] = c01 * inv;
  o[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
  o[5] = (m[2] * m[3] - m[0] * m[5]) * inv;
  o[6] = c02 * inv;
  o[7] = (m[1] * m[6] - m[0] * m[7]) * inv;
  o[8] = (m[0] * m[4] - m[1] * m[3]) * inv;
}

// structure.cpp:135-165: the smallest in-plane height of the cell faces.
double single_sweep_cutoff(const double *c) {
  const double *v1 = c, *v2 = c + 3, *v3 = c + 6;
  double ab = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
  double ac = v1[0] * v3[0] + v1[1] * v3[1] + v1[2] * v3[2];
  double bc = v2[0] * v3[0] + v2[1] * v3[1] + v2[2] * v3[2];
  double a = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  double b = std::sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
  double cc = std::sqrt(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
  double cand[6] = {a * std::sqrt(1 - std::pow(ab / (a * b), 2)),
                    b * std::sqrt(1 - std::pow(ab / (a * b), 2)),
                    a * std::sqrt(1 - std::pow(ac / (a * cc), 2)),
                    cc * std::sqrt(1 - std::pow(ac / (a * cc), 2)),
                    b * std::sqrt(1 - std::pow(bc / (b * cc), 2)),
                    cc * std::sqrt(1 - std::pow(bc / (b * cc), 2))};
  double best = cand[0];
  for (int i = 0; i < 6; i++)
    if (cand[i] < best) best = cand[i];
  return best;
}

void cross3(const double *u, const double *v, double *o) {
  o[0] = u[1] * v[2] - u[2] * v[1];
  o[1] = u[2] * v[0] - u[0] * v[2];
  o[2] = u[0] * v[1] - u[1] * v[0];
}

// The geometry of one structure: structure.cpp:7-43 cell algebra and sweep,
// plus the cell-list bins (width >= cutoff perpendicular to each face).
void make_cellargs(const double *cell, double cutoff, fm::CellArgs &ca, double &volume,
                   double &sscut) {
  for (int i = 0; i < 9; i++) ca.cell[i] = cell[i];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) ca.ct[3 * r + c] = cell[3 * c + r];
  inverse3(ca.ct, ca.cti);
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++)
      ca.cd[3 * r + c] = (cell[3 * r] * cell[3 * c] + cell[3 * r + 1] * cell[3 * c + 1]) +
                         cell[3 * r + 2] * cell[3 * c + 2];
  inverse3(ca.cd, ca.cdi);
  volume = std::fabs(det3(cell));
  sscut = single_sweep_cutoff(cell);
  FM_CHECK(std::isfinite(volume) && volume > 0, FM_ERR_INVALID, "cell is singular");
  FM_CHECK(cutoff > 0 && std::isfinite(cutoff), FM_ERR_INVALID, "cutoff must be positive");
  double sw = std::ceil(cutoff / sscut);
  FM_CHECK(sw >= 1 && sw <= 64, FM_ERR_CAPACITY,
           "cutoff spans too many periodic images of the cell");
  ca.cutoff = cutoff;
  ca.sweep = (int)sw;
  double cr[3][3];
  cross3(cell + 3, cell + 6, cr[0]);
  cross3(cell + 6, cell, cr[1]);
  cross3(cell, cell + 3, cr[2]);
  double rc = cutoff * (1 + 1e-9) + 1e-9;
  for (int ax = 0; ax < 3; ax++) {
    double h = volume / std::sqrt(cr[ax][0] * cr[ax][0] + cr[ax][1] * cr[ax][1] +
                                  cr[ax][2] * cr[ax][2]);
    int nb = std::max(1, (int)std::floor(h / rc));
    nb = std::min(nb, 1024);
    ca.nb[ax] = nb;
    ca.K[ax] = (int)std::ceil(rc / (h / nb));
  }
}

template <typename T>
void d2h(T *dst, const fm::DevBuf &src, size_t count, hipStream_t st) {
  if (dst && count) FM_HIP(hipMemcpyAsync(dst, src.as<T>(), count * sizeof(T), hipMemcpyDeviceToHost, st));
}

// flags[0]: bit 1 non-finite position, bit 2 negative species;
// flags[1]: largest species.
__global__ void k_check_positions(int n, const double *pos, const int *species, int *flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 3; k++)
    if (!isfinite(pos[3 * i + k])) atomicOr(&flags[0], 1);
  if (species[i] < 0) atomicOr(&flags[0], 2);
  atomicMax(&flags[1], species[i]);
}

}  // namespace

namespace fm {

Pool::~Pool() {
  for (auto &kv : free_)
    for (void *p : kv.second) (void)hipFree(p);
}

void *Pool::get(size_t bytes, size_t *cls) {
  size_t c = 256;
  while (c < bytes) c <<= 1;
  *cls = c;
  auto it = free_.find(c);
  if (it != free_.end() && !it->second.empty()) {
    void *p = it->second.back();
    it->second.pop_back();
    cached_ -= c;
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, c);
  if (e != hipSuccess) {
    // Drop the cache and retry once before giving up.
    for (auto &kv : free_)
      for (void *q : kv.second) (void)hipFree(q);
    free_.clear();
    cached_ = 0;
    FM_HIP(hipMalloc(&p, c));
  }
  return p;
}

void Pool::put(void *p, size_t cls) {
  free_[cls].push_back(p);
  cached_ += cls;
}

}  // namespace fm

void fm_context_s::mark_begin(int s, hipStream_t on) {
  (void)hipEventRecord(ev_begin[s], on ? on : stream);
  launches[s] = 0;
}
void fm_context_s::mark_end(int s, hipStream_t on) {
  (void)hipEventRecord(ev_end[s], on ? on : stream);
  recorded[s] = true;
}
void fm_context_s::kt_clear(int first, int last) {
  for (int k = first; k <= last; k++) kt[k].used = 0;
}
void fm_context_s::kt_begin(int k, hipStream_t on) {
  if (!kt_on) return;
  KTimes &t = kt[k];
  while ((int)t.ev.size() < 2 * (t.used + 1)) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    t.ev.push_back(e);
  }
  (void)hipEventRecord(t.ev[2 * t.used], on ? on : stream);
}
void fm_context_s::kt_end(int k, hipStream_t on) {
  if (!kt_on) return;
  KTimes &t = kt[k];
  if ((int)t.ev.size() < 2 * (t.used + 1)) return;
  (void)hipEventRecord(t.ev[2 * t.used + 1], on ? on : stream);
  t.used++;
}

extern "C" {

const char *fm_last_error(void) { return g_last_error.c_str(); }

int fm_context_create(int device, fm_context *out) {
  return guarded([&] {
    FM_CHECK(out, FM_ERR_INVALID, "null output handle");
    FM_HIP(hipSetDevice(device));
    fm::scratch_setup(device);
    auto *c = new fm_context_s();
    c->device = device;
    FM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    FM_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    FM_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    FM_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    FM_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->pinned),
                         fm_context_s::kPinnedWords * sizeof(int), hipHostMallocDefault));
    for (int s = 0; s < fm::ST_COUNT; s++) {
      FM_HIP(hipEventCreate(&c->ev_begin[s]));
      FM_HIP(hipEventCreate(&c->ev_end[s]));
    }
    *out = c;
  });
}

// Releases the streams, events and pinned words of a context and deletes it.
static void context_delete(fm_context ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int s = 0; s < fm::ST_COUNT; s++) {
      (void)hipEventDestroy(ctx->ev_begin[s]);
      (void)hipEventDestroy(ctx->ev_end[s]);
    }
    for (auto &t : ctx->kt)
      for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
    hipStream_t st = ctx->stream, side = ctx->side;
    (void)hipStreamSynchronize(side);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    (void)hipEventDestroy(ctx->ev_fork);
    (void)hipEventDestroy(ctx->ev_join);
    delete ctx;
    (void)hipStreamDestroy(side);
    (void)hipStreamDestroy(st);
}

// An object made on ctx: counted until ctx_release.
static void ctx_acquire(fm_context ctx) { ctx->life.fetch_add(2); }

// An object made on ctx is gone: delete ctx if it was the last one and the
// context's destroy call came before.
static void ctx_release(fm_context ctx) {
  if (ctx->life.fetch_sub(2) == 3) context_delete(ctx);
}

int fm_context_destroy(fm_context ctx) {
  return guarded([&] {
    if (!ctx) return;
    int64_t old = ctx->life.fetch_or(1);
    FM_CHECK(!(old & 1), FM_ERR_STATE, "context destroyed twice");
    if (old == 0) context_delete(ctx);
  });
}

int fm_context_set_row_tile(fm_context ctx, int64_t rows) {
  return guarded([&] {
    FM_CHECK(ctx, FM_ERR_INVALID, "null context");
    FM_CHECK(rows >= 0 && rows <= (1ll << 30), FM_ERR_INVALID, "row tile out of range");
    ctx->row_tile = rows;
  });
}

int fm_context_synchronize(fm_context ctx) {
  return guarded([&] { FM_HIP(hipStreamSynchronize(ctx->stream)); });
}

int fm_context_stream(fm_context ctx, void **stream) {
  return guarded([&] { *stream = (void *)ctx->stream; });
}

static int structure_create(fm_context ctx, int n_struct, const double *cells,
                            const int32_t *n_atoms, const int32_t *species,
                            const double *positions, double cutoff, bool device_ptrs,
                            fm_structure *out) {
  return guarded([&] {
    FM_CHECK(ctx && out && cells && n_atoms, FM_ERR_INVALID, "null argument");
    FM_CHECK(n_struct >= 1, FM_ERR_INVALID, "a batch needs at least one structure");
    long long ntot = 0;
    for (int k = 0; k < n_struct; k++) {
      FM_CHECK(n_atoms[k] >= 0, FM_ERR_INVALID, "negative atom count");
      ntot += n_atoms[k];
    }
    FM_CHECK(ntot < (1LL << 31), FM_ERR_CAPACITY, "too many atoms");
    int n = (int)ntot;
    FM_CHECK(n == 0 || (species && positions), FM_ERR_INVALID, "null species or positions");
    FM_HIP(hipSetDevice(ctx->device));
    auto *s = new fm_structure_s();
    s->ctx = ctx;
    s->n = n;
    s->n_struct = n_struct;
    try {
      fm::Pool *pool = &ctx->pool;
      hipStream_t st = ctx->stream;
      s->atom_off.assign(n_struct + 1, 0);
      s->cellargs.resize(n_struct);
      s->volumes.resize(n_struct);
      s->sweep_cutoffs.resize(n_struct);
      std::vector<int> ast(n);
      std::vector<double> vol_inv(n_struct);
      for (int k = 0; k < n_struct; k++) {
        make_cellargs(cells + 9 * (size_t)k, cutoff, s->cellargs[k], s->volumes[k],
                      s->sweep_cutoffs[k]);
        vol_inv[k] = 1.0 / s->volumes[k];
        s->atom_off[k + 1] = s->atom_off[k] + n_atoms[k];
        for (int i = s->atom_off[k]; i < s->atom_off[k + 1]; i++) ast[i] = k;
      }
      const fm::CellArgs &c0 = s->cellargs[0];
      for (int i = 0; i < 9; i++) {
        s->cell[i] = c0.cell[i];
        s->cell_t_inv[i] = c0.cti[i];
        s->cell_dot[i] = c0.cd[i];
        s->cell_dot_inv[i] = c0.cdi[i];
      }
      s->volume = s->volumes[0];
      s->single_sweep_cutoff = s->sweep_cutoffs[0];
      s->cutoff = cutoff;
      s->sweep = c0.sweep;
      auto up = [&](fm::DevBuf &buf, const void *src, size_t bytes) {
        buf.alloc(pool, std::max<size_t>(1, bytes));
        if (bytes) FM_HIP(hipMemcpyAsync(buf.as<void>(), src, bytes, hipMemcpyHostToDevice, st));
      };
      up(s->d_cellargs, s->cellargs.data(), s->cellargs.size() * sizeof(fm::CellArgs));
      up(s->atom_struct, ast.data(), ast.size() * sizeof(int));
      up(s->d_atom_off, s->atom_off.data(), s->atom_off.size() * sizeof(int));
      up(s->d_vol_inv, vol_inv.data(), vol_inv.size() * sizeof(double));
      // mean_efs offsets and reduction chunks per structure.
      s->mean_off.assign(n_struct + 1, 0);
      std::vector<int> cb, ce, sco(n_struct + 1, 0);
      for (int k = 0; k < n_struct; k++) {
        s->mean_off[k + 1] = s->mean_off[k] + 1 + 3 * (int64_t)n_atoms[k] + 6;
        for (int a = s->atom_off[k]; a < s->atom_off[k + 1]; a += 256) {
          cb.push_back(a);
          ce.push_back(std::min(a + 256, s->atom_off[k + 1]));
        }
        sco[k + 1] = (int)cb.size();
      }
      s->n_chunks = (int)cb.size();
      std::vector<long long> mo(s->mean_off.begin(), s->mean_off.end());
      up(s->d_mean_off, mo.data(), mo.size() * sizeof(long long));
      up(s->d_chunk_begin, cb.data(), cb.size() * sizeof(int));
      up(s->d_chunk_end, ce.data(), ce.size() * sizeof(int));
      up(s->d_struct_chunk_off, sco.data(), sco.size() * sizeof(int));
      s->species.alloc(pool, std::max<size_t>(1, (size_t)n * sizeof(int)));
      s->positions.alloc(pool, std::max<size_t>(1, 3 * (size_t)n * sizeof(double)));
      hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
      if (n > 0) {
        FM_HIP(hipMemcpyAsync(s->species.as<int>(), species, (size_t)n * sizeof(int), kind, st));
        FM_HIP(hipMemcpyAsync(s->positions.as<double>(), positions, 3 * (size_t)n * sizeof(double),
                              kind, st));
        // The position / species flags are read back with the neighbour
        // build's size readback (one synchronisation); the neighbour kernels
        // only bin (clamped) and copy species, so bad input cannot take
        // them out of bounds before it is reported.
        fm::DevBuf flags(pool, 2 * sizeof(int));
        FM_HIP(hipMemsetAsync(flags.as<int>(), 0, 2 * sizeof(int), st));
        hipLaunchKernelGGL(k_check_positions, dim3((n + 255) / 256), dim3(256), 0, st, n,
                           s->positions.as<double>(), s->species.as<int>(), flags.as<int>());
        FM_HIP(hipMemcpyAsync(ctx->pinned, flags.as<int>(), 2 * sizeof(int),
                              hipMemcpyDeviceToHost, st));
      }
      ctx->mark_begin(fm::ST_NEIGH);
      if (n > 0) {
        fm::build_neighbors(s);  // synchronises: ctx->pinned[0, 2) is valid after
        int hf0 = ctx->pinned[0], hf1 = ctx->pinned[1];
        FM_CHECK(!(hf0 & 1), FM_ERR_INVALID, "positions contain NaN or Inf");
        FM_CHECK(!(hf0 & 2), FM_ERR_INVALID, "species must be non-negative");
        s->max_species = hf1;
      } else {
        s->cumulative.alloc(pool, sizeof(int));
        FM_HIP(hipMemsetAsync(s->cumulative.as<int>(), 0, sizeof(int), st));
      }
      ctx->mark_end(fm::ST_NEIGH);
    } catch (...) {
      delete s;
      throw;
    }
    ctx_acquire(ctx);
    *out = s;
  });
}

int fm_structure_create(fm_context ctx, const double *cell, int n_atoms, const int32_t *species,
                        const double *positions, double cutoff, fm_structure *out) {
  return structure_create(ctx, 1, cell, &n_atoms, species, positions, cutoff, false, out);
}

int fm_structure_create_device(fm_context ctx, const double *cell, int n_atoms,
                               const int32_t *species_dev, const double *positions_dev,
                               double cutoff, fm_structure *out) {
  return structure_create(ctx, 1, cell, &n_atoms, species_dev, positions_dev, cutoff, true, out);
}

int fm_batch_create(fm_context ctx, int n_structures, const double *cells, const int32_t *n_atoms,
                    const int32_t *species, const double *positions, double cutoff,
                    fm_structure *out) {
  return structure_create(ctx, n_structures, cells, n_atoms, species, positions, cutoff, false,
                          out);
}

int fm_batch_create_device(fm_context ctx, int n_structures, const double *cells,
                           const int32_t *n_atoms, const int32_t *species_dev,
                           const double *positions_dev, double cutoff, fm_structure *out) {
  return structure_create(ctx, n_structures, cells, n_atoms, species_dev, positions_dev, cutoff,
                          true, out);
}

int fm_batch_info(fm_structure s, int k, int *atom_offset, int *sweep, double *volume) {
  return guarded([&] {
    FM_CHECK(s, FM_ERR_INVALID, "null structure");
    FM_CHECK(k >= 0 && k < s->n_struct, FM_ERR_INVALID, "structure index out of range");
    if (atom_offset) *atom_offset = s->atom_off[k];
    if (sweep) *sweep = s->cellargs[k].sweep;
    if (volume) *volume = s->volumes[k];
  });
}

int fm_structure_destroy(fm_structure s) {
  return guarded([&] {
    if (!s) return;
    fm_context ctx = s->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    fm::free_b3_state(s);
    delete s;
    ctx_release(ctx);
  });
}

int fm_structure_info(fm_structure s, int *sweep, int64_t *n_neighbors, double *volume,
                      double *single_sweep_cutoff) {
  return guarded([&] {
    FM_CHECK(s, FM_ERR_INVALID, "null structure");
    if (sweep) *sweep = s->sweep;
    if (n_neighbors) *n_neighbors = s->n_nb;
    if (volume) *volume = s->volume;
    if (single_sweep_cutoff) *single_sweep_cutoff = s->single_sweep_cutoff;
  });
}

int fm_structure_get_neighbors(fm_structure s, double *wrapped, int32_t *count, int32_t *cumulative,
                               int32_t *structure_indices, int32_t *neighbor_species,
                               double *relative) {
  return guarded([&] {
    FM_CHECK(s, FM_ERR_INVALID, "null structure");
    hipStream_t st = s->ctx->stream;
    if (s->n > 0) {
      d2h(wrapped, s->wrapped, 3 * (size_t)s->n, st);
      d2h(count, s->count, (size_t)s->n, st);
    }
    d2h(cumulative, s->cumulative, (size_t)s->n + 1, st);
    d2h(structure_indices, s->sidx, (size_t)s->n_nb, st);
    d2h(neighbor_species, s->nspecies, (size_t)s->n_nb, st);
    d2h(relative, s->rel, 4 * (size_t)s->n_nb, st);
    FM_HIP(hipStreamSynchronize(st));
  });
}

// B2 (kind 2) or B3 (kind 3) descriptors of a structure; B3 keeps the B2
// bond pass (c only) and adds the triple products (fm_b3.hip).
static void compute_descriptor(fm_structure s, const fm_b2_settings *st_in, int form, int kind) {
  {
    FM_CHECK(s && st_in && st_in->cutoffs, FM_ERR_INVALID, "null argument");
    FM_CHECK(form == FM_FORM_FP64 || form == FM_FORM_CHOSEN, FM_ERR_INVALID, "unknown form");
    FM_CHECK(st_in->n_species >= 1 && st_in->n_max >= 1 && st_in->l_max >= 0, FM_ERR_INVALID,
             "descriptor settings must be positive");
    FM_CHECK(st_in->cutoff_function == FM_CUTOFF_QUADRATIC ||
                 st_in->cutoff_function == FM_CUTOFF_COSINE,
             FM_ERR_INVALID, "cutoff function must be quadratic or cosine");
    fm_context ctx = s->ctx;
    FM_HIP(hipSetDevice(ctx->device));
    hipStream_t stream = ctx->stream;
    fm::Pool *pool = &ctx->pool;
    s->nos = st_in->n_species;
    s->N = st_in->n_max;
    s->lmax = st_in->l_max;
    s->cutkind = st_in->cutoff_function;
    s->r1 = st_in->radial_r1;
    s->form = form;
    s->cutoffs.assign(st_in->cutoffs, st_in->cutoffs + s->nos * s->nos);
    for (double c : s->cutoffs)
      FM_CHECK(c > s->r1 && std::isfinite(c), FM_ERR_INVALID, "pair cutoffs must exceed r1");
    int nr = s->nos * s->N;
    fm::free_b3_state(s);
    s->desc_kind = kind;
    if (kind == 3) {
      FM_CHECK(s->lmax <= 4, FM_ERR_INVALID,
               "B3 needs l_max <= 4 (the Wigner 3j table of wigner3j.cpp)");
      s->nd = fm::b3_size(s->nos, s->N, s->lmax);
    } else {
      s->nd = nr * (nr + 1) / 2 * (s->lmax + 1);
    }
    if (s->n > 0) {
      // Species must lie in [0, n_species) for this descriptor.
      FM_CHECK(s->max_species < s->nos, FM_ERR_INVALID, "species outside [0, n_species)");
    }
    std::vector<int> pairs;
    for (int a = 0; a < nr; a++)
      for (int b = a; b < nr; b++) {
        pairs.push_back(a);
        pairs.push_back(b);
      }
    s->pair_table.alloc(pool, pairs.size() * sizeof(int));
    FM_HIP(hipMemcpyAsync(s->pair_table.as<int>(), pairs.data(), pairs.size() * sizeof(int),
                          hipMemcpyHostToDevice, stream));
    ctx->mark_begin(fm::ST_DESC);
    ctx->kt_clear(fm::KT_DESC, fm::KT_DESC);
    if (s->n > 0) fm::build_bonds(s);
    else {
      s->n_by_type.assign(s->nos, 0);
      s->type_off.assign(s->nos + 1, 0);
      s->n_bonds = 0;
    }
    size_t esz = form == FM_FORM_FP64 ? sizeof(double) : sizeof(float);
    // B3 rows too large to keep are formed in row tiles inside fm_predict.
    s->row_tile = 0;
    if (kind == 3 && s->n > 0) {
      size_t per_row = (size_t)s->nd * esz;
      if (ctx->row_tile > 0) {
        s->row_tile = ctx->row_tile;
      } else if ((size_t)s->n * per_row > (32ull << 30)) {
        // B, G and (chosen form) the int8 slices of B, per row.
        size_t tile_row = (size_t)s->nd * (2 * esz + (form == FM_FORM_CHOSEN ? fm::kI8Slices : 0));
        s->row_tile = std::max<int64_t>(256, (int64_t)((8ull << 30) / tile_row) / 256 * 256);
      }
    }
    size_t rows_kept = s->row_tile > 0 ? 0 : (size_t)s->n;
    s->B.alloc(pool, std::max<size_t>(1, rows_kept * s->nd * esz));
    s->norm.alloc(pool, std::max<size_t>(1, (size_t)s->n * sizeof(double)));
    s->nn.alloc(pool, std::max<size_t>(1, (size_t)s->n * sizeof(double)));
    if (s->n > 0) {
      fm::launch_b2(s, kind == 2);
      if (kind == 3 && s->row_tile == 0) fm::launch_b3(s);
    }
    ctx->mark_end(fm::ST_DESC);
    // Per-structure species-sorted position of every row (batches only).
    fm::build_var_pos(s);
    s->nb_by_type.clear();  // filled on first export (nb_by_type_of)
    s->has_dervs = false;    // dense derivatives are rebuilt on demand
    s->DB.release();
    s->has_b2 = true;
  }
}

int fm_structure_compute_b2(fm_structure s, const fm_b2_settings *st_in, int form) {
  return guarded([&] { compute_descriptor(s, st_in, form, 2); });
}

int fm_structure_compute_b3(fm_structure s, const fm_b2_settings *st_in, int form) {
  return guarded([&] { compute_descriptor(s, st_in, form, 3); });
}

// Neighbour counts by species (b2.cpp:106-112), computed on first use.
static void nb_by_type_of(fm_structure s) {
  if (!s->nb_by_type.empty()) return;
  s->nb_by_type.assign(s->nos, 0);
  if (s->n == 0) return;
  hipStream_t st = s->ctx->stream;
  std::vector<int> bo(s->n + 1), sp(s->n);
  FM_HIP(hipMemcpyAsync(bo.data(), s->bond_off.as<int>(), (s->n + 1) * sizeof(int),
                        hipMemcpyDeviceToHost, st));
  FM_HIP(hipMemcpyAsync(sp.data(), s->species.as<int>(), s->n * sizeof(int),
                        hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < s->n; i++) s->nb_by_type[sp[i]] += bo[i + 1] - bo[i];
}

int fm_structure_b2_sizes(fm_structure s, int *n_descriptors, int32_t *n_clusters_by_type,
                          int32_t *n_neighbors_by_type) {
  return guarded([&] {
    FM_CHECK(s && s->has_b2, FM_ERR_STATE, "descriptors have not been computed");
    nb_by_type_of(s);
    if (n_descriptors) *n_descriptors = s->nd;
    for (int t = 0; t < s->nos; t++) {
      if (n_clusters_by_type) n_clusters_by_type[t] = s->n_by_type[t];
      if (n_neighbors_by_type) n_neighbors_by_type[t] = s->nb_by_type[t];
    }
  });
}

int fm_structure_get_b2(fm_structure s, int t, double *descriptors, double *norms,
                        int32_t *neighbor_counts, int32_t *cumulative_neighbor_counts,
                        int32_t *atom_indices, int32_t *neighbor_indices,
                        double *neighbor_coordinates) {
  return guarded([&] {
    FM_CHECK(s && s->has_b2, FM_ERR_STATE, "descriptors have not been computed");
    FM_CHECK(s->n_struct == 1, FM_ERR_INVALID, "descriptor export is per single structure");
    FM_CHECK(s->row_tile == 0 || (!descriptors && !norms), FM_ERR_STATE,
             "this B3 structure forms its descriptors in row tiles during prediction; they "
             "are not kept (see fm_context_set_row_tile)");
    FM_CHECK(t >= 0 && t < s->nos, FM_ERR_INVALID, "species out of range");
    hipStream_t st = s->ctx->stream;
    int ns = s->n_by_type[t], ro = s->type_off[t], nd = s->nd;
    if (ns == 0) return;
    if (descriptors) {
      // B3 rows are stored l-set major on the device: restore the
      // reference's order (b3.cpp:230-242) in a device copy first.
      size_t esz = s->form == FM_FORM_FP64 ? sizeof(double) : sizeof(float);
      fm::DevBuf ref;
      const void *src = (const char *)s->B.as<char>() + (size_t)ro * nd * esz;
      if (s->desc_kind == 3) {
        ref.alloc(&s->ctx->pool, (size_t)ns * nd * esz);
        fm::b3_rows_to_reference(s, ro, ns, ref.as<void>());
        src = ref.as<void>();
      }
      if (s->form == FM_FORM_FP64) {
        FM_HIP(hipMemcpyAsync(descriptors, src, (size_t)ns * nd * sizeof(double),
                              hipMemcpyDeviceToHost, st));
        FM_HIP(hipStreamSynchronize(st));
      } else {
        std::vector<float> tmp((size_t)ns * nd);
        FM_HIP(hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(float),
                              hipMemcpyDeviceToHost, st));
        FM_HIP(hipStreamSynchronize(st));
        for (size_t k = 0; k < tmp.size(); k++) descriptors[k] = tmp[k];
      }
    }
    if (norms)
      FM_HIP(hipMemcpyAsync(norms, s->norm.as<double>() + ro, (size_t)ns * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    std::vector<int> aor(ns), bo(s->n + 1);
    FM_HIP(hipMemcpyAsync(aor.data(), s->atom_of_row.as<int>() + ro, ns * sizeof(int),
                          hipMemcpyDeviceToHost, st));
    FM_HIP(hipMemcpyAsync(bo.data(), s->bond_off.as<int>(), (s->n + 1) * sizeof(int),
                          hipMemcpyDeviceToHost, st));
    std::vector<int> bq(s->n_bonds), sidx(s->n_nb);
    std::vector<double> rel(4 * (size_t)s->n_nb);
    if (neighbor_indices || neighbor_coordinates) {
      if (s->n_bonds)
        FM_HIP(hipMemcpyAsync(bq.data(), s->bond_q.as<int>(), bq.size() * sizeof(int),
                              hipMemcpyDeviceToHost, st));
      if (s->n_nb) {
        FM_HIP(hipMemcpyAsync(sidx.data(), s->sidx.as<int>(), sidx.size() * sizeof(int),
                              hipMemcpyDeviceToHost, st));
        FM_HIP(hipMemcpyAsync(rel.data(), s->rel.as<double>(), rel.size() * sizeof(double),
                              hipMemcpyDeviceToHost, st));
      }
    }
    FM_HIP(hipStreamSynchronize(st));
    int cum = 0;
    for (int r = 0; r < ns; r++) {
      int i = aor[r];
      int nb = bo[i + 1] - bo[i];
      if (neighbor_counts) neighbor_counts[r] = nb;
      if (cumulative_neighbor_counts) cumulative_neighbor_counts[r] = cum;
      if (atom_indices) atom_indices[r] = i;
      for (int b = 0; b < nb; b++) {
        int q = bq.empty() ? 0 : bq[bo[i] + b];
        if (neighbor_indices) neighbor_indices[cum + b] = sidx[q];
        if (neighbor_coordinates)
          for (int k = 0; k < 3; k++) neighbor_coordinates[3 * (size_t)(cum + b) + k] = rel[4 * (size_t)q + 1 + k];
      }
      cum += nb;
    }
  });
}

int fm_model_create(fm_context ctx, const fm_model_desc *d, fm_model *out) {
  return guarded([&] {
    FM_CHECK(ctx && d && out, FM_ERR_INVALID, "null argument");
    FM_CHECK(d->n_species >= 1 && d->n_descriptors >= 1, FM_ERR_INVALID, "empty model");
    FM_CHECK(d->n_sparse_by_type && d->alpha, FM_ERR_INVALID, "null sparse counts or alpha");
    FM_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    auto *m = new fm_model_s();
    std::unique_ptr<fm_model_s> owner(m);  // deleted if a check below throws
    m->ctx = ctx;
    m->sigma = d->sigma;
    m->power = d->power;
    m->nos = d->n_species;
    m->nd = d->n_descriptors;
    int nd = m->nd, M = 0;
    for (int t = 0; t < m->nos; t++) {
      FM_CHECK(d->n_sparse_by_type[t] >= 0, FM_ERR_INVALID, "negative sparse count");
      m->off.push_back(M);
      m->M_by_type.push_back(d->n_sparse_by_type[t]);
      M += d->n_sparse_by_type[t];
    }
    m->M = M;
    FM_CHECK(M == 0 || (d->sparse_descriptors && d->sparse_norms), FM_ERR_INVALID,
             "null sparse descriptors");
    m->host_alpha.assign(d->alpha, d->alpha + M);
    m->has_L = d->L_inv != nullptr;
    fm::Pool *pool = &ctx->pool;
    for (int t = 0; t < m->nos; t++) {
      int Ms = m->M_by_type[t], c0 = m->off[t];
      std::vector<double> X((size_t)Ms * nd, 0.0), XT((size_t)nd * Ms, 0.0), a(Ms, 0.0);
      size_t ldd = fm::split_ld(nd), ldm = fm::split_ld(Ms);
      std::vector<float> X32((size_t)Ms * ldd, 0.0f), XT32((size_t)nd * ldm, 0.0f);
      for (int j = 0; j < Ms; j++) {
        double nrm = d->sparse_norms[c0 + j];
        // Empty sparse environments contribute nothing
        // (normalized_dot_product.cpp:158-162).
        if (nrm < 1e-8) continue;
        for (int k = 0; k < nd; k++) {
          double v = d->sparse_descriptors[(size_t)(c0 + j) * nd + k] / nrm;
          X[(size_t)j * nd + k] = v;
          XT[(size_t)k * Ms + j] = v;
        }
        a[j] = d->alpha[c0 + j];
      }
      for (int j = 0; j < Ms; j++)
        for (int k = 0; k < nd; k++) X32[j * ldd + k] = (float)X[(size_t)j * nd + k];
      for (int k = 0; k < nd; k++)
        for (int j = 0; j < Ms; j++) XT32[k * ldm + j] = (float)XT[(size_t)k * Ms + j];
      std::vector<uint16_t> xs[2], xts[2];
      for (auto &v : xs) v.resize(X32.size());
      for (auto &v : xts) v.resize(XT32.size());
      uint16_t *xsp[2] = {xs[0].data(), xs[1].data()};
      uint16_t *xtsp[2] = {xts[0].data(), xts[1].data()};
      fm::split_bf16_host(X32.data(), X32.size(), 2, xsp);
      fm::split_bf16_host(XT32.data(), XT32.size(), 2, xtsp);
      auto up = [&](fm::DevBuf &buf, const void *src, size_t bytes) {
        buf.alloc(pool, std::max<size_t>(1, bytes));
        if (bytes) FM_HIP(hipMemcpyAsync(buf.as<void>(), src, bytes, hipMemcpyHostToDevice, st));
      };
      m->X.emplace_back();
      m->XT.emplace_back();
      m->Xs0.emplace_back();
      m->Xs1.emplace_back();
      m->XTs0.emplace_back();
      m->XTs1.emplace_back();
      m->alpha.emplace_back();
      m->L.emplace_back();
      up(m->X.back(), X.data(), X.size() * sizeof(double));
      up(m->XT.back(), XT.data(), XT.size() * sizeof(double));
      up(m->Xs0.back(), xs[0].data(), xs[0].size() * sizeof(uint16_t));
      up(m->Xs1.back(), xs[1].data(), xs[1].size() * sizeof(uint16_t));
      up(m->XTs0.back(), xts[0].data(), xts[0].size() * sizeof(uint16_t));
      up(m->XTs1.back(), xts[1].data(), xts[1].size() * sizeof(uint16_t));
      up(m->alpha.back(), a.data(), a.size() * sizeof(double));
      int rows = 0, lower = 1;
      if (m->has_L && Ms > 0) {
        // Rows [c0, r_end) of the L_inv column block; the reference takes
        // the whole kernel block (sparse_gp.cpp:121-130), whose rows outside
        // this range are zero in these columns.
        int r_begin = c0, r_end = c0 + Ms;
        for (int r = 0; r < c0; r++)
          for (int j = 0; j < Ms; j++)
            if (d->L_inv[(size_t)r * M + c0 + j] != 0.0) r_begin = std::min(r_begin, r);
        for (int r = c0 + Ms; r < M; r++)
          for (int j = 0; j < Ms; j++)
            if (d->L_inv[(size_t)r * M + c0 + j] != 0.0) r_end = std::max(r_end, r + 1);
        FM_CHECK(r_begin == c0, FM_ERR_INVALID,
                 "L_inv has entries above the diagonal block of a species");
        for (int r = 0; r < Ms; r++)
          for (int j = r + 1; j < Ms; j++)
            if (d->L_inv[(size_t)(c0 + r) * M + c0 + j] != 0.0) lower = 0;
        rows = r_end - c0;
        std::vector<double> L((size_t)rows * Ms);
        for (int r = 0; r < rows; r++)
          for (int j = 0; j < Ms; j++) L[(size_t)r * Ms + j] = d->L_inv[(size_t)(c0 + r) * M + c0 + j];
        up(m->L.back(), L.data(), L.size() * sizeof(double));
        FM_HIP(hipStreamSynchronize(st));
      }
      m->L_rows.push_back(rows);
      m->L_lower.push_back(lower);
      if (rows != Ms) m->L_block_diag = false;
      m->xnorm.emplace_back(d->sparse_norms ? d->sparse_norms + c0 : nullptr,
                            d->sparse_norms ? d->sparse_norms + c0 + Ms : nullptr);
      FM_HIP(hipStreamSynchronize(st));
    }
    m->alpha_full.alloc(pool, std::max<size_t>(1, (size_t)M * sizeof(double)));
    if (M)
      FM_HIP(hipMemcpyAsync(m->alpha_full.as<double>(), m->host_alpha.data(), M * sizeof(double),
                            hipMemcpyHostToDevice, st));
    m->has_R = d->R_inv != nullptr && M > 0;
    if (m->has_R) {
      // R_inv^T, the left operand of |R_inv^T k|^2; lower triangular when
      // R_inv is upper triangular (the QR factor's inverse).
      std::vector<double> RT((size_t)M * M);
      int upper = 1;
      for (int r = 0; r < M; r++)
        for (int c = 0; c < M; c++) {
          double v = d->R_inv[(size_t)r * M + c];
          RT[(size_t)c * M + r] = v;
          if (c < r && v != 0.0) upper = 0;
        }
      m->R_lower = upper;
      m->RinvT.alloc(pool, RT.size() * sizeof(double));
      FM_HIP(hipMemcpyAsync(m->RinvT.as<double>(), RT.data(), RT.size() * sizeof(double),
                            hipMemcpyHostToDevice, st));
    }
    FM_HIP(hipStreamSynchronize(st));
    owner.release();
    ctx_acquire(ctx);
    *out = m;
  });
}

int fm_model_destroy(fm_model m) {
  return guarded([&] {
    if (!m) return;
    fm_context ctx = m->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    delete m;
    ctx_release(ctx);
  });
}

int fm_predict(fm_model m, fm_structure s, int flags, int form, double *mean_efs, double *local_var,
               double *local_energy) {
  return guarded([&] {
    FM_CHECK(m && s, FM_ERR_INVALID, "null argument");
    FM_CHECK(m->ctx == s->ctx, FM_ERR_INVALID, "model and structure belong to different contexts");
    FM_HIP(hipSetDevice(s->ctx->device));
    // Stages this prediction does not run (the row-tiled path has no
    // separate gradient or variance stage) report zero, not an older call.
    for (int st = fm::ST_SGEMM; st < fm::ST_COUNT; st++) s->ctx->recorded[st] = false;
    fm::predict(m, s, flags, form, mean_efs, local_var, local_energy);
  });
}

int fm_structure_get_b2_dervs(fm_structure s, int t, double *force_dervs, double *force_dots) {
  return guarded([&] {
    FM_CHECK(s && s->has_b2, FM_ERR_STATE, "descriptors have not been computed");
    FM_CHECK(s->n_struct == 1, FM_ERR_INVALID, "descriptor export is per single structure");
    FM_CHECK(t >= 0 && t < s->nos, FM_ERR_INVALID, "species out of range");
    FM_HIP(hipSetDevice(s->ctx->device));
    fm::compute_dervs(s);
    hipStream_t st = s->ctx->stream;
    size_t f0 = s->nbt_off[t], nb = s->nbt_off[t + 1] - f0;
    if (nb == 0) return;
    if (force_dervs)
      FM_HIP(hipMemcpyAsync(force_dervs, s->DB.as<double>() + 3 * f0 * s->nd,
                            3 * nb * s->nd * sizeof(double), hipMemcpyDeviceToHost, st));
    if (force_dots)
      FM_HIP(hipMemcpyAsync(force_dots, s->fdots.as<double>() + 3 * f0, 3 * nb * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    FM_HIP(hipStreamSynchronize(st));
  });
}

static void dense_check(fm_model m, fm_structure s) {
  FM_CHECK(m && m->b3_ntri == 0, FM_ERR_INVALID,
           "the model's sparse set is in B3 order (it was used with a B3 descriptor)");
  FM_CHECK(m && s, FM_ERR_INVALID, "null argument");
  FM_CHECK(m->ctx == s->ctx, FM_ERR_INVALID, "model and structure belong to different contexts");
  FM_CHECK(s->has_b2, FM_ERR_STATE, "descriptors have not been computed");
  FM_CHECK(m->nos == s->nos && m->nd == s->nd, FM_ERR_INVALID,
           "model and descriptors have different species counts or sizes");
  FM_HIP(hipSetDevice(s->ctx->device));
}

int fm_envs_struc(fm_model m, fm_structure s, double *K) {
  return guarded([&] {
    dense_check(m, s);
    FM_CHECK(K, FM_ERR_INVALID, "null output");
    fm::envs_struc(m, s, K);
  });
}

int fm_self_kernel_struc(fm_model m, fm_structure s, double *kself) {
  return guarded([&] {
    dense_check(m, s);
    FM_CHECK(kself, FM_ERR_INVALID, "null output");
    fm::self_kernel_struc(m, s, kself);
  });
}

int fm_predict_dense(fm_model m, fm_structure s, int kind, double *mean_efs, double *var_efs) {
  return guarded([&] {
    dense_check(m, s);
    FM_CHECK(kind == FM_DENSE_DTC || kind == FM_DENSE_SOR, FM_ERR_INVALID,
             "kind must be FM_DENSE_DTC or FM_DENSE_SOR");
    FM_CHECK(kind != FM_DENSE_DTC || m->L_block_diag, FM_ERR_INVALID,
             "the DTC variance needs an L_inv that is block diagonal over species");
    fm::predict_dense(m, s, kind, mean_efs, var_efs);
  });
}

int fm_model_mapping_coefficients(fm_model m, double *beta, int64_t *beta_size) {
  return guarded([&] {
    FM_CHECK(m && beta_size, FM_ERR_INVALID, "null argument");
    FM_CHECK(m->b3_ntri == 0, FM_ERR_INVALID,
             "mapping coefficients are for B2 models (this sparse set is in B3 order)");
    int d = m->nd;
    FM_CHECK(m->power == 1.0 || m->power == 2.0, FM_ERR_INVALID,
             "Mapping coefficients of the normalized dot product kernel are implemented for "
             "power 1 and 2 only");
    *beta_size = m->power == 1.0 ? d : (int64_t)d * (d + 1) / 2;
    if (!beta) return;
    FM_HIP(hipSetDevice(m->ctx->device));
    std::vector<double> b;
    int64_t bs;
    fm::mapping_coefficients(m, b, bs);
    std::memcpy(beta, b.data(), b.size() * sizeof(double));
  });
}

static void mapped_check(fm_mapped mp) {
  FM_CHECK(mp->power == 1 || mp->power == 2, FM_ERR_INVALID, "Power should be 1 or 2.");
  FM_CHECK(mp->nos >= 1 && mp->N >= 1 && mp->N <= 16 && mp->lmax >= 0 && mp->lmax <= 10,
           FM_ERR_INVALID, "unsupported descriptor settings");
  FM_CHECK(mp->cutkind == FM_CUTOFF_QUADRATIC || mp->cutkind == FM_CUTOFF_COSINE,
           FM_ERR_INVALID, "Please use quadratic or cosine cutoff function.");
  int nr = mp->N * mp->nos;
  mp->nd = nr * (nr + 1) / 2 * (mp->lmax + 1);
}

int fm_mapped_create(fm_context ctx, const fm_mapped_desc *d, fm_mapped *out) {
  return guarded([&] {
    FM_CHECK(ctx && d && out && d->cutoffs && d->beta, FM_ERR_INVALID, "null argument");
    FM_HIP(hipSetDevice(ctx->device));
    auto *mp = new fm_mapped_s();
    mp->ctx = ctx;
    mp->power = d->power;
    mp->normalized = d->normalized ? 1 : 0;
    mp->nos = d->n_species;
    mp->N = d->n_max;
    mp->lmax = d->l_max;
    mp->cutkind = d->cutoff_function;
    try {
      mapped_check(mp);
      mp->cutoffs.assign(d->cutoffs, d->cutoffs + mp->nos * mp->nos);
      fm::mapped_upload(mp, d->beta);
    } catch (...) {
      delete mp;
      throw;
    }
    ctx_acquire(ctx);
    *out = mp;
  });
}

int fm_mapped_read_file(fm_context ctx, const char *path, fm_mapped *out) {
  return guarded([&] {
    FM_CHECK(ctx && path && out, FM_ERR_INVALID, "null argument");
    FM_HIP(hipSetDevice(ctx->device));
    auto *mp = new fm_mapped_s();
    mp->ctx = ctx;
    try {
      std::vector<double> beta;
      fm::mapped_parse(path, mp, beta);
      mapped_check(mp);
      fm::mapped_upload(mp, beta.data());
    } catch (...) {
      delete mp;
      throw;
    }
    ctx_acquire(ctx);
    *out = mp;
  });
}

int fm_mapped_info(fm_mapped mp, int *power, int *normalized, int *n_species, int *n_max,
                   int *l_max, int *cutoff_function, double *cutoffs) {
  return guarded([&] {
    FM_CHECK(mp, FM_ERR_INVALID, "null potential");
    if (power) *power = mp->power;
    if (normalized) *normalized = mp->normalized;
    if (n_species) *n_species = mp->nos;
    if (n_max) *n_max = mp->N;
    if (l_max) *l_max = mp->lmax;
    if (cutoff_function) *cutoff_function = mp->cutkind;
    if (cutoffs) std::memcpy(cutoffs, mp->cutoffs.data(), mp->cutoffs.size() * sizeof(double));
  });
}

int fm_mapped_destroy(fm_mapped mp) {
  return guarded([&] {
    if (!mp) return;
    fm_context ctx = mp->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    delete mp;
    ctx_release(ctx);
  });
}

int fm_mapped_compute(fm_mapped mp, fm_structure s, double *atom_energy, double *forces,
                      double *virial) {
  return guarded([&] {
    FM_CHECK(mp && s, FM_ERR_INVALID, "null argument");
    FM_CHECK(mp->ctx == s->ctx, FM_ERR_INVALID, "potential and structure belong to different contexts");
    FM_HIP(hipSetDevice(s->ctx->device));
    fm::mapped_compute(mp, s, atom_energy, forces, virial);
  });
}

int fm_mapped_compute_noalloc(fm_mapped mp, fm_structure s, double *atom_energy,
                              double *forces, double *virial) {
  return guarded([&] {
    FM_CHECK(mp && s, FM_ERR_INVALID, "null argument");
    FM_CHECK(mp->ctx == s->ctx, FM_ERR_INVALID, "potential and structure belong to different contexts");
    FM_HIP(hipSetDevice(s->ctx->device));
    fm::mapped_compute_noalloc(mp, s, atom_energy, forces, virial);
  });
}

int fm_last_kernel_timings(fm_context ctx, int n, double *ms_out, int32_t *launches) {
  return guarded([&] {
    FM_CHECK(ctx && n >= 0, FM_ERR_INVALID, "null argument");
    FM_HIP(hipStreamSynchronize(ctx->stream));
    FM_HIP(hipStreamSynchronize(ctx->side));
    for (int k = 0; k < n; k++) {
      double total = 0;
      int used = k < fm::KT_COUNT ? ctx->kt[k].used : 0;
      for (int i = 0; i < used; i++) {
        float ms = 0;
        FM_HIP(hipEventElapsedTime(&ms, ctx->kt[k].ev[2 * i], ctx->kt[k].ev[2 * i + 1]));
        total += ms;
      }
      if (ms_out) ms_out[k] = total;
      if (launches) launches[k] = used;
    }
  });
}

int fm_last_timings(fm_context ctx, double *ms_out, int32_t *launches) {
  return guarded([&] {
    FM_CHECK(ctx, FM_ERR_INVALID, "null argument");
    FM_HIP(hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < fm::ST_COUNT; s++) {
      float ms = 0;
      if (ctx->recorded[s]) FM_HIP(hipEventElapsedTime(&ms, ctx->ev_begin[s], ctx->ev_end[s]));
      if (ms_out) ms_out[s] = ms;
      if (launches) launches[s] = ctx->launches[s];
    }
  });
}

}  // extern "C"
