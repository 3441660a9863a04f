// Host-side internals of libflare_mi: error handling, the per-context device
// buffer pool, and the structure / model objects behind the C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <atomic>
#include <vector>

#include "../../include/flare_mi.h"

namespace fm {

struct B3State;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define FM_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::fm::Error(FM_ERR_HIP, std::string(#call) + ": " +              \
                                        hipGetErrorString(e_));               \
  } while (0)

// Sets the calling thread's fm_last_error() message (entry points that do
// not go through the C ABI's error wrapper).
void set_last_error(const std::string &msg);

#define FM_CHECK(cond, code, msg)                     \
  do {                                                \
    if (!(cond)) throw ::fm::Error((code), (msg));    \
  } while (0)

// Size-class caching pool over hipMalloc. Buffers return to the pool when
// the owning object dies; the stream order of the single context stream
// keeps reuse safe.
class Pool {
 public:
  ~Pool();
  void *get(size_t bytes, size_t *cls);
  void put(void *p, size_t cls);
  size_t bytes_cached() const { return cached_; }

 private:
  std::map<size_t, std::vector<void *>> free_;
  size_t cached_ = 0;
};

class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(Pool *pool, size_t bytes) { alloc(pool, bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept { *this = std::move(o); }
  DevBuf &operator=(DevBuf &&o) noexcept {
    release();
    p_ = o.p_;
    cls_ = o.cls_;
    pool_ = o.pool_;
    bytes_ = o.bytes_;
    o.p_ = nullptr;
    return *this;
  }
  void alloc(Pool *pool, size_t bytes) {
    release();
    pool_ = pool;
    bytes_ = bytes;
    p_ = pool->get(bytes ? bytes : 1, &cls_);
  }
  void ensure(Pool *pool, size_t bytes) {
    if (!p_ || bytes_ < bytes || cls_ < bytes) alloc(pool, bytes);
    bytes_ = bytes;
  }
  void release() {
    if (p_ && pool_) pool_->put(p_, cls_);
    p_ = nullptr;
  }
  template <typename T>
  T *as() const {
    return reinterpret_cast<T *>(p_);
  }
  size_t bytes() const { return bytes_; }
  bool empty() const { return p_ == nullptr; }

 private:
  void *p_ = nullptr;
  size_t cls_ = 0, bytes_ = 0;
  Pool *pool_ = nullptr;
};

// Geometry of one periodic structure for the neighbour kernels: cell (rows
// are lattice vectors), its transpose, (cell^T)^-1, cell cell^T and its
// inverse, the cutoff and sweep, and the cell-list bins per axis (nb) with
// the stencil reach (K).
struct CellArgs {
  double cell[9], ct[9], cti[9], cd[9], cdi[9];
  double cutoff;
  int sweep;
  int nb[3], K[3];
};

enum Stage { ST_NEIGH = 0, ST_DESC, ST_SGEMM, ST_GGEMM, ST_VAR, ST_ADJ, ST_COUNT };
// Per-kernel timers (fm_last_kernel_timings): each launch of these kernels
// is bracketed by an event pair on the stream it runs on.
enum KernelTimer {
  KT_DESC = 0,  // B2 descriptor kernel (bond basis, c, B)
  KT_SLICE,     // int8 slicing of descriptor rows (chosen form)
  KT_SGEMM,     // kernel contraction S = B X_hat^T with its epilogue
  KT_GGEMM,     // gradient contraction G = W X_hat
  KT_VAR,       // variance contraction Q = L_s^-1 K^T, sum of squares
  KT_DEDC,      // adjoint u = dE/dc
  KT_BOND,      // per-bond forces and virials
  KT_GATHER,    // per-atom force gather and reductions
  KT_COUNT
};

}  // namespace fm

struct fm_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  fm::Pool pool;
  // Pinned host words for small device-to-host readbacks (sizes, flags), so
  // several can be queued and waited for with one stream synchronisation.
  // Calls on one context are sequential (they share its stream).
  static constexpr int kPinnedWords = 64;
  int *pinned = nullptr;
  // Second stream (default priority) for work that overlaps the main one:
  // the variance contraction runs there while the gradient contraction and
  // the adjoint run on `stream`. fork / join order them with events.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // Row tile for B3 descriptors computed on this context (0: automatic).
  int64_t row_tile = 0;
  hipEvent_t ev_begin[fm::ST_COUNT], ev_end[fm::ST_COUNT];
  bool recorded[fm::ST_COUNT] = {};
  int launches[fm::ST_COUNT] = {};
  void mark_begin(int s, hipStream_t on = nullptr);
  void mark_end(int s, hipStream_t on = nullptr);
  // Kernel timers: event pairs of the launches since the last kt_clear.
  struct KTimes {
    std::vector<hipEvent_t> ev;
    int used = 0;
  };
  KTimes kt[fm::KT_COUNT];
  // Off while a call that may not create events runs (kt_begin / kt_end do
  // nothing then).
  bool kt_on = true;
  void kt_clear(int first, int last);
  void kt_begin(int k, hipStream_t on = nullptr);
  void kt_end(int k, hipStream_t on = nullptr);
  // Lifetime: 2 x (structures, models and potentials made on the context)
  // + 1 once fm_context_destroy was called. The context is deleted by
  // whichever comes last: the destroy call with no objects left, or the
  // destruction of the last object after the destroy call. Objects return
  // their buffers to the context's pool, so the order the caller releases
  // them in does not matter.
  std::atomic<int64_t> life{0};
};

namespace fm {
// Brackets the launches made in its scope with a kernel timer.
struct KScope {
  fm_context c;
  int k;
  hipStream_t s;
  KScope(fm_context c_, int k_, hipStream_t s_ = nullptr) : c(c_), k(k_), s(s_) {
    c->kt_begin(k, s);
  }
  ~KScope() { c->kt_end(k, s); }
};
}  // namespace fm

struct fm_structure_s {
  fm_context ctx = nullptr;
  int n = 0;
  // B3 only: > 0 when the descriptor rows are not kept but formed in tiles
  // of this many species-sorted rows inside fm_predict.
  int64_t row_tile = 0;
  // A structure object holds a batch of n_struct periodic structures; atoms
  // of structure k are [atom_off[k], atom_off[k+1]). Single structures have
  // n_struct = 1 and the fields below (cell ... sweep) describe structure 0.
  int n_struct = 1;
  int max_species = -1;
  std::vector<int> atom_off;
  std::vector<fm::CellArgs> cellargs;
  std::vector<double> volumes, sweep_cutoffs;
  fm::DevBuf d_cellargs, atom_struct, d_bin_off, d_atom_off, d_mean_off, d_vol_inv;
  // Output offsets of each structure in mean_efs (1 + 3 N_k + 6 per structure)
  // and the reduction chunks (<= 256 atoms, never crossing a structure).
  std::vector<int64_t> mean_off;
  int n_chunks = 0;
  fm::DevBuf d_chunk_begin, d_chunk_end, d_struct_chunk_off;
  // Position of each species-sorted row in the per-structure species-sorted
  // order of the local variances (device ints; empty for a single
  // structure, whose order is the identity).
  fm::DevBuf var_pos;
  double cell[9], cell_t_inv[9], cell_dot[9], cell_dot_inv[9];
  double volume = 0, single_sweep_cutoff = 0, cutoff = 0;
  int sweep = 0;
  int64_t n_nb = 0;
  fm::DevBuf species, positions, wrapped, count, cumulative, sidx, nspecies, rel;
  // Descriptor state (B2, or B3 built from the same bond coefficients)
  bool has_b2 = false;
  int desc_kind = 2;           // 2: B2, 3: B3
  fm::B3State *b3 = nullptr;   // B3 index tables (fm_b3.hip)
  int form = FM_FORM_FP64;
  int nos = 0, N = 0, lmax = 0, cutkind = 0, nd = 0;
  double r1 = 0;
  std::vector<double> cutoffs;
  int64_t n_bonds = 0;
  std::vector<int> n_by_type, nb_by_type, type_off;
  fm::DevBuf d_cutoffs, bond_off, bond_q, in_off, in_bond, row_of_atom,
      atom_of_row, pair_table, B, norm, nn;
  // Density coefficients c (atom-major, n x nrad H, descriptor precision),
  // kept for the adjoint.
  fm::DevBuf C;
  // Chosen form, B2: the FP64 descriptor rows as int8 slices with a
  // power-of-two scale per species-sorted row, written by the descriptor
  // pass; the kernel contraction's operand. The slices are in the tiled
  // layout (fm_gemm.hpp) with ldq = the padded reduction length and each
  // species starting a row tile: species t's rows are slice rows
  // q_tile_row[t] + (row - type_off[t]) (q_shift[t] = q_tile_row[t] -
  // type_off[t] on the device).
  fm::DevBuf Bq, Bq_scale, q_shift;
  std::vector<long long> q_tile_row;
  std::vector<int> q_shift_host;
  long long ldq = 0;
  bool has_q = false;
  // Dense descriptor derivatives (FP64, single structure): bonds in the
  // species-sorted order of DescriptorValues, f in [sb_off[row], sb_off[row+1])
  // for species-sorted row `row`; bonds of species t are [nbt_off[t],
  // nbt_off[t+1]). DB is 3 n_bonds x d (rows 3 f + component), fdots 3 n_bonds.
  bool has_dervs = false;
  std::vector<int> nbt_off;
  fm::DevBuf sb_off, sb_pos, sb_row, sb_bond, DB, fdots;
};

struct fm_model_s {
  fm_context ctx = nullptr;
  double sigma = 1, power = 2;
  int nos = 0, nd = 0, M = 0;
  std::vector<int> M_by_type, off;
  bool has_L = false;
  std::vector<int> L_rows;   // rows of the L_inv column block of each species
  std::vector<int> L_lower;  // 1 when that block is lower triangular
  // Per species: X-hat (M_s x d) and X-hat^T (d x M_s) in FP64 and as two
  // BF16 pieces of their FP32 values (Xs: M_s x split_ld(d), for the kernel
  // contraction; XTs: d x split_ld(M_s), for the gradient contraction; zero
  // padded), alpha (zero for empty sparse environments) and the L_inv block
  // (L_rows x M_s, starting at the species' first row).
  std::vector<fm::DevBuf> X, XT, Xs0, Xs1, XTs0, XTs1, alpha, L;
  // Chosen form: X-hat rows as int8 slices with their power-of-two scales
  // (the kernel contraction's operand), made from X on first use and again
  // after the B3 reordering of X.
  std::vector<fm::DevBuf> Xq, Xq_scale;
  bool xq_ready = false;
  std::vector<double> host_alpha;
  // Non-zero once the sparse set has been permuted to the device order of
  // B3 descriptors (l-set major) for (n_tri, n_ls) = (b3_ntri, b3_nls).
  int b3_ntri = 0, b3_nls = 0;
  // Dense (DTC / SOR) predictions: sparse norms per species, alpha in one
  // vector, R_inv^T (M x M) and whether it is lower triangular, and whether
  // L_inv is block diagonal over species.
  std::vector<std::vector<double>> xnorm;
  fm::DevBuf alpha_full, RinvT;
  bool has_R = false;
  int R_lower = 0;
  bool L_block_diag = true;
};

struct fm_mapped_s {
  fm_context ctx = nullptr;
  int power = 2, normalized = 1, nos = 0, N = 0, lmax = 0, cutkind = 0, nd = 0;
  std::vector<double> cutoffs;       // nos x nos
  std::vector<fm::DevBuf> beta;      // per species: d x d (power 2) or d
};

namespace fm {
// fm_mapped.hip
void mapping_coefficients(fm_model m, std::vector<double> &beta, int64_t &beta_size);
void mapped_compute(fm_mapped mp, fm_structure s, double *atom_energy, double *forces,
                    double *virial);
// The same evaluation with every working buffer carved from the library's
// static device scratch (kNoAllocScratchBytes per GPU); no allocation.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
void mapped_compute_noalloc(fm_mapped mp, fm_structure s, double *atom_energy, double *forces,
                            double *virial);
// Looks up the scratch array's device address on the current device (once
// per device; called when a context is created).
void scratch_setup(int device);
void mapped_upload(fm_mapped mp, const double *beta);
void mapped_parse(const std::string &path, fm_mapped mp, std::vector<double> &beta);
// Launchers (fm_neighbors.hip)
void build_neighbors(fm_structure s);
void build_bonds(fm_structure s);
// fm_descriptors.hip (want_b2 = false computes and keeps only c)
void launch_b2(fm_structure s, bool want_b2 = true);
// Device var_pos of a batch (after build_bonds); releases it for one structure.
void build_var_pos(fm_structure s);
// fm_b3.hip
int b3_size(int nos, int N, int lmax);
void launch_b3(fm_structure s);
// B3 rows of one call: all atoms (tile false: atom order, each written at its
// species-sorted row of Bout) or the species-sorted rows [r0, r0 + n) (tile
// true: written as rows 0.. n - 1 of Bout, a row tile).
void launch_b3_rows(fm_structure s, int r0, int n, bool tile, void *Bout);
// The model's sparse set in the device order of the structure's B3 rows
// (permuted once; an error if it is already in another order).
void b3_order_model(fm_model m, fm_structure s);
// Rows [r0, r0 + rows) of a B3 structure's descriptors in the reference's
// order into `out` (device memory, rows x nd, the descriptor precision).
void b3_rows_to_reference(fm_structure s, int r0, int rows, void *out);
void launch_dEdc_b3_rows(fm_structure s, int r0, int n, bool tile, const void *B,
                         const void *G, int64_t g_ld, int project, double empty_norm,
                         double empty_nn, void *U);
void launch_dEdc_b3(fm_structure s, const void *G, int64_t g_ld, int project,
                    double empty_norm, double empty_nn, void *U);
void free_b3_state(fm_structure s);
// Bond forces and per-atom virials from u = dE/dc (n x n_rad H, atom
// order) for all atoms (atom_list null, n_list = n) or the listed atoms.
void launch_bond_forces(fm_structure s, const void *U, void *fbond, double *virial_atom,
                        const int *atom_list, int n_list, double empty_norm, double empty_nn);
// With project, G . B comes from `energy` (per species-sorted row; the
// rows' local energies of a normalized dot product kernel of this power)
// when given, else from a pass over the G and B rows.
void launch_adjoint(fm_structure s, const void *G, int64_t g_ld, void *fbond,
                    double *virial_atom, int project, double empty_norm, double empty_nn,
                    const double *energy = nullptr, double power = 0, void *u_buf = nullptr);
// Bytes of the u = dE/dc buffer launch_adjoint needs (u_buf).
size_t adjoint_u_bytes(fm_structure s);
// fm_predict.cpp
void predict(fm_model m, fm_structure s, int flags, int form, double *mean_efs,
             double *local_var, double *local_energy);
// fm_dtc.hip
void compute_dervs(fm_structure s);
void envs_struc(fm_model m, fm_structure s, double *K_out);
void self_kernel_struc(fm_model m, fm_structure s, double *kself_out);
void predict_dense(fm_model m, fm_structure s, int kind, double *mean_efs, double *var_efs);
// fm_descriptors.hip: normalisation constants of the real spherical harmonics
void knorm_values(int lmax, double *out);
// scans (fm_neighbors.hip)
int scan_exclusive(fm_context ctx, const int *in, int *out, int n, bool want_total);
}  // namespace fm
