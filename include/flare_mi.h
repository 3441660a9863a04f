/*
 * flare_mi.h — C ABI of libflare_mi, the MI355X (gfx950) implementation of
 * FLARE's sparse-GP force evaluation: periodic neighbour list, B2
 * descriptor, NormalizedDotProduct kernel against the sparse set, and
 * SparseGP mean / local-uncertainty prediction.
 *
 * Plain pointers and sizes only. Every call returns an int status (FM_OK on
 * success); fm_last_error() gives a message for the calling thread. Output
 * buffers are owned by the caller. Handles are not thread-safe: one host
 * thread drives a context at a time.
 *
 * Reference interfaces cited as file:line of the mir-group/flare tree.
 */
#ifndef FLARE_MI_H
#define FLARE_MI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The reference raises std::runtime_error / invalid_argument
 * (sparse_gp.cpp:97-101, normalized_dot_product.cpp:289-292); the ABI
 * returns these codes instead. */
#define FM_OK 0
#define FM_ERR_INVALID 1   /* bad argument or unsupported setting */
#define FM_ERR_HIP 2       /* HIP runtime error */
#define FM_ERR_STATE 3     /* call made in the wrong state (stale model) */
#define FM_ERR_CAPACITY 4  /* request larger than a fixed limit */

/* Arithmetic forms of a prediction. */
#define FM_FORM_FP64 0     /* FP64 everywhere: the drop-in form */
#define FM_FORM_CHOSEN 1   /* FP32 descriptors/adjoint, split-BF16x3 MFMA for
                              the kernel contraction (FP32 accumulate), FP64
                              variance solve */

/* Cutoff functions (cutoffs.cpp:44-73). */
#define FM_CUTOFF_QUADRATIC 0
#define FM_CUTOFF_COSINE 1

/* Prediction flags. */
#define FM_PREDICT_MEAN 1   /* energy, forces, stress (sparse_gp.cpp:643-665) */
#define FM_PREDICT_LOCAL 2  /* local variances (sparse_gp.cpp:93-139,731-757) */

typedef struct fm_context_s *fm_context;
typedef struct fm_structure_s *fm_structure;
typedef struct fm_model_s *fm_model;
typedef struct fm_mapped_s *fm_mapped;

const char *fm_last_error(void);

/* Device context: one HIP stream plus a caching pool of device buffers.
 * No reference counterpart (the reference is a CPU library).
 * fm_context_destroy releases the caller's handle: structures, models and
 * potentials made on the context stay valid, and the context is deleted
 * when the last of them is destroyed (or at once if there are none). */
int fm_context_create(int device, fm_context *out);
int fm_context_destroy(fm_context ctx);
int fm_context_synchronize(fm_context ctx);
/* The hipStream_t the context launches on. */
int fm_context_stream(fm_context ctx, void **stream);

/* Structure with its neighbour list, built on the device.
 * Replaces Structure(cell, species, positions, cutoff, descriptors)
 * (src/flare_pp/structure.cpp:27-43; wrapping :120-133, sweep :135-165,
 * neighbours :52-118). cell is 3x3 row-major with lattice vectors as rows;
 * positions N x 3. The neighbour list equals the reference's bit for bit:
 * order (j, s1, s2, s3), dist < cutoff and dist != 0. */
int fm_structure_create(fm_context ctx, const double *cell, int n_atoms,
                        const int32_t *species, const double *positions,
                        double cutoff, fm_structure *out);
/* As above with species and positions already in device memory. */
int fm_structure_create_device(fm_context ctx, const double *cell, int n_atoms,
                               const int32_t *species_dev,
                               const double *positions_dev, double cutoff,
                               fm_structure *out);
/* A batch of periodic structures in one object (one launch per stage for
 * the whole batch): cells is n_structures x 9, n_atoms per structure,
 * species and positions concatenated in structure order. Every structure
 * has its own periodic images; atoms never see atoms of another structure.
 * Structure objects of one structure and of a batch are used alike:
 * fm_predict writes the mean_efs of each structure one after another and
 * each structure's local variances in its own species-sorted order. No
 * reference counterpart (the reference evaluates one Structure per call). */
int fm_batch_create(fm_context ctx, int n_structures, const double *cells,
                    const int32_t *n_atoms, const int32_t *species,
                    const double *positions, double cutoff, fm_structure *out);
int fm_batch_create_device(fm_context ctx, int n_structures, const double *cells,
                           const int32_t *n_atoms, const int32_t *species_dev,
                           const double *positions_dev, double cutoff,
                           fm_structure *out);
/* Atom offset, sweep and volume of structure k of a batch. */
int fm_batch_info(fm_structure s, int k, int *atom_offset, int *sweep, double *volume);
int fm_structure_destroy(fm_structure s);
/* sweep (structure.cpp:35), total neighbours, volume, single_sweep_cutoff. */
int fm_structure_info(fm_structure s, int *sweep, int64_t *n_neighbors,
                      double *volume, double *single_sweep_cutoff);
/* Copies of Structure::wrapped_positions, neighbor_count,
 * cumulative_neighbor_count, structure_indices, neighbor_species,
 * relative_positions (structure.h:16-45). Any pointer may be NULL. */
int fm_structure_get_neighbors(fm_structure s, double *wrapped, int32_t *count,
                               int32_t *cumulative, int32_t *structure_indices,
                               int32_t *neighbor_species, double *relative);

/* B2 descriptor settings: B2(radial_basis="chebyshev", cutoff_function,
 * radial_hyps=[r1, rc], cutoff_hyps=[], [n_species, N, lmax], cutoffs)
 * (descriptors/b2.cpp:13-50). cutoffs is n_species^2, [central][neighbour]. */
typedef struct {
  int n_species;
  int n_max;
  int l_max;
  int cutoff_function;
  double radial_r1;
  const double *cutoffs;
} fm_b2_settings;

/* B2::compute_struc (descriptors/b2.cpp:76-399) on the device; the
 * descriptors stay resident. form is FM_FORM_FP64 or FM_FORM_CHOSEN (FP32). */
int fm_structure_compute_b2(fm_structure s, const fm_b2_settings *settings,
                            int form);
/* B3::compute_struc (descriptors/b3.cpp:80-443) with the same settings
 * (l_max <= 4, the reference's Wigner 3j table); the descriptors stay resident
 * and the prediction entry points accept them as they accept B2. */
int fm_structure_compute_b3(fm_structure s, const fm_b2_settings *settings,
                            int form);
/* DescriptorValues sizes (descriptor.h:31-57): n_descriptors, and per species
 * n_clusters_by_type and n_neighbors_by_type (arrays of n_species). */
int fm_structure_b2_sizes(fm_structure s, int *n_descriptors,
                          int32_t *n_clusters_by_type,
                          int32_t *n_neighbors_by_type);
/* DescriptorValues fields of one species (descriptor.h:37-43, b2.cpp:148-176):
 * descriptors (n_s x d), descriptor_norms, neighbor_counts,
 * cumulative_neighbor_counts, atom_indices, neighbor_indices,
 * neighbor_coordinates (nb_s x 3). Any pointer may be NULL. */
int fm_structure_get_b2(fm_structure s, int species, double *descriptors,
                        double *norms, int32_t *neighbor_counts,
                        int32_t *cumulative_neighbor_counts,
                        int32_t *atom_indices, int32_t *neighbor_indices,
                        double *neighbor_coordinates);

/* Sparse-GP prediction state for one kernel (one descriptor):
 * NormalizedDotProduct(sigma, power) (kernels/normalized_dot_product.cpp:15-26),
 * its ClusterDescriptor sparse set (descriptor.h:61-90: raw rows and norms
 * per species, species-major), the kernel's slice of alpha and the diagonal
 * block of L_inv (sparse_gp.cpp:607-641; M x M row-major, lower triangular;
 * NULL disables local variances). */
typedef struct {
  double sigma;
  double power;
  int n_species;
  int n_descriptors;
  const int32_t *n_sparse_by_type;
  const double *sparse_descriptors; /* M x d, species-major */
  const double *sparse_norms;       /* M */
  const double *alpha;              /* M */
  const double *L_inv;              /* M x M or NULL */
  const double *R_inv;              /* M x M or NULL (DTC / SOR variances) */
} fm_model_desc;

int fm_model_create(fm_context ctx, const fm_model_desc *desc, fm_model *out);
int fm_model_destroy(fm_model m);

/* SparseGP::predict_mean / predict_local_uncertainties for one kernel
 * (sparse_gp.cpp:643-665, 731-757; kernel columns of
 * normalized_dot_product.cpp:117-214 contracted with alpha by the adjoint,
 * so dB/dr is never formed). Outputs, any may be NULL:
 *   mean_efs     1 + 3N + 6: [E, F (atom-major xyz), S (xx,xy,xz,yy,yz,zz)]
 *                (for a batch, one such block per structure, concatenated)
 *   local_var    N, species-sorted order (compute_cluster_uncertainties),
 *                per structure for a batch
 *   local_energy N, atom order (this kernel's per-atom energies). */
int fm_predict(fm_model m, fm_structure s, int flags, int form,
               double *mean_efs, double *local_var, double *local_energy);

/* Dense descriptor derivatives of one species (DescriptorValues
 * descriptor_force_dervs and descriptor_force_dots, descriptor.h:37-43,
 * b2.cpp:181-242): dB_i/dr_bond for every bond of the species' atoms in the
 * order of fm_structure_get_b2's neighbour lists, rows 3 bond + component
 * (3 nb_s x d), and their dots with B_i (3 nb_s). FP64 descriptors of a
 * single structure only; built on the first call and kept with the
 * structure. Either pointer may be NULL. */
int fm_structure_get_b2_dervs(fm_structure s, int species, double *force_dervs,
                              double *force_dots);

/* NormalizedDotProduct::envs_struc (normalized_dot_product.cpp:117-214):
 * the M x (1 + 3N + 6) kernel matrix between the sparse set and the
 * energy, forces and stress of a single structure (FP64 descriptors). */
int fm_envs_struc(fm_model m, fm_structure s, double *K);
/* NormalizedDotProduct::self_kernel_struc (normalized_dot_product.cpp:
 * 450-586): the 1 + 3N + 6 diagonal of the structure's own kernel. */
int fm_self_kernel_struc(fm_model m, fm_structure s, double *kself);

/* SparseGP::predict_DTC / predict_SOR (sparse_gp.cpp:667-729) for a model
 * with one kernel: mean_efs = K^T alpha and var_efs (1 + 3N + 6) =
 * K_self - diag(K^T Kuu^-1 K) + diag(K^T Sigma K) (DTC) or
 * diag(K^T R_inv R_inv^T K) (SOR), with Kuu^-1 = L_inv^T L_inv and
 * Sigma = R_inv R_inv^T. Needs R_inv in the model (and L_inv for DTC). */
#define FM_DENSE_DTC 1
#define FM_DENSE_SOR 2
int fm_predict_dense(fm_model m, fm_structure s, int kind, double *mean_efs,
                     double *var_efs);

/* Mapping coefficients of a power-1 or power-2 model
 * (NormalizedDotProduct::compute_mapping_coefficients,
 * normalized_dot_product.cpp:631-767): per species beta = sum_j sigma^2
 * alpha_j x_hat_j (power 1, beta_size = d) or the upper triangle of
 * sum_j sigma^2 alpha_j x_hat_j x_hat_j^T with off-diagonal entries doubled
 * (power 2, beta_size = d(d+1)/2). beta is n_species x beta_size; pass NULL
 * to query beta_size. */
int fm_model_mapping_coefficients(fm_model m, double *beta, int64_t *beta_size);

/* The pair_style flare potential (lammps_plugins/pair_flare.cpp), from a
 * coefficient file (PairFLARE::read_file, pair_flare.cpp:262-419) or from
 * arrays in the same layout: power 1 or 2, NormalizedDotProduct
 * (normalized = 1) or DotProduct (0), chebyshev radial basis, cutoff matrix
 * [central][neighbour], beta n_species x beta_size. */
typedef struct {
  int power;
  int normalized;
  int n_species;
  int n_max;
  int l_max;
  int cutoff_function;
  const double *cutoffs;
  const double *beta;
} fm_mapped_desc;

int fm_mapped_create(fm_context ctx, const fm_mapped_desc *desc, fm_mapped *out);
int fm_mapped_read_file(fm_context ctx, const char *path, fm_mapped *out);
/* Settings of a potential; cutoffs receives n_species^2 values (may be NULL). */
int fm_mapped_info(fm_mapped mp, int *power, int *normalized, int *n_species,
                   int *n_max, int *l_max, int *cutoff_function, double *cutoffs);
int fm_mapped_destroy(fm_mapped mp);
/* PairFLARE::compute (pair_flare.cpp:68-184) on a periodic structure whose
 * FP64 B2 descriptors were computed with the potential's settings
 * (radial_r1 = 0): per-atom energies (atom order), forces N x 3, and for
 * each structure the virial sum_bonds -f (x) r (xx, xy, xz, yy, yz, zz).
 * Any output may be NULL. */
int fm_mapped_compute(fm_mapped mp, fm_structure s, double *atom_energy,
                      double *forces, double *virial);
/* fm_mapped_compute without any allocation: no reference counterpart; it is
 * for callers that cannot allocate device memory (an MD loop that holds its
 * structures, code run where hipMalloc is forbidden). Results equal
 * fm_mapped_compute's bit for bit. The working buffers come from one static
 * device array of FM_NOALLOC_SCRATCH_BYTES per GPU (64 MiB, held until the
 * process ends); a structure that needs more returns FM_ERR_CAPACITY.
 * Outputs may be host or device memory; the call returns when they are
 * written. Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU (from several threads or contexts) race. Kernel timers are not
 * recorded by this call. */
#define FM_NOALLOC_SCRATCH_BYTES (64ull << 20)
int fm_mapped_compute_noalloc(fm_mapped mp, fm_structure s, double *atom_energy,
                              double *forces, double *virial);

/* Device-side durations (HIP events on the context stream) of the stages
 * last run on a context, in milliseconds, and the number of kernel launches
 * in each: [0] neighbour list, [1] descriptors (bond list + B2), [2] kernel
 * contraction S = B X_hat^T with its epilogue, [3] gradient contraction
 * G = W X_hat, [4] variance contraction, [5] adjoint + force gather.
 * Either pointer may be NULL. No reference counterpart. */
#define FM_N_STAGES 6
int fm_last_timings(fm_context ctx, double *ms, int32_t *launches);

/* Device-side durations of individual kernels in the last calls on a
 * context (HIP event pairs around each launch, on the stream it runs on):
 * the sum over launches in milliseconds and the number of launches, for
 * n of the kernels [0] B2 descriptor (bond basis, c, B; last descriptor
 * computation), [1] int8 slicing of descriptor rows (chosen form), [2]
 * kernel contraction S with its epilogue, [3] gradient contraction G, [4]
 * variance contraction, [5] adjoint u = dE/dc, [6] bond forces, [7] force
 * gather ([1]-[7]: last prediction). No reference counterpart. */
#define FM_N_KERNEL_TIMERS 8
int fm_last_kernel_timings(fm_context ctx, int n, double *ms, int32_t *launches);

/* Row tiles for B3 structures whose descriptors are too large to keep
 * (B3 at 10^6 atoms is 296 GB in FP32): for B3 descriptors computed on this
 * context afterwards, rows > 0 forces tiles of that many species-sorted rows;
 * 0 (the default) tiles only when the descriptor rows would exceed 32 GB,
 * with tiles of about 8 GB of working buffers. A tiled structure keeps its
 * coefficients c; its B3 rows are formed tile by tile inside fm_predict
 * (results equal the untiled ones bit for bit) and cannot be read back with
 * fm_structure_get_b2. No reference counterpart (b3.cpp keeps every row). */
int fm_context_set_row_tile(fm_context ctx, int64_t rows);

/* Serving: the batched and serving entry points of the drop-in boundary.
 * No reference counterpart (the reference evaluates one Structure per call,
 * sparse_gp.cpp:643-757; this serves the same predictions to many callers).
 * A server keeps every kernel's model resident on n_workers contexts (HIP
 * streams, one worker thread each) and batches waiting requests
 * continuously: a worker takes the oldest requests whose device footprint
 * fits hbm_budget beside the batches in flight on all workers, evaluates
 * them as one batch (descriptor calculator k feeds kernel k, as
 * Structure::descriptors and SparseGP::kernels pair up) and completes them.
 * A request is sized from its atom count and the neighbours within cutoff
 * at max_density atoms / A^3 (or its own density if higher); once its batch
 * is built the measured neighbour count replaces the estimate. Results do
 * not depend on the other requests of a batch. fm_submit and fm_poll may be
 * called from any host threads. */
#define FM_DESCRIPTOR_B2 2
#define FM_DESCRIPTOR_B3 3
typedef struct fm_server_s *fm_server;
typedef struct {
  int n_kernels;
  const fm_model_desc *models;          /* one per kernel */
  const fm_b2_settings *descriptors;    /* one per kernel */
  const int32_t *descriptor_kinds;      /* FM_DESCRIPTOR_B2 / _B3 per kernel; NULL: B2 */
  double cutoff;                        /* neighbour-list cutoff */
  int form;                             /* FM_FORM_FP64 or FM_FORM_CHOSEN */
  int flags;                            /* FM_PREDICT_MEAN [| FM_PREDICT_LOCAL] */
  int n_workers;                        /* 1 .. 16 */
  int64_t hbm_budget;                   /* device bytes of the batches in flight */
  int32_t max_batch_atoms;              /* atoms of one batch at most; 0: no cap */
  double max_density;                   /* atoms / A^3 for sizing requests; 0: 0.1 */
} fm_server_desc;
int fm_server_create(int device, const fm_server_desc *desc, fm_server *out);
/* Finishes the queued requests, then stops the workers and frees the server
 * (results not yet polled are discarded). */
int fm_server_destroy(fm_server srv);
/* Queues one periodic structure (inputs are copied; cell 3 x 3 row-major,
 * positions N x 3). FM_ERR_CAPACITY when the request alone exceeds the
 * budget. */
int fm_submit(fm_server srv, const double *cell, int n_atoms, const int32_t *species,
              const double *positions, int64_t *request);
/* Waits up to timeout_ms (0: returns at once, < 0: until done) for a
 * request. *done = 1 once it has finished: its outputs are copied (mean_efs
 * 1 + 3N + 6; local_var n_kernels x N, each kernel's variances in the
 * structure's species-sorted order; either may be NULL), the request is
 * forgotten, and the return value is FM_OK or the error the evaluation
 * raised. *done = 0: still queued or running, FM_OK. */
int fm_poll(fm_server srv, int64_t request, int timeout_ms, int *done, double *mean_efs,
            double *local_var);

#ifdef __cplusplus
}
#endif
#endif /* FLARE_MI_H */
