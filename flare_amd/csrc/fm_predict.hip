// Prediction pipeline on the device: per species, the kernel contraction
// (S = B_s X_hat^T with its fused epilogue), the gradient contraction
// G = W X_hat, the optional variance contraction Q = L_s^-1 K^T, then the
// adjoint bond forces and a deterministic gather of forces, energy and
// stress (sparse_gp.cpp:643-665 and 731-757; output layout of
// normalized_dot_product.cpp:117-214).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fm_device.hpp"
#include "fm_gemm.hpp"
#include "fm_internal.hpp"

namespace fm {

// Per row (species-sorted atom): E = sum of the energy partials over column
// tiles; var = sigma^2 (B.B / |B|^2)^p - sum of the |Q|^2 partials over row
// tiles (the diagonal of envs_envs(cluster, cluster) minus diag(Q1^T Q1),
// sparse_gp.cpp:116-136). Partials are summed in tile order.
// Rows [ro, ro + n_rows); partial t of row r at t * part_ld + r.
__device__ __forceinline__ double row_variance(int r, const double *Qpart, int n_qtiles,
                                               long long part_ld, double nrm, double nnr,
                                               double sig2, double power) {
  double v = 0;
  if (nrm >= kEmptyThresh) {
    double nd = nnr / (nrm * nrm);
    double ks = sig2 * (power == 2.0 ? nd * nd : pow(nd, power));
    double q = 0;
    for (int t = 0; t < n_qtiles; t++) q += Qpart[(size_t)t * part_ld + r];
    v = ks - q;
  }
  return v;
}

__global__ void k_species_final(int n_rows, int ro, const double *Epart, int n_etiles,
                                const double *Qpart, int n_qtiles, long long part_ld,
                                const double *norm, const double *nn, double sig2, double power,
                                double *E_row, double *var_row) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  double e = 0;
  for (int t = 0; t < n_etiles; t++) e += Epart[(size_t)t * part_ld + r];
  E_row[ro + r] = e;
  if (var_row)
    var_row[ro + r] =
        row_variance(r, Qpart, n_qtiles, part_ld, norm[ro + r], nn[ro + r], sig2, power);
}

// The variance half alone (run on the side stream).
__global__ void k_var_final(int n_rows, int ro, const double *Qpart, int n_qtiles,
                            long long part_ld, const double *norm, const double *nn, double sig2,
                            double power, double *var_row) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  var_row[ro + r] =
      row_variance(r, Qpart, n_qtiles, part_ld, norm[ro + r], nn[ro + r], sig2, power);
}

// Forces: F_i = sum_{bonds b of i} f_b - sum_{bonds b with neighbour i} f_b
// (normalized_dot_product.cpp:197-199 after K^T alpha), each list in fixed
// order, written at the atom's slot of its structure's mean_efs. One block
// per chunk of <= 256 atoms of one structure; the block also reduces the
// chunk's energies and virials.
template <typename TF>
__global__ void k_gather(const int *chunk_begin, const int *chunk_end, const int *atom_struct,
                         const int *atom_off, const long long *mean_off, const int *bond_off,
                         const int *in_off, const int *in_bond, const TF *fbond,
                         const int *row_of_atom, const double *E_row, const double *vir,
                         double *out, double *blk_part) {
  __shared__ double sh[7][256];
  int c = blockIdx.x;
  int i = chunk_begin[c] + threadIdx.x;
  double e = 0, v[6] = {0, 0, 0, 0, 0, 0};
  if (i < chunk_end[c]) {
    double f[3] = {0, 0, 0};
    for (int b = bond_off[i]; b < bond_off[i + 1]; b++)
      for (int q = 0; q < 3; q++) f[q] += (double)fbond[3 * (size_t)b + q];
    for (int a = in_off[i]; a < in_off[i + 1]; a++) {
      int b = in_bond[a];
      for (int q = 0; q < 3; q++) f[q] -= (double)fbond[3 * (size_t)b + q];
    }
    int k = atom_struct[i];
    double *F = out + mean_off[k] + 1 + 3 * (size_t)(i - atom_off[k]);
    for (int q = 0; q < 3; q++) F[q] = f[q];
    e = E_row[row_of_atom[i]];
    for (int q = 0; q < 6; q++) v[q] = vir[6 * (size_t)i + q];
  }
  sh[0][threadIdx.x] = e;
  for (int q = 0; q < 6; q++) sh[1 + q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 7; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 7) blk_part[(size_t)c * 7 + threadIdx.x] = sh[threadIdx.x][0];
}

// Per structure, the sum of its chunk partials in chunk order: energy and
// stress -1/V sum_b f (x) r (normalized_dot_product.cpp:201-206).
__global__ void k_final_sum(int n_struct, const int *struct_chunk_off, const double *blk_part,
                            const double *vol_inv, const long long *mean_off,
                            const int *atom_off, double *out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_struct) return;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int c = struct_chunk_off[k]; c < struct_chunk_off[k + 1]; c++)
    for (int q = 0; q < 7; q++) acc[q] += blk_part[(size_t)c * 7 + q];
  double *o = out + mean_off[k];
  int nk = atom_off[k + 1] - atom_off[k];
  o[0] = acc[0];
  for (int q = 0; q < 6; q++) o[1 + 3 * nk + q] = acc[1 + q] * vol_inv[k];
}

__global__ void k_fill(double *p, long long n, double v) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Per-atom forces (own bonds minus incoming bonds), energies and virials
// into `o` (one mean_efs block per structure).
static void gather_forces(fm_structure s, bool f64, DevBuf &fbond, DevBuf &vir, DevBuf &E_row,
                          double *o) {
  hipStream_t st = s->ctx->stream;
  Pool *pool = &s->ctx->pool;
  DevBuf blk(pool, (size_t)s->n_chunks * 7 * sizeof(double));
  KScope timer(s->ctx, KT_GATHER);
  if (f64)
    hipLaunchKernelGGL(k_gather<double>, dim3(s->n_chunks), dim3(256), 0, st,
                       s->d_chunk_begin.as<int>(), s->d_chunk_end.as<int>(),
                       s->atom_struct.as<int>(), s->d_atom_off.as<int>(),
                       s->d_mean_off.as<long long>(), s->bond_off.as<int>(), s->in_off.as<int>(),
                       s->in_bond.as<int>(), fbond.as<double>(), s->row_of_atom.as<int>(),
                       E_row.as<double>(), vir.as<double>(), o, blk.as<double>());
  else
    hipLaunchKernelGGL(k_gather<float>, dim3(s->n_chunks), dim3(256), 0, st,
                       s->d_chunk_begin.as<int>(), s->d_chunk_end.as<int>(),
                       s->atom_struct.as<int>(), s->d_atom_off.as<int>(),
                       s->d_mean_off.as<long long>(), s->bond_off.as<int>(), s->in_off.as<int>(),
                       s->in_bond.as<int>(), fbond.as<float>(), s->row_of_atom.as<int>(),
                       E_row.as<double>(), vir.as<double>(), o, blk.as<double>());
  hipLaunchKernelGGL(k_final_sum, dim3((s->n_struct + 63) / 64), dim3(64), 0, st, s->n_struct,
                     s->d_struct_chunk_off.as<int>(), blk.as<double>(), s->d_vol_inv.as<double>(),
                     s->d_mean_off.as<long long>(), s->d_atom_off.as<int>(), o);
}

// out[pos[r]] = in[r]: species-sorted rows to another order on the device.
__global__ void k_scatter_rows(int n, const int *pos, const double *in, double *out) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) out[pos[r]] = in[r];
}

// Atoms of each (species, structure) of a batch, species-major. A wavefront
// whose atoms all lie in one structure (atoms of a structure are
// consecutive) adds one count per species from lane 0.
__global__ void k_struct_species_count(int n, const int *species, const int *atom_struct,
                                       int n_struct, int nos, int *cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = i < n;
  int k = ok ? atom_struct[i] : -1, t = ok ? species[i] : -1;
  int k0 = __shfl(k, 0, 64);
  if (nos <= 8 && __all(!ok || k == k0)) {
    for (int u = 0; u < nos; u++) {
      unsigned long long m = __ballot(ok && t == u);
      if (m != 0 && (threadIdx.x & 63) == 0)
        atomicAdd(&cnt[(size_t)u * n_struct + k0], (int)__popcll(m));
    }
  } else if (ok) {
    atomicAdd(&cnt[(size_t)t * n_struct + k], 1);
  }
}

// Position of species-sorted row r in its structure's species-sorted order.
// Rows of one species keep atom order, so structure k's atoms of species t
// are the rows [ex[t][k], ex[t][k] + cnt[t][k]), ex the exclusive scan of
// the species-major counts.
__global__ void k_var_pos(int n, const int *atom_of_row, const int *species,
                          const int *atom_struct, const int *atom_off, const int *cnt,
                          const int *ex, int n_struct, int *var_pos) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int i = atom_of_row[r], t = species[i], k = atom_struct[i];
  int before = 0;
  for (int u = 0; u < t; u++) before += cnt[(size_t)u * n_struct + k];
  var_pos[r] = atom_off[k] + before + (r - ex[(size_t)t * n_struct + k]);
}

void build_var_pos(fm_structure s) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  int n = s->n, ns = s->n_struct;
  if (ns <= 1 || n == 0) {
    s->var_pos.release();
    return;
  }
  int m = s->nos * ns;
  DevBuf cnt(&ctx->pool, (size_t)m * sizeof(int)), ex(&ctx->pool, (size_t)(m + 1) * sizeof(int));
  FM_HIP(hipMemsetAsync(cnt.as<int>(), 0, (size_t)m * sizeof(int), st));
  hipLaunchKernelGGL(k_struct_species_count, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     s->species.as<int>(), s->atom_struct.as<int>(), ns, s->nos, cnt.as<int>());
  scan_exclusive(ctx, cnt.as<int>(), ex.as<int>(), m, false);
  s->var_pos.alloc(&ctx->pool, (size_t)n * sizeof(int));
  hipLaunchKernelGGL(k_var_pos, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     s->atom_of_row.as<int>(), s->species.as<int>(), s->atom_struct.as<int>(),
                     s->d_atom_off.as<int>(), cnt.as<int>(), ex.as<int>(), ns,
                     s->var_pos.as<int>());
  FM_HIP(hipGetLastError());
}

// Results to the caller's host buffers (any may be null); returns after the
// copies have finished.
static void read_outputs(fm_structure s, const double *o, size_t n_out, const double *var_row,
                         const double *E_row, double *mean_efs, double *local_var,
                         double *local_energy) {
  hipStream_t st = s->ctx->stream;
  Pool *pool = &s->ctx->pool;
  int n = s->n;
  DevBuf vperm, eperm;
  if (mean_efs)
    FM_HIP(hipMemcpyAsync(mean_efs, o, (size_t)n_out * sizeof(double), hipMemcpyDeviceToHost, st));
  if (local_var && var_row && n > 0) {
    // Species-sorted rows of the batch to each structure's species-sorted
    // order (compute_cluster_uncertainties order).
    const double *src = var_row;
    if (s->n_struct > 1) {
      vperm.alloc(pool, (size_t)n * sizeof(double));
      hipLaunchKernelGGL(k_scatter_rows, dim3((n + 255) / 256), dim3(256), 0, st, n,
                         s->var_pos.as<int>(), var_row, vperm.as<double>());
      src = vperm.as<double>();
    }
    FM_HIP(hipMemcpyAsync(local_var, src, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                          st));
  }
  if (local_energy && n > 0) {
    // E_row is in species-sorted order; local energies are in atom order.
    eperm.alloc(pool, (size_t)n * sizeof(double));
    hipLaunchKernelGGL(k_scatter_rows, dim3((n + 255) / 256), dim3(256), 0, st, n,
                       s->atom_of_row.as<int>(), E_row, eperm.as<double>());
    FM_HIP(hipMemcpyAsync(local_energy, eperm.as<double>(), (size_t)n * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  }
  FM_HIP(hipGetLastError());
  FM_HIP(hipStreamSynchronize(st));
}


// Rows per tile of one species in the chosen form: at most T and, when T
// allows, a whole number of rounds of the kernel contraction's 128 x 64
// tiles over the GPU's workgroup slots (one such workgroup per CU: its two
// LDS buffers take 122 KB), so the last round of each launch is not left
// nearly empty. Rows of a tile do not change any result.
static int species_tile_rows(fm_context ctx, int T, int Ms, int form) {
  if (form != FM_FORM_CHOSEN || Ms == 0) return T;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
          hipSuccess ||
      cus <= 0)
    return T;
  int per_round = cus / ((Ms + kI8TileN - 1) / kI8TileN);
  int rounds = per_round > 0 ? (T / kI8TileA) / per_round : 0;
  return rounds > 0 ? rounds * per_round * kI8TileA : T;
}

// The model's sparse rows as int8 slices for the chosen form's kernel
// contraction, made on the context stream from the FP64 rows on first use.
static void model_slices(fm_model m, hipStream_t st) {
  if (m->xq_ready) return;
  int nd = m->nd;
  long long ldq = i8_ld(nd);
  Pool *pool = &m->ctx->pool;
  m->Xq.resize(m->nos);
  m->Xq_scale.resize(m->nos);
  for (int t = 0; t < m->nos; t++) {
    int Ms = m->M_by_type[t];
    if (Ms == 0) continue;
    m->Xq[t].alloc(pool, i8_tiled_bytes(Ms, kI8TileN, ldq));
    m->Xq_scale[t].alloc(pool, (size_t)Ms * sizeof(double));
    slice_rows(st, m->X[t].as<double>(), nd, Ms, nd, m->Xq[t].as<int8_t>(), ldq, kI8TileN,
               m->Xq_scale[t].as<double>());
  }
  m->xq_ready = true;
}

// Chosen form: `cnt` FP32 descriptor rows (leading dimension nd) as int8
// slices in Aq / Aq_scale (for rows without slices of their FP64 values;
// Aq holds at least i8_tiled_bytes(cnt, kI8TileA, i8_ld(nd)) bytes).
static SliceOperand slice_fp32_rows(fm_context ctx, hipStream_t st, const float *rows, int cnt,
                                    int nd, DevBuf &Aq, DevBuf &Aq_scale) {
  long long ldq = i8_ld(nd);
  {
    KScope timer(ctx, KT_SLICE, st);
    slice_rows(st, rows, nd, cnt, nd, Aq.as<int8_t>(), ldq, kI8TileA, Aq_scale.as<double>());
  }
  SliceOperand A;
  A.q = Aq.as<int8_t>();
  A.Kq = ldq;
  A.tr = kI8TileA;
  A.scale = Aq_scale.as<double>();
  return A;
}

// Chosen form: kernel contraction of `cnt` sliced descriptor rows against
// species t's sparse set.
static void kernel_contraction_i8(fm_model m, hipStream_t st, int t, const SliceOperand &A,
                                  int cnt, int nd, const EpArgs &ek) {
  long long ldq = i8_ld(nd);
  int Ms = m->M_by_type[t];
  SliceOperand Bx;
  Bx.q = m->Xq[t].as<int8_t>();
  Bx.Kq = ldq;
  Bx.tr = kI8TileN;
  Bx.scale = m->Xq_scale[t].as<double>();
  KScope timer(m->ctx, KT_SGEMM, st);
  gemm_i8_kernel(st, cnt, Ms, (int)ldq, A, Bx, ek);
}

// B3 structures too large to keep their descriptors (s->row_tile > 0): each
// species' rows are processed in tiles of row_tile rows. Per tile: B3 rows
// (from the coefficients c kept for every atom), the kernel contraction with
// its energy partials, the variance contraction, the gradient contraction
// and the adjoint of the tile's atoms (their u and bond forces); the force
// gather runs once at the end. Every output element sums its terms in the
// order of the untiled path, so the results are the same bit for bit.
static void predict_tiled(fm_model m, fm_structure s, int flags, int form, double *mean_efs,
                          double *local_var, double *local_energy) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  Pool *pool = &ctx->pool;
  bool want_var = (flags & FM_PREDICT_LOCAL) != 0;
  bool f64 = form == FM_FORM_FP64;
  size_t esz = f64 ? sizeof(double) : sizeof(float);
  int n = s->n, nd = s->nd, nos = s->nos;
  size_t n_out = (size_t)s->mean_off[s->n_struct];
  int T = (int)std::min<int64_t>(s->row_tile, std::max(1, n));
  std::vector<size_t> eo(nos + 1, 0), qo(nos + 1, 0);
  int Mmax = 0;
  for (int t = 0; t < nos; t++) {
    size_t ns = s->n_by_type[t];
    int Ms = m->M_by_type[t];
    Mmax = std::max(Mmax, Ms);
    eo[t + 1] = eo[t] + ns * (size_t)std::max(1, gemm_col_tiles(form, Ms));
    int Lr = want_var ? m->L_rows[t] : 0;
    qo[t + 1] = qo[t] + ns * (size_t)gemm_row_tiles(FM_FORM_FP64, Lr);
  }
  // Whole-structure buffers.
  DevBuf Epart(pool, std::max<size_t>(1, eo[nos] * sizeof(double)));
  DevBuf Qpart(pool, std::max<size_t>(1, qo[nos] * sizeof(double)));
  DevBuf E_row(pool, std::max<size_t>(1, (size_t)n * sizeof(double)));
  DevBuf var_row(pool, std::max<size_t>(1, (size_t)n * sizeof(double)));
  DevBuf out(pool, (size_t)n_out * sizeof(double));
  size_t nc = (size_t)s->nos * s->N * (s->lmax + 1) * (s->lmax + 1);
  DevBuf U(pool, std::max<size_t>(1, (size_t)n * nc * esz));
  DevBuf fbond(pool, std::max<size_t>(1, (size_t)3 * s->n_bonds * esz));
  DevBuf vir(pool, std::max<size_t>(1, (size_t)6 * n * sizeof(double)));
  // Tile buffers. k is FP64 in both forms (rows padded to 128 bytes).
  long long ldm = split_ld(std::max(Mmax, 1));
  size_t kld = ((size_t)Mmax + 15) / 16 * 16;  // rows of whole 128-byte lines
  DevBuf Bt(pool, std::max<size_t>(1, (size_t)T * nd * esz));
  DevBuf Gt(pool, std::max<size_t>(1, (size_t)T * nd * esz));
  DevBuf Tq, Tq_scale, Wt, Wh, Wl, Kt;
  if (f64) {
    Wt.alloc(pool, std::max<size_t>(1, (size_t)T * Mmax * sizeof(double)));
  } else {
    model_slices(m, st);
    Tq.alloc(pool, std::max<size_t>(1, i8_tiled_bytes(T, kI8TileA, i8_ld(nd))));
    Tq_scale.alloc(pool, std::max<size_t>(1, (size_t)T * sizeof(double)));
    Wh.alloc(pool, std::max<size_t>(1, (size_t)T * ldm * sizeof(uint16_t)));
    Wl.alloc(pool, std::max<size_t>(1, (size_t)T * ldm * sizeof(uint16_t)));
  }
  if (want_var) Kt.alloc(pool, std::max<size_t>(1, (size_t)T * kld * sizeof(double)));
  double sig2 = m->sigma * m->sigma;
  ctx->mark_begin(ST_SGEMM);
  for (int t = 0; t < nos; t++) {
    int ns = s->n_by_type[t], Ms = m->M_by_type[t], ro = s->type_off[t];
    int n_et = Ms == 0 ? 1 : gemm_col_tiles(form, Ms);
    int n_qt = (want_var && Ms > 0) ? gemm_row_tiles(FM_FORM_FP64, m->L_rows[t]) : 0;
    int Ts = species_tile_rows(ctx, T, Ms, form);
    for (int c0 = 0; c0 < ns; c0 += Ts) {
      int cnt = std::min(Ts, ns - c0), r0 = ro + c0;
      launch_b3_rows(s, r0, cnt, true, Bt.as<void>());
      double *ep = Epart.as<double>() + eo[t] + c0;
      if (Ms == 0) {
        hipLaunchKernelGGL(k_fill, dim3((cnt + 255) / 256), dim3(256), 0, st, ep,
                           (long long)cnt, 0.0);
        FM_HIP(hipMemsetAsync(Gt.as<void>(), 0, (size_t)cnt * nd * esz, st));
      } else {
        // Kernel contraction of the tile.
        EpArgs ek;
        ek.sig2 = sig2;
        ek.power = m->power;
        ek.row_norm = s->norm.as<double>() + r0;
        ek.alpha = m->alpha[t].as<double>();
        if (f64) {
          ek.W = Wt.as<double>();
          ek.ldw = Ms;
        } else {
          ek.Wh = Wh.as<uint16_t>();
          ek.Wl = Wl.as<uint16_t>();
          ek.ldw = split_ld(Ms);
        }
        size_t kl = ((size_t)Ms + 1) / 2 * 2;
        if (want_var) {
          ek.K = Kt.as<void>();
          ek.ldk = (long long)kl;
        }
        ek.part = ep;
        ek.part_ld = ns;
        if (f64)
          gemm_f64(st, EP_KERNEL, cnt, Ms, nd, Bt.as<double>(), nd, m->X[t].as<double>(), nd, 0,
                   ek);
        else
          kernel_contraction_i8(m, st, t,
                                slice_fp32_rows(ctx, st, Bt.as<float>(), cnt, nd, Tq, Tq_scale),
                                cnt, nd, ek);
        // Variance contraction of the tile.
        if (want_var && m->L_rows[t] > 0) {
          EpArgs eq;
          eq.part = Qpart.as<double>() + qo[t] + c0;
          eq.part_ld = ns;
          gemm_f64(st, EP_SUMSQ, m->L_rows[t], cnt, Ms, m->L[t].as<double>(), Ms,
                   Kt.as<double>(), (long long)kl, m->L_lower[t], eq);
        }
        // Gradient contraction of the tile.
        EpArgs eg;
        eg.W = Gt.as<void>();
        eg.ldw = nd;
        if (f64) {
          gemm_f64(st, EP_STORE, cnt, nd, Ms, Wt.as<double>(), Ms, m->XT[t].as<double>(), Ms, 0,
                   eg);
        } else {
          long long lm = split_ld(Ms);
          SplitOperand A, Bx;
          A.piece[0] = Wh.as<uint16_t>();
          A.piece[1] = Wl.as<uint16_t>();
          A.ld = lm;
          Bx.piece[0] = m->XTs0[t].as<uint16_t>();
          Bx.piece[1] = m->XTs1[t].as<uint16_t>();
          Bx.ld = lm;
          gemm_split(st, EP_STORE, 2, cnt, nd, (int)lm, A, Bx, eg);
        }
      }
      hipLaunchKernelGGL(k_species_final, dim3((cnt + 255) / 256), dim3(256), 0, st, cnt, r0, ep,
                         n_et, Qpart.as<double>() + qo[t] + c0, n_qt, (long long)ns,
                         s->norm.as<double>(), s->nn.as<double>(), sig2, m->power,
                         E_row.as<double>(), want_var ? var_row.as<double>() : nullptr);
      // Adjoint of the tile's atoms: u, then their bonds' forces.
      launch_dEdc_b3_rows(s, r0, cnt, true, Bt.as<void>(), Gt.as<void>(), nd, 1, kEmptyThresh,
                          -1.0, U.as<void>());
      launch_bond_forces(s, U.as<void>(), fbond.as<void>(), vir.as<double>(),
                         s->atom_of_row.as<int>() + r0, cnt, kEmptyThresh, -1.0);
      FM_HIP(hipGetLastError());
    }
  }
  ctx->mark_end(ST_SGEMM);
  ctx->mark_begin(ST_ADJ);
  double *o = out.as<double>();
  FM_HIP(hipMemsetAsync(o, 0, (size_t)n_out * sizeof(double), st));
  if (n > 0) gather_forces(s, f64, fbond, vir, E_row, o);
  FM_HIP(hipGetLastError());
  ctx->mark_end(ST_ADJ);
  read_outputs(s, o, n_out, want_var ? var_row.as<double>() : nullptr, E_row.as<double>(),
               mean_efs, local_var, local_energy);
  FM_HIP(hipStreamSynchronize(st));
}

void predict(fm_model m, fm_structure s, int flags, int form, double *mean_efs,
             double *local_var, double *local_energy) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  Pool *pool = &ctx->pool;
  FM_CHECK(s->has_b2, FM_ERR_STATE, "descriptors have not been computed for this structure");
  FM_CHECK(s->form == form, FM_ERR_STATE, "structure descriptors were computed in another form");
  FM_CHECK(m->nos == s->nos && m->nd == s->nd, FM_ERR_INVALID,
           "model and descriptor settings disagree (species or descriptor size)");
  if (s->desc_kind == 3)
    b3_order_model(m, s);  // the sparse set in the rows' (l-set major) order
  else
    FM_CHECK(m->b3_ntri == 0, FM_ERR_INVALID,
             "the model's sparse set is in B3 order (it was used with a B3 descriptor)");
  bool want_var = (flags & FM_PREDICT_LOCAL) != 0;
  FM_CHECK(!want_var || m->has_L, FM_ERR_STATE,
           "L_inv must be up to date to evaluate uncertainties");
  ctx->kt_clear(KT_SLICE, KT_GATHER);
  if (s->desc_kind == 3 && s->row_tile > 0)
    return predict_tiled(m, s, flags, form, mean_efs, local_var, local_energy);
  bool f64 = form == FM_FORM_FP64;
  size_t esz = f64 ? sizeof(double) : sizeof(float);
  int n = s->n, nd = s->nd, nos = s->nos;
  // mean_efs of every structure of the batch, concatenated.
  size_t n_out = (size_t)s->mean_off[s->n_struct];
  // Per-species slices of the atom-major w and k matrices (n_s x M_s each)
  // and of the energy / variance partials.
  // The chosen form keeps w as BF16 pairs with padded rows (ws = split_ld).
  // k is FP64 in both forms, rows padded to 16 values (whole 128-byte lines).
  std::vector<size_t> wo(nos + 1, 0), eo(nos + 1, 0), qo(nos + 1, 0), so(nos + 1, 0),
      ko(nos + 1, 0), kld(nos, 0);
  for (int t = 0; t < nos; t++) {
    size_t ns = s->n_by_type[t], Ms = m->M_by_type[t];
    wo[t + 1] = wo[t] + ns * Ms;
    kld[t] = (Ms + 15) / 16 * 16;  // rows of whole 128-byte lines
    ko[t + 1] = ko[t] + ns * kld[t];
    so[t + 1] = so[t] + ns * (size_t)split_ld((long long)Ms);
    eo[t + 1] = eo[t] + ns * (size_t)std::max(1, gemm_col_tiles(form, (int)Ms));
    int Lr = want_var ? m->L_rows[t] : 0;
    qo[t + 1] = qo[t] + ns * (size_t)gemm_row_tiles(FM_FORM_FP64, Lr);
  }
  DevBuf W, Wh, Wl;
  if (f64) {
    W.alloc(pool, std::max<size_t>(1, wo[nos] * esz));
  } else {
    Wh.alloc(pool, std::max<size_t>(1, so[nos] * sizeof(uint16_t)));
    Wl.alloc(pool, std::max<size_t>(1, so[nos] * sizeof(uint16_t)));
  }
  DevBuf Kb;
  size_t ksz = sizeof(double);
  if (want_var) Kb.alloc(pool, std::max<size_t>(1, ko[nos] * ksz));
  // Chosen form: the descriptor rows' int8 slices, made by the descriptor
  // pass from the FP64 rows (B2), or here from the FP32 rows one species at
  // a time (B3).
  DevBuf Aq, Aq_scale;
  if (!f64) {
    model_slices(m, st);
    if (!s->has_q) {
      int nmax = 0;
      for (int t = 0; t < nos; t++) nmax = std::max(nmax, s->n_by_type[t]);
      Aq.alloc(pool, std::max<size_t>(1, i8_tiled_bytes(nmax, kI8TileA, i8_ld(nd))));
      Aq_scale.alloc(pool, std::max<size_t>(1, (size_t)nmax * sizeof(double)));
    }
  }
  DevBuf G(pool, std::max<size_t>(1, (size_t)n * nd * esz));
  DevBuf Epart(pool, std::max<size_t>(1, eo[nos] * sizeof(double)));
  DevBuf Qpart(pool, std::max<size_t>(1, qo[nos] * sizeof(double)));
  DevBuf E_row(pool, std::max<size_t>(1, (size_t)n * sizeof(double)));
  DevBuf var_row(pool, std::max<size_t>(1, (size_t)n * sizeof(double)));
  DevBuf out(pool, (size_t)n_out * sizeof(double));
  double sig2 = m->sigma * m->sigma;

  // Kernel contraction S = B_s X_hat_s^T with the fused epilogue.
  ctx->mark_begin(ST_SGEMM);
  for (int t = 0; t < nos; t++) {
    int ns = s->n_by_type[t], Ms = m->M_by_type[t], ro = s->type_off[t];
    if (ns == 0) continue;
    if (Ms == 0) {
      hipLaunchKernelGGL(k_fill, dim3((ns + 255) / 256), dim3(256), 0, st,
                         Epart.as<double>() + eo[t], (long long)ns, 0.0);
      continue;
    }
    EpArgs ek;
    ek.sig2 = sig2;
    ek.power = m->power;
    ek.row_norm = s->norm.as<double>() + ro;
    ek.alpha = m->alpha[t].as<double>();
    if (f64) {
      ek.W = W.as<double>() + wo[t];
      ek.ldw = Ms;
    } else {
      ek.Wh = Wh.as<uint16_t>() + so[t];
      ek.Wl = Wl.as<uint16_t>() + so[t];
      ek.ldw = split_ld(Ms);
    }
    if (want_var) {
      ek.K = Kb.as<char>() + ko[t] * ksz;
      ek.ldk = (long long)kld[t];
    }
    ek.part = Epart.as<double>() + eo[t];
    ek.part_ld = ns;
    if (f64) {
      KScope timer(ctx, KT_SGEMM);
      gemm_f64(st, EP_KERNEL, ns, Ms, nd, s->B.as<double>() + (size_t)ro * nd, nd,
               m->X[t].as<double>(), nd, 0, ek);
    } else {
      SliceOperand A;
      if (s->has_q) {
        // Species t's rows start a tile of the structure's slices.
        A.q = s->Bq.as<int8_t>() + i8_tiled_bytes(s->q_tile_row[t], kI8TileA, s->ldq);
        A.Kq = s->ldq;
        A.tr = kI8TileA;
        A.scale = s->Bq_scale.as<double>() + ro;
      } else {
        A = slice_fp32_rows(ctx, st, s->B.as<float>() + (size_t)ro * nd, ns, nd, Aq, Aq_scale);
      }
      kernel_contraction_i8(m, st, t, A, ns, nd, ek);
    }
    ctx->launches[ST_SGEMM]++;
  }
  ctx->mark_end(ST_SGEMM);

  // Gradient contraction G = W X_hat_s.
  ctx->mark_begin(ST_GGEMM);
  for (int t = 0; t < nos; t++) {
    int ns = s->n_by_type[t], Ms = m->M_by_type[t], ro = s->type_off[t];
    if (ns == 0) continue;
    void *Grow = (char *)G.as<char>() + (size_t)ro * nd * esz;
    if (Ms == 0) {
      FM_HIP(hipMemsetAsync(Grow, 0, (size_t)ns * nd * esz, st));
      continue;
    }
    EpArgs eg;
    eg.W = Grow;
    eg.ldw = nd;
    KScope timer(ctx, KT_GGEMM);
    if (f64)
      gemm_f64(st, EP_STORE, ns, nd, Ms, W.as<double>() + wo[t], Ms, m->XT[t].as<double>(), Ms,
               0, eg);
    else {
      long long ldm = split_ld(Ms);
      SplitOperand A, Bx;
      A.piece[0] = Wh.as<uint16_t>() + so[t];
      A.piece[1] = Wl.as<uint16_t>() + so[t];
      A.ld = ldm;
      Bx.piece[0] = m->XTs0[t].as<uint16_t>();
      Bx.piece[1] = m->XTs1[t].as<uint16_t>();
      Bx.ld = ldm;
      gemm_split(st, EP_STORE, 2, ns, nd, (int)ldm, A, Bx, eg);
    }
    ctx->launches[ST_GGEMM]++;
  }
  ctx->mark_end(ST_GGEMM);

  // Local energies from the energy partials (main stream).
  for (int t = 0; t < nos; t++) {
    int ns = s->n_by_type[t], Ms = m->M_by_type[t], ro = s->type_off[t];
    if (ns == 0) continue;
    int n_et = Ms == 0 ? 1 : gemm_col_tiles(form, Ms);
    hipLaunchKernelGGL(k_species_final, dim3((ns + 255) / 256), dim3(256), 0, st, ns, ro,
                       Epart.as<double>() + eo[t], n_et, nullptr, 0, (long long)ns,
                       s->norm.as<double>(), s->nn.as<double>(), sig2, m->power,
                       E_row.as<double>(), nullptr);
  }
  FM_HIP(hipGetLastError());

  // Variance contraction Q = L_s^-1 K^T (FP64 in both forms) on the side
  // stream, after the kernel contraction that wrote k; it overlaps the
  // gradient contraction and the adjoint, and is joined before the results
  // are read back.
  // Joins the side stream into the main one when predict leaves this scope
  // by any path (declared after the buffers the side stream uses, so it runs
  // before they return to the pool).
  struct SideJoin {
    fm_context c;
    bool active = false;
    ~SideJoin() {
      if (active) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
    }
  } join{ctx};
  if (want_var) {
    hipStream_t sd = ctx->side;
    FM_HIP(hipEventRecord(ctx->ev_fork, st));
    FM_HIP(hipStreamWaitEvent(sd, ctx->ev_fork, 0));
    ctx->mark_begin(ST_VAR, sd);
    for (int t = 0; t < nos; t++) {
      int ns = s->n_by_type[t], Ms = m->M_by_type[t];
      if (ns == 0 || Ms == 0 || m->L_rows[t] == 0) continue;
      EpArgs eq;
      eq.part = Qpart.as<double>() + qo[t];
      eq.part_ld = ns;
      KScope timer(ctx, KT_VAR, sd);
      gemm_f64(sd, EP_SUMSQ, m->L_rows[t], ns, Ms, m->L[t].as<double>(), Ms,
               Kb.as<double>() + ko[t], (long long)kld[t], m->L_lower[t], eq);
      ctx->launches[ST_VAR]++;
    }
    for (int t = 0; t < nos; t++) {
      int ns = s->n_by_type[t], Ms = m->M_by_type[t], ro = s->type_off[t];
      if (ns == 0) continue;
      int n_qt = Ms > 0 ? gemm_row_tiles(FM_FORM_FP64, m->L_rows[t]) : 0;
      hipLaunchKernelGGL(k_var_final, dim3((ns + 255) / 256), dim3(256), 0, sd, ns, ro,
                         Qpart.as<double>() + qo[t], n_qt, (long long)ns, s->norm.as<double>(),
                         s->nn.as<double>(), sig2, m->power, var_row.as<double>());
    }
    FM_HIP(hipGetLastError());
    ctx->mark_end(ST_VAR, sd);
    FM_HIP(hipEventRecord(ctx->ev_join, sd));
    join.active = true;
  }

  // Adjoint bond forces and the deterministic gather.
  ctx->mark_begin(ST_ADJ);
  DevBuf fbond(pool, std::max<size_t>(1, (size_t)3 * s->n_bonds * esz));
  DevBuf vir(pool, std::max<size_t>(1, (size_t)6 * n * sizeof(double)));
  double *o = out.as<double>();
  FM_HIP(hipMemsetAsync(o, 0, (size_t)n_out * sizeof(double), st));
  if (n > 0) {
    // FP64: G . B from the local energies (G . B = p |B| E, exact to
    // rounding). Split form: from the G and B rows, since G carries the
    // BF16-pair rounding of w that E does not, and the projection
    // G - (G . B_hat) B_hat must see the G it is applied to.
    launch_adjoint(s, G.as<void>(), nd, fbond.as<void>(), vir.as<double>(), 1, kEmptyThresh,
                   -1.0, f64 ? E_row.as<double>() : nullptr, m->power);
    gather_forces(s, f64, fbond, vir, E_row, o);
    ctx->launches[ST_ADJ] += 3;
  }
  FM_HIP(hipGetLastError());
  ctx->mark_end(ST_ADJ);
  if (join.active) {
    FM_HIP(hipStreamWaitEvent(st, ctx->ev_join, 0));
    join.active = false;
  }

  read_outputs(s, o, n_out, want_var ? var_row.as<double>() : nullptr, E_row.as<double>(),
               mean_efs, local_var, local_energy);
  FM_HIP(hipStreamSynchronize(st));
}

}  // namespace fm
