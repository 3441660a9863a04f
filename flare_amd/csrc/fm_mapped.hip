// The mapped model: mapping coefficients of a NormalizedDotProduct sparse GP
// (normalized_dot_product.cpp:631-767) and the pair_style flare evaluation
// from them (pair_flare.cpp:68-184, 262-419; lammps_descriptor.cpp:209-315).
//
// beta_s = sum_j sigma^2 alpha_j x_hat_j x_hat_j^T is one FP64 contraction
// over the sparse set per species. Evaluation: P = B_s beta_s on the matrix
// cores, then per atom E = B.P / |B|^2 (power 2, normalized) and dE/dB, and
// the same adjoint kernel as the sparse-GP path turns dE/dB into bond
// forces.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "fm_device.hpp"
#include "fm_gemm.hpp"
#include "fm_internal.hpp"

namespace fm {

// XTa[k][j] = sigma^2 alpha_j XT[k][j]
__global__ void k_scale_cols(const double *XT, const double *alpha, double sig2, int d, int Ms,
                             double *out) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)d * Ms) return;
  int j = (int)(t % Ms);
  out[t] = XT[t] * sig2 * alpha[j];
}

void mapping_coefficients(fm_model m, std::vector<double> &beta, int64_t &beta_size) {
  fm_context ctx = m->ctx;
  hipStream_t st = ctx->stream;
  int p = (int)m->power;
  FM_CHECK(m->power == 1.0 || m->power == 2.0, FM_ERR_INVALID,
           "Mapping coefficients of the normalized dot product kernel are implemented for "
           "power 1 and 2 only");
  int d = m->nd;
  beta_size = p == 1 ? d : (int64_t)d * (d + 1) / 2;
  beta.assign((size_t)m->nos * beta_size, 0.0);
  double sig2 = m->sigma * m->sigma;
  std::vector<double> full((size_t)d * (p == 1 ? 1 : d));
  for (int t = 0; t < m->nos; t++) {
    int Ms = m->M_by_type[t];
    if (Ms == 0) continue;
    DevBuf XTa(&ctx->pool, (size_t)d * Ms * sizeof(double));
    long long tot = (long long)d * Ms;
    hipLaunchKernelGGL(k_scale_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       m->XT[t].as<double>(), m->alpha[t].as<double>(), sig2, d, Ms,
                       XTa.as<double>());
    DevBuf C(&ctx->pool, full.size() * sizeof(double));
    EpArgs e;
    e.W = C.as<void>();
    if (p == 2) {
      e.ldw = d;
      gemm_f64(st, EP_STORE, d, d, Ms, XTa.as<double>(), Ms, m->XT[t].as<double>(), Ms, 0, e);
    } else {
      // beta = sum_j sigma^2 alpha_j x_hat_j: a ones-vector contraction with
      // the alpha-scaled X_hat^T (1 x d).
      std::vector<double> ones(Ms, 1.0);
      DevBuf a(&ctx->pool, (size_t)Ms * sizeof(double));
      FM_HIP(hipMemcpyAsync(a.as<double>(), ones.data(), Ms * sizeof(double),
                            hipMemcpyHostToDevice, st));
      FM_HIP(hipStreamSynchronize(st));
      e.ldw = d;
      gemm_f64(st, EP_STORE, 1, d, Ms, a.as<double>(), Ms, XTa.as<double>(), Ms, 0, e);
      FM_HIP(hipStreamSynchronize(st));
    }
    FM_HIP(hipMemcpyAsync(full.data(), C.as<double>(), full.size() * sizeof(double),
                          hipMemcpyDeviceToHost, st));
    FM_HIP(hipStreamSynchronize(st));
    double *row = &beta[(size_t)t * beta_size];
    if (p == 1) {
      for (int k = 0; k < d; k++) row[k] = full[k];
    } else {
      int64_t c = 0;
      for (int k = 0; k < d; k++)
        for (int l = k; l < d; l++) row[c++] = (k == l ? 1.0 : 2.0) * full[(size_t)k * d + l];
    }
  }
}

// Per atom: energy and G (the input of the adjoint) from P = B beta.
//   normalized p=2: E = B.P/|B|^2, G = 2P/|B| (projected in the adjoint)
//   normalized p=1: E = B.beta/|B|, G = beta
//   unnormalised p=2: E = B.P, w = 2P;  p=1: E = B.beta, w = beta
// (lammps_descriptor.cpp:259-282). Atoms with |B|^2 < 1e-8 get E = 0
// (pair_flare.cpp:134-136).
__global__ void k_mapped_energy(int ns, int ro, int d, const double *B, const double *P,
                                const double *beta, int power, int normalized,
                                const double *norm, const double *nn, double *E_row,
                                double *G) {
  int lane = threadIdx.x & 63;
  int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= ns) return;
  int row = ro + r;
  const double *b = B + (size_t)row * d;
  double *g = G + (size_t)row * d;
  double n2 = nn[row], nrm = norm[row];
  if (n2 < 1e-8) {
    for (int k = lane; k < d; k += 64) g[k] = 0;
    if (lane == 0) E_row[row] = 0;
    return;
  }
  const double *v = power == 2 ? P + (size_t)r * d : beta;
  double acc = 0;
  for (int k = lane; k < d; k += 64) acc += b[k] * v[k];
  acc = wave_sum(acc);
  double e;
  if (normalized) e = power == 2 ? acc / n2 : acc / nrm;
  else e = acc;
  double scale = power == 2 ? (normalized ? 2.0 / nrm : 2.0) : 1.0;
  for (int k = lane; k < d; k += 64) g[k] = scale * v[k];
  if (lane == 0) E_row[row] = e;
}

// Forces (atom order), per-atom energies, and per-chunk virial partials;
// one block per chunk of <= 256 atoms of one structure.
__global__ void k_mapped_gather(const int *chunk_begin, const int *chunk_end, const int *bond_off,
                                const int *in_off, const int *in_bond, const double *fbond,
                                const int *row_of_atom, const double *E_row, const double *vir,
                                double *F, double *Eatom, double *blk_part) {
  __shared__ double sh[6][256];
  int c = blockIdx.x;
  int i = chunk_begin[c] + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (i < chunk_end[c]) {
    double f[3] = {0, 0, 0};
    for (int b = bond_off[i]; b < bond_off[i + 1]; b++)
      for (int q = 0; q < 3; q++) f[q] += fbond[3 * (size_t)b + q];
    for (int a = in_off[i]; a < in_off[i + 1]; a++) {
      int b = in_bond[a];
      for (int q = 0; q < 3; q++) f[q] -= fbond[3 * (size_t)b + q];
    }
    for (int q = 0; q < 3; q++) F[3 * (size_t)i + q] = f[q];
    Eatom[i] = E_row[row_of_atom[i]];
    for (int q = 0; q < 6; q++) v[q] = vir[6 * (size_t)i + q];
  }
  for (int q = 0; q < 6; q++) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 6; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 6) blk_part[(size_t)c * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_mapped_sum(int n_struct, const int *struct_chunk_off, const double *blk_part,
                             double *vir) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_struct) return;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int c = struct_chunk_off[k]; c < struct_chunk_off[k + 1]; c++)
    for (int q = 0; q < 6; q++) acc[q] += blk_part[(size_t)c * 6 + q];
  for (int q = 0; q < 6; q++) vir[6 * (size_t)k + q] = acc[q];
}

__global__ void k_zero(double *x, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

// Working buffers of one evaluation (device memory, at least the sizes
// mapped_layout gives).
struct MappedBufs {
  double *G, *E_row, *P, *fbond, *vir, *F, *Ea, *vsum, *blk;
  void *U;
};

static void mapped_checks(fm_mapped mp, fm_structure s) {
  FM_CHECK(s->has_b2 && s->form == FM_FORM_FP64, FM_ERR_STATE,
           "the structure needs FP64 B2 descriptors for the mapped potential");
  FM_CHECK(s->desc_kind == 2, FM_ERR_INVALID, "pair_style flare evaluates the B2 descriptor");
  FM_CHECK(s->nos == mp->nos && s->N == mp->N && s->lmax == mp->lmax &&
               s->cutkind == mp->cutkind && s->r1 == 0.0,
           FM_ERR_INVALID, "structure descriptors do not match the potential's settings");
  for (size_t k = 0; k < mp->cutoffs.size(); k++)
    FM_CHECK(s->cutoffs[k] == mp->cutoffs[k], FM_ERR_INVALID,
             "structure cutoffs do not match the potential's cutoffs");
}

// Byte sizes of the buffers, in MappedBufs order.
static void mapped_layout(fm_mapped mp, fm_structure s, size_t sz[10]) {
  size_t n = s->n, d = s->nd, maxns = 0;
  for (int t = 0; t < s->nos; t++) maxns = std::max(maxns, (size_t)s->n_by_type[t]);
  size_t w = sizeof(double);
  sz[0] = n * d * w;                                // G
  sz[1] = n * w;                                    // E_row
  sz[2] = mp->power == 2 ? maxns * d * w : 0;       // P
  sz[3] = 3 * (size_t)s->n_bonds * w;               // fbond
  sz[4] = 6 * n * w;                                // vir
  sz[5] = 3 * n * w;                                // F
  sz[6] = n * w;                                    // Ea
  sz[7] = 6 * (size_t)s->n_struct * w;              // vsum
  sz[8] = 6 * (size_t)s->n_chunks * w;              // blk
  sz[9] = adjoint_u_bytes(s);                       // U
  for (int k = 0; k < 10; k++) sz[k] = std::max<size_t>(sz[k], 1);
}

// pair_flare.cpp:68-184 on the given buffers; outputs (host or device
// memory, any may be null) are copied on the context stream, then the
// stream is synchronised.
static void mapped_run(fm_mapped mp, fm_structure s, const MappedBufs &b, double *atom_energy,
                       double *forces, double *virial) {
  hipStream_t st = s->ctx->stream;
  int n = s->n, d = s->nd;
  for (int t = 0; t < s->nos; t++) {
    int ns = s->n_by_type[t], ro = s->type_off[t];
    if (ns == 0) continue;
    const double *Bs = s->B.as<double>() + (size_t)ro * d;
    if (mp->power == 2) {
      EpArgs e;
      e.W = b.P;
      e.ldw = d;
      gemm_f64(st, EP_STORE, ns, d, d, Bs, d, mp->beta[t].as<double>(), d, 0, e);
    }
    hipLaunchKernelGGL(k_mapped_energy, dim3((ns + 3) / 4), dim3(256), 0, st, ns, ro, d,
                       s->B.as<double>(), mp->power == 2 ? b.P : nullptr,
                       mp->beta[t].as<double>(), mp->power, mp->normalized,
                       s->norm.as<double>(), s->nn.as<double>(), b.E_row, b.G);
  }
  long long nv = 6 * (long long)s->n_struct;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, b.vsum, nv);
  if (n > 0) {
    launch_adjoint(s, b.G, d, b.fbond, b.vir, mp->normalized ? 1 : 0, -1.0, 1e-8, nullptr, 0,
                   b.U);
    hipLaunchKernelGGL(k_mapped_gather, dim3(s->n_chunks), dim3(256), 0, st,
                       s->d_chunk_begin.as<int>(), s->d_chunk_end.as<int>(), s->bond_off.as<int>(),
                       s->in_off.as<int>(), s->in_bond.as<int>(), b.fbond,
                       s->row_of_atom.as<int>(), b.E_row, b.vir, b.F, b.Ea, b.blk);
    hipLaunchKernelGGL(k_mapped_sum, dim3((s->n_struct + 63) / 64), dim3(64), 0, st, s->n_struct,
                       s->d_struct_chunk_off.as<int>(), b.blk, b.vsum);
  }
  FM_HIP(hipGetLastError());
  if (n > 0 && atom_energy)
    FM_HIP(hipMemcpyAsync(atom_energy, b.Ea, n * sizeof(double), hipMemcpyDefault, st));
  if (n > 0 && forces)
    FM_HIP(hipMemcpyAsync(forces, b.F, 3 * (size_t)n * sizeof(double), hipMemcpyDefault, st));
  if (virial)
    FM_HIP(hipMemcpyAsync(virial, b.vsum, 6 * (size_t)s->n_struct * sizeof(double),
                          hipMemcpyDefault, st));
  FM_HIP(hipStreamSynchronize(st));
}

void mapped_compute(fm_mapped mp, fm_structure s, double *atom_energy, double *forces,
                    double *virial) {
  mapped_checks(mp, s);
  Pool *pool = &s->ctx->pool;
  size_t sz[10];
  mapped_layout(mp, s, sz);
  DevBuf buf[10];
  for (int k = 0; k < 10; k++) buf[k].alloc(pool, sz[k]);
  MappedBufs b{buf[0].as<double>(), buf[1].as<double>(), buf[2].as<double>(),
               buf[3].as<double>(), buf[4].as<double>(), buf[5].as<double>(),
               buf[6].as<double>(), buf[7].as<double>(), buf[8].as<double>(),
               buf[9].as<void>()};
  mapped_run(mp, s, b, atom_energy, forces, virial);
}

// ---------------------------------------------------------------------------
// Allocation-free evaluation. The working buffers live in one static device
// array per GPU (kNoAllocScratchBytes, held until the process ends); host
// code reaches it only through the address hipGetSymbolAddress returns on
// each device, never through the array's name. Calls on one GPU share the
// array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<unsigned char *> g_scratch;  // per device; null until set up

void scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  FM_CHECK(device >= 0, FM_ERR_INVALID, "negative device index");
  if ((size_t)device < g_scratch.size() && g_scratch[device]) return;
  void *p = nullptr;
  hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  FM_CHECK(rc == hipSuccess && p, FM_ERR_HIP,
           std::string("hipGetSymbolAddress of the scratch array failed: ") +
               hipGetErrorString(rc));
  if ((size_t)device >= g_scratch.size()) g_scratch.resize(device + 1, nullptr);
  g_scratch[device] = static_cast<unsigned char *>(p);
}

static unsigned char *scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  FM_CHECK(device >= 0 && (size_t)device < g_scratch.size() && g_scratch[device], FM_ERR_STATE,
           "the scratch array of this device was not set up (create a context on it first)");
  return g_scratch[device];
}

void mapped_compute_noalloc(fm_mapped mp, fm_structure s, double *atom_energy, double *forces,
                            double *virial) {
  mapped_checks(mp, s);
  size_t sz[10], off[10], total = 0;
  mapped_layout(mp, s, sz);
  for (int k = 0; k < 10; k++) {
    off[k] = total;
    total += (sz[k] + 255) / 256 * 256;
  }
  FM_CHECK(total <= kNoAllocScratchBytes, FM_ERR_CAPACITY,
           "the structure needs " + std::to_string(total) +
               " bytes of working memory; the allocation-free entry point holds " +
               std::to_string(kNoAllocScratchBytes) + " (use fm_mapped_compute)");
  unsigned char *base = scratch_ptr(s->ctx->device);
  auto at = [&](int k) { return static_cast<void *>(base + off[k]); };
  MappedBufs b{(double *)at(0), (double *)at(1), (double *)at(2), (double *)at(3),
               (double *)at(4), (double *)at(5), (double *)at(6), (double *)at(7),
               (double *)at(8), at(9)};
  fm_context ctx = s->ctx;
  ctx->kt_on = false;  // the timers would create events
  try {
    mapped_run(mp, s, b, atom_energy, forces, virial);
  } catch (...) {
    ctx->kt_on = true;
    throw;
  }
  ctx->kt_on = true;
}

// Builds the device beta matrices: power 2 expands the upper triangle into a
// symmetric matrix with off-diagonal entries halved (pair_flare.cpp:396-418).
void mapped_upload(fm_mapped mp, const double *beta) {
  fm_context ctx = mp->ctx;
  hipStream_t st = ctx->stream;
  int d = mp->nd;
  int64_t bs = mp->power == 1 ? d : (int64_t)d * (d + 1) / 2;
  for (int t = 0; t < mp->nos; t++) {
    const double *src = beta + (size_t)t * bs;
    std::vector<double> full;
    if (mp->power == 1) {
      full.assign(src, src + d);
    } else {
      full.assign((size_t)d * d, 0.0);
      int64_t c = 0;
      for (int i = 0; i < d; i++)
        for (int j = i; j < d; j++) {
          if (i == j) full[(size_t)i * d + j] = src[c];
          else {
            double h = src[c] / 2;
            full[(size_t)i * d + j] = h;
            full[(size_t)j * d + i] = h;
          }
          c++;
        }
    }
    mp->beta.emplace_back(&ctx->pool, full.size() * sizeof(double));
    FM_HIP(hipMemcpyAsync(mp->beta.back().as<double>(), full.data(), full.size() * sizeof(double),
                          hipMemcpyHostToDevice, st));
    FM_HIP(hipStreamSynchronize(st));
  }
}

// PairFLARE::read_file (pair_flare.cpp:262-419) with grab (:427-439): header
// lines, then whitespace-separated cutoffs and beta values across lines.
void mapped_parse(const std::string &path, fm_mapped mp, std::vector<double> &beta) {
  std::ifstream f(path);
  FM_CHECK(f.good(), FM_ERR_INVALID, "Cannot open potential file " + path);
  std::string line;
  auto next = [&](const char *what) {
    FM_CHECK((bool)std::getline(f, line), FM_ERR_INVALID, std::string("truncated file: ") + what);
  };
  next("date line");
  next("kernel line");
  std::string kernel;
  {
    std::istringstream ls(line);
    ls >> mp->power >> kernel;
  }
  next("radial line");
  std::string radial;
  {
    std::istringstream ls(line);
    ls >> radial;
  }
  next("sizes line");
  int64_t beta_size = 0;
  {
    std::istringstream ls(line);
    ls >> mp->nos >> mp->N >> mp->lmax >> beta_size;
  }
  next("cutoff line");
  std::string cut;
  {
    std::istringstream ls(line);
    ls >> cut;
  }
  FM_CHECK(mp->nos >= 1 && mp->N >= 1 && mp->lmax >= 0, FM_ERR_INVALID, "bad descriptor sizes");
  mp->cutoffs.clear();
  double v;
  while ((int)mp->cutoffs.size() < mp->nos * mp->nos && f >> v) mp->cutoffs.push_back(v);
  FM_CHECK((int)mp->cutoffs.size() == mp->nos * mp->nos, FM_ERR_INVALID, "truncated cutoffs");
  int nr = mp->N * mp->nos;
  mp->nd = nr * (nr + 1) / 2 * (mp->lmax + 1);
  int64_t check;
  if (mp->power == 1) check = mp->nd;
  else if (mp->power == 2) check = (int64_t)mp->nd * (mp->nd + 1) / 2;
  else throw Error(FM_ERR_INVALID, "Power should be 1 or 2.");
  FM_CHECK(check == beta_size, FM_ERR_INVALID, "Beta size doesn't match the number of descriptors.");
  FM_CHECK(radial == "chebyshev", FM_ERR_INVALID, "Please use chebyshev radial basis function.");
  if (cut == "quadratic") mp->cutkind = FM_CUTOFF_QUADRATIC;
  else if (cut == "cosine") mp->cutkind = FM_CUTOFF_COSINE;
  else throw Error(FM_ERR_INVALID, "Please use quadratic or cosine cutoff function.");
  if (kernel == "NormalizedDotProduct") mp->normalized = 1;
  else if (kernel == "DotProduct") mp->normalized = 0;
  else throw Error(FM_ERR_INVALID, "Kernel string not recognized, expected <power> <kernel string>");
  beta.clear();
  beta.reserve((size_t)beta_size * mp->nos);
  while ((int64_t)beta.size() < beta_size * mp->nos && f >> v) beta.push_back(v);
  FM_CHECK((int64_t)beta.size() == beta_size * mp->nos, FM_ERR_INVALID, "truncated beta block");
}

}  // namespace fm
