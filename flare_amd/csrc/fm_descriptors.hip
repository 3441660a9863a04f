// Fused bond-basis kernels for the B2 descriptor (descriptors/b2.cpp:181-399)
// and its adjoint (the arithmetic of lammps_descriptor.cpp:250-315 and
// pair_flare.cpp:141-174, driven by dE/dB from the kernel contraction).
//
// One wavefront per atom; lane L owns the harmonics lm = L + 64 c. Per
// neighbour species the lane keeps c[n, lm] for all radial n in registers
// and accumulates g_n(r) Y_lm(r_hat) over that species' bonds in list order
// (so each coefficient sums its terms in the reference's order); the
// coefficients then sit in LDS for the power spectrum B[n1, n2, l] or the
// adjoint u = dE/dc. Dense dc/dr and dB/dr are never formed: the adjoint
// recomputes the bond basis and contracts it with u bond by bond.
#include <hip/hip_runtime.h>

#include <cmath>

#include "fm_device.hpp"
#include "fm_gemm.hpp"
#include "fm_internal.hpp"

namespace fm {

constexpr int kMaxHC = 2;  // harmonic chunks of 64 lanes (lmax <= 10)
constexpr int kAtomsPerBlock = 4;

struct B2Args {
  int n, nos, N, lmax, H, nrad, nd, cutkind;
  double r1;
  const int *species, *bond_off, *bond_q, *nspec, *row_of_atom;
  const double *rel, *cutoffs;
  const int2 *pairs;  // n1, n2 of each radial pair in B2 order
};

__constant__ double c_knorm[kMaxH];

template <typename T>
__device__ __forceinline__ void lane_lm(int lm, int &l, int &m) {
  l = (int)sqrtf((float)lm);
  while (l * l > lm) l--;
  while ((l + 1) * (l + 1) <= lm) l++;
  m = lm - l * l - l;
}

// c[(t N + n) H + lm] for all species t into LDS.
template <typename T>
__device__ void accumulate_c(const B2Args &a, int i, int lane, T *c) {
  int si = a.species[i];
  int b0 = a.bond_off[i], b1 = a.bond_off[i + 1];
  int ls[kMaxHC], ms[kMaxHC];
  T kn[kMaxHC];
  bool act[kMaxHC];
#pragma unroll
  for (int ch = 0; ch < kMaxHC; ch++) {
    int lm = lane + 64 * ch;
    act[ch] = lm < a.H;
    lane_lm<T>(act[ch] ? lm : 0, ls[ch], ms[ch]);
    kn[ch] = (T)c_knorm[act[ch] ? lm : 0];
  }
  for (int t = 0; t < a.nos; t++) {
    T acc[kMaxHC][kMaxN];
#pragma unroll
    for (int ch = 0; ch < kMaxHC; ch++)
#pragma unroll
      for (int nn = 0; nn < kMaxN; nn++) acc[ch][nn] = 0;
    for (int b = b0; b < b1; b++) {
      int q = a.bond_q[b];
      int sj = a.nspec[q];
      if (sj != t) continue;
      const double *rp = a.rel + 4 * (size_t)q;
      T r = (T)rp[0], x = (T)rp[1], y = (T)rp[2], z = (T)rp[3];
      T rc = (T)a.cutoffs[si * a.nos + sj];
      T g[kMaxN], dg[kMaxN];
      radial_basis<T>(r, rc, (T)a.r1, a.N, a.cutkind, g, dg);
      T rinv = T(1) / r;
      T ux = x * rinv, uy = y * rinv, uz = z * rinv;
#pragma unroll
      for (int ch = 0; ch < kMaxHC; ch++) {
        if (ch * 64 >= a.H) break;
        T Y, d0, d1, d2;
        sph_harm_lane<T, false>(ls[ch], ms[ch], kn[ch], ux, uy, uz, rinv, Y, d0, d1, d2);
#pragma unroll
        for (int nn = 0; nn < kMaxN; nn++) acc[ch][nn] += g[nn] * Y;
      }
    }
#pragma unroll
    for (int ch = 0; ch < kMaxHC; ch++) {
      if (!act[ch]) continue;
      int lm = lane + 64 * ch;
#pragma unroll
      for (int nn = 0; nn < kMaxN; nn++)
        if (nn < a.N) c[(t * a.N + nn) * a.H + lm] = acc[ch][nn];
    }
  }
}

// B2 (b2.cpp:209-231) per atom: B[k] = sum_m c[n1, lm] c[n2, lm], written at
// the atom's species-sorted row, with |B| and B.B. The FP32 form also writes
// each row as BF16 pairs (hi = bf16(B), lo = bf16(B - hi)) into Bh / Bl with
// leading dimension ldp and zero padding, the operand of the kernel
// contraction.
template <typename T>
__global__ __launch_bounds__(64 * kAtomsPerBlock) void k_b2(B2Args a, T *B, double *norm,
                                                              double *nn, __bf16 *Bh, __bf16 *Bl,
                                                              int ldp) {
  extern __shared__ __align__(16) unsigned char smem[];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int i = blockIdx.x * (blockDim.x >> 6) + w;
  if (i >= a.n) return;
  T *c = reinterpret_cast<T *>(smem) + (size_t)w * a.nrad * a.H;
  accumulate_c<T>(a, i, lane, c);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int row = a.row_of_atom[i];
  T *Brow = B + (size_t)row * a.nd;
  int L1 = a.lmax + 1;
  T s2 = 0;
  for (int k = lane; k < a.nd; k += 64) {
    int p = k / L1, l = k - p * L1;
    int2 pr = a.pairs[p];
    const T *c1 = c + pr.x * a.H + l * l, *c2 = c + pr.y * a.H + l * l;
    T acc = 0;
    for (int m = 0; m < 2 * l + 1; m++) acc += c1[m] * c2[m];
    Brow[k] = acc;
    if constexpr (sizeof(T) == 4) {
      __bf16 h = (__bf16)acc;
      Bh[(size_t)row * ldp + k] = h;
      Bl[(size_t)row * ldp + k] = (__bf16)(acc - (float)h);
    }
    s2 += acc * acc;
  }
  if constexpr (sizeof(T) == 4)
    for (int k = a.nd + lane; k < ldp; k += 64) {
      Bh[(size_t)row * ldp + k] = (__bf16)0.0f;
      Bl[(size_t)row * ldp + k] = (__bf16)0.0f;
    }
  s2 = wave_sum(s2);
  if (lane == 0) {
    norm[row] = (double)sqrt(s2);
    nn[row] = (double)s2;
  }
}

// Adjoint per atom. With `project` set, w = dE/dB = (G - (G.B_hat) B_hat)/|B|
// from G = dE/dB_hat |B| (for the sparse GP, G = sum_j sigma^2 alpha_j p
// S_j^(p-1) x_hat_j, the closed form of normalized_dot_product.cpp:191-195);
// without it w = G (unnormalised kernels). Then u = dE/dc
// (lammps_descriptor.cpp:286-314) and per bond f = dE_i/dr_ij =
// sum_{n,lm} u[s_j n, lm] grad(g_n Y_lm) (pair_flare.cpp:141-150). Writes f
// for every bond and the atom's virial -sum_b f (x) r_b (xx, xy, xz, yy, yz,
// zz). An atom with |B| < empty_norm or B.B < empty_nn contributes nothing.
struct AdjArgs {
  int project;
  double empty_norm, empty_nn;
};

template <typename T>
__global__ __launch_bounds__(64 * kAtomsPerBlock) void k_adjoint(B2Args a, const T *B,
                                                                   const double *norm,
                                                                   const double *nnv,
                                                                   const T *G, long long g_ld,
                                                                   AdjArgs aa, T *fbond,
                                                                   double *vir) {
  extern __shared__ __align__(16) unsigned char smem[];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int i = blockIdx.x * (blockDim.x >> 6) + w;
  if (i >= a.n) return;
  size_t per_wave = (size_t)2 * a.nrad * a.H + a.nd;
  T *c = reinterpret_cast<T *>(smem) + (size_t)w * per_wave;
  T *u = c + (size_t)a.nrad * a.H;
  T *wv = u + (size_t)a.nrad * a.H;
  int row = a.row_of_atom[i];
  int b0 = a.bond_off[i], b1 = a.bond_off[i + 1];
  double nrm = norm[row];
  if (nrm < aa.empty_norm || nnv[row] < aa.empty_nn) {
    // Empty environment: no force contribution (normalized_dot_product.cpp:171,
    // pair_flare.cpp:134-136).
    for (int b = b0 + lane; b < b1; b += 64) {
      fbond[3 * (size_t)b] = 0;
      fbond[3 * (size_t)b + 1] = 0;
      fbond[3 * (size_t)b + 2] = 0;
    }
    if (lane < 6) vir[6 * (size_t)i + lane] = 0;
    return;
  }
  accumulate_c<T>(a, i, lane, c);
  const T *Brow = B + (size_t)row * a.nd;
  const T *Grow = G + (size_t)row * g_ld;
  if (aa.project) {
    T ninv = T(1) / (T)nrm;
    T gb = 0;
    for (int k = lane; k < a.nd; k += 64) gb += Grow[k] * Brow[k];
    gb = wave_sum(gb) * ninv;
    for (int k = lane; k < a.nd; k += 64) wv[k] = (Grow[k] - gb * Brow[k] * ninv) * ninv;
  } else {
    for (int k = lane; k < a.nd; k += 64) wv[k] = Grow[k];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int L1 = a.lmax + 1, nr = a.nrad;
#pragma unroll
  for (int ch = 0; ch < kMaxHC; ch++) {
    int lm = lane + 64 * ch;
    if (lm >= a.H) continue;
    int l, m;
    lane_lm<T>(lm, l, m);
    for (int n1 = 0; n1 < nr; n1++) {
      T acc = 0;
      for (int n2 = 0; n2 < nr; n2++) {
        int lo = n1 < n2 ? n1 : n2, hi = n1 < n2 ? n2 : n1;
        int p = lo * nr - (lo * (lo - 1)) / 2 + (hi - lo);
        T wk = wv[p * L1 + l];
        T f = (n1 == n2) ? T(2) : T(1);
        acc += f * wk * c[n2 * a.H + lm];
      }
      u[n1 * a.H + lm] = acc;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int si = a.species[i];
  int ls[kMaxHC], ms[kMaxHC];
  T kn[kMaxHC];
#pragma unroll
  for (int ch = 0; ch < kMaxHC; ch++) {
    int lm = lane + 64 * ch;
    lane_lm<T>(lm < a.H ? lm : 0, ls[ch], ms[ch]);
    kn[ch] = (T)c_knorm[lm < a.H ? lm : 0];
  }
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int b = b0; b < b1; b++) {
    int q = a.bond_q[b];
    int sj = a.nspec[q];
    const double *rp = a.rel + 4 * (size_t)q;
    T r = (T)rp[0], x = (T)rp[1], y = (T)rp[2], z = (T)rp[3];
    T rc = (T)a.cutoffs[si * a.nos + sj];
    T g[kMaxN], dg[kMaxN];
    radial_basis<T>(r, rc, (T)a.r1, a.N, a.cutkind, g, dg);
    T rinv = T(1) / r;
    T ux = x * rinv, uy = y * rinv, uz = z * rinv;
    T fx = 0, fy = 0, fz = 0;
#pragma unroll
    for (int ch = 0; ch < kMaxHC; ch++) {
      int lm = lane + 64 * ch;
      if (ch * 64 >= a.H) break;
      if (lm < a.H) {
        T Y, Yx, Yy, Yz;
        sph_harm_lane<T, true>(ls[ch], ms[ch], kn[ch], ux, uy, uz, rinv, Y, Yx, Yy, Yz);
        const T *ub = u + (size_t)(sj * a.N) * a.H + lm;
        T s0 = 0, s1 = 0;
#pragma unroll
        for (int nn = 0; nn < kMaxN; nn++) {
          if (nn < a.N) {
            T uu = ub[(size_t)nn * a.H];
            s0 += uu * g[nn];
            s1 += uu * dg[nn];
          }
        }
        T radial = s1 * Y;
        fx += radial * ux + s0 * Yx;
        fy += radial * uy + s0 * Yy;
        fz += radial * uz + s0 * Yz;
      }
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane == 0) {
      fbond[3 * (size_t)b] = fx;
      fbond[3 * (size_t)b + 1] = fy;
      fbond[3 * (size_t)b + 2] = fz;
    }
    double f3[3] = {(double)fx, (double)fy, (double)fz};
    double r3[3] = {rp[1], rp[2], rp[3]};
    int sc = 0;
    for (int c1 = 0; c1 < 3; c1++)
      for (int c2 = c1; c2 < 3; c2++) v[sc++] -= f3[c1] * r3[c2];
  }
  if (lane == 0)
    for (int k = 0; k < 6; k++) vir[6 * (size_t)i + k] = v[k];
}

// Atoms (wavefronts) per block so that their LDS fits in 160 KB.
static int atoms_per_block(size_t lds_per_atom) {
  const size_t cap = 160 * 1024;
  FM_CHECK(lds_per_atom <= cap, FM_ERR_CAPACITY, "per-atom descriptor workspace exceeds LDS");
  int apb = kAtomsPerBlock;
  while (apb > 1 && apb * lds_per_atom > cap) apb >>= 1;
  return apb;
}

static B2Args make_args(fm_structure s) {
  B2Args a;
  a.n = s->n;
  a.nos = s->nos;
  a.N = s->N;
  a.lmax = s->lmax;
  a.H = (s->lmax + 1) * (s->lmax + 1);
  a.nrad = s->nos * s->N;
  a.nd = s->nd;
  a.cutkind = s->cutkind;
  a.r1 = s->r1;
  a.species = s->species.as<int>();
  a.bond_off = s->bond_off.as<int>();
  a.bond_q = s->bond_q.as<int>();
  a.nspec = s->nspecies.as<int>();
  a.row_of_atom = s->row_of_atom.as<int>();
  a.rel = s->rel.as<double>();
  a.cutoffs = s->d_cutoffs.as<double>();
  a.pairs = s->pair_table.as<int2>();
  return a;
}

void knorm_values(int lmax, double *out) {
  for (int l = 0; l <= lmax; l++)
    for (int m = -l; m <= l; m++) {
      int mm = m < 0 ? -m : m;
      double ratio = 1;
      for (int k = l - mm + 1; k <= l + mm; k++) ratio /= k;
      double K = std::sqrt((2 * l + 1) / (4 * 3.14159265358979323846) * ratio);
      if (mm > 0) K *= std::sqrt(2.0);
      out[l * l + m + l] = K;
    }
}

static void upload_knorm(hipStream_t st, int lmax) {
  static thread_local double host[kMaxH];
  knorm_values(lmax, host);
  FM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_knorm), host, sizeof(double) * (lmax + 1) * (lmax + 1),
                                0, hipMemcpyHostToDevice, st));
}

void launch_b2(fm_structure s) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  FM_CHECK(s->N <= kMaxN, FM_ERR_INVALID, "n_max above 16 is not supported");
  FM_CHECK(s->lmax <= kMaxL, FM_ERR_INVALID, "l_max above 10 is not supported");
  upload_knorm(st, s->lmax);
  B2Args a = make_args(s);
  size_t esz = s->form == FM_FORM_FP64 ? sizeof(double) : sizeof(float);
  size_t per = (size_t)a.nrad * a.H * esz;
  int apb = atoms_per_block(per);
  size_t lds = apb * per;
  int nblk = (s->n + apb - 1) / apb;
  if (nblk == 0) return;
  if (s->form == FM_FORM_FP64)
    hipLaunchKernelGGL(k_b2<double>, dim3(nblk), dim3(64 * apb), lds, st, a,
                       s->B.as<double>(), s->norm.as<double>(), s->nn.as<double>(),
                       (__bf16 *)nullptr, (__bf16 *)nullptr, 0);
  else
    hipLaunchKernelGGL(k_b2<float>, dim3(nblk), dim3(64 * apb), lds, st, a,
                       s->B.as<float>(), s->norm.as<double>(), s->nn.as<double>(),
                       s->Bh.as<__bf16>(), s->Bl.as<__bf16>(), (int)split_ld(s->nd));
  FM_HIP(hipGetLastError());
}

void launch_adjoint(fm_structure s, const void *G, int64_t g_ld, void *fbond,
                    double *virial_atom, int project, double empty_norm, double empty_nn) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  B2Args a = make_args(s);
  size_t esz = s->form == FM_FORM_FP64 ? sizeof(double) : sizeof(float);
  size_t per = ((size_t)2 * a.nrad * a.H + a.nd) * esz;
  int apb = atoms_per_block(per);
  size_t lds = apb * per;
  int nblk = (s->n + apb - 1) / apb;
  if (nblk == 0) return;
  AdjArgs aa{project, empty_norm, empty_nn};
  if (s->form == FM_FORM_FP64)
    hipLaunchKernelGGL(k_adjoint<double>, dim3(nblk), dim3(64 * apb), lds, st, a,
                       s->B.as<double>(), s->norm.as<double>(), s->nn.as<double>(),
                       (const double *)G, (long long)g_ld, aa, (double *)fbond, virial_atom);
  else
    hipLaunchKernelGGL(k_adjoint<float>, dim3(nblk), dim3(64 * apb), lds, st, a,
                       s->B.as<float>(), s->norm.as<double>(), s->nn.as<double>(),
                       (const float *)G, (long long)g_ld, aa, (float *)fbond, virial_atom);
  FM_HIP(hipGetLastError());
}

}  // namespace fm
