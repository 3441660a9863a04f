// Fused bond-basis kernels for the B2 descriptor (descriptors/b2.cpp:181-399)
// and its adjoint (the arithmetic of lammps_descriptor.cpp:250-315 and
// pair_flare.cpp:141-174, driven by dE/dB from the kernel contraction).
//
// Descriptor: one wavefront per atom; lane L owns the harmonics
// lm = L + 64 c. Per neighbour species the lane keeps c[n, lm] for all radial
// n in registers and accumulates g_n(r) Y_lm(r_hat) over that species' bonds
// in list order (so each coefficient sums its terms in the reference's
// order); the coefficients then sit in LDS for the power spectrum
// B[n1, n2, l] and are also written to HBM for the adjoint.
// Adjoint: u = dE/dc per atom (lanes over lm), then the bond forces with
// lanes over bonds, each lane sweeping all (l, m) of its bond by recursion.
// Dense dc/dr and dB/dr are never formed.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// 1 / k for the associated-Legendre recursions (k <= lmax + 1).
__constant__ double c_rcp[kMaxL + 2];

// Every real Y_lm (l <= L) of one unit vector u by an (m, l) sweep:
// Y_lm = K_lm Q_l^|m|(u_z) {Re, Im}(u_x + i u_y)^|m| (no Condon-Shortley
// phase; l = 1 gives y, z, x as y_grad.cpp:718-736) with Q_l^m = d^m P_l/dt^m
// from the upward recursion in l. Calls f(lm, Y, px, py, pz); with kGrad,
// (px, py, pz) is the gradient in u of the degree-0 extension
// (dQ_l^m/du_z = Q_l^(m+1)), so grad Y_lm(r) = (p - (u . p) u) / r.
template <typename T, bool kGrad, typename F>
__device__ __forceinline__ void ylm_sweep(int L, T ux, T uy, T uz, F &&f) {
  T cm = 1, sm = 0, cm1 = 0, sm1 = 0;  // (u_x + i u_y)^mm and ^(mm - 1)
  T dfact = 1;                         // (2 mm - 1)!!
  for (int mm = 0; mm <= L; mm++) {
    if (mm > 0) {
      cm1 = cm;
      sm1 = sm;
      cm = cm1 * ux - sm1 * uy;
      sm = sm1 * ux + cm1 * uy;
      dfact *= (T)(2 * mm - 1);
    }
    T q1 = dfact, q2 = 0, p1 = 0, p2 = 0;
    for (int l = mm; l <= L; l++) {
      T ql, pl = 0;
      if (l == mm) {
        ql = dfact;
      } else {
        ql = ((T)(2 * l - 1) * uz * q1 - (T)(l + mm - 1) * q2) * (T)c_rcp[l - mm];
        q2 = q1;
        q1 = ql;
        if (kGrad) {
          if (l == mm + 1)
            pl = dfact * (T)(2 * mm + 1);
          else
            pl = ((T)(2 * l - 1) * uz * p1 - (T)(l + mm) * p2) * (T)c_rcp[l - mm - 1];
          p2 = p1;
          p1 = pl;
        }
      }
      int base = l * l + l;
#pragma unroll
      for (int sgn = 0; sgn < 2; sgn++) {
        if (sgn == 1 && mm == 0) break;
        int lm = sgn ? base - mm : base + mm;
        T K = (T)c_knorm[lm];
        T trig, dtx = 0, dty = 0;
        if (mm == 0) {
          trig = 1;
        } else if (sgn == 0) {
          trig = cm;
          dtx = (T)mm * cm1;
          dty = -(T)mm * sm1;
        } else {
          trig = sm;
          dtx = (T)mm * sm1;
          dty = (T)mm * cm1;
        }
        T Kq = K * ql;
        if (kGrad)
          f(lm, Kq * trig, Kq * dtx, Kq * dty, K * pl * trig);
        else
          f(lm, Kq * trig, T(0), T(0), T(0));
      }
    }
  }
}

__device__ __forceinline__ float readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane(double v, int l) {
  long long x = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)x, l);
  int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// c[(t N + n) H + lm] of atom i for all neighbour species t into LDS
// (b2.cpp:367-395): for each chunk of up to 64 bonds, lane k takes bond k,
// evaluates its radial basis g_n (registers) and every Y_lm (into the LDS
// table Yt[k][lm]); then lanes over lm add g_n(k) Y_lm(k) over the chunk's
// bonds of species t in list order, reading g_n(k) from lane k.
template <typename T, int NG>
__device__ void accumulate_c(const B2Args &a, int i, int lane, T *c, T *Yt) {
  int si = a.species[i];
  int b0 = a.bond_off[i], b1 = a.bond_off[i + 1];
  int nchunks = (b1 - b0 + 63) / 64;
  T g[kMaxN], dg[kMaxN];
  int sj = -1;
  // NG lane groups of 64 / NG lanes each (NG > 1 when H <= 64 / NG): group
  // grp accumulates the radial functions n = grp, grp + NG, ... (slot nn of
  // acc holds n = nn NG + grp).
  constexpr int hp = 64 / NG;
  int grp = lane / hp, lmb = lane % hp;
  for (int t = 0; t < a.nos; t++) {
    T acc[kMaxHC][kMaxN];
#pragma unroll
    for (int ch = 0; ch < kMaxHC; ch++)
#pragma unroll
      for (int nn = 0; nn < kMaxN; nn++) acc[ch][nn] = 0;
    for (int ck = 0; ck < nchunks; ck++) {
      int bb = b0 + 64 * ck;
      int nk = min(64, b1 - bb);
      if (nchunks > 1 || t == 0) {
        if (ck > 0 || t > 0) {
          // Lanes still read the previous chunk's table.
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (lane < nk) {
          int q = a.bond_q[bb + lane];
          sj = a.nspec[q];
          const double *rp = a.rel + 4 * (size_t)q;
          T r = (T)rp[0];
          T rc = (T)a.cutoffs[si * a.nos + sj];
          radial_basis<T>(r, rc, (T)a.r1, a.N, a.cutkind, g, dg);
          T rinv = T(1) / r;
          T *yrow = Yt + (size_t)lane * a.H;
          ylm_sweep<T, false>(a.lmax, (T)rp[1] * rinv, (T)rp[2] * rinv, (T)rp[3] * rinv,
                              [&](int lm, T Y, T, T, T) { yrow[lm] = Y; });
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      for (int k = 0; k < nk; k++) {
        if (__builtin_amdgcn_readlane(sj, k) != t) continue;
        T y[kMaxHC];
#pragma unroll
        for (int ch = 0; ch < kMaxHC; ch++) {
          int lm = lmb + 64 * ch;
          y[ch] = lm < a.H ? Yt[(size_t)k * a.H + lm] : T(0);
        }
#pragma unroll
        for (int nn = 0; nn < kMaxN / NG; nn++) {
          if (nn * NG < a.N) {
            T gk = 0;
#pragma unroll
            for (int gg = 0; gg < NG; gg++) {
              T v = readlane(g[nn * NG + gg], k);
              if (gg == grp) gk = v;
            }
#pragma unroll
            for (int ch = 0; ch < kMaxHC; ch++) acc[ch][nn] += gk * y[ch];
          }
        }
      }
    }
#pragma unroll
    for (int ch = 0; ch < kMaxHC; ch++) {
      int lm = lmb + 64 * ch;
      if (lm >= a.H) continue;
#pragma unroll
      for (int nn = 0; nn < kMaxN / NG; nn++) {
        int n = nn * NG + grp;
        if (n < a.N) c[(t * a.N + n) * a.H + lm] = acc[ch][nn];
      }
    }
  }
}

// B2 (b2.cpp:209-231) per atom: B[k] = sum_m c[n1, lm] c[n2, lm], written at
// the atom's species-sorted row, with |B| and B.B. The FP32 form also writes
// each row as two BF16 pieces (b0 = bf16(B), b1 = bf16(B - b0)) with leading
// dimension ldp and zero padding, the operand of the kernel contraction. The
// coefficients c also go to HBM (Cg) for the adjoint.
struct SplitOut {
  __bf16 *p[2];
};

template <typename T, int NG>
__global__ __launch_bounds__(64 * kAtomsPerBlock) void k_b2(B2Args a, T *B, double *norm,
                                                              double *nn, SplitOut Bs, int ldp,
                                                              T *Cg) {
  extern __shared__ __align__(16) unsigned char smem[];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int i = blockIdx.x * (blockDim.x >> 6) + w;
  if (i >= a.n) return;
  int nc = a.nrad * a.H;
  T *c = reinterpret_cast<T *>(smem) + (size_t)w * ((size_t)nc + 64 * a.H);
  T *Yt = c + nc;
  accumulate_c<T, NG>(a, i, lane, c, Yt);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int k = lane; k < nc; k += 64) Cg[(size_t)i * nc + k] = c[k];
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
      float r = acc;
#pragma unroll
      for (int pc = 0; pc < 2; pc++) {
        __bf16 h = (__bf16)r;
        Bs.p[pc][(size_t)row * ldp + k] = h;
        r -= (float)h;
      }
    }
    s2 += acc * acc;
  }
  if constexpr (sizeof(T) == 4)
    for (int k = a.nd + lane; k < ldp; k += 64)
#pragma unroll
      for (int pc = 0; pc < 2; pc++) Bs.p[pc][(size_t)row * ldp + k] = (__bf16)0.0f;
  s2 = wave_sum(s2);
  if (lane == 0) {
    norm[row] = (double)sqrt(s2);
    nn[row] = (double)s2;
  }
}

// Adjoint, first half, per atom (one wavefront): with `project` set,
// w = dE/dB = (G - (G.B_hat) B_hat)/|B| from G = dE/dB_hat |B| (for the
// sparse GP, G = sum_j sigma^2 alpha_j p S_j^(p-1) x_hat_j, the closed form of
// normalized_dot_product.cpp:191-195); without it w = G (unnormalised
// kernels). Then u = dE/dc (lammps_descriptor.cpp:286-314):
// u[n1, lm] = sum_n2 (1 + [n1 = n2]) w[(min, max)(n1, n2), l] c[n2, lm], lanes
// over lm, four n1 per pass. An atom with |B| < empty_norm or
// B.B < empty_nn contributes nothing (its u is not read).
struct AdjArgs {
  int project;
  double empty_norm, empty_nn;
};

__device__ __forceinline__ int pair_index(int lo, int hi, int nr) {
  return lo * nr - (lo * (lo - 1)) / 2 + (hi - lo);
}

template <typename T>
__global__ __launch_bounds__(64 * kAtomsPerBlock) void k_dEdc(B2Args a, const T *B,
                                                                const double *norm,
                                                                const double *nnv, const T *G,
                                                                long long g_ld, AdjArgs aa,
                                                                const T *Cg, T *Ug) {
  extern __shared__ __align__(16) unsigned char smem[];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int i = blockIdx.x * (blockDim.x >> 6) + w;
  if (i >= a.n) return;
  int nc = a.nrad * a.H;
  T *c = reinterpret_cast<T *>(smem) + (size_t)w * ((size_t)nc + a.nd);
  T *wv = c + nc;
  int row = a.row_of_atom[i];
  double nrm = norm[row];
  if (nrm < aa.empty_norm || nnv[row] < aa.empty_nn) return;
  for (int k = lane; k < nc; k += 64) c[k] = Cg[(size_t)i * nc + k];
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
  T *urow = Ug + (size_t)i * nc;
  // Lane groups: with H <= 32 the wave splits into ngrp = 64 / hp groups of
  // hp lanes (hp = 16 or 32), group grp taking rows n1 = grp, grp + ngrp, ...
  int hp = a.H <= 16 ? 16 : (a.H <= 32 ? 32 : 64);
  int ngrp = 64 / hp, grp = lane / hp;
#pragma unroll
  for (int ch = 0; ch < kMaxHC; ch++) {
    int lm = (lane % hp) + 64 * ch;
    if (lm >= a.H) continue;
    int l, m;
    lane_lm<T>(lm, l, m);
    const T *wl = wv + l;   // w[p, l] = wl[p L1]
    const T *cl = c + lm;   // c[n, lm] = cl[n H]
    // Pair p(lo, hi) = base(lo) + hi - lo, base(lo) = lo nr - lo (lo - 1) / 2.
    for (int n1 = grp; n1 < nr; n1 += ngrp) {
      T acc = 0;
      int base = 0;
      for (int n2 = 0; n2 < n1; n2++) {
        acc += wl[(base + n1 - n2) * L1] * cl[n2 * a.H];
        base += nr - n2;
      }
      acc += T(2) * wl[base * L1] * cl[n1 * a.H];
      for (int n2 = n1 + 1; n2 < nr; n2++) acc += wl[(base + n2 - n1) * L1] * cl[n2 * a.H];
      urow[n1 * a.H + lm] = acc;
    }
  }
}

// Adjoint, second half: per bond f = dE_i/dr_ij = sum_{n,lm} u[s_j n, lm]
// grad(g_n Y_lm) (pair_flare.cpp:141-150, the adjoint of b2.cpp:219-228).
// One wavefront per atom, lanes over its bonds; each lane sweeps every
// (l, m) of its bond (ylm_sweep with gradients) and contracts with u staged
// in LDS:
//   f = a u_hat + (P - (u_hat . P) u_hat) / r,
//   a = sum_lm (sum_n u dg_n) Y_lm,  P = sum_lm (sum_n u g_n) grad_u Y_lm,
// with grad_u the gradient of the degree-0 extension in u. Writes f for every
// bond and the atom's virial -sum_b f (x) r_b (xx, xy, xz, yy, yz, zz).
template <typename T>
__global__ __launch_bounds__(64 * kAtomsPerBlock) void k_bond_forces(B2Args a,
                                                                       const double *norm,
                                                                       const double *nnv,
                                                                       AdjArgs aa, const T *Ug,
                                                                       T *fbond, double *vir) {
  extern __shared__ __align__(16) unsigned char smem[];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int i = blockIdx.x * (blockDim.x >> 6) + w;
  if (i >= a.n) return;
  int nc = a.nrad * a.H;
  T *u = reinterpret_cast<T *>(smem) + (size_t)w * nc;
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
  for (int k = lane; k < nc; k += 64) u[k] = Ug[(size_t)i * nc + k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int si = a.species[i];
  int L = a.lmax, H = a.H, N = a.N;
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int bb = b0; bb < b1; bb += 64) {
    int b = bb + lane;
    if (b < b1) {
      int q = a.bond_q[b];
      int sj = a.nspec[q];
      const double *rp = a.rel + 4 * (size_t)q;
      T r = (T)rp[0], x = (T)rp[1], y = (T)rp[2], z = (T)rp[3];
      T rc = (T)a.cutoffs[si * a.nos + sj];
      T g[kMaxN], dg[kMaxN];
      radial_basis<T>(r, rc, (T)a.r1, N, a.cutkind, g, dg);
      T rinv = T(1) / r;
      T ux = x * rinv, uy = y * rinv, uz = z * rinv;
      const T *ub = u + (size_t)sj * N * H;
      T acc_a = 0, Px = 0, Py = 0, Pz = 0;
      ylm_sweep<T, true>(L, ux, uy, uz, [&](int lm, T Y, T px, T py, T pz) {
        T s0 = 0, s1 = 0;
#pragma unroll
        for (int nn = 0; nn < kMaxN; nn++) {
          if (nn < N) {
            T uu = ub[nn * H + lm];
            s0 += uu * g[nn];
            s1 += uu * dg[nn];
          }
        }
        acc_a += s1 * Y;
        Px += s0 * px;
        Py += s0 * py;
        Pz += s0 * pz;
      });
      T up = ux * Px + uy * Py + uz * Pz;
      T fx = acc_a * ux + (Px - up * ux) * rinv;
      T fy = acc_a * uy + (Py - up * uy) * rinv;
      T fz = acc_a * uz + (Pz - up * uz) * rinv;
      fbond[3 * (size_t)b] = fx;
      fbond[3 * (size_t)b + 1] = fy;
      fbond[3 * (size_t)b + 2] = fz;
      double f3[3] = {(double)fx, (double)fy, (double)fz};
      double r3[3] = {rp[1], rp[2], rp[3]};
      int sc = 0;
      for (int c1 = 0; c1 < 3; c1++)
        for (int c2 = c1; c2 < 3; c2++) v[sc++] -= f3[c1] * r3[c2];
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = wave_sum(v[k]);
  if (lane == 0)
    for (int k = 0; k < 6; k++) vir[6 * (size_t)i + k] = v[k];
}

// Atoms (wavefronts) per block: of 1, 2 and 4, the one that keeps the most
// atoms resident on a CU within its 160 KB of LDS (larger on ties).
static int atoms_per_block(size_t lds_per_atom) {
  const size_t cap = 160 * 1024;
  FM_CHECK(lds_per_atom <= cap, FM_ERR_CAPACITY, "per-atom descriptor workspace exceeds LDS");
  int best = 1;
  size_t best_res = 0;
  for (int apb = 1; apb <= kAtomsPerBlock; apb <<= 1) {
    size_t blocks = cap / (apb * lds_per_atom);
    size_t res = std::min<size_t>(blocks, 8) * apb;  // at most 8 blocks of ours per CU here
    if (blocks > 0 && res >= best_res) {
      best = apb;
      best_res = res;
    }
  }
  return best;
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

// The whole table (l <= 10) is written every time: entries do not depend on
// the descriptor's lmax, so uploads from several streams write equal values.
static void upload_knorm(hipStream_t st, int) {
  static thread_local double host[kMaxH];
  knorm_values(kMaxL, host);
  FM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_knorm), host, sizeof(host), 0, hipMemcpyHostToDevice,
                                st));
}

static void upload_rcp(hipStream_t st) {
  static thread_local double host[kMaxL + 2];
  host[0] = 0;
  for (int k = 1; k < kMaxL + 2; k++) host[k] = 1.0 / k;
  FM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_rcp), host, sizeof(host), 0, hipMemcpyHostToDevice,
                                st));
}

void launch_b2(fm_structure s) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  FM_CHECK(s->N <= kMaxN, FM_ERR_INVALID, "n_max above 16 is not supported");
  FM_CHECK(s->lmax <= kMaxL, FM_ERR_INVALID, "l_max above 10 is not supported");
  upload_knorm(st, s->lmax);
  upload_rcp(st);
  B2Args a = make_args(s);
  size_t esz = s->form == FM_FORM_FP64 ? sizeof(double) : sizeof(float);
  size_t per = ((size_t)a.nrad * a.H + 64 * (size_t)a.H) * esz;  // c and the Y table
  int apb = atoms_per_block(per);
  size_t lds = apb * per;
  int nblk = (s->n + apb - 1) / apb;
  s->C.alloc(&ctx->pool, std::max<size_t>(1, (size_t)s->n * a.nrad * a.H * esz));
  if (nblk == 0) return;
  int ng = a.H <= 16 ? 4 : (a.H <= 32 ? 2 : 1);
  SplitOut none{{nullptr, nullptr}};
  SplitOut pieces{{s->Bs[0].as<__bf16>(), s->Bs[1].as<__bf16>()}};
  int ldp = (int)split_ld(s->nd);
  if (s->form == FM_FORM_FP64) {
    double *Bp = s->B.as<double>(), *Cp = s->C.as<double>();
    if (ng == 4)
      hipLaunchKernelGGL((k_b2<double, 4>), dim3(nblk), dim3(64 * apb), lds, st, a, Bp,
                         s->norm.as<double>(), s->nn.as<double>(), none, 0, Cp);
    else if (ng == 2)
      hipLaunchKernelGGL((k_b2<double, 2>), dim3(nblk), dim3(64 * apb), lds, st, a, Bp,
                         s->norm.as<double>(), s->nn.as<double>(), none, 0, Cp);
    else
      hipLaunchKernelGGL((k_b2<double, 1>), dim3(nblk), dim3(64 * apb), lds, st, a, Bp,
                         s->norm.as<double>(), s->nn.as<double>(), none, 0, Cp);
  } else {
    float *Bp = s->B.as<float>(), *Cp = s->C.as<float>();
    if (ng == 4)
      hipLaunchKernelGGL((k_b2<float, 4>), dim3(nblk), dim3(64 * apb), lds, st, a, Bp,
                         s->norm.as<double>(), s->nn.as<double>(), pieces, ldp, Cp);
    else if (ng == 2)
      hipLaunchKernelGGL((k_b2<float, 2>), dim3(nblk), dim3(64 * apb), lds, st, a, Bp,
                         s->norm.as<double>(), s->nn.as<double>(), pieces, ldp, Cp);
    else
      hipLaunchKernelGGL((k_b2<float, 1>), dim3(nblk), dim3(64 * apb), lds, st, a, Bp,
                         s->norm.as<double>(), s->nn.as<double>(), pieces, ldp, Cp);
  }
  FM_HIP(hipGetLastError());
}

void launch_adjoint(fm_structure s, const void *G, int64_t g_ld, void *fbond,
                    double *virial_atom, int project, double empty_norm, double empty_nn) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  B2Args a = make_args(s);
  FM_CHECK(a.lmax + 1 < kMaxL + 2, FM_ERR_INVALID, "l_max above 10 is not supported");
  upload_knorm(st, s->lmax);
  upload_rcp(st);
  bool f64 = s->form == FM_FORM_FP64;
  size_t esz = f64 ? sizeof(double) : sizeof(float);
  size_t nc = (size_t)a.nrad * a.H;
  DevBuf U(&ctx->pool, std::max<size_t>(1, (size_t)s->n * nc * esz));
  AdjArgs aa{project, empty_norm, empty_nn};
  // dE/dc: c and w in LDS per atom.
  size_t per1 = (nc + a.nd) * esz;
  int apb1 = atoms_per_block(per1);
  int nblk1 = (s->n + apb1 - 1) / apb1;
  // Bond forces: u in LDS per atom.
  size_t per2 = nc * esz;
  int apb2 = atoms_per_block(per2);
  int nblk2 = (s->n + apb2 - 1) / apb2;
  if (s->n == 0) return;
  if (f64) {
    hipLaunchKernelGGL(k_dEdc<double>, dim3(nblk1), dim3(64 * apb1), apb1 * per1, st, a,
                       s->B.as<double>(), s->norm.as<double>(), s->nn.as<double>(),
                       (const double *)G, (long long)g_ld, aa, s->C.as<double>(),
                       U.as<double>());
    hipLaunchKernelGGL(k_bond_forces<double>, dim3(nblk2), dim3(64 * apb2), apb2 * per2, st, a,
                       s->norm.as<double>(), s->nn.as<double>(), aa, U.as<double>(),
                       (double *)fbond, virial_atom);
  } else {
    hipLaunchKernelGGL(k_dEdc<float>, dim3(nblk1), dim3(64 * apb1), apb1 * per1, st, a,
                       s->B.as<float>(), s->norm.as<double>(), s->nn.as<double>(),
                       (const float *)G, (long long)g_ld, aa, s->C.as<float>(), U.as<float>());
    hipLaunchKernelGGL(k_bond_forces<float>, dim3(nblk2), dim3(64 * apb2), apb2 * per2, st, a,
                       s->norm.as<double>(), s->nn.as<double>(), aa, U.as<float>(),
                       (float *)fbond, virial_atom);
  }
  FM_HIP(hipGetLastError());
}

}  // namespace fm
