// Kernel contraction on the matrix cores. Every contraction of the path has
// the form C = A . B^T with both operands contiguous along the reduction
// index; a workgroup of 4 wavefronts (2 x 2) computes a tile, each wave
// TM x TN MFMA tiles of 16 x 16, with a fused epilogue:
//
//   EpKernel  C = B_s . X_hat_s^T (atoms x sparse): S = C / |B_i|,
//             k = sigma^2 S^p, w = sigma^2 alpha_j p S^(p-1); writes w and
//             (optionally) k atom-major and per-tile partial energies
//             sum_j k alpha_j (normalized_dot_product.cpp:117-177 contracted
//             with alpha as sparse_gp.cpp:664).
//   EpStore   plain store: G = W . X_hat_s (the kernel-gradient contraction).
//   EpSumSq   Q = L_s^-1 . K^T with per-tile column sums of Q^2, the local
//             variance term |L_inv k|^2 (sparse_gp.cpp:121-131); A is lower
//             triangular so each row tile stops at its last nonzero column.
//
// FP64 form: v_mfma_f64_16x16x4f64, 64 x 64 tiles. Chosen form: split-BF16
// x3 on v_mfma_f32_16x16x32_bf16 (a = a_hi + a_lo, ab ~ a_hi b_hi +
// a_hi b_lo + a_lo b_hi, FP32 accumulation), 128 x 128 tiles, on operands
// split once where they are produced; the three products of a k-step run
// over all 16 accumulators of a wave before the next, so consecutive MFMAs
// are independent.
#include <hip/hip_runtime.h>

#include <cstring>

#include "fm_device.hpp"
#include "fm_gemm.hpp"
#include "fm_internal.hpp"

namespace fm {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Epilogues. Lane l holds column l % 16 of each 16 x 16 output tile; the
// row of its value i (i < 4) is 4 (l / 16) + i for the FP32-output MFMAs and
// l / 16 + 4 i for v_mfma_f64_16x16x4f64.
// ---------------------------------------------------------------------------
template <typename Acc, int TM, int TN>
struct TileOut {
  Acc v[TM][TN][4];
};

template <typename Acc>
__device__ __forceinline__ int tile_row(int lane, int i) {
  return sizeof(Acc) == 8 ? (lane >> 4) + 4 * i : 4 * (lane >> 4) + i;
}

__device__ __forceinline__ double fpow(double x, double p, int pi) {
  if (pi == 2) return x * x;
  if (pi == 1) return x;
  return pow(x, p);
}
__device__ __forceinline__ float fpow(float x, float p, int pi) {
  if (pi == 2) return x * x;
  if (pi == 1) return x;
  return powf(x, p);
}

template <typename Acc, typename TW, int TM, int TN>
__device__ __forceinline__ void ep_kernel(const EpArgs &e, const TileOut<Acc, TM, TN> &t, int row0, int col0,
                          int M, int N, int lane, int wr, int wc, double *red) {
  constexpr int BMc = 32 * TM;
  // rows: atoms of this species block, cols: sparse environments.
  int pi = e.power == 2.0 ? 2 : (e.power == 1.0 ? 1 : 0);
  double rs[TM][4];
#pragma unroll
  for (int ti = 0; ti < TM; ti++)
#pragma unroll
    for (int i = 0; i < 4; i++) rs[ti][i] = 0;
#pragma unroll
  for (int ti = 0; ti < TM; ti++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int r = row0 + wr * 16 * TM + ti * 16 + tile_row<Acc>(lane, i);
      if (r >= M) continue;
      double nrm = e.row_norm[r];
      bool empty = nrm < kEmptyThresh;
      Acc ninv = empty ? Acc(0) : Acc(1.0 / nrm);
#pragma unroll
      for (int tj = 0; tj < TN; tj++) {
        int c = col0 + wc * 16 * TN + tj * 16 + (lane & 15);
        if (c >= N) continue;
        Acc s = t.v[ti][tj][i] * ninv;
        Acc a = (Acc)e.alpha[c];
        Acc kv = empty ? Acc(0) : Acc(e.sig2) * fpow(s, (Acc)e.power, pi);
        Acc pm1 = pi == 2 ? s : (pi == 1 ? Acc(1) : fpow(s, (Acc)(e.power - 1), 0));
        Acc wv = empty ? Acc(0) : Acc(e.sig2) * a * (Acc)e.power * pm1;
        if constexpr (sizeof(Acc) == 4) {  // split-BF16 form: w as BF16 pairs
          float wf = (float)wv;
          __bf16 h = (__bf16)wf;
          reinterpret_cast<__bf16 *>(e.Wh)[(size_t)r * e.ldw + c] = h;
          reinterpret_cast<__bf16 *>(e.Wl)[(size_t)r * e.ldw + c] = (__bf16)(wf - (float)h);
        } else {
          e.W_as<TW>()[(size_t)r * e.ldw + c] = (TW)wv;
        }
        if (e.K) e.K[(size_t)r * e.ldk + c] = (double)kv;
        rs[ti][i] += (double)(kv * a);
      }
    }
  if constexpr (sizeof(Acc) == 4) {
    // Padding columns [N, ldw) of the split output are zero: they are the
    // reduction tail of the next contraction.
#pragma unroll
    for (int ti = 0; ti < TM; ti++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int r = row0 + wr * 16 * TM + ti * 16 + tile_row<Acc>(lane, i);
        if (r >= M) continue;
#pragma unroll
        for (int tj = 0; tj < TN; tj++) {
          int c = col0 + wc * 16 * TN + tj * 16 + (lane & 15);
          if (c >= N && c < e.ldw) {
            reinterpret_cast<__bf16 *>(e.Wh)[(size_t)r * e.ldw + c] = (__bf16)0.0f;
            reinterpret_cast<__bf16 *>(e.Wl)[(size_t)r * e.ldw + c] = (__bf16)0.0f;
          }
        }
      }
  }
  // Sum over the 16 columns of each lane group, then over the two wave
  // columns through LDS; one partial per (column tile, row).
#pragma unroll
  for (int ti = 0; ti < TM; ti++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      double v = rs[ti][i];
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
      rs[ti][i] = v;
    }
  if ((lane & 15) == 0) {
#pragma unroll
    for (int ti = 0; ti < TM; ti++)
#pragma unroll
      for (int i = 0; i < 4; i++)
        red[wc * BMc + wr * 16 * TM + ti * 16 + tile_row<Acc>(lane, i)] = rs[ti][i];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < BMc; q += blockDim.x) {
    int r = row0 + q;
    if (r < M) e.part[(size_t)blockIdx.x * e.part_ld + r] = red[q] + red[BMc + q];
  }
}

template <typename Acc, typename TW, int TM, int TN>
__device__ __forceinline__ void ep_store(const EpArgs &e, const TileOut<Acc, TM, TN> &t, int row0, int col0,
                         int M, int N, int lane, int wr, int wc) {
#pragma unroll
  for (int ti = 0; ti < TM; ti++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int r = row0 + wr * 16 * TM + ti * 16 + tile_row<Acc>(lane, i);
      if (r >= M) continue;
#pragma unroll
      for (int tj = 0; tj < TN; tj++) {
        int c = col0 + wc * 16 * TN + tj * 16 + (lane & 15);
        if (c < N) e.W_as<TW>()[(size_t)r * e.ldw + c] = (TW)t.v[ti][tj][i];
      }
    }
}

template <typename Acc, int TM, int TN>
__device__ __forceinline__ void ep_sumsq(const EpArgs &e, const TileOut<Acc, TM, TN> &t, int row0, int col0,
                         int M, int N, int lane, int wr, int wc, double *red) {
  constexpr int BNc = 32 * TN;
  // rows: L rows, cols: atoms. Column sums of squares over this row tile.
  double cs[TN];
#pragma unroll
  for (int tj = 0; tj < TN; tj++) {
    cs[tj] = 0;
#pragma unroll
    for (int ti = 0; ti < TM; ti++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int r = row0 + wr * 16 * TM + ti * 16 + tile_row<Acc>(lane, i);
        double v = r < M ? (double)t.v[ti][tj][i] : 0.0;
        cs[tj] += v * v;
      }
  }
#pragma unroll
  for (int tj = 0; tj < TN; tj++) {
    double v = cs[tj];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    cs[tj] = v;
  }
  if (lane < 16) {
#pragma unroll
    for (int tj = 0; tj < TN; tj++) red[wr * BNc + wc * 16 * TN + tj * 16 + lane] = cs[tj];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < BNc; q += blockDim.x) {
    int c = col0 + q;
    if (c < N) e.part[(size_t)blockIdx.y * e.part_ld + c] = red[q] + red[BNc + q];
  }
}

template <int EP, typename Acc, typename TW, int TM, int TN>
__device__ __forceinline__ void run_epilogue(const EpArgs &e, const TileOut<Acc, TM, TN> &t,
                                             int row0, int col0, int M, int N, int lane, int wr,
                                             int wc, double *red) {
  if (EP == EP_KERNEL) ep_kernel<Acc, TW, TM, TN>(e, t, row0, col0, M, N, lane, wr, wc, red);
  else if (EP == EP_STORE) ep_store<Acc, TW, TM, TN>(e, t, row0, col0, M, N, lane, wr, wc);
  else ep_sumsq<Acc, TM, TN>(e, t, row0, col0, M, N, lane, wr, wc, red);
}

// ---------------------------------------------------------------------------
// FP64 MFMA GEMM, 64 x 64 tiles. Grid: x = column tiles (rows of B),
// y = row tiles (rows of A). LDS holds k-major tiles so a fragment read
// (16 rows at one k) is contiguous.
// ---------------------------------------------------------------------------
constexpr int F64_TM = 2, F64_TN = 2, F64_BM = 32 * F64_TM, F64_BN = 32 * F64_TN;
constexpr int BK64 = 16;
constexpr int PAD64 = 4;

template <int EP>
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, const double *A, long long lda,
                                                  const double *B, long long ldb, int lower_tri,
                                                  EpArgs e) {
  __shared__ double As[BK64][F64_BM + PAD64];
  __shared__ double Bs[BK64][F64_BN + PAD64];
  __shared__ double red[2 * (F64_BM > F64_BN ? F64_BM : F64_BN)];
  int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int wr = wid >> 1, wc = wid & 1;
  int row0 = blockIdx.y * F64_BM, col0 = blockIdx.x * F64_BN;
  int kend = K;
  if (lower_tri) kend = min(K, row0 + F64_BM);
  v4d acc[F64_TM][F64_TN];
#pragma unroll
  for (int a = 0; a < F64_TM; a++)
#pragma unroll
    for (int b = 0; b < F64_TN; b++) acc[a][b] = v4d{0, 0, 0, 0};
  int lr = tid >> 2, lk = (tid & 3) * 4;
  double ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int k = k0 + lk + q;
      int r = row0 + lr, c = col0 + lr;
      ra[q] = (r < M && k < kend) ? A[(size_t)r * lda + k] : 0.0;
      rb[q] = (c < N && k < kend) ? B[(size_t)c * ldb + k] : 0.0;
    }
  };
  gload(0);
  for (int k0 = 0; k0 < kend; k0 += BK64) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      As[lk + q][lr] = ra[q];
      Bs[lk + q][lr] = rb[q];
    }
    __syncthreads();
    if (k0 + BK64 < kend) gload(k0 + BK64);
#pragma unroll
    for (int kk = 0; kk < BK64; kk += 4) {
      double af[F64_TM], bf[F64_TN];
#pragma unroll
      for (int t = 0; t < F64_TM; t++) af[t] = As[kk + (lane >> 4)][wr * 16 * F64_TM + t * 16 + (lane & 15)];
#pragma unroll
      for (int t = 0; t < F64_TN; t++) bf[t] = Bs[kk + (lane >> 4)][wc * 16 * F64_TN + t * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < F64_TM; a++)
#pragma unroll
        for (int b = 0; b < F64_TN; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
  }
  TileOut<double, F64_TM, F64_TN> t;
#pragma unroll
  for (int a = 0; a < F64_TM; a++)
#pragma unroll
    for (int b = 0; b < F64_TN; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) t.v[a][b][i] = acc[a][b][i];
  run_epilogue<EP, double, double, F64_TM, F64_TN>(e, t, row0, col0, M, N, lane, wr, wc, red);
}

// ---------------------------------------------------------------------------
// Split-BF16 x3 GEMM (chosen form), 128 x 128 tiles, on operands stored as
// BF16 pairs: x = hi + lo with hi = bf16(x), lo = bf16(x - hi) of the FP32
// value, split once where the operand is produced (the sparse set at model
// upload, descriptors in the descriptor kernel, w in the kernel-contraction
// epilogue). Each thread copies 16 consecutive k of one row of A and of B
// per k-step as 16-byte vectors into k-contiguous LDS rows; lane l then
// reads 8 consecutive k at row l % 16, k offset 8 (l / 16).
// ---------------------------------------------------------------------------
constexpr int B16_TM = 4, B16_TN = 4, B16_BM = 32 * B16_TM, B16_BN = 32 * B16_TN;
constexpr int BK16 = 32;
constexpr int PADB = 8;
constexpr int LDK = BK16 + PADB;
static_assert(BK16 == kSplitPad, "k-step equals the operand padding");

template <int EP>
__global__ __launch_bounds__(256) void k_gemm_bf16x3(int M, int N, int Kpad, const __bf16 *Ahg,
                                                     const __bf16 *Alg, long long lda,
                                                     const __bf16 *Bhg, const __bf16 *Blg,
                                                     long long ldb, EpArgs e) {
  __shared__ __align__(16) __bf16 Ah[B16_BM * LDK];
  __shared__ __align__(16) __bf16 Al[B16_BM * LDK];
  __shared__ __align__(16) __bf16 Bh[B16_BN * LDK];
  __shared__ __align__(16) __bf16 Bl[B16_BN * LDK];
  __shared__ double red[2 * (B16_BM > B16_BN ? B16_BM : B16_BN)];
  int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int wr = wid >> 1, wc = wid & 1;
  int row0 = blockIdx.y * B16_BM, col0 = blockIdx.x * B16_BN;
  v4f acc[B16_TM][B16_TN];
#pragma unroll
  for (int a = 0; a < B16_TM; a++)
#pragma unroll
    for (int b = 0; b < B16_TN; b++) acc[a][b] = v4f{0, 0, 0, 0};
  // 128 rows x 32 k per operand: thread -> row tid / 2, k half (tid & 1) * 16.
  int lr = tid >> 1, lk = (tid & 1) * 16;
  bool va = row0 + lr < M, vb = col0 + lr < N;
  const v8bf *pah = reinterpret_cast<const v8bf *>(Ahg + (size_t)(va ? row0 + lr : 0) * lda + lk);
  const v8bf *pal = reinterpret_cast<const v8bf *>(Alg + (size_t)(va ? row0 + lr : 0) * lda + lk);
  const v8bf *pbh = reinterpret_cast<const v8bf *>(Bhg + (size_t)(vb ? col0 + lr : 0) * ldb + lk);
  const v8bf *pbl = reinterpret_cast<const v8bf *>(Blg + (size_t)(vb ? col0 + lr : 0) * ldb + lk);
  const v8bf zero = {};
  v8bf ra[4], rb[4];
  auto gload = [&](int k0) {
    int q = k0 / 8;
    ra[0] = va ? pah[q] : zero;
    ra[1] = va ? pah[q + 1] : zero;
    ra[2] = va ? pal[q] : zero;
    ra[3] = va ? pal[q + 1] : zero;
    rb[0] = vb ? pbh[q] : zero;
    rb[1] = vb ? pbh[q + 1] : zero;
    rb[2] = vb ? pbl[q] : zero;
    rb[3] = vb ? pbl[q + 1] : zero;
  };
  gload(0);
  for (int k0 = 0; k0 < Kpad; k0 += BK16) {
    __syncthreads();
    *reinterpret_cast<v8bf *>(&Ah[lr * LDK + lk]) = ra[0];
    *reinterpret_cast<v8bf *>(&Ah[lr * LDK + lk + 8]) = ra[1];
    *reinterpret_cast<v8bf *>(&Al[lr * LDK + lk]) = ra[2];
    *reinterpret_cast<v8bf *>(&Al[lr * LDK + lk + 8]) = ra[3];
    *reinterpret_cast<v8bf *>(&Bh[lr * LDK + lk]) = rb[0];
    *reinterpret_cast<v8bf *>(&Bh[lr * LDK + lk + 8]) = rb[1];
    *reinterpret_cast<v8bf *>(&Bl[lr * LDK + lk]) = rb[2];
    *reinterpret_cast<v8bf *>(&Bl[lr * LDK + lk + 8]) = rb[3];
    __syncthreads();
    if (k0 + BK16 < Kpad) gload(k0 + BK16);
    v8bf ah[B16_TM], al[B16_TM], bh[B16_TN], bl[B16_TN];
    int kk = 8 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < B16_TM; t++) {
      int ar = wr * 16 * B16_TM + t * 16 + (lane & 15);
      ah[t] = *reinterpret_cast<const v8bf *>(&Ah[ar * LDK + kk]);
      al[t] = *reinterpret_cast<const v8bf *>(&Al[ar * LDK + kk]);
    }
#pragma unroll
    for (int t = 0; t < B16_TN; t++) {
      int br = wc * 16 * B16_TN + t * 16 + (lane & 15);
      bh[t] = *reinterpret_cast<const v8bf *>(&Bh[br * LDK + kk]);
      bl[t] = *reinterpret_cast<const v8bf *>(&Bl[br * LDK + kk]);
    }
#pragma unroll
    for (int a = 0; a < B16_TM; a++)
#pragma unroll
      for (int b = 0; b < B16_TN; b++)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
#pragma unroll
    for (int a = 0; a < B16_TM; a++)
#pragma unroll
      for (int b = 0; b < B16_TN; b++)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
#pragma unroll
    for (int a = 0; a < B16_TM; a++)
#pragma unroll
      for (int b = 0; b < B16_TN; b++)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
  }
  TileOut<float, B16_TM, B16_TN> t;
#pragma unroll
  for (int a = 0; a < B16_TM; a++)
#pragma unroll
    for (int b = 0; b < B16_TN; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) t.v[a][b][i] = acc[a][b][i];
  run_epilogue<EP, float, float, B16_TM, B16_TN>(e, t, row0, col0, M, N, lane, wr, wc, red);
}

int gemm_col_tiles(int form, int N) {
  int bn = form == FM_FORM_FP64 ? F64_BN : B16_BN;
  return (N + bn - 1) / bn;
}

int gemm_row_tiles(int form, int M) {
  int bm = form == FM_FORM_FP64 ? F64_BM : B16_BM;
  return (M + bm - 1) / bm;
}

void gemm_f64(hipStream_t st, int ep, int M, int N, int K, const double *A, long long lda,
              const double *B, long long ldb, int lower_tri, const EpArgs &e) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((N + F64_BN - 1) / F64_BN, (M + F64_BM - 1) / F64_BM);
  if (ep == EP_KERNEL)
    hipLaunchKernelGGL(k_gemm_f64<EP_KERNEL>, grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb,
                       lower_tri, e);
  else if (ep == EP_STORE)
    hipLaunchKernelGGL(k_gemm_f64<EP_STORE>, grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb,
                       lower_tri, e);
  else
    hipLaunchKernelGGL(k_gemm_f64<EP_SUMSQ>, grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb,
                       lower_tri, e);
  FM_HIP(hipGetLastError());
}

void gemm_bf16x3(hipStream_t st, int ep, int M, int N, int Kpad, const uint16_t *Ah,
                 const uint16_t *Al, long long lda, const uint16_t *Bh, const uint16_t *Bl,
                 long long ldb, const EpArgs &e) {
  if (M <= 0 || N <= 0) return;
  FM_CHECK(Kpad % kSplitPad == 0 && lda % kSplitPad == 0 && ldb % kSplitPad == 0 &&
               Kpad <= lda && Kpad <= ldb,
           FM_ERR_INVALID, "split operands must be padded to the k-step");
  FM_CHECK(((uintptr_t)Ah | (uintptr_t)Al | (uintptr_t)Bh | (uintptr_t)Bl) % 16 == 0,
           FM_ERR_INVALID, "split operands must be 16-byte aligned");
  dim3 grid((N + B16_BN - 1) / B16_BN, (M + B16_BM - 1) / B16_BM);
  auto A1 = reinterpret_cast<const __bf16 *>(Ah), A2 = reinterpret_cast<const __bf16 *>(Al);
  auto B1 = reinterpret_cast<const __bf16 *>(Bh), B2 = reinterpret_cast<const __bf16 *>(Bl);
  if (ep == EP_KERNEL)
    hipLaunchKernelGGL(k_gemm_bf16x3<EP_KERNEL>, grid, dim3(256), 0, st, M, N, Kpad, A1, A2, lda,
                       B1, B2, ldb, e);
  else if (ep == EP_STORE)
    hipLaunchKernelGGL(k_gemm_bf16x3<EP_STORE>, grid, dim3(256), 0, st, M, N, Kpad, A1, A2, lda,
                       B1, B2, ldb, e);
  else
    hipLaunchKernelGGL(k_gemm_bf16x3<EP_SUMSQ>, grid, dim3(256), 0, st, M, N, Kpad, A1, A2, lda,
                       B1, B2, ldb, e);
  FM_HIP(hipGetLastError());
}

static uint16_t bf16_rne(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

void split_bf16_host(const float *x, size_t n, uint16_t *hi, uint16_t *lo) {
  for (size_t k = 0; k < n; k++) {
    uint16_t h = bf16_rne(x[k]);
    uint32_t hb = (uint32_t)h << 16;
    float hf;
    std::memcpy(&hf, &hb, 4);
    hi[k] = h;
    lo[k] = bf16_rne(x[k] - hf);
  }
}

}  // namespace fm
