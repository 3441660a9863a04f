// Kernel contraction on the matrix cores. Every contraction of the path has
// the form C = A . B^T with both operands contiguous along the reduction
// index, tiled 64 x 64 per workgroup (4 wavefronts of 32 x 32), with a fused
// epilogue:
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
// FP64 form: v_mfma_f64_16x16x4f64. Chosen form: split-BF16 x3 on
// v_mfma_f32_16x16x32_bf16 (a = a_hi + a_lo, ab ~ a_hi b_hi + a_hi b_lo +
// a_lo b_hi, FP32 accumulation).
#include <hip/hip_runtime.h>

#include "fm_device.hpp"
#include "fm_gemm.hpp"
#include "fm_internal.hpp"

namespace fm {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int BM = 64, BN = 64;

// ---------------------------------------------------------------------------
// Epilogues. Each receives the wave's 2 x 2 tiles of 16 x 16 values with the
// MFMA output layout: lane l holds rows 4 (l / 16) + i, column l % 16.
// ---------------------------------------------------------------------------
template <typename Acc>
struct TileOut {
  Acc v[2][2][4];
};

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

template <typename Acc, typename TW>
__device__ void ep_kernel(const EpArgs &e, const TileOut<Acc> &t, int row0, int col0, int M,
                          int N, int lane, int wr, int wc, double *red) {
  // rows: atoms of this species block, cols: sparse environments.
  int pi = e.power == 2.0 ? 2 : (e.power == 1.0 ? 1 : 0);
  double rs[2][4];
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int i = 0; i < 4; i++) rs[ti][i] = 0;
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int r = row0 + wr * 32 + ti * 16 + 4 * (lane >> 4) + i;
      if (r >= M) continue;
      double nrm = e.row_norm[r];
      bool empty = nrm < kEmptyThresh;
      Acc ninv = empty ? Acc(0) : Acc(1.0 / nrm);
#pragma unroll
      for (int tj = 0; tj < 2; tj++) {
        int c = col0 + wc * 32 + tj * 16 + (lane & 15);
        if (c >= N) continue;
        Acc s = t.v[ti][tj][i] * ninv;
        Acc a = (Acc)e.alpha[c];
        Acc kv = empty ? Acc(0) : Acc(e.sig2) * fpow(s, (Acc)e.power, pi);
        Acc pm1 = pi == 2 ? s : (pi == 1 ? Acc(1) : fpow(s, (Acc)(e.power - 1), 0));
        Acc wv = empty ? Acc(0) : Acc(e.sig2) * a * (Acc)e.power * pm1;
        e.W_as<TW>()[(size_t)r * e.ldw + c] = (TW)wv;
        if (e.K) e.K[(size_t)r * e.ldk + c] = (double)kv;
        rs[ti][i] += (double)(kv * a);
      }
    }
  // Sum over the 16 columns of each lane group, then over the two wave
  // columns through LDS; one partial per (column tile, row).
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      double v = rs[ti][i];
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
      rs[ti][i] = v;
    }
  if ((lane & 15) == 0) {
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
      for (int i = 0; i < 4; i++) red[wc * 64 + wr * 32 + ti * 16 + 4 * (lane >> 4) + i] = rs[ti][i];
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    int r = row0 + threadIdx.x;
    if (r < M) e.part[(size_t)blockIdx.x * e.part_ld + r] = red[threadIdx.x] + red[64 + threadIdx.x];
  }
}

template <typename Acc, typename TW>
__device__ void ep_store(const EpArgs &e, const TileOut<Acc> &t, int row0, int col0, int M, int N,
                         int lane, int wr, int wc) {
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int r = row0 + wr * 32 + ti * 16 + 4 * (lane >> 4) + i;
      if (r >= M) continue;
#pragma unroll
      for (int tj = 0; tj < 2; tj++) {
        int c = col0 + wc * 32 + tj * 16 + (lane & 15);
        if (c < N) e.W_as<TW>()[(size_t)r * e.ldw + c] = (TW)t.v[ti][tj][i];
      }
    }
}

template <typename Acc>
__device__ void ep_sumsq(const EpArgs &e, const TileOut<Acc> &t, int row0, int col0, int M,
                         int N, int lane, int wr, int wc, double *red) {
  // rows: L rows, cols: atoms. Column sums of squares over this row tile.
  double cs[2] = {0, 0};
#pragma unroll
  for (int tj = 0; tj < 2; tj++)
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int r = row0 + wr * 32 + ti * 16 + 4 * (lane >> 4) + i;
        double v = r < M ? (double)t.v[ti][tj][i] : 0.0;
        cs[tj] += v * v;
      }
#pragma unroll
  for (int tj = 0; tj < 2; tj++) {
    double v = cs[tj];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    cs[tj] = v;
  }
  if (lane < 16) {
#pragma unroll
    for (int tj = 0; tj < 2; tj++) red[wr * 64 + wc * 32 + tj * 16 + lane] = cs[tj];
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    int c = col0 + threadIdx.x;
    if (c < N) e.part[(size_t)blockIdx.y * e.part_ld + c] = red[threadIdx.x] + red[64 + threadIdx.x];
  }
}

// ---------------------------------------------------------------------------
// FP64 MFMA GEMM. Grid: x = column tiles (of B rows), y = row tiles (of A).
// For EpKernel the block x index selects the energy partial slot.
// ---------------------------------------------------------------------------
constexpr int BK64 = 16;
constexpr int PAD64 = 4;

template <int EP>
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, const double *A, long long lda,
                                                  const double *B, long long ldb, int lower_tri,
                                                  EpArgs e) {
  __shared__ double As[BK64][BM + PAD64];
  __shared__ double Bs[BK64][BN + PAD64];
  __shared__ double red[128];
  int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int wr = wid >> 1, wc = wid & 1;
  int row0 = blockIdx.y * BM, col0 = blockIdx.x * BN;
  int kend = K;
  if (lower_tri) kend = min(K, row0 + BM);
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = v4d{0, 0, 0, 0};
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
      double af[2], bf[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        af[t] = As[kk + (lane >> 4)][wr * 32 + t * 16 + (lane & 15)];
        bf[t] = Bs[kk + (lane >> 4)][wc * 32 + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
  }
  TileOut<double> t;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) t.v[a][b][i] = acc[a][b][i];
  if (EP == EP_KERNEL) ep_kernel<double, double>(e, t, row0, col0, M, N, lane, wr, wc, red);
  else if (EP == EP_STORE) ep_store<double, double>(e, t, row0, col0, M, N, lane, wr, wc);
  else ep_sumsq<double>(e, t, row0, col0, M, N, lane, wr, wc, red);
}

// ---------------------------------------------------------------------------
// Split-BF16 x3 GEMM on FP32 operands (chosen form). LDS holds hi and lo
// BF16 tiles row-major along k (lane l reads 8 consecutive k at row l % 16,
// k offset 8 (l / 16)).
// ---------------------------------------------------------------------------
constexpr int BK16 = 32;
constexpr int PADB = 8;

__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <int EP>
__global__ __launch_bounds__(256) void k_gemm_bf16x3(int M, int N, int K, const float *A,
                                                     long long lda, const float *B, long long ldb,
                                                     int lower_tri, EpArgs e) {
  __shared__ __align__(16) __bf16 Ah[BM][BK16 + PADB];
  __shared__ __align__(16) __bf16 Al[BM][BK16 + PADB];
  __shared__ __align__(16) __bf16 Bh[BN][BK16 + PADB];
  __shared__ __align__(16) __bf16 Bl[BN][BK16 + PADB];
  __shared__ double red[128];
  int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int wr = wid >> 1, wc = wid & 1;
  int row0 = blockIdx.y * BM, col0 = blockIdx.x * BN;
  int kend = K;
  if (lower_tri) kend = min(K, row0 + BM);
  v4f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = v4f{0, 0, 0, 0};
  // Each thread loads 8 consecutive k of one row of A and of B.
  int lr = tid >> 2, lk = (tid & 3) * 8;
  float ra[8], rb[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int k = k0 + lk + q;
      int r = row0 + lr, c = col0 + lr;
      ra[q] = (r < M && k < kend) ? A[(size_t)r * lda + k] : 0.0f;
      rb[q] = (c < N && k < kend) ? B[(size_t)c * ldb + k] : 0.0f;
    }
  };
  gload(0);
  for (int k0 = 0; k0 < kend; k0 += BK16) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) {
      __bf16 h, l;
      split_bf16(ra[q], h, l);
      Ah[lr][lk + q] = h;
      Al[lr][lk + q] = l;
      split_bf16(rb[q], h, l);
      Bh[lr][lk + q] = h;
      Bl[lr][lk + q] = l;
    }
    __syncthreads();
    if (k0 + BK16 < kend) gload(k0 + BK16);
    v8bf ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      int ar = wr * 32 + t * 16 + (lane & 15), br = wc * 32 + t * 16 + (lane & 15);
      int kk = 8 * (lane >> 4);
      ah[t] = *reinterpret_cast<const v8bf *>(&Ah[ar][kk]);
      al[t] = *reinterpret_cast<const v8bf *>(&Al[ar][kk]);
      bh[t] = *reinterpret_cast<const v8bf *>(&Bh[br][kk]);
      bl[t] = *reinterpret_cast<const v8bf *>(&Bl[br][kk]);
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
  }
  TileOut<float> t;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) t.v[a][b][i] = acc[a][b][i];
  if (EP == EP_KERNEL) ep_kernel<float, float>(e, t, row0, col0, M, N, lane, wr, wc, red);
  else if (EP == EP_STORE) ep_store<float, float>(e, t, row0, col0, M, N, lane, wr, wc);
  else ep_sumsq<float>(e, t, row0, col0, M, N, lane, wr, wc, red);
}

void gemm_f64(hipStream_t st, int ep, int M, int N, int K, const double *A, long long lda,
              const double *B, long long ldb, int lower_tri, const EpArgs &e) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
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

void gemm_bf16x3(hipStream_t st, int ep, int M, int N, int K, const float *A, long long lda,
                 const float *B, long long ldb, int lower_tri, const EpArgs &e) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  if (ep == EP_KERNEL)
    hipLaunchKernelGGL(k_gemm_bf16x3<EP_KERNEL>, grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb,
                       lower_tri, e);
  else if (ep == EP_STORE)
    hipLaunchKernelGGL(k_gemm_bf16x3<EP_STORE>, grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb,
                       lower_tri, e);
  else
    hipLaunchKernelGGL(k_gemm_bf16x3<EP_SUMSQ>, grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb,
                       lower_tri, e);
  FM_HIP(hipGetLastError());
}

}  // namespace fm
