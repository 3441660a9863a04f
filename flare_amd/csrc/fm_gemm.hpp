// Arguments of the fused-epilogue contractions (fm_gemm.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace fm {

enum { EP_KERNEL = 0, EP_STORE = 1, EP_SUMSQ = 2 };

struct EpArgs {
  // EP_KERNEL
  double sig2 = 1, power = 2;
  const double *row_norm = nullptr;  // |B_i| per row (atom) of the block
  const double *alpha = nullptr;     // alpha_j per column (sparse env)
  double *K = nullptr;               // optional k matrix (rows x ldk), FP64
  long long ldk = 0;
  // EP_KERNEL writes w, EP_STORE writes C, into W (rows x ldw).
  void *W = nullptr;
  long long ldw = 0;
  // EP_KERNEL: energy partials [column tile][row]; EP_SUMSQ: [row tile][col].
  double *part = nullptr;
  long long part_ld = 0;
  template <typename T>
  __device__ __forceinline__ T *W_as() const {
    return reinterpret_cast<T *>(W);
  }
};

// Number of column / row tiles a contraction of the given form uses (the
// energy partials have one slot per column tile, the variance partials one
// per row tile).
int gemm_col_tiles(int form, int N);
int gemm_row_tiles(int form, int M);

void gemm_f64(hipStream_t st, int ep, int M, int N, int K, const double *A, long long lda,
              const double *B, long long ldb, int lower_tri, const EpArgs &e);
void gemm_bf16x3(hipStream_t st, int ep, int M, int N, int K, const float *A, long long lda,
                 const float *B, long long ldb, int lower_tri, const EpArgs &e);

}  // namespace fm
