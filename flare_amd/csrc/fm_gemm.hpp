// Arguments of the fused-epilogue contractions (fm_gemm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fm {

enum { EP_KERNEL = 0, EP_STORE = 1, EP_SUMSQ = 2 };

struct EpArgs {
  // EP_KERNEL
  double sig2 = 1, power = 2;
  const double *row_norm = nullptr;  // |B_i| per row (atom) of the block
  const double *alpha = nullptr;     // alpha_j per column (sparse env)
  // Optional k matrix (rows x ldk), FP64.
  void *K = nullptr;
  long long ldk = 0;
  // EP_KERNEL writes w, EP_STORE writes C, into W (rows x ldw). In the int8
  // contraction EP_KERNEL instead writes w as the BF16 pair (hi, lo) of its
  // FP32 value into Wh / Wl (rows x ldw), zero in the padding columns
  // [N, ldw), the operand of the split-BF16 gradient contraction.
  void *W = nullptr;
  void *Wh = nullptr, *Wl = nullptr;
  long long ldw = 0;
  // EP_KERNEL: energy partials [column tile][row]; EP_SUMSQ: [row tile][col].
  double *part = nullptr;
  long long part_ld = 0;
  template <typename T>
  __device__ __forceinline__ T *W_as() const {
    return reinterpret_cast<T *>(W);
  }
  template <typename T>
  __device__ __forceinline__ T *K_as() const {
    return reinterpret_cast<T *>(K);
  }
};

// Number of column / row tiles a contraction of the given form uses (the
// energy partials have one slot per column tile, the variance partials one
// per row tile).
int gemm_col_tiles(int form, int N);
int gemm_row_tiles(int form, int M);

void gemm_f64(hipStream_t st, int ep, int M, int N, int K, const double *A, long long lda,
              const double *B, long long ldb, int lower_tri, const EpArgs &e);
// Split-BF16 x3 contraction (EP_STORE: the gradient contraction G = W X_hat
// of the chosen form) on operands stored as two BF16 pieces of FP32
// values, x = x_0 + x_1 with x_0 = bf16(x), x_1 = bf16(x - x_0); it sums
// x_1 y_0 + x_0 y_1 + x_0 y_0 (3 MFMAs per product, about 16 significant
// bits; FP32 accumulation). Rows of A and B are K-contiguous with leading
// dimensions that are multiples of kSplitPad, padded with zeros up to Kpad (a
// multiple of kSplitPad), and 16-byte aligned. np must be 2 (the number of
// pieces each operand carries).
constexpr int kSplitPad = 32;
inline long long split_ld(long long k) { return (k + kSplitPad - 1) / kSplitPad * kSplitPad; }
struct SplitOperand {
  const uint16_t *piece[3] = {nullptr, nullptr, nullptr};
  long long ld = 0;
};
void gemm_split(hipStream_t st, int ep, int np, int M, int N, int Kpad, const SplitOperand &A,
                const SplitOperand &B, const EpArgs &e);
// Exact-integer kernel contraction (the chosen form's S = B X_hat^T with the
// EP_KERNEL epilogue, k in FP64 and w as BF16 pairs): both operands as
// kI8Slices int8 slices of their rows with a power-of-two scale per row, in
// the tiled slice layout below. The slices come from slice_rows (or the
// chosen form's descriptor pass).
constexpr int kI8Slices = 5;
constexpr int kI8Pad = 64;     // k step; the reduction length is padded to it with zeros
constexpr int kI8TileA = 128;  // row tile of the descriptor operand
constexpr int kI8TileN = 64;   // column tile (sparse set): the energy partials are per column tile
// Row tiles per group in the integer contraction's XCD tile order: 32
// resident workgroups per XCD then cover 4 row tiles x 8 column tiles
// (measured 3% faster per launch than 2 row tiles x 16 column tiles; 8 row
// tiles per group measured the same as 4).
constexpr int kI8Group = 4;
inline long long i8_ld(long long k) { return (k + kI8Pad - 1) / kI8Pad * kI8Pad; }
// Tiled slice layout of n rows with reduction length Kq = nk kI8Pad: rows in
// tiles of tr, and tile (row tile rt, k step ks) is kI8Slices consecutive
// blocks of tr x kI8Pad bytes (slice, then row, then k), so one k step of a
// row tile is a single contiguous run of kI8Slices tr kI8Pad bytes. Rows past
// n in the last tile are never read into a kept result.
__host__ __device__ inline long long i8_tile_offset(int tr, long long nk, long long row, int k,
                                                    int p) {
  return (((row / tr) * nk + (k >> 6)) * kI8Slices + p) * (long long)tr * kI8Pad +
         (row % tr) * kI8Pad + (k & (kI8Pad - 1));
}
inline size_t i8_tiled_bytes(long long n, int tr, long long Kq) {
  return (size_t)((n + tr - 1) / tr) * tr * Kq * kI8Slices;
}
struct SliceOperand {
  const int8_t *q = nullptr;  // the operand's first tile (row 0 starts a tile)
  long long Kq = 0;           // padded reduction length, a multiple of kI8Pad
  int tr = 0;                 // rows per tile
  const double *scale = nullptr;
};
void slice_rows(hipStream_t st, const float *X, long long ldx, int n, int nd, int8_t *Q,
                long long Kq, int tr, double *scale);
void slice_rows(hipStream_t st, const double *X, long long ldx, int n, int nd, int8_t *Q,
                long long Kq, int tr, double *scale);
void gemm_i8_kernel(hipStream_t st, int M, int N, int Kq, const SliceOperand &A,
                    const SliceOperand &B, const EpArgs &e);
// Host-side split of FP32 values into np BF16 pieces, rounding to nearest
// even as the device conversion does.
void split_bf16_host(const float *x, size_t n, int np, uint16_t *const *pieces);

}  // namespace fm
