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
// FP64 form: v_mfma_f64_16x16x4f64, 64 x 64 tiles, for all three. Chosen
// form: EpKernel is exact on int8 MFMA (v_mfma_i32_32x32x32_i8) over five
// power-of-two-scaled slices of each operand row, int32 accumulation per
// slice-pair diagonal and an FP64 epilogue (k_gemm_i8); EpStore is
// split-BF16 x3 on v_mfma_f32_16x16x32_bf16 (a = a_hi + a_lo, ab ~ a_hi b_hi
// + a_hi b_lo + a_lo b_hi, FP32 accumulation), 128 x 128 tiles, on operands
// split once where they are produced, the three products of a k-step run
// over all 16 accumulators of a wave before the next, so consecutive MFMAs
// are independent; EpSumSq stays FP64.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

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

// Output tile of this workgroup. Workgroups are dealt to the 8 XCDs round
// robin (id % 8); the dispatch id is remapped so each XCD gets a contiguous
// run of logical tiles. With column index fastest (the default) the
// workgroups sharing a row tile of A share that XCD's L2; with row index
// fastest those sharing a column tile of B do.
__device__ __forceinline__ int xcd_id() {
  constexpr int kXcd = 8;
  int ntx = gridDim.x, nty = gridDim.y, nblk = ntx * nty;
  int pid = blockIdx.y * ntx + blockIdx.x;
  int per = (nblk + kXcd - 1) / kXcd;
  int tall = nblk % kXcd == 0 ? kXcd : nblk % kXcd;  // XCDs that get `per` ids
  int xcd = pid % kXcd, local = pid / kXcd;
  return xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
}

__device__ __forceinline__ void xcd_tile(int &tx, int &ty, bool rows_fastest = false) {
  int ntx = gridDim.x, nty = gridDim.y;
  int id = xcd_id();
  if (rows_fastest) {
    ty = id % nty;
    tx = id / nty;
  } else {
    tx = id % ntx;
    ty = id / ntx;
  }
}

// As xcd_tile, with the logical tiles of each XCD's run taken in groups of
// G row tiles, row index fastest within a group: the workgroups an XCD holds
// at once then cover G row tiles by (resident / G) column tiles instead of
// (resident / columns) row tiles by every column tile, so fewer A and B tile
// bytes share that XCD's L2 per k step.
__device__ __forceinline__ void xcd_tile_grouped(int &tx, int &ty, int G) {
  int ntx = gridDim.x, nty = gridDim.y;
  int id = xcd_id();
  int gw = G * ntx, first = id / gw * G, gs = min(G, nty - first), r = id % gw;
  tx = r / gs;
  ty = first + r % gs;
}

template <typename Acc>
__device__ __forceinline__ int tile_row(int lane, int i) {
  return sizeof(Acc) == 8 ? (lane >> 4) + 4 * i : 4 * (lane >> 4) + i;
}

__device__ __forceinline__ double fpow(double x, double p, int pi) {
  if (pi == 2) return x * x;
  if (pi == 1) return x;
  return pow(x, p);
}
// Kept out of line: inlined into every element of the unrolled FP32
// epilogue, powf makes the loop too large to unroll.
__device__ __noinline__ float powf_general(float x, float p) { return powf(x, p); }
__device__ __forceinline__ float fpow(float x, float p, int pi) {
  if (pi == 2) return x * x;
  if (pi == 1) return x;
  return powf_general(x, p);
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
        if (e.K) e.K_as<Acc>()[(size_t)r * e.ldk + c] = kv;
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
    if (r < M) e.part[(size_t)(col0 / (32 * TN)) * e.part_ld + r] = red[q] + red[BMc + q];
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
    if (c < N) e.part[(size_t)(row0 / (32 * TM)) * e.part_ld + c] = red[q] + red[BNc + q];
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
// FP64 MFMA GEMM, (32 TM) x (32 TN) tiles; 4 waves (2 x 2), each TM x TN
// MFMA tiles of 16 x 16. LDS holds k-major tiles so a fragment read (16 rows
// at one k) is contiguous. Each thread loads 4 consecutive k of rows
// tid / 4 + 64 j.
// ---------------------------------------------------------------------------
// Base shape: the partial arrays (gemm_col_tiles / gemm_row_tiles) follow it.
constexpr int F64_TM = 2, F64_TN = 2, F64_BM = 32 * F64_TM, F64_BN = 32 * F64_TN;
constexpr int BK64 = 16;
constexpr int PAD64 = 1;  // row stride 65 or 129 doubles: the transposed stores spread over the banks

// Four consecutive k of one row, widened to FP64: one or two 16-byte loads
// when `vec` (the row start is 16-byte aligned and k + 4 <= kend), else
// element by element with zeros past the row count or kend.
template <typename T>
__device__ __forceinline__ void load_row4(const T *P, long long ld, int r, int rows, int k,
                                          int kend, bool vec, double (&o)[4]) {
  const T *p = P + (size_t)(r < rows ? r : 0) * ld + k;
  if (r < rows && vec && k + 4 <= kend) {
    if constexpr (sizeof(T) == 4) {
      float4 v = *reinterpret_cast<const float4 *>(p);
      o[0] = v.x, o[1] = v.y, o[2] = v.z, o[3] = v.w;
    } else {
      double2 a = *reinterpret_cast<const double2 *>(p);
      double2 b = *reinterpret_cast<const double2 *>(p + 2);
      o[0] = a.x, o[1] = a.y, o[2] = b.x, o[3] = b.y;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++) o[q] = (r < rows && k + q < kend) ? (double)p[q] : 0.0;
  }
}

template <int EP, int TM, int TN, typename TB = double>
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, const double *A, long long lda,
                                                  const TB *B, long long ldb, int lower_tri,
                                                  int vec, EpArgs e) {
  constexpr int BM = 32 * TM, BN = 32 * TN, RA = BM / 64, RB = BN / 64;
  static_assert(BM % 64 == 0 && BN % 64 == 0, "tiles are multiples of 64");
  __shared__ double As[BK64][BM + PAD64];
  __shared__ double Bs[BK64][BN + PAD64];
  __shared__ double red[2 * (BM > BN ? BM : BN)];
  int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int wr = wid >> 1, wc = wid & 1;
  // A lower-triangular A makes row tiles uneven (row tile r reads r + 1
  // k-tiles): keep the hardware's round-robin spread of tiles there.
  int tx, ty;
  if (lower_tri) {
    tx = blockIdx.x;
    ty = blockIdx.y;
  } else {
    xcd_tile(tx, ty);
  }
  int row0 = ty * BM, col0 = tx * BN;
  int kend = K;
  if (lower_tri) kend = min(K, row0 + BM);
  v4d acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++) acc[a][b] = v4d{0, 0, 0, 0};
  int lr = tid >> 2, lk = (tid & 3) * 4;
  double ra[RA][4], rb[RB][4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int j = 0; j < RA; j++)
      load_row4(A, lda, row0 + lr + 64 * j, M, k0 + lk, kend, vec & 1, ra[j]);
#pragma unroll
    for (int j = 0; j < RB; j++)
      load_row4(B, ldb, col0 + lr + 64 * j, N, k0 + lk, kend, vec & 2, rb[j]);
  };
  gload(0);
  for (int k0 = 0; k0 < kend; k0 += BK64) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
#pragma unroll
      for (int j = 0; j < RA; j++) As[lk + q][lr + 64 * j] = ra[j][q];
#pragma unroll
      for (int j = 0; j < RB; j++) Bs[lk + q][lr + 64 * j] = rb[j][q];
    }
    __syncthreads();
    if (k0 + BK64 < kend) gload(k0 + BK64);
#pragma unroll
    for (int kk = 0; kk < BK64; kk += 4) {
      double af[TM], bf[TN];
#pragma unroll
      for (int t = 0; t < TM; t++) af[t] = As[kk + (lane >> 4)][wr * 16 * TM + t * 16 + (lane & 15)];
#pragma unroll
      for (int t = 0; t < TN; t++) bf[t] = Bs[kk + (lane >> 4)][wc * 16 * TN + t * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
  }
  TileOut<double, TM, TN> t;
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) t.v[a][b][i] = acc[a][b][i];
  run_epilogue<EP, double, double, TM, TN>(e, t, row0, col0, M, N, lane, wr, wc, red);
}

// ---------------------------------------------------------------------------
// Split-BF16 x3 GEMM (chosen form), (32 TM) x (32 TN) tiles, on operands stored as
// two BF16 pieces of FP32 values (x = x_0 + x_1, split once where the operand
// is produced: the sparse set at model upload, descriptors in the descriptor
// kernel, w in the kernel-contraction epilogue). The products x_1 y_0,
// x_0 y_1 and x_0 y_0 are summed in that order (smallest first). Threads copy
// 16-byte vectors of consecutive k into k-contiguous LDS rows (SplitLoader);
// lane l then reads 8 consecutive k at row l % 16, k offset 8 (l / 16).
// ---------------------------------------------------------------------------
// Base shape of the split contraction (partials follow it).
constexpr int B16_TM = 4, B16_TN = 4, B16_BM = 32 * B16_TM, B16_BN = 32 * B16_TN;
constexpr int BK16 = 32;
constexpr int PADB = 8;
constexpr int LDK = BK16 + PADB;
static_assert(BK16 == kSplitPad, "k-step equals the operand padding");

struct SplitPtrs {
  const __bf16 *p[3];
};

// Loader of one operand's rows [row0, row0 + BR) x 32 k, NP pieces: with
// TPR = 256 / BR threads per row, thread tid takes row tid / TPR and
// 32 / TPR consecutive k (one or two 16-byte vectors). Rows past `rows` read
// row 0 (in bounds) and load zeros.
template <int NP, int BR>
struct SplitLoader {
  static constexpr int TPR = 256 / BR, VPT = 4 / TPR;
  int lr, lk;
  bool valid;
  size_t off;
  v8bf r[NP][VPT];
  __device__ __forceinline__ SplitLoader(int tid, int row0, int rows, long long ld) {
    lr = tid / TPR;
    lk = (tid % TPR) * 8 * VPT;
    valid = row0 + lr < rows;
    off = (size_t)(valid ? row0 + lr : 0) * ld + lk;
  }
  __device__ __forceinline__ void load(const SplitPtrs &g, int k0) {
    const v8bf zero = {};
    int q = k0 / 8;
#pragma unroll
    for (int pc = 0; pc < NP; pc++) {
      const v8bf *p = reinterpret_cast<const v8bf *>(g.p[pc] + off);
#pragma unroll
      for (int v = 0; v < VPT; v++) r[pc][v] = valid ? p[q + v] : zero;
    }
  }
  __device__ __forceinline__ void store(__bf16 (*lds)[BR * LDK]) const {
#pragma unroll
    for (int pc = 0; pc < NP; pc++)
#pragma unroll
      for (int v = 0; v < VPT; v++)
        *reinterpret_cast<v8bf *>(&lds[pc][lr * LDK + lk + 8 * v]) = r[pc][v];
  }
};

// With DB the LDS tiles are double-buffered: the next k-step's tile is
// stored while the current one feeds the MFMAs, one barrier per k-step (the
// products and their order are the same, so results are too). It pays on
// long k (one workgroup of 82 KB per CU, against three single-buffered).
template <int EP, int NP, int TM, int TN, bool DB>
__global__ __launch_bounds__(256) void k_gemm_split(int M, int N, int Kpad, SplitPtrs Ag,
                                                    long long lda, SplitPtrs Bg, long long ldb,
                                                    EpArgs e) {
  constexpr int BM = 32 * TM, BN = 32 * TN, NB = DB ? 2 : 1;
  __shared__ __align__(16) __bf16 As[NB][NP][BM * LDK];
  __shared__ __align__(16) __bf16 Bs[NB][NP][BN * LDK];
  __shared__ double red[2 * (BM > BN ? BM : BN)];
  int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int wr = wid >> 1, wc = wid & 1;
  int tx, ty;
  xcd_tile(tx, ty);
  int row0 = ty * BM, col0 = tx * BN;
  v4f acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++) acc[a][b] = v4f{0, 0, 0, 0};
  SplitLoader<NP, BM> la(tid, row0, M, lda);
  SplitLoader<NP, BN> lb(tid, col0, N, ldb);
  la.load(Ag, 0);
  lb.load(Bg, 0);
  if (DB) {
    la.store(As[0]);
    lb.store(Bs[0]);
    __syncthreads();
  }
  int cur = 0;
  for (int k0 = 0; k0 < Kpad; k0 += BK16) {
    if (!DB) {
      __syncthreads();
      la.store(As[0]);
      lb.store(Bs[0]);
      __syncthreads();
    }
    bool more = k0 + BK16 < Kpad;
    if (more) {
      la.load(Ag, k0 + BK16);
      lb.load(Bg, k0 + BK16);
    }
    v8bf af[NP][TM], bf[NP][TN];
    int kk = 8 * (lane >> 4);
#pragma unroll
    for (int pc = 0; pc < NP; pc++) {
#pragma unroll
      for (int t = 0; t < TM; t++) {
        int ar = wr * 16 * TM + t * 16 + (lane & 15);
        af[pc][t] = *reinterpret_cast<const v8bf *>(&As[cur][pc][ar * LDK + kk]);
      }
#pragma unroll
      for (int t = 0; t < TN; t++) {
        int br = wc * 16 * TN + t * 16 + (lane & 15);
        bf[pc][t] = *reinterpret_cast<const v8bf *>(&Bs[cur][pc][br * LDK + kk]);
      }
    }
    // Products of piece orders i + j <= NP - 1, smallest terms first.
    auto prod = [&](const v8bf (&x)[TM], const v8bf (&y)[TN]) {
#pragma unroll
      for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x[a], y[b], acc[a][b], 0, 0, 0);
    };
    static_assert(NP == 2, "two pieces: products lo.hi, hi.lo, hi.hi");
    prod(af[1], bf[0]);
    prod(af[0], bf[1]);
    prod(af[0], bf[0]);
    if (DB) {
      if (more) {
        la.store(As[cur ^ 1]);
        lb.store(Bs[cur ^ 1]);
      }
      __syncthreads();
      cur ^= 1;
    }
  }
  TileOut<float, TM, TN> t;
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) t.v[a][b][i] = acc[a][b][i];
  run_epilogue<EP, float, float, TM, TN>(e, t, row0, col0, M, N, lane, wr, wc, red);
}

// ---------------------------------------------------------------------------
// Exact-integer kernel contraction (chosen form). Every row x of either
// operand is scaled by a power of two s = 2^E with max|x| < s, and r = x / s
// is cut into kI8Slices signed 8-bit slices, r = sum_p a_p 2^-(6 + 7p) + tail
// (|a_p| <= 64, |tail| <= 2^-(7 kI8Slices)): the first slice takes 6 bits and
// every further one 7 (rounding to nearest, so each residual is at most half
// a unit of the previous slice). The product x . y is then
//   s_x s_y sum_{p + q < kI8Slices} 2^-(12 + 7(p + q)) (A_p . B_q),
// each A_p . B_q exact on the int8 matrix cores (int32 accumulation; the
// terms of equal p + q share one accumulator). Dropping the products with
// p + q >= kI8Slices leaves about 34 significant bits, so the kernel values
// k = sigma^2 S^p that the local variance amplifies by L_inv carry FP64-like
// accuracy, where FP32 accumulation of BF16 pieces carries about 16 bits.
// ---------------------------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// One wave per row: the row maximum, its power-of-two scale and the slices
// of 4 consecutive values per lane (zeros past nd up to Kq), stored in the
// tiled slice layout (tr rows per tile).
template <typename T>
__global__ __launch_bounds__(256) void k_slice_rows(const T *X, long long ldx, int n, int nd,
                                                    int8_t *Q, long long nk, int tr,
                                                    double *scale) {
  int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  const T *x = X + (size_t)row * ldx;
  double mx = 0;
  for (int k = lane; k < nd; k += 64) mx = fmax(mx, fabs((double)x[k]));
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  int E = 0;
  if (mx > 0) frexp(mx, &E);  // mx = m 2^E, m in [0.5, 1)
  double inv = ldexp(1.0, -E);
  if (lane == 0) scale[row] = ldexp(1.0, E);
  long long Kq = nk * kI8Pad;
  for (int k0 = 4 * lane; k0 < Kq; k0 += 256) {
    int packed[kI8Slices];
#pragma unroll
    for (int p = 0; p < kI8Slices; p++) packed[p] = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int k = k0 + j;
      double r = k < nd ? (double)x[k] * inv : 0.0;
#pragma unroll
      for (int p = 0; p < kI8Slices; p++) {
        double a = rint(ldexp(r, 6 + 7 * p));
        r -= ldexp(a, -(6 + 7 * p));
        packed[p] |= ((int)a & 0xff) << (8 * j);
      }
    }
#pragma unroll
    for (int p = 0; p < kI8Slices; p++)
      *reinterpret_cast<int *>(Q + i8_tile_offset(tr, nk, row, k0, p)) = packed[p];
  }
}

// An int32 accumulator sums at most kI8Slices products per k of |a b| <=
// 64^2 = 4096: 65536 k keep it below 2^31 (5 x 4096 x 65536 = 1.3e9).
constexpr int kI8Flush = 65536;

// 16-byte chunk position of k chunk c (0..3) in row r of an LDS slice block
// (rows of 64 bytes, no padding): an XOR swizzle that keeps the fragment
// reads (32 rows at one chunk) and the contiguous stores free of bank
// conflicts for 8- and 16-lane phases.
__device__ __forceinline__ int i8_chunk(int r, int c) { return c ^ (((r >> 1) ^ (r >> 3)) & 3); }

// 128 x 64 tiles (kI8TileA x kI8TileN), 8 waves (4 x 2), each 32 x 32 of
// v_mfma_i32_32x32x32_i8 (lane l: operand row l % 32, bytes k = 16 (l / 32)
// .. + 15; result column l % 32, rows 8 (j / 4) + 4 (l / 32) + j % 4); two
// waves per SIMD, so one wave's loads, LDS traffic and barrier waits overlap
// the other's MFMAs. One k step (64) of the workgroup's row tile and column
// tile is one contiguous run of each operand (the tiled slice layout), loaded
// with 16-byte vectors into registers one step ahead and stored to one of two
// LDS buffers after the step before it computes: one barrier per k step. The slice
// products run in rounds (round r: slice r of A against each slice of B it
// pairs with), so consecutive MFMAs accumulate into different diagonals.
// Epilogue as EpKernel, in FP64 from the exact sums: S, k (FP64), w as BF16
// pairs, energy partials per column tile; k and w are staged in LDS and
// stored as whole rows of the tile.
template <bool LONG>
__global__ __launch_bounds__(512) void k_gemm_i8(int M, int N, int nk, const int8_t *Ag,
                                                 const double *sa, const int8_t *Bg,
                                                 const double *sb, EpArgs e) {
  constexpr int S = kI8Slices, BM = kI8TileA, BN = kI8TileN, NT = 512;
  constexpr int A_SL = BM * kI8Pad, B_SL = BN * kI8Pad;  // bytes of one slice block
  constexpr int A_ST = S * A_SL, B_ST = S * B_SL;        // bytes of one k step
  constexpr int NA = A_ST / (16 * NT);                   // 16-byte loads per thread (A)
  constexpr int NB = (B_ST + 16 * NT - 1) / (16 * NT);   // (B; the last by some threads)
  static_assert(A_ST % (16 * NT) == 0 && B_ST % (16 * NT / 2) == 0, "load split");
  static_assert(BM * BN * (8 + 2 + 2) <= 2 * (A_ST + B_ST), "epilogue staging fits");
  __shared__ __align__(16) int8_t smem[2 * (A_ST + B_ST)];
  __shared__ double red[2 * BM];
  int8_t *As = smem, *Bs = smem + 2 * A_ST;  // [2][A_ST], [2][B_ST]
  int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int wr = wid >> 1, wc = wid & 1;
  int tx, ty;
  xcd_tile_grouped(tx, ty, kI8Group);
  int row0 = ty * BM, col0 = tx * BN;
  const int8_t *Ab = Ag + (size_t)ty * nk * A_ST + 16 * tid;
  const int8_t *Bb = Bg + (size_t)tx * nk * B_ST + 16 * tid;
  v16i acc[S];
  // With LONG (Kq > kI8Flush) the int32 sums of the slice products
  // (p + q = d) are folded into one FP64 sum every kI8Flush k, before they
  // could overflow; otherwise once, after the loop.
  constexpr int SJ = LONG ? 16 : 1;
  double sacc[SJ];
  auto clear = [&]() {
#pragma unroll
    for (int d = 0; d < S; d++) acc[d] = v16i{};
  };
  auto diag_sum = [&](int j) {
    double sum = 0;
#pragma unroll
    for (int d = S - 1; d >= 0; d--) sum += ldexp((double)acc[d][j], -(12 + 7 * d));
    return sum;
  };
  auto fold = [&]() {
    if constexpr (LONG) {
#pragma unroll
      for (int j = 0; j < 16; j++) sacc[j] += diag_sum(j);
    }
  };
#pragma unroll
  for (int j = 0; j < SJ; j++) sacc[j] = 0;
  clear();
  v4i ra0[NA], rb0[NB];
  auto gload = [&](auto &ra, auto &rb, int ks) {
#pragma unroll
    for (int j = 0; j < NA; j++)
      ra[j] = *reinterpret_cast<const v4i *>(Ab + (size_t)ks * A_ST + 16 * NT * j);
#pragma unroll
    for (int j = 0; j < NB; j++)
      if (16 * NT * j + 16 * tid < B_ST)
        rb[j] = *reinterpret_cast<const v4i *>(Bb + (size_t)ks * B_ST + 16 * NT * j);
  };
  // Thread tid's 16 bytes at offset 16 tid + 16 NT j of a step: slice block,
  // row (of 64 bytes) and k chunk; stored at the swizzled chunk.
  auto lstore = [&](const auto &ra, const auto &rb, int buf) {
#pragma unroll
    for (int j = 0; j < NA; j++) {
      int o = 16 * tid + 16 * NT * j, p = o / A_SL, w = o % A_SL, r = w >> 6;
      *reinterpret_cast<v4i *>(
          &As[buf * A_ST + p * A_SL + r * 64 + 16 * i8_chunk(r, (w >> 4) & 3)]) = ra[j];
    }
#pragma unroll
    for (int j = 0; j < NB; j++) {
      int o = 16 * tid + 16 * NT * j, p = o / B_SL, w = o % B_SL, r = w >> 6;
      if (o < B_ST)
        *reinterpret_cast<v4i *>(
            &Bs[buf * B_ST + p * B_SL + r * 64 + 16 * i8_chunk(r, (w >> 4) & 3)]) = rb[j];
    }
  };
  auto compute = [&](int buf) {
    int ra_ = wr * 32 + (lane & 31), rb_ = wc * 32 + (lane & 31);
#pragma unroll
    for (int kk = 0; kk < kI8Pad; kk += 32) {
      int c = (kk >> 4) + (lane >> 5);
      v4i af[S], bf[S];
#pragma unroll
      for (int p = 0; p < S; p++) {
        af[p] = *reinterpret_cast<const v4i *>(
            &As[buf * A_ST + p * A_SL + ra_ * 64 + 16 * i8_chunk(ra_, c)]);
        bf[p] = *reinterpret_cast<const v4i *>(
            &Bs[buf * B_ST + p * B_SL + rb_ * 64 + 16 * i8_chunk(rb_, c)]);
      }
#pragma unroll
      for (int rd = 0; rd < S; rd++)
#pragma unroll
        for (int d = S - 1; d >= rd; d--)
          acc[d] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[rd], bf[d - rd], acc[d], 0, 0, 0);
    }
  };
  auto flush = [&](int ks) {
    if constexpr (LONG) {
      if (ks > 0 && (ks * kI8Pad) % kI8Flush == 0) {
        fold();
        clear();
      }
    }
  };
  // Before step ks: LDS buffer ks % 2 holds step ks, the registers step ks + 1
  // (issued after the previous step's barrier; the other wave on each SIMD
  // computes while they land).
  gload(ra0, rb0, 0);
  lstore(ra0, rb0, 0);
  if (nk > 1) gload(ra0, rb0, 1);
  __syncthreads();
#pragma unroll 1
  for (int ks = 0; ks < nk; ks++) {
    flush(ks);
    compute(ks & 1);
    if (ks + 1 < nk) lstore(ra0, rb0, (ks + 1) & 1);  // last read in step ks - 1
    __syncthreads();
    if (ks + 2 < nk) gload(ra0, rb0, ks + 2);
  }
  fold();
  // Epilogue: values in FP64 per element, staged in LDS (over the operand
  // buffers, free after the loop's last barrier) as k (BM x BN FP64) and the
  // two BF16 pieces of w (BM x BN each).
  double *Kst = reinterpret_cast<double *>(smem);
  uint16_t *Hst = reinterpret_cast<uint16_t *>(smem + BM * BN * 8);
  uint16_t *Lst = Hst + BM * BN;
  int pi = e.power == 2.0 ? 2 : (e.power == 1.0 ? 1 : 0);
  int cc = wc * 32 + (lane & 31), c = col0 + cc;
  bool cin = c < N;
  double a_c = cin ? e.alpha[c] : 0.0, s_c = cin ? sb[c] : 0.0;
  double rsum[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    int rr = wr * 32 + 8 * (j / 4) + 4 * (lane >> 5) + (j & 3);
    int r = row0 + rr;
    double kv = 0.0;
    float wf = 0.0f;
    if (r < M && cin) {
      double sum;
      if constexpr (LONG) sum = sacc[j];
      else sum = 0.0 + diag_sum(j);
      double nrm = e.row_norm[r];
      bool empty = nrm < kEmptyThresh;
      double s = empty ? 0.0 : sum * sa[r] * s_c / nrm;
      kv = empty ? 0.0 : e.sig2 * fpow(s, e.power, pi);
      double pm1 = pi == 2 ? s : (pi == 1 ? 1.0 : fpow(s, e.power - 1, 0));
      wf = (float)(empty ? 0.0 : e.sig2 * a_c * e.power * pm1);
    }
    __bf16 h = (__bf16)wf, l = (__bf16)(wf - (float)h);
    Kst[rr * BN + cc] = kv;
    Hst[rr * BN + cc] = __builtin_bit_cast(uint16_t, h);
    Lst[rr * BN + cc] = __builtin_bit_cast(uint16_t, l);
    rsum[j] = kv * a_c;
  }
  // Row sums over the 32 columns of each half-wave, then over the two wave
  // columns through LDS: one energy partial per (column tile, row).
#pragma unroll
  for (int j = 0; j < 16; j++) {
    double v = rsum[j];
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    rsum[j] = v;
  }
  if ((lane & 31) == 0) {
#pragma unroll
    for (int j = 0; j < 16; j++)
      red[wc * BM + wr * 32 + 8 * (j / 4) + 4 * (lane >> 5) + (j & 3)] = rsum[j];
  }
  __syncthreads();
  // k: rows of the tile as 16-byte pairs of columns (ldk even, 16-byte base).
  if (e.K) {
    double *K = e.K_as<double>();
    for (int q = tid; q < BM * (BN / 2); q += NT) {
      int rr = q / (BN / 2), c2 = 2 * (q % (BN / 2)), r = row0 + rr, cg = col0 + c2;
      if (r >= M || cg >= N) continue;
      double *dst = K + (size_t)r * e.ldk + cg;
      if (cg + 1 < N)
        *reinterpret_cast<double2 *>(dst) = *reinterpret_cast<const double2 *>(&Kst[rr * BN + c2]);
      else
        dst[0] = Kst[rr * BN + c2];
    }
  }
  // w pieces: 8 columns per 16-byte store; the padding columns [N, ldw) are
  // zero (the reduction tail of the gradient contraction), none past ldw
  // (a multiple of 32) is written.
  for (int q = tid; q < BM * (BN / 8); q += NT) {
    int rr = q / (BN / 8), c8 = 8 * (q % (BN / 8)), r = row0 + rr, cg = col0 + c8;
    if (r >= M || cg >= e.ldw) continue;
    *reinterpret_cast<v4i *>(reinterpret_cast<uint16_t *>(e.Wh) + (size_t)r * e.ldw + cg) =
        *reinterpret_cast<const v4i *>(&Hst[rr * BN + c8]);
    *reinterpret_cast<v4i *>(reinterpret_cast<uint16_t *>(e.Wl) + (size_t)r * e.ldw + cg) =
        *reinterpret_cast<const v4i *>(&Lst[rr * BN + c8]);
  }
  for (int q = tid; q < BM; q += NT) {
    int r = row0 + q;
    if (r < M) e.part[(size_t)(col0 / BN) * e.part_ld + r] = red[q] + red[BM + q];
  }
}

int gemm_col_tiles(int form, int N) {
  int bn = form == FM_FORM_FP64 ? F64_BN : kI8TileN;
  return (N + bn - 1) / bn;
}

int gemm_row_tiles(int form, int M) {
  int bm = form == FM_FORM_FP64 ? F64_BM : B16_BM;
  return (M + bm - 1) / bm;
}

// ---------------------------------------------------------------------------
// Tile autotuner. The first contraction of a given shape (epilogue, M and N
// rounded up to powers of two, K, device) times every admissible tile shape
// (the faster of two launches after a warm-up one) on the caller's stream
// and keeps the fastest for later calls. Admissible shapes never change results: the energy partials
// of EP_KERNEL are per column tile and the |Q|^2 partials of EP_SUMSQ per
// row tile, so EP_KERNEL only varies TM and EP_SUMSQ only TN; every output
// element sums its k terms in the same order whatever the tile.
// ---------------------------------------------------------------------------
namespace {

struct TileShape {
  int tm, tn, db;
};
constexpr TileShape kF64Shapes[] = {{2, 2, 0}, {4, 2, 0}, {2, 4, 0}, {4, 4, 0}};
// 8 x 4 and 4 x 8: 64 x 128 / 128 x 64 per wavefront, a third fewer LDS
// fragment reads per MFMA than 64 x 64.
constexpr TileShape kSplitShapes[] = {{4, 4, 0}, {2, 4, 0}, {4, 2, 0}, {4, 4, 1},
                                      {4, 2, 1}, {8, 4, 0}, {4, 8, 0}};

bool admissible(int ep, TileShape t, TileShape base) {
  if (ep == EP_KERNEL) return t.tn == base.tn;
  if (ep == EP_SUMSQ) return t.tm == base.tm;
  return true;
}

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int, int, int>, int> g_tune;

int next_pow2(int x) {
  int p = 64;
  while (p < x) p <<= 1;
  return p;
}

// Returns the tuned shape index for `key`, timing `launch(i)` for each
// admissible index when the key is new. `launch` enqueues one contraction.
template <typename Launch>
int tuned(int family, int ep, int M, int N, int K, hipStream_t st, const TileShape *shapes,
          int n_shapes, Launch &&launch) {
  // FLARE_MI_TILE_SHAPE=i pins shape i for every contraction that admits it
  // (tests check that each shape gives the same results).
  if (const char *f = std::getenv("FLARE_MI_TILE_SHAPE")) {
    int i = std::atoi(f);
    if (i >= 0 && i < n_shapes && admissible(ep, shapes[i], shapes[0])) return i;
  }
  int dev = 0;
  FM_HIP(hipGetDevice(&dev));
  // M and N up to powers of two: the atom count of a new batch (rows of the
  // kernel and gradient contractions, columns of the variance one) reuses the
  // choice for its size class instead of timing every shape again.
  auto key = std::make_tuple(family, ep, next_pow2(M), next_pow2(N), K, dev, 0);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) return it->second;
  }
  hipEvent_t e0, e1;
  FM_HIP(hipEventCreate(&e0));
  FM_HIP(hipEventCreate(&e1));
  int best = -1;
  float best_ms = 0;
  for (int i = 0; i < n_shapes; i++) {
    if (!admissible(ep, shapes[i], shapes[0])) continue;
    launch(i);  // warm-up (code object load, caches)
    float ms = 0;
    for (int r = 0; r < 2; r++) {  // the faster of two timed launches
      FM_HIP(hipEventRecord(e0, st));
      launch(i);
      FM_HIP(hipEventRecord(e1, st));
      FM_HIP(hipEventSynchronize(e1));
      float t = 0;
      FM_HIP(hipEventElapsedTime(&t, e0, e1));
      ms = r == 0 ? t : std::min(ms, t);
    }
    if (best < 0 || ms < best_ms) {
      best = i;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune[key] = best;
  return best;
}

template <int EP, int TM, int TN, typename TB>
void f64_launch(hipStream_t st, int M, int N, int K, const double *A, long long lda,
                const TB *B, long long ldb, int lower_tri, const EpArgs &e) {
  dim3 grid((N + 32 * TN - 1) / (32 * TN), (M + 32 * TM - 1) / (32 * TM));
  // 16-byte loads where every row start is 16-byte aligned.
  auto aligned = [](const void *p, long long ld, size_t esz) {
    return (uintptr_t)p % 16 == 0 && (ld * (long long)esz) % 16 == 0;
  };
  int vec = (aligned(A, lda, sizeof(double)) ? 1 : 0) | (aligned(B, ldb, sizeof(TB)) ? 2 : 0);
  hipLaunchKernelGGL((k_gemm_f64<EP, TM, TN, TB>), grid, dim3(256), 0, st, M, N, K, A, lda, B,
                     ldb, lower_tri, vec, e);
}

template <int EP, typename TB>
void f64_shape(int i, hipStream_t st, int M, int N, int K, const double *A, long long lda,
               const TB *B, long long ldb, int lower_tri, const EpArgs &e) {
  switch (i) {
    case 0: f64_launch<EP, 2, 2>(st, M, N, K, A, lda, B, ldb, lower_tri, e); break;
    case 1: f64_launch<EP, 4, 2>(st, M, N, K, A, lda, B, ldb, lower_tri, e); break;
    case 2: f64_launch<EP, 2, 4>(st, M, N, K, A, lda, B, ldb, lower_tri, e); break;
    default: f64_launch<EP, 4, 4>(st, M, N, K, A, lda, B, ldb, lower_tri, e); break;
  }
}

template <int EP, int TM, int TN, bool DB>
void split_launch(hipStream_t st, int M, int N, int Kpad, const SplitPtrs &a, long long lda,
                  const SplitPtrs &b, long long ldb, const EpArgs &e) {
  dim3 grid((N + 32 * TN - 1) / (32 * TN), (M + 32 * TM - 1) / (32 * TM));
  hipLaunchKernelGGL((k_gemm_split<EP, 2, TM, TN, DB>), grid, dim3(256), 0, st, M, N, Kpad, a,
                     lda, b, ldb, e);
}

template <int EP>
void split_shape(int i, hipStream_t st, int M, int N, int Kpad, const SplitPtrs &a,
                 long long lda, const SplitPtrs &b, long long ldb, const EpArgs &e) {
  switch (i) {
    case 0: split_launch<EP, 4, 4, false>(st, M, N, Kpad, a, lda, b, ldb, e); break;
    case 1: split_launch<EP, 2, 4, false>(st, M, N, Kpad, a, lda, b, ldb, e); break;
    case 2: split_launch<EP, 4, 2, false>(st, M, N, Kpad, a, lda, b, ldb, e); break;
    case 3: split_launch<EP, 4, 4, true>(st, M, N, Kpad, a, lda, b, ldb, e); break;
    case 4: split_launch<EP, 4, 2, true>(st, M, N, Kpad, a, lda, b, ldb, e); break;
    case 5: split_launch<EP, 8, 4, false>(st, M, N, Kpad, a, lda, b, ldb, e); break;
    default: split_launch<EP, 4, 8, false>(st, M, N, Kpad, a, lda, b, ldb, e); break;
  }
}

// The integer contraction runs 128 x 64 tiles (kI8TileA x kI8TileN): the
// column tile is fixed at 64 (the energy partials are per column tile).
void i8_launch(hipStream_t st, int M, int N, int nk, const int8_t *a, const double *sa,
               const int8_t *b, const double *sb, const EpArgs &e) {
  dim3 grid((N + kI8TileN - 1) / kI8TileN, (M + kI8TileA - 1) / kI8TileA);
  if ((long long)nk * kI8Pad > kI8Flush)
    hipLaunchKernelGGL((k_gemm_i8<true>), grid, dim3(512), 0, st, M, N, nk, a, sa, b, sb, e);
  else
    hipLaunchKernelGGL((k_gemm_i8<false>), grid, dim3(512), 0, st, M, N, nk, a, sa, b, sb, e);
}

}  // namespace

void slice_rows(hipStream_t st, const float *X, long long ldx, int n, int nd, int8_t *Q,
                long long Kq, int tr, double *scale) {
  if (n <= 0) return;
  FM_CHECK(Kq % kI8Pad == 0 && Kq >= nd && tr > 0, FM_ERR_INVALID, "bad slice layout");
  hipLaunchKernelGGL(k_slice_rows<float>, dim3((n + 3) / 4), dim3(256), 0, st, X, ldx, n, nd, Q,
                     Kq / kI8Pad, tr, scale);
  FM_HIP(hipGetLastError());
}

void slice_rows(hipStream_t st, const double *X, long long ldx, int n, int nd, int8_t *Q,
                long long Kq, int tr, double *scale) {
  if (n <= 0) return;
  FM_CHECK(Kq % kI8Pad == 0 && Kq >= nd && tr > 0, FM_ERR_INVALID, "bad slice layout");
  hipLaunchKernelGGL(k_slice_rows<double>, dim3((n + 3) / 4), dim3(256), 0, st, X, ldx, n, nd,
                     Q, Kq / kI8Pad, tr, scale);
  FM_HIP(hipGetLastError());
}

void gemm_i8_kernel(hipStream_t st, int M, int N, int Kq, const SliceOperand &A,
                    const SliceOperand &B, const EpArgs &e) {
  if (M <= 0 || N <= 0) return;
  FM_CHECK(Kq > 0 && Kq % kI8Pad == 0 && A.Kq == Kq && B.Kq == Kq, FM_ERR_INVALID,
           "sliced operands must share a reduction length padded to 64");
  FM_CHECK(A.tr == kI8TileA && B.tr == kI8TileN, FM_ERR_INVALID,
           "sliced operands must be tiled by 128 (rows) and 64 (columns)");
  FM_CHECK(A.q && B.q && A.scale && B.scale &&
               ((uintptr_t)A.q | (uintptr_t)B.q) % 16 == 0,
           FM_ERR_INVALID, "sliced operands must be 16-byte aligned");
  FM_CHECK(e.Wh && e.Wl && e.alpha && e.row_norm && e.part, FM_ERR_INVALID,
           "integer contraction: missing epilogue outputs");
  FM_CHECK(e.ldw % 32 == 0 && e.ldw >= N && ((uintptr_t)e.Wh | (uintptr_t)e.Wl) % 16 == 0 &&
               (!e.K || (e.ldk % 2 == 0 && e.ldk >= N && (uintptr_t)e.K % 16 == 0)),
           FM_ERR_INVALID, "integer contraction: k and w must be 16-byte aligned rows");
  i8_launch(st, M, N, Kq / kI8Pad, A.q, A.scale, B.q, B.scale, e);
  FM_HIP(hipGetLastError());
}

void gemm_f64(hipStream_t st, int ep, int M, int N, int K, const double *A, long long lda,
              const double *B, long long ldb, int lower_tri, const EpArgs &e) {
  if (M <= 0 || N <= 0) return;
  auto launch = [&](int i) {
    if (ep == EP_KERNEL) f64_shape<EP_KERNEL>(i, st, M, N, K, A, lda, B, ldb, lower_tri, e);
    else if (ep == EP_STORE) f64_shape<EP_STORE>(i, st, M, N, K, A, lda, B, ldb, lower_tri, e);
    else f64_shape<EP_SUMSQ>(i, st, M, N, K, A, lda, B, ldb, lower_tri, e);
    FM_HIP(hipGetLastError());
  };
  launch(tuned(0, ep, M, N, K, st, kF64Shapes, 4, launch));
}

void gemm_split(hipStream_t st, int ep, int np, int M, int N, int Kpad, const SplitOperand &A,
                const SplitOperand &B, const EpArgs &e) {
  if (M <= 0 || N <= 0) return;
  // The kernel contraction runs on int8 slices (gemm_i8_kernel); the split
  // form serves the gradient contraction.
  FM_CHECK(ep == EP_STORE, FM_ERR_INVALID, "split-BF16 contraction: EP_STORE only");
  FM_CHECK(np == 2, FM_ERR_INVALID, "split operands have 2 pieces");
  FM_CHECK(Kpad % kSplitPad == 0 && A.ld % kSplitPad == 0 && B.ld % kSplitPad == 0 &&
               Kpad <= A.ld && Kpad <= B.ld,
           FM_ERR_INVALID, "split operands must be padded to the k-step");
  SplitPtrs a{}, b{};
  for (int pc = 0; pc < np; pc++) {
    FM_CHECK(A.piece[pc] && B.piece[pc], FM_ERR_INVALID, "missing split operand piece");
    FM_CHECK(((uintptr_t)A.piece[pc] | (uintptr_t)B.piece[pc]) % 16 == 0, FM_ERR_INVALID,
             "split operands must be 16-byte aligned");
    a.p[pc] = reinterpret_cast<const __bf16 *>(A.piece[pc]);
    b.p[pc] = reinterpret_cast<const __bf16 *>(B.piece[pc]);
  }
  auto launch = [&](int i) {
    split_shape<EP_STORE>(i, st, M, N, Kpad, a, A.ld, b, B.ld, e);
    FM_HIP(hipGetLastError());
  };
  launch(tuned(1, ep, M, N, Kpad, st, kSplitShapes, 7, launch));
}

static uint16_t bf16_rne(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

void split_bf16_host(const float *x, size_t n, int np, uint16_t *const *pieces) {
  for (size_t k = 0; k < n; k++) {
    float r = x[k];
    for (int pc = 0; pc < np; pc++) {
      uint16_t h = bf16_rne(r);
      uint32_t hb = (uint32_t)h << 16;
      float hf;
      std::memcpy(&hf, &hb, 4);
      pieces[pc][k] = h;
      r = r - hf;
    }
  }
}

}  // namespace fm
