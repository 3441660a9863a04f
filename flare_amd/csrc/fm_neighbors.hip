// Periodic neighbour list on the device, bit-identical to the reference
// Structure::compute_neighbors (src/flare_pp/structure.cpp:52-118): same
// image vector ((d + s1 a) + s2 b) + s3 c, same test dist < cutoff && dist != 0,
// same order (j, s1, s2, s3). Built through a cell list: atoms are binned
// along the fractional axes (bin width >= cutoff perpendicular to each face),
// every candidate in the stencil is tested with the exact expression, images
// outside [-sweep, sweep]^3 are dropped (the reference loop never visits them)
// and each atom's list is sorted by its (j, s) key in LDS.
//
// This file is compiled with -ffp-contract=off so the FP64 expressions round
// exactly as written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fm_device.hpp"
#include "fm_internal.hpp"

namespace fm {

// ---------------------------------------------------------------------------
// Scans
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 1024;

__device__ __forceinline__ int block_excl_scan(int v, int *sh, int &total) {
  // 1024 threads = 16 waves.
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  if (wid == 0) {
    int w = lane < 16 ? sh[lane] : 0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      int y = __shfl_up(w, off, 64);
      if (lane >= off) w += y;
    }
    if (lane < 16) sh[lane] = w;
  }
  __syncthreads();
  int base = wid ? sh[wid - 1] : 0;
  total = sh[15];
  __syncthreads();
  return base + x - v;
}

__global__ void k_scan_block_sums(const int *in, int n, int *bsum) {
  __shared__ int sh[16];
  int i = blockIdx.x * kScanBlock + threadIdx.x;
  int v = i < n ? in[i] : 0;
  int tot;
  block_excl_scan(v, sh, tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void k_scan_sums(int *bsum, int nb) {
  __shared__ int sh[16];
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kScanBlock) {
    int i = b0 + threadIdx.x;
    int v = i < nb ? bsum[i] : 0;
    int tot;
    int e = block_excl_scan(v, sh, tot);
    if (i < nb) bsum[i] = carry + e;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ void k_scan_apply(const int *in, int n, const int *bsum, int *out) {
  __shared__ int sh[16];
  int i = blockIdx.x * kScanBlock + threadIdx.x;
  int v = i < n ? in[i] : 0;
  int tot;
  int e = block_excl_scan(v, sh, tot);
  if (i < n) out[i] = bsum[blockIdx.x] + e;
  if (i == n - 1) out[n] = bsum[blockIdx.x] + e + v;
}

// out[0..n] = exclusive scan of in[0..n); out[n] = total. With want_total
// it returns the total (synchronising the context stream), otherwise 0.
int scan_exclusive(fm_context ctx, const int *in, int *out, int n, bool want_total) {
  if (n == 0) {
    FM_HIP(hipMemsetAsync(out, 0, sizeof(int), ctx->stream));
    return 0;
  }
  int nb = (n + kScanBlock - 1) / kScanBlock;
  DevBuf bsum(&ctx->pool, (nb + 1) * sizeof(int));
  hipLaunchKernelGGL(k_scan_block_sums, dim3(nb), dim3(kScanBlock), 0, ctx->stream, in, n,
                     bsum.as<int>());
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, ctx->stream, bsum.as<int>(), nb);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, ctx->stream, in, n,
                     bsum.as<int>(), out);
  FM_HIP(hipGetLastError());
  if (!want_total) return 0;
  int total = 0;
  FM_HIP(hipMemcpyAsync(&total, out + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FM_HIP(hipStreamSynchronize(ctx->stream));
  return total;
}

// ---------------------------------------------------------------------------
// Per-structure geometry (CellArgs, built on the host by fm_capi.hip). A
// structure object may hold a batch of structures: atoms of structure k are
// the contiguous range [atom_off[k], atom_off[k+1]) and only see each other.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void row3(const double *a, const double *B, double *o) {
  for (int k = 0; k < 3; k++) o[k] = (a[0] * B[k] + a[1] * B[3 + k]) + a[2] * B[6 + k];
}

// structure.cpp:120-133 per atom, plus the global bin of the wrapped position.
__global__ void k_wrap(int n, const double *pos, const CellArgs *cas, const int *atom_struct,
                       const int *bin_off, double *wrapped, int *bin) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CellArgs &ca = cas[atom_struct[i]];
  double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  double r1[3], fr[3], w1[3], w[3];
  row3(p, ca.ct, r1);
  row3(r1, ca.cdi, fr);
  for (int k = 0; k < 3; k++) fr[k] = fr[k] - floor(fr[k]);
  row3(fr, ca.cd, w1);
  row3(w1, ca.cti, w);
  for (int k = 0; k < 3; k++) wrapped[3 * i + k] = w[k];
  // Bin from the fractional coordinates of the wrapped position.
  row3(w, ca.ct, r1);
  row3(r1, ca.cdi, fr);
  int q[3];
  for (int k = 0; k < 3; k++) {
    int b = (int)floor(fr[k] * ca.nb[k]);
    q[k] = min(max(b, 0), ca.nb[k] - 1);
  }
  bin[i] = bin_off[atom_struct[i]] + (q[0] * ca.nb[1] + q[1]) * ca.nb[2] + q[2];
}

__global__ void k_bin_count(int n, const int *bin, int *bcount) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&bcount[bin[i]], 1);
}

__global__ void k_bin_fill(int n, const int *bin, const int *bstart, int *cursor, int *order) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    int b = bin[i];
    int slot = atomicAdd(&cursor[b], 1);
    order[bstart[b] + slot] = i;
  }
}

__device__ __forceinline__ int floor_div(int a, int b) {
  int q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
  return q;
}

// structure.cpp:72-75.
__device__ __forceinline__ double image_dist(const double *wi, const double *wj,
                                             const CellArgs &ca, int s1, int s2, int s3,
                                             double *im) {
  for (int k = 0; k < 3; k++) {
    double d = wj[k] - wi[k];
    im[k] = ((d + (double)s1 * ca.cell[k]) + (double)s2 * ca.cell[3 + k]) +
            (double)s3 * ca.cell[6 + k];
  }
  return sqrt(im[0] * im[0] + im[1] * im[1] + im[2] * im[2]);
}

__device__ __forceinline__ long long nbr_key(int j, int s1, int s2, int s3, int sw) {
  int span = 2 * sw + 1;
  long long key = ((long long)j * span + (s1 + sw)) * span;
  return (key + (s2 + sw)) * span + (s3 + sw);
}

__device__ __forceinline__ void nbr_unkey(long long key, int sw, int &j, int &s1, int &s2,
                                          int &s3) {
  int span = 2 * sw + 1;
  s3 = (int)(key % span) - sw;
  key /= span;
  s2 = (int)(key % span) - sw;
  key /= span;
  s1 = (int)(key % span) - sw;
  j = (int)(key / span);
}

// One wavefront per atom. kFill = false counts the neighbours; kFill = true
// collects the keys in LDS, sorts them and writes the list. The stencil's
// bins (up to kStencilCap per pass) are laid out in LDS with a prefix sum of
// their sizes, so the wave's 64 lanes take consecutive candidates across
// bins instead of one bin at a time.
// Keys per atom held in LDS: lists of up to kNbrCapSmall (every common
// setting) use the small-LDS instance (seven workgroups per CU instead of
// two), up to kNbrCap the large one, longer lists the fallback below.
constexpr int kNbrCapSmall = 512;
constexpr int kNbrCap = 2048;
constexpr int kNbrWaves = 4;
constexpr int kStencilCap = 128;

// With kFill and `slots` set (one pass instead of count + fill): each atom's
// count goes to count[i] and, when it fits CAP, its sorted keys to
// slots[i CAP, ...), for k_neighbors_emit to write at the scanned offsets.
template <bool kFill, int CAP = kNbrCap>
__global__ __launch_bounds__(64 * kNbrWaves) void k_neighbors(
    int n, const double *wrapped, const int *bin_of, const int *bstart, const int *order,
    const int *species, const CellArgs *cas, const int *atom_struct, const int *bin_off,
    int *count, const int *cum, int *sidx, int *nspec, double *rel, int *max_count,
    long long *slots = nullptr) {
  __shared__ long long keys[kNbrWaves][kFill ? CAP : 1];
  __shared__ int s_off[kNbrWaves][kStencilCap + 1];
  __shared__ int s_p0[kNbrWaves][kStencilCap];
  __shared__ int s_img[kNbrWaves][kStencilCap];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int i = blockIdx.x * kNbrWaves + w;
  if (i >= n) return;
  int st = atom_struct[i];
  const CellArgs &ca = cas[st];
  int boff = bin_off[st];
  double wi[3] = {wrapped[3 * i], wrapped[3 * i + 1], wrapped[3 * i + 2]};
  int b = bin_of[i] - boff;
  int bx = b / (ca.nb[1] * ca.nb[2]), by = (b / ca.nb[2]) % ca.nb[1], bz = b % ca.nb[2];
  int sw = ca.sweep, span = 2 * sw + 1;
  int e0 = 2 * ca.K[0] + 1, e1 = 2 * ca.K[1] + 1, e2 = 2 * ca.K[2] + 1;
  int nst = e0 * e1 * e2;
  int found = 0;
  for (int sbase = 0; sbase < nst; sbase += kStencilCap) {
    int scount = min(kStencilCap, nst - sbase);
    if (sbase > 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    for (int e = lane; e < scount; e += 64) {
      int sidx_ = sbase + e;
      int o2 = sidx_ % e2 - ca.K[2], o1 = (sidx_ / e2) % e1 - ca.K[1],
          o0 = sidx_ / (e1 * e2) - ca.K[0];
      int t0 = bx + o0, s1 = floor_div(t0, ca.nb[0]), b0 = t0 - s1 * ca.nb[0];
      int t1 = by + o1, s2 = floor_div(t1, ca.nb[1]), b1 = t1 - s2 * ca.nb[1];
      int t2 = bz + o2, s3 = floor_div(t2, ca.nb[2]), b2 = t2 - s3 * ca.nb[2];
      bool valid = s1 >= -sw && s1 <= sw && s2 >= -sw && s2 <= sw && s3 >= -sw && s3 <= sw;
      int size = 0, p0 = 0, img = 0;
      if (valid) {
        int q = boff + (b0 * ca.nb[1] + b1) * ca.nb[2] + b2;
        p0 = bstart[q];
        size = bstart[q + 1] - p0;
        img = ((s1 + sw) * span + (s2 + sw)) * span + (s3 + sw);
      }
      s_p0[w][e] = p0;
      s_img[w][e] = img;
      s_off[w][e + 1] = size;
    }
    if (lane == 0) s_off[w][0] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // Inclusive prefix sum of the sizes, 64 entries at a time.
    int carry = 0;
    for (int cb = 0; cb < scount; cb += 64) {
      int e = cb + lane;
      int v = e < scount ? s_off[w][e + 1] : 0;
      for (int d = 1; d < 64; d <<= 1) {
        int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
      }
      if (e < scount) s_off[w][e + 1] = v + carry;
      carry += __shfl(v, 63, 64);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int total = s_off[w][scount];
    for (int c0 = 0; c0 < total; c0 += 64) {
      int t = c0 + lane;
      bool ok = false;
      int j = 0, s1 = 0, s2 = 0, s3 = 0;
      if (t < total) {
        // The bin holding candidate t: s_off[lo] <= t < s_off[lo + 1].
        int lo = 0, hi = scount;
        while (hi - lo > 1) {
          int mid = (lo + hi) >> 1;
          if (s_off[w][mid] <= t) lo = mid;
          else hi = mid;
        }
        j = order[s_p0[w][lo] + (t - s_off[w][lo])];
        int img = s_img[w][lo];
        s3 = img % span - sw;
        s2 = (img / span) % span - sw;
        s1 = img / (span * span) - sw;
        double wj[3] = {wrapped[3 * j], wrapped[3 * j + 1], wrapped[3 * j + 2]};
        double im[3];
        double dist = image_dist(wi, wj, ca, s1, s2, s3, im);
        ok = (dist < ca.cutoff) && (dist != 0);
      }
      unsigned long long mask = __ballot(ok);
      if (kFill && ok) {
        int pre = __popcll(mask & ((1ull << lane) - 1ull));
        int slot = found + pre;
        if (slot < CAP) keys[w][slot] = nbr_key(j, s1, s2, s3, sw);
      }
      found += __popcll(mask);
    }
  }
  if (!kFill || slots) {
    if (lane == 0) {
      count[i] = found;
      // Only lists longer than the small LDS capacity need another instance.
      if (found > kNbrCapSmall) atomicMax(max_count, found);
    }
    if (!kFill || found > CAP) return;
  }
  // Bitonic sort of the keys (padded to a power of two with the largest key).
  int np2 = 1;
  while (np2 < found) np2 <<= 1;
  for (int t = found + lane; t < np2; t += 64) keys[w][t] = 0x7fffffffffffffffLL;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int k = 2; k <= np2; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int t = lane; t < np2; t += 64) {
        int partner = t ^ jj;
        if (partner > t) {
          long long a = keys[w][t], c = keys[w][partner];
          bool up = (t & k) == 0;
          if ((a > c) == up) {
            keys[w][t] = c;
            keys[w][partner] = a;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  if (slots) {
    for (int t = lane; t < found; t += 64) slots[(size_t)i * CAP + t] = keys[w][t];
    return;
  }
  int base = cum[i];
  for (int t = lane; t < found; t += 64) {
    int j, s1, s2, s3;
    nbr_unkey(keys[w][t], sw, j, s1, s2, s3);
    double wj[3] = {wrapped[3 * j], wrapped[3 * j + 1], wrapped[3 * j + 2]};
    double im[3];
    double dist = image_dist(wi, wj, ca, s1, s2, s3, im);
    int o = base + t;
    sidx[o] = j;
    nspec[o] = species[j];
    rel[4 * (size_t)o] = dist;
    rel[4 * (size_t)o + 1] = im[0];
    rel[4 * (size_t)o + 2] = im[1];
    rel[4 * (size_t)o + 3] = im[2];
  }
}

// The lists parked by the one-pass k_neighbors: one wavefront per atom writes
// its sorted keys at cum[i] with the same image expression as the fill.
__global__ __launch_bounds__(64 * kNbrWaves) void k_neighbors_emit(
    int n, const double *wrapped, const int *species, const CellArgs *cas,
    const int *atom_struct, const int *count, const int *cum, const long long *slots, int cap,
    int *sidx, int *nspec, double *rel) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int i = blockIdx.x * kNbrWaves + w;
  if (i >= n) return;
  const CellArgs &ca = cas[atom_struct[i]];
  int sw = ca.sweep, cnt = count[i], base = cum[i];
  double wi[3] = {wrapped[3 * i], wrapped[3 * i + 1], wrapped[3 * i + 2]};
  for (int t = lane; t < cnt; t += 64) {
    int j, s1, s2, s3;
    nbr_unkey(slots[(size_t)i * cap + t], sw, j, s1, s2, s3);
    double wj[3] = {wrapped[3 * j], wrapped[3 * j + 1], wrapped[3 * j + 2]};
    double im[3];
    double dist = image_dist(wi, wj, ca, s1, s2, s3, im);
    int o = base + t;
    sidx[o] = j;
    nspec[o] = species[j];
    rel[4 * (size_t)o] = dist;
    rel[4 * (size_t)o + 1] = im[0];
    rel[4 * (size_t)o + 2] = im[1];
    rel[4 * (size_t)o + 3] = im[2];
  }
}

// Fallback for atoms with more than kNbrCap neighbours (tiny cells with a
// large sweep): one thread per atom, keys sorted in place in global memory.
__global__ void k_neighbors_big(int n, const double *wrapped, const int *bin_of,
                                const int *bstart, const int *order, const int *species,
                                const CellArgs *cas, const int *atom_struct, const int *bin_off,
                                const int *cum, long long *scratch, int *sidx, int *nspec,
                                double *rel) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int st = atom_struct[i];
  const CellArgs &ca = cas[st];
  int boff = bin_off[st];
  double wi[3] = {wrapped[3 * i], wrapped[3 * i + 1], wrapped[3 * i + 2]};
  int b = bin_of[i] - boff;
  int bx = b / (ca.nb[1] * ca.nb[2]), by = (b / ca.nb[2]) % ca.nb[1], bz = b % ca.nb[2];
  int sw = ca.sweep;
  long long *keys = scratch + cum[i];
  int cnt = cum[i + 1] - cum[i];
  int found = 0;
  for (int o0 = -ca.K[0]; o0 <= ca.K[0]; o0++) {
    int t0 = bx + o0, s1 = floor_div(t0, ca.nb[0]), b0 = t0 - s1 * ca.nb[0];
    if (s1 < -sw || s1 > sw) continue;
    for (int o1 = -ca.K[1]; o1 <= ca.K[1]; o1++) {
      int t1 = by + o1, s2 = floor_div(t1, ca.nb[1]), b1 = t1 - s2 * ca.nb[1];
      if (s2 < -sw || s2 > sw) continue;
      for (int o2 = -ca.K[2]; o2 <= ca.K[2]; o2++) {
        int t2 = bz + o2, s3 = floor_div(t2, ca.nb[2]), b2 = t2 - s3 * ca.nb[2];
        if (s3 < -sw || s3 > sw) continue;
        int q = boff + (b0 * ca.nb[1] + b1) * ca.nb[2] + b2;
        for (int p = bstart[q]; p < bstart[q + 1]; p++) {
          int j = order[p];
          double wj[3] = {wrapped[3 * j], wrapped[3 * j + 1], wrapped[3 * j + 2]};
          double im[3];
          double dist = image_dist(wi, wj, ca, s1, s2, s3, im);
          if ((dist < ca.cutoff) && (dist != 0) && found < cnt)
            keys[found++] = nbr_key(j, s1, s2, s3, sw);
        }
      }
    }
  }
  for (int a = 1; a < found; a++) {  // insertion sort
    long long v = keys[a];
    int c = a - 1;
    while (c >= 0 && keys[c] > v) {
      keys[c + 1] = keys[c];
      c--;
    }
    keys[c + 1] = v;
  }
  for (int t = 0; t < found; t++) {
    int j, s1, s2, s3;
    nbr_unkey(keys[t], sw, j, s1, s2, s3);
    double wj[3] = {wrapped[3 * j], wrapped[3 * j + 1], wrapped[3 * j + 2]};
    double im[3];
    double dist = image_dist(wi, wj, ca, s1, s2, s3, im);
    int o = cum[i] + t;
    sidx[o] = j;
    nspec[o] = species[j];
    rel[4 * (size_t)o] = dist;
    rel[4 * (size_t)o + 1] = im[0];
    rel[4 * (size_t)o + 2] = im[1];
    rel[4 * (size_t)o + 3] = im[2];
  }
}

void build_neighbors(fm_structure s) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  int n = s->n;
  Pool *pool = &ctx->pool;
  // Global bin offsets of the structures of the batch.
  std::vector<int> bin_off(s->n_struct + 1, 0);
  for (int k = 0; k < s->n_struct; k++) {
    const CellArgs &ca = s->cellargs[k];
    long long nb = (long long)ca.nb[0] * ca.nb[1] * ca.nb[2];
    FM_CHECK(bin_off[k] + nb < (1LL << 30), FM_ERR_CAPACITY, "cell list too large");
    bin_off[k + 1] = bin_off[k] + (int)nb;
  }
  int ncell = bin_off[s->n_struct];
  s->d_bin_off.alloc(pool, bin_off.size() * sizeof(int));
  FM_HIP(hipMemcpyAsync(s->d_bin_off.as<int>(), bin_off.data(), bin_off.size() * sizeof(int),
                        hipMemcpyHostToDevice, st));
  const CellArgs *cas = s->d_cellargs.as<CellArgs>();
  const int *ast = s->atom_struct.as<int>();
  const int *boffp = s->d_bin_off.as<int>();
  s->wrapped.alloc(pool, 3 * (size_t)n * sizeof(double));
  s->count.alloc(pool, (size_t)n * sizeof(int));
  s->cumulative.alloc(pool, ((size_t)n + 1) * sizeof(int));
  DevBuf bin(pool, (size_t)n * sizeof(int));
  DevBuf bcount(pool, (size_t)ncell * sizeof(int));
  DevBuf bstart(pool, ((size_t)ncell + 1) * sizeof(int));
  DevBuf cursor(pool, (size_t)ncell * sizeof(int));
  DevBuf order(pool, (size_t)n * sizeof(int));
  DevBuf maxc(pool, sizeof(int));
  int tpb = 256, nblk = (n + tpb - 1) / tpb;
  hipLaunchKernelGGL(k_wrap, dim3(nblk), dim3(tpb), 0, st, n, s->positions.as<double>(), cas, ast,
                     boffp, s->wrapped.as<double>(), bin.as<int>());
  FM_HIP(hipMemsetAsync(bcount.as<int>(), 0, (size_t)ncell * sizeof(int), st));
  FM_HIP(hipMemsetAsync(cursor.as<int>(), 0, (size_t)ncell * sizeof(int), st));
  FM_HIP(hipMemsetAsync(maxc.as<int>(), 0, sizeof(int), st));
  hipLaunchKernelGGL(k_bin_count, dim3(nblk), dim3(tpb), 0, st, n, bin.as<int>(), bcount.as<int>());
  scan_exclusive(ctx, bcount.as<int>(), bstart.as<int>(), ncell, false);
  hipLaunchKernelGGL(k_bin_fill, dim3(nblk), dim3(tpb), 0, st, n, bin.as<int>(), bstart.as<int>(),
                     cursor.as<int>(), order.as<int>());
  int wblk = (n + kNbrWaves - 1) / kNbrWaves;
  // One pass over the candidates that also parks each atom's sorted keys
  // (kNbrCapSmall per atom, while that area stays within 16 GB); else a
  // counting pass first.
  bool one_pass = (size_t)n * kNbrCapSmall * sizeof(long long) <= (16ull << 30);
  DevBuf slots;
  if (one_pass) {
    slots.alloc(pool, std::max<size_t>(1, (size_t)n * kNbrCapSmall * sizeof(long long)));
    hipLaunchKernelGGL((k_neighbors<true, kNbrCapSmall>), dim3(wblk), dim3(64 * kNbrWaves), 0, st,
                       n, s->wrapped.as<double>(), bin.as<int>(), bstart.as<int>(),
                       order.as<int>(), s->species.as<int>(), cas, ast, boffp,
                       s->count.as<int>(), nullptr, nullptr, nullptr, nullptr, maxc.as<int>(),
                       slots.as<long long>());
  } else {
    hipLaunchKernelGGL(k_neighbors<false>, dim3(wblk), dim3(64 * kNbrWaves), 0, st, n,
                       s->wrapped.as<double>(), bin.as<int>(), bstart.as<int>(), order.as<int>(),
                       s->species.as<int>(), cas, ast, boffp, s->count.as<int>(), nullptr,
                       nullptr, nullptr, nullptr, maxc.as<int>());
  }
  FM_HIP(hipGetLastError());
  scan_exclusive(ctx, s->count.as<int>(), s->cumulative.as<int>(), n, false);
  // Total and largest count in pinned words [2, 4), one synchronisation
  // (which also completes any readback queued before, e.g. words [0, 2)).
  int *hw = ctx->pinned;
  FM_HIP(hipMemcpyAsync(hw + 2, s->cumulative.as<int>() + n, sizeof(int), hipMemcpyDeviceToHost,
                        st));
  FM_HIP(hipMemcpyAsync(hw + 3, maxc.as<int>(), sizeof(int), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  int total = hw[2], max_count = hw[3];
  s->n_nb = total;
  s->sidx.alloc(pool, std::max<size_t>(1, (size_t)total * sizeof(int)));
  s->nspecies.alloc(pool, std::max<size_t>(1, (size_t)total * sizeof(int)));
  s->rel.alloc(pool, std::max<size_t>(1, 4 * (size_t)total * sizeof(double)));
  if (max_count <= kNbrCapSmall && one_pass) {
    hipLaunchKernelGGL(k_neighbors_emit, dim3(wblk), dim3(64 * kNbrWaves), 0, st, n,
                       s->wrapped.as<double>(), s->species.as<int>(), cas, ast,
                       s->count.as<int>(), s->cumulative.as<int>(), slots.as<long long>(),
                       kNbrCapSmall, s->sidx.as<int>(), s->nspecies.as<int>(),
                       s->rel.as<double>());
  } else if (max_count <= kNbrCapSmall) {
    hipLaunchKernelGGL((k_neighbors<true, kNbrCapSmall>), dim3(wblk), dim3(64 * kNbrWaves), 0, st,
                       n, s->wrapped.as<double>(), bin.as<int>(), bstart.as<int>(),
                       order.as<int>(), s->species.as<int>(), cas, ast, boffp, nullptr,
                       s->cumulative.as<int>(), s->sidx.as<int>(), s->nspecies.as<int>(),
                       s->rel.as<double>(), nullptr);
  } else if (max_count <= kNbrCap) {
    hipLaunchKernelGGL((k_neighbors<true, kNbrCap>), dim3(wblk), dim3(64 * kNbrWaves), 0, st, n,
                       s->wrapped.as<double>(), bin.as<int>(), bstart.as<int>(), order.as<int>(),
                       s->species.as<int>(), cas, ast, boffp, nullptr, s->cumulative.as<int>(),
                       s->sidx.as<int>(), s->nspecies.as<int>(), s->rel.as<double>(), nullptr);
  } else {
    DevBuf scratch(pool, (size_t)total * sizeof(long long));
    hipLaunchKernelGGL(k_neighbors_big, dim3(nblk), dim3(tpb), 0, st, n, s->wrapped.as<double>(),
                       bin.as<int>(), bstart.as<int>(), order.as<int>(), s->species.as<int>(), cas,
                       ast, boffp, s->cumulative.as<int>(), scratch.as<long long>(),
                       s->sidx.as<int>(), s->nspecies.as<int>(), s->rel.as<double>());
    FM_HIP(hipStreamSynchronize(st));
  }
  FM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Descriptor bond list (b2.cpp:262-304: r <= rc(s_i, s_j), list order kept),
// its transpose by neighbour atom (for the deterministic force gather), and
// the species-sorted row of each atom (b2.cpp:148-176).
// ---------------------------------------------------------------------------
// Bonds are the neighbours within the species-pair cutoff, kept in neighbour
// order. One wave per atom walks its neighbour list 64 entries at a time; a
// ballot gives each kept entry its place.
__global__ __launch_bounds__(256) void k_bond_count(int n, const int *species, const int *cum,
                                                    const int *nspec, const double *rel,
                                                    const double *cutoffs, int nos, int *bcount) {
  int lane = threadIdx.x & 63;
  int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  int si = species[i], c = 0;
  for (int q0 = cum[i]; q0 < cum[i + 1]; q0 += 64) {
    int q = q0 + lane;
    bool keep = q < cum[i + 1] && rel[4 * (size_t)q] <= cutoffs[si * nos + nspec[q]];
    c += __popcll(__ballot(keep));
  }
  if (lane == 0) bcount[i] = c;
}

__global__ __launch_bounds__(256) void k_bond_fill(int n, const int *species, const int *cum,
                                                   const int *nspec, const double *rel,
                                                   const double *cutoffs, int nos,
                                                   const int *boff, const int *sidx, int *bond_q,
                                                   int *in_count) {
  int lane = threadIdx.x & 63;
  int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  int si = species[i], o = boff[i];
  unsigned long long below = (1ull << lane) - 1ull;
  for (int q0 = cum[i]; q0 < cum[i + 1]; q0 += 64) {
    int q = q0 + lane;
    bool keep = q < cum[i + 1] && rel[4 * (size_t)q] <= cutoffs[si * nos + nspec[q]];
    unsigned long long m = __ballot(keep);
    if (keep) {
      bond_q[o + __popcll(m & below)] = q;
      atomicAdd(&in_count[sidx[q]], 1);
    }
    o += __popcll(m);
  }
}

__global__ void k_in_fill(int nbonds, const int *bond_q, const int *sidx, const int *in_off,
                          int *cursor, int *in_bond) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbonds) return;
  int j = sidx[bond_q[b]];
  int slot = atomicAdd(&cursor[j], 1);
  in_bond[in_off[j] + slot] = b;
}

// Incoming bonds of atom j in ascending bond order (so the force gather sums
// them in a fixed order): one wave per atom; each lane takes elements of the
// unordered list and writes each to its rank, the number of smaller bond ids
// in the list (ids are distinct). The list is read with wave-uniform loads.
__global__ __launch_bounds__(256) void k_in_rank(int n, const int *in_off, const int *unordered,
                                                 int *in_bond) {
  int lane = threadIdx.x & 63;
  int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  int a0 = in_off[j], cnt = in_off[j + 1] - a0;
  const int *u = unordered + a0;
  for (int e = lane; e < cnt; e += 64) {
    int v = u[e], r = 0;
    for (int k = 0; k < cnt; k++) r += u[k] < v;
    in_bond[a0 + r] = v;
  }
}

__global__ void k_species_flag(int n, const int *species, int t, int *flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = species[i] == t ? 1 : 0;
}

__global__ void k_species_rows(int n, const int *species, int t, const int *rank, int off,
                               int *row_of_atom, int *atom_of_row) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && species[i] == t) {
    int r = off + rank[i];
    row_of_atom[i] = r;
    atom_of_row[r] = i;
  }
}

void build_bonds(fm_structure s) {
  fm_context ctx = s->ctx;
  hipStream_t st = ctx->stream;
  Pool *pool = &ctx->pool;
  int n = s->n, tpb = 256, nblk = (n + tpb - 1) / tpb;
  FM_CHECK(s->nos + 1 <= fm_context_s::kPinnedWords, FM_ERR_INVALID, "too many species");
  s->d_cutoffs.alloc(pool, s->cutoffs.size() * sizeof(double));
  FM_HIP(hipMemcpyAsync(s->d_cutoffs.as<double>(), s->cutoffs.data(),
                        s->cutoffs.size() * sizeof(double), hipMemcpyHostToDevice, st));
  // Species-sorted rows: one scan of species flags per species.
  s->row_of_atom.alloc(pool, (size_t)n * sizeof(int));
  s->atom_of_row.alloc(pool, (size_t)n * sizeof(int));
  DevBuf flag(pool, (size_t)n * sizeof(int));
  DevBuf rank(pool, (size_t)s->nos * ((size_t)n + 1) * sizeof(int));
  for (int t = 0; t < s->nos; t++) {
    hipLaunchKernelGGL(k_species_flag, dim3(nblk), dim3(tpb), 0, st, n, s->species.as<int>(), t,
                       flag.as<int>());
    scan_exclusive(ctx, flag.as<int>(), rank.as<int>() + (size_t)t * (n + 1), n, false);
  }
  DevBuf bcount(pool, (size_t)n * sizeof(int));
  s->bond_off.alloc(pool, ((size_t)n + 1) * sizeof(int));
  hipLaunchKernelGGL(k_bond_count, dim3((n + 3) / 4), dim3(256), 0, st, n, s->species.as<int>(),
                     s->cumulative.as<int>(), s->nspecies.as<int>(), s->rel.as<double>(),
                     s->d_cutoffs.as<double>(), s->nos, bcount.as<int>());
  scan_exclusive(ctx, bcount.as<int>(), s->bond_off.as<int>(), n, false);
  // Bond total and per-species atom counts: one readback, one synchronisation.
  int *hw = ctx->pinned;
  FM_HIP(hipMemcpyAsync(hw, s->bond_off.as<int>() + n, sizeof(int), hipMemcpyDeviceToHost, st));
  for (int t = 0; t < s->nos; t++)
    FM_HIP(hipMemcpyAsync(hw + 1 + t, rank.as<int>() + (size_t)t * (n + 1) + n, sizeof(int),
                          hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  int nbonds = hw[0];
  s->n_bonds = nbonds;
  s->n_by_type.assign(s->nos, 0);
  s->type_off.assign(s->nos + 1, 0);
  for (int t = 0; t < s->nos; t++) {
    s->n_by_type[t] = hw[1 + t];
    s->type_off[t + 1] = s->type_off[t] + hw[1 + t];
    hipLaunchKernelGGL(k_species_rows, dim3(nblk), dim3(tpb), 0, st, n, s->species.as<int>(), t,
                       rank.as<int>() + (size_t)t * (n + 1), s->type_off[t],
                       s->row_of_atom.as<int>(), s->atom_of_row.as<int>());
  }
  s->bond_q.alloc(pool, (size_t)nbonds * sizeof(int));
  DevBuf in_count(pool, (size_t)n * sizeof(int));
  FM_HIP(hipMemsetAsync(in_count.as<int>(), 0, (size_t)n * sizeof(int), st));
  hipLaunchKernelGGL(k_bond_fill, dim3((n + 3) / 4), dim3(256), 0, st, n, s->species.as<int>(),
                     s->cumulative.as<int>(), s->nspecies.as<int>(), s->rel.as<double>(),
                     s->d_cutoffs.as<double>(), s->nos, s->bond_off.as<int>(), s->sidx.as<int>(),
                     s->bond_q.as<int>(), in_count.as<int>());
  s->in_off.alloc(pool, ((size_t)n + 1) * sizeof(int));
  scan_exclusive(ctx, in_count.as<int>(), s->in_off.as<int>(), n, false);
  s->in_bond.alloc(pool, (size_t)nbonds * sizeof(int));
  DevBuf cursor(pool, (size_t)n * sizeof(int));
  DevBuf unordered(pool, (size_t)nbonds * sizeof(int));
  FM_HIP(hipMemsetAsync(cursor.as<int>(), 0, (size_t)n * sizeof(int), st));
  if (nbonds > 0) {
    hipLaunchKernelGGL(k_in_fill, dim3((nbonds + tpb - 1) / tpb), dim3(tpb), 0, st, nbonds,
                       s->bond_q.as<int>(), s->sidx.as<int>(), s->in_off.as<int>(),
                       cursor.as<int>(), unordered.as<int>());
    hipLaunchKernelGGL(k_in_rank, dim3((n + 3) / 4), dim3(256), 0, st, n, s->in_off.as<int>(),
                       unordered.as<int>(), s->in_bond.as<int>());
  }
  FM_HIP(hipGetLastError());
}

}  // namespace fm
