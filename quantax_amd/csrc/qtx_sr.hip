// SR / MinSR: Omean, Ebar, the MFMA Gram of Obar, rocSOLVER Cholesky / eigh,
// the update vector and the guarded parameter update.
//
// Obar = (O - mean) * sqrt(r / ns) is never materialised: the Gram kernel
// centres each O tile in FP64 while staging it into LDS (then rounds to FP32
// or splits it into bf16 hi/lo), and the update vector x = Obar^T y reads O
// once more with FP64 accumulation.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "common.hpp"

namespace qtx {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------- helpers
template <typename T>
__device__ __forceinline__ double ld_as_double(const T* p, int64_t i) {
  return (double)p[i];
}

__global__ void fill_f64_kernel(double* p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// column sums over a chunk of rows, accumulated with FP64 atomics
template <typename T>
__global__ __launch_bounds__(256) void colsum_kernel(const T* __restrict__ O, int64_t ns, int64_t np,
                                                     int64_t rows_per_chunk, double* __restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  int64_t r1 = r0 + rows_per_chunk < ns ? r0 + rows_per_chunk : ns;
  double acc = 0.0;
  for (int64_t r = r0; r < r1; ++r) acc += (double)O[r * np + k];
  atomicAdd(out + k, acc);
}

__global__ void scale_f64_kernel(double* p, int64_t n, double a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] *= a;
}

// single-block: energy, VarE and Ebar (sr.py:180-195)
__global__ __launch_bounds__(1024) void ebar_kernel(const double* __restrict__ E, const double* __restrict__ r,
                                                    int64_t ns, double* __restrict__ ebar, double* __restrict__ out2) {
  __shared__ double red[3][16];
  int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE, nw = blockDim.x / WAVE;
  double s_er = 0, s_e = 0;
  for (int64_t i = tid; i < ns; i += blockDim.x) {
    double ri = r ? r[i] : 1.0;
    s_er += E[i] * ri;
    s_e += E[i];
  }
  s_er = wave_sum(s_er);
  s_e = wave_sum(s_e);
  if (lane == 0) {
    red[0][wv] = s_er;
    red[1][wv] = s_e;
  }
  __syncthreads();
  double tot_er = 0, tot_e = 0;
  for (int w = 0; w < nw; ++w) {
    tot_er += red[0][w];
    tot_e += red[1][w];
  }
  double Emean = tot_er / (double)ns, Eplain = tot_e / (double)ns;
  double s_v = 0;
  for (int64_t i = tid; i < ns; i += blockDim.x) {
    double ri = r ? r[i] : 1.0;
    double d = E[i] - Emean;
    s_v += d * d * ri;
    ebar[i] = (E[i] - Eplain) * sqrt(ri / (double)ns);
  }
  s_v = wave_sum(s_v);
  __syncthreads();
  if (lane == 0) red[2][wv] = s_v;
  __syncthreads();
  if (tid == 0) {
    double tv = 0;
    for (int w = 0; w < nw; ++w) tv += red[2][w];
    out2[0] = Emean;
    out2[1] = tv / (double)ns;
  }
}

// ---------------------------------------------------------------- Gram T = A A^T
// A = Obar (n x np) from O rows.  Workgroup: 256 threads = 2 x 2 waves, tile
// 128 x 128 (each wave 64 x 64 = 4 x 4 MFMA 16x16 blocks); only tiles with
// bi >= bj are computed and mirrored.  Output T is FP64 (n x n, symmetric).
constexpr int GT = 128;

template <int G>
struct GramCfg;
template <>
struct GramCfg<QTX_GRAM_FP32> {
  static constexpr int KC = 32;   // K elements staged per chunk
  static constexpr int PAD = 1;   // LDS row padding (elements)
};
template <>
struct GramCfg<QTX_GRAM_FP64> {
  static constexpr int KC = 16;
  static constexpr int PAD = 1;
};

__device__ __forceinline__ void tile_of(int bid, int& bi, int& bj) {
  // bid -> (bi, bj) with bj <= bi, row-major over the lower triangle
  int b = (int)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
  while ((b + 1) * (b + 2) / 2 <= bid) ++b;
  while (b * (b + 1) / 2 > bid) --b;
  bi = b;
  bj = bid - b * (b + 1) / 2;
}

template <typename TI, int G>
__global__ __launch_bounds__(256) void gram_nt_kernel(const TI* __restrict__ O, const double* __restrict__ mean,
                                                      const double* __restrict__ r, int64_t n, int64_t np,
                                                      double inv_ns, double* __restrict__ T) {
  constexpr int KC = GramCfg<G>::KC;
  constexpr int LD = KC + GramCfg<G>::PAD;
  using ST = typename std::conditional<G == QTX_GRAM_FP64, double, float>::type;
  __shared__ ST As[1][GT * LD];
  __shared__ ST Bs[1][GT * LD];

  int bi, bj;
  tile_of(blockIdx.x, bi, bj);
  const int64_t i0 = (int64_t)bi * GT, j0 = (int64_t)bj * GT;
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  const int wr = (wv / 2) * 64, wc = (wv % 2) * 64;  // wave's sub-tile origin

  using ACC = typename std::conditional<G == QTX_GRAM_FP64, v4d, v4f>::type;
  ACC acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = ACC{0, 0, 0, 0};
  // the FP32 mode flushes its MFMA accumulators into FP64 every FLUSH chunks,
  // so the rounding error does not grow with np
  constexpr int FLUSH = 8;
  constexpr bool SHADOW = (G != QTX_GRAM_FP64);
  v4d dacc[SHADOW ? 4 : 1][SHADOW ? 4 : 1];
  if constexpr (SHADOW) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) dacc[a][b] = v4d{0, 0, 0, 0};
  }
  int chunk = 0;

  for (int64_t k0 = 0; k0 < np; k0 += KC) {
    // stage A rows [i0, i0+128) and B rows [j0, j0+128), columns [k0, k0+KC)
    for (int e = tid; e < GT * KC; e += 256) {
      int row = e / KC, kk = e % KC;
      int64_t k = k0 + kk;
#pragma unroll
      for (int which = 0; which < 2; ++which) {
        int64_t gr = (which == 0 ? i0 : j0) + row;
        double v = 0.0;
        if (gr < n && k < np) {
          double sc = sqrt((r ? r[gr] : 1.0) * inv_ns);
          v = ((double)O[gr * np + k] - mean[k]) * sc;
        }
        ((which == 0) ? As[0] : Bs[0])[row * LD + kk] = (ST)v;
      }
    }
    __syncthreads();
    {
#pragma unroll
      for (int ks = 0; ks < KC; ks += 4) {
        ST av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = As[0][(wr + a * 16 + lane % 16) * LD + ks + lane / 16];
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = Bs[0][(wc + b * 16 + lane % 16) * LD + ks + lane / 16];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            if constexpr (G == QTX_GRAM_FP64)
              acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
            else
              acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
          }
      }
    }
    __syncthreads();
    if constexpr (SHADOW) {
      if (++chunk == FLUSH || k0 + KC >= np) {
        chunk = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
#pragma unroll
            for (int q = 0; q < 4; ++q) dacc[a][b][q] += (double)acc[a][b][q];
            acc[a][b] = ACC{0, 0, 0, 0};
          }
      }
    }
  }
  // store, column lane%16 of each 16x16 block; rows of element q:
  //   f32 / bf16 MFMA: 4 * (lane / 16) + q ;  f64 MFMA: lane / 16 + 4 * q
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rq = (G == QTX_GRAM_FP64) ? (lane / 16 + 4 * q) : (4 * (lane / 16) + q);
        int64_t gi = i0 + wr + a * 16 + rq;
        int64_t gj = j0 + wc + b * 16 + lane % 16;
        // off-diagonal tiles are mirrored; inside diagonal tiles the lower
        // triangle is written to both halves so T is exactly symmetric
        if (gi < n && gj < n && (bi != bj || gi >= gj)) {
          double v;
          if constexpr (SHADOW)
            v = dacc[a][b][q];
          else
            v = (double)acc[a][b][q];
          T[gi * n + gj] = v;
          T[gj * n + gi] = v;
        }
      }
}

// ---------------------------------------------------------------- split-bf16 Gram
// Pre-pass: Obar = (O - mean) sqrt(r / ns) in FP64, rounded to FP32 x and
// split into bf16 planes hi = bf16(x), lo = bf16(x - hi), row-major with rows
// padded to a multiple of GT and K to a multiple of G3_KC (zeros).  The Gram
// kernel then streams the planes straight into LDS and does only MFMA work:
// T_ij = sum_k lo_i hi_j + hi_i lo_j + hi_i hi_j (lo * lo dropped).
constexpr int G3_KC = 64;            // K per LDS stage (two 16x16x32 MFMA steps)
constexpr int G3_FLUSH = 4;          // stages per FP32 -> FP64 flush (256 K)
constexpr int G3_PLANE = GT * G3_KC * 2;  // bytes of one bf16 plane of one stage (16 KB)

static inline int64_t g3_rows(int64_t n) { return (n + GT - 1) / GT * GT; }
static inline int64_t g3_kpad(int64_t np) { return (np + G3_KC - 1) / G3_KC * G3_KC; }
static inline size_t g3_bytes(int64_t n, int64_t np) { return (size_t)g3_rows(n) * g3_kpad(np) * 2 * sizeof(__bf16); }

// Grid-stride over (row, 256-column block) pairs: a launch's work-item count
// is a 32-bit quantity, and rows * kpad passes 2^32 at 32768 x 286976.
template <typename TI>
__global__ __launch_bounds__(256) void split_bf16x3_kernel(const TI* __restrict__ O, const double* __restrict__ mean,
                                                           const double* __restrict__ r, int64_t n, int64_t np,
                                                           int64_t kpad, int64_t bpr, int64_t nblocks, double inv_ns,
                                                           __bf16* __restrict__ Hi, __bf16* __restrict__ Lo) {
  for (int64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
    const int64_t row = blk / bpr;
    const int64_t k = (blk % bpr) * 256 + threadIdx.x;
    if (k >= kpad) continue;
    double v = 0.0;
    if (row < n && k < np) v = ((double)O[row * np + k] - mean[k]) * sqrt((r ? r[row] : 1.0) * inv_ns);
    const float f = (float)v;
    const __bf16 hi = (__bf16)f;
    Hi[row * kpad + k] = hi;
    Lo[row * kpad + k] = (__bf16)(f - (float)hi);
  }
}

// lower-triangle tile t of the XCD-contiguous order: workgroups are dealt to
// the 8 XCDs round-robin, so block b runs on XCD b % 8; giving each XCD a
// contiguous run of tiles (shared row panels) keeps them in its L2
__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
  const int per = ntiles / 8, rem = ntiles % 8, x = bid % 8, loc = bid / 8;
  return x * per + (x < rem ? x : rem) + loc;
}

typedef __attribute__((address_space(3))) void lds_void;

// 128 x 128 tile, 4 waves (2 x 2, 64 x 64 each).  Operands move HBM -> LDS by
// LDS DMA (global_load_lds_dwordx4: lane l of a wave writes 16 bytes at the
// wave's LDS base + 16 l), two 64 KB stages so chunk c + 1 is in flight while
// chunk c is multiplied.  Wave w fetches plane w (A hi, A lo, B hi, B lo).  A
// plane is 128 rows x 128 bytes; the 16-byte segment s of row r sits in slot
// s ^ (r & 7), so the MFMA fragment reads (16 rows, one segment) hit 8
// distinct slots per 8 lanes: no bank conflicts.  The stage barriers are raw
// s_barrier with explicit vmcnt waits: a __syncthreads() fence would drain
// the next chunk's DMA as well.
__global__ __launch_bounds__(256) void gram_bf16x3_kernel(const __bf16* __restrict__ Hi,
                                                          const __bf16* __restrict__ Lo, int64_t n, int64_t kpad,
                                                          int ntiles, double* __restrict__ T) {
  __shared__ __attribute__((aligned(16))) char smem[2 * 4 * G3_PLANE];
  int bi, bj;
  tile_of(xcd_tile(blockIdx.x, ntiles), bi, bj);
  const int64_t i0 = (int64_t)bi * GT, j0 = (int64_t)bj * GT;
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  const int wr = (wv / 2) * 64, wc = (wv % 2) * 64;
  const __bf16* src = (wv == 0) ? Hi + i0 * kpad : (wv == 1) ? Lo + i0 * kpad : (wv == 2) ? Hi + j0 * kpad : Lo + j0 * kpad;
  // one DMA instruction covers rows 8q .. 8q+7: lane l -> row 8q + l/8, slot l%8,
  // which holds logical segment (l%8) ^ (l/8)
  const int lrow = lane / 8, lseg = (lane % 8) ^ (lane / 8);
  auto issue = [&](int64_t k0, int stage) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const __bf16* g = src + (int64_t)(q * 8 + lrow) * kpad + k0 + lseg * 8;
      char* dst = smem + (stage * 4 + wv) * G3_PLANE + q * 1024;
      __builtin_amdgcn_global_load_lds((const void*)g, (lds_void*)dst, 16, 0, 0);
    }
  };

  v4f acc[4][4];
  v4d dacc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      acc[a][b] = v4f{0, 0, 0, 0};
      dacc[a][b] = v4d{0, 0, 0, 0};
    }
  const int nchunks = (int)(kpad / G3_KC);
  issue(0, 0);
  if (nchunks > 1) issue(G3_KC, 1);
  for (int c = 0; c < nchunks; ++c) {
    // this wave's DMA for chunk c is done (16 newer ones may be in flight), then
    // every wave's is
    if (c + 1 < nchunks)
      asm volatile("s_waitcnt vmcnt(16)\n\ts_barrier" ::: "memory");
    else
      asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
    const char* stg = smem + (c & 1) * 4 * G3_PLANE;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      v8bf ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int row = wr + a * 16 + lane % 16, slot = (ks * 4 + lane / 16) ^ (row & 7);
        ah[a] = *(const v8bf*)(stg + 0 * G3_PLANE + row * 128 + slot * 16);
        al[a] = *(const v8bf*)(stg + 1 * G3_PLANE + row * 128 + slot * 16);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int row = wc + b * 16 + lane % 16, slot = (ks * 4 + lane / 16) ^ (row & 7);
        bh[b] = *(const v8bf*)(stg + 2 * G3_PLANE + row * 128 + slot * 16);
        bl[b] = *(const v8bf*)(stg + 3 * G3_PLANE + row * 128 + slot * 16);
      }
      // the three products of a block are 16 MFMAs apart (no back-to-back dependency)
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
    }
    if ((c + 1) % G3_FLUSH == 0 || c + 1 == nchunks) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
#pragma unroll
          for (int q = 0; q < 4; ++q) dacc[a][b][q] += (double)acc[a][b][q];
          acc[a][b] = v4f{0, 0, 0, 0};
        }
    }
    // every wave has read this stage (its fragments fed the MFMAs above)
    // before the DMA of chunk c + 2 overwrites it
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    if (c + 2 < nchunks) issue((int64_t)(c + 2) * G3_KC, c & 1);
  }
  // element q of a 16x16 block: row 4 * (lane / 16) + q, column lane % 16
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t gi = i0 + wr + a * 16 + 4 * (lane / 16) + q;
        const int64_t gj = j0 + wc + b * 16 + lane % 16;
        if (gi < n && gj < n && (bi != bj || gi >= gj)) {
          T[gi * n + gj] = dacc[a][b][q];
          T[gj * n + gi] = dacc[a][b][q];
        }
      }
}

// S = Obar^T Obar (np x np), FP64, small np (SR branch, ns >= np)
template <typename TI>
__global__ __launch_bounds__(256) void gram_tn_kernel(const TI* __restrict__ O, const double* __restrict__ mean,
                                                      const double* __restrict__ r, int64_t ns, int64_t np,
                                                      double inv_ns, double* __restrict__ S) {
  __shared__ double As[32][17], Bs[32][17];
  int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  int64_t k = (int64_t)blockIdx.y * 16 + ty, l = (int64_t)blockIdx.x * 16 + tx;
  double acc = 0.0;
  for (int64_t s0 = 0; s0 < ns; s0 += 32) {
    for (int e = threadIdx.x; e < 32 * 16; e += 256) {
      int ss = e / 16, c = e % 16;
      int64_t s = s0 + ss;
      int64_t ka = (int64_t)blockIdx.y * 16 + c, kb = (int64_t)blockIdx.x * 16 + c;
      double sc = (s < ns) ? sqrt((r ? r[s] : 1.0) * inv_ns) : 0.0;
      As[ss][c] = (s < ns && ka < np) ? ((double)O[s * np + ka] - mean[ka]) * sc : 0.0;
      Bs[ss][c] = (s < ns && kb < np) ? ((double)O[s * np + kb] - mean[kb]) * sc : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int ss = 0; ss < 32; ++ss) acc += As[ss][ty] * Bs[ss][tx];
    __syncthreads();
  }
  if (k < np && l < np) S[k * np + l] = acc;
}

// out[s, k] = (O[s, k] - mean[k]) sqrt(r_s / ns) in FP64 (grid-stride)
template <typename TI>
__global__ void form_obar_kernel(const TI* __restrict__ O, const double* __restrict__ mean,
                                 const double* __restrict__ r, int64_t ns, int64_t np, double inv_ns,
                                 double* __restrict__ out) {
  const int64_t total = ns * np;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = e / np, k = e - s * np;
    out[e] = ((double)O[e] - mean[k]) * sqrt((r ? r[s] : 1.0) * inv_ns);
  }
}

// x_k = sum_s (O_sk - mean_k) sqrt(r_s / ns) y_s  (FP64 accumulation)
template <typename TI>
__global__ __launch_bounds__(256) void obar_t_y_kernel(const TI* __restrict__ O, const double* __restrict__ mean,
                                                       const double* __restrict__ r, const double* __restrict__ y,
                                                       int64_t ns, int64_t np, int64_t rows_per_chunk,
                                                       double inv_ns, double* __restrict__ x) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  int64_t r1 = r0 + rows_per_chunk < ns ? r0 + rows_per_chunk : ns;
  double acc = 0.0, m = mean[k];
  for (int64_t s = r0; s < r1; ++s) {
    double w = y[s] * sqrt((r ? r[s] : 1.0) * inv_ns);
    acc += ((double)O[s * np + k] - m) * w;
  }
  atomicAdd(x + k, acc);
}

// y_s = beta y_s + alpha sqrt(r_s / ns) sum_k (O_sk - mean_k) v_k: the product
// Obar v, one wave per sample row (FP64 accumulation)
template <typename TI>
__global__ __launch_bounds__(256) void obar_matvec_kernel(const TI* __restrict__ O, const double* __restrict__ mean,
                                                          const double* __restrict__ r, int64_t ns, int64_t np,
                                                          double inv_ns, double alpha, const double* __restrict__ v,
                                                          double beta, double* __restrict__ y) {
  const int lane = threadIdx.x % WAVE;
  const int64_t s = (int64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE;
  if (s >= ns) return;
  const TI* row = O + s * np;
  double a0 = 0.0, a1 = 0.0;
  int64_t k = lane;
  for (; k + WAVE < np; k += 2 * WAVE) {
    a0 += ((double)row[k] - mean[k]) * v[k];
    a1 += ((double)row[k + WAVE] - mean[k + WAVE]) * v[k + WAVE];
  }
  if (k < np) a0 += ((double)row[k] - mean[k]) * v[k];
  const double acc = wave_sum(a0 + a1);
  if (lane == 0) y[s] = (beta == 0.0 ? 0.0 : beta * y[s]) + alpha * sqrt((r ? r[s] : 1.0) * inv_ns) * acc;
}

// O[:, k] *= d_k (grid-stride over the ns * np elements), mean_k *= d_k
template <typename TI>
__global__ void scale_columns_kernel(TI* __restrict__ O, int64_t ns, int64_t np, const double* __restrict__ d) {
  const int64_t total = ns * np;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
    O[e] = (TI)((double)O[e] * d[e % np]);
}

__global__ void mul_f64_kernel(double* __restrict__ x, const double* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= d[i];
}

// single-block: scalars[0] = trace, then T_ii += rshift * trace + ashift
__global__ __launch_bounds__(1024) void shift_diag_kernel(double* __restrict__ T, int64_t n, double rshift,
                                                          double ashift, double* __restrict__ scalars) {
  __shared__ double red[16];
  int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE, nw = blockDim.x / WAVE;
  double acc = 0.0;
  for (int64_t i = tid; i < n; i += blockDim.x) acc += T[i * n + i];
  acc = wave_sum(acc);
  if (lane == 0) red[wv] = acc;
  __syncthreads();
  double tr = 0.0;
  for (int w = 0; w < nw; ++w) tr += red[w];
  double shift = rshift * tr + ashift;
  for (int64_t i = tid; i < n; i += blockDim.x) T[i * n + i] += shift;
  if (tid == 0) scalars[0] = tr;
}

// single block: scale[0] = 2^k with 2^k * max_i |T_ii| in [0.5, 1) (exact scaling)
__global__ __launch_bounds__(1024) void pow2_scale_kernel(const double* __restrict__ T, int64_t n,
                                                          double* __restrict__ scale) {
  __shared__ double red[16];
  int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE, nw = blockDim.x / WAVE;
  double m = 0.0;
  for (int64_t i = tid; i < n; i += blockDim.x) m = fmax(m, fabs(T[i * n + i]));
  m = wave_max(m);
  if (lane == 0) red[wv] = m;
  __syncthreads();
  if (tid == 0) {
    double mx = 0.0;
    for (int w = 0; w < nw; ++w) mx = fmax(mx, red[w]);
    int e = 0;
    if (mx > 0.0 && isfinite(mx)) frexp(mx, &e);  // mx = f * 2^e, f in [0.5, 1)
    scale[0] = ldexp(1.0, -e);
  }
}

__global__ void scale_by_kernel(double* __restrict__ T, int64_t n2, const double* __restrict__ scale) {
  const double sc = scale[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    T[i] *= sc;
}

// soft inverse of eigenvalues times rho (solver.py:94-101), single block
__global__ __launch_bounds__(1024) void eig_inv_kernel(const double* __restrict__ vals, int64_t n, double rtol,
                                                       double atol, const double* __restrict__ scale,
                                                       double* __restrict__ rho) {
  const double inv_scale = 1.0 / scale[0];  // eigenvalues of the scaled matrix -> eigenvalues of T
  __shared__ double red[16];
  int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE, nw = blockDim.x / WAVE;
  double m = 0.0;
  for (int64_t i = tid; i < n; i += blockDim.x) m = fmax(m, fabs(vals[i] * inv_scale));
  m = wave_max(m);
  if (lane == 0) red[wv] = m;
  __syncthreads();
  double vmax = 0.0;
  for (int w = 0; w < nw; ++w) vmax = fmax(vmax, red[w]);
  double cut = rtol * vmax + atol;
  for (int64_t i = tid; i < n; i += blockDim.x) {
    double v = vals[i] * inv_scale, va = fabs(v);
    double inv = 0.0;
    if (va > 0.0) {
      double q = cut / va;
      double q2 = q * q;
      inv = 1.0 / (v * (1.0 + q2 * q2 * q2));
    }
    rho[i] *= inv;
  }
}

__global__ void copy_f64_kernel(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

template <typename T>
__global__ __launch_bounds__(256) void finite_check_kernel(const double* __restrict__ step, int64_t n,
                                                           int32_t* __restrict__ ok) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = (i < n) && !isfinite(step[i]);
  if (__any(bad) && (threadIdx.x % WAVE) == 0) atomicAnd(ok, 0);
}

template <typename T>
__global__ __launch_bounds__(256) void apply_update_kernel(T* __restrict__ theta, const double* __restrict__ step,
                                                           double lr, int64_t n, const int32_t* __restrict__ ok) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || *ok == 0) return;
  T d = (T)(lr * step[i]);
  theta[i] = theta[i] - d;
}

// ---------------------------------------------------------------- soft inverse by partial fractions
// The soft inverse of solver.py:94-101, g(v) = 1 / (v (1 + (c/|v|)^6)), equals
// v^5 / (v^6 + c^6).  Its six simple poles p_k = c exp(i pi (2k+1)/6) all have
// residue 1/6, so pairing conjugates
//   g(v) = (1/3) sum_{k=0..2} Re 1 / (v - p_k),
// and g(T) b = (1/3) sum_k Re (T - p_k I)^{-1} b needs no eigendecomposition.
// Each complex system is solved by iterative refinement whose inner solve is
// (T - conj(p) I) ((T - a)^2 + b^2)^{-1}, p = a + ib, with the real SPD
// quadratic factored once by Cholesky; the residual uses T itself, so the
// result has the accuracy of the well-conditioned complex system.

// 16 bytes of a matrix streamed once per pass: a non-temporal load, which
// leaves the caches to the vector operand
typedef double v2d_nt __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 stream_load2(const double2* p) {
  const v2d_nt v = __builtin_nontemporal_load(reinterpret_cast<const v2d_nt*>(p));
  return make_double2(v.x, v.y);
}

// One wave's partial sums of row[j] * x[j] over j in [lo, hi): 16-byte loads
// (1 KB per wave instruction; the row's non-temporal) where row and x share
// their 16-byte alignment phase, after peeling one element to reach it;
// 8-byte loads otherwise.  At n = 16384 a triangle pass (1.07 GB) takes
// ~180 us, a full-matrix pass ~330 us (6.5 TB/s).
__device__ __forceinline__ double wave_row_dot(const double* __restrict__ row, const double* __restrict__ x,
                                               int64_t lo, int64_t hi, int lane) {
  double acc = 0.0;
  if (((reinterpret_cast<uintptr_t>(row) ^ reinterpret_cast<uintptr_t>(x)) & 15) == 0) {
    if ((reinterpret_cast<uintptr_t>(row + lo) & 15) && lo < hi) {
      if (lane == 0) acc = row[lo] * x[lo];
      ++lo;
    }
    const double2* r2 = reinterpret_cast<const double2*>(row + lo);
    const double2* x2 = reinterpret_cast<const double2*>(x + lo);
    const int64_t m = (hi - lo) / 2;
    int64_t k = lane;
    for (; k + 3 * WAVE < m; k += 4 * WAVE) {
      const double2 a0 = stream_load2(r2 + k), a1 = stream_load2(r2 + k + WAVE),
                    a2 = stream_load2(r2 + k + 2 * WAVE), a3 = stream_load2(r2 + k + 3 * WAVE);
      const double2 b0 = x2[k], b1 = x2[k + WAVE], b2 = x2[k + 2 * WAVE], b3 = x2[k + 3 * WAVE];
      acc += a0.x * b0.x + a0.y * b0.y + a1.x * b1.x + a1.y * b1.y + a2.x * b2.x + a2.y * b2.y + a3.x * b3.x +
             a3.y * b3.y;
    }
    for (; k < m; k += WAVE) {
      const double2 a0 = stream_load2(r2 + k), b0 = x2[k];
      acc += a0.x * b0.x + a0.y * b0.y;
    }
    if (((hi - lo) & 1) && lane == 0) acc += row[hi - 1] * x[hi - 1];
    return acc;
  }
  int64_t j = lo + lane;
  for (; j + 3 * WAVE < hi; j += 4 * WAVE)
    acc += row[j] * x[j] + row[j + WAVE] * x[j + WAVE] + row[j + 2 * WAVE] * x[j + 2 * WAVE] +
           row[j + 3 * WAVE] * x[j + 3 * WAVE];
  for (; j < hi; j += WAVE) acc += row[j] * x[j];
  return acc;
}

// Y = T X for NV vectors (n x NV, column-major) with T symmetric n x n stored
// in full: one wave per row, coalesced reads of the row, X from cache.
// An HBM stream of T at close to copy bandwidth (library GEMV / GEMM with
// one or two columns reach a fraction of it).
template <int NV>
__global__ __launch_bounds__(256) void matvec_rows_kernel(const double* __restrict__ T, int64_t n,
                                                          const double* __restrict__ X, double* __restrict__ Y) {
  const int lane = threadIdx.x % WAVE;
  const int64_t i = (int64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE;
  if (i >= n) return;
  const double* row = T + i * n;
  if constexpr (NV == 1) {
    const double s = wave_sum(wave_row_dot(row, X, 0, n, lane));
    if (lane == 0) Y[i] = s;
    return;
  }
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  int64_t j = lane;
  for (; j + 3 * WAVE < n; j += 4 * WAVE) {
    const double t0 = row[j], t1 = row[j + WAVE], t2 = row[j + 2 * WAVE], t3 = row[j + 3 * WAVE];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double* x = X + v * n + j;
      acc[v] += t0 * x[0] + t1 * x[WAVE] + t2 * x[2 * WAVE] + t3 * x[3 * WAVE];
    }
  }
  for (; j < n; j += WAVE) {
    const double tj = row[j];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += tj * X[v * n + j];
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double s = wave_sum(acc[v]);
    if (lane == 0) Y[v * n + i] = s;
  }
}

// sum over the block (every thread gets the total; fixed summation order)
__device__ __forceinline__ double block_sum(double x, double* red) {
  int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE, nw = blockDim.x / WAVE;
  x = wave_sum(x);
  __syncthreads();  // red may still be read by a previous call
  if (lane == 0) red[wv] = x;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += red[w];
  return t;
}

// single block: V[0..n) = hashed start vector / norm, V[n..3n) = 0, beta[0] = 0
__global__ __launch_bounds__(1024) void lanczos_init_kernel(int64_t n, double* __restrict__ V,
                                                            double* __restrict__ beta) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;  // splitmix64
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    double v = (double)(z >> 11) * 0x1.0p-52 - 1.0;  // [-1, 1)
    V[i] = v;
    V[n + i] = 0.0;
    V[2 * n + i] = 0.0;
    acc += v * v;
  }
  double inv = 1.0 / sqrt(block_sum(acc, red));
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) V[i] *= inv;
  if (threadIdx.x == 0) beta[0] = 0.0;
}

// single block, after w = T v_j: alpha_j = w.v_j, w -= alpha_j v_j + beta_j v_{j-1},
// one re-orthogonalisation against v_j, beta_{j+1} = |w|, v_{j+1} = w / beta_{j+1}
__global__ __launch_bounds__(1024) void lanczos_step_kernel(int64_t n, int j, double* __restrict__ w,
                                                            const double* __restrict__ v,
                                                            const double* __restrict__ vp, double* __restrict__ vn,
                                                            double* __restrict__ alpha, double* __restrict__ beta) {
  __shared__ double red[16];
  const double bj = beta[j];
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) a += w[i] * v[i];
  a = block_sum(a, red);
  double a2 = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    double t = w[i] - a * v[i] - bj * vp[i];
    w[i] = t;
    a2 += t * v[i];
  }
  a2 = block_sum(a2, red);
  double nn = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    double t = w[i] - a2 * v[i];
    w[i] = t;
    nn += t * t;
  }
  double b = sqrt(block_sum(nn, red));
  double inv = b > 0.0 ? 1.0 / b : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) vn[i] = w[i] * inv;
  if (threadIdx.x == 0) {
    alpha[j] = a + a2;
    beta[j + 1] = b;
  }
}

// A = T2 - 2a T + c^2 I on the lower triangle (column-major, i >= j), 0 above
__global__ __launch_bounds__(256) void form_quad_kernel(const double* T2,
                                                        const double* __restrict__ T, int64_t n, double two_a,
                                                        double c2, double* A) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t j = e / n, i = e - j * n;
    A[e] = (i < j) ? 0.0 : T2[e] - two_a * T[e] + (i == j ? c2 : 0.0);
  }
}

// Shifted Cholesky route for small cutoffs (c < 1e-5 lambda_max, the
// reference default rtol = 1e-12): G = T + (mu / n) 1 1^T - sigma I, lower
// triangle (the upper part zero), factored once; x = T^+ b' is refined against
// T + (mu / n) 1 1^T, which deflates the null direction 1 of the centred MinSR
// Gram (mu = 0 for the SR branch), by conjugate gradients preconditioned
// with that factor.
__global__ __launch_bounds__(256) void form_shifted_kernel(const double* __restrict__ T, int64_t n, double mu_n,
                                                           double sigma, double* __restrict__ G) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / n, i = e - j * n;
    G[e] = (i < j) ? 0.0 : T[e] + mu_n - (i == j ? sigma : 0.0);
  }
}

__device__ double block_sum_1024(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += red[w];
  return s;
}

// single block: b' = b - mean(b) 1 (deflate) or b; x = 0; r = z = b'
__global__ __launch_bounds__(1024) void shifted_init_kernel(const double* __restrict__ b, int64_t n, int deflate,
                                                            double* __restrict__ bp, double* __restrict__ x,
                                                            double* __restrict__ r, double* __restrict__ z) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += b[i];
  const double mean = deflate ? block_sum_1024(s, red) / (double)n : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = b[i] - mean;
    bp[i] = v;
    r[i] = v;
    z[i] = v;
    x[i] = 0.0;
  }
}

// Conjugate gradients on T_d x = b' preconditioned by the factor of
// G = T_d - sigma I (the preconditioned operator is I + sigma G^-1: eigenvalues
// 1 + sigma / (v - sigma), clustered at 1 except for the few v near sigma).
// s[0] holds r.z.  Single-block steps around the library's triangular solves.
// step 1: q = T p (given) + (mu / n) sum(p); alpha = rz / p.q; x += alpha p;
// r -= alpha q; z = r (then z <- G^-1 z by two triangular solves)
__global__ __launch_bounds__(1024) void pcg_step1_kernel(double* __restrict__ q, const double* __restrict__ p,
                                                         int64_t n, double mu_n, double* __restrict__ x,
                                                         double* __restrict__ r, double* __restrict__ z,
                                                         const double* __restrict__ s) {
  __shared__ double red[16];
  double sp = 0.0;
  if (mu_n != 0.0)
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) sp += p[i];
  const double add = (mu_n != 0.0) ? mu_n * block_sum_1024(sp, red) : 0.0;
  double pq = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double qi = q[i] + add;
    q[i] = qi;
    pq += p[i] * qi;
  }
  pq = block_sum_1024(pq, red);
  const double alpha = (pq > 0.0 && std::isfinite(pq)) ? s[0] / pq : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    z[i] = ri;
  }
}

// step 2 (first = 1: p = z): rz' = r.z; beta = rz' / rz; p = z + beta p; s[0] = rz'
__global__ __launch_bounds__(1024) void pcg_step2_kernel(const double* __restrict__ r, const double* __restrict__ z,
                                                         int64_t n, int first, double* __restrict__ p,
                                                         double* __restrict__ s) {
  __shared__ double red[16];
  double rz = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) rz += r[i] * z[i];
  rz = block_sum_1024(rz, red);
  const double beta = (!first && s[0] > 0.0 && std::isfinite(rz)) ? rz / s[0] : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = z[i] + beta * p[i];
  __syncthreads();
  if (threadIdx.x == 0) s[0] = rz;
}

// the true residual of the result: norms <- |b' - T_d x|^2, |x|^2 (Tx given)
__global__ __launch_bounds__(1024) void pcg_check_kernel(const double* __restrict__ bp, const double* __restrict__ Tx,
                                                         const double* __restrict__ x, int64_t n, double mu_n,
                                                         double* __restrict__ norms) {
  __shared__ double red[16];
  double sx = 0.0;
  if (mu_n != 0.0)
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) sx += x[i];
  const double add = (mu_n != 0.0) ? mu_n * block_sum_1024(sx, red) : 0.0;
  double rr = 0.0, xx = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double ri = bp[i] - Tx[i] - add;
    rr += ri * ri;
    xx += x[i] * x[i];
  }
  rr = block_sum_1024(rr, red);
  xx = block_sum_1024(xx, red);
  if (threadIdx.x == 0) {
    norms[0] = rr;
    norms[1] = xx;
  }
}

// R = [b, 0], Z = 0   (R, Z: n x 2 column-major)
__global__ void refine_init_kernel(const double* __restrict__ rhs, int64_t n, double* __restrict__ R,
                                   double* __restrict__ Z) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R[i] = rhs[i];
  R[n + i] = 0.0;
  Z[i] = 0.0;
  Z[n + i] = 0.0;
}

// Z += (T - conj(p)) W with W = A^{-1} R, TW = T W
__global__ void refine_update_kernel(const double* __restrict__ W, const double* __restrict__ TW, int64_t n,
                                     double a, double b, double* __restrict__ Z) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double wr = W[i], wi = W[n + i];
  Z[i] += TW[i] - a * wr - b * wi;
  Z[n + i] += TW[n + i] - a * wi + b * wr;
}

// R = b - (T - p) Z with TZ = T Z
__global__ void refine_residual_kernel(const double* __restrict__ rhs, const double* __restrict__ Z,
                                       const double* __restrict__ TZ, int64_t n, double a, double b,
                                       double* __restrict__ R) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double zr = Z[i], zi = Z[n + i];
  R[i] = rhs[i] - (TZ[i] - a * zr + b * zi);
  R[n + i] = -(TZ[n + i] - a * zi - b * zr);
}

// x (+)= Re Z / 3
__global__ void refine_accum_kernel(const double* __restrict__ Z, int64_t n, int first, double* __restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = Z[i] * (1.0 / 3.0);
  x[i] = first ? v : x[i] + v;
}

// ---------------------------------------------------------------- soft inverse by a Chebyshev expansion
// With G = T + sigma I = L L^T and P = L^-1 (rocSOLVER's lower triangle, i.e.
// row r of the row-major n x n array holds P_{c r} at columns c >= r), M =
// T G^-1 = I - sigma P^T P is symmetric with the eigenvectors of T and
// eigenvalues mu = v / (v + sigma) in [0, mu_max]; g(T) = h(M) with h(mu) =
// g(sigma mu / (1 - mu)), a function analytic on [0, mu_max] whose nearest
// poles (from those of g at v = c e^{i pi (2k+1)/6}) stay ~0.13 away for
// sigma = c, so its Chebyshev series converges by ~1.3 per term.  Each term
// costs one P product and one P^T product, both half-row streams of the
// array once the strictly lower part (row-major) holds the transpose of the
// upper one.

// S[r n + c] = S[c n + r] for c < r: 64 x 64 tiles through LDS (rows padded
// to 65 doubles against bank conflicts); one workgroup per lower tile
__global__ __launch_bounds__(256) void tri_mirror_kernel(double* __restrict__ S, int64_t n, int nt) {
  __shared__ double buf[64][65];
  // lower tile (tr, tc), tr >= tc, from the linear index
  int64_t q = blockIdx.x;
  int tr = (int)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while ((int64_t)(tr + 1) * (tr + 2) / 2 <= q) ++tr;
  while ((int64_t)tr * (tr + 1) / 2 > q) --tr;
  const int tc = (int)(q - (int64_t)tr * (tr + 1) / 2);
  if (tr >= nt) return;
  const int64_t r0 = (int64_t)tr * 64, c0 = (int64_t)tc * 64;
  const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
  // source tile: rows c0.., columns r0.. (the upper part)
  for (int a = ty; a < 64; a += 4) {
    const int64_t sr = c0 + a, sc = r0 + tx;
    buf[a][tx] = (sr < n && sc < n) ? S[sr * n + sc] : 0.0;
  }
  __syncthreads();
  for (int b = ty; b < 64; b += 4) {
    const int64_t r = r0 + b, c = c0 + tx;
    if (r < n && c < n && c < r) S[r * n + c] = buf[tx][b];
  }
}

// Y = P X (LOWER: row i, columns c <= i) or P^T X (upper: columns c >= i):
// one wave per row, coalesced reads along the row
template <bool LOWER>
__global__ __launch_bounds__(256) void tri_rows_kernel(const double* __restrict__ S, int64_t n,
                                                       const double* __restrict__ X, double* __restrict__ Y) {
  const int lane = threadIdx.x % WAVE;
  const int64_t i = (int64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE;
  if (i >= n) return;
  const double* row = S + i * n;
  const double acc = wave_sum(wave_row_dot(row, X, LOWER ? 0 : i, LOWER ? i + 1 : n, lane));
  if (lane == 0) Y[i] = acc;
}

// One Chebyshev term: w = P^T u (upper half rows) fused with the recurrence
// t_{k+1} = 2 (alpha t_k - beta w) - t_{k-1} (first: t_1 = alpha t_0 - beta w,
// x = a0 t_0 + a t_1), x += a t_{k+1}; A = alpha I - beta P^T P maps the
// spectrum of M onto [-1, 1].  t_{k+1} replaces t_{k-1} in tprev.
__global__ __launch_bounds__(256) void cheb_rows_kernel(const double* __restrict__ S, int64_t n,
                                                        const double* __restrict__ u, const double* __restrict__ tcur,
                                                        double* __restrict__ tprev, double alpha, double beta,
                                                        double a0, double a, int first, double* __restrict__ x) {
  const int lane = threadIdx.x % WAVE;
  const int64_t i = (int64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE;
  if (i >= n) return;
  const double w = wave_sum(wave_row_dot(S + i * n, u, i, n, lane));
  if (lane == 0) {
    const double tk = tcur[i];
    const double tn = first ? alpha * tk - beta * w : 2.0 * (alpha * tk - beta * w) - tprev[i];
    tprev[i] = tn;
    x[i] = first ? a0 * tk + a * tn : x[i] + a * tn;
  }
}

// single block: out[0] = |a|^2, out[1] = |b|^2
__global__ __launch_bounds__(1024) void sq_norms_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                        int64_t n, double* __restrict__ out) {
  __shared__ double red[16];
  double sa = 0.0, sb = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    sa += a[i] * a[i];
    sb += b[i] * b[i];
  }
  sa = block_sum_1024(sa, red);
  sb = block_sum_1024(sb, red);
  if (threadIdx.x == 0) {
    out[0] = sa;
    out[1] = sb;
  }
}

// ---------------------------------------------------------------- SNR truncation
// _sum_without_noise (optimizer/solver.py:114-125): rho_s = sum_t c_ts is
// divided by 1 + (tol / snr_s)^6, snr_s = |mean_s| / sqrt(var_s / m), with the
// variance about the mean as the reference forms it (two passes: rho is known
// before the sums of squares are taken).
// var2[s] = sum_t (U[t, s] y[t] - rho[s] / m)^2, U column-major n x n, one
// block per column (MinSR and minsr_pinv_eig: c_ts = U_ts b_t)
__global__ __launch_bounds__(256) void snr_colsq_kernel(const double* __restrict__ U, int64_t n,
                                                        const double* __restrict__ y, const double* __restrict__ rho,
                                                        double m, double* __restrict__ var2) {
  __shared__ double red[4];
  const int64_t s = blockIdx.x;
  const double* col = U + s * n;
  const double mu = rho[s] / m;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = col[i] * y[i] - mu;
    acc += v * v;
  }
  acc = wave_sum(acc);
  if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) var2[s] = red[0] + red[1] + red[2] + red[3];
}

// var2[k] (+)= sum_s (C[k + s np] e[s] - rho[k] / m)^2 over a block of rows s
// (C = Obar V of the block, column-major np x rows: SR, c_sk = (Obar V)_sk Ebar_s)
__global__ __launch_bounds__(256) void snr_rowsq_kernel(const double* __restrict__ Cb, int64_t np, int64_t rows,
                                                        const double* __restrict__ e, const double* __restrict__ rho,
                                                        double m, int first, double* __restrict__ var2) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const double mu = rho[k] / m;
  double acc = 0.0;
  for (int64_t s = 0; s < rows; ++s) {
    const double v = Cb[k + s * np] * e[s] - mu;
    acc += v * v;
  }
  var2[k] = (first ? 0.0 : var2[k]) + acc;
}

__global__ void snr_scale_kernel(double* __restrict__ rho, const double* __restrict__ var2, int64_t n, double m,
                                 double tol) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // rho = 0 stays 0 (the reference's 0 / (1 + (tol / (0/0))^6) is NaN there)
  if (rho[i] == 0.0) return;
  const double mean = rho[i] / m;
  const double snr = fabs(mean) / sqrt(var2[i] / m / m);
  const double q = tol / snr, q2 = q * q;
  rho[i] = rho[i] / (1.0 + q2 * q2 * q2);
}

// ---------------------------------------------------------------- rocBLAS handle
// One handle per (thread, device, stream).  A handle carries the device
// workspace rocSOLVER uses for its temporaries, so one handle switched between
// streams would let two factorisations in flight on different streams share
// that scratch; a handle bound to a single stream cannot.  Two caches: handles
// whose scratch rocBLAS allocates and keeps (the default), and handles that
// only ever use scratch the caller brings in its workspace (QTX_SOLVE_BLAS_WS:
// rocblas_set_workspace before every use, so rocBLAS allocates nothing).  Each
// keeps at most kMaxHandles per thread and device; the oldest is retired after
// a device-wide sync.
constexpr size_t kMaxHandles = 64;
struct BlasEntry {
  hipStream_t stream;
  rocblas_handle handle;
};
static thread_local std::vector<BlasEntry> g_blas[64];
static thread_local std::vector<BlasEntry> g_blas_user[64];

static int get_blas(hipStream_t st, rocblas_handle* h, void* user_ws = nullptr, size_t user_bytes = 0,
                    bool user = false) {
  int dev = 0;
  QTX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(QTX_EHIP, "device index out of range");
  std::vector<BlasEntry>& cache = user ? g_blas_user[dev] : g_blas[dev];
  rocblas_handle found = nullptr;
  for (const BlasEntry& e : cache)
    if (e.stream == st) found = e.handle;
  if (!found) {
    if (cache.size() >= kMaxHandles) {
      QTX_HIP(hipDeviceSynchronize());
      rocblas_destroy_handle(cache.front().handle);
      cache.erase(cache.begin());
    }
    if (rocblas_create_handle(&found) != rocblas_status_success)
      return fail(QTX_ESOLVER, "rocblas_create_handle failed");
    if (rocblas_set_stream(found, st) != rocblas_status_success) {
      rocblas_destroy_handle(found);
      return fail(QTX_ESOLVER, "rocblas_set_stream failed");
    }
    cache.push_back(BlasEntry{st, found});
  }
  if (user && rocblas_set_workspace(found, user_ws, user_bytes) != rocblas_status_success)
    return fail(QTX_ESOLVER, "rocblas_set_workspace failed");
  *h = found;
  return QTX_OK;
}

static int rb(rocblas_status s, const char* what) {
  if (s == rocblas_status_success) return QTX_OK;
  if (s == rocblas_status_memory_error)
    return fail(QTX_EWORKSPACE, std::string(what) + ": rocBLAS/rocSOLVER scratch does not fit (with "
                                "QTX_SOLVE_BLAS_WS, size the workspace tail by qtx_minsr_blas_ws_bytes)");
  return fail(QTX_ESOLVER, std::string(what) + " failed: " + rocblas_status_to_string(s));
}

}  // namespace qtx

using namespace qtx;

static int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }
// grid cap of the grid-stride element-wise kernels (256-thread blocks): the
// work-item count of one launch must stay below 2^32
constexpr int64_t kGridCap = int64_t(1) << 20;

constexpr int kLanczosMax = 512;  // Lanczos steps for lambda_max before giving up on convergence
#ifndef QTX_CHOL_BLOCK
#define QTX_CHOL_BLOCK 1024
#endif
constexpr int64_t kCholBlock = QTX_CHOL_BLOCK;  // diagonal block of the blocked Cholesky
constexpr int kInfoPerFactor = (int)((131072 + kCholBlock - 1) / kCholBlock);  // status slots: n <= 131072

// Rows of FP64 Obar formed per DGEMM chunk in the SR branch (<= 256 MiB)
static int64_t obar_rows(int64_t ns, int64_t np) {
  return std::max<int64_t>(1, std::min<int64_t>(ns, (int64_t(1) << 28) / (8 * std::max<int64_t>(np, 1))));
}

// solver mode without its flags, and the branch: T = Obar Obar^T (MinSR) or
// S = Obar^T Obar (SR), by the flags or, without them, by the shape
static inline int base_mode(int mode) { return mode & 0xff; }
static inline bool use_minnorm(int64_t ns, int64_t np, int mode) {
  if (mode & QTX_SOLVE_MINNORM) return true;
  if (mode & QTX_SOLVE_LSTSQ) return false;
  return ns < np;
}
constexpr double kSnrMin = 1e-6;  // tol_snr at or below it: no truncation (solver.py:124)
// Chebyshev route (cutoffs c >= 1e-5 lambda_max): taken from n >= kChebMinN
// on, where its one factorisation and triangular inverse plus ~2 n^2 x 8 B
// of streaming per term cost less than the partial fractions' T^2 and three
// factorisations (measured: 6.4 against 7.1 ms at n = 2048, 15 against 18 ms
// at 4096, 2.4 against 2.0 ms at 512), or whenever QTX_SOLVE_CHEB asks for
// it.  Terms up to kChebMax; the series is cut where 8 consecutive
// coefficients fall below kChebTol of the largest (about 158 terms).
constexpr int64_t kChebMinN = 2048;
constexpr int kChebMax = 400;
constexpr int kChebNodes = 1024;
constexpr double kChebTol = 1e-15;
// The series covers mu down to -kChebMargin (eigenvalues of T down to -c / 6,
// which rounding in the split-bf16 Gram can produce), and the last term's norm
// is checked against the right-hand side's: a Chebyshev polynomial is bounded
// by 1 on the interval, so growth means T has eigenvalues below it, and the
// solve takes the partial fractions instead.
constexpr double kChebMargin = 0.2;
constexpr double kChebGrowth = 1.5;
// shifted Cholesky route: G = T_d - kShiftCut c I must factor, which puts every
// eigenvalue of T_d above 150 c, where the soft inverse differs from 1/v by
// (c/v)^6 <= 8.8e-14; kShiftIters preconditioned CG steps, accepted when the
// true residual's normwise backward error |b' - T_d x| / (lambda_max |x|) is
// below kShiftAccept (the accuracy a backward-stable eigensolver attains)
constexpr double kShiftCut = 150.0;
constexpr int kShiftIters = 40;
constexpr double kShiftAccept = 1e-15;

struct SolveWs {
  double* T;       // n x n
  double* y;       // max(ns, np)
  double* z;       // max(ns, np)
  double* D;       // n
  double* E;       // n
  double* scalars; // 8
  int32_t* info;   // 4 (+pad)
  double* dtile;   // 3 x 64 x 64: factored diagonal tiles of the blocked Cholesky, staged
  // partial-fraction soft inverse (EIGH_PINV modes only)
  double* T2;      // 3 x (n x n): T^2 (lower), then the three (T - a_k)^2 + b_k^2
                   // (the first in place of T^2) and their Cholesky factors
  double* R;       // n x 6: residuals, then A_k^{-1} residuals (pole k: columns 2k, 2k+1)
  double* TW;      // n x 6: T times those vectors
  double* Z;       // n x 2: complex solution
  double* x;       // n: accumulated soft-inverse solution
  double* V;       // 3n: Lanczos vectors
  double* w;       // n: Lanczos work vector
  double* ab;      // alpha[kLanczosMax], beta[kLanczosMax + 1]
  void* split;     // bf16 hi / lo planes of Obar (split-bf16 Gram only)
  double* obar;    // ns x np FP64 Obar (SR branch, ns >= np: S = Obar^T Obar by DGEMM)
  double* var2;    // n: SNR truncation, sums of squared deviations (QTX_SOLVE_SNR)
  double* cbuf;    // obar_rows x np: Obar V of a row block (SR branch with QTX_SOLVE_SNR)
};

static size_t ws_layout(int64_t ns, int64_t np, int mode, int gram, SolveWs* w, char* base) {
  const bool mn = use_minnorm(ns, np, mode);
  int64_t n = mn ? ns : np;
  int64_t m = ns > np ? ns : np;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return base ? base + o : nullptr;
  };
  SolveWs s{};
  s.T = (double*)take((size_t)n * n * 8);
  s.y = (double*)take((size_t)m * 8);
  s.z = (double*)take((size_t)m * 8);
  s.D = (double*)take((size_t)n * 8);
  s.E = (double*)take((size_t)n * 8);
  s.scalars = (double*)take(64);
  s.info = (int32_t*)take(3 * kInfoPerFactor * sizeof(int32_t));
  s.dtile = (double*)take((size_t)3 * 64 * 64 * 8);
  if (base_mode(mode) == QTX_SOLVE_EIGH_PINV) {
    s.T2 = (double*)take((size_t)3 * n * n * 8);
    s.R = (double*)take((size_t)n * 48);
    s.TW = (double*)take((size_t)n * 48);
    s.Z = (double*)take((size_t)n * 48);
    s.x = (double*)take((size_t)n * 8);
    s.V = (double*)take((size_t)n * 24);
    s.w = (double*)take((size_t)n * 8);
    s.ab = (double*)take((size_t)(2 * kLanczosMax + 1) * 8);
  }
  if (gram == QTX_GRAM_BF16X3 && mn) s.split = take(g3_bytes(ns, np));
  if (!mn) s.obar = (double*)take((size_t)obar_rows(ns, np) * np * 8);
  if (mode & QTX_SOLVE_SNR) {
    s.var2 = (double*)take((size_t)n * 8);
    if (!mn) s.cbuf = (double*)take((size_t)obar_rows(ns, np) * np * 8);
  }
  if (w) *w = s;
  return off;
}

// what the last qtx_minsr_solve on this thread did
static thread_local int g_last_path = -1;
static thread_local int g_last_lanczos = 0;
static thread_local double g_last_lmax = 0.0;
static thread_local int g_last_fallback = 0;  // why the partial-fraction route was left (0: it was not)

// largest eigenvalue of the m x m symmetric tridiagonal (diag a[0..m), off-diagonal
// e[1..m) between rows i-1 and i) by Sturm-count bisection
static double tridiag_max_eig(const double* a, const double* e, int m) {
  double lo = a[0], hi = a[0];
  for (int i = 0; i < m; ++i) {
    double r = (i > 0 ? fabs(e[i]) : 0.0) + (i + 1 < m ? fabs(e[i + 1]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
  }
  auto count_below = [&](double x) {
    int c = 0;
    double q = a[0] - x;
    for (int i = 0;;) {
      if (q == 0.0) q = -1e-300;
      if (q < 0.0) ++c;
      if (++i >= m) break;
      q = a[i] - x - e[i] * e[i] / q;
    }
    return c;
  };
  for (int it = 0; it < 256 && hi - lo > 2e-16 * fmax(fabs(lo), fabs(hi)); ++it) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count_below(mid) == m) hi = mid;
    else lo = mid;
  }
  return hi;
}

template <typename TI>
static int launch_gram_nt(hipStream_t st, int gram, const TI* O, const double* mean, const double* r, int64_t n,
                          int64_t np, double* T, void* split_ws) {
  int64_t nt = (n + GT - 1) / GT;
  int64_t ntiles = nt * (nt + 1) / 2;
  QTX_REQUIRE(ntiles < 0x7fffffff, "Gram too large");
  dim3 g((unsigned)ntiles), bl(256);
  double inv_ns = 1.0 / (double)n;
  if (gram == QTX_GRAM_BF16X3) {
    QTX_REQUIRE(split_ws != nullptr, "split-bf16 Gram needs its workspace");
    const int64_t rows = g3_rows(n), kpad = g3_kpad(np), bpr = (kpad + 255) / 256;
    const int64_t nblocks = rows * bpr;
    __bf16* Hi = (__bf16*)split_ws;
    __bf16* Lo = Hi + rows * kpad;
    hipLaunchKernelGGL(split_bf16x3_kernel<TI>, dim3((unsigned)std::min<int64_t>(nblocks, kGridCap)), dim3(256), 0,
                       st, O, mean, r, n, np, kpad, bpr, nblocks, inv_ns, Hi, Lo);
    hipLaunchKernelGGL(gram_bf16x3_kernel, g, bl, 0, st, (const __bf16*)Hi, (const __bf16*)Lo, n, kpad, (int)ntiles,
                       T);
  } else if (gram == QTX_GRAM_FP32) {
    hipLaunchKernelGGL((gram_nt_kernel<TI, QTX_GRAM_FP32>), g, bl, 0, st, O, mean, r, n, np, inv_ns, T);
  } else {
    hipLaunchKernelGGL((gram_nt_kernel<TI, QTX_GRAM_FP64>), g, bl, 0, st, O, mean, r, n, np, inv_ns, T);
  }
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

template <typename TI>
static int launch_obar_t_y(hipStream_t st, const TI* O, const double* mean, const double* r, const double* y,
                           int64_t ns, int64_t np, double* x) {
  hipLaunchKernelGGL(fill_f64_kernel, dim3(nblk(np, 256)), dim3(256), 0, st, x, np, 0.0);
  int64_t chunks = (ns + 255) / 256;
  if (chunks > 256) chunks = 256;
  int64_t rpc = (ns + chunks - 1) / chunks;
  chunks = (ns + rpc - 1) / rpc;
  dim3 g(nblk(np, 256), (unsigned)chunks);
  hipLaunchKernelGGL(obar_t_y_kernel<TI>, g, dim3(256), 0, st, O, mean, r, y, ns, np, rpc, 1.0 / (double)ns, x);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

// What the SNR truncation of the SR branch needs besides the eigenvectors:
// the per-sample contributions are (Obar V)_sk Ebar_s, Obar formed again a
// row block at a time
struct SnrRows {
  int o_dtype;
  const void* O;
  const double* mean;
  const double* r;
  const double* ebar;
  int64_t ns, np;
};

template <typename TI>
static int snr_rows_var2(hipStream_t st, rocblas_handle h, const SnrRows& L, SolveWs& w) {
  const int64_t rows = obar_rows(L.ns, L.np);
  const double one = 1.0, zero = 0.0;
  const TI* O = (const TI*)L.O;
  for (int64_t s0 = 0; s0 < L.ns; s0 += rows) {
    const int64_t nr = std::min(rows, L.ns - s0), tot = nr * L.np;
    hipLaunchKernelGGL(form_obar_kernel<TI>, dim3((unsigned)std::min<int64_t>((tot + 255) / 256, kGridCap)), dim3(256),
                       0, st, O + s0 * L.np, L.mean, L.r ? L.r + s0 : nullptr, nr, L.np, 1.0 / (double)L.ns, w.obar);
    QTX_LAUNCH_CHECK();
    // (Obar V)^T of the block, column-major np x nr: V^T B with B the block
    // as a column-major np x nr matrix
    int rc = rb(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)L.np,
                              (rocblas_int)nr, (rocblas_int)L.np, &one, w.T, (rocblas_int)L.np, w.obar,
                              (rocblas_int)L.np, &zero, w.cbuf, (rocblas_int)L.np),
                "rocblas_dgemm");
    if (rc) return rc;
    hipLaunchKernelGGL(snr_rowsq_kernel, dim3(nblk(L.np, 256)), dim3(256), 0, st, w.cbuf, L.np, nr, L.ebar + s0, w.z,
                       (double)L.ns, s0 == 0 ? 1 : 0, w.var2);
    QTX_LAUNCH_CHECK();
  }
  return QTX_OK;
}

// y <- U g(D) U^T y through rocSOLVER dsyevd (T is overwritten); with
// tol_snr > 1e-6 the components of U^T y are truncated by their SNR (rows ==
// nullptr: c_ts = U_ts y_t over the n rows of T; else the SR branch's
// (Obar V)_sk Ebar_s over the samples); returns 1 or an error
static int pinv_syevd(hipStream_t st, rocblas_handle h, int64_t n, double rtol, double atol, SolveWs& w,
                      double tol_snr = 0.0, const SnrRows* rows = nullptr) {
  const rocblas_int ni = (rocblas_int)n;
  // exact power-of-two rescaling of T (eigenvectors unchanged, eigenvalues
  // scaled by 2^k and undone in eig_inv_kernel)
  hipLaunchKernelGGL(pow2_scale_kernel, dim3(1), dim3(1024), 0, st, w.T, n, w.scalars + 1);
  hipLaunchKernelGGL(scale_by_kernel, dim3((unsigned)std::min<int64_t>((n * n + 255) / 256, kGridCap)), dim3(256), 0,
                     st, w.T, n * n,
                     w.scalars + 1);
  QTX_LAUNCH_CHECK();
  int rc = rb(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, ni, w.T, ni, w.D, w.E, w.info),
              "rocsolver_dsyevd");
  if (rc) return rc;
  const double one = 1.0, zero = 0.0;
  // z = U^T y ; z *= eig_inv ; y = U z   (U column-major in T)
  rc = rb(rocblas_dgemv(h, rocblas_operation_transpose, ni, ni, &one, w.T, ni, w.y, 1, &zero, w.z, 1),
          "rocblas_dgemv");
  if (rc) return rc;
  if (tol_snr > kSnrMin) {
    QTX_REQUIRE(w.var2 != nullptr, "tol_snr > 0 needs the QTX_SOLVE_SNR workspace");
    double m = (double)n;
    if (rows == nullptr) {
      hipLaunchKernelGGL(snr_colsq_kernel, dim3((unsigned)n), dim3(256), 0, st, w.T, n, w.y, w.z, m, w.var2);
      QTX_LAUNCH_CHECK();
    } else {
      QTX_REQUIRE(w.cbuf != nullptr, "tol_snr > 0 needs the QTX_SOLVE_SNR workspace");
      rc = rows->o_dtype == QTX_F32 ? snr_rows_var2<float>(st, h, *rows, w) : snr_rows_var2<double>(st, h, *rows, w);
      if (rc) return rc;
      m = (double)rows->ns;
    }
    hipLaunchKernelGGL(snr_scale_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, w.z, w.var2, n, m, tol_snr);
    QTX_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(eig_inv_kernel, dim3(1), dim3(1024), 0, st, w.D, n, rtol, atol, w.scalars + 1, w.z);
  QTX_LAUNCH_CHECK();
  rc = rb(rocblas_dgemv(h, rocblas_operation_none, ni, ni, &one, w.T, ni, w.z, 1, &zero, w.y, 1), "rocblas_dgemv");
  if (rc) return rc;
  int32_t info_h = 0;
  QTX_HIP(hipMemcpyAsync(&info_h, w.info, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  QTX_HIP(hipStreamSynchronize(st));
  if (info_h != 0) return fail(QTX_ESOLVER, "syevd info = " + std::to_string(info_h));
  g_last_path = 2;
  return 1;
}

// ---------------------------------------------------------------- Cholesky
// Right-looking blocked Cholesky (lower, column-major, leading dimension n) of
// `batch` matrices `stride` elements apart.  Each kCholBlock-wide diagonal
// block is factored in CT-column panels by two kernels of this file (panel:
// factor the CT x CT diagonal tile in LDS and solve the tiles below it inside
// the block; update: the rank-CT update of the block's remaining tiles), all
// matrices in one launch, so the latency-bound diagonal work of one matrix
// overlaps that of the others; the rows below a diagonal block are then solved
// by rocBLAS dtrsm and the trailing matrix updated by dsyrk, matrix by matrix,
// so nearly all flops run in level-3 BLAS kernels.  info[b * info_stride + blk]
// receives 1 + the column (within diagonal block blk) of the first
// non-positive pivot of matrix b, 0 if none; as in LAPACK, that matrix's
// factor is not usable then (NaNs propagate, the other matrices are
// unaffected).
constexpr int CT = 64;

// Thread (r, q) = (lane, wave) holds row r, columns 16q .. 16q+15 of a 64 x 64
// tile in registers.  Column step j: the wave owning column j publishes it to
// LDS (two alternating buffers, one barrier per step), every thread reads the
// pivot and the column entries of its own rows/columns and updates its 16
// values.  The column loop runs as 4 groups of 16 unrolled steps, so every
// register index and every "column after j" test is static inside a group; a
// wave's role in a group (its columns all after the group, the group's own
// columns, or all before it) is wave-uniform.  Every workgroup factors the
// panel's diagonal tile; workgroup 0 stores it, workgroup t >= 1 solves the
// t-th tile below it inside the diagonal block, X = A_t L^{-T}, by forward
// substitution along the columns in the same pattern.
__global__ __launch_bounds__(256) void chol_panel_kernel(double* A, int64_t lda, int64_t stride, int64_t d0,
                                                         int64_t kb, int kp, int32_t* info, int64_t info_stride,
                                                         double* __restrict__ dtile) {
  __shared__ double Ls[CT][CT + 1];
  __shared__ double col[2][CT];
  __shared__ double rdg[CT];
  double* Ab = A + (int64_t)blockIdx.y * stride;
  const int tid = threadIdx.x, r = tid % CT;
  const int q = __builtin_amdgcn_readfirstlane(tid / CT), c0 = 16 * q;
  const int64_t p0 = d0 + (int64_t)kp * CT;
  const int w = (int)(kb - (int64_t)kp * CT < CT ? kb - (int64_t)kp * CT : CT);
  double a[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int c = c0 + i;
    double v = (r == c) ? 1.0 : 0.0;  // padding: identity
    if (r < w && c < w && r >= c) v = Ab[(p0 + r) + (p0 + c) * lda];
    a[i] = v;
  }
  int bad = -1;
  for (int jb = 0; jb < 4; ++jb) {
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
      const int j = 16 * jb + jj;
      double* cb = col[jj & 1];
      if (q == jb) cb[r] = a[jj];
      __syncthreads();
      const double djj = cb[j];
      if (!(djj > 0.0) && bad < 0 && j < w) bad = j;  // NaN fails too
      const double rs = 1.0 / sqrt(djj);
      const double lr = cb[r] * (rs * rs);  // L_rj L_cj = A_rj A_cj / d_jj
      if (q > jb) {
#pragma unroll
        for (int i = 0; i < 16; ++i) a[i] -= lr * cb[c0 + i];
      } else if (q == jb) {
#pragma unroll
        for (int i = jj + 1; i < 16; ++i) a[i] -= lr * cb[c0 + i];
        a[jj] *= rs;
        if (r == 0) rdg[j] = rs;
      }
    }
  }
  if (blockIdx.x == 0) {
    if (tid == 0 && bad >= 0) {
      int32_t* slot = info + (int64_t)blockIdx.y * info_stride;
      if (*slot == 0) *slot = kp * CT + bad + 1;
    }
    // staged, not stored into A: the other workgroups of this launch read the
    // unfactored tile from A, possibly after this one is done
    double* dt = dtile + (int64_t)blockIdx.y * CT * CT;
#pragma unroll
    for (int i = 0; i < 16; ++i) dt[(c0 + i) * CT + r] = a[i];
    return;
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) Ls[r][c0 + i] = a[i];
  const int64_t q0 = p0 + (int64_t)blockIdx.x * CT;
  const int64_t hrem = kb - (int64_t)(kp + blockIdx.x) * CT;
  const int h = (int)(hrem < CT ? hrem : CT);
  double x[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int c = c0 + i;
    x[i] = (r < h && c < w) ? Ab[(q0 + r) + (p0 + c) * lda] : 0.0;
  }
  __syncthreads();  // Ls and rdg complete; the column buffers are free again
  for (int jb = 0; jb < 4; ++jb) {
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
      const int c = 16 * jb + jj;
      double* cb = col[jj & 1];
      if (q == jb) {
        x[jj] *= rdg[c];
        cb[r] = x[jj];
      }
      __syncthreads();
      const double xc = cb[r];
      if (q > jb) {
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] -= xc * Ls[c0 + i][c];
      } else if (q == jb) {
#pragma unroll
        for (int i = jj + 1; i < 16; ++i) x[i] -= xc * Ls[c0 + i][c];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int c = c0 + i;
    if (r < h && c < w) Ab[(q0 + r) + (p0 + c) * lda] = x[i];
  }
}

// Workgroup 0: the panel's factored diagonal tile from its staging copy into
// A.  Workgroups 1..: A_ij -= L_i L_j^T for the tiles (i, j), kp < j <= i, of
// the diagonal block (lower triangle only on diagonal tiles); one 64 x 64 tile
// per workgroup, 4 x 4 outputs per thread
__global__ __launch_bounds__(256) void chol_update_kernel(double* A, int64_t lda, int64_t stride, int64_t d0,
                                                          int64_t kb, int kp, const double* __restrict__ dtile) {
  __shared__ double Li[CT][CT + 1];
  __shared__ double Lj[CT][CT + 1];
  double* Ab = A + (int64_t)blockIdx.y * stride;
  const int tid = threadIdx.x;
  if (blockIdx.x == 0) {  // the panel's factored diagonal tile into A (no tile below reads it)
    const int64_t p0 = d0 + (int64_t)kp * CT;
    const int w = (int)(kb - (int64_t)kp * CT < CT ? kb - (int64_t)kp * CT : CT);
    const double* dt = dtile + (int64_t)blockIdx.y * CT * CT;
    for (int e = tid; e < CT * CT; e += 256) {
      const int r = e % CT, c = e / CT;
      if (r < w && c < w && r >= c) Ab[(p0 + r) + (p0 + c) * lda] = dt[c * CT + r];
    }
    return;
  }
  int ti, tj;
  tile_of(blockIdx.x - 1, ti, tj);
  const int i = kp + 1 + ti, j = kp + 1 + tj;
  const int64_t p0 = d0 + (int64_t)kp * CT;
  const int w = (int)(kb - (int64_t)kp * CT < CT ? kb - (int64_t)kp * CT : CT);
  const int64_t ri = d0 + (int64_t)i * CT, rj = d0 + (int64_t)j * CT;
  const int hi = (int)(kb - (int64_t)i * CT < CT ? kb - (int64_t)i * CT : CT);
  const int hj = (int)(kb - (int64_t)j * CT < CT ? kb - (int64_t)j * CT : CT);
  for (int e = tid; e < CT * CT; e += 256) {
    const int r = e % CT, l = e / CT;
    Li[r][l] = (r < hi && l < w) ? Ab[(ri + r) + (p0 + l) * lda] : 0.0;
    Lj[r][l] = (r < hj && l < w) ? Ab[(rj + r) + (p0 + l) * lda] : 0.0;
  }
  __syncthreads();
  const int r0 = (tid % 16) * 4, c0 = (tid / 16) * 4;
  double acc[4][4] = {};
  for (int l = 0; l < w; ++l) {
    double a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = Li[r0 + q][l];
      b[q] = Lj[c0 + q][l];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[q][u] += a[q] * b[u];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int c = c0 + u;
    if (c >= hj) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + q;
      if (r < hi && (i != j || r >= c)) Ab[(ri + r) + (rj + c) * lda] -= acc[q][u];
    }
  }
}

static int blocked_potrf(rocblas_handle h, hipStream_t st, int64_t n, double* A, int64_t stride, int batch,
                         int32_t* info, int64_t info_stride, double* dtile) {
  const double one = 1.0, minus_one = -1.0;
  const rocblas_int ld = (rocblas_int)n;
  int b = 0;
  for (int64_t k = 0; k < n; k += kCholBlock, ++b) {
    const int64_t kb = (n - k < kCholBlock) ? n - k : kCholBlock;
    const int nt = (int)((kb + CT - 1) / CT);
    for (int kp = 0; kp < nt; ++kp) {
      hipLaunchKernelGGL(chol_panel_kernel, dim3((unsigned)(nt - kp), (unsigned)batch), dim3(256), 0, st, A,
                         (int64_t)n, stride, k, kb, kp, info + b, info_stride, dtile);
      const int m = nt - kp - 1;
      hipLaunchKernelGGL(chol_update_kernel, dim3((unsigned)(1 + m * (m + 1) / 2), (unsigned)batch), dim3(256), 0,
                         st, A, (int64_t)n, stride, k, kb, kp, (const double*)dtile);
    }
    QTX_LAUNCH_CHECK();
    const int64_t m = n - k - kb;
    if (m == 0) break;
    // one call per matrix: the strided-batched forms of these two routines
    // run on much slower non-GEMM kernels at these sizes
    for (int bi = 0; bi < batch; ++bi) {
      double* Akk = A + bi * stride + k * n + k;
      double* Apan = Akk + kb;
      int rc = rb(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                rocblas_diagonal_non_unit, (rocblas_int)m, (rocblas_int)kb, &one, Akk, ld, Apan, ld),
                  "rocblas_dtrsm");
      if (rc) return rc;
      double* A22 = A + bi * stride + (k + kb) * n + (k + kb);
      rc = rb(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, (rocblas_int)m, (rocblas_int)kb,
                            &minus_one, Apan, ld, &one, A22, ld),
              "rocblas_dsyrk");
      if (rc) return rc;
    }
  }
  return QTX_OK;
}

// ---------------------------------------------------------------- triangular inverse
// L^-1 in place of a lower-triangular L (column-major, leading dimension ld).
// The kTriLeaf-wide diagonal blocks are inverted by one batched rocSOLVER
// call; the blocks are then joined bottom-up: for L = [A 0; B C] with A and C
// already inverted, the lower-left block becomes -C^-1 (B A^-1).  Both
// triangular products run as DGEMMs over a leaf-aligned split of the triangle
// (a leaf diagonal block multiplies as a full block, zeroed above its
// diagonal first), so nearly all of the n^3 / 3 flops are large DGEMMs: 29 ms
// at n = 16384 on MI355X against 43 ms for rocsolver_dtrtri, with the same
// results to rounding (max difference 6e-16 relative).  W holds the products'
// intermediate B A^-1: at least (n - n/2) x n/2 doubles at leading dimension
// ld.
constexpr int64_t kTriLeaf = 1024;

__global__ __launch_bounds__(256) void zero_leaf_upper_kernel(double* A, int64_t n, int64_t ld) {
  const int64_t o = (int64_t)blockIdx.y * kTriLeaf;
  const int64_t m = (n - o < kTriLeaf) ? n - o : kTriLeaf;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < m * m; e += (int64_t)gridDim.x * 256) {
    const int64_t r = e % m, c = e / m;
    if (r < c) A[(o + r) + (o + c) * ld] = 0.0;
  }
}

// first failure of a chain of library calls; a device-memory size query
// answers with the size statuses, which count as success there
struct BlasChain {
  rocblas_handle h;
  rocblas_status st = rocblas_status_success;
  const char* what = nullptr;
  void operator()(rocblas_status s, const char* w) {
    const bool ok = s == rocblas_status_success || s == rocblas_status_size_increased ||
                    s == rocblas_status_size_unchanged;
    if (!ok && st == rocblas_status_success) {
      st = s;
      what = w;
    }
  }
};

// leaf-aligned split of m > kTriLeaf rows: the first part holds ceil(leaves / 2) leaves
static int64_t tri_split(int64_t m) {
  const int64_t nb = (m + kTriLeaf - 1) / kTriLeaf;
  return (nb + 1) / 2 * kTriLeaf;
}

static void tri_gemm(BlasChain& b, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                     const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  b(rocblas_dgemm(b.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m, (rocblas_int)n,
                  (rocblas_int)k, &alpha, A, (rocblas_int)lda, B, (rocblas_int)ldb, &beta, C, (rocblas_int)ldc),
    "rocblas_dgemm");
}

// W = alpha X T, X: r x t, T: t x t lower triangular starting at a leaf boundary
static void tri_mul_right(BlasChain& b, int64_t r, int64_t t, double alpha, const double* X, int64_t ldx,
                          const double* T, int64_t ldt, double* W, int64_t ldw) {
  if (t <= kTriLeaf) {
    tri_gemm(b, r, t, t, alpha, X, ldx, T, ldt, 0.0, W, ldw);
    return;
  }
  const int64_t t1 = tri_split(t), t2 = t - t1;  // T = [P 0; Q R], X = [X1 X2]
  tri_mul_right(b, r, t1, alpha, X, ldx, T, ldt, W, ldw);                         // X1 P
  tri_gemm(b, r, t1, t2, alpha, X + t1 * ldx, ldx, T + t1, ldt, 1.0, W, ldw);     // + X2 Q
  tri_mul_right(b, r, t2, alpha, X + t1 * ldx, ldx, T + t1 + t1 * ldt, ldt, W + t1 * ldw, ldw);  // X2 R
}

// W = alpha T X, T: t x t lower triangular starting at a leaf boundary, X: t x c
static void tri_mul_left(BlasChain& b, int64_t t, int64_t c, double alpha, const double* T, int64_t ldt,
                         const double* X, int64_t ldx, double* W, int64_t ldw) {
  if (t <= kTriLeaf) {
    tri_gemm(b, t, c, t, alpha, T, ldt, X, ldx, 0.0, W, ldw);
    return;
  }
  const int64_t t1 = tri_split(t), t2 = t - t1;  // T = [P 0; Q R], X = [X1; X2]
  tri_mul_left(b, t1, c, alpha, T, ldt, X, ldx, W, ldw);                                  // P X1
  tri_mul_left(b, t2, c, alpha, T + t1 + t1 * ldt, ldt, X + t1, ldx, W + t1, ldw);        // R X2
  tri_gemm(b, t2, c, t1, alpha, T + t1, ldt, X, ldx, 1.0, W + t1, ldw);                   // + Q X1
}

static void tri_join(BlasChain& b, double* A, int64_t ld, int64_t m, double* W) {
  if (m <= kTriLeaf) return;
  const int64_t m1 = tri_split(m), m2 = m - m1;
  tri_join(b, A, ld, m1, W);
  tri_join(b, A + m1 + m1 * ld, ld, m2, W);
  tri_mul_right(b, m2, m1, 1.0, A + m1, ld, A, ld, W, ld);                    // W = B A^-1
  tri_mul_left(b, m2, m1, -1.0, A + m1 + m1 * ld, ld, W, ld, A + m1, ld);     // B = -C^-1 W
}

// info[i] (i < number of leaves) receives rocSOLVER's status of leaf i.  With
// query set, only the library calls are made (inside a device-memory size
// query): no kernel runs
static int tri_inverse(rocblas_handle h, hipStream_t st, int64_t n, double* A, int64_t ld, double* W,
                       int32_t* info, bool query) {
  BlasChain b{h};
  const int64_t full = n / kTriLeaf, rem = n % kTriLeaf;
  if (full)
    b(rocsolver_dtrtri_strided_batched(h, rocblas_fill_lower, rocblas_diagonal_non_unit, (rocblas_int)kTriLeaf, A,
                                       (rocblas_int)ld, (rocblas_stride)(kTriLeaf * ld + kTriLeaf), info,
                                       (rocblas_int)full),
      "rocsolver_dtrtri_strided_batched");
  if (rem)
    b(rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, (rocblas_int)rem,
                       A + full * kTriLeaf * (ld + 1), (rocblas_int)ld, info + full),
      "rocsolver_dtrtri");
  if (full && !query) {
    hipLaunchKernelGGL(zero_leaf_upper_kernel, dim3(64, (unsigned)(full + (rem ? 1 : 0))), dim3(256), 0, st, A, n,
                       ld);
    QTX_LAUNCH_CHECK();
  }
  tri_join(b, A, ld, n, W);
  if (b.st != rocblas_status_success) return rb(b.st, b.what);
  return QTX_OK;
}

// ---------------------------------------------------------------- phased solve
// A solve runs as three phases so that many solves on different streams can
// be driven from one host thread with batched host syncs
// (qtx_minsr_solve_many): every task's phase p is enqueued before any task's
// phase p + 1 reads results back.
//   phase 1: Gram (and for SR the right-hand side Obar^T ebar); CHOL_SHIFT:
//            shift + potrf + potrs; EIGH_PINV: the first Lanczos steps.
//   phase 2: EIGH_PINV: read the Lanczos coefficients (sync), extend until the
//            largest Ritz value settles, then syrk + three (potrf, refinement)
//            for the partial fractions; dsyevd where that route does not apply.
//   phase 3: read the factorisation status (sync); dsyevd fallback; the update
//            vector x = Obar^T y.
// All phases of different tasks overlap on their streams (the factorisations
// too, unless QTX_SERIAL_FACTOR=1).  A panel kernel of the blocked Cholesky
// once stored its factored diagonal tile into the matrix while other
// workgroups of the same launch could still be reading the unfactored tile;
// delayed workgroups — common when other tasks occupy the GPU — then factored
// garbage (non-positive pivots in 249 of 640 overlapped solves); the tile is
// now staged and stored by the next launch.
// Chebyshev coefficients of h(mu) = g(sigma mu / (1 - mu)), g(v) = v^5 /
// (v^6 + c^6), on [mu_lo, mu_max] (mu = mu_lo + (mu_max - mu_lo) (u + 1) / 2,
// u in [-1, 1]) from kChebNodes Chebyshev points (a discrete cosine
// transform); returns the number of terms kept, 0 when the series does not
// settle within kChebMax.
static int cheb_coeffs_raw(double c, double sigma, double mu_lo, double mu_max, std::vector<double>& a) {
  const int N = kChebNodes;
  std::vector<double> f(N);
  const double pi = 3.14159265358979323846;
  for (int j = 0; j < N; ++j) {
    const double u = cos(pi * (j + 0.5) / N), mu = mu_lo + (mu_max - mu_lo) * (u + 1.0) * 0.5;
    const double tt = sigma * mu / ((1.0 - mu) * c);  // v / c, negative below mu = 0
    const double t5 = tt * tt * tt * tt * tt;
    f[j] = t5 / (t5 * tt + 1.0) / c;
  }
  const int K = kChebMax + 8;
  a.assign(K, 0.0);
  for (int j = 0; j < N; ++j) {  // cos(k theta_j) by the three-term recurrence
    const double c1 = cos(pi * (j + 0.5) / N);
    double ckm = 1.0, ck = c1;
    a[0] += f[j];
    a[1] += f[j] * c1;
    for (int k = 2; k < K; ++k) {
      const double cn = 2.0 * c1 * ck - ckm;
      a[k] += f[j] * cn;
      ckm = ck;
      ck = cn;
    }
  }
  double amax = 0.0;
  for (int k = 0; k < K; ++k) {
    a[k] *= (k == 0 ? 1.0 : 2.0) / N;
    amax = std::max(amax, fabs(a[k]));
  }
  for (int k = 8; k < kChebMax; ++k) {
    bool small = true;
    for (int m = k; m < k + 8; ++m) small = small && fabs(a[m]) < kChebTol * amax;
    if (small) {
      a.resize(k);
      return k;
    }
  }
  return 0;
}

// sigma = c: the series of c h depends on mu_max = 1 / (1 + rel) alone (rel =
// c / lambda_max), so it is computed once per rel and thread
static int cheb_coeffs(double c, double rel, std::vector<double>& a) {
  static thread_local std::vector<std::pair<double, std::vector<double>>> cache;
  for (const auto& e : cache)
    if (e.first == rel) {
      a = e.second;
      for (double& v : a) v /= c;
      return (int)a.size();
    }
  std::vector<double> u;
  const int k = cheb_coeffs_raw(1.0, 1.0, -kChebMargin, 1.0 / (1.0 + rel), u);
  if (k == 0) return 0;
  if (cache.size() >= 16) cache.erase(cache.begin());
  cache.emplace_back(rel, u);
  a = u;
  for (double& v : a) v /= c;
  return k;
}

struct Task {
  qtx_minsr_task* io = nullptr;
  hipStream_t st = nullptr;
  SolveWs w{};
  char* blas_ws = nullptr;  // QTX_SOLVE_BLAS_WS: rocBLAS scratch = the workspace past the library's own part
  size_t blas_bytes = 0;
  bool minnorm = false;
  int64_t n = 0;
  bool failed = false;
  bool need_syevd = false;  // phase 2 runs the dsyevd route
  bool pf = false;          // partial-fraction factorisations enqueued
  bool cf = false;          // shifted Cholesky route enqueued (small cutoffs)
  bool cheb = false;        // Chebyshev route enqueued
  int cheb_terms = 0;
  double c = 0.0, rel = 0.0;  // the cutoff and c / lambda_max
  int mcap = 0, done = 0;   // Lanczos steps available / enqueued
  double lmax = 0.0;
  int lanczos = 0;
  int path = -1, fallback = 0;
};

static int task_fail(Task& t, int rc) {
  t.failed = true;
  t.io->status = rc;
  return rc;
}

static int lanczos_enqueue(Task& t, rocblas_handle h, int to) {
  (void)h;
  double* alpha = t.w.ab;
  double* beta = t.w.ab + kLanczosMax;
  const unsigned gm = (unsigned)((t.n + 3) / 4);
  for (int j = t.done; j < to; ++j) {
    double* v = t.w.V + (j % 3) * t.n;
    double* vp = t.w.V + ((j + 2) % 3) * t.n;
    double* vn = t.w.V + ((j + 1) % 3) * t.n;
    hipLaunchKernelGGL(matvec_rows_kernel<1>, dim3(gm), dim3(256), 0, t.st, t.w.T, t.n, v, t.w.w);
    hipLaunchKernelGGL(lanczos_step_kernel, dim3(1), dim3(1024), 0, t.st, t.n, j, t.w.w, v, vp, vn, alpha, beta);
  }
  QTX_LAUNCH_CHECK();
  t.done = to;
  return QTX_OK;
}

template <typename TI>
static int phase1(Task& t, rocblas_handle h) {
  qtx_minsr_task& io = *t.io;
  const TI* O = (const TI*)io.O;
  int rc;
  if (t.minnorm) {
    rc = launch_gram_nt<TI>(t.st, io.gram, O, io.mean, io.r, io.ns, io.np, t.w.T, t.w.split);
    if (rc) return rc;
    hipLaunchKernelGGL(copy_f64_kernel, dim3(nblk(io.ns, 256)), dim3(256), 0, t.st, io.ebar, t.w.y, io.ns);
  } else {
    // S = Obar^T Obar: Obar formed in FP64 a row block at a time, each block
    // accumulated by one rocBLAS DGEMM (a row-major rows x np block is the
    // column-major np x rows matrix B, S += B B^T)
    const int64_t rows = obar_rows(io.ns, io.np);
    const double one = 1.0, zero = 0.0;
    for (int64_t s0 = 0; s0 < io.ns; s0 += rows) {
      const int64_t nr = std::min(rows, io.ns - s0), tot = nr * io.np;
      hipLaunchKernelGGL(form_obar_kernel<TI>, dim3((unsigned)std::min<int64_t>((tot + 255) / 256, kGridCap)),
                         dim3(256), 0, t.st, O + s0 * io.np, io.mean, io.r ? io.r + s0 : nullptr, nr, io.np,
                         1.0 / (double)io.ns, t.w.obar);
      QTX_LAUNCH_CHECK();
      rc = rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)io.np,
                            (rocblas_int)io.np, (rocblas_int)nr, &one, t.w.obar, (rocblas_int)io.np, t.w.obar,
                            (rocblas_int)io.np, s0 ? &one : &zero, t.w.T, (rocblas_int)io.np),
              "rocblas_dgemm");
      if (rc) return rc;
    }
    rc = launch_obar_t_y<TI>(t.st, O, io.mean, io.r, io.ebar, io.ns, io.np, t.w.y);  // F = Obar^T ebar
    if (rc) return rc;
  }
  const rocblas_int ni = (rocblas_int)t.n;
  const int mode = base_mode(io.mode);
  if (mode == QTX_SOLVE_CHOL_SHIFT) {
    hipLaunchKernelGGL(shift_diag_kernel, dim3(1), dim3(1024), 0, t.st, t.w.T, t.n, io.rshift, io.ashift,
                       t.w.scalars);
    QTX_LAUNCH_CHECK();
    hipLaunchKernelGGL(fill_i32_kernel, dim3(nblk(3 * kInfoPerFactor, 256)), dim3(256), 0, t.st, t.w.info,
                       (int64_t)(3 * kInfoPerFactor), 0);
    rc = blocked_potrf(h, t.st, t.n, t.w.T, 0, 1, t.w.info, kInfoPerFactor, t.w.dtile);
    if (rc) return rc;
    return rb(rocsolver_dpotrs(h, rocblas_fill_lower, ni, 1, t.w.T, ni, t.w.y, ni), "rocsolver_dpotrs");
  }
  if (mode == QTX_SOLVE_EIGH_PINV_DIRECT) {
    t.need_syevd = true;
    return QTX_OK;
  }
  if (io.tol_snr > kSnrMin) {  // the truncation needs the eigenvectors
    t.fallback = 5;
    t.need_syevd = true;
    return QTX_OK;
  }
  // lambda_max = max |eigenvalue| of the PSD T by Lanczos (three-term
  // recurrence; the largest Ritz value is checked every 32 steps)
  hipLaunchKernelGGL(lanczos_init_kernel, dim3(1), dim3(1024), 0, t.st, t.n, t.w.V, t.w.ab + kLanczosMax);
  t.mcap = (int)(t.n < kLanczosMax ? t.n : kLanczosMax);
  // convergence checks every 16 steps when one product streams more than
  // 512 MB of T (n > 8192: 16 products cost far more than the host round
  // trip of a check; a VMC Gram's top eigenvalue, well separated, settles
  // within 8 steps), else after 48 and then every 32
  const int first = t.n > 8192 ? 16 : 48;
  return lanczos_enqueue(t, h, t.mcap < first ? t.mcap : first);
}

static int snr_syevd(Task& t, rocblas_handle h) {
  qtx_minsr_task& io = *t.io;
  SnrRows rows{io.o_dtype, io.O, io.mean, io.r, io.ebar, io.ns, io.np};
  return pinv_syevd(t.st, h, t.n, io.rtol, io.atol, t.w, io.tol_snr, t.minnorm ? nullptr : &rows);
}

// Partial-fraction route: T^2 by dsyrk, the three quadratics (T - a_k)^2 + b_k^2
// factored together, and iterative refinement of the complex shifted systems
// against T itself.
static int pf_enqueue(Task& t, rocblas_handle h, double c, double rel) {
  // refinement contracts by about 4 u / rel^2 per pass (4.4e-8 at rel = 1e-4,
  // 4.4e-6 at 1e-5): two passes reach ~1e-15 for rel >= 1e-4, three for 1e-5
  const int iters = rel >= 1e-4 ? 2 : (rel >= 1e-5 ? 3 : 4);
  const rocblas_int ni = (rocblas_int)t.n;
  const double one = 1.0, zero = 0.0;
  const unsigned gb = (unsigned)nblk(t.n, 256);
  const unsigned gm = (unsigned)((t.n + 3) / 4);
  int rc = rb(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, ni, ni, &one, t.w.T, ni, &zero,
                            t.w.T2, ni),
              "rocblas_dsyrk");
  if (rc) return rc;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(nblk(3 * kInfoPerFactor, 256)), dim3(256), 0, t.st, t.w.info,
                     (int64_t)(3 * kInfoPerFactor), 0);
  // pole k = c e^{i pi (2k+1)/6}: quadratic k into matrix k of the T2 block,
  // k = 0 last because it overwrites T^2 in place
  const int64_t nn = t.n * t.n;
  double pa[3], pb[3];
  for (int k = 0; k < 3; ++k) {
    const double th = 3.14159265358979323846 * (2 * k + 1) / 6.0;
    pa[k] = (k == 1) ? 0.0 : c * cos(th);
    pb[k] = c * sin(th);
  }
  const unsigned gq = (unsigned)std::min<int64_t>((nn + 255) / 256, kGridCap);
  for (int k : {1, 2, 0}) {
    hipLaunchKernelGGL(form_quad_kernel, dim3(gq), dim3(256), 0, t.st, t.w.T2, t.w.T, t.n, 2.0 * pa[k], c * c,
                       t.w.T2 + k * nn);
    QTX_LAUNCH_CHECK();
  }
  rc = blocked_potrf(h, t.st, t.n, t.w.T2, nn, 3, t.w.info, kInfoPerFactor, t.w.dtile);
  if (rc) return rc;
  // iterative refinement of the three complex systems (T - p_k) z = y,
  // the inner solves batched over the three factors, the products with T
  // over all six vectors in one pass
  for (int k = 0; k < 3; ++k)
    hipLaunchKernelGGL(refine_init_kernel, dim3(gb), dim3(256), 0, t.st, t.w.y, t.n, t.w.R + 2 * k * t.n,
                       t.w.Z + 2 * k * t.n);
  for (int it = 0; it < iters; ++it) {
    rc = rb(rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                                          rocblas_diagonal_non_unit, ni, 2, &one, t.w.T2, ni, nn, t.w.R, ni,
                                          2 * t.n, 3),
            "rocblas_dtrsm_strided_batched");
    if (rc) return rc;
    rc = rb(rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                                          rocblas_diagonal_non_unit, ni, 2, &one, t.w.T2, ni, nn, t.w.R, ni,
                                          2 * t.n, 3),
            "rocblas_dtrsm_strided_batched");
    if (rc) return rc;
    hipLaunchKernelGGL(matvec_rows_kernel<6>, dim3(gm), dim3(256), 0, t.st, t.w.T, t.n, t.w.R, t.w.TW);
    for (int k = 0; k < 3; ++k)
      hipLaunchKernelGGL(refine_update_kernel, dim3(gb), dim3(256), 0, t.st, t.w.R + 2 * k * t.n,
                         t.w.TW + 2 * k * t.n, t.n, pa[k], pb[k], t.w.Z + 2 * k * t.n);
    if (it + 1 < iters) {
      hipLaunchKernelGGL(matvec_rows_kernel<6>, dim3(gm), dim3(256), 0, t.st, t.w.T, t.n, t.w.Z, t.w.TW);
      for (int k = 0; k < 3; ++k)
        hipLaunchKernelGGL(refine_residual_kernel, dim3(gb), dim3(256), 0, t.st, t.w.y, t.w.Z + 2 * k * t.n,
                           t.w.TW + 2 * k * t.n, t.n, pa[k], pb[k], t.w.R + 2 * k * t.n);
    }
  }
  for (int k = 0; k < 3; ++k)
    hipLaunchKernelGGL(refine_accum_kernel, dim3(gb), dim3(256), 0, t.st, t.w.Z + 2 * k * t.n, t.n, k == 0 ? 1 : 0,
                       t.w.x);
  QTX_LAUNCH_CHECK();
  t.pf = true;
  return QTX_OK;
}

static int phase2(Task& t, rocblas_handle h) {
  qtx_minsr_task& io = *t.io;
  if (base_mode(io.mode) == QTX_SOLVE_CHOL_SHIFT) return QTX_OK;
  if (!t.need_syevd) {
    std::vector<double> ha(kLanczosMax), hb(kLanczosMax + 1);
    double lmax = 0.0;
    int msteps = 0;
    bool bad = false;
    for (;;) {
      QTX_HIP(hipMemcpyAsync(ha.data(), t.w.ab, (size_t)t.done * 8, hipMemcpyDeviceToHost, t.st));
      QTX_HIP(hipMemcpyAsync(hb.data(), t.w.ab + kLanczosMax, (size_t)(t.done + 1) * 8, hipMemcpyDeviceToHost, t.st));
      QTX_HIP(hipStreamSynchronize(t.st));
      int m = t.done;
      bool exact = false;
      for (int j = 0; j < t.done; ++j) {
        if (!std::isfinite(ha[j]) || !std::isfinite(hb[j + 1])) {
          bad = true;
          break;
        }
        if (hb[j + 1] == 0.0) {  // invariant Krylov subspace: the Ritz values are exact
          m = j + 1;
          exact = true;
          break;
        }
      }
      if (bad) break;
      lmax = tridiag_max_eig(ha.data(), hb.data(), m);
      msteps = m;
      if (exact || t.done >= t.mcap) break;
      double prev = tridiag_max_eig(ha.data(), hb.data(), m - 8);
      if (fabs(lmax - prev) <= 1e-14 * fabs(lmax)) break;
      const int more = t.n > 8192 ? 16 : 32;
      int rc = lanczos_enqueue(t, h, t.done + more < t.mcap ? t.done + more : t.mcap);
      if (rc) return rc;
    }
    t.lanczos = msteps;
    t.lmax = lmax;
    const double c = io.rtol * lmax + io.atol, rel = io.rtol + io.atol / lmax;  // c / lmax, exact for atol = 0
    if (bad) {
      t.fallback = 2;
    } else if (!(lmax > 0.0) || !std::isfinite(lmax)) {
      t.fallback = 3;
    } else if (!std::isfinite(c) || (!(rel >= 1e-5) && t.minnorm && io.r != nullptr)) {
      t.fallback = 1;  // (reweighted MinSR: the null direction of T is not 1)
    }
    if (t.fallback) {
      t.need_syevd = true;
    } else if (!(rel >= 1e-5)) {
      // small cutoff: T_d x = b' by the shifted factorisation and refinement
      const rocblas_int ni = (rocblas_int)t.n;
      const double one = 1.0;
      const double mu_n = t.minnorm ? lmax / (double)t.n : 0.0;
      const int64_t nn = t.n * t.n;
      const unsigned gq = (unsigned)std::min<int64_t>((nn + 255) / 256, kGridCap);
      const unsigned gm = (unsigned)((t.n + 3) / 4);
      hipLaunchKernelGGL(fill_i32_kernel, dim3(nblk(3 * kInfoPerFactor, 256)), dim3(256), 0, t.st, t.w.info,
                         (int64_t)(3 * kInfoPerFactor), 0);
      hipLaunchKernelGGL(form_shifted_kernel, dim3(gq), dim3(256), 0, t.st, t.w.T, t.n, mu_n, kShiftCut * c, t.w.T2);
      QTX_LAUNCH_CHECK();
      int rc = blocked_potrf(h, t.st, t.n, t.w.T2, nn, 1, t.w.info, kInfoPerFactor, t.w.dtile);
      if (rc) return rc;
      double* r = t.w.R;
      double* z = t.w.R + t.n;
      double* p = t.w.R + 2 * t.n;
      double* q = t.w.TW;
      double* s = t.w.scalars + 4;
      hipLaunchKernelGGL(shifted_init_kernel, dim3(1), dim3(1024), 0, t.st, t.w.y, t.n, t.minnorm ? 1 : 0, t.w.z,
                         t.w.x, r, z);
      QTX_LAUNCH_CHECK();
      // L^-1 in place of L: each preconditioner application is then two
      // bandwidth-bound triangular products instead of two triangular solves
      // (a single-vector dtrsm at n = 16384 ran at ~10 ms)
      if ((rc = tri_inverse(h, t.st, t.n, t.w.T2, t.n, t.w.T2 + nn, t.w.info + kInfoPerFactor, false))) return rc;
      auto precondition = [&]() -> int {  // z <- G^-1 z = L^-T L^-1 z
        int rc2 = rb(rocblas_dtrmv(h, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit, ni,
                                   t.w.T2, ni, z, 1),
                     "rocblas_dtrmv");
        if (rc2) return rc2;
        return rb(rocblas_dtrmv(h, rocblas_fill_lower, rocblas_operation_transpose, rocblas_diagonal_non_unit, ni,
                                t.w.T2, ni, z, 1),
                  "rocblas_dtrmv");
      };
      if ((rc = precondition())) return rc;
      hipLaunchKernelGGL(pcg_step2_kernel, dim3(1), dim3(1024), 0, t.st, r, z, t.n, 1, p, s);
      for (int it = 0; it < kShiftIters; ++it) {
        hipLaunchKernelGGL(matvec_rows_kernel<1>, dim3(gm), dim3(256), 0, t.st, t.w.T, t.n, p, q);
        hipLaunchKernelGGL(pcg_step1_kernel, dim3(1), dim3(1024), 0, t.st, q, p, t.n, mu_n, t.w.x, r, z, s);
        QTX_LAUNCH_CHECK();
        if ((rc = precondition())) return rc;
        hipLaunchKernelGGL(pcg_step2_kernel, dim3(1), dim3(1024), 0, t.st, r, z, t.n, 0, p, s);
      }
      hipLaunchKernelGGL(matvec_rows_kernel<1>, dim3(gm), dim3(256), 0, t.st, t.w.T, t.n, t.w.x, q);
      hipLaunchKernelGGL(pcg_check_kernel, dim3(1), dim3(1024), 0, t.st, t.w.z, q, t.w.x, t.n, mu_n,
                         t.w.scalars + 6);
      QTX_LAUNCH_CHECK();
      t.cf = true;
    }
    t.c = c;
    t.rel = rel;
    std::vector<double> ca;
    // sigma = c; mu_max from rel slightly reduced: the Ritz value may sit
    // just below lambda_max
    const double sigma = c;
    const double rel_s = rel / (1.0 + 1e-9);
    const double mu_max = 1.0 / (1.0 + rel_s), mu_lo = -kChebMargin;
    if (!t.need_syevd && !t.cf && (t.n >= kChebMinN || (io.mode & QTX_SOLVE_CHEB)) &&
        (t.cheb_terms = cheb_coeffs(c, rel_s, ca)) > 0) {
      // Chebyshev route: G = T + sigma I = L L^T, P = L^-1, then
      // x = sum_k a_k T_k(A) y with A = (2 / mu_max) (I - sigma P^T P) - I
      const int64_t nn = t.n * t.n;
      const unsigned gq = (unsigned)std::min<int64_t>((nn + 255) / 256, kGridCap);
      const unsigned gm = (unsigned)((t.n + 3) / 4);
      hipLaunchKernelGGL(fill_i32_kernel, dim3(nblk(3 * kInfoPerFactor, 256)), dim3(256), 0, t.st, t.w.info,
                         (int64_t)(3 * kInfoPerFactor), 0);
      hipLaunchKernelGGL(form_shifted_kernel, dim3(gq), dim3(256), 0, t.st, t.w.T, t.n, 0.0, -sigma, t.w.T2);
      QTX_LAUNCH_CHECK();
      int rc = blocked_potrf(h, t.st, t.n, t.w.T2, nn, 1, t.w.info, kInfoPerFactor, t.w.dtile);
      if (rc) return rc;
      if ((rc = tri_inverse(h, t.st, t.n, t.w.T2, t.n, t.w.T2 + nn, t.w.info + kInfoPerFactor, false))) return rc;
      const int nt = (int)((t.n + 63) / 64);
      hipLaunchKernelGGL(tri_mirror_kernel, dim3((unsigned)((int64_t)nt * (nt + 1) / 2)), dim3(256), 0, t.st, t.w.T2,
                         t.n, nt);
      double* B[2] = {t.w.R, t.w.R + t.n};  // t_k, t_{k +- 1}
      double* u = t.w.R + 2 * t.n;
      hipLaunchKernelGGL(copy_f64_kernel, dim3(nblk(t.n, 256)), dim3(256), 0, t.st, t.w.y, B[0], t.n);
      const double span = mu_max - mu_lo;
      const double alpha = (2.0 - mu_max - mu_lo) / span, beta = 2.0 * sigma / span;
      for (int k = 0; k + 1 < t.cheb_terms; ++k) {
        const double* tcur = B[k & 1];
        double* tprev = B[(k + 1) & 1];
        hipLaunchKernelGGL(tri_rows_kernel<true>, dim3(gm), dim3(256), 0, t.st, t.w.T2, t.n, tcur, u);
        hipLaunchKernelGGL(cheb_rows_kernel, dim3(gm), dim3(256), 0, t.st, t.w.T2, t.n, u, tcur, tprev, alpha, beta,
                           ca[0], ca[k + 1], k == 0 ? 1 : 0, t.w.x);
      }
      // |t_last|^2 and |y|^2 for the growth check
      hipLaunchKernelGGL(sq_norms_kernel, dim3(1), dim3(1024), 0, t.st, B[(t.cheb_terms - 1) & 1], t.w.y, t.n,
                         t.w.scalars + 4);
      QTX_LAUNCH_CHECK();
      t.cheb = true;
    } else if (!t.need_syevd && !t.cf) {
      int rc = pf_enqueue(t, h, c, rel);
      if (rc) return rc;
    }
  }
  if (t.need_syevd) {
    int rc = snr_syevd(t, h);
    if (rc < 0) return rc;
    t.path = 2;
  }
  return QTX_OK;
}

template <typename TI>
static int phase3(Task& t, rocblas_handle h) {
  qtx_minsr_task& io = *t.io;
  const bool chol = base_mode(io.mode) == QTX_SOLVE_CHOL_SHIFT;
  if (t.cheb) {
    std::vector<int32_t> info_h(2 * kInfoPerFactor);  // the factorisation's blocks, then the inverse's leaves
    double nrm[2];                                     // |t_last|^2, |y|^2
    QTX_HIP(hipMemcpyAsync(info_h.data(), t.w.info, info_h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, t.st));
    QTX_HIP(hipMemcpyAsync(nrm, t.w.scalars + 4, sizeof(nrm), hipMemcpyDeviceToHost, t.st));
    QTX_HIP(hipStreamSynchronize(t.st));
    bool ok = true;
    for (int32_t v : info_h) ok = ok && (v == 0);
    const bool bounded = std::isfinite(nrm[0]) && nrm[0] <= kChebGrowth * kChebGrowth * nrm[1];
    if (ok && bounded) {
      hipLaunchKernelGGL(copy_f64_kernel, dim3(nblk(t.n, 256)), dim3(256), 0, t.st, t.w.x, t.w.y, t.n);
      QTX_LAUNCH_CHECK();
      t.path = 4;
    } else if (ok) {  // T has eigenvalues below the series' interval: partial fractions
      t.fallback = 8;
      int rc = pf_enqueue(t, h, t.c, t.rel);
      if (rc) return rc;
    } else {  // T and y are intact: take the dsyevd route
      t.fallback = 4;
      int rc = snr_syevd(t, h);
      if (rc < 0) return rc;
      t.path = 2;
    }
  }
  if (chol || t.pf) {
    std::vector<int32_t> info_h(3 * kInfoPerFactor);
    QTX_HIP(hipMemcpyAsync(info_h.data(), t.w.info, info_h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, t.st));
    QTX_HIP(hipStreamSynchronize(t.st));
    int bad = -1;  // first diagonal block whose factorisation failed
    const int nslots = chol ? kInfoPerFactor : 3 * kInfoPerFactor;
    for (int i = 0; i < nslots && bad < 0; ++i)
      if (info_h[i] != 0) bad = i;
    if (chol) {
      if (bad >= 0)
        return fail(QTX_ESOLVER, "potrf failed in diagonal block " + std::to_string(bad) +
                                     " (info " + std::to_string(info_h[bad]) + ")");
      t.path = 0;
    } else if (bad >= 0) {  // T and y are intact: take the dsyevd route
      t.fallback = 4;
      int rc = snr_syevd(t, h);
      if (rc < 0) return rc;
      t.path = 2;
    } else {
      hipLaunchKernelGGL(copy_f64_kernel, dim3(nblk(t.n, 256)), dim3(256), 0, t.st, t.w.x, t.w.y, t.n);
      QTX_LAUNCH_CHECK();
      t.path = 1;
    }
  }
  if (t.cf) {
    std::vector<int32_t> info_h(2 * kInfoPerFactor);  // the factorisation's blocks, then the inverse's leaves
    double nrm[2];
    QTX_HIP(hipMemcpyAsync(info_h.data(), t.w.info, info_h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, t.st));
    QTX_HIP(hipMemcpyAsync(nrm, t.w.scalars + 6, sizeof(nrm), hipMemcpyDeviceToHost, t.st));
    QTX_HIP(hipStreamSynchronize(t.st));
    bool ok = true;
    for (int32_t v : info_h) ok = ok && (v == 0);
    // nrm: |b' - T_d x|^2, |x|^2 of the result
    const double res = sqrt(nrm[0]), xn = sqrt(nrm[1]);
    const bool converged = std::isfinite(res) && std::isfinite(xn) && res <= kShiftAccept * t.lmax * xn;
    if (ok && converged) {
      hipLaunchKernelGGL(copy_f64_kernel, dim3(nblk(t.n, 256)), dim3(256), 0, t.st, t.w.x, t.w.y, t.n);
      QTX_LAUNCH_CHECK();
      t.path = 3;
    } else {  // T and y are intact: take the dsyevd route
      t.fallback = ok ? 7 : 6;
      int rc = snr_syevd(t, h);
      if (rc < 0) return rc;
      t.path = 2;
    }
  }
  if (t.minnorm) {
    int rc = launch_obar_t_y<TI>(t.st, (const TI*)io.O, io.mean, io.r, t.w.y, io.ns, io.np, io.step);
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL(copy_f64_kernel, dim3(nblk(io.np, 256)), dim3(256), 0, t.st, t.w.y, io.step, io.np);
    QTX_LAUNCH_CHECK();
  }
  return QTX_OK;
}

static int check_task(const qtx_minsr_task& io) {
  QTX_REQUIRE(io.o_dtype == QTX_F32 || io.o_dtype == QTX_F64, "o_dtype must be QTX_F32 or QTX_F64");
  QTX_REQUIRE(io.ns > 0 && io.np > 0 && io.O && io.mean && io.ebar && io.step, "bad arguments");
  const int mode = base_mode(io.mode);
  QTX_REQUIRE(mode == QTX_SOLVE_CHOL_SHIFT || mode == QTX_SOLVE_EIGH_PINV || mode == QTX_SOLVE_EIGH_PINV_DIRECT,
              "unknown solver mode");
  QTX_REQUIRE((io.mode & ~(0xff | QTX_SOLVE_MINNORM | QTX_SOLVE_LSTSQ | QTX_SOLVE_SNR | QTX_SOLVE_BLAS_WS |
                            QTX_SOLVE_CHEB)) == 0 &&
                  !((io.mode & QTX_SOLVE_MINNORM) && (io.mode & QTX_SOLVE_LSTSQ)),
              "bad solver mode flags");
  QTX_REQUIRE(io.gram == QTX_GRAM_FP64 || io.gram == QTX_GRAM_FP32 || io.gram == QTX_GRAM_BF16X3,
              "unknown gram mode");
  QTX_REQUIRE((use_minnorm(io.ns, io.np, io.mode) ? io.ns : io.np) <= 131072, "matrix too large (n > 131072)");
  QTX_REQUIRE(!(io.tol_snr > kSnrMin) || mode == QTX_SOLVE_CHOL_SHIFT || (io.mode & QTX_SOLVE_SNR),
              "tol_snr > 0 needs QTX_SOLVE_SNR in the mode (and its workspace)");
  size_t need = ws_layout(io.ns, io.np, io.mode, io.gram, nullptr, nullptr);
  if (io.mode & QTX_SOLVE_BLAS_WS) need += 256;  // a non-empty rocBLAS scratch tail
  if (io.ws == nullptr || io.ws_bytes < need)
    return fail(QTX_EWORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  return QTX_OK;
}

// rocBLAS / rocSOLVER scratch of one solve: a device-memory size query (no
// kernels run, nothing is read or written) over every library call the solve
// can make, with its dimensions and with pointers into the caller's workspace
static int blas_query(rocblas_handle h, int64_t ns, int64_t np, int mode, SolveWs& w, size_t* bytes) {
  const bool mn = use_minnorm(ns, np, mode);
  const int64_t n = mn ? ns : np, nn = n * n;
  const rocblas_int ni = (rocblas_int)n;
  const int bm = base_mode(mode);
  const double one = 1.0, zero = 0.0, minus_one = -1.0;
  if (rocblas_start_device_memory_size_query(h) != rocblas_status_success)
    return fail(QTX_ESOLVER, "rocblas_start_device_memory_size_query failed");
  auto ok = [](rocblas_status s) {
    return s == rocblas_status_success || s == rocblas_status_size_increased || s == rocblas_status_size_unchanged;
  };
  bool good = true;
  if (!mn) {  // SR branch: S by row blocks of Obar, and the SNR products
    const rocblas_int npi = (rocblas_int)np, rows = (rocblas_int)obar_rows(ns, np);
    good &= ok(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, npi, npi, rows, &one, w.obar,
                             npi, w.obar, npi, &zero, w.T, npi));
    if (mode & QTX_SOLVE_SNR)
      good &= ok(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, npi, rows, npi, &one, w.T, npi,
                               w.obar, npi, &zero, w.cbuf, npi));
  }
  if (bm == QTX_SOLVE_CHOL_SHIFT || bm == QTX_SOLVE_EIGH_PINV) {  // panels of the blocked Cholesky
    double* A = (bm == QTX_SOLVE_CHOL_SHIFT) ? w.T : w.T2;
    for (int64_t k = 0; k < n; k += kCholBlock) {
      const int64_t kb = (n - k < kCholBlock) ? n - k : kCholBlock, m = n - k - kb;
      if (m == 0) break;
      double* Akk = A + k * n + k;
      good &= ok(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                               rocblas_diagonal_non_unit, (rocblas_int)m, (rocblas_int)kb, &one, Akk, ni, Akk + kb,
                               ni));
      good &= ok(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, (rocblas_int)m, (rocblas_int)kb,
                               &minus_one, Akk + kb, ni, &one, A + (k + kb) * n + (k + kb), ni));
    }
  }
  if (bm == QTX_SOLVE_CHOL_SHIFT) {
    good &= ok(rocsolver_dpotrs(h, rocblas_fill_lower, ni, 1, w.T, ni, w.y, ni));
  } else {  // the dsyevd route (EIGH_PINV: its fallback)
    good &= ok(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, ni, w.T, ni, w.D, w.E, w.info));
    good &= ok(rocblas_dgemv(h, rocblas_operation_transpose, ni, ni, &one, w.T, ni, w.y, 1, &zero, w.z, 1));
    good &= ok(rocblas_dgemv(h, rocblas_operation_none, ni, ni, &one, w.T, ni, w.z, 1, &zero, w.y, 1));
  }
  if (bm == QTX_SOLVE_EIGH_PINV) {  // small-cutoff and Chebyshev routes: inverse factor and its products
    good &= tri_inverse(h, nullptr, n, w.T2, n, w.T2 + nn, w.info, true) == QTX_OK;
    for (rocblas_operation op : {rocblas_operation_none, rocblas_operation_transpose})
      good &= ok(rocblas_dtrmv(h, rocblas_fill_lower, op, rocblas_diagonal_non_unit, ni, w.T2, ni, w.R, 1));
  }
  if (bm == QTX_SOLVE_EIGH_PINV) {  // partial fractions: T^2 and the refinement's triangular solves
    good &= ok(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, ni, ni, &one, w.T, ni, &zero, w.T2, ni));
    for (rocblas_operation op : {rocblas_operation_none, rocblas_operation_transpose})
      good &= ok(rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_lower, op, rocblas_diagonal_non_unit,
                                               ni, 2, &one, w.T2, ni, nn, w.R, ni, 2 * n, 3));
  }
  size_t sz = 0;
  const rocblas_status st = rocblas_stop_device_memory_size_query(h, &sz);
  if (!good || st != rocblas_status_success) return fail(QTX_ESOLVER, "rocBLAS device memory size query failed");
  // at least 256 bytes: a zero-byte workspace would put the handle back under
  // rocBLAS's own memory management
  *bytes = std::max<size_t>(256, (sz + 255) & ~(size_t)255);
  return QTX_OK;
}

// QTX_SERIAL_FACTOR=1 runs the factorisations of different tasks one after
// another (default: they overlap on their streams)
static bool serial_factorisations() {
  static const bool v = [] {
    const char* e = getenv("QTX_SERIAL_FACTOR");
    return e && e[0] == '1';
  }();
  return v;
}

static int solve_many(qtx_minsr_task* tasks, int ntasks) {
  std::vector<Task> ts((size_t)ntasks);
  int first_rc = QTX_OK;
  std::string first_msg;
  auto note = [&](Task& t, int rc) {
    task_fail(t, rc);
    if (first_rc == QTX_OK) {
      first_rc = rc;
      first_msg = qtx_last_error();
    }
  };
  for (int i = 0; i < ntasks; ++i) {
    Task& t = ts[i];
    t.io = &tasks[i];
    t.io->status = QTX_OK;
    t.io->path = -1;
    t.io->fallback = 0;
    int rc = check_task(tasks[i]);
    if (rc) {
      note(t, rc);
      continue;
    }
    t.st = as_stream(tasks[i].stream);
    t.minnorm = use_minnorm(tasks[i].ns, tasks[i].np, tasks[i].mode);
    t.n = t.minnorm ? tasks[i].ns : tasks[i].np;
    const size_t own = ws_layout(tasks[i].ns, tasks[i].np, tasks[i].mode, tasks[i].gram, &t.w, (char*)tasks[i].ws);
    if (tasks[i].mode & QTX_SOLVE_BLAS_WS) {
      t.blas_ws = (char*)tasks[i].ws + own;
      t.blas_bytes = tasks[i].ws_bytes - own;
    }
  }
  rocblas_handle h = nullptr;
  for (int phase = 1; phase <= 3; ++phase) {
    for (Task& t : ts) {
      if (t.failed) continue;
      const bool user = (t.io->mode & QTX_SOLVE_BLAS_WS) != 0;
      int rc = get_blas(t.st, &h, t.blas_ws, t.blas_bytes, user);
      if (rc == QTX_OK) {
        rocblas_pointer_mode pm;
        rocblas_get_pointer_mode(h, &pm);
        rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
        const bool f32 = t.io->o_dtype == QTX_F32;
        if (phase == 1) rc = f32 ? phase1<float>(t, h) : phase1<double>(t, h);
        else if (phase == 2) rc = phase2(t, h);
        else rc = f32 ? phase3<float>(t, h) : phase3<double>(t, h);
        // factorisations of one task finish before the next task's start
        const bool factored = (phase == 1 && base_mode(t.io->mode) == QTX_SOLVE_CHOL_SHIFT) || (phase == 2 && t.pf);
        if (rc == QTX_OK && factored && ntasks > 1 && serial_factorisations())
          rc = hip_check(hipStreamSynchronize(t.st), "hipStreamSynchronize");
        rocblas_set_pointer_mode(h, pm);
      }
      if (rc) note(t, rc);
    }
  }
  for (Task& t : ts) {
    t.io->path = t.path;
    t.io->fallback = t.fallback;
  }
  if (ntasks > 0) {
    const Task& t = ts[ntasks - 1];
    g_last_path = t.failed ? -1 : t.path;
    g_last_lanczos = t.lanczos;
    g_last_lmax = t.lmax;
    g_last_fallback = t.fallback;
  }
  if (first_rc != QTX_OK) return fail(first_rc, first_msg);
  return QTX_OK;
}

extern "C" {

int qtx_colmean(void* stream, int o_dtype, const void* O, int64_t ns, int64_t np, double* mean) {
  QTX_REQUIRE(o_dtype == QTX_F32 || o_dtype == QTX_F64, "o_dtype must be QTX_F32 or QTX_F64");
  QTX_REQUIRE(ns > 0 && np > 0 && O && mean, "bad arguments");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(fill_f64_kernel, dim3(nblk(np, 256)), dim3(256), 0, st, mean, np, 0.0);
  int64_t chunks = (ns + 127) / 128;
  if (chunks > 128) chunks = 128;
  int64_t rpc = (ns + chunks - 1) / chunks;
  chunks = (ns + rpc - 1) / rpc;
  dim3 g(nblk(np, 256), (unsigned)chunks);
  if (o_dtype == QTX_F32)
    hipLaunchKernelGGL(colsum_kernel<float>, g, dim3(256), 0, st, (const float*)O, ns, np, rpc, mean);
  else
    hipLaunchKernelGGL(colsum_kernel<double>, g, dim3(256), 0, st, (const double*)O, ns, np, rpc, mean);
  hipLaunchKernelGGL(scale_f64_kernel, dim3(nblk(np, 256)), dim3(256), 0, st, mean, np, 1.0 / (double)ns);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_obar_matvec(void* stream, int o_dtype, const void* O, const double* mean, const double* r, int64_t ns,
                    int64_t np, double alpha, const double* v, double beta, double* y) {
  QTX_REQUIRE(o_dtype == QTX_F32 || o_dtype == QTX_F64, "o_dtype must be QTX_F32 or QTX_F64");
  QTX_REQUIRE(ns > 0 && np > 0 && O && mean && v && y, "bad arguments");
  hipStream_t st = as_stream(stream);
  dim3 g(nblk(ns, 256 / WAVE));
  if (o_dtype == QTX_F32)
    hipLaunchKernelGGL(obar_matvec_kernel<float>, g, dim3(256), 0, st, (const float*)O, mean, r, ns, np,
                       1.0 / (double)ns, alpha, v, beta, y);
  else
    hipLaunchKernelGGL(obar_matvec_kernel<double>, g, dim3(256), 0, st, (const double*)O, mean, r, ns, np,
                       1.0 / (double)ns, alpha, v, beta, y);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_scale_columns(void* stream, int o_dtype, void* O, double* mean, int64_t ns, int64_t np, const double* d) {
  QTX_REQUIRE(o_dtype == QTX_F32 || o_dtype == QTX_F64, "o_dtype must be QTX_F32 or QTX_F64");
  QTX_REQUIRE(ns > 0 && np > 0 && O && d, "bad arguments");
  hipStream_t st = as_stream(stream);
  const unsigned grid = (unsigned)std::min<int64_t>((ns * np + 255) / 256, kGridCap);
  if (o_dtype == QTX_F32)
    hipLaunchKernelGGL(scale_columns_kernel<float>, dim3(grid), dim3(256), 0, st, (float*)O, ns, np, d);
  else
    hipLaunchKernelGGL(scale_columns_kernel<double>, dim3(grid), dim3(256), 0, st, (double*)O, ns, np, d);
  if (mean) hipLaunchKernelGGL(mul_f64_kernel, dim3(nblk(np, 256)), dim3(256), 0, st, mean, d, np);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_ebar(void* stream, const double* eloc, const double* r, int64_t ns, double* ebar, double* out2) {
  QTX_REQUIRE(ns > 0 && eloc && ebar && out2, "bad arguments");
  hipLaunchKernelGGL(ebar_kernel, dim3(1), dim3(1024), 0, as_stream(stream), eloc, r, ns, ebar, out2);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

size_t qtx_minsr_ws_bytes(int64_t ns, int64_t np, int mode, int gram) {
  if (ns <= 0 || np <= 0) return 0;
  return ws_layout(ns, np, mode, gram, nullptr, nullptr);
}

int qtx_minsr_solve(void* stream, int o_dtype, const void* O, const double* mean, const double* r,
                    const double* ebar, int64_t ns, int64_t np, int mode, int gram, double rshift, double ashift,
                    double rtol, double atol, double* step, void* ws, size_t ws_bytes) {
  qtx_minsr_task task{};
  task.stream = stream;
  task.o_dtype = o_dtype;
  task.mode = mode;
  task.gram = gram;
  task.O = O;
  task.mean = mean;
  task.r = r;
  task.ebar = ebar;
  task.ns = ns;
  task.np = np;
  task.rshift = rshift;
  task.ashift = ashift;
  task.rtol = rtol;
  task.atol = atol;
  task.step = step;
  task.ws = ws;
  task.ws_bytes = ws_bytes;
  g_last_path = -1;
  g_last_lanczos = 0;
  g_last_lmax = 0.0;
  g_last_fallback = 0;
  return solve_many(&task, 1);
}

int qtx_minsr_solve_many(qtx_minsr_task* tasks, int32_t ntasks) {
  QTX_REQUIRE(ntasks >= 0 && (ntasks == 0 || tasks != nullptr), "bad task array");
  return solve_many(tasks, ntasks);
}

int qtx_minsr_last_path(int* path, int* lanczos_steps, double* lambda_max, int* fallback) {
  QTX_REQUIRE(path && lanczos_steps && lambda_max, "bad arguments");
  if (fallback) *fallback = g_last_fallback;
  *path = g_last_path;
  *lanczos_steps = g_last_lanczos;
  *lambda_max = g_last_lmax;
  return QTX_OK;
}

size_t qtx_gram_ws_bytes(int64_t ns, int64_t np, int gram) {
  if (ns <= 0 || np <= 0 || gram != QTX_GRAM_BF16X3 || ns >= np) return 0;
  return g3_bytes(ns, np);
}

int qtx_gram(void* stream, int o_dtype, const void* O, const double* mean, const double* r, int64_t ns,
             int64_t np, int gram, double* T, void* ws, size_t ws_bytes) {
  QTX_REQUIRE(o_dtype == QTX_F32 || o_dtype == QTX_F64, "o_dtype must be QTX_F32 or QTX_F64");
  QTX_REQUIRE(ns > 0 && np > 0 && O && mean && T, "bad arguments");
  QTX_REQUIRE(gram == QTX_GRAM_FP64 || gram == QTX_GRAM_FP32 || gram == QTX_GRAM_BF16X3, "unknown gram mode");
  const size_t need = qtx_gram_ws_bytes(ns, np, gram);
  if (need > 0 && (ws == nullptr || ws_bytes < need))
    return fail(QTX_EWORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  hipStream_t st = as_stream(stream);
  if (ns < np) {
    if (o_dtype == QTX_F32) return launch_gram_nt<float>(st, gram, (const float*)O, mean, r, ns, np, T, ws);
    return launch_gram_nt<double>(st, gram, (const double*)O, mean, r, ns, np, T, ws);
  }
  dim3 g(nblk(np, 16), nblk(np, 16));
  if (o_dtype == QTX_F32)
    hipLaunchKernelGGL(gram_tn_kernel<float>, g, dim3(256), 0, st, (const float*)O, mean, r, ns, np,
                       1.0 / (double)ns, T);
  else
    hipLaunchKernelGGL(gram_tn_kernel<double>, g, dim3(256), 0, st, (const double*)O, mean, r, ns, np,
                       1.0 / (double)ns, T);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_apply_update(void* stream, int dtype, void* theta, const double* step, double lr, int64_t np,
                     int32_t* applied) {
  QTX_REQUIRE(dtype == QTX_F32 || dtype == QTX_F64, "dtype must be QTX_F32 or QTX_F64");
  QTX_REQUIRE(np > 0 && theta && step && applied, "bad arguments");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, st, applied, (int64_t)1, 1);
  hipLaunchKernelGGL(finite_check_kernel<double>, dim3(nblk(np, 256)), dim3(256), 0, st, step, np, applied);
  if (dtype == QTX_F32)
    hipLaunchKernelGGL(apply_update_kernel<float>, dim3(nblk(np, 256)), dim3(256), 0, st, (float*)theta, step, lr,
                       np, applied);
  else
    hipLaunchKernelGGL(apply_update_kernel<double>, dim3(nblk(np, 256)), dim3(256), 0, st, (double*)theta, step,
                       lr, np, applied);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_minsr_blas_ws_bytes(void* stream, int64_t ns, int64_t np, int mode, int gram, void* ws, size_t ws_bytes,
                            size_t* bytes) {
  QTX_REQUIRE(ns > 0 && np > 0 && bytes, "bad arguments");
  const int bm = base_mode(mode);
  QTX_REQUIRE(bm == QTX_SOLVE_CHOL_SHIFT || bm == QTX_SOLVE_EIGH_PINV || bm == QTX_SOLVE_EIGH_PINV_DIRECT,
              "unknown solver mode");
  const size_t own = ws_layout(ns, np, mode, gram, nullptr, nullptr);
  if (ws == nullptr || ws_bytes < own)
    return fail(QTX_EWORKSPACE, "workspace too small: need " + std::to_string(own) + " bytes");
  SolveWs w{};
  ws_layout(ns, np, mode, gram, &w, (char*)ws);
  rocblas_handle h = nullptr;
  int rc = get_blas(as_stream(stream), &h, nullptr, 0, true);
  if (rc) return rc;
  rocblas_pointer_mode pm;
  rocblas_get_pointer_mode(h, &pm);
  rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
  rc = blas_query(h, ns, np, mode, w, bytes);
  rocblas_set_pointer_mode(h, pm);
  return rc;
}

int qtx_minsr_solve_snr(void* stream, int o_dtype, const void* O, const double* mean, const double* r,
                        const double* ebar, int64_t ns, int64_t np, int mode, int gram, double rtol, double atol,
                        double tol_snr, double* step, void* ws, size_t ws_bytes) {
  QTX_REQUIRE(base_mode(mode) == QTX_SOLVE_EIGH_PINV || base_mode(mode) == QTX_SOLVE_EIGH_PINV_DIRECT,
              "SNR truncation applies to the EIGH_PINV modes");
  QTX_REQUIRE(tol_snr >= 0.0, "tol_snr must be >= 0");
  qtx_minsr_task task{};
  task.stream = stream;
  task.o_dtype = o_dtype;
  task.mode = mode | QTX_SOLVE_SNR;
  task.gram = gram;
  task.O = O;
  task.mean = mean;
  task.r = r;
  task.ebar = ebar;
  task.ns = ns;
  task.np = np;
  task.rtol = rtol;
  task.atol = atol;
  task.tol_snr = tol_snr;
  task.step = step;
  task.ws = ws;
  task.ws_bytes = ws_bytes;
  g_last_path = -1;
  g_last_lanczos = 0;
  g_last_lmax = 0.0;
  g_last_fallback = 0;
  return solve_many(&task, 1);
}

// the MinSR-branch layout of an n x n problem: T, y, z, D, E, scalars and var2
static constexpr int kPinvEigMode = QTX_SOLVE_EIGH_PINV_DIRECT | QTX_SOLVE_MINNORM | QTX_SOLVE_SNR;

size_t qtx_pinv_eig_ws_bytes(int64_t n) {
  if (n <= 0) return 0;
  return ws_layout(n, n, kPinvEigMode, QTX_GRAM_FP64, nullptr, nullptr);
}

int qtx_pinv_eig(void* stream, const double* T, int64_t n, const double* b, double rtol, double atol, double tol_snr,
                 double* x, void* ws, size_t ws_bytes) {
  QTX_REQUIRE(n > 0 && T && b && x, "bad arguments");
  QTX_REQUIRE(n <= 131072, "matrix too large (n > 131072)");
  QTX_REQUIRE(rtol >= 0.0 && atol >= 0.0 && tol_snr >= 0.0, "tolerances must be >= 0");
  const size_t need = qtx_pinv_eig_ws_bytes(n);
  if (ws == nullptr || ws_bytes < need)
    return fail(QTX_EWORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  hipStream_t st = as_stream(stream);
  SolveWs w{};
  ws_layout(n, n, kPinvEigMode, QTX_GRAM_FP64, &w, (char*)ws);
  QTX_HIP(hipMemcpyAsync(w.T, T, (size_t)n * n * 8, hipMemcpyDeviceToDevice, st));
  QTX_HIP(hipMemcpyAsync(w.y, b, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  rocblas_handle h = nullptr;
  int rc = get_blas(st, &h);
  if (rc) return rc;
  rocblas_pointer_mode pm;
  rocblas_get_pointer_mode(h, &pm);
  rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
  rc = pinv_syevd(st, h, n, rtol, atol, w, tol_snr, nullptr);
  rocblas_set_pointer_mode(h, pm);
  if (rc < 0) return rc;
  QTX_HIP(hipMemcpyAsync(x, w.y, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  return QTX_OK;
}

}  // extern "C"
