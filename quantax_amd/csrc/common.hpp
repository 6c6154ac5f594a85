// Shared device helpers and error plumbing for libqtxhip (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/qtx_hip.h"

namespace qtx {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);

#define QTX_REQUIRE(cond, msg)                         \
  do {                                                 \
    if (!(cond)) return ::qtx::fail(QTX_EINVAL, msg); \
  } while (0)

#define QTX_HIP(call)                                   \
  do {                                                  \
    int _rc = ::qtx::hip_check((call), #call);          \
    if (_rc != QTX_OK) return _rc;                      \
  } while (0)

#define QTX_LAUNCH_CHECK() QTX_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int WAVE = 64;

// ---------------------------------------------------------------- reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
  return v;
}

// ---------------------------------------------------------------- math
__device__ __forceinline__ float logcosh(float x) {
  float ax = fabsf(x);
  return ax + log1pf(expf(-2.0f * ax)) - 0.69314718055994531f;
}
__device__ __forceinline__ double logcosh(double x) {
  double ax = fabs(x);
  return ax + log1p(exp(-2.0 * ax)) - 0.69314718055994530942;
}

// ---------------------------------------------------------------- XCD-aware block order
// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Remapping b to x * per + min(x, rem) + b / 8 (x = b % 8) gives each XCD a
// contiguous range of virtual blocks, so data that neighbouring virtual
// blocks share (cache lines of per-chain streams, row panels) stays in one
// XCD's L2.  A bijection of [0, nblocks).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblocks) {
  const int64_t per = nblocks / 8, rem = nblocks % 8, x = b % 8, loc = b / 8;
  return x * per + (x < rem ? x : rem) + loc;
}

// ---------------------------------------------------------------- Philox4x32-10
struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

// multiply-shift map of a 32-bit word to [0, n)
__host__ __device__ __forceinline__ uint32_t bounded(uint32_t w, uint32_t n) {
  return (uint32_t)(((uint64_t)w * (uint64_t)n) >> 32);
}

// u = ((w2 << 32 | w3) >> 11) * 2^-53 in [0, 1)
__host__ __device__ __forceinline__ double words_to_uniform(uint32_t w2, uint32_t w3) {
  uint64_t v = ((uint64_t)w2 << 32) | (uint64_t)w3;
  return (double)(v >> 11) * 1.1102230246251565404e-16;
}

constexpr uint32_t TAG_SWEEP = 0x51A7u;
constexpr uint32_t TAG_INIT = 0x1D17u;

// ---------------------------------------------------------------- proposals
// Picks the connected move of one chain.  Spins of the chain live in `sl`
// (LDS, N bytes).  Called by all 64 lanes of a wave with uniform arguments;
// returns the flipped sites (fa, fb; fb = -1 if one site) with `updated`
// false for no-op moves (swap of equal spins / padded neighbour / no hopping
// site).  n_hop is the number of sites holding `hopping` (conserved).
__device__ __forceinline__ void propose_move(int kind, const int8_t* sl, int N, const int32_t* nbr,
                                             int width, int hopping, int n_hop, uint32_t w0, uint32_t w1,
                                             int lane, int& fa, int& fb, bool& updated) {
  if (kind == QTX_LOCALFLIP) {
    fa = (int)bounded(w0, (uint32_t)N);
    fb = -1;
    updated = true;
    return;
  }
  fa = -1;
  fb = -1;
  updated = false;
  if (n_hop <= 0) return;
  uint32_t k = bounded(w0, (uint32_t)n_hop);
  int site = -1;
  for (int base = 0; base < N; base += WAVE) {
    int i = base + lane;
    bool is_hop = (i < N) && (sl[i] == hopping);
    uint64_t m = __ballot(is_hop);
    uint32_t cnt = (uint32_t)__popcll(m);
    if (k < cnt) {
      uint64_t below = (lane == 0) ? 0ull : (m & ((~0ull) >> (64 - lane)));
      bool hit = is_hop && ((uint32_t)__popcll(below) == k);
      uint64_t hm = __ballot(hit);
      site = base + (__ffsll((unsigned long long)hm) - 1);
      break;
    }
    k -= cnt;
  }
  if (site < 0) return;
  uint32_t slot = bounded(w1, (uint32_t)width);
  int nb = nbr[site * width + (int)slot];
  if (nb < 0) nb = site;
  int8_t a = sl[site], b = sl[nb];
  if (nb == site || a == b) return;
  fa = site;
  fb = nb;
  updated = true;
}

// ---------------------------------------------------------------- operator tables
__device__ __forceinline__ double diag_energy(const qtx_ham& H, const int8_t* sl, int lane) {
  double acc = 0.0;
  for (int t = lane; t < H.ndiag; t += WAVE) {
    int kind = H.diag_kind[t];
    double c = H.diag_c[t];
    if (kind == 0)
      acc += c * (double)(sl[H.diag_a[t]] * sl[H.diag_b[t]]);
    else if (kind == 1)
      acc += c * (double)sl[H.diag_a[t]];
    else
      acc += c;
  }
  return wave_sum(acc);
}

__device__ __forceinline__ bool cand_valid(const qtx_ham& H, const int8_t* sl, int k) {
  int a = H.off_a[k], b = H.off_b[k];
  int ra = H.off_ra[k], rb = H.off_rb[k];
  bool ok = (ra == 0) || (sl[a] == ra);
  if (b >= 0) ok = ok && ((rb == 0) || (sl[b] == rb));
  return ok;
}

}  // namespace qtx
