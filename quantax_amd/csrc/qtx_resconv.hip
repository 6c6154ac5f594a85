// ResConv(nblocks, C, 3) on a periodic L0 x L1 grid: batched forward on MFMA
// (FP32 16x16x4, or split-bf16 16x16x32 with FP32 accumulation), Metropolis
// with one full forward per proposal, local energy over connected
// configurations generated in LDS, and a per-sample backward for O whose
// three convolution products per layer are FP32 MFMA GEMMs.
//
// Forward work unit ("item") = one 16-position tile of one configuration.
// Convolution as an implicit GEMM per item: D[co, p] = sum_k W[co, k] X[k, p]
// with k = (tap, ci); A = weights (rows = output channels, register
// resident per layer), B = activations read from LDS at the tap-shifted
// position.  The residual stream stays in registers in the MFMA output
// layout (lane l: channel 16 ct + 4 (l / 16) + q, position 16 t + l % 16).
//
// LDS activations: per configuration, per position, C/4 chunks of 16 bytes,
// XOR-swizzled so 16 consecutive positions hit distinct banks.
//   FP32   : chunk j = channels 4j .. 4j+3 (float)
//   BF16X3 : chunk j < C/8 = hi(channels 8j .. 8j+7), chunk C/8 + j = lo(...)
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>

#include "common.hpp"

namespace qtx {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int RC_WAVES = 4;
constexpr int RC_THREADS = RC_WAVES * WAVE;
constexpr int RC_MAXN = 1024;

// ---------------------------------------------------------------- layout
struct RcLayout {
  int nb, C, L0, L1, N, mode, actk;
  int win;                   // 16 x 16, C = 16: sliding-row-window forward and its K-step order
  int64_t nparams;
  int64_t off_l0w, off_l0b;  // first conv (Cin = 1): W (C, 9), b (C) in floats
  int64_t off_frag;          // start of per-layer fragment blocks (bytes)
  int64_t frag_bytes;        // bytes of one layer's fragments
  int64_t bias_off;          // start of per-layer biases (floats, (2nb) x C)
  int64_t total_bytes;
};

__host__ __device__ inline int64_t rc_nparams(int nb, int C) {
  int64_t n = 0;
  for (int i = 0; i < nb; ++i) {
    int cin = (i == 0) ? 1 : C;
    n += (int64_t)C * cin * 9 + C;                 // conv1 (+ bias)
    n += (int64_t)C * C * 9 + ((i == nb - 1) ? 0 : C);  // conv2 (+ bias except last)
  }
  return n;
}

// offsets of layer l's weight/bias in the ravel_pytree vector
__host__ __device__ inline void rc_theta_offsets(int nb, int C, int l, int64_t& w, int64_t& b) {
  int64_t off = 0;
  for (int ll = 0; ll < 2 * nb; ++ll) {
    int i = ll / 2;
    int cin = (ll == 0) ? 1 : C;
    bool hb = !((ll % 2 == 1) && (i == nb - 1));
    if (ll == l) {
      w = off;
      b = hb ? off + (int64_t)C * cin * 9 : -1;
      return;
    }
    off += (int64_t)C * cin * 9 + (hb ? C : 0);
  }
  w = b = -1;
}

template <int C, int MODE>
struct RcK {
  static constexpr int CT = C / 16;                                   // output-channel tiles
  static constexpr int KS = (MODE == QTX_CONV_FP32) ? (9 * C / 4) : (C == 16 ? 5 : 9);  // K steps
  static constexpr int CH = C / 4;                                    // 16-byte chunks per position
  static constexpr int FRAG_WORDS = (MODE == QTX_CONV_FP32) ? 1 : 8;  // 4-byte words per lane per (ct, ks)
};

inline RcLayout rc_layout(const qtx_resconv_desc* d) {
  RcLayout L;
  L.nb = d->nblocks;
  L.C = d->channels;
  L.L0 = d->L0;
  L.L1 = d->L1;
  L.N = d->L0 * d->L1;
  L.mode = d->precision;
  L.actk = d->activation;
  L.win = (L.C == 16 && L.L0 == 16 && L.L1 == 16) ? 1 : 0;
  L.nparams = rc_nparams(L.nb, L.C);
  int64_t off = L.nparams * 4;  // raw theta copy first
  off = (off + 255) & ~(int64_t)255;
  L.off_l0w = off / 4;
  off += (int64_t)L.C * 9 * 4;
  L.off_l0b = off / 4;
  off += (int64_t)L.C * 4;
  off = (off + 255) & ~(int64_t)255;
  L.bias_off = off / 4;
  off += (int64_t)2 * L.nb * L.C * 4;
  off = (off + 255) & ~(int64_t)255;
  L.off_frag = off;
  int ct = L.C / 16;
  int ks = (L.mode == QTX_CONV_FP32) ? 9 * L.C / 4 : (L.C == 16 ? (L.win ? 6 : 5) : 9);
  int words = (L.mode == QTX_CONV_FP32) ? 1 : 8;
  L.frag_bytes = (int64_t)ct * ks * WAVE * words * 4;
  off += L.frag_bytes * 2 * L.nb;  // slot per layer index (slot 0 unused)
  L.total_bytes = off;
  return L;
}

// ---------------------------------------------------------------- packing
// one thread per (layer, ct, ks, lane) fragment for layers 1 .. 2nb-1
template <int C, int MODE, int WIN>
__global__ void rc_pack_frag_kernel(const float* __restrict__ theta, int nb, char* __restrict__ packed,
                                    int64_t off_frag, int64_t frag_bytes) {
  using K0 = RcK<C, MODE>;
  constexpr int KCT = K0::CT;
  constexpr int KKS = (WIN && MODE == QTX_CONV_BF16X3) ? 6 : K0::KS;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t per_layer = (int64_t)KCT * KKS * WAVE;
  int64_t total = per_layer * (2 * nb - 1);
  if (idx >= total) return;
  int l = 1 + (int)(idx / per_layer);
  int64_t rem = idx % per_layer;
  int ct = (int)(rem / (KKS * WAVE));
  int ks = (int)((rem / WAVE) % KKS);
  int lane = (int)(rem % WAVE);
  int64_t w0, b0;
  rc_theta_offsets(nb, C, l, w0, b0);
  const int co = 16 * ct + lane % 16, g = lane / 16;
  // window layout, split-bf16: conv1 layers l >= 2 carry the scale of the
  // conv2 before them, which then outputs act(s x) / s
  const float wsc = (WIN && MODE == QTX_CONV_BF16X3 && l >= 2 && (l % 2) == 0) ? 1.0f / sqrtf((float)(l / 2 + 1)) : 1.0f;
  auto W = [&](int ci, int tap) -> float { return theta[w0 + ((int64_t)co * C + ci) * 9 + tap] * wsc; };
  char* dst = packed + off_frag + (int64_t)l * frag_bytes;
  if constexpr (MODE == QTX_CONV_FP32) {
    // ks = (tap, h, j'): ci = 16 h + 4 g + j'
    int tap = ks / (C / 4), rr = ks % (C / 4), h = rr / 4, jp = rr % 4;
    float v = W(16 * h + 4 * g + jp, tap);
    reinterpret_cast<float*>(dst)[((int64_t)ct * KKS + ks) * WAVE + lane] = v;
  } else {
    int tap, c0;
    // window layout: per K-step a plane of 64 x 16 B slot-0 fragments, then
    // one of slot-1 fragments (16-byte lane stride for the LDS reads);
    // other layouts: the two slots of a lane side by side
    __bf16* o = reinterpret_cast<__bf16*>(dst) + (((int64_t)ct * KKS + ks) * WAVE + lane) * 16;
    __bf16* o1 = o + 8;
    if (WIN) {
      o = reinterpret_cast<__bf16*>(dst) + ((((int64_t)ct * KKS + ks) * 2) * WAVE + lane) * 8;
      o1 = o + WAVE * 8;
    }
    if (WIN && (ks % 2) == 1) {
      // tap (dy, 2) holds 16 of a K-step's 32 rows, so its three split
      // products share instructions: slot 0 of K-step 2 dy + 1 is
      // [hi | lo] of tap (dy, 2) (against [X hi; X hi]); slot 1 carries the
      // hi-times-X-lo halves: [hi (0, 2) | hi (1, 2)] in K-step 1,
      // [hi (2, 2) | 0] in K-step 3, zeros in K-step 5
      const int dy = ks / 2, up = g >> 1;
      c0 = 8 * (g & 1);
      for (int q = 0; q < 8; ++q) {
        const float v = W(c0 + q, 3 * dy + 2);
        const __bf16 hi = (__bf16)v;
        o[q] = up ? (__bf16)(v - (float)hi) : hi;
        float v1 = 0.0f;
        if (dy == 0) v1 = W(c0 + q, up ? 5 : 2);
        else if (dy == 1 && !up) v1 = W(c0 + q, 8);
        o1[q] = (__bf16)v1;
      }
      return;
    }
    if (WIN) {
      // ks = 2 dy: taps (dy, 0), (dy, 1)
      tap = 3 * (ks / 2) + (g >> 1);
      c0 = 8 * (g & 1);
    } else if (C == 16) {
      tap = 2 * ks + (g >> 1);
      c0 = 8 * (g & 1);
    } else {
      tap = ks;
      c0 = 8 * g;
    }
    for (int q = 0; q < 8; ++q) {
      float v = (tap < 9) ? W(c0 + q, tap) : 0.0f;
      __bf16 hi = (__bf16)v;
      __bf16 lo = (__bf16)(v - (float)hi);
      o[q] = hi;
      o1[q] = lo;
    }
  }
}

__global__ void rc_pack_misc_kernel(const float* __restrict__ theta, int nb, int C, int64_t nparams,
                                    float* __restrict__ packed, int64_t off_l0w, int64_t off_l0b, int64_t bias_off) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < nparams) packed[idx] = theta[idx];
  int64_t j = idx - nparams;
  if (j >= 0 && j < (int64_t)C * 9) packed[off_l0w + j] = theta[j];  // layer 0 weights start at 0
  j -= (int64_t)C * 9;
  if (j >= 0 && j < C) packed[off_l0b + j] = theta[(int64_t)C * 9 + j];
  j -= C;
  if (j >= 0 && j < (int64_t)2 * nb * C) {
    int l = (int)(j / C), co = (int)(j % C);
    int64_t w0, b0;
    rc_theta_offsets(nb, C, l, w0, b0);
    packed[bias_off + j] = (b0 >= 0) ? theta[b0 + co] : 0.0f;
  }
}

// ---------------------------------------------------------------- math
// Activations (qtx_resconv_desc.activation): QTX_ACT_GELU_TANH, the
// reference's jax.nn.gelu default (model/conv_nets.py:85, 87),
// QTX_ACT_GELU_ERF, and QTX_ACT_RELU (max(x, 0), derivative 0 at x <= 0, as
// jax.nn.relu), an opt-in variant with no counterpart in the reference.
// tanh form: 0.5 x (1 + tanh(k (x + 0.044715 x^3))) = x / (1 + 2^(x (A + B x^2)))
// with A = -2 k log2(e), B = -2 k 0.044715 log2(e), k = sqrt(2 / pi); native exp2 and rcp
__device__ __forceinline__ float act_f(float x, int actk) {
  if (actk == QTX_ACT_RELU) return fmaxf(x, 0.0f);
  if (actk == QTX_ACT_GELU_ERF) return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
  constexpr float A = -2.0f * 0.79788456080286536f * 1.4426950408889634f;
  constexpr float Bc = A * 0.044715f;
  float e = __builtin_amdgcn_exp2f(x * fmaf(x * x, Bc, A));
  return x * __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ __forceinline__ float act_grad_f(float x, int actk) {
  if (actk == QTX_ACT_RELU) return x > 0.0f ? 1.0f : 0.0f;
  if (actk == QTX_ACT_GELU_ERF)
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * __expf(-0.5f * x * x) * 0.39894228040143268f;
  float u = 0.79788456080286536f * (x + 0.044715f * x * x * x);
  float t = 1.0f - 2.0f / (1.0f + __expf(2.0f * u));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * 0.79788456080286536f * (1.0f + 0.134145f * x * x);
}

// v = hi + lo + O(2^-16 |v|): hi = bf16(v) (round to nearest even), lo =
// bf16(v - hi).  v - hi is formed by v_dot2c_f32_bf16 from the packed hi
// pair, (hi0, hi1) . (-1, 0) + v0 and (hi0, hi1) . (0, -1) + v1, exact
// (bf16 products are exact in FP32 and v - hi is representable), so hi is
// never widened back to FP32 by bit moves.
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split_bf16x4(v4f v, v4bf& hi, v4bf& lo) {
  const v2bf h01 = {(__bf16)v[0], (__bf16)v[1]}, h23 = {(__bf16)v[2], (__bf16)v[3]};
  // the selectors come from registers: an inline constant here (the
  // compiler encodes e0 as "-1.0") is not read as (bf16 -1, 0) by the hardware
  uint32_t u0 = 0x0000bf80u, u1 = 0xbf800000u;
  asm("" : "+v"(u0), "+v"(u1));
  const v2bf e0 = __builtin_bit_cast(v2bf, u0), e1 = __builtin_bit_cast(v2bf, u1);
  const float l0 = __builtin_amdgcn_fdot2_f32_bf16(h01, e0, v[0], false);
  const float l1 = __builtin_amdgcn_fdot2_f32_bf16(h01, e1, v[1], false);
  const float l2 = __builtin_amdgcn_fdot2_f32_bf16(h23, e0, v[2], false);
  const float l3 = __builtin_amdgcn_fdot2_f32_bf16(h23, e1, v[3], false);
  const v2bf l01 = {(__bf16)l0, (__bf16)l1}, l23 = {(__bf16)l2, (__bf16)l3};
  hi = v4bf{h01[0], h01[1], h23[0], h23[1]};
  lo = v4bf{l01[0], l01[1], l23[0], l23[1]};
}

// ---------------------------------------------------------------- forward core
template <int C>
__device__ __forceinline__ int swz_chunk(int p, int j) {
  // a ds_read_b128 is served eight lanes (128 bytes, 32 banks) at a time and
  // the eight lanes read eight consecutive sites: the key makes their chunks
  // fall on distinct 16-byte bank groups
  constexpr int CH = C / 4;
  return j ^ ((p / (8 / CH)) & (CH - 1));
}

struct RcGeom {
  int L0, L1, N, NT;
  double log_cn;  // log(C N): the head's normalisation
};

__device__ __forceinline__ int tap_pos(const RcGeom& g, int y, int x, int tap) {
  int dy = tap / 3 - 1, dx = tap % 3 - 1;
  int ys = y + dy, xs = x + dx;
  ys = ys < 0 ? ys + g.L0 : (ys >= g.L0 ? ys - g.L0 : ys);
  xs = xs < 0 ? xs + g.L1 : (xs >= g.L1 ? xs - g.L1 : xs);
  return ys * g.L1 + xs;
}

// Write 4 consecutive channels (c4 .. c4+3, c4 % 4 == 0) of position p.
template <int C, int MODE>
__device__ __forceinline__ void act_store4(char* act_cfg, int p, int c4, v4f v) {
  constexpr int CH = C / 4;
  if constexpr (MODE == QTX_CONV_FP32) {
    int j = c4 / 4;
    *reinterpret_cast<v4f*>(act_cfg + ((int64_t)p * CH + swz_chunk<C>(p, j)) * 16) = v;
  } else {
    v4bf hi, lo;
    split_bf16x4(v, hi, lo);
    int j = c4 / 8, half = (c4 % 8) / 4;
    char* base = act_cfg + (int64_t)p * CH * 16;
    *reinterpret_cast<v4bf*>(base + swz_chunk<C>(p, j) * 16 + half * 8) = hi;
    *reinterpret_cast<v4bf*>(base + swz_chunk<C>(p, C / 8 + j) * 16 + half * 8) = lo;
  }
}

// Accumulate one conv layer (C -> C) for one item into acc[CT].
template <int C, int MODE>
__device__ __forceinline__ void conv_item(const char* act_cfg, const RcGeom& g, int p, int y, int x, int lane,
                                          const float* wf, const v8bf* wh, const v8bf* wl, v4f* acc) {
  using K = RcK<C, MODE>;
  constexpr int CH = K::CH;
  const int gq = lane / 16;
  if constexpr (MODE == QTX_CONV_FP32) {
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      int ps = tap_pos(g, y, x, tap);
      const char* base = act_cfg + (int64_t)ps * CH * 16;
#pragma unroll
      for (int h = 0; h < C / 16; ++h) {
        int j = 4 * h + gq;
        v4f bv = *reinterpret_cast<const v4f*>(base + swz_chunk<C>(ps, j) * 16);
#pragma unroll
        for (int jp = 0; jp < 4; ++jp) {
          int ks = tap * (C / 4) + 4 * h + jp;
#pragma unroll
          for (int ct = 0; ct < K::CT; ++ct)
            acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(wf[ct * K::KS + ks], bv[jp], acc[ct], 0, 0, 0);
        }
      }
    }
  } else {
#pragma unroll
    for (int ks = 0; ks < K::KS; ++ks) {
      int tap, c0;
      if (C == 16) {
        tap = 2 * ks + (gq >> 1);
        c0 = 8 * (gq & 1);
      } else {
        tap = ks;
        c0 = 8 * gq;
      }
      int ps = (tap < 9) ? tap_pos(g, y, x, tap) : p;
      const char* base = act_cfg + (int64_t)ps * CH * 16;
      int jh = c0 / 8;
      v8bf bh = *reinterpret_cast<const v8bf*>(base + swz_chunk<C>(ps, jh) * 16);
      v8bf bl = *reinterpret_cast<const v8bf*>(base + swz_chunk<C>(ps, C / 8 + jh) * 16);
#pragma unroll
      for (int ct = 0; ct < K::CT; ++ct) {
        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[ct * K::KS + ks], bh, acc[ct], 0, 0, 0);
        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[ct * K::KS + ks], bl, acc[ct], 0, 0, 0);
        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[ct * K::KS + ks], bh, acc[ct], 0, 0, 0);
      }
    }
  }
}

// Forward of B configurations whose spins are in sp[b * N + i] (LDS int8).
// Result: lp_out[b] (LDS double).  All threads of the block must call it.
// act: two LDS activation buffers of B * N * C * 4 bytes each (ping-pong: a
// layer reads one and writes the other, so one barrier per layer suffices).
// red: >= 2 * B * W floats.  Items of wave w: w + W * k, k < MAXT.
template <int C, int MODE, int B, int MAXT, int W>
__device__ __attribute__((always_inline)) inline void rc_forward(const RcGeom& g, int nb, int actk, const char* __restrict__ packed,
                           int64_t off_l0w, int64_t off_l0b, int64_t bias_off, int64_t off_frag, int64_t frag_bytes,
                           const int8_t* sp, char* act, float* red, double* lp_out) {
  using K = RcK<C, MODE>;
  constexpr int CT = K::CT;
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));  // per-lane addresses formed inside the forward
  const int lane = tid % WAVE, wv = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int gq = lane / 16;
  const int nitems = B * g.NT;
  const int64_t act_cfg_bytes = (int64_t)g.N * C * 4;
  const int64_t buf_bytes = act_cfg_bytes * B;
  const float* pf = reinterpret_cast<const float*>(packed);

  v4f res[MAXT][CT];

  // ---- layer 0: conv (Cin = 1) of s / sqrt(2) on VALU, GELU, into buffer 0
  {
    const float* W0 = pf + off_l0w;
    const float* b0 = pf + off_l0b;
#pragma unroll
    for (int k = 0; k < MAXT; ++k) {
      int it = wv + W * k;
      if (it < nitems) {
        int b = it / g.NT, t = it % g.NT;
        int p = 16 * t + lane % 16;
        bool pv = p < g.N;
        int pc = pv ? p : 0;
        int y = pc / g.L1, x = pc % g.L1;
        const int8_t* s = sp + b * g.N;
        float sv[9];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) sv[tap] = (float)s[tap_pos(g, y, x, tap)] * 0.70710678118654752f;
        float spin = (float)s[pc];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          v4f z;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            int co = 16 * ct + 4 * gq + q;
            float a = b0[co];
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) a += W0[co * 9 + tap] * sv[tap];
            z[q] = act_f(a, actk);
          }
          res[k][ct] = v4f{spin, spin, spin, spin};
          if (pv) act_store4<C, MODE>(act + b * act_cfg_bytes, p, 16 * ct + 4 * gq, z);
        }
      }
    }
  }
  __syncthreads();

  // ---- layers 1 .. 2nb-1 on MFMA
  int cur = 0;
  for (int l = 1; l < 2 * nb; ++l) {
    float wf[(MODE == QTX_CONV_FP32) ? CT * K::KS : 1];
    v8bf wh[(MODE == QTX_CONV_FP32) ? 1 : CT * K::KS];
    v8bf wl[(MODE == QTX_CONV_FP32) ? 1 : CT * K::KS];
    const char* fr = packed + off_frag + (int64_t)l * frag_bytes;
    if constexpr (MODE == QTX_CONV_FP32) {
#pragma unroll
      for (int i = 0; i < CT * K::KS; ++i) wf[i] = reinterpret_cast<const float*>(fr)[i * WAVE + lane];
    } else {
#pragma unroll
      for (int i = 0; i < CT * K::KS; ++i) {
        const v8bf* src = reinterpret_cast<const v8bf*>(fr) + ((int64_t)i * WAVE + lane) * 2;
        wh[i] = src[0];
        wl[i] = src[1];
      }
    }
    const float* bias = pf + bias_off + (int64_t)l * C;
    const char* in = act + cur * buf_bytes;
    char* out = act + (cur ^ 1) * buf_bytes;
    const bool is_conv2 = (l % 2) == 1;
    const int blk = l / 2;
    const bool last = (l == 2 * nb - 1);
    const float scale = last ? rsqrtf((float)(nb + 1)) : (is_conv2 ? rsqrtf((float)(blk + 2)) : 1.0f);
    float m_b[B], s_b[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      m_b[b] = -INFINITY;
      s_b[b] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < MAXT; ++k) {
      int it = wv + W * k;
      if (it < nitems) {
        int b = it / g.NT, t = it % g.NT;
        int p = 16 * t + lane % 16;
        int pc = p < g.N ? p : 0;
        int y = pc / g.L1, x = pc % g.L1;
        v4f acc[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          int cb = 16 * ct + 4 * gq;
          acc[ct] = v4f{bias[cb], bias[cb + 1], bias[cb + 2], bias[cb + 3]};
        }
        conv_item<C, MODE>(in + b * act_cfg_bytes, g, pc, y, x, lane, wf, wh, wl, acc);
        if (!last) {
#pragma unroll
          for (int ct = 0; ct < CT; ++ct) {
            v4f o;
            if (is_conv2) {
              v4f xr = acc[ct] + res[k][ct];
              res[k][ct] = xr;
#pragma unroll
              for (int q = 0; q < 4; ++q) o[q] = act_f(xr[q] * scale, actk);
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q) o[q] = act_f(acc[ct][q], actk);
            }
            if (p < g.N) act_store4<C, MODE>(out + b * act_cfg_bytes, p, 16 * ct + 4 * gq, o);
          }
        } else {
          // head: h = (z2 + r) / sqrt(nb + 1); running logsumexp per configuration
          float m = -INFINITY;
          float hv[CT * 4];
#pragma unroll
          for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              float h = (acc[ct][q] + res[k][ct][q]) * scale;
              hv[ct * 4 + q] = h;
              if (p < g.N) m = fmaxf(m, h);
            }
          m = wave_max(m);
          float sacc = 0.0f;
          if (p < g.N) {
#pragma unroll
            for (int i = 0; i < CT * 4; ++i) sacc += __expf(hv[i] - m);
          }
          sacc = wave_sum(sacc);
#pragma unroll
          for (int bb = 0; bb < B; ++bb) {
            if (bb == b) {
              float mm = fmaxf(m_b[bb], m);
              s_b[bb] = s_b[bb] * __expf(m_b[bb] - mm) + sacc * __expf(m - mm);
              m_b[bb] = mm;
            }
          }
        }
      }
    }
    if (!last) {
      __syncthreads();
      cur ^= 1;
    } else {
      if (lane == 0) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
          red[(b * W + wv) * 2] = m_b[b];
          red[(b * W + wv) * 2 + 1] = s_b[b];
        }
      }
      __syncthreads();
      if (threadIdx.x < B) {
        // waves' (max, sum) merged in FP32 (the sums are FP32 already): FP64
        // exp / log here would hold the workgroup at the barrier below
        int b = threadIdx.x;
        float m = -INFINITY;
        for (int w = 0; w < W; ++w) m = fmaxf(m, red[(b * W + w) * 2]);
        float ssum = 0.0f;
        for (int w = 0; w < W; ++w) {
          const float mw = red[(b * W + w) * 2];
          if (mw > -INFINITY) ssum += red[(b * W + w) * 2 + 1] * __expf(mw - m);
        }
        lp_out[b] = (double)m + (double)logf(ssum) - g.log_cn;
      }
      __syncthreads();
    }
  }
}

}  // namespace qtx

namespace qtx {

// Row-tile forward for C = 16 on L0 x 16 grids (L0 <= 16): an item is one
// lattice row of one configuration, so the 16 lanes of a fragment are the 16
// columns and every tap address is a wave-uniform row offset plus a per-lane
// column constant computed once.  Same LDS layout and results as rc_forward.
template <int MODE, int B, int MAXT, int W>
__device__ __attribute__((always_inline)) inline void rc_forward_rows(const RcGeom& g, int nb, int actk, const char* __restrict__ packed,
                                int64_t off_l0w, int64_t off_l0b, int64_t bias_off, int64_t off_frag,
                                int64_t frag_bytes, const int8_t* sp, char* act, float* red, double* lp_out) {
  constexpr int C = 16;
  using K = RcK<C, MODE>;
  constexpr int ROWB = 16 * 64;  // bytes of one lattice row (16 positions x 16 channels x 4 B)
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const int gq = lane / 16, x = lane % 16;
  const int L0 = g.L0;
  const int nitems = B * L0;
  const int cfg_bytes = L0 * ROWB;
  const int buf_bytes = cfg_bytes * B;
  const float* pf = reinterpret_cast<const float*>(packed);

  // per-lane column constants (swizzle key of column c is (c >> 1) & 3)
  int xoff[3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx) {
    int xs = (x + dx + 15) & 15;
    xoff[dx] = xs * 64 + ((gq ^ ((xs >> 1) & 3)) << 4);
  }
  int kdy[5], khi[5], klo[5];
#pragma unroll
  for (int ks = 0; ks < 5; ++ks) {
    int tap = 2 * ks + (gq >> 1);
    bool valid = tap < 9;
    int dy = valid ? tap / 3 : 1, dx = valid ? tap % 3 : 1;
    int xs = (x + dx + 15) & 15, key = (xs >> 1) & 3, jh = gq & 1;
    kdy[ks] = dy;
    khi[ks] = xs * 64 + ((jh ^ key) << 4);
    klo[ks] = xs * 64 + (((2 + jh) ^ key) << 4);
  }
  const int key0 = (x >> 1) & 3;
  const int st_f32 = x * 64 + ((gq ^ key0) << 4);
  const int st_hi = x * 64 + (((gq >> 1) ^ key0) << 4) + (gq & 1) * 8;
  const int st_lo = x * 64 + (((2 + (gq >> 1)) ^ key0) << 4) + (gq & 1) * 8;

  auto store4 = [&](char* row_base, v4f v) {
    if constexpr (MODE == QTX_CONV_FP32) {
      *reinterpret_cast<v4f*>(row_base + st_f32) = v;
    } else {
      v4bf hi, lo;
      split_bf16x4(v, hi, lo);
      *reinterpret_cast<v4bf*>(row_base + st_hi) = hi;
      *reinterpret_cast<v4bf*>(row_base + st_lo) = lo;
    }
  };
  auto wrap = [&](int r) { return r < 0 ? r + L0 : (r >= L0 ? r - L0 : r); };

  v4f res[MAXT];

  // ---- layer 0: conv (Cin = 1) of s / sqrt(2) on VALU, GELU, into buffer 0
  {
    const float* W0 = pf + off_l0w;
    const float* b0 = pf + off_l0b;
#pragma unroll
    for (int k = 0; k < MAXT; ++k) {
      int it = wv + W * k;
      if (it < nitems) {
        int b = it / L0, y = it % L0;
        const int8_t* s = sp + b * 16 * L0;
        float sv[9];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap)
          sv[tap] = (float)s[wrap(y + tap / 3 - 1) * 16 + ((x + tap % 3 + 15) & 15)] * 0.70710678118654752f;
        float spin = (float)s[y * 16 + x];
        v4f z;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int co = 4 * gq + q;
          float a = b0[co];
#pragma unroll
          for (int tap = 0; tap < 9; ++tap) a += W0[co * 9 + tap] * sv[tap];
          z[q] = act_f(a, actk);
        }
        res[k] = v4f{spin, spin, spin, spin};
        store4(act + b * cfg_bytes + y * ROWB, z);
      }
    }
  }
  __syncthreads();

  int cur = 0;
  for (int l = 1; l < 2 * nb; ++l) {
    float wf[(MODE == QTX_CONV_FP32) ? K::KS : 1];
    v8bf wh[(MODE == QTX_CONV_FP32) ? 1 : K::KS];
    v8bf wl[(MODE == QTX_CONV_FP32) ? 1 : K::KS];
    const char* fr = packed + off_frag + (int64_t)l * frag_bytes;
    if constexpr (MODE == QTX_CONV_FP32) {
#pragma unroll
      for (int i = 0; i < K::KS; ++i) wf[i] = reinterpret_cast<const float*>(fr)[i * WAVE + lane];
    } else {
#pragma unroll
      for (int i = 0; i < K::KS; ++i) {
        const v8bf* src = reinterpret_cast<const v8bf*>(fr) + ((int64_t)i * WAVE + lane) * 2;
        wh[i] = src[0];
        wl[i] = src[1];
      }
    }
    const float* bias = pf + bias_off + (int64_t)l * C;
    const v4f bias4 = v4f{bias[4 * gq], bias[4 * gq + 1], bias[4 * gq + 2], bias[4 * gq + 3]};
    const char* in = act + cur * buf_bytes;
    char* out = act + (cur ^ 1) * buf_bytes;
    const bool is_conv2 = (l % 2) == 1;
    const int blk = l / 2;
    const bool last = (l == 2 * nb - 1);
    const float scale = last ? rsqrtf((float)(nb + 1)) : (is_conv2 ? rsqrtf((float)(blk + 2)) : 1.0f);
    float m_b[B], s_b[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      m_b[b] = -INFINITY;
      s_b[b] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < MAXT; ++k) {
      int it = wv + W * k;
      if (it < nitems) {
        int b = it / L0, y = it % L0;
        const char* base = in + b * cfg_bytes;
        v4f acc = bias4;
        if constexpr (MODE == QTX_CONV_FP32) {
#pragma unroll
          for (int dy = 0; dy < 3; ++dy) {
            const char* rb = base + wrap(y + dy - 1) * ROWB;
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
              v4f bv = *reinterpret_cast<const v4f*>(rb + xoff[dx]);
              const int tap = 3 * dy + dx;
#pragma unroll
              for (int jp = 0; jp < 4; ++jp)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(wf[tap * 4 + jp], bv[jp], acc, 0, 0, 0);
            }
          }
        } else {
#pragma unroll
          for (int ks = 0; ks < 5; ++ks) {
            const char* rb = base + wrap(y + kdy[ks] - 1) * ROWB;
            v8bf bh = *reinterpret_cast<const v8bf*>(rb + khi[ks]);
            v8bf bl = *reinterpret_cast<const v8bf*>(rb + klo[ks]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[ks], bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[ks], bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[ks], bh, acc, 0, 0, 0);
          }
        }
        if (!last) {
          v4f o;
          if (is_conv2) {
            v4f xr = acc + res[k];
            res[k] = xr;
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = act_f(xr[q] * scale, actk);
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = act_f(acc[q], actk);
          }
          store4(out + b * cfg_bytes + y * ROWB, o);
          // keep one item's fragments live at a time (bounds VGPR use)
          __builtin_amdgcn_sched_barrier(0);
        } else {
          float hv[4], m = -INFINITY;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            hv[q] = (acc[q] + res[k][q]) * scale;
            m = fmaxf(m, hv[q]);
          }
          m = wave_max(m);
          float sacc = 0.0f;
#pragma unroll
          for (int q = 0; q < 4; ++q) sacc += __expf(hv[q] - m);
          sacc = wave_sum(sacc);
#pragma unroll
          for (int bb = 0; bb < B; ++bb) {
            if (bb == b) {
              float mm = fmaxf(m_b[bb], m);
              s_b[bb] = s_b[bb] * __expf(m_b[bb] - mm) + sacc * __expf(m - mm);
              m_b[bb] = mm;
            }
          }
        }
      }
    }
    if (!last) {
      __syncthreads();
      cur ^= 1;
    } else {
      if (lane == 0) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
          red[(b * W + wv) * 2] = m_b[b];
          red[(b * W + wv) * 2 + 1] = s_b[b];
        }
      }
      __syncthreads();
      if (threadIdx.x < B) {
        // waves' (max, sum) merged in FP32 (the sums are FP32 already): FP64
        // exp / log here would hold the workgroup at the barrier below
        int b = threadIdx.x;
        float m = -INFINITY;
        for (int w = 0; w < W; ++w) m = fmaxf(m, red[(b * W + w) * 2]);
        float ssum = 0.0f;
        for (int w = 0; w < W; ++w) {
          const float mw = red[(b * W + w) * 2];
          if (mw > -INFINITY) ssum += red[(b * W + w) * 2 + 1] * __expf(mw - m);
        }
        lp_out[b] = (double)m + (double)logf(ssum) - g.log_cn;
      }
      __syncthreads();
    }
  }
}

// Sliding-row-window forward for C = 16 on a 16 x 16 grid.  Workgroup:
// WPC B waves, B configurations; wave w owns the RPW = 16 / WPC rows
// RPW (w % WPC) .. of configuration w / WPC and walks them in order keeping
// the B fragments of input rows y - 1, y, y + 1 in registers: every fragment
// depends only on its input row (split-bf16 K-steps 2 dy + pr: pr 0 = taps
// (dy, 0) + (dy, 1), pr 1 = tap (dy, 2) in both K halves, whose split
// products over the three dy take five instructions per output row, see the
// packing; FP32 K-steps = taps), so each row is read from LDS once per wave
// and layer instead of three times.  The hi-times-X-lo halves of tap (dy, 2)
// pair rows y - 1 and y in one instruction: the lo fragment of row r is
// loaded with its upper K half (lanes 32-63) from row r + 1, so
// [X lo (y - 1); X lo (y)] is row y - 1's fragment as loaded, with no lane
// select per row.  With split-bf16, a conv1 layer's packed weights carry the
// scale s of the conv2 before it, whose epilogue then writes act(s x) / s
// (s folds into the activation's constants: one multiply per value less).
//
// Issue structure: the activation and the conv kind (conv1 / conv2 with
// residual) are compile-time inside a layer (one branch per layer), and the
// rows are software-pipelined: row i's MFMA chain and row i - 1's epilogue
// (residual, activation, bf16 split, LDS store) are one basic block, so VALU
// and MFMA work of the same wave can overlap.  Every wave reads layer l's
// weight fragments from one LDS stage (wst, RC_WSTAGE bytes) into registers
// at the layer start, together with its first two input rows and the row
// after its last (the latter two belong to the configuration's other waves);
// after a barrier, layer l + 1's fragments travel HBM/L2 -> LDS by DMA into
// the same stage while layer l runs, and each wave writes its output rows in
// place of its input rows (one activation buffer per configuration).  The
// residual stream x (FP32, 4 values per lane and row) stays in registers.
// Dispatched with WPC = 2, B = 2 (8 rows per wave, 46-52 KB of LDS and at
// most 256 VGPRs, two workgroups per CU): against WPC = 4, B = 1 (4 rows,
// ping-pong buffers, three workgroups per CU at 168 VGPRs) the per-layer
// prologue and pipeline drain are spread over twice the rows, and E_loc at
// 16 x 16 runs at 290 against 283 TFLOP/s (352 against 334 with ReLU).
constexpr int RC_WSTAGE = 12288;  // one layer's fragments: bf16x3 6 x 64 x 32 B; FP32 9216 B
constexpr int RC_WBIAS = 64;     // the layer's 16 FP32 biases, after the fragments
constexpr int rc_win_lds(int B) { return RC_WSTAGE + RC_WBIAS; }  // one weight stage
typedef __attribute__((address_space(3))) void rc_lds_void;

template <int ACT>
__device__ __forceinline__ float act_t(float x) {
  return act_f(x, ACT);
}

// act(s x) / s for s > 0, with the layer constants k0 = s A, k1 = s^3 B (tanh
// GELU) or k0 = s / sqrt(2) (erf GELU)
template <int ACT>
__device__ __forceinline__ float act_over_s(float x, float k0, float k1) {
  if constexpr (ACT == QTX_ACT_RELU) return fmaxf(x, 0.0f);
  else if constexpr (ACT == QTX_ACT_GELU_ERF) return 0.5f * x * (1.0f + erff(x * k0));
  else return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(x * fmaf(x * x, k1, k0)));
}

template <int MODE, int B, int WPC>
__device__ __attribute__((always_inline)) inline void rc_forward_win(
    const RcGeom& g, int nb, int actk, const char* __restrict__ packed, int64_t off_l0w, int64_t off_l0b,
    int64_t bias_off, int64_t off_frag, int64_t frag_bytes, const int8_t* sp, char* act, char* wst, float* red,
    double* lp_out) {
  constexpr int C = 16, W = WPC * B, L0 = 16, RPW = L0 / WPC;
  constexpr int ROWB = 16 * 64;
  constexpr int CFGB = L0 * ROWB;
  constexpr int KSW = (MODE == QTX_CONV_FP32) ? 36 : 6;
  // the lane index through an opaque move: every per-lane address below is
  // then computed inside the forward instead of being hoisted into
  // registers held (or spilled) across a caller's whole loop
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid % WAVE, wv = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int gq = lane / 16, x = lane % 16;
  const int b = wv / WPC, y0 = RPW * (wv % WPC);
  const float* pf = reinterpret_cast<const float*>(packed);

  // the 16-byte chunks of column c (64 bytes per column) are XORed with
  // (c >> 1) & 3: the eight lanes a ds_read_b128 serves per cycle read eight
  // consecutive columns, which then fall on distinct 16-byte bank groups
  // (with (c >> 2) & 3 columns c and c + 2 collide: the sweep is 6.6 % slower)
  int xoff[3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx) {
    int xs = (x + dx + 15) & 15;
    xoff[dx] = xs * 64 + ((gq ^ ((xs >> 1) & 3)) << 4);
  }
  int khi[2], klo[2];
#pragma unroll
  for (int pr = 0; pr < 2; ++pr) {
    int dx = (pr == 0) ? (gq >> 1) : 2;
    int xs = (x + dx + 15) & 15, key = (xs >> 1) & 3, jh = gq & 1;
    khi[pr] = xs * 64 + ((jh ^ key) << 4);
    klo[pr] = xs * 64 + (((2 + jh) ^ key) << 4);
  }
  const int key0 = (x >> 1) & 3;
  const int st_f32 = x * 64 + ((gq ^ key0) << 4);
  const int st_hi = x * 64 + (((gq >> 1) ^ key0) << 4) + (gq & 1) * 8;
  const int st_lo = x * 64 + (((2 + (gq >> 1)) ^ key0) << 4) + (gq & 1) * 8;
  auto store4 = [&](char* row_base, v4f v) {
    if constexpr (MODE == QTX_CONV_FP32) {
      *reinterpret_cast<v4f*>(row_base + st_f32) = v;
    } else {
      v4bf hi, lo;
      split_bf16x4(v, hi, lo);
      *reinterpret_cast<v4bf*>(row_base + st_hi) = hi;
      *reinterpret_cast<v4bf*>(row_base + st_lo) = lo;
    }
  };

  // B fragments of one input row
  struct Frag {
    v4f f[(MODE == QTX_CONV_FP32) ? 3 : 1];
    v8bf h[(MODE == QTX_CONV_FP32) ? 1 : 2];
    v8bf l[(MODE == QTX_CONV_FP32) ? 1 : 2];
  };
  const int klo1n = klo[1] + (lane >= 32 ? ROWB : 0);
  // nxt: the upper K half of F.l[1] comes from row r + 1 (no wrap), else row r
  auto load_row = [&](const char* cfg, int r, bool nxt) {
    Frag F;
    const char* rb = cfg + ((r + L0) & (L0 - 1)) * ROWB;
    if constexpr (MODE == QTX_CONV_FP32) {
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) F.f[dx] = *reinterpret_cast<const v4f*>(rb + xoff[dx]);
    } else {
#pragma unroll
      for (int pr = 0; pr < 2; ++pr) F.h[pr] = *reinterpret_cast<const v8bf*>(rb + khi[pr]);
      F.l[0] = *reinterpret_cast<const v8bf*>(rb + klo[0]);
      F.l[1] = *reinterpret_cast<const v8bf*>(rb + (nxt ? klo1n : klo[1]));
    }
    return F;
  };
  // layer l's fragment block -> the weight stage, frag_bytes / 1024 wave-wide DMA instructions
  auto stage_weights = [&](int l) {
    const char* src = packed + off_frag + (int64_t)l * frag_bytes;
    char* dst = wst;
    const int nchunk = (int)(frag_bytes >> 10);
    for (int c = wv; c < nchunk; c += W)
      __builtin_amdgcn_global_load_lds((const void*)(src + c * 1024 + lane * 16), (rc_lds_void*)(dst + c * 1024), 16,
                                       0, 0);
    // the biases travel with the fragments: read from global memory at the
    // layer start, the first MFMA would wait for the whole DMA as well
    if (wv == W - 1 && lane < C / 4)
      __builtin_amdgcn_global_load_lds((const void*)(packed + bias_off * 4 + (int64_t)l * C * 4 + lane * 16),
                                       (rc_lds_void*)(dst + RC_WSTAGE), 16, 0, 0);
  };

  v4f resr[RPW];  // residual stream of the wave's rows, MFMA output layout
  stage_weights(1);
  // ---- layer 0 on VALU
  {
    // this lane's 4 output channels 4 gq + q: weights from one base
    const float* W0 = pf + off_l0w + 4 * gq * 9;
    const float* b0 = pf + off_l0b + 4 * gq;
    // an opaque zero keeps the 9 x RPW spin addresses from being hoisted out
    // of the forward (they would be held, or spilled, for the whole kernel)
    int z0 = 0;
    asm volatile("" : "+s"(z0));
    const int8_t* s = sp + b * 256 + z0;
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int y = y0 + i;
      float sv[9];
#pragma unroll
      for (int tap = 0; tap < 9; ++tap)
        sv[tap] = (float)s[((y + tap / 3 - 1 + L0) & (L0 - 1)) * 16 + ((x + tap % 3 + 15) & 15)] *
                  0.70710678118654752f;
      float spin = (float)s[y * 16 + x];
      v4f z;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float a = b0[q];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) a += W0[q * 9 + tap] * sv[tap];
        z[q] = act_f(a, actk);
      }
      resr[i] = v4f{spin, spin, spin, spin};
      store4(act + b * CFGB + y * ROWB, z);
    }
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's weight DMA has landed
  __syncthreads();

  // one layer: weights from the LDS stage, rows software-pipelined; EPI is
  // called with (row, accumulator)
  auto run_layer = [&](int l, const char* in, auto&& epi) {
    // rows y0 - 1 and y0 + RPW belong to the configuration's other waves:
    // read before the barrier below, after which every wave overwrites its
    // own rows in place
    Frag Fm = load_row(in, y0 - 1, false);
    if constexpr (MODE != QTX_CONV_FP32) {
      // [X lo (y0 - 1); X lo (y0)]: row y0 - 1 may wrap
      const char* pm = (lane < 32) ? in + ((y0 - 1 + L0) & (L0 - 1)) * ROWB : in + y0 * ROWB;
      Fm.l[1] = *reinterpret_cast<const v8bf*>(pm + klo[1]);
    }
    Frag F0 = load_row(in, y0, true);
    Frag Fe = load_row(in, y0 + RPW, false);
    float wf[(MODE == QTX_CONV_FP32) ? KSW : 1];
    v8bf wh[(MODE == QTX_CONV_FP32) ? 1 : KSW];
    v8bf wl[(MODE == QTX_CONV_FP32) ? 1 : KSW];
    const char* fr = wst;
    if constexpr (MODE == QTX_CONV_FP32) {
#pragma unroll
      for (int i = 0; i < KSW; ++i) wf[i] = reinterpret_cast<const float*>(fr)[i * WAVE + lane];
    } else {
#pragma unroll
      for (int i = 0; i < KSW; ++i) {
        const v8bf* src = reinterpret_cast<const v8bf*>(fr) + 2 * i * WAVE + lane;
        wh[i] = src[0];
        wl[i] = src[WAVE];
      }
    }
    const v4f bias4 = *reinterpret_cast<const v4f*>(fr + RC_WSTAGE + 16 * gq);
    // every wave holds this layer's fragments: the next layer's travel into
    // the stage while this one runs
    __syncthreads();
    if (l + 1 < 2 * nb) stage_weights(l + 1);
    v4f accp = bias4;
#pragma unroll
    for (int i = 0; i <= RPW; ++i) {
      v4f acc = bias4;
      Frag Fp;
      if (i < RPW) {
        // the row's MFMA chain at raised wave priority, its epilogue at the
        // normal one: of the two waves on a SIMD, the one with MFMAs to issue
        // wins the arbiter, so one wave's epilogue runs beside the other's
        // chain (-5 % on the forward kernels)
        __builtin_amdgcn_s_setprio(1);
        Fp = (i + 1 < RPW) ? load_row(in, y0 + i + 1, i + 2 < RPW) : Fe;
        if constexpr (MODE == QTX_CONV_FP32) {
#pragma unroll
          for (int dy = 0; dy < 3; ++dy) {
            const Frag& R = (dy == 0) ? Fm : (dy == 1 ? F0 : Fp);
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
              for (int jp = 0; jp < 4; ++jp)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(wf[(3 * dy + dx) * 4 + jp], R.f[dx][jp], acc, 0, 0, 0);
          }
        } else {
          // taps (dy, 0), (dy, 1): three split products per K-step; tap
          // (dy, 2): [hi | lo] x [X hi; X hi] per dy, then the hi x X lo
          // halves two dy to an instruction (14 MFMAs per row instead of 18)
#pragma unroll
          for (int dy = 0; dy < 3; ++dy) {
            const Frag& R = (dy == 0) ? Fm : (dy == 1 ? F0 : Fp);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[2 * dy], R.h[0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[2 * dy], R.l[0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[2 * dy], R.h[0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[2 * dy + 1], R.h[1], acc, 0, 0, 0);
          }
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[1], Fm.l[1], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[3], Fp.l[1], acc, 0, 0, 0);
        }
      }
      __builtin_amdgcn_s_setprio(0);
      if (i > 0) epi(i - 1, accp);
      if (i < RPW) {
        accp = acc;
        Fm = F0;
        F0 = Fp;
      }
    }
  };

  char* const buf = act + b * CFGB;  // one activation buffer per configuration, updated in place
  auto layers = [&](auto act_c) {
    constexpr int ACT = decltype(act_c)::value;
    // layers 1 .. 2 nb - 2 in pairs: conv2 of block l / 2 (residual), conv1 of block l / 2 + 1
    for (int l = 1; l < 2 * nb - 1; l += 2) {
      const float scale = rsqrtf((float)(l / 2 + 2));
      constexpr float GA = -2.0f * 0.79788456080286536f * 1.4426950408889634f;
      const float k0 = (ACT == QTX_ACT_GELU_ERF) ? scale * 0.70710678118654752f : scale * GA;
      const float k1 = scale * scale * scale * (GA * 0.044715f);
      run_layer(l, buf, [&](int i, v4f acc) {
        v4f o;
        v4f xr = acc + resr[i];
        resr[i] = xr;
        if constexpr (MODE == QTX_CONV_FP32) {
#pragma unroll
          for (int q = 0; q < 4; ++q) o[q] = act_t<ACT>(xr[q] * scale);
        } else {
          // the next layer's weights carry the scale
#pragma unroll
          for (int q = 0; q < 4; ++q) o[q] = act_over_s<ACT>(xr[q], k0, k1);
        }
        store4(buf + (y0 + i) * ROWB, o);
      });
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      run_layer(l + 1, buf, [&](int i, v4f acc) {
        v4f o;
#pragma unroll
        for (int q = 0; q < 4; ++q) o[q] = act_t<ACT>(acc[q]);
        store4(buf + (y0 + i) * ROWB, o);
      });
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
    }
  };
  if (actk == QTX_ACT_GELU_ERF) layers(std::integral_constant<int, QTX_ACT_GELU_ERF>{});
  else if (actk == QTX_ACT_RELU) layers(std::integral_constant<int, QTX_ACT_RELU>{});
  else layers(std::integral_constant<int, QTX_ACT_GELU_TANH>{});
  // ---- last layer (conv2 of the last block, no bias, no GELU) and the head
  {
    const int l = 2 * nb - 1;
    const char* in = buf;
    const float scale = rsqrtf((float)(nb + 1));
    // log-sum-exp of the head: each lane keeps a running (max, sum) over its
    // own values of the wave's rows; one wave reduction at the end
    float m_l = -INFINITY, s_l = 0.0f;
    run_layer(l, in, [&](int i, v4f acc) {
      float hv[4];
      const v4f r = resr[i];
#pragma unroll
      for (int q = 0; q < 4; ++q) hv[q] = (acc[q] + r[q]) * scale;
      const float mn = fmaxf(m_l, fmaxf(fmaxf(hv[0], hv[1]), fmaxf(hv[2], hv[3])));
      float sacc = s_l * __expf(m_l - mn);
#pragma unroll
      for (int q = 0; q < 4; ++q) sacc += __expf(hv[q] - mn);
      s_l = sacc;
      m_l = mn;
    });
    const float m_w = wave_max(m_l);
    const float s_w = wave_sum(s_l * __expf(m_l - m_w));
    if (lane == 0) {
      red[wv * 2] = m_w;
      red[wv * 2 + 1] = s_w;
    }
    __syncthreads();
    if (threadIdx.x < B) {
      // merge the four waves' (max, sum) in FP32 (the sums are FP32 already);
      // FP64 exp / log here would hold every wave at the barrier below
      constexpr double kLogCN = 8.317766166719343;  // log(16 channels * 256 sites)
      const int bb = threadIdx.x;
      float m = -INFINITY;
      for (int w = WPC * bb; w < WPC * bb + WPC; ++w) m = fmaxf(m, red[w * 2]);
      float ssum = 0.0f;
      for (int w = WPC * bb; w < WPC * bb + WPC; ++w) ssum += red[w * 2 + 1] * __expf(red[w * 2] - m);
      lp_out[bb] = (double)m + (double)logf(ssum) - kLogCN;
    }
    __syncthreads();
  }
}

// Split-bf16 forward for C = 32 on grids of up to 400 sites (the 20 x 20
// benchmark lattice): one configuration per workgroup of W = 4 waves, two
// workgroups per CU (2 waves per SIMD, at most 256 VGPRs each).  Items are
// 16-site tiles in row-major order; wave w owns tiles w, w + 4, ... (at most
// MAXT = 7 of the 25 of 20 x 20) and computes both 16-channel output tiles
// (CTW = 2) of each, so one B fragment read from LDS feeds six MFMAs.  Per
// layer the K loop runs over the 9 taps outside the tile loop: the A
// fragments of one tap (2 output tiles x hi/lo, 16 VGPRs, prefetched one tap
// ahead from L2) serve every tile of the wave, and the wave's 2 x MAXT
// accumulators are independent MFMA chains.  Activations sit once in LDS
// (bf16 hi/lo, 128 bytes per site, 16-byte chunks XORed with site & 7); a
// layer computes into registers and its outputs replace its inputs between
// two barriers, which halves the LDS of a ping-pong pair so two workgroups
// fit on a CU.  The epilogue runs after the first barrier, one tile at a
// time (GELU, split, store), so the GELU outputs never coexist with the
// accumulators; the next layer's first A fragments are fetched after it.
// That keeps every C = 32 kernel within 256 VGPRs without scratch (GELU
// before the barrier, with all outputs held, spilled 200-370 bytes per lane
// and ran the 20 x 20 sweep at 300 against 378 TFLOP/s).  The B fragment of
// (tap, tile) is found through a table of byte offsets of the tap's source
// site (swizzle key in bits 4-6), built once per launch.  The residual stream
// stays in registers in the MFMA output layout.  CTW = 1 (six waves, wave pairs sharing tiles, three waves per
// SIMD) ran the 20 x 20 sweep at 186-221 TFLOP/s against 301 for CTW = 2:
// it reads every B fragment twice and issues three MFMAs per LDS round trip.
constexpr int RC32_MAXN = 400;
constexpr int RC32_CTW = 2;  // output tiles per wave
constexpr int rc32_tab_bytes(int nmax) { return 9 * nmax * 2; }

template <int MAXT, int W, int CTW, int ACT>
__device__ __attribute__((always_inline)) inline void rc_forward_t32(
    const RcGeom& g, int nb, const char* __restrict__ packed, int64_t off_l0w, int64_t off_l0b,
    int64_t bias_off, int64_t off_frag, int64_t frag_bytes, const int8_t* sp, char* act, const uint16_t* tab,
    float* red, double* lp_out) {
  // CTW output tiles per wave: 1 (wave pairs share tiles, ct = w & 1) or 2
  constexpr int C = 32, NP = (CTW == 1) ? W / 2 : W;  // NP groups of waves own disjoint tiles
  const int lane = threadIdx.x % WAVE;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int gq = lane / 16, x16 = lane % 16;
  const int ct0 = (CTW == 1) ? (wv & 1) : 0, m = (CTW == 1) ? (wv >> 1) : wv;
  const int N = g.N, NT = g.NT;
  const int nti = (NT - m + NP - 1) / NP;  // tiles of this wave (wave-uniform)
  const float* pf = reinterpret_cast<const float*>(packed);
  // the B fragment of lane group gq is chunk gq (hi) / 4 + gq (lo)
  const uint32_t xhi = (uint32_t)gq << 4, xlo = (uint32_t)(gq ^ 4) << 4;
  // output stores: channel 16 ct + 4 gq + i of site p lies in chunk
  // 2 ct + gq / 2 (hi) or 4 + 2 ct + gq / 2 (lo), half gq & 1; the key p & 7
  // equals lane & 7 for every tile (tiles start at multiples of 16)
  const int key = lane & 7;
  int st_hi[CTW], st_lo[CTW];
#pragma unroll
  for (int j = 0; j < CTW; ++j) {
    const int ct = ct0 + j;
    st_hi[j] = x16 * 128 + (((2 * ct + (gq >> 1)) ^ key) << 4) + (gq & 1) * 8;
    st_lo[j] = x16 * 128 + (((4 + 2 * ct + (gq >> 1)) ^ key) << 4) + (gq & 1) * 8;
  }
  auto site = [&](int k) { return 16 * (m + NP * k) + x16; };  // may be >= N
  auto store4 = [&](int k, int j, v4f v) {
    const int t = m + NP * k;
    if (16 * t + x16 < N) {
      v4bf hi, lo;
      split_bf16x4(v, hi, lo);
      char* base = act + t * 16 * 128;
      *reinterpret_cast<v4bf*>(base + st_hi[j]) = hi;
      *reinterpret_cast<v4bf*>(base + st_lo[j]) = lo;
    }
  };
  // A fragments (hi, lo) of (layer, tap) for this wave's output tiles
  auto load_a = [&](int l, int ks, v8bf (&a)[CTW][2]) {
#pragma unroll
    for (int j = 0; j < CTW; ++j) {
      const v8bf* q = reinterpret_cast<const v8bf*>(packed + off_frag + (int64_t)l * frag_bytes) +
                      (((ct0 + j) * 9 + ks) * WAVE + lane) * 2;
      a[j][0] = q[0];
      a[j][1] = q[1];
    }
  };

  v4f res[MAXT][CTW];

  // ---- layer 0: conv (Cin = 1) of s / sqrt(2) on VALU, GELU, into act
  {
    // this lane's 4 CTW output channels: weights co * 9 + tap from one base
    const float* W0 = pf + off_l0w + (16 * ct0 + 4 * gq) * 9;
    const float* b0 = pf + off_l0b + 16 * ct0 + 4 * gq;
#pragma unroll
    for (int k = 0; k < MAXT; ++k) {
      if (k < nti) {
        const int p = site(k);
        const int pc = p < N ? p : N - 1;
        float sv[9];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) sv[tap] = (float)sp[tab[tap * RC32_MAXN + pc] >> 7] * 0.70710678118654752f;
        const float spin = (float)sp[pc];
#pragma unroll
        for (int j = 0; j < CTW; ++j) {
          v4f z;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int co = 16 * j + q;
            float a = b0[co];
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) a += W0[co * 9 + tap] * sv[tap];
            z[q] = act_t<ACT>(a);
          }
          res[k][j] = v4f{spin, spin, spin, spin};
          store4(k, j, z);
        }
      }
    }
  }
  v8bf an[CTW][2];
  load_a(1, 0, an);
  __syncthreads();

  for (int l = 1; l < 2 * nb; ++l) {
    const float* bias = pf + bias_off + (int64_t)l * C;
    v4f acc[MAXT][CTW];
#pragma unroll
    for (int j = 0; j < CTW; ++j) {
      const int cb = 16 * (ct0 + j) + 4 * gq;
      const v4f b4 = v4f{bias[cb], bias[cb + 1], bias[cb + 2], bias[cb + 3]};
#pragma unroll
      for (int k = 0; k < MAXT; ++k) acc[k][j] = b4;
    }
    const bool last = (l == 2 * nb - 1);
    // the tap offsets are the same in every layer: an opaque per-layer zero
    // keeps the compiler from hoisting all 9 x MAXT of them out of the layer
    // loop into registers
    int tz = 0;
    asm volatile("" : "+s"(tz));
    // the layer's MFMA loop at raised wave priority, the epilogue at the
    // normal one, so that the other workgroup's epilogue runs beside it
    __builtin_amdgcn_s_setprio(1);
#pragma unroll 1
    for (int ks = 0; ks < 9; ++ks) {
      v8bf a[CTW][2];
#pragma unroll
      for (int j = 0; j < CTW; ++j) {
        a[j][0] = an[j][0];
        a[j][1] = an[j][1];
      }
      if (ks < 8) load_a(l, ks + 1, an);
      const uint16_t* tk = tab + ks * RC32_MAXN + tz;
      // tiles that exist for every wave run unconditionally (sites past N
      // read the table's padding, offset 0, and are never stored); only the
      // last is guarded
#pragma unroll
      for (int k = 0; k < MAXT; ++k) {
        if (k < MAXT - 1 || k < nti) {
          const uint32_t off = tk[site(k)];  // < RC32_MAXN for every tile read here
          const v8bf bh = *reinterpret_cast<const v8bf*>(act + (off ^ xhi));
          const v8bf bl = *reinterpret_cast<const v8bf*>(act + (off ^ xlo));
#pragma unroll
          for (int j = 0; j < CTW; ++j) {
            acc[k][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[j][1], bh, acc[k][j], 0, 0, 0);
            acc[k][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[j][0], bl, acc[k][j], 0, 0, 0);
            acc[k][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[j][0], bh, acc[k][j], 0, 0, 0);
          }
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
    if (!last) {
      const bool is_conv2 = (l % 2) == 1;
      const float scale = is_conv2 ? rsqrtf((float)(l / 2 + 2)) : 1.0f;
      __syncthreads();  // every wave has read this layer's input
#pragma unroll
      for (int k = 0; k < MAXT; ++k) {
        if (k < nti) {
#pragma unroll
          for (int j = 0; j < CTW; ++j) {
            v4f o;
            if (is_conv2) {
              const v4f xr = acc[k][j] + res[k][j];
              res[k][j] = xr;
#pragma unroll
              for (int q = 0; q < 4; ++q) o[q] = act_t<ACT>(xr[q] * scale);
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q) o[q] = act_t<ACT>(acc[k][j][q]);
            }
            store4(k, j, o);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      load_a(l + 1, 0, an);
      __syncthreads();
    } else {
      // head: h = (z2 + r) / sqrt(nb + 1); per-lane running logsumexp, one
      // wave reduction, then the waves' (max, sum) merged in FP32
      const float scale = rsqrtf((float)(nb + 1));
      float m_l = -INFINITY, s_l = 0.0f;
#pragma unroll
      for (int k = 0; k < MAXT; ++k) {
        if (k >= nti || site(k) >= N) continue;
#pragma unroll
        for (int j = 0; j < CTW; ++j) {
          float hv[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) hv[q] = (acc[k][j][q] + res[k][j][q]) * scale;
          const float mn = fmaxf(m_l, fmaxf(fmaxf(hv[0], hv[1]), fmaxf(hv[2], hv[3])));
          float sacc = s_l * __expf(m_l - mn);
#pragma unroll
          for (int q = 0; q < 4; ++q) sacc += __expf(hv[q] - mn);
          s_l = sacc;
          m_l = mn;
        }
      }
      const float m_w = wave_max(m_l);
      const float s_w = wave_sum(m_l > -INFINITY ? s_l * __expf(m_l - m_w) : 0.0f);
      if (lane == 0) {
        red[wv * 2] = m_w;
        red[wv * 2 + 1] = s_w;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        float mx = -INFINITY;
        for (int w = 0; w < W; ++w) mx = fmaxf(mx, red[w * 2]);
        float ssum = 0.0f;
        for (int w = 0; w < W; ++w)
          if (red[w * 2] > -INFINITY) ssum += red[w * 2 + 1] * __expf(red[w * 2] - mx);
        lp_out[0] = (double)mx + (double)logf(ssum) - g.log_cn;
      }
      __syncthreads();
    }
  }
}

// byte offsets (source site * 128 | (source site & 7) << 4) of the 9 taps of
// every site, tab[tap * RC32_MAXN + p], padded with offset 0 up to
// RC32_MAXN sites; all threads of the block call it
__device__ __forceinline__ void rc32_build_tab(const RcGeom& g, uint16_t* tab) {
  for (int e = threadIdx.x; e < 9 * RC32_MAXN; e += blockDim.x) {
    const int tap = e / RC32_MAXN, p = e - tap * RC32_MAXN;
    uint16_t v = 0;
    if (p < g.N) {
      const int src = tap_pos(g, p / g.L1, p % g.L1, tap);
      v = (uint16_t)((src << 7) | ((src & 7) << 4));
    }
    tab[e] = v;
  }
}

// ---------------------------------------------------------------- kernels
struct RcArgs {
  RcGeom g;
  int nb, actk;
  const char* packed;
  int64_t off_l0w, off_l0b, bias_off, off_frag, frag_bytes;
};

#define RC_FWD(SP, LP)                                                                                        \
  do {                                                                                                        \
    if constexpr (ROWS == 3) {                                                                                \
      const uint16_t* tab_ = reinterpret_cast<const uint16_t*>(wst);                                          \
      if (a.actk == QTX_ACT_GELU_ERF)                                                                         \
        rc_forward_t32<MAXT, W, RC32_CTW, QTX_ACT_GELU_ERF>(a.g, a.nb, a.packed, a.off_l0w, a.off_l0b,         \
                                                            a.bias_off, a.off_frag, a.frag_bytes, SP, act,     \
                                                            tab_, red, LP);                                    \
      else if (a.actk == QTX_ACT_RELU)                                                                        \
        rc_forward_t32<MAXT, W, RC32_CTW, QTX_ACT_RELU>(a.g, a.nb, a.packed, a.off_l0w, a.off_l0b, a.bias_off, \
                                                        a.off_frag, a.frag_bytes, SP, act, tab_, red, LP);     \
      else                                                                                                    \
        rc_forward_t32<MAXT, W, RC32_CTW, QTX_ACT_GELU_TANH>(a.g, a.nb, a.packed, a.off_l0w, a.off_l0b,        \
                                                             a.bias_off, a.off_frag, a.frag_bytes, SP, act,    \
                                                             tab_, red, LP);                                   \
    } else if constexpr (ROWS == 2)                                                                           \
      rc_forward_win<MODE, B, W / B>(a.g, a.nb, a.actk, a.packed, a.off_l0w, a.off_l0b, a.bias_off, a.off_frag, \
                           a.frag_bytes, SP, act, wst, red, LP);                                              \
    else if constexpr (ROWS == 1)                                                                             \
      rc_forward_rows<MODE, B, MAXT, W>(a.g, a.nb, a.actk, a.packed, a.off_l0w, a.off_l0b, a.bias_off,   \
                                        a.off_frag, a.frag_bytes, SP, act, red, LP);                          \
    else                                                                                                      \
      rc_forward<C, MODE, B, MAXT, W>(a.g, a.nb, a.actk, a.packed, a.off_l0w, a.off_l0b, a.bias_off,     \
                                      a.off_frag, a.frag_bytes, SP, act, red, LP);                            \
  } while (0)

template <int C, int MODE, int B, int NMAX, int MAXT, int W, int ROWS>
__global__ __launch_bounds__(W * 64, (ROWS == 2 ? (W == 4 * B ? 3 : 2) : (ROWS == 3 ? 2 : (W == 8 ? 2 : 1)))) void rc_logpsi_kernel(RcArgs a, const int8_t* __restrict__ s, int64_t ns,
                                                               double* __restrict__ logabs) {
  __shared__ __attribute__((aligned(16))) char act[(ROWS >= 2 ? 1 : 2) * B * NMAX * C * 4];
  __shared__ __attribute__((aligned(16))) char wst[ROWS == 2 ? rc_win_lds(B) : (ROWS == 3 ? rc32_tab_bytes(NMAX) : 16)];
  __shared__ int8_t sp[B * NMAX];
  __shared__ float red[2 * B * W];
  __shared__ double lp[B];
  const int N = a.g.N;
  const int64_t c0 = xcd_block(blockIdx.x, gridDim.x) * B;  // per-chain Philox words of a step share lines
  for (int i = threadIdx.x; i < B * N; i += (W * WAVE)) {
    int b = i / N;
    sp[i] = (c0 + b < ns) ? s[(c0 + b) * N + (i % N)] : (int8_t)1;
  }
  if constexpr (ROWS == 3) rc32_build_tab(a.g, reinterpret_cast<uint16_t*>(wst));
  __syncthreads();
  RC_FWD(sp, lp);
  if (threadIdx.x < B && c0 + threadIdx.x < ns) logabs[c0 + threadIdx.x] = lp[threadIdx.x];
}

// log(u) of the acceptance test (metropolis.py:291-322) in FP64, kept out of
// line: inlined, the constants of FP64 log1p would be hoisted into VGPRs for
// the whole kernel
__device__ __noinline__ double accept_threshold(uint32_t w2, uint32_t w3) {
  return log1p(-words_to_uniform(w2, w3));
}

template <int C, int MODE, int B, int NMAX, int MAXT, int W, int ROWS>
__global__ __launch_bounds__(W * 64, (ROWS == 2 ? (W == 4 * B ? 3 : 2) : (ROWS == 3 ? 2 : (W == 8 ? 2 : 1)))) void rc_metropolis_kernel(
    RcArgs a, int kind, const int32_t* __restrict__ nbr, int width, int hopping, const uint32_t* __restrict__ words,
    int nsteps, int64_t ns, double reweight, double tie_tol, int8_t* __restrict__ s, double* __restrict__ logabs,
    int32_t* __restrict__ n_accept, int32_t* __restrict__ n_near) {
  __shared__ __attribute__((aligned(16))) char act[(ROWS >= 2 ? 1 : 2) * B * NMAX * C * 4];
  __shared__ __attribute__((aligned(16))) char wst[ROWS == 2 ? rc_win_lds(B) : (ROWS == 3 ? rc32_tab_bytes(NMAX) : 16)];
  __shared__ int8_t cur[B * NMAX];
  __shared__ int8_t prop[B * NMAX];
  __shared__ float red[2 * B * W];
  __shared__ double lp_new[B];
  __shared__ double lp_cur[B];
  __shared__ int mv_a[B], mv_b[B], mv_upd[B], nhop[B], nacc[B], nnear[B];
  __shared__ double mv_thr[B];
  const int N = a.g.N;
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const int64_t c0 = xcd_block(blockIdx.x, gridDim.x) * B;  // per-chain Philox words of a step share lines
  for (int i = threadIdx.x; i < B * N; i += (W * WAVE)) {
    int b = i / N;
    cur[i] = (c0 + b < ns) ? s[(c0 + b) * N + (i % N)] : (int8_t)1;
  }
  if (threadIdx.x < B) {
    int b = threadIdx.x;
    lp_cur[b] = (c0 + b < ns) ? logabs[c0 + b] : 0.0;
    nacc[b] = 0;
    nnear[b] = 0;
  }
  if constexpr (ROWS == 3) rc32_build_tab(a.g, reinterpret_cast<uint16_t*>(wst));
  __syncthreads();
  for (int b = wv; b < B; b += W) {
    int cnt = 0;
    for (int i = lane; i < N; i += WAVE) cnt += (cur[b * N + i] == hopping);
    cnt = wave_sum(cnt);
    if (lane == 0) nhop[b] = cnt;
  }
  __syncthreads();
  // with at least two waves per chain, the FP64 acceptance threshold of
  // chain b is computed by wave B + b while wave b draws the move
  constexpr bool kThrWave = (2 * B <= W);
  static_assert(W >= B, "one wave per chain at most");
  // the step's Philox words are fetched one step ahead (their HBM latency
  // hidden by the forward): waves 0 .. B - 1 draw chain wv's move, waves
  // B .. 2 B - 1 its acceptance threshold
  const int cbw = (wv < B) ? wv : ((kThrWave && wv < 2 * B) ? wv - B : -1);
  const int64_t cw = c0 + cbw;
  const bool wok = cbw >= 0 && cw < ns;
  uint4 wn = make_uint4(0u, 0u, 0u, 0u);
  if (wok && nsteps > 0) wn = reinterpret_cast<const uint4*>(words)[cw];
  for (int t = 0; t < nsteps; ++t) {
    const uint4 wcur = wn;
    if (wok && t + 1 < nsteps) wn = reinterpret_cast<const uint4*>(words)[(int64_t)(t + 1) * ns + cw];
    // the lane index through an opaque move, as in the forwards: the step's
    // per-lane LDS addresses are formed here, not held across the loop
    int lane = threadIdx.x % WAVE;
    asm volatile("" : "+v"(lane));
    // proposals: one wave per chain
    for (int b = wv; b < B; b += W) {
      int64_t c = c0 + b;
      int fa = -1, fb = -1;
      bool upd = false;
      double thr = 0.0;
      if (c < ns) {
        const uint4 w = wcur;
        propose_move(kind, cur + b * N, N, nbr, width, hopping, nhop[b], w.x, w.y, lane, fa, fb, upd);
        if constexpr (!kThrWave) thr = accept_threshold(w.z, w.w);
      }
      for (int i = lane; i < N; i += WAVE) {
        int8_t v = cur[b * N + i];
        if (upd && (i == fa || i == fb)) v = -v;
        prop[b * N + i] = v;
      }
      if (lane == 0) {
        mv_a[b] = fa;
        mv_b[b] = fb;
        mv_upd[b] = upd ? 1 : 0;
        if constexpr (!kThrWave) mv_thr[b] = thr;
      }
    }
    if constexpr (kThrWave) {
      if (wv >= B && wv < 2 * B) {
        const int b = wv - B;
        const int64_t c = c0 + b;
        double thr = 0.0;
        if (c < ns) thr = accept_threshold(wcur.z, wcur.w);
        if (lane == 0) mv_thr[b] = thr;
      }
    }
    __syncthreads();
    RC_FWD(prop, lp_new);
    // accept / reject (rc_forward ends with a barrier)
    for (int b = wv; b < B; b += W) {
      if (!mv_upd[b]) continue;
      double d = reweight * (lp_new[b] - lp_cur[b]);
      double thr = mv_thr[b];
      bool acc = (d > thr) || (lp_cur[b] == -INFINITY);
      if (acc) {
        int fa = mv_a[b], fb = mv_b[b];
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
          cur[b * N + fa] = -cur[b * N + fa];
          if (fb >= 0) cur[b * N + fb] = -cur[b * N + fb];
          lp_cur[b] = lp_new[b];
          nacc[b] += 1;
        }
      }
      if (lane == 0 && fabs(d - thr) < tie_tol) nnear[b] += 1;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < B * N; i += (W * WAVE)) {
    int b = i / N;
    if (c0 + b < ns) s[(c0 + b) * N + (i % N)] = cur[i];
  }
  if (threadIdx.x < B && c0 + threadIdx.x < ns) {
    int b = threadIdx.x;
    logabs[c0 + b] = lp_cur[b];
    n_accept[c0 + b] += nacc[b];
    n_near[c0 + b] += nnear[b];
  }
}

// psi(s') / psi(s) = e^d for the local energy: 2^t, t = d log2(e), split
// into an integer k applied exactly by ldexp and f = t - k in [-1/2, 1/2]
// whose 2^f is one FP32 exp2 (about 1 ulp); the FP64 product d log2(e) and
// the FP32 rounding of f keep the relative error near 1e-7 for any d, where
// an FP32 exp of d itself would lose |d| * 2^-24 to the rounding of d alone.
// |t| >= 2100 is exp's overflow (+inf) or underflow (0) outright, and NaN
// stays NaN: no FP64 exp, whose polynomial constants would otherwise occupy
// VGPRs for the whole kernel.
__device__ __forceinline__ double exp_ratio(double d) {
  const double t = d * 1.4426950408889634;
  if (!(fabs(t) < 2100.0)) return d > 0.0 ? HUGE_VAL : (d < 0.0 ? 0.0 : d);  // exp's overflow / underflow; NaN stays
  const double k = rint(t);
  return ldexp((double)__builtin_amdgcn_exp2f((float)(t - k)), (int)k);
}

constexpr int RC_MAXCAND = 4096;
constexpr int RC_MAXCAND_WIN = 3072;  // 16 x 16 window kernels: keeps two workgroups' LDS within a CU

template <int C, int MODE, int B, int NMAX, int MAXT, int W, int ROWS>
__global__ __launch_bounds__(W * 64, (ROWS == 2 ? (W == 4 * B ? 3 : 2) : (ROWS == 3 ? 2 : (W == 8 ? 2 : 1)))) void rc_eloc_kernel(RcArgs a, qtx_ham H, const int8_t* __restrict__ s,
                                                             const double* __restrict__ logabs, int64_t ns,
                                                             double* __restrict__ eloc, int32_t* __restrict__ nconn) {
  __shared__ __attribute__((aligned(16))) char act[(ROWS >= 2 ? 1 : 2) * B * NMAX * C * 4];
  __shared__ __attribute__((aligned(16))) char wst[ROWS == 2 ? rc_win_lds(B) : (ROWS == 3 ? rc32_tab_bytes(NMAX) : 16)];
  __shared__ int8_t s0[NMAX];
  __shared__ int8_t prop[B * NMAX];
  __shared__ float red[2 * B * W];
  __shared__ double lp[B];
  __shared__ int16_t vlist[ROWS == 2 ? RC_MAXCAND_WIN : RC_MAXCAND];
  __shared__ int nvalid_s;
  __shared__ double diag_s, e_s;
  __shared__ int16_t mv[4 * B];  // slot b: flips applied (2), flips of the next candidate (2)
  const int N = a.g.N;
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const int64_t n = blockIdx.x;
  for (int i = threadIdx.x; i < N; i += (W * WAVE)) s0[i] = s[n * N + i];
  if constexpr (ROWS == 3) rc32_build_tab(a.g, reinterpret_cast<uint16_t*>(wst));
  __syncthreads();
  if (wv == 0) {
    // diagonal part and the ordered list of valid candidates (ballot compaction)
    double dg = diag_energy(H, s0, lane);
    int pos = 0;
    for (int base = 0; base < H.noff; base += WAVE) {
      int k = base + lane;
      bool v = (k < H.noff) && cand_valid(H, s0, k);
      uint64_t m = __ballot(v);
      uint64_t below = (lane == 0) ? 0ull : (m & ((~0ull) >> (64 - lane)));
      if (v) vlist[pos + __popcll(below)] = (int16_t)k;
      pos += __popcll(m);
    }
    if (lane == 0) {
      nvalid_s = pos;
      diag_s = dg;
    }
  }
  __syncthreads();
  const int nvalid = nvalid_s;
  const double lp0 = logabs[n];
  // prop[b] holds s0 with candidate start + b applied: thread b < B undoes
  // the previous candidate's (at most two) flips and applies the next one's,
  // whose sites were fetched before the forward.  These flips and the sum
  // live in LDS, not in registers held across the forward.
  for (int i = threadIdx.x; i < B * N; i += (W * WAVE)) prop[i] = s0[i % N];
  if (threadIdx.x < B) {
    const int b = threadIdx.x;
    int na = -1, nb = -1;
    if (b < nvalid) {
      const int k = vlist[b];
      na = H.off_a[k];
      nb = H.off_b[k];
    }
    mv[4 * b] = (int16_t)na;
    mv[4 * b + 1] = (int16_t)nb;
  }
  if (threadIdx.x == 0) e_s = 0.0;
  __syncthreads();
  if (threadIdx.x < B) {
    const int b = threadIdx.x;
    int8_t* pb = prop + b * N;
    const int na = mv[4 * b], nb = mv[4 * b + 1];
    if (na >= 0) pb[na] = -s0[na];
    if (nb >= 0) pb[nb] = -s0[nb];
  }
  __syncthreads();
  for (int start = 0; start < nvalid; start += B) {
    // thread b < B: the next candidate's sites and this candidate's coupling,
    // fetched before the forward and held in registers across it (their
    // global-memory latency hidden by the forward)
    int na = -1, nb = -1;
    double cb = 0.0;
    if (threadIdx.x < B) {
      const int b = threadIdx.x;
      if (start + B + b < nvalid) {
        const int k = vlist[start + B + b];
        na = H.off_a[k];
        nb = H.off_b[k];
      }
      if (start + b < nvalid) cb = H.off_c[vlist[start + b]];
    }
    RC_FWD(prop, lp);
    if (wv == 0) {
      // sum in candidate order by thread 0, as the couplings' owners pass them on
      double term = 0.0;
      if (lane < B && start + lane < nvalid) term = cb * exp_ratio(lp[lane] - lp0);
      double e = e_s;
      for (int b = 0; b < B && start + b < nvalid; ++b) e += __shfl(term, b);
      if (lane == 0) e_s = e;
    }
    if (threadIdx.x < B) {
      const int b = threadIdx.x;
      int8_t* pb = prop + b * N;
      const int fa = mv[4 * b], fb = mv[4 * b + 1];
      if (fa >= 0) pb[fa] = s0[fa];
      if (fb >= 0) pb[fb] = s0[fb];
      if (na >= 0) pb[na] = -s0[na];
      if (nb >= 0) pb[nb] = -s0[nb];
      mv[4 * b] = (int16_t)na;
      mv[4 * b + 1] = (int16_t)nb;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    eloc[n] = diag_s + e_s;
    if (nconn) nconn[n] = nvalid;
  }
}
#undef RC_FWD


// ---------------------------------------------------------------- Jacobian
// One workgroup (4 waves) per sample.  Every convolution of the forward
// recompute and of the backward pass is an FP32-MFMA GEMM (16x16x4 f32) whose
// operands are gathered from channel-major [c][p] activations in LDS through a
// table of the 9 periodic neighbours of every site:
//   forward   Z[co][p]     = sum_{ci,tap} W[co][ci][tap] X[ci][nb(p, tap)]
//   input grad dX[ci][q]   = sum_{co,tap} W[co][ci][tap] dZ[co][nb(q, 8 - tap)]
//   weight grad G[co][ci,tap] = sum_p dZ[co][p] X[ci][nb(p, tap)]
// (MFMA lane l holds A[m0 + l%16][k0 + l/16], B[k0 + l/16][n0 + l%16]; result
// element i of lane l is D[m0 + 4(l/16) + i][n0 + l%16].)  The forward records,
// per block i, x_i (block input), y_i (conv1 input), z1_i (conv1 output) and
// a1_i (conv2 input) in HBM workspace; the backward produces the O row in
// ravel_pytree order (per layer weight (co, ci, tap), bias (co)).
constexpr int JAC_THREADS = 256;

__device__ __forceinline__ v4f mfma_f32_4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// out[co][p] for all co < C, p < N; K = cin * 9; A fragments (weights) of a
// 16-row tile held in registers across the wave's column tiles.
template <int C, int KS, typename Epi>
__device__ __forceinline__ void jac_conv_fwd_mfma(const float* __restrict__ in, const float* __restrict__ W,
                                                  const int16_t* __restrict__ nbr, int N, int nmax, Epi epi) {
  constexpr int K = (KS == 3) ? 9 : C * 9;  // cin * 9
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const int ntile = (N + 15) / 16;
#pragma unroll
  for (int m0 = 0; m0 < C; m0 += 16) {
    float af[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = 4 * ks + lane / 16;
      af[ks] = (k < K) ? W[(m0 + lane % 16) * K + k] : 0.0f;
    }
    for (int t = wv; t < ntile; t += JAC_THREADS / WAVE) {
      const int p = 16 * t + lane % 16;
      const int pc = p < N ? p : N - 1;
      v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int k = 4 * ks + lane / 16;
        const int ci = k / 9, tap = k - 9 * ci;
        const float b = (k < K) ? in[ci * N + nbr[tap * nmax + pc]] : 0.0f;
        acc = mfma_f32_4(af[ks], b, acc);
      }
      const int pw = 16 * t + lane % 16;
      if (pw < N) {
#pragma unroll
        for (int i = 0; i < 4; ++i) epi(m0 + 4 * (lane / 16) + i, pw, acc[i]);
      }
    }
  }
}

// din[ci][q] for all ci < C (input channels of a C -> C convolution)
template <int C, typename Epi>
__device__ __forceinline__ void jac_conv_bwd_in_mfma(const float* __restrict__ dz, const float* __restrict__ W,
                                                     const int16_t* __restrict__ nbr, int N, int nmax, Epi epi) {
  constexpr int KS = C * 9 / 4;
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const int ntile = (N + 15) / 16;
#pragma unroll
  for (int m0 = 0; m0 < C; m0 += 16) {
    float af[KS];
    const int ci = m0 + lane % 16;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = 4 * ks + lane / 16;
      const int co = k / 9, tap = k - 9 * co;
      af[ks] = W[(co * C + ci) * 9 + tap];
    }
    for (int t = wv; t < ntile; t += JAC_THREADS / WAVE) {
      const int q = 16 * t + lane % 16;
      const int qc = q < N ? q : N - 1;
      v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int k = 4 * ks + lane / 16;
        const int co = k / 9, tap = k - 9 * co;
        acc = mfma_f32_4(af[ks], dz[co * N + nbr[(8 - tap) * nmax + qc]], acc);
      }
      if (q < N) {
#pragma unroll
        for (int i = 0; i < 4; ++i) epi(m0 + 4 * (lane / 16) + i, q, acc[i]);
      }
    }
  }
}

// G[co][n], n = ci * 9 + tap < cin * 9, written to the O row; bias grads by VALU
template <int C>
__device__ __forceinline__ void jac_conv_grad_w_mfma(const float* __restrict__ dz, const float* __restrict__ x, int cin,
                                                     const int16_t* __restrict__ nbr, int N, int nmax,
                                                     float* __restrict__ Orow, int64_t w0, int64_t b0) {
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const int K9 = cin * 9;
  const int ncol = (K9 + 15) / 16;
  const int ntiles = (C / 16) * ncol;
  const int ksteps = (N + 3) / 4;
  for (int t = wv; t < ntiles; t += JAC_THREADS / WAVE) {
    const int m0 = 16 * (t % (C / 16)), n0 = 16 * (t / (C / 16));
    const int co = m0 + lane % 16;
    const int n = n0 + lane % 16;
    const bool nok = n < K9;
    const int ci = nok ? n / 9 : 0, tap = nok ? n - 9 * (n / 9) : 0;
    const float* xr = x + ci * N;
    const int16_t* nr = nbr + tap * nmax;
    const float* dr = dz + co * N;
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
    for (int ks = 0; ks < ksteps; ++ks) {
      const int p = 4 * ks + lane / 16;
      const bool pok = p < N;
      const float a = pok ? dr[p] : 0.0f;
      const float b = (pok && nok) ? xr[nr[pok ? p : 0]] : 0.0f;
      acc = mfma_f32_4(a, b, acc);
    }
    if (nok) {
#pragma unroll
      for (int i = 0; i < 4; ++i) Orow[w0 + (int64_t)(m0 + 4 * (lane / 16) + i) * K9 + n] = acc[i];
    }
  }
  if (b0 >= 0) {
    for (int c = wv; c < C; c += JAC_THREADS / WAVE) {
      float acc = 0.0f;
      for (int p = lane; p < N; p += WAVE) acc += dz[c * N + p];
      acc = wave_sum(acc);
      if (lane == 0) Orow[b0 + c] = acc;
    }
  }
}

template <int C, int NMAX>
__global__ __launch_bounds__(JAC_THREADS) void rc_jacobian_kernel(RcGeom g, int nb, int actk,
                                                                  const float* __restrict__ theta,
                                                                  const int8_t* __restrict__ s, int64_t s0,
                                                                  int64_t ns_chunk, int64_t np,
                                                                  float* __restrict__ O, float* __restrict__ ws) {
  __shared__ float buf0[C * NMAX];
  __shared__ float buf1[C * NMAX];
  __shared__ float buf2[C * NMAX];
  __shared__ int16_t nbr[9 * NMAX];
  __shared__ float red[2 * JAC_THREADS / WAVE];
  const int64_t nl = blockIdx.x;
  if (nl >= ns_chunk) return;
  const int64_t n = s0 + nl;
  const int N = g.N;
  const int64_t blk_f = (int64_t)4 * C * N;
  float* rec = ws + nl * blk_f * nb;
  float* Orow = O + n * np;
  const float isq2 = 0.70710678118654752f;
  float* bX = buf0;  // block input / running gradient
  float* bA = buf1;
  float* bB = buf2;

  for (int e = threadIdx.x; e < 9 * N; e += JAC_THREADS) {
    const int tap = e / N, p = e - tap * N;
    nbr[tap * NMAX + p] = (int16_t)tap_pos(g, p / g.L1, p % g.L1, tap);
  }
  // ---- forward with record
  for (int p = threadIdx.x; p < N; p += JAC_THREADS) bX[p] = (float)s[n * N + p];
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    const int cin = (i == 0) ? 1 : C;
    float* rx = rec + i * blk_f;
    float* ry = rx + C * N;
    float* rz = ry + C * N;
    float* ra = rz + C * N;
    const float sc = rsqrtf((float)(i + 1));
    for (int e = threadIdx.x; e < cin * N; e += JAC_THREADS) {
      float xv = bX[e];
      rx[e] = xv;
      float xin = xv * sc;
      float yv = (i == 0) ? xin * isq2 : act_f(xin, actk);
      bA[e] = yv;
      ry[e] = yv;
    }
    __syncthreads();
    int64_t w1, b1, w2, b2;
    rc_theta_offsets(nb, C, 2 * i, w1, b1);
    rc_theta_offsets(nb, C, 2 * i + 1, w2, b2);
    {
      const float* bias = theta + b1;
      auto epi = [&](int co, int p, float v) {
        v += bias[co];
        rz[co * N + p] = v;
        const float a = act_f(v, actk);
        bB[co * N + p] = a;
        ra[co * N + p] = a;
      };
      if (i == 0)
        jac_conv_fwd_mfma<C, 3>(bA, theta + w1, nbr, N, NMAX, epi);
      else
        jac_conv_fwd_mfma<C, C * 9 / 4>(bA, theta + w1, nbr, N, NMAX, epi);
    }
    __syncthreads();
    {
      const float* bias = (b2 >= 0) ? theta + b2 : nullptr;
      auto epi = [&](int co, int p, float v) {
        if (bias) v += bias[co];
        v += (cin == 1) ? bX[p] : bX[co * N + p];  // residual (block 0: the spin, broadcast)
        bA[co * N + p] = v;
      };
      jac_conv_fwd_mfma<C, C * 9 / 4>(bB, theta + w2, nbr, N, NMAX, epi);
    }
    __syncthreads();
    float* t = bX;
    bX = bA;
    bA = t;
  }
  // ---- head: dX = softmax(h) / sqrt(nb + 1), h = x / sqrt(nb + 1)
  const float hs = rsqrtf((float)(nb + 1));
  {
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    float m = -INFINITY;
    for (int e = threadIdx.x; e < C * N; e += JAC_THREADS) m = fmaxf(m, bX[e] * hs);
    m = wave_max(m);
    if (lane == 0) red[wv] = m;
    __syncthreads();
    m = -INFINITY;
    for (int w = 0; w < JAC_THREADS / WAVE; ++w) m = fmaxf(m, red[w]);
    float ssum = 0.0f;
    for (int e = threadIdx.x; e < C * N; e += JAC_THREADS) ssum += expf(bX[e] * hs - m);
    ssum = wave_sum(ssum);
    __syncthreads();
    if (lane == 0) red[wv] = ssum;
    __syncthreads();
    ssum = 0.0f;
    for (int w = 0; w < JAC_THREADS / WAVE; ++w) ssum += red[w];
    float inv = 1.0f / ssum;
    for (int e = threadIdx.x; e < C * N; e += JAC_THREADS) bX[e] = expf(bX[e] * hs - m) * inv * hs;
    __syncthreads();
  }
  // ---- backward
  for (int i = nb - 1; i >= 0; --i) {
    const int cin = (i == 0) ? 1 : C;
    float* rx = rec + i * blk_f;
    float* ry = rx + C * N;
    float* rz = ry + C * N;
    float* ra = rz + C * N;
    int64_t w1, b1, w2, b2;
    rc_theta_offsets(nb, C, 2 * i, w1, b1);
    rc_theta_offsets(nb, C, 2 * i + 1, w2, b2);
    // conv2: dz2 = dX; input a1
    for (int e = threadIdx.x; e < C * N; e += JAC_THREADS) bA[e] = ra[e];
    __syncthreads();
    jac_conv_grad_w_mfma<C>(bX, bA, C, nbr, N, NMAX, Orow, w2, b2);
    // da1 -> dz1 = da1 * act'(z1), into bB
    jac_conv_bwd_in_mfma<C>(bX, theta + w2, nbr, N, NMAX, [&](int ci, int q, float v) {
      bB[ci * N + q] = v * act_grad_f(rz[ci * N + q], actk);
    });
    __syncthreads();
    // conv1: input y_i
    for (int e = threadIdx.x; e < cin * N; e += JAC_THREADS) bA[e] = ry[e];
    __syncthreads();
    jac_conv_grad_w_mfma<C>(bB, bA, cin, nbr, N, NMAX, Orow, w1, b1);
    __syncthreads();
    if (i > 0) {
      const float sc = rsqrtf((float)(i + 1));
      jac_conv_bwd_in_mfma<C>(bB, theta + w1, nbr, N, NMAX, [&](int ci, int q, float v) {
        const float xin = rx[ci * N + q] * sc;
        bA[ci * N + q] = v * act_grad_f(xin, actk) * sc + bX[ci * N + q];
      });
      __syncthreads();
      float* t = bX;
      bX = bA;
      bA = t;
    }
  }
}

}  // namespace qtx

using namespace qtx;

// ================================================================ C ABI
static int rc_check(const qtx_resconv_desc* d) {
  QTX_REQUIRE(d != nullptr, "null qtx_resconv_desc");
  QTX_REQUIRE(d->channels == 16 || d->channels == 32, "channels must be 16 or 32");
  QTX_REQUIRE(d->nblocks >= 1 && d->nblocks <= 64, "nblocks out of range (1..64)");
  QTX_REQUIRE(d->L0 >= 1 && d->L1 >= 1 && d->L0 * d->L1 <= RC_MAXN, "L0 * L1 out of range (1..1024)");
  QTX_REQUIRE(d->precision == QTX_CONV_FP32 || d->precision == QTX_CONV_BF16X3, "unknown precision");
  QTX_REQUIRE(d->activation == QTX_ACT_GELU_TANH || d->activation == QTX_ACT_GELU_ERF ||
                  d->activation == QTX_ACT_RELU,
              "unknown activation");
  if (d->channels == 16) QTX_REQUIRE(d->L0 * d->L1 <= 512, "channels = 16 needs L0 * L1 <= 512");
  if (d->channels == 32) QTX_REQUIRE(d->L0 * d->L1 <= 400, "channels = 32 needs L0 * L1 <= 400");
  return QTX_OK;
}

static RcArgs rc_args(const RcLayout& L, const void* packed) {
  RcArgs a;
  a.g = RcGeom{L.L0, L.L1, L.N, (L.N + 15) / 16, std::log((double)L.C * (double)L.N)};
  a.nb = L.nb;
  a.actk = L.actk;
  a.packed = (const char*)packed;
  a.off_l0w = L.off_l0w;
  a.off_l0b = L.off_l0b;
  a.bias_off = L.bias_off;
  a.off_frag = L.off_frag;
  a.frag_bytes = L.frag_bytes;
  return a;
}

// Kernel configurations: (C, B, NMAX, MAXT) with B * ceil(NMAX/16) <= 4 * MAXT
// (C = 16: N <= 256 -> 2 configurations per workgroup, N <= 512 -> 1; C = 32: N <= 400 -> 1).
#define RC_DISPATCH(KERNEL, NS_PER_BLOCK_OUT, GRID_FROM_B, ...)                                        \
  do {                                                                                                 \
    if (L.win) {                                                                                       \
      constexpr int B_ = 2;                                                                            \
      NS_PER_BLOCK_OUT = B_;                                                                           \
      dim3 grid_ = GRID_FROM_B;                                                                        \
      if (L.mode == QTX_CONV_FP32)                                                                     \
        hipLaunchKernelGGL((KERNEL<16, QTX_CONV_FP32, 2, 256, 4, 4, 2>), grid_, dim3(256), 0, st, __VA_ARGS__); \
      else                                                                                             \
        hipLaunchKernelGGL((KERNEL<16, QTX_CONV_BF16X3, 2, 256, 4, 4, 2>), grid_, dim3(256), 0, st, __VA_ARGS__); \
    } else if (L.C == 16 && L.L1 == 16 && L.L0 <= 16) {                                                \
      constexpr int B_ = 2;                                                                            \
      NS_PER_BLOCK_OUT = B_;                                                                           \
      dim3 grid_ = GRID_FROM_B;                                                                        \
      if (L.mode == QTX_CONV_FP32)                                                                     \
        hipLaunchKernelGGL((KERNEL<16, QTX_CONV_FP32, 2, 256, 4, 8, 1>), grid_, dim3(512), 0, st, __VA_ARGS__); \
      else                                                                                             \
        hipLaunchKernelGGL((KERNEL<16, QTX_CONV_BF16X3, 2, 256, 4, 8, 1>), grid_, dim3(512), 0, st, __VA_ARGS__); \
    } else if (L.C == 16 && L.N <= 256) {                                                              \
      constexpr int B_ = 2;                                                                            \
      NS_PER_BLOCK_OUT = B_;                                                                           \
      dim3 grid_ = GRID_FROM_B;                                                                        \
      if (L.mode == QTX_CONV_FP32)                                                                     \
        hipLaunchKernelGGL((KERNEL<16, QTX_CONV_FP32, 2, 256, 8, 4, 0>), grid_, dim3(256), 0, st, __VA_ARGS__); \
      else                                                                                             \
        hipLaunchKernelGGL((KERNEL<16, QTX_CONV_BF16X3, 2, 256, 8, 4, 0>), grid_, dim3(256), 0, st, __VA_ARGS__); \
    } else if (L.C == 16) {                                                                            \
      constexpr int B_ = 1;                                                                            \
      NS_PER_BLOCK_OUT = B_;                                                                           \
      dim3 grid_ = GRID_FROM_B;                                                                        \
      if (L.mode == QTX_CONV_FP32)                                                                     \
        hipLaunchKernelGGL((KERNEL<16, QTX_CONV_FP32, 1, 512, 8, 4, 0>), grid_, dim3(256), 0, st, __VA_ARGS__); \
      else                                                                                             \
        hipLaunchKernelGGL((KERNEL<16, QTX_CONV_BF16X3, 1, 512, 8, 4, 0>), grid_, dim3(256), 0, st, __VA_ARGS__); \
    } else {                                                                                           \
      constexpr int B_ = 1;                                                                            \
      NS_PER_BLOCK_OUT = B_;                                                                           \
      dim3 grid_ = GRID_FROM_B;                                                                        \
      if (L.mode == QTX_CONV_FP32)                                                                     \
        hipLaunchKernelGGL((KERNEL<32, QTX_CONV_FP32, 1, 400, 7, 4, 0>), grid_, dim3(256), 0, st, __VA_ARGS__); \
      else                                                                                             \
        hipLaunchKernelGGL((KERNEL<32, QTX_CONV_BF16X3, 1, RC32_MAXN, 7, 4, 3>), grid_, dim3(256), 0, st, __VA_ARGS__); \
    }                                                                                                  \
  } while (0)

extern "C" {

int64_t qtx_resconv_nparams(const qtx_resconv_desc* d) {
  if (!d || d->nblocks < 1 || d->channels < 1) return -1;
  return rc_nparams(d->nblocks, d->channels);
}

size_t qtx_resconv_packed_bytes(const qtx_resconv_desc* d) {
  if (rc_check(d) != QTX_OK) return 0;
  return (size_t)rc_layout(d).total_bytes;
}

int qtx_resconv_pack(void* stream, const qtx_resconv_desc* d, const float* theta, void* packed) {
  int rc = rc_check(d);
  if (rc) return rc;
  QTX_REQUIRE(theta && packed, "null pointer");
  RcLayout L = rc_layout(d);
  hipStream_t st = as_stream(stream);
  int64_t misc = L.nparams + (int64_t)L.C * 10 + (int64_t)2 * L.nb * L.C;
  hipLaunchKernelGGL(rc_pack_misc_kernel, dim3((unsigned)((misc + 255) / 256)), dim3(256), 0, st, theta, L.nb, L.C,
                     L.nparams, (float*)packed, L.off_l0w, L.off_l0b, L.bias_off);
  int ct = L.C / 16;
  int ks = (L.mode == QTX_CONV_FP32) ? 9 * L.C / 4 : (L.C == 16 ? (L.win ? 6 : 5) : 9);
  int64_t total = (int64_t)ct * ks * WAVE * (2 * L.nb - 1);
  if (total > 0) {
    dim3 g((unsigned)((total + 255) / 256)), b(256);
    if (L.C == 16 && L.mode == QTX_CONV_FP32)
      hipLaunchKernelGGL((rc_pack_frag_kernel<16, QTX_CONV_FP32, 0>), g, b, 0, st, theta, L.nb, (char*)packed,
                         L.off_frag, L.frag_bytes);
    else if (L.C == 16 && L.win)
      hipLaunchKernelGGL((rc_pack_frag_kernel<16, QTX_CONV_BF16X3, 1>), g, b, 0, st, theta, L.nb, (char*)packed,
                         L.off_frag, L.frag_bytes);
    else if (L.C == 16)
      hipLaunchKernelGGL((rc_pack_frag_kernel<16, QTX_CONV_BF16X3, 0>), g, b, 0, st, theta, L.nb, (char*)packed,
                         L.off_frag, L.frag_bytes);
    else if (L.mode == QTX_CONV_FP32)
      hipLaunchKernelGGL((rc_pack_frag_kernel<32, QTX_CONV_FP32, 0>), g, b, 0, st, theta, L.nb, (char*)packed,
                         L.off_frag, L.frag_bytes);
    else
      hipLaunchKernelGGL((rc_pack_frag_kernel<32, QTX_CONV_BF16X3, 0>), g, b, 0, st, theta, L.nb, (char*)packed,
                         L.off_frag, L.frag_bytes);
  }
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_resconv_logpsi(void* stream, const qtx_resconv_desc* d, const void* packed, const int8_t* s, int64_t ns,
                       double* logabs) {
  int rc = rc_check(d);
  if (rc) return rc;
  QTX_REQUIRE(packed && ns >= 0 && (ns == 0 || (s && logabs)), "bad arguments");
  if (ns == 0) return QTX_OK;
  RcLayout L = rc_layout(d);
  RcArgs a = rc_args(L, packed);
  hipStream_t st = as_stream(stream);
  int per = 1;
  RC_DISPATCH(rc_logpsi_kernel, per, dim3((unsigned)((ns + B_ - 1) / B_)), a, s, ns, logabs);
  (void)per;
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_resconv_metropolis(void* stream, const qtx_resconv_desc* d, const void* packed, int kind,
                           const int32_t* nbr, int32_t width, int32_t hopping, const uint32_t* words,
                           int32_t nsteps, int64_t ns, double reweight, double tie_tol, int8_t* s, double* logabs,
                           int32_t* n_accept, int32_t* n_near) {
  int rc = rc_check(d);
  if (rc) return rc;
  QTX_REQUIRE(kind == QTX_LOCALFLIP || kind == QTX_EXCHANGE, "kind must be QTX_LOCALFLIP or QTX_EXCHANGE");
  QTX_REQUIRE(kind == QTX_LOCALFLIP || (nbr != nullptr && width > 0), "EXCHANGE needs a neighbour table");
  QTX_REQUIRE(hopping == 1 || hopping == -1, "hopping must be +1 or -1");
  QTX_REQUIRE(nsteps >= 0 && ns >= 0, "negative size");
  QTX_REQUIRE(packed && (ns == 0 || (s && logabs && n_accept && n_near && (nsteps == 0 || words))),
              "null pointer");
  if (ns == 0) return QTX_OK;
  RcLayout L = rc_layout(d);
  RcArgs a = rc_args(L, packed);
  hipStream_t st = as_stream(stream);
  int per = 1;
  RC_DISPATCH(rc_metropolis_kernel, per, dim3((unsigned)((ns + B_ - 1) / B_)), a, kind, nbr, width, hopping, words,
              nsteps, ns, reweight, tie_tol, s, logabs, n_accept, n_near);
  (void)per;
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_resconv_eloc(void* stream, const qtx_resconv_desc* d, const void* packed, const qtx_ham* H,
                     const int8_t* s, const double* logabs, int64_t ns, double* eloc, int32_t* nconn) {
  int rc = rc_check(d);
  if (rc) return rc;
  QTX_REQUIRE(H != nullptr, "null qtx_ham");
  QTX_REQUIRE(H->N == d->L0 * d->L1, "qtx_ham.N != L0 * L1");
  RcLayout L = rc_layout(d);
  QTX_REQUIRE(H->noff <= (L.win ? RC_MAXCAND_WIN : RC_MAXCAND),
              L.win ? "too many off-diagonal candidates (max 3072 on a 16 x 16 lattice with 16 channels)"
                    : "too many off-diagonal candidates (max 4096)");
  QTX_REQUIRE(H->noff == 0 || (H->off_a && H->off_b && H->off_ra && H->off_rb && H->off_c), "null table");
  QTX_REQUIRE(H->ndiag == 0 || (H->diag_c && H->diag_a && H->diag_b && H->diag_kind), "null table");
  QTX_REQUIRE(packed && ns >= 0 && (ns == 0 || (s && logabs && eloc)), "bad arguments");
  if (ns == 0) return QTX_OK;
  QTX_REQUIRE(ns <= 0x7fffffff, "ns too large");
  RcArgs a = rc_args(L, packed);
  hipStream_t st = as_stream(stream);
  int per = 1;
  RC_DISPATCH(rc_eloc_kernel, per, dim3((unsigned)ns), a, *H, s, logabs, ns, eloc, nconn);
  (void)per;
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

static int64_t jac_chunk(int64_t ns) { return ns < 2048 ? ns : 2048; }

size_t qtx_resconv_jacobian_ws_bytes(const qtx_resconv_desc* d, int64_t ns) {
  if (rc_check(d) != QTX_OK || ns <= 0) return 0;
  int64_t N = (int64_t)d->L0 * d->L1;
  return (size_t)(jac_chunk(ns) * (int64_t)d->nblocks * 4 * d->channels * N * 4);
}

int qtx_resconv_jacobian(void* stream, const qtx_resconv_desc* d, const void* packed, const int8_t* s, int64_t ns,
                         float* O, void* ws, size_t ws_bytes) {
  int rc = rc_check(d);
  if (rc) return rc;
  QTX_REQUIRE(packed && ns >= 0 && (ns == 0 || (s && O)), "bad arguments");
  if (ns == 0) return QTX_OK;
  size_t need = qtx_resconv_jacobian_ws_bytes(d, ns);
  if (ws == nullptr || ws_bytes < need)
    return fail(QTX_EWORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  RcLayout L = rc_layout(d);
  RcGeom g{L.L0, L.L1, L.N, (L.N + 15) / 16, std::log((double)L.C * (double)L.N)};
  hipStream_t st = as_stream(stream);
  const float* theta = (const float*)packed;  // raw theta copy at offset 0
  int64_t chunk = jac_chunk(ns);
  for (int64_t s0 = 0; s0 < ns; s0 += chunk) {
    int64_t nc = (ns - s0) < chunk ? (ns - s0) : chunk;
    if (L.C == 16 && L.N <= 256)
      hipLaunchKernelGGL((rc_jacobian_kernel<16, 256>), dim3((unsigned)nc), dim3(JAC_THREADS), 0, st, g, L.nb,
                         L.actk, theta, s, s0, nc, L.nparams, O, (float*)ws);
    else if (L.C == 16)
      hipLaunchKernelGGL((rc_jacobian_kernel<16, 512>), dim3((unsigned)nc), dim3(JAC_THREADS), 0, st, g, L.nb,
                         L.actk, theta, s, s0, nc, L.nparams, O, (float*)ws);
    else
      hipLaunchKernelGGL((rc_jacobian_kernel<32, 400>), dim3((unsigned)nc), dim3(JAC_THREADS), 0, st, g, L.nb,
                         L.actk, theta, s, s0, nc, L.nparams, O, (float*)ws);
    QTX_LAUNCH_CHECK();
  }
  return QTX_OK;
}

// ---------------------------------------------------------------- allocation-free Jacobian
// The per-sample forward records of qtx_resconv_jacobian_noalloc live in this
// file-scope device array (one copy per GPU, held until the process ends)
// instead of a caller workspace; samples run in launches of as many as the
// array holds.  Host code uses the per-device address from
// hipGetSymbolAddress, never the array's name (a host-side stand-in).
constexpr size_t RC_JAC_SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) char scratch[RC_JAC_SCRATCH_BYTES];

namespace {
constexpr int RC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[RC_MAX_DEVICES] = {};

// the calling thread's current device's copy of scratch
int jac_scratch(void** out) {
  int dev = 0;
  QTX_HIP(hipGetDevice(&dev));
  QTX_REQUIRE(dev >= 0 && dev < RC_MAX_DEVICES, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    QTX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QTX_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the Jacobian scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return QTX_OK;
}
}  // namespace

int qtx_resconv_jacobian_noalloc(void* stream, const qtx_resconv_desc* d, const void* packed, const int8_t* s,
                                 int64_t ns, float* O) {
  int rc = rc_check(d);
  if (rc) return rc;
  QTX_REQUIRE(packed && ns >= 0 && (ns == 0 || (s && O)), "bad arguments");
  void* ws = nullptr;
  rc = jac_scratch(&ws);
  if (rc) return rc;
  if (ns == 0) return QTX_OK;
  const size_t per_sample = qtx_resconv_jacobian_ws_bytes(d, 1);
  if (per_sample > RC_JAC_SCRATCH_BYTES)
    return fail(QTX_EINVAL, "one sample's Jacobian records need " + std::to_string(per_sample) +
                                " bytes, more than the " + std::to_string(RC_JAC_SCRATCH_BYTES) +
                                "-byte scratch of the allocation-free entry point");
  const int64_t per_launch = (int64_t)(RC_JAC_SCRATCH_BYTES / per_sample);
  for (int64_t s0 = 0; s0 < ns; s0 += per_launch) {
    const int64_t nc = (ns - s0) < per_launch ? (ns - s0) : per_launch;
    // the kernel indexes samples from the start of the pointers it is given
    rc = qtx_resconv_jacobian(stream, d, packed, s + s0 * (int64_t)d->L0 * d->L1, nc,
                              O + s0 * qtx_resconv_nparams(d), ws, per_launch * per_sample);
    if (rc) return rc;
  }
  return QTX_OK;
}

}  // extern "C"
