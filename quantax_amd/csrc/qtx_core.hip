// Streams, initial chains, operator tables and the RBM_Dense path.
//
// RBM layout: packed = [Wt (N, M) | b (M)] in the parameter dtype, so the
// column W[:, j] touched by a flip of site j is one contiguous row of Wt.
// One wavefront owns one chain / sample; lane l holds hidden units
// l, l + 64, ... (R = ceil(M / 64) registers per lane).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "common.hpp"
#include "rbm_device.hpp"

namespace qtx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return QTX_OK;
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return QTX_EHIP;
}

// ================================================================ streams
__global__ void philox_fill_kernel(uint64_t seed, uint32_t tag, int64_t step0, int64_t nsteps,
                                   int64_t chain0, int64_t nchains, uint32_t* __restrict__ words) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = nsteps * nchains;
  if (idx >= total) return;
  int64_t t = idx / nchains, c = idx % nchains;
  u32x4 ctr{(uint32_t)(chain0 + c), (uint32_t)(step0 + t), tag, 0u};
  u32x4 r = philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
  reinterpret_cast<uint4*>(words)[idx] = make_uint4(r.x, r.y, r.z, r.w);
}

// one thread per chain: uniform spins from Philox bits, or Fisher-Yates of
// nup up spins (the oracle's restatement of utils/basis.py:121-133).
__global__ void init_spins_kernel(uint64_t seed, int64_t nchains, int N, int nup, int8_t* __restrict__ s) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchains) return;
  int8_t* row = s + c * N;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  if (nup < 0) {
    for (int blk = 0; blk * 128 < N; ++blk) {
      u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)blk, TAG_INIT, 1u}, k0, k1);
      uint32_t wv[4] = {r.x, r.y, r.z, r.w};
      for (int b = 0; b < 128 && blk * 128 + b < N; ++b) {
        uint32_t bit = (wv[b / 32] >> (b % 32)) & 1u;
        row[blk * 128 + b] = bit ? 1 : -1;
      }
    }
    return;
  }
  for (int i = 0; i < N; ++i) row[i] = (i < nup) ? 1 : -1;
  for (int i = N - 1; i > 0; --i) {
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)i, TAG_INIT, 0u}, k0, k1);
    int j = (int)bounded(r.x, (uint32_t)(i + 1));
    int8_t a = row[i];
    row[i] = row[j];
    row[j] = a;
  }
}

// ================================================================ operator
__global__ __launch_bounds__(256) void ham_diag_kernel(qtx_ham H, const int8_t* __restrict__ s, int64_t ns,
                                                       double* __restrict__ out) {
  __shared__ int8_t sbuf[WAVES_PER_BLOCK][MAXN];
  int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  int64_t i = (int64_t)blockIdx.x * WAVES_PER_BLOCK + wv;
  if (i >= ns) return;
  load_spins(sbuf[wv], s + i * H.N, H.N, lane);
  __builtin_amdgcn_wave_barrier();
  double e = diag_energy(H, sbuf[wv], lane);
  if (lane == 0) out[i] = e;
}

__global__ __launch_bounds__(256) void conn_count_kernel(qtx_ham H, const int8_t* __restrict__ s, int64_t ns,
                                                         int32_t* __restrict__ count) {
  __shared__ int8_t sbuf[WAVES_PER_BLOCK][MAXN];
  int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  int64_t i = (int64_t)blockIdx.x * WAVES_PER_BLOCK + wv;
  if (i >= ns) return;
  load_spins(sbuf[wv], s + i * H.N, H.N, lane);
  __builtin_amdgcn_wave_barrier();
  int cnt = 0;
  for (int k = lane; k < H.noff; k += WAVE) cnt += cand_valid(H, sbuf[wv], k) ? 1 : 0;
  cnt = wave_sum(cnt);
  if (lane == 0) count[i] = cnt;
}

__global__ __launch_bounds__(256) void conn_list_kernel(qtx_ham H, const int8_t* __restrict__ s, int64_t ns,
                                                        const int64_t* __restrict__ offsets,
                                                        int32_t* __restrict__ segment, int32_t* __restrict__ cand,
                                                        int8_t* __restrict__ s_conn, double* __restrict__ h_conn) {
  __shared__ int8_t sbuf[WAVES_PER_BLOCK][MAXN];
  int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  int64_t i = (int64_t)blockIdx.x * WAVES_PER_BLOCK + wv;
  if (i >= ns) return;
  const int N = H.N;
  int8_t* sl = sbuf[wv];
  load_spins(sl, s + i * N, N, lane);
  __builtin_amdgcn_wave_barrier();
  int64_t pos = offsets[i];
  for (int base = 0; base < H.noff; base += WAVE) {
    int k = base + lane;
    bool v = (k < H.noff) && cand_valid(H, sl, k);
    uint64_t m = __ballot(v);
    uint64_t below = (lane == 0) ? 0ull : (m & ((~0ull) >> (64 - lane)));
    int64_t my = pos + __popcll(below);
    if (v) {
      segment[my] = (int32_t)i;
      cand[my] = k;
      h_conn[my] = H.off_c[k];
    }
    // write the configurations row by row (wave-cooperative, coalesced)
    uint64_t mm = m;
    int64_t row = pos;
    while (mm) {
      int bit = __ffsll((unsigned long long)mm) - 1;
      mm &= mm - 1;
      int kk = base + bit;
      int a = H.off_a[kk], b = H.off_b[kk];
      int8_t* dst = s_conn + row * N;
      for (int j = lane; j < N; j += WAVE) {
        int8_t v8 = sl[j];
        if (j == a || j == b) v8 = -v8;
        dst[j] = v8;
      }
      ++row;
    }
    pos += __popcll(m);
  }
}

// ================================================================ RBM
template <typename T>
__global__ void rbm_pack_kernel(const T* __restrict__ theta, int M, int N, T* __restrict__ packed) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)M * N + M;
  if (idx >= total) return;
  if (idx < (int64_t)M * N) {
    int j = (int)(idx / M), i = (int)(idx % M);  // packed[j*M + i] = W[i, j]
    packed[idx] = theta[(int64_t)i * N + j];
  } else {
    packed[idx] = theta[idx];
  }
}

template <typename T, int R>
__global__ __launch_bounds__(256) void rbm_logpsi_kernel(const T* __restrict__ packed, int M, int N,
                                                         const int8_t* __restrict__ s, int64_t ns,
                                                         double* __restrict__ logabs, T* __restrict__ h_out) {
  __shared__ int8_t sbuf[WAVES_PER_BLOCK][MAXN];
  int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  int64_t n = (int64_t)blockIdx.x * WAVES_PER_BLOCK + wv;
  if (n >= ns) return;
  rbm_chain_logpsi<T, R>(packed, M, N, s + n * N, sbuf[wv], lane, logabs + n, h_out ? h_out + n * M : nullptr);
}

template <typename T, int R>
__global__ __launch_bounds__(256) void rbm_metropolis_kernel(
    const T* __restrict__ packed, int M, int N, int kind, const int32_t* __restrict__ nbr, int width,
    int hopping, const uint32_t* __restrict__ words, int nsteps, int64_t ns, double reweight, double tie_tol,
    int8_t* __restrict__ s, double* __restrict__ logabs, int32_t* __restrict__ n_accept,
    int32_t* __restrict__ n_near) {
  __shared__ int8_t sbuf[WAVES_PER_BLOCK][MAXN];
  int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  int64_t c = xcd_block(blockIdx.x, gridDim.x) * WAVES_PER_BLOCK + wv;
  if (c >= ns) return;
  auto getw = [&](int t) { return reinterpret_cast<const uint4*>(words)[(int64_t)t * ns + c]; };
  rbm_chain_metropolis<T, R>(packed, M, N, kind, nbr, width, hopping, nsteps, getw, reweight, tie_tol, s + c * N,
                             sbuf[wv], lane, logabs + c, n_accept + c, n_near + c);
}

template <typename T, int R>
__global__ __launch_bounds__(256) void rbm_eloc_kernel(const T* __restrict__ packed, int M, int N, qtx_ham H,
                                                       const int8_t* __restrict__ s, int64_t ns,
                                                       double* __restrict__ eloc, int32_t* __restrict__ nconn) {
  __shared__ int8_t sbuf[WAVES_PER_BLOCK][MAXN];
  int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  int64_t n = (int64_t)blockIdx.x * WAVES_PER_BLOCK + wv;
  if (n >= ns) return;
  rbm_sample_eloc<T, R>(packed, M, N, H, s + n * N, sbuf[wv], lane, eloc + n, nconn ? nconn + n : nullptr);
}

template <typename T, typename TO>
__global__ __launch_bounds__(256) void rbm_jacobian_kernel(const T* __restrict__ packed, int M, int N,
                                                           const int8_t* __restrict__ s, int64_t ns,
                                                           TO* __restrict__ O) {
  __shared__ float tl[MAXN];
  __shared__ double tld[MAXN];
  __shared__ int8_t sl[MAXN];
  int64_t n = blockIdx.x;
  if (n >= ns) return;
  rbm_jacobian_row<T, TO>(packed, M, N, s + n * N, O + n * ((int64_t)M * N + M), tl, tld, sl);
}

}  // namespace qtx

using namespace qtx;

// ================================================================ C ABI
static int blocks_for(int64_t n, int per) { return (int)((n + per - 1) / per); }

static int check_ham(const qtx_ham* H) {
  QTX_REQUIRE(H != nullptr, "null qtx_ham");
  QTX_REQUIRE(H->N > 0 && H->N <= MAXN, "qtx_ham.N out of range (1..1024)");
  QTX_REQUIRE(H->ndiag >= 0 && H->noff >= 0, "negative table size");
  QTX_REQUIRE(H->ndiag == 0 || (H->diag_c && H->diag_a && H->diag_b && H->diag_kind), "null diag table");
  QTX_REQUIRE(H->noff == 0 || (H->off_a && H->off_b && H->off_ra && H->off_rb && H->off_c), "null off-diag table");
  return QTX_OK;
}

extern "C" {

const char* qtx_last_error(void) { return g_last_error.c_str(); }

int qtx_device_sync(void) {
  QTX_HIP(hipDeviceSynchronize());
  return QTX_OK;
}

int qtx_philox_fill(void* stream, uint64_t seed, uint32_t tag, int64_t step0, int64_t nsteps, int64_t chain0,
                    int64_t nchains, uint32_t* words) {
  QTX_REQUIRE(nsteps >= 0 && nchains >= 0, "negative size");
  QTX_REQUIRE(words != nullptr || nsteps * nchains == 0, "null words");
  int64_t total = nsteps * nchains;
  if (total == 0) return QTX_OK;
  hipLaunchKernelGGL(philox_fill_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, as_stream(stream), seed, tag,
                     step0, nsteps, chain0, nchains, words);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_init_spins(void* stream, uint64_t seed, int64_t nchains, int32_t N, int32_t nup, int8_t* s) {
  QTX_REQUIRE(N > 0 && nchains >= 0 && s != nullptr, "bad init_spins arguments");
  QTX_REQUIRE(nup <= N, "nup > N");
  if (nchains == 0) return QTX_OK;
  hipLaunchKernelGGL(init_spins_kernel, dim3(blocks_for(nchains, 128)), dim3(128), 0, as_stream(stream), seed,
                     nchains, N, nup, s);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_ham_diag(void* stream, const qtx_ham* H, const int8_t* s, int64_t ns, double* diag) {
  int rc = check_ham(H);
  if (rc) return rc;
  QTX_REQUIRE(ns >= 0 && (ns == 0 || (s && diag)), "bad arguments");
  if (ns == 0) return QTX_OK;
  hipLaunchKernelGGL(ham_diag_kernel, dim3(blocks_for(ns, WAVES_PER_BLOCK)), dim3(256), 0, as_stream(stream), *H,
                     s, ns, diag);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_conn_count(void* stream, const qtx_ham* H, const int8_t* s, int64_t ns, int32_t* count) {
  int rc = check_ham(H);
  if (rc) return rc;
  QTX_REQUIRE(ns >= 0 && (ns == 0 || (s && count)), "bad arguments");
  if (ns == 0) return QTX_OK;
  hipLaunchKernelGGL(conn_count_kernel, dim3(blocks_for(ns, WAVES_PER_BLOCK)), dim3(256), 0, as_stream(stream), *H,
                     s, ns, count);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_conn_list(void* stream, const qtx_ham* H, const int8_t* s, int64_t ns, const int64_t* offsets,
                  int32_t* segment, int32_t* cand, int8_t* s_conn, double* h_conn) {
  int rc = check_ham(H);
  if (rc) return rc;
  QTX_REQUIRE(ns >= 0 && (ns == 0 || (s && offsets && segment && cand && s_conn && h_conn)), "bad arguments");
  if (ns == 0) return QTX_OK;
  hipLaunchKernelGGL(conn_list_kernel, dim3(blocks_for(ns, WAVES_PER_BLOCK)), dim3(256), 0, as_stream(stream), *H,
                     s, ns, offsets, segment, cand, s_conn, h_conn);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_rbm_pack(void* stream, int dtype, const void* theta, int32_t M, int32_t N, void* packed) {
  QTX_REQUIRE(M > 0 && M <= MAXN && N > 0 && N <= MAXN, "RBM M/N out of range (1..1024)");
  QTX_REQUIRE(theta && packed && theta != packed, "bad pointers");
  int64_t total = (int64_t)M * N + M;
  dim3 g(blocks_for(total, 256)), bl(256);
  if (dtype == QTX_F32)
    hipLaunchKernelGGL(rbm_pack_kernel<float>, g, bl, 0, as_stream(stream), (const float*)theta, M, N,
                       (float*)packed);
  else if (dtype == QTX_F64)
    hipLaunchKernelGGL(rbm_pack_kernel<double>, g, bl, 0, as_stream(stream), (const double*)theta, M, N,
                       (double*)packed);
  else
    return fail(QTX_EINVAL, "dtype must be QTX_F32 or QTX_F64");
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

}  // extern "C"

// run CALL with a compile-time RR = next power of two >= ceil(M / 64)
#define QTX_WITH_R(Rrt, ...)                 \
  do {                                       \
    if ((Rrt) <= 1) {                        \
      constexpr int RR = 1;                  \
      __VA_ARGS__;                           \
    } else if ((Rrt) <= 2) {                 \
      constexpr int RR = 2;                  \
      __VA_ARGS__;                           \
    } else if ((Rrt) <= 4) {                 \
      constexpr int RR = 4;                  \
      __VA_ARGS__;                           \
    } else if ((Rrt) <= 8) {                 \
      constexpr int RR = 8;                  \
      __VA_ARGS__;                           \
    } else {                                 \
      constexpr int RR = 16;                 \
      __VA_ARGS__;                           \
    }                                        \
  } while (0)

static int check_rbm(int dtype, const void* packed, int M, int N) {
  QTX_REQUIRE(dtype == QTX_F32 || dtype == QTX_F64, "dtype must be QTX_F32 or QTX_F64");
  QTX_REQUIRE(packed != nullptr, "null packed parameters");
  QTX_REQUIRE(M > 0 && M <= MAXN && N > 0 && N <= MAXN, "RBM M/N out of range (1..1024)");
  return QTX_OK;
}

template <typename T>
static void launch_rbm_logpsi(hipStream_t st, const void* packed, int M, int N, const int8_t* s, int64_t ns,
                              double* logabs, void* h_out) {
  dim3 g(blocks_for(ns, WAVES_PER_BLOCK)), bl(256);
  int R = (M + WAVE - 1) / WAVE;
  QTX_WITH_R(R, hipLaunchKernelGGL((rbm_logpsi_kernel<T, RR>), g, bl, 0, st, (const T*)packed, M, N, s, ns,
                                   logabs, (T*)h_out));
}

template <typename T>
static void launch_rbm_metropolis(hipStream_t st, const void* packed, int M, int N, int kind, const int32_t* nbr,
                                  int width, int hopping, const uint32_t* words, int nsteps, int64_t ns,
                                  double reweight, double tie_tol, int8_t* s, double* logabs, int32_t* n_accept,
                                  int32_t* n_near) {
  dim3 g(blocks_for(ns, WAVES_PER_BLOCK)), bl(256);
  int R = (M + WAVE - 1) / WAVE;
  QTX_WITH_R(R, hipLaunchKernelGGL((rbm_metropolis_kernel<T, RR>), g, bl, 0, st, (const T*)packed, M, N, kind,
                                   nbr, width, hopping, words, nsteps, ns, reweight, tie_tol, s, logabs, n_accept,
                                   n_near));
}

template <typename T>
static void launch_rbm_eloc(hipStream_t st, const void* packed, int M, int N, const qtx_ham& H, const int8_t* s,
                            int64_t ns, double* eloc, int32_t* nconn) {
  dim3 g(blocks_for(ns, WAVES_PER_BLOCK)), bl(256);
  int R = (M + WAVE - 1) / WAVE;
  QTX_WITH_R(R, hipLaunchKernelGGL((rbm_eloc_kernel<T, RR>), g, bl, 0, st, (const T*)packed, M, N, H, s, ns, eloc,
                                   nconn));
}

extern "C" {

int qtx_rbm_logpsi(void* stream, int dtype, const void* packed, int32_t M, int32_t N, const int8_t* s,
                   int64_t ns, double* logabs, void* h_out) {
  int rc = check_rbm(dtype, packed, M, N);
  if (rc) return rc;
  QTX_REQUIRE(ns >= 0 && (ns == 0 || (s && logabs)), "bad arguments");
  if (ns == 0) return QTX_OK;
  if (dtype == QTX_F32)
    launch_rbm_logpsi<float>(as_stream(stream), packed, M, N, s, ns, logabs, h_out);
  else
    launch_rbm_logpsi<double>(as_stream(stream), packed, M, N, s, ns, logabs, h_out);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_rbm_metropolis(void* stream, int dtype, const void* packed, int32_t M, int32_t N, int kind,
                       const int32_t* nbr, int32_t width, int32_t hopping, const uint32_t* words, int32_t nsteps,
                       int64_t ns, double reweight, double tie_tol, int8_t* s, double* logabs, int32_t* n_accept,
                       int32_t* n_near) {
  int rc = check_rbm(dtype, packed, M, N);
  if (rc) return rc;
  QTX_REQUIRE(kind == QTX_LOCALFLIP || kind == QTX_EXCHANGE, "kind must be QTX_LOCALFLIP or QTX_EXCHANGE");
  QTX_REQUIRE(kind == QTX_LOCALFLIP || (nbr != nullptr && width > 0), "EXCHANGE needs a neighbour table");
  QTX_REQUIRE(hopping == 1 || hopping == -1, "hopping must be +1 or -1");
  QTX_REQUIRE(nsteps >= 0 && ns >= 0, "negative size");
  QTX_REQUIRE(ns == 0 || (s && logabs && n_accept && n_near && (nsteps == 0 || words)), "null pointer");
  if (ns == 0) return QTX_OK;
  if (dtype == QTX_F32)
    launch_rbm_metropolis<float>(as_stream(stream), packed, M, N, kind, nbr, width, hopping, words, nsteps, ns,
                                 reweight, tie_tol, s, logabs, n_accept, n_near);
  else
    launch_rbm_metropolis<double>(as_stream(stream), packed, M, N, kind, nbr, width, hopping, words, nsteps, ns,
                                  reweight, tie_tol, s, logabs, n_accept, n_near);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_rbm_eloc(void* stream, int dtype, const void* packed, int32_t M, int32_t N, const qtx_ham* H,
                 const int8_t* s, int64_t ns, double* eloc, int32_t* nconn) {
  int rc = check_rbm(dtype, packed, M, N);
  if (rc) return rc;
  rc = check_ham(H);
  if (rc) return rc;
  QTX_REQUIRE(H->N == N, "qtx_ham.N != RBM N");
  QTX_REQUIRE(ns >= 0 && (ns == 0 || (s && eloc)), "bad arguments");
  if (ns == 0) return QTX_OK;
  if (dtype == QTX_F32)
    launch_rbm_eloc<float>(as_stream(stream), packed, M, N, *H, s, ns, eloc, nconn);
  else
    launch_rbm_eloc<double>(as_stream(stream), packed, M, N, *H, s, ns, eloc, nconn);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_rbm_jacobian(void* stream, int dtype, const void* packed, int32_t M, int32_t N, const int8_t* s,
                     int64_t ns, int o_dtype, void* O) {
  int rc = check_rbm(dtype, packed, M, N);
  if (rc) return rc;
  QTX_REQUIRE(o_dtype == QTX_F32 || o_dtype == QTX_F64, "o_dtype must be QTX_F32 or QTX_F64");
  QTX_REQUIRE(ns >= 0 && (ns == 0 || (s && O)), "bad arguments");
  if (ns == 0) return QTX_OK;
  QTX_REQUIRE(ns <= 0x7fffffff, "ns too large for one launch");
  dim3 g((unsigned)ns), bl(256);
  hipStream_t st = as_stream(stream);
  if (dtype == QTX_F32 && o_dtype == QTX_F32)
    hipLaunchKernelGGL((rbm_jacobian_kernel<float, float>), g, bl, 0, st, (const float*)packed, M, N, s, ns,
                       (float*)O);
  else if (dtype == QTX_F32)
    hipLaunchKernelGGL((rbm_jacobian_kernel<float, double>), g, bl, 0, st, (const float*)packed, M, N, s, ns,
                       (double*)O);
  else if (o_dtype == QTX_F32)
    hipLaunchKernelGGL((rbm_jacobian_kernel<double, float>), g, bl, 0, st, (const double*)packed, M, N, s, ns,
                       (float*)O);
  else
    hipLaunchKernelGGL((rbm_jacobian_kernel<double, double>), g, bl, 0, st, (const double*)packed, M, N, s, ns,
                       (double*)O);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

}  // extern "C"
