// RBM_Dense device bodies shared by the single-job kernels (qtx_core.hip) and
// the grouped serving kernels (qtx_serve.hip), so a chain evolves through the
// same instructions whichever launch carries it.
//
// packed = [Wt (N, M) | b (M)]: the column W[:, j] touched by a flip of site j
// is one contiguous row of Wt.  One wavefront owns one chain / sample; lane l
// holds hidden units l, l + 64, ... (R = ceil(M / 64) registers per lane).
#pragma once
#include "common.hpp"

namespace qtx {

constexpr int MAXN = 1024;
constexpr int WAVES_PER_BLOCK = 4;

// copy one sample's spins into LDS (wave-cooperative)
__device__ __forceinline__ void load_spins(int8_t* sl, const int8_t* src, int N, int lane) {
  for (int i = lane; i < N; i += WAVE) sl[i] = src[i];
}

// h for lane-owned hidden units: h[r] = b[i] + sum_j Wt[j, i] s_j,  i = lane + 64 r
template <typename T, int R>
__device__ __forceinline__ void rbm_hidden(const T* __restrict__ Wt, const T* __restrict__ b, int M, int N,
                                           const int8_t* sl, int lane, T (&h)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int i = lane + WAVE * r;
    h[r] = (i < M) ? b[i] : T(0);
  }
  for (int j = 0; j < N; ++j) {
    T sj = (T)sl[j];
    const T* row = Wt + (int64_t)j * M;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      int i = lane + WAVE * r;
      if (i < M) h[r] += row[i] * sj;
    }
  }
}

template <typename T, int R>
__device__ __forceinline__ double rbm_sum_logcosh(const T (&h)[R], int M, int lane) {
  T acc = T(0);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int i = lane + WAVE * r;
    if (i < M) acc += logcosh(h[r]);
  }
  return wave_sum((double)acc);
}

// log|psi| of one sample (and optionally its h row)
template <typename T, int R>
__device__ __forceinline__ void rbm_chain_logpsi(const T* __restrict__ packed, int M, int N, const int8_t* src,
                                                 int8_t* sl, int lane, double* logabs, T* h_out) {
  load_spins(sl, src, N, lane);
  __builtin_amdgcn_wave_barrier();
  const T* Wt = packed;
  const T* b = packed + (int64_t)M * N;
  T h[R];
  rbm_hidden<T, R>(Wt, b, M, N, sl, lane, h);
  double la = rbm_sum_logcosh<T, R>(h, M, lane);
  if (lane == 0) *logabs = la;
  if (h_out) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      int i = lane + WAVE * r;
      if (i < M) h_out[i] = h[r];
    }
  }
}

// nsteps Metropolis steps of one chain with h resident in registers.
// getw(t) -> uint4 Philox block of step t.  s is read and written back;
// *logabs receives the local-update value; counters are incremented.
template <typename T, int R, typename GetW>
__device__ __forceinline__ void rbm_chain_metropolis(const T* __restrict__ packed, int M, int N, int kind,
                                                     const int32_t* __restrict__ nbr, int width, int hopping,
                                                     int nsteps, GetW getw, double reweight, double tie_tol,
                                                     int8_t* s, int8_t* sl, int lane, double* logabs,
                                                     int32_t* n_accept, int32_t* n_near) {
  load_spins(sl, s, N, lane);
  __builtin_amdgcn_wave_barrier();
  const T* Wt = packed;
  const T* b = packed + (int64_t)M * N;
  T h[R];
  rbm_hidden<T, R>(Wt, b, M, N, sl, lane, h);
  double lp = rbm_sum_logcosh<T, R>(h, M, lane);
  int n_hop = 0;
  if (kind == QTX_EXCHANGE) {
    int cnt = 0;
    for (int i = lane; i < N; i += WAVE) cnt += (sl[i] == hopping);
    n_hop = wave_sum(cnt);
  }
  int acc_count = 0, near_count = 0;
  for (int t = 0; t < nsteps; ++t) {
    uint4 w = getw(t);
    int fa, fb;
    bool updated;
    propose_move(kind, sl, N, nbr, width, hopping, n_hop, w.x, w.y, lane, fa, fb, updated);
    if (!updated) continue;  // s' == s is never accepted (metropolis.py:314-316)
    T sa = (T)sl[fa];
    T sb = (fb >= 0) ? (T)sl[fb] : T(0);
    const T* ra = Wt + (int64_t)fa * M;
    const T* rb = Wt + (int64_t)(fb >= 0 ? fb : fa) * M;
    T hn[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      int i = lane + WAVE * r;
      T v = h[r];
      if (i < M) {
        v -= T(2) * ra[i] * sa;
        if (fb >= 0) v -= T(2) * rb[i] * sb;
      }
      hn[r] = v;
    }
    double lpn = rbm_sum_logcosh<T, R>(hn, M, lane);
    double u = words_to_uniform(w.z, w.w);
    double thr = log1p(-u);
    double d = reweight * (lpn - lp);
    bool accept = (d > thr) || (lp == -INFINITY);
    near_count += (fabs(d - thr) < tie_tol) ? 1 : 0;
    if (accept) {
      acc_count += 1;
#pragma unroll
      for (int r = 0; r < R; ++r) h[r] = hn[r];
      lp = lpn;
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) {
        sl[fa] = -sl[fa];
        if (fb >= 0) sl[fb] = -sl[fb];
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  __builtin_amdgcn_wave_barrier();
  for (int i = lane; i < N; i += WAVE) s[i] = sl[i];
  if (lane == 0) {
    *logabs = lp;
    *n_accept += acc_count;
    *n_near += near_count;
  }
}

// E_loc of one sample with local updates of h over the valid candidates
template <typename T, int R>
__device__ __forceinline__ void rbm_sample_eloc(const T* __restrict__ packed, int M, int N, const qtx_ham& H,
                                                const int8_t* src, int8_t* sl, int lane, double* eloc,
                                                int32_t* nconn) {
  load_spins(sl, src, N, lane);
  __builtin_amdgcn_wave_barrier();
  const T* Wt = packed;
  const T* b = packed + (int64_t)M * N;
  T h[R], lc0[R];
  rbm_hidden<T, R>(Wt, b, M, N, sl, lane, h);
#pragma unroll
  for (int r = 0; r < R; ++r) lc0[r] = (lane + WAVE * r < M) ? logcosh(h[r]) : T(0);
  double e = diag_energy(H, sl, lane);
  int cnt = 0;
  for (int k = 0; k < H.noff; ++k) {
    if (!cand_valid(H, sl, k)) continue;
    ++cnt;
    int fa = H.off_a[k], fb = H.off_b[k];
    T sa = (T)sl[fa];
    T sb = (fb >= 0) ? (T)sl[fb] : T(0);
    const T* ra = Wt + (int64_t)fa * M;
    const T* rb = Wt + (int64_t)(fb >= 0 ? fb : fa) * M;
    T dl = T(0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      int i = lane + WAVE * r;
      if (i < M) {
        T v = h[r] - T(2) * ra[i] * sa;
        if (fb >= 0) v -= T(2) * rb[i] * sb;
        dl += logcosh(v) - lc0[r];
      }
    }
    double dlp = wave_sum((double)dl);
    e += H.off_c[k] * exp(dlp);
  }
  if (lane == 0) {
    *eloc = e;
    if (nconn) *nconn = cnt;
  }
}

// one O row [tanh(h) (x) s, tanh(h)] by a whole block: tanh(h) into LDS, then a
// coalesced write of the row.  tl / tld / sl: block-shared scratch of MAXN.
template <typename T, typename TO>
__device__ __forceinline__ void rbm_jacobian_row(const T* __restrict__ packed, int M, int N, const int8_t* src,
                                                 TO* __restrict__ row, float* tl, double* tld, int8_t* sl) {
  const T* Wt = packed;
  const T* b = packed + (int64_t)M * N;
  for (int j = threadIdx.x; j < N; j += blockDim.x) sl[j] = src[j];
  __syncthreads();
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    T h = b[i];
    for (int j = 0; j < N; ++j) h += Wt[(int64_t)j * M + i] * (T)sl[j];
    T t = tanh(h);
    tl[i] = (float)t;
    tld[i] = (double)t;
  }
  __syncthreads();
  int64_t np = (int64_t)M * N + M;
  for (int64_t k = threadIdx.x; k < np; k += blockDim.x) {
    TO v;
    if (k < (int64_t)M * N) {
      int i = (int)(k / N), j = (int)(k % N);
      if (sizeof(TO) == 8)
        v = (TO)(tld[i] * (double)sl[j]);
      else
        v = (TO)(tl[i] * (float)sl[j]);
    } else {
      int i = (int)(k - (int64_t)M * N);
      v = (sizeof(TO) == 8) ? (TO)tld[i] : (TO)tl[i];
    }
    row[k] = v;
  }
  __syncthreads();  // the scratch is reused by the block's next row
}

}  // namespace qtx
