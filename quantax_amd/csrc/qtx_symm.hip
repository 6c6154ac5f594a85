// Symmetry projection of any wavefunction on the device, and the generic
// Metropolis step and local-energy combination it needs.
//
// psi_sym(s) = sum_g c_g psi(T_g s), c_g = chi_g chi_0 / |G|
// (symmetry/symmetry.py:343-392); the images T_g s = s[perm_g] and, with spin
// inversion, -s[perm_g] after them (symmetry.py:324-341).  A symmetrised state
// has no fused local-update kernel, so a Metropolis step is propose (one
// wavefront per chain, the same proposal code as the fused kernels) ->
// images -> model forward -> symmetrise -> accept.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "common.hpp"

namespace qtx {

constexpr int MAXN_SYMM = 1024;  // sites per configuration (LDS row of the proposal kernel)

// out[(s * nimg + g) * N + i] = (g < nperm ? 1 : -1) * s[s * N + perm[(g mod nperm) * N + i]]
__global__ void symm_images_kernel(const int8_t* __restrict__ s, int64_t ns, int N, const int32_t* __restrict__ perm,
                                   int nperm, int nimg, int8_t* __restrict__ out) {
  const int64_t row = blockIdx.x;  // sample * nimg + image
  if (row >= ns * nimg) return;
  const int64_t smp = row / nimg;
  const int g = (int)(row % nimg);
  const int gp = g % nperm;
  const int8_t sg = (g < nperm) ? 1 : -1;
  const int8_t* src = s + smp * N;
  const int32_t* p = perm + (int64_t)gp * N;
  int8_t* dst = out + row * N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = (int8_t)(sg * src[p[i]]);
}

// one wavefront per sample: psi_sym = sum_g c_g sign_g exp(l_g), in log form,
// and (optionally) w_g = c_g sign_g exp(l_g) / psi_sym, the weight of image g
// in d log|psi_sym| / d theta
__global__ __launch_bounds__(256) void symmetrize_kernel(const double* __restrict__ la, const double* __restrict__ sg_in,
                                                         const double* __restrict__ coef, int64_t ns, int nimg,
                                                         double* __restrict__ sg_out, double* __restrict__ la_out,
                                                         double* __restrict__ w) {
  const int lane = threadIdx.x % WAVE;
  const int64_t smp = (int64_t)blockIdx.x * 4 + threadIdx.x / WAVE;
  if (smp >= ns) return;
  const double* l = la + smp * nimg;
  double m = -INFINITY;
  for (int g = lane; g < nimg; g += WAVE)
    if (coef[g] != 0.0) m = fmax(m, l[g]);
  m = wave_max(m);
  double acc = 0.0;
  if (m > -INFINITY) {
    for (int g = lane; g < nimg; g += WAVE) {
      const double sgn = sg_in ? sg_in[smp * nimg + g] : 1.0;
      acc += coef[g] * sgn * exp(l[g] - m);
    }
  }
  const double S = wave_sum(acc);
  if (lane == 0) {
    sg_out[smp] = (S > 0.0) ? 1.0 : (S < 0.0 ? -1.0 : 0.0);
    la_out[smp] = (S != 0.0) ? m + log(fabs(S)) : -INFINITY;
  }
  if (w) {
    for (int g = lane; g < nimg; g += WAVE) {
      const double sgn = sg_in ? sg_in[smp * nimg + g] : 1.0;
      w[smp * nimg + g] = (S != 0.0) ? coef[g] * sgn * exp(l[g] - m) / S : 0.0;
    }
  }
}

// O[s, k] = sum_g w[s, g] Oimg[s * nimg + g, k]  (FP64 accumulation)
template <typename T>
__global__ void rows_combine_kernel(const T* __restrict__ Oimg, const double* __restrict__ w, int64_t ns, int nimg,
                                    int64_t np, T* __restrict__ O) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t smp = blockIdx.y;
  if (k >= np) return;
  double acc = 0.0;
  for (int g = 0; g < nimg; ++g) acc += w[smp * nimg + g] * (double)Oimg[(smp * nimg + g) * np + k];
  O[smp * np + k] = (T)acc;
}

// one wavefront per chain: s' = s with the proposal of this step applied
// (metropolis.py:291-322 requires s' != s for acceptance: updated = 0 otherwise)
__global__ __launch_bounds__(256) void propose_kernel(int kind, const int32_t* __restrict__ nbr, int width, int hopping,
                                                      const uint4* __restrict__ words, const int8_t* __restrict__ s,
                                                      int64_t ns, int N, int8_t* __restrict__ s_prop,
                                                      int32_t* __restrict__ updated) {
  __shared__ int8_t sh[4][MAXN_SYMM];
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const int64_t c = (int64_t)blockIdx.x * 4 + wv;
  if (c >= ns) return;
  int8_t* sl = sh[wv];
  for (int i = lane; i < N; i += WAVE) sl[i] = s[c * N + i];
  __builtin_amdgcn_wave_barrier();
  int n_hop = 0;
  if (kind == QTX_EXCHANGE) {
    int cnt = 0;
    for (int i = lane; i < N; i += WAVE) cnt += (sl[i] == hopping);
    n_hop = wave_sum(cnt);
  }
  const uint4 w4 = words[c];
  int fa, fb;
  bool upd;
  propose_move(kind, sl, N, nbr, width, hopping, n_hop, w4.x, w4.y, lane, fa, fb, upd);
  for (int i = lane; i < N; i += WAVE) {
    int8_t v = sl[i];
    if (upd && (i == fa || i == fb)) v = (int8_t)(-v);
    s_prop[c * N + i] = v;
  }
  if (lane == 0) updated[c] = upd ? 1 : 0;
}

// one wavefront per chain: accept iff |psi'/psi|^reweight > 1 - u or psi = 0
// (log domain: reweight (l' - l) > log1p(-u) or l = -inf), only for s' != s
__global__ __launch_bounds__(256) void accept_kernel(const uint4* __restrict__ words, const int8_t* __restrict__ s_prop,
                                                     const double* __restrict__ lp_prop,
                                                     const double* __restrict__ sg_prop,
                                                     const int32_t* __restrict__ updated, int64_t ns, int N,
                                                     double reweight, double tie_tol, int8_t* __restrict__ s,
                                                     double* __restrict__ lp, double* __restrict__ sg,
                                                     int32_t* __restrict__ n_accept, int32_t* __restrict__ n_near) {
  const int lane = threadIdx.x % WAVE;
  const int64_t c = (int64_t)blockIdx.x * 4 + threadIdx.x / WAVE;
  if (c >= ns || !updated[c]) return;
  const uint4 w4 = words[c];
  const double thr = log1p(-words_to_uniform(w4.z, w4.w));
  const double l0 = lp[c];
  const double d = reweight * (lp_prop[c] - l0);
  const bool acc = (d > thr) || (l0 == -INFINITY);
  if (acc)
    for (int i = lane; i < N; i += WAVE) s[c * N + i] = s_prop[c * N + i];
  if (lane == 0) {
    if (fabs(d - thr) < tie_tol) n_near[c] += 1;
    if (acc) {
      lp[c] = lp_prop[c];
      if (sg) sg[c] = sg_prop ? sg_prop[c] : 1.0;
      n_accept[c] += 1;
    }
  }
}

// one wavefront per sample: E_loc = diag + sum_k H_k (sign_k / sign) exp(l_k - l)
// over the sample's connected configurations [off[s], off[s+1])
__global__ __launch_bounds__(256) void conn_combine_kernel(const double* __restrict__ diag, const int64_t* __restrict__ off,
                                                           const double* __restrict__ hc, const double* __restrict__ sg_c,
                                                           const double* __restrict__ la_c, const double* __restrict__ sg,
                                                           const double* __restrict__ la, int64_t ns,
                                                           double* __restrict__ eloc) {
  const int lane = threadIdx.x % WAVE;
  const int64_t smp = (int64_t)blockIdx.x * 4 + threadIdx.x / WAVE;
  if (smp >= ns) return;
  const double l0 = la[smp];
  const double s0 = sg ? sg[smp] : 1.0;
  double acc = 0.0;
  for (int64_t k = off[smp] + lane; k < off[smp + 1]; k += WAVE)
    acc += hc[k] * (sg_c ? sg_c[k] : 1.0) * s0 * exp(la_c[k] - l0);
  acc = wave_sum(acc);
  if (lane == 0) eloc[smp] = diag[smp] + acc;
}

}  // namespace qtx

using namespace qtx;

static unsigned wave_blocks(int64_t n) { return (unsigned)((n + 3) / 4); }

extern "C" {

int qtx_symm_images(void* stream, const int8_t* s, int64_t ns, int32_t N, const int32_t* perm, int32_t nperm,
                    int32_t z2, int8_t* out) {
  QTX_REQUIRE(s && perm && out && ns >= 0 && N > 0 && nperm > 0, "bad arguments");
  const int nimg = nperm * (z2 ? 2 : 1);
  QTX_REQUIRE(ns * nimg < 0x7fffffffLL, "too many images in one call");
  if (ns == 0) return QTX_OK;
  hipLaunchKernelGGL(symm_images_kernel, dim3((unsigned)(ns * nimg)), dim3(N < 256 ? 64 : 256), 0,
                     as_stream(stream), s, ns, (int)N, perm, (int)nperm, nimg, out);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_symmetrize(void* stream, const double* logabs, const double* sign_in, const double* coef, int64_t ns,
                   int32_t nimg, double* sign, double* logabs_out, double* w) {
  QTX_REQUIRE(logabs && coef && sign && logabs_out && ns >= 0 && nimg > 0, "bad arguments");
  if (ns == 0) return QTX_OK;
  hipLaunchKernelGGL(symmetrize_kernel, dim3(wave_blocks(ns)), dim3(256), 0, as_stream(stream), logabs, sign_in, coef,
                     ns, (int)nimg, sign, logabs_out, w);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_rows_combine(void* stream, int o_dtype, const void* Oimg, const double* w, int64_t ns, int32_t nimg,
                     int64_t np, void* O) {
  QTX_REQUIRE(o_dtype == QTX_F32 || o_dtype == QTX_F64, "o_dtype must be QTX_F32 or QTX_F64");
  QTX_REQUIRE(Oimg && w && O && ns >= 0 && nimg > 0 && np > 0 && ns < 65536 * 32768LL, "bad arguments");
  if (ns == 0) return QTX_OK;
  hipStream_t st = as_stream(stream);
  for (int64_t s0 = 0; s0 < ns; s0 += 65535) {
    const int64_t n = (ns - s0) < 65535 ? (ns - s0) : 65535;
    dim3 g((unsigned)((np + 255) / 256), (unsigned)n);
    if (o_dtype == QTX_F32)
      hipLaunchKernelGGL(rows_combine_kernel<float>, g, dim3(256), 0, st, (const float*)Oimg + s0 * nimg * np,
                         w + s0 * nimg, n, (int)nimg, np, (float*)O + s0 * np);
    else
      hipLaunchKernelGGL(rows_combine_kernel<double>, g, dim3(256), 0, st, (const double*)Oimg + s0 * nimg * np,
                         w + s0 * nimg, n, (int)nimg, np, (double*)O + s0 * np);
  }
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_propose(void* stream, int kind, const int32_t* nbr, int32_t width, int32_t hopping, const uint32_t* words,
                const int8_t* s, int64_t ns, int32_t N, int8_t* s_prop, int32_t* updated) {
  QTX_REQUIRE(kind == QTX_LOCALFLIP || kind == QTX_EXCHANGE, "kind must be QTX_LOCALFLIP or QTX_EXCHANGE");
  QTX_REQUIRE(words && s && s_prop && updated && ns >= 0 && N > 0 && N <= MAXN_SYMM, "bad arguments (N <= 1024)");
  QTX_REQUIRE(kind == QTX_LOCALFLIP || (nbr && width > 0), "SpinExchange needs the neighbour table");
  if (ns == 0) return QTX_OK;
  hipLaunchKernelGGL(propose_kernel, dim3(wave_blocks(ns)), dim3(256), 0, as_stream(stream), kind, nbr, (int)width,
                     (int)hopping, (const uint4*)words, s, ns, (int)N, s_prop, updated);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_accept(void* stream, const uint32_t* words, const int8_t* s_prop, const double* lp_prop, const double* sg_prop,
               const int32_t* updated, int64_t ns, int32_t N, double reweight, double tie_tol, int8_t* s, double* lp,
               double* sg, int32_t* n_accept, int32_t* n_near) {
  QTX_REQUIRE(words && s_prop && lp_prop && updated && s && lp && n_accept && n_near && ns >= 0 && N > 0,
              "bad arguments");
  if (ns == 0) return QTX_OK;
  hipLaunchKernelGGL(accept_kernel, dim3(wave_blocks(ns)), dim3(256), 0, as_stream(stream), (const uint4*)words,
                     s_prop, lp_prop, sg_prop, updated, ns, (int)N, reweight, tie_tol, s, lp, sg, n_accept, n_near);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

int qtx_conn_combine(void* stream, const double* diag, const int64_t* off, const double* hc, const double* sign_c,
                     const double* logabs_c, const double* sign, const double* logabs, int64_t ns, double* eloc) {
  QTX_REQUIRE(diag && off && logabs && eloc && ns >= 0, "bad arguments");
  if (ns == 0) return QTX_OK;
  hipLaunchKernelGGL(conn_combine_kernel, dim3(wave_blocks(ns)), dim3(256), 0, as_stream(stream), diag, off, hc,
                     sign_c, logabs_c, sign, logabs, ns, eloc);
  QTX_LAUNCH_CHECK();
  return QTX_OK;
}

}  // extern "C"
