// Grouped RBM_Dense launches for the serving loop: many independent VMC jobs
// (different N, M, sample counts, couplings) advanced by a handful of
// persistent launches instead of one launch per job and stage.
//
// A stage call receives the jobs as a host array.  Jobs are split by register
// class R = ceil(M / 64) in {1, 2, 4, 8, 16}; each class is one launch whose
// virtual blocks (4 chains, one per wave; or one sample per block for the
// Jacobian) are dealt grid-stride over a persistent grid, so small and large
// jobs share the CUs.  The per-chain code is the single-job code
// (rbm_device.hpp) and every job's Philox stream is keyed by its own seed and
// step counter, so a job's chains and energies do not depend on which other
// jobs share the launch.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "rbm_device.hpp"

namespace qtx {

constexpr int G_LOGPSI = 0, G_METRO = 1, G_ELOC = 2, G_JAC = 3;
constexpr int G_CLASSES = 5;  // R = 1, 2, 4, 8, 16
constexpr int64_t G_MAX_GRID = 4096;

__host__ __device__ inline int g_class_of(int M) {
  int R = (M + WAVE - 1) / WAVE;
  return R <= 1 ? 0 : R <= 2 ? 1 : R <= 4 ? 2 : R <= 8 ? 3 : 4;
}

// the job slot j of a launch with vstart[j] <= v < vstart[j + 1]
__device__ __forceinline__ int g_find(const int64_t* __restrict__ vstart, int cnt, int64_t v) {
  int lo = 0, hi = cnt - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (vstart[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int R>
__global__ __launch_bounds__(256) void rbm_group_logpsi_kernel(const qtx_rbm_job* __restrict__ jobs,
                                                               const int32_t* __restrict__ idx,
                                                               const int64_t* __restrict__ vstart, int cnt,
                                                               int64_t nv) {
  __shared__ int8_t sbuf[WAVES_PER_BLOCK][MAXN];
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  for (int64_t v = blockIdx.x; v < nv; v += gridDim.x) {
    const int j = g_find(vstart, cnt, v);
    const qtx_rbm_job& J = jobs[idx[j]];
    const int64_t c = (v - vstart[j]) * WAVES_PER_BLOCK + wv;
    if (c < J.ns)
      rbm_chain_logpsi<float, R>(J.packed, J.M, J.N, J.s + c * J.N, sbuf[wv], lane, J.logabs + c, (float*)nullptr);
  }
}

template <int R>
__global__ __launch_bounds__(256) void rbm_group_metropolis_kernel(const qtx_rbm_job* __restrict__ jobs,
                                                                   const int32_t* __restrict__ idx,
                                                                   const int64_t* __restrict__ vstart, int cnt,
                                                                   int64_t nv) {
  __shared__ int8_t sbuf[WAVES_PER_BLOCK][MAXN];
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  for (int64_t v = blockIdx.x; v < nv; v += gridDim.x) {
    const int j = g_find(vstart, cnt, v);
    const qtx_rbm_job& J = jobs[idx[j]];
    const int64_t c = (v - vstart[j]) * WAVES_PER_BLOCK + wv;
    if (c >= J.ns) continue;
    const uint32_t k0 = (uint32_t)J.seed, k1 = (uint32_t)(J.seed >> 32);
    const int64_t step0 = J.step0;
    // the Philox block qtx_philox_fill(seed, TAG_SWEEP, step0, ...) stores for (step, chain)
    auto getw = [&](int t) {
      u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)(step0 + t), TAG_SWEEP, 0u}, k0, k1);
      return make_uint4(r.x, r.y, r.z, r.w);
    };
    rbm_chain_metropolis<float, R>(J.packed, J.M, J.N, J.kind, J.nbr, J.width, J.hopping, J.nsteps, getw,
                                   J.reweight, J.tie_tol, J.s + c * J.N, sbuf[wv], lane, J.logabs + c,
                                   J.n_accept + c, J.n_near + c);
  }
}

template <int R>
__global__ __launch_bounds__(256) void rbm_group_eloc_kernel(const qtx_rbm_job* __restrict__ jobs,
                                                             const int32_t* __restrict__ idx,
                                                             const int64_t* __restrict__ vstart, int cnt,
                                                             int64_t nv) {
  __shared__ int8_t sbuf[WAVES_PER_BLOCK][MAXN];
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  for (int64_t v = blockIdx.x; v < nv; v += gridDim.x) {
    const int j = g_find(vstart, cnt, v);
    const qtx_rbm_job& J = jobs[idx[j]];
    const int64_t n = (v - vstart[j]) * WAVES_PER_BLOCK + wv;
    if (n < J.ns)
      rbm_sample_eloc<float, R>(J.packed, J.M, J.N, J.H, J.s + n * J.N, sbuf[wv], lane, J.eloc + n,
                                J.nconn ? J.nconn + n : nullptr);
  }
}

__global__ __launch_bounds__(256) void rbm_group_jacobian_kernel(const qtx_rbm_job* __restrict__ jobs,
                                                                 const int32_t* __restrict__ idx,
                                                                 const int64_t* __restrict__ vstart, int cnt,
                                                                 int64_t nv) {
  __shared__ float tl[MAXN];
  __shared__ double tld[MAXN];
  __shared__ int8_t sl[MAXN];
  for (int64_t v = blockIdx.x; v < nv; v += gridDim.x) {
    const int j = g_find(vstart, cnt, v);
    const qtx_rbm_job& J = jobs[idx[j]];
    const int64_t n = v - vstart[j];
    rbm_jacobian_row<float, float>(J.packed, J.M, J.N, J.s + n * J.N,
                                   J.O + n * ((int64_t)J.M * J.N + J.M), tl, tld, sl);
  }
}

// ---------------------------------------------------------------- host side
// per-thread pinned staging of the job table; the event guards its reuse
struct GroupStaging {
  char* host = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  int device = -1;
  bool pending = false;
};
static thread_local GroupStaging g_stage;

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t group_ws_bytes(int njobs) {
  return align256((size_t)njobs * sizeof(qtx_rbm_job)) + align256((size_t)njobs * sizeof(int32_t)) +
         align256((size_t)(njobs + G_CLASSES) * sizeof(int64_t));
}

static int check_job(const qtx_rbm_job& J, int op, int j) {
  const std::string at = " (job " + std::to_string(j) + ")";
  if (J.packed == nullptr) return fail(QTX_EINVAL, "null packed parameters" + at);
  if (J.M < 1 || J.M > MAXN || J.N < 1 || J.N > MAXN) return fail(QTX_EINVAL, "RBM M/N out of range (1..1024)" + at);
  if (J.ns < 0) return fail(QTX_EINVAL, "negative ns" + at);
  if (J.ns > 0 && J.s == nullptr) return fail(QTX_EINVAL, "null spins" + at);
  if (J.ns == 0) return QTX_OK;
  if (op == G_LOGPSI && J.logabs == nullptr) return fail(QTX_EINVAL, "null logabs" + at);
  if (op == G_METRO) {
    if (J.nsteps < 0) return fail(QTX_EINVAL, "negative nsteps" + at);
    if (J.kind != QTX_LOCALFLIP && J.kind != QTX_EXCHANGE) return fail(QTX_EINVAL, "unknown proposal kind" + at);
    if (J.kind == QTX_EXCHANGE && (J.nbr == nullptr || J.width < 1 || (J.hopping != 1 && J.hopping != -1)))
      return fail(QTX_EINVAL, "EXCHANGE needs nbr, width >= 1 and hopping = +-1" + at);
    if (!J.logabs || !J.n_accept || !J.n_near) return fail(QTX_EINVAL, "null Metropolis output" + at);
  }
  if (op == G_ELOC) {
    if (J.eloc == nullptr) return fail(QTX_EINVAL, "null eloc" + at);
    if (J.H.N != J.N) return fail(QTX_EINVAL, "qtx_ham.N != N" + at);
    if (J.H.noff > 0 && !(J.H.off_a && J.H.off_b && J.H.off_ra && J.H.off_rb && J.H.off_c))
      return fail(QTX_EINVAL, "null off-diagonal table" + at);
    if (J.H.ndiag > 0 && !(J.H.diag_c && J.H.diag_a && J.H.diag_b && J.H.diag_kind))
      return fail(QTX_EINVAL, "null diagonal table" + at);
  }
  if (op == G_JAC && J.O == nullptr) return fail(QTX_EINVAL, "null O" + at);
  return QTX_OK;
}

static int group_call(void* stream, const qtx_rbm_job* jobs, int njobs, void* ws, size_t ws_bytes, int op) {
  QTX_REQUIRE(njobs >= 0 && (njobs == 0 || jobs != nullptr), "bad job array");
  if (njobs == 0) return QTX_OK;
  const size_t need = group_ws_bytes(njobs);
  if (ws == nullptr || ws_bytes < need)
    return fail(QTX_EWORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  for (int j = 0; j < njobs; ++j) {
    int rc = check_job(jobs[j], op, j);
    if (rc) return rc;
  }
  // partition: slots of class k are [cstart[k], cstart[k] + ccount[k]); the
  // Jacobian is block-per-sample and uses one class
  std::vector<int> members[G_CLASSES];
  for (int j = 0; j < njobs; ++j) {
    const qtx_rbm_job& J = jobs[j];
    if (J.ns == 0 || (op == G_METRO && J.nsteps == 0)) continue;
    members[op == G_JAC ? 0 : g_class_of(J.M)].push_back(j);
  }
  hipStream_t st = as_stream(stream);
  int dev = 0;
  QTX_HIP(hipGetDevice(&dev));
  GroupStaging& S = g_stage;
  if (S.ev != nullptr && S.device != dev) {
    if (S.pending) QTX_HIP(hipEventSynchronize(S.ev));
    QTX_HIP(hipEventDestroy(S.ev));
    S.ev = nullptr;
    S.pending = false;
  }
  if (S.ev == nullptr) {
    QTX_HIP(hipEventCreateWithFlags(&S.ev, hipEventDisableTiming));
    S.device = dev;
  }
  if (S.pending) {  // the previous copy out of the staging buffer has finished
    QTX_HIP(hipEventSynchronize(S.ev));
    S.pending = false;
  }
  if (S.cap < need) {
    if (S.host) QTX_HIP(hipHostFree(S.host));
    S.host = nullptr;
    S.cap = 0;
    QTX_HIP(hipHostMalloc((void**)&S.host, need, hipHostMallocDefault));
    S.cap = need;
  }
  const size_t off_idx = align256((size_t)njobs * sizeof(qtx_rbm_job));
  const size_t off_vs = off_idx + align256((size_t)njobs * sizeof(int32_t));
  std::memcpy(S.host, jobs, (size_t)njobs * sizeof(qtx_rbm_job));
  int32_t* idx = reinterpret_cast<int32_t*>(S.host + off_idx);
  int64_t* vs = reinterpret_cast<int64_t*>(S.host + off_vs);
  int cstart[G_CLASSES], vstart_pos[G_CLASSES];
  int64_t nvk[G_CLASSES];
  int slot = 0, vpos = 0;
  for (int k = 0; k < G_CLASSES; ++k) {
    cstart[k] = slot;
    vstart_pos[k] = vpos;
    int64_t acc = 0;
    for (int j : members[k]) {
      idx[slot++] = j;
      vs[vpos++] = acc;
      const int64_t ns = jobs[j].ns;
      acc += (op == G_JAC) ? ns : (ns + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    }
    vs[vpos++] = acc;  // sentinel
    nvk[k] = acc;
  }
  char* wsb = static_cast<char*>(ws);
  QTX_HIP(hipMemcpyAsync(wsb, S.host, need, hipMemcpyHostToDevice, st));
  QTX_HIP(hipEventRecord(S.ev, st));
  S.pending = true;
  const qtx_rbm_job* djobs = reinterpret_cast<const qtx_rbm_job*>(wsb);
  const int32_t* didx = reinterpret_cast<const int32_t*>(wsb + off_idx);
  const int64_t* dvs = reinterpret_cast<const int64_t*>(wsb + off_vs);
  for (int k = 0; k < G_CLASSES; ++k) {
    const int cnt = (int)members[k].size();
    if (cnt == 0 || nvk[k] == 0) continue;
    const int32_t* ci = didx + cstart[k];
    const int64_t* cv = dvs + vstart_pos[k];
    const int64_t nv = nvk[k];
    dim3 grid((unsigned)(nv < G_MAX_GRID ? nv : G_MAX_GRID)), block(256);
#define G_LAUNCH(KERNEL)                                                                            \
  do {                                                                                              \
    switch (k) {                                                                                    \
      case 0: hipLaunchKernelGGL((KERNEL<1>), grid, block, 0, st, djobs, ci, cv, cnt, nv); break;   \
      case 1: hipLaunchKernelGGL((KERNEL<2>), grid, block, 0, st, djobs, ci, cv, cnt, nv); break;   \
      case 2: hipLaunchKernelGGL((KERNEL<4>), grid, block, 0, st, djobs, ci, cv, cnt, nv); break;   \
      case 3: hipLaunchKernelGGL((KERNEL<8>), grid, block, 0, st, djobs, ci, cv, cnt, nv); break;   \
      default: hipLaunchKernelGGL((KERNEL<16>), grid, block, 0, st, djobs, ci, cv, cnt, nv); break; \
    }                                                                                               \
  } while (0)
    if (op == G_LOGPSI) G_LAUNCH(rbm_group_logpsi_kernel);
    else if (op == G_METRO) G_LAUNCH(rbm_group_metropolis_kernel);
    else if (op == G_ELOC) G_LAUNCH(rbm_group_eloc_kernel);
    else hipLaunchKernelGGL(rbm_group_jacobian_kernel, grid, block, 0, st, djobs, ci, cv, cnt, nv);
#undef G_LAUNCH
    QTX_LAUNCH_CHECK();
  }
  return QTX_OK;
}

}  // namespace qtx

using namespace qtx;

extern "C" {

size_t qtx_rbm_group_ws_bytes(int32_t njobs) { return njobs > 0 ? group_ws_bytes(njobs) : 0; }

int qtx_rbm_group_logpsi(void* stream, const qtx_rbm_job* jobs, int32_t njobs, void* ws, size_t ws_bytes) {
  return group_call(stream, jobs, njobs, ws, ws_bytes, G_LOGPSI);
}

int qtx_rbm_group_metropolis(void* stream, const qtx_rbm_job* jobs, int32_t njobs, void* ws, size_t ws_bytes) {
  return group_call(stream, jobs, njobs, ws, ws_bytes, G_METRO);
}

int qtx_rbm_group_eloc(void* stream, const qtx_rbm_job* jobs, int32_t njobs, void* ws, size_t ws_bytes) {
  return group_call(stream, jobs, njobs, ws, ws_bytes, G_ELOC);
}

int qtx_rbm_group_jacobian(void* stream, const qtx_rbm_job* jobs, int32_t njobs, void* ws, size_t ws_bytes) {
  return group_call(stream, jobs, njobs, ws, ws_bytes, G_JAC);
}

}  // extern "C"
