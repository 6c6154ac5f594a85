/*
 * qtx_hip.h — C ABI of the MI355X (gfx950) VMC-step library (libqtxhip.so).
 *
 * Every entry point:
 *   - takes device pointers (unless stated otherwise), plain sizes and a HIP
 *     stream handle passed as `void*` (NULL = the default stream);
 *   - enqueues its work on that stream and returns 0 on success or a negative
 *     QTX_E* code; `qtx_last_error()` returns the thread-local message;
 *   - is reentrant: lattices, operators, models and random seeds are explicit
 *     arguments (the reference keeps them in process globals).
 *
 * Spins are int8 +-1, row-major (ns, N).  Log-amplitudes are float64.  The
 * sign of psi is +1 for both RBM_Dense (cosh > 0) and ResConv (exp head).
 * Parameter vectors are in the reference's ravel_pytree order.
 */
#ifndef QTX_HIP_H
#define QTX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define QTX_OK 0
#define QTX_EINVAL (-1)   /* bad argument (shape, size, enum, null pointer) */
#define QTX_EHIP (-2)     /* HIP runtime error */
#define QTX_ESOLVER (-3)  /* rocSOLVER/rocBLAS failure or non-positive pivot */
#define QTX_EWORKSPACE (-4) /* workspace too small */

/* dtypes */
#define QTX_F32 0
#define QTX_F64 1

/* proposal kinds */
#define QTX_LOCALFLIP 0
#define QTX_EXCHANGE 1

/* Gram precision for the SR/MinSR solve */
#define QTX_GRAM_FP64 0   /* FP64 products on FP64 MFMA                     */
#define QTX_GRAM_FP32 1   /* FP32 products on FP32 MFMA                     */
#define QTX_GRAM_BF16X3 2 /* split-bf16: hi*hi + hi*lo + lo*hi, FP32 accum. */

/* solver modes */
#define QTX_SOLVE_CHOL_SHIFT 0        /* (T + shift I) y = b by Cholesky                */
#define QTX_SOLVE_EIGH_PINV 1         /* soft pseudo-inverse of T, fastest exact route  */
#define QTX_SOLVE_EIGH_PINV_DIRECT 2  /* the same soft pseudo-inverse, always by dsyevd */
/* flags OR'd into a solver mode: the branch of the minnorm_* (T = Obar Obar^T)
 * or lstsq_* (S = Obar^T Obar) factories whatever the shape (default: MinSR
 * when ns < np, else SR, as auto_*), and the workspace for SNR truncation */
#define QTX_SOLVE_MINNORM 0x100
#define QTX_SOLVE_LSTSQ 0x200
#define QTX_SOLVE_SNR 0x400
/* flag: the workspace is qtx_minsr_ws_bytes(...) bytes for this library
 * followed by the rocBLAS / rocSOLVER scratch of the solve (at least
 * qtx_minsr_blas_ws_bytes(...) bytes): the library's handles then allocate
 * nothing.  Without it that scratch is allocated by rocBLAS, once per
 * (thread, device, stream) handle, and kept. */
#define QTX_SOLVE_BLAS_WS 0x800
/* flag: EIGH_PINV takes the Chebyshev route whenever the cutoff allows it (c
 * >= 1e-5 lambda_max), whatever n; without it that route is taken from
 * n = 2048 on.  The route: one Cholesky factorisation of T + c I, its
 * triangular inverse P, then the soft inverse as a Chebyshev series in
 * (T + c I)^-1 T = I - c P^T P (about 158 terms at 1e-15 of the largest
 * coefficient). */
#define QTX_SOLVE_CHEB 0x1000

/* ResConv forward precision */
#define QTX_CONV_FP32 0   /* FP32 products on FP32 MFMA */
#define QTX_CONV_BF16X3 1 /* split-bf16 products on bf16 MFMA, FP32 accum. */

const char* qtx_last_error(void);
int qtx_device_sync(void);

/* ---------------------------------------------------------------- streams
 * Philox4x32-10 blocks: words[(t * nchains + c) * 4 + w] for step t in
 * [0, nsteps), chain c in [0, nchains); counter (chain0 + c, step0 + t, tag, 0),
 * key (seed & 0xffffffff, seed >> 32).
 * Replaces the host key splitting of quantax/global_defs.py:75-91 and the
 * draws of sampler/common_samplers.py:31,65-73 / sampler/metropolis.py:303. */
int qtx_philox_fill(void* stream, uint64_t seed, uint32_t tag, int64_t step0, int64_t nsteps,
                    int64_t chain0, int64_t nchains, uint32_t* words);

/* Initial chains: uniform +-1 (nup < 0) or a uniform permutation of nup up
 * spins (Fisher-Yates on Philox draws).  Replaces utils/basis.py:160-219. */
int qtx_init_spins(void* stream, uint64_t seed, int64_t nchains, int32_t N, int32_t nup, int8_t* s);

/* ---------------------------------------------------------------- operator
 * Compiled spin Hamiltonian (tables are device arrays).  Replaces the
 * jax_op_list of quantax/operator/operator.py:219-236.
 * diag term t : kind 0 -> c*s[a]*s[b], 1 -> c*s[a], 2 -> c.
 * candidate k : flips site a (and b if b >= 0); valid iff (ra == 0 || s[a] == ra)
 *               and (b < 0 || rb == 0 || s[b] == rb); matrix element c. */
typedef struct qtx_ham {
  int32_t N;
  int32_t ndiag;
  const double* diag_c;
  const int32_t* diag_a;
  const int32_t* diag_b;
  const int32_t* diag_kind;
  int32_t noff;
  const int32_t* off_a;
  const int32_t* off_b;
  const int8_t* off_ra;
  const int8_t* off_rb;
  const double* off_c;
} qtx_ham;

/* Diagonal part of H on each sample (operator.py:81-93). */
int qtx_ham_diag(void* stream, const qtx_ham* H, const int8_t* s, int64_t ns, double* diag);
/* Number of valid connected configurations per sample (operator.py:122-141). */
int qtx_conn_count(void* stream, const qtx_ham* H, const int8_t* s, int64_t ns, int32_t* count);
/* Compacted connected configurations, sample-major, candidate order, at the
 * exclusive-scan offsets of qtx_conn_count (operator.py:96-119, 144-165).
 * Outputs: segment[K], cand[K], s_conn[K*N], h_conn[K]. */
int qtx_conn_list(void* stream, const qtx_ham* H, const int8_t* s, int64_t ns, const int64_t* offsets,
                  int32_t* segment, int32_t* cand, int8_t* s_conn, double* h_conn);

/* ---------------------------------------------------------------- RBM_Dense
 * Packed parameters: W transposed to (N, M) then b (M), in `dtype`.
 * Build it from the ravel_pytree vector [W (M, N) row-major, b (M)] with
 * qtx_rbm_pack (model/shallow_nets.py:35-126).  M <= 1024, N <= 1024. */
int qtx_rbm_pack(void* stream, int dtype, const void* theta, int32_t M, int32_t N, void* packed);

/* log|psi| = sum_i log cosh(h_i), h = W s + b; optionally h (ns, M) in dtype.
 * Replaces Variational.__call__ / init_internal for SingleDense
 * (state/variational.py:325-356, shallow_nets.py:80-85). */
int qtx_rbm_logpsi(void* stream, int dtype, const void* packed, int32_t M, int32_t N, const int8_t* s,
                   int64_t ns, double* logabs, void* h_out);

/* nsteps Metropolis steps for all ns chains with h resident on chip.
 * words: (nsteps, ns, 4) Philox blocks; nbr (N, width) -1-padded neighbour
 * table (EXCHANGE only); hopping = +1 or -1 species moved by EXCHANGE.
 * In/out: s.  Out: logabs (local-update value).  n_accept / n_near (ns,)
 * are incremented (acceptances; decisions with |d - log1p(-u)| < tie_tol).
 * Replaces Metropolis._partial_sweep/_single_sweep/_update with LocalFlip or
 * SpinExchange proposals and SingleDense.ref_forward (metropolis.py:246-322,
 * common_samplers.py:14-162, shallow_nets.py:87-108). */
int qtx_rbm_metropolis(void* stream, int dtype, const void* packed, int32_t M, int32_t N, int kind,
                       const int32_t* nbr, int32_t width, int32_t hopping, const uint32_t* words,
                       int32_t nsteps, int64_t ns, double reweight, double tie_tol, int8_t* s,
                       double* logabs, int32_t* n_accept, int32_t* n_near);

/* Local energy E_loc(s) = diag + sum_k c_k psi(s_k')/psi(s) with local updates
 * of h (operator.py:510-562 with Variational.ref_forward, variational.py:387-422).
 * nconn (nullable) receives the number of valid candidates per sample. */
int qtx_rbm_eloc(void* stream, int dtype, const void* packed, int32_t M, int32_t N, const qtx_ham* H,
                 const int8_t* s, int64_t ns, double* eloc, int32_t* nconn);

/* Per-sample log-derivatives O (ns, M*N + M) in o_dtype: [tanh(h) (x) s, tanh(h)]
 * (variational.py:424-511 for SingleDense). */
int qtx_rbm_jacobian(void* stream, int dtype, const void* packed, int32_t M, int32_t N, const int8_t* s,
                     int64_t ns, int o_dtype, void* O);

/* ---------------------------------------------------------------- serving
 * Independent RBM_Dense VMC jobs (different N, M, sample counts, couplings)
 * advanced together: each stage is one persistent launch per register class
 * R = ceil(M / 64) in {1, 2, 4, 8, 16} covering every job of that class.  A
 * job's chains use the single-job code and its own Philox stream, so its
 * results are the same alone and in any mix.  No counterpart in the
 * reference, which runs one lattice per process (sites/sites.py:46-51); the
 * per-job stages are the ones qtx_rbm_* replace (float32 parameters).
 * jobs: host array, read before the call returns (it is staged through a
 * pinned per-thread buffer, the one allocation these entry points make
 * themselves: hipHostMalloc once per thread, grown to the largest job table
 * seen and kept, 224 bytes per job, so the host-to-device copy of the table
 * is asynchronous).  ws: qtx_rbm_group_ws_bytes(njobs) bytes of
 * device memory holding the launch's job table; calls that share ws must be
 * issued on one stream. */
typedef struct qtx_rbm_job {
  const float* packed; /* [Wt (N, M) | b (M)] float32 (qtx_rbm_pack) */
  int32_t M, N;
  int32_t kind;        /* QTX_LOCALFLIP | QTX_EXCHANGE */
  int32_t width;       /* row width of nbr */
  const int32_t* nbr;  /* (N, width) -1-padded neighbour table (EXCHANGE) */
  int32_t hopping;     /* species moved by EXCHANGE: +1 or -1 */
  int32_t nsteps;      /* Metropolis steps in a grouped sweep (0: the job sits out) */
  uint64_t seed;       /* key of the job's sweep stream (qtx_philox_fill, TAG_SWEEP) */
  int64_t step0;       /* stream step counter of the sweep's first step */
  int64_t ns;          /* chains = samples */
  int8_t* s;           /* (ns, N) in/out */
  double* logabs;      /* (ns) */
  int32_t* n_accept;   /* (ns) incremented */
  int32_t* n_near;     /* (ns) incremented */
  double reweight;
  double tie_tol;
  qtx_ham H;           /* local-energy tables (device pointers) */
  double* eloc;        /* (ns) out */
  int32_t* nconn;      /* (ns) out, nullable */
  float* O;            /* (ns, M N + M) out */
} qtx_rbm_job;

size_t qtx_rbm_group_ws_bytes(int32_t njobs);
/* logabs of every job's chains (as qtx_rbm_logpsi). */
int qtx_rbm_group_logpsi(void* stream, const qtx_rbm_job* jobs, int32_t njobs, void* ws, size_t ws_bytes);
/* nsteps Metropolis steps per job (as qtx_rbm_metropolis with the words of
 * qtx_philox_fill(seed, TAG_SWEEP, step0, nsteps, 0, ns) generated in-kernel). */
int qtx_rbm_group_metropolis(void* stream, const qtx_rbm_job* jobs, int32_t njobs, void* ws, size_t ws_bytes);
/* E_loc of every job's samples (as qtx_rbm_eloc). */
int qtx_rbm_group_eloc(void* stream, const qtx_rbm_job* jobs, int32_t njobs, void* ws, size_t ws_bytes);
/* O rows of every job (as qtx_rbm_jacobian with o_dtype QTX_F32). */
int qtx_rbm_group_jacobian(void* stream, const qtx_rbm_job* jobs, int32_t njobs, void* ws, size_t ws_bytes);

/* ---------------------------------------------------------------- ResConv
 * ResConv(nblocks, channels, kernel 3) on an L0 x L1 periodic grid
 * (model/conv_nets.py:26-183, nn/conv.py:13-68).  channels 16 (L0 * L1 <= 512)
 * or 32 (L0 * L1 <= 400), nblocks <= 64.  activation: QTX_ACT_GELU_TANH is
 * the reference's jax.nn.gelu default (model/conv_nets.py:85, 87);
 * QTX_ACT_GELU_ERF the exact GELU; QTX_ACT_RELU an opt-in ReLU variant (no
 * counterpart in the reference; derivative 0 at x <= 0, as jax.nn.relu). */
#define QTX_ACT_GELU_TANH 0
#define QTX_ACT_GELU_ERF 1
#define QTX_ACT_RELU 2
typedef struct qtx_resconv_desc {
  int32_t nblocks;
  int32_t channels;
  int32_t L0;
  int32_t L1;
  int32_t precision;  /* QTX_CONV_* */
  int32_t activation; /* QTX_ACT_* */
} qtx_resconv_desc;

int64_t qtx_resconv_nparams(const qtx_resconv_desc* d);
size_t qtx_resconv_packed_bytes(const qtx_resconv_desc* d);
/* theta: float32 ravel_pytree vector (qtx_resconv_nparams entries). */
int qtx_resconv_pack(void* stream, const qtx_resconv_desc* d, const float* theta, void* packed);

/* log psi for a batch (Variational.__call__ for ResConv). */
int qtx_resconv_logpsi(void* stream, const qtx_resconv_desc* d, const void* packed, const int8_t* s,
                       int64_t ns, double* logabs);

/* Metropolis steps with one full forward per proposal (ResConv is not a
 * RefModel: variational.py:383-384).  logabs: in = current, out = final. */
int qtx_resconv_metropolis(void* stream, const qtx_resconv_desc* d, const void* packed, int kind,
                           const int32_t* nbr, int32_t width, int32_t hopping, const uint32_t* words,
                           int32_t nsteps, int64_t ns, double reweight, double tie_tol, int8_t* s,
                           double* logabs, int32_t* n_accept, int32_t* n_near);

/* Local energy: one full forward per valid connected configuration, never
 * materialised.  logabs = log psi(s) of the samples.  H->noff <= 4096
 * (<= 3072 for the 16 x 16, 16-channel network). */
int qtx_resconv_eloc(void* stream, const qtx_resconv_desc* d, const void* packed, const qtx_ham* H,
                     const int8_t* s, const double* logabs, int64_t ns, double* eloc, int32_t* nconn);

/* Per-sample log-derivatives O (ns, nparams) float32 via a per-sample
 * backward pass.  ws: qtx_resconv_jacobian_ws_bytes(d, ns) bytes. */
size_t qtx_resconv_jacobian_ws_bytes(const qtx_resconv_desc* d, int64_t ns);
int qtx_resconv_jacobian(void* stream, const qtx_resconv_desc* d, const void* packed, const int8_t* s,
                         int64_t ns, float* O, void* ws, size_t ws_bytes);

/* The same O as qtx_resconv_jacobian with no workspace argument and no
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The per-sample records live in a
 * 128 MiB file-scope device array of the library, one per GPU, held until the
 * process ends; samples run in launches of as many as it holds (e.g. 256 at
 * 16 x 16, 8 blocks, 16 channels).  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (different streams or threads) race.
 * Returns QTX_EINVAL when one sample's records exceed the array. */
int qtx_resconv_jacobian_noalloc(void* stream, const qtx_resconv_desc* d, const void* packed, const int8_t* s,
                                 int64_t ns, float* O);

/* ---------------------------------------------------------------- SR / MinSR
 * Column means of O (ns, np) in o_dtype, accumulated in float64
 * (the Omean of optimizer/sr.py:79-88). */
int qtx_colmean(void* stream, int o_dtype, const void* O, int64_t ns, int64_t np, double* mean);

/* Ebar = (E - mean E) sqrt(r/ns); out2 = {energy, VarE} (sr.py:180-195).
 * r nullable (= 1). */
int qtx_ebar(void* stream, const double* eloc, const double* r, int64_t ns, double* ebar, double* out2);

/* y = beta y + alpha Obar v, Obar = (O - mean) sqrt(r/ns) formed on the fly
 * (the Obar @ last_step / Obar @ m products of the momentum optimizers,
 * optimizer/sr.py:249, :320, :411).  r nullable (= 1); beta = 0 ignores y's
 * contents. */
int qtx_obar_matvec(void* stream, int o_dtype, const void* O, const double* mean, const double* r, int64_t ns,
                    int64_t np, double alpha, const double* v, double beta, double* y);

/* O[:, k] *= d[k] in place (in o_dtype), and mean[k] *= d[k] when mean is
 * given: the column scaling Obar /= V[None, :] of optimizer/sr.py:327, :416,
 * applied to O and its mean because Obar is formed inside the solve. */
int qtx_scale_columns(void* stream, int o_dtype, void* O, double* mean, int64_t ns, int64_t np, const double* d);

/* step = solver(Obar, Ebar) with Obar = (O - mean) sqrt(r/ns) formed on the
 * fly.  ns < np: MinSR, T = Obar Obar^T (ns x ns); else SR, S = Obar^T Obar.
 * mode CHOL_SHIFT: shift = rshift * trace + ashift (solver.py:50-91);
 * mode EIGH_PINV : soft inverse g(v) = 1/(v (1 + (c/|v|)^6)), c = rtol max|v| + atol
 *                  (solver.py:94-201, tol_snr = 0).  When c >= 1e-5 lambda_max it is
 *                  evaluated without eigenvectors: g(v) = v^5/(v^6 + c^6) =
 *                  (1/3) sum_k Re 1/(v - p_k), p_k = c exp(i pi (2k+1)/6), each complex
 *                  shifted system solved by Cholesky of (T - Re p)^2 + (Im p)^2 (the
 *                  three factored together by this library's blocked Cholesky) plus
 *                  iterative refinement against T; lambda_max by Lanczos.  Smaller
 *                  cutoffs (the reference default rtol = 1e-12): if T_d - 150 c I
 *                  factors, T_d = T + lambda_max 1 1^T / ns for MinSR (the centred
 *                  Gram's null direction 1 deflated; b projected off it), T_d = S for
 *                  SR, then every other eigenvalue exceeds 150 c, where g(v) = 1/v to
 *                  (c/v)^6 <= 8.8e-14, and T_d^-1 b is refined with that factor; a
 *                  failed or non-converging factorisation (and reweighted MinSR)
 *                  uses rocSOLVER dsyevd, as does a failed partial-fraction one.
 * mode EIGH_PINV_DIRECT: the same g(T) b through dsyevd always.
 * gram: QTX_GRAM_* (the ns >= np branch always uses FP64 products: Obar formed
 *       in FP64 by row blocks of at most 256 MiB, S accumulated by rocBLAS DGEMM).
 * ws: qtx_minsr_ws_bytes(...) bytes of device memory. */
size_t qtx_minsr_ws_bytes(int64_t ns, int64_t np, int mode, int gram);
/* The Gram matrix alone, FP64 row-major: T = Obar Obar^T (ns x ns) when
 * ns < np, else S = Obar^T Obar (np x np) (solver.py:134-139, 157).
 * ws: qtx_gram_ws_bytes(...) bytes (the bf16 hi/lo planes of the split-bf16
 * mode; 0 bytes, ws may be NULL, otherwise). */
size_t qtx_gram_ws_bytes(int64_t ns, int64_t np, int gram);
int qtx_gram(void* stream, int o_dtype, const void* O, const double* mean, const double* r, int64_t ns,
             int64_t np, int gram, double* T, void* ws, size_t ws_bytes);
int qtx_minsr_solve(void* stream, int o_dtype, const void* O, const double* mean, const double* r,
                    const double* ebar, int64_t ns, int64_t np, int mode, int gram, double rshift,
                    double ashift, double rtol, double atol, double* step, void* ws, size_t ws_bytes);
/* Many independent solves (different shapes, streams and modes) driven from
 * one host thread: each phase of every task is enqueued on that task's stream
 * before any task's next phase reads results back, so the tasks overlap on
 * the device.  Results are those of qtx_minsr_solve on each task.  Returns 0
 * when every task succeeded, else the first failing task's code (its message
 * in qtx_last_error); status / path / fallback are set per task.  No
 * counterpart in the reference (one solve per process step); it serves the
 * job stream of qtx_rbm_group_*. */
typedef struct qtx_minsr_task {
  void* stream;
  int32_t o_dtype;      /* QTX_F32 | QTX_F64 */
  int32_t mode;         /* QTX_SOLVE_* */
  int32_t gram;         /* QTX_GRAM_* */
  int32_t status;       /* out: 0 or the QTX_E* code of this task */
  const void* O;        /* (ns, np) */
  const double* mean;   /* (np) column means */
  const double* r;      /* (ns) reweight factors, nullable */
  const double* ebar;   /* (ns) */
  int64_t ns, np;
  double rshift, ashift, rtol, atol;
  double* step;         /* (np) out */
  void* ws;             /* qtx_minsr_ws_bytes(ns, np, mode, gram) bytes */
  size_t ws_bytes;
  int32_t path;         /* out: as qtx_minsr_last_path */
  int32_t fallback;     /* out: as qtx_minsr_last_path */
  double tol_snr;       /* SNR truncation of the eigen-components (EIGH_PINV modes; > 1e-6 takes dsyevd and
                           needs QTX_SOLVE_SNR in the mode of qtx_minsr_ws_bytes), 0: none */
} qtx_minsr_task;
int qtx_minsr_solve_many(qtx_minsr_task* tasks, int32_t ntasks);

/* The rocBLAS / rocSOLVER scratch bytes of a solve of this shape and mode on
 * this stream, by a rocBLAS device-memory size query over every library call
 * the solve can make (no kernels run).  ws: a workspace of at least
 * qtx_minsr_ws_bytes(ns, np, mode, gram) bytes (pointers into it are handed to
 * the query; nothing is read or written).  *bytes is at least 256 (a solve
 * with QTX_SOLVE_BLAS_WS needs a non-empty tail).  A workspace of
 * qtx_minsr_ws_bytes + *bytes bytes with QTX_SOLVE_BLAS_WS in the mode makes
 * the solve allocate nothing; a tail too small for a rocSOLVER call fails
 * that solve with QTX_EWORKSPACE.  No counterpart in the reference (XLA owns its scratch). */
int qtx_minsr_blas_ws_bytes(void* stream, int64_t ns, int64_t np, int mode, int gram, void* ws, size_t ws_bytes,
                            size_t* bytes);

/* qtx_minsr_solve with the signal-to-noise truncation of minnorm_pinv_eig /
 * lstsq_pinv_eig (optimizer/solver.py:114-164, _sum_without_noise): each
 * eigen-component rho_s = sum_t c_ts of U^T b is divided by
 * 1 + (tol_snr / snr_s)^6, snr_s = |mean_t c_ts| / sqrt(var_t c_ts / n_t), with
 * c_ts = U_ts Ebar_t (MinSR; t over samples) or (Obar V)_ts Ebar_t (SR).
 * mode: QTX_SOLVE_EIGH_PINV or _DIRECT, with optional branch flags; ws:
 * qtx_minsr_ws_bytes(ns, np, mode | QTX_SOLVE_SNR, gram) bytes.  tol_snr <= 1e-6
 * is no truncation, as in the reference. */
int qtx_minsr_solve_snr(void* stream, int o_dtype, const void* O, const double* mean, const double* r,
                        const double* ebar, int64_t ns, int64_t np, int mode, int gram, double rtol, double atol,
                        double tol_snr, double* step, void* ws, size_t ws_bytes);

/* minsr_pinv_eig (optimizer/solver.py:262-294): x = U g(D) rho for a given
 * symmetric T (n x n, FP64, either triangle layout: T must be symmetric),
 * T = U D U^T by rocSOLVER dsyevd, g the soft inverse with rtol, atol, and rho
 * = U^T b with the SNR truncation above (c_ts = U_ts b_t) when tol_snr > 1e-6.
 * ws: qtx_pinv_eig_ws_bytes(n) bytes. */
size_t qtx_pinv_eig_ws_bytes(int64_t n);
int qtx_pinv_eig(void* stream, const double* T, int64_t n, const double* b, double rtol, double atol, double tol_snr,
                 double* x, void* ws, size_t ws_bytes);

/* What the calling thread's last qtx_minsr_solve did: path 0 Cholesky shift,
 * 1 partial fractions, 2 dsyevd, 3 shifted Cholesky (small cutoffs), 4
 * Chebyshev series (-1: none
 * or failed); Lanczos steps and the lambda_max they gave (0 when not run);
 * fallback (nullable): why EIGH_PINV took dsyevd: 0 it did not, 1 non-finite
 * cutoff or a small cutoff with reweighted MinSR, 2 non-finite Lanczos
 * coefficients (non-finite T), 3 lambda_max <= 0, 4 a partial-fraction or
 * Chebyshev-route Cholesky factorisation (or triangular inverse) failed, 5 SNR
 * truncation (needs the eigenvectors), 6
 * the shifted factorisation failed (an eigenvalue of T_d at or below 150 c),
 * 7 its refinement did not converge, 8 the Chebyshev series grew (T has
 * eigenvalues below about -0.048 c, from rounding in the Gram): the partial
 * fractions ran instead (path 1).  Diagnostic only: no counterpart in the
 * reference, which always runs eigh. */
int qtx_minsr_last_path(int* path, int* lanczos_steps, double* lambda_max, int* fallback);

/* theta <- theta - lr * step in dtype, skipped entirely when any step entry is
 * non-finite (state/variational.py:558-579).  *applied (device int32) = 1/0. */
int qtx_apply_update(void* stream, int dtype, void* theta, const double* step, double lr, int64_t np,
                     int32_t* applied);

/* ---------------------------------------------------------------- symmetry projection
 * Images of each sample under a symmetry group, Symmetry.get_symm_spins
 * (symmetry/symmetry.py:324-341): out[s * nimg + g] = s[perm[g]] for g < nperm
 * and -s[perm[g - nperm]] after them when z2 != 0 (nimg = nperm * (z2 ? 2 : 1));
 * perm: (nperm, N) int32 site indices. */
int qtx_symm_images(void* stream, const int8_t* s, int64_t ns, int32_t N, const int32_t* perm, int32_t nperm,
                    int32_t z2, int8_t* out);

/* Symmetry.symmetrize (symmetry.py:343-392) in log form: psi_sym(s) =
 * sum_g coef[g] sign_g exp(logabs[s * nimg + g]) -> (sign, logabs_out) per
 * sample; coef[g] = chi_g chi_0 / nimg; sign_in nullable (= +1).  w (nullable):
 * the weights coef_g psi_g / psi_sym of the images in d log|psi_sym| / d theta
 * (the per-copy vjp sum of variational.py:436-491). */
int qtx_symmetrize(void* stream, const double* logabs, const double* sign_in, const double* coef, int64_t ns,
                   int32_t nimg, double* sign, double* logabs_out, double* w);

/* O[s] = sum_g w[s, g] Oimg[s * nimg + g] (rows of np, o_dtype; FP64 sums):
 * the symmetrised Jacobian from the images' Jacobians. */
int qtx_rows_combine(void* stream, int o_dtype, const void* Oimg, const double* w, int64_t ns, int32_t nimg,
                     int64_t np, void* O);

/* One proposal per chain from this step's Philox words (words: ns x 4 uint32,
 * the slice of one step), LocalFlip / SpinExchange as in the fused Metropolis
 * kernels (sampler/common_samplers.py:14-162): s_prop = s with the move applied,
 * updated[c] = 0 when the move leaves s unchanged.  N <= 1024. */
int qtx_propose(void* stream, int kind, const int32_t* nbr, int32_t width, int32_t hopping, const uint32_t* words,
                const int8_t* s, int64_t ns, int32_t N, int8_t* s_prop, int32_t* updated);

/* Metropolis._update (sampler/metropolis.py:291-322) for the proposals of
 * qtx_propose: accept iff reweight (lp' - lp) > log1p(-u) or lp = -inf, u from
 * the same words; copies s_prop into s, lp_prop into lp (and sign_prop into
 * sign, both nullable); counts acceptances and |d - log1p(-u)| < tie_tol. */
int qtx_accept(void* stream, const uint32_t* words, const int8_t* s_prop, const double* lp_prop, const double* sg_prop,
               const int32_t* updated, int64_t ns, int32_t N, double reweight, double tie_tol, int8_t* s, double* lp,
               double* sg, int32_t* n_accept, int32_t* n_near);

/* E_loc(s) = diag(s) + sum_{k in [off[s], off[s+1])} H_k (sign_k / sign_s)
 * exp(logabs_k - logabs_s) over compacted connected configurations
 * (operator/operator.py:168-184, 510-562); sign_c, sign nullable (= +1). */
int qtx_conn_combine(void* stream, const double* diag, const int64_t* off, const double* hc, const double* sign_c,
                     const double* logabs_c, const double* sign, const double* logabs, int64_t ns, double* eloc);

#ifdef __cplusplus
}
#endif
#endif /* QTX_HIP_H */
