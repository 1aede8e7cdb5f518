/*
 * hipmc: C ABI of the MI355X-native tfp.mcmc gradient-based sampling path.
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream
 * (as uint64), returns 0 on success or a negative HIPMC_E* code, and never
 * throws.  `hipmc_last_error()` returns a thread-local message for the last
 * failure.  The library allocates no device memory: callers own every
 * buffer (the Python layer passes torch / DLPack tensors).
 *
 * Reference interfaces replaced (paths relative to
 * tensorflow_probability/python/):
 *   hipmc_hmc_run        HamiltonianMonteCarlo.one_step  (mcmc/hmc.py:501-529)
 *                        = MetropolisHastings.one_step (metropolis_hastings.py:160-254)
 *                        over UncalibratedHamiltonianMonteCarlo.one_step
 *                        (hmc.py:661-729); with n_iters > 1 also the
 *                        sample_chain loop (mcmc/sample.py:311-383).
 *   hipmc_nuts_run       NoUTurnSampler.one_step (mcmc/nuts.py:321-445) and
 *                        its tree doubling (nuts.py:541-1010); n_iters > 1
 *                        fuses the sample_chain loop.
 *   hipmc_target_eval    target_log_prob_fn + autodiff gradient
 *                        (mcmc/internal/util.py:246-308 ->
 *                        math/gradient.py:28) for the target specs, with the
 *                        TTK bijector composition
 *                        (mcmc/transformed_kernel.py:86-128).
 *   hipmc_da_update      DualAveragingStepSizeAdaptation._one_step_part
 *                        (mcmc/dual_averaging_step_size_adaptation.py:353-475).
 *   hipmc_simple_step_update
 *                        SimpleStepSizeAdaptation.one_step's step-size rule
 *                        (mcmc/simple_step_size_adaptation.py:344-438).
 *   hipmc_moments_update RunningCovariance.update / RunningMean.update
 *                        (experimental/stats/sample_stats.py:167-235,
 *                        483-520), the state update of the streaming
 *                        reducers (experimental/mcmc/covariance_reducer.py,
 *                        expectations_reducer.py,
 *                        potential_scale_reduction_reducer.py).
 *   hipmc_mh_accept      MetropolisHastings accept/reject
 *                        (mcmc/metropolis_hastings.py:204-246).
 *   hipmc_leapfrog_run   UncalibratedHamiltonianMonteCarlo.one_step proposal
 *                        (mcmc/hmc.py:661-729, leapfrog_integrator.py:222-362).
 *   hipmc_bijector       Bijector.forward / inverse for TTK state mapping
 *                        (bijectors/softplus.py:139-147, power_transform.py:85-100).
 *   hipmc_glm_hmc_run    HamiltonianMonteCarlo.one_step / sample_chain for
 *                        logistic regression at large batch: the
 *                        autodiff gradient of the Bernoulli-logits GLM
 *                        (bernoulli.py:119-135) as fused BF16 MFMA GEMMs.
 *   hipmc_variance_update
 *                        RunningVariance.update inside
 *                        DiagonalMassMatrixAdaptation.one_step
 *                        (experimental/mcmc/diagonal_mass_matrix_adaptation.py:186-228).
 *   hipmc_draws          the momentum / uniform draws of samplers.normal /
 *                        samplers.uniform (internal/samplers.py:308-367) under
 *                        this library's counter-keyed Philox schedule.
 */
#ifndef HIPMC_H_
#define HIPMC_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIPMC_ABI_VERSION 4

/* error codes */
#define HIPMC_OK 0
#define HIPMC_E_INVALID -1   /* bad argument (shape, dtype, null pointer) */
#define HIPMC_E_UNSUPPORTED -2 /* valid but not built (e.g. dim too large) */
#define HIPMC_E_HIP -3       /* a HIP runtime call failed */

/* dtypes */
#define HIPMC_F32 0
#define HIPMC_F64 1

/* target kinds */
#define HIPMC_TARGET_LOGISTIC 1     /* beta ~ N(0,s^2), y ~ Bern(sigmoid(X beta)) */
#define HIPMC_TARGET_LINEAR 2       /* [beta, sigma]: y ~ N(X beta, sigma), sigma ~ HalfNormal */
#define HIPMC_TARGET_HIER_NORMAL 3  /* [mu, tau, z]: non-centred hierarchical normal */
#define HIPMC_TARGET_DIAG_NORMAL 4  /* independent N(loc_i, scale_i) */
#define HIPMC_TARGET_FUNNEL 5       /* Neal's funnel [v, x] */
#define HIPMC_TARGET_NAN_EXP 6      /* sum(where(x<0, NaN, -x)) */
#define HIPMC_TARGET_LOG_GAMMA 7    /* x = log Gamma(param[0], rate param[1]) */

/* bijector codes */
#define HIPMC_BIJ_IDENTITY 0
#define HIPMC_BIJ_SOFTPLUS 1
#define HIPMC_BIJ_EXP 2

#define HIPMC_MAX_BIJ_RANGES 4

/* reductions of the chains' log accept probabilities (hipmc_da_update) */
#define HIPMC_REDUCE_MEAN 0
#define HIPMC_REDUCE_HARMONIC 1

/*
 * One posterior ("problem").  Lives in DEVICE memory as an array; the
 * layout is fixed (offsets asserted in the library).  Data pointers are in
 * the kernel dtype.
 *   LOGISTIC / LINEAR: x = X [n_obs, dim_x] row-major (row pitch x_pitch),
 *                      xt = X^T [dim_x, n_obs] (row pitch xt_pitch) or
 *                      NULL: read only when dim_x <= 64 (F32) / 32 (F64),
 *                      for the logits; otherwise logits and gradient come
 *                      from one pass over the rows of x,
 *                      y [n_obs]; param[0] = prior scale of beta,
 *                      param[1] = HalfNormal scale of sigma (LINEAR).
 *   HIER_NORMAL:       y [J], aux = sigma [J], n_obs = J;
 *                      param[0] = mu prior scale, param[1] = tau scale.
 *   DIAG_NORMAL:       y = loc [dim], aux = scale [dim].
 *   FUNNEL / NAN_EXP:  no data.
 *   LOG_GAMMA:         iid elements; param[0] = concentration,
 *                      param[1] = rate.
 * Bijector ranges [off, off+len) over the flat state, code per range.
 */
typedef struct hipmc_problem {
  int32_t kind;
  int32_t dim;      /* flat state size D (sum of part sizes) */
  int32_t n_obs;
  int32_t n_chains; /* chains of this problem */
  int64_t chain_offset; /* first row of this problem in the state arrays */
  const void* x;
  const void* xt;
  const void* y;
  const void* aux;
  double param[4];
  double log_norm;  /* constant added to the log density (host-computed) */
  int32_t n_bij;
  int32_t bij_off[HIPMC_MAX_BIJ_RANGES];
  int32_t bij_len[HIPMC_MAX_BIJ_RANGES];
  int32_t bij_code[HIPMC_MAX_BIJ_RANGES];
  uint32_t key; /* problem word of the Philox counter (stable id) */
  /* LOGISTIC / LINEAR: row pitches (elements) of x (>= dim_x) and of a
   * non-NULL xt (>= n_obs), multiples of 4 with x / xt 16-byte aligned:
   * rows are read 16 bytes per lane (padding is read and ignored).  An xt
   * whose pitch breaks these rules is not used.  Other kinds: unused. */
  int32_t x_pitch;
  int32_t xt_pitch;
} hipmc_problem;

/*
 * Optional per-problem sampling stream (device array indexed like the
 * problem table).  When hipmc_run_args.streams is set, problem p's chains
 * run min(n_iters, streams[p].n_iters) iterations of the launch, iteration
 * i uses fold_in(streams[p].seed, streams[p].iter0 + i), and if
 * streams[p].trace is set the post-step state of recorded iterations
 * (g - trace_start a non-negative multiple of trace_every, row < n_trace)
 * goes to trace[row][local_chain][0:dim] (dense per problem).  This lets
 * posteriors at different points of their runs share one launch.
 */
typedef struct hipmc_stream {
  uint32_t seed[2];
  int64_t iter0;
  int32_t n_iters;
  int32_t trace_every;
  void* trace;
  int64_t trace_start;
  int32_t n_trace;
  int32_t pad_;
} hipmc_stream;

/*
 * Arguments of the sampling kernels.  Rows of the state arrays are chains
 * (stride `dim_stride` elements).  Chain c belongs to problem
 * chain_problem[c] (NULL: problem 0) and is that problem's chain
 * chain_local[c] (NULL: c); the pair keys its random draws, so results do
 * not depend on which other chains share the launch.
 *
 * Seeds: if fold_seed != 0, iteration i uses fold_in(seed, iter0 + i);
 * otherwise `seed` is the step seed itself (n_iters must be 1).
 *
 * Per-iteration outputs are [rows, n_chains] arrays; iteration i of the
 * launch writes row out_row0 + i.  Any may be NULL.  trace_state, if set,
 * is [n_trace, n_chains, dim_stride]: global iteration g = iter0 + i
 * writes its post-step (unconstrained) state to row
 * (g - trace_start) / trace_every when g >= trace_start, that difference is
 * a multiple of trace_every and the row is < n_trace (sample_chain's
 * burn-in and thinning).
 */
typedef struct hipmc_run_args {
  int32_t dtype;
  int32_t n_chains;
  int32_t dim_stride;
  int32_t max_dim;        /* max problem dim in the launch (<= dim_stride) */
  int32_t n_iters;
  int32_t fold_seed;
  int64_t iter0;
  uint32_t seed[2];
  /* HMC */
  int32_t num_leapfrog_steps;
  /* NUTS */
  int32_t max_tree_depth;
  int32_t unrolled_leapfrog_steps;
  /* kind of the launch's problems; LOGISTIC and LINEAR may be mixed */
  int32_t target_kind;
  double max_energy_diff;
  /* problems and chain maps (device) */
  const hipmc_problem* problems;
  int32_t n_problems;
  /* LOGISTIC / LINEAR launches of hipmc_nuts_run / hipmc_hmc_run: waves that
   * share one chain's gradient (each takes every waves_per_chain-th block of
   * rows; partial sums meet in LDS).  0 or 1: one wave per chain.  At most
   * 8 for hipmc_nuts_run and 4 for hipmc_hmc_run (larger values are
   * clamped), and fewer where the chain's LDS would exceed 160 KiB (large
   * slot counts in FP64); other target kinds ignore it. */
  int32_t waves_per_chain;
  const int32_t* chain_problem;
  const int32_t* chain_local;
  /* step size (device, dtype): step_size[chain_step[c]] (chain_step NULL: 0) */
  const void* step_size;
  const int32_t* chain_step;
  /* state in/out (device, dtype) */
  void* state;   /* [n_chains, dim_stride] */
  void* tlp;     /* [n_chains] */
  void* grad;    /* [n_chains, dim_stride] */
  /* per-iteration outputs (device), any may be NULL */
  void* out_log_accept_ratio; /* dtype */
  int32_t* out_leapfrogs;
  uint8_t* out_is_accepted;
  uint8_t* out_reach_max_depth;
  uint8_t* out_has_divergence;
  void* out_energy;           /* dtype; HMC: log_acceptance_correction */
  int32_t* out_tree_depth;
  int64_t out_row0;
  void* trace_state;
  int64_t trace_start;
  int32_t trace_every;
  int32_t n_trace;
  /* HMC extras (last iteration only, any may be NULL), [n_chains, dim_stride] */
  void* out_proposed_state;
  void* out_proposed_tlp;     /* [n_chains] */
  void* out_initial_momentum;
  void* out_final_momentum;
  /* optional per-problem streams (device, [n_problems]); NULL: launch-level
   * seed / iter0 / trace for every problem */
  const hipmc_stream* streams;
  /* optional row of chain c in state / tlp / grad (device, [n_chains]);
   * NULL: row c.  Lets a launch cover a subset of a resident chain pool. */
  const int32_t* chain_row;
  /* optional per-iteration target log-prob after the step (dtype,
   * [rows, n_chains], written like out_log_accept_ratio) */
  void* out_target_log_prob;
  /* optional diagonal momentum precision (device, dtype,
   * [n_problems, dim_stride], row chain_problem[c]): the precision of the
   * momentum distribution N(0, diag(1 / prec)), i.e. the estimated
   * posterior variance of each state element, as set by
   * DiagonalMassMatrixAdaptation
   * (experimental/mcmc/diagonal_mass_matrix_adaptation.py:227-240,
   * preconditioning_utils.py:70-88).  Momentum p = z / sqrt(prec),
   * velocity prec * p, kinetic energy 0.5 sum(prec p^2) (the preconditioned
   * kernels, experimental/mcmc/preconditioned_hmc.py,
   * preconditioned_nuts.py:804-1010); NUTS then reports energy including
   * the momentum density's normalising terms.  NULL: identity. */
  const void* momentum_precision;
  /* hipmc_nuts_run on LOGISTIC / LINEAR problems: chains that share each
   * pass over X (0 or 1: none).  With 2 or 4, blocks of waves_per_chain
   * waves run that many consecutive chains (fewer where the LDS does not
   * fit, and none for states above 192 elements or more than one leapfrog
   * per leaf), and every group of chain_problem must lie within one
   * problem: each problem's chains are consecutive and their count a
   * multiple of chains_per_block.  Results equal each chain's own run up to
   * the order of floating-point sums (the group's partial sums are added in
   * a fixed order, so a chain's results depend on its group size only). */
  int32_t chains_per_block;
  int32_t pad0_;
} hipmc_run_args;

/*
 * Large-batch HMC on Bayesian logistic regression with X stored in BF16 and
 * the all-chain gradient on MFMA (BF16 hi/lo split of positions and
 * residuals, FP32 accumulation).  float32 state; dim must be 256; n_obs a
 * multiple of 32; n_chains a multiple of 64.  Same momentum / MH draw
 * schedule as hipmc_hmc_run (problem word = key, chain word = row).
 * Workspace arrays are caller-owned (see field comments).
 */
typedef struct hipmc_glm_hmc_args {
  int32_t n_obs;
  int32_t dim;
  int32_t n_chains;
  int32_t n_split;          /* row splits of each gradient launch (>= 1) */
  int32_t num_leapfrog_steps;
  int32_t n_iters;
  int32_t fold_seed;
  uint32_t key;
  int64_t iter0;
  uint32_t seed[2];
  double prior_scale;
  const void* x_bf16;       /* [n_obs, dim] */
  const void* xt_bf16;      /* [dim, n_obs] */
  const float* y;           /* [n_obs] */
  const float* step_size;   /* [1] or [n_chains] */
  int32_t per_chain_step;
  int32_t want_grad_only;   /* 1: only compute tlp/grad of `state` */
  float* state;             /* [n_chains, dim] in/out */
  float* tlp;               /* [n_chains] in/out */
  float* grad;              /* [n_chains, dim] in/out */
  float* ws_g;              /* [n_split, n_chains, dim] */
  float* ws_ll;             /* [n_split, n_chains] */
  float* ws_q;              /* [n_chains, dim] */
  float* ws_p;              /* [n_chains, dim] */
  void* ws_b;               /* [2, n_chains, dim] bf16 */
  double* ws_ke0;           /* [n_chains] */
  float* out_log_accept_ratio; /* [n_iters, n_chains] or NULL */
  uint8_t* out_is_accepted;    /* [n_iters, n_chains] or NULL */
  float* trace_state;          /* [n_trace, n_chains, dim] or NULL */
  int64_t trace_start;
  int32_t trace_every;
  int32_t n_trace;
} hipmc_glm_hmc_args;

int hipmc_glm_hmc_run(const hipmc_glm_hmc_args* args, uint64_t stream);

/*
 * hipmc_glm_hmc_run without caller workspaces, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It has no counterpart in the
 * reference.  The ws_* fields of `args` are ignored: the workspaces are
 * carved from a static device array of HIPMC_NOALLOC_SCRATCH_BYTES that
 * each GPU the library runs on holds until the process ends; the entry point
 * calls no allocator.  Needs 256-byte-rounded sizes of n_split*n_chains*256
 * float32 + n_split*n_chains float32 + 2*n_chains*256 float32 +
 * 2*n_chains*256 bf16 + n_chains float64 to fit, else HIPMC_E_UNSUPPORTED.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams or threads) race.  Outputs equal hipmc_glm_hmc_run's.
 */
#define HIPMC_NOALLOC_SCRATCH_BYTES (64u << 20)
int hipmc_glm_hmc_run_noalloc(const hipmc_glm_hmc_args* args,
                              uint64_t stream);

const char* hipmc_last_error(void);
int hipmc_abi_version(void);

/* Host-side seed derivation with the same Philox block the kernels use
 * (role of internal/samplers.py:79-258 sanitize_seed / split_seed):
 *   split:   out = philox(key=seed, ctr=(i, 0, 0, 'SPLT'))[0:2]
 *   fold_in: out = philox(key=seed, ctr=(t_lo, t_hi, 0, 'FOLD'))[0:2]
 *   salt:    out = philox(key=seed, ctr=(salt_hash, 0, 0, 'SALT'))[0:2] */
int hipmc_seed_split(const uint32_t seed[2], uint32_t i, uint32_t out[2]);
int hipmc_seed_fold_in(const uint32_t seed[2], uint64_t t, uint32_t out[2]);
int hipmc_seed_salt(const uint32_t seed[2], uint32_t salt_hash,
                    uint32_t out[2]);
/* Philox4x32-10 block on the host (known-answer tests of the ABI). */
int hipmc_philox_host(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]);
int hipmc_sizeof_problem(void);
int hipmc_sizeof_run_args(void);
int hipmc_sizeof_stream(void);
int hipmc_sizeof_glm_hmc_args(void);

/* Fused calibrated HMC (leapfrog + MH) for n_iters iterations. */
int hipmc_hmc_run(const hipmc_run_args* args, uint64_t stream);

/* Uncalibrated HMC proposal (n_iters must be 1): state/tlp/grad become the
 * proposal; momenta and log_acceptance_correction (out_energy) written. */
int hipmc_leapfrog_run(const hipmc_run_args* args, uint64_t stream);

/* NUTS for n_iters iterations (per-chain wavefront state machine). */
int hipmc_nuts_run(const hipmc_run_args* args, uint64_t stream);

/* target log-prob and gradient at `x` [n_chains, dim_stride]. */
int hipmc_target_eval(int32_t dtype, int32_t target_kind,
                      const hipmc_problem* problems,
                      const int32_t* chain_problem, int32_t n_chains,
                      int32_t dim_stride, int32_t max_dim, const void* x,
                      void* out_tlp, void* out_grad, uint64_t stream);

/*
 * Dual-averaging update of n_steps step-size elements, applied for n_rows
 * successive DA steps (row r reads log_accept_ratio + r*row_stride and is DA
 * step `step + r`).  Element s keeps its DA state and step size at index
 * state_index[s] of the state arrays (NULL: s).  Element s reads the
 * chains of group g = step_group[s] (NULL: g = s), i.e. entries
 * [group_start[g], group_start[g+1]) of `log_accept_ratio` (dtype; chains of
 * a group contiguous).  DA state arrays are float64 [n_steps]: error_sum,
 * log_averaging_step, log_shrinkage_target, target_accept_prob.  `step` is
 * the DA step count before this update.  Writes step_size (dtype
 * [n_steps]) and, if non-NULL, new_step_size_f64 (which also supplies the
 * previous step size when set).  reduce_mode: HIPMC_REDUCE_MEAN reduces the
 * chains' log accept probabilities with reduce_logmeanexp (the class
 * default), HIPMC_REDUCE_HARMONIC with reduce_log_harmonic_mean_exp
 * (math/generic.py:387-432), as windowed adaptation configures it
 * (experimental/mcmc/windowed_sampling.py:869-873).  elem_step and
 * elem_num_adaptation_steps (int64 [n_steps], or NULL for the scalars
 * `step` and `num_adaptation_steps`) give each element its own DA step
 * count and adaptation length, so problems at different points of their
 * adaptation (requests admitted at different times) update in one launch.
 */
int hipmc_da_update(int32_t dtype, int32_t n_steps, const int32_t* step_group,
                    const int32_t* state_index,
                    const int32_t* group_start, const void* log_accept_ratio,
                    int32_t n_rows, int64_t row_stride,
                    double* error_sum, double* log_averaging_step,
                    const double* log_shrinkage_target,
                    const double* target_accept_prob, int64_t step,
                    int64_t num_adaptation_steps, double exploration_shrinkage,
                    double step_count_smoothing, double decay_rate,
                    int32_t reduce_mode, const int64_t* elem_step,
                    const int64_t* elem_num_adaptation_steps, void* step_size,
                    double* new_step_size_f64, uint64_t stream);

/*
 * SimpleStepSizeAdaptation update of n_elems step-size elements over n_rows
 * successive steps (row r reads log_accept_ratio + r*row_stride and is step
 * `step + r`).  Element s reads the chains of group g = step_group[s] (NULL:
 * g = s), entries [group_start[g], group_start[g+1]) (dtype, contiguous),
 * whose log-mean-exp of min(0, lar) (non-finite lar as -inf) is taken in
 * float64; while step + r < num_adaptation_steps the element is multiplied
 * by (1 + adaptation_rate) if that exceeds log(target_accept_prob[s]) and
 * divided by it otherwise, in the step-size dtype.  step_size: dtype
 * [n_elems], read and written; target_accept_prob: float64 [n_elems].
 */
int hipmc_simple_step_update(int32_t dtype, int32_t n_elems,
                             const int32_t* step_group,
                             const int32_t* group_start,
                             const void* log_accept_ratio, int32_t n_rows,
                             int64_t row_stride,
                             const double* target_accept_prob,
                             double adaptation_rate, int64_t step,
                             int64_t num_adaptation_steps, void* step_size,
                             uint64_t stream);

/*
 * One-pass (Pebay 2008) update of running moments by a chunk of chunk_n
 * samples.  sample: dtype [chunk_n, n_batch, event_size], contiguous; mean:
 * float64 [n_batch, event_size], read from mean_in and written to mean_out
 * (ssr likewise; out may equal in for an in-place update, other overlaps
 * are not allowed); num_samples: samples already absorbed (shared by every
 * element).  The chunk's mean m_c and residual sum
 * S_c = sum (x - m_c)(x - m_c)' are formed in float64, then, with
 * n' = num_samples + chunk_n and d = m_c - mean:
 *   mean <- mean + chunk_n * d / n'
 *   ssr  <- ssr + S_c + (num_samples * chunk_n / n') * d d'
 * mode 0: mean only (ssr_in / ssr_out may be NULL); mode 1: ssr float64
 * [n_batch, event_size] (per-element variance sums); mode 2: ssr float64
 * [n_batch, event_size, event_size] (cross products within each batch
 * member; event_size <= 46340 in this mode).  Replaces
 * RunningCovariance.update / RunningMean.update
 * (experimental/stats/sample_stats.py:167-235, 483-520), computed there in
 * the sample dtype.
 */
int hipmc_moments_update(int32_t dtype, int32_t mode, int64_t n_batch,
                         int64_t event_size, int64_t chunk_n,
                         const void* sample, double num_samples,
                         const double* mean_in, const double* ssr_in,
                         double* mean_out, double* ssr_out, uint64_t stream);

/*
 * Running variance of groups of chains: group g (chains
 * [group_start[g], group_start[g+1]), row chain_row[c] of x, or c when
 * chain_row is NULL) adds its chains' states x [rows, dim_stride] (dtype,
 * first `dim` elements) as one chunk of observations to (num_samples[g],
 * mean[g, :], ssr[g, :]) (float64; mean and ssr [n_groups, dim_stride]):
 * chunk mean and residual sum in float64, then Pebay's merge.  Replaces the
 * RunningVariance.update of DiagonalMassMatrixAdaptation.one_step
 * (experimental/mcmc/diagonal_mass_matrix_adaptation.py:186-228), whose
 * variance ssr / num_samples becomes the momentum precision at the end of
 * each slow window.
 */
int hipmc_variance_update(int32_t dtype, int32_t n_groups,
                          const int32_t* group_start, const int32_t* chain_row,
                          const void* x, int32_t dim_stride, int32_t dim,
                          double* num_samples, double* mean, double* ssr,
                          uint64_t stream);

/*
 * Metropolis-Hastings accept for n_chains chains:
 *   lar = safe_sum(proposed_tlp - current_tlp [+ correction]);
 *   accept = log(U) < lar, U from the MH purpose keyed by `seed` (counter
 *   problem word = problems[chain_problem[c]].key).
 * Rows of `current` (state [n, dim_stride], tlp, grad) are overwritten by the
 * proposal where accepted.  correction may be NULL.
 */
int hipmc_mh_accept(int32_t dtype, int32_t n_chains, int32_t dim_stride,
                    const uint32_t seed[2], const hipmc_problem* problems,
                    const int32_t* chain_problem,
                    const int32_t* chain_local, const void* proposed_state,
                    const void* proposed_tlp, const void* proposed_grad,
                    const void* correction, void* state, void* tlp,
                    void* grad, void* out_log_accept_ratio,
                    uint8_t* out_is_accepted, uint64_t stream);

/* Elementwise bijector on x [n_rows, dim_stride] per the problem's bijector
 * ranges: direction 0 = forward, 1 = inverse.  Row r is chain
 * (chain_period > 0 ? r % chain_period : r) of chain_problem (a trace of
 * [draws, chains] rows uses chain_period = chains). */
int hipmc_bijector(int32_t dtype, const hipmc_problem* problems,
                   const int32_t* chain_problem, int32_t n_rows,
                   int32_t dim_stride, int32_t chain_period,
                   int32_t direction, const void* x, void* out,
                   uint64_t stream);

/*
 * Raw draws of the schedule, for tests and for callers that need the same
 * stream on the host side.  kind 0: momentum [n_chains, dim];
 * kind 1: MH uniform [n_chains]; kind 2: NUTS tree draws for `depth`
 * ([n_chains] direction as 0/1 in out_aux (uint8) and uniform in out);
 * kind 3: NUTS leaf uniform for (`depth`, `index`) [n_chains];
 * kind 4: initial positions uniform on (-2, 2) [n_chains, dim] (the role of
 * init_near_unconstrained_zero, experimental/mcmc/initialization.py, for
 * windowed adaptation): element e of chain c is
 * 4 (w / 2^8 + 0.5) / 2^24 - 2 with w word e % 4 of the block
 * (e / 4, c, problem, 'INIT').
 */
int hipmc_draws(int32_t dtype, int32_t kind, const uint32_t seed[2],
                int32_t problem, int32_t n_chains, int32_t dim, int32_t depth,
                int32_t index, void* out, uint8_t* out_aux, uint64_t stream);

#ifdef __cplusplus
}
#endif

#endif /* HIPMC_H_ */
