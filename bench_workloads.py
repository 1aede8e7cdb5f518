"""The non-headline workloads of bench.py (one JSON line each).

* logistic_1024x25    plain HMC on logistic regression, X [1024, 25], 64 chains,
                      L = 10, eps = 0.05, 500 burn-in + 1000 results.
* logistic_65536x256  HMC (L = 20) + dual averaging (target 0.75, 500 burn-in
                      steps) + 1000 results on X [65536, 256] stored BF16,
                      4096 chains; the all-chain gradient runs as BF16 MFMA
                      GEMMs with a hi/lo split of beta.
* ragged_glms_512     (the default) 512 GLMs (half logistic, half Gaussian
                      linear with sigma ~ HalfNormal(1) under an Exp
                      bijector), N_k ~ LogUniform[128, 8192],
                      D_k ~ U{4..128}, 32 chains each, NUTS with per-problem
                      windowed adaptation (step size and diagonal mass
                      matrix; --c3-dual-averaging: step size only), 500
                      adaptation + 1000 draws, all problems in shared
                      launches; beside it the same problems one per launch.
* reducers_1024x258   the streaming reducers at the headline's state shape
                      (1024 chains x 258 parameters, float32 draws):
                      running mean, per-element variance, per-chain
                      covariance (event_ndims = 1) and streaming R-hat, one
                      update per draw, 64 draws per step.
* serving_poisson200  the continuous-batching server under Poisson arrivals
                      (200 req/s); half the requests are GLMs from the ragged
                      distribution, half hierarchical normals; 16 chains,
                      500 adaptation + 2000 draws each.

Each line carries the headline fields of bench.py (metric, value, unit, ...),
a roofline entry for the dominant kernel and, where the oracle can run a
bounded sample on the host, a cpu_baseline entry.
"""

import math
import sys
import time

import numpy as np

FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
# The per-chain sampling kernels issue no matrix instructions: one wave per
# chain runs FP32 FMA chains, cross-lane reductions and FP64 energy sums, so
# they sit under the compute roof, priced at the FP32 dense peak, and are
# held far below it by instruction latency within each wave.
VALU_LIMITER = ('VALU/latency: one wave per chain, FP32 FMA and cross-lane '
                'reductions, no MFMA; priced against the FP32 dense peak')


def _base(metric, value, unit, args, ms, dtype, config, higher=True,
          steps=None, warmup=None):
  """The headline fields; steps/warmup are the timed and untimed steps the
  workload actually ran (the arguments' when not given)."""
  return {'metric': metric, 'value': value, 'unit': unit, 'n_gpus': 1,
          'steps': args.steps if steps is None else steps,
          'warmup': args.warmup if warmup is None else warmup,
          'ms_per_step': ms,
          'higher_is_better': higher, 'scaling': 'weak', 'vs_baseline': None,
          'dtype': dtype, 'data': 'synthetic (seeded NumPy PCG64)',
          'config': config}


def _ess_rhat(torch, pa, flat):
  """(ESS [D], split R-hat [D]) of draws [R, C, D] (chains cross-pooled)."""
  ess, rhat = [], []
  for s in range(0, flat.shape[-1], 64):
    blk = flat[..., s:s + 64]
    ess.append(pa.mcmc.effective_sample_size(blk, cross_chain_dims=1))
    rhat.append(pa.mcmc.potential_scale_reduction(blk, split_chains=True))
  return torch.cat(ess), torch.cat(rhat)


def _say(msg):
  print(f'[bench] {msg}', file=sys.stderr, flush=True)


def _timed(torch, fn, steps, warmup):
  for _ in range(warmup):
    fn(0)
  torch.cuda.synchronize()
  out, times = None, []
  for k in range(steps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn(k + 1)
    torch.cuda.synchronize()
    times.append(time.perf_counter() - t0)
  return out, 1000.0 * float(np.mean(times))


# ------------------------------------------------------------ C0 plain HMC
def logistic_1024x25(args, torch, pa):
  from probability_amd import workloads
  wl = workloads.logistic_1024x25()
  c, d, n = wl['chains'], wl['x'].shape[1], wl['x'].shape[0]
  spec = pa.targets.LogisticRegression(wl['x'], wl['y'], dtype=torch.float32,
                                       device='cuda')
  hmc = pa.mcmc.HamiltonianMonteCarlo(
      spec, step_size=np.float32(wl['step_size']),
      num_leapfrog_steps=wl['num_leapfrog_steps'])
  init = torch.zeros((c, d), device='cuda')
  nb, nr, L = wl['num_burnin_steps'], wl['num_results'], wl[
      'num_leapfrog_steps']

  def run(k, fused=True, stats=None):
    return pa.mcmc.sample_chain(nr, init, kernel=hmc, num_burnin_steps=nb,
                                trace_fn=None, seed=args.seed + k,
                                experimental_fused=fused,
                                experimental_stats=stats)

  draws, ms = _timed(torch, run, args.steps, args.warmup)
  ess, rhat = _ess_rhat(torch, pa, draws)
  min_ess = float(ess.min())
  stats = {'time_kernels': True}
  run(0, stats=stats)
  kern_ms = sum(l[0] for l in stats['launches'])
  flop = 4.0 * n * d * c * (nb + nr) * L
  achieved = flop / (kern_ms * 1e-3) / 1e12
  line = _base('min-ESS/sec at split-Rhat<1.01 (plain HMC, logistic '
               'regression 1024x25, 64 chains)', min_ess / (ms / 1e3),
               'min-ESS/s', args, ms, 'f32',
               {'workload': 'logistic_1024x25: HamiltonianMonteCarlo(L=10, '
                            'eps=0.05), 64 chains at 0, 500 burn-in + 1000 '
                            'results', 'chains': c, 'n_obs': n, 'dim': d})
  line.update({
      'min_ess': min_ess, 'rhat_max': float(rhat.max()),
      'rhat_ok': float(rhat.max()) < 1.01,
      'chain_grad_evals_per_s': c * (nb + nr) * L / (ms / 1e3),
      'roofline': {'bound': 'mfma', 'achieved': achieved,
                   'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                   'frac': achieved / FP32_PEAK_TFLOPS, 'traffic': None,
                   'limiter': VALU_LIMITER,
                   'kernel': 'hmc_kernel<float,1,glm,calibrated> (all 1500 '
                             'transitions in one launch)',
                   'kernel_ms': kern_ms,
                   'flop_per_chain_leapfrog': 4 * n * d}})
  if not args.no_plain:
    _, pms = _timed(torch, lambda k: run(k, fused=False), 1, 0)
    line['plain'] = {'value': min_ess / (pms / 1e3), 'ms_per_step': pms,
                     'what': 'host step loop, one hmc launch per transition'}
  if not args.no_cpu:
    line['cpu_baseline'] = _cpu_glm_hmc(
        wl['x'], wl['y'], c, L, wl['step_size'], nb + nr, min_ess,
        args.cpu_seconds, 'the 64-chain batch')
  return line


def _cpu_glm_hmc(x, y, chains, L, eps, transitions, min_ess, budget_s,
                 what, sample_chains=None, one_thread_chains=None):
  """TFP's HMC restated in NumPy (oracle.mcmc.hmc_one_step) over the
  logistic-regression log-density with float32 design-matrix products
  (_CpuGlm), timed on the host: transitions of `sample_chains` chains
  (default all) within a time budget, the median of 3 repeats after a
  warm-up, with BLAS on every core of the process's affinity mask and on one
  (`one_thread_chains` chains).  The cost scales linearly in chains (the
  products are GEMMs over chains) to the run's `chains` x `transitions`."""
  import os
  from threadpoolctl import threadpool_limits
  from oracle import mcmc as omcmc
  from oracle import rng as orng
  cores = len(os.sched_getaffinity(0))
  tg = _CpuGlm(x, y, False)
  rep_s = max(0.5, budget_s / 8.0)

  def per_transition(threads, c):
    with threadpool_limits(limits=threads):
      q = np.zeros((c, x.shape[1]))
      t, g = tg.log_prob_and_grad(q)
      reps = []
      for rep in range(4):   # first is warm-up
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < rep_s or n == 0:
          q, r = omcmc.hmc_one_step(tg, eps, L, q, t, g, orng.fold_in(7, n),
                                    orng.PhiloxDraws(np.float32))
          t, g = r.target_log_prob, r.grads_target_log_prob
          n += 1
        if rep:
          reps.append((time.perf_counter() - t0) / n)
    return float(np.median(reps)) * chains / c

  ca = sample_chains or chains
  co = one_thread_chains or ca
  s_all = per_transition(cores, ca) * transitions
  s_one = per_transition(1, co) * transitions
  scaled = ('' if ca == chains else
            f' ({ca} of the {chains} chains timed, scaled by {chains / ca:g})')
  return {'value': float(min_ess / s_all), 'unit': 'min-ESS/s',
          'cores': cores, 'kind': 'port',
          'sample': (f'HMC transitions of {what}{scaled}, median of 3 after '
                     f'a warm-up, NumPy with float32 design-matrix products '
                     f'on {cores} BLAS threads, extrapolated to '
                     f'{transitions} transitions = {s_all:.0f} s with the GPU '
                     'run\'s min-ESS'),
          'seconds_per_run': s_all,
          'one_thread': {'value': float(min_ess / s_one),
                         'seconds_per_run': s_one, 'cores': 1,
                         'chains_timed': co}}


# ------------------------------------------------ C2 large-batch MFMA HMC
# HBM bytes per all-chain gradient of glm_grad_mfma<8> at 65536 x 256 x 4096:
# rocprofv3 FETCH_SIZE 265 300 KiB x 2.00 + WRITE_SIZE 32 863 KiB x 1.00 per
# dispatch, the factors being what the counters read for a 256 MiB device
# copy (profiles/glm_grad_mfma_logistic_65536x256_counters.md)
C2_HBM_BYTES_PER_GRADIENT = (265300 * 2.0 + 32863 * 1.0) * 1024
def logistic_65536x256(args, torch, pa):
  from probability_amd import workloads
  n, d, c, L = 65536, 256, 4096, 20
  nb = 500 if args.adapt is None else args.adapt
  nr = 1000 if args.draws is None else args.draws
  x, y, _ = workloads.logistic_data(n, d, seed=5, x_scale=1.0 / 16.0)
  spec = pa.targets.LogisticRegressionBF16(x, y, device='cuda')
  hmc = pa.mcmc.HamiltonianMonteCarlo(spec, step_size=np.float32(0.02),
                                      num_leapfrog_steps=L)
  da = pa.mcmc.DualAveragingStepSizeAdaptation(
      hmc, num_adaptation_steps=nb, target_accept_prob=0.75)
  init = torch.zeros((c, d), device='cuda')

  def run(k, stats=None):
    return pa.mcmc.sample_chain(
        nr, init, kernel=da, num_burnin_steps=nb,
        trace_fn=pa.mcmc.TraceFields('is_accepted'), seed=args.seed + k,
        return_final_kernel_results=True, experimental_stats=stats)

  # one step is one complete run (500 + 1000 transitions of 4096 chains)
  steps, warm = 1, min(args.warmup, 1)
  _say('mfma: runs')
  stats = {'time_kernels': True}
  res, ms = _timed(torch, lambda k: run(k, stats=stats), steps, warm)
  _say(f'mfma: {ms:.0f} ms; diagnostics')
  draws = res.all_states
  acc = float(res.trace.float().mean())
  ess, rhat = _ess_rhat(torch, pa, draws)
  min_ess = float(ess.min())
  # posterior means vs an FP64 reference: the float64 oracle gradient at the
  # posterior mean must vanish up to the prior/likelihood scale (MAP check is
  # not the mean; report the mean's max MCSE-normalised change between the
  # two halves of the draws instead)
  half = nr // 2
  m1 = draws[:half].double().mean((0, 1))
  m2 = draws[half:].double().mean((0, 1))
  sd = draws.double().std((0, 1))
  mcse = sd / torch.sqrt(ess.double() / 2)
  z_halves = float(((m1 - m2).abs() / (math.sqrt(2) * mcse)).max())
  # error of the MFMA log-density and gradient against float64 on the same
  # BF16-rounded X, at the final states of all chains
  q = draws[-1]
  tlp_m, grad_m = spec.log_prob_and_grad_flat(q)
  xr, y64, q64 = spec.x.double(), spec.y.double(), q.double()
  sc = q64 @ xr.t()
  tlp_r = ((y64 * sc - torch.nn.functional.softplus(sc)).sum(1)
           - 0.5 * (q64 ** 2).sum(1) - d * 0.5 * math.log(2 * math.pi))
  grad_r = (y64 - torch.sigmoid(sc)) @ xr - q64
  del sc
  nw = lambda a, b: float((a.double() - b).abs().max() / b.abs().max())
  err64 = {'target_log_prob': nw(tlp_m, tlp_r), 'grad': nw(grad_m, grad_r),
           'what': 'normwise max|err| / max|value| of the MFMA path vs '
                   'float64 on the BF16-rounded X at the final states of '
                   'all 4096 chains'}
  # kernel time of the gradient GEMMs: HIP events per launch
  launches = stats['launches']
  sample_ms = sum(l[0] for l in launches if l[2] > 1)
  sample_iters = sum(l[2] for l in launches if l[2] > 1)
  grads = sample_iters * L
  flop_grad = 4.0 * n * d * c  # two GEMMs per all-chain gradient
  achieved = grads * flop_grad / (sample_ms * 1e-3) / 1e12
  line = _base('min-ESS/sec at split-Rhat<1.01 (HMC+DA, logistic regression '
               '65536x256 BF16 MFMA, 4096 chains)', min_ess / (ms / 1e3),
               'min-ESS/s', args, ms, 'bf16 x, f32 state',
               {'workload': 'logistic_65536x256: DualAveraging(HMC(L=20), '
                            'target 0.75, 500 steps) + 1000 results, X '
                            'N(0,1)/16 in BF16, 4096 chains at 0; one step '
                            '= one complete run',
                'chains': c, 'n_obs': n, 'dim': d},
               steps=steps, warmup=warm)
  line.update({
      'min_ess': min_ess, 'rhat_max': float(rhat.max()),
      'rhat_ok': float(rhat.max()) < 1.01, 'accept_rate': acc,
      'final_step_size': float(res.final_kernel_results.new_step_size),
      'mean_halves_max_z': z_halves,
      'error_vs_fp64': err64,
      'chain_grad_evals_per_s': c * (nb + nr) * L / (ms / 1e3),
      'roofline': {'bound': 'mfma', 'achieved': achieved,
                   'peak': BF16_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                   'frac': achieved / BF16_PEAK_TFLOPS,
                   'traffic': C2_HBM_BYTES_PER_GRADIENT,
                   'traffic_what': 'HBM bytes per gradient launch '
                                   '(FETCH_SIZE + WRITE_SIZE, '
                                   'copy-calibrated); X and X^T are 64 MiB',
                   'kernel': 'glm_grad_mfma<8> + mfma_chain_step (sampling '
                             'launch, all 1000 transitions)',
                   'algorithmic_flop_per_gradient': flop_grad,
                   'executed_mfma_flop_per_gradient': 2 * flop_grad,
                   'sampling_ms': sample_ms}})
  if not args.no_cpu:
    line['cpu_baseline'] = _cpu_glm_hmc(
        x, y, c, L, float(res.final_kernel_results.new_step_size), nb + nr,
        min_ess, args.cpu_seconds, 'the 65536 x 256 problem',
        sample_chains=256, one_thread_chains=32)
  return line


# ------------------------------------------------------ C3 ragged batched
def _ragged_specs(torch, pa, n_problems):
  from probability_amd import workloads
  specs, bij = [], []
  for prob in workloads.ragged_glms(n_problems):
    if prob['kind'] == 'logistic':
      specs.append(pa.targets.LogisticRegression(prob['x'], prob['y'],
                                                 device='cuda'))
      bij.append(None)
    else:
      specs.append(pa.targets.LinearRegression(prob['x'], prob['y'],
                                               device='cuda'))
      bij.append([pa.bijectors.Identity(), pa.bijectors.Exp()])
  return specs, bij


C3_STEP_TRANSITIONS = 75     # one bench step: 1/20 of a 1500-transition run
# transitions per sampling launch: the whole sampling phase in one launch per
# size class (each launch ends with a tail in which its last groups finish
# alone; 100 per launch took 33.7 s per run against 31.1 s for 1000)
C3_ITERS_PER_LAUNCH = 1000
# HBM bytes the GLM NUTS kernels (one chain per block and the chain
# groups) move per design-matrix entry streamed: rocprofv3 FETCH_SIZE
# (907 401 286 KiB) x 2.00 + WRITE_SIZE (169 728 112 KiB) x 1.00 over a
# shortened run of this workload that streamed 4.763e12 entries, the
# factors being what the counters read for a 256 MiB device copy
# (profiles/nuts_kernel_glm_ragged512.md); most X reads hit L2.
C3_HBM_BYTES_PER_ENTRY = (907401286 * 2.0 + 169728112 * 1.0) * 1024 / 4.763e12


# sampling launches of the 8- and 4-wave classes run chain groups of 4
# (one pass over X per gradient for 4 chains of a problem)
C3_CHAINS_PER_BLOCK = {8: 4, 4: 4}
# size classes split into independent batches (by --c3-class-splits); none
# by default: with the runtime's 4 hardware queues per process the extra
# streams share queues and the run slows (8:4 48 s against 34.6 s unsplit;
# with 8 queues 8:4 takes 32.2 s), because every adaptation step is set by
# its longest trajectory, which a split does not shorten
C3_CLASS_SPLITS = {}


def _class_splits(args):
  """--c3-class-splits 'w:n,...' or the default."""
  text = getattr(args, 'c3_class_splits', None)
  if not text:
    return dict(C3_CLASS_SPLITS)
  out = {}
  for part in text.split(','):
    w, n = part.split(':')
    out[int(w)] = int(n)
  return out


def _init_steps(specs, adaptation):
  """Initial step sizes: windowed adaptation starts each problem at the
  reference's windowed default 0.5 D^-1/4 (`_get_step_size`); dual
  averaging alone at 0.05."""
  if adaptation == 'windowed':
    return [0.5 * float(s.dim) ** -0.25 for s in specs]
  return 0.05


def _ragged_run(pa, specs, bij, seed, adaptation, na, nr, tap,
                group_adaptation=False, class_splits=None):
  from probability_amd.mcmc import batched
  return batched.RaggedRun(
      specs, 32, nr, bijectors=bij, slot_dim=192, seed=seed,
      num_burnin_steps=na, num_adaptation_steps=na,
      step_size=_init_steps(specs, adaptation),
      target_accept_prob=tap, max_tree_depth=10,
      max_iters_per_launch=C3_ITERS_PER_LAUNCH,
      adaptation=adaptation, stats={'time_kernels': False},
      chains_per_block=C3_CHAINS_PER_BLOCK,
      group_adaptation=group_adaptation, class_splits=class_splits)


def _ragged_gate(torch, pa, draws):
  """Per-problem (min ESS, max split R-hat); a problem counts toward the
  metric only at split R-hat < 1.01."""
  ess, rhat = [], []
  for dr in draws:
    e, r = _ess_rhat(torch, pa, dr)
    ess.append(float(e.min()))
    rhat.append(float(r.max()))
  return np.array(ess), np.array(rhat)


def ragged_glms_512(args, torch, pa):
  """512 ragged GLM posteriors (C3), the default workload.  One step is
  C3_STEP_TRANSITIONS transitions of all 512 problems, so 20 steps are one
  complete run (500 adaptation + 1000 draws); the timed region holds exactly
  `steps` steps: complete runs with seeds seed, seed + 1, ... (a last partial
  run is timed but its draws are not counted), each enqueued up to its last
  timed transition at once, so that launches are not cut at the steps'
  boundaries.  Warm-up steps run a separate run's first transitions.
  value = sum of min-ESS over the problems at split R-hat < 1.01 (failing
  problems count zero) / timed seconds."""
  from probability_amd import _lib
  from probability_amd.mcmc import batched
  p, chains = 512, 32
  na = 500 if args.adapt is None else args.adapt
  nr = 1000 if args.draws is None else args.draws
  total = na + nr
  per_step = max(1, total // 20) if (args.adapt or args.draws) else \
      C3_STEP_TRANSITIONS
  adaptation = 'dual_averaging' if args.c3_dual_averaging else 'windowed'
  tap = 0.8 if adaptation == 'dual_averaging' else 0.85
  specs, bij = _ragged_specs(torch, pa, p)
  _say(f'ragged: {adaptation} adaptation, {per_step} transitions per step')
  if args.warmup:
    wr = _ragged_run(pa, specs, bij, args.seed + 999, adaptation, na, nr, tap,
                     args.c3_group_adaptation, _class_splits(args))
    wr.advance_to(min(total, args.warmup * per_step))
    torch.cuda.synchronize()
    del wr
  done_runs = []
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  # the timed steps' transitions, run after run
  timed = args.steps * per_step
  left, k = timed, 0
  while left > 0:
    n = min(total, left)
    run = _ragged_run(pa, specs, bij, args.seed + k, adaptation, na, nr, tap,
                      args.c3_group_adaptation, _class_splits(args))
    run.advance_to(n)
    if n >= total:
      done_runs.append((run.finish(), run.stats))
    torch.cuda.synchronize()
    left -= n
    k += 1
    _say(f'ragged: run {k} done, {timed - left} of {timed} timed '
         'transitions')
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  ms = 1000.0 * wall / max(args.steps, 1)
  _say(f'ragged: timed {wall:.1f} s; diagnostics')
  gates = [_ragged_gate(torch, pa, res.draws) for res, _ in done_runs]
  if gates:
    ess, rhat = gates[-1]
  else:
    ess, rhat = np.zeros(p), np.full(p, np.inf)
  ok = rhat < 1.01
  counted = float(sum(float(e[r < 1.01].sum()) for e, r in gates))
  value = counted / wall
  # one posterior per launch: a stratified sample of problems, each alone
  # with the same code, keys, size class and adaptation, one chain per
  # block (chain groups only pay when many problems share the launches; a
  # problem alone is faster without them, and its draws then differ from
  # the batch's only by the groups' summation order)
  idx = np.linspace(0, p - 1, max(1, args.plain_sample)).round().astype(int)
  plain_s = []
  for i in idx:
    one = batched.RaggedRun([specs[i]], chains, nr, bijectors=[bij[i]],
                            keys=[int(i)], slot_dim=192, seed=args.seed,
                            num_burnin_steps=na, num_adaptation_steps=na,
                            step_size=_init_steps([specs[i]], adaptation),
                            target_accept_prob=tap,
                            max_tree_depth=10,
                            max_iters_per_launch=C3_ITERS_PER_LAUNCH,
                            adaptation=adaptation)
    torch.cuda.synchronize()
    ta = time.perf_counter()
    one.advance_to(total)
    torch.cuda.synchronize()
    plain_s.append(time.perf_counter() - ta)
  plain_total = float(np.mean(plain_s)) * p
  per_run_s = wall / max(len(done_runs), 1) if done_runs else wall
  line = _base(
      'min-ESS/sec at split-Rhat<1.01 over 512 ragged GLM posteriors '
      '(NUTS, per-problem step-size and mass-matrix adaptation, 32 chains '
      'each), batched vs one posterior per launch',
      value, 'min-ESS/s', args, ms, 'f32',
      {'workload': 'ragged_glms_512: 256 logistic + 256 linear (Exp on '
                   'sigma) GLMs, N~LogU[128,8192], D~U{4..128}, X iid N(0,1) '
                   'f32, 32 chains each, NUTS(max_tree_depth 10) with '
                   f'per-problem {adaptation} adaptation (target {tap}), '
                   f'{na} adaptation + {nr} draws; one step = {per_step} '
                   'transitions of all 512 problems',
       'problems': p, 'chains_per_problem': chains,
       'adaptation': adaptation, 'transitions_per_step': per_step,
       'complete_runs_timed': len(done_runs),
       'parallelism': 'single GPU; problems batched in size classes on '
                      'concurrent streams'})
  line.update({
      'problems_passing_rhat': int(ok.sum()),
      'problems_failing_rhat': int((~ok).sum()),
      'rhat_max': float(rhat.max()),
      'min_ess_passing_sum': float(ess[ok].sum()),
      'min_ess_per_problem_min_median': [float(ess.min()),
                                         float(np.median(ess))],
      'run_seconds': per_run_s,
      'plain': {'value': float(ess[ok].sum()) / plain_total,
                'seconds_per_run': plain_total,
                'what': f'each problem alone in its own launches '
                        f'({len(idx)} problems timed, mean x 512)'},
      'batched_over_plain': plain_total / per_run_s,
  })
  if done_runs:
    line['roofline'] = _ragged_roofline(specs, done_runs[-1][1], per_run_s)
  if args.c3_compare and adaptation == 'windowed':
    rr = _ragged_run(pa, specs, bij, args.seed, 'dual_averaging', na, nr, 0.8,
                     class_splits=_class_splits(args))
    torch.cuda.synchronize()
    ta = time.perf_counter()
    rr.advance_to(total)
    res = rr.finish()
    torch.cuda.synchronize()
    td = time.perf_counter() - ta
    e2, r2 = _ragged_gate(torch, pa, res.draws)
    line['dual_averaging_only'] = {
        'value': float(e2[r2 < 1.01].sum()) / td, 'seconds': td,
        'problems_passing_rhat': int((r2 < 1.01).sum()),
        'rhat_max': float(r2.max())}
  if not args.no_cpu:
    line['cpu_baseline'] = _ragged_cpu(args, specs, line, ess, ok)
  return line


def _ragged_roofline(specs, stats, run_s):
  """The GLM kernel's arithmetic over one run: every chain-leapfrog takes
  its problem's logits and gradient in one pass over X, two FMAs (4 FLOP)
  per design-matrix entry, summed over the run's leapfrogs, over the run's
  wall time (the size classes' kernels overlap on their streams and cover
  nearly all of it), against the FP32 peak.  The adaptation steps run one
  chain per block on FP32 VALU passes; the sampling launches of the 8- and
  4-wave classes run groups of four chains on v_mfma_f32_4x4x1_16b_f32.
  The design-matrix bytes those leapfrogs read (N_k D_k 4 B each, ~95% from
  L2) are reported beside it, with the HBM traffic from counters."""
  lpp = stats['leapfrogs_per_problem']
  entries = float(sum(float(lpp[i]) * s.n_obs * s.dim
                      for i, s in enumerate(specs)))
  achieved = 4.0 * entries / run_s / 1e12
  out = {'bound': 'mfma', 'achieved': achieved, 'peak': FP32_PEAK_TFLOPS,
         'unit': 'TFLOP/s', 'frac': achieved / FP32_PEAK_TFLOPS,
         'kernel': 'nuts_kernel<float,3,glm> (all size classes)',
         'what': 'GLM logit and gradient FLOP (4 per design-matrix entry '
                 'per chain-leapfrog) over the run\'s wall time',
         'limiter': 'adaptation steps of the 8-wave class: each lasts '
                    'as long as its longest trajectory at one CU\'s rate '
                    '(one chain per block); sampling launches: chain '
                    'groups of 4 on FP32 MFMA, bound by fetching X from L2 '
                    '(~15 B/clk per CU); timed alone at 8192 x 128 with '
                    '2048 chains: 0.17 us per chain-leapfrog one chain per '
                    'block, 0.11 us in groups',
         'design_matrix_GBps': entries * 4.0 / run_s / 1e9,
         'chain_leapfrogs': float(lpp.sum()),
         'design_matrix_entries': entries,
         'traffic': None}
  if C3_HBM_BYTES_PER_ENTRY is not None:
    out['traffic'] = entries * C3_HBM_BYTES_PER_ENTRY
    out['traffic_what'] = ('HBM bytes over the run from FETCH_SIZE + '
                           'WRITE_SIZE per design-matrix entry, '
                           'copy-calibrated')
  return out


class _CpuGlm:
  """The GLM log-density and gradient in NumPy with the design-matrix
  products in float32 (the CPU restatement's target; oracle.targets holds
  the float64 one)."""

  def __init__(self, spec_x, spec_y, linear):
    self.x = np.ascontiguousarray(spec_x, np.float32)
    self.y = np.asarray(spec_y, np.float32)
    self.linear = linear

  def log_prob_and_grad(self, q):
    q = np.asarray(q, np.float64)
    if self.linear:
      beta, ls = q[:, :-1], q[:, -1]
      sig = np.exp(ls)
      s = (self.x @ beta.T.astype(np.float32)).T.astype(np.float64)
      res = self.y[None, :] - s
      n = self.x.shape[0]
      lp = (-0.5 * (res ** 2).sum(1) / sig ** 2 - n * np.log(sig)
            - 0.5 * (beta ** 2).sum(1) - 0.5 * sig ** 2 + ls)
      gb = (res.astype(np.float32) @ self.x).astype(np.float64) / \
          sig[:, None] ** 2 - beta
      gs = ((res ** 2).sum(1) / sig ** 3 - n / sig - sig) * sig + 1.0
      return lp, np.concatenate([gb, gs[:, None]], 1)
    s = (self.x @ q.T.astype(np.float32)).T.astype(np.float64)
    lp = (self.y[None] * s - np.logaddexp(0, s)).sum(1) - 0.5 * (q ** 2).sum(1)
    r = (self.y[None] - 1 / (1 + np.exp(-s))).astype(np.float32)
    return lp, (r @ self.x).astype(np.float64) - q


class _CpuHier:
  """The non-centred hierarchical normal with Softplus on tau, in float32
  NumPy over unconstrained [mu, tau_u, z] (the CPU restatement's target;
  oracle.targets holds the float64 one): mu ~ N(0, 5), tau ~ HalfCauchy(0,
  5), z ~ N(0, 1), y ~ N(mu + tau z, sigma), plus log sigmoid(tau_u), the
  Softplus log-Jacobian."""

  def __init__(self, y, sigma, mu_scale=5.0, tau_scale=5.0):
    self.y = np.asarray(y, np.float32)
    self.s2 = np.asarray(sigma, np.float32) ** 2
    j = len(self.y)
    # normalising constants: the likelihood's and z's Gaussians, mu's
    # Gaussian, tau's half-Cauchy
    self.c = np.float32(-np.log(np.asarray(sigma, np.float64)).sum()
                        - j * np.log(2 * np.pi)
                        - 0.5 * np.log(2 * np.pi) - np.log(mu_scale)
                        + np.log(2.0 / (np.pi * tau_scale)))
    self.ms, self.ts = np.float32(mu_scale), np.float32(tau_scale)

  def log_prob_and_grad(self, q):
    q = np.asarray(q, np.float32)
    mu, tu, z = q[:, 0], q[:, 1], q[:, 2:]
    tau = np.logaddexp(np.float32(0), tu).astype(np.float32)
    sg = (1 / (1 + np.exp(-tu))).astype(np.float32)
    r = (self.y - (mu[:, None] + tau[:, None] * z)) / self.s2
    lp = (-0.5 * (mu / self.ms) ** 2
          - np.log1p((tau / self.ts) ** 2) + np.log(sg)
          - 0.5 * (z * z).sum(1)
          - 0.5 * (r * r * self.s2).sum(1) + self.c)
    g_tau = -2 * tau / (self.ts ** 2 + tau * tau) + (r * z).sum(1)
    g = np.empty_like(q)
    g[:, 0] = -mu / self.ms ** 2 + r.sum(1)
    g[:, 1] = g_tau * sg + (1 - sg)
    g[:, 2:] = -z + tau[:, None] * r
    return lp.astype(np.float64), g.astype(np.float64)


def _cpu_nuts_worker(target, q, eps, budget_s, seed):
  """NUTS transitions (oracle.mcmc.nuts_one_step) of the chains q on one
  core for budget_s seconds after one warm-up transition; returns
  (transitions, seconds, chain-leapfrogs)."""
  from threadpoolctl import threadpool_limits
  from oracle import mcmc as omcmc
  from oracle import rng as orng
  with threadpool_limits(limits=1):
    t, g = target.log_prob_and_grad(q)
    n, lf, t0 = 0, 0, None
    while t0 is None or time.perf_counter() - t0 < budget_s:
      q, r = omcmc.nuts_one_step(target, eps, q, t, g, orng.fold_in(seed, n),
                                 orng.PhiloxDraws(np.float32), max_depth=10)
      t, g = r.target_log_prob, r.grads_target_log_prob
      if t0 is None:
        t0 = time.perf_counter()
      else:
        n += 1
        lf += int(r.leapfrogs_taken.sum())
    return n, time.perf_counter() - t0, lf


def _cpu_nuts_one(target, q, eps, budget_s):
  """Seconds per NUTS transition of the chains q on one core: the median of
  3 timed stretches of budget_s / 3 seconds."""
  reps = [_cpu_nuts_worker(target, q, eps, budget_s / 3, 5) for _ in range(3)]
  return float(np.median([s / max(n, 1) for n, s, _ in reps]))


def _cpu_nuts_all(jobs, budget_s):
  """The jobs (target, chains q, step size), one process per job, all at
  once (spawned children, one per core of the affinity mask); seconds per
  transition of each job under that load."""
  import concurrent.futures
  import multiprocessing
  ctx = multiprocessing.get_context('spawn')
  with concurrent.futures.ProcessPoolExecutor(len(jobs),
                                              mp_context=ctx) as ex:
    futs = [ex.submit(_cpu_nuts_worker, tg, q, eps, budget_s, 5 + k)
            for k, (tg, q, eps) in enumerate(jobs)]
    return [s / max(n, 1) for n, s, _ in (f.result() for f in futs)]


def _ragged_cpu(args, specs, line, ess, ok):
  """TFP's batched NUTS restated in NumPy (oracle.mcmc.nuts_one_step over
  the float32 GLM above), timed on the host: three problems (smallest,
  median and largest N x D), post-adaptation transitions of their 32
  chains within a time budget, as the median of 3 repeats after a warm-up,
  with BLAS on every core of the process's affinity mask and on one.  The
  per-entry cost (seconds per chain-leapfrog per design-matrix entry) times
  the GPU run's entries gives the CPU time of the whole run."""
  import os
  from threadpoolctl import threadpool_limits
  from oracle import mcmc as omcmc
  from oracle import rng as orng
  cores = len(os.sched_getaffinity(0))
  sizes = np.array([s.n_obs * s.dim for s in specs])
  pick = [int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2]),
          int(np.argmax(sizes))]
  budget = max(1.0, args.cpu_seconds / 6.0)

  def cost(threads):
    per_entry = []
    with threadpool_limits(limits=threads):
      for i in pick:
        s = specs[i]
        lin = s.kind == 2
        tg = _CpuGlm(s.x.cpu().numpy()[:, :s.dim - (1 if lin else 0)],
                     s.y.cpu().numpy(), lin)
        q = np.zeros((32, s.dim))
        t, g = tg.log_prob_and_grad(q)
        reps = []
        for rep in range(4):   # first is warm-up
          lf, n_it, t0 = 0, 0, time.perf_counter()
          while time.perf_counter() - t0 < budget / 4 or n_it == 0:
            q, rr = omcmc.nuts_one_step(tg, 0.02, q, t, g,
                                        orng.fold_in(5, n_it),
                                        orng.PhiloxDraws(np.float32),
                                        max_depth=10)
            t, g = rr.target_log_prob, rr.grads_target_log_prob
            lf += int(rr.leapfrogs_taken.sum())
            n_it += 1
          if rep:
            reps.append((time.perf_counter() - t0) / max(lf, 1) /
                        (s.n_obs * s.dim))
        per_entry.append(float(np.median(reps)))
    return float(np.mean(per_entry))

  entries = line.get('roofline', {}).get('design_matrix_entries', 0.0)
  c_all = cost(cores)
  c_one = cost(1)
  cpu_s = entries * c_all
  return {'value': float(ess[ok].sum()) / cpu_s, 'unit': 'min-ESS/s',
          'cores': cores, 'kind': 'port',
          'sample': (f'NUTS transitions of 32 chains on 3 of the 512 problems '
                     f'(N x D = {", ".join(str(int(sizes[i])) for i in pick)}),'
                     ' median of 3 after a warm-up, NumPy with float32 '
                     f'design-matrix products on {cores} BLAS threads; per '
                     'design-matrix entry cost x the GPU run\'s entries = '
                     f'{cpu_s:.0f} s for the run'),
          'seconds_per_run': cpu_s,
          'one_thread': {'value': float(ess[ok].sum()) / (entries * c_one),
                         'seconds_per_run': entries * c_one, 'cores': 1}}


# ---------------------------------------------------------- C4 serving
def serving_poisson200(args, torch, pa):
  from probability_amd import serving
  from probability_amd import workloads
  n_req = args.requests
  rate = 200.0
  arrivals, models = workloads.serving_requests(n_req, seed=11, rate=rate)
  cfgs = []
  for prob in models:
    if prob['kind'] == 'logistic':
      spec = pa.targets.LogisticRegression(prob['x'], prob['y'],
                                           device='cuda')
      bij = None
    elif prob['kind'] == 'linear':
      spec = pa.targets.LinearRegression(prob['x'], prob['y'],
                                         device='cuda')
      bij = [pa.bijectors.Identity(), pa.bijectors.Exp()]
    else:
      spec = pa.targets.HierarchicalNormal(prob['y'], prob['sigma'],
                                           device='cuda')
      bij = [pa.bijectors.Identity(), pa.bijectors.Softplus(),
             pa.bijectors.Identity()]
    cfgs.append((spec, bij))
  # one MI355X holds 288 GB of HBM: the server may keep 192 GiB of admitted
  # requests' data and draws resident
  srv = serving.SamplerServer(iters_per_tick=16,
                              hbm_budget_bytes=args.hbm_budget_gib << 30)
  reqs = [serving.Request(spec, num_chains=16, num_adaptation_steps=500,
                          num_results=2000, seed=args.seed + k, bijectors=bij,
                          step_size=0.05, key=k)
          for k, (spec, bij) in enumerate(cfgs)]
  lat, ess_min, rhat_max, lf_total = [], [], [], 0
  t0 = time.perf_counter()
  nxt = 0
  last_print = t0
  while nxt < n_req or not srv.idle():
    now = time.perf_counter() - t0
    while nxt < n_req and arrivals[nxt] <= now:
      reqs[nxt].arrival = t0 + arrivals[nxt]
      srv.submit(reqs[nxt])
      nxt += 1
    if srv.idle():
      time.sleep(max(0.0, min(arrivals[nxt] - now, 0.01)))
      continue
    for req in srv.tick():
      lat.append(req.done - req.arrival)
      e, rh = _ess_rhat(torch, pa, req.draws)
      ess_min.append(float(e.min()))
      rhat_max.append(float(rh.max()))
      req.draws = None
    if time.perf_counter() - last_print > 20:
      last_print = time.perf_counter()
      _say(f'serving: {len(lat)}/{n_req} done, {len(srv.active)} active, '
           f'{len(srv.queue)} queued, t={last_print - t0:.0f}s')
  wall = time.perf_counter() - t0
  lat = np.array(lat)
  ess_min = np.array(ess_min)
  rhat_max = np.array(rhat_max)
  ok = rhat_max < 1.01
  line = _base('served posterior-sampling requests: throughput with p50/p99 '
               'latency (Poisson 200 req/s, NUTS+DA, 16 chains)',
               n_req / wall, 'req/s', args, 1000.0 * wall, 'f32',
               {'workload': f'serving_poisson200: {n_req} requests, half '
                            'ragged GLMs (N~LogU[128,8192], D~U{4..128}), '
                            'half hierarchical normals (J~U{8..1024}, Softplus '
                            'on tau); 16 chains, 500 adaptation + 2000 draws; '
                            'one step = the whole arrival process',
                'requests': n_req, 'arrival_rate': rate,
                'hbm_budget_gib': args.hbm_budget_gib},
               steps=1, warmup=0)
  line.update({
      'latency_s_p50_p99': [float(np.percentile(lat, 50)),
                            float(np.percentile(lat, 99))],
      'min_ess_per_request_min_median': [float(ess_min.min()),
                                         float(np.median(ess_min))],
      'aggregate_min_ess_per_s': float(ess_min[ok].sum()) / wall,
      'aggregate_counts': 'requests at split-Rhat < 1.01 only',
      'requests_failing_rhat': int((~ok).sum()),
      'rhat_ok_fraction': float(ok.mean()),
      'rhat_max_median': float(np.median(rhat_max)),
      'ticks': srv.ticks})
  if not args.no_cpu:
    from probability_amd import _lib
    glm = next(r for r in reqs if r.spec.kind != _lib.TARGET_HIER_NORMAL)
    hier = next(r for r in reqs if r.spec.kind == _lib.TARGET_HIER_NORMAL)
    line['cpu_baseline'] = _serving_cpu(args, [glm, hier], 2500)
  return line


def _serving_cpu(args, samples, transitions):
  """Requests served by TFP's NUTS restated in NumPy (oracle.mcmc over the
  float32 targets above): post-adaptation NUTS transitions of 16 chains of a
  sampled GLM and hierarchical request at their adapted step sizes, on one
  core (median of 3) and with every core busy (one process per core, half
  on each request).  A request costs its seconds per transition x
  `transitions` of one core; the cores serve the half-and-half mix at
  cores / (mean cost) requests per second."""
  import os
  from probability_amd import _lib
  cores = len(os.sched_getaffinity(0))
  jobs, what = [], []
  for r in samples:
    spec = r.spec
    if spec.kind == _lib.TARGET_HIER_NORMAL:
      tg = _CpuHier(spec.y.cpu().numpy(), spec.sigma.cpu().numpy())
    else:
      lin = spec.kind == _lib.TARGET_LINEAR
      tg = _CpuGlm(spec.x.cpu().numpy()[:, :spec.dim - (1 if lin else 0)],
                   spec.y.cpu().numpy(), lin)
    jobs.append((tg, np.zeros((r.num_chains, spec.dim)),
                 float(r.final_step_size)))
    what.append(f'kind {spec.kind} dim {spec.dim}')
  budget = max(1.0, args.cpu_seconds / 4.0)
  one = [_cpu_nuts_one(tg, q, eps, budget / 2) for tg, q, eps in jobs]
  loaded = _cpu_nuts_all([jobs[k % len(jobs)] for k in range(cores)], budget)
  per_kind = [float(np.mean(loaded[k::len(jobs)])) for k in range(len(jobs))]
  s_one = float(np.mean(one)) * transitions
  s_all = float(np.mean(per_kind)) * transitions
  return {'value': cores / s_all, 'unit': 'req/s', 'cores': cores,
          'kind': 'port',
          'sample': ('NUTS transitions of a GLM and a hierarchical request '
                     '(' + ', '.join(what) + '; 16 chains at their adapted '
                     'step sizes), float32 NumPy targets; one core (median '
                     'of 3) and one process per core; x ' + str(transitions) +
                     ' transitions per request'),
          'core_seconds_per_request_loaded': s_all,
          'one_thread': {'value': 1.0 / s_one, 'cores': 1,
                         'seconds_per_request': s_one}}


# ------------------------------------------------------------ reducers
def reducers_1024x258(args, torch, pa):
  """Streaming reducer updates (hipmc_moments_update) per draw at the
  headline's shape.  Algorithmic HBM bytes per update: the draw (C D 4 B)
  plus reading and writing the float64 sums: C D 16 B for the mean, C D 16 B
  more for per-element variances, C D^2 16 B more for per-chain
  covariances."""
  from probability_amd.experimental import mcmc as emcmc
  c, d, draws = 1024, 258, 64
  g = torch.Generator(device='cuda').manual_seed(args.seed)
  xs = torch.randn(draws, c, d, device='cuda', generator=g)
  kinds = {
      'mean': (emcmc.ExpectationsReducer(), c * d * (4 + 16)),
      'variance': (emcmc.VarianceReducer(), c * d * (4 + 32)),
      'covariance': (emcmc.CovarianceReducer(event_ndims=1),
                     c * d * (4 + 16) + c * d * d * 16),
      'rhat': (emcmc.PotentialScaleReductionReducer(), c * d * (4 + 32)),
  }
  out = {}
  for name, (red, nbytes) in kinds.items():

    def step(k, red=red):
      s = red.initialize(xs[0])
      for t in range(draws):
        s = red.one_step(xs[t], s)
      return s
    _, ms = _timed(torch, step, args.steps, args.warmup)
    # kernel time alone, HIP events on the stream the updates run on
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(
        enable_timing=True)
    st = red.initialize(xs[0])
    e0.record()
    for t in range(draws):
      st = red.one_step(xs[t], st)
    e1.record()
    torch.cuda.synchronize()
    kms = e0.elapsed_time(e1) / draws
    out[name] = {'ms_per_update': ms / draws, 'event_ms_per_update': kms,
                 'bytes_per_update': nbytes,
                 'achieved_GBps': nbytes / (kms * 1e-3) / 1e9}
    _say(f'reducers: {name} {kms * 1e3:.1f} us per update, '
         f'{out[name]["achieved_GBps"]:.0f} GB/s')
  cov = out['covariance']
  ms = sum(v['ms_per_update'] for v in out.values()) * draws
  line = _base('draws reduced per second (mean + variance + covariance + '
               'R-hat reducers, 1024 chains x 258)', draws / (ms / 1e3),
               'draws/s', args, ms, 'f64 sums of f32 draws',
               {'workload': 'reducers_1024x258: ExpectationsReducer, '
                            'VarianceReducer, CovarianceReducer(event_ndims=1) '
                            'and PotentialScaleReductionReducer, one update '
                            'per draw, 64 draws per step', 'chains': c,
                'dim': d, 'draws_per_step': draws})
  line.update({
      'reducers': out,
      'roofline': {'bound': 'hbm', 'achieved': cov['achieved_GBps'],
                   'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                   'frac': cov['achieved_GBps'] / HBM_PEAK_GBS,
                   'traffic': None,
                   'kernel': 'moments_cross_kernel + moments_elem_kernel '
                             '(covariance update)',
                   'bytes_per_update': cov['bytes_per_update']}})
  return line


def run(args, torch, pa):
  fn = {'logistic_1024x25': logistic_1024x25,
        'logistic_65536x256': logistic_65536x256,
        'ragged_glms_512': ragged_glms_512,
        'serving_poisson200': serving_poisson200,
        'reducers_1024x258': reducers_1024x258}[args.workload]
  return fn(args, torch, pa)
