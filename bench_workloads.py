"""The non-headline workloads of bench.py (one JSON line each).

* logistic_1024x25    plain HMC on logistic regression, X [1024, 25], 64 chains,
                      L = 10, eps = 0.05, 500 burn-in + 1000 results.
* logistic_65536x256  HMC (L = 20) + dual averaging (target 0.75, 500 burn-in
                      steps) + 1000 results on X [65536, 256] stored BF16,
                      4096 chains; the all-chain gradient runs as BF16 MFMA
                      GEMMs with a hi/lo split of beta.
* ragged_glms_512     512 GLMs (half logistic, half Gaussian linear with
                      sigma ~ HalfNormal(1) under an Exp bijector),
                      N_k ~ LogUniform[128, 8192], D_k ~ U{4..128}, 32 chains
                      each, NUTS + per-problem dual averaging (target 0.8),
                      500 adaptation + 1000 draws, all problems in shared
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


def _base(metric, value, unit, args, ms, dtype, config, higher=True):
  return {'metric': metric, 'value': value, 'unit': unit, 'n_gpus': 1,
          'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': ms,
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
                   'kernel': 'hmc_kernel<float,1,glm,calibrated> (all 1500 '
                             'transitions in one launch)',
                   'kernel_ms': kern_ms,
                   'flop_per_chain_leapfrog': 4 * n * d}})
  if not args.no_plain:
    _, pms = _timed(torch, lambda k: run(k, fused=False), 1, 0)
    line['plain'] = {'value': min_ess / (pms / 1e3), 'ms_per_step': pms,
                     'what': 'host step loop, one hmc launch per transition'}
  if not args.no_cpu:
    line['cpu_baseline'] = _cpu_hmc(wl, min_ess, args.cpu_seconds)
  return line


def _cpu_hmc(wl, min_ess, budget_s):
  from oracle import mcmc as omcmc
  from oracle import rng as orng
  from oracle import targets as otg
  ot = otg.LogisticRegression(wl['x'].astype(np.float64),
                              wl['y'].astype(np.float64))
  c, d = wl['chains'], wl['x'].shape[1]
  q = np.zeros((c, d))
  t, g = ot.log_prob_and_grad(q)
  draws = orng.PhiloxDraws(np.float32)
  n_iter = 0
  t0 = time.perf_counter()
  while time.perf_counter() - t0 < budget_s:
    q, r = omcmc.hmc_one_step(ot, wl['step_size'], wl['num_leapfrog_steps'],
                              q, t, g, orng.fold_in(7, n_iter), draws)
    t, g = r.target_log_prob, r.grads_target_log_prob
    n_iter += 1
  el = time.perf_counter() - t0
  total = wl['num_burnin_steps'] + wl['num_results']
  return {'value': float(min_ess / (el / n_iter * total)), 'unit': 'min-ESS/s',
          'cores': 1, 'kind': 'port',
          'sample': (f'{n_iter} HMC transitions of the 64-chain batch '
                     f'({el:.1f} s), extrapolated to {total} transitions with '
                     'the GPU run\'s min-ESS; NumPy float64, 1 thread'),
          'chain_grads_per_s': float(n_iter * c * wl['num_leapfrog_steps'] /
                                     el)}


# ------------------------------------------------ C2 large-batch MFMA HMC
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

  steps = max(1, min(args.steps, 1))
  _say('mfma: runs')
  stats = {'time_kernels': True}
  res, ms = _timed(torch, lambda k: run(k, stats=stats), steps,
                   min(args.warmup, 1))
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
                            'N(0,1)/16 in BF16, 4096 chains at 0',
                'chains': c, 'n_obs': n, 'dim': d})
  line.update({
      'min_ess': min_ess, 'rhat_max': float(rhat.max()),
      'rhat_ok': float(rhat.max()) < 1.01, 'accept_rate': acc,
      'final_step_size': float(res.final_kernel_results.new_step_size),
      'mean_halves_max_z': z_halves,
      'chain_grad_evals_per_s': c * (nb + nr) * L / (ms / 1e3),
      'roofline': {'bound': 'mfma', 'achieved': achieved,
                   'peak': BF16_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                   'frac': achieved / BF16_PEAK_TFLOPS, 'traffic': None,
                   'kernel': 'glm_grad_mfma + mfma_chain_step (sampling '
                             'launch, all 1000 transitions)',
                   'algorithmic_flop_per_gradient': flop_grad,
                   'executed_mfma_flop_per_gradient': 2 * flop_grad,
                   'sampling_ms': sample_ms}})
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


def ragged_glms_512(args, torch, pa):
  from probability_amd import _lib
  from probability_amd.mcmc import batched
  p, chains = 512, 32
  na = 500 if args.adapt is None else args.adapt
  nr = 1000 if args.draws is None else args.draws
  specs, bij = _ragged_specs(torch, pa, p)
  slot = 192
  # the sampling phase runs 100 transitions per launch so that progress is
  # visible between launches
  kw = dict(num_results=nr, num_burnin_steps=na, num_adaptation_steps=na,
            step_size=0.05, target_accept_prob=0.8, max_tree_depth=10,
            max_iters_per_launch=100)
  mark = [0]

  def report(t, n):
    if t >= mark[0] + 100 or t == n:
      mark[0] = 0 if t == n else t
      torch.cuda.synchronize()
      _say(f'ragged: {t}/{n} transitions')

  if args.glm_waves is None:
    # size classes: each problem's chains are shared by the waves its own
    # N x D asks for; the classes run concurrently on their own streams
    def run(k, stats=None):
      return batched.sample_ragged(specs, chains, bijectors=bij,
                                   slot_dim=slot, seed=args.seed + k,
                                   stats=stats, progress=report, **kw)

    def waves_of(i):
      return _lib.glm_waves_per_chain(specs[i].n_obs * specs[i].dim)
  else:
    batch = batched.PosteriorBatch(specs, chains, bijectors=bij,
                                   slot_dim=slot, waves_per_chain=args.glm_waves)

    def run(k, stats=None):
      return batched.sample_posteriors(batch, seed=args.seed + k, stats=stats,
                                       progress=report, **kw)

    def waves_of(i):
      return batch.waves_per_chain

  _say('ragged: batched runs')
  stats = {'time_kernels': True}
  res, ms = _timed(torch, lambda k: run(k, stats=stats),
                   max(1, min(args.steps, 1)), min(args.warmup, 1))
  _say(f'ragged: batched {ms:.0f} ms; diagnostics')
  min_ess, rhat_max = [], []
  for dr in res.draws:
    e, r = _ess_rhat(torch, pa, dr)
    min_ess.append(float(e.min()))
    rhat_max.append(float(r.max()))
  min_ess = np.array(min_ess)
  rhat_max = np.array(rhat_max)
  lf_total = stats['leapfrogs_total']
  nd = sum(s.n_obs * s.dim for s in specs)
  # one posterior per launch on a stratified sample of problems (the same
  # code and keys, so each problem's draws equal its draws in the batch)
  idx = np.linspace(0, p - 1, max(1, args.plain_sample)).round().astype(int)
  plain_ms = []
  for i in idx:
    _say(f'ragged: problem {i} alone')
    one = batched.PosteriorBatch([specs[i]], chains, bijectors=[bij[i]],
                                 keys=[i], slot_dim=slot,
                                 waves_per_chain=waves_of(i))
    _, pm = _timed(torch, lambda k: batched.sample_posteriors(
        one, seed=args.seed + k, progress=report, **kw), 1, 0)
    plain_ms.append(pm)
  plain_total_ms = float(np.mean(plain_ms)) * p
  agg = float(min_ess.sum()) / (ms / 1e3)
  classes = stats.get('classes', [(p, waves_of(0))])
  cls_stats = stats.get('class_stats', [stats])
  line = _base('aggregate min-ESS/sec over 512 ragged GLM posteriors (NUTS+DA, '
               '32 chains each), batched vs one posterior per launch',
               agg, 'min-ESS/s', args, ms, 'f32',
               {'workload': 'ragged_glms_512: 256 logistic + 256 linear '
                            '(Exp on sigma), N~LogU[128,8192], D~U{4..128}, '
                            '32 chains, NUTS(depth 10) + per-problem DA '
                            f'(target 0.8), {na} adaptation + {nr} draws',
                'problems': p, 'chains_per_problem': chains,
                'size_classes': [
                    {'problems': n, 'waves_per_chain': w,
                     'entries_total': cs['cost_total'],
                     'entries_adaptation': cs['cost_adaptation'],
                     'entries_critical_path': cs['cost_critical'],
                     'entries_critical_path_adaptation':
                         cs['cost_critical_adaptation'],
                     'adaptation_kernel_ms': sum(
                         l[0] for l in cs['launches'] if l[2] == 1),
                     'sampling_kernel_ms': sum(
                         l[0] for l in cs['launches'] if l[2] > 1)}
                    for (n, w), cs in zip(classes, cls_stats)],
                'sum_n_times_d': nd})
  line.update({
      'min_ess_per_problem_min_median': [float(min_ess.min()),
                                         float(np.median(min_ess))],
      'rhat_max': float(rhat_max.max()),
      'rhat_ok_fraction': float((rhat_max < 1.01).mean()),
      'chain_grad_evals_per_s': lf_total / (ms / 1e3),
      'plain': {'value': float(min_ess.sum()) / (plain_total_ms / 1e3),
                'ms_per_step': plain_total_ms,
                'what': f'each problem alone in its own launches '
                        f'({len(idx)} problems timed, mean x 512)'},
      'batched_over_plain': plain_total_ms / ms,
      'roofline': _ragged_roofline(stats, specs, lf_total, ms)})
  return line


def _ragged_roofline(stats, specs, lf_total, wall_ms):
  # algorithmic FLOP: 4 N_k D_k per chain-leapfrog, weighted by each
  # problem's share of the leapfrogs (per-problem totals), over the whole
  # run's wall time (the size classes' kernels overlap on their streams)
  lf = stats['leapfrogs_per_problem']
  flop = float(sum(4.0 * s.n_obs * s.dim * float(lf[i])
                   for i, s in enumerate(specs)))
  achieved = flop / (wall_ms * 1e-3) / 1e12
  return {'bound': 'mfma', 'achieved': achieved, 'peak': FP32_PEAK_TFLOPS,
          'unit': 'TFLOP/s', 'frac': achieved / FP32_PEAK_TFLOPS,
          'traffic': None, 'kernel': 'nuts_kernel<float,3,glm> (all size '
                                     'classes, wall time)',
          'chain_leapfrogs': lf_total}


# ---------------------------------------------------------- C4 serving
def serving_poisson200(args, torch, pa):
  from probability_amd import serving
  from probability_amd import workloads
  n_req = args.requests
  rate = 200.0
  r = np.random.default_rng(11)
  arrivals = np.cumsum(r.exponential(1.0 / rate, size=n_req))
  glms = workloads.ragged_glms(n_req, seed=13)
  cfgs = []
  for k in range(n_req):
    if k % 2 == 0:
      prob = glms[k]
      if prob['kind'] == 'logistic':
        spec = pa.targets.LogisticRegression(prob['x'], prob['y'],
                                             device='cuda')
        bij = None
      else:
        spec = pa.targets.LinearRegression(prob['x'], prob['y'],
                                           device='cuda')
        bij = [pa.bijectors.Identity(), pa.bijectors.Exp()]
    else:
      j = int(r.integers(8, 511))
      y, sig = workloads.hier_normal(j, seed=1000 + k)
      spec = pa.targets.HierarchicalNormal(y, sig, device='cuda')
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
      print(f'serving: {len(lat)}/{n_req} done, {len(srv.active)} active, '
            f'{len(srv.queue)} queued, t={last_print - t0:.0f}s', flush=True)
  wall = time.perf_counter() - t0
  lat = np.array(lat)
  ess_min = np.array(ess_min)
  rhat_max = np.array(rhat_max)
  line = _base('served posterior-sampling requests: throughput with p50/p99 '
               'latency (Poisson 200 req/s, NUTS+DA, 16 chains)',
               n_req / wall, 'req/s', args, 1000.0 * wall, 'f32',
               {'workload': f'serving_poisson200: {n_req} requests, half '
                            'ragged GLMs (N~LogU[128,8192], D~U{4..128}), '
                            'half hierarchical normals (J~U{8..510}, Softplus '
                            'on tau); 16 chains, 500 adaptation + 2000 draws',
                'requests': n_req, 'arrival_rate': rate,
                'hbm_budget_gib': args.hbm_budget_gib})
  line.update({
      'latency_s_p50_p99': [float(np.percentile(lat, 50)),
                            float(np.percentile(lat, 99))],
      'min_ess_per_request_min_median': [float(ess_min.min()),
                                         float(np.median(ess_min))],
      'aggregate_min_ess_per_s': float(ess_min.sum()) / wall,
      'rhat_ok_fraction': float((rhat_max < 1.01).mean()),
      'rhat_max_median': float(np.median(rhat_max)),
      'ticks': srv.ticks})
  return line


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
