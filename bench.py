"""Benchmark: min-ESS/s at split-R-hat < 1.01, batched vs one posterior per
launch, on one MI355X.

Default workload (ragged_glms_512): 512 independent Bayesian GLMs (256
logistic, 256 Gaussian linear with an Exp bijector on sigma),
N_k ~ LogUniform[128, 8192], D_k ~ U{4..128}, X iid N(0,1) float32, 32
chains each, NUTS (max_tree_depth 10) with per-problem windowed adaptation
(dual-averaging step size and diagonal mass matrix, target 0.85), 500
adaptation + 1000 draws, the problems batched in size classes on concurrent
streams with the data in HBM.  One step is 75 transitions of all 512
problems, so --steps 20 times exactly one complete run; value = the sum of
min-ESS over the problems that reach split R-hat < 1.01 (the others count
zero) divided by the timed seconds.  Beside it: the same problems one per
launch (a stratified sample, timed alone), the GPU kernels' design-matrix
read rate against HBM bandwidth, and the CPU restatement of TFP's batched
NUTS timed on the host cores.

Other workloads (one JSON line each, same fields):
  --workload hier_normal_j256     NUTS+DA+TTK(Softplus), hierarchical normal
                                  J=256, 1024 chains, 1000 + 2000; one step
                                  is one sampling run
  --workload logistic_1024x25     plain HMC, logistic regression 1024x25,
                                  64 chains, L=10, 500 + 1000 steps
  --workload logistic_65536x256   HMC + dual averaging on the BF16 MFMA
                                  gradient path, 4096 chains x 256, L=20
  --workload serving_poisson200   continuous-batching server, Poisson
                                  arrivals at 200 req/s
  --workload reducers_1024x258    streaming reducers at 1024 x 258

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload W]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3
# HBM bytes of one adaptation launch of nuts_kernel on this workload:
# rocprofv3 FETCH_SIZE (1330 KiB) x 2.00 + WRITE_SIZE (2851 KiB) x 1.00, the
# factors being what the two counters read for a 256 MiB device copy.
NUTS_LAUNCH_TRAFFIC_BYTES = (1330 * 2 + 2851) * 1024
# Algorithmic FP32 work of one chain-leapfrog on the hierarchical normal:
# 13 FLOP per group for log-prob + gradient (theta, residual, two scalings,
# the two quadratic terms, grad_z, the two reductions), 8 FLOP per state
# element for the leapfrog update and 2 per element for the kinetic energy.
def flops_per_leapfrog(j):
  return 13 * j + 10 * (j + 2)


def parse():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=20)
  ap.add_argument('--warmup', type=int, default=5)
  ap.add_argument('--cpu-seconds', type=float, default=24.0)
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--no-plain', action='store_true')
  ap.add_argument('--seed', type=int, default=20241)
  ap.add_argument('--adapt', type=int, default=None,
                  help='override the adaptation length (profiling runs)')
  ap.add_argument('--draws', type=int, default=None,
                  help='override the number of draws (profiling runs)')
  ap.add_argument('--calib-copy', action='store_true',
                  help='run one 256 MiB device copy (counter calibration)')
  ap.add_argument('--c3-dual-averaging', action='store_true',
                  help='ragged GLMs: step-size adaptation only (no mass '
                       'matrix)')
  ap.add_argument('--c3-group-adaptation', action='store_true',
                  help='ragged GLMs: chain groups in the adaptation steps '
                       'too')
  ap.add_argument('--c3-class-splits', default=None,
                  help="ragged GLMs: 'w:n,...' deals the w-wave class to n "
                       'independent batches (default: none)')
  ap.add_argument('--no-c3-compare', dest='c3_compare', action='store_false',
                  help='ragged GLMs: skip the dual-averaging-only comparison '
                       'run')
  ap.add_argument('--workload', default='ragged_glms_512',
                  choices=['hier_normal_j256', 'logistic_1024x25',
                           'logistic_65536x256', 'ragged_glms_512',
                           'serving_poisson200', 'reducers_1024x258'])
  ap.add_argument('--requests', type=int, default=2000,
                  help='serving: number of requests')
  ap.add_argument('--hbm-budget-gib', type=int, default=192,
                  help='serving: device memory the server may hold for '
                       'admitted requests')
  ap.add_argument('--plain-sample', type=int, default=16,
                  help='ragged GLMs: problems timed one per launch')
  ap.add_argument('--glm-waves', type=int, default=None,
                  help='ragged GLMs: one batch with this many waves '
                       'sharing each chain (default: size classes, each '
                       'problem by its own N x D, on concurrent streams)')
  return ap.parse_args()


def build_kernel(wl, torch, pa):
  spec = pa.targets.HierarchicalNormal(wl['y'], wl['sigma'],
                                       dtype=torch.float32, device='cuda')
  nuts = pa.mcmc.NoUTurnSampler(spec, step_size=np.float32(wl['step_size']),
                                max_tree_depth=wl['max_tree_depth'])
  ttk = pa.mcmc.TransformedTransitionKernel(
      nuts, [pa.bijectors.Identity(), pa.bijectors.Softplus(),
             pa.bijectors.Identity()])
  da = pa.mcmc.DualAveragingStepSizeAdaptation(
      ttk, num_adaptation_steps=wl['num_adaptation_steps'],
      target_accept_prob=wl['target_accept_prob'])
  c, j = wl['chains'], len(wl['y'])
  init = [torch.zeros(c, device='cuda'),
          torch.full((c,), float(np.log(2.0)), device='cuda'),
          torch.zeros((c, j), device='cuda')]
  return spec, da, init


def run_once(pa, da, init, wl, seed, fused=True, stats=None):
  return pa.mcmc.sample_chain(
      wl['num_results'], init, kernel=da,
      num_burnin_steps=wl['num_burnin_steps'],
      trace_fn=pa.mcmc.TraceFields('step_size'), seed=seed,
      return_final_kernel_results=True, experimental_fused=fused,
      experimental_stats=stats)


def ess_rhat(torch, pa, states):
  mu, tau, z = states
  flat = torch.cat([mu[..., None], tau[..., None], z], -1)  # [R, C, D]
  ess = []
  rhat = []
  for s in range(0, flat.shape[-1], 32):
    blk = flat[..., s:s + 32]
    ess.append(pa.mcmc.effective_sample_size(blk, cross_chain_dims=1))
    rhat.append(pa.mcmc.potential_scale_reduction(blk, split_chains=True))
  ess = torch.cat(ess)
  rhat = torch.cat(rhat)
  return flat, ess, rhat


def cpu_baseline(wl, final_step, final_x, min_ess, budget_s):
  """TFP's batched NUTS algorithm restated in NumPy (oracle.mcmc over a
  float32 NumPy target), timed on the host from the GPU run's final
  (adapted) states and step size: the 1024-chain batch on one core (median
  of 3), and the batch split over one process per core, all at once.
  Seconds per transition x the run's transitions gives the CPU run time."""
  import os
  import bench_workloads as bw
  cores = len(os.sched_getaffinity(0))
  tg = bw._CpuHier(wl['y'], wl['sigma'])
  x = final_x.astype(np.float64)
  total = wl['num_adaptation_steps'] + wl['num_results']
  budget = max(1.0, budget_s / 2.0)
  t_one = bw._cpu_nuts_one(tg, x, final_step, budget)
  parts = np.array_split(x, cores)
  t_all = max(bw._cpu_nuts_all([(tg, p, final_step) for p in parts], budget))
  return dict(value=float(min_ess / (t_all * total)), unit='min-ESS/s',
              cores=cores, kind='port',
              sample=(f'post-adaptation NUTS transitions of the '
                      f'{wl["chains"]}-chain batch, float32 NumPy target; '
                      f'{cores} processes with {len(parts[0])} chains each '
                      f'({t_all:.2f} s per transition), extrapolated to the '
                      f'{total}-transition run with the GPU run\'s min-ESS'),
              seconds_per_run=t_all * total,
              one_thread=dict(value=float(min_ess / (t_one * total)),
                              seconds_per_run=t_one * total, cores=1))


def main():
  args = parse()
  if args.gpus != 1:
    print(f'bench.py: this build runs on one GPU; --gpus {args.gpus} is not '
          'supported (chains and posteriors shard across GPUs with no '
          'communication, so N GPUs run N independent shares)',
          file=sys.stderr)
    sys.exit(2)
  os.environ.setdefault('OMP_NUM_THREADS', '1')
  import torch
  import probability_amd as pa
  from probability_amd import workloads
  if args.calib_copy:
    src = torch.ones(64 << 20, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
  if args.workload != 'hier_normal_j256':  # the workloads of bench_workloads
    import bench_workloads
    line = bench_workloads.run(args, torch, pa)
    print(json.dumps(line), flush=True)
    return

  wl = workloads.hier_normal_j256()
  if args.adapt is not None:
    wl['num_adaptation_steps'] = wl['num_burnin_steps'] = args.adapt
  if args.draws is not None:
    wl['num_results'] = args.draws
  j, c = len(wl['y']), wl['chains']
  spec, da, init = build_kernel(wl, torch, pa)

  for w in range(args.warmup):
    run_once(pa, da, init, wl, args.seed + 1000 + w)
  torch.cuda.synchronize()
  times = []
  res = None
  for k in range(args.steps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = run_once(pa, da, init, wl, args.seed + k)
    torch.cuda.synchronize()
    times.append(time.perf_counter() - t0)
  ms = 1000 * float(np.mean(times))
  flat, ess, rhat = ess_rhat(torch, pa, res.all_states)
  min_ess = float(ess.min())
  rhat_max = float(rhat.max())

  # posterior check against float64 quadrature (mu, tau)
  mom = workloads.hier_posterior_moments(wl['y'], wl['sigma'])
  draws_mu = flat[..., 0].double()
  draws_tau = flat[..., 1].double()
  z_mu = (float(draws_mu.mean()) - mom['mean_mu']) / np.sqrt(
      mom['var_mu'] / float(ess[0]))
  z_tau = (float(draws_tau.mean()) - mom['mean_tau']) / np.sqrt(
      mom['var_tau'] / float(ess[1]))

  # kernel timing (untimed extra run with HIP events around each launch)
  stats = {'time_kernels': True}
  run_once(pa, da, init, wl, args.seed, stats=stats)
  launches = stats['launches']
  lf_per_step = stats['leapfrogs_per_step'].double().cpu().numpy()
  kern_ms = np.array([l[0] for l in launches])
  flops = []
  for (ms_k, t0_, n_) in launches:
    flops.append(lf_per_step[t0_:t0_ + n_].sum() * flops_per_leapfrog(j))
  flops = np.array(flops)
  achieved = float(flops.sum() / (kern_ms.sum() * 1e-3) / 1e12)
  adapt_ms = kern_ms[:-1] if len(kern_ms) > 1 else kern_ms
  # latency of one leapfrog on the critical path of an adaptation launch
  # (one transition per launch: the launch lasts as long as its longest
  # trajectory)
  lf_max = stats['leapfrogs_max_per_step'].double().cpu().numpy()
  crit_us = [1000.0 * ms_k / max(lf_max[t0_], 1.0)
             for (ms_k, t0_, n_) in launches if n_ == 1]
  grad_evals = float(stats['leapfrogs_total'])

  line = {
      'metric': 'min-ESS/sec at split-Rhat<1.01 (NUTS+DA+TTK, hierarchical '
                'normal J=256, 1024 chains)',
      'value': min_ess / (ms / 1000.0),
      'unit': 'min-ESS/s',
      'n_gpus': 1,
      'steps': args.steps,
      'warmup': args.warmup,
      'ms_per_step': ms,
      'higher_is_better': True,
      'scaling': 'weak',
      'vs_baseline': None,
      'dtype': 'f32',
      'data': 'synthetic (seeded NumPy PCG64)',
      'config': {
          'workload': 'hier_normal_j256: NUTS(max_tree_depth=10) + '
                      'DualAveraging(target 0.8, 1000 steps) + TTK(Softplus '
                      'on tau), non-centred hierarchical normal J=256, 1024 '
                      'chains, 1000 adaptation + 2000 draws',
          'chains': c, 'groups': j, 'num_adaptation_steps': 1000,
          'num_results': 2000, 'parallelism': 'single GPU, chains batched'},
      'min_ess': min_ess,
      'rhat_max': rhat_max,
      'rhat_ok': rhat_max < 1.01,
      'posterior_z_vs_quadrature': {'mu': z_mu, 'tau': z_tau},
      'chain_grad_evals_per_s': grad_evals / (ms / 1000.0),
      'final_step_size': float(res.final_kernel_results.new_step_size),
      'adapt_iteration_ms_p50_p99': [float(np.percentile(adapt_ms, 50)),
                                     float(np.percentile(adapt_ms, 99))],
      'adapt_longest_trajectory_p50': float(np.percentile(
          lf_max[:len(crit_us)], 50)) if crit_us else None,
      'adapt_us_per_critical_leapfrog_p50': (
          float(np.percentile(crit_us, 50)) if crit_us else None),
      'roofline': {
          'bound': 'mfma', 'achieved': achieved, 'peak': FP32_PEAK_TFLOPS,
          'unit': 'TFLOP/s', 'frac': achieved / FP32_PEAK_TFLOPS,
          'limiter': ('VALU/latency: one wave per chain, FP32 FMA and '
                      'cross-lane reductions, no MFMA; priced against the '
                      'FP32 dense peak; adaptation launches are paced by '
                      'the longest trajectory of 1024 chains'),
          'traffic': NUTS_LAUNCH_TRAFFIC_BYTES,
          'traffic_what': 'HBM bytes per adaptation launch (rocprofv3 '
                          'FETCH_SIZE/WRITE_SIZE, copy-calibrated; '
                          'profiles/nuts_kernel_hier_normal_j256_traffic.md)',
          'kernel': 'nuts_kernel<float,5,hier>',
          'avg_launch_ms': float(kern_ms.mean()),
          'launches': len(kern_ms),
          'flop_per_chain_leapfrog': flops_per_leapfrog(j)},
  }
  if not args.no_plain:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_once(pa, da, init, wl, args.seed + args.steps - 1, fused=False)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    line['plain'] = {'value': min_ess / tp, 'ms_per_step': 1000 * tp,
                     'what': 'same chain via the host step loop (one NUTS + '
                             'one DA launch per transition)'}
    line['batched_over_plain'] = (min_ess / (ms / 1000.0)) / (min_ess / tp)
  if not args.no_cpu:
    final = res.final_kernel_results
    xs = spec.pack_state(final.inner_results.transformed_state).cpu().numpy()
    line['cpu_baseline'] = cpu_baseline(
        wl, float(final.new_step_size), xs, min_ess, args.cpu_seconds)
  print(json.dumps(line), flush=True)


if __name__ == '__main__':
  main()
