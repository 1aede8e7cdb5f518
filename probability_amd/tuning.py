"""Launch-shape autotuner: waves per GLM chain chosen from timings.

The waves that share one GLM chain's gradient (1, 2, 4 or 8; each wave takes
a share of the design-matrix rows and the partial sums meet in LDS) trade
per-leapfrog latency against the number of chains resident at once.  The
best choice depends on N, D, the chain count of the launch and the sampler,
so it is measured: `tune()` times the sampling kernels on synthetic
problems over a grid of (sampler, kind, N, D, chains) and writes, per grid
point, the time per chain-leapfrog of each wave count and the fastest one,
to `tuning_table.json` beside this file.  `glm_waves()` returns the entry
nearest in (log N, log D, log chains) for the sampler and kind, or the
fixed N x D rule (`_lib.glm_waves_per_chain`) when the table has none.

Timing: NUTS at a step size small enough that no trajectory turns or
diverges runs 2^depth - 1 leapfrogs per transition (max_tree_depth 5: 31),
HMC runs L = 10; each point is the launch time over the chain-leapfrogs
done, the median of three HIP-event-timed launches after a warm-up launch.

Usage: python -m probability_amd.tuning [--quick]   (on the GPU; rewrites
the table).
"""

import json
import math
import os

import numpy as np
import torch

from probability_amd import _lib

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                          'tuning_table.json')

# the many-chain points of the grid (one launch of many chains of a
# problem); the ragged batches and the serving lanes look a problem up at
# its own chain count, so that its wave count does not depend on which
# other problems share its launches
BATCH_CHAINS = 2048

_table = None


def _load():
  global _table
  if _table is None:
    try:
      with open(TABLE_PATH) as f:
        _table = json.load(f)
    except (OSError, ValueError):
      _table = {}
  return _table


def glm_waves(kind, n_obs, dim, chains, sampler='nuts', dtype=torch.float32):
  """Waves per chain for a GLM launch of `chains` chains of an N x D
  problem: the tuned choice of the nearest measured point, else the rule."""
  rows = _load().get('glm_waves', {}).get(str(dtype).replace('torch.', ''), [])
  name = 'linear' if kind == _lib.TARGET_LINEAR else 'logistic'
  best, dist = None, None
  for r in rows:
    if r['sampler'] != sampler or r['kind'] != name:
      continue
    dd = ((math.log2(r['n_obs']) - math.log2(max(n_obs, 1))) ** 2 +
          (math.log2(r['dim']) - math.log2(max(dim, 1))) ** 2 +
          0.25 * (math.log2(r['chains']) - math.log2(max(chains, 1))) ** 2)
    if dist is None or dd < dist:
      best, dist = r, dd
  if best is None:
    w = _lib.glm_waves_per_chain(n_obs * dim)
  else:
    w = int(best['best'])
  return min(w, 4) if sampler == 'hmc' else w


# ------------------------------------------------------------------ tuner
def _problem(kind, n, d, seed, dtype):
  from probability_amd import targets
  r = np.random.default_rng(seed)
  x = r.normal(size=(n, d)).astype(np.float32)
  beta = r.normal(size=d) * 0.3
  if kind == 'logistic':
    y = (r.uniform(size=n) < 1 / (1 + np.exp(-(x @ beta)))).astype(np.float32)
    return targets.LogisticRegression(x, y, dtype=dtype)
  y = (x @ beta + 0.5 * r.normal(size=n)).astype(np.float32)
  return targets.LinearRegression(x, y, dtype=dtype)


def _time_launch(spec, chains, waves, sampler, n_iters=2, chains_per_block=1):
  """Microseconds per chain-leapfrog of one sampling launch with `waves`
  waves per chain (NUTS GLMs: and chain groups of `chains_per_block`;
  median of 3 after a warm-up)."""
  from probability_amd.mcmc import _runner
  dev, dt = spec.device, spec.dtype
  x = torch.zeros((chains, spec.dim), dtype=dt, device=dev)
  if spec.kind == _lib.TARGET_LINEAR:
    x[:, -1] = 1.0  # sigma
  table = spec.table(spec.bijectors_for_eval())
  tlp = torch.zeros(chains, dtype=dt, device=dev)
  grad = torch.zeros_like(x)
  _lib.check(_lib.lib().hipmc_target_eval(
      _lib.dtype_code(dt), spec.kind, _lib.ptr(table), None, chains,
      spec.dim, spec.dim, _lib.ptr(x), _lib.ptr(tlp), _lib.ptr(grad),
      _lib.stream_handle(dev)))
  # at this step size a trajectory of 31 leapfrogs moves ~ p t (t = 3e-5):
  # no U-turn and no divergence, so every NUTS transition runs all 31
  step = torch.full((1,), 1e-6, dtype=dt, device=dev)
  times = []
  for rep in range(4):
    run = _runner.Run(spec, table, x.clone(), tlp.clone(), grad.clone(),
                      step, [1, rep], n_iters=n_iters, fold_seed=True)
    a = run.args
    a.waves_per_chain = int(waves)
    a.chains_per_block = int(chains_per_block)
    if sampler == 'nuts':
      a.max_tree_depth, a.unrolled_leapfrog_steps = 5, 1
      a.max_energy_diff = 1000.0
      fn, per = _lib.lib().hipmc_nuts_run, 31
      lf = run.want('out_leapfrogs', (n_iters, chains), torch.int32)
    else:
      a.num_leapfrog_steps = 10
      fn, per, lf = _lib.lib().hipmc_hmc_run, 10, None
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    run.call(fn)
    e1.record()
    torch.cuda.synchronize(dev)
    # per chain-leapfrog of the work done (every chain's leapfrogs)
    work = int(lf.sum()) if lf is not None else n_iters * per * chains
    if rep:
      times.append(e0.elapsed_time(e1) * 1e3 / max(work, 1))
  return float(np.median(times))


def tune(quick=False, dtype=torch.float32, log=print):
  grid = []
  ns = [128, 512, 2048, 8192] if quick else [128, 256, 512, 1024, 2048,
                                             4096, 8192]
  ds = [4, 32, 128] if quick else [4, 16, 64, 128]
  for kind in ('logistic', 'linear'):
    for n in ns:
      for d in ds:
        for c in (32, BATCH_CHAINS):
          grid.append(('nuts', kind, n, d, c, (1, 2, 4, 8)))
  hn = [1024, 16384] if quick else [256, 1024, 4096, 16384, 65536]
  hd = [25, 256] if quick else [8, 25, 64, 256]
  for n in hn:
    for d in hd:
      for c in (64, 1024):
        grid.append(('hmc', 'logistic', n, d, c, (1, 2, 4)))
  rows = []
  for i, (sampler, kind, n, d, c, ws) in enumerate(grid):
    dim = d + (1 if kind == 'linear' else 0)
    spec = _problem(kind, n, d, 100 + i, dtype)
    t = {str(w): _time_launch(spec, c, w, sampler) for w in ws}
    best = min(t, key=t.get)
    rows.append(dict(sampler=sampler, kind=kind, n_obs=n, dim=dim, chains=c,
                     us_per_chain_leapfrog=t, best=int(best)))
    log(f'{sampler} {kind} N={n} D={dim} chains={c}: ' +
        ' '.join(f'W{w}={v:.3f}' for w, v in t.items()) + f' -> W{best}')
    del spec
  return rows


def main():
  import argparse
  ap = argparse.ArgumentParser()
  ap.add_argument('--quick', action='store_true')
  ap.add_argument('--out', default=TABLE_PATH)
  ap.add_argument('--time', nargs=5, default=None,
                  metavar=('KIND', 'N', 'D', 'CHAINS', 'WAVES'),
                  help='time one configuration (logistic or linear, N x D, '
                       'chains, waves per chain) and write no table')
  ap.add_argument('--chains-per-block', type=int, default=1,
                  help='with --time: NUTS chain groups of this size')
  ap.add_argument('--sampler', default='nuts', choices=['nuts', 'hmc'])
  args = ap.parse_args()
  _lib.require_cuda()
  if args.time:
    kind, n, d, c, w = args.time
    spec = _problem(kind, int(n), int(d), 1, torch.float32)
    us = _time_launch(spec, int(c), int(w), args.sampler,
                      chains_per_block=args.chains_per_block)
    print(json.dumps({'sampler': args.sampler, 'kind': kind, 'n_obs': int(n),
                      'dim': spec.dim, 'chains': int(c), 'waves': int(w),
                      'chains_per_block': args.chains_per_block,
                      'us_per_chain_leapfrog': us}), flush=True)
    return
  rows = tune(quick=args.quick, log=lambda s: print(s, flush=True))
  table = {'device': torch.cuda.get_device_name(0),
           'what': 'microseconds per chain-leapfrog of the sampling kernels '
                   'by waves per GLM chain; best = the fastest',
           'glm_waves': {'float32': rows}}
  with open(args.out, 'w') as f:
    json.dump(table, f, indent=1)
  print(f'wrote {args.out} ({len(rows)} points)', flush=True)


if __name__ == '__main__':
  main()
