"""The ragged-GLM workload's own problems at full size against the oracle: a
stratified subset of workloads.ragged_glms(512) — the largest logistic
design (of the 8-wave class) and the smallest problem of each size
class, both families — batched by sample_ragged for a few NUTS + dual
averaging transitions, each problem against the oracle's
sample_chain(DualAveraging(TTK(NUTS))) with the same draws: identical
leapfrog counts and accepts, states within tolerance (float64 to 1e-9;
float32 over several transitions to 2e-4)."""

import numpy as np
import pytest
import torch

from oracle import mcmc as omcmc
from oracle import rng as orng
from oracle import targets as otg
from probability_amd import _lib
from probability_amd import bijectors as pbij
from probability_amd import targets as ptg
from probability_amd import tuning
from probability_amd import workloads
from probability_amd.mcmc import batched

pytestmark = pytest.mark.gpu

CHAINS = 6


def _subset(dtype):
  probs = workloads.ragged_glms(512)
  waves = [tuning.glm_waves(_lib.TARGET_LOGISTIC if p['kind'] ==
                            'logistic' else _lib.TARGET_LINEAR,
                            p['x'].shape[0], p['x'].shape[1] +
                            (p['kind'] == 'linear'), 32, 'nuts', dtype)
           for p in probs]
  size = [p['x'].size for p in probs]
  # the largest logistic design (from zero a linear regression of this size
  # only diverges until its step size has adapted)
  pick = [max((k for k, p in enumerate(probs) if p['kind'] == 'logistic'),
              key=lambda k: size[k])]
  for w in sorted(set(waves)):
    for kind in ('logistic', 'linear'):
      ks = [k for k, p in enumerate(probs) if waves[k] == w and
            p['kind'] == kind and k not in pick]
      if ks:
        pick.append(min(ks, key=lambda k: size[k]))
  specs, bij, oracles = [], [], []
  for k in pick:
    x, y = probs[k]['x'], probs[k]['y']
    if probs[k]['kind'] == 'logistic':
      specs.append(ptg.LogisticRegression(x, y, dtype=dtype))
      bij.append(None)
      oracles.append(otg.LogisticRegression(x.astype(np.float64),
                                            y.astype(np.float64)))
    else:
      specs.append(ptg.LinearRegression(x, y, dtype=dtype))
      bij.append([pbij.Identity(), pbij.Exp()])
      codes = np.zeros(x.shape[1] + 1, np.int32)
      codes[-1] = 2
      oracles.append(otg.Transformed(
          otg.LinearRegression(x.astype(np.float64), y.astype(np.float64)),
          codes))
  return pick, specs, bij, oracles


@pytest.mark.parametrize('dtype,ndt,tol', [(torch.float64, np.float64, 1e-9),
                                           (torch.float32, np.float32, 2e-4)])
def test_ragged_workload_problems_match_oracle(dtype, ndt, tol):
  pick, specs, bij, oracles = _subset(dtype)
  # the workload's largest logistic design, of the 8-wave class
  assert specs[0].n_obs * specs[0].dim > 500000
  # step 0.5 / sqrt(N) per problem; a short tree keeps the float64 oracle
  # quick at the largest size
  steps = [0.5 / np.sqrt(s.n_obs) for s in specs]
  kw = dict(num_results=2, num_burnin_steps=3, num_adaptation_steps=3,
            step_size=steps, target_accept_prob=0.8, max_tree_depth=6,
            seed=5)
  stats = {}
  res = batched.sample_ragged(specs, CHAINS, bijectors=bij, keys=pick,
                              slot_dim=192, stats=stats, **kw)
  assert len(stats['classes']) >= 3
  for i, (k, ot) in enumerate(zip(pick, oracles)):
    d = specs[i].dim
    ox, otrace, _ = omcmc.sample_nuts_da(
        ot, np.zeros((CHAINS, d)), np.asarray(steps[i], ndt), 3, 2, 3,
        seed=orng.sanitize_seed(5), target_accept_prob=0.8, max_depth=6,
        dtype=ndt, problem=k)
    want = np.stack([ot.forward(v) if hasattr(ot, 'forward') else v
                     for v in ox])
    lf = np.stack([r.leapfrogs_taken for r in otrace])
    acc = np.stack([r.is_accepted for r in otrace])
    np.testing.assert_array_equal(res.leapfrogs_taken[i].cpu().numpy(), lf)
    np.testing.assert_array_equal(res.is_accepted[i].cpu().numpy(), acc)
    if i == 0:
      # the largest problem really moves (the check is not vacuous)
      assert acc.any() and lf.max() > 1
    got = res.draws[i].double().cpu().numpy()
    err = np.abs(got - want).max() / max(np.abs(want).max(), 1.0)
    print(f'ERR ragged problem {k} N={specs[i].n_obs} D={d} '
          f'{np.dtype(ndt).name} draws rel={err:.3e}', flush=True)
    assert err <= tol, (k, err)
