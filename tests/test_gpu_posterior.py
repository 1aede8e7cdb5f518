"""Posterior means and variances of the float32 runs of the benchmark
workloads against float64 long runs of the same sampler (whose every
transition equals the float64 oracle to 1e-11 in the parity tests), within
3 Monte Carlo standard errors, following `mcmc/nuts_test.py:358-421`.

MCSE: batch means (batches of 50 draws per chain) of x for the mean and of
(x - mean)^2 for the variance; a difference is scored against the two runs'
combined MCSE.  With p parameters the bound on the largest |z| is the
two-sided normal quantile at 0.001 / p (a 1e-3 chance that any parameter of
a correct sampler exceeds it), and at least 95% of the parameters must lie
within 3 MCSE.
"""

import numpy as np
import pytest
import torch
from scipy import stats as sstats

from probability_amd import bijectors as pbij
from probability_amd import mcmc as pmcmc
from probability_amd import targets as ptg
from probability_amd import workloads
from probability_amd.mcmc import batched

pytestmark = pytest.mark.gpu


def _moments(draws, batch=50):
  """draws [n, chains, D] (device) -> float64 numpy mean, var and their
  batch-means MCSE."""
  x = draws.double()
  n = x.shape[0] - x.shape[0] % batch
  x = x[:n]
  m = x.mean((0, 1))
  sq = (x - m) ** 2
  v = sq.mean((0, 1))

  def se(y):
    bm = y.reshape((n // batch, batch) + tuple(y.shape[1:])).mean(1)
    k = bm.shape[0] * bm.shape[1]
    return bm.reshape(k, -1).std(0, unbiased=True) / np.sqrt(k)

  return [t.cpu().numpy() for t in (m, v, se(x), se(sq))]


def _check(label, draws32, draws64):
  m32, v32, sm32, sv32 = _moments(draws32)
  m64, v64, sm64, sv64 = _moments(draws64)
  zm = (m32 - m64) / np.sqrt(sm32 ** 2 + sm64 ** 2)
  zv = (v32 - v64) / np.sqrt(sv32 ** 2 + sv64 ** 2)
  z = np.abs(np.concatenate([zm.ravel(), zv.ravel()]))
  zmax = sstats.norm.ppf(1 - 0.0005 / z.size)
  print(f'POSTERIOR {label}: {z.size // 2} params, max|z| mean '
        f'{np.abs(zm).max():.2f} var {np.abs(zv).max():.2f} (bound '
        f'{zmax:.2f}), within 3 MCSE {np.mean(z <= 3):.4f}', flush=True)
  assert z.max() <= zmax
  assert np.mean(z <= 3) >= 0.95


def test_plain_hmc_logistic_1024x25_float32_vs_float64_long_run():
  # the plain-HMC workload as configured (64 chains at 0, L = 10,
  # eps = 0.05, 500 burn-in + 1000 results) against 8000 float64 results
  wl = workloads.logistic_1024x25()
  out = []
  for dt, nr, seed in ((torch.float32, wl['num_results'], 1),
                       (torch.float64, 8000, 2)):
    spec = ptg.LogisticRegression(wl['x'], wl['y'], dtype=dt)
    hmc = pmcmc.HamiltonianMonteCarlo(spec, step_size=wl['step_size'],
                                      num_leapfrog_steps=10)
    init = torch.zeros((wl['chains'], 25), dtype=dt, device='cuda')
    out.append(pmcmc.sample_chain(nr, init, kernel=hmc, trace_fn=None,
                                  num_burnin_steps=wl['num_burnin_steps'],
                                  seed=seed))
  _check('logistic_1024x25', out[0], out[1])


def test_nuts_hier_normal_j256_float32_vs_float64_long_run():
  # NUTS + dual averaging + Softplus TTK on the J = 256 hierarchical normal
  # (1024 chains, 1000 adaptation + 2000 draws) against 256 float64 chains
  # of 1000 adaptation + 8000 draws
  wl = workloads.hier_normal_j256()
  out = []
  for dt, c, nr, seed in ((torch.float32, 1024, 2000, 1),
                          (torch.float64, 256, 8000, 2)):
    spec = ptg.HierarchicalNormal(wl['y'], wl['sigma'], dtype=dt)
    nuts = pmcmc.NoUTurnSampler(spec, step_size=wl['step_size'])
    ttk = pmcmc.TransformedTransitionKernel(
        nuts, [pbij.Identity(), pbij.Softplus(), pbij.Identity()])
    da = pmcmc.DualAveragingStepSizeAdaptation(
        ttk, num_adaptation_steps=wl['num_adaptation_steps'],
        target_accept_prob=wl['target_accept_prob'])
    init = [torch.zeros(c, dtype=dt, device='cuda'),
            torch.full((c,), float(np.log(2.0)), dtype=dt, device='cuda'),
            torch.zeros((c, 256), dtype=dt, device='cuda')]
    mu, tau, z = pmcmc.sample_chain(nr, init, kernel=da, trace_fn=None,
                                    num_burnin_steps=wl['num_burnin_steps'],
                                    seed=seed)
    out.append(torch.cat([mu[..., None], tau[..., None], z], -1))
  _check('hier_normal_j256', out[0], out[1])


def _ragged_subset(dtype):
  """Four of the ragged GLM workload's problems small enough for a quick
  float64 long run: the smallest and a mid-size logistic and linear
  regression."""
  probs = workloads.ragged_glms(512)
  size = lambda k: probs[k]['x'].size
  pick = []
  for kind in ('logistic', 'linear'):
    ks = sorted((k for k, p in enumerate(probs) if p['kind'] == kind),
                key=size)
    pick += [ks[0], ks[len(ks) // 3]]
  specs, bij = [], []
  for k in pick:
    p = probs[k]
    if p['kind'] == 'logistic':
      specs.append(ptg.LogisticRegression(p['x'], p['y'], dtype=dtype))
      bij.append(None)
    else:
      specs.append(ptg.LinearRegression(p['x'], p['y'], dtype=dtype))
      bij.append([pbij.Identity(), pbij.Exp()])
  return pick, specs, bij


def test_ragged_glms_windowed_float32_vs_float64_long_run():
  # the ragged-GLM protocol (windowed adaptation, 32 chains, 500 adaptation
  # + 1000 draws) against float64 runs of 500 adaptation + 6000 draws
  res = []
  for dt, nr, seed in ((torch.float32, 1000, 1), (torch.float64, 6000, 2)):
    pick, specs, bij = _ragged_subset(dt)
    res.append(batched.sample_ragged(
        specs, 32, nr, bijectors=bij, keys=pick, seed=seed,
        num_burnin_steps=500, num_adaptation_steps=500, step_size=0.05,
        target_accept_prob=0.85, max_tree_depth=10, adaptation='windowed'))
  for i, k in enumerate(pick):
    _check(f'ragged problem {k}', res[0].draws[i], res[1].draws[i])
