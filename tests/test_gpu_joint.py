"""Compiled JointDistribution models on the GPU: the target spec's log
density equals the model's, summed from scipy.stats closed forms of each
variable, and the compiled model samples under TTK + NUTS."""

import numpy as np
import pytest
import torch
from scipy import stats

import probability_amd as pa
from probability_amd import distributions as tfd

pytestmark = pytest.mark.gpu

SIGMA = np.array([15., 10., 16., 11., 9., 11., 10., 18.])
Y8 = np.array([28., 8., -3., 7., -1., 1., 18., 12.])


def _eight_schools():
  return tfd.JointDistributionSequential([
      tfd.Normal(0., 5.), tfd.HalfCauchy(0., 5.),
      tfd.Sample(tfd.Normal(0., 1.), 8),
      lambda z, tau, mu: tfd.Independent(tfd.Normal(mu + tau * z, SIGMA), 1)])


def _lp(t):
  return t.cpu().numpy().astype(np.float64)


def test_eight_schools_density():
  ct = tfd.compile_target(_eight_schools(), dtype=torch.float64, x3=Y8)
  r = np.random.default_rng(0)
  mu, tau, z = r.normal(size=6) * 3, r.uniform(0.1, 8, 6), r.normal(size=(6, 8))
  got = _lp(ct.target(torch.as_tensor(mu), torch.as_tensor(tau),
                      torch.as_tensor(z)))
  want = (stats.norm.logpdf(mu, 0, 5) + stats.halfcauchy.logpdf(tau, 0, 5)
          + stats.norm.logpdf(z).sum(-1)
          + stats.norm.logpdf(Y8, mu[:, None] + tau[:, None] * z,
                              SIGMA).sum(-1))
  np.testing.assert_allclose(got, want, rtol=1e-11)


def test_regression_densities():
  r = np.random.default_rng(1)
  x = r.normal(size=(60, 3))
  yb = (r.uniform(size=60) < 0.4) * 1.
  logistic = tfd.JointDistributionNamed(dict(
      beta=tfd.Sample(tfd.Normal(0., 2.), 3),
      y=lambda beta: tfd.Independent(tfd.Bernoulli(logits=x @ beta), 1)))
  ct = tfd.compile_target(logistic, dtype=torch.float64, y=yb)
  beta = r.normal(size=(5, 3))
  want = (stats.norm.logpdf(beta, 0, 2).sum(-1)
          + stats.bernoulli.logpmf(yb, 1 / (1 + np.exp(-beta @ x.T))).sum(-1))
  np.testing.assert_allclose(_lp(ct.target(torch.as_tensor(beta))), want,
                             rtol=1e-10)
  yl = x @ np.array([1., -2., .5]) + r.normal(size=60) * .3
  linear = tfd.JointDistributionSequential([
      tfd.Independent(tfd.Normal(np.zeros(3), 1.5), 1), tfd.HalfNormal(0.7),
      lambda sigma, beta: tfd.Independent(
          tfd.Normal(tfd.matvec(x, beta), sigma), 1)])
  ct = tfd.compile_target(linear, dtype=torch.float64, x2=yl)
  sig = r.uniform(0.2, 1.5, 5)
  want = (stats.norm.logpdf(beta, 0, 1.5).sum(-1)
          + stats.halfnorm.logpdf(sig, 0, 0.7)
          + stats.norm.logpdf(yl, beta @ x.T, sig[:, None]).sum(-1))
  got = _lp(ct.target(torch.as_tensor(beta), torch.as_tensor(sig)))
  np.testing.assert_allclose(got, want, rtol=1e-10)


def test_compiled_eight_schools_samples_under_ttk_nuts():
  ct = tfd.compile_target(_eight_schools(), dtype=torch.float32, x3=Y8)
  kern = pa.mcmc.TransformedTransitionKernel(
      pa.mcmc.NoUTurnSampler(ct.target, step_size=0.3), ct.bijectors)
  c = 64
  init = [torch.zeros(c, device='cuda'), torch.ones(c, device='cuda'),
          torch.zeros(c, 8, device='cuda')]
  states = pa.mcmc.sample_chain(200, init, kernel=kern, seed=4,
                                num_burnin_steps=100, trace_fn=None)
  mu, tau, z = states
  assert torch.all(tau > 0)
  assert all(torch.isfinite(s).all() for s in states)
  # posterior mean of mu in the eight-schools model is about 4.4 (sd ~3.3)
  assert abs(float(mu.double().mean()) - 4.4) < 1.5
