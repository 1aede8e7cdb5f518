"""JointDistribution models matched onto the target families (host logic,
no GPU): the traced structure, the variable names TFP's conventions give,
the extracted parameters, and refusals of models that are not built."""

import numpy as np
import pytest

from probability_amd import bijectors as bij
from probability_amd import distributions as tfd
from probability_amd import targets


def _eight_schools_seq(sigma):
  return tfd.JointDistributionSequential([
      tfd.Normal(0., 5.),                                    # mu
      tfd.HalfCauchy(0., 5.),                                # tau
      tfd.Sample(tfd.Normal(0., 1.), 8),                     # z
      lambda z, tau, mu: tfd.Independent(
          tfd.Normal(mu + tau * z, sigma), 1),               # y
  ])


def test_sequential_names_and_hierarchical_match():
  sigma = np.array([15., 10., 16., 11., 9., 11., 10., 18.])
  y = np.array([28., 8., -3., 7., -1., 1., 18., 12.])
  model = _eight_schools_seq(sigma)
  assert [n for n, _ in model._trace()] == ['mu', 'tau', 'z', 'x3']
  cls, kw, bijs, names = tfd.match(model, x3=y)
  assert cls is targets.HierarchicalNormal
  assert names == ['mu', 'tau', 'z']
  assert kw['mu_scale'] == 5. and kw['tau_scale'] == 5.
  np.testing.assert_array_equal(kw['y'], y)
  np.testing.assert_array_equal(kw['sigma'], sigma)
  assert [type(b) for b in bijs] == [bij.Identity, bij.Softplus, bij.Identity]


def test_named_hierarchical_either_operand_order():
  sigma, y = np.full(4, 2.), np.arange(4.)
  model = tfd.JointDistributionNamed(dict(
      mu=tfd.Normal(0., 3.),
      tau=tfd.HalfCauchy(0., 2.),
      z=tfd.Independent(tfd.Normal(np.zeros(4), 1.), 1),
      y=lambda mu, tau, z: tfd.Independent(tfd.Normal(z * tau + mu, sigma), 1),
  ))
  cls, kw, _, names = tfd.match(model, y=y)
  assert cls is targets.HierarchicalNormal and names == ['mu', 'tau', 'z']
  assert (kw['mu_scale'], kw['tau_scale']) == (3., 2.)


def test_logistic_and_linear_regression_match():
  r = np.random.default_rng(0)
  x, y = r.normal(size=(50, 3)), (r.uniform(size=50) < .5) * 1.
  model = tfd.JointDistributionNamed(dict(
      beta=tfd.Sample(tfd.Normal(0., 2.), 3),
      y=lambda beta: tfd.Independent(tfd.Bernoulli(logits=x @ beta), 1)))
  cls, kw, bijs, names = tfd.match(model, y=y)
  assert cls is targets.LogisticRegression and names == ['beta']
  assert kw['prior_scale'] == 2. and np.array_equal(kw['x'], x)
  model = tfd.JointDistributionSequential([
      tfd.Independent(tfd.Normal(np.zeros(3), 1.5), 1),
      tfd.HalfNormal(0.7),
      lambda sigma, beta: tfd.Independent(
          tfd.Normal(tfd.matvec(x, beta), sigma), 1)])
  cls, kw, bijs, names = tfd.match(model, x2=y)
  assert cls is targets.LinearRegression and names == ['beta', 'sigma']
  assert (kw['prior_scale'], kw['sigma_scale']) == (1.5, 0.7)
  assert [type(b) for b in bijs] == [bij.Identity, bij.Softplus]


def test_independent_normals_match():
  model = tfd.JointDistributionSequential(
      [tfd.Independent(tfd.Normal([1., 2.], [3., 4.]), 1)])
  cls, kw, _, _ = tfd.match(model)
  assert cls is targets.DiagNormal
  np.testing.assert_array_equal(kw['loc'], [1., 2.])
  np.testing.assert_array_equal(kw['scale'], [3., 4.])


def test_models_that_are_not_built_raise():
  sigma, y = np.ones(8), np.zeros(8)
  centred = tfd.JointDistributionSequential([
      tfd.Normal(0., 5.), tfd.HalfCauchy(0., 5.),
      lambda tau, mu: tfd.Sample(tfd.Normal(mu, tau), 8),
      lambda theta: tfd.Independent(tfd.Normal(theta, sigma), 1)])
  with pytest.raises(NotImplementedError):
    tfd.match(centred, x3=y)
  reordered = tfd.JointDistributionNamed(dict(
      tau=tfd.HalfCauchy(0., 5.), mu=tfd.Normal(0., 5.),
      z=tfd.Sample(tfd.Normal(0., 1.), 8),
      y=lambda mu, tau, z: tfd.Independent(tfd.Normal(mu + tau * z, sigma), 1)))
  with pytest.raises(NotImplementedError):
    tfd.match(reordered, y=y)
  with pytest.raises(ValueError):
    tfd.match(_eight_schools_seq(sigma), nope=y)
  with pytest.raises(NotImplementedError):
    tfd.Bernoulli(probs=0.5)
  x = np.ones((10, 3))
  wrong_width = tfd.JointDistributionNamed(dict(
      beta=tfd.Sample(tfd.Normal(0., 1.), 4),
      y=lambda beta: tfd.Independent(tfd.Bernoulli(logits=x @ beta), 1)))
  with pytest.raises(NotImplementedError):
    tfd.match(wrong_width, y=np.zeros(10))
  vector_mu = tfd.JointDistributionSequential([
      tfd.Sample(tfd.Normal(0., 5.), 2), tfd.HalfCauchy(0., 5.),
      tfd.Sample(tfd.Normal(0., 1.), 8),
      lambda z, tau, mu: tfd.Independent(tfd.Normal(mu + tau * z, sigma), 1)])
  with pytest.raises(NotImplementedError):
    tfd.match(vector_mu, x3=y)
