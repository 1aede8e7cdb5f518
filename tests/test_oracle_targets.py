"""Pins the oracle's analytic gradients (central differences, FP64) and its
log-densities (closed forms / scipy)."""

import numpy as np
import pytest
from scipy import stats

from oracle import bijectors as bij
from oracle import targets as tg


def _rng(s=0):
  return np.random.default_rng(s)


def _targets():
  r = _rng(1)
  x = r.normal(size=(40, 5))
  yb = (r.uniform(size=40) < 0.5).astype(float)
  yl = x @ r.normal(size=5) + 0.3 * r.normal(size=40)
  return {
      'logistic': (tg.LogisticRegression(x, yb, prior_scale=1.3),
                   r.normal(size=(4, 5))),
      'linear': (tg.LinearRegression(x, yl, 1.0, 1.0),
                 np.concatenate([r.normal(size=(4, 5)),
                                 r.uniform(0.5, 2, size=(4, 1))], 1)),
      'hier': (tg.HierarchicalNormal(r.normal(size=7) * 10,
                                     r.uniform(5, 20, size=7)),
               np.concatenate([r.normal(size=(4, 1)),
                               r.uniform(0.5, 4, size=(4, 1)),
                               r.normal(size=(4, 7))], 1)),
      'diag': (tg.DiagNormal(r.normal(size=6), r.uniform(0.5, 2, size=6)),
               r.normal(size=(4, 6))),
      'funnel': (tg.NealsFunnel(10), r.normal(size=(4, 10))),
  }


@pytest.mark.parametrize('name', ['logistic', 'linear', 'hier', 'diag',
                                  'funnel'])
def test_gradient_matches_central_differences(name):
  target, x = _targets()[name]
  _, g = target.log_prob_and_grad(x)
  fd = tg.finite_difference_grad(target, x, h=1e-6)
  np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


def test_transformed_gradient_matches_central_differences():
  target, x = _targets()['hier']
  codes = np.zeros(x.shape[1], np.int32)
  codes[1] = bij.SOFTPLUS
  t = tg.Transformed(target, codes)
  xu = t.inverse(x)
  _, g = t.log_prob_and_grad(xu)
  np.testing.assert_allclose(g, tg.finite_difference_grad(t, xu), rtol=1e-6,
                             atol=1e-6)
  codes[1] = bij.EXP
  t = tg.Transformed(target, codes)
  xu = t.inverse(x)
  _, g = t.log_prob_and_grad(xu)
  np.testing.assert_allclose(g, tg.finite_difference_grad(t, xu), rtol=1e-6,
                             atol=1e-6)


def test_log_densities_match_scipy():
  r = _rng(2)
  x = r.normal(size=(3, 5))
  y = (r.uniform(size=30) < 0.4).astype(float)
  xx = r.normal(size=(30, 5))
  lr = tg.LogisticRegression(xx, y, 2.0)
  lp, _ = lr.log_prob_and_grad(x)
  p = 1 / (1 + np.exp(-(x @ xx.T)))
  ref = (stats.bernoulli.logpmf(y, p).sum(-1)
         + stats.norm.logpdf(x, 0, 2.0).sum(-1))
  np.testing.assert_allclose(lp, ref, rtol=1e-10)

  ys, sg = r.normal(size=6) * 5, r.uniform(5, 20, size=6)
  h = tg.HierarchicalNormal(ys, sg)
  st = np.concatenate([r.normal(size=(3, 1)), r.uniform(0.1, 3, (3, 1)),
                       r.normal(size=(3, 6))], 1)
  lp, _ = h.log_prob_and_grad(st)
  mu, tau, z = st[:, 0], st[:, 1], st[:, 2:]
  ref = (stats.norm.logpdf(mu, 0, 5) + stats.halfcauchy.logpdf(tau, 0, 5)
         + stats.norm.logpdf(z).sum(-1)
         + stats.norm.logpdf(ys, mu[:, None] + tau[:, None] * z, sg).sum(-1))
  np.testing.assert_allclose(lp, ref, rtol=1e-10)

  yl = r.normal(size=30)
  l = tg.LinearRegression(xx, yl, 1.5, 0.7)
  st = np.concatenate([x, r.uniform(0.2, 2, (3, 1))], 1)
  lp, _ = l.log_prob_and_grad(st)
  ref = (stats.norm.logpdf(yl, st[:, :5] @ xx.T, st[:, 5:]).sum(-1)
         + stats.norm.logpdf(st[:, :5], 0, 1.5).sum(-1)
         + stats.halfnorm.logpdf(st[:, 5], scale=0.7))
  np.testing.assert_allclose(lp, ref, rtol=1e-10)


def test_softplus_bijector_round_trip_and_fldj():
  x = np.linspace(-30, 30, 121)
  y = bij.softplus_forward(x)
  np.testing.assert_allclose(bij.softplus_inverse(y), x, rtol=1e-9, atol=1e-9)
  # fldj = log dy/dx = log sigmoid(x)
  np.testing.assert_allclose(bij.fldj(bij.SOFTPLUS, x), np.log(bij.sigmoid(x)),
                             rtol=1e-12, atol=1e-12)
  np.testing.assert_allclose(bij.inverse(bij.EXP, bij.forward(bij.EXP, x)), x)


def test_log_gamma_density_gradient_and_moments():
  # log of a Gamma(c, rate r) variable: scipy's loggamma is the log of a
  # Gamma(c, 1) variable, so x + log r follows it.
  t = tg.LogGamma(4, 5.0, 10.0)
  x = _rng(3).normal(-0.7, 0.5, size=(6, 4))
  lp, g = t.log_prob_and_grad(x)
  ref = stats.loggamma.logpdf(x + np.log(10.0), 5.0).sum(-1)
  np.testing.assert_allclose(lp, ref, rtol=1e-12)
  np.testing.assert_allclose(g, tg.finite_difference_grad(t, x, h=1e-6),
                             rtol=1e-6, atol=1e-6)
  # E[x] = digamma(c) - log r and E[exp x] = c / r, by quadrature of the
  # oracle's own density (the expectations hmc_test.py checks by sampling).
  from scipy import integrate, special
  dens = lambda v: np.exp(tg.LogGamma(1, 5.0, 10.0).log_prob_and_grad(
      np.array([[v]]))[0][0])
  assert abs(integrate.quad(dens, -12, 6)[0] - 1) < 1e-9
  ex = integrate.quad(lambda v: v * dens(v), -12, 6)[0]
  assert abs(ex - (special.digamma(5.0) - np.log(10.0))) < 1e-9
  eex = integrate.quad(lambda v: np.exp(v) * dens(v), -12, 6)[0]
  assert abs(eex - 0.5) < 1e-9
