"""Analytic log-densities and gradients of the targets on the hot path.

Every target maps a flat per-chain state x [C, D] (state parts concatenated
in order) to (target_log_prob [C], grad [C, D]) in float64.  The reference
gets gradients from TF autodiff (`mcmc/internal/util.py:246-308` ->
`math/gradient.py:28`), which cannot run here; these closed forms are pinned
instead by central finite differences and closed-form posteriors.

Log-density formulas follow TFP's own:
  Normal      `distributions/normal.py:182-188`
  Bernoulli   `distributions/bernoulli.py:119-135` (logits parameterisation,
              log p1 = -softplus(-s), log p0 = -softplus(s), multiply_no_nan)
  HalfCauchy  `distributions/half_cauchy.py:138-145`
  HalfNormal  `distributions/half_normal.py:146-153`
TTK composition follows `mcmc/transformed_kernel.py:86-128`.

Test infrastructure only.
"""

import numpy as np

from oracle import bijectors as bij

_HALF_LOG_2PI = 0.5 * np.log(2.0 * np.pi)


def normal_lp(x, loc, scale):
  return -0.5 * (x / scale - loc / scale) ** 2 - (_HALF_LOG_2PI + np.log(scale))


def half_cauchy_lp(x, scale):
  lp = np.log(2.0 / np.pi) - np.log(scale) - np.log1p((np.abs(x) / scale) ** 2)
  return np.where(x < 0, -np.inf, lp)


def half_normal_lp(x, scale):
  lp = -0.5 * (x / scale) ** 2 - (np.log(scale) + 0.5 * np.log(np.pi / 2.0))
  return np.where(x >= 0, lp, -np.inf)


def _mul_no_nan(x, y):
  return np.where(y == 0, 0.0, x * y)


class Target:
  """Base class: subclasses define `dim`, `part_sizes`, `log_prob_and_grad`."""

  part_sizes = ()

  @property
  def dim(self):
    return int(sum(self.part_sizes))

  def log_prob(self, x):
    return self.log_prob_and_grad(x)[0]


class LogisticRegression(Target):
  """Bayesian logistic regression: beta ~ N(0, s^2 I), y ~ Bern(sigmoid(X beta))."""

  def __init__(self, x, y, prior_scale=1.0):
    self.x = np.asarray(x, np.float64)
    self.y = np.asarray(y, np.float64)
    self.prior_scale = float(prior_scale)
    self.part_sizes = (self.x.shape[1],)

  def log_prob_and_grad(self, beta):
    beta = np.asarray(beta, np.float64)
    s = beta @ self.x.T
    ll = (_mul_no_nan(-bij.softplus(s), 1.0 - self.y)
          + _mul_no_nan(-bij.softplus(-s), self.y))
    tlp = ll.sum(-1) + normal_lp(beta, 0.0, self.prior_scale).sum(-1)
    r = self.y - bij.sigmoid(s)
    grad = r @ self.x - beta / self.prior_scale ** 2
    return tlp, grad


class LinearRegression(Target):
  """Gaussian linear GLM: parts [beta (D), sigma ()]; beta ~ N(0, s^2 I),
  sigma ~ HalfNormal(sigma_scale), y ~ N(X beta, sigma)."""

  def __init__(self, x, y, prior_scale=1.0, sigma_scale=1.0):
    self.x = np.asarray(x, np.float64)
    self.y = np.asarray(y, np.float64)
    self.prior_scale = float(prior_scale)
    self.sigma_scale = float(sigma_scale)
    self.part_sizes = (self.x.shape[1], 1)

  def log_prob_and_grad(self, state):
    state = np.asarray(state, np.float64)
    d = self.x.shape[1]
    beta, sigma = state[:, :d], state[:, d]
    mu = beta @ self.x.T
    resid = self.y - mu
    with np.errstate(invalid='ignore', divide='ignore'):
      ll = normal_lp(self.y, mu, sigma[:, None]).sum(-1)
      tlp = (ll + normal_lp(beta, 0.0, self.prior_scale).sum(-1)
             + half_normal_lp(sigma, self.sigma_scale))
      gb = (resid / sigma[:, None] ** 2) @ self.x - beta / self.prior_scale ** 2
      n = self.x.shape[0]
      gs = ((resid ** 2).sum(-1) / sigma ** 3 - n / sigma
            - sigma / self.sigma_scale ** 2)
    return tlp, np.concatenate([gb, gs[:, None]], axis=1)


class HierarchicalNormal(Target):
  """Non-centred hierarchical normal (eight-schools family).

  parts [mu (), tau (), z (J)]: mu ~ N(0, mu_scale), tau ~ HalfCauchy(0,
  tau_scale), z_j ~ N(0, 1), y_j ~ N(mu + tau z_j, sigma_j).
  """

  def __init__(self, y, sigma, mu_scale=5.0, tau_scale=5.0):
    self.y = np.asarray(y, np.float64)
    self.sigma = np.asarray(sigma, np.float64)
    self.mu_scale = float(mu_scale)
    self.tau_scale = float(tau_scale)
    self.part_sizes = (1, 1, self.y.shape[0])

  def log_prob_and_grad(self, state):
    state = np.asarray(state, np.float64)
    mu, tau, z = state[:, 0], state[:, 1], state[:, 2:]
    theta = mu[:, None] + tau[:, None] * z
    r = (self.y - theta) / self.sigma ** 2
    with np.errstate(invalid='ignore', divide='ignore'):
      tlp = (normal_lp(mu, 0.0, self.mu_scale)
             + half_cauchy_lp(tau, self.tau_scale)
             + normal_lp(z, 0.0, 1.0).sum(-1)
             + normal_lp(self.y, theta, self.sigma).sum(-1))
    s2 = self.tau_scale ** 2
    g_mu = -mu / self.mu_scale ** 2 + r.sum(-1)
    g_tau = -2.0 * tau / (s2 + tau ** 2) + (r * z).sum(-1)
    g_z = -z + tau[:, None] * r
    return tlp, np.concatenate([g_mu[:, None], g_tau[:, None], g_z], axis=1)


class DiagNormal(Target):
  """Independent normals N(loc_i, scale_i) (closed-form test target)."""

  def __init__(self, loc, scale):
    self.loc = np.asarray(loc, np.float64)
    self.scale = np.asarray(scale, np.float64)
    self.part_sizes = (self.loc.shape[0],)

  def log_prob_and_grad(self, x):
    x = np.asarray(x, np.float64)
    tlp = normal_lp(x, self.loc, self.scale).sum(-1)
    return tlp, -(x - self.loc) / self.scale ** 2


class NealsFunnel(Target):
  """parts [v (), x (dim-1)]: v ~ N(0, 3), x_i ~ N(0, exp(v/2)).

  The divergence test target of `mcmc/nuts_test.py:445-474`.
  """

  def __init__(self, dim=10):
    self.part_sizes = (1, dim - 1)

  def log_prob_and_grad(self, state):
    state = np.asarray(state, np.float64)
    v, x = state[:, 0], state[:, 1:]
    s = np.exp(v / 2.0)
    with np.errstate(over='ignore', invalid='ignore', divide='ignore'):
      tlp = normal_lp(v, 0.0, 3.0) + normal_lp(x, 0.0, s[:, None]).sum(-1)
      k = x.shape[1]
      g_v = -v / 9.0 + 0.5 * ((x ** 2).sum(-1) / s ** 2 - k)
      g_x = -x / s[:, None] ** 2
    return tlp, np.concatenate([g_v[:, None], g_x], axis=1)


class LogGamma(Target):
  """x_i = log of iid Gamma(concentration, rate): the log-gamma target of
  `mcmc/hmc_test.py:146-200` (written there with TransformedDistribution(
  Gamma, Log)); density c x - r e^x + c log r - lgamma(c) per element."""

  def __init__(self, dim, concentration, rate):
    self.part_sizes = (dim,)
    self.c, self.r = float(concentration), float(rate)

  def log_prob_and_grad(self, x):
    import math
    x = np.asarray(x, np.float64)
    with np.errstate(over='ignore'):
      ex = np.exp(x)
      tlp = (self.c * x - self.r * ex).sum(-1) + x.shape[-1] * (
          self.c * math.log(self.r) - math.lgamma(self.c))
      return tlp, self.c - self.r * ex


class NaNExponential(Target):
  """log p = sum(where(x < 0, NaN, -x)): the NaN-rejection target of
  `mcmc/hmc_test.py:398-431`."""

  def __init__(self, dim):
    self.part_sizes = (dim,)

  def log_prob_and_grad(self, x):
    x = np.asarray(x, np.float64)
    tlp = np.where(x < 0, np.nan, -x).sum(-1)
    return tlp, np.where(x < 0, 0.0, -1.0)


class Transformed(Target):
  """tlp_x(x) = tlp(f(x)) + sum fldj(x)  (`mcmc/transformed_kernel.py:86-128`).

  `codes` holds one bijector code per flat element (0 identity, 1 softplus,
  2 exp).
  """

  def __init__(self, base, codes):
    self.base = base
    self.codes = np.asarray(codes, np.int32)
    self.part_sizes = base.part_sizes

  def forward(self, x):
    x = np.asarray(x, np.float64)
    out = np.array(x, copy=True)
    for code in (bij.SOFTPLUS, bij.EXP):
      m = self.codes == code
      if m.any():
        out[:, m] = bij.forward(code, x[:, m])
    return out

  def inverse(self, y):
    y = np.asarray(y, np.float64)
    out = np.array(y, copy=True)
    for code in (bij.SOFTPLUS, bij.EXP):
      m = self.codes == code
      if m.any():
        out[:, m] = bij.inverse(code, y[:, m])
    return out

  def log_prob_and_grad(self, x):
    x = np.asarray(x, np.float64)
    theta = self.forward(x)
    tlp, g = self.base.log_prob_and_grad(theta)
    g = np.array(g, copy=True)
    for code in (bij.SOFTPLUS, bij.EXP):
      m = self.codes == code
      if m.any():
        tlp = tlp + bij.fldj(code, x[:, m]).sum(-1)
        g[:, m] = (g[:, m] * bij.forward_grad(code, x[:, m])
                   + bij.fldj_grad(code, x[:, m]))
    return tlp, g


def finite_difference_grad(target, x, h=1e-6):
  """Central differences of target.log_prob at x [C, D] (float64)."""
  x = np.asarray(x, np.float64)
  g = np.zeros_like(x)
  for i in range(x.shape[1]):
    e = np.zeros_like(x)
    e[:, i] = h
    g[:, i] = (target.log_prob(x + e) - target.log_prob(x - e)) / (2 * h)
  return g
