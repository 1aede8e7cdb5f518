"""Seeded synthetic workloads (NumPy PCG64) for the benchmark configurations.

* logistic_1024x25       plain HMC on Bayesian logistic regression,
                         X [1024, 25] iid N(0,1), beta* ~ N(0, I),
                         y ~ Bernoulli(sigmoid(X beta*)), prior N(0, I);
                         64 chains at 0, L = 10, eps = 0.05, 500 + 1000 steps.
* hier_normal_j256       NUTS + dual averaging (target 0.8) + Softplus TTK on
                         tau, non-centred hierarchical normal: J = 256,
                         sigma_j ~ U(5, 20), theta_j = 4 + 3 z_j*,
                         y_j ~ N(theta_j, sigma_j); mu ~ N(0, 5),
                         tau ~ HalfCauchy(0, 5); 1024 chains, init mu = 0,
                         tau = softplus(0), z = 0; eps0 = 0.1; 1000 adaptation
                         + 2000 draws; max_tree_depth 10.
* logistic_65536x256     large-batch HMC: X [65536, 256] iid N(0,1)/16,
                         4096 chains x 256, L = 20.
* ragged_glms            512 GLMs (half logistic, half Gaussian linear with
                         sigma ~ HalfNormal(1)), N_k ~ LogUniform[128, 8192],
                         D_k ~ U{4..128}, 32 chains each.
* serving_requests       Poisson arrivals at 200 requests/s, alternately a
                         GLM of ragged_glms's size distribution and a
                         hierarchical normal with J ~ U{8..1024}.
"""

import math

import numpy as np


def logistic_data(n, d, seed, x_scale=1.0):
  r = np.random.default_rng(seed)
  x = (r.normal(size=(n, d)) * x_scale).astype(np.float32)
  beta = r.normal(size=d)
  s = x.astype(np.float64) @ beta
  y = (r.uniform(size=n) < 1 / (1 + np.exp(-s))).astype(np.float32)
  return x, y, beta


def logistic_1024x25(seed=0):
  x, y, beta = logistic_data(1024, 25, seed)
  return dict(x=x, y=y, beta=beta, chains=64, num_leapfrog_steps=10,
              step_size=0.05, num_burnin_steps=500, num_results=1000)


def hier_normal(j=256, seed=1):
  r = np.random.default_rng(seed)
  sigma = r.uniform(5, 20, size=j)
  z = r.normal(size=j)
  theta = 4.0 + 3.0 * z
  y = theta + sigma * r.normal(size=j)
  return y.astype(np.float32), sigma.astype(np.float32)


def hier_normal_j256(seed=1):
  y, sigma = hier_normal(256, seed)
  return dict(y=y, sigma=sigma, chains=1024, step_size=0.1,
              target_accept_prob=0.8, num_adaptation_steps=1000,
              num_burnin_steps=1000, num_results=2000, max_tree_depth=10)


def hier_posterior_moments(y, sigma, mu_scale=5.0, tau_scale=5.0, n_mu=801,
                           n_tau=1601):
  """Float64 posterior means/variances of mu and tau by 2-D quadrature.

  Integrating z out of the non-centred model gives
  y_j | mu, tau ~ N(mu, sqrt(sigma_j^2 + tau^2)), so
  p(mu, tau | y) ~ N(mu; 0, mu_scale) HalfCauchy(tau; tau_scale)
                   prod_j N(y_j; mu, sqrt(sigma_j^2 + tau^2)).
  """
  y = np.asarray(y, np.float64)
  s2 = np.asarray(sigma, np.float64) ** 2
  mu = np.linspace(-25, 35, n_mu)
  # tau on a stretched grid to cover the Cauchy tail
  u = np.linspace(0, 1, n_tau)[1:-1]
  tau = tau_scale * np.tan(0.5 * np.pi * u)
  dtau = tau_scale * 0.5 * np.pi / np.cos(0.5 * np.pi * u) ** 2 * (u[1] - u[0])
  v = s2[None, :] + tau[:, None] ** 2                         # [T, J]
  lp_tau = np.log1p((tau / tau_scale) ** 2) * -1.0
  ll = np.zeros((len(tau), len(mu)))
  for i in range(len(tau)):
    ll[i] = (-0.5 * ((y[None, :] - mu[:, None]) ** 2 / v[i][None, :]).sum(1)
             - 0.5 * np.log(v[i]).sum())
  lp = ll + lp_tau[:, None] - 0.5 * (mu[None, :] / mu_scale) ** 2
  w = np.exp(lp - lp.max()) * dtau[:, None]
  w /= w.sum()
  e_mu = (w.sum(0) * mu).sum()
  e_tau = (w.sum(1) * tau).sum()
  v_mu = (w.sum(0) * (mu - e_mu) ** 2).sum()
  v_tau = (w.sum(1) * (tau - e_tau) ** 2).sum()
  return dict(mean_mu=e_mu, var_mu=v_mu, mean_tau=e_tau, var_tau=v_tau)


def ragged_glms(n_problems=512, seed=3, n_min=128, n_max=8192, d_min=4,
                d_max=128):
  r = np.random.default_rng(seed)
  out = []
  for k in range(n_problems):
    n = int(round(math.exp(r.uniform(math.log(n_min), math.log(n_max)))))
    d = int(r.integers(d_min, d_max + 1))
    x = r.normal(size=(n, d)).astype(np.float32)
    beta = r.normal(size=d)
    if k < n_problems // 2:
      s = x.astype(np.float64) @ beta
      y = (r.uniform(size=n) < 1 / (1 + np.exp(-s))).astype(np.float32)
      out.append(dict(kind='logistic', x=x, y=y))
    else:
      sig = abs(r.normal())
      y = (x.astype(np.float64) @ beta + sig * r.normal(size=n)).astype(
          np.float32)
      out.append(dict(kind='linear', x=x, y=y))
  return out


def serving_requests(n_requests, seed=11, rate=200.0, j_min=8, j_max=1024):
  """Arrival times (seconds, a Poisson process of `rate` per second) and
  models of the serving workload: even requests the GLMs of
  ragged_glms(n_requests, seed + 2) (logistic in the first half, linear in
  the second), odd ones a hierarchical normal (hier_normal) with
  J ~ U{j_min..j_max}."""
  r = np.random.default_rng(seed)
  arrivals = np.cumsum(r.exponential(1.0 / rate, size=n_requests))
  glms = ragged_glms(n_requests, seed=seed + 2)
  models = []
  for k in range(n_requests):
    if k % 2 == 0:
      models.append(glms[k])
    else:
      j = int(r.integers(j_min, j_max + 1))
      y, sigma = hier_normal(j, seed=1000 + k)
      models.append(dict(kind='hier', y=y, sigma=sigma))
  return arrivals, models
