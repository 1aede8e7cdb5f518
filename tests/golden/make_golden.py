"""Writes the small golden fixtures of tests/golden/ from the oracle.

Each fixture holds inputs and expected outputs (float64) of one piece of the
path on a tiny, seeded case:

  leapfrog_tiny.npz   10 leapfrog steps of a 5-parameter logistic regression
                      (4 chains, 16 observations) from injected momenta
                      (`mcmc/internal/leapfrog_integrator.py:280-309`).
  hmc_step_c0.npz     one calibrated HMC step (L = 10, eps = 0.05) of the
                      plain-HMC logistic workload restricted to its first 128
                      observations, 64 chains, Philox draws of this library's
                      schedule (`mcmc/hmc.py:661-875`,
                      `mcmc/metropolis_hastings.py:160-254`).
  nuts_step_tiny.npz  one NUTS transition of a J = 8 hierarchical normal with
                      Softplus on tau, 16 chains (`mcmc/nuts.py:321-1010`).
  da_updates.npz      the dual-averaging broadcasting table of
                      `dual_averaging_step_size_adaptation_test.py:605-671`
                      (two updates, state parts [2,3] and [2,3,4]).
  ess_rhat_ar1.npz    ESS and split R-hat of seeded AR(1) chains
                      (`mcmc/diagnostic.py:38-568`).

Usage: python tests/golden/make_golden.py   (run from the repository root)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import diagnostic as odiag  # noqa: E402
from oracle import mcmc as omcmc  # noqa: E402
from oracle import rng as orng  # noqa: E402
from oracle import targets as otg  # noqa: E402


def leapfrog_tiny():
  r = np.random.default_rng(101)
  x = r.normal(size=(16, 5))
  y = (r.uniform(size=16) < 0.5).astype(np.float64)
  t = otg.LogisticRegression(x, y, 1.0)
  q0 = r.normal(size=(4, 5)) * 0.3
  p0 = r.normal(size=(4, 5))
  t0, g0 = t.log_prob_and_grad(q0)
  p1, q1, t1, g1 = omcmc.leapfrog(t, 0.1, 10, p0, q0, t0, g0)
  return dict(x=x, y=y, q0=q0, p0=p0, eps=np.float64(0.1),
              num_steps=np.int64(10), q1=q1, p1=p1, t1=t1, g1=g1)


def hmc_step_c0():
  sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
  from probability_amd import workloads
  wl = workloads.logistic_1024x25()
  x = wl['x'][:128].astype(np.float64)
  y = wl['y'][:128].astype(np.float64)
  t = otg.LogisticRegression(x, y, 1.0)
  r = np.random.default_rng(102)
  q = r.normal(size=(64, 25)) * 0.2
  t0, g0 = t.log_prob_and_grad(q)
  seed = np.asarray(orng.sanitize_seed(2024), np.uint32)
  q1, res = omcmc.hmc_one_step(t, 0.05, 10, q, t0, g0, seed,
                               orng.PhiloxDraws(np.float64))
  return dict(x=x, y=y, q0=q, seed=seed, eps=np.float64(0.05),
              num_steps=np.int64(10), q1=q1,
              log_accept_ratio=res.log_accept_ratio,
              is_accepted=res.is_accepted,
              target_log_prob=res.target_log_prob,
              initial_momentum=res.initial_momentum)


def nuts_step_tiny():
  r = np.random.default_rng(103)
  j = 8
  sig = r.uniform(5, 20, size=j)
  yv = 4 + 3 * r.normal(size=j) + sig * r.normal(size=j)
  codes = np.zeros(j + 2, np.int32)
  codes[1] = 1
  t = otg.Transformed(otg.HierarchicalNormal(yv, sig), codes)
  q = r.normal(size=(16, j + 2)) * 0.5
  t0, g0 = t.log_prob_and_grad(q)
  seed = np.asarray(orng.sanitize_seed(2025), np.uint32)
  q1, res = omcmc.nuts_one_step(t, 0.3, q, t0, g0, seed,
                                orng.PhiloxDraws(np.float64), max_depth=10)
  return dict(y=yv, sigma=sig, q0=q, seed=seed, eps=np.float64(0.3),
              q1=q1, log_accept_ratio=res.log_accept_ratio,
              leapfrogs_taken=res.leapfrogs_taken,
              is_accepted=res.is_accepted,
              reach_max_depth=res.reach_max_depth,
              has_divergence=res.has_divergence,
              target_log_prob=res.target_log_prob, energy=res.energy,
              tree_depth=res.tree_depth)


def da_updates():
  lar = np.log([[0.70, 0.76, 0.73], [0.76, 0.76, 0.73]])
  steps, outs = [], []
  for shape in ([3, 1], [2, 3, 1], [2, 1, 1], [1, 3, 1], [1, 1, 1], [1, 1],
                [1]):
    st = omcmc.da_bootstrap(np.ones(shape))
    for _ in range(2):
      st = omcmc.da_update(st, lar, state_rank=3, num_adaptation_steps=1,
                           target_accept_prob=0.75)
    steps.append(np.ones(shape).ravel())
    outs.append(np.asarray(st.step_size).ravel())
  return dict(log_accept_ratio=lar,
              shapes=np.array([len(s) for s in steps]),
              step_size_after_two=np.concatenate(outs))


def ess_rhat_ar1():
  r = np.random.default_rng(104)
  n, c, d = 500, 8, 3
  phi = np.array([0.0, 0.5, 0.9])
  x = np.zeros((n, c, d))
  x[0] = r.normal(size=(c, d)) / np.sqrt(1 - phi ** 2)
  for i in range(1, n):
    x[i] = phi * x[i - 1] + r.normal(size=(c, d))
  return dict(states=x, ess=odiag.effective_sample_size(x),
              rhat=odiag.potential_scale_reduction(x, split_chains=True))


FIXTURES = {
    'leapfrog_tiny': leapfrog_tiny,
    'hmc_step_c0': hmc_step_c0,
    'nuts_step_tiny': nuts_step_tiny,
    'da_updates': da_updates,
    'ess_rhat_ar1': ess_rhat_ar1,
}


def main():
  for name, fn in FIXTURES.items():
    path = os.path.join(HERE, name + '.npz')
    np.savez_compressed(path, **fn())
    print(path, os.path.getsize(path))


if __name__ == '__main__':
  main()
