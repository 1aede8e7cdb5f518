"""The oracle's MALA step (`mcmc/langevin.py`): pinned by the identity with
one-leapfrog HMC at step sqrt(eps) (same proposal and the same correction,
written differently), by a planted sign error failing it, and by
stationarity on a Gaussian (the reference's own test targets)."""

import numpy as np

from oracle import mcmc as omcmc
from oracle import rng as orng
from oracle import targets as otg


def _gauss(d, seed=0):
  r = np.random.default_rng(seed)
  return otg.DiagNormal(r.normal(size=d), r.uniform(0.5, 2.0, d))


def test_mala_equals_one_step_hmc_at_sqrt_eps():
  tg = _gauss(5)
  r = np.random.default_rng(1)
  q = r.normal(size=(16, 5))
  t, g = tg.log_prob_and_grad(q)
  for eps in (0.01, 0.3, 1.7):
    seed = orng.sanitize_seed(int(eps * 100))
    qm, rm = omcmc.mala_one_step(tg, eps, q, t, g, seed)
    qh, rh = omcmc.hmc_one_step(tg, np.sqrt(eps), 1, q, t, g, seed)
    np.testing.assert_allclose(rm.proposed_state, rh.proposed_state,
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(rm.log_acceptance_correction,
                               rh.log_acceptance_correction, rtol=1e-9,
                               atol=1e-9)
    np.testing.assert_array_equal(rm.is_accepted, rh.is_accepted)
    np.testing.assert_allclose(qm, qh, rtol=1e-12, atol=1e-12)


def test_identity_fails_for_a_sign_error_in_the_correction():
  # a planted mistake (backward drift with the wrong sign) breaks the
  # identity the previous test relies on
  tg = _gauss(5)
  q = np.random.default_rng(2).normal(size=(16, 5))
  t, g = tg.log_prob_and_grad(q)
  eps = 0.3
  seed = orng.sanitize_seed(3)
  _, rm = omcmc.mala_one_step(tg, eps, q, t, g, seed)
  q1 = rm.proposed_state
  t1, g1 = tg.log_prob_and_grad(q1)
  vol = np.sqrt(eps)
  wrong = (np.sum(0.5 * ((q1 - q - 0.5 * eps * g) / vol) ** 2, -1) -
           np.sum(0.5 * ((q1 - q - 0.5 * eps * g1) / vol) ** 2, -1))
  assert np.max(np.abs(wrong - rm.log_acceptance_correction)) > 1e-3


def test_correction_is_the_normal_proposal_density_ratio():
  # closed form: log N(q | q1 + drift(q1), eps) - log N(q1 | q + drift(q), eps)
  tg = _gauss(3, seed=5)
  q = np.random.default_rng(6).normal(size=(8, 3))
  t, g = tg.log_prob_and_grad(q)
  eps = 0.2
  _, rm = omcmc.mala_one_step(tg, eps, q, t, g, orng.sanitize_seed(9))
  q1 = rm.proposed_state
  _, g1 = tg.log_prob_and_grad(q1)

  def lognorm(x, mu):
    return np.sum(-0.5 * (x - mu) ** 2 / eps - 0.5 * np.log(2 * np.pi * eps),
                  -1)
  want = lognorm(q, q1 + 0.5 * eps * g1) - lognorm(q1, q + 0.5 * eps * g)
  np.testing.assert_allclose(rm.log_acceptance_correction, want, rtol=1e-10,
                             atol=1e-10)


def test_stationary_on_a_normal():  # langevin_test.py:35-58
  tg = otg.DiagNormal(np.zeros(1), np.ones(1))
  c = 256
  q = np.zeros((c, 1))
  t, g = tg.log_prob_and_grad(q)
  draws = []
  for k in range(700):
    q, r = omcmc.mala_one_step(tg, 0.75, q, t, g, orng.fold_in(11, k))
    t, g = r.target_log_prob, r.grads_target_log_prob
    if k >= 200:
      draws.append(q[:, 0].copy())
  x = np.stack(draws)
  assert abs(x.mean()) < 0.05
  assert abs(x.std() - 1.0) < 0.05
