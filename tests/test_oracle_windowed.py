"""Oracle checks for windowed adaptation and the preconditioned kernels
(experimental/mcmc/windowed_sampling.py, diagonal_mass_matrix_adaptation.py,
preconditioned_nuts.py, preconditioned_hmc.py).

Pins: the reference's window-size known answers
(windowed_sampling_test.py:277-294); the mass matrix equals the population
variance of the slow window's pooled states
(diagonal_mass_matrix_adaptation_test.py:233-249 in windowed form); the
preconditioned kernels equal the identity-metric kernels on the linearly
rescaled target (a change of variables u = q / sqrt(prec), exact for a
diagonal normal); windowed NUTS recovers a diagonal normal's variances
(windowed_sampling_test.py:404-434).
"""

import numpy as np
import pytest

from oracle import mcmc as om
from oracle import rng as orng
from oracle import targets as otg


@pytest.mark.parametrize('n', [0, 1, 500, 499, 100, 10000])
def test_window_sizes_known_answers(n):
  first, slow, last = om.window_sizes(n)
  assert first + slow + 2 * slow + 4 * slow + 8 * slow + last == n
  if n == 500:
    assert (first, slow, last) == (75, 25, 50)


def test_window_schedule():
  opts = {s: om.window_option(s, 500) for s in range(600)}
  starts = {s: o for s, o in opts.items() if o}
  assert starts == {0: 1, 75: 2, 100: 3, 150: 3, 250: 3, 450: 4}
  # fewer than 20 adaptation steps: every start coincides at step 0 and the
  # clipped option is the last-window update
  assert om.window_option(0, 10) == 4
  assert all(om.window_option(s, 10) == 0 for s in range(1, 30))


def test_log_harmonic_mean_exp():
  x = np.log(np.array([0.5, 1.0, 0.25]))
  want = np.log(1.0 / np.mean(1.0 / np.exp(x)))
  assert np.isclose(om.reduce_log_harmonic_mean_exp(x), want, rtol=1e-14)
  assert om.reduce_log_harmonic_mean_exp(np.array([-np.inf, 0.0])) == -np.inf


def _diag(d=5, seed=0):
  r = np.random.default_rng(seed)
  loc = r.normal(size=d)
  scale = np.exp(r.uniform(-2.5, 2.5, size=d))
  return loc, scale


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_preconditioned_nuts_is_rescaled_identity_nuts(seed):
  loc, scale = _diag(6, seed)
  prec = np.exp(np.random.default_rng(seed + 10).uniform(-2, 2, size=6))
  s = np.sqrt(prec)
  c = 16
  q = np.random.default_rng(seed + 20).normal(size=(c, 6)) * scale + loc
  tq = otg.DiagNormal(loc, scale)
  tu = otg.DiagNormal(loc / s, scale / s)
  t0, g0 = tq.log_prob_and_grad(q)
  u = q / s
  tu0, gu0 = tu.log_prob_and_grad(u)
  sd = orng.sanitize_seed(seed)
  qn, rq = om.nuts_one_step(tq, 0.4, q, t0, g0, sd, precision=prec)
  un, ru = om.nuts_one_step(tu, 0.4, u, tu0, gu0, sd)
  np.testing.assert_array_equal(rq.leapfrogs_taken, ru.leapfrogs_taken)
  np.testing.assert_array_equal(rq.is_accepted, ru.is_accepted)
  np.testing.assert_allclose(qn, un * s, rtol=1e-10, atol=1e-12)
  np.testing.assert_allclose(rq.log_accept_ratio, ru.log_accept_ratio,
                             rtol=1e-9, atol=1e-12)
  np.testing.assert_allclose(rq.energy, ru.energy - 0.5 * 6 * np.log(2 * np.pi),
                             rtol=1e-10)


@pytest.mark.parametrize('seed', [0, 1])
def test_preconditioned_hmc_is_rescaled_identity_hmc(seed):
  loc, scale = _diag(5, seed)
  prec = scale ** 2 * np.exp(np.random.default_rng(seed).uniform(-1, 1, 5))
  s = np.sqrt(prec)
  c = 32
  q = np.random.default_rng(seed + 5).normal(size=(c, 5)) * scale + loc
  tq = otg.DiagNormal(loc, scale)
  tu = otg.DiagNormal(loc / s, scale / s)
  t0, g0 = tq.log_prob_and_grad(q)
  tu0, gu0 = tu.log_prob_and_grad(q / s)
  sd = orng.sanitize_seed(seed + 3)
  qn, rq = om.hmc_one_step(tq, 0.3, 7, q, t0, g0, sd, prec=prec)
  un, ru = om.hmc_one_step(tu, 0.3, 7, q / s, tu0, gu0, sd)
  np.testing.assert_array_equal(rq.is_accepted, ru.is_accepted)
  np.testing.assert_allclose(qn, un * s, rtol=1e-10, atol=1e-12)
  np.testing.assert_allclose(rq.log_accept_ratio, ru.log_accept_ratio,
                             rtol=1e-9, atol=1e-11)


def test_mass_matrix_is_population_variance_of_slow_window():
  loc, scale = _diag(4, 3)
  tq = otg.DiagNormal(loc, scale)
  c, a = 8, 60   # windows: fast 9, slow 3, 6, 12, 24, last 6
  init = np.random.default_rng(1).uniform(-2, 2, size=(c, 4))
  _, _, _, info = om.sample_windowed(tq, init, 0.1, a, 5, seed=11,
                                     max_depth=6)
  first, slow, _ = om.window_sizes(a)
  st, used = info['states'], info['precision_used']
  lo = first
  for size in (slow, 2 * slow, 4 * slow, 8 * slow):
    hi = lo + size
    pooled = st[lo:hi].reshape(-1, 4)
    want = pooled.var(axis=0)            # population variance (ddof 0)
    np.testing.assert_allclose(used[hi][0], want, rtol=1e-10)
    # the window's own steps ran with the previous estimate
    assert np.array_equal(used[hi - 1], used[lo])
    lo = hi


def test_windowed_nuts_recovers_diagonal_variances():
  loc, scale = _diag(4, 7)
  tq = otg.DiagNormal(loc, scale)
  init = np.random.default_rng(2).uniform(-2, 2, size=(8, 4))
  draws, _, ws, _ = om.sample_windowed(tq, init, None or 0.5, 400, 200,
                                       seed=5, max_depth=8)
  np.testing.assert_allclose(ws.precision[0], scale ** 2, rtol=0.2)
  z = (draws.reshape(-1, 4).mean(0) - loc) / (scale / np.sqrt(200.0))
  assert np.all(np.abs(z) < 5.0)
