"""The oracle's SimpleStepSizeAdaptation rule against the reference's own
known answers (`mcmc/simple_step_size_adaptation_test.py`)."""

import numpy as np
import pytest

from oracle import mcmc as omcmc

RATE = 1.01


def _run(step, lar, state_rank, n_steps=2, num_adapt=1, **kw):
  """The step size the inner kernel holds after n_steps one_step calls."""
  s = np.asarray(step, np.float64)
  for t in range(n_steps - 1):
    s = omcmc.simple_step_update(s, lar, state_rank, t, num_adapt,
                                 adaptation_rate=kw.get('rate', RATE - 1.),
                                 target_accept_prob=kw.get('target', 0.75))
  return s


def test_step_size_increases():  # :164-185
  lar = np.log([0.74, 0.76, 0.76])
  np.testing.assert_allclose(_run(0.1, lar, 1), 0.1 * RATE)


def test_step_size_decreases_and_clips_above_zero():  # :187-209
  lar = np.log([0.24, 1., 100.])
  np.testing.assert_allclose(_run(0.1, lar, 1), 0.1 / RATE)


def test_adaptation_steps():  # :211-233
  got = [_run(0.1, 0., 0, n_steps=k, num_adapt=2) for k in (1, 2, 3, 4)]
  np.testing.assert_allclose(got, [0.1, 0.1 * RATE, 0.1 * RATE**2,
                                   0.1 * RATE**2])


def test_adaptive_adaptation():  # :235-258
  lar = np.log([0.49, 0.49, 0.51])
  np.testing.assert_allclose(_run(0.1, lar, 1, target=0.5, rate=0.1),
                             0.1 / 1.1)


def test_list_step():  # :260-280
  lar = np.log([0.74, 0.76, 0.76])
  np.testing.assert_allclose(_run([0.1, 0.2, 0.3], lar, 1),
                             [0.1 / RATE, 0.2 * RATE, 0.3 * RATE])


def test_chain_targets():  # :303-346
  lar = np.log([0.74, 0.76])
  np.testing.assert_allclose(_run([0.1, 0.2], lar, 1),
                             [0.1 / RATE, 0.2 * RATE])
  np.testing.assert_allclose(_run([0.1, 0.2], lar, 1, target=[0.7, 0.8]),
                             [0.1 * RATE, 0.2 / RATE])


LAR2 = np.log([[0.72, 0.76, 0.73], [0.77, 0.77, 0.73]])
R = RATE


@pytest.mark.parametrize('old,new', [  # :450-508 (second state part)
    (np.ones([3, 1]), [[1 / R], [R], [1 / R]]),
    (np.ones([2, 3, 1]), [[[1 / R], [R], [1 / R]], [[R], [R], [1 / R]]]),
    (np.ones([2, 1, 1]), [[[1 / R]], [[R]]]),
    (np.ones([1, 3, 1]), [[[1 / R], [R], [1 / R]]]),
    (np.ones([1, 1, 1]), [[[1 / R]]]),
    (np.ones([1, 1]), [[1 / R]]),
    (np.ones([1]), [1 / R]),
])
def test_broadcasting_table(old, new):
  np.testing.assert_allclose(_run(old, LAR2, 3), new)
  # the scalar first part reduces over every chain
  np.testing.assert_allclose(_run(1., LAR2, 2), 1 / R)


def test_reduction_is_log_mean_exp_not_mean_of_logs():
  # mean of the log accept ratios would be below log(0.75); the mean of the
  # probabilities is above it
  lar = np.log([0.3, 0.99, 0.99])
  assert np.mean(lar) < np.log(0.75) < np.log(np.mean(np.exp(lar)))
  np.testing.assert_allclose(_run(0.1, lar, 1), 0.1 * RATE)
