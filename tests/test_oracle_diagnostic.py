"""Pins ESS / split-R-hat (oracle and the product's device code run on CPU
tensors) against closed forms."""

import numpy as np
import pytest
import torch

from oracle import diagnostic as odiag
from probability_amd.mcmc import diagnostic as pdiag


def _ar1(n, c, phi, seed):
  r = np.random.default_rng(seed)
  x = np.zeros((n, c))
  x[0] = r.normal(size=c) / np.sqrt(1 - phi**2)
  e = r.normal(size=(n, c))
  for t in range(1, n):
    x[t] = phi * x[t - 1] + e[t]
  return x


def test_ess_iid_is_about_n_times_chains():
  x = np.random.default_rng(0).normal(size=(2000, 8))
  ess = odiag.effective_sample_size(x, cross_chain_dims=1)
  assert abs(ess / 16000 - 1) < 0.1


@pytest.mark.parametrize('phi', [0.5, 0.9])
def test_ess_ar1_matches_closed_form(phi):
  x = _ar1(20000, 4, phi, 1)
  ess = odiag.effective_sample_size(x, cross_chain_dims=1)
  expected = 80000 * (1 - phi) / (1 + phi)
  assert abs(ess / expected - 1) < 0.15


def test_rhat_near_one_for_identical_chains_and_large_for_shifted():
  x = np.random.default_rng(2).normal(size=(1000, 6))
  assert abs(odiag.potential_scale_reduction(x) - 1) < 0.01
  y = x + np.arange(6)[None] * 2.0
  assert odiag.potential_scale_reduction(y) > 1.5


def test_rhat_closed_form_two_chains():
  # m chains of n draws: R = ((m+1)/m) s2+/W - (n-1)/(m n)
  x = np.random.default_rng(3).normal(size=(10, 3))
  n, m = 10, 3
  w = x.var(0, ddof=1).mean()
  b = x.mean(0).var(ddof=1)
  s2 = (n - 1) / n * w + b
  ref = (m + 1) / m * s2 / w - (n - 1) / (m * n)
  np.testing.assert_allclose(
      odiag.potential_scale_reduction(x, split_chains=False), ref, rtol=1e-12)


def test_product_diagnostics_equal_oracle():
  x = _ar1(3000, 5, 0.7, 4)[:, :, None] * np.array([1.0, 3.0])
  for cross in (1, None):
    a = odiag.effective_sample_size(x, cross_chain_dims=cross)
    b = pdiag.effective_sample_size(torch.as_tensor(x), cross_chain_dims=cross)
    np.testing.assert_allclose(b.numpy(), a, rtol=1e-10)
  for split in (True, False):
    a = odiag.potential_scale_reduction(x, split_chains=split)
    b = pdiag.potential_scale_reduction(torch.as_tensor(x), split_chains=split)
    np.testing.assert_allclose(b.numpy(), a, rtol=1e-10)


# ------------------------------------------- the reference's own test cases
# (`mcmc/diagnostic_test.py`), each run through the oracle and the product.
def _both_ess(x, cross):
  a = odiag.effective_sample_size(x, cross_chain_dims=cross)
  b = pdiag.effective_sample_size(torch.as_tensor(x), cross_chain_dims=cross)
  np.testing.assert_allclose(b.numpy(), a, rtol=1e-10)
  return a


def _both_rhat(x, icn=1, split=False):
  a = odiag.potential_scale_reduction(x, independent_chain_ndims=icn,
                                      split_chains=split)
  b = pdiag.potential_scale_reduction(torch.as_tensor(x),
                                      independent_chain_ndims=icn,
                                      split_chains=split)
  np.testing.assert_allclose(b.numpy(), a, rtol=1e-10)
  return np.asarray(a)


def test_ess_length10_correlation_is_one_tenth():
  # diagnostic_test.py:148-161: blocks of 10 equal draws -> ESS = N / 10
  iid = np.random.default_rng(10).normal(size=(5000, 1)).astype(np.float32)
  x = (iid * np.ones((5000, 10), np.float32)).reshape(50000)[:, None]
  ess = _both_ess(x, None)
  np.testing.assert_allclose(ess, 5000, rtol=0.1)


def test_cross_chain_ess_well_and_poorly_mixing():
  # diagnostic_test.py:298-331
  r = np.random.default_rng(11)
  x = r.normal(size=(500, 4)).astype(np.float32)
  per_chain = _both_ess(x, None)
  np.testing.assert_allclose(per_chain.sum(), _both_ess(x, 1), rtol=0.05)
  y = x + np.array([-10., -5., 5., 10.], np.float32)
  np.testing.assert_allclose(_both_ess(y, 1), 4.0, rtol=0.05)


def test_cross_chain_ess_multiple_chain_dims():
  # diagnostic_test.py:333-343
  x = np.random.default_rng(12).normal(size=(500, 2, 2)).astype(np.float32)
  per_chain = _both_ess(x, None)
  np.testing.assert_allclose(per_chain.sum(), _both_ess(x, [1, 2]), rtol=0.05)
  # flattening the chain dims is the same statistic as one chain axis
  np.testing.assert_allclose(_both_ess(x, [1, 2]),
                             _both_ess(x.reshape(500, 4), 1), rtol=1e-12)


def test_cross_chain_ess_needs_two_chains():
  x = np.zeros((500, 1))
  with pytest.raises(ValueError, match='there must be > 1 chain'):
    pdiag.effective_sample_size(torch.as_tensor(x), cross_chain_dims=1)


def test_rhat_manual_computation():
  # diagnostic_test.py:405-420
  x = np.arange(15, dtype=np.float32).reshape((5, 3))
  b_div_n = np.var(np.mean(x, axis=0), ddof=1)
  w = np.mean(np.var(x, axis=0, ddof=1))
  s2 = 0.8 * w + b_div_n
  ref = (4. / 3) * s2 / w - 4. / 15
  assert abs(_both_rhat(x) - ref) < 1e-6


@pytest.mark.parametrize('shape,icn,split', [
    ((10000, 4), 1, False), ((10000, 3, 7), 1, False),
    ((10000, 3, 7), 1, True), ((10001, 3, 7), 1, True),
    ((10000, 2, 5, 7), 1, False), ((10000, 2, 3, 5, 7), 2, False)])
def test_rhat_iid_chains_pass(shape, icn, split):
  # diagnostic_test.py:470-549
  r = np.random.default_rng(13)
  x = r.normal(size=shape)
  other = shape[1 + icn:]
  if other:
    x = x * r.uniform(size=other)
  rh = _both_rhat(x, icn, split)
  assert rh.shape == other
  np.testing.assert_allclose(rh, np.ones_like(rh), rtol=0.02, atol=0)


def test_rhat_offset_chains_fail():
  # diagnostic_test.py:551-568
  x = np.random.default_rng(14).normal(size=(10000, 2, 5))
  x += np.linspace(0, 2, 2).reshape(1, 2, 1)
  assert (_both_rhat(x) > 1.2).all()


def test_rhat_linear_trend_needs_split():
  # diagnostic_test.py:570-593: plain R-hat misses a trend shared by all
  # chains; split R-hat catches it.
  r = np.random.default_rng(15)
  x = r.normal(size=(1000, 10)) + np.linspace(0, 1, 1000)[:, None]
  np.testing.assert_allclose(_both_rhat(x), 1.0, rtol=0.02, atol=0)
  x = r.normal(size=(10000, 10)) + np.linspace(0, 10, 10000)[:, None]
  assert _both_rhat(x, split=True) > 1.2


def test_split_rhat_is_rhat_of_the_halves_as_chains():
  # The split puts the first and second halves side by side as chains
  # (diagnostic.py:520-540), not alternating draws: on a drifting series the
  # two differ, so an interleaving mistake fails here.
  r = np.random.default_rng(16)
  for n in (400, 401):
    x = r.normal(size=(n, 3)) + np.linspace(0, 3, n)[:, None]
    h = n // 2
    halves = np.concatenate([x[:h], x[h:2 * h]], axis=1)
    np.testing.assert_allclose(_both_rhat(x, split=True),
                               _both_rhat(halves), rtol=1e-12)
    interleaved = np.concatenate([x[0:2 * h:2], x[1:2 * h:2]], axis=1)
    assert abs(_both_rhat(interleaved) - _both_rhat(halves)) > 0.05


def test_rhat_argument_checks():
  # diagnostic_test.py:595-640
  with pytest.raises(ValueError, match='at least 2 samples'):
    pdiag.potential_scale_reduction(torch.zeros(1, 10))
  with pytest.raises(ValueError, match='at least 4 samples'):
    pdiag.potential_scale_reduction(torch.zeros(3, 10), split_chains=True)
  with pytest.raises(ValueError, match='independent_chain_ndims'):
    pdiag.potential_scale_reduction(torch.zeros(5, 10),
                                    independent_chain_ndims=0)
