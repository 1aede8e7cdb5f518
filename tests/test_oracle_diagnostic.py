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
