"""Large-batch HMC with the BF16 MFMA contraction against the float64 oracle
on the same BF16-rounded X.

Tolerance: the chain state enters the products as a BF16 hi/lo pair
(~2^-17 relative) with FP32 accumulation.  Gradients, log-probs and the
states after one HMC transition are held to 1e-5 normwise (max error over
max magnitude; measured 2.3e-6, 4.7e-7 and 1.6e-7), the log accept ratio
to 1e-3 absolute (measured 7.8e-5), and accept decisions must agree wherever
the oracle's |log U - log_accept_ratio| exceeds 1e-3.
"""

import numpy as np
import pytest
import torch

from oracle import mcmc as omcmc
from oracle import rng as orng
from oracle import targets as otg
from probability_amd import mcmc as pmcmc
from probability_amd import targets as ptg

pytestmark = pytest.mark.gpu


def _err(label, got, want):
  """Prints the measured normwise error (max |got - want| / max |want|)."""
  got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
  d = np.abs(got - want).max()
  print(f'ERR {label} abs={d:.3e} rel={d / max(np.abs(want).max(), 1e-300):.3e}',
        flush=True)


def _normwise(got, want, tol):
  got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
  assert np.abs(got - want).max() <= tol * max(np.abs(want).max(), 1.0)


def _case(n=512, c=64, seed=0):
  r = np.random.default_rng(seed)
  x = (r.normal(size=(n, 256)) / 16).astype(np.float32)
  beta = r.normal(size=256)
  y = (r.uniform(size=n) < 1 / (1 + np.exp(-(x @ beta)))).astype(np.float32)
  spec = ptg.LogisticRegressionBF16(x, y)
  xr = spec.x.double().cpu().numpy()  # the BF16-rounded X
  o = otg.LogisticRegression(xr, y.astype(np.float64))
  q = (r.normal(size=(c, 256)) * 0.3).astype(np.float32)
  return spec, o, q


def test_mfma_layout_and_gradient_match_oracle():
  spec, o, q = _case()
  tlp, grad = spec.log_prob_and_grad_flat(torch.as_tensor(q, device='cuda'))
  rt, rg = o.log_prob_and_grad(q.astype(np.float64))
  _err('mfma eval tlp', tlp.cpu().numpy(), rt)
  _err('mfma eval grad', grad.cpu().numpy(), rg)
  _normwise(tlp.cpu().numpy(), rt, 1e-5)
  _normwise(grad.cpu().numpy(), rg, 1e-5)


@pytest.mark.parametrize('n,c', [(512, 64), (1024, 128)])
def test_mfma_hmc_step_matches_oracle(n, c):
  spec, o, q = _case(n, c, seed=n)
  kern = pmcmc.HamiltonianMonteCarlo(spec, step_size=0.05,
                                     num_leapfrog_steps=6)
  x = torch.as_tensor(q, device='cuda')
  pkr = kern.bootstrap_results(x)
  seed = orng.sanitize_seed(55)
  nxt, res = kern.one_step(x, pkr, seed=seed)
  t0, g0 = o.log_prob_and_grad(q.astype(np.float64))
  oq, ores = omcmc.hmc_one_step(o, np.float32(0.05), 6, q.astype(np.float64),
                                t0, g0, seed, orng.PhiloxDraws(np.float32))
  u = orng.mh_uniform(orng.split_seed(seed, 2)[1], np.arange(c), np.float32)
  margin = np.abs(np.log(u.astype(np.float64)) - ores.log_accept_ratio)
  ok = margin > 1e-3
  got_acc = res.is_accepted.cpu().numpy()
  np.testing.assert_array_equal(got_acc[ok], ores.is_accepted[ok])
  np.testing.assert_allclose(res.log_accept_ratio.cpu().numpy(),
                             ores.log_accept_ratio, atol=1e-3, rtol=0)
  sel = got_acc == ores.is_accepted
  _err(f'mfma hmc {n}x{c} lar', res.log_accept_ratio.cpu().numpy(),
       ores.log_accept_ratio)
  _err(f'mfma hmc {n}x{c} state', nxt.cpu().numpy()[sel], oq[sel])
  _normwise(nxt.cpu().numpy()[sel], oq[sel], 1e-5)


def test_mfma_gradient_at_full_size_matches_float64_torch():
  """65536 x 256 BF16 X, 4096 chains, eight row splits: the MFMA gradient
  and log-prob against a float64 torch evaluation of the same BF16-rounded X
  (same tolerance as the small case)."""
  n, c = 65536, 4096
  g = torch.Generator(device='cuda').manual_seed(3)
  x = torch.randn(n, 256, device='cuda', generator=g) / 16
  beta = torch.randn(256, device='cuda', generator=g)
  y = (torch.rand(n, device='cuda', generator=g) <
       torch.sigmoid(x @ beta)).float()
  spec = ptg.LogisticRegressionBF16(x, y)
  q = torch.randn(c, 256, device='cuda', generator=g) * 0.3
  tlp, grad = spec.log_prob_and_grad_flat(q)
  xr, y64, q64 = spec.x.double(), y.double(), q.double()
  s = q64 @ xr.t()                                      # [C, N] logits
  ll = (y64 * s - torch.nn.functional.softplus(s)).sum(1)
  lp = ll - 0.5 * (q64 ** 2).sum(1) - 256 * 0.5 * np.log(2 * np.pi)
  rg = (y64 - torch.sigmoid(s)) @ xr - q64
  del s
  _normwise(tlp.double().cpu().numpy(), lp.cpu().numpy(), 1e-5)
  rg = rg.cpu().numpy()
  _err('mfma full tlp', tlp.double().cpu().numpy(), lp.cpu().numpy())
  _err('mfma full grad', grad.double().cpu().numpy(), rg)
  _normwise(grad.double().cpu().numpy(), rg, 1e-5)
