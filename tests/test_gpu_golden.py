"""The HIP path (float64 kernels, through the TFP-named Python layer)
reproduces the committed golden fixtures of tests/golden/ to 1e-11 relative,
with identical accept flags, leapfrog counts and tree depths."""

import collections
import os

import numpy as np
import pytest
import torch

from probability_amd import bijectors as pbij
from probability_amd import mcmc as pmcmc
from probability_amd import targets as ptg
from probability_amd.mcmc import kernel as pk

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
TOL = 1e-11


def _load(name):
  with np.load(os.path.join(HERE, name + '.npz')) as f:
    return {k: f[k] for k in f.files}


def _np(t):
  return t.detach().cpu().numpy()


def test_hmc_step_matches_golden():
  g = _load('hmc_step_c0')
  spec = ptg.LogisticRegression(g['x'], g['y'], dtype=torch.float64)
  kern = pmcmc.HamiltonianMonteCarlo(spec, step_size=float(g['eps']),
                                     num_leapfrog_steps=int(g['num_steps']))
  x = torch.as_tensor(g['q0'], dtype=torch.float64, device='cuda')
  pkr = kern.bootstrap_results(x)
  nxt, res = kern.one_step(x, pkr, seed=g['seed'])
  np.testing.assert_array_equal(_np(res.is_accepted), g['is_accepted'])
  np.testing.assert_allclose(_np(nxt), g['q1'], rtol=TOL, atol=TOL)
  np.testing.assert_allclose(_np(res.log_accept_ratio), g['log_accept_ratio'],
                             rtol=1e-9, atol=1e-9)
  np.testing.assert_allclose(
      _np(res.accepted_results.target_log_prob), g['target_log_prob'],
      rtol=TOL)
  np.testing.assert_allclose(_np(res.proposed_results.initial_momentum[0]),
                             g['initial_momentum'], rtol=TOL, atol=TOL)


def test_nuts_step_matches_golden():
  g = _load('nuts_step_tiny')
  spec = ptg.HierarchicalNormal(g['y'], g['sigma'], dtype=torch.float64)
  tspec = ptg.TransformedTarget(
      spec, [pbij.Identity(), pbij.Softplus(), pbij.Identity()])
  kern = pmcmc.NoUTurnSampler(tspec, step_size=float(g['eps']),
                              max_tree_depth=10)
  x = torch.as_tensor(g['q0'], dtype=torch.float64, device='cuda')
  parts = spec.unpack_state(x, (x.shape[0],))
  pkr = kern.bootstrap_results(parts)
  nxt, res = kern.one_step(parts, pkr, seed=g['seed'])
  np.testing.assert_array_equal(_np(res.leapfrogs_taken), g['leapfrogs_taken'])
  np.testing.assert_array_equal(_np(res.is_accepted), g['is_accepted'])
  np.testing.assert_array_equal(_np(res.reach_max_depth), g['reach_max_depth'])
  np.testing.assert_array_equal(_np(res.has_divergence), g['has_divergence'])
  np.testing.assert_allclose(_np(spec.pack_state(nxt)), g['q1'], rtol=1e-10,
                             atol=1e-10)
  np.testing.assert_allclose(_np(res.target_log_prob), g['target_log_prob'],
                             rtol=1e-10)
  fin = np.isfinite(g['log_accept_ratio'])
  np.testing.assert_allclose(_np(res.log_accept_ratio)[fin],
                             g['log_accept_ratio'][fin], rtol=1e-9, atol=1e-9)


def test_da_matches_golden():
  g = _load('da_updates')
  lar = torch.as_tensor(g['log_accept_ratio'], device='cuda')
  R = collections.namedtuple('R', 'log_accept_ratio, step_size')

  class Fake(pk.TransitionKernel):
    def __init__(self, step_size, store_parameters_in_results=False):
      self._parameters = dict(step_size=step_size,
                              store_parameters_in_results=True)

    def one_step(self, s, pkr, seed=None):
      return s, pkr

    def bootstrap_results(self, s):
      return R(lar, torch.as_tensor(self._parameters['step_size'],
                                    dtype=torch.float64, device='cuda'))

    @property
    def is_calibrated(self):
      return True

  out = []
  for shape in ([3, 1], [2, 3, 1], [2, 1, 1], [1, 3, 1], [1, 1, 1], [1, 1],
                [1]):
    k = pmcmc.DualAveragingStepSizeAdaptation(Fake(np.ones(shape)),
                                              num_adaptation_steps=1,
                                              target_accept_prob=0.75)
    state = torch.zeros((2, 3, 4), dtype=torch.float64, device='cuda')
    pkr = k.bootstrap_results(state)
    for _ in range(2):
      _, pkr = k.one_step(state, pkr)
    out.append(_np(pkr.inner_results.step_size).ravel())
  np.testing.assert_allclose(np.concatenate(out), g['step_size_after_two'],
                             rtol=1e-12)


def test_diagnostics_match_golden():
  g = _load('ess_rhat_ar1')
  x = torch.as_tensor(g['states'], device='cuda')
  ess = pmcmc.effective_sample_size(x, cross_chain_dims=1)
  rhat = pmcmc.potential_scale_reduction(x, split_chains=True)
  np.testing.assert_allclose(_np(ess), g['ess'], rtol=1e-10)
  np.testing.assert_allclose(_np(rhat), g['rhat'], rtol=1e-12)
