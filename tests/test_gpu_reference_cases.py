"""The reference's own known-answer tests for the path, on the GPU:
MetropolisHastings over a fake proposal kernel
(`mcmc/metropolis_hastings_test.py:137-330`), the log-gamma expectations and
the thinning equivalence of sample_chain (`mcmc/hmc_test.py:91-200`)."""

import collections
import warnings

import numpy as np
import pytest
import torch
from scipy import special

from probability_amd import mcmc as pmcmc
from probability_amd import targets as ptg
from probability_amd.mcmc import kernel as kernel_base

pytestmark = pytest.mark.gpu

ResultsNoCorrection = collections.namedtuple(
    'ResultsNoCorrection',
    ['target_log_prob', 'grads_target_log_prob', 'extraneous'])
ResultsWithCorrection = collections.namedtuple(
    'ResultsWithCorrection',
    ['log_acceptance_correction', 'target_log_prob', 'grads_target_log_prob',
     'extraneous'])


class FakeKernel(kernel_base.TransitionKernel):
  """Proposes part i scaled by (i + 2) and halves every results field
  (`metropolis_hastings_test.py:51-118`)."""

  def __init__(self, init_results):
    self._init = init_results
    self.call_count = collections.Counter()

  @property
  def is_calibrated(self):
    self.call_count['is_calibrated'] += 1
    return False

  def one_step(self, current_state, previous_kernel_results, seed=None):
    self.call_count['one_step'] += 1
    nxt = [s * float(i + 2) for i, s in enumerate(current_state)]
    kr = {}
    for fn in previous_kernel_results._fields:
      v = getattr(previous_kernel_results, fn)
      if fn == 'grads_target_log_prob':
        kr[fn] = [0.5 * g for g in v]
      elif fn == 'extraneous':
        kr[fn] = v
      else:
        kr[fn] = 0.5 * v
    return nxt, type(previous_kernel_results)(**kr)

  def bootstrap_results(self, init_state):
    self.call_count['bootstrap_results'] += 1
    return type(self._init)(**{
        fn: (list(v) if fn == 'grads_target_log_prob' else v)
        for fn, v in self._init._asdict().items()})


def _t(x):
  return torch.tensor(x, dtype=torch.float32, device='cuda')


def _run_fake(correction):
  state = [_t([1., 2.]), _t([3., 4.])]
  fields = dict(target_log_prob=_t([100., -100.]),
                grads_target_log_prob=[_t([1.25, 1.5]), _t([2.25, 2.5])],
                extraneous=_t([1.75, 2.]))
  if correction:
    init = ResultsWithCorrection(log_acceptance_correction=_t([300., -300.]),
                                 **fields)
  else:
    init = ResultsNoCorrection(**fields)
  mh = pmcmc.MetropolisHastings(FakeKernel(init))
  pkr = mh.bootstrap_results(state)
  with warnings.catch_warnings(record=True) as rec:
    warnings.simplefilter('always')
    nxt, kr = mh.one_step(state, pkr, seed=[7, 11])
  # a missing correction is taken as 0 with a warning
  warned = [w for w in rec if 'log_acceptance_correction' in str(w.message)]
  assert bool(warned) == (not correction)
  # the untouched field is passed through as the same object
  assert kr.accepted_results.extraneous is pkr.accepted_results.extraneous
  assert kr.proposed_results.extraneous is pkr.accepted_results.extraneous
  assert dict(mh.inner_kernel.call_count) == dict(
      is_calibrated=1, one_step=1, bootstrap_results=1)
  for r in (pkr.accepted_results, pkr.proposed_results, kr.accepted_results,
            kr.proposed_results):
    assert type(r) is type(init)
  np.testing.assert_allclose([p.cpu().numpy() for p in kr.proposed_state],
                             [[2., 4.], [9., 12.]], rtol=1e-6, atol=0)
  assert float(kr.proposed_results.target_log_prob[0]) == 50.
  np.testing.assert_allclose(
      kr.proposed_results.grads_target_log_prob[1].cpu().numpy(),
      [1.125, 1.25])
  return nxt, kr


def test_mh_fake_kernel_without_correction():
  # a very likely current state rejects, a very unlikely one accepts
  nxt, kr = _run_fake(False)
  assert kr.is_accepted.cpu().tolist() == [False, True]
  np.testing.assert_array_equal(kr.log_accept_ratio.cpu().numpy(),
                                [0.5 * 100. - 100., 0.5 * -100. + 100.])
  np.testing.assert_allclose([p.cpu().numpy() for p in nxt],
                             [[1., 2. * 2.], [3., 4. * 3.]])
  # accepted results: chain 1 took the halved proposal, chain 0 kept its own
  np.testing.assert_array_equal(
      kr.accepted_results.target_log_prob.cpu().numpy(), [100., -50.])


def test_mh_fake_kernel_with_correction():
  # the correction overrides the target: accept the first, reject the second
  nxt, kr = _run_fake(True)
  assert kr.is_accepted.cpu().tolist() == [True, False]
  np.testing.assert_array_equal(
      kr.log_accept_ratio.cpu().numpy(),
      [0.5 * 100. - 100. + 0.5 * 300., 0.5 * -100. + 100. + 0.5 * -300.])
  np.testing.assert_allclose([p.cpu().numpy() for p in nxt],
                             [[1. * 2., 2.], [3. * 3., 4.]])


# ------------------------------------------------------------- log-gamma
def _log_gamma_chains(shape, chain_ndims, seed=0):
  """The reference's state x = rand(shape) split into chain and event dims
  (`hmc_test.py:192-194`): chains = prod(shape[:chain_ndims])."""
  x = np.random.default_rng(seed).uniform(size=shape).astype(np.float32)
  c = int(np.prod(shape[:chain_ndims]))
  return torch.as_tensor(x.reshape(c, -1), device='cuda')


@pytest.mark.parametrize('chain_ndims', [0, 1, 2])
def test_hmc_log_gamma_expectations(chain_ndims):
  # hmc_test.py:146-200: 150 burn-in + 150 draws, step 0.05, 2 leapfrogs;
  # E[x] = digamma(5) - log(10) within 0.045, E[exp x] = 0.5 within 0.02 and
  # every acceptance probability above 0.5.  chain_ndims 0 is one chain of
  # 1000 elements.
  x = _log_gamma_chains((50, 10, 2), chain_ndims)
  spec = ptg.LogGamma(x.shape[1], 5.0, 10.0)
  kern = pmcmc.HamiltonianMonteCarlo(spec, step_size=0.05,
                                     num_leapfrog_steps=2)
  for fused in (True, False):
    samples, lar = pmcmc.sample_chain(
        150, x, kernel=kern, num_burnin_steps=150, seed=[5, 17],
        trace_fn=pmcmc.TraceFields('log_accept_ratio'),
        experimental_fused=fused)
    s = samples.double().cpu().numpy()
    assert s.shape == (150,) + tuple(x.shape)
    assert abs(s.mean() - (special.digamma(5.0) - np.log(10.0))) < 0.045
    assert abs(np.exp(s).mean() - 0.5) < 0.02
    acc = np.exp(np.minimum(lar.double().cpu().numpy(), 0.))
    assert (acc > 0.5).all() and (acc <= 1.).all()


def test_hmc_thinning_equals_subsampled_run():
  # hmc_test.py:91-140: 20 draws with no thinning, taken every other one,
  # equal 10 draws with one step between them under the same seed.
  x = _log_gamma_chains((4, 3, 2), 1, seed=1)
  spec = ptg.LogGamma(6, 5.0, 10.0)
  kern = pmcmc.HamiltonianMonteCarlo(spec, step_size=0.1,
                                     num_leapfrog_steps=2)
  for fused in (True, False):
    s0, t0 = pmcmc.sample_chain(20, x, kernel=kern, num_burnin_steps=150,
                                seed=[3, 9], experimental_fused=fused,
                                trace_fn=pmcmc.TraceFields('target_log_prob'))
    s1, t1 = pmcmc.sample_chain(10, x, kernel=kern, num_burnin_steps=150,
                                num_steps_between_results=1, seed=[3, 9],
                                experimental_fused=fused,
                                trace_fn=pmcmc.TraceFields('target_log_prob'))
    np.testing.assert_array_equal(s0[::2].cpu().numpy(), s1.cpu().numpy())
    np.testing.assert_array_equal(t0[::2].cpu().numpy(), t1.cpu().numpy())
    if fused:
      fs, ft = s1, t1
  # and the device loop equals the step loop
  np.testing.assert_allclose(fs.cpu().numpy(), s1.cpu().numpy(), rtol=1e-6,
                             atol=1e-6)
  np.testing.assert_allclose(ft.cpu().numpy(), t1.cpu().numpy(), rtol=1e-6,
                             atol=1e-5)
