"""Metropolis-adjusted Langevin algorithm (`mcmc/langevin.py`).

With the default (constant, unit) volatility the Langevin proposal
state + 0.5 * eps * grad + sqrt(eps) * xi (`_euler_method`, `:579-635`;
`_get_drift`, `:637-690`) is one leapfrog step of size sqrt(eps) from the
momentum xi, and its proposal-density correction
(`_compute_log_acceptance_correction`, `:692-823`) equals that leapfrog's
kinetic-energy difference.  The kernels therefore run the HMC launches of
libhipmc with one step of size sqrt(eps) (`hipmc_hmc_run` for the
calibrated kernel, `hipmc_leapfrog_run` for the proposal alone) and present
the reference's Langevin results.  A `volatility_fn` other than the default
is not built (arbitrary Python callables do not run on the device).
"""

import collections
import math

import torch

from probability_amd import _lib
from probability_amd import random as samplers
from probability_amd import targets as targets_lib
from probability_amd.mcmc import _runner
from probability_amd.mcmc import kernel as kernel_base
from probability_amd.mcmc import metropolis_hastings as mh
from probability_amd.mcmc.internal import util as mcmc_util

__all__ = ['MetropolisAdjustedLangevinAlgorithm', 'UncalibratedLangevin',
           'UncalibratedLangevinKernelResults']


class UncalibratedLangevinKernelResults(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple(
        'UncalibratedLangevinKernelResults',
        ['grads_target_log_prob', 'log_acceptance_correction',
         'target_log_prob', 'volatility', 'grads_volatility',
         'diffusion_drift', 'seed'])):
  """`mcmc/langevin.py:36-52` (results are for the next state)."""
  __slots__ = ()


def _as_parts(state):
  return list(state) if mcmc_util.is_list_like(state) else [state]


def _maybe_flatten(parts, like):
  return parts if mcmc_util.is_list_like(like) else parts[0]


def _leapfrog_step(step_size, dtype, device):
  """sqrt(eps), the leapfrog step that makes the Langevin proposal."""
  if isinstance(step_size, torch.Tensor):
    return torch.sqrt(step_size.to(device=device, dtype=torch.float64))
  if mcmc_util.is_list_like(step_size):
    raise NotImplementedError('per-part step sizes are not built')
  return math.sqrt(float(step_size))


def _results(spec, chain_shape, eps, tlp, grad, corr, seed):
  grads = spec.unpack_state(grad, chain_shape)
  e = (eps.to(grad.dtype) if isinstance(eps, torch.Tensor) else eps)
  return UncalibratedLangevinKernelResults(
      grads_target_log_prob=grads,
      log_acceptance_correction=corr.reshape(chain_shape),
      target_log_prob=tlp.reshape(chain_shape),
      volatility=[torch.ones_like(g) for g in grads],
      grads_volatility=[torch.zeros_like(g) for g in grads],
      diffusion_drift=[0.5 * e * g for g in grads],
      seed=seed)


class UncalibratedLangevin(kernel_base.TransitionKernel):
  """Langevin proposal without the MH step (`mcmc/langevin.py:315-577`)."""

  def __init__(self, target_log_prob_fn, step_size, volatility_fn=None,
               parallel_iterations=10, compute_acceptance=True,
               experimental_shard_axis_names=None, name=None):
    if volatility_fn is not None:
      raise NotImplementedError('only the default (unit) volatility is built')
    if experimental_shard_axis_names is not None:
      raise NotImplementedError('state sharding across devices is not built')
    self._parameters = dict(
        target_log_prob_fn=target_log_prob_fn, step_size=step_size,
        volatility_fn=volatility_fn, parallel_iterations=parallel_iterations,
        compute_acceptance=compute_acceptance,
        experimental_shard_axis_names=experimental_shard_axis_names,
        name=name)

  target_log_prob_fn = property(lambda s: s._parameters['target_log_prob_fn'])
  step_size = property(lambda s: s._parameters['step_size'])
  volatility_fn = property(lambda s: s._parameters['volatility_fn'])
  compute_acceptance = property(lambda s: s._parameters['compute_acceptance'])
  name = property(lambda s: s._parameters['name'])

  @property
  def is_calibrated(self):
    return False

  def one_step(self, current_state, previous_kernel_results, seed=None):
    spec = targets_lib.as_target(self.target_log_prob_fn)
    pkr = previous_kernel_results
    state, tlp, grad = _runner.flat_inputs(
        spec, _as_parts(current_state), pkr.target_log_prob,
        _as_parts(pkr.grads_target_log_prob))
    seed = samplers.sanitize_seed(seed)
    h = _leapfrog_step(self.step_size, spec.dtype, spec.device)
    run = _runner.Run(spec, spec.table(spec.bijectors_for_eval()), state, tlp,
                      grad, _runner.step_tensor(h, spec.dtype, spec.device),
                      seed, sampler='hmc')
    run.args.num_leapfrog_steps = 1
    c, _ = state.shape
    corr = run.want('out_energy', (1, c), spec.dtype)
    run.call(_lib.lib().hipmc_leapfrog_run)
    cs = tuple(pkr.target_log_prob.shape)
    if not self.compute_acceptance:
      corr = torch.zeros_like(corr)
    res = _results(spec, cs, self.step_size, tlp, grad, corr, seed)
    return _maybe_flatten(spec.unpack_state(state, cs), current_state), res

  def bootstrap_results(self, init_state):
    spec = targets_lib.as_target(self.target_log_prob_fn)
    parts = _as_parts(init_state)
    flat = spec.pack_state(parts)
    cs = spec.chain_shape(
        [targets_lib._to_device(p, spec.dtype, spec.device) for p in parts])
    tlp, grad = spec.log_prob_and_grad_flat(flat, spec.bijectors_for_eval())
    return _results(spec, cs, self.step_size, tlp, grad,
                    torch.zeros_like(tlp), samplers.zeros_seed())


class MetropolisAdjustedLangevinAlgorithm(kernel_base.TransitionKernel):
  """MALA (`mcmc/langevin.py:54-312`): MetropolisHastings over
  UncalibratedLangevin, run as one fused launch per step."""

  def __init__(self, target_log_prob_fn, step_size, volatility_fn=None,
               parallel_iterations=10, experimental_shard_axis_names=None,
               name=None):
    self._impl = mh.MetropolisHastings(
        inner_kernel=UncalibratedLangevin(
            target_log_prob_fn=target_log_prob_fn, step_size=step_size,
            volatility_fn=volatility_fn,
            parallel_iterations=parallel_iterations,
            experimental_shard_axis_names=experimental_shard_axis_names,
            name=name))
    self._parameters = self._impl.inner_kernel.parameters.copy()
    self._parameters.pop('compute_acceptance')

  target_log_prob_fn = property(
      lambda s: s._impl.inner_kernel.target_log_prob_fn)
  step_size = property(lambda s: s._impl.inner_kernel.step_size)
  volatility_fn = property(lambda s: s._impl.inner_kernel.volatility_fn)
  name = property(lambda s: s._parameters['name'])

  @property
  def is_calibrated(self):
    return True

  def bootstrap_results(self, init_state):
    return self._impl.bootstrap_results(init_state)

  def one_step(self, current_state, previous_kernel_results, seed=None):
    """One hipmc_hmc_run launch: proposal, correction and MH accept."""
    spec = targets_lib.as_target(self.target_log_prob_fn)
    pkr = previous_kernel_results
    acc = pkr.accepted_results
    state, tlp, grad = _runner.flat_inputs(
        spec, _as_parts(current_state), acc.target_log_prob,
        _as_parts(acc.grads_target_log_prob))
    seed = samplers.sanitize_seed(seed)
    h = _leapfrog_step(self.step_size, spec.dtype, spec.device)
    run = _runner.Run(spec, spec.table(spec.bijectors_for_eval()), state, tlp,
                      grad, _runner.step_tensor(h, spec.dtype, spec.device),
                      seed, sampler='hmc')
    run.args.num_leapfrog_steps = 1
    c, d = state.shape
    lar = run.want('out_log_accept_ratio', (1, c), spec.dtype)
    is_acc = run.want('out_is_accepted', (1, c), torch.uint8)
    corr = run.want('out_energy', (1, c), spec.dtype)
    pst = run.want('out_proposed_state', (c, d), spec.dtype)
    ptlp = run.want('out_proposed_tlp', (c,), spec.dtype)
    run.call(_lib.lib().hipmc_hmc_run)
    cs = tuple(acc.target_log_prob.shape)
    accepted = is_acc.reshape(cs).bool()
    # the proposal's gradient is not kept by the fused launch (the accepted
    # chains' gradient is in `grad`)
    proposed = acc._replace(
        log_acceptance_correction=corr.reshape(cs),
        target_log_prob=ptlp.reshape(cs), grads_target_log_prob=None,
        diffusion_drift=None, seed=samplers.split_seed(seed, 2)[0])
    new_acc = _results(spec, cs, self.step_size, tlp, grad,
                       torch.where(accepted, corr.reshape(cs),
                                   acc.log_acceptance_correction), [])
    results = mh.MetropolisHastingsKernelResults(
        accepted_results=new_acc, is_accepted=accepted,
        log_accept_ratio=lar.reshape(cs),
        proposed_state=_maybe_flatten(spec.unpack_state(pst, cs),
                                      current_state),
        proposed_results=proposed, extra=[], seed=seed)
    return (_maybe_flatten(spec.unpack_state(state, cs), current_state),
            results)
