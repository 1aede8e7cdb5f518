"""Hamiltonian Monte Carlo (`mcmc/hmc.py`).

`HamiltonianMonteCarlo` keeps the reference's structure -- a
`MetropolisHastings` over `UncalibratedHamiltonianMonteCarlo` -- and its
results types, but `one_step` runs the whole calibrated step (momentum draw,
leapfrog, kinetic-energy correction, MH accept) as one fused HIP launch.
The composition `MetropolisHastings(UncalibratedHamiltonianMonteCarlo)`
is also available and produces the same chain from the same seed, since
both split the step seed into (proposal, acceptance) identically
(`metropolis_hastings.py:183`).
"""

import collections

import torch

from probability_amd import _lib
from probability_amd import random as samplers
from probability_amd import targets as targets_lib
from probability_amd.mcmc import kernel as kernel_base
from probability_amd.mcmc import metropolis_hastings as mh
from probability_amd.mcmc import _runner
from probability_amd.mcmc.internal import util as mcmc_util

__all__ = ['HamiltonianMonteCarlo', 'UncalibratedHamiltonianMonteCarlo',
           'UncalibratedHamiltonianMonteCarloKernelResults']


class UncalibratedHamiltonianMonteCarloKernelResults(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple(
        'UncalibratedHamiltonianMonteCarloKernelResults',
        ['log_acceptance_correction', 'target_log_prob',
         'grads_target_log_prob', 'initial_momentum', 'final_momentum',
         'step_size', 'num_leapfrog_steps', 'seed'])):
  """`mcmc/hmc.py:40-57`."""
  __slots__ = ()


def _as_parts(state):
  return list(state) if mcmc_util.is_list_like(state) else [state]


def _maybe_flatten(parts, like):
  return parts if mcmc_util.is_list_like(like) else parts[0]


class UncalibratedHamiltonianMonteCarlo(kernel_base.TransitionKernel):
  """Proposal half of HMC (`mcmc/hmc.py:538-770`)."""

  def __init__(self, target_log_prob_fn, step_size, num_leapfrog_steps,
               state_gradients_are_stopped=False,
               store_parameters_in_results=False,
               experimental_shard_axis_names=None, name=None):
    if experimental_shard_axis_names is not None:
      raise NotImplementedError('state sharding across devices is not built')
    self._parameters = dict(
        target_log_prob_fn=target_log_prob_fn, step_size=step_size,
        num_leapfrog_steps=num_leapfrog_steps,
        state_gradients_are_stopped=state_gradients_are_stopped,
        store_parameters_in_results=store_parameters_in_results,
        experimental_shard_axis_names=experimental_shard_axis_names,
        name=name)

  @property
  def target_log_prob_fn(self):
    return self._parameters['target_log_prob_fn']

  @property
  def step_size(self):
    return self._parameters['step_size']

  @property
  def num_leapfrog_steps(self):
    return self._parameters['num_leapfrog_steps']

  @property
  def name(self):
    return self._parameters['name']

  @property
  def is_calibrated(self):
    return False

  def _step_and_steps(self, pkr):
    if self._parameters['store_parameters_in_results']:
      return pkr.step_size, int(pkr.num_leapfrog_steps)
    return self.step_size, int(self.num_leapfrog_steps)

  def one_step(self, current_state, previous_kernel_results, seed=None):
    spec = targets_lib.as_target(self.target_log_prob_fn)
    parts = _as_parts(current_state)
    state, tlp, grad = _runner.flat_inputs(
        spec, parts, previous_kernel_results.target_log_prob,
        _as_parts(previous_kernel_results.grads_target_log_prob))
    step_size, nsteps = self._step_and_steps(previous_kernel_results)
    seed = samplers.sanitize_seed(seed)
    run = _runner.Run(spec, spec.table(spec.bijectors_for_eval()), state, tlp,
                      grad, _runner.step_tensor(step_size, spec.dtype,
                                                spec.device), seed,
                      sampler='hmc')
    run.args.num_leapfrog_steps = nsteps
    c, d = state.shape
    corr = run.want('out_energy', (1, c), spec.dtype)
    p0 = run.want('out_initial_momentum', (c, d), spec.dtype)
    p1 = run.want('out_final_momentum', (c, d), spec.dtype)
    run.call(_lib.lib().hipmc_leapfrog_run)
    chain_shape = tuple(previous_kernel_results.target_log_prob.shape)
    next_parts = spec.unpack_state(state, chain_shape)
    results = previous_kernel_results._replace(
        log_acceptance_correction=corr.reshape(chain_shape),
        target_log_prob=tlp.reshape(chain_shape),
        grads_target_log_prob=spec.unpack_state(grad, chain_shape),
        initial_momentum=spec.unpack_state(p0, chain_shape),
        final_momentum=spec.unpack_state(p1, chain_shape),
        seed=seed)
    return _maybe_flatten(next_parts, current_state), results

  def bootstrap_results(self, init_state):
    spec = targets_lib.as_target(self.target_log_prob_fn)
    parts = _as_parts(init_state)
    flat = spec.pack_state(parts)
    chain_shape = spec.chain_shape(
        [targets_lib._to_device(p, spec.dtype, spec.device) for p in parts])
    tlp, grad = spec.log_prob_and_grad_flat(flat, spec.bijectors_for_eval())
    zeros = [torch.zeros_like(p) for p in spec.unpack_state(flat, chain_shape)]
    res = UncalibratedHamiltonianMonteCarloKernelResults(
        log_acceptance_correction=torch.zeros_like(tlp).reshape(chain_shape),
        target_log_prob=tlp.reshape(chain_shape),
        grads_target_log_prob=spec.unpack_state(grad, chain_shape),
        initial_momentum=zeros, final_momentum=[z.clone() for z in zeros],
        step_size=[], num_leapfrog_steps=[], seed=samplers.zeros_seed())
    if self._parameters['store_parameters_in_results']:
      res = res._replace(
          step_size=torch.as_tensor(self.step_size, dtype=spec.dtype,
                                    device=spec.device),
          num_leapfrog_steps=torch.tensor(int(self.num_leapfrog_steps),
                                          dtype=torch.int32))
    return res


class HamiltonianMonteCarlo(kernel_base.TransitionKernel):
  """Calibrated HMC (`mcmc/hmc.py:175-535`)."""

  def __init__(self, target_log_prob_fn, step_size, num_leapfrog_steps,
               state_gradients_are_stopped=False, step_size_update_fn=None,
               store_parameters_in_results=False,
               experimental_shard_axis_names=None, name=None):
    if step_size_update_fn is not None:
      raise NotImplementedError('step_size_update_fn is deprecated upstream')
    self._impl = mh.MetropolisHastings(
        inner_kernel=UncalibratedHamiltonianMonteCarlo(
            target_log_prob_fn=target_log_prob_fn, step_size=step_size,
            num_leapfrog_steps=num_leapfrog_steps,
            state_gradients_are_stopped=state_gradients_are_stopped,
            store_parameters_in_results=store_parameters_in_results,
            experimental_shard_axis_names=experimental_shard_axis_names,
            name=name))
    self._parameters = self._impl.inner_kernel.parameters.copy()
    self._parameters['step_size_update_fn'] = step_size_update_fn

  @property
  def target_log_prob_fn(self):
    return self._impl.inner_kernel.target_log_prob_fn

  @property
  def step_size(self):
    return self._impl.inner_kernel.step_size

  @property
  def num_leapfrog_steps(self):
    return self._impl.inner_kernel.num_leapfrog_steps

  @property
  def is_calibrated(self):
    return True

  @property
  def name(self):
    return self._parameters['name']

  def bootstrap_results(self, init_state):
    return self._impl.bootstrap_results(init_state)

  def one_step(self, current_state, previous_kernel_results, seed=None):
    """Fused calibrated step: one hipmc_hmc_run launch (or, for
    LogisticRegressionBF16 at large batch, the MFMA path)."""
    inner = self._impl.inner_kernel
    spec = targets_lib.as_target(inner.target_log_prob_fn)
    acc = previous_kernel_results.accepted_results
    if (isinstance(spec, targets_lib.LogisticRegressionBF16) and
        spec.mfma_ok(int(acc.target_log_prob.numel()))):
      return self._one_step_mfma(spec, current_state, previous_kernel_results,
                                 seed)
    parts = _as_parts(current_state)
    state, tlp, grad = _runner.flat_inputs(
        spec, parts, acc.target_log_prob, _as_parts(acc.grads_target_log_prob))
    step_size, nsteps = inner._step_and_steps(acc)
    seed = samplers.sanitize_seed(seed)
    run = _runner.Run(spec, spec.table(spec.bijectors_for_eval()), state, tlp,
                      grad, _runner.step_tensor(step_size, spec.dtype,
                                                spec.device), seed,
                      sampler='hmc')
    run.args.num_leapfrog_steps = nsteps
    c, d = state.shape
    lar = run.want('out_log_accept_ratio', (1, c), spec.dtype)
    is_acc = run.want('out_is_accepted', (1, c), torch.uint8)
    corr = run.want('out_energy', (1, c), spec.dtype)
    pst = run.want('out_proposed_state', (c, d), spec.dtype)
    ptlp = run.want('out_proposed_tlp', (c,), spec.dtype)
    p0 = run.want('out_initial_momentum', (c, d), spec.dtype)
    p1 = run.want('out_final_momentum', (c, d), spec.dtype)
    run.call(_lib.lib().hipmc_hmc_run)
    cs = tuple(acc.target_log_prob.shape)
    accepted = is_acc.reshape(cs).bool()
    proposed_results = acc._replace(
        log_acceptance_correction=corr.reshape(cs),
        target_log_prob=ptlp.reshape(cs),
        grads_target_log_prob=None,
        initial_momentum=spec.unpack_state(p0, cs),
        final_momentum=spec.unpack_state(p1, cs),
        seed=samplers.split_seed(seed, 2)[0])
    new_acc = acc._replace(
        log_acceptance_correction=torch.where(
            accepted, proposed_results.log_acceptance_correction,
            acc.log_acceptance_correction),
        target_log_prob=tlp.reshape(cs),
        grads_target_log_prob=spec.unpack_state(grad, cs),
        initial_momentum=mcmc_util.choose(
            accepted, proposed_results.initial_momentum, acc.initial_momentum),
        final_momentum=mcmc_util.choose(
            accepted, proposed_results.final_momentum, acc.final_momentum),
        seed=[])
    results = mh.MetropolisHastingsKernelResults(
        accepted_results=new_acc, is_accepted=accepted,
        log_accept_ratio=lar.reshape(cs),
        proposed_state=_maybe_flatten(spec.unpack_state(pst, cs),
                                      current_state),
        proposed_results=proposed_results, extra=[], seed=seed)
    return _maybe_flatten(spec.unpack_state(state, cs), current_state), results

  def mfma_runner(self, spec, n_chains):
    key = (id(spec), n_chains)
    r = getattr(self, '_mfma', None)
    if r is None or r[0] != key:
      from probability_amd.mcmc import glm_mfma
      self._mfma = (key, glm_mfma.GlmHmcRunner(spec, n_chains))
    return self._mfma[1]

  def _one_step_mfma(self, spec, current_state, pkr, seed):
    inner = self._impl.inner_kernel
    acc = pkr.accepted_results
    parts = _as_parts(current_state)
    state, tlp, grad = _runner.flat_inputs(
        spec, parts, acc.target_log_prob, _as_parts(acc.grads_target_log_prob))
    step_size, nsteps = inner._step_and_steps(acc)
    seed = samplers.sanitize_seed(seed)
    c = state.shape[0]
    run = self.mfma_runner(spec, c)
    lar = torch.empty((1, c), dtype=torch.float32, device=spec.device)
    isacc = torch.empty((1, c), dtype=torch.uint8, device=spec.device)
    run.run(state, tlp, grad, _runner.step_tensor(step_size, spec.dtype,
                                                  spec.device),
            seed, nsteps, out_lar=lar, out_acc=isacc)
    cs = tuple(acc.target_log_prob.shape)
    accepted = isacc.reshape(cs).bool()
    new_acc = acc._replace(target_log_prob=tlp.reshape(cs),
                           grads_target_log_prob=spec.unpack_state(grad, cs),
                           seed=[])
    results = mh.MetropolisHastingsKernelResults(
        accepted_results=new_acc, is_accepted=accepted,
        log_accept_ratio=lar.reshape(cs), proposed_state=None,
        proposed_results=None, extra=[], seed=seed)
    return _maybe_flatten(spec.unpack_state(state, cs), current_state), results
