"""No-U-Turn Sampler (`mcmc/nuts.py`).

`one_step` is one launch of the per-chain wavefront NUTS state machine
(hipmc_nuts_run): multinomial sampling, generalized U-turn with the
reference's memory schedule, max_energy_diff divergence, exactly the
semantics of `nuts.py:321-1010`, without the lockstep masking.
"""

import collections

import numpy as np
import torch

from probability_amd import _lib
from probability_amd import random as samplers
from probability_amd import targets as targets_lib
from probability_amd.mcmc import kernel as kernel_base
from probability_amd.mcmc import _runner
from probability_amd.mcmc.internal import util as mcmc_util

__all__ = ['NoUTurnSampler', 'NUTSKernelResults',
           'build_tree_uturn_instruction',
           'generate_efficient_write_read_instruction']


class NUTSKernelResults(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple('NUTSKernelResults', [
        'target_log_prob', 'grads_target_log_prob', 'step_size',
        'log_accept_ratio', 'leapfrogs_taken', 'is_accepted',
        'reach_max_depth', 'has_divergence', 'energy', 'seed'])):
  """`mcmc/nuts.py:74-91`."""
  __slots__ = ()


def uturn_instructions(max_depth):
  """(write [2^d], read [2^d, 2]) U-turn memory tables.

  Same tables as `nuts.py:1013-1071`, from the closed form the kernel uses:
  leaf i (0-based) stores its momentum and the cumulative momentum before it
  in slot popcount(i) when i is even (odd leaves write the spare slot
  `max_depth`), and an odd leaf checks slots [popcount(i) - tz(i+1),
  popcount(i)).
  """
  n = 1 << max_depth
  i = np.arange(n)
  pc = np.array([bin(v).count('1') for v in range(n)], np.int32)
  tz = np.array([((v + 1) & -(v + 1)).bit_length() - 1 for v in range(n)],
                np.int32)
  write = np.where(i % 2 == 0, pc, max_depth).astype(np.int32)
  read = np.where((i % 2 == 1)[:, None],
                  np.stack([pc - tz, pc], axis=1), 0).astype(np.int32)
  return write, read


def build_tree_uturn_instruction(max_depth, init_memory=0):
  """(left, right) leaf pairs of every subtree of size >= 2, sorted
  (`nuts.py:1013-1030`)."""
  n = 1 << max_depth
  base = init_memory + 1
  out = []
  size = 2
  while size <= n:
    for start in range(0, n, size):
      out.append((base + start, base + start + size - 1))
    size *= 2
  return np.array(sorted(out), dtype=np.int32)


def generate_efficient_write_read_instruction(instruction_array):
  """Tables for a depth-d instruction array (see `uturn_instructions`)."""
  n = int(np.max(instruction_array)) + 1
  return uturn_instructions(int(round(np.log2(n))))


def _as_parts(state):
  return list(state) if mcmc_util.is_list_like(state) else [state]


class NoUTurnSampler(kernel_base.TransitionKernel):
  """`mcmc/nuts.py:165-503` with a device-side per-chain state machine."""

  def __init__(self, target_log_prob_fn, step_size, max_tree_depth=10,
               max_energy_diff=1000., unrolled_leapfrog_steps=1,
               parallel_iterations=10, experimental_shard_axis_names=None,
               name=None):
    if not isinstance(max_tree_depth, (int, np.integer)) or max_tree_depth < 1:
      raise ValueError('max_tree_depth must be known statically and >= 1 but '
                       'was {}'.format(max_tree_depth))
    if experimental_shard_axis_names is not None:
      raise NotImplementedError('state sharding across devices is not built')
    self._write_instruction, self._read_instruction = uturn_instructions(
        int(max_tree_depth))
    self._parameters = dict(
        target_log_prob_fn=target_log_prob_fn, step_size=step_size,
        max_tree_depth=int(max_tree_depth), max_energy_diff=max_energy_diff,
        unrolled_leapfrog_steps=unrolled_leapfrog_steps,
        parallel_iterations=parallel_iterations,
        experimental_shard_axis_names=experimental_shard_axis_names,
        name=name or 'NoUTurnSampler')

  target_log_prob_fn = property(lambda s: s._parameters['target_log_prob_fn'])
  step_size = property(lambda s: s._parameters['step_size'])
  max_tree_depth = property(lambda s: s._parameters['max_tree_depth'])
  max_energy_diff = property(lambda s: s._parameters['max_energy_diff'])
  unrolled_leapfrog_steps = property(
      lambda s: s._parameters['unrolled_leapfrog_steps'])
  parallel_iterations = property(lambda s: s._parameters['parallel_iterations'])
  name = property(lambda s: s._parameters['name'])
  write_instruction = property(lambda s: s._write_instruction)
  read_instruction = property(lambda s: s._read_instruction)

  @property
  def is_calibrated(self):
    return True

  def _configure(self, run):
    run.args.max_tree_depth = self.max_tree_depth
    run.args.unrolled_leapfrog_steps = int(self.unrolled_leapfrog_steps)
    run.args.max_energy_diff = float(self.max_energy_diff)

  def one_step(self, current_state, previous_kernel_results, seed=None):
    seed = samplers.sanitize_seed(seed)
    if (not mcmc_util.is_list_like(current_state) and
        isinstance(current_state, dict)):
      raise TypeError('NUTS does not currently support nested or '
                      'non-list-like state structures')
    spec = targets_lib.as_target(self.target_log_prob_fn)
    parts = _as_parts(current_state)
    pkr = previous_kernel_results
    state, tlp, grad = _runner.flat_inputs(
        spec, parts, pkr.target_log_prob, _as_parts(pkr.grads_target_log_prob))
    run = _runner.Run(spec, spec.table(spec.bijectors_for_eval()), state, tlp,
                      grad, _runner.step_tensor(pkr.step_size, spec.dtype,
                                                spec.device), seed)
    self._configure(run)
    c = state.shape[0]
    lar = run.want('out_log_accept_ratio', (1, c), spec.dtype)
    lf = run.want('out_leapfrogs', (1, c), torch.int32)
    acc = run.want('out_is_accepted', (1, c), torch.uint8)
    rmd = run.want('out_reach_max_depth', (1, c), torch.uint8)
    div = run.want('out_has_divergence', (1, c), torch.uint8)
    energy = run.want('out_energy', (1, c), spec.dtype)
    run.call(_lib.lib().hipmc_nuts_run)
    cs = tuple(pkr.target_log_prob.shape)
    results = NUTSKernelResults(
        target_log_prob=tlp.reshape(cs),
        grads_target_log_prob=spec.unpack_state(grad, cs),
        step_size=pkr.step_size,
        log_accept_ratio=lar.reshape(cs),
        leapfrogs_taken=lf.reshape(cs),
        is_accepted=acc.reshape(cs).bool(),
        reach_max_depth=rmd.reshape(cs).bool(),
        has_divergence=div.reshape(cs).bool(),
        energy=energy.reshape(cs),
        seed=seed)
    next_parts = spec.unpack_state(state, cs)
    out = next_parts if mcmc_util.is_list_like(current_state) else next_parts[0]
    return out, results

  def bootstrap_results(self, init_state):
    spec = targets_lib.as_target(self.target_log_prob_fn)
    parts = _as_parts(init_state)
    dev_parts = [targets_lib._to_device(p, spec.dtype, spec.device)
                 for p in parts]
    cs = spec.chain_shape(dev_parts)
    tlp, grad = spec.log_prob_and_grad_flat(spec.pack_state(dev_parts),
                                            spec.bijectors_for_eval())
    _runner.step_tensor(self.step_size, spec.dtype, spec.device)
    tlp = tlp.reshape(cs)
    step = self.step_size
    if mcmc_util.is_list_like(step):
      step = [torch.as_tensor(s, dtype=spec.dtype, device=spec.device)
              for s in step]
    else:
      step = torch.as_tensor(step, dtype=spec.dtype, device=spec.device)
    return NUTSKernelResults(
        target_log_prob=tlp,
        grads_target_log_prob=spec.unpack_state(grad, cs),
        step_size=step,
        log_accept_ratio=torch.zeros_like(tlp),
        leapfrogs_taken=torch.zeros(cs, dtype=torch.int32, device=spec.device),
        is_accepted=torch.zeros(cs, dtype=torch.bool, device=spec.device),
        reach_max_depth=torch.zeros(cs, dtype=torch.bool, device=spec.device),
        has_divergence=torch.zeros(cs, dtype=torch.bool, device=spec.device),
        # compute_hamiltonian with unit momenta (`nuts.py:461-500`)
        energy=tlp - 0.5 * spec.dim,
        seed=samplers.zeros_seed())
