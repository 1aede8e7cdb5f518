"""Metropolis-Hastings over an uncalibrated proposal kernel
(`mcmc/metropolis_hastings.py`).

The accept/reject (`lar = safe_sum([tlp', -tlp, correction])`,
`log U < lar`, masked select; `:204-246`) runs in libhipmc
(hipmc_mh_accept) on the packed state of the inner kernel's target spec, or,
for an inner kernel over any other target, on its states flattened to one
row per chain.
"""

import collections
import warnings

import torch

from probability_amd import _lib
from probability_amd import random as samplers
from probability_amd import targets as targets_lib
from probability_amd.mcmc import kernel as kernel_base
from probability_amd.mcmc.internal import util as mcmc_util

__all__ = ['MetropolisHastings', 'MetropolisHastingsKernelResults']


class MetropolisHastingsKernelResults(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple('MetropolisHastingsKernelResults', [
        'accepted_results', 'is_accepted', 'log_accept_ratio',
        'proposed_state', 'proposed_results', 'extra', 'seed'])):
  """`mcmc/metropolis_hastings.py:41-56`."""
  __slots__ = ()


def _find_spec(kernel):
  """The target spec of the innermost kernel, or None when no nested kernel
  names a target_log_prob_fn (any kernel with the TransitionKernel protocol
  may propose)."""
  k = kernel
  while True:
    params = getattr(k, 'parameters', None) or {}
    if 'target_log_prob_fn' in params:
      return targets_lib.as_target(params['target_log_prob_fn'])
    if 'inner_kernel' not in params:
      return None
    k = params['inner_kernel']


def _rows(parts, c, dtype):
  """Parts [chains..., event...] -> contiguous rows [C, sum(event sizes)]."""
  return torch.cat([p.to(dtype).reshape(c, -1) for p in parts], 1).contiguous()


def _unrows(rows, like_parts):
  out, off = [], 0
  for p in like_parts:
    size = int(p.numel() // max(1, rows.shape[0]))
    out.append(rows[:, off:off + size].reshape(p.shape).to(p.dtype))
    off += size
  return out


class MetropolisHastings(kernel_base.TransitionKernel):
  """Calibrates an uncalibrated proposal with the MH accept step."""

  def __init__(self, inner_kernel, name=None):
    if inner_kernel.is_calibrated:
      warnings.warn('Supplied `TransitionKernel` is already calibrated. '
                    'Composing `MetropolisHastings` `TransitionKernel` may not '
                    'be required.')
    self._parameters = dict(inner_kernel=inner_kernel, name=name)

  @property
  def inner_kernel(self):
    return self._parameters['inner_kernel']

  @property
  def name(self):
    return self._parameters['name']

  @property
  def is_calibrated(self):
    return True

  def one_step(self, current_state, previous_kernel_results, seed=None):
    seed = samplers.sanitize_seed(seed)
    proposal_seed, acceptance_seed = samplers.split_seed(seed, 2)
    proposed_state, proposed_results = self.inner_kernel.one_step(
        current_state, previous_kernel_results.accepted_results,
        seed=proposal_seed)
    prev = previous_kernel_results.accepted_results
    if (getattr(proposed_results, 'target_log_prob', None) is None or
        getattr(prev, 'target_log_prob', None) is None):
      raise ValueError('"target_log_prob" must be a member of `inner_kernel` '
                       'results.')
    spec = _find_spec(self.inner_kernel)
    if spec is None:
      return self._one_step_any(current_state, proposed_state,
                                proposed_results, prev, acceptance_seed, seed)
    is_list = mcmc_util.is_list_like(current_state)
    cur_parts = list(current_state) if is_list else [current_state]
    prop_parts = list(proposed_state) if is_list else [proposed_state]
    cs = tuple(prev.target_log_prob.shape)
    state = spec.pack_state(cur_parts).clone()
    tlp = prev.target_log_prob.reshape(-1).to(spec.dtype).clone()
    grads = prev.grads_target_log_prob
    grad = spec.pack_state(grads if mcmc_util.is_list_like(grads)
                           else [grads]).clone()
    pstate = spec.pack_state(prop_parts)
    ptlp = proposed_results.target_log_prob.reshape(-1).to(spec.dtype)
    pgr = proposed_results.grads_target_log_prob
    pgrad = spec.pack_state(pgr if mcmc_util.is_list_like(pgr) else [pgr])
    corr = getattr(proposed_results, 'log_acceptance_correction', None)
    if corr is None:
      warnings.warn('Supplied inner `TransitionKernel` does not have a '
                    '`log_acceptance_correction`. Assuming its value is `0.`')
    elif isinstance(corr, torch.Tensor):
      corr = corr.reshape(-1).to(spec.dtype).contiguous()
    else:
      corr = None
    c, d = state.shape
    lar = torch.empty(c, dtype=spec.dtype, device=spec.device)
    acc = torch.empty(c, dtype=torch.uint8, device=spec.device)
    seed_arr = acceptance_seed.copy()
    _lib.check(_lib.lib().hipmc_mh_accept(
        _lib.dtype_code(spec.dtype), c, d,
        seed_arr.ctypes.data_as(_lib.ctypes.c_void_p),
        _lib.ptr(spec.table(spec.bijectors_for_eval())), None, None,
        _lib.ptr(pstate), _lib.ptr(ptlp.contiguous()), _lib.ptr(pgrad),
        _lib.ptr(corr), _lib.ptr(state), _lib.ptr(tlp), _lib.ptr(grad),
        _lib.ptr(lar), _lib.ptr(acc), _lib.stream_handle(spec.device)))
    is_accepted = acc.reshape(cs).bool()
    stripped = mcmc_util.strip_seeds(proposed_results)
    accepted_results = mcmc_util.choose(is_accepted, stripped,
                                        mcmc_util.strip_seeds(prev))
    accepted_results = accepted_results._replace(
        target_log_prob=tlp.reshape(cs),
        grads_target_log_prob=spec.unpack_state(grad, cs))
    next_parts = spec.unpack_state(state, cs)
    next_state = next_parts if is_list else next_parts[0]
    return next_state, MetropolisHastingsKernelResults(
        accepted_results=accepted_results, is_accepted=is_accepted,
        log_accept_ratio=lar.reshape(cs), proposed_state=proposed_state,
        proposed_results=proposed_results, extra=[], seed=seed)

  def _one_step_any(self, current_state, proposed_state, proposed_results,
                    prev, acceptance_seed, seed):
    """Accept/reject for an inner kernel over any target: the proposal and
    current states go to hipmc_mh_accept as rows (one per chain) with a
    one-row problem table that keys the uniforms."""
    is_list = mcmc_util.is_list_like(current_state)
    cur = list(current_state) if is_list else [current_state]
    prop = list(proposed_state) if is_list else [proposed_state]
    ptlp = proposed_results.target_log_prob
    cs = tuple(prev.target_log_prob.shape)
    dev = prev.target_log_prob.device
    if dev.type != 'cuda' or any(not torch.as_tensor(p).is_cuda
                                 for p in cur + prop):
      raise TypeError('MetropolisHastings runs on the GPU: states and '
                      'target log-probs must be CUDA tensors')
    dt = prev.target_log_prob.dtype
    c = int(prev.target_log_prob.numel())
    state, pstate = _rows(cur, c, dt), _rows(prop, c, dt)
    tlp = prev.target_log_prob.reshape(-1).to(dt).clone()
    grads = getattr(prev, 'grads_target_log_prob', None)
    pgr = getattr(proposed_results, 'grads_target_log_prob', None)
    if grads is not None and pgr is not None:
      grad = _rows(list(grads) if mcmc_util.is_list_like(grads) else [grads],
                   c, dt).clone()
      pgrad = _rows(list(pgr) if mcmc_util.is_list_like(pgr) else [pgr], c, dt)
    else:
      grad = torch.zeros_like(state)
      pgrad = torch.zeros_like(state)
    corr = getattr(proposed_results, 'log_acceptance_correction', None)
    if corr is None:
      warnings.warn('Supplied inner `TransitionKernel` does not have a '
                    '`log_acceptance_correction`. Assuming its value is `0.`')
    else:
      corr = torch.as_tensor(corr, device=dev).reshape(-1).to(dt).contiguous()
    prob = _lib.Problem()
    prob.kind = _lib.TARGET_DIAG_NORMAL
    prob.dim = state.shape[1]
    table = torch.frombuffer(bytearray(bytes(prob)), dtype=torch.uint8).to(dev)
    lar = torch.empty(c, dtype=dt, device=dev)
    acc = torch.empty(c, dtype=torch.uint8, device=dev)
    seed_arr = acceptance_seed.copy()
    ptlp_flat = ptlp.reshape(-1).to(dt).contiguous()
    _lib.check(_lib.lib().hipmc_mh_accept(
        _lib.dtype_code(dt), c, state.shape[1],
        seed_arr.ctypes.data_as(_lib.ctypes.c_void_p), _lib.ptr(table), None,
        None, _lib.ptr(pstate), _lib.ptr(ptlp_flat), _lib.ptr(pgrad),
        _lib.ptr(corr), _lib.ptr(state), _lib.ptr(tlp), _lib.ptr(grad),
        _lib.ptr(lar), _lib.ptr(acc), _lib.stream_handle(dev)))
    is_accepted = acc.reshape(cs).bool()
    accepted_results = mcmc_util.choose(
        is_accepted, mcmc_util.strip_seeds(proposed_results),
        mcmc_util.strip_seeds(prev))
    nxt = _unrows(state, cur)
    return (nxt if is_list else nxt[0]), MetropolisHastingsKernelResults(
        accepted_results=accepted_results, is_accepted=is_accepted,
        log_accept_ratio=lar.reshape(cs), proposed_state=proposed_state,
        proposed_results=proposed_results, extra=[], seed=seed)

  def bootstrap_results(self, init_state):
    pkr = self.inner_kernel.bootstrap_results(init_state)
    if getattr(pkr, 'target_log_prob', None) is None:
      raise ValueError('"target_log_prob" must be a member of `inner_kernel` '
                       'results.')
    x = pkr.target_log_prob
    return MetropolisHastingsKernelResults(
        accepted_results=mcmc_util.strip_seeds(pkr),
        is_accepted=torch.ones_like(x, dtype=torch.bool),
        log_accept_ratio=torch.zeros_like(x),
        proposed_state=init_state, proposed_results=pkr, extra=[],
        seed=samplers.zeros_seed())
