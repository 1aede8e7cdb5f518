"""TransformedTransitionKernel (`mcmc/transformed_kernel.py`).

The inner kernel is rebuilt around `TransformedTarget(spec, bijectors)`
(the role of `make_transformed_log_prob`, `:86-128`), whose log-density and
gradient kernels add each part's forward_log_det_jacobian in-kernel.  The
state maps (forward after each step, inverse at bootstrap) run in libhipmc.
"""

import collections

import torch

from probability_amd import _lib
from probability_amd import bijectors as bij_lib
from probability_amd import targets as targets_lib
from probability_amd.mcmc import kernel as kernel_base
from probability_amd.mcmc.internal import util as mcmc_util

__all__ = ['TransformedTransitionKernel', 'TransformedTransitionKernelResults']


class TransformedTransitionKernelResults(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple('TransformedTransitionKernelResults',
                           ['transformed_state', 'inner_results'])):
  """`mcmc/transformed_kernel.py:33-40`."""
  __slots__ = ()


def _make_kernel_stack(kernel):
  stack = [kernel]
  while 'target_log_prob_fn' not in kernel.parameters:
    if 'inner_kernel' not in kernel.parameters:
      raise ValueError('"None of the nested `inner_kernel`s contains a '
                       '`target_log_prob_fn`."')
    kernel = kernel.parameters['inner_kernel']
    stack.append(kernel)
  return stack


def _update_target_log_prob(kernel, new_target):
  """`mcmc/transformed_kernel.py:147-162`."""
  stack = _make_kernel_stack(kernel)
  prev = stack.pop().copy(target_log_prob_fn=new_target)
  while stack:
    prev = stack.pop().copy(inner_kernel=prev)
  return prev


def _map_state(spec, bijectors, parts, direction):
  """Elementwise forward (0) / inverse (1) of every part, on the device."""
  dev_parts = [targets_lib._to_device(p, spec.dtype, spec.device) for p in parts]
  cs = spec.chain_shape(dev_parts)
  flat = spec.pack_state(dev_parts)
  out = torch.empty_like(flat)
  _lib.check(_lib.lib().hipmc_bijector(
      _lib.dtype_code(spec.dtype), _lib.ptr(spec.table(bijectors)), None,
      flat.shape[0], flat.shape[1], 0, direction, _lib.ptr(flat),
      _lib.ptr(out),
      _lib.stream_handle(spec.device)))
  return spec.unpack_state(out, cs)


class TransformedTransitionKernel(kernel_base.TransitionKernel):
  """Runs the inner kernel in the bijectors' unconstrained space."""

  def __init__(self, inner_kernel, bijector, name=None):
    self._parameters = dict(inner_kernel=inner_kernel, bijector=bijector,
                            name=name or 'transformed_kernel')
    stack = _make_kernel_stack(inner_kernel)
    base = targets_lib.as_target(
        stack[-1].parameters['target_log_prob_fn'])
    bijs = list(bijector) if mcmc_util.is_list_like(bijector) else [bijector]
    if len(bijs) == 1 and len(base.event_shapes) > 1:
      bijs = bijs * len(base.event_shapes)
    self._bijectors = [bij_lib.as_bijector(b) for b in bijs]
    self._base = base
    self._transformed_target = targets_lib.TransformedTarget(base,
                                                            self._bijectors)
    self._inner_kernel = _update_target_log_prob(inner_kernel,
                                                 self._transformed_target)

  @property
  def inner_kernel(self):
    return self._parameters['inner_kernel']

  @property
  def bijector(self):
    return self._parameters['bijector']

  @property
  def name(self):
    return self._parameters['name']

  @property
  def is_calibrated(self):
    return self._inner_kernel.is_calibrated

  def _forward(self, parts):
    return _map_state(self._base, self._bijectors, parts, 0)

  def _inverse(self, parts):
    return _map_state(self._base, self._bijectors, parts, 1)

  def one_step(self, current_state, previous_kernel_results, seed=None):
    kw = {} if seed is None else dict(seed=seed)
    t_prev = previous_kernel_results.transformed_state
    t_next, inner = self._inner_kernel.one_step(
        t_prev, previous_kernel_results.inner_results, **kw)
    is_list = mcmc_util.is_list_like(t_next)
    nxt = self._forward(list(t_next) if is_list else [t_next])
    return (nxt if is_list else nxt[0]), TransformedTransitionKernelResults(
        transformed_state=t_next, inner_results=inner)

  def bootstrap_results(self, init_state=None, transformed_init_state=None):
    if (init_state is None) == (transformed_init_state is None):
      raise ValueError('Must specify exactly one of `init_state` '
                       'or `transformed_init_state`.')
    if transformed_init_state is None:
      is_list = mcmc_util.is_list_like(init_state)
      t = self._inverse(list(init_state) if is_list else [init_state])
      transformed_init_state = t if is_list else t[0]
    else:
      if mcmc_util.is_list_like(transformed_init_state):
        transformed_init_state = [
            targets_lib._to_device(s, self._base.dtype, self._base.device)
            for s in transformed_init_state]
      else:
        transformed_init_state = targets_lib._to_device(
            transformed_init_state, self._base.dtype, self._base.device)
    return TransformedTransitionKernelResults(
        transformed_state=transformed_init_state,
        inner_results=self._inner_kernel.bootstrap_results(
            transformed_init_state))
