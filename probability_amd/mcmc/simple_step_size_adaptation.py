"""SimpleStepSizeAdaptation (`mcmc/simple_step_size_adaptation.py`) and the
step-size getters/setters shared by the adaptation kernels (`:32-60`)."""

import collections

import numpy as np
import torch

from probability_amd import _lib
from probability_amd.mcmc import kernel as kernel_base
from probability_amd.mcmc.internal import util as mcmc_util


def hmc_like_step_size_setter_fn(kernel_results, new_step_size):
  return mcmc_util.replace_innermost(kernel_results, step_size=new_step_size)


def hmc_like_step_size_getter_fn(kernel_results):
  return mcmc_util.get_innermost(kernel_results, 'step_size')


def hmc_like_log_accept_prob_getter_fn(kernel_results):
  """min(0, lar) with non-finite -> -inf.  The DA kernel applies the same
  map on the device, so this returns log_accept_ratio unchanged (the map is
  idempotent)."""
  return mcmc_util.get_innermost(kernel_results, 'log_accept_ratio')


class SimpleStepSizeAdaptationResults(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple(
        'SimpleStepSizeAdaptationResults',
        'inner_results, target_accept_prob, adaptation_rate, step, '
        'new_step_size')):
  """`simple_step_size_adaptation.py:63-89`."""
  __slots__ = ()


class SimpleStepSizeAdaptation(kernel_base.TransitionKernel):
  """Adapts the inner kernel's `step_size` multiplicatively
  (`simple_step_size_adaptation.py:91-438`).

  After each inner step, every step-size element whose chains' log-mean-exp
  accept probability exceeds log(target_accept_prob) grows by a factor
  (1 + adaptation_rate), the others shrink by it, for the first
  num_adaptation_steps steps.  The chains that feed each element follow the
  reference's reduction rules (leading chain dimensions, then the
  dimensions where log_accept_prob and the step size differ); the update
  runs in libhipmc (hipmc_simple_step_update).  Only the default
  reduce_logmeanexp is built.
  """

  def __init__(self, inner_kernel, num_adaptation_steps,
               target_accept_prob=0.75, adaptation_rate=0.01,
               step_size_setter_fn=hmc_like_step_size_setter_fn,
               step_size_getter_fn=hmc_like_step_size_getter_fn,
               log_accept_prob_getter_fn=hmc_like_log_accept_prob_getter_fn,
               reduce_fn=None, experimental_reduce_chain_axis_names=None,
               validate_args=False, name=None):
    if reduce_fn is not None:
      raise NotImplementedError('only the default reduce_logmeanexp is built')
    if experimental_reduce_chain_axis_names is not None:
      raise NotImplementedError('cross-device chain reduction is not built')
    if validate_args:
      t = np.asarray(target_accept_prob, np.float64)
      if np.any(t <= 0):
        raise ValueError('`target_accept_prob` must be > 0.')
      if np.any(t >= 1):
        raise ValueError('`target_accept_prob` must be < 1.')
    inner_kernel = mcmc_util.enable_store_parameters_in_results(inner_kernel)
    self._parameters = dict(
        inner_kernel=inner_kernel, num_adaptation_steps=num_adaptation_steps,
        target_accept_prob=target_accept_prob,
        adaptation_rate=adaptation_rate,
        step_size_setter_fn=step_size_setter_fn,
        step_size_getter_fn=step_size_getter_fn,
        log_accept_prob_getter_fn=log_accept_prob_getter_fn,
        reduce_fn=reduce_fn,
        experimental_reduce_chain_axis_names=experimental_reduce_chain_axis_names,
        validate_args=validate_args, name=name)
    self._plans = {}

  inner_kernel = property(lambda s: s._parameters['inner_kernel'])
  num_adaptation_steps = property(
      lambda s: s._parameters['num_adaptation_steps'])
  name = property(lambda s: s._parameters['name'])

  @property
  def is_calibrated(self):
    return self.inner_kernel.is_calibrated

  def step_size_setter_fn(self, kernel_results, new_step_size):
    return self._parameters['step_size_setter_fn'](kernel_results,
                                                   new_step_size)

  def step_size_getter_fn(self, kernel_results):
    return self._parameters['step_size_getter_fn'](kernel_results)

  def log_accept_prob_getter_fn(self, kernel_results):
    return self._parameters['log_accept_prob_getter_fn'](kernel_results)

  def _plan(self, lap_shape, step_shape, state_rank, device):
    from probability_amd.mcmc import dual_averaging_step_size_adaptation as da
    key = (tuple(lap_shape), tuple(step_shape), state_rank, str(device))
    p = self._plans.get(key)
    if p is None:
      p = da._GroupPlan(lap_shape, step_shape, state_rank, device)
      self._plans[key] = p
    return p

  def one_step(self, current_state, previous_kernel_results, seed=None):
    pkr = previous_kernel_results
    inner_results = self.step_size_setter_fn(pkr.inner_results,
                                             pkr.new_step_size)
    kw = {} if seed is None else dict(seed=seed)
    new_state, new_inner = self.inner_kernel.one_step(current_state,
                                                      inner_results, **kw)
    lap = self.log_accept_prob_getter_fn(new_inner)
    if not isinstance(lap, torch.Tensor):
      raise TypeError('log_accept_prob must be a device tensor')
    step_size = self.step_size_getter_fn(new_inner)
    is_list = mcmc_util.is_list_like(step_size)
    step_parts = list(step_size) if is_list else [step_size]
    states = (list(current_state) if mcmc_util.is_list_like(current_state)
              else [current_state])
    # one lone step size broadcasts over every part (rank of any part)
    if len(step_parts) == 1:
      states = states[:1]
    new_steps = [self._update_part(ss, st, lap, pkr)
                 for ss, st in zip(step_parts, states)]
    return new_state, pkr._replace(
        inner_results=new_inner, step=pkr.step + 1,
        new_step_size=new_steps if is_list else new_steps[0])

  def _update_part(self, step_size, state, lap, pkr):
    device = lap.device
    if not isinstance(step_size, torch.Tensor):
      step_size = torch.as_tensor(np.asarray(step_size), device=device)
    dtype = step_size.dtype if step_size.is_floating_point() else lap.dtype
    if dtype not in (torch.float32, torch.float64):
      dtype = lap.dtype
    rank = state.dim() if isinstance(state, torch.Tensor) else np.ndim(state)
    plan = self._plan(lap.shape, step_size.shape, rank, device)
    lap_flat = lap.reshape(-1).to(dtype)
    if not plan.identity:
      lap_flat = lap_flat.index_select(0, plan.perm)
    lap_flat = lap_flat.contiguous()
    shape = step_size.shape
    tgt = _f64(pkr.target_accept_prob, shape, device).contiguous()
    new_step = step_size.to(device=device, dtype=dtype).reshape(-1).clone()
    _lib.check(_lib.lib().hipmc_simple_step_update(
        _lib.dtype_code(dtype), plan.n_steps, _lib.ptr(plan.step_group),
        _lib.ptr(plan.group_start), _lib.ptr(lap_flat), 1, 0, _lib.ptr(tgt),
        float(pkr.adaptation_rate), int(pkr.step),
        int(self.num_adaptation_steps), _lib.ptr(new_step),
        _lib.stream_handle(device)))
    return new_step.reshape(shape)

  def bootstrap_results(self, init_state):
    inner_results = self.inner_kernel.bootstrap_results(init_state)
    return SimpleStepSizeAdaptationResults(
        inner_results=inner_results,
        step=0,
        target_accept_prob=self._parameters['target_accept_prob'],
        adaptation_rate=self._parameters['adaptation_rate'],
        new_step_size=self.step_size_getter_fn(inner_results))


def _f64(v, shape, device):
  # Python / NumPy scalars go through float64 NumPy: torch.as_tensor would
  # round them to float32 first.
  if isinstance(v, torch.Tensor):
    t = v.to(device=device, dtype=torch.float64)
  else:
    t = torch.as_tensor(np.asarray(v, np.float64), device=device)
  return t.expand(shape).reshape(-1) if t.dim() <= len(shape) else t.reshape(-1)
