"""DualAveragingStepSizeAdaptation
(`mcmc/dual_averaging_step_size_adaptation.py`).

The update of `_one_step_part` (`:353-475`; Hoffman & Gelman 2014 sec. 3.2)
runs in libhipmc (hipmc_da_update) with float64 accumulators.  The host only
works out, from the shapes, which chains feed which step-size element (the
reduction rules of `:399-427`) and gathers each group's chains contiguously.
"""

import collections
import math

import numpy as np
import torch

from probability_amd import _lib
from probability_amd import math as pmath
from probability_amd.mcmc import kernel as kernel_base
from probability_amd.mcmc import simple_step_size_adaptation as ssa
from probability_amd.mcmc.internal import util as mcmc_util

__all__ = ['DualAveragingStepSizeAdaptation',
           'DualAveragingStepSizeAdaptationResults']


class DualAveragingStepSizeAdaptationResults(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple(
        'DualAveragingStepSizeAdaptationResults',
        'inner_results, target_accept_prob, log_shrinkage_target, '
        'exploration_shrinkage, step_count_smoothing, decay_rate, error_sum, '
        'log_averaging_step, step, num_adaptation_steps, new_step_size')):
  """`dual_averaging_step_size_adaptation.py:35-71`.

  error_sum has the shape of its step-size part (one copy per element).
  """
  __slots__ = ()


def reduce_mode(reduce_fn):
  """The device reduction for a `reduce_fn` argument: None or
  `reduce_logmeanexp` (the reference's default) or
  `reduce_log_harmonic_mean_exp` (what windowed adaptation passes,
  `experimental/mcmc/windowed_sampling.py:869-873`).  Other callables are
  not built and raise."""
  if reduce_fn is None or reduce_fn is pmath.reduce_logmeanexp:
    return _lib.REDUCE_MEAN
  fn = getattr(reduce_fn, 'func', reduce_fn)  # functools.partial
  if fn is pmath.reduce_logmeanexp:
    return _lib.REDUCE_MEAN
  if fn is pmath.reduce_log_harmonic_mean_exp:
    return _lib.REDUCE_HARMONIC
  raise NotImplementedError(
      'reduce_fn: only reduce_logmeanexp and reduce_log_harmonic_mean_exp '
      'are built')


def chain_groups(lap_shape, step_shape, state_rank):
  """Group structure of one step-size part.

  Returns (group_of_chain [prod(lap_shape)], step_group [prod(step_shape)],
  n_groups): chain c feeds log-mean-exp group group_of_chain[c]; step-size
  element s reads group step_group[s].
  """
  lap_shape, step_shape = tuple(lap_shape), tuple(step_shape)
  num_reduce = max(min(len(lap_shape), state_rank - len(step_shape)), 0)
  red = lap_shape[num_reduce:]
  keep = [i for i in range(len(red))
          if i < len(step_shape) and red[i] == step_shape[i]]
  err_shape = tuple(red[i] if i in keep else 1 for i in range(len(red)))
  err_pad = err_shape + (1,) * max(len(step_shape) - len(err_shape), 0)
  nb = int(np.prod(lap_shape)) if lap_shape else 1
  idx = np.indices(lap_shape).reshape(len(lap_shape), nb) if lap_shape else \
      np.zeros((0, 1), np.int64)
  gidx = [idx[num_reduce + i] if i in keep else np.zeros(nb, np.int64)
          for i in range(len(red))]
  n_groups = int(np.prod(err_shape)) if err_shape else 1
  group_of_chain = (np.ravel_multi_index(gidx, err_shape) if err_shape
                    else np.zeros(nb, np.int64))
  ns = int(np.prod(step_shape)) if step_shape else 1
  sidx = np.indices(step_shape).reshape(len(step_shape), ns) if step_shape \
      else np.zeros((0, 1), np.int64)
  if err_shape:
    sg = [np.where(np.asarray(err_pad)[i] == 1, 0, sidx[i])
          for i in range(len(err_shape))]
    step_group = np.ravel_multi_index(sg, err_shape)
  else:
    step_group = np.zeros(ns, np.int64)
  return group_of_chain.astype(np.int64), step_group.astype(np.int32), n_groups


class _GroupPlan:
  """Device-side gather plan for one step-size part."""

  def __init__(self, lap_shape, step_shape, state_rank, device):
    goc, sg, ng = chain_groups(lap_shape, step_shape, state_rank)
    perm = np.argsort(goc, kind='stable')
    counts = np.bincount(goc, minlength=ng)
    starts = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    self.identity = bool(np.all(perm == np.arange(len(perm))))
    self.perm = torch.as_tensor(perm, device=device)
    self.group_start = torch.as_tensor(starts, device=device)
    self.step_group = torch.as_tensor(sg, device=device)
    self.n_steps = len(sg)


class DualAveragingStepSizeAdaptation(kernel_base.TransitionKernel):
  """Adapts the inner kernel's `step_size` by dual averaging."""

  def __init__(self, inner_kernel, num_adaptation_steps,
               target_accept_prob=0.75, exploration_shrinkage=0.05,
               shrinkage_target=None, step_count_smoothing=10,
               decay_rate=0.75,
               step_size_setter_fn=ssa.hmc_like_step_size_setter_fn,
               step_size_getter_fn=ssa.hmc_like_step_size_getter_fn,
               log_accept_prob_getter_fn=ssa.hmc_like_log_accept_prob_getter_fn,
               reduce_fn=None, experimental_reduce_chain_axis_names=None,
               validate_args=False, name=None):
    reduce_mode(reduce_fn)  # raises for a reduction that is not built
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
        exploration_shrinkage=exploration_shrinkage,
        shrinkage_target=shrinkage_target,
        step_count_smoothing=step_count_smoothing, decay_rate=decay_rate,
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
    key = (tuple(lap_shape), tuple(step_shape), state_rank, str(device))
    p = self._plans.get(key)
    if p is None:
      p = _GroupPlan(lap_shape, step_shape, state_rank, device)
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
    step_size = self.step_size_getter_fn(new_inner)
    is_list = mcmc_util.is_list_like(step_size)
    step_parts = list(step_size) if is_list else [step_size]
    state_parts = (list(current_state) if mcmc_util.is_list_like(current_state)
                   else [current_state])[:len(step_parts)]
    err_parts = _flat(pkr.error_sum)
    avg_parts = _flat(pkr.log_averaging_step)
    lst_parts = _flat(pkr.log_shrinkage_target)
    new_steps, new_errs, new_avgs = [], [], []
    step_no = int(pkr.step)
    for ss, st, err, avg, lst in zip(step_parts, state_parts, err_parts,
                                     avg_parts, lst_parts):
      ns, ne, na = self._update_part(ss, st, err, avg, lst, lap,
                                     pkr.target_accept_prob, step_no, pkr)
      new_steps.append(ns)
      new_errs.append(ne)
      new_avgs.append(na)
    return new_state, pkr._replace(
        inner_results=new_inner,
        error_sum=new_errs if _is_seq(pkr.error_sum) else new_errs[0],
        step=pkr.step + 1,
        log_averaging_step=(new_avgs if _is_seq(pkr.log_averaging_step)
                            else new_avgs[0]),
        new_step_size=new_steps if is_list else new_steps[0])

  def _update_part(self, step_size, state, err, avg, lst, lap, target, step_no,
                   pkr):
    if not isinstance(step_size, torch.Tensor):
      step_size = torch.as_tensor(np.asarray(step_size), device=lap.device)
    device = lap.device
    dtype = step_size.dtype if step_size.is_floating_point() else lap.dtype
    if dtype not in (torch.float32, torch.float64):
      dtype = lap.dtype
    plan = self._plan(lap.shape, step_size.shape, state.dim(), device)
    lap_flat = lap.reshape(-1).to(dtype)
    if not plan.identity:
      lap_flat = lap_flat.index_select(0, plan.perm)
    lap_flat = lap_flat.contiguous()
    n = plan.n_steps
    shape = step_size.shape
    err = _f64(err, shape, device).clone()
    avg = _f64(avg, shape, device).clone()
    lst = _f64(lst, shape, device).contiguous()
    tgt = _f64(target, shape, device).contiguous()
    new_step = step_size.to(device=device, dtype=dtype).reshape(-1).clone()
    step64 = step_size.to(device=device, dtype=torch.float64).reshape(-1).clone()
    _lib.check(_lib.lib().hipmc_da_update(
        _lib.dtype_code(dtype), n, _lib.ptr(plan.step_group), None,
        _lib.ptr(plan.group_start), _lib.ptr(lap_flat), 1, 0, _lib.ptr(err),
        _lib.ptr(avg), _lib.ptr(lst), _lib.ptr(tgt), step_no,
        int(pkr.num_adaptation_steps), float(pkr.exploration_shrinkage),
        float(pkr.step_count_smoothing), float(pkr.decay_rate),
        reduce_mode(self.parameters['reduce_fn']), None, None,
        _lib.ptr(new_step),
        _lib.ptr(step64), _lib.stream_handle(device)))
    return new_step.reshape(shape), err.reshape(shape), avg.reshape(shape)

  def bootstrap_results(self, init_state):
    inner_results = self.inner_kernel.bootstrap_results(init_state)
    step_size = self.step_size_getter_fn(inner_results)
    lap = self.log_accept_prob_getter_fn(inner_results)
    is_list = mcmc_util.is_list_like(step_size)
    step_parts = list(step_size) if is_list else [step_size]
    shrink = self._parameters['shrinkage_target']
    if shrink is None:
      shrink_parts = [None] * len(step_parts)
    else:
      shrink_parts = list(shrink) if mcmc_util.is_list_like(shrink) else [shrink]
      if len(shrink_parts) not in (1, len(step_parts)):
        raise ValueError(
            '`shrinkage_target` should be a Tensor or list of tensors of '
            'same length as `step_size`. Found len(`step_size`) = {} and '
            'len(shrinkage_target) = {}'.format(len(step_parts),
                                                len(shrink_parts)))
      if len(shrink_parts) < len(step_parts):
        shrink_parts = shrink_parts * len(step_parts)
    device = lap.device
    errs, avgs, lsts = [], [], []
    for ss, sh in zip(step_parts, shrink_parts):
      ss = (ss.to(device=device, dtype=torch.float64)
            if isinstance(ss, torch.Tensor) else
            torch.as_tensor(np.asarray(ss, np.float64), device=device))
      errs.append(torch.zeros_like(ss))
      avgs.append(torch.zeros_like(ss))
      if sh is None:
        lsts.append(math.log(10.0) + torch.log(ss))
      else:
        lsts.append(torch.log(torch.as_tensor(np.asarray(sh, np.float64),
                                              device=device)) +
                    torch.zeros_like(ss))
    pack = (lambda v: v) if is_list else (lambda v: v[0])
    return DualAveragingStepSizeAdaptationResults(
        inner_results=inner_results,
        step=0,
        target_accept_prob=self._parameters['target_accept_prob'],
        log_shrinkage_target=pack(lsts),
        exploration_shrinkage=self._parameters['exploration_shrinkage'],
        step_count_smoothing=self._parameters['step_count_smoothing'],
        decay_rate=self._parameters['decay_rate'],
        error_sum=pack(errs),
        log_averaging_step=pack(avgs),
        new_step_size=step_size,
        num_adaptation_steps=int(self.num_adaptation_steps))


def _is_seq(x):
  return isinstance(x, (list, tuple))


def _flat(x):
  return list(x) if _is_seq(x) else [x]


def _f64(v, shape, device):
  # Python / NumPy scalars go through float64 NumPy: torch.as_tensor would
  # round them to float32 first.
  if isinstance(v, torch.Tensor):
    t = v.to(device=device, dtype=torch.float64)
  else:
    t = torch.as_tensor(np.asarray(v, np.float64), device=device)
  return t.expand(shape).reshape(-1) if t.dim() <= len(shape) else t.reshape(-1)
