"""sample_chain (`mcmc/sample.py:81-383`).

Two executions of the same chain:

* the step loop: `kernel.one_step` per step on the host, one (or two, with
  dual averaging) launches per step -- any kernel stack, any `trace_fn`;
* the fused loop, used when the kernel stack is HMC or NUTS on a target spec,
  optionally inside TransformedTransitionKernel and/or
  DualAveragingStepSizeAdaptation (default getters/setters), and `trace_fn`
  is None or a `TraceFields`: the device runs all post-adaptation steps in
  one launch (one NUTS/HMC launch plus one DA launch per adaptation step).

Step t of either loop uses seed fold_in(sanitize_seed(seed, salt), t), so
both produce the same chain.
"""

import collections
import warnings

import numpy as np
import torch

from probability_amd import _lib
from probability_amd import random as samplers
from probability_amd import targets as targets_lib
from probability_amd.mcmc import dual_averaging_step_size_adaptation as da_lib
from probability_amd.mcmc import hmc as hmc_lib
from probability_amd.mcmc import kernel as kernel_base
from probability_amd.mcmc import metropolis_hastings as mh_lib
from probability_amd.mcmc import nuts as nuts_lib
from probability_amd.mcmc import simple_step_size_adaptation as ssa
from probability_amd.mcmc import transformed_kernel as ttk_lib
from probability_amd.mcmc import _runner
from probability_amd.mcmc.internal import util as mcmc_util

__all__ = ['sample_chain', 'CheckpointableStatesAndTrace', 'StatesAndTrace',
           'TraceFields']


class StatesAndTrace(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple('StatesAndTrace', ['all_states', 'trace'])):
  __slots__ = ()


class CheckpointableStatesAndTrace(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple('CheckpointableStatesAndTrace',
                           ['all_states', 'trace', 'final_kernel_results'])):
  """`mcmc/sample.py:60-77`."""
  __slots__ = ()


_FUSED_FIELDS = ('log_accept_ratio', 'is_accepted', 'leapfrogs_taken',
                 'reach_max_depth', 'has_divergence', 'energy', 'step_size',
                 'target_log_prob')


class TraceFields:
  """trace_fn that returns innermost kernel-results fields by name.

  Works in the step loop like any trace_fn, and lets the fused loop trace
  the same fields from device buffers.
  """

  def __init__(self, *names):
    for n in names:
      if n not in _FUSED_FIELDS:
        raise ValueError(f'cannot trace {n!r}; choose from {_FUSED_FIELDS}')
    self.names = names

  def __call__(self, state, kernel_results):
    out = tuple(mcmc_util.get_innermost(kernel_results, n) for n in self.names)
    return out[0] if len(out) == 1 else out


def _trace_kernel_results(current_state, kernel_results):
  return kernel_results


def _stack(xs):
  if isinstance(xs[0], torch.Tensor):
    return torch.stack(xs)
  if isinstance(xs[0], np.ndarray):
    return np.stack(xs)
  if mcmc_util.is_namedtuple_like(xs[0]):
    return type(xs[0])(*[_stack([getattr(x, f) for x in xs])
                         for f in xs[0]._fields])
  if isinstance(xs[0], (list, tuple)):
    if len(xs[0]) == 0:
      return xs[0]
    return type(xs[0])(_stack([x[i] for x in xs]) for i in range(len(xs[0])))
  if isinstance(xs[0], (int, float, bool, np.generic)):
    return np.asarray(xs)
  return xs


def sample_chain(num_results, current_state, previous_kernel_results=None,
                 kernel=None, num_burnin_steps=0, num_steps_between_results=0,
                 trace_fn=_trace_kernel_results,
                 return_final_kernel_results=False, parallel_iterations=10,
                 seed=None, name=None, experimental_fused=True,
                 experimental_stats=None):
  """Runs `num_burnin_steps + num_results*(1+num_steps_between_results)`
  steps of `kernel` and returns the post-burn-in states."""
  del parallel_iterations, name
  if kernel is None:
    raise ValueError('kernel is required')
  if not kernel.is_calibrated:
    warnings.warn('supplied `TransitionKernel` is not calibrated. Markov '
                  'chain may not converge to intended target distribution.')
  if trace_fn is _trace_kernel_results:
    warnings.warn('Tracing all kernel results by default is deprecated. Set '
                  'the `trace_fn` argument to None (the future default '
                  'value) or an explicit callback that traces the values '
                  'you are interested in.')
  base_seed = samplers.sanitize_seed(seed, salt='mcmc.sample_chain')
  # with no step to take the step loop returns the bootstrap results as
  # they are, as the reference's loop of zero iterations does
  if experimental_fused and num_results + num_burnin_steps > 0:
    plan = _fused_plan(kernel, trace_fn)
    if plan is not None:
      out = _run_fused(plan, num_results, current_state, num_burnin_steps,
                       num_steps_between_results, trace_fn,
                       return_final_kernel_results, base_seed,
                       stats=experimental_stats,
                       previous_kernel_results=previous_kernel_results)
      if out is not None:
        return out
  return _run_steps(kernel, num_results, current_state,
                    previous_kernel_results, num_burnin_steps,
                    num_steps_between_results, trace_fn,
                    return_final_kernel_results, base_seed)


def _run_steps(kernel, num_results, current_state, pkr, num_burnin_steps,
               thin, trace_fn, return_final, base_seed):
  if pkr is None:
    pkr = kernel.bootstrap_results(current_state)
  state = current_state
  t = 0
  states, traces = [], []
  for r in range(num_results):
    n = (1 + num_burnin_steps) if r == 0 else (1 + thin)
    for _ in range(n):
      state, pkr = kernel.one_step(state, pkr, seed=samplers.fold_in(base_seed, t))
      t += 1
    states.append(state)
    if trace_fn is not None:
      traces.append(trace_fn(state, pkr))
  all_states = _stack(states) if states else []
  trace = _stack(traces) if traces else ()
  if return_final:
    return CheckpointableStatesAndTrace(all_states, trace, pkr)
  if trace_fn is None:
    return all_states
  return StatesAndTrace(all_states, trace)


# ---------------------------------------------------------------- fused loop
class _Plan:
  pass


def _fused_plan(kernel, trace_fn):
  if trace_fn is not None and not isinstance(trace_fn, TraceFields):
    return None
  p = _Plan()
  p.outer = kernel
  p.da = None
  p.ttk = None
  k = kernel
  if isinstance(k, da_lib.DualAveragingStepSizeAdaptation):
    prm = k.parameters
    if (prm['step_size_setter_fn'] is not ssa.hmc_like_step_size_setter_fn or
        prm['step_size_getter_fn'] is not ssa.hmc_like_step_size_getter_fn or
        prm['log_accept_prob_getter_fn'] is not
        ssa.hmc_like_log_accept_prob_getter_fn):
      return None
    p.da = k
    k = k.inner_kernel
  if isinstance(k, ttk_lib.TransformedTransitionKernel):
    p.ttk = k
    k = k._inner_kernel
  if isinstance(k, nuts_lib.NoUTurnSampler):
    p.kind = 'nuts'
  elif isinstance(k, hmc_lib.HamiltonianMonteCarlo):
    p.kind = 'hmc'
  else:
    return None
  p.base = k
  if not isinstance(k.target_log_prob_fn, targets_lib.TargetSpec):
    return None
  return p


def _inner_results_of(plan, pkr):
  r = pkr
  if plan.da is not None:
    r = r.inner_results
  if plan.ttk is not None:
    r = r.inner_results
  return r


def _run_fused(plan, num_results, current_state, num_burnin_steps, thin,
               trace_fn, return_final, base_seed, max_iters_per_launch=4096,
               stats=None, previous_kernel_results=None):
  """`stats` (a dict, optional) receives leapfrog totals and, if
  stats['time_kernels'] is set, per-launch HIP-event durations of the
  sampling kernel.

  With `previous_kernel_results` the run resumes from them as the step loop
  does (`mcmc/sample.py:330`): the chains start from their target log-probs
  and gradients (and, under TTK, their transformed state), and dual
  averaging continues from its accumulators and step counter; the
  transitions' seeds count from 0 in each call, as in the step loop."""
  kernel = plan.outer
  pkr = (previous_kernel_results if previous_kernel_results is not None
         else kernel.bootstrap_results(current_state))
  base_res = _inner_results_of(plan, pkr)
  spec = plan.base.target_log_prob_fn
  if plan.ttk is not None:
    ttk_res = pkr.inner_results if plan.da is not None else pkr
    tstate = ttk_res.transformed_state
  else:
    tstate = current_state
  is_list = mcmc_util.is_list_like(tstate)
  parts = list(tstate) if is_list else [tstate]
  acc = base_res.accepted_results if plan.kind == 'hmc' else base_res
  grads = acc.grads_target_log_prob
  state, tlp, grad = _runner.flat_inputs(
      spec, parts, acc.target_log_prob,
      list(grads) if mcmc_util.is_list_like(grads) else [grads])
  cs = tuple(acc.target_log_prob.shape)
  c, d = state.shape
  dev, dt = spec.device, spec.dtype
  if plan.da is not None:
    step0 = pkr.new_step_size
  elif plan.kind == 'hmc':
    step0 = plan.base.step_size
  else:
    step0 = base_res.step_size
  # a working copy: the launches update it in place, and the caller's
  # kernel results may hold the same tensor
  step = _runner.step_tensor(step0, dt, dev).clone()
  # traced step sizes keep the shape the step loop reports (a scalar step
  # size traces as [num_results]); list step sizes stay flat
  step_shape = (None if isinstance(step0, (list, tuple))
                else tuple(torch.as_tensor(step0).shape))
  n_total = (num_burnin_steps + 1 + (num_results - 1) * (1 + thin)
             if num_results else num_burnin_steps)
  table = spec.table(spec.bijectors_for_eval())
  want = trace_fn.names if isinstance(trace_fn, TraceFields) else ()
  rows = max(n_total, 1)
  lar_all = torch.empty((rows, c), dtype=dt, device=dev)
  outs = {}
  if 'is_accepted' in want or return_final:
    outs['out_is_accepted'] = torch.empty((rows, c), dtype=torch.uint8, device=dev)
  if plan.kind == 'nuts':
    for name, f, tdt in (('leapfrogs_taken', 'out_leapfrogs', torch.int32),
                         ('reach_max_depth', 'out_reach_max_depth', torch.uint8),
                         ('has_divergence', 'out_has_divergence', torch.uint8),
                         ('energy', 'out_energy', dt)):
      if name in want or return_final or (stats is not None and
                                          f == 'out_leapfrogs'):
        outs[f] = torch.empty((rows, c), dtype=tdt, device=dev)
  events = [] if (stats is not None and stats.get('time_kernels')) else None
  trace_state = torch.empty((num_results, c, d), dtype=dt, device=dev)
  step_hist = (torch.empty((rows,) + tuple(step.shape), dtype=dt, device=dev)
               if 'step_size' in want else None)
  tlp_hist = None
  if 'target_log_prob' in want:
    # every launch writes each iteration's post-step target log-prob
    tlp_hist = outs['out_target_log_prob'] = torch.empty(
        (rows, c), dtype=dt, device=dev)

  # dual-averaging step number of this call's first transition
  da0 = int(pkr.step) if plan.da is not None else 0
  if plan.da is not None:
    n_adapt = int(pkr.num_adaptation_steps)
    step_shape = tuple(torch.as_tensor(step0).shape)
    err = da_lib._f64(pkr.error_sum, step_shape, dev).clone()
    avg = da_lib._f64(pkr.log_averaging_step, step_shape, dev).clone()
    lst = da_lib._f64(pkr.log_shrinkage_target, step_shape, dev).contiguous()
    tgt = da_lib._f64(pkr.target_accept_prob, step_shape, dev).contiguous()
    step64 = da_lib._f64(step0, step_shape, dev).clone()
    state_rank = len(spec.event_shapes[0]) + len(cs)
    gplan = plan.da._plan(cs, step_shape, state_rank, dev)
    if not gplan.identity:
      return None
  else:
    n_adapt = 0

  fn = (_lib.lib().hipmc_nuts_run if plan.kind == 'nuts'
        else _lib.lib().hipmc_hmc_run)
  # Large-batch logistic HMC runs its all-chain gradients as MFMA GEMMs.
  mfma = (plan.kind == 'hmc' and plan.ttk is None and
          isinstance(spec, targets_lib.LogisticRegressionBF16) and
          spec.mfma_ok(c) and step.numel() == 1)
  if mfma and tlp_hist is not None:
    return None
  stream = _lib.stream_handle(dev)
  t = 0
  while t < n_total:
    if plan.da is not None and da0 + t < n_adapt:
      n = 1
    else:
      n = min(n_total - t, max_iters_per_launch)
    if step_hist is not None:
      step_hist[t:t + n] = step.reshape(step_hist.shape[1:])
    if mfma:
      if events is not None:
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
      plan.base.mfma_runner(spec, c).run(
          state, tlp, grad, step, base_seed, plan.base.num_leapfrog_steps,
          n_iters=n, fold_seed=True, iter0=t, out_lar=lar_all[t],
          out_acc=(outs['out_is_accepted'][t] if 'out_is_accepted' in outs
                   else None),
          trace=trace_state, trace_start=num_burnin_steps,
          trace_every=1 + thin, n_trace=num_results)
      if events is not None:
        e1.record()
        events.append((e0, e1, t, n))
    else:
      _launch_fused(plan, spec, table, state, tlp, grad, step, base_seed, n,
                    t, lar_all, outs, trace_state, num_burnin_steps, thin,
                    num_results, fn, stream, events)
    if plan.da is not None:
      _lib.check(_lib.lib().hipmc_da_update(
          _lib.dtype_code(dt), gplan.n_steps, _lib.ptr(gplan.step_group),
          None,
          _lib.ptr(gplan.group_start), _lib.ptr(lar_all[t]), n, c,
          _lib.ptr(err), _lib.ptr(avg), _lib.ptr(lst), _lib.ptr(tgt),
          da0 + t, n_adapt, float(pkr.exploration_shrinkage),
          float(pkr.step_count_smoothing), float(pkr.decay_rate),
          da_lib.reduce_mode(plan.da.parameters['reduce_fn']), None, None,
          _lib.ptr(step),
          _lib.ptr(step64), stream))
    t += n
  if stats is not None:
    if 'out_leapfrogs' in outs:
      stats['leapfrogs_total'] = int(outs['out_leapfrogs'][:n_total].sum())
      stats['leapfrogs_per_step'] = outs['out_leapfrogs'][:n_total].sum(1)
      stats['leapfrogs_max_per_step'] = (
          outs['out_leapfrogs'][:n_total].max(1).values)
    stats['n_steps'] = n_total
    stats['n_chains'] = c
    if events is not None:
      torch.cuda.synchronize(dev)
      stats['launches'] = [(e0.elapsed_time(e1), t0, n)
                           for e0, e1, t0, n in events]
  # ---- assemble outputs in the reference's structures
  flat_states = trace_state.reshape(num_results * c, d)
  if plan.ttk is not None:
    out_flat = torch.empty_like(flat_states)
    if flat_states.numel():
      _lib.check(_lib.lib().hipmc_bijector(
          _lib.dtype_code(dt), _lib.ptr(table), None, num_results * c, d, 0,
          0, _lib.ptr(flat_states), _lib.ptr(out_flat), stream))
  else:
    out_flat = flat_states
  rec_idx = torch.arange(num_burnin_steps, n_total, 1 + thin, device=dev)
  state_parts = spec.unpack_state(out_flat, (num_results,) + cs)
  all_states = state_parts if is_list else state_parts[0]
  trace = ()
  if isinstance(trace_fn, TraceFields):
    vals = []
    for name in trace_fn.names:
      if name == 'log_accept_ratio':
        v = lar_all[rec_idx]
      elif name == 'is_accepted':
        v = outs['out_is_accepted'][rec_idx].bool()
      elif name == 'leapfrogs_taken':
        v = outs['out_leapfrogs'][rec_idx]
      elif name == 'reach_max_depth':
        v = outs['out_reach_max_depth'][rec_idx].bool()
      elif name == 'has_divergence':
        v = outs['out_has_divergence'][rec_idx].bool()
      elif name == 'energy':
        v = outs['out_energy'][rec_idx]
      elif name == 'step_size':
        v = step_hist[rec_idx]
      elif name == 'target_log_prob':
        v = tlp_hist[rec_idx]
      if name != 'step_size':
        v = v.reshape((num_results,) + cs + tuple(v.shape[2:]))
      elif step_shape is not None:
        v = v.reshape((num_results,) + step_shape)
      vals.append(v)
    trace = vals[0] if len(vals) == 1 else tuple(vals)
  if return_final:
    final = _final_results(plan, pkr, spec, state, tlp, grad, step, cs,
                           lar_all, outs, n_total, base_seed,
                           (err, avg, step64) if plan.da is not None else None,
                           da0)
    return CheckpointableStatesAndTrace(all_states, trace, final)
  if trace_fn is None:
    return all_states
  return StatesAndTrace(all_states, trace)


def _launch_fused(plan, spec, table, state, tlp, grad, step, base_seed, n, t,
                  lar_all, outs, trace_state, num_burnin_steps, thin,
                  num_results, fn, stream, events):
  """One sampler launch of the fused loop over iterations [t, t + n)."""
  run = _runner.Run(spec, table, state, tlp, grad, step, base_seed,
                    n_iters=n, fold_seed=True, iter0=t,
                    sampler='nuts' if plan.kind == 'nuts' else 'hmc')
  a = run.args
  if plan.kind == 'nuts':
    plan.base._configure(run)
  else:
    a.num_leapfrog_steps = int(plan.base.num_leapfrog_steps)
  a.out_row0 = t
  a.out_log_accept_ratio = lar_all.data_ptr()
  for f, buf in outs.items():
    setattr(a, f, buf.data_ptr())
  a.trace_state = trace_state.data_ptr()
  a.trace_start = num_burnin_steps
  a.trace_every = 1 + thin
  a.n_trace = num_results
  if events is not None:
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
  _lib.check(fn(a, stream))
  if events is not None:
    e1.record()
    events.append((e0, e1, t, n))


def _final_results(plan, pkr, spec, state, tlp, grad, step, cs, lar_all, outs,
                   n_total, base_seed, da_tensors, da0=0):
  """Kernel results after the last step, in the reference's nesting."""
  last = n_total - 1
  step_val = step.reshape(torch.as_tensor(
      plan.base.step_size if plan.da is None else pkr.new_step_size).shape)
  last_seed = samplers.fold_in(base_seed, last) if n_total else None
  if plan.kind == 'nuts':
    inner = nuts_lib.NUTSKernelResults(
        target_log_prob=tlp.reshape(cs),
        grads_target_log_prob=spec.unpack_state(grad, cs),
        step_size=step_val,
        log_accept_ratio=lar_all[last].reshape(cs),
        leapfrogs_taken=outs['out_leapfrogs'][last].reshape(cs),
        is_accepted=outs['out_is_accepted'][last].reshape(cs).bool(),
        reach_max_depth=outs['out_reach_max_depth'][last].reshape(cs).bool(),
        has_divergence=outs['out_has_divergence'][last].reshape(cs).bool(),
        energy=outs['out_energy'][last].reshape(cs),
        seed=last_seed)
  else:
    acc = _inner_results_of(plan, pkr).accepted_results
    inner = mh_lib.MetropolisHastingsKernelResults(
        accepted_results=acc._replace(
            target_log_prob=tlp.reshape(cs),
            grads_target_log_prob=spec.unpack_state(grad, cs),
            step_size=step_val if plan.da is not None else acc.step_size),
        is_accepted=outs['out_is_accepted'][last].reshape(cs).bool(),
        log_accept_ratio=lar_all[last].reshape(cs),
        proposed_state=None, proposed_results=None, extra=[],
        seed=last_seed)
  res = inner
  if plan.ttk is not None:
    res = ttk_lib.TransformedTransitionKernelResults(
        transformed_state=spec.unpack_state(state, cs), inner_results=res)
  if plan.da is not None:
    err, avg, step64 = da_tensors
    shape = step_val.shape
    res = pkr._replace(inner_results=res, error_sum=err.reshape(shape),
                       log_averaging_step=avg.reshape(shape),
                       step=da0 + n_total,
                       new_step_size=step_val)
  return res
