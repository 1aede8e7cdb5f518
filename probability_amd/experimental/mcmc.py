"""Streaming reductions over MCMC draws (`experimental/mcmc/`).

`WithReductions` (`with_reductions.py:41-190`) wraps a transition kernel and,
after each inner step, feeds the new state to a (possibly nested) structure
of `Reducer`s; the reducers keep running statistics on the GPU
(`probability_amd.experimental.stats`, one `hipmc_moments_update` launch per
reducer and state part per step), so a long run needs no `[R, C, D]` trace:

* `ExpectationsReducer` (`expectations_reducer.py:42-160`): running mean of
  `transform_fn(state, kernel_results)`.
* `CovarianceReducer` / `VarianceReducer` (`covariance_reducer.py:43-298`):
  running covariance over the trailing `event_ndims` dimensions, finalized
  with `ddof`.
* `PotentialScaleReductionReducer`
  (`potential_scale_reduction_reducer.py:36-150`): streaming R-hat with the
  leading `independent_chain_ndims` dimensions as chains.

`step_kernel` (`step.py:32-110`) runs a kernel for a number of steps and
`SampleDiscardingKernel` (`sample_discarding_kernel.py:41-170`) keeps one
draw per call, skipping burn-in and thinned draws.  `sample_fold`
(`sample_fold.py:35-170`) runs a kernel for `num_steps`
results, with burn-in and thinning as `SampleDiscardingKernel`
(`sample_discarding_kernel.py:81-130`: the first result skips
`num_burnin_steps + num_steps_between_results` draws, later ones
`num_steps_between_results`), and returns the finalized reductions.
"""

import collections

from probability_amd import random as samplers
from probability_amd.experimental import stats as running
from probability_amd.mcmc import kernel as kernel_base
from probability_amd.mcmc.internal import util as mcmc_util

__all__ = ['CovarianceReducer', 'ExpectationsReducer',
           'PotentialScaleReductionReducer', 'Reducer',
           'SampleDiscardingKernel', 'SampleDiscardingKernelResults',
           'VarianceReducer', 'WithReductions', 'WithReductionsKernelResults',
           'default_hmc_trace_fn', 'default_nuts_trace_fn', 'sample_fold',
           'step_kernel', 'windowed_adaptive_hmc', 'windowed_adaptive_nuts']


# ------------------------------------------------------------- structures

def _map(fn, structure, *rest):
  """map over lists, tuples and dicts of `structure`; `rest` either mirror
  it or are leaves broadcast to every position."""
  if isinstance(structure, dict):
    return {k: _map(fn, structure[k],
                    *[r[k] if isinstance(r, dict) else r for r in rest])
            for k in structure}
  if isinstance(structure, (list, tuple)) and \
      not mcmc_util.is_namedtuple_like(structure):
    out = [_map(fn, s, *[r[i] if isinstance(r, (list, tuple)) and
                         not mcmc_util.is_namedtuple_like(r) else r
                         for r in rest])
           for i, s in enumerate(structure)]
    return type(structure)(out) if isinstance(structure, tuple) else out
  return fn(structure, *rest)


def _get_sample(current_state, kernel_results):
  del kernel_results
  return current_state


# ---------------------------------------------------------------- reducers
class Reducer:
  """`reducer.py:29-100`: `initialize`, `one_step`, `finalize`."""

  def initialize(self, initial_chain_state, initial_kernel_results=None):
    raise NotImplementedError()

  def one_step(self, new_chain_state, current_reducer_state,
               previous_kernel_results=None, axis=None):
    raise NotImplementedError()

  def finalize(self, final_reducer_state):
    return final_reducer_state


ExpectationsReducerState = collections.namedtuple(
    'ExpectationsReducerState', 'expectation_state')


class ExpectationsReducer(Reducer):
  """Running mean of `transform_fn(state, kernel_results)` (a function or a
  structure of functions)."""

  def __init__(self, transform_fn=_get_sample, name=None):
    self.transform_fn = transform_fn
    self.name = name or 'expectations_reducer'

  def _apply(self, state, kr):
    return _map(lambda fn: fn(state, kr), self.transform_fn)

  def initialize(self, initial_chain_state, initial_kernel_results=None):
    example = self._apply(initial_chain_state, initial_kernel_results)
    return ExpectationsReducerState(
        _map(lambda e: running.RunningMean.from_example(e), example))

  def one_step(self, new_chain_state, current_reducer_state,
               previous_kernel_results=None, axis=None):
    values = self._apply(new_chain_state, previous_kernel_results)
    return ExpectationsReducerState(_map(
        lambda rm, v, ax: rm.update(v, axis=ax),
        current_reducer_state.expectation_state, values, axis))

  def finalize(self, final_reducer_state):
    return _map(lambda rm: rm.mean, final_reducer_state.expectation_state)


CovarianceReducerState = collections.namedtuple(
    'CovarianceReducerState', 'cov_state')


class CovarianceReducer(Reducer):
  """Running covariance of `transform_fn(state, kernel_results)` over the
  trailing `event_ndims` dimensions (None: all), `ddof` at finalize."""

  def __init__(self, event_ndims=None, transform_fn=_get_sample, ddof=0,
               name=None):
    self.event_ndims = event_ndims
    self.transform_fn = transform_fn
    self.ddof = ddof
    self.name = name or 'covariance_reducer'

  def _apply(self, state, kr):
    return _map(lambda fn: fn(state, kr), self.transform_fn)

  def initialize(self, initial_chain_state, initial_kernel_results=None):
    example = self._apply(initial_chain_state, initial_kernel_results)
    return CovarianceReducerState(_map(
        lambda e, k: running.RunningCovariance.from_example(e, k), example,
        self.event_ndims))

  def one_step(self, new_chain_state, current_reducer_state,
               previous_kernel_results=None, axis=None):
    values = self._apply(new_chain_state, previous_kernel_results)
    return CovarianceReducerState(_map(
        lambda rc, v, ax: rc.update(v, axis=ax),
        current_reducer_state.cov_state, values, axis))

  def finalize(self, final_reducer_state):
    return _map(lambda rc, d: rc.covariance(d), final_reducer_state.cov_state,
                self.ddof)


class VarianceReducer(CovarianceReducer):
  """`covariance_reducer.py:280-298`: per-element variance."""

  def __init__(self, transform_fn=_get_sample, ddof=0, name=None):
    super().__init__(event_ndims=0, transform_fn=transform_fn, ddof=ddof,
                     name=name or 'variance_reducer')


PotentialScaleReductionReducerState = collections.namedtuple(
    'PotentialScaleReductionReducerState', 'rhat_state')


class PotentialScaleReductionReducer(Reducer):
  """Streaming R-hat of every state part (no chain splitting)."""

  def __init__(self, independent_chain_ndims=1, name=None):
    if independent_chain_ndims < 1:
      raise ValueError('Must specify at least 1 dimension to represent '
                       'independent Markov chains.')
    self.independent_chain_ndims = independent_chain_ndims
    self.name = name or 'potential_scale_reduction_reducer'

  def initialize(self, initial_chain_state, initial_kernel_results=None):
    return PotentialScaleReductionReducerState(_map(
        lambda s: running.RunningPotentialScaleReduction.from_example(
            s, self.independent_chain_ndims), initial_chain_state))

  def one_step(self, new_chain_state, current_reducer_state,
               previous_kernel_results=None, axis=None):
    if axis is not None:
      raise NotImplementedError('PotentialScaleReductionReducer does not '
                                'take chunked samples')
    return PotentialScaleReductionReducerState(_map(
        lambda r, s: r.update(s), current_reducer_state.rhat_state,
        new_chain_state))

  def finalize(self, final_reducer_state):
    return _map(lambda r: r.potential_scale_reduction(),
                final_reducer_state.rhat_state)


# ---------------------------------------------------------- WithReductions
class WithReductionsKernelResults(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple('WithReductionsKernelResults',
                           'reduction_results, inner_results')):
  """`with_reductions.py:32-38`."""
  __slots__ = ()


def _reducer_map(fn, reducer, *rest):
  """Maps over the reducer structure only (a Reducer is a leaf)."""
  if isinstance(reducer, Reducer):
    return fn(reducer, *rest)
  if isinstance(reducer, dict):
    return {k: _reducer_map(fn, reducer[k], *[r[k] for r in rest])
            for k in reducer}
  out = [_reducer_map(fn, r, *[x[i] for x in rest])
         for i, r in enumerate(reducer)]
  return type(reducer)(out) if isinstance(reducer, tuple) else out


class WithReductions(kernel_base.TransitionKernel):
  """Runs `inner_kernel.one_step`, then every reducer's `one_step` on the
  new state (`with_reductions.py:84-124`)."""

  def __init__(self, inner_kernel, reducer, adjust_kr_fn=lambda x: x,
               name=None):
    self.inner_kernel = inner_kernel
    self.reducer = reducer
    self.adjust_kr_fn = adjust_kr_fn
    self.name = name or 'reduced_kernel'

  @property
  def is_calibrated(self):
    return self.inner_kernel.is_calibrated

  def one_step(self, current_state, previous_kernel_results, seed=None):
    new_state, inner_results = self.inner_kernel.one_step(
        current_state, previous_kernel_results.inner_results, seed=seed)
    adjusted = self.adjust_kr_fn(inner_results)
    reductions = _reducer_map(
        lambda r, s: r.one_step(new_state, s, previous_kernel_results=adjusted),
        self.reducer, previous_kernel_results.reduction_results)
    return new_state, WithReductionsKernelResults(reductions, inner_results)

  def bootstrap_results(self, init_state, inner_results=None,
                        previous_reducer_state=None):
    """`with_reductions.py:126-168`: the initial state is not a sample."""
    if inner_results is None:
      inner_results = self.inner_kernel.bootstrap_results(init_state)
    if previous_reducer_state is None:
      adjusted = self.adjust_kr_fn(inner_results)
      previous_reducer_state = _reducer_map(
          lambda r: r.initialize(init_state, adjusted), self.reducer)
    return WithReductionsKernelResults(previous_reducer_state, inner_results)


# ------------------------------------------------ step_kernel / discarding
def step_kernel(num_steps, current_state, previous_kernel_results=None,
                kernel=None, return_final_kernel_results=False,
                parallel_iterations=10, seed=None, name=None):
  """Runs `kernel` `num_steps` times; returns the last state (and the last
  kernel results).  Step k uses seed fold_in(sanitize_seed(seed), k); with
  seed None the kernel draws its own."""
  del parallel_iterations, name
  if kernel is None:
    raise ValueError('step_kernel needs a kernel')
  if previous_kernel_results is None:
    previous_kernel_results = kernel.bootstrap_results(current_state)
  base = (samplers.sanitize_seed(seed, salt='experimental.mcmc.step_kernel')
          if seed is not None else None)
  state, kr = current_state, previous_kernel_results
  for k in range(int(num_steps)):
    state, kr = kernel.one_step(
        state, kr, seed=None if base is None else samplers.fold_in(base, k))
  if return_final_kernel_results:
    return state, kr
  return state


class SampleDiscardingKernelResults(
    mcmc_util.PrettyNamedTupleMixin,
    collections.namedtuple('SampleDiscardingKernelResults',
                           'call_counter, inner_results')):
  """`sample_discarding_kernel.py:33-38`."""
  __slots__ = ()


class SampleDiscardingKernel(kernel_base.TransitionKernel):
  """Each call runs the inner kernel `num_steps_between_results + 1`
  times (plus `num_burnin_steps` on the first call) and returns the last
  state (`sample_discarding_kernel.py:81-130`)."""

  def __init__(self, inner_kernel, num_burnin_steps=0,
               num_steps_between_results=0, name=None):
    if int(num_burnin_steps) < 0 or int(num_steps_between_results) < 0:
      raise ValueError('num_burnin_steps and num_steps_between_results '
                       'must be >= 0')
    self.inner_kernel = inner_kernel
    self.num_burnin_steps = int(num_burnin_steps)
    self.num_steps_between_results = int(num_steps_between_results)
    self.name = name or 'sample_discarding_kernel'

  @property
  def is_calibrated(self):
    return self.inner_kernel.is_calibrated

  def _num_samples_to_skip(self, call_counter):
    return (self.num_steps_between_results +
            (self.num_burnin_steps if int(call_counter) == 0 else 0))

  def one_step(self, current_state, previous_kernel_results, seed=None):
    state, inner = step_kernel(
        self._num_samples_to_skip(previous_kernel_results.call_counter) + 1,
        current_state, previous_kernel_results.inner_results,
        kernel=self.inner_kernel, return_final_kernel_results=True,
        seed=seed)
    return state, SampleDiscardingKernelResults(
        int(previous_kernel_results.call_counter) + 1, inner)

  def bootstrap_results(self, init_state, inner_results=None):
    if inner_results is None:
      inner_results = self.inner_kernel.bootstrap_results(init_state)
    return SampleDiscardingKernelResults(0, inner_results)


# -------------------------------------------------------------- sample_fold
def sample_fold(num_steps, current_state, previous_kernel_results=None,
                kernel=None, reducer=None, previous_reducer_state=None,
                return_final_reducer_states=False, num_burnin_steps=0,
                num_steps_between_results=0, parallel_iterations=10,
                seed=None, name=None):
  """Runs `kernel` for `num_steps` kept draws and reduces them.

  Returns `(reduction_results, end_state, final_kernel_results)` and, with
  `return_final_reducer_states`, the reducer states last
  (`sample_fold.py:150-170`).  Step t of the inner kernel (burn-in and
  thinned steps included) uses seed fold_in(sanitize_seed(seed), t); with
  seed None the inner kernels draw their own.
  """
  del parallel_iterations, name
  if kernel is None:
    raise ValueError('sample_fold needs a kernel')
  num_steps = int(num_steps)
  num_burnin_steps = int(num_burnin_steps)
  num_steps_between_results = int(num_steps_between_results)
  if num_steps < 0 or num_burnin_steps < 0 or num_steps_between_results < 0:
    raise ValueError('num_steps, num_burnin_steps and '
                     'num_steps_between_results must be >= 0')
  if previous_kernel_results is None:
    previous_kernel_results = kernel.bootstrap_results(current_state)
  if reducer is None:
    reducer, reducing = [], False
  else:
    reducing = True
  wr = WithReductions(kernel, reducer)
  wr_results = wr.bootstrap_results(
      current_state, inner_results=previous_kernel_results,
      previous_reducer_state=previous_reducer_state)
  base = (samplers.sanitize_seed(seed, salt='experimental.mcmc.sample_fold')
          if seed is not None else None)
  t = 0

  def seed_for(t):
    return None if base is None else samplers.fold_in(base, t)

  state, inner = current_state, wr_results.inner_results
  reductions = wr_results.reduction_results
  for i in range(num_steps):
    skip = num_steps_between_results + (num_burnin_steps if i == 0 else 0)
    for _ in range(skip):
      state, inner = kernel.one_step(state, inner, seed=seed_for(t))
      t += 1
    state, wr_results = wr.one_step(
        state, WithReductionsKernelResults(reductions, inner),
        seed=seed_for(t))
    t += 1
    inner, reductions = wr_results.inner_results, wr_results.reduction_results
  finalized = (_reducer_map(lambda r, s: r.finalize(s), reducer, reductions)
               if reducing else None)
  if return_final_reducer_states:
    return finalized, state, inner, reductions
  return finalized, state, inner


# windowed adaptation (`experimental/mcmc/windowed_sampling.py`)
from probability_amd.experimental.windowed_sampling import (  # noqa: E402
    default_hmc_trace_fn, default_nuts_trace_fn, windowed_adaptive_hmc,
    windowed_adaptive_nuts)
