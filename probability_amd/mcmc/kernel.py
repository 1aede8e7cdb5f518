"""TransitionKernel protocol (`mcmc/kernel.py:26-164`)."""

import abc


class TransitionKernel(abc.ABC):
  """Base class: `one_step`, `bootstrap_results`, `is_calibrated`, `copy`."""

  @abc.abstractmethod
  def one_step(self, current_state, previous_kernel_results, seed=None):
    raise NotImplementedError

  @abc.abstractmethod
  def bootstrap_results(self, init_state):
    raise NotImplementedError

  @property
  def is_calibrated(self):
    return False

  @property
  def parameters(self):
    return self._parameters

  def copy(self, **override_parameter_kwargs):
    params = dict(self.parameters)
    params.update(override_parameter_kwargs)
    return type(self)(**params)

  @property
  def experimental_shard_axis_names(self):
    return self.parameters.get('experimental_shard_axis_names')

  def experimental_with_shard_axes(self, shard_axis_names):
    if shard_axis_names is not None:
      raise NotImplementedError('state sharding across devices is not built')
    return self.copy()
