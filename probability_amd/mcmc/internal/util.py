"""Host-side helpers mirroring `mcmc/internal/util.py` and
`internal/unnest.py` of the reference (nested-results access, list
handling).  No numerics happen here."""

import collections

import numpy as np
import torch

_NESTED = ('inner_results', 'accepted_results')


class PrettyNamedTupleMixin:
  """Readable repr for kernel-results namedtuples."""
  __slots__ = ()

  def __repr__(self):
    body = ',\n'.join(f'  {k}={v!r}'.replace('\n', '\n  ')
                      for k, v in self._asdict().items())
    return f'{type(self).__name__}(\n{body}\n)'


def is_list_like(x):
  return isinstance(x, (tuple, list))


def is_namedtuple_like(x):
  return hasattr(x, '_fields') and hasattr(x, '_replace')


def _nested(obj):
  return [(a, getattr(obj, a)) for a in _NESTED if hasattr(obj, a)]


def has_nested(obj, attr):
  if hasattr(obj, attr):
    return True
  return any(has_nested(v, attr) for _, v in _nested(obj))


def get_innermost(obj, attr, default=None, _sentinel=object()):
  """`internal/unnest.py:147-181`: innermost object's `attr`."""
  for _, v in _nested(obj):
    if has_nested(v, attr):
      return get_innermost(v, attr)
  if hasattr(obj, attr):
    return getattr(obj, attr)
  if default is not None:
    return default
  raise AttributeError(f'no nested attribute {attr!r}')


def get_outermost(obj, attr):
  if hasattr(obj, attr):
    return getattr(obj, attr)
  for _, v in _nested(obj):
    if has_nested(v, attr):
      return get_outermost(v, attr)
  raise AttributeError(f'no nested attribute {attr!r}')


def replace_innermost(ntuple, **kw):
  """`internal/unnest.py:221-260`: replace the innermost occurrences."""
  updates = {}
  for a, v in _nested(ntuple):
    sub = {k: val for k, val in kw.items() if has_nested(v, k)}
    if sub:
      updates[a] = replace_innermost(v, **sub)
      for k in sub:
        kw.pop(k)
  for k in list(kw):
    if k in getattr(ntuple, '_fields', ()):
      updates[k] = kw.pop(k)
  if kw:
    raise ValueError(f'unable to replace {list(kw)}')
  return ntuple._replace(**updates)


def strip_seeds(obj):
  """`mcmc/internal/util.py:198-202`."""
  if not is_namedtuple_like(obj):
    return obj
  return type(obj)(**{k: (strip_seeds(v) if k != 'seed' else [])
                      for k, v in obj._asdict().items()})


def choose(is_accepted, proposed, current):
  """`mcmc/internal/util.py:103-165`: masked select over nested results.

  Device-side torch.where on already-computed tensors (left-justified
  broadcast of the chain mask).
  """
  if is_namedtuple_like(proposed):
    return type(proposed)(**{
        k: choose(is_accepted, getattr(proposed, k), getattr(current, k))
        for k in proposed._fields})
  if is_list_like(proposed):
    return type(proposed)(choose(is_accepted, p, c)
                          for p, c in zip(proposed, current))
  if not isinstance(proposed, torch.Tensor) or proposed is current:
    return proposed
  m = is_accepted
  while m.dim() < proposed.dim():
    m = m.unsqueeze(-1)
  return torch.where(m, proposed, current)


def enable_store_parameters_in_results(kernel):
  """`mcmc/internal/util.py:311-345`."""
  kernel_stack = []
  while hasattr(kernel, 'parameters') and 'inner_kernel' in kernel.parameters:
    kernel_stack.append(kernel)
    kernel = kernel.parameters['inner_kernel']
  def _recreate(k):
    return type(k)(**k.parameters)
  if (hasattr(kernel, 'parameters') and
      'store_parameters_in_results' in kernel.parameters):
    params = dict(kernel.parameters)
    params['store_parameters_in_results'] = True
    kernel = type(kernel)(**params)
  for outer in reversed(kernel_stack):
    params = dict(outer.parameters)
    params['inner_kernel'] = kernel
    kernel = type(outer)(**params)
  return kernel


def seed_array(seed):
  if seed is None:
    return None
  return np.asarray(seed, np.uint32).reshape(2)
