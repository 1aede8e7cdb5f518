"""Log-space reductions of `tfp.math` used by the step-size adaptations
(`math/generic.py:221-273` reduce_logmeanexp, `:387-432`
reduce_log_harmonic_mean_exp), on torch tensors.  The dual-averaging kernel
computes the same reductions in float64 on the device; passing one of these
functions as `reduce_fn` selects which."""

import math

import torch

__all__ = ['reduce_logmeanexp', 'reduce_log_harmonic_mean_exp']


def reduce_logmeanexp(input_tensor, axis=None, keepdims=False, **unused):
  """log(mean(exp(x))) over `axis` (all dimensions if None)."""
  x = torch.as_tensor(input_tensor)
  dims = tuple(range(x.ndim)) if axis is None else (
      (axis,) if isinstance(axis, int) else tuple(axis))
  if not dims:
    return x
  m = torch.amax(x, dim=dims, keepdim=True)
  m = torch.where(torch.isfinite(m), m, torch.zeros_like(m))
  lse = torch.log(torch.sum(torch.exp(x - m), dim=dims, keepdim=True)) + m
  n = 1
  for d in dims:
    n *= x.shape[d]
  out = lse - math.log(n)
  return out if keepdims else out.squeeze(dims) if dims else out


def reduce_log_harmonic_mean_exp(input_tensor, axis=None, keepdims=False,
                                 **unused):
  """log(1 / mean(1 / exp(x))) = -reduce_logmeanexp(-x)."""
  return -reduce_logmeanexp(-torch.as_tensor(input_tensor), axis=axis,
                            keepdims=keepdims)
