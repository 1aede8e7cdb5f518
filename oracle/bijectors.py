"""Softplus / Exp / Identity bijectors for the TTK path (float64 oracle).

Restates `bijectors/softplus.py:36-50` (forward with the log1p(exp(x)) branch
below -20 in float64 / -9 in float32), `:139-172` (forward_log_det_jacobian
= -softplus(-x)), `math/generic.py:531-584` (softplus_inverse), and
`bijectors/power_transform.py:85-112` with power=0 (Exp: forward exp(x),
fldj x).  Test infrastructure only.
"""

import numpy as np

IDENTITY, SOFTPLUS, EXP = 0, 1, 2


def softplus(x):
  x = np.asarray(x)
  return np.log1p(np.exp(-np.abs(x))) + np.maximum(x, 0)


def sigmoid(x):
  x = np.asarray(x)
  e = np.exp(-np.abs(x))
  return np.where(x >= 0, 1.0 / (1.0 + e), e / (1.0 + e))


def softplus_forward(x):
  x = np.asarray(x)
  cutoff = -9.0 if x.dtype == np.float32 else -20.0
  return np.where(x < cutoff, np.log1p(np.exp(x)), softplus(x))


def softplus_inverse(y):
  """`math/generic.py:571-584`."""
  y = np.asarray(y)
  eps = np.finfo(y.dtype if y.dtype.kind == 'f' else np.float64).eps
  threshold = np.log(eps) + 2.0
  too_small = y < np.exp(threshold)
  too_large = y > -threshold
  safe = np.where(too_small | too_large, np.ones_like(y), y)
  mid = safe + np.log(-np.expm1(-safe))
  with np.errstate(divide='ignore', invalid='ignore'):
    small_val = np.log(y)
  return np.where(too_small, small_val, np.where(too_large, y, mid))


def forward(code, x):
  if code == IDENTITY:
    return np.array(x, copy=True)
  if code == SOFTPLUS:
    return softplus_forward(x)
  if code == EXP:
    return np.exp(x)
  raise ValueError(code)


def inverse(code, y):
  if code == IDENTITY:
    return np.array(y, copy=True)
  if code == SOFTPLUS:
    return softplus_inverse(y)
  if code == EXP:
    return np.log(y)
  raise ValueError(code)


def fldj(code, x):
  """forward_log_det_jacobian, elementwise (event_ndims summed by caller)."""
  if code == IDENTITY:
    return np.zeros_like(x)
  if code == SOFTPLUS:
    return -softplus(-x)
  if code == EXP:
    return np.array(x, copy=True)
  raise ValueError(code)


def forward_grad(code, x):
  """d forward / dx."""
  if code == IDENTITY:
    return np.ones_like(x)
  if code == SOFTPLUS:
    return sigmoid(x)
  if code == EXP:
    return np.exp(x)
  raise ValueError(code)


def fldj_grad(code, x):
  """d fldj / dx."""
  if code == IDENTITY:
    return np.zeros_like(x)
  if code == SOFTPLUS:
    return sigmoid(-x)
  if code == EXP:
    return np.ones_like(x)
  raise ValueError(code)
