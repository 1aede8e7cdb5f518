"""Streaming moments of MCMC draws, float64 NumPy restatement.

`experimental/stats/sample_stats.py:167-235` (RunningCovariance.update: the
one-pass chunk merge of Pebay 2008, with `_batch_outer_product` over the
trailing `event_ndims` dimensions, `:281-312`), `:237-250` (covariance =
ssr / (n - ddof)), `:483-520` (RunningMean.update), `:847-883`
(RunningPotentialScaleReduction.potential_scale_reduction, Brooks and Gelman
1998 with unbiased between-chain variance).

Pinned by closed forms: streaming over a sequence (one sample or a chunk at a
time, in any split) equals np.mean / np.cov of the whole sequence; the
streaming R-hat equals the batch non-split R-hat of `oracle.diagnostic`;
and the reference tests' known answers (means [2.5, 3.5] and [2.5]*3, zero
covariance of zeros).

Test infrastructure only.
"""

import numpy as np


class Moments:
  """num_samples, mean [B..., E...], ssr (None | [B..., E...] | [B..., E..., E...])."""

  def __init__(self, shape, event_ndims, mode):
    self.num_samples = 0.0
    self.mean = np.zeros(shape, np.float64)
    self.event_ndims = event_ndims
    self.mode = mode  # 0 mean, 1 per-element variance, 2 cross products
    ev = tuple(shape[len(shape) - event_ndims:]) if event_ndims else ()
    if mode == 0:
      self.ssr = None
    elif mode == 1:
      self.ssr = np.zeros(shape, np.float64)
    else:
      self.ssr = np.zeros(tuple(shape) + ev, np.float64)


def _outer(d, event_ndims):
  """`_batch_outer_product` (`sample_stats.py:281-312`) for event_ndims k."""
  if event_ndims == 0:
    return d * d
  b = d.shape[:d.ndim - event_ndims]
  e = d.shape[d.ndim - event_ndims:]
  flat = d.reshape(b + (-1,))
  return np.einsum('...i,...j->...ij', flat, flat).reshape(b + e + e)


def update(m, sample, axis=None):
  """One RunningCovariance / RunningMean update (`sample_stats.py:167-235`)."""
  x = np.asarray(sample, np.float64)
  if axis is not None:
    cn = float(x.shape[axis])
    cm = x.mean(axis=axis)
    if m.ssr is not None:
      dev = x - np.expand_dims(cm, axis)
      dev = np.moveaxis(dev, axis, 0)
      if m.mode == 2:
        cs = sum(_outer(dev[r], m.event_ndims) for r in range(dev.shape[0]))
      else:
        cs = (dev * dev).sum(0)
  else:
    cn = 1.0
    cm = x
    cs = 0.0
  new_n = m.num_samples + cn
  delta = cm - m.mean
  if m.ssr is not None:
    adj = m.num_samples * cn / new_n
    pair = _outer(delta, m.event_ndims) if m.mode == 2 else delta * delta
    m.ssr = m.ssr + cs + adj * pair
  m.mean = m.mean + cn * delta / new_n
  m.num_samples = new_n
  return m


def covariance(m, ddof=0):
  """`sample_stats.py:237-250`."""
  return m.ssr / (m.num_samples - ddof)


def potential_scale_reduction(m, independent_chain_ndims):
  """Streaming R-hat from per-element running variances
  (`sample_stats.py:847-883`)."""
  n = m.num_samples
  k = independent_chain_ndims
  axes = tuple(range(k))
  shape = m.mean.shape
  chains = float(np.prod(shape[:k]))
  b_div_n = m.mean.var(axis=axes, ddof=1)
  w = m.ssr.sum(axis=axes) / (chains * (n - 1))
  sigma2_plus = ((n - 1) / n) * w + b_div_n
  return ((chains + 1.0) / chains) * sigma2_plus / w - (n - 1.0) / (chains * n)
