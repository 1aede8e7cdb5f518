"""Streaming statistics on the device (`experimental/stats/sample_stats.py`).

`RunningMean`, `RunningVariance`, `RunningCovariance` and
`RunningPotentialScaleReduction` with the reference's interface
(`from_shape`, `from_example`, `update(new_sample, axis=None)`, and
`covariance` / `variance` / `potential_scale_reduction`).  Every update is
one `hipmc_moments_update` launch (Pebay's one-pass chunk merge,
`sample_stats.py:167-235`) on the current stream; the sums are held in
float64 on the GPU whatever the sample dtype (the reference accumulates in
the sample dtype), and the finalizers return the sample dtype.  Updates
return a new object whose sums the kernel writes into fresh buffers while
reading the old ones (no copy), so the old object stays valid.
"""

import math

import torch

from probability_amd import _lib

__all__ = ['RunningCovariance', 'RunningMean', 'RunningPotentialScaleReduction',
           'RunningVariance']

_MODE_MEAN, _MODE_VAR, _MODE_COV = 0, 1, 2


def _float_dtype_like(dtype):
  """`sample_stats.py:314-320`: int64 samples give float64 results, other
  integers float32."""
  if dtype in (torch.float32, torch.float64):
    return dtype
  if dtype == torch.int64:
    return torch.float64
  if not dtype.is_floating_point:
    return torch.float32
  raise TypeError(f'unsupported sample dtype {dtype}')


def _device():
  _lib.require_cuda()
  return torch.device('cuda', torch.cuda.current_device())


def _as_sample(x, dtype, device):
  x = torch.as_tensor(x, device=device)
  if x.dtype != dtype:
    x = x.to(dtype)
  return x.contiguous()


def _moments_update(mode, mean, ssr, num_samples, sample, event_ndims,
                    axis):
  """Launches one update reading (mean, ssr) and writing fresh buffers;
  returns (chunk size, new mean, new ssr)."""
  shape = tuple(mean.shape)
  if axis is not None:
    if sample.dim() == 0:
      raise ValueError('a chunked update needs a sample with the chunk axis')
    axis = axis % sample.dim()
    chunk_n = sample.shape[axis]
    sample = torch.movedim(sample, axis, 0).contiguous()
    body = tuple(sample.shape[1:])
  else:
    chunk_n = 1
    body = tuple(sample.shape)
  if body != shape:
    raise ValueError(f'sample shape {body} does not match the running '
                     f'statistic\'s shape {shape}')
  if chunk_n < 1:
    return 0, mean, ssr
  if mode == _MODE_COV:  # [n_batch, ev] samples, [n_batch, ev, ev] sums
    ev = math.prod(shape[len(shape) - event_ndims:])
    n_batch = math.prod(shape[:len(shape) - event_ndims])
    ssr_shape = shape + shape[len(shape) - event_ndims:]
  else:  # element-wise
    ev, n_batch = math.prod(shape), 1
    ssr_shape = shape
  # the kernels read float64, contiguous sums on the sample's device
  dev = sample.device
  mean = mean.to(device=dev, dtype=torch.float64).contiguous()
  if ssr is not None:
    ssr = ssr.to(device=dev, dtype=torch.float64).contiguous()
    if tuple(ssr.shape) != tuple(ssr_shape):
      raise ValueError(f'sum of squared residuals has shape '
                       f'{tuple(ssr.shape)}, expected {tuple(ssr_shape)}')
  if ev == 0 or n_batch == 0:
    return chunk_n, mean.clone(), None if ssr is None else ssr.clone()
  new_mean = torch.empty_like(mean)
  new_ssr = None if ssr is None else torch.empty_like(ssr)
  _lib.check(_lib.lib().hipmc_moments_update(
      _lib.dtype_code(sample.dtype), mode, n_batch, ev, chunk_n,
      _lib.ptr(sample), float(num_samples), _lib.ptr(mean), _lib.ptr(ssr),
      _lib.ptr(new_mean), _lib.ptr(new_ssr),
      _lib.stream_handle(mean.device)))
  return chunk_n, new_mean, new_ssr


class RunningMean:
  """`sample_stats.py:414-527`."""

  def __init__(self, num_samples, mean, dtype=None):
    self.num_samples = float(num_samples)
    self._mean = mean
    self._dtype = dtype or torch.float32

  @classmethod
  def from_shape(cls, shape, dtype=torch.float32):
    dtype = _float_dtype_like(dtype)
    return cls(0., torch.zeros(tuple(shape), dtype=torch.float64,
                               device=_device()), dtype)

  @classmethod
  def from_example(cls, example):
    example = torch.as_tensor(example)
    return cls.from_shape(tuple(example.shape), example.dtype)

  @property
  def mean(self):
    return self._mean.to(self._dtype)

  def update(self, new_sample, axis=None):
    """`sample_stats.py:483-520`."""
    x = _as_sample(new_sample, self._dtype, self._mean.device)
    cn, mean, _ = _moments_update(_MODE_MEAN, self._mean, None,
                                  self.num_samples, x, 0, axis)
    return RunningMean(self.num_samples + cn, mean, self._dtype)

  def __repr__(self):
    return (f'RunningMean(num_samples={self.num_samples!r}, '
            f'mean={self.mean!r})')


def _default_event_ndims(event_ndims, rank):
  """`sample_stats.py:260-278`."""
  if event_ndims is None:
    event_ndims = rank
  if event_ndims > rank:
    raise ValueError(f'Cannot calculate cross-products in {event_ndims} '
                     f'dimensions for samples of rank {rank}')
  if event_ndims < 0:
    raise ValueError('event_ndims must be >= 0')
  return event_ndims


class RunningCovariance:
  """`sample_stats.py:54-257`: covariance over the trailing `event_ndims`
  dimensions (None: all of them), batched over the leading ones."""

  def __init__(self, num_samples, mean, sum_squared_residuals, event_ndims,
               dtype=None):
    self.num_samples = float(num_samples)
    self._mean = mean
    self._ssr = sum_squared_residuals
    self.event_ndims = event_ndims
    self._dtype = dtype or torch.float32

  def _mode(self):
    return _MODE_COV if self.event_ndims > 0 else _MODE_VAR

  @classmethod
  def from_shape(cls, shape=(), dtype=torch.float32, event_ndims=None):
    shape = tuple(shape)
    dtype = _float_dtype_like(dtype)
    event_ndims = _default_event_ndims(event_ndims, len(shape))
    dev = _device()
    ev = shape[len(shape) - event_ndims:] if event_ndims else ()
    mean = torch.zeros(shape, dtype=torch.float64, device=dev)
    ssr = torch.zeros(shape + ev, dtype=torch.float64, device=dev)
    return cls(0., mean, ssr, event_ndims, dtype)

  @classmethod
  def from_example(cls, example, event_ndims=None):
    example = torch.as_tensor(example)
    return cls.from_shape(tuple(example.shape), example.dtype, event_ndims)

  @property
  def mean(self):
    return self._mean.to(self._dtype)

  @property
  def sum_squared_residuals(self):
    return self._ssr.to(self._dtype)

  def update(self, new_sample, axis=None):
    """`sample_stats.py:167-235`."""
    x = _as_sample(new_sample, self._dtype, self._mean.device)
    cn, mean, ssr = _moments_update(self._mode(), self._mean, self._ssr,
                                    self.num_samples, x, self.event_ndims,
                                    axis)
    return type(self)(self.num_samples + cn, mean, ssr, self.event_ndims,
                      self._dtype)

  def covariance(self, ddof=0):
    """`sample_stats.py:237-250`: ssr / (num_samples - ddof)."""
    return (self._ssr / (self.num_samples - ddof)).to(self._dtype)

  def __repr__(self):
    return (f'{type(self).__name__}(num_samples={self.num_samples!r}, '
            f'mean={self.mean!r}, event_ndims={self.event_ndims!r})')


class RunningVariance(RunningCovariance):
  """`sample_stats.py:323-411`: per-element variance (event_ndims 0)."""

  def __init__(self, num_samples, mean, sum_squared_residuals,
               event_ndims=0, dtype=None):
    super().__init__(num_samples, mean, sum_squared_residuals, 0, dtype)

  @classmethod
  def from_shape(cls, shape=(), dtype=torch.float32):
    return super().from_shape(shape, dtype, event_ndims=0)

  @classmethod
  def from_example(cls, example):
    example = torch.as_tensor(example)
    return cls.from_shape(tuple(example.shape), example.dtype)

  @classmethod
  def from_stats(cls, num_samples, mean, variance):
    """`sample_stats.py:383-404`."""
    mean = torch.as_tensor(mean, device=_device())
    dtype = _float_dtype_like(mean.dtype)
    variance = torch.as_tensor(variance, device=mean.device)
    m64 = mean.to(torch.float64).contiguous()
    ssr = (variance.to(torch.float64) * float(num_samples)).contiguous()
    ssr = torch.broadcast_to(ssr, m64.shape).contiguous()
    return cls(float(num_samples), m64, ssr, 0, dtype)

  def variance(self, ddof=0):
    return self.covariance(ddof)


class RunningPotentialScaleReduction:
  """`sample_stats.py:714-890`: streaming R-hat (Brooks and Gelman 1998,
  no chain splitting) from per-element running variances; the leading
  `independent_chain_ndims` dimensions index chains."""

  def __init__(self, chain_variances, independent_chain_ndims):
    self.chain_variances = chain_variances
    self.independent_chain_ndims = independent_chain_ndims

  @classmethod
  def from_shape(cls, shape=(), independent_chain_ndims=1,
                 dtype=torch.float32):
    shape = tuple(shape)
    if independent_chain_ndims < 1 or independent_chain_ndims > len(shape):
      raise ValueError('independent_chain_ndims must be in [1, rank]')
    return cls(RunningVariance.from_shape(shape, dtype),
               independent_chain_ndims)

  @classmethod
  def from_example(cls, example, independent_chain_ndims=1):
    example = torch.as_tensor(example)
    return cls.from_shape(tuple(example.shape), independent_chain_ndims,
                          example.dtype)

  def update(self, new_sample):
    return type(self)(self.chain_variances.update(new_sample),
                      self.independent_chain_ndims)

  def potential_scale_reduction(self):
    """`sample_stats.py:847-883`, in float64."""
    cv = self.chain_variances
    n = cv.num_samples
    k = self.independent_chain_ndims
    dims = tuple(range(k))
    m = float(math.prod(cv._mean.shape[:k]))
    b_div_n = cv._mean.var(dim=dims, unbiased=True)
    w = cv._ssr.sum(dim=dims) / (m * (n - 1))
    sigma2_plus = ((n - 1) / n) * w + b_div_n
    rhat = ((m + 1.) / m) * sigma2_plus / w - (n - 1.) / (m * n)
    return rhat.to(cv._dtype)
