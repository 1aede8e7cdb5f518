"""ESS and (split) R-hat, float64 NumPy restatement.

`mcmc/diagnostic.py:203-336` (cross-chain ESS, Vehtari et al. 2021 eq. 10,
filter_threshold=0 truncation), `:476-568` (potential_scale_reduction with
split_chains), `stats/sample_stats.py:44-216` (FFT auto-covariance, centred,
zero-padded to 2^ceil(log2(2n)), divided by n - k).

Pinned by closed forms: iid draws give ESS ~ C*N, AR(1) with coefficient phi
gives ESS ~ C*N*(1-phi)/(1+phi), identical chains give R-hat ~ 1.

Test infrastructure only.
"""

import numpy as np


def auto_covariance(x, axis=0):
  """Unnormalised auto-covariance along `axis` (`sample_stats.py:120-216`)."""
  x = np.moveaxis(np.asarray(x, np.float64), axis, -1)
  x = x - x.mean(-1, keepdims=True)
  n = x.shape[-1]
  m = int(2 ** np.ceil(np.log2(2 * n)))
  f = np.fft.rfft(x, n=m, axis=-1)
  ac = np.fft.irfft(f * np.conj(f), n=m, axis=-1)[..., :n]
  ac = ac / (n - np.arange(n))
  return np.moveaxis(ac, -1, axis)


def effective_sample_size(states, cross_chain_dims=1, filter_threshold=0.0):
  """states [N, C, ...] -> ESS [...] (cross-chain when cross_chain_dims=1)."""
  states = np.asarray(states, np.float64)
  if cross_chain_dims is not None and not isinstance(cross_chain_dims, int):
    # several chain axes: the same statistic over their flattened product
    dims = [int(d) for d in cross_chain_dims]
    rest = [d for d in range(1, states.ndim) if d not in dims]
    x = np.transpose(states, [0] + dims + rest)
    states = x.reshape((x.shape[0], -1) + tuple(states.shape[d] for d in rest))
  n = states.shape[0]
  auto_cov = auto_covariance(states, axis=0)
  if cross_chain_dims is not None:
    c = states.shape[1]
    chain_means = states.mean(0)
    b_div_n = chain_means.var(axis=0, ddof=1)
    w_biased = auto_cov[0].mean(axis=0)
    approx_var = w_biased + b_div_n
    mean_auto_cov = auto_cov.mean(axis=1)
    auto_corr = 1.0 - (w_biased - mean_auto_cov) / approx_var
  else:
    c = 1
    auto_corr = auto_cov / auto_cov[:1]
  k = np.arange(auto_corr.shape[0], dtype=np.float64)
  nk = ((n - k) / n).reshape((-1,) + (1,) * (auto_corr.ndim - 1))
  weighted = nk * auto_corr
  if filter_threshold is not None:
    mask = (auto_corr < filter_threshold).astype(np.float64)
    mask = np.maximum(1.0 - np.cumsum(mask, axis=0), 0.0)
    weighted = weighted * mask
  return c * n / (-1.0 + 2.0 * weighted.sum(0))


def potential_scale_reduction(states, independent_chain_ndims=1,
                              split_chains=True):
  """states [N, C, ...] -> R-hat [...]."""
  x = np.asarray(states, np.float64)
  icn = independent_chain_ndims
  if split_chains:
    n = x.shape[0]
    x = x[:n - n % 2]
    x = x.reshape((2, n // 2) + x.shape[1:])
    x = np.swapaxes(x, 0, 1)
    icn += 1
  n = x.shape[0]
  chain_axes = tuple(range(1, 1 + icn))
  sample_chain_axes = tuple(range(0, 1 + icn))
  m = int(np.prod([x.shape[a] for a in chain_axes]))
  means = x.mean(0, keepdims=True)
  b_div_n = means.var(axis=sample_chain_axes, ddof=1)
  w = x.var(axis=0, ddof=1, keepdims=True).mean(axis=sample_chain_axes)
  sigma2_plus = ((n - 1) / n) * w + b_div_n
  return ((m + 1.0) / m) * sigma2_plus / w - (n - 1.0) / (m * n)
