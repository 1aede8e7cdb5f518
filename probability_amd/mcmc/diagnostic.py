"""ESS and potential scale reduction on the device
(`mcmc/diagnostic.py:38-568`, `stats/sample_stats.py:44-216`).

Same formulas as the reference: centred FFT auto-covariance zero-padded to
2^ceil(log2(2n)) divided by (n - k); cross-chain ESS (Vehtari et al. 2021
eq. 10) truncated at the first negative auto-correlation; split R-hat.
Computed in float64 on the GPU with torch.fft (these run once per sampling
run, on the traced draws).
"""

import math

import torch

__all__ = ['effective_sample_size', 'potential_scale_reduction']


def _as64(x):
  if isinstance(x, torch.Tensor):
    return x.to(torch.float64)
  return torch.as_tensor(x, dtype=torch.float64)


def auto_covariance(x):
  """x [n, ...] -> auto-covariance along axis 0 (unnormalised)."""
  x = _as64(x)
  n = x.shape[0]
  x = x - x.mean(0, keepdim=True)
  m = int(2 ** math.ceil(math.log2(2 * n)))
  f = torch.fft.rfft(x, n=m, dim=0)
  ac = torch.fft.irfft(f * f.conj(), n=m, dim=0)[:n]
  k = torch.arange(n, dtype=torch.float64, device=x.device)
  return ac / (n - k).reshape((-1,) + (1,) * (x.dim() - 1))


def _flatten_chain_dims(states, cross_chain_dims):
  """Moves the chain dimensions (an int or a sequence of ints, counted in
  `states`) to axis 1 as one flattened chain axis."""
  dims = ([cross_chain_dims] if isinstance(cross_chain_dims, int)
          else [int(d) for d in cross_chain_dims])
  dims = [d % states.dim() for d in dims]
  if 0 in dims:
    raise ValueError('cross_chain_dims cannot include the sample axis 0')
  rest = [d for d in range(1, states.dim()) if d not in dims]
  x = states.permute([0] + dims + rest)
  return x.reshape((x.shape[0], -1) + tuple(states.shape[d] for d in rest))


def _ess_one(states, cross_chain_dims, filter_threshold):
  states = _as64(states)
  if cross_chain_dims is not None:
    states = _flatten_chain_dims(states, cross_chain_dims)
  n = states.shape[0]
  ac = auto_covariance(states)
  if cross_chain_dims is not None:
    c = states.shape[1]
    if c < 2:
      raise ValueError('When `cross_chain_dims` is not `None`, there must be '
                       '> 1 chain in `states`.')
    b_div_n = states.mean(0).var(dim=0, unbiased=True)
    w = ac[0].mean(0)
    approx = w + b_div_n
    rho = 1.0 - (w - ac.mean(1)) / approx
  else:
    c = 1
    rho = ac / ac[:1]
  k = torch.arange(rho.shape[0], dtype=torch.float64, device=rho.device)
  nk = ((n - k) / n).reshape((-1,) + (1,) * (rho.dim() - 1))
  wr = nk * rho
  if filter_threshold is not None:
    mask = (rho < filter_threshold).to(torch.float64)
    mask = torch.clamp(1.0 - torch.cumsum(mask, 0), min=0.0)
    wr = wr * mask
  return c * n / (-1.0 + 2.0 * wr.sum(0))


def effective_sample_size(states, filter_threshold=0., filter_beyond_lag=None,
                          filter_beyond_positive_pairs=False,
                          cross_chain_dims=None, validate_args=False,
                          name=None):
  """ESS of each state part ([n, chains, ...] -> [...])."""
  if filter_beyond_lag is not None or filter_beyond_positive_pairs:
    raise NotImplementedError('only the default truncation rule is built')
  if isinstance(states, (list, tuple)):
    cdims = (cross_chain_dims if isinstance(cross_chain_dims, (list, tuple))
             else [cross_chain_dims] * len(states))
    return [_ess_one(s, cd, filter_threshold) for s, cd in zip(states, cdims)]
  return _ess_one(states, cross_chain_dims, filter_threshold)


def _rhat_one(state, independent_chain_ndims, split_chains):
  x = _as64(state)
  icn = independent_chain_ndims
  if icn < 1:
    raise ValueError('Argument `independent_chain_ndims` must be `>= 1`, '
                     'found: {}'.format(icn))
  if not split_chains and x.shape[0] < 2:
    raise ValueError('Must provide at least 2 samples. Found {}'.format(
        x.shape[0]))
  if split_chains:
    n = x.shape[0]
    if n < 4:
      raise ValueError('Must provide at least 4 samples when splitting '
                       'chains. Found {}'.format(n))
    x = x[:n - n % 2]
    x = x.reshape((2, n // 2) + tuple(x.shape[1:])).transpose(0, 1)
    icn += 1
  n = x.shape[0]
  chain_axes = tuple(range(1, 1 + icn))
  all_axes = tuple(range(0, 1 + icn))
  m = 1
  for a in chain_axes:
    m *= x.shape[a]
  means = x.mean(0, keepdim=True)
  b_div_n = means.var(dim=all_axes, unbiased=True)
  w = x.var(dim=0, unbiased=True, keepdim=True).mean(dim=all_axes)
  s2 = ((n - 1) / n) * w + b_div_n
  return ((m + 1.0) / m) * s2 / w - (n - 1.0) / (m * n)


def potential_scale_reduction(chains_states, independent_chain_ndims=1,
                              split_chains=False, validate_args=False,
                              name=None):
  """R-hat of each state part ([n, chains, ...] -> [...])."""
  if isinstance(chains_states, (list, tuple)):
    return [_rhat_one(s, independent_chain_ndims, split_chains)
            for s in chains_states]
  return _rhat_one(chains_states, independent_chain_ndims, split_chains)
