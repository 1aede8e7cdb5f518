"""Target specs: the `target_log_prob_fn` objects the GPU kernels dispatch on.

TFP kernels take any callable `target_log_prob_fn(*state_parts)` and
differentiate it with TF autodiff (`mcmc/internal/util.py:246-308`).  Here a
target is a spec object holding its data in device memory; it is callable
exactly like a TFP target (parts in, chain-shaped log-prob out) and its
log-density and analytic gradient run in libhipmc.  Formulas follow TFP's
distributions: Normal (`distributions/normal.py:182-188`), Bernoulli logits
(`bernoulli.py:119-135`), HalfCauchy (`half_cauchy.py:138-145`), HalfNormal
(`half_normal.py:146-153`).

State parts are packed into one [chains, dim] row per chain (parts
concatenated in order); `pack_state` / `unpack_state` convert.
"""

import math

import numpy as np
import torch

from probability_amd import _lib
from probability_amd import bijectors as bij_lib

_HALF_LOG_2PI = 0.5 * math.log(2.0 * math.pi)


def _to_device(x, dtype, device):
  if isinstance(x, torch.Tensor):
    return x.detach().to(device=device, dtype=dtype).contiguous()
  if hasattr(x, '__dlpack__') and not isinstance(x, np.ndarray):
    return torch.from_dlpack(x).to(device=device, dtype=dtype).contiguous()
  return torch.as_tensor(np.asarray(x), dtype=dtype, device=device).contiguous()


def _numel(shape):
  n = 1
  for s in shape:
    n *= int(s)
  return n


class TargetSpec:
  """Base class of the device-evaluated targets."""

  kind = None
  event_shapes = ()

  def __init__(self, dtype=torch.float32, device=None):
    if dtype not in (torch.float32, torch.float64):
      raise TypeError('dtype must be torch.float32 or torch.float64')
    self.dtype = dtype
    self.device = torch.device(device if device is not None else 'cuda')
    self._tables = {}

  # -- layout ---------------------------------------------------------------
  @property
  def part_sizes(self):
    return [_numel(s) for s in self.event_shapes]

  @property
  def dim(self):
    return int(sum(self.part_sizes))

  @property
  def base(self):
    return self

  @property
  def bijectors(self):
    return [bij_lib.Identity() for _ in self.event_shapes]

  def chain_shape(self, parts):
    p0, e0 = parts[0], self.event_shapes[0]
    return tuple(p0.shape[:p0.dim() - len(e0)])

  # -- device problem --------------------------------------------------------
  def _fill(self, p):
    raise NotImplementedError

  def problem(self, bijectors=None, n_chains=0, chain_offset=0, key=0):
    """hipmc_problem for this target.  `key` is the problem word of the
    Philox counter: chains with the same (seed, key, chain index) draw the
    same numbers whatever batch they run in."""
    p = _lib.Problem()
    p.kind = self.kind
    p.dim = self.dim
    p.n_chains = n_chains
    p.chain_offset = chain_offset
    p.key = key
    self._fill(p)
    ranges = []
    off = 0
    for size, b in zip(self.part_sizes,
                       bijectors or [None] * len(self.part_sizes)):
      b = bij_lib.as_bijector(b)
      if b.code != _lib.BIJ_IDENTITY:
        ranges.append((off, size, b.code))
      off += size
    if len(ranges) > _lib.MAX_BIJ_RANGES:
      raise NotImplementedError('at most 4 transformed state parts')
    p.n_bij = len(ranges)
    for i, (o, n, c) in enumerate(ranges):
      p.bij_off[i], p.bij_len[i], p.bij_code[i] = o, n, c
    return p

  def table(self, bijectors=None):
    """Device array holding this target's hipmc_problem (cached)."""
    key = tuple(bij_lib.as_bijector(b).code
                for b in (bijectors or [None] * len(self.event_shapes)))
    t = self._tables.get(key)
    if t is None:
      raw = bytes(self.problem(bijectors))
      t = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(self.device)
      self._tables[key] = t
    return t

  # -- evaluation -------------------------------------------------------------
  def pack_state(self, parts):
    return pack_state(parts, self.event_shapes, self.dtype, self.device)

  def unpack_state(self, flat, chain_shape):
    return unpack_state(flat, chain_shape, self.event_shapes)

  def log_prob_and_grad_flat(self, x, bijectors=None):
    """x [C, dim] -> (tlp [C], grad [C, dim]) on the device."""
    _lib.require_cuda()
    x = x.to(device=self.device, dtype=self.dtype).contiguous()
    c = x.shape[0]
    tlp = torch.empty(c, dtype=self.dtype, device=self.device)
    grad = torch.empty_like(x)
    _lib.check(_lib.lib().hipmc_target_eval(
        _lib.dtype_code(self.dtype), self.kind,
        _lib.ptr(self.table(bijectors)), None, c, x.shape[1], self.dim,
        _lib.ptr(x), _lib.ptr(tlp), _lib.ptr(grad),
        _lib.stream_handle(self.device)))
    return tlp, grad

  def __call__(self, *parts):
    parts = [_to_device(p, self.dtype, self.device) for p in parts]
    cs = self.chain_shape(parts)
    tlp, _ = self.log_prob_and_grad_flat(self.pack_state(parts),
                                         self.bijectors_for_eval())
    return tlp.reshape(cs)

  def bijectors_for_eval(self):
    return None


class TransformedTarget(TargetSpec):
  """tlp(f(x)) + sum_parts fldj(x)  (`mcmc/transformed_kernel.py:86-128`).

  The bijector composition is fused into the target kernels: the wrapped
  spec's data and kind are reused with per-part bijector codes.
  """

  def __init__(self, base, bijectors):
    self._base = base
    bijectors = list(bijectors)
    if len(bijectors) == 1 and len(base.event_shapes) > 1:
      raise ValueError('need one bijector per state part')
    if len(bijectors) != len(base.event_shapes):
      raise ValueError('need one bijector per state part')
    self._bijectors = [bij_lib.as_bijector(b) for b in bijectors]
    self.kind = base.kind
    self.event_shapes = base.event_shapes
    self.dtype = base.dtype
    self.device = base.device
    self._tables = base._tables

  @property
  def base(self):
    return self._base

  @property
  def bijectors(self):
    return list(self._bijectors)

  def bijectors_for_eval(self):
    return self._bijectors

  def _fill(self, p):
    self._base._fill(p)

  def table(self, bijectors=None):
    return self._base.table(bijectors if bijectors is not None
                            else self._bijectors)

  def log_prob_and_grad_flat(self, x, bijectors=None):
    return self._base.log_prob_and_grad_flat(
        x, bijectors if bijectors is not None else self._bijectors)


def pack_state(parts, event_shapes, dtype, device):
  """List of [chain..., event...] tensors -> [C, dim] contiguous tensor."""
  flat = []
  cs = None
  for p, es in zip(parts, event_shapes):
    p = _to_device(p, dtype, device)
    lead = tuple(p.shape[:p.dim() - len(es)])
    if cs is None:
      cs = lead
    elif lead != cs:
      raise ValueError(f'state parts disagree on chain shape: {lead} vs {cs}')
    flat.append(p.reshape(_numel(cs), _numel(es)))
  return torch.cat(flat, dim=1).contiguous()


def unpack_state(flat, chain_shape, event_shapes):
  out = []
  off = 0
  for es in event_shapes:
    n = _numel(es)
    out.append(flat[:, off:off + n].reshape(tuple(chain_shape) + tuple(es)))
    off += n
  return out


def _padded_rows(a, pitch_multiple=4):
  """(storage [r, pitch], view [r, c]) of a 2-D tensor whose rows start
  16-byte aligned: the pitch is c rounded up to `pitch_multiple` and the
  padding is zero."""
  r, c = a.shape
  pitch = -(-c // pitch_multiple) * pitch_multiple
  buf = torch.zeros((r, pitch), dtype=a.dtype, device=a.device)
  buf[:, :c] = a
  return buf, buf[:, :c]


# Largest X (bytes) that also keeps a transposed copy: small designs stay
# in L2, where a chain's leaf is bound by latency and instruction count
# rather than by bytes.
_XT_MAX_BYTES = 256 * 1024


class _GlmData:
  """X stored row-major with padded 16-byte rows.  Small designs of one
  coefficient block (D <= 64 in float32, D <= 32 in float64; X at most
  256 KiB) also keep X^T: the kernels then take the logits from X^T and the
  gradient from X, fewer instructions per row.  Every other design gets
  both from one pass over X, half the bytes."""

  def _set_design(self, x):
    if x.dim() != 2 or x.shape[0] < 1 or x.shape[1] < 1:
      raise ValueError('x must be a non-empty [N, D] matrix')
    self._xbuf, self.x = _padded_rows(x)
    self._xtbuf = self.xt = None
    if (x.shape[1] <= 16 * (16 // x.element_size()) and
        x.numel() * x.element_size() <= _XT_MAX_BYTES):
      self._xtbuf, self.xt = _padded_rows(x.t())

  def _fill_design(self, p):
    p.x, p.x_pitch = self._xbuf.data_ptr(), self._xbuf.shape[1]
    if self._xtbuf is None:
      p.xt, p.xt_pitch = None, 0
    else:
      p.xt, p.xt_pitch = self._xtbuf.data_ptr(), self._xtbuf.shape[1]


class LogisticRegression(_GlmData, TargetSpec):
  """Bayesian logistic regression.

  beta ~ N(0, prior_scale^2 I_D); y_n ~ Bernoulli(logits = x_n . beta).
  State: one part, beta [chains..., D].
  """

  kind = _lib.TARGET_LOGISTIC

  def __init__(self, x, y, prior_scale=1.0, dtype=torch.float32, device=None):
    super().__init__(dtype, device)
    self._set_design(_to_device(x, dtype, self.device))
    self.y = _to_device(y, dtype, self.device).reshape(-1)
    self.n_obs, d = self.x.shape
    if self.y.shape[0] != self.n_obs:
      raise ValueError('x and y disagree on the number of observations')
    self.prior_scale = float(prior_scale)
    self.event_shapes = [(d,)]

  def _fill(self, p):
    d = self.dim
    p.n_obs = self.n_obs
    self._fill_design(p)
    p.y = self.y.data_ptr()
    p.param[0] = self.prior_scale
    p.log_norm = d * (-_HALF_LOG_2PI - math.log(self.prior_scale))


class LogisticRegressionBF16(LogisticRegression):
  """Logistic regression with X stored in BF16 (large-batch HMC path).

  X is rounded to BF16 once; the model is exactly the logistic regression of
  the rounded X.  HamiltonianMonteCarlo on float32 chains with dim 256,
  chains a multiple of 64 and n_obs a multiple of 32 runs the all-chain
  gradient as fused MFMA GEMMs (hipmc_glm_hmc_run); other kernels use the
  float32 copy of the rounded X.
  """

  mfma_dim = 256

  def __init__(self, x, y, prior_scale=1.0, device=None):
    dev = torch.device(device if device is not None else 'cuda')
    xb = _to_device(x, torch.float32, dev).to(torch.bfloat16)
    super().__init__(xb.float(), y, prior_scale, dtype=torch.float32,
                     device=dev)
    self.x_bf16 = xb.contiguous()
    self.xt_bf16 = xb.t().contiguous()

  def mfma_ok(self, n_chains):
    return (self.dim == self.mfma_dim and n_chains % 64 == 0 and
            self.n_obs % 32 == 0 and self.dtype == torch.float32)

  def log_prob_and_grad_flat(self, x, bijectors=None):
    if bijectors is None and self.mfma_ok(int(x.shape[0])):
      from probability_amd.mcmc import glm_mfma
      x = x.to(device=self.device, dtype=torch.float32).contiguous()
      return glm_mfma.GlmHmcRunner(self, x.shape[0]).bootstrap(x)
    return super().log_prob_and_grad_flat(x, bijectors)


class LinearRegression(_GlmData, TargetSpec):
  """Bayesian Gaussian linear regression with unknown noise scale.

  beta ~ N(0, prior_scale^2 I_D); sigma ~ HalfNormal(sigma_scale);
  y_n ~ N(x_n . beta, sigma).  State parts: [beta (D,), sigma ()].
  """

  kind = _lib.TARGET_LINEAR

  def __init__(self, x, y, prior_scale=1.0, sigma_scale=1.0,
               dtype=torch.float32, device=None):
    super().__init__(dtype, device)
    self._set_design(_to_device(x, dtype, self.device))
    self.y = _to_device(y, dtype, self.device).reshape(-1)
    self.n_obs, d = self.x.shape
    self.prior_scale = float(prior_scale)
    self.sigma_scale = float(sigma_scale)
    self.event_shapes = [(d,), ()]

  def _fill(self, p):
    dx = self.dim - 1
    p.n_obs = self.n_obs
    self._fill_design(p)
    p.y = self.y.data_ptr()
    p.param[0] = self.prior_scale
    p.param[1] = self.sigma_scale
    p.log_norm = (self.n_obs * -_HALF_LOG_2PI
                  + dx * (-_HALF_LOG_2PI - math.log(self.prior_scale))
                  - math.log(self.sigma_scale) - 0.5 * math.log(math.pi / 2))


class HierarchicalNormal(TargetSpec):
  """Non-centred hierarchical normal (eight-schools family).

  mu ~ N(0, mu_scale); tau ~ HalfCauchy(0, tau_scale); z_j ~ N(0, 1);
  y_j ~ N(mu + tau z_j, sigma_j).  State parts: [mu (), tau (), z (J,)].
  """

  kind = _lib.TARGET_HIER_NORMAL

  def __init__(self, y, sigma, mu_scale=5.0, tau_scale=5.0,
               dtype=torch.float32, device=None):
    super().__init__(dtype, device)
    self.y = _to_device(y, dtype, self.device).reshape(-1)
    self.sigma = _to_device(sigma, dtype, self.device).reshape(-1)
    self.n_groups = self.y.shape[0]
    self.mu_scale = float(mu_scale)
    self.tau_scale = float(tau_scale)
    self.event_shapes = [(), (), (self.n_groups,)]
    self._sum_log_sigma = float(np.log(self.sigma.double().cpu().numpy()).sum())

  def _fill(self, p):
    j = self.n_groups
    p.n_obs = j
    p.y, p.aux = self.y.data_ptr(), self.sigma.data_ptr()
    p.param[0] = self.mu_scale
    p.param[1] = self.tau_scale
    p.log_norm = (-_HALF_LOG_2PI - math.log(self.mu_scale)
                  + math.log(2.0 / math.pi) - math.log(self.tau_scale)
                  + j * -_HALF_LOG_2PI
                  + j * -_HALF_LOG_2PI - self._sum_log_sigma)


class DiagNormal(TargetSpec):
  """Independent normals N(loc_i, scale_i); one part of shape [D]."""

  kind = _lib.TARGET_DIAG_NORMAL

  def __init__(self, loc, scale, dtype=torch.float32, device=None):
    super().__init__(dtype, device)
    self.loc = _to_device(loc, dtype, self.device).reshape(-1)
    self.scale = _to_device(scale, dtype, self.device).reshape(-1)
    self.event_shapes = [(self.loc.shape[0],)]
    self._sum_log_scale = float(np.log(self.scale.double().cpu().numpy()).sum())

  def _fill(self, p):
    p.n_obs = self.dim
    p.y, p.aux = self.loc.data_ptr(), self.scale.data_ptr()
    p.log_norm = self.dim * -_HALF_LOG_2PI - self._sum_log_scale


class NealsFunnel(TargetSpec):
  """v ~ N(0, 3), x_i ~ N(0, exp(v/2)), i < dim-1.  Parts [v (), x (dim-1,)]."""

  kind = _lib.TARGET_FUNNEL

  def __init__(self, dim=10, dtype=torch.float32, device=None):
    super().__init__(dtype, device)
    self.event_shapes = [(), (dim - 1,)]

  def _fill(self, p):
    p.log_norm = (-_HALF_LOG_2PI - math.log(3.0)) + (self.dim - 1) * -_HALF_LOG_2PI


class NaNExponential(TargetSpec):
  """log p = sum(where(x < 0, NaN, -x)); one part of shape [D]."""

  kind = _lib.TARGET_NAN_EXP

  def __init__(self, dim, dtype=torch.float32, device=None):
    super().__init__(dtype, device)
    self.event_shapes = [(dim,)]

  def _fill(self, p):
    p.log_norm = 0.0


class LogGamma(TargetSpec):
  """x_i = log of iid Gamma(concentration, rate) variables; one part of
  shape [D]: log p = sum(concentration x - rate exp(x)) + D (concentration
  log(rate) - lgamma(concentration)).  E[x] = digamma(concentration) -
  log(rate), E[exp(x)] = concentration / rate (`mcmc/hmc_test.py:146-200`).
  """

  kind = _lib.TARGET_LOG_GAMMA

  def __init__(self, dim, concentration, rate, dtype=torch.float32,
               device=None):
    super().__init__(dtype, device)
    self.event_shapes = [(int(dim),)]
    self.concentration, self.rate = float(concentration), float(rate)

  def _fill(self, p):
    c, r = self.concentration, self.rate
    p.param[0], p.param[1] = c, r
    p.log_norm = self.dim * (c * math.log(r) - math.lgamma(c))


def as_target(fn):
  """Returns `fn` if it is a device target spec, else raises TypeError."""
  if isinstance(fn, TargetSpec):
    return fn
  raise TypeError(
      'target_log_prob_fn must be a probability_amd.targets spec (the GPU '
      f'kernels evaluate analytic log-densities); got {type(fn).__name__}')
