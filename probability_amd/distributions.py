"""JointDistribution models compiled to device target specs.

A model written with TFP's names and calling conventions --
`JointDistributionSequential` (`distributions/joint_distribution_sequential.py`:
a list whose callables receive the previously built variables, most recent
first) or `JointDistributionNamed` (`joint_distribution_named.py`: a dict
whose callables receive variables by argument name) over `Normal`,
`HalfNormal`, `HalfCauchy`, `Bernoulli(logits=...)`, `Sample` and
`Independent` -- is traced once with symbolic variables, and the traced graph
is matched against the model families the HIP kernels evaluate:

* logistic regression: beta ~ N(0, s)^D, y ~ Bernoulli(logits = X beta);
* linear regression: beta ~ N(0, s)^D, sigma ~ HalfNormal(h),
  y ~ N(X beta, sigma);
* non-centred hierarchical normal: mu ~ N(0, a), tau ~ HalfCauchy(0, b),
  z ~ N(0, 1)^J, y ~ N(mu + tau z, sigma_j);
* independent normals: x ~ N(loc, scale) with constant loc and scale.

`compile_target(model, dtype, **pins)` returns the target spec, the
unconstraining bijector of each free part (the distributions' default event
space bijectors, `half_cauchy.py` / `half_normal.py`: Softplus; Normal:
Identity), and the free parts' names -- what `windowed_sampling.py:151-190`
(`_get_flat_unconstraining_bijector`) and `experimental_pin` derive from a
model.  The model's log density equals the spec's, constants included.
Models outside these families raise `NotImplementedError` (there is no
autodiff path on the device).
"""

import collections
import inspect

import numpy as np
import torch

from probability_amd import bijectors as bij
from probability_amd import targets

__all__ = ['Bernoulli', 'CompiledTarget', 'HalfCauchy', 'HalfNormal',
           'Independent', 'JointDistributionNamed',
           'JointDistributionSequential', 'Normal', 'Sample', 'compile_target',
           'matvec']


# ------------------------------------------------------------ tracing values
class _Expr:
  """A value computed from model variables while tracing."""

  __array_ufunc__ = None  # numpy operators defer to the reflected methods

  def __init__(self, op, *args):
    self.op, self.args = op, args

  def __add__(self, other):
    return _Expr('add', self, other)

  __radd__ = __add__

  def __mul__(self, other):
    return _Expr('mul', self, other)

  __rmul__ = __mul__

  def __rmatmul__(self, x):
    return _Expr('matvec', x, self)

  def __repr__(self):
    return f'{self.op}{self.args!r}'


class _Var(_Expr):

  def __init__(self, name):
    super().__init__('var', name)
    self.name = name

  def __repr__(self):
    return self.name


def matvec(x, v):
  """`tf.linalg.matvec(x, v)`: x @ v for a traced v."""
  if isinstance(v, _Expr):
    return _Expr('matvec', x, v)
  return np.asarray(x) @ np.asarray(v)


def _const(v):
  """float or array for a constant parameter; None when it is traced."""
  if isinstance(v, _Expr):
    return None
  if isinstance(v, torch.Tensor):
    return v.detach().cpu().numpy().astype(np.float64)
  return np.asarray(v, np.float64)


def _scalar(v):
  c = _const(v)
  if c is None:
    return None
  flat = c.reshape(-1)
  if flat.size == 0 or np.any(flat != flat[0]):
    return None
  return float(flat[0])


# ------------------------------------------------------------ distributions
class Distribution:
  batch_shape = ()


class Normal(Distribution):
  """`normal.py`: log N(x; loc, scale) = -(x-loc)^2/(2 scale^2) - log scale
  - log(2 pi)/2."""

  def __init__(self, loc, scale, name=None):
    self.loc, self.scale, self.name = loc, scale, name


class HalfNormal(Distribution):
  """`half_normal.py:146-153`."""

  def __init__(self, scale, name=None):
    self.scale, self.name = scale, name


class HalfCauchy(Distribution):
  """`half_cauchy.py:138-145`."""

  def __init__(self, loc, scale, name=None):
    self.loc, self.scale, self.name = loc, scale, name


class Bernoulli(Distribution):
  """`bernoulli.py:119-135` (logits parameterisation only)."""

  def __init__(self, logits=None, probs=None, name=None):
    if logits is None or probs is not None:
      raise NotImplementedError('Bernoulli is compiled from logits only')
    self.logits, self.name = logits, name


class Sample(Distribution):
  """`sample.py`: `sample_shape` iid draws of `distribution` as one event."""

  def __init__(self, distribution, sample_shape=(), name=None):
    self.distribution = distribution
    self.sample_shape = ((int(sample_shape),) if np.ndim(sample_shape) == 0
                         else tuple(int(s) for s in sample_shape))
    self.name = name


class Independent(Distribution):
  """`independent.py`: batch dimensions reinterpreted as event dimensions."""

  def __init__(self, distribution, reinterpreted_batch_ndims=None, name=None):
    self.distribution = distribution
    self.reinterpreted_batch_ndims = reinterpreted_batch_ndims
    self.name = name


class JointDistributionSequential:
  """A list of distributions and callables of the earlier variables."""

  def __init__(self, model, name=None):
    self.model = list(model)
    self.name = name

  def _trace(self):
    """[(name, distribution)] in model order with traced parameters."""
    n = len(self.model)
    names = [None] * n
    for k, m in enumerate(self.model):  # names from later callables' args
      if callable(m) and not isinstance(m, Distribution):
        for j, a in enumerate(inspect.signature(m).parameters):
          if k - 1 - j >= 0 and names[k - 1 - j] is None:
            names[k - 1 - j] = a
    for k, m in enumerate(self.model):
      d = m if isinstance(m, Distribution) else None
      if d is not None and getattr(d, 'name', None):
        names[k] = d.name
      if names[k] is None:
        names[k] = f'x{k}'
    nodes, vars_ = [], []
    for k, m in enumerate(self.model):
      if isinstance(m, Distribution):
        d = m
      else:
        nargs = len(inspect.signature(m).parameters)
        d = m(*list(reversed(vars_))[:nargs])
      nodes.append((names[k], d))
      vars_.append(_Var(names[k]))
    return nodes


class JointDistributionNamed:
  """A dict of distributions and callables taking earlier variables by
  argument name (in the dict's order)."""

  def __init__(self, model, name=None):
    self.model = dict(model)
    self.name = name

  def _trace(self):
    nodes = []
    for k, m in self.model.items():
      if isinstance(m, Distribution):
        d = m
      else:
        args = list(inspect.signature(m).parameters)
        unknown = [a for a in args if a not in self.model]
        if unknown:
          raise ValueError(f'{k}: unknown variables {unknown}')
        d = m(*[_Var(a) for a in args])
      nodes.append((k, d))
    return nodes


# ------------------------------------------------------------ matching
CompiledTarget = collections.namedtuple(
    'CompiledTarget', 'target, bijectors, names')


def _unwrap(d):
  """(base distribution, iid event size or None) of Sample / Independent."""
  if isinstance(d, Sample):
    base, _ = _unwrap(d.distribution)
    size = int(np.prod(d.sample_shape)) if d.sample_shape else 1
    return base, size
  if isinstance(d, Independent):
    return _unwrap(d.distribution)
  return d, None


def _event_size(d):
  base, size = _unwrap(d)
  if size is not None:
    return size
  sizes = [np.size(_const(getattr(base, p))) for p in ('loc', 'scale')
           if hasattr(base, p) and _const(getattr(base, p)) is not None]
  return max(sizes) if sizes else 1


def _cols(x):
  c = _const(x)
  return c.shape[1] if c is not None and c.ndim == 2 else -1


def _is_var(e, name):
  return isinstance(e, _Var) and e.name == name


def _std_normal_prior(d):
  """scale of an N(0, s) prior (constant zero loc, constant scale) or None."""
  base, _ = _unwrap(d)
  if not isinstance(base, Normal):
    return None
  loc, s = _const(base.loc), _scalar(base.scale)
  if loc is None or s is None or np.any(loc != 0):
    return None
  return s


def _match_glm(free, obs, pins):
  """Logistic or linear regression."""
  (bname, bd) = free[0]
  s = _std_normal_prior(bd)
  yname, yd = obs[0]
  ybase, _ = _unwrap(yd)
  y = pins[yname]
  if s is None:
    return None
  if isinstance(ybase, Bernoulli) and len(free) == 1:
    e = ybase.logits
    if (isinstance(e, _Expr) and e.op == 'matvec' and
        _is_var(e.args[1], bname) and _cols(e.args[0]) == _event_size(bd)):
      return (targets.LogisticRegression,
              dict(x=_const(e.args[0]), y=y, prior_scale=s),
              [bij.Identity()])
  if isinstance(ybase, Normal) and len(free) == 2:
    sname, sd = free[1]
    sbase, _ = _unwrap(sd)
    e = ybase.loc
    if (isinstance(sbase, HalfNormal) and _scalar(sbase.scale) is not None
        and _event_size(sd) == 1
        and isinstance(e, _Expr) and e.op == 'matvec'
        and _is_var(e.args[1], bname) and _is_var(ybase.scale, sname)
        and _cols(e.args[0]) == _event_size(bd)):
      return (targets.LinearRegression,
              dict(x=_const(e.args[0]), y=y, prior_scale=s,
                   sigma_scale=_scalar(sbase.scale)),
              [bij.Identity(), bij.Softplus()])
  return None


def _is_affine(e, mu, tau, z):
  """e == mu + tau * z (either operand order)."""
  if not (isinstance(e, _Expr) and e.op == 'add'):
    return False
  a, b = e.args
  for m, t in ((a, b), (b, a)):
    if _is_var(m, mu) and isinstance(t, _Expr) and t.op == 'mul':
      u, v = t.args
      if (_is_var(u, tau) and _is_var(v, z)) or (_is_var(u, z) and
                                                  _is_var(v, tau)):
        return True
  return False


def _match_hier(free, obs, pins):
  (mn, md), (tn, td), (zn, zd) = free
  a = _std_normal_prior(md)
  tb, _ = _unwrap(td)
  if a is None or not isinstance(tb, HalfCauchy):
    return None
  if _event_size(md) != 1 or _event_size(td) != 1:
    return None
  if _scalar(tb.loc) != 0.0 or _scalar(tb.scale) is None:
    return None
  if _std_normal_prior(zd) != 1.0:
    return None
  yname, yd = obs[0]
  yb, _ = _unwrap(yd)
  if not isinstance(yb, Normal) or not _is_affine(yb.loc, mn, tn, zn):
    return None
  sigma = _const(yb.scale)
  if sigma is None:
    return None
  y = np.asarray(pins[yname], np.float64).reshape(-1)
  sigma = np.broadcast_to(sigma, y.shape).astype(np.float64)
  if y.shape[0] != _event_size(zd):
    return None
  return (targets.HierarchicalNormal,
          dict(y=y, sigma=sigma, mu_scale=a, tau_scale=_scalar(tb.scale)),
          [bij.Identity(), bij.Softplus(), bij.Identity()])


def _match_diag(free):
  (name, d), = free
  base, size = _unwrap(d)
  if not isinstance(base, Normal):
    return None
  loc, scale = _const(base.loc), _const(base.scale)
  if loc is None or scale is None:
    return None
  n = size or max(np.size(loc), np.size(scale))
  return (targets.DiagNormal,
          dict(loc=np.broadcast_to(loc, (n,)).copy(),
               scale=np.broadcast_to(scale, (n,)).copy()),
          [bij.Identity()])


def match(model, **pins):
  """(target class, constructor kwargs, bijectors, free names) for `model`
  with the given observed variables, or NotImplementedError."""
  nodes = model._trace()
  names = [n for n, _ in nodes]
  unknown = [k for k in pins if k not in names]
  if unknown:
    raise ValueError(f'pinned names {unknown} are not in the model {names}')
  free = [(n, d) for n, d in nodes if n not in pins]
  obs = [(n, d) for n, d in nodes if n in pins]
  found = None
  if len(obs) == 1 and len(free) in (1, 2):
    found = _match_glm(free, obs, pins)
  if found is None and len(obs) == 1 and len(free) == 3:
    found = _match_hier(free, obs, pins)
  if found is None and not obs and len(free) == 1:
    found = _match_diag(free)
  if found is None:
    raise NotImplementedError(
        'this model is not one of the compiled families (logistic or linear '
        'regression, non-centred hierarchical normal with the free '
        'variables in the order mu, tau, z, independent normals); free '
        f'variables {[n for n, _ in free]}, observed {[n for n, _ in obs]}')
  cls, kwargs, bijs = found
  return cls, kwargs, bijs, [n for n, _ in free]


def compile_target(model, dtype=torch.float32, device=None, **pins):
  """The device target spec of `model` with `pins` observed."""
  cls, kwargs, bijs, names = match(model, **pins)
  return CompiledTarget(cls(dtype=dtype, device=device, **kwargs), bijs,
                        names)
