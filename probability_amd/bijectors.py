"""Bijectors understood by the fused TTK path.

`Identity`, `Softplus` and `Exp` mirror `tfp.bijectors.Identity`,
`tfp.bijectors.Softplus` (`bijectors/softplus.py:36-172`, default
hinge_softness and low) and `tfp.bijectors.Exp` (= PowerTransform(0),
`bijectors/power_transform.py:85-112`) for scalar events.  Each carries the
code the kernels use; the elementwise forward/inverse maps run in libhipmc.
"""

from probability_amd import _lib


class Bijector:
  code = _lib.BIJ_IDENTITY
  name = 'identity'

  def __repr__(self):
    return f'{type(self).__name__}()'


class Identity(Bijector):
  code = _lib.BIJ_IDENTITY
  name = 'identity'


class Softplus(Bijector):
  code = _lib.BIJ_SOFTPLUS
  name = 'softplus'

  def __init__(self, hinge_softness=None, low=None):
    if hinge_softness is not None or low is not None:
      raise NotImplementedError(
          'Softplus with hinge_softness/low is not built into the kernels')


class Exp(Bijector):
  code = _lib.BIJ_EXP
  name = 'exp'


def as_bijector(b):
  if b is None:
    return Identity()
  if isinstance(b, Bijector):
    return b
  raise TypeError(
      f'bijector {b!r} is not one of Identity, Softplus, Exp of '
      'probability_amd.bijectors')
