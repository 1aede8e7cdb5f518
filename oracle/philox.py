"""Philox4x32-10 counter-based generator, vectorised over NumPy arrays.

Algorithm: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel
random numbers: as easy as 1, 2, 3", SC'11.  Ten rounds of

    (hi0, lo0) = mulhilo(M0, c0);  (hi1, lo1) = mulhilo(M1, c2)
    c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0)

with the Weyl key bump (k0 += W0, k1 += W1) between rounds.  TF's stateless
RNG (`internal/samplers.py:308-367` calls `tf.random.stateless_*`) is built on
the same block function, but its seed-to-counter mapping is TF-internal and
not importable here, so this module is pinned by the published Random123
known-answer vectors and by rocRAND's Philox header instead.

Test infrastructure only (see `oracle/__init__.py`).
"""

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)
_SH32 = np.uint64(32)


def philox4x32_10(counter, key):
  """Returns the 4 output words for `counter` [..., 4] and `key` [..., 2].

  Inputs broadcast against each other; outputs are uint32 [..., 4].
  """
  counter = np.asarray(counter, dtype=np.uint32)
  key = np.asarray(key, dtype=np.uint32)
  shape = np.broadcast_shapes(counter.shape[:-1], key.shape[:-1])
  c0, c1, c2, c3 = [np.broadcast_to(counter[..., i], shape).astype(np.uint64)
                    for i in range(4)]
  k0 = np.broadcast_to(key[..., 0], shape).astype(np.uint32)
  k1 = np.broadcast_to(key[..., 1], shape).astype(np.uint32)
  with np.errstate(over='ignore'):
    for rnd in range(10):
      if rnd:
        k0 = k0 + W0
        k1 = k1 + W1
      p0 = M0 * c0
      p1 = M1 * c2
      hi0, lo0 = p0 >> _SH32, p0 & _MASK32
      hi1, lo1 = p1 >> _SH32, p1 & _MASK32
      c0, c1, c2, c3 = (hi1 ^ c1 ^ k0.astype(np.uint64), lo1,
                        hi0 ^ c3 ^ k1.astype(np.uint64), lo0)
  return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)
