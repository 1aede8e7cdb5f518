"""Seed plumbing: the role of `internal/samplers.py:79-258` (sanitize_seed,
split_seed) under this library's counter-keyed Philox schedule.

Seeds are uint32[2] NumPy arrays.  The derivations run in libhipmc on the
host with the same Philox block the kernels use, so a host-side loop and a
device-side fused loop derive identical per-step seeds.
"""

import ctypes
import zlib

import numpy as np

from probability_amd import _lib


def _arr(seed):
  return np.ascontiguousarray(np.asarray(seed, dtype=np.uint32).reshape(2))


def _ptr(a):
  return a.ctypes.data_as(ctypes.c_void_p)


def sanitize_seed(seed, salt=None):
  """int, uint32[2] or None -> uint32[2].  None draws fresh OS entropy."""
  if seed is None:
    seed = int(np.random.SeedSequence().entropy & 0xFFFFFFFFFFFFFFFF)
  if isinstance(seed, (int, np.integer)):
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    out = np.array([s & 0xFFFFFFFF, s >> 32], dtype=np.uint32)
  else:
    out = _arr(seed)
  if salt is not None:
    h = zlib.crc32(salt.encode('utf-8')) & 0xFFFFFFFF
    res = np.zeros(2, np.uint32)
    _lib.check(_lib.lib().hipmc_seed_salt(_ptr(out), ctypes.c_uint32(h),
                                           _ptr(res)))
    out = res
  return out


def split_seed(seed, n=2):
  """[n, 2] child seeds."""
  s = sanitize_seed(seed)
  out = np.zeros((n, 2), np.uint32)
  row = np.zeros(2, np.uint32)
  for i in range(n):
    _lib.check(_lib.lib().hipmc_seed_split(_ptr(s), ctypes.c_uint32(i),
                                            _ptr(row)))
    out[i] = row
  return out


def fold_in(seed, t):
  """Seed of step t of a sampling loop."""
  s = sanitize_seed(seed)
  out = np.zeros(2, np.uint32)
  _lib.check(_lib.lib().hipmc_seed_fold_in(_ptr(s), ctypes.c_uint64(int(t)),
                                            _ptr(out)))
  return out


def zeros_seed():
  return np.zeros(2, np.uint32)
