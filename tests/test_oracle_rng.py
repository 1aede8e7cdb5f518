"""Pins the Philox block and the draw schedule shared by oracle and GPU."""

import ctypes

import numpy as np
import pytest

from oracle import philox
from oracle import rng

# Random123 known-answer vectors for philox4x32 with 10 rounds
# (Salmon et al., SC'11; kat_vectors of the Random123 distribution).
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2,
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
     [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def _u32(v):
  return np.asarray(v, np.uint64).astype(np.uint32)


@pytest.mark.parametrize('ctr,key,out', KAT)
def test_philox_known_answers(ctr, key, out):
  np.testing.assert_array_equal(philox.philox4x32_10(_u32(ctr), _u32(key)),
                                _u32(out))


def test_philox_vectorised_matches_scalar():
  r = np.random.default_rng(0)
  ctr = r.integers(0, 2**32, size=(50, 4), dtype=np.uint64).astype(np.uint32)
  key = r.integers(0, 2**32, size=(50, 2), dtype=np.uint64).astype(np.uint32)
  v = philox.philox4x32_10(ctr, key)
  for i in range(50):
    np.testing.assert_array_equal(v[i], philox.philox4x32_10(ctr[i], key[i]))


@pytest.mark.parametrize('ctr,key,out', KAT)
def test_library_host_philox_known_answers(ctr, key, out):
  from probability_amd import _lib
  c = np.ascontiguousarray(_u32(ctr))
  k = np.ascontiguousarray(_u32(key))
  o = np.zeros(4, np.uint32)
  _lib.check(_lib.lib().hipmc_philox_host(
      c.ctypes.data_as(ctypes.c_void_p), k.ctypes.data_as(ctypes.c_void_p),
      o.ctypes.data_as(ctypes.c_void_p)))
  np.testing.assert_array_equal(o, _u32(out))


def test_library_seed_derivation_matches_oracle():
  from probability_amd import random as prand
  for s in (0, 1, 42, 2**40 + 7, [3, 4]):
    np.testing.assert_array_equal(prand.sanitize_seed(s), rng.sanitize_seed(s))
    np.testing.assert_array_equal(prand.sanitize_seed(s, salt='abc'),
                                  rng.sanitize_seed(s, salt='abc'))
    np.testing.assert_array_equal(prand.split_seed(s, 3), rng.split_seed(s, 3))
    for t in (0, 1, 999, 2**33):
      np.testing.assert_array_equal(prand.fold_in(s, t), rng.fold_in(s, t))


def test_uniform_ranges_and_moments():
  u32 = rng.mh_uniform(7, np.arange(20000), np.float32)
  u64 = rng.mh_uniform(7, np.arange(20000), np.float64)
  assert u32.dtype == np.float32 and u64.dtype == np.float64
  for u in (u32, u64):
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1 / 12) < 0.003
  # float32 uniforms are exact multiples of 2^-24
  assert np.all((u32.astype(np.float64) * 2**24) % 1 == 0)


def test_momentum_is_standard_normal_and_chain_keyed():
  m = rng.momentum(11, np.arange(4000), 7)
  assert abs(m.mean()) < 0.02
  assert abs(m.std() - 1) < 0.02
  # a chain's draws do not depend on which other chains are drawn
  sub = rng.momentum(11, np.array([5, 17, 3999]), 7)
  np.testing.assert_array_equal(sub, m[[5, 17, 3999]])
  # problem id changes the stream
  other = rng.momentum(11, np.arange(4), 7, problem=1)
  assert not np.allclose(other, m[:4])


def test_tree_and_leaf_draws_are_distinct_streams():
  d, tu = rng.tree_draws(5, np.arange(10000), 3)
  assert 0.47 < d.mean() < 0.53
  lu = rng.leaf_uniform(5, np.arange(10000), 3, 0)
  assert abs(np.corrcoef(tu, lu)[0, 1]) < 0.05
