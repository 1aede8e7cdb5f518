"""The C ABI library loads, exports every declared symbol, and its struct
layouts match the ctypes mirror.  No device calls."""

import ctypes
import os
import re

import numpy as np
import pytest

from probability_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
  text = open(os.path.join(ROOT, 'include', 'hipmc.h')).read()
  return sorted(set(re.findall(r'\b(hipmc_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
  lib = _lib.lib()
  names = _declared()
  assert len(names) >= 15
  for n in names:
    assert hasattr(lib, n), n
  assert set(names) == set(_lib.EXPORTED_SYMBOLS)


def test_struct_layouts():
  lib = _lib.lib()
  assert lib.hipmc_sizeof_problem() == ctypes.sizeof(_lib.Problem)
  assert lib.hipmc_sizeof_run_args() == ctypes.sizeof(_lib.RunArgs)
  assert lib.hipmc_sizeof_stream() == ctypes.sizeof(_lib.Stream)
  assert lib.hipmc_abi_version() == 4


def test_invalid_arguments_return_codes_without_device():
  lib = _lib.lib()
  assert lib.hipmc_nuts_run(None, 0) == _lib.E_INVALID
  assert b'NULL' in lib.hipmc_last_error()
  a = _lib.RunArgs()
  a.dtype = 7
  assert lib.hipmc_hmc_run(ctypes.byref(a), 0) == _lib.E_INVALID
  a.dtype = 0
  a.n_iters = 1
  # GLM states are built up to 512 elements, the others up to 1088
  a.target_kind = _lib.TARGET_LOGISTIC
  a.max_dim = 600
  a.dim_stride = 600
  assert lib.hipmc_nuts_run(ctypes.byref(a), 0) == _lib.E_UNSUPPORTED
  assert b'above 512' in lib.hipmc_last_error()
  a.target_kind = _lib.TARGET_HIER_NORMAL
  assert lib.hipmc_nuts_run(ctypes.byref(a), 0) == _lib.E_INVALID  # no table
  a.max_dim = a.dim_stride = 1089
  assert lib.hipmc_nuts_run(ctypes.byref(a), 0) == _lib.E_UNSUPPORTED
  assert b'above 1088' in lib.hipmc_last_error()
  with pytest.raises(ValueError):
    _lib.check(lib.hipmc_hmc_run(None, 0))


def test_targets_refuse_plain_callables():
  from probability_amd import targets
  with pytest.raises(TypeError):
    targets.as_target(lambda x: x)


def test_da_group_plans():
  from probability_amd.mcmc import dual_averaging_step_size_adaptation as da
  goc, sg, ng = da.chain_groups((2, 3), (), 3)
  assert ng == 1 and np.all(goc == 0) and list(sg) == [0]
  goc, sg, ng = da.chain_groups((2, 3), (3, 1), 3)
  assert ng == 3 and list(goc) == [0, 1, 2, 0, 1, 2] and list(sg) == [0, 1, 2]
  goc, sg, ng = da.chain_groups((2, 3), (2, 3, 1), 3)
  assert ng == 6 and list(goc) == list(range(6))
  goc, sg, ng = da.chain_groups((2, 3), (2, 1, 1), 3)
  assert ng == 2 and list(goc) == [0, 0, 0, 1, 1, 1] and list(sg) == [0, 1]
