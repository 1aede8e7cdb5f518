"""The committed launch-shape table and its lookup (host logic, no GPU)."""

import json

import torch

from probability_amd import _lib
from probability_amd import tuning


def test_table_is_complete_and_consistent():
  with open(tuning.TABLE_PATH) as f:
    table = json.load(f)
  rows = table['glm_waves']['float32']
  assert rows
  for r in rows:
    times = {int(w): t for w, t in r['us_per_chain_leapfrog'].items()}
    assert r['best'] == min(times, key=times.get)
    assert set(times) == ({1, 2, 4} if r['sampler'] == 'hmc' else
                          {1, 2, 4, 8})


def test_lookup_returns_the_measured_choice_at_grid_points():
  with open(tuning.TABLE_PATH) as f:
    rows = json.load(f)['glm_waves']['float32']
  for r in rows:
    kind = (_lib.TARGET_LINEAR if r['kind'] == 'linear'
            else _lib.TARGET_LOGISTIC)
    assert tuning.glm_waves(kind, r['n_obs'], r['dim'], r['chains'],
                            r['sampler'], torch.float32) == r['best']


def test_lookup_falls_back_to_the_rule_without_measurements():
  # float64 launches have no measured points
  assert tuning.glm_waves(_lib.TARGET_LOGISTIC, 8192, 128, 32, 'nuts',
                          torch.float64) == _lib.glm_waves_per_chain(8192 * 128)
  # HMC kernels take at most four waves per chain
  assert tuning.glm_waves(_lib.TARGET_LOGISTIC, 8192, 128, 32, 'hmc',
                          torch.float64) <= 4
