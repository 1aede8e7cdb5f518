"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle still reproduces each one, and each is checked
against what the reference's tests or the mathematics fix."""

import os

import numpy as np
import pytest

from oracle import mcmc as omcmc
from oracle import targets as otg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _load(name):
  with np.load(os.path.join(HERE, name + '.npz')) as f:
    return {k: f[k] for k in f.files}


@pytest.mark.parametrize('name', ['leapfrog_tiny', 'hmc_step_c0',
                                  'nuts_step_tiny', 'da_updates',
                                  'ess_rhat_ar1'])
def test_oracle_reproduces_fixture(name):
  import importlib.util
  spec = importlib.util.spec_from_file_location(
      'make_golden', os.path.join(HERE, 'make_golden.py'))
  mod = importlib.util.module_from_spec(spec)
  spec.loader.exec_module(mod)
  fresh = mod.FIXTURES[name]()
  gold = _load(name)
  assert set(fresh) == set(gold)
  for k in gold:
    np.testing.assert_allclose(np.asarray(fresh[k], np.float64),
                               gold[k].astype(np.float64), rtol=1e-13,
                               atol=1e-13, err_msg=k)


def test_leapfrog_fixture_is_reversible():
  g = _load('leapfrog_tiny')
  t = otg.LogisticRegression(g['x'], g['y'], 1.0)
  t1, g1 = t.log_prob_and_grad(g['q1'])
  p_back, q_back, _, _ = omcmc.leapfrog(t, float(g['eps']),
                                        int(g['num_steps']), -g['p1'],
                                        g['q1'], t1, g1)
  np.testing.assert_allclose(q_back, g['q0'], atol=1e-12)
  np.testing.assert_allclose(-p_back, g['p0'], atol=1e-12)


def test_da_fixture_matches_reference_table():
  # `dual_averaging_step_size_adaptation_test.py:45-50,605-671`
  m05, m02, m01, u0, u01 = 9.131008, 9.642897, 9.819825, 10., 10.183481
  want = np.array([m02, u01, m02,
                   m05, u01, m02, u01, u01, m02,
                   m02, u0,
                   m02, u01, m02,
                   m01, m01, m01])
  np.testing.assert_allclose(_load('da_updates')['step_size_after_two'], want,
                             rtol=1e-6)


def test_ess_fixture_matches_ar1_closed_form():
  g = _load('ess_rhat_ar1')
  n, c, _ = g['states'].shape
  phi = np.array([0.0, 0.5, 0.9])
  want = n * c * (1 - phi) / (1 + phi)
  np.testing.assert_allclose(g['ess'], want, rtol=0.35)
  assert np.all(g['rhat'] < 1.1)


def test_nuts_fixture_is_consistent():
  g = _load('nuts_step_tiny')
  codes = np.zeros(g['q0'].shape[1], np.int32)
  codes[1] = 1
  t = otg.Transformed(otg.HierarchicalNormal(g['y'], g['sigma']), codes)
  t1, _ = t.log_prob_and_grad(g['q1'])
  np.testing.assert_allclose(g['target_log_prob'], t1, rtol=1e-12)
  # a tree of depth d holds the 2^(d-1) - 1 leaves of the completed levels
  # plus 1 .. 2^(d-1) leaves of the last subtree
  d = g['tree_depth']
  lf = g['leapfrogs_taken']
  assert np.all((lf >= 2 ** (d - 1)) & (lf <= 2 ** d - 1))
