"""The streaming-moments oracle (`oracle/reducers.py`) against closed forms
and the reference tests' known answers (`experimental/mcmc/*_reducer_test.py`,
`experimental/stats/sample_stats_test.py`), plus the host-side structure
logic of the reducers (no GPU)."""

import numpy as np
import pytest

from oracle import diagnostic as odiag
from oracle import reducers as ored


def _stream(x, event_ndims, mode, chunks=None):
  """Feeds x [N, ...] one sample at a time, or in the given chunk sizes."""
  m = ored.Moments(x.shape[1:], event_ndims, mode)
  if chunks is None:
    for r in range(x.shape[0]):
      ored.update(m, x[r])
  else:
    at = 0
    for c in chunks:
      ored.update(m, x[at:at + c], axis=0)
      at += c
    assert at == x.shape[0]
  return m


@pytest.mark.parametrize('chunks', [None, [1] * 40, [7, 13, 20], [40]])
def test_variance_equals_batch(chunks):
  x = np.random.default_rng(0).normal(size=(40, 5, 3)) * 3 + 1
  m = _stream(x, 0, 1, chunks)
  np.testing.assert_allclose(m.mean, x.mean(0), rtol=1e-13, atol=1e-13)
  np.testing.assert_allclose(ored.covariance(m), x.var(0), rtol=1e-12)
  np.testing.assert_allclose(ored.covariance(m, 1), x.var(0, ddof=1),
                             rtol=1e-12)
  assert m.num_samples == 40


@pytest.mark.parametrize('chunks', [None, [3, 30, 7]])
def test_covariance_event_ndims_1_equals_np_cov(chunks):
  x = np.random.default_rng(1).normal(size=(40, 4, 3))
  x[..., 1] += 0.5 * x[..., 0]
  m = _stream(x, 1, 2, chunks)
  assert ored.covariance(m).shape == (4, 3, 3)
  for b in range(4):
    np.testing.assert_allclose(ored.covariance(m)[b],
                               np.cov(x[:, b].T, ddof=0), rtol=1e-12,
                               atol=1e-14)
    np.testing.assert_allclose(ored.covariance(m, 1)[b],
                               np.cov(x[:, b].T, ddof=1), rtol=1e-12,
                               atol=1e-14)


def test_covariance_all_dims_equals_np_cov():
  x = np.random.default_rng(2).normal(size=(30, 2, 3))
  m = _stream(x, 2, 2, [10, 20])
  flat = x.reshape(30, 6)
  np.testing.assert_allclose(ored.covariance(m).reshape(6, 6),
                             np.cov(flat.T, ddof=0), rtol=1e-12, atol=1e-14)


def test_dropping_the_merge_term_is_caught():
  """A planted mistake in the oracle (no cross-chunk term) breaks the check
  above, so that check pins the merge formula."""
  x = np.random.default_rng(3).normal(size=(20, 3))
  m = ored.Moments((3,), 0, 1)
  for r in range(20):
    cm = x[r]
    delta = cm - m.mean
    m.ssr = m.ssr + 0.0  # merge term (n * 1 / (n + 1)) * delta**2 dropped
    m.mean = m.mean + delta / (m.num_samples + 1)
    m.num_samples += 1
  assert np.all(np.abs(ored.covariance(m) - x.var(0)) > 1e-3)


def test_reference_expectations_known_answers():
  # expectations_reducer_test.py:39-55: means of 0..5 and 1..6
  a, b = ored.Moments((), 0, 0), ored.Moments((), 0, 0)
  for s in range(6):
    ored.update(a, s)
    ored.update(b, s + 1)
  assert (a.mean, b.mean) == (2.5, 3.5)
  # :57-65: chunks of nine equal samples along axis 1
  m = ored.Moments((3,), 0, 0)
  for s in range(6):
    ored.update(m, np.ones((3, 9)) * s, axis=1)
  np.testing.assert_array_equal(m.mean, [2.5, 2.5, 2.5])


def test_reference_zero_covariance():
  # covariance_reducer_test.py:46-59
  m = ored.Moments((9, 3), 1, 2)
  for _ in range(2):
    ored.update(m, np.zeros((5, 9, 3)), axis=0)
  assert m.num_samples == 10
  np.testing.assert_array_equal(m.mean, np.zeros((9, 3)))
  np.testing.assert_array_equal(ored.covariance(m), np.zeros((9, 3, 3)))


def test_streaming_rhat_equals_batch_rhat_int_samples():
  # potential_scale_reduction_reducer_test.py:33-45
  chain_state = np.arange(60).reshape((4, 5, 3))
  m = ored.Moments((5, 3), 0, 1)
  for s in chain_state:
    ored.update(m, s)
  want = odiag.potential_scale_reduction(chain_state, 1, split_chains=False)
  np.testing.assert_allclose(ored.potential_scale_reduction(m, 1), want,
                             rtol=1e-12)


@pytest.mark.parametrize('icn', [1, 2])
def test_streaming_rhat_equals_batch_rhat(icn):
  r = np.random.default_rng(4)
  x = r.normal(size=(200, 4, 3, 2)) + r.normal(size=(1, 4, 3, 2))
  m = _stream(x, 0, 1)
  want = odiag.potential_scale_reduction(x, icn, split_chains=False)
  got = ored.potential_scale_reduction(m, icn)
  assert got.shape == x.shape[1 + icn:]
  np.testing.assert_allclose(got, want, rtol=1e-10)


def test_rhat_iid_passes_offset_fails():
  # potential_scale_reduction_reducer_test.py:47-71
  r = np.random.default_rng(5)
  m = _stream(r.normal(size=(500, 5)), 0, 1)
  assert abs(ored.potential_scale_reduction(m, 1) - 1.) < 0.02
  off = r.normal(size=(500, 3, 4)) + np.array([1., -1., 2.]).reshape(3, 1)
  m = _stream(off, 0, 1)
  assert np.all(ored.potential_scale_reduction(m, 1) > 1.2)


# ------------------------------------------------- host logic (no GPU)
def test_structure_map_and_validation():
  from probability_amd.experimental import mcmc as emcmc
  out = emcmc._map(lambda a, b: a + b, {'x': [1, 2], 'y': (3,)}, 10)
  assert out == {'x': [11, 12], 'y': (13,)}
  out = emcmc._map(lambda a, b: a * b, [1, 2], [3, 4])
  assert out == [3, 8]
  r1, r2 = emcmc.ExpectationsReducer(), emcmc.VarianceReducer()
  got = emcmc._reducer_map(lambda r, s: (type(r).__name__, s),
                           {'m': r1, 'v': [r2]}, {'m': 1, 'v': [2]})
  assert got == {'m': ('ExpectationsReducer', 1),
                 'v': [('VarianceReducer', 2)]}
  assert r2.event_ndims == 0 and r2.ddof == 0
  with pytest.raises(ValueError):
    emcmc.PotentialScaleReductionReducer(independent_chain_ndims=0)
  with pytest.raises(ValueError):
    emcmc.sample_fold(3, 0.)


def test_no_cpu_fallback():
  import torch
  from probability_amd import _lib
  from probability_amd.experimental import stats as estats
  if torch.cuda.is_available():
    pytest.skip('GPU present')
  with pytest.raises(_lib.HipmcError):
    estats.RunningVariance.from_shape((3,))


def test_sample_discarding_and_step_kernel_schedule():
  """Burn-in and thinning (`sample_discarding_kernel.py:81-93`) and the
  per-step seeds of step_kernel, on a counting kernel (no GPU)."""
  from probability_amd import random as samplers
  from probability_amd.experimental import mcmc as emcmc
  from probability_amd.mcmc import kernel as kernel_base

  class Count(kernel_base.TransitionKernel):
    seeds = []

    @property
    def is_calibrated(self):
      return True

    def one_step(self, state, kr, seed=None):
      Count.seeds.append(None if seed is None else tuple(int(v) for v in seed))
      return state + 1, kr + 1

    def bootstrap_results(self, state):
      return 0

  k = emcmc.SampleDiscardingKernel(Count(), num_burnin_steps=3,
                                   num_steps_between_results=2)
  kr = k.bootstrap_results(0)
  assert kr.call_counter == 0
  s, kr = k.one_step(0, kr)
  assert (s, kr.call_counter, kr.inner_results) == (6, 1, 6)
  s, kr = k.one_step(s, kr)
  assert (s, kr.call_counter) == (9, 2)
  Count.seeds = []
  s, kr = emcmc.step_kernel(3, 0, kernel=Count(), seed=5,
                            return_final_kernel_results=True)
  base = samplers.sanitize_seed(5, salt='experimental.mcmc.step_kernel')
  want = [tuple(int(v) for v in samplers.fold_in(base, t)) for t in range(3)]
  assert (s, Count.seeds) == (3, want)
  with pytest.raises(ValueError):
    emcmc.SampleDiscardingKernel(Count(), num_burnin_steps=-1)
