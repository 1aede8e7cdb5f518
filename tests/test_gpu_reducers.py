"""Streaming reducers on the GPU (`hipmc_moments_update`) against the
oracle's one-pass moments, the reference tests' known answers, and full
traces of the same chains."""

import collections

import numpy as np
import pytest
import torch

from oracle import diagnostic as odiag
from oracle import reducers as ored
import probability_amd as pa
from probability_amd.experimental import mcmc as emcmc
from probability_amd.experimental import stats as estats

pytestmark = pytest.mark.gpu
# float64: sums in float64 on both sides, differing only in summation order
# float32: the same float32 samples; the result is rounded to float32
DTYPES = [(torch.float32, np.float32, 2e-6), (torch.float64, np.float64, 1e-12)]


def _oracle(x, event_ndims, mode, chunks):
  m = ored.Moments(x.shape[1:], event_ndims, mode)
  if chunks is None:
    for r in range(x.shape[0]):
      ored.update(m, x[r])
  else:
    at = 0
    for c in chunks:
      ored.update(m, x[at:at + c], axis=0)
      at += c
  return m


def _device(x, tdt):
  return torch.as_tensor(x, dtype=tdt, device='cuda')


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
@pytest.mark.parametrize('chunks', [None, [7, 13, 20]])
def test_running_variance_matches_oracle(tdt, ndt, tol, chunks):
  x = (np.random.default_rng(0).normal(size=(40, 5, 3)) * 3 + 1).astype(ndt)
  rv = estats.RunningVariance.from_shape((5, 3), tdt)
  if chunks is None:
    for r in range(40):
      rv = rv.update(_device(x[r], tdt))
  else:
    at = 0
    for c in chunks:
      rv = rv.update(_device(x[at:at + c], tdt), axis=0)
      at += c
  m = _oracle(x.astype(np.float64), 0, 1, chunks)
  assert rv.num_samples == 40
  assert rv.variance().dtype == tdt
  np.testing.assert_allclose(rv.mean.cpu().numpy(), m.mean, rtol=tol,
                             atol=tol)
  for ddof in (0, 1):
    np.testing.assert_allclose(rv.variance(ddof).cpu().numpy(),
                               ored.covariance(m, ddof), rtol=tol)


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
@pytest.mark.parametrize('event_ndims,chunks', [(1, None), (1, [3, 30, 7]),
                                                (2, [10, 30]), (None, None)])
def test_running_covariance_matches_oracle(tdt, ndt, tol, event_ndims,
                                           chunks):
  x = np.random.default_rng(1).normal(size=(40, 4, 3)).astype(ndt)
  x[..., 1] += 0.5 * x[..., 0]
  rc = estats.RunningCovariance.from_shape((4, 3), tdt, event_ndims)
  k = 2 if event_ndims is None else event_ndims
  assert rc.event_ndims == k
  if chunks is None:
    for r in range(40):
      rc = rc.update(_device(x[r], tdt))
  else:
    at = 0
    for c in chunks:
      rc = rc.update(_device(x[at:at + c], tdt), axis=0)
      at += c
  m = _oracle(x.astype(np.float64), k, 2, chunks)
  got = rc.covariance().cpu().numpy()
  assert got.shape == ((4, 3, 3) if k == 1 else (4, 3, 4, 3))
  np.testing.assert_allclose(got, ored.covariance(m), rtol=tol, atol=tol)
  np.testing.assert_allclose(rc.covariance(1).cpu().numpy(),
                             ored.covariance(m, 1), rtol=tol, atol=tol)
  np.testing.assert_allclose(rc.mean.cpu().numpy(), m.mean, rtol=tol,
                             atol=tol)


def test_update_leaves_the_old_state_unchanged():
  rm = estats.RunningMean.from_shape((3,), torch.float64)
  a = rm.update(_device(np.ones(3), torch.float64))
  b = a.update(_device(np.full(3, 3.), torch.float64))
  np.testing.assert_array_equal(a.mean.cpu().numpy(), [1., 1., 1.])
  np.testing.assert_array_equal(b.mean.cpu().numpy(), [2., 2., 2.])
  assert (a.num_samples, b.num_samples) == (1., 2.)


FakeKernelResults = collections.namedtuple('FakeKernelResults',
                                           'value, inner_results')
FakeInnerResults = collections.namedtuple('FakeInnerResults', 'value')


def test_reference_expectations_known_answers():
  # expectations_reducer_test.py:39-65
  red = emcmc.ExpectationsReducer(
      transform_fn=[lambda s, kr: kr.value,
                    lambda s, kr: kr.inner_results.value])
  kr = FakeKernelResults(0, FakeInnerResults(0))
  st = red.initialize(0, kr)
  for s in range(6):
    kr = FakeKernelResults(s, FakeInnerResults(s + 1))
    st = red.one_step(s, st, kr)
  assert [float(v) for v in red.finalize(st)] == [2.5, 3.5]
  red = emcmc.ExpectationsReducer()
  st = red.initialize(torch.ones(3))
  for s in range(6):
    st = red.one_step(torch.ones((3, 9), device='cuda') * s, st, axis=1)
  np.testing.assert_array_equal(red.finalize(st).cpu().numpy(), [2.5] * 3)


def test_reference_zero_covariance():
  # covariance_reducer_test.py:46-59
  red = emcmc.CovarianceReducer(event_ndims=1)
  st = red.initialize(torch.ones((9, 3)))
  for _ in range(2):
    st = red.one_step(torch.zeros((5, 9, 3), device='cuda'), st, axis=0)
  assert st.cov_state.num_samples == 10
  assert tuple(st.cov_state.mean.shape) == (9, 3)
  cov = red.finalize(st)
  assert tuple(cov.shape) == (9, 3, 3)
  assert not torch.any(cov != 0) and not torch.any(st.cov_state.mean != 0)


def test_rhat_reducer_int_samples():
  # potential_scale_reduction_reducer_test.py:33-45
  red = emcmc.PotentialScaleReductionReducer(independent_chain_ndims=1)
  st = red.initialize(torch.zeros((5, 3), dtype=torch.int64))
  chain_state = np.arange(60).reshape((4, 5, 3))
  for s in chain_state:
    st = red.one_step(torch.as_tensor(s, device='cuda'), st)
  rhat = red.finalize(st)
  assert rhat.dtype == torch.float64
  want = odiag.potential_scale_reduction(chain_state, 1, split_chains=False)
  np.testing.assert_allclose(rhat.cpu().numpy(), want, rtol=1e-6)


def _hmc(c=16, d=4):
  r = np.random.default_rng(2)
  x = r.normal(size=(200, d))
  y = (r.uniform(size=200) < 0.5).astype(np.float64)
  spec = pa.targets.LogisticRegression(x, y, 1.0, dtype=torch.float64)
  kern = pa.mcmc.HamiltonianMonteCarlo(spec, step_size=0.05,
                                       num_leapfrog_steps=4)
  q0 = torch.as_tensor(r.normal(size=(c, d)) * 0.1, device='cuda')
  return kern, q0


def test_with_reductions_over_hmc_matches_the_trace():
  kern, q = _hmc()
  reducers = {'mean': emcmc.ExpectationsReducer(),
              'cov': emcmc.CovarianceReducer(event_ndims=1, ddof=1),
              'rhat': emcmc.PotentialScaleReductionReducer()}
  wr = emcmc.WithReductions(kern, reducers)
  assert wr.is_calibrated
  pkr = wr.bootstrap_results(q)
  plain_q, plain_kr = q, kern.bootstrap_results(q)
  trace = []
  for t in range(30):
    seed = pa.random.sanitize_seed(500 + t)
    q, pkr = wr.one_step(q, pkr, seed=seed)
    plain_q, plain_kr = kern.one_step(plain_q, plain_kr, seed=seed)
    # the reducers do not disturb the chain
    assert torch.equal(q, plain_q)
    trace.append(q.cpu().numpy())
  tr = np.stack(trace)
  out = emcmc._reducer_map(lambda r, s: r.finalize(s), reducers,
                           pkr.reduction_results)
  np.testing.assert_allclose(out['mean'].cpu().numpy(), tr.mean(0),
                             rtol=1e-12, atol=1e-14)
  cov = out['cov'].cpu().numpy()
  for c in range(tr.shape[1]):
    np.testing.assert_allclose(cov[c], np.cov(tr[:, c].T, ddof=1),
                               rtol=1e-10, atol=1e-14)
  np.testing.assert_allclose(
      out['rhat'].cpu().numpy(),
      odiag.potential_scale_reduction(tr, 1, split_chains=False),
      rtol=1e-10)


def test_sample_fold_burnin_and_thinning():
  kern, q0 = _hmc(c=8, d=3)
  res, end, fkr = emcmc.sample_fold(
      5, q0, kernel=kern, reducer=emcmc.ExpectationsReducer(),
      num_burnin_steps=3, num_steps_between_results=2, seed=7)
  # the same inner steps by hand: seed of step t is fold_in(base, t); the
  # first kept draw is step 5 (3 burn-in + 2 thinned), then every third
  base = pa.random.sanitize_seed(7, salt='experimental.mcmc.sample_fold')
  q, kr = q0, kern.bootstrap_results(q0)
  kept = []
  for t in range(18):
    q, kr = kern.one_step(q, kr, seed=pa.random.fold_in(base, t))
    if t in (5, 8, 11, 14, 17):
      kept.append(q.cpu().numpy())
  assert torch.equal(end, q)
  np.testing.assert_allclose(res.cpu().numpy(), np.mean(kept, 0),
                             rtol=1e-12, atol=1e-14)
  _, _, _, states = emcmc.sample_fold(
      2, q0, kernel=kern, reducer=emcmc.VarianceReducer(), seed=3,
      return_final_reducer_states=True)
  assert states.cov_state.num_samples == 2


def test_sample_discarding_kernel_over_hmc():
  kern, q0 = _hmc(c=8, d=3)
  sdk = emcmc.SampleDiscardingKernel(kern, num_burnin_steps=2,
                                     num_steps_between_results=1)
  kr = sdk.bootstrap_results(q0)
  q, kr = sdk.one_step(q0, kr, seed=11)
  # the same inner steps by hand: 2 burn-in + 1 thinned + 1 kept
  want = emcmc.step_kernel(4, q0, kernel=kern, seed=11)
  assert torch.equal(q, want) and kr.call_counter == 1


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
@pytest.mark.parametrize('chunks', [None, [4, 25, 11]])
def test_running_covariance_even_event_many_batch_members(tdt, ndt, tol,
                                                          chunks):
  # shape (5, 6), event_ndims 1: five batch members, an even event size
  # (the kernel's paired 16-byte path with blockIdx.y over members)
  x = np.random.default_rng(4).normal(size=(40, 5, 6)).astype(ndt)
  x[..., 3] -= 0.7 * x[..., 2]
  rc = estats.RunningCovariance.from_shape((5, 6), tdt, 1)
  if chunks is None:
    for r in range(40):
      rc = rc.update(_device(x[r], tdt))
  else:
    at = 0
    for c in chunks:
      rc = rc.update(_device(x[at:at + c], tdt), axis=0)
      at += c
  m = _oracle(x.astype(np.float64), 1, 2, chunks)
  np.testing.assert_allclose(rc.covariance().cpu().numpy(),
                             ored.covariance(m), rtol=tol, atol=tol)
  np.testing.assert_allclose(rc.mean.cpu().numpy(), m.mean, rtol=tol,
                             atol=tol)


@pytest.mark.parametrize('offset', [0, 1])
def test_moments_update_in_place_and_unaligned(offset):
  # through the C ABI: an in-place update (out == in), and sums starting
  # one float64 past a 16-byte boundary (the scalar cross-product path)
  from probability_amd import _lib
  r = np.random.default_rng(6)
  nb, ev, n = 3, 4, 9
  x = r.normal(size=(n, nb, ev))
  dev = 'cuda'
  mean_buf = torch.zeros(nb * ev + offset, dtype=torch.float64, device=dev)
  ssr_buf = torch.zeros(nb * ev * ev + offset, dtype=torch.float64,
                        device=dev)
  mean, ssr = mean_buf[offset:], ssr_buf[offset:]
  xt = torch.as_tensor(x, device=dev).contiguous()
  _lib.check(_lib.lib().hipmc_moments_update(
      _lib.dtype_code(torch.float64), 2, nb, ev, n, _lib.ptr(xt), 0.0,
      _lib.ptr(mean), _lib.ptr(ssr), _lib.ptr(mean), _lib.ptr(ssr), 0))
  m = ored.Moments((nb, ev), 1, 2)
  m = ored.update(m, x, axis=0)
  np.testing.assert_allclose(mean.cpu().numpy().reshape(nb, ev), m.mean,
                             rtol=1e-12, atol=1e-14)
  np.testing.assert_allclose(ssr.cpu().numpy().reshape(nb, ev, ev), m.ssr,
                             rtol=1e-12, atol=1e-13)
